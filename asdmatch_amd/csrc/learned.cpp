// Learned planner: features of (sub-)query graphs for the message-passing estimator, batched
// cardinality estimates, and the greedy matching order driven by them.
//
// The reference fills subsetToCard (plan/estimator.cpp:7) with sampling estimates for the
// optimizer (plan/estimator.cpp:173-291) and orders by them.  Here the same role -- a cardinality
// per vertex subset, consumed by the greedy prefix order -- is played by the network.  The greedy
// runs in lock step over a whole batch of queries, so each step is one forward launch for all
// queries' candidate subsets.  Whatever order it picks, the engine's count is exact.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <memory>
#include <unordered_map>
#include <unordered_set>

#include "engine.h"

namespace asdm {
namespace {

constexpr uint32_t kFeat = 8;

struct FeatBatch {
    std::vector<uint32_t> node_off{0};
    std::vector<float> x;
    std::vector<uint32_t> adj_off{0};
    std::vector<uint32_t> adj;
};

inline float lg(double v) { return (float)std::log2(1.0 + std::max(0.0, v)); }

// Node features of the sub-query induced by `mask` on the candidate space `cs`:
//   0 log2(1+|C(u)|)/16            candidate-set size
//   1 log2(1+|V_L(u)|)/16          label frequency in the data graph
//   2 log2(1+mean deg of L(u))/8   label degree in the data graph
//   3 deg_S(u)/8                   degree inside the sub-query
//   4 #same-label vertices in S/8  (injectivity pressure)
//   5 mean over S-neighbours w of log2(1 + W(w->u)/|C(w)|)/8   expected fan-out into u
//   6 |S|/32, 7 |E(S)|/64          size of the sub-query
void featurize(const asdm_graph *g, const asdm_cs *cs, uint64_t mask, FeatBatch &fb) {
    uint32_t local[64];
    uint32_t k = 0;
    for (uint64_t r = mask; r; r &= r - 1) local[__builtin_ctzll(r)] = k++;
    const uint32_t base = fb.node_off.back();
    uint32_t edges2 = 0;
    for (uint64_t r = mask; r; r &= r - 1) {
        const uint32_t u = (uint32_t)__builtin_ctzll(r);
        for (uint32_t a = cs->qoff[u]; a < cs->qoff[u + 1]; ++a) edges2 += (mask >> cs->qnbr[a]) & 1;
    }
    for (uint64_t r = mask; r; r &= r - 1) {
        const uint32_t u = (uint32_t)__builtin_ctzll(r);
        const uint32_t l = cs->qlab[u];
        uint32_t deg = 0, same = 0;
        double fo = 0;
        for (uint32_t a = cs->qoff[u]; a < cs->qoff[u + 1]; ++a) {
            const uint32_t w = cs->qnbr[a];
            if (!(mask >> w & 1)) continue;
            ++deg;
            const uint32_t aw = cs->arc_index(w, u);
            fo += lg((double)cs->arc_weight[aw] / std::max(1.0, (double)cs->csize[w]));
            fb.adj.push_back(base + local[w]);
        }
        for (uint64_t r2 = mask; r2; r2 &= r2 - 1) same += cs->qlab[__builtin_ctzll(r2)] == l;
        const float f[kFeat] = {lg(cs->csize[u]) / 16.f,
                                lg(g->label_count(l)) / 16.f,
                                l < g->lab_avgdeg.size() ? lg(g->lab_avgdeg[l]) / 8.f : 0.f,
                                deg / 8.f,
                                same / 8.f,
                                deg ? (float)(fo / deg) / 8.f : 0.f,
                                k / 32.f,
                                (edges2 / 2) / 64.f};
        fb.x.insert(fb.x.end(), f, f + kFeat);
        fb.adj_off.push_back((uint32_t)fb.adj.size());
    }
    fb.node_off.push_back(base + k);
}

int forward(asdm_mpnn *m, const FeatBatch &fb, std::vector<float> &y) {
    const uint32_t B = (uint32_t)fb.node_off.size() - 1;
    y.assign(B, 0.f);
    if (!B) return ASDM_OK;
    return asdm_mpnn_forward(m, B, fb.node_off.data(), fb.x.data(), fb.adj_off.data(),
                             fb.adj.empty() ? nullptr : fb.adj.data(), y.data());
}

inline double card_from_log(float y) { return std::exp2(std::min(64.f, std::max(0.f, y))) - 1.0; }

}  // namespace

// Cost-based order search over a batch of candidate spaces.  The leaf set is the plain planner's.
// The cost of an order is its summed prefix cardinalities: the estimated partial embeddings the
// engine materializes.  With multi_start, a query whose enumerated part has at most kDpVertices
// vertices gets the optimal connected order by a dynamic program over its connected subsets (the
// subset DP of plan/optimizer.cpp:380-548, over the subsets the engine can enumerate): one
// estimator call holds every such subset of every such query, then
//   cost(S) = card(S) + min over v in S, S \ {v} connected and adjacent to v, of cost(S \ {v}).
// Other queries grow greedy orders from the plain planner's first vertex (the smallest candidate
// set) or, with multi_start, from each vertex of the enumerated part: each step adds the frontier
// vertex whose extended prefix has the smallest estimated cardinality, and the cheapest order wins.
// All greedy runs of all queries advance in lock step, with one estimator call per step for every
// subset they ask about.  Singletons use the exact |C(u)|.
constexpr uint32_t kDpVertices = 12;

// The optimal order of rest by the subset DP above; cards holds every connected subset's estimate.
std::vector<uint32_t> dp_order(const asdm_cs *cs, uint64_t rest, const std::unordered_map<uint64_t, double> &cards,
                               const std::vector<uint64_t> &subsets) {
    std::unordered_map<uint64_t, std::pair<double, uint32_t>> best;  // cost, last vertex
    for (uint64_t r = rest; r; r &= r - 1) {
        const uint32_t u = (uint32_t)__builtin_ctzll(r);
        best[1ull << u] = {(double)cs->csize[u], u};
    }
    for (uint64_t S : subsets) {  // by increasing size
        std::pair<double, uint32_t> b{0.0, UINT32_MAX};
        for (uint64_t r = S; r; r &= r - 1) {
            const uint32_t v = (uint32_t)__builtin_ctzll(r);
            const uint64_t P = S & ~(1ull << v);
            const auto it = best.find(P);
            if (it == best.end() || !(frontier_of(cs, P, rest) >> v & 1)) continue;
            const double c = it->second.first + cards.at(S);
            if (b.second == UINT32_MAX || c < b.first) b = {c, v};
        }
        best[S] = b;
    }
    std::vector<uint32_t> order;
    for (uint64_t S = rest; S;) {
        const uint32_t v = best.at(S).second;
        order.push_back(v);
        S &= ~(1ull << v);
    }
    std::reverse(order.begin(), order.end());
    return order;
}
int cost_orders(const std::vector<const asdm_cs *> &css, uint32_t flags, bool multi_start, const BatchEstimator &est,
                std::vector<std::vector<uint32_t>> &orders, std::vector<uint32_t> &Ds,
                std::vector<std::unordered_map<uint64_t, double>> &cards) {
    const size_t Q = css.size();
    orders.assign(Q, {});
    Ds.assign(Q, 0);
    cards.assign(Q, {});
    struct Run {
        uint32_t q;
        uint64_t S;
        std::vector<uint32_t> order;
        double cost;
    };
    std::vector<Run> runs;
    std::vector<uint64_t> rest(Q), leaves(Q);
    std::vector<std::vector<uint64_t>> dp_sets(Q);  // connected subsets (size >= 2) by size, DP queries
    std::vector<char> dp(Q, 0);
    std::vector<std::pair<uint32_t, uint64_t>> dp_reqs;
    for (size_t i = 0; i < Q; ++i) {
        const asdm_cs *cs = css[i];
        const uint64_t all = cs->nq == 64 ? ~0ull : ((1ull << cs->nq) - 1);
        leaves[i] = choose_leaves(cs, flags);
        rest[i] = all & ~leaves[i];
        for (uint64_t r = rest[i]; r; r &= r - 1) cards[i][r & -r] = cs->csize[__builtin_ctzll(r)];
        if (multi_start && __builtin_popcountll(rest[i]) <= kDpVertices) {
            dp[i] = 1;
            std::unordered_set<uint64_t> seen;
            std::vector<uint64_t> cur;
            for (uint64_t r = rest[i]; r; r &= r - 1) cur.push_back(r & -r);
            while (!cur.empty()) {
                std::vector<uint64_t> nxt;
                for (uint64_t S : cur)
                    for (uint64_t f = frontier_of(cs, S, rest[i]); f; f &= f - 1) {
                        const uint64_t T = S | (f & -f);
                        if (seen.insert(T).second) nxt.push_back(T);
                    }
                for (uint64_t T : nxt) {
                    dp_sets[i].push_back(T);
                    dp_reqs.emplace_back((uint32_t)i, T);
                }
                cur.swap(nxt);
            }
            continue;
        }
        const uint64_t starts = multi_start ? rest[i] : (1ull << smallest_candidates(cs, rest[i]));
        for (uint64_t r = starts; r; r &= r - 1) {
            const uint32_t u = (uint32_t)__builtin_ctzll(r);
            runs.push_back({(uint32_t)i, 1ull << u, {u}, (double)cs->csize[u]});
            cards[i][1ull << u] = cs->csize[u];
        }
    }
    std::vector<std::pair<uint32_t, uint64_t>> reqs;  // (query, subset) asked this step
    std::vector<std::pair<size_t, uint32_t>> who;     // (run, vertex) of each request
    std::vector<double> y;
    if (!dp_reqs.empty()) {
        if (int rc = est(dp_reqs, y)) return rc;
        for (size_t k = 0; k < dp_reqs.size(); ++k) cards[dp_reqs[k].first][dp_reqs[k].second] = y[k];
    }
    for (size_t i = 0; i < Q; ++i)
        if (dp[i]) orders[i] = dp_order(css[i], rest[i], cards[i], dp_sets[i]);
    for (;;) {
        reqs.clear();
        who.clear();
        for (size_t k = 0; k < runs.size(); ++k) {
            Run &r = runs[k];
            if (r.S == rest[r.q]) continue;
            for (uint64_t f = frontier_of(css[r.q], r.S, rest[r.q]); f; f &= f - 1) {
                const uint32_t u = (uint32_t)__builtin_ctzll(f);
                reqs.emplace_back(r.q, r.S | (1ull << u));
                who.emplace_back(k, u);
            }
        }
        if (reqs.empty()) break;
        if (int rc = est(reqs, y)) return rc;
        std::vector<uint32_t> best(runs.size(), UINT32_MAX);
        std::vector<double> bestc(runs.size(), 0);
        for (size_t k = 0; k < who.size(); ++k) {
            const size_t ri = who[k].first;
            const double c = y[k];
            cards[reqs[k].first][reqs[k].second] = c;
            if (best[ri] == UINT32_MAX || c < bestc[ri]) {
                best[ri] = who[k].second;
                bestc[ri] = c;
            }
        }
        for (size_t k = 0; k < runs.size(); ++k) {
            if (best[k] == UINT32_MAX) continue;
            runs[k].order.push_back(best[k]);
            runs[k].S |= 1ull << best[k];
            runs[k].cost += bestc[k];
        }
    }
    std::vector<double> bestcost(Q, 0);
    std::vector<char> have(Q, 0);
    for (const Run &r : runs)  // first run (smallest start vertex) on ties
        if (!dp[r.q] && (!have[r.q] || r.cost < bestcost[r.q])) {
            have[r.q] = 1;
            bestcost[r.q] = r.cost;
            orders[r.q] = r.order;
        }
    for (size_t i = 0; i < Q; ++i) {
        Ds[i] = (uint32_t)orders[i].size();
        for (uint64_t l = leaves[i]; l; l &= l - 1) orders[i].push_back((uint32_t)__builtin_ctzll(l));
    }
    return ASDM_OK;
}

// The independence model's cardinality of the sub-query induced by mask (the plain planner's
// estimate): prod_u |C(u)| * prod over edges (u, w) of W(u->w) / (|C(u)| |C(w)|).
double indep_card(const asdm_cs *cs, uint64_t mask) {
    double e = 1.0;
    uint64_t S = 0;
    for (uint64_t r = mask; r; r &= r - 1) {
        const uint32_t u = (uint32_t)__builtin_ctzll(r);
        e *= growth(cs, S, u);
        S |= 1ull << u;
    }
    return e;
}

// Plans for a batch of candidate spaces from the learned estimates; a query whose learned plan
// cannot be built (or every query, when the network cannot run) gets the plain plan.
int learned_plans(asdm_mpnn *m, const asdm_graph *g, const std::vector<const asdm_cs *> &css, uint32_t flags,
                  std::vector<asdm_plan *> &out) {
    const size_t Q = css.size();
    out.assign(Q, nullptr);
    auto t0 = std::chrono::steady_clock::now();
    std::vector<std::vector<uint32_t>> orders;
    std::vector<uint32_t> Ds;
    std::vector<std::unordered_map<uint64_t, double>> cards;
    BatchEstimator learned = [&](const std::vector<std::pair<uint32_t, uint64_t>> &reqs, std::vector<double> &out) {
        FeatBatch fb;
        for (const auto &r : reqs) featurize(g, css[r.first], r.second, fb);
        std::vector<float> y;
        if (int rc = forward(m, fb, y)) return rc;
        out.resize(reqs.size());
        for (size_t k = 0; k < reqs.size(); ++k) out[k] = card_from_log(y[k]);
        return ASDM_OK;
    };
    const bool have = m && cost_orders(css, flags, (flags & ASDM_PLAN_COST) != 0, learned, orders, Ds, cards) == ASDM_OK;
    const double share = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() /
                         std::max<size_t>(Q, 1);
    for (size_t i = 0; i < Q; ++i) {
        auto t1 = std::chrono::steady_clock::now();
        auto p = std::make_unique<asdm_plan>();
        std::string err;
        int rc = ASDM_E_INVALID;
        if (have) {
            const auto &cm = cards[i];
            rc = finish_plan(
                css[i], orders[i], Ds[i], flags,
                [&cm](uint64_t mask, double fallback) {
                    auto it = cm.find(mask);
                    return it == cm.end() ? fallback : it->second;
                },
                p.get(), err);
        }
        if (rc) {
            p = std::make_unique<asdm_plan>();
            rc = build_plan(nullptr, css[i], nullptr, flags, [](uint64_t, double fb) { return fb; }, p.get(), err);
        }
        if (rc) {
            for (auto &q : out) {
                asdm_plan_free(q);
                q = nullptr;
            }
            css[i]->ctx->err = err;
            return rc;
        }
        p->plan_ms = share + std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t1).count();
        out[i] = p.release();
    }
    return ASDM_OK;
}

}  // namespace asdm

using namespace asdm;

extern "C" {

int asdm_mpnn_features(const asdm_graph *g, const asdm_cs *cs, const uint64_t *masks, uint32_t n, uint32_t *node_off,
                       float *x, uint32_t *adj_off, uint32_t *adj, uint32_t *num_nodes, uint32_t *num_adj) try {
    if (!g || !cs || (!masks && n) || !num_nodes || !num_adj) return ASDM_E_INVALID;
    FeatBatch fb;
    for (uint32_t i = 0; i < n; ++i) {
        const uint64_t valid = cs->nq == 64 ? ~0ull : ((1ull << cs->nq) - 1);
        if (!masks[i] || (masks[i] & ~valid) || __builtin_popcountll(masks[i]) > 32) return ASDM_E_INVALID;
        featurize(g, cs, masks[i], fb);
    }
    const uint32_t N = fb.node_off.back(), E = (uint32_t)fb.adj.size();
    if (node_off && x && adj_off && (adj || !E)) {
        if (*num_nodes < N || *num_adj < E) return ASDM_E_INVALID;
        std::copy(fb.node_off.begin(), fb.node_off.end(), node_off);
        std::copy(fb.x.begin(), fb.x.end(), x);
        std::copy(fb.adj_off.begin(), fb.adj_off.end(), adj_off);
        std::copy(fb.adj.begin(), fb.adj.end(), adj);
    }
    *num_nodes = N;
    *num_adj = E;
    return ASDM_OK;
} ASDM_HOST_CATCH

int asdm_mpnn_estimate(asdm_mpnn *m, const asdm_graph *g, const asdm_cs *cs, const uint64_t *masks, uint32_t n,
                       double *cards) try {
    if (!m || !g || !cs || (!masks && n) || (!cards && n)) return ASDM_E_INVALID;
    FeatBatch fb;
    const uint64_t valid = cs->nq == 64 ? ~0ull : ((1ull << cs->nq) - 1);
    for (uint32_t i = 0; i < n; ++i) {
        if (!masks[i] || (masks[i] & ~valid)) return ASDM_E_INVALID;
        featurize(g, cs, masks[i], fb);
    }
    std::vector<float> y;
    if (int rc = forward(m, fb, y)) return rc;
    for (uint32_t i = 0; i < n; ++i) cards[i] = card_from_log(y[i]);
    return ASDM_OK;
} ASDM_HOST_CATCH

int asdm_plain_estimate(const asdm_cs *cs, const uint64_t *masks, uint32_t n, double *cards) try {
    if (!cs || (n && (!masks || !cards))) return ASDM_E_INVALID;
    const uint64_t valid = cs->nq == 64 ? ~0ull : ((1ull << cs->nq) - 1);
    for (uint32_t i = 0; i < n; ++i) {
        if (!masks[i] || (masks[i] & ~valid)) return ASDM_E_INVALID;
        double est = 1.0;
        uint64_t S = 0;
        for (uint64_t r = masks[i]; r; r &= r - 1) {
            const uint32_t u = (uint32_t)__builtin_ctzll(r);
            est *= growth(cs, S, u);
            S |= 1ull << u;
        }
        cards[i] = est;
    }
    return ASDM_OK;
} ASDM_HOST_CATCH

int asdm_plan_create_learned(asdm_mpnn *m, const asdm_graph *g, const asdm_graph *q, const asdm_cs *cs,
                             uint32_t flags, asdm_plan **out) try {
    if (!m || !g || !q || !cs || !out || q->n != cs->nq) return ASDM_E_INVALID;
    std::vector<asdm_plan *> plans;
    int rc = learned_plans(m, g, {cs}, flags, plans);
    *out = rc ? nullptr : plans[0];
    return rc;
} ASDM_HOST_CATCH

}  // extern "C"
