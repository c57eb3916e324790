// Host planner: CFL root selection and BFS order for the filter, and the matching plan (order,
// leaf set, device plan tables) for the enumeration engine.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <functional>
#include <memory>
#include <queue>

#include "engine.h"

namespace asdm {

// selectCFLRoot (relation/candidate_space.cpp:723-763): keep the three query vertices of smallest
// rank |V_L(u)| / deg(u) in a std::priority_queue (so equal ranks leave in the same order as in the
// reference), then pick the smallest |B(u)| / deg(u), first one on ties.
uint32_t cfl_root(const asdm_graph *q, const asdm_graph *g, const std::vector<uint32_t> &base_count) {
    auto cmp = [](const std::pair<uint32_t, double> &l, const std::pair<uint32_t, double> &r) { return l.second < r.second; };
    std::priority_queue<std::pair<uint32_t, double>, std::vector<std::pair<uint32_t, double>>, decltype(cmp)> pq(cmp);
    for (uint32_t u = 0; u < q->n; ++u)
        pq.emplace(u, (double)g->label_count(q->lab[u]) / (double)q->deg(u));
    while (pq.size() > 3) pq.pop();
    uint32_t root = 0;
    double best = (double)g->n;
    while (!pq.empty()) {
        uint32_t u = pq.top().first;
        double score = (double)base_count[u] / (double)q->deg(u);
        if (score < best) {
            root = u;
            best = score;
        }
        pq.pop();
    }
    return root;
}

// buildCFLBFSTree's vertex order (relation/candidate_space.cpp:765-816): BFS from the root with
// neighbours in query adjacency order.  Vertices unreachable from the root are not included.
std::vector<uint32_t> bfs_order(const asdm_graph *q, uint32_t root) {
    std::vector<uint32_t> order{root};
    std::vector<char> seen(q->n, 0);
    seen[root] = 1;
    for (size_t h = 0; h < order.size(); ++h) {
        uint32_t u = order[h];
        for (uint32_t a = q->off[u]; a < q->off[u + 1]; ++a) {
            uint32_t w = q->nbr[a];
            if (!seen[w]) {
                seen[w] = 1;
                order.push_back(w);
            }
        }
    }
    return order;
}

namespace {

bool adjacent(const asdm_cs *cs, uint32_t u, uint32_t w) { return cs->arc_index(u, w) != UINT32_MAX; }

// Is the subgraph induced by `keep` (bitmask) connected?
bool connected(const asdm_cs *cs, uint64_t keep) {
    if (!keep) return false;
    uint32_t start = (uint32_t)__builtin_ctzll(keep);
    uint64_t seen = 1ull << start, frontier = seen;
    while (frontier) {
        uint64_t next = 0;
        for (uint64_t f = frontier; f; f &= f - 1) {
            uint32_t u = (uint32_t)__builtin_ctzll(f);
            for (uint32_t a = cs->qoff[u]; a < cs->qoff[u + 1]; ++a) {
                uint32_t w = cs->qnbr[a];
                if ((keep >> w & 1) && !(seen >> w & 1)) next |= 1ull << w;
            }
        }
        seen |= next;
        frontier = next;
    }
    return seen == keep;
}

// Expected list length when extending to u from matched neighbour p: W(p -> u) / |C(p)|.
double fanout(const asdm_cs *cs, uint32_t p, uint32_t u) {
    uint32_t a = cs->arc_index(p, u);
    return (double)cs->arc_weight[a] / std::max<double>(1.0, cs->csize[p]);
}

void partitions(uint32_t k, std::vector<std::vector<uint32_t>> &out, std::vector<int64_t> &coef) {
    // restricted growth strings over k elements
    std::vector<uint32_t> a(k, 0);
    std::function<void(uint32_t, uint32_t)> rec = [&](uint32_t i, uint32_t m) {
        if (i == k) {
            std::vector<uint32_t> blocks(m, 0);
            for (uint32_t j = 0; j < k; ++j) blocks[a[j]] |= 1u << j;
            int64_t c = 1;
            for (uint32_t b : blocks) {
                int s = __builtin_popcount(b);
                int64_t f = 1;
                for (int t = 2; t < s; ++t) f *= t;  // (s-1)!
                c *= (s % 2 == 1 ? 1 : -1) * f;      // (-1)^(s-1) (s-1)!
            }
            out.push_back(blocks);
            coef.push_back(c);
            return;
        }
        for (uint32_t v = 0; v <= m; ++v) {
            a[i] = v;
            rec(i + 1, std::max(m, v + 1));
        }
    };
    rec(0, 0);
}

}  // namespace

// Estimated number of partial embeddings when u joins the matched set S (independence model):
// |C(u)| * prod over matched neighbours p of W(p->u) / (|C(p)| |C(u)|).
double growth(const asdm_cs *cs, uint64_t S, uint32_t u) {
    double g = cs->csize[u];
    for (uint32_t a = cs->qoff[u]; a < cs->qoff[u + 1]; ++a) {
        uint32_t p = cs->qnbr[a];
        if (S >> p & 1) g *= (double)cs->arc_weight[cs->arc_index(p, u)] / std::max(1.0, (double)cs->csize[p] * cs->csize[u]);
    }
    return g;
}

// Leaves: an independent set whose removal keeps the rest connected, preferring large expected
// fan-out (cheap to count, expensive to enumerate).
uint64_t choose_leaves(const asdm_cs *cs, uint32_t flags) {
    const uint32_t nq = cs->nq;
    const bool nofactor = flags & ASDM_PLAN_NO_FACTOR;
    const uint32_t max_group = (flags & ASDM_PLAN_LOCAL) ? 1 : kMaxGroup;
    const uint64_t all = nq == 64 ? ~0ull : ((1ull << nq) - 1);
    uint64_t leaves = 0;
    std::vector<std::pair<double, uint32_t>> cand;
    for (uint32_t u = 0; u < nq; ++u) {
        double f = 1e300;
        for (uint32_t a = cs->qoff[u]; a < cs->qoff[u + 1]; ++a) f = std::min(f, fanout(cs, cs->qnbr[a], u));
        cand.emplace_back(-f, u);
    }
    std::sort(cand.begin(), cand.end());
    // Pass 0 keeps only leaves that the fused last-level kernel counts with a few lookups -- one
    // query neighbour, or two adjacent ones (a triangle, counted from a per-candidate-edge table) --
    // with pairwise distinct labels, so no inclusion-exclusion pair intersections are needed per
    // partial embedding.  Pass 1 (taken when pass 0 admits only the last vertex) is the plain rule:
    // any independent set within the label-group limit.
    auto cheap_leaf = [&](uint32_t u) {
        const uint32_t d = cs->qoff[u + 1] - cs->qoff[u];
        return d == 1 || (d == 2 && adjacent(cs, cs->qnbr[cs->qoff[u]], cs->qnbr[cs->qoff[u] + 1]));
    };
    for (int pass = 0; pass < 2 && nq > 1; ++pass) {
        leaves = 0;
        for (auto &c : cand) {
            if (nofactor && leaves) break;
            uint32_t u = c.second;
            if (pass == 0 && !cheap_leaf(u)) continue;
            bool indep = true;
            for (uint32_t a = cs->qoff[u]; a < cs->qoff[u + 1]; ++a) indep &= !(leaves >> cs->qnbr[a] & 1);
            if (!indep) continue;
            uint32_t same = 1;
            for (uint64_t l = leaves; l; l &= l - 1) same += cs->qlab[__builtin_ctzll(l)] == cs->qlab[u];
            if (same > (pass == 0 ? 1u : max_group)) continue;
            uint64_t rest = all & ~(leaves | (1ull << u));
            if (!connected(cs, rest)) continue;
            leaves |= 1ull << u;
        }
        if (__builtin_popcountll(leaves) >= 2 || nofactor) break;
    }
    return leaves;
}

uint64_t frontier_of(const asdm_cs *cs, uint64_t S, uint64_t rest) {
    uint64_t f = 0;
    for (uint64_t r = rest & ~S; r; r &= r - 1) {
        uint32_t u = (uint32_t)__builtin_ctzll(r);
        for (uint32_t a = cs->qoff[u]; a < cs->qoff[u + 1]; ++a)
            if (S >> cs->qnbr[a] & 1) {
                f |= 1ull << u;
                break;
            }
    }
    return f;
}

uint32_t smallest_candidates(const asdm_cs *cs, uint64_t rest) {
    uint32_t first = UINT32_MAX;
    for (uint64_t r = rest; r; r &= r - 1) {
        uint32_t u = (uint32_t)__builtin_ctzll(r);
        if (first == UINT32_MAX || cs->csize[u] < cs->csize[first]) first = u;
    }
    return first;
}

int build_plan(const asdm_graph *q, const asdm_cs *cs, const uint32_t *given_order, uint32_t flags,
               const CardFn &card_of, asdm_plan *out, std::string &err) {
    (void)q;
    const uint32_t nq = cs->nq;
    const bool nofactor = flags & ASDM_PLAN_NO_FACTOR;
    const uint32_t max_group = (flags & ASDM_PLAN_LOCAL) ? 1 : kMaxGroup;
    std::vector<uint32_t> order;
    uint32_t D = 0;

    if (given_order) {
        order.assign(given_order, given_order + nq);
        std::vector<char> seen(nq, 0);
        for (uint32_t u : order) {
            if (u >= nq || seen[u]) {
                err = "order is not a permutation of the query vertices";
                return ASDM_E_INVALID;
            }
            seen[u] = 1;
        }
        // leaves: the longest suffix that is an independent set with all neighbours before it,
        // within the label-group limits
        D = nq;
        while (D > 1 && !(nofactor && D < nq)) {
            uint32_t u = order[D - 1];
            bool ok = true;
            for (uint32_t j = D; j < nq && ok; ++j) ok = !adjacent(cs, u, order[j]);
            uint32_t same = 1;
            for (uint32_t j = D; j < nq; ++j) same += cs->qlab[order[j]] == cs->qlab[u];
            if (!ok || same > max_group) break;
            --D;
        }
        if (D == nq) D = nq - 1 > 0 ? nq - 1 : 1;
        if (nq == 1) D = 1;
    } else if (flags & ASDM_PLAN_COST) {
        // multi-start cost-based search with the independence model's cardinalities
        std::vector<std::vector<uint32_t>> orders;
        std::vector<uint32_t> Ds;
        std::vector<std::unordered_map<uint64_t, double>> cards;
        BatchEstimator plain = [cs](const std::vector<std::pair<uint32_t, uint64_t>> &reqs, std::vector<double> &o) {
            o.resize(reqs.size());
            for (size_t k = 0; k < reqs.size(); ++k) o[k] = indep_card(cs, reqs[k].second);
            return ASDM_OK;
        };
        int rc = cost_orders({cs}, flags, true, plain, orders, Ds, cards);
        if (rc) return rc;
        order = orders[0];
        D = Ds[0];
    } else {
        const uint64_t all = nq == 64 ? ~0ull : ((1ull << nq) - 1);
        const uint64_t leaves = choose_leaves(cs, flags);
        // prefix order over the rest: smallest candidate set first, then greedy smallest
        // estimated cardinality among vertices adjacent to the matched set
        const uint64_t rest = all & ~leaves;
        const uint32_t first = smallest_candidates(cs, rest);
        order.push_back(first);
        uint64_t S = 1ull << first;
        double est = cs->csize[first];
        while (S != rest) {
            uint32_t best = UINT32_MAX;
            double bestg = 0;
            for (uint64_t r = frontier_of(cs, S, rest); r; r &= r - 1) {
                uint32_t u = (uint32_t)__builtin_ctzll(r);
                double gr = card_of(S | (1ull << u), est * growth(cs, S, u));
                if (best == UINT32_MAX || gr < bestg) {
                    best = u;
                    bestg = gr;
                }
            }
            order.push_back(best);
            S |= 1ull << best;
            est = bestg;
        }
        D = (uint32_t)order.size();
        for (uint64_t l = leaves; l; l &= l - 1) order.push_back((uint32_t)__builtin_ctzll(l));
    }
    return finish_plan(cs, std::move(order), D, flags, card_of, out, err);
}

int finish_plan(const asdm_cs *cs, std::vector<uint32_t> order, uint32_t D, uint32_t flags, const CardFn &card_of,
                asdm_plan *out, std::string &err) {
    const uint32_t nq = cs->nq;
    const bool local = flags & ASDM_PLAN_LOCAL;
    // check connectivity of the enumerated prefix (every depth >= 1 needs a backward neighbour)
    for (uint32_t d = 1; d < nq; ++d) {
        bool has = false;
        for (uint32_t j = 0; j < std::min(d, D); ++j) has |= adjacent(cs, order[j], order[d]);
        if (!has) {
            err = "matching order is not connected (depth " + std::to_string(d) + " has no earlier neighbour)";
            return ASDM_E_UNSUPPORTED;
        }
    }

    // estimated frontier sizes along the final order, and the prefix cardinalities the plan
    // relies on (its subsetToCard entries)
    {
        double est = cs->csize[order[0]], tot = est;
        uint64_t S = 1ull << order[0];
        out->prefix_cards.clear();
        out->prefix_cards.emplace_back(S, est);
        for (uint32_t d = 1; d < D; ++d) {
            est = card_of(S | (1ull << order[d]), est * growth(cs, S, order[d]));
            S |= 1ull << order[d];
            tot += est;
            out->prefix_cards.emplace_back(S, est);
        }
        out->est_leaf = est;
        out->est_total = tot;
    }

    DPlan &P = out->h;
    std::memset(&P, 0, sizeof(P));
    P.nq = nq;
    P.D = D;
    for (uint32_t d = 0; d < nq; ++d) {
        uint32_t u = order[d];
        P.u[d] = u;
        P.coff[d] = cs->cand_off[u];
        P.csz[d] = cs->csize[u];
        for (uint32_t j = 0; j < std::min(d, D); ++j) {
            uint32_t a = cs->arc_index(order[j], u);
            if (a != UINT32_MAX) {
                P.bdep[d][P.nb[d]] = j;
                P.bbase[d][P.nb[d]] = cs->ce_base[a];
                P.nb[d]++;
            }
            if (cs->qlab[order[j]] == cs->qlab[u]) P.sdep[d][P.ns[d]++] = j;
        }
    }
    // leaf groups by label
    std::vector<char> done(nq, 0);
    for (uint32_t d = D; d < nq; ++d) {
        if (done[d]) continue;
        uint32_t g = P.ngroups++;
        for (uint32_t e = d; e < nq; ++e)
            if (!done[e] && cs->qlab[order[e]] == cs->qlab[order[d]]) {
                if (P.gsize[g] >= kMaxGroup) {
                    err = "leaf label group larger than supported";
                    return ASDM_E_UNSUPPORTED;
                }
                P.gleaf[g][P.gsize[g]++] = e;
                done[e] = 1;
            }
    }
    // On a data graph without self-loops a vertex adjacent to v_b is never v_b itself, so a
    // same-label depth that is a backward neighbour of an enumerated vertex (of every leaf of a
    // group, for a leaf group) needs no injectivity test against it.
    if (cs->loop_free) {
        auto is_back = [&](uint32_t e, uint32_t j) {
            for (uint32_t b = 0; b < P.nb[e]; ++b)
                if (P.bdep[e][b] == j) return true;
            return false;
        };
        auto prune = [&](uint32_t e, const std::function<bool(uint32_t)> &skip) {
            uint32_t k = 0;
            for (uint32_t s = 0; s < P.ns[e]; ++s)
                if (!skip(P.sdep[e][s])) P.sdep[e][k++] = P.sdep[e][s];
            P.ns[e] = k;
        };
        for (uint32_t d = 0; d < D; ++d) prune(d, [&](uint32_t j) { return is_back(d, j); });
        for (uint32_t g = 0; g < P.ngroups; ++g) {
            auto common = [&](uint32_t j) {
                for (uint32_t i = 0; i < P.gsize[g]; ++i)
                    if (!is_back(P.gleaf[g][i], j)) return false;
                return true;
            };
            for (uint32_t i = 0; i < P.gsize[g]; ++i) prune(P.gleaf[g][i], common);
        }
    }
    for (uint32_t e = D; e < nq; ++e)
        for (uint32_t j = 0; j < P.nb[e]; ++j)
            if (P.bdep[e][j] + 1 == D) {
                P.nadj_last++;
                break;
            }
    for (uint32_t e = D; e < nq; ++e)
        if (P.nb[e] == 2) {
            const uint32_t a = cs->arc_index(order[P.bdep[e][0]], order[P.bdep[e][1]]);
            if (a != UINT32_MAX) {
                P.tri[e] = 1;
                P.tri_base[e] = cs->ce_base[a];
            }
        }
    for (uint32_t k = 2; k <= kMaxGroup; ++k) {
        std::vector<std::vector<uint32_t>> parts;
        std::vector<int64_t> coef;
        partitions(k, parts, coef);
        P.nparts[k] = (uint32_t)parts.size();
        for (size_t i = 0; i < parts.size(); ++i) {
            P.pnblk[k][i] = (uint32_t)parts[i].size();
            for (size_t b = 0; b < parts[i].size(); ++b) P.pblk[k][i][b] = parts[i][b];
            P.pcoef[k][i] = coef[i];
        }
    }
    P.bits = cs->d_bits;
    P.rank = cs->d_rank;
    P.W = cs->W;
    out->order = order;
    out->D = D;
    out->local = local;
    return ASDM_OK;
}

// The automatic planner considers the bag-table engine when the frontier plan's estimated rows
// exceed kAutoBagRows or it enumerates at least kAutoBagDepth vertices (the independence model
// underestimates deep plans on power-law graphs by orders of magnitude, while the bag plan's cost
// is known exactly from its table sizes), and takes it when its cost (element operations) is at
// most kAutoBagCost.
constexpr double kAutoBagRows = 16777216.0;
constexpr uint32_t kAutoBagDepth = 8;
constexpr double kAutoBagCost = 1.0e14;

// Attaches a bag plan when flags ask for one (ASDM_PLAN_BAGS), or when the frontier plan is large
// and the bag plan is cheap (automatic choice, unless ASDM_PLAN_FRONTIER or an explicit order).
int attach_bags(const asdm_cs *cs, uint32_t flags, bool explicit_order, asdm_plan *p, std::string &err) {
    const bool local = flags & ASDM_PLAN_LOCAL;
    if (flags & ASDM_PLAN_BAGS) {
        auto b = std::make_shared<BagPlan>();
        std::string why;
        if (!build_bag_plan(cs, local, *b, why)) {
            err = "bag-table plan: " + why;
            return ASDM_E_UNSUPPORTED;
        }
        p->bags = b;
        return ASDM_OK;
    }
    if (explicit_order || (flags & (ASDM_PLAN_FRONTIER | ASDM_PLAN_NO_FACTOR)) ||
        (p->est_total <= kAutoBagRows && p->D < kAutoBagDepth))
        return ASDM_OK;
    auto b = std::make_shared<BagPlan>();
    std::string why;
    if (build_bag_plan(cs, local, *b, why) && b->cost <= kAutoBagCost) p->bags = b;
    return ASDM_OK;
}

}  // namespace asdm

using namespace asdm;

extern "C" {

int asdm_plan_create_with_cards(const asdm_graph *q, const asdm_cs *cs, const uint64_t *masks, const double *cards,
                                uint32_t ncards, uint32_t flags, asdm_plan **out) try {
    if (!q || !cs || !out || q->n != cs->nq) return ASDM_E_INVALID;
    *out = nullptr;
    auto t0 = std::chrono::steady_clock::now();
    auto p = std::make_unique<asdm_plan>();
    std::string err;
    if (ncards && (!masks || !cards)) return ASDM_E_INVALID;
    auto card_of = [&](uint64_t mask, double fallback) {
        for (uint32_t i = 0; i < ncards; ++i)
            if (masks[i] == mask) return cards[i];
        return fallback;
    };
    int rc = build_plan(q, cs, nullptr, flags, card_of, p.get(), err);
    if (!rc) rc = attach_bags(cs, flags, false, p.get(), err);
    if (rc) {
        cs->ctx->err = err;
        return rc;
    }
    p->plan_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    *out = p.release();
    return ASDM_OK;
} ASDM_HOST_CATCH

int asdm_plan_create(const asdm_graph *q, const asdm_cs *cs, const uint32_t *order, uint32_t flags, asdm_plan **out) try {
    if (!q || !cs || !out || q->n != cs->nq) return ASDM_E_INVALID;
    *out = nullptr;
    auto t0 = std::chrono::steady_clock::now();
    auto p = std::make_unique<asdm_plan>();
    std::string err;
    // an explicit order with cartesian steps is planned on cs extended with cartesian arcs
    const asdm_cs *pcs = cs;
    if (order) {
        for (uint32_t i = 0; i < cs->nq; ++i)
            if (order[i] >= cs->nq) return cs->ctx->fail(ASDM_E_INVALID, "order is not a permutation of the query vertices");
        asdm_cs *ext = nullptr;
        if (int rc = cartesian_cs(cs, std::vector<uint32_t>(order, order + cs->nq), &ext, err)) {
            cs->ctx->err = err;
            return rc;
        }
        if (ext) {
            p->cart_cs.reset(ext, asdm_cs_free);
            pcs = ext;
        }
    }
    int rc = build_plan(q, pcs, order, flags, [](uint64_t, double fallback) { return fallback; }, p.get(), err);
    if (!rc) rc = attach_bags(pcs, flags, order != nullptr, p.get(), err);
    if (rc) {
        cs->ctx->err = err;
        return rc;
    }
    p->plan_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    *out = p.release();
    return ASDM_OK;
} ASDM_HOST_CATCH

int asdm_plan_create_ordered(const asdm_graph *q, const asdm_cs *cs, const uint32_t *order, uint32_t num_enumerated,
                             uint32_t flags, asdm_plan **out) try {
    if (!q || !cs || !order || !out || q->n != cs->nq || num_enumerated < 1 || num_enumerated > cs->nq)
        return ASDM_E_INVALID;
    *out = nullptr;
    const uint32_t nq = cs->nq, D = num_enumerated;
    const uint32_t max_group = (flags & ASDM_PLAN_LOCAL) ? 1 : kMaxGroup;
    std::vector<uint32_t> o(order, order + nq);
    std::vector<char> seen(nq, 0);
    for (uint32_t u : o) {
        if (u >= nq || seen[u]) return cs->ctx->fail(ASDM_E_INVALID, "order is not a permutation of the query vertices");
        seen[u] = 1;
    }
    // the leaves (depths D..nq-1) must be pairwise non-adjacent, within the label-group limit
    for (uint32_t i = D; i < nq; ++i) {
        uint32_t same = 0;
        for (uint32_t j = D; j < nq; ++j) {
            if (j != i && adjacent(cs, o[i], o[j]))
                return cs->ctx->fail(ASDM_E_UNSUPPORTED, "leaves of the order are adjacent");
            same += cs->qlab[o[j]] == cs->qlab[o[i]];
        }
        if (same > max_group) return cs->ctx->fail(ASDM_E_UNSUPPORTED, "leaf label group larger than supported");
    }
    if (D == nq && nq > 1) return cs->ctx->fail(ASDM_E_INVALID, "at least one vertex must be a leaf");
    auto t0 = std::chrono::steady_clock::now();
    auto p = std::make_unique<asdm_plan>();
    std::string err;
    // cartesian steps (depths with no earlier neighbour) get cartesian arcs in an extended space
    const asdm_cs *pcs = cs;
    asdm_cs *ext = nullptr;
    if (int rc0 = cartesian_cs(cs, o, &ext, err)) {
        cs->ctx->err = err;
        return rc0;
    }
    if (ext) {
        p->cart_cs.reset(ext, asdm_cs_free);
        pcs = ext;
    }
    int rc = finish_plan(pcs, o, D, flags, [](uint64_t, double fb) { return fb; }, p.get(), err);
    if (rc) {
        cs->ctx->err = err;
        return rc;
    }
    p->plan_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    *out = p.release();
    return ASDM_OK;
} ASDM_HOST_CATCH

void asdm_plan_free(asdm_plan *p) { delete p; }

int asdm_plan_uses_bags(const asdm_plan *p) { return p && p->bags ? 1 : 0; }

int asdm_plan_bag_info(const asdm_plan *p, uint32_t *terms, double *cost, double *peak_entries) try {
    if (!p) return ASDM_E_INVALID;
    if (terms) *terms = p->bags ? (uint32_t)p->bags->terms.size() : 0;
    if (cost) *cost = p->bags ? p->bags->cost : 0;
    if (peak_entries) *peak_entries = p->bags ? p->bags->peak_table : 0;
    return ASDM_OK;
} ASDM_HOST_CATCH

int asdm_plan_cards(const asdm_plan *p, uint64_t *masks, double *cards, uint32_t cap, uint32_t *n) try {
    if (!p || !n) return ASDM_E_INVALID;
    *n = (uint32_t)p->prefix_cards.size();
    if (!masks || !cards) return ASDM_OK;
    if (cap < *n) return ASDM_E_INVALID;
    for (uint32_t i = 0; i < *n; ++i) {
        masks[i] = p->prefix_cards[i].first;
        cards[i] = p->prefix_cards[i].second;
    }
    return ASDM_OK;
} ASDM_HOST_CATCH

int asdm_plan_estimate(const asdm_plan *p, double *rows_leaf, double *rows_total) try {
    if (!p) return ASDM_E_INVALID;
    if (rows_leaf) *rows_leaf = p->est_leaf;
    if (rows_total) *rows_total = p->est_total;
    return ASDM_OK;
} ASDM_HOST_CATCH

int asdm_plan_order(const asdm_plan *p, uint32_t *order, uint32_t *num_enumerated) try {
    if (!p) return ASDM_E_INVALID;
    if (order) std::memcpy(order, p->order.data(), sizeof(uint32_t) * p->order.size());
    if (num_enumerated) *num_enumerated = p->D;
    return ASDM_OK;
} ASDM_HOST_CATCH

}  // extern "C"
