// Host planner of the bag-table engine: partition terms, quotient graphs and width-two
// elimination orders.
#include <algorithm>
#include <cmath>
#include <functional>
#include <map>
#include <set>

#include "bags.h"
#include "engine.h"

namespace asdm {

namespace {

// Partitions of `elems` into blocks of pairwise non-adjacent query vertices (restricted growth).
void independent_partitions(const asdm_cs *cs, const std::vector<uint32_t> &elems,
                            std::vector<std::vector<std::vector<uint32_t>>> &out, uint64_t cap, bool &overflow) {
    std::vector<std::vector<uint32_t>> cur;
    std::function<void(size_t)> rec = [&](size_t i) {
        if (overflow) return;
        if (i == elems.size()) {
            out.push_back(cur);
            if (out.size() > cap) overflow = true;
            return;
        }
        const uint32_t u = elems[i];
        for (size_t b = 0; b < cur.size(); ++b) {
            bool ok = true;
            for (uint32_t w : cur[b]) ok &= cs->arc_index(u, w) == UINT32_MAX;
            if (!ok) continue;
            cur[b].push_back(u);
            rec(i + 1);
            cur[b].pop_back();
        }
        cur.push_back({u});
        rec(i + 1);
        cur.pop_back();
    };
    rec(0);
}

int64_t mobius(const std::vector<std::vector<uint32_t>> &blocks) {
    int64_t c = 1;
    for (auto &b : blocks) {
        const int64_t s = (int64_t)b.size();
        int64_t f = 1;
        for (int64_t t = 2; t < s; ++t) f *= t;
        c *= (s % 2 == 1 ? 1 : -1) * f;
    }
    return c;
}

using Sig = std::vector<uint32_t>;

std::vector<uint32_t> merged(std::vector<uint32_t> a, const std::vector<uint32_t> &b) {
    a.insert(a.end(), b.begin(), b.end());
    std::sort(a.begin(), a.end());
    return a;
}

// The key under which the engine caches a factor result for the rest of a query: its kind (1 edge
// weights, 2 dense table, 3 vector), the member masks of its two blocks (0 for a vector's second)
// and the sorted member masks of the blocks eliminated into it.  The result depends on nothing
// else, so a step whose key an earlier term or marginal of the query produced is a lookup.
Sig factor_sig(uint32_t kind, uint32_t ma, uint32_t mb, const std::vector<uint32_t> &S) {
    Sig k{kind, std::min(ma, mb), std::max(ma, mb)};
    k.insert(k.end(), S.begin(), S.end());
    return k;
}

// Greedy elimination of the quotient graph with tables of at most two variables.  Each step
// eliminates a live block (never `keep`) of live degree <= 2, the cheapest first: degree 0 and 1
// cost a pass over a vector or a table.  Degree 2 joins the two factors of v into one factor on
// its neighbours (a, b): when (a, b) already carry an edge factor, the product is only needed on
// that factor's edges, so the step costs per edge and the factor stays sparse; otherwise it
// writes a |D(a)| x |D(b)| table (and multiplies two dense tables when both incident factors are
// dense).  nnz holds the estimated entries of each edge relation.
// new_table_weight scales the cost of a step that starts a new dense table from two edge
// relations: tables started in several places of one cycle later meet at a vertex whose both
// factors are dense, a dense x dense product, so a larger weight keeps extending one table.
// A step whose factor key is in `seen` (if given: made by the query's earlier eliminations) costs only the
// pass that applies the cached result; `made` receives the keys this order produces.  `cost` adds
// the order's modelled work with those lookups, `full` the same work with none.
bool eliminate_once(uint32_t nb, const std::vector<std::pair<uint32_t, uint32_t>> &edges, const std::vector<double> &nnz_in,
                    const std::vector<double> &dom, const std::vector<uint32_t> &mask, const std::set<Sig> *seen,
                    uint32_t keep, double new_table_weight, std::vector<BagStep> &steps, std::vector<Sig> &made,
                    double &cost, double &full, double &peak) {
    std::vector<std::vector<int>> kind(nb, std::vector<int>(nb, 0));  // 0 none, 1 edge relation, 2 dense table
    std::vector<std::vector<double>> nnz(nb, std::vector<double>(nb, 0));
    std::vector<std::vector<std::vector<uint32_t>>> fs(nb, std::vector<std::vector<uint32_t>>(nb));  // eliminated into (a, b)
    std::vector<std::vector<uint32_t>> us(nb);  // eliminated into the vector of b
    for (size_t i = 0; i < edges.size(); ++i) {
        const uint32_t a = edges[i].first, b = edges[i].second;
        kind[a][b] = kind[b][a] = 1;
        nnz[a][b] = nnz[b][a] = nnz_in[i];
    }
    // mean list length of a's rows in the relation (a, v)
    auto list_len = [&](uint32_t a, uint32_t v) { return nnz[a][v] / std::max(1.0, dom[a]); };
    auto S = [&](uint32_t a, uint32_t b) -> std::vector<uint32_t> & { return fs[std::min(a, b)][std::max(a, b)]; };
    auto known = [&](const Sig &k) { return seen && seen->count(k) > 0; };
    std::vector<char> alive(nb, 1);
    uint32_t left = nb - (keep < nb ? 1 : 0);
    steps.clear();
    made.clear();
    while (left) {
        double best = 0, best_charged = 0, best_full = 0;
        int bv = -1;
        std::vector<uint32_t> bn;
        for (uint32_t v = 0; v < nb; ++v) {
            if (!alive[v] || v == keep) continue;
            std::vector<uint32_t> n;
            for (uint32_t w = 0; w < nb; ++w)
                if (w != v && alive[w] && kind[v][w]) n.push_back(w);
            if (n.size() > 2) continue;
            double c, charged, cfull;
            if (n.empty()) c = charged = cfull = dom[v];
            else if (n.size() == 1) {
                const uint32_t a = n[0];
                c = charged = kind[v][a] == 2 ? dom[v] * dom[a] : dom[v] + dom[a];
                cfull = charged;
                if (known(factor_sig(3, mask[a], 0, merged(merged(S(a, v), us[v]), {mask[v]})))) c = charged = dom[a];
            } else {
                const uint32_t x = n[0], y = n[1];
                const bool d0 = kind[v][x] == 2, d1 = kind[v][y] == 2;
                const std::vector<uint32_t> Snew = merged(merged(merged(S(v, x), S(v, y)), us[v]), {mask[v]});
                const double out = dom[x] * dom[y];
                if (kind[x][y] == 1) {  // sparse join on the edges of (x, y)
                    const double e = nnz[x][y];
                    if (d0 && d1) c = e * dom[v];
                    else if (d0) c = e * (1 + list_len(y, v));
                    else if (d1) c = e * (1 + list_len(x, v));
                    else c = e * (1 + std::min(list_len(x, v), list_len(y, v)));
                    charged = cfull = c;
                    if (known(factor_sig(1, mask[x], mask[y], merged(S(x, y), Snew)))) c = charged = 0;
                } else {
                    if (d0 && d1) c = out * (1 + dom[v]);
                    else if (d0 || d1) c = out * 4;
                    else c = out * new_table_weight;
                    // the cost charged is the step's modelled work, whatever weight steered the choice
                    charged = cfull = !d0 && !d1 ? out : c;
                    if (kind[x][y] == 2 && known(factor_sig(2, mask[x], mask[y], merged(S(x, y), Snew)))) c = charged = 0;
                    else if (known(factor_sig(2, mask[x], mask[y], Snew))) c = charged = out;  // a transpose or the join
                }
            }
            // degree 0 and 1 steps first: they shrink the graph without writing a table (a graph of
            // treewidth <= 2 reduces to nothing by such steps in any order, and always has two
            // vertices of degree <= 2, so keeping one block live never blocks the reduction)
            if (bv < 0 || n.size() < bn.size() || (n.size() == bn.size() && c < best)) {
                bv = (int)v;
                bn = n;
                best = c;
                best_charged = charged;
                best_full = cfull;
            }
        }
        if (bv < 0) return false;
        const uint32_t v = (uint32_t)bv;
        BagStep s{};
        s.v = v;
        s.nn = (uint32_t)bn.size();
        for (uint32_t i = 0; i < s.nn; ++i) s.nb[i] = bn[i];
        steps.push_back(s);
        cost += best_charged;
        full += best_full;
        // the keys the engine forms for this step, and the blocks each factor then holds
        if (s.nn == 1) {
            const uint32_t a = bn[0];
            const std::vector<uint32_t> St = merged(merged(S(a, v), us[v]), {mask[v]});
            made.push_back(factor_sig(3, mask[a], 0, St));
            us[a] = merged(us[a], St);
        } else if (s.nn == 2) {
            const uint32_t x = bn[0], y = bn[1];
            const std::vector<uint32_t> Snew = merged(merged(merged(S(v, x), S(v, y)), us[v]), {mask[v]});
            if (kind[x][y] == 1) {
                S(x, y) = merged(S(x, y), Snew);
                made.push_back(factor_sig(1, mask[x], mask[y], S(x, y)));
            } else {
                const Sig kj = factor_sig(2, mask[x], mask[y], merged(S(x, y), Snew));
                if (kind[x][y] != 2 || !known(kj)) made.push_back(factor_sig(2, mask[x], mask[y], Snew));
                if (kind[x][y] == 2) made.push_back(kj);
                S(x, y) = kind[x][y] == 2 ? merged(S(x, y), Snew) : Snew;
                peak = std::max(peak, dom[x] * dom[y]);
                kind[x][y] = kind[y][x] = 2;
            }
        }
        alive[v] = 0;
        for (uint32_t w = 0; w < nb; ++w) {
            kind[v][w] = kind[w][v] = 0;
            S(v, w).clear();
        }
        --left;
    }
    return true;
}

// The cheapest of a few greedy orders (different weights on starting a new table); the chosen
// order's factor results join `seen` (if given).  Each greedy step counts the results in `seen` as lookups,
// but the orders are compared on their work without lookups: compared with lookups, config 4's
// batch (452 queries) took more dense x dense steps and ran 8 % slower on six streams.  The cost
// added is the chosen order's work with lookups.
bool eliminate(uint32_t nb, const std::vector<std::pair<uint32_t, uint32_t>> &edges, const std::vector<double> &nnz,
               const std::vector<double> &dom, const std::vector<uint32_t> &mask, std::set<Sig> *seen, uint32_t keep,
               std::vector<BagStep> &steps, double &cost, double &peak) {
    bool found = false;
    double best_cost = 0, best_full = 0, best_peak = 0;
    std::vector<Sig> best_made;
    for (double w : {1.0, 8.0, 64.0, 4096.0}) {
        std::vector<BagStep> st;
        std::vector<Sig> mk;
        double c = 0, fc = 0, pk = 0;
        if (!eliminate_once(nb, edges, nnz, dom, mask, seen, keep, w, st, mk, c, fc, pk)) return false;  // width > 2 for every weight
        if (!found || fc < best_full) {
            found = true;
            best_cost = c;
            best_full = fc;
            best_peak = pk;
            steps = std::move(st);
            best_made = std::move(mk);
        }
    }
    if (seen) seen->insert(best_made.begin(), best_made.end());
    cost += best_cost;
    peak = std::max(peak, best_peak);
    return true;
}

}  // namespace

bool build_bag_plan(const asdm_cs *cs, bool local, BagPlan &out, std::string &why) {
    const uint32_t nq = cs->nq;
    out = BagPlan{};
    out.local = local;
    if (!cs->loop_free) {
        why = "the data graph has self-loops";
        return false;
    }
    // label classes with more than one vertex
    std::map<uint32_t, std::vector<uint32_t>> classes;
    for (uint32_t u = 0; u < nq; ++u) classes[cs->qlab[u]].push_back(u);
    std::vector<std::vector<std::vector<std::vector<uint32_t>>>> per;  // per class: its partitions
    uint64_t nterms = 1;
    for (auto &c : classes) {
        std::vector<std::vector<std::vector<uint32_t>>> parts;
        bool overflow = false;
        independent_partitions(cs, c.second, parts, kMaxBagTerms, overflow);
        nterms *= parts.size();
        if (overflow || nterms > kMaxBagTerms) {
            why = "too many partition terms";
            return false;
        }
        per.push_back(std::move(parts));
    }
    std::vector<size_t> idx(per.size(), 0);
    // factor keys produced by the eliminations planned so far.  Marginal plans are made without them:
    // on config 3 (12 vertices, 200,000 x 12 matrix) the lookups steered the greedy steps of the
    // marginals into more vector and dense x dense work, 21.3 against 18.5 ms per count.
    std::set<Sig> seen_keys;
    std::set<Sig> *seen = local ? nullptr : &seen_keys;
    for (uint64_t t = 0; t < nterms; ++t) {
        BagTerm T;
        std::vector<std::vector<uint32_t>> blocks;
        for (size_t c = 0; c < per.size(); ++c)
            for (auto &b : per[c][idx[c]]) blocks.push_back(b);
        T.mu = mobius(blocks);
        const uint32_t nb = (uint32_t)blocks.size();
        std::vector<uint32_t> bof(nq);
        std::vector<double> dom(nb);
        for (uint32_t b = 0; b < nb; ++b) {
            uint32_t rep = blocks[b][0];
            for (uint32_t u : blocks[b]) {
                bof[u] = b;
                if (cs->csize[u] < cs->csize[rep]) rep = u;
            }
            T.rep.push_back(rep);
            dom[b] = std::max(1.0, (double)cs->csize[rep]);
        }
        std::map<std::pair<uint32_t, uint32_t>, std::pair<uint32_t, uint32_t>> em;
        for (uint32_t a = 0; a < cs->narcs; ++a) {
            const uint32_t A = bof[cs->arc_u[a]], B = bof[cs->arc_u2[a]];
            if (A == B) continue;  // cannot happen: blocks are independent sets
            if (A < B && !em.count({A, B})) em[{A, B}] = {a, cs->arc_index(cs->arc_u2[a], cs->arc_u[a])};
        }
        for (auto &e : em) {
            T.edges.push_back(e.first);
            T.arc_ab.push_back(e.second.first);
            T.arc_ba.push_back(e.second.second);
        }
        // entries of each quotient edge's relation: at most the arc's candidate edges
        std::vector<double> nnz(T.edges.size());
        for (size_t i = 0; i < T.edges.size(); ++i)
            nnz[i] = std::min((double)cs->arc_weight[T.arc_ab[i]], dom[T.edges[i].first] * dom[T.edges[i].second]);
        std::vector<uint32_t> mask(nb, 0);
        for (uint32_t b = 0; b < nb; ++b)
            for (uint32_t q : blocks[b]) mask[b] |= 1u << q;
        T.blocks = std::move(blocks);
        double peak = 0;
        if (!eliminate(nb, T.edges, nnz, dom, mask, seen, UINT32_MAX, T.count_steps, out.cost, peak)) {
            why = "a quotient of the query needs tables over three variables";
            return false;
        }
        if (local) {
            T.local_steps.resize(nb);
            for (uint32_t k = 0; k < nb; ++k)
                if (!eliminate(nb, T.edges, nnz, dom, mask, seen, k, T.local_steps[k], out.cost, peak)) {
                    why = "a quotient of the query needs tables over three variables for a marginal";
                    return false;
                }
        }
        out.peak_table = std::max(out.peak_table, peak);
        out.terms.push_back(std::move(T));
        for (size_t c = 0; c < per.size(); ++c) {
            if (++idx[c] < per[c].size()) break;
            idx[c] = 0;
        }
    }
    if (out.peak_table > kMaxBagTable) {
        why = "a bag table would exceed " + std::to_string((uint64_t)kMaxBagTable) + " entries";
        return false;
    }
    return true;
}

}  // namespace asdm
