// Seeded synthetic graphs and random-walk queries.  All randomness comes from one portable
// xoshiro256** stream per call (std::*_distribution is implementation-defined), so the same seed
// gives the same graph on every machine.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <numeric>
#include <vector>

#include "host.h"

namespace asdm {

static uint64_t splitmix64(uint64_t &x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

Rng::Rng(uint64_t seed) {
    uint64_t x = seed;
    for (auto &w : s) w = splitmix64(x);
}

uint64_t Rng::next() {
    const uint64_t r = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
}

double Rng::uniform() { return (double)(next() >> 11) * 0x1.0p-53; }

uint64_t Rng::below(uint64_t bound) {
    if (bound <= 1) return 0;
    uint64_t lim = (~0ull) - ((~0ull) % bound);
    uint64_t x;
    do { x = next(); } while (x >= lim);
    return x % bound;
}

namespace {

// Walker/Vose alias table for sampling an index with probability proportional to w[i].
struct Alias {
    std::vector<double> prob;
    std::vector<uint32_t> alias;
    explicit Alias(const std::vector<double> &w) {
        size_t n = w.size();
        prob.resize(n);
        alias.resize(n);
        double sum = 0;
        for (double x : w) sum += x;
        std::vector<double> p(n);
        std::vector<uint32_t> small, large;
        for (size_t i = 0; i < n; ++i) {
            p[i] = w[i] * (double)n / sum;
            (p[i] < 1.0 ? small : large).push_back((uint32_t)i);
        }
        while (!small.empty() && !large.empty()) {
            uint32_t s = small.back(), l = large.back();
            small.pop_back();
            prob[s] = p[s];
            alias[s] = l;
            p[l] = (p[l] + p[s]) - 1.0;
            if (p[l] < 1.0) {
                large.pop_back();
                small.push_back(l);
            }
        }
        for (uint32_t i : large) { prob[i] = 1.0; alias[i] = i; }
        for (uint32_t i : small) { prob[i] = 1.0; alias[i] = i; }
    }
    uint32_t sample(Rng &r) const {
        uint32_t i = (uint32_t)r.below(prob.size());
        return r.uniform() < prob[i] ? i : alias[i];
    }
};

// Open-addressing set of undirected edge keys (min << 32 | max).
struct EdgeSet {
    std::vector<uint64_t> slot;
    uint64_t mask;
    explicit EdgeSet(uint64_t expect) {
        uint64_t cap = 16;
        while (cap < expect * 2) cap <<= 1;
        slot.assign(cap, ~0ull);
        mask = cap - 1;
    }
    bool insert(uint64_t key) {
        uint64_t h = key * 0x9E3779B97F4A7C15ull;
        uint64_t i = (h ^ (h >> 29)) & mask;
        while (slot[i] != ~0ull) {
            if (slot[i] == key) return false;
            i = (i + 1) & mask;
        }
        slot[i] = key;
        return true;
    }
};

std::vector<uint32_t> draw_labels(uint32_t n, uint32_t nlabels, double s, Rng &r) {
    std::vector<uint32_t> lab(n, 0);
    if (nlabels <= 1) return lab;
    std::vector<double> cdf(nlabels);
    double acc = 0;
    for (uint32_t k = 0; k < nlabels; ++k) {
        acc += s == 0.0 ? 1.0 : std::pow((double)(k + 1), -s);
        cdf[k] = acc;
    }
    for (uint32_t v = 0; v < n; ++v) {
        double x = r.uniform() * acc;
        uint32_t k = (uint32_t)(std::upper_bound(cdf.begin(), cdf.end(), x) - cdf.begin());
        lab[v] = std::min(k, nlabels - 1);
    }
    return lab;
}

// Random relabelling of vertex ids, then the graph in edge-record order.
asdm_graph *finish_random(uint32_t n, std::vector<uint32_t> &edges, uint32_t nlabels, double zipf_s, Rng &r) {
    std::vector<uint32_t> perm(n);
    std::iota(perm.begin(), perm.end(), 0u);
    for (uint32_t i = n; i > 1; --i) std::swap(perm[i - 1], perm[r.below(i)]);
    for (auto &x : edges) x = perm[x];
    std::vector<uint32_t> lab = draw_labels(n, nlabels, zipf_s, r);
    return graph_from_edges(n, lab.data(), edges.size() / 2, edges.data());
}

}  // namespace
}  // namespace asdm

using asdm::Rng;

extern "C" {

int asdm_gen_chung_lu(uint32_t n, uint64_t ne, double beta, uint32_t nlabels, double zipf_s, uint64_t seed,
                      asdm_graph **out) try {
    if (!out || n < 2 || beta <= 1.0 || nlabels == 0) return ASDM_E_INVALID;
    *out = nullptr;
    if (ne > (uint64_t)n * (n - 1) / 4 || 2 * ne >= (1ull << 32)) return ASDM_E_INVALID;
    Rng r(seed);
    std::vector<double> w(n);
    const double ex = -1.0 / (beta - 1.0);
    for (uint32_t i = 0; i < n; ++i) w[i] = std::pow((double)i + 1.0, ex);
    asdm::Alias al(w);
    asdm::EdgeSet set(ne);
    std::vector<uint32_t> edges;
    edges.reserve(2 * ne);
    uint64_t tries = 0, limit = 200 * ne + 1000;
    while (edges.size() < 2 * ne) {
        if (++tries > limit) return ASDM_E_INVALID;
        uint32_t a = al.sample(r), b = al.sample(r);
        if (a == b) continue;
        uint64_t key = a < b ? ((uint64_t)a << 32 | b) : ((uint64_t)b << 32 | a);
        if (!set.insert(key)) continue;
        edges.push_back(a);
        edges.push_back(b);
    }
    *out = asdm::finish_random(n, edges, nlabels, zipf_s, r);
    return ASDM_OK;
} ASDM_HOST_CATCH

int asdm_gen_rmat(uint32_t scale, uint32_t edge_factor, double a, double b, double c, uint32_t nlabels,
                  double zipf_s, uint64_t seed, asdm_graph **out) try {
    if (!out || scale == 0 || scale > 30 || nlabels == 0 || a < 0 || b < 0 || c < 0 || a + b + c > 1.0)
        return ASDM_E_INVALID;
    *out = nullptr;
    const uint32_t n = 1u << scale;
    const uint64_t draws = (uint64_t)edge_factor << scale;
    if (2 * draws >= (1ull << 32)) return ASDM_E_UNSUPPORTED;
    Rng r(seed);
    asdm::EdgeSet set(draws);
    std::vector<uint32_t> edges;
    edges.reserve(2 * draws);
    const double ab = a + b, abc = a + b + c;
    for (uint64_t i = 0; i < draws; ++i) {
        uint32_t s = 0, d = 0;
        for (uint32_t bit = 0; bit < scale; ++bit) {
            double x = r.uniform();
            uint32_t si = 0, di = 0;
            if (x < a) {
            } else if (x < ab) {
                di = 1;
            } else if (x < abc) {
                si = 1;
            } else {
                si = 1;
                di = 1;
            }
            s = (s << 1) | si;
            d = (d << 1) | di;
        }
        if (s == d) continue;
        uint64_t key = s < d ? ((uint64_t)s << 32 | d) : ((uint64_t)d << 32 | s);
        if (!set.insert(key)) continue;
        edges.push_back(s);
        edges.push_back(d);
    }
    *out = asdm::finish_random(n, edges, nlabels, zipf_s, r);
    return ASDM_OK;
} ASDM_HOST_CATCH

int asdm_gen_query(const asdm_graph *g, uint32_t k, double density, uint32_t max_extra, uint64_t seed,
                   asdm_graph **out) try {
    if (!g || !out || k == 0 || k > 64 || density < 0 || density > 1) return ASDM_E_INVALID;
    *out = nullptr;
    Rng r(seed);
    std::vector<uint32_t> starts;
    for (uint32_t v = 0; v < g->n; ++v)
        if (g->deg(v) > 0) starts.push_back(v);
    if (starts.empty()) return ASDM_E_INVALID;
    std::vector<uint32_t> verts;
    std::vector<std::pair<uint32_t, uint32_t>> tree;  // (query id, query id)
    bool ok = false;
    for (int attempt = 0; attempt < 1000 && !ok; ++attempt) {
        verts.clear();
        tree.clear();
        uint32_t x = starts[r.below(starts.size())];
        verts.push_back(x);
        uint32_t cur = 0;  // query id of x
        for (uint64_t step = 0; step < 200ull * k && verts.size() < k; ++step) {
            uint32_t d = g->deg(x);
            uint32_t y = g->nbr[g->off[x] + (uint32_t)r.below(d)];
            auto it = std::find(verts.begin(), verts.end(), y);
            uint32_t qy;
            if (it == verts.end()) {
                qy = (uint32_t)verts.size();
                verts.push_back(y);
                tree.emplace_back(cur, qy);
            } else {
                qy = (uint32_t)(it - verts.begin());
            }
            x = y;
            cur = qy;
        }
        ok = verts.size() == k;
    }
    if (!ok) return ASDM_E_INVALID;
    std::vector<std::pair<uint32_t, uint32_t>> extra;
    for (uint32_t i = 0; i < k; ++i)
        for (uint32_t j = i + 1; j < k; ++j) {
            if (!g->has_edge(verts[i], verts[j])) continue;
            bool in_tree = false;
            for (auto &t : tree)
                if ((t.first == i && t.second == j) || (t.first == j && t.second == i)) in_tree = true;
            if (!in_tree) extra.emplace_back(i, j);
        }
    for (size_t i = extra.size(); i > 1; --i) std::swap(extra[i - 1], extra[r.below(i)]);
    size_t keep = (size_t)std::llround(density * (double)extra.size());
    keep = std::min<size_t>(keep, max_extra);
    std::vector<uint32_t> edges;
    for (auto &t : tree) {
        edges.push_back(t.first);
        edges.push_back(t.second);
    }
    for (size_t i = 0; i < keep; ++i) {
        edges.push_back(extra[i].first);
        edges.push_back(extra[i].second);
    }
    std::vector<uint32_t> lab(k);
    for (uint32_t i = 0; i < k; ++i) lab[i] = g->lab[verts[i]];
    *out = asdm::graph_from_edges(k, lab.data(), edges.size() / 2, edges.data());
    return ASDM_OK;
} ASDM_HOST_CATCH

}  // extern "C"
