// Graph I/O and host-side layout.  Format and in-memory layout follow the reference loader
// (relation/graph.cpp:68-148) and text writer (relation/graph.cpp:451-471).
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <memory>
#include <numeric>
#include <thread>

#include "host.h"

namespace asdm {

void finish_graph(asdm_graph &g) {
    // Per-vertex sort by (label(nbr), nbr), split across host threads for large graphs.
    const uint32_t *lab = g.lab.data();
    auto sort_range = [&](uint32_t v0, uint32_t v1) {
        for (uint32_t v = v0; v < v1; ++v)
            std::sort(g.nbr.begin() + g.off[v], g.nbr.begin() + g.off[v + 1], [lab](uint32_t a, uint32_t b) {
                return lab[a] != lab[b] ? lab[a] < lab[b] : a < b;
            });
    };
    unsigned nt = g.m > (1u << 22) ? std::max(1u, std::min(16u, std::thread::hardware_concurrency())) : 1u;
    if (nt <= 1) {
        sort_range(0, g.n);
    } else {
        std::vector<std::thread> th;
        for (unsigned t = 0; t < nt; ++t) {
            uint32_t v0 = (uint32_t)((uint64_t)g.n * t / nt), v1 = (uint32_t)((uint64_t)g.n * (t + 1) / nt);
            th.emplace_back(sort_range, v0, v1);
        }
        for (auto &t : th) t.join();
    }
    g.loop_free = true;
    for (uint32_t v = 0; v < g.n && g.loop_free; ++v)
        for (uint32_t a = g.off[v]; a < g.off[v + 1]; ++a)
            if (g.nbr[a] == v) {
                g.loop_free = false;
                break;
            }
    g.vbl_off.assign(g.nlab + 1, 0);
    for (uint32_t v = 0; v < g.n; ++v) g.vbl_off[g.lab[v] + 1]++;
    for (uint32_t l = 0; l < g.nlab; ++l) g.vbl_off[l + 1] += g.vbl_off[l];
    g.vbl.resize(g.n);
    std::vector<uint32_t> pos(g.vbl_off.begin(), g.vbl_off.end());
    for (uint32_t v = 0; v < g.n; ++v) g.vbl[pos[g.lab[v]]++] = v;
    std::vector<double> degsum(g.nlab, 0.0);
    for (uint32_t v = 0; v < g.n; ++v) degsum[g.lab[v]] += g.off[v + 1] - g.off[v];
    g.lab_avgdeg.assign(g.nlab, 0.f);
    for (uint32_t l = 0; l < g.nlab; ++l)
        if (g.vbl_off[l + 1] > g.vbl_off[l]) g.lab_avgdeg[l] = (float)(degsum[l] / (g.vbl_off[l + 1] - g.vbl_off[l]));
}

asdm_graph *graph_from_edges(uint32_t n, const uint32_t *labels, uint64_t ne, const uint32_t *edges) {
    auto g = std::make_unique<asdm_graph>();
    g->n = n;
    g->m = 2 * ne;
    g->lab.assign(labels, labels + n);
    uint32_t maxlab = 0;
    for (uint32_t v = 0; v < n; ++v) maxlab = std::max(maxlab, labels[v]);
    g->nlab = n ? maxlab + 1 : 0;
    g->off.assign((size_t)n + 1, 0);
    for (uint64_t e = 0; e < ne; ++e) {
        g->off[edges[2 * e] + 1]++;
        g->off[edges[2 * e + 1] + 1]++;
    }
    for (uint32_t v = 0; v < n; ++v) g->off[v + 1] += g->off[v];
    g->nbr.resize(g->m);
    std::vector<uint32_t> fill(n, 0);
    for (uint64_t e = 0; e < ne; ++e) {
        uint32_t a = edges[2 * e], b = edges[2 * e + 1];
        g->nbr[g->off[a] + fill[a]++] = b;
        g->nbr[g->off[b] + fill[b]++] = a;
    }
    finish_graph(*g);
    return g.release();
}

}  // namespace asdm

void asdm_graph::label_slice(uint32_t v, uint32_t l, uint32_t &b, uint32_t &cnt) const {
    const uint32_t *lo = nbr.data() + off[v], *hi = nbr.data() + off[v + 1];
    const uint32_t *L = lab.data();
    const uint32_t *s = std::lower_bound(lo, hi, l, [L](uint32_t x, uint32_t lv) { return L[x] < lv; });
    const uint32_t *e = std::upper_bound(s, hi, l, [L](uint32_t lv, uint32_t x) { return lv < L[x]; });
    b = (uint32_t)(s - nbr.data());
    cnt = (uint32_t)(e - s);
}

bool asdm_graph::has_edge(uint32_t v, uint32_t w) const {
    uint32_t b, c;
    label_slice(v, lab[w], b, c);
    return std::binary_search(nbr.begin() + b, nbr.begin() + b + c, w);
}

namespace {

// Whole-file reader with the reference's token semantics: the record type is one non-space
// character (`infile >> type`), numbers are unsigned decimals, other characters are skipped.
struct Scanner {
    std::vector<char> buf;
    size_t p = 0;
    bool load(const char *path) {
        FILE *f = std::fopen(path, "rb");
        if (!f) return false;
        std::fseek(f, 0, SEEK_END);
        long sz = std::ftell(f);
        std::fseek(f, 0, SEEK_SET);
        if (sz < 0) { std::fclose(f); return false; }
        buf.resize((size_t)sz);
        size_t got = sz ? std::fread(buf.data(), 1, (size_t)sz, f) : 0;
        std::fclose(f);
        return got == (size_t)sz;
    }
    static bool space(char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '\v' || c == '\f'; }
    int next_char() {
        while (p < buf.size() && space(buf[p])) ++p;
        return p < buf.size() ? (unsigned char)buf[p++] : -1;
    }
    bool next_u32(uint32_t &out) {
        while (p < buf.size() && space(buf[p])) ++p;
        if (p >= buf.size() || buf[p] < '0' || buf[p] > '9') return false;
        uint64_t x = 0;
        while (p < buf.size() && buf[p] >= '0' && buf[p] <= '9') x = x * 10 + (uint64_t)(buf[p++] - '0');
        out = (uint32_t)x;
        return true;
    }
};

}  // namespace

extern "C" {

const char *asdm_version(void) { return "asdm-mi355x 0.1 (gfx950)"; }

int asdm_graph_load_text(const char *path, asdm_graph **out) {
    if (!path || !out) return ASDM_E_INVALID;
    *out = nullptr;
    Scanner s;
    if (!s.load(path)) return ASDM_E_IO;
    auto g = std::make_unique<asdm_graph>();
    uint32_t n = 0, me = 0;
    if (s.next_char() < 0 || !s.next_u32(n) || !s.next_u32(me)) return ASDM_E_IO;
    g->n = n;
    g->m = 2ull * me;
    if (g->m >= (1ull << 32)) return ASDM_E_UNSUPPORTED;
    g->lab.assign(n, 0);
    g->off.assign((size_t)n + 1, 0);
    g->nbr.assign(g->m, 0);
    std::vector<uint32_t> fill(n, 0);
    uint32_t maxlab = 0;
    for (;;) {
        int c = s.next_char();
        if (c < 0) break;
        if (c == 'v') {
            uint32_t id, l, d;
            if (!s.next_u32(id) || !s.next_u32(l) || !s.next_u32(d) || id >= n) return ASDM_E_IO;
            g->lab[id] = l;
            g->off[id + 1] = g->off[id] + d;
            maxlab = std::max(maxlab, l);
        } else if (c == 'e') {
            uint32_t a, b;
            if (!s.next_u32(a) || !s.next_u32(b) || a >= n || b >= n) return ASDM_E_IO;
            if (g->off[a] + fill[a] >= g->off[a + 1] || g->off[b] + fill[b] >= g->off[b + 1]) return ASDM_E_IO;
            g->nbr[g->off[a] + fill[a]++] = b;
            g->nbr[g->off[b] + fill[b]++] = a;
        }
    }
    for (uint32_t v = 0; v < n; ++v)
        if (fill[v] != g->off[v + 1] - g->off[v]) return ASDM_E_IO;
    if (g->off[n] != g->m) return ASDM_E_IO;
    g->nlab = n ? maxlab + 1 : 0;
    asdm::finish_graph(*g);
    *out = g.release();
    return ASDM_OK;
}

int asdm_graph_from_edges(uint32_t n, const uint32_t *labels, uint64_t ne, const uint32_t *edges, asdm_graph **out) {
    if (!out || (n && !labels) || (ne && !edges)) return ASDM_E_INVALID;
    *out = nullptr;
    if (2 * ne >= (1ull << 32)) return ASDM_E_UNSUPPORTED;
    for (uint64_t i = 0; i < 2 * ne; ++i)
        if (edges[i] >= n) return ASDM_E_INVALID;
    *out = asdm::graph_from_edges(n, labels, ne, edges);
    return ASDM_OK;
}

int asdm_graph_write_text(const asdm_graph *g, const char *path) {
    if (!g || !path) return ASDM_E_INVALID;
    FILE *f = std::fopen(path, "w");
    if (!f) return ASDM_E_IO;
    std::vector<char> wbuf(1 << 20);
    std::setvbuf(f, wbuf.data(), _IOFBF, wbuf.size());
    std::fprintf(f, "t %u %llu\n", g->n, (unsigned long long)(g->m / 2));
    for (uint32_t v = 0; v < g->n; ++v) std::fprintf(f, "v %u %u %u\n", v, g->lab[v], g->deg(v));
    for (uint32_t v = 0; v < g->n; ++v)
        for (uint32_t i = g->off[v]; i < g->off[v + 1]; ++i)
            if (v < g->nbr[i]) std::fprintf(f, "e %u %u\n", v, g->nbr[i]);
    int rc = std::fclose(f) == 0 ? ASDM_OK : ASDM_E_IO;
    return rc;
}

void asdm_graph_free(asdm_graph *g) {
    if (!g) return;
    asdm::release_device_graph(*g);
    delete g;
}

int asdm_graph_info(const asdm_graph *g, uint32_t *n, uint64_t *m, uint32_t *nlabels) {
    if (!g) return ASDM_E_INVALID;
    if (n) *n = g->n;
    if (m) *m = g->m;
    if (nlabels) *nlabels = g->nlab;
    return ASDM_OK;
}

int asdm_graph_csr(const asdm_graph *g, uint32_t *off, uint32_t *nbr, uint32_t *lab) {
    if (!g) return ASDM_E_INVALID;
    if (off) std::memcpy(off, g->off.data(), sizeof(uint32_t) * ((size_t)g->n + 1));
    if (nbr && g->m) std::memcpy(nbr, g->nbr.data(), sizeof(uint32_t) * g->m);
    if (lab && g->n) std::memcpy(lab, g->lab.data(), sizeof(uint32_t) * g->n);
    return ASDM_OK;
}

int asdm_graph_edges(const asdm_graph *g, uint32_t *edges, uint64_t cap, uint64_t *ne) {
    if (!g || !ne) return ASDM_E_INVALID;
    std::vector<std::pair<uint32_t, uint32_t>> es;
    es.reserve(g->m / 2);
    for (uint32_t v = 0; v < g->n; ++v)
        for (uint32_t i = g->off[v]; i < g->off[v + 1]; ++i)
            if (v < g->nbr[i]) es.emplace_back(v, g->nbr[i]);
    std::sort(es.begin(), es.end());
    *ne = es.size();
    if (!edges) return ASDM_OK;
    if (cap < es.size()) return ASDM_E_INVALID;
    for (size_t i = 0; i < es.size(); ++i) {
        edges[2 * i] = es[i].first;
        edges[2 * i + 1] = es[i].second;
    }
    return ASDM_OK;
}

}  // extern "C"
