// Graph I/O and host-side layout.  Format and in-memory layout follow the reference loader
// (relation/graph.cpp:68-148) and text writer (relation/graph.cpp:451-471).
#include <algorithm>
#include <atomic>
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
    index_graph(g);
}

void index_graph(asdm_graph &g) {
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
    const char *lo = nullptr;  // a range of another scanner's buffer (parse_parallel), else buf
    size_t hi = 0;
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
    const char *data() const { return lo ? lo : buf.data(); }
    size_t size() const { return lo ? hi : buf.size(); }
    int next_char() {
        const char *b = data();
        const size_t n = size();
        while (p < n && space(b[p])) ++p;
        return p < n ? (unsigned char)b[p++] : -1;
    }
    bool next_u32(uint32_t &out) {
        const char *b = data();
        const size_t n = size();
        while (p < n && space(b[p])) ++p;
        if (p >= n || b[p] < '0' || b[p] > '9') return false;
        uint64_t x = 0;
        while (p < n && b[p] >= '0' && b[p] <= '9') x = x * 10 + (uint64_t)(b[p++] - '0');
        out = (uint32_t)x;
        return true;
    }
};

// Parallel parse of a large graph file.  Every 'v' or 'e' character is a record type wherever the
// sequential scanner meets it (numbers hold digits only), so cutting the records at such
// characters parses each chunk exactly as the sequential scan would, and a record cut short by a
// chunk end fails in both.  Each thread collects its chunk's vertex and edge records; vertex records
// are applied in file order, then each thread places the edge endpoints of its own vertex range
// (the order inside an adjacency list does not matter: finish_graph sorts it).  Returns false, untouched, when
// an edge record precedes a vertex record (the sequential scan then decides), and sets rc on a
// malformed file.
constexpr size_t kParallelParseBytes = 32u << 20;

bool parse_parallel(Scanner &s, asdm_graph &g, uint32_t n, uint32_t &maxlab, int &rc) {
    const size_t p0 = s.p, end = s.buf.size();
    const char *B = s.buf.data();
    const unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<size_t> cut(nt + 1, end);
    cut[0] = p0;
    for (unsigned k = 1; k < nt; ++k) {
        size_t q = std::max(cut[k - 1], p0 + (end - p0) / nt * k);
        while (q < end && !((B[q] == 'v' || B[q] == 'e') && q > 0 && Scanner::space(B[q - 1]))) ++q;
        cut[k] = q;
    }
    struct Part {
        std::vector<uint32_t> v, e;  // (id, label, degree) and (a, b) records
        size_t first_e = SIZE_MAX, last_v = 0;
        bool bad = false;
    };
    std::vector<Part> parts(nt);
    auto scan = [&](unsigned k) {
        Scanner c;
        c.lo = B + cut[k];
        c.hi = cut[k + 1] - cut[k];
        Part &P = parts[k];
        for (;;) {
            const size_t at = c.p;
            const int ch = c.next_char();
            if (ch < 0) break;
            if (ch == 'v') {
                uint32_t id, l, d;
                if (!c.next_u32(id) || !c.next_u32(l) || !c.next_u32(d)) { P.bad = true; break; }
                P.v.insert(P.v.end(), {id, l, d});
                P.last_v = cut[k] + at;
            } else if (ch == 'e') {
                uint32_t a, b;
                if (!c.next_u32(a) || !c.next_u32(b)) { P.bad = true; break; }
                P.e.insert(P.e.end(), {a, b});
                if (P.first_e == SIZE_MAX) P.first_e = cut[k] + at;
            }
        }
    };
    {
        std::vector<std::thread> th;
        for (unsigned k = 0; k < nt; ++k) th.emplace_back(scan, k);
        for (auto &t : th) t.join();
    }
    size_t first_e = SIZE_MAX, last_v = 0;
    for (auto &P : parts) {
        first_e = std::min(first_e, P.first_e);
        if (!P.v.empty()) last_v = std::max(last_v, P.last_v);
    }
    if (first_e != SIZE_MAX && last_v > first_e) return false;
    // records up to the first malformed one are the ones the sequential scan would apply
    for (auto &P : parts) {
        for (size_t i = 0; i < P.v.size(); i += 3) {
            const uint32_t id = P.v[i], l = P.v[i + 1], d = P.v[i + 2];
            if (id >= n) { rc = ASDM_E_IO; return true; }
            if (l > asdm::kMaxLabel) { rc = ASDM_E_UNSUPPORTED; return true; }
            g.lab[id] = l;
            g.off[id + 1] = g.off[id] + d;
            maxlab = std::max(maxlab, l);
        }
        if (P.bad) { rc = ASDM_E_IO; return true; }
    }
    // the offsets are final here: non-decreasing up to m keeps every placement inside nbr
    for (uint32_t v = 0; v < n; ++v)
        if (g.off[v] > g.off[v + 1]) { rc = ASDM_E_IO; return true; }
    if (g.off[n] != g.m) { rc = ASDM_E_IO; return true; }
    // thread k owns the adjacency lists of vertices [v0, v1): it reads every edge record and places
    // the endpoints in its range, so no cursor is shared
    std::vector<uint32_t> fill(n, 0);
    std::atomic<bool> bad{false};
    auto place = [&](unsigned k) {
        const uint32_t v0 = (uint32_t)((uint64_t)n * k / nt), v1 = (uint32_t)((uint64_t)n * (k + 1) / nt);
        for (const Part &P : parts) {
            const std::vector<uint32_t> &E = P.e;
            for (size_t i = 0; i < E.size(); i += 2) {
                const uint32_t a = E[i], b = E[i + 1];
                if (a >= n || b >= n) { bad = true; return; }
                if (a >= v0 && a < v1) {
                    if (g.off[a] + fill[a] >= g.off[a + 1]) { bad = true; return; }
                    g.nbr[g.off[a] + fill[a]++] = b;
                }
                if (b >= v0 && b < v1) {
                    if (g.off[b] + fill[b] >= g.off[b + 1]) { bad = true; return; }
                    g.nbr[g.off[b] + fill[b]++] = a;
                }
            }
            if (bad.load(std::memory_order_relaxed)) return;
        }
    };
    {
        std::vector<std::thread> th;
        for (unsigned k = 0; k < nt; ++k) th.emplace_back(place, k);
        for (auto &t : th) t.join();
    }
    if (bad) { rc = ASDM_E_IO; return true; }
    for (uint32_t v = 0; v < n; ++v)
        if (fill[v] != g.off[v + 1] - g.off[v]) { rc = ASDM_E_IO; return true; }
    rc = ASDM_OK;
    return true;
}

}  // namespace

extern "C" {

const char *asdm_version(void) { return "asdm-mi355x 0.1 (gfx950)"; }

// Binary CSR cache: "ASDMCSR1", u32 n, u32 nlabels, u64 m, then lab[n], off[n+1], nbr[m] (u32, the
// label-sorted adjacency as held in memory).  No text graph can start with these bytes (its header
// needs a number after the first character).
constexpr char kBinMagic[8] = {'A', 'S', 'D', 'M', 'C', 'S', 'R', '1'};

static int load_binary(const std::vector<char> &buf, asdm_graph **out) {
    const char *b = buf.data();
    const size_t sz = buf.size();
    if (sz < 24) return ASDM_E_IO;
    uint32_t n, nlab;
    uint64_t m;
    std::memcpy(&n, b + 8, 4);
    std::memcpy(&nlab, b + 12, 4);
    std::memcpy(&m, b + 16, 8);
    if (m >= (1ull << 32) || sz != 24 + 4ull * ((uint64_t)n + (uint64_t)n + 1 + m)) return ASDM_E_IO;
    auto g = std::make_unique<asdm_graph>();
    g->n = n;
    g->m = m;
    g->nlab = nlab;
    const char *p = b + 24;
    g->lab.resize(n);
    g->off.resize((size_t)n + 1);
    g->nbr.resize(m);
    std::memcpy(g->lab.data(), p, 4ull * n);
    p += 4ull * n;
    std::memcpy(g->off.data(), p, 4ull * (n + 1));
    p += 4ull * (n + 1);
    if (m) std::memcpy(g->nbr.data(), p, 4ull * m);
    // the indexes trust the layout: check it (labels in range, offsets, ids, sorted adjacency)
    if (nlab > asdm::kMaxLabel + 1) return ASDM_E_UNSUPPORTED;
    if (g->off[0] != 0 || g->off[n] != m) return ASDM_E_IO;
    for (uint32_t v = 0; v < n; ++v)
        if (g->lab[v] >= nlab || g->off[v] > g->off[v + 1]) return ASDM_E_IO;
    const uint32_t *L = g->lab.data();
    std::atomic<bool> bad{false};
    auto check = [&](uint32_t v0, uint32_t v1) {
        for (uint32_t v = v0; v < v1 && !bad.load(std::memory_order_relaxed); ++v)
            for (uint32_t a = g->off[v]; a < g->off[v + 1]; ++a) {
                const uint32_t w = g->nbr[a];
                if (w >= n) { bad = true; return; }
                if (a > g->off[v]) {
                    const uint32_t u = g->nbr[a - 1];
                    if (L[u] > L[w] || (L[u] == L[w] && u >= w)) { bad = true; return; }
                }
            }
    };
    const unsigned nt = m > (1u << 22) ? std::max(1u, std::min(16u, std::thread::hardware_concurrency())) : 1u;
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; ++t)
        th.emplace_back(check, (uint32_t)((uint64_t)n * t / nt), (uint32_t)((uint64_t)n * (t + 1) / nt));
    for (auto &t : th) t.join();
    if (bad) return ASDM_E_IO;
    asdm::index_graph(*g);
    *out = g.release();
    return ASDM_OK;
}

int asdm_graph_save_binary(const asdm_graph *g, const char *path) try {
    if (!g || !path) return ASDM_E_INVALID;
    FILE *f = std::fopen(path, "wb");
    if (!f) return ASDM_E_IO;
    const uint64_t m = g->m;
    bool ok = std::fwrite(kBinMagic, 1, 8, f) == 8 && std::fwrite(&g->n, 4, 1, f) == 1 &&
              std::fwrite(&g->nlab, 4, 1, f) == 1 && std::fwrite(&m, 8, 1, f) == 1;
    ok = ok && (g->n == 0 || std::fwrite(g->lab.data(), 4, g->n, f) == g->n);
    ok = ok && std::fwrite(g->off.data(), 4, (size_t)g->n + 1, f) == (size_t)g->n + 1;
    ok = ok && (m == 0 || std::fwrite(g->nbr.data(), 4, m, f) == m);
    ok = (std::fclose(f) == 0) && ok;
    return ok ? ASDM_OK : ASDM_E_IO;
} ASDM_HOST_CATCH

int asdm_graph_load_text(const char *path, asdm_graph **out) try {
    if (!path || !out) return ASDM_E_INVALID;
    *out = nullptr;
    Scanner s;
    if (!s.load(path)) return ASDM_E_IO;
    if (s.buf.size() >= 8 && std::memcmp(s.buf.data(), kBinMagic, 8) == 0) return load_binary(s.buf, out);
    auto g = std::make_unique<asdm_graph>();
    uint32_t n = 0, me = 0;
    if (s.next_char() < 0 || !s.next_u32(n) || !s.next_u32(me)) return ASDM_E_IO;
    g->n = n;
    g->m = 2ull * me;
    if (g->m >= (1ull << 32)) return ASDM_E_UNSUPPORTED;
    if ((uint64_t)me > s.buf.size() / 5) return ASDM_E_IO;  // each edge needs a record of at least 5 bytes
    g->lab.assign(n, 0);
    g->off.assign((size_t)n + 1, 0);
    g->nbr.assign(g->m, 0);
    uint32_t maxlab = 0;
    if (s.buf.size() - s.p >= kParallelParseBytes) {
        int rc = ASDM_OK;
        if (parse_parallel(s, *g, n, maxlab, rc)) {
            if (rc) return rc;
            if (g->off[n] != g->m) return ASDM_E_IO;
            g->nlab = n ? maxlab + 1 : 0;
            asdm::finish_graph(*g);
            *out = g.release();
            return ASDM_OK;
        }
        maxlab = 0;  // an edge record precedes a vertex record: the sequential scan decides
    }
    std::vector<uint32_t> fill(n, 0);
    for (;;) {
        int c = s.next_char();
        if (c < 0) break;
        if (c == 'v') {
            uint32_t id, l, d;
            if (!s.next_u32(id) || !s.next_u32(l) || !s.next_u32(d) || id >= n) return ASDM_E_IO;
            if (l > asdm::kMaxLabel) return ASDM_E_UNSUPPORTED;
            g->lab[id] = l;
            g->off[id + 1] = g->off[id] + d;
            maxlab = std::max(maxlab, l);
        } else if (c == 'e') {
            uint32_t a, b;
            if (!s.next_u32(a) || !s.next_u32(b) || a >= n || b >= n) return ASDM_E_IO;
            // every write stays below off[v + 1] <= m, whatever the degrees declared so far
            if (g->off[a + 1] > g->m || g->off[b + 1] > g->m) return ASDM_E_IO;
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
} ASDM_HOST_CATCH

int asdm_graph_from_edges(uint32_t n, const uint32_t *labels, uint64_t ne, const uint32_t *edges, asdm_graph **out) try {
    if (!out || (n && !labels) || (ne && !edges)) return ASDM_E_INVALID;
    *out = nullptr;
    if (2 * ne >= (1ull << 32)) return ASDM_E_UNSUPPORTED;
    for (uint64_t i = 0; i < 2 * ne; ++i)
        if (edges[i] >= n) return ASDM_E_INVALID;
    for (uint32_t v = 0; v < n; ++v)
        if (labels[v] > asdm::kMaxLabel) return ASDM_E_UNSUPPORTED;
    *out = asdm::graph_from_edges(n, labels, ne, edges);
    return ASDM_OK;
} ASDM_HOST_CATCH

int asdm_graph_write_text(const asdm_graph *g, const char *path) try {
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
} ASDM_HOST_CATCH

void asdm_graph_free(asdm_graph *g) {
    if (!g) return;
    asdm::release_device_graph(*g);
    delete g;
}

int asdm_graph_info(const asdm_graph *g, uint32_t *n, uint64_t *m, uint32_t *nlabels) try {
    if (!g) return ASDM_E_INVALID;
    if (n) *n = g->n;
    if (m) *m = g->m;
    if (nlabels) *nlabels = g->nlab;
    return ASDM_OK;
} ASDM_HOST_CATCH

int asdm_graph_csr(const asdm_graph *g, uint32_t *off, uint32_t *nbr, uint32_t *lab) try {
    if (!g) return ASDM_E_INVALID;
    if (off) std::memcpy(off, g->off.data(), sizeof(uint32_t) * ((size_t)g->n + 1));
    if (nbr && g->m) std::memcpy(nbr, g->nbr.data(), sizeof(uint32_t) * g->m);
    if (lab && g->n) std::memcpy(lab, g->lab.data(), sizeof(uint32_t) * g->n);
    return ASDM_OK;
} ASDM_HOST_CATCH

int asdm_graph_edges(const asdm_graph *g, uint32_t *edges, uint64_t cap, uint64_t *ne) try {
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
} ASDM_HOST_CATCH

}  // extern "C"
