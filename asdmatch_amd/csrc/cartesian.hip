// Query-graph paths and cartesian-step candidates (the candidates of a query vertex with no
// matched neighbour, reached from a matched vertex along a path of the query).
//
// Paths (host): the reference weighs each query arc with the total length of its candidate-edge
// lists, held in a 32-bit `ui` (relation/candidate_space.cpp:507-524), and runs Floyd-Warshall
// minimising the product of the weights in size_t, dist[i][i] starting at the maximum
// (relation/graph.cpp:341-372); reconstructPath follows the successor matrix
// (relation/candidate_space.cpp:529-537).  The products wrap mod 2^64 exactly as the reference's,
// so the successor matrix is the reference's, including the cycles it can hold when they wrap.
//
// Candidates (device): the reference walks every path of candidate edges from v depth first and
// merges the lists of the last arc (plan/estimator.cpp:41-102).  The union of their ends is the
// set reached layer by layer, so the walk here keeps one bitmap over data ids per layer: a wave
// per set word of layer t expands each set vertex's candidate-edge list into layer t+1 with
// atomic ORs, and the last layer, without v, is compacted into ascending ids.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <limits>

#include "engine.h"
#include "kernels_common.h"

namespace asdm {

namespace {

constexpr int kCartThreads = 256;

void floyd_warshall(const asdm_cs *cs, std::vector<uint64_t> &dist, std::vector<uint32_t> &next) {
    const uint32_t nq = cs->nq;
    const uint64_t INF = std::numeric_limits<uint64_t>::max();
    dist.assign((size_t)nq * nq, INF);
    next.assign((size_t)nq * nq, UINT32_MAX);
    for (uint32_t i = 0; i < nq; ++i) next[(size_t)i * nq + i] = i;
    for (uint32_t a = 0; a < cs->narcs; ++a) {
        const size_t k = (size_t)cs->arc_u[a] * nq + cs->arc_u2[a];
        dist[k] = (uint32_t)cs->arc_weight[a];  // the reference's 32-bit weight
        next[k] = cs->arc_u2[a];
    }
    // dist[i][k] is read again for every j, as the reference does: once a product has wrapped to a
    // small value (even 0), dist[i][k] itself can drop within iteration k (at j = k), and the
    // later j of the same row see the new value
    for (uint32_t k = 0; k < nq; ++k)
        for (uint32_t i = 0; i < nq; ++i)
            for (uint32_t j = 0; j < nq; ++j) {
                const uint64_t a = dist[(size_t)i * nq + k], b = dist[(size_t)k * nq + j];
                if (a == INF || b == INF) continue;
                if (a * b < dist[(size_t)i * nq + j]) {
                    dist[(size_t)i * nq + j] = a * b;
                    next[(size_t)i * nq + j] = next[(size_t)i * nq + k];
                }
            }
}

// Layer 0: v's bit, when v is a candidate of the path's first vertex.
__global__ void k_cart_seed(const uint64_t *__restrict__ bits0, uint32_t v, uint64_t *__restrict__ layer) {
    if (threadIdx.x == 0 && blockIdx.x == 0 && ((bits0[v >> 6] >> (v & 63)) & 1)) layer[v >> 6] = 1ull << (v & 63);
}

// One wave per word of `in` (bitmap over data ids, all of them candidates of the arc's source):
// each set vertex x's list (positions in C(b)) goes into `out` as data ids.
__global__ void __launch_bounds__(kCartThreads) k_cart_step(const uint64_t *__restrict__ in, uint32_t W,
                                                           const uint64_t *__restrict__ bits_a,
                                                           const uint32_t *__restrict__ rank_a,
                                                           const uint64_t *__restrict__ ce_off,
                                                           const uint32_t *__restrict__ ce_val,
                                                           const uint32_t *__restrict__ cand_b, uint64_t *__restrict__ out) {
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t nw = (uint64_t)gridDim.x * (kCartThreads / 64);
    for (uint64_t w = (uint64_t)blockIdx.x * (kCartThreads / 64) + threadIdx.x / 64; w < W; w += nw) {
        uint64_t m = in[w];
        while (m) {  // uniform across the wave
            const uint32_t x = (uint32_t)(w * 64 + __builtin_ctzll(m));
            m &= m - 1;
            const uint32_t p = rank_a[x >> 6] + (uint32_t)__popcll(bits_a[x >> 6] & ((1ull << (x & 63)) - 1));
            const uint64_t lb = ce_off[p], le = ce_off[p + 1];
            for (uint64_t e = lb + lane; e < le; e += 64) {
                const uint32_t y = cand_b[ce_val[e]];
                atomicOr((unsigned long long *)&out[y >> 6], 1ull << (y & 63));
            }
        }
    }
}

__global__ void k_cart_count(const uint64_t *__restrict__ layer, uint32_t W, uint32_t v, uint32_t *__restrict__ cnt) {
    for (uint32_t w = blockIdx.x * blockDim.x + threadIdx.x; w < W; w += gridDim.x * blockDim.x) {
        uint64_t m = layer[w];
        if (w == (v >> 6)) m &= ~(1ull << (v & 63));
        cnt[w] = (uint32_t)__popcll(m);
    }
}

__global__ void k_cart_fill(const uint64_t *__restrict__ layer, uint32_t W, uint32_t v, const uint64_t *__restrict__ off,
                            uint32_t *__restrict__ ids) {
    for (uint32_t w = blockIdx.x * blockDim.x + threadIdx.x; w < W; w += gridDim.x * blockDim.x) {
        uint64_t m = layer[w];
        if (w == (v >> 6)) m &= ~(1ull << (v & 63));
        uint64_t o = off[w];
        while (m) {
            ids[o++] = (uint32_t)(w * 64 + __builtin_ctzll(m));
            m &= m - 1;
        }
    }
}

}  // namespace

}  // namespace asdm

using namespace asdm;

extern "C" {

int asdm_cs_query_paths(const asdm_cs *cs, uint64_t *dist, uint32_t *next) try {
    if (!cs || (!dist && !next)) return ASDM_E_INVALID;
    std::vector<uint64_t> d;
    std::vector<uint32_t> n;
    floyd_warshall(cs, d, n);
    if (dist) std::memcpy(dist, d.data(), 8 * d.size());
    if (next) std::memcpy(next, n.data(), 4 * n.size());
    return ASDM_OK;
} ASDM_HOST_CATCH

int asdm_cs_query_path(const asdm_cs *cs, uint32_t i, uint32_t j, uint32_t *path, uint32_t cap, uint32_t *len) try {
    if (!cs || !len || i >= cs->nq || j >= cs->nq || (!path && cap)) return ASDM_E_INVALID;
    std::vector<uint64_t> d;
    std::vector<uint32_t> n;
    floyd_warshall(cs, d, n);
    const uint32_t nq = cs->nq;
    *len = 0;
    if (n[(size_t)i * nq + j] == UINT32_MAX) return ASDM_OK;  // no path
    std::vector<uint32_t> p{i};
    for (uint32_t a = i; a != j;) {
        a = n[(size_t)a * nq + j];
        if (a == UINT32_MAX || p.size() > nq) {
            // wrapped products left a cycle in the successor matrix: the reference's
            // reconstructPath would not return
            return cs->ctx ? cs->ctx->fail(ASDM_E_UNSUPPORTED, "the successor matrix holds a cycle toward this vertex "
                                                               "(wrapped path products); the path does not end")
                           : ASDM_E_UNSUPPORTED;
        }
        p.push_back(a);
    }
    if (p.size() > cap) return ASDM_E_BUDGET;
    std::copy(p.begin(), p.end(), path);
    *len = (uint32_t)p.size();
    return ASDM_OK;
} ASDM_HOST_CATCH

int asdm_cartesian_candidates(asdm_ctx *ctx, const asdm_cs *cs, uint32_t v, const uint32_t *path, uint32_t len,
                              uint32_t *out, uint32_t cap, uint32_t *n) try {
    if (!ctx || !cs || !path || !n || len < 2 || (!out && cap)) return ASDM_E_INVALID;
    for (uint32_t t = 0; t < len; ++t)
        if (path[t] >= cs->nq) return ctx->fail(ASDM_E_INVALID, "path vertex out of range");
    *n = 0;
    if (v >= cs->nv) return ASDM_OK;  // not a candidate of path[0]: reaches nothing
    (void)hipSetDevice(ctx->device);
    const uint32_t W = cs->W;
    uint64_t *layers = nullptr, *off = nullptr;
    uint32_t *cnt = nullptr, *ids = nullptr;
    std::vector<void *> tmp;
    auto alloc = [&](void **p, size_t bytes) {
        const int rc = dev_alloc(ctx, p, std::max<size_t>(bytes, 8));
        if (!rc) tmp.push_back(*p);
        return rc;
    };
    auto done = [&](int rc) {
        for (void *p : tmp) dev_free(ctx, p);
        return rc;
    };
    int rc = alloc((void **)&layers, 16ull * W);  // two layers, alternating
    if (!rc) rc = alloc((void **)&cnt, 4ull * W);
    if (!rc) rc = alloc((void **)&off, 8ull * (W + 1));
    if (!rc) rc = fill_u64(ctx, layers, 2ull * W, 0);
    if (rc) return done(rc);
    uint64_t *cur = layers, *nxt = layers + W;
    const uint32_t u0 = path[0];
    hipLaunchKernelGGL(k_cart_seed, dim3(1), dim3(64), 0, ctx->stream, cs->d_bits + (size_t)u0 * W, v, cur);
    const unsigned wgrid = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((W + 3) / 4, 8ull * ctx->num_cus));
    bool reach = true;
    for (uint32_t t = 0; t + 1 < len && reach; ++t) {
        const uint32_t a = path[t], b = path[t + 1];
        const uint32_t arc = cs->arc_index(a, b);
        if (arc == UINT32_MAX) {
            reach = false;  // the reference's list of a non-arc is empty
            break;
        }
        if ((rc = fill_u64(ctx, nxt, W, 0))) return done(rc);
        KTimer kt(ctx, "cartesian_step");
        hipLaunchKernelGGL(k_cart_step, dim3(wgrid), dim3(kCartThreads), 0, ctx->stream, cur, W,
                           cs->d_bits + (size_t)a * W, cs->d_rank + (size_t)a * W, cs->d_ce_off + cs->ce_base[arc],
                           cs->d_ce_val, cs->d_cand + cs->cand_off[b], nxt);
        kt.stop(8ull * W * 2 + 16ull * cs->csize[a] + 8ull * cs->arc_weight[arc]);
        if ((rc = check_hip(ctx, hipGetLastError(), "cartesian step"))) return done(rc);
        std::swap(cur, nxt);
    }
    if (!reach) return done(ASDM_OK);
    const unsigned cgrid = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((W + 255) / 256, 8ull * ctx->num_cus));
    hipLaunchKernelGGL(k_cart_count, dim3(cgrid), dim3(256), 0, ctx->stream, cur, W, v, cnt);
    if ((rc = check_hip(ctx, hipGetLastError(), "cartesian count"))) return done(rc);
    if ((rc = scan_u32_to_u64(ctx, cnt, off, W))) return done(rc);
    uint64_t total = 0;
    rc = check_hip(ctx, hipMemcpyAsync(&total, off + W, 8, hipMemcpyDeviceToHost, ctx->stream), "D2H");
    if (!rc) rc = check_hip(ctx, hipStreamSynchronize(ctx->stream), "sync");
    if (rc) return done(rc);
    *n = (uint32_t)total;
    if (total > cap) return done(ctx->fail(ASDM_E_BUDGET, "more cartesian candidates than the output holds"));
    if (!total) return done(ASDM_OK);
    if ((rc = alloc((void **)&ids, 4 * total))) return done(rc);
    hipLaunchKernelGGL(k_cart_fill, dim3(cgrid), dim3(256), 0, ctx->stream, cur, W, v, off, ids);
    if ((rc = check_hip(ctx, hipGetLastError(), "cartesian fill"))) return done(rc);
    rc = check_hip(ctx, hipMemcpyAsync(out, ids, 4 * total, hipMemcpyDeviceToHost, ctx->stream), "D2H");
    if (!rc) rc = check_hip(ctx, hipStreamSynchronize(ctx->stream), "sync");
    return done(rc);
} ASDM_CTX_CATCH(ctx)

}  // extern "C"
