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

// ---- cartesian arcs: the candidates of the path's last vertex reached from every candidate of its
// first vertex at once.  Rows r0 .. r0 + R of C(p0) in one chunk; a layer is R bitmaps over the
// positions of C(p_t), W_t words each.

// Layer 1: row r's candidate-edge list of the first arc (positions in C(p1)).  One wave per row.
__global__ void __launch_bounds__(kCartThreads) k_vr_first(uint32_t r0, uint32_t R, const uint64_t *__restrict__ ce_off,
                                                          const uint32_t *__restrict__ ce_val, uint32_t W1,
                                                          uint64_t *__restrict__ out) {
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t nw = (uint64_t)gridDim.x * (kCartThreads / 64);
    for (uint64_t r = (uint64_t)blockIdx.x * (kCartThreads / 64) + threadIdx.x / 64; r < R; r += nw) {
        const uint64_t lb = ce_off[r0 + r], le = ce_off[r0 + r + 1];
        uint64_t *row = out + r * W1;
        for (uint64_t e = lb + lane; e < le; e += 64) {
            const uint32_t q = ce_val[e];
            atomicOr((unsigned long long *)&row[q >> 6], 1ull << (q & 63));
        }
    }
}

// Layer t + 1 from layer t: one wave per (row, word) of layer t; each set position x expands to
// x's candidate-edge list of the arc p_t -> p_{t+1}.
__global__ void __launch_bounds__(kCartThreads) k_vr_hop(uint32_t R, const uint64_t *__restrict__ in, uint32_t Win,
                                                        const uint64_t *__restrict__ ce_off,
                                                        const uint32_t *__restrict__ ce_val, uint32_t Wout,
                                                        uint64_t *__restrict__ out) {
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t items = (uint64_t)R * Win, nw = (uint64_t)gridDim.x * (kCartThreads / 64);
    for (uint64_t it = (uint64_t)blockIdx.x * (kCartThreads / 64) + threadIdx.x / 64; it < items; it += nw) {
        const uint64_t r = it / Win, w = it % Win;
        uint64_t m = in[it];
        uint64_t *row = out + r * Wout;
        while (m) {  // uniform across the wave
            const uint32_t x = (uint32_t)(w * 64 + __builtin_ctzll(m));
            m &= m - 1;
            for (uint64_t e = ce_off[x] + lane; e < ce_off[x + 1]; e += 64) {
                const uint32_t q = ce_val[e];
                atomicOr((unsigned long long *)&row[q >> 6], 1ull << (q & 63));
            }
        }
    }
}

// Position in C(pk) of the candidate of p0 at row r when it is also a candidate of pk (its bit is
// left out of the row), else UINT32_MAX.
__device__ inline uint32_t own_position(const uint32_t *__restrict__ cand0, uint32_t r, const uint64_t *__restrict__ bits_l,
                                        const uint32_t *__restrict__ rank_l) {
    const uint32_t v = cand0[r];
    if (!((bits_l[v >> 6] >> (v & 63)) & 1)) return UINT32_MAX;
    return rank_l[v >> 6] + (uint32_t)__popcll(bits_l[v >> 6] & ((1ull << (v & 63)) - 1));
}

// cnt[r0 + r] = positions in row r other than the row's own vertex.  One wave per row.
__global__ void __launch_bounds__(kCartThreads) k_vr_count(uint32_t r0, uint32_t R, const uint64_t *__restrict__ layer,
                                                          uint32_t W, const uint32_t *__restrict__ cand0,
                                                          const uint64_t *__restrict__ bits_l,
                                                          const uint32_t *__restrict__ rank_l, uint32_t *__restrict__ cnt) {
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t nw = (uint64_t)gridDim.x * (kCartThreads / 64);
    for (uint64_t r = (uint64_t)blockIdx.x * (kCartThreads / 64) + threadIdx.x / 64; r < R; r += nw) {
        const uint64_t *row = layer + r * W;
        const uint32_t qv = own_position(cand0, r0 + (uint32_t)r, bits_l, rank_l);
        uint64_t c = 0;
        for (uint32_t w = lane; w < W; w += 64) {
            uint64_t m = row[w];
            if (w == (qv >> 6)) m &= ~(1ull << (qv & 63));
            c += __popcll(m);
        }
        c = group_sum<64>(c);
        if (lane == 0) cnt[r0 + r] = (uint32_t)c;
    }
}

// Writes row r's positions (without its own vertex) in ascending order at off[r0 + r]: the wave
// walks the row's words 64 at a time with a prefix over their popcounts.
__global__ void __launch_bounds__(kCartThreads) k_vr_fill(uint32_t r0, uint32_t R, const uint64_t *__restrict__ layer,
                                                         uint32_t W, const uint32_t *__restrict__ cand0,
                                                         const uint64_t *__restrict__ bits_l,
                                                         const uint32_t *__restrict__ rank_l, const uint64_t *__restrict__ off,
                                                         uint32_t *__restrict__ val) {
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t nw = (uint64_t)gridDim.x * (kCartThreads / 64);
    for (uint64_t r = (uint64_t)blockIdx.x * (kCartThreads / 64) + threadIdx.x / 64; r < R; r += nw) {
        const uint64_t *row = layer + r * W;
        const uint32_t qv = own_position(cand0, r0 + (uint32_t)r, bits_l, rank_l);
        uint64_t o = off[r0 + r];
        for (uint32_t w0 = 0; w0 < W; w0 += 64) {
            const uint32_t w = w0 + lane;
            uint64_t m = w < W ? row[w] : 0;
            if (w == (qv >> 6)) m &= ~(1ull << (qv & 63));
            const uint64_t c = (uint64_t)__popcll(m);
            const uint64_t inc = wave_inclusive_scan(c);
            uint64_t k = o + inc - c;
            while (m) {
                val[k++] = w * 64 + (uint32_t)__builtin_ctzll(m);
                m &= m - 1;
            }
            o += __shfl(inc, 63, 64);
        }
    }
}

// off[i] -= lo for n offsets (rebasing a copied range of list offsets).
__global__ void k_rebase(uint64_t *__restrict__ off, uint64_t n, uint64_t lo) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) off[i] -= lo;
}

}  // namespace

// Candidate-edge lists of a cartesian arc p0 -> pk along `path`: for every position r of C(p0),
// the positions in C(pk) reached from C(p0)[r], without C(p0)[r] itself, ascending.  off[n0 + 1]
// receives global offsets starting at vbase (off[0] = vbase); the values are allocated here.
static int cartesian_lists(asdm_ctx *ctx, const asdm_cs *cs, const std::vector<uint32_t> &path, uint64_t vbase,
                           uint64_t *off, uint32_t **vals_out, uint64_t *nvals) {
    const uint32_t p0 = path[0], pk = path.back();
    const uint32_t n0 = cs->csize[p0];
    std::vector<uint32_t> arcs, W;
    for (size_t t = 0; t + 1 < path.size(); ++t) arcs.push_back(cs->arc_index(path[t], path[t + 1]));
    for (size_t t = 0; t < path.size(); ++t) W.push_back(std::max<uint32_t>(1, (cs->csize[path[t]] + 63) / 64));
    const uint32_t Wmax = *std::max_element(W.begin() + 1, W.end());
    // rows per chunk: two layers of R x Wmax words within 512 MiB
    const uint32_t R = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(n0, (256ull << 20) / (8ull * Wmax)));
    uint64_t *la = nullptr, *lb = nullptr;
    uint32_t *cnt = nullptr, *vals = nullptr;
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
    int rc = alloc((void **)&la, 8ull * R * Wmax);
    if (!rc) rc = alloc((void **)&lb, 8ull * R * Wmax);
    if (!rc) rc = alloc((void **)&cnt, 4ull * std::max<uint32_t>(n0, 1));
    if (rc) return done(rc);
    const unsigned grid = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((R + 3) / 4, 8ull * ctx->num_cus));
    // the last layer of rows [r0, r0 + R) in la (or lb), recomputed for the count and the fill passes
    auto chunk = [&](uint32_t r0, uint32_t nr, uint64_t **last) -> int {
        uint64_t *cur = la, *nxt = lb;
        int r = fill_u64(ctx, cur, (uint64_t)nr * W[1], 0);
        if (r) return r;
        hipLaunchKernelGGL(k_vr_first, dim3(grid), dim3(kCartThreads), 0, ctx->stream, r0, nr,
                           cs->d_ce_off + cs->ce_base[arcs[0]], cs->d_ce_val, W[1], cur);
        if ((r = check_hip(ctx, hipGetLastError(), "cartesian arc first layer"))) return r;
        for (size_t t = 1; t < arcs.size(); ++t) {
            if ((r = fill_u64(ctx, nxt, (uint64_t)nr * W[t + 1], 0))) return r;
            const unsigned g2 = (unsigned)std::max<uint64_t>(
                1, std::min<uint64_t>(((uint64_t)nr * W[t] + 3) / 4, 8ull * ctx->num_cus));
            hipLaunchKernelGGL(k_vr_hop, dim3(g2), dim3(kCartThreads), 0, ctx->stream, nr, cur, W[t],
                               cs->d_ce_off + cs->ce_base[arcs[t]], cs->d_ce_val, W[t + 1], nxt);
            if ((r = check_hip(ctx, hipGetLastError(), "cartesian arc layer"))) return r;
            std::swap(cur, nxt);
        }
        *last = cur;
        return ASDM_OK;
    };
    const uint32_t Wl = W.back();
    for (uint32_t r0 = 0; r0 < n0 && !rc; r0 += R) {
        const uint32_t nr = std::min(R, n0 - r0);
        uint64_t *last = nullptr;
        rc = chunk(r0, nr, &last);
        if (!rc) {
            hipLaunchKernelGGL(k_vr_count, dim3(std::max(1u, (nr + 3) / 4)), dim3(kCartThreads), 0, ctx->stream, r0, nr, last,
                               Wl, cs->d_cand + cs->cand_off[p0], cs->d_bits + (size_t)pk * cs->W,
                               cs->d_rank + (size_t)pk * cs->W, cnt);
            rc = check_hip(ctx, hipGetLastError(), "cartesian arc count");
        }
    }
    if (rc) return done(rc);
    // off = vbase + exclusive scan of the counts
    uint64_t *doff = nullptr;
    if ((rc = alloc((void **)&doff, 8ull * (n0 + 1)))) return done(rc);
    if ((rc = scan_u32_to_u64(ctx, cnt, doff, n0))) return done(rc);
    std::vector<uint64_t> hoff(n0 + 1);
    rc = check_hip(ctx, hipMemcpyAsync(hoff.data(), doff, 8ull * (n0 + 1), hipMemcpyDeviceToHost, ctx->stream), "D2H");
    if (!rc) rc = check_hip(ctx, hipStreamSynchronize(ctx->stream), "sync");
    if (rc) return done(rc);
    const uint64_t total = hoff[n0];
    for (uint32_t r = 0; r <= n0; ++r) off[r] = vbase + hoff[r];
    if ((rc = dev_alloc(ctx, (void **)&vals, 4ull * std::max<uint64_t>(total, 1)))) return done(rc);
    for (uint32_t r0 = 0; r0 < n0 && !rc; r0 += R) {
        const uint32_t nr = std::min(R, n0 - r0);
        uint64_t *last = nullptr;
        rc = chunk(r0, nr, &last);
        if (!rc) {
            hipLaunchKernelGGL(k_vr_fill, dim3(grid), dim3(kCartThreads), 0, ctx->stream, r0, nr, last, Wl,
                               cs->d_cand + cs->cand_off[p0], cs->d_bits + (size_t)pk * cs->W,
                               cs->d_rank + (size_t)pk * cs->W, doff, vals);
            rc = check_hip(ctx, hipGetLastError(), "cartesian arc fill");
        }
    }
    if (rc) {
        dev_free(ctx, vals);
        return done(rc);
    }
    *vals_out = vals;
    *nvals = total;
    return done(ASDM_OK);
}

int cartesian_cs(const asdm_cs *cs, const std::vector<uint32_t> &order, asdm_cs **out, std::string &err) {
    *out = nullptr;
    const uint32_t nq = cs->nq;
    asdm_ctx *ctx = cs->ctx;
    // cartesian steps: depths with no earlier neighbour, and their parents (plan/estimator.cpp:9-39)
    std::vector<uint64_t> dist;
    std::vector<uint32_t> next;
    struct Step {
        uint32_t parent, u;
        std::vector<uint32_t> path;
    };
    std::vector<Step> steps;
    for (uint32_t d = 1; d < nq; ++d) {
        bool has = false;
        for (uint32_t j = 0; j < d && !has; ++j) has = cs->arc_index(order[j], order[d]) != UINT32_MAX;
        if (has) continue;
        if (dist.empty()) floyd_warshall(cs, dist, next);
        const uint32_t u = order[d];
        uint64_t best = std::numeric_limits<uint64_t>::max();
        uint32_t parent = UINT32_MAX;
        for (uint32_t j = 0; j < d; ++j)
            if (dist[(size_t)u * nq + order[j]] < best) {
                best = dist[(size_t)u * nq + order[j]];
                parent = order[j];
            }
        if (parent == UINT32_MAX) {
            err = "the query is not connected: no path reaches the vertex of depth " + std::to_string(d);
            return ASDM_E_UNSUPPORTED;
        }
        std::vector<uint32_t> path{parent};
        for (uint32_t a = parent; a != u;) {
            a = next[(size_t)a * nq + u];
            if (a == UINT32_MAX || path.size() > nq) {
                err = "the path to a cartesian step's vertex does not end (wrapped path products)";
                return ASDM_E_UNSUPPORTED;
            }
            path.push_back(a);
        }
        steps.push_back({parent, u, path});
    }
    if (steps.empty()) return ASDM_OK;
    (void)hipSetDevice(ctx->device);
    // the lists of the caller's arcs: their offsets (rebased to 0) and values, copied; then one
    // block of offsets and values per cartesian arc
    const uint64_t nitems = cs->ce_base[cs->narcs];
    uint64_t lo = 0, hi = 0;
    int rc = check_hip(ctx, hipMemcpyAsync(&lo, cs->d_ce_off, 8, hipMemcpyDeviceToHost, ctx->stream), "D2H");
    if (!rc) rc = check_hip(ctx, hipMemcpyAsync(&hi, cs->d_ce_off + nitems, 8, hipMemcpyDeviceToHost, ctx->stream), "D2H");
    if (!rc) rc = check_hip(ctx, hipStreamSynchronize(ctx->stream), "sync");
    if (rc) return rc;
    std::vector<std::vector<uint64_t>> voff(steps.size());
    std::vector<uint32_t *> vval(steps.size(), nullptr);
    std::vector<uint64_t> vn(steps.size(), 0);
    auto free_vals = [&]() {
        for (uint32_t *p : vval) dev_free(ctx, p);
    };
    uint64_t vbase = hi - lo, extra_items = 0;
    for (size_t k = 0; k < steps.size() && !rc; ++k) {
        voff[k].resize(cs->csize[steps[k].parent] + 1);
        rc = cartesian_lists(ctx, cs, steps[k].path, vbase, voff[k].data(), &vval[k], &vn[k]);
        vbase += vn[k];
        extra_items += cs->csize[steps[k].parent];
    }
    if (rc) {
        free_vals();
        err = ctx->err;
        return rc;
    }
    auto store = new asdm_cs::Store();
    store->refs = 1;
    const uint64_t total_items = nitems + extra_items;
    rc = dev_alloc(ctx, (void **)&store->d_ce_off, 8ull * (total_items + 1));
    if (!rc) rc = dev_alloc(ctx, (void **)&store->d_ce_val, 4ull * std::max<uint64_t>(vbase, 1));
    if (!rc) rc = check_hip(ctx, hipMemcpyAsync(store->d_ce_off, cs->d_ce_off, 8ull * (nitems + 1), hipMemcpyDeviceToDevice,
                                                ctx->stream), "D2D offsets");
    if (!rc && hi > lo)
        rc = check_hip(ctx, hipMemcpyAsync(store->d_ce_val, cs->d_ce_val + lo, 4ull * (hi - lo), hipMemcpyDeviceToDevice,
                                           ctx->stream), "D2D values");
    if (!rc) {
        hipLaunchKernelGGL(k_rebase, dim3(std::max(1u, (unsigned)std::min<uint64_t>((nitems + 256) / 256, 1u << 16))), dim3(256),
                           0, ctx->stream, store->d_ce_off, nitems + 1, lo);
        rc = check_hip(ctx, hipGetLastError(), "rebase offsets");
    }
    uint64_t item = nitems;
    for (size_t k = 0; k < steps.size() && !rc; ++k) {
        // voff[k][0] equals the previous block's end, already in place
        const uint32_t n0 = cs->csize[steps[k].parent];
        rc = check_hip(ctx, hipMemcpyAsync(store->d_ce_off + item + 1, voff[k].data() + 1, 8ull * n0, hipMemcpyHostToDevice,
                                           ctx->stream), "H2D offsets");
        if (!rc && vn[k])
            rc = check_hip(ctx, hipMemcpyAsync(store->d_ce_val + voff[k][0], vval[k], 4ull * vn[k], hipMemcpyDeviceToDevice,
                                               ctx->stream), "D2D values");
        item += n0;
    }
    if (!rc) rc = check_hip(ctx, hipStreamSynchronize(ctx->stream), "sync");
    free_vals();
    if (rc) {
        dev_free(ctx, store->d_ce_off);
        dev_free(ctx, store->d_ce_val);
        delete store;
        err = ctx->err;
        return rc;
    }
    // the extended candidate space: the caller's candidate sets and bitmaps, the new lists, and a
    // query CSR with the cartesian arcs after each parent's own arcs
    auto ext = new asdm_cs(*cs);
    ext->store = store;
    ext->base_store = cs->store;
    if (cs->store) cs->store->refs++;
    ext->d_ce_off = store->d_ce_off;
    ext->d_ce_val = store->d_ce_val;
    ext->ce_total = vbase;
    std::vector<std::vector<std::pair<uint32_t, uint32_t>>> adj(nq);  // (neighbour, old arc or UINT32_MAX - k)
    for (uint32_t u = 0; u < nq; ++u)
        for (uint32_t a = cs->qoff[u]; a < cs->qoff[u + 1]; ++a) adj[u].push_back({cs->qnbr[a], a});
    for (size_t k = 0; k < steps.size(); ++k) adj[steps[k].parent].push_back({steps[k].u, (uint32_t)(UINT32_MAX - k)});
    ext->qoff.assign(nq + 1, 0);
    ext->qnbr.clear();
    ext->arc_u.clear();
    ext->arc_u2.clear();
    ext->arc_weight.clear();
    ext->ce_base.clear();
    for (uint32_t u = 0; u < nq; ++u) {
        for (auto &[w, a] : adj[u]) {
            ext->qnbr.push_back(w);
            ext->arc_u.push_back(u);
            ext->arc_u2.push_back(w);
            if (a < cs->narcs) {
                ext->ce_base.push_back(cs->ce_base[a]);
                ext->arc_weight.push_back(cs->arc_weight[a]);
            } else {
                const size_t k = UINT32_MAX - a;
                uint64_t base = nitems;
                for (size_t k2 = 0; k2 < k; ++k2) base += cs->csize[steps[k2].parent];
                ext->ce_base.push_back(base);
                ext->arc_weight.push_back(vn[k]);
            }
        }
        ext->qoff[u + 1] = (uint32_t)ext->qnbr.size();
    }
    ext->narcs = (uint32_t)ext->qnbr.size();
    ext->ce_base.push_back(total_items);
    *out = ext;
    return ASDM_OK;
}

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
