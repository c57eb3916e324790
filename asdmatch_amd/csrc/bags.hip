// Bag-table engine (bags.h): merged-block domains, relation CSRs, width-two variable elimination
// over dense bag tables, and the local-count scatter.  All values are uint64 mod 2^64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <map>
#include <memory>
#include <set>

#include "bags.h"
#include "engine.h"
#include "kernels_common.h"

namespace asdm {

namespace {

constexpr int kThreads = 256;
// Largest bytes of factor results kept per query for reuse (also at most a third of free HBM).
constexpr uint64_t kBagCacheBytes = 64ull << 30;

unsigned grid_for(uint64_t items, uint64_t per_block) {
    const uint64_t g = (items + per_block - 1) / per_block;
    return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(g, 1u << 20));
}

// One atomic per wave: the wave's summed algorithmic bytes into *acc.
__device__ inline void wave_add_bytes(uint64_t b, unsigned long long *acc) {
    b = group_sum<64>(b);
    if ((threadIdx.x & 63) == 0 && b) atomicAdd(acc, (unsigned long long)b);
}

__device__ inline uint32_t bit_rank(const uint64_t *__restrict__ bits, const uint32_t *__restrict__ rank, uint32_t v) {
    const uint64_t w = bits[v >> 6];
    return rank[v >> 6] + (uint32_t)__popcll(w & ((1ull << (v & 63)) - 1));
}
__device__ inline bool bit_test(const uint64_t *__restrict__ bits, uint32_t v) { return (bits[v >> 6] >> (v & 63)) & 1ull; }

struct Members {
    uint32_t n;
    uint32_t u[kMaxQ];
};

// Domain of a merged block: AND of its members' candidate bitmaps, popcount per word.
__global__ void k_dom_words(const uint64_t *__restrict__ bits, uint32_t W, Members m, uint64_t *__restrict__ out,
                            uint32_t *__restrict__ cnt) {
    for (uint32_t w = blockIdx.x * blockDim.x + threadIdx.x; w < W; w += gridDim.x * blockDim.x) {
        uint64_t x = ~0ull;
        for (uint32_t k = 0; k < m.n; ++k) x &= bits[(size_t)m.u[k] * W + w];
        out[w] = x;
        cnt[w] = (uint32_t)__popcll(x);
    }
}

// Rank prefix per word and the ascending data ids of the domain.
__global__ void k_dom_fill(const uint64_t *__restrict__ bits, uint32_t W, const uint64_t *__restrict__ off,
                           uint32_t *__restrict__ rank, uint32_t *__restrict__ ids) {
    for (uint32_t w = blockIdx.x * blockDim.x + threadIdx.x; w < W; w += gridDim.x * blockDim.x) {
        uint64_t o = off[w];
        rank[w] = (uint32_t)o;
        for (uint64_t x = bits[w]; x; x &= x - 1) ids[o++] = w * 64u + (uint32_t)__builtin_ctzll(x);
    }
}

// Relation CSRs of quotient edges A -> B from the candidate-edge lists of query arcs a -> b
// (a in A, b in B): row i of a relation (data vertex ids_a[i]) holds the D(B) positions of its
// neighbours in D(B), ascending.  One build covers the relations of many jobs: their rows are
// numbered one after the other (job j: rows row0[j] .. row0[j + 1]) and share one value array.
struct RelJob {
    const uint32_t *ids_a;   // D(A), ascending data ids
    const uint64_t *bits_a;  // membership bitmap and rank prefix of C(a)
    const uint32_t *rank_a;
    const uint64_t *ce_off;  // candidate-edge offsets of arc a -> b (indexed by position in C(a))
    const uint32_t *cand_b;  // C(b)
    const uint64_t *bits_B;  // membership bitmap and rank prefix of D(B)
    const uint32_t *rank_B;
};

// Flat relation build.  A hub's candidate-edge list can hold thousands of entries, and a lane
// group that walks one row alone sets the time of the whole launch; here the work is spread over
// entries instead.  k_rel_len: each row's list (source offset and length); a scan of the lengths
// numbers the entries of all rows; k_rel_keep: per entry, whether its data vertex is in D(B) and
// its position there; a scan of the keep flags gives every kept entry its slot, in list order
// within each row; k_rel_place writes the positions and k_rel_off the rows' offsets.
__global__ void k_rel_len(const RelJob *__restrict__ jobs, const uint64_t *__restrict__ row0, uint32_t nj,
                          uint64_t *__restrict__ src, uint32_t *__restrict__ len) {
    const uint64_t rows = row0[nj];
    for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < rows; r += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t lo = 0, hi = nj - 1;  // the job holding row r: the last j with row0[j] <= r
        while (lo < hi) {
            const uint32_t mid = (lo + hi + 1) >> 1;
            if (row0[mid] <= r) lo = mid;
            else hi = mid - 1;
        }
        const RelJob J = jobs[lo];
        const uint32_t p = bit_rank(J.bits_a, J.rank_a, J.ids_a[r - row0[lo]]);
        const uint64_t lb = J.ce_off[p], le = J.ce_off[p + 1];
        src[r] = lb;
        len[r] = (uint32_t)(le - lb);
    }
}

constexpr int kRelPer = 8;  // consecutive entries per lane in k_rel_keep
__global__ void k_rel_keep(const RelJob *__restrict__ jobs, const uint64_t *__restrict__ row0, uint32_t nj,
                           const uint64_t *__restrict__ src, const uint64_t *__restrict__ ebase,
                           const uint32_t *__restrict__ ce_val, uint32_t *__restrict__ keep, uint32_t *__restrict__ pos,
                           unsigned long long *__restrict__ nbytes) {
    const uint64_t rows = row0[nj], E = ebase[rows];
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x * kRelPer;
    uint64_t bytes = 0;
    for (uint64_t e0 = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) * kRelPer; e0 < E; e0 += stride) {
        // the row of entry e0 (last r with ebase[r] <= e0), then its job
        uint32_t lo = 0, hi = (uint32_t)rows - 1;
        while (lo < hi) {
            const uint32_t mid = (lo + hi + 1) >> 1;
            if (ebase[mid] <= e0) lo = mid;
            else hi = mid - 1;
        }
        uint64_t r = lo;
        uint32_t jl = 0, jh = nj - 1;
        while (jl < jh) {
            const uint32_t mid = (jl + jh + 1) >> 1;
            if (row0[mid] <= r) jl = mid;
            else jh = mid - 1;
        }
        const uint64_t e1 = min(E, e0 + kRelPer);
        uint32_t w[kRelPer];
        uint64_t rr[kRelPer];
        uint32_t jj[kRelPer];
#pragma unroll
        for (int q = 0; q < kRelPer; ++q) {  // the entries' rows and jobs, walking forward
            const uint64_t e = e0 + q;
            if (e < e1) {
                while (ebase[r + 1] <= e) ++r;
                while (jl + 1 < nj && row0[jl + 1] <= r) ++jl;
            }
            rr[q] = r;
            jj[q] = jl;
        }
#pragma unroll
        for (int q = 0; q < kRelPer; ++q)  // the loads of all eight entries in flight together
            w[q] = e0 + q < e1 ? jobs[jj[q]].cand_b[ce_val[src[rr[q]] + (e0 + q - ebase[rr[q]])]] : 0;
#pragma unroll
        for (int q = 0; q < kRelPer; ++q) {
            if (e0 + q >= e1) break;
            const RelJob &J = jobs[jj[q]];
            const bool ok = bit_test(J.bits_B, w[q]);
            keep[e0 + q] = ok;
            pos[e0 + q] = ok ? bit_rank(J.bits_B, J.rank_B, w[q]) : 0;
        }
        bytes += 16ull * (e1 - e0);
    }
    wave_add_bytes(bytes, nbytes);
}

__global__ void k_rel_place(const uint32_t *__restrict__ keep, const uint32_t *__restrict__ pos,
                            const uint64_t *__restrict__ kpos, uint64_t E, uint32_t *__restrict__ val) {
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < E; e += (uint64_t)gridDim.x * blockDim.x)
        if (keep[e]) val[kpos[e]] = pos[e];
}

__global__ void k_rel_off(const uint64_t *__restrict__ ebase, const uint64_t *__restrict__ kpos, uint64_t rows,
                          uint64_t *__restrict__ off) {
    for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r <= rows; r += (uint64_t)gridDim.x * blockDim.x)
        off[r] = kpos[ebase[r]];
}

__global__ void k_vec_mul(uint64_t *__restrict__ u, const uint64_t *__restrict__ t, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) u[i] *= t[i];
}

// Row r of a CSR with n rows that holds entry e (off[0] <= e < off[n]): the last r with off[r] <= e.
__device__ inline uint32_t row_of(const uint64_t *__restrict__ off, uint32_t n, uint64_t e) {
    uint32_t lo = 0, hi = n - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (off[mid] <= e) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// t[i] = sum over row i of the relation of w x[col], w the entry's weight (1 when the weight array
// is null), flat over the entries: each lane takes kSpmvPer consecutive entries, finds the row of
// its first one by binary search, and adds each row's partial sum into t (zeroed before) with one
// atomic per (lane, row) piece.  A hub row's entries are spread over many lanes, so no lane walks
// a long row alone (16 lanes per row spent most of a launch on the longest row); the u64 atomics
// wrap mod 2^64 like the sum itself.
constexpr int kSpmvPer = 8;
__global__ void k_spmv(const uint64_t *__restrict__ off, const uint32_t *__restrict__ val, const uint64_t *__restrict__ w,
                       uint32_t n, const uint64_t *__restrict__ x, uint64_t *__restrict__ t,
                       unsigned long long *__restrict__ nbytes) {
    const uint64_t base = off[0], nnz = off[n] - base;  // weights are indexed from the relation's first entry
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x * kSpmvPer;
    uint64_t bytes = 0;
    for (uint64_t s0 = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) * kSpmvPer; s0 < nnz; s0 += stride) {
        const uint32_t m = (uint32_t)min<uint64_t>(kSpmvPer, nnz - s0);
        uint64_t prod[kSpmvPer];
        uint32_t col[kSpmvPer];
#pragma unroll
        for (int q = 0; q < kSpmvPer; ++q) col[q] = val[base + s0 + (q < (int)m ? q : 0)];
#pragma unroll
        for (int q = 0; q < kSpmvPer; ++q) prod[q] = q < (int)m ? x[col[q]] * (w ? w[s0 + q] : 1ull) : 0;
        uint32_t r = row_of(off, n, base + s0);
        uint64_t rend = off[r + 1], acc = 0;
#pragma unroll
        for (int q = 0; q < kSpmvPer; ++q) {
            if (q < (int)m) {
                while (base + s0 + q >= rend) {  // the entry starts a later row: flush this one
                    if (acc) atomicAdd((unsigned long long *)&t[r], (unsigned long long)acc);
                    acc = 0;
                    rend = off[++r + 1];
                }
                acc += prod[q];
            }
        }
        if (acc) atomicAdd((unsigned long long *)&t[r], (unsigned long long)acc);
        bytes += (w ? 20 : 12) * (uint64_t)m;
    }
    wave_add_bytes(bytes, nbytes);
}

// t[i] = sum_j F[i][j] x[j]  (F row-major rows x cols): one wave per row.
__global__ void k_dense_mv(const uint64_t *__restrict__ F, uint32_t rows, uint32_t cols, const uint64_t *__restrict__ x,
                           uint64_t *__restrict__ t, unsigned long long *__restrict__ nbytes) {
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t nw = (uint64_t)gridDim.x * (blockDim.x / 64);
    uint64_t bytes = 0;
    for (uint64_t r = (uint64_t)blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64; r < rows; r += nw) {
        uint64_t s = 0;
        for (uint32_t j = lane; j < cols; j += 64) s += F[r * cols + j] * x[j];
        s = group_sum<64>(s);
        if (lane == 0) t[r] = s;
        bytes += lane == 0 ? 16ull * cols + 8 : 0;
    }
    wave_add_bytes(bytes, nbytes);
}

// t[c] += sum over a chunk of rows r of F[r][c] x[r] (the transposed product): blocks tile the
// columns (coalesced rows) and the rows; t is zeroed before, and the u64 atomics wrap mod 2^64
// like the sum itself.
constexpr uint32_t kMvtRows = 64;
__global__ void k_dense_mv_t(const uint64_t *__restrict__ F, uint32_t rows, uint32_t cols, const uint64_t *__restrict__ x,
                             uint64_t *__restrict__ t, unsigned long long *__restrict__ nbytes) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t r0 = blockIdx.y * kMvtRows, r1 = min(rows, r0 + kMvtRows);
    uint64_t bytes = 0;
    if (c < cols && r0 < r1) {
        uint64_t s = 0;
        for (uint32_t r = r0; r < r1; ++r) s += F[(uint64_t)r * cols + c] * x[r];
        if (s) atomicAdd((unsigned long long *)&t[c], (unsigned long long)s);
        bytes = 16ull * (r1 - r0) + 8;
    }
    wave_add_bytes(bytes, nbytes);
}

// acc[0] += sum of x
__global__ void k_sum(const uint64_t *__restrict__ x, uint32_t n, unsigned long long *__restrict__ acc) {
    uint64_t s = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) s += x[i];
    s = group_sum<64>(s);
    if ((threadIdx.x & 63) == 0 && s) atomicAdd(acc, (unsigned long long)s);
}

// scal *= acc; acc = 0
__global__ void k_scal_fold(uint64_t *__restrict__ scal, uint64_t *__restrict__ acc) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        scal[0] *= acc[0];
        acc[0] = 0;
    }
}

// total += mu * scal
__global__ void k_total_add(unsigned long long *__restrict__ total, const uint64_t *__restrict__ scal, int64_t mu) {
    if (threadIdx.x == 0 && blockIdx.x == 0) total[0] += (unsigned long long)((uint64_t)mu * scal[0]);
}

// The product of two edge relations has two forms.  Its terms are the pairs (entry x -> j of
// R(X->v), entry j -> y of R(v->Y)) with a nonzero weight u[j] wx; when they are few against the
// table's cells (the tables of sparse relations are mostly zeros), the table is cleared with
// streaming stores and the terms are added into it with global atomics (k_ee_zero, k_ee_scatter);
// otherwise the rows are accumulated in LDS (k_ee).  k_ee_terms counts the terms on the device and
// every kernel of the product reads the count, so the form is chosen without a host round trip.
constexpr uint64_t kEeSparse = 64;  // scatter form while terms * kEeSparse < cells
__device__ inline bool ee_scatter_form(int form, const unsigned long long *terms, uint64_t cells) {
    return form ? form == 2 : terms[0] * kEeSparse < cells;  // form: 0 by the count, 1 LDS, 2 scatter
}

// terms[0] += the number of the product's terms (zeroed before)
__global__ void k_ee_terms(const uint64_t *__restrict__ xoff, const uint32_t *__restrict__ xval,
                           const uint64_t *__restrict__ xw, const uint64_t *__restrict__ voff,
                           const uint64_t *__restrict__ u, uint32_t nx, unsigned long long *__restrict__ terms) {
    const uint64_t base = xoff[0], nnz = xoff[nx] - base;
    uint64_t s = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nnz; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t j = xval[base + i];
        if (u[j] * (xw ? xw[i] : 1ull)) s += voff[j + 1] - voff[j];
    }
    s = group_sum<64>(s);
    if ((threadIdx.x & 63) == 0 && s) atomicAdd(terms, (unsigned long long)s);
}

typedef unsigned long long u64x2 __attribute__((ext_vector_type(2)));

// Scatter form, first step: G = 0 (G is 256-byte aligned: 16-byte streaming stores, then the odd tail).
__global__ void k_ee_zero(uint64_t *__restrict__ G, uint64_t cells, const unsigned long long *__restrict__ terms,
                          int form) {
    if (!ee_scatter_form(form, terms, cells)) return;
    const uint64_t pairs = cells / 2, tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const u64x2 z = {0ull, 0ull};
    for (uint64_t i = tid; i < pairs; i += (uint64_t)gridDim.x * blockDim.x) __builtin_nontemporal_store(z, (u64x2 *)G + i);
    if (tid == 0 && (cells & 1)) G[cells - 1] = 0;
}

// Scatter form, second step: G[x][y] += u[j] wx wv for every term.  A wave takes 64 consecutive
// entries of R(X->v), one per lane, and then the terms of their lists together, 64 at a time: the
// lane of term t finds the entry that owns it by a binary search over the lanes' running list
// lengths (cross-lane reads), so a long list is spread over the wave instead of held by one lane.
__global__ void k_ee_scatter(const uint64_t *__restrict__ xoff, const uint32_t *__restrict__ xval,
                             const uint64_t *__restrict__ xw, const uint64_t *__restrict__ voff,
                             const uint32_t *__restrict__ vval, const uint64_t *__restrict__ vw,
                             const uint64_t *__restrict__ u, uint32_t nx, uint32_t ny, uint64_t *__restrict__ G,
                             const unsigned long long *__restrict__ terms, int form,
                             unsigned long long *__restrict__ nbytes) {
    if (!ee_scatter_form(form, terms, (uint64_t)nx * ny)) return;
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t xbase = xoff[0], vbase = voff[0], nnz = xoff[nx] - xbase;
    const uint64_t nwaves = (uint64_t)gridDim.x * (blockDim.x / 64);
    uint64_t bytes = 0;
    for (uint64_t c0 = ((uint64_t)blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64) * 64; c0 < nnz; c0 += nwaves * 64) {
        const uint64_t i = c0 + lane;
        uint64_t len = 0, b2 = 0, w = 0;
        uint32_t row = 0;
        if (i < nnz) {
            const uint32_t j = xval[xbase + i];
            w = u[j] * (xw ? xw[i] : 1ull);
            if (w) {
                b2 = voff[j];
                len = voff[j + 1] - b2;
                row = row_of(xoff, nx, xbase + i);
            }
        }
        const uint64_t inc = wave_inclusive_scan(len);
        const uint64_t total = __shfl(inc, 63, 64);
        for (uint64_t t0 = 0; t0 < total; t0 += 64) {
            const uint64_t t = t0 + lane;
            const uint64_t tq = t < total ? t : total - 1;
            uint32_t lo = 0, hi = 63;  // the first lane whose running length exceeds tq
            while (lo < hi) {
                const uint32_t mid = (lo + hi) >> 1;
                if (__shfl(inc, (int)mid, 64) > tq) hi = mid;
                else lo = mid + 1;
            }
            const uint64_t o_inc = __shfl(inc, (int)lo, 64), o_len = __shfl(len, (int)lo, 64);
            const uint64_t o_b2 = __shfl(b2, (int)lo, 64), o_w = __shfl(w, (int)lo, 64);
            const uint32_t o_row = __shfl(row, (int)lo, 64);
            if (t < total) {
                const uint64_t f = o_b2 + (t - (o_inc - o_len));
                const uint64_t wf = vw ? o_w * vw[f - vbase] : o_w;
                if (wf) atomicAdd((unsigned long long *)&G[(uint64_t)o_row * ny + vval[f]], (unsigned long long)wf);
                bytes += vw ? 28 : 20;
            }
        }
        if (i < nnz) bytes += 28;
    }
    wave_add_bytes(bytes, nbytes);
}

// G[x][y] = sum_{j in R(X->v)(x)} u[j] wx * wv [y in R(v->Y)(j)]: both factors are edge relations
// (wx, wv their edge weights, 1 where the weight array is null).  The work items are (row x, chunk
// of kEeCols columns) pairs, taken in turn by a resident grid.  A workgroup accumulates one item's
// columns in LDS (64 KiB, two workgroups per CU) with 64-bit LDS atomics: its waves take the row's
// entries j in turn and the lanes of a wave split j's list (ascending, so when a row has several
// chunks, a chunk's part of the list is found by two binary searches).  It then writes the chunk
// with 16-byte non-temporal stores, clearing the LDS for the next item in the same pass.  On the
// config-4 batch (tables of ~7.5k x 7k) 8,192 columns per chunk took 464 ms per batch against 790
// with 2,560 (eight workgroups per CU), 766 with 4,096 and 564 with 12,288 or 16,384 (1,024
// threads): most rows fit one chunk, so the searches and the per-item latency go.  (One
// workgroup per whole row with 80 KiB of LDS, a separate clearing pass and 8-byte stores wrote the
// table at about 1 TB/s.)  Two workgroups of 512 threads per CU (the LDS allows two): 344 ms per
// batch against 395 with 256 threads and 368 with 1,024 (whose registers allow one per CU).
constexpr int kEeThreads = 512;
constexpr uint32_t kEeCols = 8192;
__global__ void __launch_bounds__(kEeThreads) k_ee(const uint64_t *__restrict__ xoff, const uint32_t *__restrict__ xval,
                                                 const uint64_t *__restrict__ xw, const uint64_t *__restrict__ voff,
                                                 const uint32_t *__restrict__ vval, const uint64_t *__restrict__ vw,
                                                 const uint64_t *__restrict__ u, uint32_t nx, uint32_t ny,
                                                 uint64_t *__restrict__ G, const unsigned long long *__restrict__ terms,
                                                 int form, unsigned long long *__restrict__ next,
                                                 unsigned long long *__restrict__ nbytes) {
    if (ee_scatter_form(form, terms, (uint64_t)nx * ny)) return;  // the scatter form makes this table
    __shared__ __align__(16) unsigned long long acc[kEeCols];
    __shared__ uint64_t s_it;
    for (uint32_t y = threadIdx.x; y < kEeCols; y += kEeThreads) acc[y] = 0;
    __syncthreads();
    const uint32_t nchunk = (ny + kEeCols - 1) / kEeCols;
    const uint64_t items = (uint64_t)nx * nchunk;
    const uint64_t xbase = xoff[0], vbase = voff[0];  // weights are indexed from each relation's first entry
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint64_t bytes = 0;
    for (;;) {
        // the next item from the launch's queue (zeroed before the launch): workgroups that start
        // late, behind other streams' kernels, take what is left instead of a fixed share
        if (threadIdx.x == 0) s_it = atomicAdd(next, 1ull);
        __syncthreads();
        const uint64_t it = s_it;
        if (it >= items) break;
        const uint32_t x = (uint32_t)(it / nchunk), c = (uint32_t)(it % nchunk);
        const uint32_t y0 = c * kEeCols, n = min(ny - y0, kEeCols);
        const uint64_t lb = xoff[x], le = xoff[x + 1];
        for (uint64_t e = lb + wave; e < le; e += kEeThreads / 64) {
            const uint32_t j = xval[e];
            const uint64_t w = u[j] * (xw ? xw[e - xbase] : 1ull);
            uint64_t b2 = voff[j], e2 = voff[j + 1];
            if (lane == 0) bytes += 28 + 4 * (e2 - b2);
            if (!w) continue;
            if (nchunk > 1) {  // the entries of j's list in [y0, y0 + n)
                const uint32_t len = (uint32_t)(e2 - b2);
                const uint32_t lo = lower_bound_u32(vval + b2, len, y0);
                const uint32_t hi = lo + lower_bound_u32(vval + b2 + lo, len - lo, y0 + n);
                e2 = b2 + hi;
                b2 += lo;
            }
            for (uint64_t f = b2 + lane; f < e2; f += 64) {
                const uint64_t wf = vw ? w * vw[f - vbase] : w;
                atomicAdd(&acc[vval[f] - y0], (unsigned long long)wf);
            }
        }
        __syncthreads();
        // write [y0, y0 + n) of row x: one 8-byte store to reach a 16-byte boundary, pairs, and one
        // 8-byte store for an odd tail; every read LDS word is cleared for the next item
        const uint64_t base = (uint64_t)x * ny + y0;
        const uint32_t head = (uint32_t)(base & 1) < n ? (uint32_t)(base & 1) : 0;
        const uint32_t pairs = (n - head) / 2, tail = head + 2 * pairs;
        if (threadIdx.x == 0 && head) {
            __builtin_nontemporal_store((uint64_t)acc[0], G + base);
            acc[0] = 0;
        }
        if (threadIdx.x == kEeThreads - 1 && tail < n) {
            __builtin_nontemporal_store((uint64_t)acc[tail], G + base + tail);
            acc[tail] = 0;
        }
        u64x2 *dst = (u64x2 *)(G + base + head);
        for (uint32_t q = threadIdx.x; q < pairs; q += kEeThreads) {
            const uint32_t y = head + 2 * q;
            u64x2 v;
            v.x = acc[y];
            v.y = acc[y + 1];
            acc[y] = 0;
            acc[y + 1] = 0;
            __builtin_nontemporal_store(v, dst + q);
        }
        __syncthreads();
        if (threadIdx.x == 0) bytes += 16 + 8ull * n;
    }
    wave_add_bytes(bytes, nbytes);
}

// G[x][y] = sum_{j in R(x)} u[j] w[e] F[j][y]: an edge relation (rows x, columns j; w its edge
// weights, 1 where null) times a dense table whose rows are the relation's columns.
// A work unit is (strip of 64 columns, block of kEdsRows consecutive rows), one wave, lane l
// owning column l of the strip, so every dense-row load of the wave reads 64 consecutive u64.
// The wave reads the block's entries 64 at a time (one per lane: column j, weight) and walks them
// in order, broadcasting each one's j and weight through scalar registers and gathering
// kEdInFlight dense rows before it uses the first; a row's sum is stored when the walk passes the
// row's last entry, and rows without entries get zeros.  The strips are dealt to the eight XCDs
// (workgroup b runs on XCD b % 8; the grid is a multiple of 8 and fits the GPU at once), and each
// XCD walks its strips in turn, so the rows of an F strip, re-read once per relation entry that
// names them, are re-read within one XCD's L2.  The table is written with non-temporal stores,
// which keep it out of L2.  Measured on the config-4 batch (tables of ~7.5k x 7k), one item per
// (strip, row) instead: 64 lanes x 1 column beat 32 x 1 by 23 %, and 16 x 4, 32 x 2, 32 x 4, 64 x 2
// and 16 loads in flight were slower; this streamed form is 2 % faster than that one, with 32 rows
// per unit (8, 16 and 64 slower) and 16 loads in flight (4, 8 and 12 slower).  The push form (each dense row read once, its entries added into
// an LDS row over the relation's lists with 64-bit LDS atomics) was four times slower: ~450M random
// LDS atomics per table.  SIDE only names the instantiation after the step that launches it (0:
// edge (X, v) x dense (v, Y); 1: the dense (X, v) x edge (v, Y) step computed on its transposed
// side), so that profilers report the two apart.
constexpr int kEdInFlight = 16;
constexpr uint32_t kEdsRows = 32;
template <int SIDE>
__global__ void __launch_bounds__(kThreads) k_eds(const uint64_t *__restrict__ xoff, const uint32_t *__restrict__ xval,
                                                  const uint64_t *__restrict__ xw, const uint64_t *__restrict__ u,
                                                  const uint64_t *__restrict__ F, uint32_t nx, uint32_t ny,
                                                  uint64_t *__restrict__ G, unsigned long long *__restrict__ nbytes) {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t nstrips = (ny + 63) / 64, nrb = (nx + kEdsRows - 1) / kEdsRows;
    const uint32_t nxcd = nstrips >= 8 && gridDim.x % 8 == 0 ? 8 : 1;
    const uint32_t xcd = blockIdx.x % nxcd;
    const uint64_t my_strips = xcd < nstrips ? (nstrips - xcd + nxcd - 1) / nxcd : 0;
    const uint32_t units = (uint32_t)(my_strips * nrb);
    const uint64_t xbase = xoff[0];  // weights are indexed from the relation's first entry
    uint64_t bytes = 0;
    // each XCD's units are dealt to its waves in turn (a queue per XCD, one atomic per unit, was 3-7 %
    // slower alone and no faster beside other streams)
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    constexpr uint32_t wpb = kThreads / 64;
    const uint32_t lblk = blockIdx.x / nxcd, nblk = gridDim.x / nxcd;
    for (uint32_t it = lblk * wpb + wave; it < units; it += nblk * wpb) {
        const uint32_t strip = xcd + nxcd * (uint32_t)(it / nrb), rb = (uint32_t)(it % nrb);
        const uint32_t c = strip * 64 + lane;
        const bool ok = c < ny;
        const uint64_t *fc = F + (ok ? c : 0);
        const uint32_t r0 = rb * kEdsRows, nr = min(nx - r0, kEdsRows);
        // lane l < nr holds the end of row r0 + l
        const uint64_t rend_l = lane < nr ? xoff[r0 + lane + 1] : 0;
        const uint64_t e0 = xoff[r0], e1 = xoff[r0 + nr];
        uint64_t *gr = G + (uint64_t)r0 * ny + c;
        uint32_t cur = 0;  // the row being summed, relative to r0
        uint64_t rend = __builtin_amdgcn_readfirstlane((uint32_t)rend_l) |
                        ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(rend_l >> 32)) << 32);
        uint64_t acc = 0;
        for (uint64_t k0 = e0; k0 < e1; k0 += 64) {
            const uint64_t ke = k0 + lane;
            uint32_t jl = 0;
            uint64_t wl = 0;
            if (ke < e1) {
                jl = xval[ke];
                wl = xw ? u[jl] * xw[ke - xbase] : u[jl];
            }
            const uint32_t cnt = (uint32_t)min<uint64_t>(64, e1 - k0);
            for (uint32_t t = 0; t < cnt; t += kEdInFlight) {
                uint64_t jr[kEdInFlight];  // element offsets of the dense rows (tables may pass 2^32 entries)
                uint64_t fv[kEdInFlight];
#pragma unroll
                for (int q = 0; q < kEdInFlight; ++q)
                    jr[q] = (uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)jl, (int)((t + q) & 63)) * ny;
#pragma unroll
                for (int q = 0; q < kEdInFlight; ++q) fv[q] = t + q < cnt && ok ? fc[jr[q]] : 0;
#pragma unroll
                for (int q = 0; q < kEdInFlight; ++q) {
                    if (t + q >= cnt) break;
                    const uint64_t e = k0 + t + q;
                    while (e >= rend) {  // entry e is past the row: store the row, move on
                        if (ok) __builtin_nontemporal_store(acc, gr + (uint64_t)cur * ny);
                        acc = 0;
                        ++cur;
                        rend = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)rend_l, (int)cur) |
                               ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(rend_l >> 32), (int)cur) << 32);
                    }
                    const uint64_t w = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)wl, (int)(t + q)) |
                                       ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(wl >> 32), (int)(t + q)) << 32);
                    acc += w * fv[q];
                }
            }
        }
        for (; cur < nr; ++cur) {  // the last summed row, then rows with no entries left
            if (ok) __builtin_nontemporal_store(acc, gr + (uint64_t)cur * ny);
            acc = 0;
        }
        if (lane == 0) bytes += 8ull * (nr + 1) + (e1 - e0) * (12 + 8ull * 64) + 8ull * 64 * nr;
    }
    wave_add_bytes(bytes, nbytes);
}

// G = A diag(u) B with A (nx x nv) and B (nv x ny) dense: 64 x 64 output tiles, 16 x 16 threads
// with 4 x 4 outputs each, K-chunks of 16 staged in LDS.
constexpr int kT = 64, kK = 16;
__global__ void __launch_bounds__(256) k_dd(const uint64_t *__restrict__ A, const uint64_t *__restrict__ u,
                                            const uint64_t *__restrict__ B, uint32_t nx, uint32_t nv, uint32_t ny,
                                            uint64_t *__restrict__ G, unsigned long long *__restrict__ nbytes) {
    __shared__ uint64_t As[kK][kT];
    __shared__ uint64_t Bs[kK][kT];
    const uint32_t tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const uint64_t tiles_y = (ny + kT - 1) / kT;
    const uint64_t ntiles = ((uint64_t)(nx + kT - 1) / kT) * tiles_y;
    uint64_t bytes = 0;
    for (uint64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const uint32_t x0 = (uint32_t)(t / tiles_y) * kT, y0 = (uint32_t)(t % tiles_y) * kT;
        uint64_t c[4][4] = {};
        for (uint32_t k0 = 0; k0 < nv; k0 += kK) {
            for (int i = threadIdx.x; i < kK * kT; i += 256) {
                const int kk = i / kT, xx = i % kT;
                const uint32_t gx = x0 + xx, gk = k0 + kk;
                As[kk][xx] = (gx < nx && gk < nv) ? A[(uint64_t)gx * nv + gk] * u[gk] : 0;
                const uint32_t gy = y0 + xx;
                Bs[kk][xx] = (gk < nv && gy < ny) ? B[(uint64_t)gk * ny + gy] : 0;
            }
            __syncthreads();
#pragma unroll
            for (int kk = 0; kk < kK; ++kk) {
                uint64_t a[4], b[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) a[i] = As[kk][ty * 4 + i];
#pragma unroll
                for (int i = 0; i < 4; ++i) b[i] = Bs[kk][tx * 4 + i];
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int j = 0; j < 4; ++j) c[i][j] += a[i] * b[j];
            }
            __syncthreads();
        }
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const uint32_t gx = x0 + ty * 4 + i, gy = y0 + tx * 4 + j;
                if (gx < nx && gy < ny) G[(uint64_t)gx * ny + gy] = c[i][j];
            }
        if (threadIdx.x == 0) bytes += 8ull * kT * (2ull * nv + kT);
    }
    wave_add_bytes(bytes, nbytes);
}

// ---- dense x dense on the matrix cores.  G = A diag(u) B mod 2^64 is computed from balanced base-256
// digits: every u64 value v is written v = sum_i d_i 256^i (mod 2^64) with d_i in [-128, 127] (byte i of
// v plus the carry from below, less 256 when that reaches 128), so
//   G = sum_{i + j <= 7} 256^(i + j) (A_i B_j)  (mod 2^64)
// with A_i, B_j int8 matrices of digits; pairs with i + j >= 8 vanish mod 2^64.  Each A_i B_j is an
// exact int8 product on v_mfma_i32_32x32x32_i8: the partial sums of one shift s = i + j over
// kDdFold values of k stay below 8 * 2^14 * 8192 = 2^30 in magnitude, so they are folded into the
// u64 result (sign-extended, shifted by 8 s) every kDdFold values of k.
// Digit layout, for both operands (rows of G's rows for A, of G's columns for B, K = nv padded to
// kDdK): per row and per block of 16 values of k, 128 contiguous bytes, the 16 digits of plane 0, then
// of plane 1, ... plane 7.  A lane's operands for all eight planes of one MFMA step are then one
// 128-byte line.
constexpr uint32_t kDdK = 64, kDdFold = 8192;

// The balanced digits of v, digit i in byte i (two's complement).
__device__ inline uint64_t balanced_digits(uint64_t v) {
    uint64_t out = 0;
    uint32_t carry = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const uint32_t t = (uint32_t)((v >> (8 * i)) & 255) + carry;  // 0..256
        carry = t >= 128;
        out |= (uint64_t)(t & 255) << (8 * i);  // t, or t - 256 as an int8
    }
    return out;
}

// Writes the 128-byte digit block of 16 packed values p[0..15] to dst.
__device__ inline void store_digit_block(const uint64_t (&p)[16], int8_t *__restrict__ dst) {
    u64x2 *o = (u64x2 *)dst;
#pragma unroll
    for (int d = 0; d < 8; ++d) {
        uint64_t lo = 0, hi = 0;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            lo |= ((p[q] >> (8 * d)) & 255) << (8 * q);
            hi |= ((p[q + 8] >> (8 * d)) & 255) << (8 * q);
        }
        u64x2 v;
        v.x = lo;
        v.y = hi;
        o[d] = v;
    }
}

// Digits of M (rows x K, row-major, k contiguous) times u[k] (null: 1), for rows_p x Kp (zeros in
// the padding): one thread per (row, block of 16 k).
__global__ void k_dd_digits_rows(const uint64_t *__restrict__ M, const uint64_t *__restrict__ u, uint32_t rows,
                                 uint32_t K, uint32_t rows_p, uint32_t Kp, int8_t *__restrict__ dst) {
    const uint32_t kbs = Kp / 16;
    const uint64_t n = (uint64_t)rows_p * kbs;
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t r = (uint32_t)(t / kbs), kb = (uint32_t)(t % kbs);
        uint64_t p[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const uint32_t k = kb * 16 + q;
            const uint64_t v = r < rows && k < K ? M[(uint64_t)r * K + k] * (u ? u[k] : 1ull) : 0;
            p[q] = balanced_digits(v);
        }
        store_digit_block(p, dst + t * 128);
    }
}

// The same for M stored k-major (K x rows, row-major: M[k * rows + r]): 64 x 64 tiles transposed in LDS.
__global__ void __launch_bounds__(256) k_dd_digits_cols(const uint64_t *__restrict__ M, const uint64_t *__restrict__ u,
                                                        uint32_t rows, uint32_t K, uint32_t rows_p, uint32_t Kp,
                                                        int8_t *__restrict__ dst) {
    __shared__ uint64_t tile[64][65];
    const uint32_t r0 = blockIdx.x * 64, k0 = blockIdx.y * 64;
    {
        const uint32_t rr = threadIdx.x & 63;
        for (uint32_t kk = threadIdx.x >> 6; kk < 64; kk += 4) {
            const uint32_t r = r0 + rr, k = k0 + kk;
            const uint64_t v = r < rows && k < K ? M[(uint64_t)k * rows + r] * (u ? u[k] : 1ull) : 0;
            tile[kk][rr] = balanced_digits(v);
        }
    }
    __syncthreads();
    const uint32_t rr = threadIdx.x >> 2, part = threadIdx.x & 3;
    uint64_t p[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) p[q] = tile[part * 16 + q][rr];
    const uint32_t kbs = Kp / 16;
    store_digit_block(p, dst + ((uint64_t)(r0 + rr) * kbs + k0 / 16 + part) * 128);
}

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

// G (nx x ny) from the digit blocks of A (rows of G) and B (columns of G): a workgroup of four
// waves per 64 x 64 tile, one 32 x 32 block per wave.  Lane l holds row l % 32 of its A block and
// column l % 32 of its B block, and k block 2 s + l / 32 of each 32-wide step s; the result D[i][j]
// of the 32 x 32 block is in lane j + 32 ((i / 4) % 2), register (i % 4) + 4 (i / 8).
__global__ void __launch_bounds__(256) k_dd_mfma(const int8_t *__restrict__ Ad, const int8_t *__restrict__ Bd,
                                                 uint32_t nx, uint32_t ny, uint32_t Kp, uint64_t *__restrict__ G,
                                                 unsigned long long *__restrict__ nbytes) {
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t m0 = blockIdx.y * 64 + (w >> 1) * 32, n0 = blockIdx.x * 64 + (w & 1) * 32;
    const uint32_t kbs = Kp / 16;
    const int8_t *arow = Ad + ((uint64_t)(m0 + (lane & 31)) * kbs + (lane >> 5)) * 128;
    const int8_t *brow = Bd + ((uint64_t)(n0 + (lane & 31)) * kbs + (lane >> 5)) * 128;
    uint64_t res[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) res[r] = 0;
    for (uint32_t kc = 0; kc < Kp; kc += kDdFold) {
        v16i acc[8];
#pragma unroll
        for (int s = 0; s < 8; ++s)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[s][r] = 0;
        const uint32_t ke = min(Kp, kc + kDdFold);
        // the next step's operands are loaded while this step's products run
        v4i a[8], b[8];
        {
            const v4i *ap = (const v4i *)(arow + (uint64_t)(kc / 16) * 128);
            const v4i *bp = (const v4i *)(brow + (uint64_t)(kc / 16) * 128);
#pragma unroll
            for (int d = 0; d < 8; ++d) a[d] = ap[d];
#pragma unroll
            for (int d = 0; d < 8; ++d) b[d] = bp[d];
        }
        for (uint32_t k = kc; k < ke; k += 32) {
            v4i an[8], bn[8];
            const uint32_t kn = k + 32 < ke ? k + 32 : k;  // the last step reloads its own operands
            const v4i *ap = (const v4i *)(arow + (uint64_t)(kn / 16) * 128);
            const v4i *bp = (const v4i *)(brow + (uint64_t)(kn / 16) * 128);
#pragma unroll
            for (int d = 0; d < 8; ++d) an[d] = ap[d];
#pragma unroll
            for (int d = 0; d < 8; ++d) bn[d] = bp[d];
#pragma unroll
            for (int s = 0; s < 8; ++s)
#pragma unroll
                for (int i = 0; i <= s; ++i) acc[s] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a[i], b[s - i], acc[s], 0, 0, 0);
#pragma unroll
            for (int d = 0; d < 8; ++d) {
                a[d] = an[d];
                b[d] = bn[d];
            }
        }
#pragma unroll
        for (int s = 0; s < 8; ++s)
#pragma unroll
            for (int r = 0; r < 16; ++r) res[r] += (uint64_t)(int64_t)acc[s][r] << (8 * s);
    }
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const uint32_t m = m0 + 8 * (r / 4) + 4 * (lane >> 5) + (r % 4), n = n0 + (lane & 31);
        if (m < nx && n < ny) G[(uint64_t)m * ny + n] = res[r];
    }
    if (threadIdx.x == 0) atomicAdd(nbytes, (unsigned long long)(16ull * 64 * Kp + 8ull * 64 * 64));
}

// G = A diag(u) B (mod 2^64) on the matrix cores.  A (nx x nv) is stored row-major with rows D(X)
// (a_kmajor false) or as its transpose (nv x nx, a_kmajor true); B (nv x ny) row-major (b_kmajor true)
// or as its transpose (ny x nv); u may be null (all ones).  The digit blocks are temporaries.
int dense_product_mfma(asdm_ctx *ctx, const uint64_t *A, bool a_kmajor, const uint64_t *u, const uint64_t *B,
                       bool b_kmajor, uint32_t nx, uint32_t nv, uint32_t ny, uint64_t *G, unsigned long long *nbytes,
                       KTimer &kt) {
    auto up = [](uint64_t x, uint64_t m) { return (std::max<uint64_t>(x, 1) + m - 1) / m * m; };
    const uint64_t Kp = up(nv, kDdK), nxp = up(nx, 64), nyp = up(ny, 64);
    if (nxp / 64 > 65535 || Kp / 64 > 65535) return ctx->fail(ASDM_E_UNSUPPORTED, "dense product beyond the tile grid");
    int8_t *Ad = nullptr, *Bd = nullptr;
    int rc = dev_alloc(ctx, (void **)&Ad, nxp * Kp * 8);
    if (!rc) rc = dev_alloc(ctx, (void **)&Bd, nyp * Kp * 8);
    if (!rc) {
        auto digits = [&](const uint64_t *M, bool kmajor, const uint64_t *scale, uint32_t rows, uint64_t rows_p, int8_t *dst) {
            if (kmajor)
                kt.launch(k_dd_digits_cols, dim3((unsigned)(rows_p / 64), (unsigned)(Kp / 64)), dim3(256), 0, M, scale, rows,
                          nv, (uint32_t)rows_p, (uint32_t)Kp, dst);
            else
                kt.launch(k_dd_digits_rows, dim3(grid_for(rows_p * (Kp / 16), 256)), dim3(256), 0, M, scale, rows, nv,
                          (uint32_t)rows_p, (uint32_t)Kp, dst);
        };
        digits(A, a_kmajor, u, nx, nxp, Ad);
        digits(B, b_kmajor, nullptr, ny, nyp, Bd);
        kt.launch(k_dd_mfma, dim3((unsigned)(nyp / 64), (unsigned)(nxp / 64)), dim3(256), 0, (const int8_t *)Ad,
                  (const int8_t *)Bd, nx, ny, (uint32_t)Kp, G, nbytes);
        rc = check_hip(ctx, hipGetLastError(), "dense product");
    }
    if (Ad) dev_free(ctx, Ad);
    if (Bd) dev_free(ctx, Bd);
    return rc;
}

// ---- sparse joins: the product of v's two factors evaluated only on the edges of the relation
// R(X->Y) that already joins X and Y, times each edge's own weight (1 where the relation has
// none), written to wout.  The work items are the relation's edges, kJoinLanes lanes per edge
// (a power-law hub's row holds many edges, so rows would leave most lanes idle behind a few);
// an edge finds its row by binary search in the offsets (row_of).
// Weight arrays are indexed by entry - off[0]: a relation over single query vertices is a slice
// of the candidate-edge store, whose offsets do not start at zero.
constexpr int kJoinLanes = 8;

// Both factors edge relations: the product at (x, y) is the sum of u[j] wa wb over the common
// entries j of the rows R(X->v)(x) and R(Y->v)(y) (ascending D(v) positions).  The lanes split
// the shorter row and binary-search the longer.
__global__ void k_join_ee(const uint64_t *__restrict__ roff, const uint32_t *__restrict__ rval,
                          const uint64_t *__restrict__ rw, uint32_t nx, uint64_t nnz, const uint64_t *__restrict__ aoff,
                          const uint32_t *__restrict__ aval, const uint64_t *__restrict__ aw,
                          const uint64_t *__restrict__ boff, const uint32_t *__restrict__ bval,
                          const uint64_t *__restrict__ bw, const uint64_t *__restrict__ u, uint64_t *__restrict__ wout,
                          unsigned long long *__restrict__ nbytes) {
    constexpr int G = kJoinLanes;
    const int lg = threadIdx.x & (G - 1);
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x / G;
    const uint64_t rbase = roff[0], abase = aoff[0], bbase = boff[0];
    nnz = min(nnz, roff[nx] - rbase);  // nnz: the host's bound on the entries
    uint64_t bytes = 0;
    for (uint64_t k = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / G; k < nnz; k += stride) {
        const uint64_t e = rbase + k;
        const uint32_t x = row_of(roff, nx, e), y = rval[e];
        const uint64_t a0 = aoff[x], na = aoff[x + 1] - a0;
        const uint64_t b0 = boff[y], nb = boff[y + 1] - b0;
        const bool a_short = na <= nb;
        const uint64_t s0 = a_short ? a0 : b0, ns = a_short ? na : nb;
        const uint64_t l0 = a_short ? b0 : a0, nl = a_short ? nb : na;
        const uint32_t *sv = a_short ? aval : bval, *lv = a_short ? bval : aval;
        const uint64_t *sw = a_short ? aw : bw, *lw = a_short ? bw : aw;
        const uint64_t sbase = a_short ? abase : bbase, lbase = a_short ? bbase : abase;
        uint64_t acc = 0;
        for (uint64_t t = lg; t < ns; t += G) {
            const uint32_t j = sv[s0 + t];
            const uint32_t p = lower_bound_u32(lv + l0, (uint32_t)nl, j);
            if (p < nl && lv[l0 + p] == j) {
                uint64_t m = u[j];
                if (sw) m *= sw[s0 + t - sbase];
                if (lw) m *= lw[l0 + p - lbase];
                acc += m;
            }
        }
        acc = group_sum<G>(acc);
        if (lg == 0) {
            wout[k] = rw ? acc * rw[k] : acc;
            bytes += 32 + 12 * (ns + nl);
        }
    }
    wave_add_bytes(bytes, nbytes);
}

// One factor an edge relation L (rows: the list owner o, x when own_x else y) and the other a
// dense table F over (the other end a, v) read as F[a * sa + j * sj]: the product at (x, y) is
// the sum over the entries j of L(o) of u[j] wl F(a, j).
__global__ void k_join_ed(const uint64_t *__restrict__ roff, const uint32_t *__restrict__ rval,
                          const uint64_t *__restrict__ rw, uint32_t nx, uint64_t nnz, const uint64_t *__restrict__ loff,
                          const uint32_t *__restrict__ lval, const uint64_t *__restrict__ lw, int own_x,
                          const uint64_t *__restrict__ F, uint64_t sa, uint64_t sj, const uint64_t *__restrict__ u,
                          uint64_t *__restrict__ wout, unsigned long long *__restrict__ nbytes,
                          unsigned long long *__restrict__ ngather) {
    constexpr int G = kJoinLanes;
    const int lg = threadIdx.x & (G - 1);
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x / G;
    const uint64_t rbase = roff[0], lbase = loff[0];
    nnz = min(nnz, roff[nx] - rbase);  // nnz: the host's bound on the entries
    uint64_t bytes = 0, gathered = 0;
    for (uint64_t k = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / G; k < nnz; k += stride) {
        const uint64_t e = rbase + k;
        const uint32_t x = row_of(roff, nx, e), y = rval[e];
        const uint64_t o = own_x ? x : y, a = own_x ? y : x;
        const uint64_t l0 = loff[o], l1 = loff[o + 1];
        const uint64_t *fa = F + a * sa;
        uint64_t acc = 0;
        // four entries per lane per pass (a hub's list is long and every one of its edges walks it)
        constexpr int Q = 4;
        for (uint64_t t0 = l0 + lg; t0 < l1; t0 += Q * G) {
            uint32_t j[Q];
            bool in[Q];
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                const uint64_t t = t0 + q * G;
                in[q] = t < l1;
                j[q] = lval[in[q] ? t : l0];
            }
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                uint64_t m = u[j[q]] * fa[(uint64_t)j[q] * sj];
                if (lw) m *= lw[(in[q] ? t0 + q * G : l0) - lbase];
                acc += in[q] ? m : 0;
            }
        }
        acc = group_sum<G>(acc);
        if (lg == 0) {
            wout[k] = rw ? acc * rw[k] : acc;
            bytes += 32 + 28 * (l1 - l0);
            gathered += l1 - l0;
        }
    }
    wave_add_bytes(bytes, nbytes);
    wave_add_bytes(gathered, ngather);
}

// Both factors dense, A(x, j) = A[x * sax + j * saj] and B(j, y) = B[j * sbj + y * sby]: the
// product at (x, y) is sum_j A(x, j) u[j] B(j, y), one wave per edge, lanes over j.
__global__ void k_join_dd(const uint64_t *__restrict__ roff, const uint32_t *__restrict__ rval,
                          const uint64_t *__restrict__ rw, uint32_t nx, uint64_t nnz, const uint64_t *__restrict__ A,
                          uint64_t sax, uint64_t saj, const uint64_t *__restrict__ B, uint64_t sbj, uint64_t sby,
                          uint32_t nv, const uint64_t *__restrict__ u, uint64_t *__restrict__ wout,
                          unsigned long long *__restrict__ nbytes) {
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t nw = (uint64_t)gridDim.x * (blockDim.x / 64);
    const uint64_t rbase = roff[0];
    nnz = min(nnz, roff[nx] - rbase);  // nnz: the host's bound on the entries
    uint64_t bytes = 0;
    for (uint64_t k = (uint64_t)blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64; k < nnz; k += nw) {
        const uint64_t e = rbase + k;
        const uint64_t x = row_of(roff, nx, e), y = rval[e];
        uint64_t acc = 0;
        for (uint32_t j = lane; j < nv; j += 64) acc += A[x * sax + j * saj] * u[j] * B[j * sbj + y * sby];
        acc = group_sum<64>(acc);
        if (lane == 0) {
            wout[k] = rw ? acc * rw[k] : acc;
            bytes += 32 + 24ull * nv;
        }
    }
    wave_add_bytes(bytes, nbytes);
}

// Weights of the reverse relation R(Y->X) from those of R(X->Y): entry (y, x) takes the weight of
// entry (x, y), found by binary search in row x.  One lane per entry.
__global__ void k_join_reverse(const uint64_t *__restrict__ foff, const uint32_t *__restrict__ fval,
                               const uint64_t *__restrict__ wf, const uint64_t *__restrict__ roff,
                               const uint32_t *__restrict__ rval, uint32_t ny, uint64_t nnz, uint64_t *__restrict__ wr) {
    const uint64_t fbase = foff[0], rbase = roff[0];
    nnz = min(nnz, roff[ny] - rbase);  // nnz: the host's bound on the entries
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < nnz; k += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t e = rbase + k;
        const uint32_t y = row_of(roff, ny, e), x = rval[e];
        const uint64_t f0 = foff[x];
        const uint32_t n = (uint32_t)(foff[x + 1] - f0);
        const uint32_t p = lower_bound_u32(fval + f0, n, y);
        wr[k] = p < n && fval[f0 + p] == y ? wf[f0 + p - fbase] : 0;  // the entry is always there
    }
}

// out (cols x rows) = transpose of in (rows x cols), 32 x 32 tiles through LDS.
__global__ void k_transpose(const uint64_t *__restrict__ in, uint32_t rows, uint32_t cols, uint64_t *__restrict__ out) {
    __shared__ uint64_t tile[32][33];
    const uint64_t tiles_c = (cols + 31) / 32;
    const uint64_t ntiles = ((uint64_t)(rows + 31) / 32) * tiles_c;
    const uint32_t tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 256 threads: 8 rows per pass
    for (uint64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const uint32_t r0 = (uint32_t)(t / tiles_c) * 32, c0 = (uint32_t)(t % tiles_c) * 32;
        for (uint32_t k = ty; k < 32; k += 8)
            if (r0 + k < rows && c0 + tx < cols) tile[k][tx] = in[(uint64_t)(r0 + k) * cols + c0 + tx];
        __syncthreads();
        for (uint32_t k = ty; k < 32; k += 8)
            if (c0 + k < cols && r0 + tx < rows) out[(uint64_t)(c0 + k) * rows + r0 + tx] = tile[tx][k];
        __syncthreads();
    }
}

// out = a * b elementwise
__global__ void k_hadamard(uint64_t *__restrict__ out, const uint64_t *__restrict__ a, const uint64_t *__restrict__ b,
                           uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        out[i] = a[i] * b[i];
}

// mat[ids[i]][u] += mu * scal * marg[i] for every query vertex u of the block.  Domain ids are
// distinct, so every cell has one writer per launch and no atomics are needed.
__global__ void k_scatter_local(const uint32_t *__restrict__ ids, uint32_t n, const uint64_t *__restrict__ marg,
                                const uint64_t *__restrict__ scal, int64_t mu, Members m, uint32_t nq,
                                uint64_t *__restrict__ mat) {
    const uint64_t c = (uint64_t)mu * scal[0];
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint64_t add = c * marg[i];
        for (uint32_t k = 0; k < m.n; ++k) mat[(uint64_t)ids[i] * nq + m.u[k]] += add;
    }
}

// ------------------------------------------------------------------------------ executor
struct Dom {
    const uint32_t *ids = nullptr;
    const uint64_t *bits = nullptr;
    const uint32_t *rank = nullptr;
    uint32_t n = 0;
};
struct Rel {
    const uint64_t *off = nullptr;
    const uint32_t *val = nullptr;
    uint64_t nnz = 0;  // entries, or an upper bound for a relation built on the device (exact: off[rows] - off[0])
    const uint64_t *w = nullptr;  // per-entry weights, w[e - off[0]] for entry e (null: every weight is 1)
};
struct Table {  // dense factor of a block pair: row-major, rows indexed by D(rows), columns by the other
    uint64_t *p = nullptr;
    uint32_t rows = 0;
};
// Live factor of a block pair (lo, hi), lo < hi: an edge relation in both orientations (rows D(lo)
// and rows D(hi)), possibly weighted, or a dense table.
// S (sorted member masks) names the blocks eliminated into the factor.  A factor over (X, Y) is
// the sum over the blocks of S of the product of every quotient edge inside S + {X, Y}, and the
// quotient's edges are those of the query between members: the factor is a function of the member
// masks of S, X and Y alone, which is what lets partition terms and marginals share it.
struct Fac {
    int kind = 0;  // 1 edge relation, 2 dense table
    Table t;
    Rel lo, hi;
    uint64_t *w[2] = {nullptr, nullptr};  // weight arrays of lo and hi
    bool owned = true;                    // t.p / w are this factor's (not the cache's)
    std::vector<uint32_t> S;
};
// A factor result kept for the rest of the query: a vector over one block, the weights of an
// edge relation (both orientations) or a dense table (rows: the member mask of its row block).
struct CacheEnt {
    uint64_t *p = nullptr, *p2 = nullptr;
    uint32_t rows = 0;
};

struct BagRun {
    asdm_ctx *ctx;
    const asdm_cs *cs;
    const BagPlan *plan;
    unsigned long long *d_acc = nullptr;    // ctx counters: [0] total, [4] scratch sum, [5] term scalar,
                                            // [6] relation and [7] vector algorithmic bytes, [14..17]
                                            // bytes the table kernels' loops touch (ee, de, ed, dd),
                                            // [18..19] those of the sparse joins (ee; ed and dd),
                                            // [20..21] a relation product's term count and item
                                            // queue, [30] the table entries the ed joins gathered
    uint64_t *d_mat = nullptr;
    std::vector<void *> owned;              // per-term device buffers
    // Domains of merged blocks and relations between blocks depend only on the blocks' members,
    // and the same blocks recur across the partition terms: both are built once per query.
    std::vector<void *> kept;               // device buffers of the caches, freed at the end
    std::map<std::vector<uint32_t>, Dom> dom_cache;
    std::map<std::pair<std::vector<uint32_t>, std::vector<uint32_t>>, Rel> rel_cache;
    bool keep_allocs = false;               // allocations go to `kept` while building cached items
    // Factor results by signature (kind, boundary masks, eliminated masks), kept for the query;
    // cache_cap bounds their bytes (beyond it results are computed and dropped as before).
    std::map<std::vector<uint32_t>, CacheEnt> fcache;
    std::vector<std::vector<RelJob>> host_jobs;     // host sources of queued job uploads
    std::vector<std::vector<uint64_t>> host_row0;
    uint64_t cache_bytes = 0, cache_cap = 0, reuse_n = 0, reuse_bytes = 0;
    std::chrono::steady_clock::time_point t0;

    unsigned long long *bytes_slot(int k) { return d_acc + 6 + k; }

    int alloc(void **p, size_t bytes) {
        int rc = dev_alloc(ctx, p, std::max<size_t>(bytes, 8));
        if (!rc) (keep_allocs ? kept : owned).push_back(*p);
        return rc;
    }
    void release(void *p) {
        for (auto *v : {&owned, &kept}) {
            auto it = std::find(v->begin(), v->end(), p);
            if (it != v->end()) v->erase(it);
        }
        dev_free(ctx, p);
    }
    void release_all() {
        for (void *p : owned) dev_free(ctx, p);
        owned.clear();
    }
    void release_cache() {
        release_all();
        for (void *p : kept) dev_free(ctx, p);
        kept.clear();
        dom_cache.clear();
        rel_cache.clear();
        fcache.clear();
        cache_bytes = 0;
    }
    // Moves a per-term buffer into the query-lifetime set.
    void adopt(void *p) {
        auto it = std::find(owned.begin(), owned.end(), p);
        if (it != owned.end()) owned.erase(it);
        kept.push_back(p);
    }
    static std::vector<uint32_t> sig(uint32_t kind, uint32_t ma, uint32_t mb, std::vector<uint32_t> S) {
        std::sort(S.begin(), S.end());
        std::vector<uint32_t> k{kind, std::min(ma, mb), std::max(ma, mb)};
        k.insert(k.end(), S.begin(), S.end());
        return k;
    }
    const CacheEnt *lookup(const std::vector<uint32_t> &k, uint64_t bytes) {
        auto it = fcache.find(k);
        if (it == fcache.end()) return nullptr;
        ++reuse_n;
        reuse_bytes += bytes;
        return &it->second;
    }
    // Keeps a freshly computed result under k when the budget allows; true if kept.
    bool store(const std::vector<uint32_t> &k, const CacheEnt &e, uint64_t bytes) {
        if (cache_bytes + bytes > cache_cap) return false;
        cache_bytes += bytes;
        adopt(e.p);
        if (e.p2) adopt(e.p2);
        fcache[k] = e;
        return true;
    }
    int launched(const char *what) { return check_hip(ctx, hipGetLastError(), what); }
    bool timed_out() const {
        return ctx->time_limit_ms > 0 &&
               std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() > ctx->time_limit_ms;
    }

    int cached_dom(const std::vector<uint32_t> &members, Dom &d) {
        if (members.size() == 1) return build_dom(members, d);
        std::vector<uint32_t> key(members);
        std::sort(key.begin(), key.end());
        auto it = dom_cache.find(key);
        if (it != dom_cache.end()) {
            d = it->second;
            return ASDM_OK;
        }
        keep_allocs = true;
        const int rc = build_dom(members, d);
        keep_allocs = false;
        if (!rc) dom_cache[key] = d;
        return rc;
    }

    // One relation the executor needs: rows D(A) through query arc `arc` (u_a in A, u_b in B).
    struct RelNeed {
        uint32_t arc;
        std::vector<uint32_t> ma, mb;  // members of A and B
        Dom A, B;
    };
    static std::pair<std::vector<uint32_t>, std::vector<uint32_t>> rel_key(const std::vector<uint32_t> &ma,
                                                                        const std::vector<uint32_t> &mb) {
        std::vector<uint32_t> ka(ma), kb(mb);
        std::sort(ka.begin(), ka.end());
        std::sort(kb.begin(), kb.end());
        return {ka, kb};
    }

    int cached_rel(uint32_t arc, const std::vector<uint32_t> &ma, const Dom &A, const std::vector<uint32_t> &mb,
                   const Dom &B, Rel &r) {
        if (ma.size() == 1 && mb.size() == 1) {  // the candidate-edge lists are the relation
            r.off = cs->d_ce_off + cs->ce_base[arc];
            r.val = cs->d_ce_val;
            r.nnz = cs->arc_weight[arc];
            r.w = nullptr;
            return ASDM_OK;
        }
        const auto key = rel_key(ma, mb);
        auto it = rel_cache.find(key);
        if (it == rel_cache.end()) {
            int rc = build_rels({RelNeed{arc, ma, mb, A, B}});
            if (rc) return rc;
            it = rel_cache.find(key);
        }
        r = it->second;
        return ASDM_OK;
    }

    // Builds the relations of `needs` (none of them over two single vertices) in one count launch,
    // one scan and one fill launch, without a host synchronization; they go to the cache.
    int build_rels(const std::vector<RelNeed> &needs) {
        const uint32_t nj = (uint32_t)needs.size();
        if (!nj) return ASDM_OK;
        std::vector<RelJob> jobs(nj);
        std::vector<uint64_t> row0(nj + 1, 0);
        uint64_t ub = 0;  // each relation is a subset of its arc's candidate edges
        for (uint32_t j = 0; j < nj; ++j) {
            const RelNeed &n = needs[j];
            const uint32_t ua = cs->arc_u[n.arc], ub_ = cs->arc_u2[n.arc];
            jobs[j] = RelJob{n.A.ids, cs->d_bits + (size_t)ua * cs->W, cs->d_rank + (size_t)ua * cs->W,
                             cs->d_ce_off + cs->ce_base[n.arc], cs->d_cand + cs->cand_off[ub_], n.B.bits, n.B.rank};
            row0[j + 1] = row0[j] + n.A.n;
            ub += cs->arc_weight[n.arc];
        }
        const uint64_t rows = row0[nj];
        keep_allocs = true;  // offsets and values live as long as the cache
        uint64_t *off = nullptr;
        uint32_t *val = nullptr;
        int rc = alloc((void **)&off, 8ull * (rows + 1));
        if (!rc) rc = alloc((void **)&val, 4ull * std::max<uint64_t>(ub, 1));
        keep_allocs = false;
        RelJob *d_jobs = nullptr;
        uint64_t *d_row0 = nullptr, *src = nullptr, *ebase = nullptr, *kpos = nullptr;
        uint32_t *len = nullptr, *keep = nullptr, *pos = nullptr;
        const uint64_t R1 = std::max<uint64_t>(rows, 1), U1 = std::max<uint64_t>(ub, 1);
        if (!rc) rc = alloc((void **)&d_jobs, sizeof(RelJob) * nj);
        if (!rc) rc = alloc((void **)&d_row0, 8ull * (nj + 1));
        if (!rc) rc = alloc((void **)&src, 8ull * R1);
        if (!rc) rc = alloc((void **)&len, 4ull * R1);
        if (!rc) rc = alloc((void **)&ebase, 8ull * (R1 + 1));
        if (!rc) rc = alloc((void **)&keep, 4ull * U1);
        if (!rc) rc = alloc((void **)&pos, 4ull * U1);
        if (!rc) rc = alloc((void **)&kpos, 8ull * (U1 + 1));
        if (rc) return rc;
        // the host copies stay alive until the queued copies have run (the run's end)
        host_jobs.push_back(std::move(jobs));
        host_row0.push_back(std::move(row0));
        rc = check_hip(ctx, hipMemcpyAsync(d_jobs, host_jobs.back().data(), sizeof(RelJob) * nj, hipMemcpyHostToDevice, ctx->stream),
                       "H2D relation jobs");
        if (!rc)
            rc = check_hip(ctx, hipMemcpyAsync(d_row0, host_row0.back().data(), 8ull * (nj + 1), hipMemcpyHostToDevice, ctx->stream),
                           "H2D relation rows");
        // the entries of all rows are numbered up to ub (every relation is a subset of its arc's
        // candidate edges), so no size has to come back to the host: flags past the last entry
        // stay zero
        if (!rc) rc = fill_u32(ctx, keep, U1, 0);
        if (rc) return rc;
        {
            KTimer kt(ctx, "bag_relation", KTimer::kDispatch);
            kt.launch(k_rel_len, dim3(grid_for(R1, kThreads)), dim3(kThreads), 0, d_jobs, d_row0, nj, src, len);
            rc = launched("bag relation lengths");
            if (!rc) rc = scan_u32_to_u64(ctx, len, ebase, rows);
            if (!rc) {
                kt.launch(k_rel_keep, dim3(grid_for((U1 + kRelPer - 1) / kRelPer, kThreads)), dim3(kThreads), 0, d_jobs,
                          d_row0, nj, src, ebase, cs->d_ce_val, keep, pos, bytes_slot(0));
                rc = launched("bag relation keep");
            }
            if (!rc) rc = scan_u32_to_u64(ctx, keep, kpos, ub);
            if (!rc) {
                kt.launch(k_rel_place, dim3(grid_for(U1, kThreads)), dim3(kThreads), 0, keep, pos, kpos, ub, val);
                kt.launch(k_rel_off, dim3(grid_for(R1 + 1, kThreads)), dim3(kThreads), 0, ebase, kpos, rows, off);
                rc = launched("bag relation place");
            }
            kt.stop();
        }
        if (rc) return rc;
        for (void *p : {(void *)d_jobs, (void *)d_row0, (void *)src, (void *)len, (void *)ebase, (void *)keep, (void *)pos,
                        (void *)kpos})
            release(p);
        const std::vector<uint64_t> &r0 = host_row0.back();
        for (uint32_t j = 0; j < nj; ++j) {
            Rel r;
            r.off = off + r0[j];
            r.val = val;
            r.nnz = cs->arc_weight[needs[j].arc];
            rel_cache[rel_key(needs[j].ma, needs[j].mb)] = r;
        }
        return ASDM_OK;
    }

    // The arcs of quotient edge e of term T, oriented: ab from a member of T.edges[e].first.
    std::pair<uint32_t, uint32_t> edge_arcs(const BagTerm &T, size_t e) const {
        uint32_t ab = T.arc_ab[e], ba = T.arc_ba[e];
        bool a_from_A = false;
        for (uint32_t u : T.blocks[T.edges[e].first]) a_from_A |= cs->arc_u[ab] == u;
        if (!a_from_A) std::swap(ab, ba);
        return {ab, ba};
    }

    // Builds, before the first term, every merged-block relation that some term will use, in one
    // batched build (the terms of a query share most of them).
    int prebuild() {
        std::vector<RelNeed> needs;
        std::set<std::pair<std::vector<uint32_t>, std::vector<uint32_t>>> seen;
        for (const BagTerm &T : plan->terms) {
            const uint32_t nb = (uint32_t)T.blocks.size();
            std::vector<Dom> dom(nb);
            int rc = ASDM_OK;
            for (uint32_t b = 0; b < nb && !rc; ++b) rc = cached_dom(T.blocks[b], dom[b]);
            if (rc) return rc;
            bool empty = false;
            for (uint32_t b = 0; b < nb; ++b) empty |= dom[b].n == 0;
            if (empty) continue;  // the term is 0 and builds nothing
            for (size_t e = 0; e < T.edges.size(); ++e) {
                const uint32_t A = T.edges[e].first, B = T.edges[e].second;
                if (T.blocks[A].size() == 1 && T.blocks[B].size() == 1) continue;
                const auto arcs = edge_arcs(T, e);
                for (int side = 0; side < 2; ++side) {
                    const uint32_t P = side ? B : A, Q = side ? A : B;
                    const auto key = rel_key(T.blocks[P], T.blocks[Q]);
                    if (rel_cache.count(key) || !seen.insert(key).second) continue;
                    needs.push_back(RelNeed{side ? arcs.second : arcs.first, T.blocks[P], T.blocks[Q], dom[P], dom[Q]});
                }
            }
        }
        return build_rels(needs);
    }

    int build_dom(const std::vector<uint32_t> &members, Dom &d) {
        if (members.size() == 1) {
            const uint32_t u = members[0];
            d.ids = cs->d_cand + cs->cand_off[u];
            d.bits = cs->d_bits + (size_t)u * cs->W;
            d.rank = cs->d_rank + (size_t)u * cs->W;
            d.n = cs->csize[u];
            return ASDM_OK;
        }
        const uint32_t W = cs->W;
        Members m{};
        m.n = (uint32_t)members.size();
        for (uint32_t k = 0; k < m.n; ++k) m.u[k] = members[k];
        uint64_t *bits = nullptr, *off = nullptr;
        uint32_t *cnt = nullptr, *rank = nullptr, *ids = nullptr;
        int rc = alloc((void **)&bits, 8ull * W);
        if (!rc) rc = alloc((void **)&cnt, 4ull * W);
        if (!rc) rc = alloc((void **)&off, 8ull * (W + 1));
        if (!rc) rc = alloc((void **)&rank, 4ull * W);
        if (rc) return rc;
        // two timed spans, so that the host synchronization between them is not counted as
        // kernel time
        KTimer kt(ctx, "bag_domain", KTimer::kDispatch);
        kt.launch(k_dom_words, dim3(grid_for(W, kThreads)), dim3(kThreads), 0, cs->d_bits, W, m, bits, cnt);
        if ((rc = launched("bag domain words"))) return rc;
        if ((rc = scan_u32_to_u64(ctx, cnt, off, W))) return rc;
        kt.stop(8ull * W * m.n + 12ull * W + 8ull * (W + 1));
        uint64_t n = 0;
        rc = check_hip(ctx, hipMemcpyAsync(&n, off + W, 8, hipMemcpyDeviceToHost, ctx->stream), "D2H");
        if (!rc) rc = check_hip(ctx, hipStreamSynchronize(ctx->stream), "sync");
        if (!rc) rc = alloc((void **)&ids, 4ull * std::max<uint64_t>(n, 1));
        if (rc) return rc;
        KTimer kf(ctx, "bag_domain", KTimer::kDispatch);
        kf.launch(k_dom_fill, dim3(grid_for(W, kThreads)), dim3(kThreads), 0, bits, W, off, rank, ids);
        kf.stop(8ull * W + 8ull * W + 4ull * W + 4 * n);
        if ((rc = launched("bag domain fill"))) return rc;
        release(cnt);
        release(off);
        d.ids = ids;
        d.bits = bits;
        d.rank = rank;
        d.n = (uint32_t)n;
        return ASDM_OK;
    }

    // A dense table of the block pair {t.rows, other} with rows D(want): the table itself or its
    // transpose (temp = true: the caller releases it).
    int oriented(const Table &t, uint32_t other, uint32_t want, const std::vector<Dom> &dom, uint64_t **out, bool &temp) {
        temp = false;
        if (t.rows == want) {
            *out = t.p;
            return ASDM_OK;
        }
        const uint32_t r = dom[t.rows].n, c = dom[other].n;
        int rc = alloc((void **)out, 8ull * std::max<uint64_t>((uint64_t)r * c, 1));
        if (rc) return rc;
        temp = true;
        const uint64_t tiles = ((uint64_t)(r + 31) / 32) * ((c + 31) / 32);
        KTimer kt(ctx, "bag_transpose", KTimer::kDispatch);
        kt.launch(k_transpose, dim3(grid_for(tiles, 1)), dim3(256), 0, t.p, r, c, *out);
        kt.stop(16ull * r * c);
        return launched("bag transpose");
    }

    // Runs one elimination sequence.  On return *scal (device) holds the product of the sums of
    // the blocks eliminated with no live neighbour, and u[keep] the marginal of `keep`; the
    // vectors u[b] share one allocation, u[0], which the caller releases.
    int eliminate(const BagTerm &T, const std::vector<BagStep> &steps, const std::vector<Dom> &dom,
                  const std::map<std::pair<uint32_t, uint32_t>, std::pair<Rel, Rel>> &rels, std::vector<uint64_t *> &u,
                  uint64_t *scal) {
        const uint32_t nb = (uint32_t)T.blocks.size();
        std::vector<uint32_t> mask(nb, 0);
        for (uint32_t b = 0; b < nb; ++b)
            for (uint32_t q : T.blocks[b]) mask[b] |= 1u << q;
        std::vector<std::vector<uint32_t>> usig(nb);  // blocks eliminated into u[b]
        // the vectors of all blocks in one buffer (u[0] is its start), set to 1 by one fill
        u.assign(nb, nullptr);
        uint64_t total = 0;
        for (uint32_t b = 0; b < nb; ++b) total += std::max<uint32_t>(dom[b].n, 1);
        int rc = alloc((void **)&u[0], 8 * total);
        if (rc) return rc;
        for (uint32_t b = 1; b < nb; ++b) u[b] = u[b - 1] + std::max<uint32_t>(dom[b - 1].n, 1);
        if (!rc) rc = fill_u64(ctx, u[0], total, 1);
        if (!rc) rc = fill_u64(ctx, scal, 1, 1);
        if (rc) return rc;
        std::map<std::pair<uint32_t, uint32_t>, Fac> fac;
        for (auto &e : T.edges) {
            Fac f;
            f.kind = 1;
            f.owned = false;  // the relation caches own them
            f.lo = rels.at(e).first;
            f.hi = rels.at(e).second;
            fac[e] = f;
        }
        auto key = [](uint32_t a, uint32_t b) { return std::make_pair(std::min(a, b), std::max(a, b)); };
        auto rel = [&](uint32_t a, uint32_t b) -> Rel {  // rows D(a)
            const Fac &f = fac.at(key(a, b));
            return a < b ? f.lo : f.hi;
        };
        auto drop = [&](Fac &f) {
            if (!f.owned) return;
            if (f.kind == 2) release(f.t.p);
            for (uint64_t *w : f.w)
                if (w) release(w);
        };
        auto join_sets = [](std::vector<uint32_t> a, const std::vector<uint32_t> &b) {
            a.insert(a.end(), b.begin(), b.end());
            std::sort(a.begin(), a.end());
            return a;
        };
        for (const BagStep &s : steps) {
            if (timed_out()) return ctx->fail(ASDM_E_TIMEOUT, "per-query time limit exceeded");
            const uint32_t v = s.v;
            const Dom &dv = dom[v];
            if (s.nn == 0) {
                hipLaunchKernelGGL(k_sum, dim3(grid_for(dv.n, kThreads)), dim3(kThreads), 0, ctx->stream, u[v], dv.n, d_acc + 4);
                hipLaunchKernelGGL(k_scal_fold, dim3(1), dim3(64), 0, ctx->stream, scal, (uint64_t *)(d_acc + 4));
                if ((rc = launched("bag sum"))) return rc;
            } else if (s.nn == 1) {
                const uint32_t a = s.nb[0];
                const Fac &f = fac.at(key(a, v));
                const uint32_t na = dom[a].n;
                std::vector<uint32_t> St = join_sets(join_sets(f.S, usig[v]), {mask[v]});
                const auto k = sig(3, mask[a], 0, St);
                const CacheEnt *ce = lookup(k, 8ull * na);
                uint64_t *t = ce ? ce->p : nullptr;
                if (!ce) {
                    if ((rc = alloc((void **)&t, 8ull * std::max<uint32_t>(na, 1)))) return rc;
                    KTimer kt(ctx, "bag_vector", KTimer::kDispatch);
                    if (f.kind == 1) {
                        const Rel r = rel(a, v);
                        if ((rc = fill_u64(ctx, t, na, 0))) return rc;
                        const uint64_t lanes = (r.nnz + kSpmvPer - 1) / kSpmvPer;
                        kt.launch(k_spmv, dim3(std::max(1u, grid_for(lanes, kThreads))), dim3(kThreads), 0, r.off,
                                  r.val, r.w, na, u[v], t, bytes_slot(1));
                    } else if (f.t.rows == a) {
                        kt.launch(k_dense_mv, dim3(grid_for(na, kThreads / 64)), dim3(kThreads), 0, f.t.p, na, dv.n,
                                  u[v], t, bytes_slot(1));
                    } else {  // table rows D(v): the transposed product, accumulated with atomics
                        if ((rc = fill_u64(ctx, t, na, 0))) return rc;
                        const dim3 grid((na + kThreads - 1) / kThreads, std::max<uint32_t>(1, (dv.n + kMvtRows - 1) / kMvtRows));
                        kt.launch(k_dense_mv_t, grid, dim3(kThreads), 0, f.t.p, dv.n, na, u[v], t, bytes_slot(1));
                    }
                    kt.stop();
                    if ((rc = launched("bag vector step"))) return rc;
                }
                hipLaunchKernelGGL(k_vec_mul, dim3(grid_for(na, kThreads)), dim3(kThreads), 0, ctx->stream, u[a], t, na);
                if ((rc = launched("bag vector step"))) return rc;
                usig[a] = join_sets(usig[a], St);
                if (!ce && !store(k, CacheEnt{t, nullptr, 0}, 8ull * na)) release(t);
            } else {
                const uint32_t X = std::min(s.nb[0], s.nb[1]), Y = std::max(s.nb[0], s.nb[1]);
                Fac &fxv = fac.at(key(X, v)), &fvy = fac.at(key(v, Y));
                const std::vector<uint32_t> Snew = join_sets(join_sets(join_sets(fxv.S, fvy.S), usig[v]), {mask[v]});
                auto it = fac.find(key(X, Y));
                if (it != fac.end() && it->second.kind == 1)
                    rc = join_sparse(X, Y, v, dom, mask, fxv, fvy, it->second, Snew, u[v]);
                else
                    rc = join_dense(X, Y, v, dom, mask, fxv, fvy, fac, Snew, u[v]);
                if (rc) return rc;
            }
            // drop v's factors
            for (auto it = fac.begin(); it != fac.end();) {
                if (it->first.first == v || it->first.second == v) {
                    drop(it->second);
                    it = fac.erase(it);
                } else {
                    ++it;
                }
            }
        }
        return ASDM_OK;
    }

    // Eliminates v between X < Y when (X, Y) already carry the edge factor fxy: the product of
    // v's factors is evaluated on fxy's edges only and multiplied into their weights, so the
    // factor stays an edge relation.
    int join_sparse(uint32_t X, uint32_t Y, uint32_t v, const std::vector<Dom> &dom, const std::vector<uint32_t> &mask,
                    const Fac &fxv, const Fac &fvy, Fac &fxy, const std::vector<uint32_t> &Snew, const uint64_t *uv) {
        const uint32_t nx = dom[X].n, ny = dom[Y].n, nv = dom[v].n;
        const Rel rxy = fxy.lo, ryx = fxy.hi;  // rows D(X), rows D(Y)
        std::vector<uint32_t> S(fxy.S);
        S.insert(S.end(), Snew.begin(), Snew.end());
        std::sort(S.begin(), S.end());
        const auto k = sig(1, mask[X], mask[Y], S);
        const uint64_t bytes = 8ull * (rxy.nnz + ryx.nnz);
        auto install = [&](uint64_t *wf, uint64_t *wr, bool owned) {
            if (fxy.owned)
                for (uint64_t *w : fxy.w)
                    if (w) release(w);
            fxy.w[0] = wf;
            fxy.w[1] = wr;
            fxy.lo.w = wf;
            fxy.hi.w = wr;
            fxy.owned = owned;
            fxy.S = S;
        };
        if (const CacheEnt *ce = lookup(k, bytes)) {  // p holds the rows of block mask ce->rows
            if (ce->rows == mask[X]) install(ce->p, ce->p2, false);
            else install(ce->p2, ce->p, false);
            return ASDM_OK;
        }
        uint64_t *wf = nullptr, *wr = nullptr;
        int rc = alloc((void **)&wf, 8ull * std::max<uint64_t>(rxy.nnz, 1));
        if (!rc) rc = alloc((void **)&wr, 8ull * std::max<uint64_t>(ryx.nnz, 1));
        if (rc) return rc;
        const dim3 grid(grid_for(rxy.nnz * kJoinLanes, kThreads));
        auto relv = [&](const Fac &f, uint32_t a) { return a < v ? f.lo : f.hi; };  // rows D(a) of (a, v)
        bool reversed = false;  // the join wrote wr (entries of R(Y->X)) instead of wf
        if (fxv.kind == 1 && fvy.kind == 1) {
            const Rel ra = relv(fxv, X), rb = relv(fvy, Y);
            KTimer kt(ctx, "bag_join_ee", KTimer::kDispatch);
            kt.launch(k_join_ee, grid, dim3(kThreads), 0, rxy.off, rxy.val, rxy.w, nx, rxy.nnz, ra.off, ra.val, ra.w,
                      rb.off, rb.val, rb.w, uv, wf, bytes_slot(12));
            kt.stop(8ull * (nx + 1) + 12 * rxy.nnz + 8ull * (nx + ny + 2) + 12 * (ra.nnz + rb.nnz) + 8ull * nv);
        } else if (fxv.kind == 1 || fvy.kind == 1) {
            const bool own_x = fxv.kind == 1;  // the list side is X: F is the (v, Y) table
            const uint32_t o = own_x ? X : Y, a = own_x ? Y : X;
            const Rel rl = relv(own_x ? fxv : fvy, o);
            const Table &td = (own_x ? fvy : fxv).t;
            // F(a, j): rows D(a) -> a * nv + j; rows D(v) -> j * |D(a)| + a
            const uint64_t sa = td.rows == a ? nv : 1, sj = td.rows == a ? 1 : dom[a].n;
            // walk the edges grouped by a (the orientation of the joining relation whose rows are
            // D(a)), so that consecutive edges gather from the same part of F; the list owner o is
            // then the column of each entry
            const Rel rw_ = a == X ? rxy : ryx;
            const uint32_t nrows = a == X ? nx : ny;
            reversed = a != X;
            KTimer kt(ctx, "bag_join_ed", KTimer::kDispatch);
            kt.launch(k_join_ed, dim3(grid_for(rw_.nnz * kJoinLanes, kThreads)), dim3(kThreads), 0, rw_.off, rw_.val,
                      rw_.w, nrows, rw_.nnz, rl.off, rl.val, rl.w, 0, td.p, sa, sj, uv, reversed ? wr : wf,
                      bytes_slot(13), d_acc + 30);
            // the relations and lists once; the table entries it gathers (8 B each, one per list
            // entry of each joined edge) are counted on the device and added when the query ends
            kt.stop(8ull * (nrows + 1) + 12 * rw_.nnz + 8ull * (dom[o].n + 1) + 12 * rl.nnz + 8ull * nv);
        } else {
            // A(x, j) and B(j, y) read with j contiguous: A with rows D(X) (x * nv + j) and B with rows
            // D(Y) (y * nv + j), transposed first when stored the other way (a wave reads one row of
            // each per edge; a strided row would touch one cache line per element)
            uint64_t *Ar = nullptr, *Br = nullptr;
            bool ta = false, tb = false;
            if ((rc = oriented(fxv.t, fxv.t.rows == X ? v : X, X, dom, &Ar, ta))) return rc;
            if ((rc = oriented(fvy.t, fvy.t.rows == Y ? v : Y, Y, dom, &Br, tb))) return rc;
            KTimer kt(ctx, "bag_join_dd", KTimer::kDispatch);
            kt.launch(k_join_dd, dim3(grid_for(rxy.nnz, kThreads / 64)), dim3(kThreads), 0, rxy.off, rxy.val, rxy.w, nx,
                      rxy.nnz, Ar, (uint64_t)nv, 1ull, Br, 1ull, (uint64_t)nv, nv, uv, wf, bytes_slot(13));
            kt.stop(8ull * (nx + 1) + 12 * rxy.nnz + 8ull * nv * (nx + ny + 1));
            if (ta) release(Ar);
            if (tb) release(Br);
        }
        if ((rc = launched("bag sparse join"))) return rc;
        // the other orientation's weights: entry (y, x) of R(Y->X) takes the weight of (x, y)
        if (reversed)
            hipLaunchKernelGGL(k_join_reverse, dim3(grid_for(rxy.nnz, kThreads)), dim3(kThreads), 0, ctx->stream, ryx.off,
                               ryx.val, wr, rxy.off, rxy.val, nx, rxy.nnz, wf);
        else
            hipLaunchKernelGGL(k_join_reverse, dim3(grid_for(ryx.nnz, kThreads)), dim3(kThreads), 0, ctx->stream, rxy.off,
                               rxy.val, wf, ryx.off, ryx.val, ny, ryx.nnz, wr);
        if ((rc = launched("bag sparse join reverse"))) return rc;
        install(wf, wr, !store(k, CacheEnt{wf, wr, mask[X]}, bytes));
        return ASDM_OK;
    }

    // Eliminates v between X < Y into a new dense table on (X, Y), multiplied with the dense table
    // already there if any.
    int join_dense(uint32_t X, uint32_t Y, uint32_t v, const std::vector<Dom> &dom, const std::vector<uint32_t> &mask,
                   const Fac &fxv, const Fac &fvy, std::map<std::pair<uint32_t, uint32_t>, Fac> &fac,
                   const std::vector<uint32_t> &Snew, const uint64_t *uv) {
        const uint32_t nx = dom[X].n, ny = dom[Y].n, nv = dom[v].n;
        const uint64_t tbytes = 8ull * std::max<uint64_t>((uint64_t)nx * ny, 1);
        auto it = fac.find(std::make_pair(X, Y));
        auto row_block = [&](uint32_t m) { return m == mask[X] ? X : Y; };
        auto install = [&](Fac &f, uint64_t *p, uint32_t rows, bool owned, const std::vector<uint32_t> &S) {
            f.kind = 2;
            f.t = Table{p, rows};
            f.owned = owned;
            f.S = S;
        };
        // the joined result may be known already
        std::vector<uint32_t> Sj;
        std::vector<uint32_t> kj;
        if (it != fac.end()) {
            Sj = it->second.S;
            Sj.insert(Sj.end(), Snew.begin(), Snew.end());
            std::sort(Sj.begin(), Sj.end());
            kj = sig(2, mask[X], mask[Y], Sj);
            if (const CacheEnt *ce = lookup(kj, tbytes)) {
                Fac &f = it->second;
                if (f.owned) release(f.t.p);
                install(f, ce->p, row_block(ce->rows), false, Sj);
                return ASDM_OK;
            }
        }
        // the new table G
        const auto kg = sig(2, mask[X], mask[Y], Snew);
        Table out{};
        bool g_cached = false;
        int rc = ASDM_OK;
        if (const CacheEnt *ce = lookup(kg, tbytes)) {
            out = Table{ce->p, row_block(ce->rows)};
            g_cached = true;
        } else {
            if ((rc = product(X, Y, v, dom, fxv, fvy, uv, out))) return rc;
            g_cached = store(kg, CacheEnt{out.p, nullptr, mask[out.rows]}, tbytes);
        }
        if (it == fac.end()) {
            Fac f;
            install(f, out.p, out.rows, !g_cached, Snew);
            fac[std::make_pair(X, Y)] = f;
            return ASDM_OK;
        }
        // join with the dense table already on (X, Y): J = G o H in H's orientation
        Fac &f = it->second;
        const uint32_t ocol = out.rows == X ? Y : X;
        uint64_t *Gor = out.p;
        bool tmp = false;
        if (f.t.rows != out.rows && (rc = oriented(out, ocol, f.t.rows, dom, &Gor, tmp))) return rc;
        uint64_t *J = nullptr;
        if ((rc = alloc((void **)&J, tbytes))) return rc;
        const uint64_t n = (uint64_t)nx * ny;
        hipLaunchKernelGGL(k_hadamard, dim3(grid_for(n, kThreads)), dim3(kThreads), 0, ctx->stream, J, f.t.p, Gor, n);
        if ((rc = launched("bag join"))) return rc;
        if (tmp) release(Gor);
        if (!g_cached) release(out.p);
        const uint32_t rows = f.t.rows;
        if (f.owned) release(f.t.p);
        install(f, J, rows, !store(kj, CacheEnt{J, nullptr, mask[rows]}, tbytes), Sj);
        return ASDM_OK;
    }

    // The new dense table on (X, Y) from v's two factors (not both edge relations joined to an
    // existing edge: that is join_sparse).  out.rows says which block indexes its rows.
    int product(uint32_t X, uint32_t Y, uint32_t v, const std::vector<Dom> &dom, const Fac &fxv, const Fac &fvy,
                const uint64_t *uv, Table &out) {
        const uint32_t nx = dom[X].n, ny = dom[Y].n, nv = dom[v].n;
        uint64_t *G = nullptr;
        int rc = alloc((void **)&G, 8ull * std::max<uint64_t>((uint64_t)nx * ny, 1));
        if (rc) return rc;
        out = Table{G, X};
        const uint64_t out_bytes = 8ull * nx * ny;  // the table written once
        const char *kname = fxv.kind == 1 ? (fvy.kind == 1 ? "bag_table_ee" : "bag_table_ed")
                                          : (fvy.kind == 1 ? "bag_table_de" : "bag_table_dd");
        auto relv = [&](const Fac &f, uint32_t a) { return a < v ? f.lo : f.hi; };  // rows D(a) of (a, v)
        auto relfrom_v = [&](const Fac &f, uint32_t a) { return a < v ? f.hi : f.lo; };  // rows D(v) of (a, v)
        if (fxv.kind == 1 && fvy.kind == 1) {
            const Rel rx = relv(fxv, X), rv = relfrom_v(fvy, Y);
            KTimer kt(ctx, kname, KTimer::kDispatch);
            // count the terms, then both forms: each kernel returns at once unless the count picks
            // its form (the LDS form for many terms per cell, the scatter form for few)
            unsigned long long *terms = d_acc + 20;  // [0] the term count, [1] the LDS form's item queue
            if ((rc = fill_u64(ctx, (uint64_t *)terms, 2, 0))) return rc;
            const uint64_t cells = (uint64_t)nx * ny, resident = 8ull * ctx->num_cus;
            auto capped = [&](uint64_t items, uint64_t cap) {  // grid-stride kernels: at most cap workgroups
                return dim3((unsigned)std::max<uint64_t>(1, std::min<uint64_t>((items + kThreads - 1) / kThreads, cap)));
            };
            kt.launch(k_ee_terms, capped(rx.nnz, resident), dim3(kThreads), 0, rx.off, rx.val, rx.w, rv.off, uv, nx,
                      terms);
            kt.launch(k_ee_zero, capped(cells / 2, 2 * resident), dim3(kThreads), 0, G, cells, terms, ctx->table_form);
            kt.launch(k_ee_scatter, capped(rx.nnz, resident), dim3(kThreads), 0, rx.off, rx.val, rx.w, rv.off, rv.val,
                      rv.w, uv, nx, ny, G, terms, ctx->table_form, bytes_slot(8));
            // the LDS form: as many workgroups as fit the GPU at once (two per CU), fed by a queue
            const uint64_t items = (uint64_t)nx * ((ny + kEeCols - 1) / kEeCols);
            const unsigned grid = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(items, 2ull * ctx->num_cus));
            kt.launch(k_ee, dim3(grid), dim3(kEeThreads), 0, rx.off, rx.val, rx.w, rv.off, rv.val, rv.w, uv, nx, ny, G,
                      terms, ctx->table_form, terms + 1, bytes_slot(8));
            kt.stop(8ull * (nx + 1) + 4 * rx.nnz + 8ull * (nv + 1) + 4 * rv.nnz + 8ull * nv + out_bytes);
        } else if (fxv.kind == 1 || fvy.kind == 1) {
            // one dense side: G = R(e -> v) diag(u) F(v, d) with e the edge side and d the dense side;
            // the table comes out with rows D(e)
            const bool dense_x = fxv.kind == 2;
            const uint32_t e = dense_x ? Y : X, d = dense_x ? X : Y;
            const Table &td = dense_x ? fxv.t : fvy.t;
            uint64_t *F = nullptr;
            bool tmp = false;
            if ((rc = oriented(td, td.rows == v ? d : v, v, dom, &F, tmp))) return rc;
            const Rel re = relv(dense_x ? fvy : fxv, e);
            const uint32_t ne = dom[e].n, nd = dom[d].n;
            KTimer kt(ctx, kname, KTimer::kDispatch);
            // a resident grid, a multiple of 8 (one share per XCD): at most 8 workgroups per CU
            const uint64_t units = (uint64_t)((nd + 63) / 64) * ((ne + kEdsRows - 1) / kEdsRows);
            const uint64_t want = (units + kThreads / 64 - 1) / (kThreads / 64);
            const dim3 grid((unsigned)(8 * std::max<uint64_t>(1, std::min<uint64_t>((want + 7) / 8, ctx->num_cus))));
            if (dense_x)
                kt.launch(k_eds<1>, grid, dim3(kThreads), 0, re.off, re.val, re.w, uv, F, ne, nd, G, bytes_slot(9));
            else
                kt.launch(k_eds<0>, grid, dim3(kThreads), 0, re.off, re.val, re.w, uv, F, ne, nd, G, bytes_slot(10));
            kt.stop(8ull * (ne + 1) + 4 * re.nnz + 8ull * nv + 8ull * nv * nd + out_bytes);
            if (tmp) release(F);
            out.rows = e;
        } else if (nx >= 32 && ny >= 32 && nv >= 64 && (uint64_t)nx + 63 <= 64ull * 65535 &&
                   (uint64_t)nv + 63 <= 64ull * 65535) {
            // on the matrix cores (within the tile grid's 65,535 rows of workgroups), each operand
            // read in the orientation it is stored in
            KTimer kt(ctx, kname, KTimer::kDispatch);
            if ((rc = dense_product_mfma(ctx, fxv.t.p, fxv.t.rows != X, uv, fvy.t.p, fvy.t.rows == v, nx, nv, ny, G,
                                         bytes_slot(11), kt)))
                return rc;
            kt.stop(8ull * nx * nv + 8ull * nv * ny + 8ull * nv + out_bytes);
        } else {
            uint64_t *Fa = nullptr, *Fb = nullptr;
            bool ta = false, tb = false;
            if ((rc = oriented(fxv.t, fxv.t.rows == X ? v : X, X, dom, &Fa, ta))) return rc;
            if ((rc = oriented(fvy.t, fvy.t.rows == v ? Y : v, v, dom, &Fb, tb))) return rc;
            const uint64_t tiles = ((uint64_t)(nx + kT - 1) / kT) * ((ny + kT - 1) / kT);
            KTimer kt(ctx, kname, KTimer::kDispatch);
            kt.launch(k_dd, dim3(grid_for(tiles, 1)), dim3(256), 0, Fa, uv, Fb, nx, nv, ny, G, bytes_slot(11));
            kt.stop(8ull * nx * nv + 8ull * nv * ny + 8ull * nv + out_bytes);
            if (ta) release(Fa);
            if (tb) release(Fb);
        }
        return launched("bag table");
    }

    int run_term(const BagTerm &T) {
        const uint32_t nb = (uint32_t)T.blocks.size();
        std::vector<Dom> dom(nb);
        int rc = ASDM_OK;
        for (uint32_t b = 0; b < nb && !rc; ++b) rc = cached_dom(T.blocks[b], dom[b]);
        if (rc) return rc;
        for (uint32_t b = 0; b < nb; ++b)
            if (dom[b].n == 0) return ASDM_OK;  // no data vertex fits the merged block: the term is 0
        std::map<std::pair<uint32_t, uint32_t>, std::pair<Rel, Rel>> rels;
        for (size_t e = 0; e < T.edges.size() && !rc; ++e) {
            const uint32_t A = T.edges[e].first, B = T.edges[e].second;
            // the arcs run from the member of one block to the member of the other: orient them
            const auto [ab, ba] = edge_arcs(T, e);
            Rel r1, r2;
            rc = cached_rel(ab, T.blocks[A], dom[A], T.blocks[B], dom[B], r1);
            if (!rc) rc = cached_rel(ba, T.blocks[B], dom[B], T.blocks[A], dom[A], r2);
            rels[{A, B}] = {r1, r2};
        }
        if (rc) return rc;
        uint64_t *scal = (uint64_t *)(d_acc + 5);
        std::vector<uint64_t *> u;
        if ((rc = eliminate(T, T.count_steps, dom, rels, u, scal))) return rc;
        hipLaunchKernelGGL(k_total_add, dim3(1), dim3(64), 0, ctx->stream, d_acc, scal, T.mu);
        if ((rc = launched("bag total"))) return rc;
        release(u[0]);
        if (d_mat) {
            for (uint32_t t = 0; t < nb; ++t) {
                if ((rc = eliminate(T, T.local_steps[t], dom, rels, u, scal))) return rc;
                Members m{};
                m.n = (uint32_t)T.blocks[t].size();
                for (uint32_t k = 0; k < m.n; ++k) m.u[k] = T.blocks[t][k];
                KTimer kt(ctx, "bag_local_scatter", KTimer::kDispatch);
                kt.launch(k_scatter_local, dim3(grid_for(dom[t].n, kThreads)), dim3(kThreads), 0, dom[t].ids, dom[t].n,
                          u[t], scal, T.mu, m, cs->nq, d_mat);
                kt.stop(12ull * dom[t].n + 16ull * dom[t].n * m.n);
                if ((rc = launched("bag local scatter"))) return rc;
                release(u[0]);
            }
        }
        return ASDM_OK;
    }
};

}  // namespace

int run_bags(asdm_ctx *ctx, const asdm_plan *plan, const asdm_cs *cs, uint64_t *mat_host, uint64_t *count) {
    const BagPlan &B = *plan->bags;
    if (mat_host && !B.local) return ctx->fail(ASDM_E_INVALID, "local counting needs a plan made with ASDM_PLAN_LOCAL");
    (void)hipSetDevice(ctx->device);
    auto t0 = std::chrono::steady_clock::now();
    BagRun R{ctx, cs, &B};
    R.t0 = t0;
    {
        size_t freeb = 0, totalb = 0;
        (void)hipMemGetInfo(&freeb, &totalb);
        R.cache_cap = (uint64_t)((double)std::min<uint64_t>(freeb / 3, kBagCacheBytes) * ctx->mem_share);
    }
    R.d_acc = (unsigned long long *)ctx->d_counters;
    int rc = fill_u64(ctx, ctx->d_counters, kNumCounters, 0);
    const uint64_t matn = (uint64_t)cs->nv * cs->nq;
    if (!rc && mat_host) {
        rc = dev_alloc(ctx, (void **)&R.d_mat, 8ull * std::max<uint64_t>(matn, 1));
        if (!rc) rc = fill_u64(ctx, R.d_mat, matn, 0);
    }
    hipEvent_t ea, eb;  // the call's device time: timing only, no system-scope fence on record
    (void)hipEventCreateWithFlags(&ea, hipEventDisableSystemFence);
    (void)hipEventCreateWithFlags(&eb, hipEventDisableSystemFence);
    (void)hipEventRecord(ea, ctx->stream);
    if (!rc) rc = R.prebuild();
    for (size_t t = 0; t < B.terms.size() && !rc; ++t) {
        rc = R.run_term(B.terms[t]);
        R.release_all();
    }
    (void)hipEventRecord(eb, ctx->stream);
    R.release_cache();
    uint64_t res[kNumCounters] = {};
    if (!rc) rc = check_hip(ctx, hipMemcpyAsync(res, ctx->d_counters, 8 * kNumCounters, hipMemcpyDeviceToHost, ctx->stream), "D2H");
    if (!rc && mat_host)
        rc = check_hip(ctx, hipMemcpyAsync(mat_host, R.d_mat, 8 * matn, hipMemcpyDeviceToHost, ctx->stream), "D2H matrix");
    if (!rc) rc = check_hip(ctx, hipStreamSynchronize(ctx->stream), "sync");
    float kms = 0;
    (void)hipEventElapsedTime(&kms, ea, eb);
    (void)hipEventDestroy(ea);
    (void)hipEventDestroy(eb);
    R.release_cache();
    if (R.d_mat) dev_free(ctx, R.d_mat);
    (void)hipStreamSynchronize(ctx->stream);
    if (rc) return rc;
    *count = res[0];
    resolve_timers(ctx);
    // table kernels report their minimal traffic (each operand read once, the table written once)
    // through their timers; the in-kernel counter of the bytes their loops touch (slot 12,
    // including lists re-read per table row) is kept apart
    add_kernel_bytes(ctx, "bag_relation", res[6]);
    add_kernel_bytes(ctx, "bag_vector", res[7]);
    const char *touched[6] = {"bag_table_ee_touched", "bag_table_de_touched", "bag_table_ed_touched",
                              "bag_table_dd_touched", "bag_join_ee_touched", "bag_join_ed_touched"};
    for (int k = 0; k < 6; ++k)
        if (res[14 + k]) add_kernel_bytes(ctx, touched[k], res[14 + k]);
    if (res[30]) add_kernel_bytes(ctx, "bag_join_ed", 8 * res[30]);  // the table entries the ed joins gathered
    // factor results taken from the per-query cache instead of computed: entry "bag_cache_reuse",
    // launches = results reused, bytes = their sizes
    if (R.reuse_n) add_kernel_count(ctx, "bag_cache_reuse", R.reuse_n, R.reuse_bytes);
    ctx->stats.exec_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    ctx->stats.kernel_ms = kms;
    ctx->stats.alg_bytes = res[6] + res[7];
    ctx->stats.rows_total = 0;
    ctx->stats.launches = 0;
    ctx->stats.plan_ms = plan->plan_ms;
    ctx->stats.filter_ms = cs->filter_ms;
    return ASDM_OK;
}

}  // namespace asdm

using namespace asdm;

extern "C" {

int asdm_dense_product(asdm_ctx *ctx, const uint64_t *A, int a_kmajor, const uint64_t *u, const uint64_t *B, int b_kmajor,
                       uint32_t nx, uint32_t nv, uint32_t ny, uint64_t *G, int method) try {
    if (!ctx || !A || !B || !G || (method != 1 && method != 2)) return ASDM_E_INVALID;
    if (method == 1 && (a_kmajor || !b_kmajor)) return ctx->fail(ASDM_E_INVALID, "the LDS-tile kernel takes A by rows, B k-major");
    (void)hipSetDevice(ctx->device);
    const uint64_t na = (uint64_t)nx * nv, nb = (uint64_t)nv * ny, ng = (uint64_t)nx * ny;
    uint64_t *dA = nullptr, *dB = nullptr, *du = nullptr, *dG = nullptr;
    int rc = dev_alloc(ctx, (void **)&dA, 8 * std::max<uint64_t>(na, 1));
    if (!rc) rc = dev_alloc(ctx, (void **)&dB, 8 * std::max<uint64_t>(nb, 1));
    if (!rc) rc = dev_alloc(ctx, (void **)&du, 8 * std::max<uint64_t>(nv, 1));
    if (!rc) rc = dev_alloc(ctx, (void **)&dG, 8 * std::max<uint64_t>(ng, 1));
    if (!rc && na) rc = check_hip(ctx, hipMemcpyAsync(dA, A, 8 * na, hipMemcpyHostToDevice, ctx->stream), "copy");
    if (!rc && nb) rc = check_hip(ctx, hipMemcpyAsync(dB, B, 8 * nb, hipMemcpyHostToDevice, ctx->stream), "copy");
    if (!rc && nv) {
        if (u) rc = check_hip(ctx, hipMemcpyAsync(du, u, 8 * nv, hipMemcpyHostToDevice, ctx->stream), "copy");
        else rc = fill_u64(ctx, du, nv, 1);
    }
    unsigned long long *nbytes = (unsigned long long *)ctx->d_counters + 1;
    if (!rc && ng) {
        if (method == 2) {
            KTimer kt(ctx, "dense_product", KTimer::kDispatch);
            rc = dense_product_mfma(ctx, dA, a_kmajor != 0, du, dB, b_kmajor != 0, nx, nv, ny, dG, nbytes, kt);
            kt.stop();
        } else {
            const uint64_t tiles = ((uint64_t)(nx + kT - 1) / kT) * ((ny + kT - 1) / kT);
            hipLaunchKernelGGL(k_dd, dim3(grid_for(tiles, 1)), dim3(256), 0, ctx->stream, dA, du, dB, nx, nv, ny, dG, nbytes);
            rc = check_hip(ctx, hipGetLastError(), "dense product");
        }
    }
    if (!rc && ng) rc = check_hip(ctx, hipMemcpyAsync(G, dG, 8 * ng, hipMemcpyDeviceToHost, ctx->stream), "copy");
    if (!rc) rc = check_hip(ctx, hipStreamSynchronize(ctx->stream), "sync");
    (void)hipStreamSynchronize(ctx->stream);
    for (uint64_t *p : {dA, dB, du, dG})
        if (p) dev_free(ctx, p);
    return rc;
} ASDM_CTX_CATCH(ctx)

}  // extern "C"
