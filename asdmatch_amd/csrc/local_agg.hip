// Owner-partitioned LOCAL_COUNT aggregation: the (cell, value) pairs of one batch of leaf rows
// (cell = data vertex * |Vq| + query vertex) are grouped by bucket, a range of kAggCells
// consecutive cells, with a counting pass, a scan and a scatter; then one workgroup per bucket
// sums its pairs in LDS and adds each nonzero cell of its range into the matrix with one plain
// read-modify-write.  Every matrix cell has a single writer, so the matrix needs no atomics.
//
// The counting and scatter passes take the pairs in tiles of kAggTile; a tile counts its pairs per
// bucket in LDS and reserves each bucket's slots with one global atomic, so a bucket that many
// pairs hit (a hub vertex) costs one atomic per tile, not one per pair.
#include <hip/hip_runtime.h>

#include "engine.h"
#include "kernels_common.h"

namespace asdm {

namespace {

constexpr uint32_t kAggShift = 12;                 // 4096 cells (32 KiB of u64 sums) per bucket
constexpr uint32_t kAggCells = 1u << kAggShift;
constexpr uint32_t kAggLdsBuckets = 4096;          // buckets counted in LDS per tile (beyond: global atomics)
constexpr uint32_t kAggTile = 8192;                // pairs per tile of the counting and scatter passes
constexpr int kAggThreads = 256;

// hist[b] += pairs of bucket b.
__global__ void __launch_bounds__(kAggThreads) k_agg_count(const uint64_t *__restrict__ keys, uint64_t n, uint32_t nb,
                                                           uint32_t *__restrict__ hist) {
    __shared__ uint32_t h[kAggLdsBuckets];
    const bool lds = nb <= kAggLdsBuckets;
    const uint64_t ntiles = (n + kAggTile - 1) / kAggTile;
    for (uint64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const uint64_t a = t * kAggTile, e = min(n, a + kAggTile);
        if (lds) {
            for (uint32_t i = threadIdx.x; i < nb; i += kAggThreads) h[i] = 0;
            __syncthreads();
        }
        for (uint64_t i = a + threadIdx.x; i < e; i += kAggThreads) {
            const uint32_t b = (uint32_t)(keys[i] >> kAggShift);
            if (lds) atomicAdd(&h[b], 1u);
            else atomicAdd(&hist[b], 1u);
        }
        if (lds) {
            __syncthreads();
            for (uint32_t i = threadIdx.x; i < nb; i += kAggThreads)
                if (h[i]) atomicAdd(&hist[i], h[i]);
            __syncthreads();
        }
    }
}

// Pair i of bucket b goes to boff[b] + (its slot among b's pairs): cells relative to the bucket
// start (32 bits) and values.  cursor[nb] starts at zero.
__global__ void __launch_bounds__(kAggThreads) k_agg_scatter(const uint64_t *__restrict__ keys,
                                                             const uint64_t *__restrict__ vals, uint64_t n, uint32_t nb,
                                                             const uint64_t *__restrict__ boff, uint32_t *__restrict__ cursor,
                                                             uint32_t *__restrict__ ocell, uint64_t *__restrict__ oval) {
    __shared__ uint32_t cnt[kAggLdsBuckets];
    __shared__ uint32_t base[kAggLdsBuckets];
    const bool lds = nb <= kAggLdsBuckets;
    const uint64_t ntiles = (n + kAggTile - 1) / kAggTile;
    for (uint64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const uint64_t a = t * kAggTile, e = min(n, a + kAggTile);
        if (lds) {
            for (uint32_t i = threadIdx.x; i < nb; i += kAggThreads) cnt[i] = 0;
            __syncthreads();
            for (uint64_t i = a + threadIdx.x; i < e; i += kAggThreads) atomicAdd(&cnt[(uint32_t)(keys[i] >> kAggShift)], 1u);
            __syncthreads();
            for (uint32_t i = threadIdx.x; i < nb; i += kAggThreads) {
                base[i] = cnt[i] ? atomicAdd(&cursor[i], cnt[i]) : 0;
                cnt[i] = 0;
            }
            __syncthreads();
        }
        for (uint64_t i = a + threadIdx.x; i < e; i += kAggThreads) {
            const uint64_t k = keys[i];
            const uint32_t b = (uint32_t)(k >> kAggShift);
            const uint32_t slot = lds ? base[b] + atomicAdd(&cnt[b], 1u) : atomicAdd(&cursor[b], 1u);
            const uint64_t p = boff[b] + slot;
            ocell[p] = (uint32_t)(k & (kAggCells - 1));
            oval[p] = vals[i];
        }
        if (lds) __syncthreads();
    }
}

// One workgroup per bucket: sums its pairs per cell in LDS, then mat[cell] += sum for the
// bucket's nonzero cells (the only writer of those cells in this launch).
__global__ void __launch_bounds__(kAggThreads) k_agg_sum(const uint32_t *__restrict__ ocell,
                                                         const uint64_t *__restrict__ oval,
                                                         const uint64_t *__restrict__ boff, uint32_t nb, uint64_t cells,
                                                         unsigned long long *__restrict__ mat) {
    __shared__ unsigned long long acc[kAggCells];
    for (uint32_t b = blockIdx.x; b < nb; b += gridDim.x) {
        const uint64_t a = boff[b], e = boff[b + 1];
        if (a == e) continue;  // uniform across the workgroup
        for (uint32_t i = threadIdx.x; i < kAggCells; i += kAggThreads) acc[i] = 0;
        __syncthreads();
        for (uint64_t i = a + threadIdx.x; i < e; i += kAggThreads) atomicAdd(&acc[ocell[i]], (unsigned long long)oval[i]);
        __syncthreads();
        const uint64_t c0 = (uint64_t)b << kAggShift;
        for (uint32_t i = threadIdx.x; i < kAggCells && c0 + i < cells; i += kAggThreads)
            if (acc[i]) mat[c0 + i] += acc[i];
        __syncthreads();
    }
}

}  // namespace

int aggregate_pairs(asdm_ctx *ctx, const uint64_t *keys, const uint64_t *vals, uint64_t n, uint64_t cells,
                    uint64_t *mat) {
    if (n == 0) return ASDM_OK;
    const uint64_t nb64 = (cells + kAggCells - 1) >> kAggShift;
    if (nb64 > 0xffffffffull || n > 0xffffffffull)
        return ctx->fail(ASDM_E_UNSUPPORTED, "local-count aggregation beyond 2^32 buckets or pairs per batch");
    const uint32_t nb = (uint32_t)nb64;
    uint32_t *hist = nullptr, *cursor = nullptr, *ocell = nullptr;
    uint64_t *boff = nullptr, *oval = nullptr;
    std::vector<void *> tmp;
    auto alloc = [&](void **p, size_t bytes) {
        const int rc = dev_alloc(ctx, p, std::max<size_t>(bytes, 8));
        if (!rc) tmp.push_back(*p);
        return rc;
    };
    int rc = alloc((void **)&hist, 4ull * nb);
    if (!rc) rc = alloc((void **)&cursor, 4ull * nb);
    if (!rc) rc = alloc((void **)&boff, 8ull * (nb + 1));
    if (!rc) rc = alloc((void **)&ocell, 4ull * n);
    if (!rc) rc = alloc((void **)&oval, 8ull * n);
    if (!rc) rc = fill_u32(ctx, hist, nb, 0);
    if (!rc) rc = fill_u32(ctx, cursor, nb, 0);
    if (!rc) {
        const uint64_t ntiles = (n + kAggTile - 1) / kAggTile;
        const unsigned grid = (unsigned)std::min<uint64_t>(ntiles, 8ull * ctx->num_cus);
        {
            KTimer kt(ctx, "local_agg_bucket", KTimer::kDispatch);
            kt.launch(k_agg_count, dim3(grid), dim3(kAggThreads), 0, keys, n, nb, hist);
            rc = check_hip(ctx, hipGetLastError(), "aggregation count");
            if (!rc) rc = scan_u32_to_u64(ctx, hist, boff, nb);
            if (!rc) {
                kt.launch(k_agg_scatter, dim3(grid), dim3(kAggThreads), 0, keys, vals, n, nb, boff, cursor, ocell,
                          oval);
                rc = check_hip(ctx, hipGetLastError(), "aggregation scatter");
            }
            kt.stop(8 * n + 16 * n + 12 * n + 16ull * nb);
        }
        if (!rc) {
            KTimer kt(ctx, "local_agg_sum", KTimer::kDispatch);
            const unsigned g2 = (unsigned)std::min<uint64_t>(nb, 8ull * ctx->num_cus);
            kt.launch(k_agg_sum, dim3(g2), dim3(kAggThreads), 0, ocell, oval, boff, nb, cells,
                      (unsigned long long *)mat);
            rc = check_hip(ctx, hipGetLastError(), "aggregation sum");
            kt.stop(12 * n + 8ull * nb + 16ull * std::min<uint64_t>(n, cells));
        }
    }
    for (void *p : tmp) dev_free(ctx, p);
    return rc;
}

}  // namespace asdm
