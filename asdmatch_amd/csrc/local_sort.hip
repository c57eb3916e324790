// Sort-and-reduce step of the atomic-free LOCAL_COUNT aggregation: (key, value) pairs of one
// batch of leaf rows are radix-sorted by key and reduced by key (hipCUB on rocPRIM), so that every
// matrix cell receives one plain read-modify-write per batch.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include "engine.h"

namespace asdm {

int sort_reduce_pairs(asdm_ctx *ctx, const uint64_t *keys, const uint64_t *vals, uint64_t n, int end_bit,
                      uint64_t *keys_sorted, uint64_t *vals_sorted, uint64_t *ukeys, uint64_t *usums,
                      uint32_t *nruns) {
    if (n == 0) return fill_u32(ctx, nruns, 1, 0);
    if (n > (1ull << 31) - 1) return ctx->fail(ASDM_E_BUDGET, "local-count pair batch too large");
    const int ni = (int)n;
    size_t sort_bytes = 0, red_bytes = 0;
    if (check_hip(ctx, hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, keys, keys_sorted, vals, vals_sorted, ni, 0,
                                                          end_bit, ctx->stream), "sort size"))
        return ASDM_E_HIP;
    if (check_hip(ctx, hipcub::DeviceReduce::ReduceByKey(nullptr, red_bytes, keys_sorted, ukeys, vals_sorted, usums, nruns,
                                                         hipcub::Sum(), ni, ctx->stream), "reduce size"))
        return ASDM_E_HIP;
    void *temp = nullptr;
    const size_t tb = std::max<size_t>(std::max(sort_bytes, red_bytes), 256);
    int rc = dev_alloc(ctx, &temp, tb);
    if (rc) return rc;
    size_t b1 = sort_bytes, b2 = red_bytes;
    rc = check_hip(ctx, hipcub::DeviceRadixSort::SortPairs(temp, b1, keys, keys_sorted, vals, vals_sorted, ni, 0, end_bit,
                                                           ctx->stream), "radix sort");
    if (!rc)
        rc = check_hip(ctx, hipcub::DeviceReduce::ReduceByKey(temp, b2, keys_sorted, ukeys, vals_sorted, usums, nruns,
                                                              hipcub::Sum(), ni, ctx->stream), "reduce by key");
    dev_free(ctx, temp);
    return rc;
}

}  // namespace asdm
