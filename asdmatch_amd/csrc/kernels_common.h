// Device helpers shared by the kernels: wave64 scans and reductions, lane groups, searches.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace asdm {

// Inclusive scan across the 64 lanes of a wave.
__device__ inline uint64_t wave_inclusive_scan(uint64_t v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint64_t t = __shfl_up(v, (unsigned)o, 64);
        if (lane >= o) v += t;
    }
    return v;
}

// Exclusive scan of one value per thread over a block of NT threads (NT multiple of 64).
// lds must hold NT/64 + 1 entries.  Returns the exclusive prefix; total receives the block sum.
template <int NT>
__device__ inline uint64_t block_exclusive_scan(uint64_t v, uint64_t *lds, uint64_t &total) {
    constexpr int NW = NT / 64;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint64_t inc = wave_inclusive_scan(v);
    if (lane == 63) lds[wave] = inc;
    __syncthreads();
    if (wave == 0) {
        uint64_t w = lane < NW ? lds[lane] : 0;
        uint64_t wi = wave_inclusive_scan(w);
        if (lane < NW) lds[lane] = wi - w;
        if (lane == NW - 1) lds[NW] = wi;
    }
    __syncthreads();
    uint64_t r = lds[wave] + inc - v;
    total = lds[NW];
    __syncthreads();
    return r;
}

// Sum over the G lanes of an aligned lane group (G power of two <= 64).
template <int G>
__device__ inline uint64_t group_sum(uint64_t v) {
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) v += __shfl_xor(v, (unsigned)o, 64);
    return v;
}

// Bits of a wave ballot that belong to this lane's aligned group of G lanes, shifted to bit 0.
template <int G>
__device__ inline uint64_t group_bits(uint64_t ballot) {
    if constexpr (G == 64) {
        return ballot;
    } else {
        const int g0 = (threadIdx.x & 63) & ~(G - 1);
        return (ballot >> g0) & ((1ull << G) - 1);
    }
}

// First position p in a[0, n) with a[p] >= x.
__device__ inline uint32_t lower_bound_u32(const uint32_t *__restrict__ a, uint32_t n, uint32_t x) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        uint32_t mid = (lo + hi) >> 1;
        if (a[mid] < x) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

__device__ inline bool contains_u32(const uint32_t *__restrict__ a, uint32_t n, uint32_t x) {
    uint32_t p = lower_bound_u32(a, n, x);
    return p < n && a[p] == x;
}

}  // namespace asdm
