// Context, device memory, prefix scan, fills, graph upload and the k-way intersection primitive.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <mutex>

#include "engine.h"
#include "kernels_common.h"

namespace asdm {

int check_hip(asdm_ctx *ctx, hipError_t e, const char *what) {
    if (e == hipSuccess) return ASDM_OK;
    if (ctx) ctx->err = std::string(what) + ": " + hipGetErrorString(e);
    return e == hipErrorOutOfMemory ? ASDM_E_OOM : ASDM_E_HIP;
}

// Frees every cached block (after the context's stream has drained: a cached block's last use was
// on that stream).
static void trim_cache(asdm_ctx *ctx) {
    if (ctx->cache.empty()) return;
    (void)hipStreamSynchronize(ctx->stream);
    for (auto &b : ctx->cache) (void)hipFree(b.p);
    ctx->cache.clear();
    ctx->cached_bytes = 0;
}

// Frees the least recently cached blocks until at most `keep` bytes stay cached.  hipFree waits
// for the whole device, so a trim is rare and keeps the blocks most likely to be asked for again.
static void trim_cache_to(asdm_ctx *ctx, size_t keep) {
    if (ctx->cached_bytes <= keep) return;
    (void)hipStreamSynchronize(ctx->stream);
    size_t i = 0;  // the cache is in the order the blocks were freed
    while (i < ctx->cache.size() && ctx->cached_bytes > keep) {
        (void)hipFree(ctx->cache[i].p);
        ctx->cached_bytes -= ctx->cache[i].bytes;
        ++i;
    }
    ctx->cache.erase(ctx->cache.begin(), ctx->cache.begin() + i);
    ++ctx->trims;
}

// Size class of a request: 256-byte granules below 64 KiB, above that eight classes per power of
// two (at most 12.5 % over the request), so blocks freed by one query fit the next one's requests.
static size_t size_class(size_t bytes) {
    bytes = (std::max<size_t>(bytes, 16) + 255) & ~(size_t)255;
    if (bytes <= (64u << 10)) return bytes;
    int top = 63 - __builtin_clzll((unsigned long long)bytes);
    const size_t step = (size_t)1 << (top - 3);
    return (bytes + step - 1) & ~(step - 1);
}

// Bytes of cached (free) blocks the context keeps: its share of 96 GiB.
static size_t cache_limit(const asdm_ctx *ctx) { return (size_t)((double)(96ull << 30) * ctx->mem_share); }

int dev_alloc(asdm_ctx *ctx, void **p, size_t bytes) {
    *p = nullptr;
    bytes = size_class(bytes);
    // the same class, else the best fit among cached blocks no more than 2x (+1 MiB) larger
    int best = -1;
    for (size_t i = 0; i < ctx->cache.size(); ++i) {
        const size_t b = ctx->cache[i].bytes;
        if (b == bytes) {
            best = (int)i;
            break;
        }
        if (b >= bytes && b <= 2 * bytes + (1u << 20) && (best < 0 || b < ctx->cache[best].bytes)) best = (int)i;
    }
    if (best >= 0) {
        asdm_ctx::Block blk{ctx->cache[best].p, ctx->cache[best].bytes};
        ctx->cache.erase(ctx->cache.begin() + best);
        ctx->cached_bytes -= blk.bytes;
        ctx->live.push_back({blk.p, blk.bytes});
        *p = blk.p;
        return ASDM_OK;
    }
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        trim_cache(ctx);
        e = hipMalloc(p, bytes);
    }
    if (e != hipSuccess) {
        (void)hipGetLastError();
        *p = nullptr;
        ctx->err = "device allocation of " + std::to_string(bytes) + " bytes failed: " + hipGetErrorString(e);
        return ASDM_E_OOM;
    }
    ctx->live.push_back({*p, bytes});
    return ASDM_OK;
}

void dev_free(asdm_ctx *ctx, void *p) {
    if (!p) return;
    for (size_t i = 0; i < ctx->live.size(); ++i)
        if (ctx->live[i].p == p) {
            ctx->cache.push_back({p, ctx->live[i].bytes});
            ctx->cached_bytes += ctx->live[i].bytes;
            ctx->live[i] = ctx->live.back();
            ctx->live.pop_back();
            break;
        }
    const size_t lim = cache_limit(ctx);
    if (ctx->cached_bytes > lim) trim_cache_to(ctx, lim / 2);
}

// ---------------------------------------------------------------- scan
constexpr int kScanThreads = 1024;
constexpr int kScanItems = 8;
constexpr uint64_t kScanTile = (uint64_t)kScanThreads * kScanItems;

__global__ void k_scan_tile_sums(const uint32_t *__restrict__ in, uint64_t n, uint64_t *__restrict__ tile_sums) {
    __shared__ uint64_t lds[kScanThreads / 64 + 1];
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanItems;
    uint64_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k)
        if (base + k < n) s += in[base + k];
    uint64_t total;
    (void)block_exclusive_scan<kScanThreads>(s, lds, total);
    if (threadIdx.x == 0) tile_sums[blockIdx.x] = total;
}

// Single block: exclusive scan of the tile sums in place, total into *total_out.
__global__ void k_scan_tiles(uint64_t *__restrict__ tile_sums, uint64_t ntiles, uint64_t *__restrict__ total_out) {
    __shared__ uint64_t lds[kScanThreads / 64 + 1];
    uint64_t carry = 0;
    for (uint64_t b = 0; b < ntiles; b += kScanThreads) {
        uint64_t i = b + threadIdx.x;
        uint64_t v = i < ntiles ? tile_sums[i] : 0;
        uint64_t tot;
        uint64_t ex = block_exclusive_scan<kScanThreads>(v, lds, tot);
        if (i < ntiles) tile_sums[i] = carry + ex;
        carry += tot;
        __syncthreads();
    }
    if (threadIdx.x == 0) *total_out = carry;
}

__global__ void k_scan_apply(const uint32_t *__restrict__ in, uint64_t n, const uint64_t *__restrict__ tile_pref,
                             uint64_t *__restrict__ out) {
    __shared__ uint64_t lds[kScanThreads / 64 + 1];
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanItems;
    uint32_t v[kScanItems];
    uint64_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        v[k] = base + k < n ? in[base + k] : 0u;
        s += v[k];
    }
    uint64_t tot;
    uint64_t ex = block_exclusive_scan<kScanThreads>(s, lds, tot) + tile_pref[blockIdx.x];
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        if (base + k < n) out[base + k] = ex;
        ex += v[k];
    }
}

int scan_u32_to_u64(asdm_ctx *ctx, const uint32_t *in, uint64_t *out, uint64_t n) {
    if (n == 0) return fill_u64(ctx, out, 1, 0);
    uint64_t ntiles = (n + kScanTile - 1) / kScanTile;
    uint64_t *tiles = nullptr;
    int rc = dev_alloc(ctx, (void **)&tiles, sizeof(uint64_t) * ntiles);
    if (rc) return rc;
    hipLaunchKernelGGL(k_scan_tile_sums, dim3((unsigned)ntiles), dim3(kScanThreads), 0, ctx->stream, in, n, tiles);
    hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(kScanThreads), 0, ctx->stream, tiles, ntiles, out + n);
    hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)ntiles), dim3(kScanThreads), 0, ctx->stream, in, n, tiles, out);
    dev_free(ctx, tiles);
    return check_hip(ctx, hipGetLastError(), "scan");
}

// ---------------------------------------------------------------- fills
__global__ void k_fill_u64(uint64_t *p, uint64_t n, uint64_t v) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_fill_u32(uint32_t *p, uint64_t n, uint32_t v) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) p[i] = v;
}

static unsigned grid_for(uint64_t n, unsigned block) {
    uint64_t g = (n + block - 1) / block;
    return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(g, 65536));
}

int fill_u64(asdm_ctx *ctx, uint64_t *p, uint64_t n, uint64_t v) {
    hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, p, n, v);
    return check_hip(ctx, hipGetLastError(), "fill");
}
int fill_u32(asdm_ctx *ctx, uint32_t *p, uint64_t n, uint32_t v) {
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, p, n, v);
    return check_hip(ctx, hipGetLastError(), "fill");
}

// ---------------------------------------------------------------- timing
static int kstat_slot(asdm_ctx *ctx, const char *name) {
    for (size_t i = 0; i < ctx->kstats.size(); ++i)
        if (ctx->kstats[i].name == name) return (int)i;
    ctx->kstats.push_back({name, 0, 0, 0});
    return (int)ctx->kstats.size() - 1;
}

// Creates the timing events a profiled call typically needs up front, so that their creation does
// not fall inside the first timed calls.
void reserve_timer_events(asdm_ctx *ctx) {
    constexpr size_t kReserve = 1024;
    (void)hipSetDevice(ctx->device);
    while (ctx->event_pool.size() < kReserve) {
        hipEvent_t e = nullptr;
        if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess) break;
        ctx->event_pool.push_back(e);
    }
}

hipEvent_t take_timer_event(asdm_ctx *ctx) {
    hipEvent_t e = nullptr;
    if (!ctx->event_pool.empty()) {
        e = ctx->event_pool.back();
        ctx->event_pool.pop_back();
    } else {
        // timing only: no system-scope fence (cache writeback and invalidation) when it is recorded
        (void)hipEventCreateWithFlags(&e, hipEventDisableSystemFence);
    }
    return e;
}

KTimer::KTimer(asdm_ctx *c, const char *name, Mode m) : ctx(c), slot(-1), mode(m) {
    if (!ctx->profiling) return;
    slot = kstat_slot(ctx, name);
    if (mode == kDispatch) return;
    a = take_timer_event(ctx);
    b = take_timer_event(ctx);
    (void)hipEventRecord(a, ctx->stream);
}

// Records the end event only; the elapsed time is read later by resolve_timers, so timing does
// not serialize the launch sequence.
void KTimer::stop(uint64_t bytes) {
    if (!ctx->profiling || slot < 0) return;
    ctx->kstats[slot].bytes += bytes;
    if (mode == kDispatch) {
        ctx->kstats[slot].launches += 1;
    } else {
        (void)hipEventRecord(b, ctx->stream);
        ctx->pending.push_back({slot, a, b, 1});
    }
    slot = -1;
}

void resolve_timers(asdm_ctx *ctx) {
    for (auto &t : ctx->pending) {
        (void)hipEventSynchronize(t.b);
        float ms = 0;
        (void)hipEventElapsedTime(&ms, t.a, t.b);
        ctx->kstats[t.slot].ms += ms;
        ctx->kstats[t.slot].launches += t.launches;
        ctx->event_pool.push_back(t.a);
        ctx->event_pool.push_back(t.b);
    }
    ctx->pending.clear();
}

void add_kernel_bytes(asdm_ctx *ctx, const char *name, uint64_t bytes) {
    if (!ctx->profiling) return;
    ctx->kstats[kstat_slot(ctx, name)].bytes += bytes;
}

void add_kernel_count(asdm_ctx *ctx, const char *name, uint64_t count, uint64_t bytes) {
    if (!ctx->profiling) return;
    auto &k = ctx->kstats[kstat_slot(ctx, name)];
    k.launches += count;
    k.bytes += bytes;
}

void add_host_ms(asdm_ctx *ctx, const char *name, double ms) {
    if (!ctx->profiling) return;
    auto &k = ctx->kstats[kstat_slot(ctx, name)];
    k.ms += ms;
    k.launches += 1;
}

void merge_kernel_stats(asdm_ctx *into, asdm_ctx *from) {
    resolve_timers(from);
    for (const auto &k : from->kstats) {
        auto &e = into->kstats[kstat_slot(into, k.name.c_str())];
        e.ms += k.ms;
        e.launches += k.launches;
        e.bytes += k.bytes;
    }
    from->kstats.clear();
}

void release_cached_blocks(asdm_ctx *ctx) { trim_cache(ctx); }

void release_device_graph(asdm_graph &g) {
    auto &d = g.dev;
    if (d.device < 0) return;
    int cur = 0;
    (void)hipGetDevice(&cur);
    (void)hipSetDevice(d.device);
    (void)hipFree(d.off);
    (void)hipFree(d.nbr);
    (void)hipFree(d.lab);
    (void)hipFree(d.loff);
    (void)hipFree(d.vbl_off);
    (void)hipFree(d.vbl);
    (void)hipFree(d.labbits);
    (void)hipSetDevice(cur);
    d = DeviceGraph{};
}

// ---------------------------------------------------------------- intersection primitive
// The allocation-free entry point's kernel, one wave per call: the shortest list is the probe;
// each lane tests probe elements by binary search in the other lists; survivors are compacted in
// order with a ballot prefix.
__global__ void k_intersect_k(const uint32_t *__restrict__ flat, const uint64_t *__restrict__ offs, uint32_t k,
                              uint32_t *__restrict__ out, uint32_t *__restrict__ nout) {
    const int lane = threadIdx.x & 63;
    uint32_t pj = 0;
    uint64_t plen = offs[1] - offs[0];
    for (uint32_t j = 1; j < k; ++j) {
        uint64_t len = offs[j + 1] - offs[j];
        if (len < plen) { plen = len; pj = j; }
    }
    const uint32_t *probe = flat + offs[pj];
    uint32_t run = 0;
    for (uint64_t t0 = 0; t0 < plen; t0 += 64) {
        uint64_t t = t0 + lane;
        bool ok = t < plen;
        uint32_t x = ok ? probe[t] : 0;
        for (uint32_t j = 0; j < k && ok; ++j) {
            if (j == pj) continue;
            const uint32_t *l = flat + offs[j];
            uint32_t n = (uint32_t)(offs[j + 1] - offs[j]);
            uint32_t p = lower_bound_u32(l, n, x);
            ok = p < n && l[p] == x;
        }
        uint64_t bal = __ballot(ok);
        if (ok) out[run + __popcll(bal & ((1ull << lane) - 1))] = x;
        run += __popcll(bal);
    }
    if (lane == 0) *nout = run;
}

// Pairwise intersection R n L of two ascending unique lists, as the k-way call's steps (the
// shortest list first, then each other list from the shortest up), in two passes over tiles of
// kIsectTile: pass 0 counts each tile's matches into cnt[tile], pass 1 writes them, in order, from
// off[tile] (the exclusive scan of the counts).
//  - Balanced pairs (|L| <= kIsectBalanced |R|): merge path.  Tile t holds merge positions
//    [t T, (t + 1) T) of R and L merged (R first on ties); both ends are found by a binary search
//    along their diagonals, the tile's part of L (plus the next element, where R's last element can
//    meet its equal) is staged in LDS, and each element of the tile's part of R is searched there.
//  - Otherwise: tile t holds R's elements [t T, (t + 1) T), each searched in L in global memory.
constexpr uint32_t kIsectTile = 2048, kIsectBalanced = 16;

// Elements of R among the first d of the merge of R (nr) and L (nl), R first on ties.
__device__ inline uint32_t merge_split(const uint32_t *__restrict__ R, uint32_t nr, const uint32_t *__restrict__ L,
                                       uint32_t nl, uint64_t d) {
    uint32_t lo = d > nl ? (uint32_t)(d - nl) : 0, hi = (uint32_t)min<uint64_t>(d, nr);
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (R[mid] <= L[d - mid - 1]) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

template <bool MERGE>
__global__ void __launch_bounds__(256) k_isect_pair(const uint32_t *__restrict__ R, uint32_t nr,
                                                    const uint32_t *__restrict__ L, uint32_t nl, int pass,
                                                    uint32_t *__restrict__ cnt, const uint64_t *__restrict__ off,
                                                    uint32_t *__restrict__ out) {
    __shared__ uint32_t seg[kIsectTile + 1];
    __shared__ uint64_t scan_lds[256 / 64 + 1];
    uint32_t r0, r1;
    const uint32_t *lseg;
    uint32_t ln;
    if (MERGE) {
        const uint64_t d0 = (uint64_t)blockIdx.x * kIsectTile, d1 = min<uint64_t>((uint64_t)nr + nl, d0 + kIsectTile);
        r0 = merge_split(R, nr, L, nl, d0);
        r1 = merge_split(R, nr, L, nl, d1);
        const uint32_t l0 = (uint32_t)(d0 - r0), l1 = min(nl, (uint32_t)(d1 - r1) + 1);
        ln = l1 - l0;  // at most kIsectTile + 1
        for (uint32_t i = threadIdx.x; i < ln; i += 256) seg[i] = L[l0 + i];
        __syncthreads();
        lseg = seg;
    } else {
        r0 = blockIdx.x * kIsectTile;
        r1 = min(nr, r0 + kIsectTile);
        lseg = L;
        ln = nl;
    }
    uint64_t base = pass ? off[blockIdx.x] : 0, total = 0;
    for (uint32_t i0 = r0; i0 < r1; i0 += 256) {  // rounds of 256 consecutive elements keep the order
        const uint32_t i = i0 + threadIdx.x;
        bool hit = false;
        uint32_t x = 0;
        if (i < r1) {
            x = R[i];
            const uint32_t q = lower_bound_u32(lseg, ln, x);
            hit = q < ln && lseg[q] == x;
        }
        uint64_t round = 0;
        const uint64_t ex = block_exclusive_scan<256>(hit ? 1 : 0, scan_lds, round);
        if (pass && hit) out[base + ex] = x;
        base += round;
        total += round;
    }
    if (!pass && threadIdx.x == 0) cnt[blockIdx.x] = (uint32_t)total;
}

// ---------------------------------------------------------------- allocation-free scratch
// Device memory of asdm_intersect_k_noalloc: the staged lists, their offsets, the result and its
// size, in one file-scope array of a fixed size.  Every device holds its own copy for the life of
// the process; host code reaches it only through the address hipGetSymbolAddress returns for the
// current device (the array's name is a host-side stand-in, not a device address).
constexpr uint64_t kNoAllocListWords = 1ull << 21;  // list elements staged per call
constexpr uint32_t kNoAllocMaxLists = 64;
constexpr size_t kNoAllocFlatOff = 0;
constexpr size_t kNoAllocOutOff = kNoAllocFlatOff + 4 * kNoAllocListWords;
constexpr size_t kNoAllocOffsOff = kNoAllocOutOff + 4 * kNoAllocListWords;
constexpr size_t kNoAllocCountOff = kNoAllocOffsOff + 8 * (kNoAllocMaxLists + 1);
constexpr size_t kNoAllocBytes = (kNoAllocCountOff + 8 + 255) & ~(size_t)255;

static __device__ __align__(256) uint8_t scratch[kNoAllocBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint8_t *g_scratch_ptr[kMaxDevices] = {};

// Looks up the current device's scratch address once per device (the device must be current).
int noalloc_scratch_for(asdm_ctx *ctx, int device, uint8_t **out) {
    if (device < 0 || device >= kMaxDevices) return ctx->fail(ASDM_E_INVALID, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        int rc = check_hip(ctx, hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "scratch address");
        if (rc) return rc;
        g_scratch_ptr[device] = (uint8_t *)p;
    }
    *out = g_scratch_ptr[device];
    return ASDM_OK;
}

}  // namespace asdm

using namespace asdm;

extern "C" {

int asdm_ctx_create(int device, asdm_ctx **out) try {
    if (!out) return ASDM_E_INVALID;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        (void)hipGetLastError();
        return ASDM_E_NODEVICE;
    }
    if (device < 0 || device >= ndev) return ASDM_E_INVALID;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return ASDM_E_NODEVICE;
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return ASDM_E_NODEVICE;
    auto ctx = std::make_unique<asdm_ctx>();
    ctx->device = device;
    ctx->num_cus = (uint32_t)std::max(8, prop.multiProcessorCount);
    ctx->err.reserve(512);  // room for any fixed message without allocating when it is set
    if (hipSetDevice(device) != hipSuccess) return ASDM_E_HIP;
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) return ASDM_E_HIP;
    if (hipMalloc((void **)&ctx->d_counters, sizeof(uint64_t) * kNumCounters) != hipSuccess) return ASDM_E_OOM;
    if (int rc = noalloc_scratch_for(ctx.get(), device, &ctx->noalloc_scratch)) {
        (void)hipFree(ctx->d_counters);
        (void)hipStreamDestroy(ctx->stream);
        return rc;
    }
    *out = ctx.release();
    return ASDM_OK;
} ASDM_HOST_CATCH

void asdm_ctx_destroy(asdm_ctx *ctx) {
    if (!ctx) return;
    for (asdm_ctx *w : ctx->workers) asdm_ctx_destroy(w);
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (auto &b : ctx->cache) (void)hipFree(b.p);
    for (auto &b : ctx->live) (void)hipFree(b.p);
    for (hipEvent_t e : ctx->event_pool) (void)hipEventDestroy(e);
    for (auto &t : ctx->pending) {
        (void)hipEventDestroy(t.a);
        (void)hipEventDestroy(t.b);
    }
    (void)hipFree(ctx->d_counters);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char *asdm_last_error(const asdm_ctx *ctx) { return ctx ? ctx->err.c_str() : "no context"; }

int asdm_last_stats(const asdm_ctx *ctx, asdm_stats *out) try {
    if (!ctx || !out) return ASDM_E_INVALID;
    *out = ctx->stats;
    return ASDM_OK;
} ASDM_HOST_CATCH

int asdm_ctx_set_time_limit(asdm_ctx *ctx, double ms) try {
    if (!ctx || ms < 0) return ASDM_E_INVALID;
    ctx->time_limit_ms = ms;
    return ASDM_OK;
} ASDM_CTX_CATCH(ctx)

int asdm_ctx_set_row_limit(asdm_ctx *ctx, uint64_t rows) try {
    if (!ctx) return ASDM_E_INVALID;
    ctx->row_limit = rows;
    return ASDM_OK;
} ASDM_CTX_CATCH(ctx)

int asdm_ctx_set_bag_workers(asdm_ctx *ctx, uint32_t workers) try {
    if (!ctx || workers < 1 || workers > 16) return ASDM_E_INVALID;
    ctx->bag_workers = workers;
    return ASDM_OK;
} ASDM_CTX_CATCH(ctx)

int asdm_ctx_set_table_form(asdm_ctx *ctx, int form) try {
    if (!ctx || form < 0 || form > 2) return ASDM_E_INVALID;
    ctx->table_form = form;
    return ASDM_OK;
} ASDM_CTX_CATCH(ctx)

int asdm_ctx_get_bag_workers(const asdm_ctx *ctx, uint32_t *workers) try {
    if (!ctx || !workers) return ASDM_E_INVALID;
    *workers = ctx->bag_workers;
    return ASDM_OK;
} ASDM_HOST_CATCH

// Per-kernel device timing, collected while profiling is on (HIP events on the context stream).
int asdm_ctx_set_profiling(asdm_ctx *ctx, int on) try {
    if (!ctx) return ASDM_E_INVALID;
    resolve_timers(ctx);
    ctx->profiling = on != 0;
    ctx->kstats.clear();
    if (ctx->profiling) reserve_timer_events(ctx);
    for (asdm_ctx *w : ctx->workers) {  // the bag workers follow, their events made now too
        resolve_timers(w);
        w->kstats.clear();
        w->profiling = ctx->profiling;
        if (w->profiling) reserve_timer_events(w);
    }
    return ASDM_OK;
} ASDM_CTX_CATCH(ctx)

int asdm_kernel_stat(const asdm_ctx *ctx, uint32_t i, char *name, uint32_t name_cap, double *ms, uint64_t *launches,
                     uint64_t *bytes) try {
    if (!ctx) return ASDM_E_INVALID;
    resolve_timers(const_cast<asdm_ctx *>(ctx));
    if (i >= ctx->kstats.size()) return ASDM_E_INVALID;
    const auto &k = ctx->kstats[i];
    if (name && name_cap) {
        std::strncpy(name, k.name.c_str(), name_cap - 1);
        name[name_cap - 1] = 0;
    }
    if (ms) *ms = k.ms;
    if (launches) *launches = k.launches;
    if (bytes) *bytes = k.bytes;
    return ASDM_OK;
} ASDM_HOST_CATCH

int asdm_graph_upload(asdm_ctx *ctx, asdm_graph *g) try {
    if (!ctx || !g) return ASDM_E_INVALID;
    if (g->dev.device == ctx->device) return ASDM_OK;
    if (g->dev.device >= 0) release_device_graph(*g);
    (void)hipSetDevice(ctx->device);
    const uint64_t L1 = (uint64_t)g->nlab + 1;
    if ((uint64_t)g->n * L1 >= (1ull << 33))
        return ctx->fail(ASDM_E_UNSUPPORTED, "per-vertex label offset table too large (n * (labels + 1) >= 2^33)");
    // loff[v * (L + 1) + l] = first arc of label >= l in v's adjacency
    std::vector<uint32_t> loff((size_t)(g->n * L1));
    for (uint32_t v = 0; v < g->n; ++v) {
        uint32_t i = g->off[v], e = g->off[v + 1];
        uint32_t *row = loff.data() + (size_t)v * L1;
        for (uint32_t l = 0; l <= g->nlab; ++l) {
            while (i < e && g->lab[g->nbr[i]] < l) ++i;
            row[l] = i;
        }
    }
    DeviceGraph d;
    d.device = ctx->device;
    auto up = [&](uint32_t **dst, const uint32_t *src, size_t n) -> int {
        if (hipMalloc((void **)dst, sizeof(uint32_t) * std::max<size_t>(n, 1)) != hipSuccess) return ASDM_E_OOM;
        if (n && hipMemcpy(*dst, src, sizeof(uint32_t) * n, hipMemcpyHostToDevice) != hipSuccess) return ASDM_E_HIP;
        return ASDM_OK;
    };
    int rc = up(&d.off, g->off.data(), (size_t)g->n + 1);
    if (!rc) rc = up(&d.nbr, g->nbr.data(), g->m);
    if (!rc) rc = up(&d.lab, g->lab.data(), g->n);
    if (!rc) rc = up(&d.loff, loff.data(), loff.size());
    if (!rc) rc = up(&d.vbl_off, g->vbl_off.data(), g->vbl_off.size());
    if (!rc) rc = up(&d.vbl, g->vbl.data(), g->vbl.size());
    if (!rc) {  // per-label vertex bitmaps: the base filter walks only the vertices of a label
        const size_t W = ((size_t)g->n + 63) / 64;
        std::vector<uint64_t> lb(std::max<size_t>(W * g->nlab, 1), 0);
        for (uint32_t v = 0; v < g->n; ++v) lb[(size_t)g->lab[v] * W + (v >> 6)] |= 1ull << (v & 63);
        if (hipMalloc((void **)&d.labbits, 8 * lb.size()) != hipSuccess) rc = ASDM_E_OOM;
        else if (hipMemcpy(d.labbits, lb.data(), 8 * lb.size(), hipMemcpyHostToDevice) != hipSuccess) rc = ASDM_E_HIP;
    }
    g->dev = d;
    if (rc) {
        release_device_graph(*g);
        return ctx->fail(rc, "graph upload failed");
    }
    return ASDM_OK;
} ASDM_CTX_CATCH(ctx)

int asdm_intersect_k(asdm_ctx *ctx, const uint32_t *const *lists, const uint32_t *sizes, uint32_t k, uint32_t *out,
                     uint32_t cap, uint32_t *n) try {
    if (!ctx || !lists || !sizes || !n || k == 0) return ASDM_E_INVALID;
    (void)hipSetDevice(ctx->device);
    std::vector<uint64_t> offs(k + 1ull, 0);
    for (uint32_t j = 0; j < k; ++j) offs[j + 1] = offs[j] + sizes[j];
    uint64_t minsz = UINT64_MAX;
    for (uint32_t j = 0; j < k; ++j) minsz = std::min<uint64_t>(minsz, sizes[j]);
    uint32_t *dflat = nullptr, *dout = nullptr;
    int rc = dev_alloc(ctx, (void **)&dflat, sizeof(uint32_t) * std::max<uint64_t>(offs[k], 1));
    if (!rc) rc = dev_alloc(ctx, (void **)&dout, sizeof(uint32_t) * std::max<uint64_t>(minsz, 1));
    for (uint32_t j = 0; j < k && !rc; ++j)
        if (sizes[j])
            rc = check_hip(ctx, hipMemcpyAsync(dflat + offs[j], lists[j], sizeof(uint32_t) * sizes[j], hipMemcpyHostToDevice,
                                               ctx->stream), "copy");
    // the shortest list first, then the others from the shortest up, one pairwise step each; the
    // running result alternates between two buffers of the shortest list's size
    std::vector<uint32_t> order(k);
    for (uint32_t j = 0; j < k; ++j) order[j] = j;
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return sizes[a] < sizes[b]; });
    uint32_t *buf[2] = {dout, nullptr};
    uint32_t *dcnt = nullptr;
    uint64_t *doff = nullptr;
    const uint64_t max_tiles = (minsz + offs[k]) / kIsectTile + 2;
    if (!rc && k > 1) rc = dev_alloc(ctx, (void **)&buf[1], sizeof(uint32_t) * std::max<uint64_t>(minsz, 1));
    if (!rc && k > 1) rc = dev_alloc(ctx, (void **)&dcnt, sizeof(uint32_t) * max_tiles);
    if (!rc && k > 1) rc = dev_alloc(ctx, (void **)&doff, sizeof(uint64_t) * (max_tiles + 1));
    uint64_t hn = minsz;
    const uint32_t *cur = dflat + offs[order[0]];
    int which = 0;
    for (uint32_t s = 1; s < k && !rc && hn; ++s) {
        const uint32_t j = order[s];
        const uint32_t *L = dflat + offs[j];
        const uint32_t nl = sizes[j], nr = (uint32_t)hn;
        const bool merge = (uint64_t)nl <= (uint64_t)kIsectBalanced * nr;
        const uint64_t tiles = merge ? ((uint64_t)nr + nl + kIsectTile - 1) / kIsectTile : (nr + kIsectTile - 1) / kIsectTile;
        uint32_t *dst = buf[which];
        for (int pass = 0; pass < 2 && !rc; ++pass) {
            if (merge)
                hipLaunchKernelGGL(k_isect_pair<true>, dim3((unsigned)std::max<uint64_t>(tiles, 1)), dim3(256), 0, ctx->stream,
                                   cur, nr, L, nl, pass, dcnt, doff, dst);
            else
                hipLaunchKernelGGL(k_isect_pair<false>, dim3((unsigned)std::max<uint64_t>(tiles, 1)), dim3(256), 0, ctx->stream,
                                   cur, nr, L, nl, pass, dcnt, doff, dst);
            rc = check_hip(ctx, hipGetLastError(), "intersect launch");
            if (!rc && pass == 0) rc = scan_u32_to_u64(ctx, dcnt, doff, std::max<uint64_t>(tiles, 1));
        }
        if (!rc) rc = check_hip(ctx, hipMemcpyAsync(&hn, doff + std::max<uint64_t>(tiles, 1), sizeof(uint64_t),
                                                    hipMemcpyDeviceToHost, ctx->stream), "copy");
        if (!rc) rc = check_hip(ctx, hipStreamSynchronize(ctx->stream), "sync");
        cur = dst;
        which ^= 1;
    }
    if (!rc && k == 1 && hn)  // one list: itself
        rc = check_hip(ctx, hipMemcpyAsync(dout, cur, sizeof(uint32_t) * hn, hipMemcpyDeviceToDevice, ctx->stream), "copy");
    if (!rc) rc = check_hip(ctx, hipStreamSynchronize(ctx->stream), "sync");
    if (!rc) {
        *n = (uint32_t)hn;
        if (out && hn > cap) rc = ctx->fail(ASDM_E_INVALID, "output capacity too small");
        else if (out && hn)
            rc = check_hip(ctx, hipMemcpy(out, k == 1 ? dout : cur, sizeof(uint32_t) * hn, hipMemcpyDeviceToHost), "copy");
    }
    dev_free(ctx, dflat);
    dev_free(ctx, dout);
    dev_free(ctx, buf[1]);
    dev_free(ctx, dcnt);
    dev_free(ctx, doff);
    (void)hipStreamSynchronize(ctx->stream);
    return rc;
} ASDM_CTX_CATCH(ctx)

// asdm_intersect_k with every device buffer in this device's static scratch: no allocation on
// the device or the host.  Calls that overlap on one device share the scratch and race.
int asdm_intersect_k_noalloc(asdm_ctx *ctx, const uint32_t *const *lists, const uint32_t *sizes, uint32_t k,
                             uint32_t *out, uint32_t cap, uint32_t *n) try {
    if (!ctx || !lists || !sizes || !n || k == 0) return ASDM_E_INVALID;
    // error messages are literals assigned into the context's reserved error string: no allocation
    if (k > kNoAllocMaxLists) {
        ctx->err.assign("more than 64 lists for the allocation-free scratch");
        return ASDM_E_BUDGET;
    }
    uint64_t offs[kNoAllocMaxLists + 1];
    offs[0] = 0;
    for (uint32_t j = 0; j < k; ++j) {
        if (sizes[j] && !lists[j]) return ASDM_E_INVALID;
        offs[j + 1] = offs[j] + sizes[j];
    }
    static_assert(kNoAllocListWords == 2097152, "the message below states the limit");
    if (offs[k] > kNoAllocListWords) {
        ctx->err.assign("lists hold more than 2097152 elements in total, the allocation-free scratch's limit");
        return ASDM_E_BUDGET;
    }
    uint8_t *base = ctx->noalloc_scratch;
    if (!base) {
        ctx->err.assign("context has no scratch address");
        return ASDM_E_INVALID;
    }
    (void)hipSetDevice(ctx->device);
    // HIP failures record a literal message (check_hip would build a string)
    auto hip_ok = [ctx](hipError_t e, const char *what) {
        if (e == hipSuccess) return (int)ASDM_OK;
        ctx->err.assign(what);
        return (int)(e == hipErrorOutOfMemory ? ASDM_E_OOM : ASDM_E_HIP);
    };
    uint32_t *dflat = (uint32_t *)(base + kNoAllocFlatOff);
    uint32_t *dout = (uint32_t *)(base + kNoAllocOutOff);
    uint64_t *doffs = (uint64_t *)(base + kNoAllocOffsOff);
    uint32_t *dn = (uint32_t *)(base + kNoAllocCountOff);
    int rc = ASDM_OK;
    for (uint32_t j = 0; j < k && !rc; ++j)
        if (sizes[j])
            rc = hip_ok(hipMemcpyAsync(dflat + offs[j], lists[j], sizeof(uint32_t) * sizes[j], hipMemcpyHostToDevice,
                                               ctx->stream), "copy");
    if (!rc) rc = hip_ok(hipMemcpyAsync(doffs, offs, sizeof(uint64_t) * (k + 1ull), hipMemcpyHostToDevice, ctx->stream), "copy");
    if (!rc) {
        // the kernel writes the result size unconditionally, so nothing needs clearing first
        hipLaunchKernelGGL(k_intersect_k, dim3(1), dim3(64), 0, ctx->stream, dflat, doffs, k, dout, dn);
        rc = hip_ok(hipGetLastError(), "intersect launch");
    }
    uint32_t hn = 0;
    if (!rc) rc = hip_ok(hipMemcpyAsync(&hn, dn, sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream), "copy");
    if (!rc) rc = hip_ok(hipStreamSynchronize(ctx->stream), "sync");
    if (rc) return rc;
    *n = hn;
    if (out && hn > cap) {
        ctx->err.assign("output capacity too small");
        return ASDM_E_INVALID;
    }
    if (out && hn) {
        rc = hip_ok(hipMemcpyAsync(out, dout, sizeof(uint32_t) * hn, hipMemcpyDeviceToHost, ctx->stream), "copy");
        if (!rc) rc = hip_ok(hipStreamSynchronize(ctx->stream), "sync");
    }
    return rc;
} ASDM_CTX_CATCH(ctx)

}  // extern "C"
