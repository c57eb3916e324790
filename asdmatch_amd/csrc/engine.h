// Internal structures shared by the HIP translation units and the host planner.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <cstdint>
#include <functional>
#include <memory>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include "bags.h"
#include "host.h"

namespace asdm {

constexpr int kNumCounters = 32;  // device counters per context

constexpr uint32_t kMaxQ = 32;      // MAX_PATTERN_SIZE (config/config.h:31)
constexpr uint32_t kMaxGroup = 4;   // largest same-label leaf group counted by inclusion-exclusion
constexpr uint32_t kMaxParts = 15;  // Bell(4)

// Device plan: everything the enumeration kernels need about the matching order.  Lives in
// device memory; kernels read it with wave-uniform indices.
struct DPlan {
    uint32_t nq;                     // query vertices
    uint32_t D;                      // enumerated depths; depths D..nq-1 are leaves
    uint32_t u[kMaxQ];               // query vertex at each depth
    uint64_t coff[kMaxQ];            // offset of C(u_d) in the concatenated candidate array
    uint32_t csz[kMaxQ];             // |C(u_d)|
    uint32_t nb[kMaxQ];              // backward neighbours of depth d (all at depths < min(d, D))
    uint32_t bdep[kMaxQ][kMaxQ];     // their depths
    uint64_t bbase[kMaxQ][kMaxQ];    // candidate-edge offset base of arc (u_bdep -> u_d)
    uint32_t ns[kMaxQ];              // earlier enumerated depths with the same label as depth d
    uint32_t sdep[kMaxQ][kMaxQ];
    // leaf groups: leaves with equal labels, counted jointly by inclusion-exclusion
    uint32_t ngroups;
    uint32_t nadj_last;              // leaves with a matched neighbour at depth D-1
    // Triangle leaves: leaf e with exactly two matched neighbours (list indices 0 and 1) that are
    // adjacent in the query.  tri_tab[e][k] = |list0 ∩ list1| for the k-th candidate edge of the
    // arc (u_bdep[e][0] -> u_bdep[e][1]), whose offsets start at tri_base[e].
    uint32_t tri[kMaxQ];
    uint64_t tri_base[kMaxQ];
    const uint32_t *tri_tab[kMaxQ];
    uint32_t fused;                  // depth D-1 evaluated by the fused leaf path (batch engine)
    // Row-sum table of a fused level whose depth D-1 has a single backward neighbour, depth
    // tail_dep, and whose leaves put no same-label constraint on its vertex: tail[p] = sum over
    // the candidates x of D-1 adjacent to position p of C(u_tail_dep) of the product of the
    // adjacent leaves' list counts at x.  nullptr when the level has no such table.
    const uint64_t *tail;
    uint32_t tail_dep;
    // One correction factor of the table: tail_fix 1 = the adjacent leaf at depth tail_leaf has
    // same-label matched vertices (its count at x drops by one for each adjacent to x; tail_rbase:
    // list offsets of the arc (u_tail_leaf -> u_{D-1})); 2 = the hoisted leaf at depth tail_leaf
    // has the label of depth D-1 (its count drops by one at the x in its list); 0 = none.
    uint32_t tail_fix, tail_leaf;
    uint64_t tail_rbase;
    // Second table, one level up: when the table's backward neighbour is depth D-2, which has a
    // single backward neighbour tail2_dep, tail2[p] = sum over the candidates x' of D-2 adjacent
    // to position p of C(u_tail2_dep) of tail[x'], and rows of width D-2 are finished from it.
    const uint64_t *tail2;
    uint32_t tail2_dep;
    // Edge table of a fused level whose depth D-1 has two backward neighbours b1 < b2 that are
    // adjacent in the query: etail[k] = sum over x in L(b1 -> D-1) ∩ L(b2 -> D-1) of the product of
    // the adjacent leaves' list counts at x, for the k-th candidate edge (p1, p2) of the arc
    // (u_b1 -> u_b2).  nullptr when the level has none.
    const uint64_t *etail;
    uint32_t gsize[kMaxQ];
    uint32_t gleaf[kMaxQ][kMaxGroup];  // depth of each leaf in the group
    // set partitions of {0..k-1} for k = 2..4: blocks as leaf-index masks, Moebius coefficient
    uint32_t nparts[kMaxGroup + 1];
    uint32_t pnblk[kMaxGroup + 1][kMaxParts];
    uint32_t pblk[kMaxGroup + 1][kMaxParts][kMaxGroup];
    int64_t pcoef[kMaxGroup + 1][kMaxParts];
    // candidate-set membership bitmaps of the query's vertices ([nq][W] words, row = query vertex)
    // and their per-word rank prefixes: data id -> position in C(u) in two loads
    const uint64_t *bits;
    const uint32_t *rank;
    uint32_t W;
};

}  // namespace asdm

struct asdm_ctx {
    int device = 0;
    uint32_t num_cus = 256;  // compute units of the device
    hipStream_t stream = nullptr;
    std::string err;
    asdm_stats stats{};
    bool profiling = false;
    double time_limit_ms = 0;  // 0 = no limit
    uint64_t row_limit = 0;    // 0 = no limit
    // Batch-engine launch shape (asdm_ctx_set_batch_tuning / asdm_batch_autotune): lanes per
    // leaf-depth row, the factor on the list-length thresholds of the lane classes (3, 12, 48
    // expected entries for 4, 16, 64 lanes), and rows per lane group in a wave tile.
    uint32_t tune_leaf_lanes = 16;
    double tune_class_scale = 1.0;
    uint32_t tune_tile_iters = 4;
    // per-kernel device time (ms), launches and algorithmic bytes, accumulated while profiling
    struct KStat {
        std::string name;
        double ms = 0;
        uint64_t launches = 0;
        uint64_t bytes = 0;
    };
    std::vector<KStat> kstats;
    struct PendingTimer {
        int slot;
        hipEvent_t a, b;
        uint32_t launches;  // sections the pair completes (0 for all but a marker pair)
    };
    std::vector<PendingTimer> pending;  // recorded, resolved after the stream has drained
    std::vector<hipEvent_t> event_pool;  // timer events reused across launches
    // device address of this device's allocation-free scratch (asdm_intersect_k_noalloc), looked
    // up once per device when a context is created
    uint8_t *noalloc_scratch = nullptr;
    // Worker contexts of the same device (own stream, allocator and counters) that count the
    // bag-engine queries of a batched call concurrently, created on first use; bag_workers of
    // them (1: the calling context counts them one after the other).  mem_share is the share of
    // the device memory this context's caches may hold (1 / workers for a worker).
    uint32_t bag_workers = 6;
    // Form of the bag engine's relation products: 0 chosen per product from its term count, 1 the
    // LDS row accumulation, 2 the zero-fill + atomic scatter (for tests and tuning).
    int table_form = 0;
    std::vector<asdm_ctx *> workers;
    double mem_share = 1.0;
    uint64_t *d_counters = nullptr;  // [kNumCounters] device counters: 0 count, then algorithmic
                                     // bytes per kernel (single engine: 1..3; batch engine: 1..16)
    // caching allocator: freed blocks are kept for reuse (all work is ordered on `stream`, so a
    // block freed by one launch sequence may be handed to the next one immediately)
    struct Block {
        void *p;
        size_t bytes;
    };
    std::vector<Block> cache;
    std::vector<Block> live;
    size_t cached_bytes = 0;
    uint64_t trims = 0;  // partial trims of the cache (each waits for the device)
    int fail(int code, const std::string &msg) {
        err = msg;
        return code;
    }
};

namespace asdm {

// Records code and message as the context's last error where that is possible, for a handler
// that must not throw.
inline int ctx_failure(asdm_ctx *c, int code, const char *msg) noexcept {
    try {
        return c ? c->fail(code, msg) : code;
    } catch (...) {
        return code;
    }
}

}  // namespace asdm

// Handler of the function-try-block of an entry point with a context `c`: as ASDM_HOST_CATCH, and
// the failure is also the context's last error.
#define ASDM_CTX_CATCH(c)                                                                  \
    catch (const std::bad_alloc &) {                                                      \
        return asdm::ctx_failure((c), ASDM_E_OOM, "host allocation failed");             \
    }                                                                                     \
    catch (const std::exception &e) {                                                     \
        return asdm::ctx_failure((c), ASDM_E_INVALID, e.what());                          \
    }                                                                                     \
    catch (...) {                                                                         \
        return asdm::ctx_failure((c), ASDM_E_INVALID, "unexpected host exception");      \
    }

struct asdm_cs {
    // Device buffers shared by the candidate spaces of one filter call (a batch of queries); the
    // last view freed releases them.
    struct Store {
        uint32_t refs = 0;
        uint32_t *d_cand = nullptr;
        uint64_t *d_bits = nullptr;
        uint32_t *d_rank = nullptr;
        uint64_t *d_ce_off = nullptr;
        uint32_t *d_ce_val = nullptr;
    };
    Store *store = nullptr;
    // A candidate space extended with cartesian arcs (cartesian.hip) owns new candidate-edge
    // arrays in `store` and keeps the store of the space it extends, whose candidate sets and
    // bitmaps it shares, in base_store.
    Store *base_store = nullptr;
    asdm_ctx *ctx = nullptr;
    uint32_t nq = 0, nv = 0, root = 0;
    bool loop_free = false;                  // the data graph has no self-loops
    std::vector<uint32_t> qlab, qoff, qnbr;  // query CSR copy (label-sorted)
    std::vector<uint32_t> csize;             // |C(u)|
    std::vector<uint64_t> cand_off;          // [nq + 1] offsets into d_cand
    uint32_t *d_cand = nullptr;              // concatenated ascending candidate lists
    uint64_t *d_bits = nullptr;              // [nq][W] membership bitmaps of C(u)
    uint32_t *d_rank = nullptr;              // [nq][W] exclusive popcount prefix per 64-bit word
    uint32_t W = 0;                          // 64-bit words per bitmap
    uint32_t narcs = 0;
    std::vector<uint32_t> arc_u, arc_u2;     // query arcs in query CSR order
    std::vector<uint64_t> ce_base;           // [narcs + 1] first index of arc a in d_ce_off
    std::vector<uint64_t> arc_weight;        // total candidate-edge list length per arc
    uint64_t *d_ce_off = nullptr;            // [ce_base[narcs] + 1] list offsets into d_ce_val (global
                                             // offsets of the shared store)
    uint32_t *d_ce_val = nullptr;            // candidate-edge lists: positions in C(u2), ascending
    uint64_t ce_total = 0;
    double filter_ms = 0;
    uint32_t arc_index(uint32_t u, uint32_t u2) const {
        for (uint32_t a = qoff[u]; a < qoff[u + 1]; ++a)
            if (qnbr[a] == u2) return a;
        return UINT32_MAX;
    }
};

struct asdm_plan {
    asdm::DPlan h{};                 // host copy
    std::vector<uint32_t> order;
    uint32_t D = 0;
    bool local = false;
    double plan_ms = 0;
    double est_leaf = 0, est_total = 0;  // estimated rows at depth D and summed over depths 1..D
    std::vector<std::pair<uint64_t, double>> prefix_cards;  // (prefix subset mask, estimate)
    // Set when the query is counted by the bag-table engine (bags.h) instead of the frontier
    // engine: partition terms and their elimination orders.
    std::shared_ptr<asdm::BagPlan> bags;
    // Set for an explicit order with cartesian steps: the candidate space the plan was made on,
    // the caller's extended with one arc per cartesian step (cartesian.hip); the engine counts on
    // it in place of the caller's.
    std::shared_ptr<asdm_cs> cart_cs;
};

namespace asdm {

// Device helpers implemented in ctx.hip
int check_hip(asdm_ctx *ctx, hipError_t e, const char *what);
int dev_alloc(asdm_ctx *ctx, void **p, size_t bytes);
void dev_free(asdm_ctx *ctx, void *p);
// out[0..n] = exclusive prefix of in[0..n) with out[n] = total.
int scan_u32_to_u64(asdm_ctx *ctx, const uint32_t *in, uint64_t *out, uint64_t n);
int fill_u64(asdm_ctx *ctx, uint64_t *p, uint64_t n, uint64_t v);
int fill_u32(asdm_ctx *ctx, uint32_t *p, uint64_t n, uint32_t v);

// Batched CFL filter + candidate edges (filter.hip)
int filter_many(asdm_ctx *ctx, const asdm_graph *g, const asdm_graph *const *qs, uint32_t nb, asdm_cs **out,
                int *status);

// mat[key] += value for n (key, value) pairs, keys < cells, with one writer per cell
// (owner-partitioned buckets, local_agg.hip).
int aggregate_pairs(asdm_ctx *ctx, const uint64_t *keys, const uint64_t *vals, uint64_t n, uint64_t cells,
                    uint64_t *mat);

// Planner (plan.cpp).  card_of(mask, fallback) returns the cardinality estimate of the
// sub-query induced by `mask` (the subsetToCard role), or `fallback` when it has none.
using CardFn = std::function<double(uint64_t mask, double fallback)>;
int build_plan(const asdm_graph *q, const asdm_cs *cs, const uint32_t *order, uint32_t flags, const CardFn &card_of,
               asdm_plan *out, std::string &err);
// Pieces of the automatic plan, shared with the learned planner: the leaf set, the independence
// model's growth factor, the vertices of `rest` outside S adjacent to S, and the device tables of
// a chosen order whose first D vertices are enumerated.
uint64_t choose_leaves(const asdm_cs *cs, uint32_t flags);
double growth(const asdm_cs *cs, uint64_t S, uint32_t u);
uint64_t frontier_of(const asdm_cs *cs, uint64_t S, uint64_t rest);
uint32_t smallest_candidates(const asdm_cs *cs, uint64_t rest);
int finish_plan(const asdm_cs *cs, std::vector<uint32_t> order, uint32_t D, uint32_t flags, const CardFn &card_of,
                asdm_plan *out, std::string &err);
// Cost-based order search shared by the plain (ASDM_PLAN_COST) and learned planners (learned.cpp):
// est(requests, out) returns the estimated cardinality of each (query index, vertex subset).
using BatchEstimator =
    std::function<int(const std::vector<std::pair<uint32_t, uint64_t>> &reqs, std::vector<double> &out)>;
int cost_orders(const std::vector<const asdm_cs *> &css, uint32_t flags, bool multi_start, const BatchEstimator &est,
                std::vector<std::vector<uint32_t>> &orders, std::vector<uint32_t> &Ds,
                std::vector<std::unordered_map<uint64_t, double>> &cards);
// The independence model's cardinality of the sub-query induced by mask.
double indep_card(const asdm_cs *cs, uint64_t mask);
// Plans of a batch of candidate spaces from the learned estimator (learned.cpp).
int learned_plans(asdm_mpnn *m, const asdm_graph *g, const std::vector<const asdm_cs *> &css, uint32_t flags,
                  std::vector<asdm_plan *> &out);
// CFL root choice and BFS order (plan.cpp): base_count[u] = |{v : label, degree, NLF ok}|.
uint32_t cfl_root(const asdm_graph *q, const asdm_graph *g, const std::vector<uint32_t> &base_count);
std::vector<uint32_t> bfs_order(const asdm_graph *q, uint32_t root);

// For an order whose depth d >= 1 has no earlier neighbour (a cartesian step), a copy of cs with one
// more arc per such depth, from its cartesian parent (the earlier vertex nearest under the query
// paths' products, as the reference chooses it) to it, whose candidate-edge lists are the
// cartesian candidates of the path between them (cartesian.hip).  *out stays null when the order
// has no cartesian step.
int cartesian_cs(const asdm_cs *cs, const std::vector<uint32_t> &order, asdm_cs **out, std::string &err);
// Attaches a bag-table plan to p when flags ask for one, or by the automatic rule (plan.cpp).
int attach_bags(const asdm_cs *cs, uint32_t flags, bool explicit_order, asdm_plan *p, std::string &err);
// Bag-table engine (bags.hip): count (mat_host == NULL) or local counts of a plan with bags.
int run_bags(asdm_ctx *ctx, const asdm_plan *plan, const asdm_cs *cs, uint64_t *mat_host, uint64_t *count);

// Kernel timing helpers (ctx.hip)
// A timing event from the context's pool (created without a system-scope fence).
hipEvent_t take_timer_event(asdm_ctx *ctx);
// Fills the context's pool of timing events (profiling is about to start).
void reserve_timer_events(asdm_ctx *ctx);
// A timed section of launches, while the context is profiling.  By default the section is bracketed
// by two event markers on the stream.  With kDispatch, the kernels launched through launch() carry
// their own start and end events in their dispatches (no marker packets, so no pause between
// kernels), and the section's time is the sum of those kernels' durations.
struct KTimer {
    enum Mode { kMarkers, kDispatch };
    asdm_ctx *ctx;
    int slot;
    Mode mode;
    hipEvent_t a = nullptr, b = nullptr;
    KTimer(asdm_ctx *c, const char *name, Mode m = kMarkers);
    template <typename F, typename... Args>
    void launch(F kernel, const dim3 &grid, const dim3 &block, uint32_t shmem, Args... args) {
        if (mode != kDispatch || slot < 0) {
            hipLaunchKernelGGL(kernel, grid, block, shmem, ctx->stream, args...);
            return;
        }
        hipEvent_t e0 = take_timer_event(ctx), e1 = take_timer_event(ctx);
        hipExtLaunchKernelGGL(kernel, grid, block, shmem, ctx->stream, e0, e1, 0, args...);
        ctx->pending.push_back({slot, e0, e1, 0});
    }
    void stop(uint64_t bytes = 0);
};
// Resolves the recorded kernel timers (call after the stream has been synchronized).
void resolve_timers(asdm_ctx *ctx);
// Adds algorithmic bytes to the named kernel's entry (created if missing) while profiling.
void add_kernel_bytes(asdm_ctx *ctx, const char *name, uint64_t bytes);
// Adds count to an entry's launches and bytes to its bytes (entries that count events, not kernels).
void add_kernel_count(asdm_ctx *ctx, const char *name, uint64_t count, uint64_t bytes);
// Adds host wall time of a phase (entries named "host: ...", no bytes) while profiling.
void add_host_ms(asdm_ctx *ctx, const char *name, double ms);
// Adds the kernel table of `from` (timers resolved) into that of `into` and clears it.
void merge_kernel_stats(asdm_ctx *into, asdm_ctx *from);
// Frees the context's cached device blocks (after its stream has drained).
void release_cached_blocks(asdm_ctx *ctx);

// Concurrent bag-engine counting for the batched entry points (bag_pool.cpp).  Jobs are counted
// on the context's worker contexts, the largest estimated cost first; a job the workers cannot
// fit in memory is counted again on the calling context alone once they are done.
class BagPool {
  public:
    struct Job {
        size_t k;                 // index into the result arrays
        const asdm_plan *plan;    // a plan with bags
        const asdm_cs *cs;
        double cost;              // estimated element operations (BagPlan::cost)
    };
    // counts / status / lat are indexed by Job::k; per_query says whether a failure of one query
    // (ASDM_E_TIMEOUT, ASDM_E_UNSUPPORTED) is recorded in status rather than failing the call.
    BagPool(asdm_ctx *ctx, uint64_t budget, std::function<float()> clock, uint64_t *counts, int *status, float *lat,
            bool per_query);
    ~BagPool();
    int start();                  // spawns the worker threads (none with one worker)
    void push(const Job &j);      // thread-safe; before finish()
    int finish();                 // no more jobs: joins, retries, merges the kernel tables

  private:
    struct State;
    std::unique_ptr<State> s_;
};

}  // namespace asdm
