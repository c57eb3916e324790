// Concurrent bag-engine counting for the batched entry points.
//
// A bag-engine query is a chain of dependent launches (domains, relations, then one table or
// vector step per eliminated block) with a few host synchronizations, and many of its launches
// are too small to fill the GPU.  Queries are independent, so the batch hands them to worker
// contexts of the same device, each with its own stream, allocator and counters, driven by one
// host thread each: while one query waits on a synchronization or runs a small launch, the
// others' launches fill the device.  The shared work queue is ordered by the bag plan's estimated
// cost, largest first, so the longest queries do not start last.  The candidate spaces and plans
// are read-only here; every device buffer a worker allocates comes from and returns to its own
// allocator, whose blocks are only reused on its own stream.
#include <algorithm>
#include <condition_variable>
#include <mutex>
#include <queue>
#include <thread>

#include "engine.h"

namespace asdm {

namespace {
bool per_query_code(int code) { return code == ASDM_E_TIMEOUT || code == ASDM_E_UNSUPPORTED; }

// The workers together may cache (and keep factor results in) kWorkerShare times what one
// context alone may: a query's working set is a small part of its context's limits.
constexpr double kWorkerShare = 2.0;
// Cached bytes the calling context may keep when the workers start (more are returned to the device).
constexpr size_t kParentCacheKeep = 4ull << 30;

struct ByCost {
    bool operator()(const BagPool::Job &a, const BagPool::Job &b) const { return a.cost < b.cost; }
};
}  // namespace

struct BagPool::State {
    asdm_ctx *ctx;
    uint64_t budget;
    std::function<float()> clock;
    uint64_t *counts;
    int *status;
    float *lat;
    bool per_query;
    std::mutex mu;
    std::condition_variable cv;
    std::priority_queue<Job, std::vector<Job>, ByCost> q;
    std::vector<Job> retry;        // jobs a worker ran out of device memory on
    std::vector<Job> serial;       // jobs pushed when there are no worker threads
    bool closed = false;
    int error = ASDM_OK;           // first failure that fails the call
    std::string error_msg;
    std::vector<std::thread> threads;
    std::vector<uint64_t> trims0;  // workers' trim counts at the start

    void run_one(asdm_ctx *w, const Job &j) {
        uint64_t c = 0;
        const int rc = asdm_count(w, j.plan, j.cs, budget, &c);
        const float t = clock();
        std::lock_guard<std::mutex> lk(mu);
        if (rc == ASDM_E_OOM && w != ctx) {
            retry.push_back(j);
            return;
        }
        counts[j.k] = rc == ASDM_OK ? c : 0;
        status[j.k] = rc;
        lat[j.k] = t;
        if (rc != ASDM_OK && !(per_query && per_query_code(rc)) && error == ASDM_OK) {
            error = rc;
            error_msg = w->err;
        }
    }

    void worker(asdm_ctx *w) {
        (void)hipSetDevice(w->device);
        for (;;) {
            Job j;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv.wait(lk, [&] { return closed || !q.empty() || error != ASDM_OK; });
                if (error != ASDM_OK || q.empty()) return;  // failed call, or closed and drained
                j = q.top();
                q.pop();
            }
            try {
                run_one(w, j);
            } catch (...) {  // host allocation failures inside the bookkeeping
                std::lock_guard<std::mutex> lk(mu);
                if (error == ASDM_OK) {
                    error = ASDM_E_OOM;
                    error_msg = "host allocation failed in a bag worker";
                }
            }
            cv.notify_all();
        }
    }
};

BagPool::BagPool(asdm_ctx *ctx, uint64_t budget, std::function<float()> clock, uint64_t *counts, int *status, float *lat,
                 bool per_query)
    : s_(new State{ctx, budget, std::move(clock), counts, status, lat, per_query}) {}

BagPool::~BagPool() {
    if (!s_->threads.empty()) {
        {
            std::lock_guard<std::mutex> lk(s_->mu);
            s_->closed = true;
            if (s_->error == ASDM_OK) s_->error = ASDM_E_INVALID;  // abandoned: stop the workers
        }
        s_->cv.notify_all();
        for (auto &t : s_->threads) t.join();
    }
}

int BagPool::start() {
    asdm_ctx *ctx = s_->ctx;
    const uint32_t nw = std::max<uint32_t>(1, ctx->bag_workers);
    if (nw == 1) return ASDM_OK;
    while (ctx->workers.size() < nw) {
        asdm_ctx *w = nullptr;
        if (int rc = asdm_ctx_create(ctx->device, &w)) return ctx->fail(rc, "bag worker context creation failed");
        ctx->workers.push_back(w);
    }
    // the candidate spaces and plans were made on the calling context's stream: they must be
    // complete before another stream reads them
    if (int rc = check_hip(ctx, hipStreamSynchronize(ctx->stream), "sync before bag workers")) return rc;
    // blocks the calling context keeps cached from earlier calls (queries counted one at a time, say)
    // sit idle while the workers run, yet count as used memory when each worker sizes its factor
    // cache from free memory: after the plain per-query path over config 4's batch, the batched
    // call's workers reused 2 % fewer results and spent 8 % more kernel time.  The device is idle
    // here, so freeing them costs little.
    if (ctx->cached_bytes > kParentCacheKeep) release_cached_blocks(ctx);
    for (uint32_t i = 0; i < nw; ++i) {
        asdm_ctx *w = ctx->workers[i];
        w->time_limit_ms = ctx->time_limit_ms;
        w->row_limit = ctx->row_limit;
        w->table_form = ctx->table_form;
        w->mem_share = kWorkerShare / nw;
        s_->trims0.push_back(w->trims);
        if (w->profiling != ctx->profiling) {
            resolve_timers(w);
            w->kstats.clear();
            w->profiling = ctx->profiling;
            if (w->profiling) reserve_timer_events(w);
        }
    }
    for (uint32_t i = 0; i < nw; ++i) s_->threads.emplace_back([this, w = ctx->workers[i]] { s_->worker(w); });
    return ASDM_OK;
}

void BagPool::push(const Job &j) {
    {
        std::lock_guard<std::mutex> lk(s_->mu);
        if (s_->threads.empty()) {
            s_->serial.push_back(j);
            return;
        }
        s_->q.push(j);
    }
    s_->cv.notify_one();
}

int BagPool::finish() {
    asdm_ctx *ctx = s_->ctx;
    {
        std::lock_guard<std::mutex> lk(s_->mu);
        s_->closed = true;
    }
    s_->cv.notify_all();
    for (auto &t : s_->threads) t.join();
    const bool threaded = !s_->threads.empty();
    s_->threads.clear();
    if (threaded) {
        const uint32_t nw = std::max<uint32_t>(1, ctx->bag_workers);
        for (uint32_t i = 0; i < nw && i < ctx->workers.size(); ++i) {
            asdm_ctx *w = ctx->workers[i];
            if (ctx->profiling) {
                merge_kernel_stats(ctx, w);
                add_kernel_count(ctx, "device_cache_trims", w->trims - s_->trims0[i], 0);
            }
            if (!s_->retry.empty()) {
                (void)hipStreamSynchronize(w->stream);
                release_cached_blocks(w);  // make room for the retries
            }
        }
    }
    if (s_->error != ASDM_OK) {
        ctx->err = s_->error_msg;
        return s_->error;
    }
    // jobs the workers could not fit next to each other, then those of a pool without workers,
    // one at a time on the calling context
    std::vector<Job> rest = s_->retry;
    rest.insert(rest.end(), s_->serial.begin(), s_->serial.end());
    std::stable_sort(rest.begin(), rest.end(), [](const Job &a, const Job &b) { return a.cost > b.cost; });
    for (const Job &j : rest) {
        s_->run_one(ctx, j);
        if (s_->error != ASDM_OK) return s_->error;
    }
    return ASDM_OK;
}

}  // namespace asdm
