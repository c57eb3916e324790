// Multi-query entry point.  Every query runs filter -> plan -> count on the context stream against
// the resident data graph; latencies are host wall-clock completion times from the call start.
#include <hip/hip_runtime.h>

#include <chrono>

#include "engine.h"

extern "C" {

int asdm_batch_count(asdm_ctx *ctx, const asdm_graph *g, const asdm_graph *const *qs, uint32_t nq, uint64_t budget_bytes,
                     uint64_t *counts, float *lat_ms) {
    if (!ctx || !g || (!qs && nq) || (!counts && nq)) return ASDM_E_INVALID;
    if (g->dev.device != ctx->device) return ctx->fail(ASDM_E_INVALID, "data graph not uploaded to this device");
    auto t0 = std::chrono::steady_clock::now();
    for (uint32_t i = 0; i < nq; ++i) {
        asdm_cs *cs = nullptr;
        int rc = asdm_filter_cfl(ctx, qs[i], g, &cs);
        uint64_t c = 0;
        if (rc == ASDM_OK) {
            asdm_plan *p = nullptr;
            rc = asdm_plan_create(qs[i], cs, nullptr, 0, &p);
            if (rc == ASDM_OK) rc = asdm_count(ctx, p, cs, budget_bytes, &c);
            asdm_plan_free(p);
            asdm_cs_free(cs);
        } else if (rc == ASDM_E_EMPTY_CANDIDATES) {
            rc = ASDM_OK;
        }
        if (rc != ASDM_OK) return rc;
        counts[i] = c;
        if (lat_ms)
            lat_ms[i] = (float)std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    return ASDM_OK;
}

}  // extern "C"
