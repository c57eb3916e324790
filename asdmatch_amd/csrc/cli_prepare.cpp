// prepare.out -- drop-in for executable/prepare.cpp: filter + plan, then write the planner's
// subset cardinalities to a .card file (-s).  ASDMatch.out -s reads it back (and also reads the
// map of .card files written by the reference's prepare.out).
//   prepare.out -q <query> -d <data> -s <card>
#include <chrono>
#include <cstdio>
#include <iostream>
#include <vector>

#include "asdm.h"
#include "cli_common.h"

int main(int argc, char **argv) {
    Args a(argc, argv);
    const std::string qp = a.str("-q"), dp = a.str("-d"), sp = a.str("-s");
    asdm_graph *q = nullptr, *g = nullptr;
    if (asdm_graph_load_text(qp.c_str(), &q) != ASDM_OK) {
        std::cout << "Can not open file " << qp << " ." << std::endl;
        return 255;
    }
    if (asdm_graph_load_text(dp.c_str(), &g) != ASDM_OK) {
        std::cout << "Can not open file " << dp << " ." << std::endl;
        return 255;
    }
    asdm_ctx *ctx = nullptr;
    int rc = asdm_ctx_create(0, &ctx);
    if (rc != ASDM_OK) {
        std::fprintf(stderr, "prepare.out: no usable gfx950 device (%d)\n", rc);
        return 1;
    }
    asdm_graph_upload(ctx, g);
    auto t0 = std::chrono::steady_clock::now();
    asdm_cs *cs = nullptr;
    rc = asdm_filter_cfl(ctx, q, g, &cs);
    if (rc == ASDM_E_EMPTY_CANDIDATES) return 0;
    if (rc != ASDM_OK) {
        std::fprintf(stderr, "prepare.out: filter failed: %s\n", asdm_last_error(ctx));
        return 1;
    }
    asdm_plan *plan = nullptr;
    if ((rc = asdm_plan_create(q, cs, nullptr, 0, &plan)) != ASDM_OK) {
        std::fprintf(stderr, "prepare.out: plan failed: %s\n", asdm_last_error(ctx));
        return 1;
    }
    uint32_t n = 0;
    asdm_plan_cards(plan, nullptr, nullptr, 0, &n);
    std::vector<uint64_t> masks(n);
    std::vector<double> cards(n);
    asdm_plan_cards(plan, masks.data(), cards.data(), n, &n);
    const double t = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    rc = asdm_card_write(sp.c_str(), masks.data(), cards.data(), n, t);
    asdm_plan_free(plan);
    asdm_cs_free(cs);
    asdm_ctx_destroy(ctx);
    return rc == ASDM_OK ? 0 : 1;
}
