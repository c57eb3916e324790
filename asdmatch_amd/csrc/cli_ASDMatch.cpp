// ASDMatch.out -- drop-in for the reference's CLI (main.cpp:11-147) on one MI355X.
//   ASDMatch.out -q <query> -d <data> [-s <card>] [-m <KB>] [-r <out.txt>] [-c <local.csv>]
// Prints the reference's lines in its order; "Number of matches" is the exact count (mod 2^64).
// -m bounds the per-level frontier memory (KB, default 16e5 as the reference's budget).
// -c writes the |Vg| x |Vq| local-count CSV (the reference needs an ALL_LEVEL + LOCAL_COUNT
// build for it; here it is a run-time option).  Nothing is printed when a candidate set is empty.
#include <chrono>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <memory>
#include <vector>

#include "asdm.h"
#include "cli_common.h"

static double secs(std::chrono::steady_clock::time_point a) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - a).count();
}

static int die(asdm_ctx *ctx, int rc, const char *what) {
    std::fprintf(stderr, "ASDMatch.out: %s failed (%d): %s\n", what, rc, ctx ? asdm_last_error(ctx) : "");
    return 1;
}

int main(int argc, char **argv) {
    Args a(argc, argv);
    const std::string qp = a.str("-q"), dp = a.str("-d"), rp = a.str("-r"), sp = a.str("-s"), cp = a.str("-c");
    int memKB = a.num("-m");
    if (memKB < 0) {
        std::fprintf(stderr, "ASDMatch.out: -m takes a positive size in KB\n");
        return 1;
    }
    if (memKB == 0) memKB = 1600000;
    const uint64_t budget = (uint64_t)1024 * (uint64_t)memKB;

    asdm_graph *q = nullptr, *g = nullptr;
    if (asdm_graph_load_text(qp.c_str(), &q) != ASDM_OK) {
        std::cout << "Can not open file " << qp << " ." << std::endl;
        return 255;
    }
    if (asdm_graph_load_text(dp.c_str(), &g) != ASDM_OK) {
        std::cout << "Can not open file " << dp << " ." << std::endl;
        return 255;
    }
    std::ofstream file;
    if (!rp.empty()) file.open(rp);
    std::ostream &out = rp.empty() ? std::cout : file;

    asdm_ctx *ctx = nullptr;
    int rc = asdm_ctx_create(0, &ctx);
    if (rc != ASDM_OK) return die(nullptr, rc, "device context (a gfx950 GPU is required)");
    if ((rc = asdm_graph_upload(ctx, g)) != ASDM_OK) return die(ctx, rc, "graph upload");

    auto t0 = std::chrono::steady_clock::now();
    asdm_cs *cs = nullptr;
    rc = asdm_filter_cfl(ctx, q, g, &cs);
    if (rc == ASDM_E_EMPTY_CANDIDATES) return 0;  // the reference prints nothing
    if (rc != ASDM_OK) return die(ctx, rc, "filter");
    out << "Filtering time (s): " << secs(t0) << std::endl;

    uint32_t nq = 0, nv = 0;
    uint64_t m = 0;
    asdm_graph_info(q, &nq, &m, nullptr);
    asdm_graph_info(g, &nv, &m, nullptr);
    const bool local = !cp.empty();

    t0 = std::chrono::steady_clock::now();
    std::vector<uint64_t> masks;
    std::vector<double> cards;
    double dp_time = 0;
    bool have_cards = false;
    if (!sp.empty()) {
        uint32_t n = 0;
        if (asdm_card_read(sp.c_str(), nullptr, nullptr, 0, &n, &dp_time) == ASDM_OK) {
            masks.resize(n);
            cards.resize(n);
            have_cards = asdm_card_read(sp.c_str(), masks.data(), cards.data(), n, &n, &dp_time) == ASDM_OK;
        }
    }
    out << "FHD time (s): " << 0.0 << std::endl;  // no hypertree decomposition in this planner
    if (!sp.empty()) out << "dp time: " << dp_time << std::endl;
    asdm_plan *plan = nullptr;
    const uint32_t flags = local ? ASDM_PLAN_LOCAL : 0;
    rc = have_cards ? asdm_plan_create_with_cards(q, cs, masks.data(), cards.data(), (uint32_t)masks.size(), flags, &plan)
                    : asdm_plan_create(q, cs, nullptr, flags, &plan);
    if (rc != ASDM_OK) return die(ctx, rc, "plan");
    double est_leaf = 0, est_total = 0;
    asdm_plan_estimate(plan, &est_leaf, &est_total);
    out << "Planning Time: " << secs(t0) << std::endl;
    out << "cost: " << est_total << std::endl;

    t0 = std::chrono::steady_clock::now();
    uint64_t count = 0;
    std::vector<uint64_t> mat;
    if (local) {
        mat.assign((size_t)nv * nq, 0);
        rc = asdm_local_count(ctx, plan, cs, budget, mat.data(), &count);
    } else {
        rc = asdm_count(ctx, plan, cs, budget, &count);
    }
    if (rc != ASDM_OK) return die(ctx, rc, "count");
    asdm_stats st{};
    asdm_last_stats(ctx, &st);
    out << "Execution Time: " << secs(t0) << std::endl;
    out << "Number of matches: " << count << std::endl;
    out << "Number of fixed attributes: " << 0 << std::endl;
    out << "Number of intersections: " << st.rows_total << std::endl;
    out << "Number of fix cases: " << 0 << std::endl;
    if (local) {
        std::ofstream csv(cp);
        std::vector<char> buf(1 << 20);
        csv.rdbuf()->pubsetbuf(buf.data(), buf.size());
        for (uint32_t v = 0; v < nv; ++v) {
            for (uint32_t u = 0; u < nq; ++u) {
                csv << mat[(size_t)v * nq + u];
                if (u + 1 != nq) csv << ",";
            }
            csv << "\n";
        }
    }
    asdm_plan_free(plan);
    asdm_cs_free(cs);
    asdm_ctx_destroy(ctx);
    asdm_graph_free(q);
    asdm_graph_free(g);
    return 0;
}
