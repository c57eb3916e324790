// Test-infrastructure driver (not product code).  It links the reference's own libraries
// (relation/, plan/, matching/, utility/, utility/td + GLPK, all compiled from the reference's
// sources by oracle/ref.mk) and calls them in the order main.cpp:11-147 does, with one change:
// the per-bag trie roots are allocated AFTER reorderBags() rather than before it.
//
// Why: main.cpp:58 sizes `nodes` with t.numNodes, and main.cpp:72 then calls reorderBags(),
// which may append a virtual global bag (plan/decomposition.cpp:1636-1638, t.addGlobalNode).
// adaptiveShareJoin then indexes trieNodes[t.numNodes - 1] (matching/adaptive.cpp:304-310,
// :1223-1227, :2179) one past the end of the vector, which crashes most queries (878 of the
// 1,000 shipped yeast queries segfault in an unmodified build).  Allocating the roots after
// reorderBags leaves every algorithmic step of the reference untouched.
//
// Usage: ref_driver -q <query> -d <data> [-m <KB>] [-c <local.csv>]
// Prints the same lines as ASDMatch.out.  The -c CSV is written only in the build compiled with
// -DALL_LEVEL -DLOCAL_COUNT (oracle/_ref/ref_driver_local), like the reference's own -c.
#include <chrono>
#include <cstring>
#include <fstream>
#include <iostream>
#include "adaptive.h"
#include "candidate_space.h"
#include "command.h"
#include "estimator.h"
#include "join.h"
#include "optimizer.h"

int main(int argc, char **argv) {
    Command cmd(argc, argv);
    int memKB = cmd.getMemoryBudget();
    if (memKB == 0) memKB = 16e5;
    const size_t budget = (size_t)1024 * memKB;
    Graph q, g;
    q.loadGraphFromTextFile(cmd.getQueryGraphPath());
    g.loadGraphFromTextFile(cmd.getDataGraphPath());
    CandidateSpace cs(q, g);
    using clk = std::chrono::steady_clock;
    auto t0 = clk::now();
    if (!cs.buildCandCFL()) return 0;  // the reference prints nothing in this case
    std::chrono::duration<double> dt = clk::now() - t0;
    std::cout << "Filtering time (s): " << dt.count() << std::endl;
    cs.setQueryGraphWeights(q);

    t0 = clk::now();
    FHD fhd;
    q.buildFHD(fhd);
    HyperTree t;
    t.buildFromTD(fhd);
    dt = clk::now() - t0;
    const double fhdTime = dt.count();
    std::cout << "FHD time (s): " << fhdTime << std::endl;

    const ui nq = q.getNumVertices();
    bool *visited = new bool[g.getNumVertices()];
    std::memset(visited, 0, sizeof(bool) * g.getNumVertices());
    VertexID *partMatch = new VertexID[nq];
    VertexID **candidates = new VertexID *[nq];
    const ui maxSize = cs.getMaxSize();
    for (ui i = 0; i < nq; ++i) candidates[i] = new VertexID[maxSize];
    ui *candCount = new ui[nq];
    std::memset(candCount, 0, sizeof(ui) * nq);

    reorderBags(q, t);
    // Trie roots sized after the bag list is final (see the header comment).
    std::vector<TrieNode *> nodes(t.numNodes);
    for (ui i = 0; i < t.numNodes; ++i) nodes[i] = new TrieNode();

    ui numNodes = t.numNodes;
    if (t.newGlobalNode) --numNodes;
    std::vector<SubsetStructure> dp(numNodes);
    smallCard(q, cs);
    t0 = clk::now();
    for (VertexID nID = 0; nID < numNodes; ++nID) {
        std::vector<VertexID> local(t.nodes[nID].attributes,
                                    t.nodes[nID].attributes + t.nodes[nID].numAttributes);
        SubsetStructure s(local, q, false);
        s.optimalPlanDP(q, cs, visited, partMatch, candidates, candCount);
        s.reverseDP(q, cs);
        dp[nID] = s;
    }
    dt = clk::now() - t0;
    const double cardTime = dt.count() + fhdTime;

    bool traverseAll = false;
#if defined(ALL_LEVEL) || defined(COLLECT_RESULT)
    traverseAll = true;
#endif
#ifdef LOCAL_COUNT
    gLocalCount = std::vector<std::vector<size_t>>(g.getNumVertices(), std::vector<size_t>(nq, 0));
#endif
    PrefixNode *pt = nullptr;
    double minCost = 0;
    int reorder = 1;
    t0 = clk::now();
    if (numNodes != 1)
        optCostPlan(q, t, cs, pt, dp, minCost, reorder, true, false);
    else
        simplePlan(q, t, cs, pt, dp, visited, partMatch, candidates, candCount, minCost, nullptr, false);
    dt = clk::now() - t0;
    std::cout << "Planning Time: " << dt.count() + cardTime << std::endl;
    std::cout << "cost: " << minCost << std::endl;

    std::vector<std::vector<VertexID>> result;
    size_t count = 0;
    t0 = clk::now();
    adaptiveShareJoin(q, t, pt, cs, nodes, visited, result, count, traverseAll, budget, false);
    dt = clk::now() - t0;
    std::cout << "Execution Time: " << dt.count() << std::endl;
    std::cout << "Number of matches: " << count << std::endl;
    std::cout << "Number of fixed attributes: " << gMaxNewFixed << std::endl;
    std::cout << "Number of intersections: " << gNumInterSection << std::endl;
    std::cout << "Number of fix cases: " << gNumFixedCase << std::endl;
#ifdef LOCAL_COUNT
    const std::string countPath = cmd.getCountPath();
    if (!countPath.empty()) {
        std::ofstream out(countPath);
        for (VertexID v = 0; v < g.getNumVertices(); ++v) {
            for (VertexID u = 0; u < nq; ++u) {
                out << gLocalCount[v][u];
                if (u + 1 != nq) out << ",";
            }
            out << "\n";
        }
    }
#endif
    return 0;
}
