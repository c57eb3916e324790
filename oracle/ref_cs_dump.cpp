// Test-infrastructure driver (not product code): links the reference's own relation/ and
// utility/ objects and prints the candidate space that CandidateSpace::buildCandCFL builds
// (relation/candidate_space.cpp:550-721) so the restated filter can be compared set by set.
//
// Usage: ref_cs_dump <query.graph> <data.graph> [--edges]
// Output (text):
//   EMPTY                                   if buildCandCFL returned false
//   C <u> <n> <v_0> ... <v_{n-1}>           candidate set of query vertex u, ascending
//   W <u> <u2> <sum_v |candidateEdge[u][v][u2]|>   per query arc, in query CSR order
//   E <u> <u2> <v> <n> <v2...>              (--edges) each candidate-edge list
#include <cstdio>
#include <cstring>
#include <string>
#include "candidate_space.h"

int main(int argc, char **argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s query data [--edges]\n", argv[0]);
        return 2;
    }
    bool edges = argc > 3 && std::strcmp(argv[3], "--edges") == 0;
    Graph q, g;
    q.loadGraphFromTextFile(argv[1]);
    g.loadGraphFromTextFile(argv[2]);
    CandidateSpace cs(q, g);
    if (!cs.buildCandCFL()) {
        std::printf("EMPTY\n");
        return 0;
    }
    for (VertexID u = 0; u < q.getNumVertices(); ++u) {
        const auto &c = cs.candidateSet[u];
        std::printf("C %u %zu", u, c.size());
        for (VertexID v : c) std::printf(" %u", v);
        std::printf("\n");
    }
    for (VertexID u = 0; u < q.getNumVertices(); ++u) {
        ui n;
        const VertexID *nb = q.getNeighbors(u, n);
        for (ui i = 0; i < n; ++i) {
            VertexID u2 = nb[i];
            unsigned long long w = 0;
            for (VertexID v : cs.candidateSet[u]) w += cs.candidateEdge[u][v][u2].size();
            std::printf("W %u %u %llu\n", u, u2, w);
            if (edges) {
                for (VertexID v : cs.candidateSet[u]) {
                    const auto &l = cs.candidateEdge[u][v][u2];
                    std::printf("E %u %u %u %zu", u, u2, v, l.size());
                    for (VertexID x : l) std::printf(" %u", x);
                    std::printf("\n");
                }
            }
        }
    }
    return 0;
}
