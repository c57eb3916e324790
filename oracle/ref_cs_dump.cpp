// Test-infrastructure driver (not product code): links the reference's own relation/ and
// utility/ objects and prints the candidate space that CandidateSpace::buildCandCFL builds
// (relation/candidate_space.cpp:550-721) so the restated filter can be compared set by set.
//
// Usage: ref_cs_dump <query.graph> <data.graph> [--edges | --paths [max_x_path]]
// Output (text):
//   EMPTY                                   if buildCandCFL returned false
//   C <u> <n> <v_0> ... <v_{n-1}>           candidate set of query vertex u, ascending
//   W <u> <u2> <sum_v |candidateEdge[u][v][u2]|>   per query arc, in query CSR order
//   E <u> <u2> <v> <n> <v2...>              (--edges) each candidate-edge list
//   D <i> <j> <dist>                        (--paths) CandidateSpace::getDist after
//                                           setQueryGraphWeights (relation/candidate_space.cpp:507-528)
//   N <i> <j> <next>                        (--paths) the successor matrix `next` of the same call
//   P <i> <j> <len> <u_0> ... <u_{len-1}>   (--paths) reconstructPath(i, j) (:529-537); len = -1 when
//                                           following `next` from i does not reach j within |Vq| steps
//                                           (products that wrap mod 2^64 can leave a cycle in `next`,
//                                           and reconstructPath would loop without end)
//   X <i> <j> <v> <n> <v2...>               (--paths) optimizedCartesianProduct(cs, v, path(i, j))
//                                           (plan/estimator.cpp:41-102) for the first, middle and
//                                           last candidate v of C(i), paths of two or more vertices
//                                           (and at most max_x_path vertices when given: the function
//                                           walks every path of candidate edges)
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <string>
#include "candidate_space.h"
#include "estimator.h"

int main(int argc, char **argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s query data [--edges]\n", argv[0]);
        return 2;
    }
    bool edges = argc > 3 && std::strcmp(argv[3], "--edges") == 0;
    bool paths = argc > 3 && std::strcmp(argv[3], "--paths") == 0;
    const size_t max_x_path = paths && argc > 4 ? (size_t)std::atoi(argv[4]) : (size_t)-1;
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
    if (paths) {
        cs.setQueryGraphWeights(q);
        const ui nq = q.getNumVertices();
        std::vector<VertexID> buf;
        for (VertexID i = 0; i < nq; ++i)
            for (VertexID j = 0; j < nq; ++j) {
                std::printf("D %u %u %llu\n", i, j, (unsigned long long)cs.getDist(i, j));
                std::printf("N %u %u %u\n", i, j, cs.next[i][j]);
                bool ends = cs.next[i][j] != std::numeric_limits<VertexID>::max();
                if (ends) {
                    VertexID a = i;
                    ui steps = 0;
                    while (a != j && steps <= nq) {
                        a = cs.next[a][j];
                        ++steps;
                        if (a == std::numeric_limits<VertexID>::max()) break;
                    }
                    ends = a == j;
                }
                if (!ends && cs.next[i][j] != std::numeric_limits<VertexID>::max()) {
                    std::printf("P %u %u -1\n", i, j);
                    continue;
                }
                const std::vector<VertexID> path = cs.reconstructPath(i, j);
                std::printf("P %u %u %zu", i, j, path.size());
                for (VertexID u : path) std::printf(" %u", u);
                std::printf("\n");
                if (i == j || path.size() < 2 || path.size() > max_x_path) continue;
                const auto &ci = cs.candidateSet[i];
                const size_t picks[3] = {0, ci.size() / 2, ci.size() - 1};
                for (int k = 0; k < 3; ++k) {
                    if (k && picks[k] == picks[k - 1]) continue;
                    const VertexID v = ci[picks[k]];
                    buf.assign(cs.candidateSet[j].size() + 1, 0);
                    ui cnt = 0;
                    optimizedCartesianProduct(cs, v, path, buf.data(), cnt);
                    std::printf("X %u %u %u %u", i, j, v, cnt);
                    for (ui t = 0; t < cnt; ++t) std::printf(" %u", buf[t]);
                    std::printf("\n");
                }
            }
    }
    return 0;
}
