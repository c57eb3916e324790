// Host-only driver for the sanitizer fuzz of the graph and .card readers (tests/test_loader_fuzz.py).
// Built with g++ -fsanitize=address,undefined from graph_host.cpp and card_io.cpp, without the HIP
// units: the one device hook those files call is defined here as a no-op.
//   loader_fuzz g FILE...   load each file as a graph (text or binary CSR) and copy its CSR out
//   loader_fuzz c FILE...   read each file as a .card map
//   loader_fuzz s OUT IN    write IN's graph as a binary CSR to OUT
// Prints "rc <code>: <files>" per status code.  The sanitizers abort on the first memory error.
#include <cstdio>
#include <cstring>
#include <map>
#include <vector>

#include "host.h"

namespace asdm {
void release_device_graph(asdm_graph &) {}
}  // namespace asdm

int main(int argc, char **argv) {
    if (argc < 2) return 2;
    if (!std::strcmp(argv[1], "s")) {
        if (argc != 4) return 2;
        asdm_graph *g = nullptr;
        if (asdm_graph_load_text(argv[3], &g) != ASDM_OK) return 1;
        const int rc = asdm_graph_save_binary(g, argv[2]);
        asdm_graph_free(g);
        return rc == ASDM_OK ? 0 : 1;
    }
    const bool card = !std::strcmp(argv[1], "c");
    std::map<int, int> by_code;
    for (int i = 2; i < argc; ++i) {
        int rc;
        if (card) {
            std::vector<uint64_t> masks(4096);
            std::vector<double> cards(4096);
            uint32_t n = 0;
            double t = 0;
            rc = asdm_card_read(argv[i], masks.data(), cards.data(), 4096, &n, &t);
        } else {
            asdm_graph *g = nullptr;
            rc = asdm_graph_load_text(argv[i], &g);
            if (g) {
                uint32_t n = 0, nl = 0;
                uint64_t m = 0;
                asdm_graph_info(g, &n, &m, &nl);
                std::vector<uint32_t> off(n + 1ull), nbr(m), lab(n);
                asdm_graph_csr(g, off.data(), nbr.data(), lab.data());
                asdm_graph_free(g);
            }
        }
        by_code[rc]++;
    }
    for (const auto &kv : by_code) std::printf("rc %d: %d\n", kv.first, kv.second);
    return 0;
}
