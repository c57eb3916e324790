// Host-side internals of libasdm shared by the graph I/O, generators, planner and engine.
#pragma once

#include <cstdint>
#include <exception>
#include <new>
#include <string>
#include <vector>

#include "asdm.h"

// Handler of the function-try-block around every int-returning entry point: no C++ exception
// crosses the C ABI.  A failed host allocation returns ASDM_E_OOM, any other exception
// ASDM_E_INVALID.
#define ASDM_HOST_CATCH                     \
    catch (const std::bad_alloc &) {        \
        return ASDM_E_OOM;                  \
    }                                       \
    catch (...) {                           \
        return ASDM_E_INVALID;              \
    }

namespace asdm {

// Largest label value a graph may hold: labels index per-label tables of (max label + 1) entries.
constexpr uint32_t kMaxLabel = (1u << 24) - 1;

// Label-sorted CSR as the reference's Graph holds it (relation/graph.h:34-59): neighbours of v
// are nbr[off[v] .. off[v+1]) sorted by (label(nbr), nbr).
struct DeviceGraph {
    int device = -1;
    uint32_t *off = nullptr;      // [n+1]
    uint32_t *nbr = nullptr;      // [m]
    uint32_t *lab = nullptr;      // [n]
    uint32_t *loff = nullptr;     // [n * (nlab + 1)] start of each label run in nbr (absolute)
    uint32_t *vbl_off = nullptr;  // [nlab + 1]
    uint32_t *vbl = nullptr;      // [n] vertices grouped by label, ascending id
    uint64_t *labbits = nullptr;  // [nlab][W] membership bitmap of each label (W = ceil(n / 64))
};

}  // namespace asdm

struct asdm_graph {
    uint32_t n = 0;
    uint64_t m = 0;  // arcs
    uint32_t nlab = 0;
    std::vector<uint32_t> lab, off, nbr;
    std::vector<uint32_t> vbl_off, vbl;
    std::vector<float> lab_avgdeg;  // mean degree of the vertices of each label
    bool loop_free = false;         // no vertex is its own neighbour
    asdm::DeviceGraph dev;

    uint32_t deg(uint32_t v) const { return off[v + 1] - off[v]; }
    // [b, b + cnt) = label-l slice of v's adjacency
    void label_slice(uint32_t v, uint32_t l, uint32_t &b, uint32_t &cnt) const;
    bool has_edge(uint32_t v, uint32_t w) const;
    uint32_t label_count(uint32_t l) const { return l < nlab ? vbl_off[l + 1] - vbl_off[l] : 0; }
};

namespace asdm {

// Sorts adjacency by (label, id) and builds the vertices-by-label index.
void finish_graph(asdm_graph &g);
// The indexes over a sorted CSR: vertices by label, mean degree per label, self-loop flag.
void index_graph(asdm_graph &g);
asdm_graph *graph_from_edges(uint32_t n, const uint32_t *labels, uint64_t ne, const uint32_t *edges);
void release_device_graph(asdm_graph &g);

// Portable PRNG: splitmix64-seeded xoshiro256**.
struct Rng {
    uint64_t s[4];
    explicit Rng(uint64_t seed);
    uint64_t next();
    double uniform();                  // [0, 1) with 53 random bits
    uint64_t below(uint64_t bound);    // [0, bound), unbiased
};

}  // namespace asdm
