// Bag-table engine: exact counts and local counts through a tree decomposition of the query.
//
// The injective count is a signed sum of homomorphism counts.  For every partition pi of the
// query vertices into blocks of pairwise non-adjacent, same-label vertices,
//     #embeddings = sum_pi mu(pi) * hom(q / pi),   mu(pi) = prod_B (-1)^(|B|-1) (|B|-1)!
// (Moebius inversion over the partition lattice; blocks that mix labels or contain a query edge
// contribute nothing on a loop-free data graph).  hom(q / pi) is a sum-product over the quotient
// graph whose variables are the blocks, each ranging over the intersection of its members'
// candidate sets, with one 0/1 factor per quotient edge (data adjacency).  It is evaluated by
// variable elimination of width at most two: every intermediate factor is a table over one bag of
// at most two variables (a vector over a candidate domain, or a dense |D(a)| x |D(b)| table), and
// eliminating a variable joins the tables that share it.  All arithmetic is in Z / 2^64, so the
// result equals the reference's size_t count (config/config.h:46) exactly.
//
// Local counts: the marginal of block t (every other block eliminated, t last) gives, for each
// v in D(t), the homomorphisms with t -> v; summed with the Moebius coefficients over all
// partitions it is the number of embeddings with f(u) = v for every u in t.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

struct asdm_cs;

namespace asdm {

// One elimination step: eliminate block v, whose live neighbours at that point are nb[0..nn).
struct BagStep {
    uint32_t v;
    uint32_t nn;
    uint32_t nb[2];
};

struct BagTerm {
    int64_t mu = 1;                              // Moebius coefficient of the partition
    std::vector<std::vector<uint32_t>> blocks;   // query vertices of each block
    std::vector<uint32_t> rep;                   // member with the smallest candidate set
    std::vector<std::pair<uint32_t, uint32_t>> edges;  // quotient edges (A < B)
    std::vector<uint32_t> arc_ab, arc_ba;        // a query arc a -> b (and b -> a), a in A, b in B
    std::vector<BagStep> count_steps;            // eliminates every block
    std::vector<std::vector<BagStep>> local_steps;  // [t]: eliminates every block but t
};

struct BagPlan {
    bool local = false;
    std::vector<BagTerm> terms;
    double cost = 0;          // estimated element operations of the count (and marginals)
    double peak_table = 0;    // entries of the largest dense table
};

// Largest number of partition terms and of entries of one dense table the planner accepts.
constexpr uint64_t kMaxBagTerms = 1u << 14;
constexpr double kMaxBagTable = 2.0e9;

// Builds the bag plan of cs's query; false (with the reason in why) when some quotient needs a
// table over three or more variables, or the plan exceeds the limits above.
bool build_bag_plan(const asdm_cs *cs, bool local, BagPlan &out, std::string &why);

}  // namespace asdm
