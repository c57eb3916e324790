// Enumeration engine: exact uint64 embedding counts and LOCAL_COUNT matrices.
//
// The count is the number of injective, label- and edge-preserving maps q -> g (the number the
// reference's adaptiveShareJoin produces, matching/adaptive.cpp:1179-2279, independent of plan).
// Execution model (DFS over chunks, BFS inside a chunk):
//   * depths 0..D-1 of the plan are enumerated level by level into frontier tables of partial
//     embeddings (row = positions in C(u_0..u_{d-1})), in chunks that fit a per-level budget;
//   * depths D..nq-1 are leaves: an independent set whose neighbours are all matched, so their
//     completions are counted in closed form per row: a product over label groups, with
//     inclusion-exclusion over set partitions inside a group of equal-label leaves (the
//     quantity numUniqueCombine computes, utility/utils.cpp:145-201; leaves of different labels
//     can never collide).  This replaces traverse(count) (matching/join.cpp:910-1014).
// Candidates of depth d are the intersection of the candidate-edge lists of its matched
// neighbours (LeapfrogJoin's role, utility/compute_set_intersection.cpp:193-233): the shortest
// list is the probe, an aligned group of G lanes walks it and binary-searches the others, and
// survivors are compacted in order with a ballot prefix.  Injectivity is tested against the
// data ids of earlier depths with the same label (other labels cannot collide).
#include <hip/hip_runtime.h>

#include <atomic>
#include <thread>

#include <algorithm>
#include <chrono>
#include <functional>
#include <cstring>
#include <memory>
#include <type_traits>
#include <vector>

#include "engine.h"
#include "kernels_common.h"

namespace asdm {

struct ListRef {
    uint64_t lo;
    uint32_t len;
};

__device__ inline ListRef list_of(const DPlan *__restrict__ P, const uint64_t *__restrict__ ce_off,
                                  const uint32_t *__restrict__ row, uint32_t d, uint32_t j) {
    const uint64_t base = P->bbase[d][j] + row[P->bdep[d][j]];
    const uint64_t lo = ce_off[base];
    return {lo, (uint32_t)(ce_off[base + 1] - lo)};
}

// Data ids of the matched vertices that share depth d's label (at most 8 held in registers).
struct SameLabel {
    uint32_t y[8];
    uint32_t ns;
    const uint32_t *row;
    const uint32_t *sdep;
    const uint64_t *coff;
};

__device__ inline void load_same(SameLabel &S, const DPlan *__restrict__ P, const uint32_t *__restrict__ cand,
                                 const uint32_t *__restrict__ row, uint32_t d) {
    S.ns = P->ns[d];
    S.row = row;
    S.sdep = P->sdep[d];
    S.coff = P->coff;
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        uint32_t dep = P->sdep[d][s];
        S.y[s] = (uint32_t)s < S.ns ? cand[P->coff[dep] + row[dep]] : 0xffffffffu;
    }
}

__device__ inline bool injective(const SameLabel &S, const uint32_t *__restrict__ cand, uint32_t xid) {
    bool ok = true;
#pragma unroll
    for (int s = 0; s < 8; ++s) ok &= xid != S.y[s];
    for (uint32_t s = 8; s < S.ns && ok; ++s) {
        uint32_t dep = S.sdep[s];
        ok = xid != cand[S.coff[dep] + S.row[dep]];
    }
    return ok;
}

// Position of data vertex xid in C(u_d), the candidate set of depth d's query vertex, or
// 0xffffffff when xid is not a candidate: one membership-bitmap word and its rank prefix.
__device__ inline uint32_t cand_pos(const DPlan *__restrict__ P, uint32_t d, uint32_t xid) {
    const uint64_t idx = (uint64_t)P->u[d] * P->W + (xid >> 6);
    const uint64_t w = P->bits[idx];
    const uint32_t b = xid & 63;
    if (!((w >> b) & 1)) return 0xffffffffu;
    return P->rank[idx] + (uint32_t)__builtin_popcountll(w & ((1ull << b) - 1));
}

// Number of list elements (positions in C(u_d)) whose data id is one of the same-label matched
// vertices: each such vertex is located in C(u_d) through the bitmap, then in the list by binary
// search.
__device__ inline uint32_t same_hits(const SameLabel &S, const DPlan *__restrict__ P, uint32_t d,
                                     const uint32_t *__restrict__ cand, const uint32_t *__restrict__ list,
                                     uint32_t len) {
    uint32_t h = 0;
    for (uint32_t s = 0; s < S.ns; ++s) {
        uint32_t y;
        if (s < 8) {
            y = S.y[0];
#pragma unroll
            for (int t = 1; t < 8; ++t)
                if ((uint32_t)t == s) y = S.y[t];
        } else {
            uint32_t dep = S.sdep[s];
            y = cand[S.coff[dep] + S.row[dep]];
        }
        const uint32_t x = cand_pos(P, d, y);
        if (x != 0xffffffffu && contains_u32(list, len, x)) ++h;
    }
    return h;
}

// -------------------------------------------------------------------------- expand
// Row-level device functions shared by the single-query kernels and the batch kernels.  A lane
// group of G aligned lanes handles one row; all lanes of the group call these together.

// Number of valid depth-d candidates of a row (reduced over the group; every lane gets it).
template <int G>
__device__ inline uint32_t expand_count_row(const DPlan *__restrict__ P, uint32_t d, const uint32_t *__restrict__ row,
                                            const uint64_t *__restrict__ ce_off, const uint32_t *__restrict__ ce_val,
                                            const uint32_t *__restrict__ cand, uint64_t &bytes) {
    const int lg = threadIdx.x & (G - 1);
    const uint32_t nb = P->nb[d];
    const uint64_t cand_d = P->coff[d];
    uint32_t pj = 0;
    ListRef pr = list_of(P, ce_off, row, d, 0);
    for (uint32_t j = 1; j < nb; ++j) {
        ListRef r = list_of(P, ce_off, row, d, j);
        if (r.len < pr.len) {
            pr = r;
            pj = j;
        }
    }
    SameLabel S;
    load_same(S, P, cand, row, d);
    const uint32_t *probe = ce_val + pr.lo;
    uint32_t c;
    if (nb == 1) {
        c = lg == 0 ? pr.len - same_hits(S, P, d, cand, probe, pr.len) : 0;
        bytes += lg == 0 ? 4ull * d + 16 : 0;
    } else {
        c = 0;
        for (uint32_t t = lg; t < pr.len; t += G) {
            const uint32_t x = probe[t];
            bool ok = true;
            for (uint32_t j = 0; j < nb && ok; ++j) {
                if (j == pj) continue;
                ListRef r = list_of(P, ce_off, row, d, j);
                ok = contains_u32(ce_val + r.lo, r.len, x);
            }
            if (ok) ok = injective(S, cand, cand[cand_d + x]);
            c += ok;
        }
        bytes += 4ull * ((pr.len + G - 1 - lg) / G);
        if (lg == 0) bytes += 4ull * d + 16ull * nb;
    }
    return (uint32_t)group_sum<G>(c);
}

// Writes the row's children (width d + 1) at out[o ...] in candidate order.
template <int G>
__device__ inline void expand_fill_row(const DPlan *__restrict__ P, uint32_t d, const uint32_t *__restrict__ row,
                                       const uint64_t *__restrict__ ce_off, const uint32_t *__restrict__ ce_val,
                                       const uint32_t *__restrict__ cand, uint64_t o, uint64_t nout,
                                       uint32_t *__restrict__ out, uint64_t &bytes) {
    const int lg = threadIdx.x & (G - 1);
    const uint32_t nb = P->nb[d];
    const uint64_t cand_d = P->coff[d];
    uint32_t pj = 0;
    ListRef pr = list_of(P, ce_off, row, d, 0);
    for (uint32_t j = 1; j < nb; ++j) {
        ListRef r = list_of(P, ce_off, row, d, j);
        if (r.len < pr.len) {
            pr = r;
            pj = j;
        }
    }
    SameLabel S;
    load_same(S, P, cand, row, d);
    const uint32_t *probe = ce_val + pr.lo;
    for (uint32_t t0 = 0; t0 < pr.len; t0 += G) {
        const uint32_t t = t0 + lg;
        bool ok = t < pr.len;
        uint32_t x = 0;
        if (ok) {
            x = probe[t];
            for (uint32_t j = 0; j < nb && ok; ++j) {
                if (j == pj) continue;
                ListRef r = list_of(P, ce_off, row, d, j);
                ok = contains_u32(ce_val + r.lo, r.len, x);
            }
            if (ok) ok = injective(S, cand, cand[cand_d + x]);
        }
        const uint64_t gm = group_bits<G>(__ballot(ok));
        if (ok) {
            uint32_t *dst = out + (o + __popcll(gm & ((1ull << lg) - 1))) * (d + 1);
            for (uint32_t k = 0; k < d; ++k) dst[k] = row[k];
            dst[d] = x;
        }
        o += __popcll(gm);
    }
    bytes += 4ull * ((pr.len + G - 1 - lg) / G);
    if (lg == 0) bytes += 4ull * d + 16ull * nb + 4ull * (d + 1) * nout;
}

// Rows of width d -> candidates of depth d.  Count mode writes counts[row]; fill mode writes rows
// of width d + 1 at out[(outoff[row] - out_base) ...] in (row, candidate) order.
template <int G, bool FILL>
__global__ void __launch_bounds__(256) k_expand(const DPlan *__restrict__ P, uint32_t d,
                                                const uint32_t *__restrict__ rows, uint64_t nrows,
                                                const uint64_t *__restrict__ ce_off, const uint32_t *__restrict__ ce_val,
                                                const uint32_t *__restrict__ cand, uint32_t *__restrict__ counts,
                                                const uint64_t *__restrict__ outoff, uint64_t out_base,
                                                uint32_t *__restrict__ out, unsigned long long *__restrict__ bytes_ctr) {
    const int lg = threadIdx.x & (G - 1);
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x / G;
    uint64_t bytes = 0;
    for (uint64_t it = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / G; it < nrows; it += stride) {
        const uint32_t *row = rows + it * d;
        if constexpr (!FILL) {
            const uint32_t c = expand_count_row<G>(P, d, row, ce_off, ce_val, cand, bytes);
            if (lg == 0) counts[it] = c;
        } else {
            expand_fill_row<G>(P, d, row, ce_off, ce_val, cand, outoff[it] - out_base, outoff[it + 1] - outoff[it], out,
                               bytes);
        }
    }
    bytes = group_sum<64>(bytes);
    if ((threadIdx.x & 63) == 0 && bytes_ctr) atomicAdd(bytes_ctr, (unsigned long long)bytes);
}

// -------------------------------------------------------------------------- triangle leaves
// For a triangle leaf e (matched neighbours p0 = u_bdep[e][0], p1 = u_bdep[e][1], p0 ~ p1 in the
// query), the intersection of its two candidate-edge lists depends only on the candidate edge
// (f(p0), f(p1)).  It is computed once per candidate edge of arc (p0 -> p1) -- the
// materialization of the bag {p0, p1, e} -- instead of once per partial embedding.
// One lane per candidate edge k (global index base.. of the arc's lists).
// The two lists of candidate edge ek of the arc, shorter one first.
struct TriLists {
    const uint32_t *A, *B;
    uint32_t na, nb;
};

__device__ inline TriLists tri_lists(const uint64_t *__restrict__ ce_off, const uint32_t *__restrict__ ce_val,
                                     uint64_t arc_item_base, uint32_t nitems, uint64_t ek, uint64_t l0_base,
                                     uint64_t l1_base) {
    // i0 = position of f(p0): last item whose list starts at or before ek
    uint32_t lo = 0, hi = nitems;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (ce_off[arc_item_base + mid] <= ek) lo = mid;
        else hi = mid;
    }
    const uint32_t i0 = lo, i1 = ce_val[ek];
    const uint64_t a0 = ce_off[l0_base + i0], a1 = ce_off[l0_base + i0 + 1];
    const uint64_t b0 = ce_off[l1_base + i1], b1 = ce_off[l1_base + i1 + 1];
    TriLists L{ce_val + a0, ce_val + b0, (uint32_t)(a1 - a0), (uint32_t)(b1 - b0)};
    if (L.na > L.nb) {
        const uint32_t *t = L.A;
        L.A = L.B;
        L.B = t;
        const uint32_t tn = L.na;
        L.na = L.nb;
        L.nb = tn;
    }
    return L;
}

// Edges whose shorter list holds more than kTriWide entries are left to k_tri_table_wide.
constexpr uint32_t kTriWide = 64;

// Entry k of a triangle table by one lane (shorter list at most kTriWide long) ...
__device__ inline void tri_entry_narrow(const uint64_t *__restrict__ ce_off, const uint32_t *__restrict__ ce_val,
                                        uint64_t arc_item_base, uint32_t nitems, uint64_t e_first, uint64_t k,
                                        uint64_t l0_base, uint64_t l1_base, uint32_t *__restrict__ tab) {
    const TriLists L = tri_lists(ce_off, ce_val, arc_item_base, nitems, e_first + k, l0_base, l1_base);
    if (L.na > kTriWide) return;
    const uint32_t *A = L.A, *B = L.B;
    const uint32_t na = L.na, nb = L.nb;
    uint32_t c = 0, p = 0;
    for (uint32_t t = 0; t < na; ++t) {  // probe the shorter list, galloping in the longer one
        const uint32_t x = A[t];
        uint32_t step = 1, q = p;
        while (q + step < nb && B[q + step] < x) { q += step; step <<= 1; }
        uint32_t hi2 = q + step < nb ? q + step + 1 : nb;
        p = q + lower_bound_u32(B + q, hi2 - q, x);
        if (p < nb && B[p] == x) ++c;
    }
    tab[k] = c;
}

// ... and by a whole wave (the others): lanes split the shorter list and binary-search the longer.
__device__ inline void tri_entry_wide(const uint64_t *__restrict__ ce_off, const uint32_t *__restrict__ ce_val,
                                      uint64_t arc_item_base, uint32_t nitems, uint64_t e_first, uint64_t k,
                                      uint64_t l0_base, uint64_t l1_base, uint32_t *__restrict__ tab) {
    const int lane = threadIdx.x & 63;
    const TriLists L = tri_lists(ce_off, ce_val, arc_item_base, nitems, e_first + k, l0_base, l1_base);
    if (L.na <= kTriWide) return;  // wave-uniform
    uint64_t c = 0;
    for (uint32_t t = lane; t < L.na; t += 64) c += contains_u32(L.B, L.nb, L.A[t]);
    c = group_sum<64>(c);
    if (lane == 0) tab[k] = (uint32_t)c;
}

__global__ void k_tri_table(const uint64_t *__restrict__ ce_off, const uint32_t *__restrict__ ce_val,
                            uint64_t arc_item_base, uint32_t nitems /* |C(p0)| */, uint64_t nedges, uint64_t l0_base,
                            uint64_t l1_base, uint32_t *__restrict__ tab) {
    const uint64_t e_first = ce_off[arc_item_base];  // first candidate edge of the arc
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < nedges; k += (uint64_t)gridDim.x * blockDim.x)
        tri_entry_narrow(ce_off, ce_val, arc_item_base, nitems, e_first, k, l0_base, l1_base, tab);
}

// The edges k_tri_table leaves: one wave per edge.
__global__ void __launch_bounds__(256) k_tri_table_wide(const uint64_t *__restrict__ ce_off,
                                                        const uint32_t *__restrict__ ce_val, uint64_t arc_item_base,
                                                        uint32_t nitems, uint64_t nedges, uint64_t l0_base,
                                                        uint64_t l1_base, uint32_t *__restrict__ tab) {
    const uint64_t e_first = ce_off[arc_item_base];
    const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    for (uint64_t k = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; k < nedges; k += nwaves)
        tri_entry_wide(ce_off, ce_val, arc_item_base, nitems, e_first, k, l0_base, l1_base, tab);
}

// The triangle tables of many queries in one launch pair (batch engine): job j covers the items
// [pref[j], pref[j+1]) (its candidate edges).
struct TriJob {
    const uint64_t *ce_off;
    const uint32_t *ce_val;
    uint64_t arc_item_base, nedges, l0_base, l1_base;
    uint32_t nitems;
    uint32_t *tab;
};

__device__ inline uint32_t job_of(const uint64_t *__restrict__ pref, uint32_t njobs, uint64_t item) {
    uint32_t lo = 0, hi = njobs;  // last job with pref[j] <= item
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (pref[mid] <= item) lo = mid;
        else hi = mid;
    }
    return lo;
}

__global__ void k_tri_table_multi(const TriJob *__restrict__ jobs, const uint64_t *__restrict__ pref, uint32_t njobs) {
    const uint64_t total = pref[njobs];
    for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total; g += (uint64_t)gridDim.x * blockDim.x) {
        const TriJob &J = jobs[job_of(pref, njobs, g)];
        const uint64_t k = g - pref[&J - jobs];
        tri_entry_narrow(J.ce_off, J.ce_val, J.arc_item_base, J.nitems, J.ce_off[J.arc_item_base], k, J.l0_base,
                         J.l1_base, J.tab);
    }
}

__global__ void __launch_bounds__(256) k_tri_table_wide_multi(const TriJob *__restrict__ jobs,
                                                              const uint64_t *__restrict__ pref, uint32_t njobs) {
    const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    const uint64_t total = pref[njobs];
    for (uint64_t w = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; w < total; w += nwaves) {
        const uint32_t j = job_of(pref, njobs, w);  // wave-uniform
        const TriJob &J = jobs[j];
        tri_entry_wide(J.ce_off, J.ce_val, J.arc_item_base, J.nitems, J.ce_off[J.arc_item_base], w - pref[j], J.l0_base,
                       J.l1_base, J.tab);
    }
}

// Candidates of triangle leaf e given the two matched positions: |list0 ∩ list1| from the table,
// minus matched same-label vertices lying in both lists.
__device__ inline uint64_t tri_count(const DPlan *__restrict__ P, uint32_t e, uint32_t i0, uint32_t i1,
                                     ListRef l0, ListRef l1, const uint64_t *__restrict__ ce_off,
                                     const uint32_t *__restrict__ ce_val, const uint32_t *__restrict__ cand,
                                     const uint32_t *y, uint32_t ns) {
    const uint64_t lo = ce_off[P->tri_base[e] + i0];
    const uint32_t len = (uint32_t)(ce_off[P->tri_base[e] + i0 + 1] - lo);
    const uint32_t k = lower_bound_u32(ce_val + lo, len, i1);  // (f(p0), f(p1)) is a candidate edge
    uint64_t c = P->tri_tab[e][lo + k - ce_off[P->tri_base[e]]];
    for (uint32_t s = 0; s < ns; ++s) {
        const uint32_t x = cand_pos(P, e, y[s]);
        if (x != 0xffffffffu && contains_u32(ce_val + l0.lo, l0.len, x) && contains_u32(ce_val + l1.lo, l1.len, x))
            --c;
    }
    return c;
}

// -------------------------------------------------------------------------- leaves
// |∩_{leaf e in mask} S_e| where S_e = ∩ of e's candidate-edge lists minus matched same-label
// vertices, reduced over the lane group.  Lists of different leaves index different candidate
// sets, so the probe element is carried as a data id and re-located in each other leaf's C(u_e).
template <int G>
__device__ inline uint64_t leaf_set_count(const DPlan *__restrict__ P, uint32_t g, uint32_t mask,
                                          const uint32_t *__restrict__ row, const uint64_t *__restrict__ ce_off,
                                          const uint32_t *__restrict__ ce_val, const uint32_t *__restrict__ cand,
                                          const SameLabel &S, uint64_t &bytes) {
    const int lg = threadIdx.x & (G - 1);
    // probe: the shortest list among all lists of the leaves in mask
    uint32_t pe = 0, pj = 0, nlists = 0;
    ListRef pr{0, 0xffffffffu};
    for (uint32_t i = 0; i < P->gsize[g]; ++i) {
        if (!(mask >> i & 1)) continue;
        const uint32_t e = P->gleaf[g][i];
        for (uint32_t j = 0; j < P->nb[e]; ++j) {
            ListRef r = list_of(P, ce_off, row, e, j);
            ++nlists;
            if (r.len < pr.len) {
                pr = r;
                pe = e;
                pj = j;
            }
        }
    }
    const uint64_t cand_p = P->coff[pe];
    const uint32_t *probe = ce_val + pr.lo;
    if (nlists == 1) {
        if (lg == 0) {
            bytes += 16 + 8ull * S.ns;
            return pr.len - same_hits(S, P, pe, cand, probe, pr.len);
        }
        return 0;
    }

    uint64_t c = 0;
    for (uint32_t t = lg; t < pr.len; t += G) {
        const uint32_t x = probe[t];
        const uint32_t xid = cand[cand_p + x];
        bool ok = injective(S, cand, xid);
        for (uint32_t i = 0; i < P->gsize[g] && ok; ++i) {
            if (!(mask >> i & 1)) continue;
            const uint32_t e = P->gleaf[g][i];
            uint32_t pos = x;
            if (e != pe) {
                pos = cand_pos(P, e, xid);
                ok = pos != 0xffffffffu;
            }
            for (uint32_t j = 0; j < P->nb[e] && ok; ++j) {
                if (e == pe && j == pj) continue;
                ListRef r = list_of(P, ce_off, row, e, j);
                ok = contains_u32(ce_val + r.lo, r.len, pos);
            }
        }
        c += ok;
    }
    bytes += 4ull * ((pr.len + G - 1 - lg) / G);
    return group_sum<G>(c);
}

template <int G>
__device__ inline uint64_t group_count(const DPlan *__restrict__ P, uint32_t g, const uint32_t *__restrict__ row,
                                       const uint64_t *__restrict__ ce_off, const uint32_t *__restrict__ ce_val,
                                       const uint32_t *__restrict__ cand, uint64_t &bytes) {
    SameLabel S;
    load_same(S, P, cand, row, P->gleaf[g][0]);
    const uint32_t k = P->gsize[g];
    if (k == 1) return leaf_set_count<G>(P, g, 1u, row, ce_off, ce_val, cand, S, bytes);
    // intersection sizes of every nonempty subset of the group, then the partition sum.  The sizes
    // sit in LDS, one row per lane group of the 256-thread block: the partition blocks index them at
    // run time, which in registers would hold 2^kMaxGroup 64-bit values per lane.  Each size is at
    // most one list's length (32 bits).  The group leader's value is the group's (a single-list
    // subset is counted by the leader alone), so the leader writes and the group reads.
    __shared__ uint32_t sizes_lds[256 / G][1u << kMaxGroup];
    uint32_t *cs = sizes_lds[threadIdx.x / G];
    const bool leader = (threadIdx.x & (G - 1)) == 0;
    for (uint32_t m = 1; m < (1u << k); ++m) {
        const uint64_t v = leaf_set_count<G>(P, g, m, row, ce_off, ce_val, cand, S, bytes);
        if (leader) cs[m] = (uint32_t)v;
    }
    // the group's lanes are in one wave: order the leader's writes before their reads
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    uint64_t tot = 0;
    for (uint32_t p = 0; p < P->nparts[k]; ++p) {
        uint64_t term = (uint64_t)P->pcoef[k][p];
        for (uint32_t b = 0; b < P->pnblk[k][p]; ++b) term *= cs[P->pblk[k][p][b]];
        tot += term;
    }
    return tot;
}

// Product of the leaf-group counts of a row of width D (group-uniform value).
template <int G>
__device__ inline uint64_t leaf_row(const DPlan *__restrict__ P, const uint32_t *__restrict__ row,
                                    const uint64_t *__restrict__ ce_off, const uint32_t *__restrict__ ce_val,
                                    const uint32_t *__restrict__ cand, uint64_t &bytes) {
    uint64_t prod = 1;
    for (uint32_t g = 0; g < P->ngroups; ++g) {
        uint64_t c = group_count<G>(P, g, row, ce_off, ce_val, cand, bytes);
        c = __shfl(c, (int)((threadIdx.x & 63) & ~(G - 1)), 64);  // broadcast the group leader's value
        prod *= c;
        if (prod == 0) break;
    }
    return prod;
}

// Rows of width D -> sum over rows of the product of the leaf-group counts (mod 2^64).  The
// lookups are dependent binary searches, so the kernel runs at eight waves per SIMD (64 VGPRs,
// 48 B per lane spilled): config-2 queries/s 1.23 at four waves (105 VGPRs, no spills), 1.49 at
// six, 1.61 at seven, 1.64 at eight with the subset sizes of group_count in registers (about
// 200 B spilled), 1.87 with them in LDS.
template <int G>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8, 8))) k_leaf(const DPlan *__restrict__ P, const uint32_t *__restrict__ rows,
                                              uint64_t nrows, const uint64_t *__restrict__ ce_off,
                                              const uint32_t *__restrict__ ce_val, const uint32_t *__restrict__ cand,
                                              unsigned long long *__restrict__ total,
                                              unsigned long long *__restrict__ bytes_ctr) {
    const int lg = threadIdx.x & (G - 1);
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x / G;
    const uint32_t D = P->D;
    uint64_t acc = 0, bytes = 0;
    for (uint64_t it = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / G; it < nrows; it += stride) {
        const uint64_t prod = leaf_row<G>(P, rows + it * D, ce_off, ce_val, cand, bytes);
        if (lg == 0) {
            acc += prod;
            bytes += 4ull * D;
        }
    }
    acc = group_sum<64>(acc);
    bytes = group_sum<64>(bytes);
    if ((threadIdx.x & 63) == 0) {
        if (acc) atomicAdd(total, (unsigned long long)acc);
        if (bytes_ctr) atomicAdd(bytes_ctr, (unsigned long long)bytes);
    }
}

// LOCAL_COUNT leaves (every leaf group has one leaf), aggregated without atomics on the matrix.
// Per row, each matched vertex receives the number of completions and each leaf candidate x the
// product of the other leaves' counts.  A row's contributions are (cell, value) pairs: each
// matched vertex v_j gets (v_j * |Vq| + u_j, product of the leaf-group counts) and each valid
// candidate x of leaf e gets (x * |Vq| + u_e, product of the other groups' counts).  The count pass
// (FILL = false) writes the number of pairs of every row; after an exclusive scan the fill pass
// (FILL = true) writes them at the row's offset (ordered ballot compaction within the lane group).
// The pairs are then aggregated by owner bucket (local_agg.hip): every cell receives one plain add.
template <int G, bool FILL>
__global__ void __launch_bounds__(256) k_leaf_local_pairs(const DPlan *__restrict__ P, const uint32_t *__restrict__ rows,
                                                          uint64_t nrows, const uint64_t *__restrict__ ce_off,
                                                          const uint32_t *__restrict__ ce_val,
                                                          const uint32_t *__restrict__ cand,
                                                          uint32_t *__restrict__ npairs, const uint64_t *__restrict__ poff,
                                                          uint64_t pbase, uint64_t *__restrict__ keys,
                                                          uint64_t *__restrict__ vals,
                                                          unsigned long long *__restrict__ total,
                                                          uint32_t *__restrict__ overflow) {
    const int lane = threadIdx.x & 63;
    const int lg = lane & (G - 1);
    const int leader = lane & ~(G - 1);
    const uint64_t below = lg ? ((1ull << lg) - 1) : 0ull;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x / G;
    const uint32_t D = P->D, ng = P->ngroups, nq = P->nq;
    uint64_t acc = 0, dummy = 0;
    for (uint64_t it = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / G; it < nrows; it += stride) {
        const uint32_t *row = rows + it * D;
        uint64_t cnt[kMaxQ];
        uint64_t prod = 1;
        for (uint32_t g = 0; g < ng; ++g) {
            uint64_t c = group_count<G>(P, g, row, ce_off, ce_val, cand, dummy);
            c = __shfl(c, leader, 64);
            cnt[g] = c;
            prod *= c;
        }
        if constexpr (!FILL) {
            if (lg == 0) {
                uint64_t np = 0;
                if (prod) {
                    np = D;
                    for (uint32_t g = 0; g < ng; ++g) np += cnt[g];
                }
                if (np > 0xffffffffull) {
                    *overflow = 1;
                    np = 0;
                }
                npairs[it] = (uint32_t)np;
                acc += prod;
            }
            continue;
        }
        if (prod == 0) continue;
        uint64_t o = poff[it] - pbase;
        for (uint32_t j = lg; j < D; j += G) {
            keys[o + j] = (uint64_t)cand[P->coff[j] + row[j]] * nq + P->u[j];
            vals[o + j] = prod;
        }
        o += D;
        for (uint32_t g = 0; g < ng; ++g) {
            uint64_t others = 1;
            for (uint32_t h = 0; h < ng; ++h)
                if (h != g) others *= cnt[h];
            const uint32_t e = P->gleaf[g][0];
            SameLabel S;
            load_same(S, P, cand, row, e);
            uint32_t pj = 0;
            ListRef pr = list_of(P, ce_off, row, e, 0);
            for (uint32_t j = 1; j < P->nb[e]; ++j) {
                ListRef r = list_of(P, ce_off, row, e, j);
                if (r.len < pr.len) {
                    pr = r;
                    pj = j;
                }
            }
            for (uint32_t t0 = 0; t0 < pr.len; t0 += G) {
                const uint32_t t = t0 + lg;
                bool ok = t < pr.len;
                uint32_t xid = 0;
                if (ok) {
                    const uint32_t x = ce_val[pr.lo + t];
                    for (uint32_t j = 0; j < P->nb[e] && ok; ++j) {
                        if (j == pj) continue;
                        ListRef r = list_of(P, ce_off, row, e, j);
                        ok = contains_u32(ce_val + r.lo, r.len, x);
                    }
                    xid = cand[P->coff[e] + x];
                    if (ok) ok = injective(S, cand, xid);
                }
                const uint64_t m = group_bits<G>(__ballot(ok));
                if (ok) {
                    const uint64_t k = o + (uint64_t)__builtin_popcountll(m & below);
                    keys[k] = (uint64_t)xid * nq + P->u[e];
                    vals[k] = others;
                }
                o += (uint64_t)__builtin_popcountll(m);
            }
        }
    }
    if constexpr (!FILL) {
        acc = group_sum<64>(acc);
        if (lane == 0 && acc) atomicAdd(total, (unsigned long long)acc);
    }
}

// -------------------------------------------------------------------------- fused last level
// Per-lane leaf evaluation for a row of depth D-1 extended by candidate x at depth D-1: the value
// matched at depth j is (j == dx ? x : row[j]).
struct RowView {
    const uint32_t *row;
    uint32_t dx, x;
    __device__ inline uint32_t at(uint32_t j) const { return j == dx ? x : row[j]; }
};

__device__ inline ListRef list_of_v(const DPlan *__restrict__ P, const uint64_t *__restrict__ ce_off, const RowView &R,
                                    uint32_t d, uint32_t j) {
    const uint64_t base = P->bbase[d][j] + R.at(P->bdep[d][j]);
    const uint64_t lo = ce_off[base];
    return {lo, (uint32_t)(ce_off[base + 1] - lo)};
}

// Data id matched at depth dep.
__device__ inline uint32_t matched_id(const DPlan *__restrict__ P, const uint32_t *__restrict__ cand, const RowView &R,
                                      uint32_t dep) {
    return cand[P->coff[dep] + R.at(dep)];
}

// One lane: |∩_{leaf e in mask of group g} S_e|.
__device__ inline uint64_t leaf_set_count_lane(const DPlan *__restrict__ P, uint32_t g, uint32_t mask, const RowView &R,
                                               const uint64_t *__restrict__ ce_off, const uint32_t *__restrict__ ce_val,
                                               const uint32_t *__restrict__ cand, const uint32_t *y, uint32_t ns) {
    uint32_t pe = 0, pj = 0, nlists = 0;
    ListRef pr{0, 0xffffffffu};
    for (uint32_t i = 0; i < P->gsize[g]; ++i) {
        if (!(mask >> i & 1)) continue;
        const uint32_t e = P->gleaf[g][i];
        for (uint32_t j = 0; j < P->nb[e]; ++j) {
            ListRef r = list_of_v(P, ce_off, R, e, j);
            ++nlists;
            if (r.len < pr.len) {
                pr = r;
                pe = e;
                pj = j;
            }
        }
    }
    const uint64_t cand_p = P->coff[pe];
    const uint32_t *probe = ce_val + pr.lo;
    if (mask == 1u && P->tri[P->gleaf[g][0]]) {
        const uint32_t e = P->gleaf[g][0];
        return tri_count(P, e, R.at(P->bdep[e][0]), R.at(P->bdep[e][1]), list_of_v(P, ce_off, R, e, 0),
                         list_of_v(P, ce_off, R, e, 1), ce_off, ce_val, cand, y, ns);
    }
    if (nlists == 1) {
        uint32_t h = 0;
        for (uint32_t s = 0; s < ns; ++s) {
            const uint32_t x = cand_pos(P, pe, y[s]);
            if (x != 0xffffffffu && contains_u32(probe, pr.len, x)) ++h;
        }
        return pr.len - h;
    }
    uint64_t c = 0;
    for (uint32_t t = 0; t < pr.len; ++t) {
        const uint32_t x = probe[t];
        const uint32_t xid = cand[cand_p + x];
        bool ok = true;
        for (uint32_t s = 0; s < ns && ok; ++s) ok = xid != y[s];
        for (uint32_t i = 0; i < P->gsize[g] && ok; ++i) {
            if (!(mask >> i & 1)) continue;
            const uint32_t e = P->gleaf[g][i];
            uint32_t pos = x;
            if (e != pe) {
                pos = cand_pos(P, e, xid);
                ok = pos != 0xffffffffu;
            }
            for (uint32_t j = 0; j < P->nb[e] && ok; ++j) {
                if (e == pe && j == pj) continue;
                ListRef r = list_of_v(P, ce_off, R, e, j);
                ok = contains_u32(ce_val + r.lo, r.len, pos);
            }
        }
        c += ok;
    }
    return c;
}

// One lane: completions of a single-leaf group of a fusable plan (its leaf has one list, or two
// forming a triangle, and at most kFusedSame same-label matched vertices).
constexpr uint32_t kFusedSame = 8;

__device__ inline uint64_t single_leaf_lane(const DPlan *__restrict__ P, uint32_t g, const RowView &R,
                                            const uint64_t *__restrict__ ce_off, const uint32_t *__restrict__ ce_val,
                                            const uint32_t *__restrict__ cand) {
    const uint32_t e = P->gleaf[g][0];
    const uint32_t ns = P->ns[e];
    uint32_t y[kFusedSame];
#pragma unroll
    for (uint32_t s = 0; s < kFusedSame; ++s) y[s] = s < ns ? matched_id(P, cand, R, P->sdep[e][s]) : 0xffffffffu;
    return leaf_set_count_lane(P, g, 1u, R, ce_off, ce_val, cand, y, ns);
}

// Up to kRowLeaves leaf groups whose leaf has no matched neighbour at depth D-1 are counted once per
// row instead of once per candidate x: their lists do not depend on x, and x can only matter to
// them as a same-label vertex, which one membership test per candidate corrects.
constexpr uint32_t kRowLeaves = 4;

// Is data vertex xid in the candidate list set S_e of the single-leaf group g?
__device__ inline bool in_leaf_lists(const DPlan *__restrict__ P, uint32_t g, const RowView &R,
                                     const uint64_t *__restrict__ ce_off, const uint32_t *__restrict__ ce_val,
                                     uint32_t xid) {
    const uint32_t e = P->gleaf[g][0];
    const uint32_t pos = cand_pos(P, e, xid);
    bool in = pos != 0xffffffffu;
    for (uint32_t j = 0; j < P->nb[e] && in; ++j) {
        const ListRef r = list_of_v(P, ce_off, R, e, j);
        in = contains_u32(ce_val + r.lo, r.len, pos);
    }
    return in;
}

// The leaf groups of a fused plan that are counted once per row: those whose leaf has no matched
// neighbour at the last enumerated depth d (plan-level, the same for every row).
struct HoistedGroups {
    uint32_t g[kRowLeaves];
    uint32_t n, mask, same;  // count, bit per group, bit k: depth d shares the label of group g[k]
};

__device__ inline HoistedGroups hoisted_groups(const DPlan *__restrict__ P, uint32_t d) {
    HoistedGroups H{};
    for (uint32_t g = 0; g < P->ngroups && H.n < kRowLeaves; ++g) {
        const uint32_t e = P->gleaf[g][0];
        bool adj = false, same_d = false;
        for (uint32_t j = 0; j < P->nb[e]; ++j) adj |= P->bdep[e][j] == d;
        if (adj) continue;
        for (uint32_t s2 = 0; s2 < P->ns[e]; ++s2) same_d |= P->sdep[e][s2] == d;
        H.g[H.n] = g;
        H.same |= (same_d ? 1u : 0u) << H.n;
        H.mask |= 1u << g;
        ++H.n;
    }
    return H;
}

// Per-row state of a fused row: the probe list at depth d and the hoisted groups' counts, with
// depth d left out of their same-label sets.
struct FusedRow {
    ListRef pr;
    uint32_t pj;
    uint64_t hc[kRowLeaves];  // entries beyond the hoisted count stay 0
};

__device__ inline void fused_setup(const DPlan *__restrict__ P, const uint32_t *__restrict__ row, uint32_t d,
                                   const HoistedGroups &H, const uint64_t *__restrict__ ce_off,
                                   const uint32_t *__restrict__ ce_val, const uint32_t *__restrict__ cand,
                                   FusedRow &F) {
    const uint32_t nb = P->nb[d];
    F.pj = 0;
    F.pr = list_of(P, ce_off, row, d, 0);
    for (uint32_t j = 1; j < nb; ++j) {
        const ListRef r = list_of(P, ce_off, row, d, j);
        if (r.len < F.pr.len) {
            F.pr = r;
            F.pj = j;
        }
    }
    const RowView R0{row, d, 0};  // depth d is read by none of the hoisted leaves' lists
    for (uint32_t k = 0; k < H.n; ++k) {
        const uint32_t g = H.g[k], e = P->gleaf[g][0];
        uint32_t y[kFusedSame], ns = 0;
        for (uint32_t s2 = 0; s2 < P->ns[e]; ++s2) {
            const uint32_t dep = P->sdep[e][s2];
            if (dep != d) y[ns++] = matched_id(P, cand, R0, dep);
        }
        F.hc[k] = F.pr.len ? leaf_set_count_lane(P, g, 1u, R0, ce_off, ce_val, cand, y, ns) : 0;
    }
}

// Leaf product of (row, x) for candidate x (a position in C(u_d) from the probe list), or 0 when x
// is not a valid extension (missing from another backward list, or a same-label matched vertex).
__device__ inline uint64_t fused_candidate(const DPlan *__restrict__ P, const uint32_t *__restrict__ row, uint32_t d,
                                           uint32_t pj, const uint64_t *hc, const HoistedGroups &H,
                                           const SameLabel &S, uint32_t x, const uint64_t *__restrict__ ce_off,
                                           const uint32_t *__restrict__ ce_val, const uint32_t *__restrict__ cand,
                                           uint64_t &bytes) {
    bool ok = true;
    for (uint32_t j = 0; j < P->nb[d] && ok; ++j) {
        if (j == pj) continue;
        const ListRef r = list_of(P, ce_off, row, d, j);
        ok = contains_u32(ce_val + r.lo, r.len, x);
    }
    const uint32_t xid = cand[P->coff[d] + x];
    if (ok) ok = injective(S, cand, xid);
    if (!ok) return 0;
    const RowView R{row, d, x};
    uint64_t prod = 1;
#pragma unroll
    for (uint32_t k = 0; k < kRowLeaves; ++k) {
        if (k >= H.n || !prod) continue;
        uint64_t c = hc[k];
        if ((H.same >> k & 1) && c && in_leaf_lists(P, H.g[k], R, ce_off, ce_val, xid)) --c;
        prod *= c;
    }
    for (uint32_t g = 0; g < P->ngroups && prod; ++g)
        if (!(H.mask >> g & 1)) prod *= single_leaf_lane(P, g, R, ce_off, ce_val, cand);
    bytes += 4ull + 16ull * P->nadj_last;
    return prod;
}

// Leaves of a fused level with a row-sum table (DPlan::tail): each leaf adjacent to depth d = D-1
// is a pendant leaf of d (count: its list length at x) or a triangle leaf on (tail_dep, d)
// (count: its triangle-table entry at the candidate edge (p, x)).
constexpr uint32_t kTailLeaves = 8;
struct TailSpec {
    uint64_t arc_base;  // list offsets of the arc (u_tail_dep -> u_d) start at ce_off[arc_base]
    uint32_t np;        // |C(u_tail_dep)|
    uint32_t nadj;      // adjacent leaves
    uint64_t pend_base[kTailLeaves];       // pendant leaf: list offsets of the arc (u_d -> u_e)
    const uint32_t *tri_tab[kTailLeaves];  // triangle leaf: its table (nullptr for a pendant leaf)
    const uint64_t *inner;                 // second table: the entry of x is inner[x] (no leaves)
};

// tail[p] for every position p of C(u_tail_dep): a wave per position, lanes split its list.  With
// T.inner set (the second table), the entry of a candidate x is inner[x].
// One table entry, by the whole wave: the sum over position p's list of the product of the leaves'
// counts (times inner[x] for a second table).
__device__ inline void tail_sum_at(const TailSpec &T, uint64_t p, const uint64_t *__restrict__ ce_off,
                                   const uint32_t *__restrict__ ce_val, uint64_t *__restrict__ tail) {
    const int lane = threadIdx.x & 63;
    const uint64_t e_first = ce_off[T.arc_base];  // the arc's first candidate edge (triangle-table index 0)
    const uint64_t lo = ce_off[T.arc_base + p], hi = ce_off[T.arc_base + p + 1];
    uint64_t s = 0;
    for (uint64_t i = lo + lane; i < hi; i += 64) {
        const uint32_t x = ce_val[i];
        uint64_t prod = T.inner ? T.inner[x] : 1;
        for (uint32_t k = 0; k < T.nadj; ++k)
            prod *= T.tri_tab[k] ? (uint64_t)T.tri_tab[k][i - e_first]
                                 : ce_off[T.pend_base[k] + x + 1] - ce_off[T.pend_base[k] + x];
        s += prod;
    }
    s = group_sum<64>(s);
    if (lane == 0) tail[p] = s;
}

__global__ void __launch_bounds__(256) k_tail_sums(TailSpec T, const uint64_t *__restrict__ ce_off,
                                                   const uint32_t *__restrict__ ce_val, uint64_t *__restrict__ tail) {
    const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    for (uint64_t p = wave; p < T.np; p += nwaves) tail_sum_at(T, p, ce_off, ce_val, tail);
}

// The tables of many queries in one launch (batch engine): job j covers waves [pref[j], pref[j+1]),
// one per position of its table.
struct TailJob {
    TailSpec T;
    const uint64_t *ce_off;
    const uint32_t *ce_val;
    uint64_t *out;
};

__global__ void __launch_bounds__(256) k_tail_sums_multi(const TailJob *__restrict__ jobs,
                                                         const uint64_t *__restrict__ pref, uint32_t njobs) {
    const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    const uint64_t total = pref[njobs];
    for (uint64_t w = wave; w < total; w += nwaves) {
        uint32_t lo = 0, hi = njobs;  // last job with pref[j] <= w (wave-uniform)
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (pref[mid] <= w) lo = mid;
            else hi = mid;
        }
        const TailJob &J = jobs[lo];
        tail_sum_at(J.T, w - pref[lo], J.ce_off, J.ce_val, J.out);
    }
}

constexpr uint32_t kTailFixLeaf = 1, kTailFixHoist = 2;

// f0(x): the product of the adjacent leaves' list counts at x = L[t] (i = L.lo + t), the value the
// row-sum table holds for x; the leaf at depth skip (if any) is left out.
__device__ inline uint64_t tail_clean(const DPlan *__restrict__ P, uint32_t x, uint64_t i, const HoistedGroups &H,
                                      const uint64_t *__restrict__ ce_off, uint32_t skip = 0xffffffffu) {
    uint64_t prod = 1;
    for (uint32_t g = 0; g < P->ngroups; ++g) {
        if (H.mask >> g & 1) continue;
        const uint32_t e = P->gleaf[g][0];
        if (e == skip) continue;
        prod *= P->tri[e] ? (uint64_t)P->tri_tab[e][i - ce_off[P->tri_base[e]]]
                          : ce_off[P->bbase[e][0] + x + 1] - ce_off[P->bbase[e][0] + x];
    }
    return prod;
}

// One row of a fused level with a row-sum table: H0 * tail[row[b]] (H0: the product of the row's
// hoisted counts), less H0 * f0(x) for each candidate x that is the matched vertex of a
// same-label depth (not an injective extension; there are at most ns[d] of them, and the table
// counted them).  Exact modulo 2^64, as the per-candidate sum is.  For the correction factor's
// pass it also returns the row's list L from depth tail_dep and coef: H0, or for a hoisted
// correction leaf the product of the other hoisted counts.
__device__ inline uint64_t tail_row(const DPlan *__restrict__ P, const uint32_t *__restrict__ row, uint32_t d,
                                    const HoistedGroups &H, const uint64_t *__restrict__ ce_off,
                                    const uint32_t *__restrict__ ce_val, const uint32_t *__restrict__ cand,
                                    uint64_t &coef, ListRef &L, uint64_t &bytes) {
    uint64_t s0 = P->tail[row[P->tail_dep]];
    uint64_t H0 = 1, Hx = 1;
    const RowView R0{row, d, 0};
    const bool need = s0 || P->ns[d] || P->tail_fix;
    for (uint32_t k = 0; k < H.n && need; ++k) {
        const uint32_t g = H.g[k], e = P->gleaf[g][0];
        uint32_t y[kFusedSame], ns = 0;
        for (uint32_t s2 = 0; s2 < P->ns[e]; ++s2) {
            const uint32_t dep = P->sdep[e][s2];
            if (dep != d) y[ns++] = matched_id(P, cand, R0, dep);
        }
        const uint64_t c = leaf_set_count_lane(P, g, 1u, R0, ce_off, ce_val, cand, y, ns);
        H0 *= c;
        if (!(P->tail_fix == kTailFixHoist && e == P->tail_leaf)) Hx *= c;
    }
    coef = P->tail_fix == kTailFixHoist ? Hx : H0;
    bytes += 4ull * d + 8ull + 16ull * H.n;
    if (!P->ns[d] && !P->tail_fix) return s0 * H0;
    L = list_of(P, ce_off, row, d, 0);
    const uint32_t *Lv = ce_val + L.lo;
    for (uint32_t s2 = 0; s2 < P->ns[d]; ++s2) {
        const uint32_t x = cand_pos(P, d, matched_id(P, cand, R0, P->sdep[d][s2]));
        if (x == 0xffffffffu) continue;
        const uint32_t t = lower_bound_u32(Lv, L.len, x);
        if (t < L.len && Lv[t] == x) s0 -= tail_clean(P, x, L.lo + t, H, ce_off);
    }
    return s0 * H0;
}

// Edge table (DPlan::etail) of a fused level with two adjacent backward neighbours b1 < b2 of d:
// a wave per candidate edge (p1, p2) of the arc (u_b1 -> u_b2); lanes split the shorter of
// L1 = L(b1 -> d)(p1) and L2 = L(b2 -> d)(p2) and binary-search the other.  A triangle leaf on
// (b_j, d) reads its table at x's candidate edge in L_j.
struct EdgeTailSpec {
    uint64_t arc_base;   // list offsets of the arc (u_b1 -> u_b2)
    uint32_t np1;        // |C(u_b1)|
    uint64_t nedges;     // candidate edges of the arc
    uint64_t l1_base, l2_base;  // list offsets of the arcs (u_b1 -> u_d) and (u_b2 -> u_d)
    uint32_t nadj;
    uint64_t pend_base[kTailLeaves];
    const uint32_t *tri_tab[kTailLeaves];
    uint32_t tri_list[kTailLeaves];  // triangle leaf on (b1, d): 0; on (b2, d): 1
};

__global__ void __launch_bounds__(256) k_edge_tail_sums(EdgeTailSpec T, const uint64_t *__restrict__ ce_off,
                                                        const uint32_t *__restrict__ ce_val, uint64_t *__restrict__ tab) {
    const int lane = threadIdx.x & 63;
    const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    const uint64_t e_first = ce_off[T.arc_base];
    const uint64_t f1 = ce_off[T.l1_base], f2 = ce_off[T.l2_base];  // first candidate edges of L1's and L2's arcs
    for (uint64_t k = wave; k < T.nedges; k += nwaves) {
        uint32_t lo = 0, hi = T.np1;  // p1: the last position whose list starts at or before edge k
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (ce_off[T.arc_base + mid] <= e_first + k) lo = mid;
            else hi = mid;
        }
        const uint32_t p1 = lo, p2 = ce_val[e_first + k];
        const uint64_t a0 = ce_off[T.l1_base + p1], a1 = ce_off[T.l1_base + p1 + 1];
        const uint64_t b0 = ce_off[T.l2_base + p2], b1 = ce_off[T.l2_base + p2 + 1];
        const bool scan1 = a1 - a0 <= b1 - b0;
        const uint64_t s0 = scan1 ? a0 : b0, s1 = scan1 ? a1 : b1, o0 = scan1 ? b0 : a0;
        const uint32_t olen = (uint32_t)(scan1 ? b1 - b0 : a1 - a0);
        uint64_t s = 0;
        for (uint64_t i = s0 + lane; i < s1; i += 64) {
            const uint32_t x = ce_val[i];
            const uint32_t t = lower_bound_u32(ce_val + o0, olen, x);
            if (t >= olen || ce_val[o0 + t] != x) continue;
            const uint64_t i1 = scan1 ? i : o0 + t, i2 = scan1 ? o0 + t : i;  // x's global index in L1, L2
            uint64_t prod = 1;
            for (uint32_t j = 0; j < T.nadj; ++j)
                prod *= T.tri_tab[j] ? (uint64_t)T.tri_tab[j][T.tri_list[j] ? i2 - f2 : i1 - f1]
                                     : ce_off[T.pend_base[j] + x + 1] - ce_off[T.pend_base[j] + x];
            s += prod;
        }
        s = group_sum<64>(s);
        if (lane == 0) tab[k] = s;
    }
}

// A row of width d = D-1 of a plan with an edge table: H0 * etail[edge (row[b1], row[b2])], less
// H0 * f0(x) for each x in L1 ∩ L2 that is the matched vertex of a same-label depth of d.
__device__ inline uint64_t etail_row(const DPlan *__restrict__ P, const uint32_t *__restrict__ row, uint32_t d,
                                     const HoistedGroups &H, const uint64_t *__restrict__ ce_off,
                                     const uint32_t *__restrict__ ce_val, const uint32_t *__restrict__ cand,
                                     uint64_t &bytes) {
    const uint32_t b1 = P->bdep[d][0], b2 = P->bdep[d][1];
    uint64_t ab = 0;  // list offsets of the arc (u_b1 -> u_b2)
    for (uint32_t j = 0; j < P->nb[b2]; ++j)
        if (P->bdep[b2][j] == b1) ab = P->bbase[b2][j];
    const uint64_t lo = ce_off[ab + row[b1]];
    const uint32_t len = (uint32_t)(ce_off[ab + row[b1] + 1] - lo);
    const uint32_t k = lower_bound_u32(ce_val + lo, len, row[b2]);  // (p1, p2) is a candidate edge
    uint64_t s0 = P->etail[lo + k - ce_off[ab]];
    const RowView R0{row, d, 0};
    if (P->ns[d]) {
        const ListRef L1 = list_of(P, ce_off, row, d, 0), L2 = list_of(P, ce_off, row, d, 1);
        for (uint32_t s2 = 0; s2 < P->ns[d]; ++s2) {
            const uint32_t x = cand_pos(P, d, matched_id(P, cand, R0, P->sdep[d][s2]));
            if (x == 0xffffffffu) continue;
            const uint32_t t1 = lower_bound_u32(ce_val + L1.lo, L1.len, x);
            const uint32_t t2 = lower_bound_u32(ce_val + L2.lo, L2.len, x);
            if (t1 >= L1.len || ce_val[L1.lo + t1] != x || t2 >= L2.len || ce_val[L2.lo + t2] != x) continue;
            uint64_t f = 1;
            for (uint32_t g = 0; g < P->ngroups; ++g) {
                if (H.mask >> g & 1) continue;
                const uint32_t e = P->gleaf[g][0];
                if (P->tri[e]) {
                    const bool on1 = P->bdep[e][0] == b1;
                    f *= P->tri_tab[e][(on1 ? L1.lo + t1 : L2.lo + t2) - ce_off[P->tri_base[e]]];
                } else {
                    f *= ce_off[P->bbase[e][0] + x + 1] - ce_off[P->bbase[e][0] + x];
                }
            }
            s0 -= f;
        }
    }
    uint64_t H0 = 1;
    for (uint32_t k2 = 0; k2 < H.n && s0; ++k2) {
        const uint32_t g = H.g[k2], e = P->gleaf[g][0];
        uint32_t y[kFusedSame];
        const uint32_t ns = P->ns[e];  // d is not among them: no hoisted leaf has d's label
        for (uint32_t s2 = 0; s2 < ns; ++s2) y[s2] = matched_id(P, cand, R0, P->sdep[e][s2]);
        H0 *= leaf_set_count_lane(P, g, 1u, R0, ce_off, ce_val, cand, y, ns);
    }
    bytes += 4ull * d + 16ull + 8ull + 16ull * H.n;
    return s0 * H0;
}

// A row of width c = D-2 of a plan with a second table.  c's candidates x' depend on the row only
// through its vertex at tail2_dep, and each contributes inner(x') = (tail[x'] when the table's
// backward neighbour b is c) x (the list lengths at x' of the hoisted leaves pendant on c); tail2
// holds their sums.  The row's sum is H0 * (tail[row[b]] when b < c) * (tail2[row[tail2_dep]] less
// inner(x') at each x' that is the matched vertex of a same-label depth of c), H0 being the
// product of the other hoisted leaves' counts, which read depths below c only.
__device__ inline bool pendant_on(const DPlan *__restrict__ P, uint32_t e, uint32_t c) {
    return P->nb[e] == 1 && P->bdep[e][0] == c;
}

__device__ inline uint64_t tail2_row(const DPlan *__restrict__ P, const uint32_t *__restrict__ row, uint32_t c,
                                     const HoistedGroups &H, const uint64_t *__restrict__ ce_off,
                                     const uint32_t *__restrict__ ce_val, const uint32_t *__restrict__ cand,
                                     uint64_t &bytes) {
    const uint32_t b = P->tail_dep;
    uint64_t s0 = P->tail2[row[P->tail2_dep]];
    const RowView R0{row, c, 0};
    if (P->ns[c]) {
        const ListRef L = list_of(P, ce_off, row, c, 0);
        const uint32_t *Lv = ce_val + L.lo;
        for (uint32_t s2 = 0; s2 < P->ns[c]; ++s2) {
            const uint32_t x = cand_pos(P, c, matched_id(P, cand, R0, P->sdep[c][s2]));
            if (x == 0xffffffffu) continue;
            const uint32_t t = lower_bound_u32(Lv, L.len, x);
            if (t >= L.len || Lv[t] != x) continue;
            uint64_t f = b == c ? P->tail[x] : 1;
            for (uint32_t k = 0; k < H.n; ++k) {
                const uint32_t e = P->gleaf[H.g[k]][0];
                if (pendant_on(P, e, c)) f *= ce_off[P->bbase[e][0] + x + 1] - ce_off[P->bbase[e][0] + x];
            }
            s0 -= f;
        }
    }
    if (b < c) s0 *= P->tail[row[b]];
    uint64_t H0 = 1;
    for (uint32_t k = 0; k < H.n && s0; ++k) {
        const uint32_t g = H.g[k], e = P->gleaf[g][0];
        if (pendant_on(P, e, c)) continue;  // inside tail2
        uint32_t y[kFusedSame];
        const uint32_t ns = P->ns[e];
        for (uint32_t s2 = 0; s2 < ns; ++s2) y[s2] = matched_id(P, cand, R0, P->sdep[e][s2]);
        H0 *= leaf_set_count_lane(P, g, 1u, R0, ce_off, ce_val, cand, y, ns);
    }
    bytes += 4ull * c + 16ull + 16ull * H.n;
    return s0 * H0;
}

// Correction pass of a row-sum table with a correction factor (DPlan::tail_fix), flattened over
// the wave as in fused_tile.  For every row of the tile and every source of the factor (each
// same-label matched vertex y of the adjacent correction leaf, or the hoisted correction leaf),
// the injective candidates x of the row's list L that lie in the source's set -- the candidates
// of d adjacent to y (the reverse arc's list at y), or the hoisted leaf's lists -- each take
// coef * (product of the other adjacent leaves' counts at x) off the row's sum.  Lane k holds
// row k's scan: the shorter of L and the set (a hoisted leaf's first list) is scanned and the
// other probed (with a hoisted leaf's other lists).  Returns this
// lane's share, for rows of any lane.
__device__ inline uint64_t tail_fix_tile(const DPlan *__restrict__ P, const uint32_t *__restrict__ rows, uint64_t r0,
                                         uint32_t d, bool mine, const uint32_t *__restrict__ row, uint64_t coef,
                                         ListRef L, const HoistedGroups &H, const uint64_t *__restrict__ ce_off,
                                         const uint32_t *__restrict__ ce_val, const uint32_t *__restrict__ cand,
                                         uint64_t &bytes) {
    const int lane = threadIdx.x & 63;
    const uint32_t fix = P->tail_fix, fl = P->tail_leaf;
    const uint32_t nsrc = fix == kTailFixLeaf ? P->ns[fl] : 1u;
    const uint32_t skip = fix == kTailFixLeaf ? fl : 0xffffffffu;
    uint64_t corr = 0;
    for (uint32_t i = 0; i < nsrc; ++i) {
        ListRef Sr{0, 0};
        if (mine && L.len && coef) {
            const RowView R0{row, d, 0};
            if (fix == kTailFixLeaf) {
                const uint32_t q = cand_pos(P, fl, matched_id(P, cand, R0, P->sdep[fl][i]));
                bool ok = q != 0xffffffffu;
                for (uint32_t j = 0; j < P->nb[fl] && ok; ++j) {
                    if (P->bdep[fl][j] == d) continue;  // the other list of a triangle leaf
                    const ListRef r = list_of_v(P, ce_off, R0, fl, j);
                    ok = contains_u32(ce_val + r.lo, r.len, q);
                }
                if (ok) {
                    Sr.lo = ce_off[P->tail_rbase + q];
                    Sr.len = (uint32_t)(ce_off[P->tail_rbase + q + 1] - Sr.lo);
                }
            } else {
                Sr = list_of_v(P, ce_off, R0, fl, 0);
            }
            bytes += 16ull;
        }
        const bool scan_l = L.len <= Sr.len;
        const uint64_t n = Sr.len ? (scan_l ? L.len : Sr.len) : 0;
        const uint64_t inc = wave_inclusive_scan(n);
        const uint64_t excl = inc - n;
        const uint64_t total = __shfl(inc, 63, 64);
        for (uint64_t base = 0; base < total; base += 64) {
            const uint64_t c = base + lane;
            int k = 0;  // the row holding work item c (as in fused_tile, six steps on every lane)
#pragma unroll
            for (int step = 32; step; step >>= 1) {
                const uint64_t v = __shfl(excl, (k + step) & 63, 64);
                if (k + step < 64 && v <= c) k += step;
            }
            const uint64_t off = __shfl(excl, k, 64);
            const uint64_t llo = __shfl(L.lo, k, 64), slo = __shfl(Sr.lo, k, 64), ck = __shfl(coef, k, 64);
            const uint32_t llen = __shfl(L.len, k, 64), slen = __shfl(Sr.len, k, 64);
            const int sl = __shfl((int)scan_l, k, 64);
            if (c >= total) continue;
            const uint32_t t = (uint32_t)(c - off);
            const uint32_t *Lv = ce_val + llo;
            uint32_t x = 0xffffffffu, idx = 0;
            bool member = false;
            const uint32_t *rk = rows + (r0 + (uint32_t)k) * d;
            uint32_t q = 0xffffffffu;  // hoisted correction leaf: x's position in C(u_fl)
            if (sl) {
                idx = t;
                x = Lv[t];
                if (fix == kTailFixLeaf) {
                    member = contains_u32(ce_val + slo, slen, x);
                } else {
                    q = cand_pos(P, fl, cand[P->coff[d] + x]);
                    member = q != 0xffffffffu && contains_u32(ce_val + slo, slen, q);
                }
            } else {
                const uint32_t e = ce_val[slo + t];
                if (fix == kTailFixLeaf) {
                    x = e;
                } else {
                    q = e;
                    x = cand_pos(P, d, cand[P->coff[fl] + e]);
                }
                if (x != 0xffffffffu) {
                    idx = lower_bound_u32(Lv, llen, x);
                    member = idx < llen && Lv[idx] == x;
                }
            }
            if (fix == kTailFixHoist) {  // the hoisted leaf's other lists (a triangle leaf's second)
                const RowView Rk{rk, d, 0};
                for (uint32_t j = 1; j < P->nb[fl] && member; ++j) {
                    const ListRef r = list_of_v(P, ce_off, Rk, fl, j);
                    member = contains_u32(ce_val + r.lo, r.len, q);
                }
            }
            bytes += 8ull;
            if (!member) continue;
            const uint32_t xid = cand[P->coff[d] + x];
            bool valid = true;
            for (uint32_t s2 = 0; s2 < P->ns[d]; ++s2) {
                const uint32_t dep = P->sdep[d][s2];
                valid &= xid != cand[P->coff[dep] + rk[dep]];
            }
            if (valid) corr += ck * tail_clean(P, x, llo + idx, H, ce_off, skip);
        }
    }
    return corr;
}

// Fused rows of one query tile (at most 64 rows, one per lane), flattened: lane k sets up row k,
// a wave scan lays the rows' probe lists end to end, and the wave then takes 64 consecutive
// candidates at a time whatever rows they belong to, reading each row's state from its lane.  A
// tile of rows with short lists keeps every lane busy, and a hub row's list is split across the
// whole wave.
__device__ inline uint64_t fused_tile(const DPlan *__restrict__ P, const uint32_t *__restrict__ rows, uint64_t r0,
                                      uint32_t nrows, uint32_t d, const uint64_t *__restrict__ ce_off,
                                      const uint32_t *__restrict__ ce_val, const uint32_t *__restrict__ cand,
                                      uint32_t *__restrict__ counts, uint64_t &bytes) {
    const int lane = threadIdx.x & 63;
    const bool mine = (uint32_t)lane < nrows;
    const uint32_t *row = rows + (r0 + (mine ? (uint32_t)lane : 0u)) * d;
    if (P->tail2 && d + 2 == P->D) {  // rows of width D-2 finished from the second table
        uint64_t v = 0;
        if (mine) {
            v = tail2_row(P, row, d, hoisted_groups(P, d + 1), ce_off, ce_val, cand, bytes);
            if (counts) counts[r0 + lane] = 0;
        }
        return v;
    }
    const HoistedGroups H = hoisted_groups(P, d);
    if (P->etail) {
        uint64_t v = 0;
        if (mine) {
            v = etail_row(P, row, d, H, ce_off, ce_val, cand, bytes);
            if (counts) counts[r0 + lane] = 0;
        }
        return v;
    }
    if (P->tail) {
        uint64_t v = 0, coef = 0;
        ListRef L{0, 0};
        if (mine) {
            v = tail_row(P, row, d, H, ce_off, ce_val, cand, coef, L, bytes);
            if (counts) counts[r0 + lane] = 0;  // a fused row has no children
        }
        if (P->tail_fix) v -= tail_fix_tile(P, rows, r0, d, mine, row, coef, L, H, ce_off, ce_val, cand, bytes);
        return v;
    }
    FusedRow F{};
    SameLabel S{};
    if (mine) {
        fused_setup(P, row, d, H, ce_off, ce_val, cand, F);
        load_same(S, P, cand, row, d);
        if (counts) counts[r0 + lane] = 0;  // a fused row has no children
        bytes += 4ull * d + 16ull * P->nb[d] + 16ull * (P->ngroups - P->nadj_last) + 4ull * F.pr.len;
    }
    const uint64_t plen = mine ? F.pr.len : 0;
    const uint64_t inc = wave_inclusive_scan(plen);
    const uint64_t excl = inc - plen;
    const uint64_t total = __shfl(inc, 63, 64);
    const uint32_t ns = P->ns[d] < 8 ? P->ns[d] : 8;
    uint64_t acc = 0;
    for (uint64_t base = 0; base < total; base += 64) {
        const uint64_t c = base + lane;
        // The row whose list holds candidate c: the last k < nrows with excl_k <= c.  Six steps
        // for every lane: a shuffle must run on the whole wave (one that reads an inactive lane
        // returns 0), so the search cannot stop early per lane.
        int k = 0;
#pragma unroll
        for (int step = 32; step; step >>= 1) {
            const uint64_t v = __shfl(excl, (k + step) & 63, 64);
            if ((uint32_t)(k + step) < nrows && v <= c) k += step;
        }
        const uint64_t plo = __shfl(F.pr.lo, k, 64);
        const uint64_t off = __shfl(excl, k, 64);
        const uint32_t pj = __shfl(F.pj, k, 64);
        uint64_t hc[kRowLeaves];
#pragma unroll
        for (uint32_t h = 0; h < kRowLeaves; ++h) hc[h] = __shfl(F.hc[h], k, 64);
        SameLabel Sk = S;
#pragma unroll
        for (uint32_t s2 = 0; s2 < 8; ++s2) {
            const uint32_t v = __shfl(S.y[s2], k, 64);
            Sk.y[s2] = s2 < ns ? v : 0xffffffffu;
        }
        if (c < total) {
            const uint32_t *rk = rows + (r0 + (uint32_t)k) * d;
            Sk.row = rk;
            Sk.ns = P->ns[d];
            Sk.sdep = P->sdep[d];
            Sk.coff = P->coff;
            acc += fused_candidate(P, rk, d, pj, hc, H, Sk, ce_val[plo + (c - off)], ce_off, ce_val, cand, bytes);
        }
    }
    return acc;
}

// Rows of width d = D-1 -> sum over their valid depth-(D-1) candidates x of the leaf product of
// (row, x), in flattened tiles of 64 rows (fused_tile), one wave per tile; the depth-D frontier is
// never written.  With a second table the rows have width d = D-2.  Five waves per SIMD, as
// k_bfused_flat.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5, 8))) k_leaf_fused(const DPlan *__restrict__ P, const uint32_t *__restrict__ rows,
                                                    uint64_t nrows, const uint64_t *__restrict__ ce_off,
                                                    const uint32_t *__restrict__ ce_val,
                                                    const uint32_t *__restrict__ cand,
                                                    unsigned long long *__restrict__ total,
                                                    unsigned long long *__restrict__ bytes_ctr, uint32_t d) {
    const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    const uint64_t ntiles = (nrows + 63) / 64;
    uint64_t acc = 0, bytes = 0;
    for (uint64_t t = wave; t < ntiles; t += nwaves) {
        const uint64_t r0 = t * 64;
        const uint32_t n = (uint32_t)min<uint64_t>(64, nrows - r0);
        acc += fused_tile(P, rows, r0, n, d, ce_off, ce_val, cand, nullptr, bytes);
    }
    acc = group_sum<64>(acc);
    bytes = group_sum<64>(bytes);
    if ((threadIdx.x & 63) == 0) {
        if (acc) atomicAdd(total, (unsigned long long)acc);
        if (bytes_ctr) atomicAdd(bytes_ctr, (unsigned long long)bytes);
    }
}

// Full rows (width w = |V(q)|, positions by depth) -> data ids by query vertex:
// out[r * w + u_j] = C(u_j)[row[j]].
__global__ void k_emit_ids(const DPlan *__restrict__ P, const uint32_t *__restrict__ rows, uint64_t n, uint32_t w,
                           const uint32_t *__restrict__ cand, uint32_t *__restrict__ out) {
    const uint64_t total = n * w;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t r = i / w;
        const uint32_t j = (uint32_t)(i - r * w);
        out[r * w + P->u[j]] = cand[P->coff[j] + rows[i]];
    }
}

__global__ void k_iota(uint32_t *p, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = (uint32_t)i;
}

// -------------------------------------------------------------------------- batch kernels
// A batch frontier holds the rows of many queries at one depth d, query by query: segment s
// covers rows [seg_start[s], seg_start[s+1]) of query seg_q[s] (segments may be empty).  Plans
// are re-based onto the shared candidate arrays of the batch, so every row only needs its plan.

__device__ inline uint32_t seg_of(const uint64_t *__restrict__ seg_start, uint32_t nseg, uint64_t it) {
    uint32_t lo = 0, hi = nseg;  // last s with seg_start[s] <= it
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (seg_start[mid] <= it) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Row kinds of a batch level; one kernel instantiation per kind keeps each kernel's registers to
// what its own row function needs (a kernel holding all three paths needs twice the registers of
// the largest and runs at half the occupancy).
constexpr int kExpandRows = 0, kFusedRows = 1, kLeafRows = 2;
// Kernel-table names and algorithmic-byte counter slots (offsets from d_counters + 1) of the count
// launches, one per (row kind, lane width), so that each entry is one kernel instantiation.  Fused
// rows run in k_bfused_flat (own name and slot); their row of the table is unused.
constexpr const char *kKindTimer[3][4] = {
    {"batch_count_expand_g1", "batch_count_expand_g4", "batch_count_expand_g16", "batch_count_expand_g64"},
    {"batch_count_fused_g1", "batch_count_fused_g4", "batch_count_fused_g16", "batch_count_fused_g64"},
    {"batch_count_leaf_g1", "batch_count_leaf_g4", "batch_count_leaf_g16", "batch_count_leaf_g64"}};
constexpr const char *kFillTimer[4] = {"batch_fill_g1", "batch_fill_g4", "batch_fill_g16", "batch_fill_g64"};
inline int lane_index(uint32_t G) { return G == 1 ? 0 : G == 4 ? 1 : G == 16 ? 2 : 3; }
inline int count_slot(int kind, uint32_t G) { return kind * 4 + lane_index(G); }  // 0..11
inline int fill_slot(uint32_t G) { return 12 + lane_index(G); }                    // 12..15
constexpr int kFlatFusedSlot = 16;

// Wave tiles.  A launch covers the rows of queries [qa, qb); each query's rows are cut into tiles
// of tile_rows consecutive rows, and tile_pref[i] is the number of tiles before query i.  A wave
// takes whole tiles, so the query -- and the plan every row function reads -- is uniform across
// the wave and its fields are loaded once into scalar registers instead of once per lane.
__device__ inline uint32_t tile_query(const uint64_t *__restrict__ tile_pref, uint32_t qa, uint32_t qb, uint64_t t) {
    uint32_t lo = qa, hi = qb;  // last i in [qa, qb) with tile_pref[i] <= t
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (tile_pref[mid] <= t) lo = mid;
        else hi = mid;
    }
    return (uint32_t)__builtin_amdgcn_readfirstlane((int)lo);
}

// Per row: finish it (leaf depth) into qcount[query] and count 0 children, or count its children.
// Every row of a launch is of the given kind (kLeafRows or kExpandRows; fused rows run in
// k_bfused_flat).  At least seven waves per SIMD: the leaf-depth instantiations need 109 VGPRs
// for four waves; at seven (72 VGPRs, then about 180 B per lane spilled, up to 20 B since
// group_count keeps its subset sizes in LDS) the config-4 batch took 76.5 ms against 79.3 (77.3
// at six waves, 77.8 at eight).  The expansion kind fits seven waves unforced.
template <int G, int KIND>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(7, 8))) k_bcount(const DPlan *__restrict__ plans, const uint64_t *__restrict__ seg_start,
                                                const uint64_t *__restrict__ tile_pref, uint32_t qa, uint32_t qb,
                                                uint32_t tile_rows, uint32_t d, const uint32_t *__restrict__ rows,
                                                const uint64_t *__restrict__ ce_off, const uint32_t *__restrict__ ce_val,
                                                const uint32_t *__restrict__ cand, uint32_t *__restrict__ counts,
                                                unsigned long long *__restrict__ qcount,
                                                unsigned long long *__restrict__ bytes_ctr) {
    const int lane = threadIdx.x & 63;
    const int lg = lane & (G - 1);
    const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    const uint64_t t0 = tile_pref[qa], t1 = tile_pref[qb];
    uint64_t bytes = 0;
    // Finished rows add into a running per-lane sum for the wave's current query, which the wave
    // reduces and adds with one atomic when its query changes and at the end.
    uint32_t cur_q = 0xffffffffu;
    uint64_t acc = 0;
    auto flush = [&]() {
        if (KIND != kExpandRows && cur_q != 0xffffffffu) {
            const uint64_t tot = group_sum<64>(acc);
            if (lane == 0 && tot) atomicAdd(&qcount[cur_q], (unsigned long long)tot);
        }
        acc = 0;
    };
    for (uint64_t t = t0 + wave; t < t1; t += nwaves) {
        const uint32_t q = tile_query(tile_pref, qa, qb, t);
        if (q != cur_q) {
            flush();
            cur_q = q;
        }
        const DPlan *P = plans + q;
        const uint64_t r0 = seg_start[q] + (t - tile_pref[q]) * tile_rows;
        const uint64_t r1 = min(r0 + tile_rows, seg_start[q + 1]);
        for (uint64_t it = r0 + (uint64_t)(lane / G); it < r1; it += 64 / G) {
            const uint32_t *row = rows + it * d;
            if constexpr (KIND == kLeafRows) {
                const uint64_t v = leaf_row<G>(P, row, ce_off, ce_val, cand, bytes);
                if (lg == 0) {
                    acc += v;
                    if (counts) counts[it] = 0;  // null at a level without expansions
                }
            } else {
                const uint32_t c = expand_count_row<G>(P, d, row, ce_off, ce_val, cand, bytes);
                if (lg == 0) counts[it] = c;
            }
        }
    }
    flush();
    bytes = group_sum<64>(bytes);
    if (lane == 0 && bytes_ctr) atomicAdd(bytes_ctr, (unsigned long long)bytes);
}

// Writes the children of the rows of a level (offsets from the scan of the count pass).  The cap
// of eight waves per SIMD costs no spills (43-50 VGPRs) and takes the config-4 batch from 76.6 to
// 75.0 ms: without it the compiler settles on seven.
template <int G>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8, 8))) k_bfill(const DPlan *__restrict__ plans, const uint64_t *__restrict__ seg_start,
                                               const uint64_t *__restrict__ tile_pref, uint32_t qa, uint32_t qb,
                                               uint32_t tile_rows, uint32_t d, const uint32_t *__restrict__ rows,
                                               const uint64_t *__restrict__ ce_off, const uint32_t *__restrict__ ce_val,
                                               const uint32_t *__restrict__ cand, const uint64_t *__restrict__ offs,
                                               uint32_t *__restrict__ out, unsigned long long *__restrict__ bytes_ctr) {
    const int lane = threadIdx.x & 63;
    const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    const uint64_t t0 = tile_pref[qa], t1 = tile_pref[qb];
    uint64_t bytes = 0;
    for (uint64_t t = t0 + wave; t < t1; t += nwaves) {
        const uint32_t q = tile_query(tile_pref, qa, qb, t);
        const DPlan *P = plans + q;
        const uint64_t r0 = seg_start[q] + (t - tile_pref[q]) * tile_rows;
        const uint64_t r1 = min(r0 + tile_rows, seg_start[q + 1]);
        for (uint64_t it = r0 + (uint64_t)(lane / G); it < r1; it += 64 / G) {
            const uint64_t nout = offs[it + 1] - offs[it];
            if (nout == 0) continue;  // finished rows and rows without children (group-uniform)
            expand_fill_row<G>(P, d, rows + it * d, ce_off, ce_val, cand, offs[it], nout, out, bytes);
        }
    }
    bytes = group_sum<64>(bytes);
    if (lane == 0 && bytes_ctr) atomicAdd(bytes_ctr, (unsigned long long)bytes);
}

// Logical block index that keeps runs of C consecutive logical blocks on one XCD.  The hardware
// places block b on XCD b % 8, so plain indexing spreads consecutive tiles (rows that share their
// prefixes, and so their list and table entries) over the eight L2 caches.  Here chunks of C blocks
// rotate over the XCDs: neighbouring tiles share an L2, and every XCD still gets a fine-grained share
// of the launch.  Blocks past the last whole round of 8 * C keep their own index, so the map is a
// bijection on [0, gridDim.x).
__device__ inline uint32_t xcd_chunked_block(uint32_t C) {
    const uint32_t b = blockIdx.x, span = 8 * C, full = gridDim.x / span * span;
    if (b >= full) return b;
    const uint32_t x = b % 8, k = b / 8;  // the k-th block placed on XCD x
    return (k / C) * span + x * C + (k % C);
}

// Fused rows of the batch in flattened tiles of 64 rows (fused_tile); tiles as in k_bcount with
// tile_rows = 64, taken in XCD-chunked block order (chunks of 64 blocks: 2-3 % less time per launch
// than plain block order; one contiguous range per XCD is 10 % slower, as heavy tiles cluster).
// Five waves per SIMD (at most 96 VGPRs, about 100 B per lane spilled) hide more of the lookups'
// latency than four without spills (116 VGPRs).
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5, 8))) k_bfused_flat(const DPlan *__restrict__ plans, const uint64_t *__restrict__ seg_start,
                                                     const uint64_t *__restrict__ tile_pref, uint32_t qa, uint32_t qb,
                                                     uint32_t d, const uint32_t *__restrict__ rows,
                                                     const uint64_t *__restrict__ ce_off,
                                                     const uint32_t *__restrict__ ce_val,
                                                     const uint32_t *__restrict__ cand, uint32_t *__restrict__ counts,
                                                     unsigned long long *__restrict__ qcount,
                                                     unsigned long long *__restrict__ bytes_ctr) {
    const int lane = threadIdx.x & 63;
    const uint64_t wave = ((uint64_t)xcd_chunked_block(64) * blockDim.x + threadIdx.x) >> 6;
    const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    const uint64_t t0 = tile_pref[qa], t1 = tile_pref[qb];
    uint64_t bytes = 0, acc = 0;
    uint32_t cur_q = 0xffffffffu;
    auto flush = [&]() {
        if (cur_q != 0xffffffffu) {
            const uint64_t tot = group_sum<64>(acc);
            if (lane == 0 && tot) atomicAdd(&qcount[cur_q], (unsigned long long)tot);
        }
        acc = 0;
    };
    for (uint64_t t = t0 + wave; t < t1; t += nwaves) {
        const uint32_t q = tile_query(tile_pref, qa, qb, t);
        if (q != cur_q) {
            flush();
            cur_q = q;
        }
        const uint64_t r0 = seg_start[q] + (t - tile_pref[q]) * 64;
        const uint32_t nrows = (uint32_t)min<uint64_t>(64, seg_start[q + 1] - r0);
        acc += fused_tile(plans + q, rows, r0, nrows, d, ce_off, ce_val, cand, counts, bytes);
    }
    flush();
    bytes = group_sum<64>(bytes);
    if (lane == 0 && bytes_ctr) atomicAdd(bytes_ctr, (unsigned long long)bytes);
}

// Depth-1 rows: position it - seg_start[s] in C(u_0) of the segment's query.
__global__ void k_biota(const uint64_t *__restrict__ seg_start, uint32_t nseg, uint32_t *__restrict__ rows, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        rows[i] = (uint32_t)(i - seg_start[seg_of(seg_start, nseg, i)]);
}

__global__ void k_bgather(const uint64_t *__restrict__ offs, const uint64_t *__restrict__ seg_start, uint32_t nseg,
                          uint64_t *__restrict__ out) {
    for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s <= nseg; s += gridDim.x * blockDim.x)
        out[s] = offs[seg_start[s]];
}

// -------------------------------------------------------------------------- host driver
namespace {

constexpr int kG = 16;                      // lanes per row
constexpr uint64_t kLeafChunk = 1ull << 22;  // rows per leaf-count launch
constexpr uint64_t kFusedMinRows = 1ull << 19;  // rows per fused launch, at least
constexpr uint64_t kPairCap = 1ull << 26;    // local-count (cell, value) pairs per aggregation batch

unsigned groups_grid(uint64_t items, int G) {
    uint64_t threads = items * (uint64_t)G;
    uint64_t blocks = (threads + 255) / 256;
    return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(blocks, 1u << 16));
}

bool plan_fusable(const asdm_plan *plan);

constexpr uint64_t kEdgeTableMax = 1ull << 27;  // candidate edges of an edge table (8 B each: 1 GiB)

// Table builds deferred by the batch engine: the triangle tables of every query in one launch pair,
// then the edge tables (one launch each: they read the triangle tables), then the row-sum tables in
// one launch per table level.
struct EdgeJob {
    EdgeTailSpec T;
    const uint64_t *ce_off;
    const uint32_t *ce_val;
    uint64_t *tab;
};
struct TableJobs {
    std::vector<TriJob> tri;
    std::vector<EdgeJob> edge;
    std::vector<TailJob> first, second;
};

// Edge table of a fused plan (DPlan::etail) when depth d = D-1 has two backward neighbours b1 < b2
// that are adjacent in the query: the row's candidates are L1 ∩ L2, a function of the candidate
// edge (row[b1], row[b2]).  Every leaf group must be a pendant leaf of d or a triangle leaf on
// (b1, d) or (b2, d) with no same-label matched vertex to avoid, or a hoisted leaf (at most
// kRowLeaves) whose label is not d's; same-label depths of d are corrected per row.
int build_edge_tail_table(asdm_ctx *ctx, const asdm_plan *plan, const asdm_cs *cs, DPlan &hp,
                          std::vector<uint32_t *> &tabs, TableJobs *jobs) {
    const DPlan &P = plan->h;
    const uint32_t d = P.D - 1, b1 = P.bdep[d][0], b2 = P.bdep[d][1];
    uint32_t jb = UINT32_MAX;
    for (uint32_t j = 0; j < P.nb[b2]; ++j)
        if (P.bdep[b2][j] == b1) jb = j;
    if (jb == UINT32_MAX) return ASDM_OK;
    const uint32_t a12 = cs->arc_index(plan->order[b1], plan->order[b2]);
    if (a12 == UINT32_MAX) return ASDM_OK;
    EdgeTailSpec T{};
    T.arc_base = P.bbase[b2][jb];
    T.np1 = cs->csize[plan->order[b1]];
    T.nedges = cs->arc_weight[a12];
    if (T.nedges > kEdgeTableMax) return ASDM_OK;  // the per-candidate path needs no table memory
    T.l1_base = P.bbase[d][0];
    T.l2_base = P.bbase[d][1];
    uint32_t hoisted = 0;
    for (uint32_t g = 0; g < P.ngroups; ++g) {
        const uint32_t e = P.gleaf[g][0];
        bool adj = false;
        for (uint32_t j = 0; j < P.nb[e]; ++j) adj |= P.bdep[e][j] == d;
        if (!adj) {
            for (uint32_t s = 0; s < P.ns[e]; ++s)
                if (P.sdep[e][s] == d) return ASDM_OK;  // shares d's label
            if (++hoisted > kRowLeaves) return ASDM_OK;
            continue;
        }
        if (P.ns[e] != 0 || T.nadj == kTailLeaves) return ASDM_OK;
        if (P.nb[e] == 1) {
            T.pend_base[T.nadj] = P.bbase[e][0];
            T.tri_tab[T.nadj] = nullptr;
        } else if (P.tri[e] && P.nb[e] == 2 && P.bdep[e][1] == d && (P.bdep[e][0] == b1 || P.bdep[e][0] == b2) &&
                   hp.tri_tab[e]) {
            T.tri_tab[T.nadj] = hp.tri_tab[e];
            T.tri_list[T.nadj] = P.bdep[e][0] == b2 ? 1u : 0u;
        } else {
            return ASDM_OK;
        }
        ++T.nadj;
    }
    uint64_t *tab = nullptr;
    int rc = dev_alloc(ctx, (void **)&tab, 8ull * std::max<uint64_t>(T.nedges, 1));
    if (rc) return rc;
    tabs.push_back(reinterpret_cast<uint32_t *>(tab));
    if (T.nedges && jobs) {
        jobs->edge.push_back({T, cs->d_ce_off, cs->d_ce_val, tab});
    } else if (T.nedges) {
        KTimer kt(ctx, "edge_tail_sums");
        const unsigned grid = (unsigned)std::min<uint64_t>((T.nedges + 3) / 4, 1u << 16);
        hipLaunchKernelGGL(k_edge_tail_sums, dim3(grid), dim3(256), 0, ctx->stream, T, cs->d_ce_off, cs->d_ce_val, tab);
        kt.stop(40ull * T.nedges);
        if ((rc = check_hip(ctx, hipGetLastError(), "edge tail table"))) return rc;
    }
    hp.etail = tab;
    return ASDM_OK;
}

// Row-sum table of a fused plan (DPlan::tail) when depth d = D-1 has a single backward neighbour
// b: tail[p] = sum over the candidates x adjacent to position p of C(u_b) of the product of the
// adjacent leaves' list counts at x.  It applies when every leaf group is a pendant leaf of d or a
// triangle leaf on (b, d), or a hoisted leaf (not adjacent to d; at most kRowLeaves of them), and
// at most one leaf is a correction factor: an adjacent leaf with same-label matched vertices to
// avoid, or a hoisted leaf with d's label.  Then a row's sum needs no scan of its
// candidates: tail_row corrects injectivity against the same-label depths of d, and
// tail_fix_tile the correction factor over the intersection of the row's list with one other
// list per source.  Needs hp.tri_tab.

int build_tail_table(asdm_ctx *ctx, const asdm_plan *plan, const asdm_cs *cs, DPlan &hp, std::vector<uint32_t *> &tabs,
                     TableJobs *jobs) {
    const DPlan &P = plan->h;
    hp.tail = nullptr;
    hp.tail_dep = 0;
    hp.tail2 = nullptr;
    hp.tail2_dep = 0;
    hp.etail = nullptr;
    hp.tail_fix = 0;
    hp.tail_leaf = 0;
    hp.tail_rbase = 0;
    if (!plan_fusable(plan)) return ASDM_OK;
    const uint32_t d = P.D - 1;
    if (P.nb[d] == 2) return build_edge_tail_table(ctx, plan, cs, hp, tabs, jobs);
    if (P.nb[d] != 1) return ASDM_OK;
    const uint32_t b = P.bdep[d][0];
    TailSpec T{};
    T.arc_base = P.bbase[d][0];
    T.np = cs->csize[plan->order[b]];
    uint32_t hoisted = 0, fix = 0, fix_leaf = 0;
    uint64_t rbase = 0;
    for (uint32_t g = 0; g < P.ngroups; ++g) {
        const uint32_t e = P.gleaf[g][0];
        bool adj = false;
        for (uint32_t j = 0; j < P.nb[e]; ++j) adj |= P.bdep[e][j] == d;
        if (!adj) {
            bool same_d = false;  // shares d's label
            for (uint32_t s = 0; s < P.ns[e]; ++s) same_d |= P.sdep[e][s] == d;
            if (same_d) {
                if (fix) return ASDM_OK;
                fix = kTailFixHoist;
                fix_leaf = e;
            }
            if (++hoisted > kRowLeaves) return ASDM_OK;
            continue;
        }
        if (T.nadj == kTailLeaves) return ASDM_OK;
        if (P.ns[e] != 0) {
            const uint32_t ra = cs->arc_index(plan->order[e], plan->order[d]);
            if (fix || ra == UINT32_MAX) return ASDM_OK;
            fix = kTailFixLeaf;
            fix_leaf = e;
            rbase = cs->ce_base[ra];
        }
        if (P.nb[e] == 1) {
            T.pend_base[T.nadj] = P.bbase[e][0];
            T.tri_tab[T.nadj] = nullptr;
        } else if (P.tri[e] && P.nb[e] == 2 && P.bdep[e][0] == b && P.bdep[e][1] == d && hp.tri_tab[e]) {
            T.pend_base[T.nadj] = 0;
            T.tri_tab[T.nadj] = hp.tri_tab[e];
        } else {
            return ASDM_OK;
        }
        ++T.nadj;
    }
    uint64_t *tab = nullptr;
    int rc = dev_alloc(ctx, (void **)&tab, 8ull * std::max<uint32_t>(T.np, 1));
    if (rc) return rc;
    tabs.push_back(reinterpret_cast<uint32_t *>(tab));
    if (T.np && jobs) {
        jobs->first.push_back({T, cs->d_ce_off, cs->d_ce_val, tab});
    } else if (T.np) {
        KTimer kt(ctx, "tail_sums");
        const unsigned grid = (unsigned)std::min<uint64_t>(((uint64_t)T.np + 3) / 4, 1u << 16);
        hipLaunchKernelGGL(k_tail_sums, dim3(grid), dim3(256), 0, ctx->stream, T, cs->d_ce_off, cs->d_ce_val, tab);
        const uint32_t a = cs->arc_index(plan->order[b], plan->order[d]);
        kt.stop(24ull * T.np + (4ull + 16ull * T.nadj) * cs->arc_weight[a]);
        if ((rc = check_hip(ctx, hipGetLastError(), "tail table"))) return rc;
    }
    hp.tail = tab;
    hp.tail_dep = b;
    hp.tail_fix = fix;
    hp.tail_leaf = fix_leaf;
    hp.tail_rbase = rbase;
    // Second table: c = D-2 has a single backward neighbour b2, the table's backward neighbour b is
    // c or earlier, no correction factor and no same-label depth of d apply, and every hoisted leaf
    // that reads c (as a neighbour or a same-label vertex) is a pendant leaf of c with no same-label
    // matched vertex to avoid.  Then c's candidates x' depend on a row of width c only through its
    // vertex at b2, and each contributes (tail[x'] if b = c) times those leaves' list lengths at x'
    // (1 when there are none): depth D-1 is never materialized.
    const uint32_t c = P.D - 2;
    if (fix || P.ns[d] || P.D < 3 || b > c || P.nb[c] != 1) return ASDM_OK;
    TailSpec T2{};
    for (uint32_t g = 0; g < P.ngroups; ++g) {
        const uint32_t e = P.gleaf[g][0];
        bool adj_d = false, reads_c = false;
        for (uint32_t j = 0; j < P.nb[e]; ++j) {
            adj_d |= P.bdep[e][j] == d;
            reads_c |= P.bdep[e][j] == c;
        }
        for (uint32_t s = 0; s < P.ns[e]; ++s) reads_c |= P.sdep[e][s] == c;
        if (adj_d || !reads_c) continue;  // inside tail, or part of H0
        if (P.nb[e] != 1 || P.bdep[e][0] != c || P.ns[e] != 0 || T2.nadj == kTailLeaves) return ASDM_OK;
        T2.pend_base[T2.nadj] = P.bbase[e][0];
        T2.tri_tab[T2.nadj] = nullptr;
        ++T2.nadj;
    }
    T2.arc_base = P.bbase[c][0];
    T2.np = cs->csize[plan->order[P.bdep[c][0]]];
    T2.inner = b == c ? tab : nullptr;
    uint64_t *tab2 = nullptr;
    if ((rc = dev_alloc(ctx, (void **)&tab2, 8ull * std::max<uint32_t>(T2.np, 1)))) return rc;
    tabs.push_back(reinterpret_cast<uint32_t *>(tab2));
    if (T2.np && jobs) {
        jobs->second.push_back({T2, cs->d_ce_off, cs->d_ce_val, tab2});
    } else if (T2.np) {
        KTimer kt(ctx, "tail_sums");
        const unsigned grid = (unsigned)std::min<uint64_t>(((uint64_t)T2.np + 3) / 4, 1u << 16);
        hipLaunchKernelGGL(k_tail_sums, dim3(grid), dim3(256), 0, ctx->stream, T2, cs->d_ce_off, cs->d_ce_val, tab2);
        const uint32_t a = cs->arc_index(plan->order[P.bdep[c][0]], plan->order[c]);
        kt.stop(24ull * T2.np + 12ull * cs->arc_weight[a]);
        if ((rc = check_hip(ctx, hipGetLastError(), "second tail table"))) return rc;
    }
    hp.tail2 = tab2;
    hp.tail2_dep = P.bdep[c][0];
    return ASDM_OK;
}

// Builds the per-candidate-edge intersection table of every triangle leaf of the plan and points
// hp.tri_tab at them, then the row-sum table of the fused level where it applies (tables appended
// to tabs; the caller frees them).
int build_tri_tables(asdm_ctx *ctx, const asdm_plan *plan, const asdm_cs *cs, DPlan &hp, std::vector<uint32_t *> &tabs,
                     TableJobs *jobs = nullptr) {
    const DPlan &P = plan->h;
    for (uint32_t e = P.D; e < P.nq; ++e) {
        if (!P.tri[e]) continue;
        const uint32_t p0 = plan->order[P.bdep[e][0]], p1 = plan->order[P.bdep[e][1]];
        const uint32_t a = cs->arc_index(p0, p1);
        const uint64_t nedges = cs->arc_weight[a];
        uint32_t *tab = nullptr;
        int rc = dev_alloc(ctx, (void **)&tab, 4 * std::max<uint64_t>(nedges, 1));
        if (rc) return rc;
        tabs.push_back(tab);
        hp.tri_tab[e] = tab;
        if (nedges && jobs) {
            jobs->tri.push_back({cs->d_ce_off, cs->d_ce_val, cs->ce_base[a], nedges, P.bbase[e][0], P.bbase[e][1],
                                 cs->csize[p0], tab});
        } else if (nedges) {
            KTimer kt(ctx, "tri_table");
            const unsigned grid = (unsigned)std::min<uint64_t>((nedges + 255) / 256, 1u << 16);
            hipLaunchKernelGGL(k_tri_table, dim3(grid), dim3(256), 0, ctx->stream, cs->d_ce_off, cs->d_ce_val,
                               cs->ce_base[a], cs->csize[p0], nedges, P.bbase[e][0], P.bbase[e][1], tab);
            const unsigned wgrid = (unsigned)std::min<uint64_t>((nedges + 3) / 4, 1u << 14);
            hipLaunchKernelGGL(k_tri_table_wide, dim3(wgrid), dim3(256), 0, ctx->stream, cs->d_ce_off, cs->d_ce_val,
                               cs->ce_base[a], cs->csize[p0], nedges, P.bbase[e][0], P.bbase[e][1], tab);
            kt.stop(8ull * nedges);
            if ((rc = check_hip(ctx, hipGetLastError(), "tri table"))) return rc;
        }
    }
    return build_tail_table(ctx, plan, cs, hp, tabs, jobs);
}

// Device copies of deferred jobs and their item prefix (freed with tabs).
template <typename Job>
int upload_jobs(asdm_ctx *ctx, const std::vector<Job> &jobs, const std::vector<uint64_t> &pref, std::vector<uint32_t *> &tabs,
                Job **jd, uint64_t **pd) {
    int rc = dev_alloc(ctx, (void **)jd, sizeof(Job) * jobs.size());
    if (rc) return rc;
    tabs.push_back(reinterpret_cast<uint32_t *>(*jd));
    if ((rc = dev_alloc(ctx, (void **)pd, 8 * pref.size()))) return rc;
    tabs.push_back(reinterpret_cast<uint32_t *>(*pd));
    if ((rc = check_hip(ctx, hipMemcpy(*jd, jobs.data(), sizeof(Job) * jobs.size(), hipMemcpyHostToDevice), "H2D jobs")))
        return rc;
    return check_hip(ctx, hipMemcpy(*pd, pref.data(), 8 * pref.size(), hipMemcpyHostToDevice), "H2D jobs");
}

// The deferred triangle tables in one launch pair.
int run_tri_jobs(asdm_ctx *ctx, const std::vector<TriJob> &jobs, std::vector<uint32_t *> &tabs) {
    if (jobs.empty()) return ASDM_OK;
    std::vector<uint64_t> pref(jobs.size() + 1, 0);
    for (size_t j = 0; j < jobs.size(); ++j) pref[j + 1] = pref[j] + jobs[j].nedges;
    TriJob *jd = nullptr;
    uint64_t *pd = nullptr;
    if (int rc = upload_jobs(ctx, jobs, pref, tabs, &jd, &pd)) return rc;
    KTimer kt(ctx, "tri_table");
    const uint64_t total = pref.back();
    const unsigned grid = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((total + 255) / 256, 1u << 16));
    hipLaunchKernelGGL(k_tri_table_multi, dim3(grid), dim3(256), 0, ctx->stream, jd, pd, (uint32_t)jobs.size());
    const unsigned wgrid = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((total + 3) / 4, 1u << 14));
    hipLaunchKernelGGL(k_tri_table_wide_multi, dim3(wgrid), dim3(256), 0, ctx->stream, jd, pd, (uint32_t)jobs.size());
    kt.stop(8ull * total);
    return check_hip(ctx, hipGetLastError(), "tri tables");
}

// The deferred edge tables, one launch each.
int run_edge_jobs(asdm_ctx *ctx, const std::vector<EdgeJob> &jobs) {
    for (const EdgeJob &J : jobs) {
        KTimer kt(ctx, "edge_tail_sums");
        const unsigned grid = (unsigned)std::min<uint64_t>((J.T.nedges + 3) / 4, 1u << 16);
        hipLaunchKernelGGL(k_edge_tail_sums, dim3(grid), dim3(256), 0, ctx->stream, J.T, J.ce_off, J.ce_val, J.tab);
        kt.stop(40ull * J.T.nedges);
        if (int rc = check_hip(ctx, hipGetLastError(), "edge tail table")) return rc;
    }
    return ASDM_OK;
}

// One launch for the deferred tables of a level (job device copies are freed with tabs).
int run_tail_jobs(asdm_ctx *ctx, const std::vector<TailJob> &jobs, std::vector<uint32_t *> &tabs) {
    if (jobs.empty()) return ASDM_OK;
    std::vector<uint64_t> pref(jobs.size() + 1, 0);
    for (size_t j = 0; j < jobs.size(); ++j) pref[j + 1] = pref[j] + jobs[j].T.np;
    TailJob *jd = nullptr;
    uint64_t *pd = nullptr;
    if (int rc = upload_jobs(ctx, jobs, pref, tabs, &jd, &pd)) return rc;
    KTimer kt(ctx, "tail_sums");
    const unsigned grid = (unsigned)std::min<uint64_t>((pref.back() + 3) / 4, 1u << 16);
    hipLaunchKernelGGL(k_tail_sums_multi, dim3(std::max(1u, grid)), dim3(256), 0, ctx->stream, jd, pd, (uint32_t)jobs.size());
    kt.stop();
    return check_hip(ctx, hipGetLastError(), "tail tables");
}

// Fusion pays when each leaf count is a few lookups: every leaf group is a single leaf with one
// candidate-edge list or a triangle leaf.
bool plan_fusable(const asdm_plan *plan) {
    if (plan->D < 2 || plan->local) return false;
    for (uint32_t g = 0; g < plan->h.ngroups; ++g) {
        const uint32_t e = plan->h.gleaf[g][0];
        if (plan->h.gsize[g] != 1 || !(plan->h.nb[e] == 1 || plan->h.tri[e]) || plan->h.ns[e] > kFusedSame)
            return false;
    }
    return true;
}

struct Engine {
    asdm_ctx *ctx;
    const asdm_plan *plan;
    const asdm_cs *cs;
    DPlan *dplan = nullptr;
    uint32_t D;
    uint64_t budget;                  // bytes of row storage per level
    std::vector<uint32_t *> rows;     // [D + 1] level buffers (rows of width d)
    std::vector<uint64_t> cap;        // rows capacity per level
    std::vector<uint32_t *> counts;   // [D] per-row candidate counts
    std::vector<uint64_t *> offs;     // [D] exclusive scan of counts (+ total)
    unsigned long long *d_total = nullptr;
    unsigned long long *d_bytes = nullptr;
    unsigned long long *d_mat = nullptr;  // local mode
    uint64_t rows_total = 0, launches = 0;
    std::vector<uint64_t> hoff;       // host copy of offsets when splitting
    std::chrono::steady_clock::time_point t_start;
    bool fused = false;           // depth D-1 evaluated by k_leaf_fused
    // Enumeration (asdm_enumerate): the plan's leaf depth (at most one) is expanded like the others,
    // D = |V(q)|, and every full row is written out as data ids (enum_out may be null: count only).
    bool enumerate = false;
    uint32_t *enum_out = nullptr;
    uint64_t enum_cap = 0, enum_n = 0;
    uint32_t *enum_dev = nullptr;
    uint64_t enum_dev_rows = 0;
    uint64_t fanout_last = 1;     // expected candidate-list length at depth D-1

    bool timed_out() const {
        if (ctx->time_limit_ms <= 0) return false;
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count() >
               ctx->time_limit_ms;
    }

    DPlan hplan;                       // plan copy with the triangle-table pointers filled in
    std::vector<uint32_t *> tri_tabs;

    int init() {
        D = enumerate ? plan->h.nq : plan->D;
        hplan = plan->h;
        int rc = build_tri_tables(ctx, plan, cs, hplan, tri_tabs);
        if (!rc) rc = dev_alloc(ctx, (void **)&dplan, sizeof(DPlan));
        if (!rc) rc = check_hip(ctx, hipMemcpyAsync(dplan, &hplan, sizeof(DPlan), hipMemcpyHostToDevice, ctx->stream), "plan upload");
        if (rc) return rc;
        rows.assign(D + 1, nullptr);
        cap.assign(D + 1, 0);
        counts.assign(D + 1, nullptr);
        offs.assign(D + 1, nullptr);
        for (uint32_t d = 1; d <= D; ++d) {
            uint64_t c = std::max<uint64_t>(1, std::min<uint64_t>(budget / (4ull * d), 1ull << 28));
            if (d == 1) c = std::max<uint64_t>(c, cs->csize[plan->order[0]]);
            cap[d] = c;
        }
        if ((rc = ensure_rows(1))) return rc;
        rc = fill_u64(ctx, ctx->d_counters, kNumCounters, 0);
        d_total = (unsigned long long *)ctx->d_counters;
        d_bytes = (unsigned long long *)ctx->d_counters + 1;
        return rc;
    }

    // Level buffers are allocated on first use, at the level's full capacity.
    int ensure_rows(uint32_t d) {
        if (rows[d]) return ASDM_OK;
        return dev_alloc(ctx, (void **)&rows[d], 4ull * d * cap[d]);
    }
    int ensure_scan(uint32_t d) {
        if (counts[d]) return ASDM_OK;
        int rc = dev_alloc(ctx, (void **)&counts[d], 4ull * cap[d]);
        if (!rc) rc = dev_alloc(ctx, (void **)&offs[d], 8ull * (cap[d] + 1));
        return rc;
    }

    void release() {
        dev_free(ctx, enum_dev);
        for (auto t : tri_tabs) dev_free(ctx, t);
        dev_free(ctx, dplan);
        for (auto p : rows) dev_free(ctx, p);
        for (auto p : counts) dev_free(ctx, p);
        for (auto p : offs) dev_free(ctx, p);
    }

    // Fused last level: rows at depth D-1, leaf products summed over their depth-(D-1) candidates.
    // rows r[0, n) of width w (D-1, or D-2 with a second table)
    int leaf_fused(uint32_t *r, uint64_t n, uint32_t w) {
        // about kLeafChunk candidates per launch, but never fewer rows than fill the chip with
        // 64-row tiles (2^19 rows = 8192 waves, 32 per CU)
        const uint64_t chunk = std::max<uint64_t>(kFusedMinRows, kLeafChunk / std::max<uint64_t>(1, fanout_last));
        for (uint64_t i0 = 0; i0 < n; i0 += chunk) {
            const uint64_t m = std::min<uint64_t>(chunk, n - i0);
            if (i0 && timed_out()) return ctx->fail(ASDM_E_TIMEOUT, "per-query time limit exceeded");
            ++launches;
            KTimer kt(ctx, "leaf_fused");
            const uint32_t *rr = r + i0 * w;
            hipLaunchKernelGGL(k_leaf_fused, dim3(groups_grid((m + 63) / 64, 64)), dim3(256), 0, ctx->stream, dplan, rr, m,
                               cs->d_ce_off, cs->d_ce_val, cs->d_cand, d_total, d_bytes + 2, w);
            kt.stop();
            int rc = check_hip(ctx, hipGetLastError(), "fused leaf kernel");
            if (rc) return rc;
            if (i0 + m < n) {
                rc = check_hip(ctx, hipStreamSynchronize(ctx->stream), "sync");
                if (rc) return rc;
            }
        }
        return ASDM_OK;
    }

    // LOCAL_COUNT of leaf rows r[0, m) without atomics on the matrix: count pass, scan, then per
    // batch of at most `cap` pairs a fill pass and the owner-partitioned aggregation (one writer
    // per cell).
    int leaf_local(uint32_t *r, uint64_t m) {
        uint32_t *np = nullptr, *flag = nullptr;
        uint64_t *po = nullptr;
        std::vector<void *> tmp;
        auto alloc = [&](void **p, size_t bytes) {
            int rc = dev_alloc(ctx, p, std::max<size_t>(bytes, 8));
            if (!rc) tmp.push_back(*p);
            return rc;
        };
        auto done = [&](int rc) {
            for (void *p : tmp) dev_free(ctx, p);
            return rc;
        };
        int rc = alloc((void **)&np, 4 * m);
        if (!rc) rc = alloc((void **)&po, 8 * (m + 1));
        if (!rc) rc = alloc((void **)&flag, 4);
        if (!rc) rc = fill_u32(ctx, flag, 1, 0);
        if (rc) return done(rc);
        {
            KTimer kt(ctx, "leaf_local_count");
            hipLaunchKernelGGL((k_leaf_local_pairs<kG, false>), dim3(groups_grid(m, kG)), dim3(256), 0, ctx->stream, dplan,
                               r, m, cs->d_ce_off, cs->d_ce_val, cs->d_cand, np, nullptr, 0, nullptr, nullptr, d_total,
                               flag);
            kt.stop();
        }
        if ((rc = check_hip(ctx, hipGetLastError(), "local pair count"))) return done(rc);
        if ((rc = scan_u32_to_u64(ctx, np, po, m))) return done(rc);
        std::vector<uint64_t> hpo(m + 1);
        uint32_t hflag = 0;
        rc = check_hip(ctx, hipMemcpyAsync(hpo.data(), po, 8 * (m + 1), hipMemcpyDeviceToHost, ctx->stream), "D2H");
        if (!rc) rc = check_hip(ctx, hipMemcpyAsync(&hflag, flag, 4, hipMemcpyDeviceToHost, ctx->stream), "D2H");
        if (!rc) rc = check_hip(ctx, hipStreamSynchronize(ctx->stream), "sync");
        if (rc) return done(rc);
        if (hflag) return done(ctx->fail(ASDM_E_UNSUPPORTED, "more than 2^32 local-count pairs in one row"));
        const uint64_t cells = (uint64_t)cs->nv * cs->nq;
        // the pairs (16 bytes) and the aggregation's bucketed copy (12 bytes) within the budget
        const uint64_t cap = std::max<uint64_t>(1024, std::min<uint64_t>(kPairCap, budget / 28));
        for (uint64_t a = 0; a < m;) {
            // largest b > a with at most cap pairs in rows [a, b) (one row may exceed the cap)
            uint64_t lo = a + 1, hi = m;
            while (lo < hi) {
                const uint64_t mid = lo + (hi - lo + 1) / 2;
                if (hpo[mid] - hpo[a] <= cap) lo = mid;
                else hi = mid - 1;
            }
            const uint64_t b = lo, n = hpo[b] - hpo[a];
            if (n) {
                uint64_t *keys = nullptr, *vals = nullptr;
                rc = dev_alloc(ctx, (void **)&keys, 8 * n);
                if (!rc && (rc = dev_alloc(ctx, (void **)&vals, 8 * n))) {
                    dev_free(ctx, keys);
                    keys = nullptr;
                }
                if (!rc) {
                    KTimer kt(ctx, "leaf_local_fill");
                    hipLaunchKernelGGL((k_leaf_local_pairs<kG, true>), dim3(groups_grid(b - a, kG)), dim3(256), 0,
                                       ctx->stream, dplan, r + a * D, b - a, cs->d_ce_off, cs->d_ce_val, cs->d_cand,
                                       nullptr, po + a, hpo[a], keys, vals, nullptr, nullptr);
                    kt.stop();
                    rc = check_hip(ctx, hipGetLastError(), "local pair fill");
                }
                if (!rc) rc = aggregate_pairs(ctx, keys, vals, n, cells, (uint64_t *)d_mat);
                dev_free(ctx, keys);
                dev_free(ctx, vals);
                if (rc) return done(rc);
            }
            a = b;
        }
        return done(ASDM_OK);
    }

    // Leaf counting in launches of at most kLeafChunk rows, so that the limits are checked often.
    // Full rows r[0, n) to the caller's buffer, as data ids by query vertex.
    int emit(uint32_t *r, uint64_t n) {
        if (enum_out && enum_n + n > enum_cap)
            return ctx->fail(ASDM_E_BUDGET, "more embeddings than the output buffer holds");
        if (enum_out) {
            if (enum_dev_rows < n) {
                dev_free(ctx, enum_dev);
                enum_dev = nullptr;
                enum_dev_rows = 0;
                if (int rc = dev_alloc(ctx, (void **)&enum_dev, 4ull * D * n)) return rc;
                enum_dev_rows = n;
            }
            const unsigned grid = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((n * D + 255) / 256, 1u << 16));
            hipLaunchKernelGGL(k_emit_ids, dim3(grid), dim3(256), 0, ctx->stream, dplan, r, n, D, cs->d_cand, enum_dev);
            if (int rc = check_hip(ctx, hipGetLastError(), "emit")) return rc;
            if (int rc = check_hip(ctx, hipMemcpyAsync(enum_out + enum_n * D, enum_dev, 4ull * D * n, hipMemcpyDeviceToHost,
                                                       ctx->stream), "D2H embeddings"))
                return rc;
            if (int rc = check_hip(ctx, hipStreamSynchronize(ctx->stream), "sync")) return rc;
        }
        enum_n += n;
        return ASDM_OK;
    }

    int leaf(uint32_t *r, uint64_t n) {
        for (uint64_t i0 = 0; i0 < n; i0 += kLeafChunk) {
            const uint64_t m = std::min<uint64_t>(kLeafChunk, n - i0);
            if (i0 && timed_out()) return ctx->fail(ASDM_E_TIMEOUT, "per-query time limit exceeded");
            ++launches;
            if (d_mat) {
                if (int rc = leaf_local(r + i0 * D, m)) return rc;
            } else {
                KTimer kt(ctx, "leaf_count");
                hipLaunchKernelGGL((k_leaf<kG>), dim3(groups_grid(m, kG)), dim3(256), 0, ctx->stream, dplan, r + i0 * D,
                                   m, cs->d_ce_off, cs->d_ce_val, cs->d_cand, d_total, d_bytes + 2);
                kt.stop();
            }
            int rc = check_hip(ctx, hipGetLastError(), "leaf kernel");
            if (rc) return rc;
            if (i0 + m < n) {  // let the device catch up so the wall-clock check means something
                rc = check_hip(ctx, hipStreamSynchronize(ctx->stream), "sync");
                if (rc) return rc;
            }
        }
        return ASDM_OK;
    }

    // rows r[0, n) at depth d (width d)
    int process(uint32_t d, uint32_t *r, uint64_t n) {
        if (n == 0) return ASDM_OK;
        if (timed_out()) return ctx->fail(ASDM_E_TIMEOUT, "per-query time limit exceeded");
        rows_total += n;
        if (ctx->row_limit && rows_total > ctx->row_limit)
            return ctx->fail(ASDM_E_TIMEOUT, "per-query row limit exceeded");
        if (d == D) return enumerate ? emit(r, n) : leaf(r, n);
        // rows of width D-1 of a fused plan: the fused kernel scans (or reads from the row-sum
        // table) their depth-(D-1) candidates itself, so they need no child count, and the rows
        // it finishes are never materialized (nor counted against the row limit)
        if (fused && d + 1 == D) return leaf_fused(r, n, d);
        if (fused && hplan.tail2 && d + 2 == D) return leaf_fused(r, n, d);
        int rc = ensure_scan(d);
        if (rc) return rc;
        ++launches;
        {
            KTimer kt(ctx, "expand_count");
            hipLaunchKernelGGL((k_expand<kG, false>), dim3(groups_grid(n, kG)), dim3(256), 0, ctx->stream, dplan, d, r, n,
                               cs->d_ce_off, cs->d_ce_val, cs->d_cand, counts[d], nullptr, 0, nullptr, d_bytes);
            kt.stop();
        }
        rc = check_hip(ctx, hipGetLastError(), "expand count");
        if (!rc) rc = scan_u32_to_u64(ctx, counts[d], offs[d], n);
        uint64_t total = 0;
        if (!rc) rc = check_hip(ctx, hipMemcpyAsync(&total, offs[d] + n, 8, hipMemcpyDeviceToHost, ctx->stream), "D2H");
        if (!rc) rc = check_hip(ctx, hipStreamSynchronize(ctx->stream), "sync");
        if (rc || total == 0) return rc;
        const uint64_t c = cap[d + 1];
        if (total <= c) return fill_and_recurse(d, r, 0, n, 0, total);
        // split the rows so that each chunk's output fits the next level
        std::vector<uint64_t> h(n + 1);
        if ((rc = check_hip(ctx, hipMemcpy(h.data(), offs[d], 8 * (n + 1), hipMemcpyDeviceToHost), "D2H offsets"))) return rc;
        uint64_t i0 = 0;
        while (i0 < n) {
            // largest i1 with h[i1] - h[i0] <= c
            uint64_t i1 = (uint64_t)(std::upper_bound(h.begin() + i0, h.begin() + n + 1, h[i0] + c) - h.begin()) - 1;
            if (i1 == i0) {
                // a single row has more children than the level holds: grow that level
                i1 = i0 + 1;
                const uint64_t need = h[i1] - h[i0];
                dev_free(ctx, rows[d + 1]);
                rows[d + 1] = nullptr;
                cap[d + 1] = need;
                if (counts[d + 1]) {
                    dev_free(ctx, counts[d + 1]);
                    dev_free(ctx, offs[d + 1]);
                    counts[d + 1] = nullptr;
                    offs[d + 1] = nullptr;
                }
            }
            if ((rc = fill_and_recurse(d, r, i0, i1, h[i0], h[i1] - h[i0]))) return rc;
            i0 = i1;
        }
        return ASDM_OK;
    }

    int fill_and_recurse(uint32_t d, uint32_t *r, uint64_t i0, uint64_t i1, uint64_t out_base, uint64_t nout) {
        int rc = ensure_rows(d + 1);
        if (rc) return rc;
        ++launches;
        {
            KTimer kt(ctx, "expand_fill");
            hipLaunchKernelGGL((k_expand<kG, true>), dim3(groups_grid(i1 - i0, kG)), dim3(256), 0, ctx->stream, dplan, d,
                               r + i0 * d, i1 - i0, cs->d_ce_off, cs->d_ce_val, cs->d_cand, nullptr, offs[d] + i0,
                               out_base, rows[d + 1], d_bytes + 1);
            kt.stop();
        }
        rc = check_hip(ctx, hipGetLastError(), "expand fill");
        if (rc) return rc;
        return process(d + 1, rows[d + 1], nout);
    }

    int run() {
        const uint32_t c0 = cs->csize[plan->order[0]];
        hipLaunchKernelGGL(k_iota, dim3(groups_grid(c0, 1)), dim3(256), 0, ctx->stream, rows[1], (uint64_t)c0);
        int rc = check_hip(ctx, hipGetLastError(), "iota");
        if (rc) return rc;
        return process(1, rows[1], c0);
    }
};

uint64_t default_budget(asdm_ctx *ctx) {
    size_t freeb = 0, totalb = 0;
    (void)hipMemGetInfo(&freeb, &totalb);
    uint64_t b = freeb / 32;
    return std::max<uint64_t>(64ull << 20, std::min<uint64_t>(b, 1ull << 30));
}

int run_engine(asdm_ctx *ctx, const asdm_plan *plan, const asdm_cs *cs, uint64_t budget, uint64_t *mat_host,
               uint64_t *count, bool enumerate = false, uint32_t *enum_out = nullptr, uint64_t enum_cap = 0) {
    if (!ctx || !plan || !cs || !count) return ASDM_E_INVALID;
    if (plan->h.nq != cs->nq) return ctx->fail(ASDM_E_INVALID, "plan and candidate space disagree");
    // a plan with cartesian steps counts on the candidate space it was made on: the caller's,
    // extended with the cartesian arcs
    if (plan->cart_cs) cs = plan->cart_cs.get();
    if (plan->bags) {
        if (enumerate) return ctx->fail(ASDM_E_INVALID, "enumeration needs a plan made with ASDM_PLAN_NO_FACTOR");
        return run_bags(ctx, plan, cs, mat_host, count);
    }
    if (mat_host && !plan->local) return ctx->fail(ASDM_E_INVALID, "local counting needs a plan made with ASDM_PLAN_LOCAL");
    // a plan without factorization has at most one leaf (the last vertex, counted in closed form)
    if (enumerate && plan->D + 1 != plan->h.nq && plan->D != plan->h.nq)
        return ctx->fail(ASDM_E_INVALID, "enumeration needs a plan made with ASDM_PLAN_NO_FACTOR");
    (void)hipSetDevice(ctx->device);
    auto t0 = std::chrono::steady_clock::now();
    Engine E{ctx, plan, cs};
    E.t_start = t0;
    E.enumerate = enumerate;
    E.enum_out = enum_out;
    E.enum_cap = enum_cap;
    // Fuse the last enumerated level into leaf counting when there is one (count mode only); the
    // expected list length of that level (smallest arc fan-out) sizes the fused launches' chunks.
    if (!mat_host && !enumerate && plan_fusable(plan)) {
        E.fused = true;
        const uint32_t d = plan->D - 1;
        double best = 1e300;
        for (uint32_t j = 0; j < plan->h.nb[d]; ++j) {
            const uint32_t p = plan->order[plan->h.bdep[d][j]];
            const uint32_t a = cs->arc_index(p, plan->order[d]);
            best = std::min(best, (double)cs->arc_weight[a] / std::max(1.0, (double)cs->csize[p]));
        }
        E.fanout_last = (uint64_t)std::max(1.0, best);
    }
    E.budget = budget ? std::max<uint64_t>(budget, 1ull << 12) : default_budget(ctx);
    int rc = E.init();
    uint64_t matn = (uint64_t)cs->nv * cs->nq;
    if (!rc && mat_host) {
        rc = dev_alloc(ctx, (void **)&E.d_mat, 8ull * std::max<uint64_t>(matn, 1));
        if (!rc) rc = fill_u64(ctx, (uint64_t *)E.d_mat, matn, 0);
    }
    hipEvent_t ea, eb;  // the call's device time: timing only, no system-scope fence on record
    (void)hipEventCreateWithFlags(&ea, hipEventDisableSystemFence);
    (void)hipEventCreateWithFlags(&eb, hipEventDisableSystemFence);
    (void)hipEventRecord(ea, ctx->stream);
    if (!rc) rc = E.run();
    (void)hipEventRecord(eb, ctx->stream);
    uint64_t res[4] = {0, 0, 0, 0};
    if (!rc) rc = check_hip(ctx, hipMemcpyAsync(res, ctx->d_counters, 32, hipMemcpyDeviceToHost, ctx->stream), "D2H");
    if (!rc && mat_host) rc = check_hip(ctx, hipMemcpyAsync(mat_host, E.d_mat, 8 * matn, hipMemcpyDeviceToHost, ctx->stream), "D2H matrix");
    if (!rc) rc = check_hip(ctx, hipStreamSynchronize(ctx->stream), "sync");
    float kms = 0;
    (void)hipEventElapsedTime(&kms, ea, eb);
    (void)hipEventDestroy(ea);
    (void)hipEventDestroy(eb);
    if (E.d_mat) dev_free(ctx, E.d_mat);
    E.release();
    (void)hipStreamSynchronize(ctx->stream);
    if (rc) return rc;
    *count = enumerate ? E.enum_n : res[0];
    resolve_timers(ctx);
    add_kernel_bytes(ctx, "expand_count", res[1]);
    add_kernel_bytes(ctx, "expand_fill", res[2]);
    add_kernel_bytes(ctx, E.fused ? "leaf_fused" : (mat_host ? "leaf_local" : "leaf_count"), res[3]);
    ctx->stats.exec_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    ctx->stats.kernel_ms = kms;
    ctx->stats.alg_bytes = res[1] + res[2] + res[3];
    ctx->stats.rows_total = E.rows_total;
    ctx->stats.launches = E.launches;
    ctx->stats.plan_ms = plan->plan_ms;
    ctx->stats.filter_ms = cs->filter_ms;
    return ASDM_OK;
}

}  // namespace
}  // namespace asdm

namespace asdm {
namespace {

// Lane-group width class of a query: its expected candidate-list length at the deepest
// enumerated level (smallest fan-out among that level's matched neighbours).
uint32_t lane_class(const asdm_plan *plan, const asdm_cs *cs, double scale) {
    const uint32_t d = plan->D >= 2 ? plan->D - 1 : 0;
    double best = 1e300;
    for (uint32_t j = 0; j < plan->h.nb[d]; ++j) {
        const uint32_t p = plan->order[plan->h.bdep[d][j]];
        const uint32_t a = cs->arc_index(p, plan->order[d]);
        best = std::min(best, (double)cs->arc_weight[a] / std::max(1.0, (double)cs->csize[p]));
    }
    if (plan->h.nb[d] == 0) best = cs->csize[plan->order[0]];
    return best < 3 * scale ? 1 : best < 12 * scale ? 4 : best < 48 * scale ? 16 : 64;
}

// Rows per wave tile for lane width G: `iters` rows per lane group.
inline uint32_t tile_rows_of(uint32_t G, uint32_t iters) { return 64u / G * iters; }

// Rows of width d of this plan are finished by the fused kernel: d = D-1, or d = D-2 with a second
// row-sum table.
inline bool fused_at(const DPlan &P, uint32_t d) {
    return P.fused && (d + 1 == P.D || (P.tail2 && d + 2 == P.D));
}

inline unsigned tiles_grid(uint64_t tiles) {
    return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((tiles + 3) / 4, 1u << 16));
}

template <int G, int KIND>
void launch_bcount(asdm_ctx *ctx, const DPlan *plans, const uint64_t *seg_start, const uint64_t *tile_pref, uint32_t qa,
                   uint32_t qb, uint64_t tiles, uint32_t d, const uint32_t *rows, const asdm_cs::Store *st,
                   uint32_t *counts, unsigned long long *qcount, unsigned long long *bytes) {
    hipLaunchKernelGGL((k_bcount<G, KIND>), dim3(tiles_grid(tiles)), dim3(256), 0, ctx->stream, plans, seg_start,
                       tile_pref, qa, qb, tile_rows_of(G, ctx->tune_tile_iters), d, rows, st->d_ce_off, st->d_ce_val, st->d_cand, counts,
                       qcount, bytes);
}

template <int G>
void launch_bfill(asdm_ctx *ctx, const DPlan *plans, const uint64_t *seg_start, const uint64_t *tile_pref, uint32_t qa,
                  uint32_t qb, uint64_t tiles, uint32_t d, const uint32_t *rows, const asdm_cs::Store *st,
                  const uint64_t *offs, uint32_t *out, unsigned long long *bytes) {
    hipLaunchKernelGGL((k_bfill<G>), dim3(tiles_grid(tiles)), dim3(256), 0, ctx->stream, plans, seg_start, tile_pref,
                       qa, qb, tile_rows_of(G, ctx->tune_tile_iters), d, rows, st->d_ce_off, st->d_ce_val, st->d_cand, offs, out, bytes);
}

template <class F1, class F4, class F16, class F64>
void by_class(uint32_t g, F1 f1, F4 f4, F16 f16, F64 f64) {
    if (g == 1) f1();
    else if (g == 4) f4();
    else if (g == 16) f16();
    else f64();
}

// Level-synchronous batch engine over queries whose candidate spaces share one filter store.
// All queries advance one depth per step: one count launch, one scan, one fill launch per depth
// for the whole batch, and one host synchronization per depth.
// qstatus_out (optional) receives ASDM_OK or, for a query whose partial embeddings exceed the
// context's row limit, ASDM_E_TIMEOUT (its rows are dropped and its count is 0).  Without it a
// query over the limit fails the whole call with ASDM_E_TIMEOUT.
// lat_out[i] (optional) receives now() when query i's count became final: after the host
// synchronization of the depth at which its last rows were counted.  With done_out, a run that
// ends in ASDM_E_BUDGET still returns the counts of the queries that had finished by then
// (done_out[i] = 1), so a caller re-runs only the others.
int batch_engine(asdm_ctx *ctx, const std::vector<asdm_cs *> &css_in, const std::vector<asdm_plan *> &plans_in,
                 uint64_t budget, uint64_t *counts_out, int *qstatus_out, float *lat_out = nullptr,
                 const std::function<float()> *now = nullptr, char *done_out = nullptr) {
    const uint32_t Q = (uint32_t)css_in.size();
    if (Q == 0) return ASDM_OK;
    // queries grouped by lane class, so that each class is a contiguous row range at every depth
    std::vector<uint32_t> perm(Q), cls(Q);
    for (uint32_t i = 0; i < Q; ++i) {
        perm[i] = i;
        cls[i] = lane_class(plans_in[i], css_in[i], ctx->tune_class_scale);
    }
    std::stable_sort(perm.begin(), perm.end(), [&](uint32_t a, uint32_t b) {
        return cls[a] != cls[b] ? cls[a] < cls[b] : plans_in[a]->D < plans_in[b]->D;
    });
    std::vector<asdm_cs *> css(Q);
    std::vector<asdm_plan *> plans(Q);
    std::vector<uint32_t> qcls(Q);
    for (uint32_t i = 0; i < Q; ++i) {
        css[i] = css_in[perm[i]];
        plans[i] = plans_in[perm[i]];
        qcls[i] = cls[perm[i]];
    }
    std::vector<uint64_t> counts(Q);
    const asdm_cs::Store *st = css[0]->store;
    std::vector<DPlan> hp(Q);
    std::vector<uint32_t *> tabs;
    auto cleanup = [&](void *a, void *b, void *c, void *d2, void *e, void *f, void *g2) {
        for (auto t : tabs) dev_free(ctx, t);
        for (void *p : {a, b, c, d2, e, f, g2}) dev_free(ctx, p);
    };
    int rc = ASDM_OK;
    uint32_t maxD = 0;
    std::vector<uint64_t> seg_start_h(Q + 1, 0);
    TableJobs jobs;  // every query's triangle and row-sum tables, built in a few launches below
    for (uint32_t i = 0; i < Q; ++i) {
        const asdm_cs *cs = css[i];
        if (cs->store != st) return ctx->fail(ASDM_E_INVALID, "batch candidate spaces must come from one filter call");
        hp[i] = plans[i]->h;
        if ((rc = build_tri_tables(ctx, plans[i], cs, hp[i], tabs, &jobs))) {
            cleanup(nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
            return rc;
        }
        const uint64_t cb = (uint64_t)(cs->d_cand - st->d_cand), ib = (uint64_t)(cs->d_ce_off - st->d_ce_off);
        for (uint32_t d = 0; d < hp[i].nq; ++d) {
            hp[i].coff[d] += cb;
            for (uint32_t j = 0; j < hp[i].nb[d]; ++j) hp[i].bbase[d][j] += ib;
            if (hp[i].tri[d]) hp[i].tri_base[d] += ib;
        }
        if (hp[i].tail_fix == kTailFixLeaf) hp[i].tail_rbase += ib;
        hp[i].fused = plan_fusable(plans[i]);
        maxD = std::max(maxD, hp[i].D);
        seg_start_h[i + 1] = seg_start_h[i] + cs->csize[plans[i]->order[0]];
    }
    if ((rc = run_tri_jobs(ctx, jobs.tri, tabs)) || (rc = run_edge_jobs(ctx, jobs.edge)) ||
        (rc = run_tail_jobs(ctx, jobs.first, tabs)) || (rc = run_tail_jobs(ctx, jobs.second, tabs))) {
        cleanup(nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
        return rc;
    }
    DPlan *dplans = nullptr;
    uint64_t *seg_start = nullptr, *seg_next = nullptr;
    uint32_t *seg_q = nullptr, *counts_d = nullptr;
    uint64_t *offs = nullptr;
    uint32_t *rows = nullptr;
    rc = dev_alloc(ctx, (void **)&dplans, sizeof(DPlan) * Q);
    if (!rc) rc = dev_alloc(ctx, (void **)&seg_start, 8 * (Q + 1));
    if (!rc) rc = dev_alloc(ctx, (void **)&seg_next, 8 * (Q + 1));
    if (!rc) rc = dev_alloc(ctx, (void **)&seg_q, 4 * Q);
    if (!rc) rc = check_hip(ctx, hipMemcpyAsync(dplans, hp.data(), sizeof(DPlan) * Q, hipMemcpyHostToDevice, ctx->stream), "H2D");
    std::vector<uint32_t> seg_q_h(Q);
    for (uint32_t i = 0; i < Q; ++i) seg_q_h[i] = i;
    if (!rc) rc = check_hip(ctx, hipMemcpyAsync(seg_q, seg_q_h.data(), 4 * Q, hipMemcpyHostToDevice, ctx->stream), "H2D");
    if (!rc) rc = check_hip(ctx, hipMemcpyAsync(seg_start, seg_start_h.data(), 8 * (Q + 1), hipMemcpyHostToDevice, ctx->stream), "H2D");
    if (!rc) rc = fill_u64(ctx, ctx->d_counters, kNumCounters, 0);
    unsigned long long *qcount = nullptr;
    if (!rc) rc = dev_alloc(ctx, (void **)&qcount, 8 * Q);
    if (!rc) rc = fill_u64(ctx, (uint64_t *)qcount, Q, 0);
    uint64_t n = seg_start_h[Q];
    if (!rc) rc = dev_alloc(ctx, (void **)&rows, 4 * std::max<uint64_t>(n, 1));
    if (!rc && n) {
        hipLaunchKernelGGL(k_biota, dim3(groups_grid(n, 1)), dim3(256), 0, ctx->stream, seg_start, Q, rows, n);
        rc = check_hip(ctx, hipGetLastError(), "biota");
    }
    unsigned long long *d_bytes = (unsigned long long *)ctx->d_counters + 1;
    // Calls launch(class, first query, end query) for each maximal run of queries (in batch order)
    // that share a class; queries without rows at this depth, or whose rows need no launch (skip),
    // join the run around them.
    auto for_runs = [&](auto cls_of, auto skip, auto launch) {
        uint32_t a = 0;
        auto idle = [&](uint32_t i) { return seg_start_h[i + 1] == seg_start_h[i] || skip(i); };
        while (a < Q) {
            while (a < Q && idle(a)) ++a;
            if (a >= Q) break;
            const uint32_t c = cls_of(a);
            uint32_t b = a + 1;
            while (b < Q && (idle(b) || cls_of(b) == c)) ++b;
            launch(c, a, b);
            a = b;
        }
    };
    // Tile prefix of one phase: query i's rows in tiles of tile_rows_of(lanes(i)); skipped queries
    // get no tiles.  Uploaded into buf from host (kept alive until the next stream sync).
    auto tile_prefix = [&](auto rows_per_tile, auto skip, std::vector<uint64_t> &h, uint64_t *buf) {
        h.assign(Q + 1, 0);
        for (uint32_t i = 0; i < Q; ++i) {
            const uint64_t r = seg_start_h[i + 1] - seg_start_h[i];
            const uint64_t tr = rows_per_tile(i);
            h[i + 1] = h[i] + (skip(i) ? 0 : (r + tr - 1) / tr);
        }
        return check_hip(ctx, hipMemcpyAsync(buf, h.data(), 8 * (Q + 1), hipMemcpyHostToDevice, ctx->stream), "H2D");
    };
    std::vector<uint64_t> tp_count_h, tp_fill_h;
    uint64_t *tp_count = nullptr, *tp_fill = nullptr;
    if (!rc) rc = dev_alloc(ctx, (void **)&tp_count, 8 * (Q + 1));
    if (!rc) rc = dev_alloc(ctx, (void **)&tp_fill, 8 * (Q + 1));
    std::vector<uint64_t> qrows(Q, 0);
    std::vector<char> dropped(Q, 0);
    std::vector<char> fin(Q, 0);  // count final (known to the host)
    auto stamp = [&](uint32_t i) {
        fin[i] = 1;
        if (lat_out && now) lat_out[perm[i]] = (*now)();
    };
    for (uint32_t d = 1; d <= maxD && n && !rc; ++d) {
        dev_free(ctx, counts_d);
        dev_free(ctx, offs);
        counts_d = nullptr;
        offs = nullptr;
        // a level at which no query expands is the last: its rows (leaf-depth or fused) need no
        // child counts, no scan and no next level
        bool any_expand = false;
        for (uint32_t i = 0; i < Q && !any_expand; ++i)
            any_expand = seg_start_h[i + 1] > seg_start_h[i] && hp[i].D != d && !fused_at(hp[i], d);
        if (any_expand) {
            rc = dev_alloc(ctx, (void **)&counts_d, 4 * n);
            if (!rc) rc = dev_alloc(ctx, (void **)&offs, 8 * (n + 1));
            if (rc) break;
        }
        {
            // one launch per run of queries with the same row kind and lane width at this depth:
            // leaf-depth rows (inclusion-exclusion over leaf lists) use tune_leaf_lanes, the others
            // their class
            auto kind_of = [&](uint32_t i) { return hp[i].D == d ? kLeafRows : fused_at(hp[i], d) ? kFusedRows : kExpandRows; };
            // fused rows run flattened (lane width 0 in the key: 64-row tiles, all lanes pooled)
            auto lanes = [&](uint32_t i) {
                const int k = kind_of(i);
                return k == kLeafRows ? ctx->tune_leaf_lanes : k == kFusedRows ? 0u : qcls[i];
            };
            auto tile_rows = [&](uint32_t i) {
                return lanes(i) ? (uint64_t)tile_rows_of(lanes(i), ctx->tune_tile_iters) : (uint64_t)64;
            };
            if ((rc = tile_prefix(tile_rows, [](uint32_t) { return false; }, tp_count_h, tp_count))) break;
            for_runs(
                [&](uint32_t i) { return 256u * kind_of(i) + lanes(i); },
                [](uint32_t) { return false; },
                [&](uint32_t key, uint32_t qa, uint32_t qb) {
                    const uint32_t c = key % 256;
                    const uint64_t tiles = tp_count_h[qb] - tp_count_h[qa];
                    if (key / 256 == kFusedRows) {
                        unsigned long long *kb = d_bytes + kFlatFusedSlot;
                        KTimer kt(ctx, "batch_count_fused_flat");
                        hipLaunchKernelGGL(k_bfused_flat, dim3(tiles_grid(tiles)), dim3(256), 0, ctx->stream, dplans,
                                           seg_start, tp_count, qa, qb, d, rows, st->d_ce_off, st->d_ce_val,
                                           st->d_cand, counts_d, qcount, kb);
                        kt.stop();
                        return;
                    }
                    auto go = [&](auto kind) {
                        constexpr int K = decltype(kind)::value;
                        unsigned long long *kb = d_bytes + count_slot(K, c);
                        KTimer kt(ctx, kKindTimer[K][lane_index(c)]);
                        by_class(c,
                                 [&] { launch_bcount<1, K>(ctx, dplans, seg_start, tp_count, qa, qb, tiles, d, rows, st, counts_d, qcount, kb); },
                                 [&] { launch_bcount<4, K>(ctx, dplans, seg_start, tp_count, qa, qb, tiles, d, rows, st, counts_d, qcount, kb); },
                                 [&] { launch_bcount<16, K>(ctx, dplans, seg_start, tp_count, qa, qb, tiles, d, rows, st, counts_d, qcount, kb); },
                                 [&] { launch_bcount<64, K>(ctx, dplans, seg_start, tp_count, qa, qb, tiles, d, rows, st, counts_d, qcount, kb); });
                        kt.stop();
                    };
                    if (key / 256 == kLeafRows) go(std::integral_constant<int, kLeafRows>{});
                    else go(std::integral_constant<int, kExpandRows>{});
                });
        }
        if ((rc = check_hip(ctx, hipGetLastError(), "batch count"))) break;
        if (!any_expand) break;
        if ((rc = scan_u32_to_u64(ctx, counts_d, offs, n))) break;
        hipLaunchKernelGGL(k_bgather, dim3(1), dim3(256), 0, ctx->stream, offs, seg_start, Q, seg_next);
        std::vector<uint64_t> next_h(Q + 1);
        if ((rc = check_hip(ctx, hipMemcpyAsync(next_h.data(), seg_next, 8 * (Q + 1), hipMemcpyDeviceToHost, ctx->stream), "D2H"))) break;
        if ((rc = check_hip(ctx, hipStreamSynchronize(ctx->stream), "sync"))) break;
        if (ctx->row_limit && d < maxD) {  // per-query row limit: drop the queries over it
            std::vector<uint32_t> over;
            for (uint32_t i = 0; i < Q; ++i) {
                if (dropped[i]) continue;
                qrows[i] += next_h[i + 1] - next_h[i];
                if (qrows[i] > ctx->row_limit) over.push_back(i);
            }
            if (!over.empty()) {
                if (!qstatus_out) {
                    rc = ctx->fail(ASDM_E_TIMEOUT, "row limit reached in a batched query");
                    break;
                }
                for (uint32_t i : over) {
                    dropped[i] = 1;
                    if (seg_start_h[i + 1] > seg_start_h[i] &&
                        (rc = fill_u32(ctx, counts_d + seg_start_h[i], seg_start_h[i + 1] - seg_start_h[i], 0)))
                        break;
                }
                if (rc || (rc = scan_u32_to_u64(ctx, counts_d, offs, n))) break;
                hipLaunchKernelGGL(k_bgather, dim3(1), dim3(256), 0, ctx->stream, offs, seg_start, Q, seg_next);
                if ((rc = check_hip(ctx, hipMemcpyAsync(next_h.data(), seg_next, 8 * (Q + 1), hipMemcpyDeviceToHost, ctx->stream), "D2H"))) break;
                if ((rc = check_hip(ctx, hipStreamSynchronize(ctx->stream), "sync"))) break;
            }
        }
        // after this synchronization every query whose rows were counted at depth d, or that has
        // no rows at the next depth, has its final count
        for (uint32_t i = 0; i < Q; ++i)
            if (!fin[i] && (dropped[i] || hp[i].D == d || fused_at(hp[i], d) || next_h[i + 1] == next_h[i])) stamp(i);
        const uint64_t total = next_h[Q];
        if (total == 0 || d == maxD) break;
        if (4ull * (d + 1) * total > budget) {
            rc = ctx->fail(ASDM_E_BUDGET, "batch frontier exceeds the budget");
            break;
        }
        uint32_t *next_rows = nullptr;
        if ((rc = dev_alloc(ctx, (void **)&next_rows, 4ull * (d + 1) * total))) break;
        {
            // rows of leaf-depth and fused queries have no children: they get no tiles and join
            // any run
            auto finished = [&](uint32_t i) { return hp[i].D == d || fused_at(hp[i], d); };
            if ((rc = tile_prefix([&](uint32_t i) { return (uint64_t)tile_rows_of(qcls[i], ctx->tune_tile_iters); },
                                  finished, tp_fill_h, tp_fill)))
                break;
            for_runs(
                [&](uint32_t i) { return qcls[i]; }, finished,
                [&](uint32_t c, uint32_t qa, uint32_t qb) {
                    const uint64_t tiles = tp_fill_h[qb] - tp_fill_h[qa];
                    unsigned long long *fb = d_bytes + fill_slot(c);
                    KTimer kt(ctx, kFillTimer[lane_index(c)]);
                    by_class(c,
                             [&] { launch_bfill<1>(ctx, dplans, seg_start, tp_fill, qa, qb, tiles, d, rows, st, offs, next_rows, fb); },
                             [&] { launch_bfill<4>(ctx, dplans, seg_start, tp_fill, qa, qb, tiles, d, rows, st, offs, next_rows, fb); },
                             [&] { launch_bfill<16>(ctx, dplans, seg_start, tp_fill, qa, qb, tiles, d, rows, st, offs, next_rows, fb); },
                             [&] { launch_bfill<64>(ctx, dplans, seg_start, tp_fill, qa, qb, tiles, d, rows, st, offs, next_rows, fb); });
                    kt.stop();
                });
        }
        if ((rc = check_hip(ctx, hipGetLastError(), "batch fill"))) break;
        dev_free(ctx, rows);
        rows = next_rows;
        std::swap(seg_start, seg_next);
        seg_start_h = next_h;
        n = total;
    }
    std::vector<uint64_t> qc(Q);
    uint64_t kbytes[kNumCounters] = {0};
    const bool keep_done = rc == ASDM_E_BUDGET && done_out;
    int rc2 = rc;
    if (!rc || keep_done) {
        rc2 = check_hip(ctx, hipMemcpyAsync(qc.data(), qcount, 8 * Q, hipMemcpyDeviceToHost, ctx->stream), "D2H");
        if (!rc2) rc2 = check_hip(ctx, hipMemcpyAsync(kbytes, ctx->d_counters, 8 * kNumCounters, hipMemcpyDeviceToHost, ctx->stream), "D2H");
        if (!rc2) rc2 = check_hip(ctx, hipStreamSynchronize(ctx->stream), "sync");
        if (!rc) rc = rc2;
    }
    if (keep_done && !rc2)
        for (uint32_t i = 0; i < Q; ++i)
            if (fin[i]) {
                done_out[perm[i]] = 1;
                counts_out[perm[i]] = dropped[i] ? 0 : qc[i];
                if (qstatus_out) qstatus_out[perm[i]] = dropped[i] ? ASDM_E_TIMEOUT : ASDM_OK;
            }
    if (!rc && ctx->profiling) {
        for (int k = 0; k < 3; ++k)
            for (int gi = 0; gi < 4; ++gi)
                if (kbytes[1 + k * 4 + gi]) add_kernel_bytes(ctx, kKindTimer[k][gi], kbytes[1 + k * 4 + gi]);
        for (int gi = 0; gi < 4; ++gi)
            if (kbytes[13 + gi]) add_kernel_bytes(ctx, kFillTimer[gi], kbytes[13 + gi]);
        if (kbytes[1 + kFlatFusedSlot]) add_kernel_bytes(ctx, "batch_count_fused_flat", kbytes[1 + kFlatFusedSlot]);
    }
    if (!rc)
        for (uint32_t i = 0; i < Q; ++i) {
            if (!fin[i]) stamp(i);
            if (done_out) done_out[perm[i]] = 1;
            counts_out[perm[i]] = dropped[i] ? 0 : qc[i];
            if (qstatus_out) qstatus_out[perm[i]] = dropped[i] ? ASDM_E_TIMEOUT : ASDM_OK;
        }
    resolve_timers(ctx);
    cleanup(dplans, seg_start, seg_next, seg_q, counts_d, offs, rows);
    dev_free(ctx, qcount);
    dev_free(ctx, tp_count);
    dev_free(ctx, tp_fill);
    (void)hipStreamSynchronize(ctx->stream);
    return rc;
}

}  // namespace
}  // namespace asdm

using namespace asdm;

extern "C" {

int asdm_ctx_set_batch_tuning(asdm_ctx *ctx, uint32_t leaf_lanes, double class_scale, uint32_t tile_iters) try {
    if (!ctx || !(leaf_lanes == 1 || leaf_lanes == 4 || leaf_lanes == 16 || leaf_lanes == 64) || !(class_scale > 0) ||
        tile_iters < 1 || tile_iters > 64)
        return ASDM_E_INVALID;
    ctx->tune_leaf_lanes = leaf_lanes;
    ctx->tune_class_scale = class_scale;
    ctx->tune_tile_iters = tile_iters;
    return ASDM_OK;
} ASDM_CTX_CATCH(ctx)

int asdm_ctx_get_batch_tuning(const asdm_ctx *ctx, uint32_t *leaf_lanes, double *class_scale, uint32_t *tile_iters) try {
    if (!ctx) return ASDM_E_INVALID;
    if (leaf_lanes) *leaf_lanes = ctx->tune_leaf_lanes;
    if (class_scale) *class_scale = ctx->tune_class_scale;
    if (tile_iters) *tile_iters = ctx->tune_tile_iters;
    return ASDM_OK;
} ASDM_HOST_CATCH

// Autotuner of the batch launch shape: after one untimed run of the caller's batch, every shape of
// a small grid runs it `reps` times; the shape with the smallest median time is kept on the context.
// Every run's counts must equal the first run's (they are plan- and shape-independent).
int asdm_batch_autotune(asdm_ctx *ctx, const asdm_graph *g, const asdm_graph *const *qs, uint32_t nq,
                        uint64_t budget_bytes, asdm_mpnn *m, uint32_t reps, uint32_t *leaf_lanes, double *class_scale,
                        uint32_t *tile_iters, float *ms_best) try {
    if (!ctx || !g || (!qs && nq)) return ASDM_E_INVALID;
    reps = std::max<uint32_t>(1, reps);
    const uint32_t save_l = ctx->tune_leaf_lanes, save_t = ctx->tune_tile_iters;
    const double save_s = ctx->tune_class_scale;
    std::vector<uint64_t> ref(nq), cur(nq);
    int rc = asdm_batch_count_learned(ctx, g, qs, nq, budget_bytes, m, ref.data(), nullptr, nullptr);
    if (rc) return rc;
    double best_ms = 1e300;
    uint32_t bl = save_l, bt = save_t;
    double bs = save_s;
    for (uint32_t l : {4u, 16u, 64u})
        for (double sc : {0.25, 0.5, 1.0, 2.0})
            for (uint32_t ti : {2u, 4u, 8u}) {
                ctx->tune_leaf_lanes = l;
                ctx->tune_class_scale = sc;
                ctx->tune_tile_iters = ti;
                std::vector<double> t;
                for (uint32_t r = 0; r < reps && !rc; ++r) {
                    auto t0 = std::chrono::steady_clock::now();
                    rc = asdm_batch_count_learned(ctx, g, qs, nq, budget_bytes, m, cur.data(), nullptr, nullptr);
                    t.push_back(std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
                    if (!rc && cur != ref) rc = ctx->fail(ASDM_E_INVALID, "autotune: counts differ between launch shapes");
                }
                if (rc) break;
                std::sort(t.begin(), t.end());
                const double med = t[t.size() / 2];
                if (med < best_ms) {
                    best_ms = med;
                    bl = l;
                    bs = sc;
                    bt = ti;
                }
            }
    ctx->tune_leaf_lanes = rc ? save_l : bl;
    ctx->tune_class_scale = rc ? save_s : bs;
    ctx->tune_tile_iters = rc ? save_t : bt;
    if (rc) return rc;
    if (leaf_lanes) *leaf_lanes = bl;
    if (class_scale) *class_scale = bs;
    if (tile_iters) *tile_iters = bt;
    if (ms_best) *ms_best = (float)best_ms;
    return ASDM_OK;
} ASDM_CTX_CATCH(ctx)

int asdm_count(asdm_ctx *ctx, const asdm_plan *plan, const asdm_cs *cs, uint64_t budget_bytes, uint64_t *count) try {
    return run_engine(ctx, plan, cs, budget_bytes, nullptr, count);
} ASDM_CTX_CATCH(ctx)

int asdm_local_count(asdm_ctx *ctx, const asdm_plan *plan, const asdm_cs *cs, uint64_t budget_bytes, uint64_t *mat,
                     uint64_t *count) try {
    if (!mat) return ASDM_E_INVALID;
    return run_engine(ctx, plan, cs, budget_bytes, mat, count);
} ASDM_CTX_CATCH(ctx)

int asdm_enumerate(asdm_ctx *ctx, const asdm_plan *plan, const asdm_cs *cs, uint64_t budget_bytes, uint32_t *out,
                   uint64_t cap_rows, uint64_t *count) try {
    return run_engine(ctx, plan, cs, budget_bytes, nullptr, count, true, out, out ? cap_rows : 0);
} ASDM_CTX_CATCH(ctx)


namespace asdm {
// Learned engine choice (asdm_batch_count_learned): a query the plain rule sends to the bag engine
// runs on the frontier when its learned row estimate is at most kLearnedFrontierRows and the
// frontier's modelled time (rows x kFrontierSecondsPerRow) is below the bag plan's (cost x
// kBagSecondsPerCost, both measured on one MI355X); it falls back to bags past kLearnedRowCap rows.
constexpr double kLearnedFrontierRows = 16777216.0;
constexpr double kFrontierSecondsPerRow = 1.0e-9;
constexpr double kBagSecondsPerCost = 2.0e-13;
constexpr uint64_t kLearnedRowCap = 1ull << 26;
}  // namespace asdm

// Batched multi-query count: the filter of all queries runs as one batched pass (filter_many);
// the queries are then counted together by the level-synchronous batch engine.  A query whose
// plan is not supported by the batch engine, or a batch whose frontier exceeds the budget, is
// counted with the single-query engine instead.  lat_ms[i] is the time at which query i's count
// was known, from the call start.
int asdm_batch_count(asdm_ctx *ctx, const asdm_graph *g, const asdm_graph *const *qs, uint32_t nq, uint64_t budget_bytes,
                     uint64_t *counts, float *lat_ms) try {
    return asdm_batch_count_learned(ctx, g, qs, nq, budget_bytes, nullptr, counts, nullptr, lat_ms);
} ASDM_CTX_CATCH(ctx)

// The same with the plans of all queries ordered by the learned estimator (one lock-step greedy
// over the batch); m == NULL gives the plain planner.  With qstatus, a query that fails on its own
// (row limit, unsupported shape) gets its status there and the others are still counted.
int asdm_batch_count_learned(asdm_ctx *ctx, const asdm_graph *g, const asdm_graph *const *qs, uint32_t nq,
                             uint64_t budget_bytes, asdm_mpnn *m, uint64_t *counts, int *qstatus, float *lat_ms) try {
    if (!ctx || !g || (!qs && nq) || (!counts && nq)) return ASDM_E_INVALID;
    if (g->dev.device != ctx->device) return ctx->fail(ASDM_E_INVALID, "data graph not uploaded to this device");
    (void)hipSetDevice(ctx->device);
    auto t0 = std::chrono::steady_clock::now();
    auto now_ms = [&]() {
        return (float)std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    };
    auto per_query = [](int code) { return code == ASDM_E_TIMEOUT || code == ASDM_E_UNSUPPORTED; };
    std::vector<asdm_cs *> cs(nq, nullptr);
    std::vector<asdm_plan *> bplans;
    // frees the plans, then the candidate spaces, on every return (an exception's included)
    struct Release {
        std::vector<asdm_cs *> &cs;
        std::vector<asdm_plan *> &plans;
        ~Release() {
            for (auto p : plans) asdm_plan_free(p);
            for (auto c : cs) asdm_cs_free(c);
        }
    } release{cs, bplans};
    std::vector<int> status(nq, ASDM_E_INVALID);
    float phase = now_ms();
    auto host_phase = [&](const char *name) {
        const float t = now_ms();
        add_host_ms(ctx, name, t - phase);
        phase = t;
    };
    int rc = filter_many(ctx, g, qs, nq, cs.data(), status.data());
    if (rc) return rc;
    host_phase("host: batch filter");
    std::vector<asdm_cs *> bcs;
    std::vector<uint32_t> bidx;
    std::vector<std::shared_ptr<BagPlan>> fallback(nq);  // bag plans of queries moved to the frontier
    for (uint32_t i = 0; i < nq; ++i) {
        counts[i] = 0;
        if (qstatus) qstatus[i] = status[i];
        if (status[i] == ASDM_E_EMPTY_CANDIDATES) {
            if (lat_ms) lat_ms[i] = now_ms();
            continue;
        }
        if (status[i] != ASDM_OK) {
            if (qstatus && per_query(status[i])) continue;
            rc = status[i];
            break;
        }
        if (!m) {
            asdm_plan *p = nullptr;
            if ((rc = asdm_plan_create(qs[i], cs[i], nullptr, 0, &p))) {
                if (qstatus && per_query(rc)) {
                    qstatus[i] = rc;
                    rc = ASDM_OK;
                    continue;
                }
                break;
            }
            bplans.push_back(p);
        }
        bcs.push_back(cs[i]);
        bidx.push_back(i);
    }
    if (!rc && m) {
        host_phase("host: plans");
        rc = learned_plans(m, g, std::vector<const asdm_cs *>(bcs.begin(), bcs.end()), 0, bplans);
        host_phase("host: learned plans");
        // Engine choice.  The plain rule (asdm_plan_create's) sends deep or large plans to the bag
        // engine, because the independence estimates of the frontier's rows are unreliable there.
        // The learned estimates are not: a query whose learned row estimate makes the frontier
        // cheaper than its bag plan runs on the frontier instead, under a row cap, and falls back
        // to the bag plan if it reaches the cap.
        // host-only planning, one query per task over a few host threads; the first failing query
        // in batch order fails the call, as a serial loop would
        const size_t nb = bplans.size();
        std::vector<int> prc(nb, ASDM_OK);
        std::vector<std::string> perr(nb);
        std::atomic<size_t> next_k{0};
        auto choose = [&]() {
            for (size_t k; (k = next_k.fetch_add(1)) < nb;) {
                try {  // an exception must not leave a helper thread
                    asdm_plan plain;
                    int r = build_plan(qs[bidx[k]], bcs[k], nullptr, 0, [](uint64_t, double fb) { return fb; }, &plain,
                                       perr[k]);
                    if (!r) r = attach_bags(bcs[k], 0, false, &plain, perr[k]);
                    prc[k] = r;
                    if (r || !plain.bags) continue;
                    const double rows = bplans[k]->est_total;
                    if (rows <= kLearnedFrontierRows && rows * kFrontierSecondsPerRow < plain.bags->cost * kBagSecondsPerCost)
                        fallback[k] = plain.bags;
                    else
                        bplans[k]->bags = plain.bags;
                } catch (const std::bad_alloc &) {
                    prc[k] = ASDM_E_OOM;
                    perr[k] = "host allocation failed";
                } catch (...) {
                    prc[k] = ASDM_E_INVALID;
                    perr[k] = "unexpected host exception in planning";
                }
            }
        };
        const size_t nthreads = std::min<size_t>({(size_t)8, std::max(1u, std::thread::hardware_concurrency()), (nb + 15) / 16});
        std::vector<std::thread> helpers;
        for (size_t t = 1; t < nthreads; ++t) helpers.emplace_back(choose);
        choose();
        for (auto &h : helpers) h.join();
        for (size_t k = 0; k < nb; ++k)
            if (prc[k]) {
                rc = prc[k];
                ctx->err = perr[k];
                break;
            }
        uint64_t nmoved = 0;
        for (auto &f : fallback) nmoved += f != nullptr;
        add_kernel_count(ctx, "learned: bag queries moved to the frontier", nmoved, 0);
        host_phase("host: engine choice");
    } else {
        host_phase("host: plans");
    }
    if (!rc) {
        const uint64_t budget = budget_bytes ? budget_bytes : (16ull << 30);
        std::vector<uint64_t> bc(bcs.size(), 0);
        std::vector<int> bst(bcs.size(), ASDM_OK);
        std::vector<float> blat(bcs.size(), 0);
        const std::function<float()> clock = now_ms;
        // A (sub-)batch whose frontier exceeds the budget keeps the counts of the queries that had
        // finished and splits the others in halves; a single query that still exceeds it runs on
        // the single-query engine, which bounds memory by chunking.
        std::function<int(const std::vector<size_t> &)> run = [&](const std::vector<size_t> &idx) -> int {
            std::vector<asdm_cs *> c;
            std::vector<asdm_plan *> p;
            for (size_t k : idx) {
                c.push_back(bcs[k]);
                p.push_back(bplans[k]);
            }
            std::vector<uint64_t> cnt(idx.size(), 0);
            std::vector<int> st(idx.size(), ASDM_OK);
            std::vector<float> lat(idx.size(), 0);
            std::vector<char> done(idx.size(), 0);
            int r = batch_engine(ctx, c, p, budget, cnt.data(), qstatus ? st.data() : nullptr, lat.data(), &clock,
                                 done.data());
            std::vector<size_t> rest;
            for (size_t j = 0; j < idx.size(); ++j) {
                if (done[j]) {
                    bc[idx[j]] = cnt[j];
                    bst[idx[j]] = st[j];
                    blat[idx[j]] = lat[j];
                } else {
                    rest.push_back(idx[j]);
                }
            }
            if (r != ASDM_E_BUDGET) return r;
            if (rest.size() > 1) {
                const size_t mid = rest.size() / 2;
                if ((r = run(std::vector<size_t>(rest.begin(), rest.begin() + mid)))) return r;
                return run(std::vector<size_t>(rest.begin() + mid, rest.end()));
            }
            for (size_t k : rest) {
                bst[k] = asdm_count(ctx, bplans[k], bcs[k], budget_bytes, &bc[k]);
                blat[k] = now_ms();
                if (bst[k] != ASDM_OK) {
                    bc[k] = 0;
                    if (!(qstatus && per_query(bst[k]))) return bst[k];
                }
            }
            return ASDM_OK;
        };
        // queries whose plan uses bag tables are counted by the bag engine, the others together;
        // queries moved to the frontier by the learned estimates run in a batch of their own under
        // the row cap, and those that reach it join the bag engine's queue
        std::vector<size_t> frontier, moved, bagged;
        for (size_t k = 0; k < bcs.size(); ++k) (bplans[k]->bags ? bagged : fallback[k] ? moved : frontier).push_back(k);
        // the bag-engine queries go to the worker pool first, and are counted while this thread
        // runs the frontier batches on the context's own stream
        BagPool pool(ctx, budget_bytes, clock, bc.data(), bst.data(), blat.data(), qstatus != nullptr);
        if (!bagged.empty() || !moved.empty()) rc = pool.start();
        for (size_t k : bagged) pool.push({k, bplans[k], bcs[k], bplans[k]->bags->cost});
        if (!rc && !frontier.empty()) rc = run(frontier);
        host_phase("host: frontier batch");
        if (!rc && !moved.empty()) {
            const uint64_t saved_limit = ctx->row_limit;
            ctx->row_limit = saved_limit ? std::min(saved_limit, kLearnedRowCap) : kLearnedRowCap;
            std::vector<asdm_cs *> c;
            std::vector<asdm_plan *> p;
            for (size_t k : moved) {
                c.push_back(bcs[k]);
                p.push_back(bplans[k]);
            }
            std::vector<uint64_t> cnt(moved.size(), 0);
            std::vector<int> st(moved.size(), ASDM_OK);
            std::vector<float> lat(moved.size(), 0);
            rc = batch_engine(ctx, c, p, budget, cnt.data(), st.data(), lat.data(), &clock, nullptr);
            ctx->row_limit = saved_limit;
            for (size_t j = 0; j < moved.size() && !rc; ++j) {
                const size_t k = moved[j];
                if (st[j] == ASDM_E_TIMEOUT) {  // over the cap: count it with its bag plan
                    bplans[k]->bags = fallback[k];
                    add_kernel_count(ctx, "learned: moved queries back to bags (row cap)", 1, 0);
                    pool.push({k, bplans[k], bcs[k], fallback[k]->cost});
                } else {
                    bc[k] = cnt[j];
                    bst[k] = st[j];
                    blat[k] = lat[j];
                }
            }
            if (rc == ASDM_E_BUDGET) {  // too large to batch: each moved query on its bag plan
                rc = ASDM_OK;
                for (size_t k : moved) {
                    bplans[k]->bags = fallback[k];
                    pool.push({k, bplans[k], bcs[k], fallback[k]->cost});
                }
            }
        }
        host_phase("host: moved queries");
        const int prc = pool.finish();
        if (!rc) rc = prc;
        host_phase("host: bag pool wait");
        for (size_t k = 0; k < bcs.size(); ++k) {
            counts[bidx[k]] = bc[k];
            if (qstatus) qstatus[bidx[k]] = bst[k];
            if (lat_ms) lat_ms[bidx[k]] = blat[k];
        }
    }
    return rc;
} ASDM_CTX_CATCH(ctx)
}  // extern "C"
