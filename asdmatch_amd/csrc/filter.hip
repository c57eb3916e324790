// CFL candidate filtering and candidate-edge index on the GPU.
//
// Semantics follow CandidateSpace::buildCandCFL (relation/candidate_space.cpp:550-721) in set
// form.  With B(u) = {v : L(v) = L(u), deg(v) >= deg(u), NLF(u, v)} (utility/utils.cpp:30-40) and
// the reference's parent arrays, whose zero-filled tail makes data vertex 0 a parent candidate
// whenever |Gen(p)| < |V_L(p)|:
//     P*(p)  = Gen(p) + {0 if |Gen(p)| < |V_L(p)|}
//     Gen(u) = B(u)   ∩ { v : for every earlier neighbour p, v is adjacent to some w in P*(p) }
//     C(u)   = Gen(u) ∩ { v : for every later neighbour p,   v is adjacent to some w in P*(p) }
// where "earlier/later" is the CFL BFS order from the CFL root (:723-816).  Every set is a
// bitmap over data vertices (one 64-bit word per wave of 64 vertices, written from a ballot),
// and each test is a pull: thread v scans its own label-L(p) adjacency slice for a member of
// P*(p) and stops at the first hit, so no atomics touch the bitmaps.
//
// Candidate edges (buildCandidateEdge(false), :458-487): for query arc (u, u2) and v in C(u), the
// label-L(u2) slice of N(v) filtered by C(u2), stored as positions in C(u2) (ascending, because
// slices are ascending by id) in one CSR per arc: count pass, global scan, fill pass.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <memory>

#include "engine.h"
#include "kernels_common.h"

namespace asdm {

struct QV {
    uint32_t lab, deg, nnlc;
    uint32_t nlc_lab[kMaxQ], nlc_cnt[kMaxQ];
};

struct Par {  // one parent test of a prune pass
    uint32_t p;        // parent slot (query vertex of one query in the batch)
    uint32_t lab;      // L(p)
    uint32_t labcount; // |V_L(p)|
};

// Word-parallel filter kernels: one thread per 64-vertex bitmap word of one query vertex (slot).
// A thread visits only the set bits of its source word (the label bitmap for the base filter, the
// previous pass for a prune), assembles its output word in a register and stores it; popcounts go
// to the slot's counter with one atomic per wave.

__global__ void k_base(const uint64_t *__restrict__ labbits, const uint32_t *__restrict__ goff,
                       const uint32_t *__restrict__ loff, uint32_t L, const QV *__restrict__ qv,
                       uint64_t *__restrict__ bits, uint32_t W, uint32_t *__restrict__ cnt) {
    const uint32_t u = blockIdx.y;
    const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    const QV &q = qv[u];
    uint64_t res = 0;
    if (w < W && q.lab < L) {
        uint64_t word = labbits[(size_t)q.lab * W + w];
        while (word) {
            const int j = __ffsll((unsigned long long)word) - 1;
            word &= word - 1;
            const uint32_t v = w * 64 + (uint32_t)j;
            if (goff[v + 1] - goff[v] < q.deg) continue;
            const uint32_t *lo = loff + (size_t)v * (L + 1);
            bool pass = true;
            for (uint32_t k = 0; k < q.nnlc && pass; ++k) {
                const uint32_t l = q.nlc_lab[k];
                pass = l < L && lo[l + 1] - lo[l] >= q.nlc_cnt[k];
            }
            if (pass) res |= 1ull << j;
        }
    }
    if (w < W) bits[(size_t)u * W + w] = res;
    const uint64_t c = group_sum<64>((uint64_t)__popcll(res));
    if ((threadIdx.x & 63) == 0 && c) atomicAdd(&cnt[u], (uint32_t)c);
}

// dst(u) = src(u) ∩ every parent test listed for u.  grid.y indexes (u, parent-list) pairs.
__global__ void k_prune(const uint32_t *__restrict__ nbr, const uint32_t *__restrict__ loff, uint32_t n, uint32_t L,
                        uint32_t lab0, const uint64_t *__restrict__ src, uint64_t *__restrict__ dst,
                        const uint64_t *__restrict__ genbits, const uint32_t *__restrict__ gcnt, uint32_t W,
                        const uint32_t *__restrict__ us, const uint32_t *__restrict__ par_off,
                        const Par *__restrict__ pars, uint32_t *__restrict__ dcnt) {
    const uint32_t slot = blockIdx.y;
    const uint32_t u = us[slot];
    const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t res = 0;
    if (w < W) {
        uint64_t word = src[(size_t)u * W + w];
        const uint32_t p0 = par_off[slot], p1 = par_off[slot + 1];
        while (word) {
            const int j = __ffsll((unsigned long long)word) - 1;
            word &= word - 1;
            const uint32_t v = w * 64 + (uint32_t)j;
            const uint32_t *lo = loff + (size_t)v * (L + 1);
            // v adjacent to data vertex 0: 0 would lead v's label-L(0) slice (slices ascend by id)
            const bool adj0 = lab0 < L && lo[lab0 + 1] > lo[lab0] && nbr[lo[lab0]] == 0;
            bool pass = true;
            for (uint32_t k = p0; k < p1 && pass; ++k) {
                const Par pr = pars[k];
                const bool pad = gcnt[pr.p] < pr.labcount;
                bool hit = pad && adj0;
                if (!hit && pr.lab < L) {
                    const uint64_t *gb = genbits + (size_t)pr.p * W;
                    for (uint32_t i = lo[pr.lab], e = lo[pr.lab + 1]; i < e; ++i) {
                        const uint32_t x = nbr[i];
                        if ((gb[x >> 6] >> (x & 63)) & 1ull) {
                            hit = true;
                            break;
                        }
                    }
                }
                pass = hit;
            }
            if (pass) res |= 1ull << j;
        }
        dst[(size_t)u * W + w] = res;
    }
    const uint64_t c = group_sum<64>((uint64_t)__popcll(res));
    if ((threadIdx.x & 63) == 0 && c) atomicAdd(&dcnt[u], (uint32_t)c);
}

// Per query vertex: rank[w] = set bits before word w; cand[cand_off[u] + rank] = ascending ids.
__global__ void k_compact(const uint64_t *__restrict__ bits, uint32_t W, uint32_t *__restrict__ rank,
                          const uint64_t *__restrict__ cand_off, uint32_t *__restrict__ cand,
                          const uint32_t *__restrict__ slots) {
    __shared__ uint64_t lds[1024 / 64 + 1];
    const uint32_t u = slots[blockIdx.x];
    const uint64_t *b = bits + (size_t)u * W;
    uint32_t *out = cand + cand_off[u];
    uint64_t carry = 0;
    for (uint32_t base = 0; base < W; base += 1024) {
        const uint32_t w = base + threadIdx.x;
        const uint64_t word = w < W ? b[w] : 0;
        uint64_t tot;
        const uint64_t ex = block_exclusive_scan<1024>((uint64_t)__popcll(word), lds, tot) + carry;
        if (w < W) {
            rank[(size_t)u * W + w] = (uint32_t)ex;
            uint64_t x = word;
            uint64_t k = ex;
            while (x) {
                int j = __ffsll((unsigned long long)x) - 1;
                out[k++] = w * 64 + (uint32_t)j;
                x &= x - 1;
            }
        }
        carry += tot;
    }
}

struct ArcInfo {
    uint32_t u, u2, lab2;
    uint64_t item_base;  // first item (= position 0 of C(u)) of this arc in the flat item space
    uint64_t cu;         // offset of C(u) in the candidate array
};

__device__ inline uint32_t arc_of_item(const ArcInfo *__restrict__ arcs, uint32_t narcs, uint64_t k) {
    uint32_t lo = 0, hi = narcs;  // last a with item_base <= k
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (arcs[mid].item_base <= k) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Count (FILL = false) or write (FILL = true) the candidate-edge lists.  One aligned group of G
// lanes per (arc, candidate) item; lanes stride over the label slice and test C(u2) bits.
template <int G, bool FILL>
__global__ void k_cand_edges(const uint32_t *__restrict__ nbr, const uint32_t *__restrict__ loff, uint32_t L,
                             const uint32_t *__restrict__ cand, const ArcInfo *__restrict__ arcs, uint32_t narcs,
                             uint64_t nitems, const uint64_t *__restrict__ bits, const uint32_t *__restrict__ rank,
                             uint32_t W, uint32_t *__restrict__ counts, const uint64_t *__restrict__ offs,
                             uint32_t *__restrict__ vals) {
    const int lg = threadIdx.x & (G - 1);
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x / G;
    for (uint64_t k = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / G; k < nitems; k += stride) {
        const uint32_t a = arc_of_item(arcs, narcs, k);
        const ArcInfo ai = arcs[a];
        const uint32_t v = cand[ai.cu + (k - ai.item_base)];
        uint32_t b = 0, e = 0;
        if (ai.lab2 < L) {
            b = loff[(size_t)v * (L + 1) + ai.lab2];
            e = loff[(size_t)v * (L + 1) + ai.lab2 + 1];
        }
        const uint64_t *cb = bits + (size_t)ai.u2 * W;
        if constexpr (!FILL) {
            uint32_t c = 0;
            for (uint32_t i = b + lg; i < e; i += G) {
                uint32_t w = nbr[i];
                c += (uint32_t)((cb[w >> 6] >> (w & 63)) & 1ull);
            }
            c = (uint32_t)group_sum<G>(c);
            if (lg == 0) counts[k] = c;
        } else {
            const uint32_t *rk = rank + (size_t)ai.u2 * W;
            uint64_t o = offs[k];
            for (uint32_t i0 = b; i0 < e; i0 += G) {
                const uint32_t i = i0 + lg;
                bool ok = false;
                uint32_t pos = 0;
                if (i < e) {
                    uint32_t w = nbr[i];
                    uint64_t word = cb[w >> 6];
                    ok = (word >> (w & 63)) & 1ull;
                    pos = rk[w >> 6] + (uint32_t)__popcll(word & ((1ull << (w & 63)) - 1));
                }
                const uint64_t gm = group_bits<G>(__ballot(ok));
                if (ok) vals[o + __popcll(gm & ((1ull << lg) - 1))] = pos;
                o += __popcll(gm);
            }
        }
    }
}

// out[a] = offs[item_base[a]] for a in [0, n): the first candidate-edge offset of each arc.
__global__ void k_gather_u64(const uint64_t *__restrict__ offs, const uint64_t *__restrict__ idx, uint32_t n,
                             uint64_t *__restrict__ out) {
    for (uint32_t a = blockIdx.x * blockDim.x + threadIdx.x; a < n; a += gridDim.x * blockDim.x) out[a] = offs[idx[a]];
}

}  // namespace asdm

using namespace asdm;

namespace {

struct DevBuf {
    asdm_ctx *ctx;
    void *p = nullptr;
    explicit DevBuf(asdm_ctx *c) : ctx(c) {}
    ~DevBuf() { dev_free(ctx, p); }
    template <class T>
    T *as() { return (T *)p; }
};

int upload(asdm_ctx *ctx, void *dst, const void *src, size_t bytes) {
    return check_hip(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream), "H2D copy");
}
int download(asdm_ctx *ctx, void *dst, const void *src, size_t bytes) {
    int rc = check_hip(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream), "D2H copy");
    if (rc) return rc;
    return check_hip(ctx, hipStreamSynchronize(ctx->stream), "stream sync");
}

}  // namespace

extern "C" {

void asdm_cs_free(asdm_cs *cs) {
    if (!cs) return;
    asdm_ctx *ctx = cs->ctx;
    (void)hipSetDevice(ctx->device);
    if (cs->store) {
        if (--cs->store->refs == 0) {
            dev_free(ctx, cs->store->d_cand);
            dev_free(ctx, cs->store->d_bits);
            dev_free(ctx, cs->store->d_rank);
            dev_free(ctx, cs->store->d_ce_off);
            dev_free(ctx, cs->store->d_ce_val);
            delete cs->store;
        }
    }
    if (cs->base_store && --cs->base_store->refs == 0) {
        dev_free(ctx, cs->base_store->d_cand);
        dev_free(ctx, cs->base_store->d_bits);
        dev_free(ctx, cs->base_store->d_rank);
        dev_free(ctx, cs->base_store->d_ce_off);
        dev_free(ctx, cs->base_store->d_ce_val);
        delete cs->base_store;
    }
    (void)hipStreamSynchronize(ctx->stream);
    delete cs;
}

}  // extern "C"

namespace asdm {

// CFL filter + candidate edges for nb queries against one uploaded data graph, with every phase
// batched across the queries: one base-filter launch over all query vertices ("slots"), one
// Gen launch per BFS position, one final prune, one compaction, one candidate-edge count/scan/
// fill.  out[i] receives a view into the shared buffers, or nullptr with status[i] =
// ASDM_E_EMPTY_CANDIDATES when some candidate set of query i is empty.
int filter_many(asdm_ctx *ctx, const asdm_graph *g, const asdm_graph *const *qs, uint32_t nb, asdm_cs **out,
                int *status) {
    for (uint32_t i = 0; i < nb; ++i) {
        out[i] = nullptr;
        status[i] = ASDM_E_INVALID;
    }
    if (g->dev.device != ctx->device) return ctx->fail(ASDM_E_INVALID, "data graph not uploaded to this device");
    for (uint32_t i = 0; i < nb; ++i)
        if (!qs[i] || qs[i]->n == 0 || qs[i]->n > kMaxQ)
            return ctx->fail(ASDM_E_UNSUPPORTED, "query must have 1..32 vertices");
    (void)hipSetDevice(ctx->device);
    auto t0 = std::chrono::steady_clock::now();
    const uint32_t n = g->n, L = g->nlab;
    const uint32_t W = (n + 63) / 64;
    std::vector<uint32_t> sb(nb + 1, 0);  // first slot of each query
    for (uint32_t i = 0; i < nb; ++i) sb[i + 1] = sb[i] + qs[i]->n;
    const uint32_t S = sb[nb];

    std::vector<QV> qv(S);
    for (uint32_t i = 0; i < nb; ++i) {
        const asdm_graph *q = qs[i];
        for (uint32_t u = 0; u < q->n; ++u) {
            QV &d = qv[sb[i] + u];
            std::memset(&d, 0, sizeof(d));
            d.lab = q->lab[u];
            d.deg = q->deg(u);
            for (uint32_t k = q->off[u]; k < q->off[u + 1];) {
                uint32_t l = q->lab[q->nbr[k]], j = k;
                while (j < q->off[u + 1] && q->lab[q->nbr[j]] == l) ++j;
                d.nlc_lab[d.nnlc] = l;
                d.nlc_cnt[d.nnlc] = j - k;
                d.nnlc++;
                k = j;
            }
        }
    }
    DevBuf dqv(ctx), dbase(ctx), dgen(ctx), dfin(ctx), dcnt(ctx);
    int rc = dev_alloc(ctx, &dqv.p, sizeof(QV) * S);
    if (!rc) rc = dev_alloc(ctx, &dbase.p, sizeof(uint64_t) * (size_t)W * S);
    if (!rc) rc = dev_alloc(ctx, &dgen.p, sizeof(uint64_t) * (size_t)W * S);
    if (!rc) rc = dev_alloc(ctx, &dfin.p, sizeof(uint64_t) * (size_t)W * S);
    if (!rc) rc = dev_alloc(ctx, &dcnt.p, sizeof(uint32_t) * 3 * S);
    if (rc) return rc;
    uint32_t *cnt_base = dcnt.as<uint32_t>(), *cnt_gen = cnt_base + S, *cnt_fin = cnt_gen + S;
    if ((rc = upload(ctx, dqv.p, qv.data(), sizeof(QV) * S))) return rc;
    if ((rc = fill_u32(ctx, cnt_base, 3 * (uint64_t)S, 0))) return rc;
    if ((rc = fill_u64(ctx, dgen.as<uint64_t>(), (uint64_t)W * S, 0))) return rc;
    if ((rc = fill_u64(ctx, dfin.as<uint64_t>(), (uint64_t)W * S, 0))) return rc;
    {
        KTimer kt(ctx, "filter_base");
        hipLaunchKernelGGL(k_base, dim3((W + 255) / 256, S), dim3(256), 0, ctx->stream, g->dev.labbits, g->dev.off,
                           g->dev.loff, L, dqv.as<QV>(), dbase.as<uint64_t>(), W, cnt_base);
        kt.stop((uint64_t)W * 8 * S);
        if ((rc = check_hip(ctx, hipGetLastError(), "k_base"))) return rc;
    }
    std::vector<uint32_t> base_cnt(S);
    if ((rc = download(ctx, base_cnt.data(), cnt_base, sizeof(uint32_t) * S))) return rc;

    // CFL roots and BFS orders (host); parent lists by BFS position for the Gen launches, and one
    // list per reached slot for the final prune
    std::vector<std::vector<uint32_t>> order(nb);
    std::vector<uint32_t> roots(nb);
    uint32_t maxlen = 0;
    for (uint32_t i = 0; i < nb; ++i) {
        std::vector<uint32_t> bc(base_cnt.begin() + sb[i], base_cnt.begin() + sb[i + 1]);
        roots[i] = cfl_root(qs[i], g, bc);
        order[i] = bfs_order(qs[i], roots[i]);
        maxlen = std::max<uint32_t>(maxlen, (uint32_t)order[i].size());
    }
    std::vector<uint32_t> us, par_off{0}, launch_start;
    std::vector<Par> pars;
    auto add_list = [&](uint32_t i, uint32_t posi, bool before) {
        const asdm_graph *q = qs[i];
        const uint32_t u = order[i][posi];
        std::vector<uint32_t> pos(q->n, UINT32_MAX);
        for (uint32_t k = 0; k < order[i].size(); ++k) pos[order[i][k]] = k;
        us.push_back(sb[i] + u);
        for (uint32_t a = q->off[u]; a < q->off[u + 1]; ++a) {
            uint32_t p = q->nbr[a];
            if (pos[p] == UINT32_MAX) continue;
            if (before ? pos[p] < posi : pos[p] > posi)
                pars.push_back({sb[i] + p, q->lab[p], g->label_count(q->lab[p])});
        }
        par_off.push_back((uint32_t)pars.size());
    };
    for (uint32_t k = 0; k < maxlen; ++k) {
        launch_start.push_back((uint32_t)us.size());
        for (uint32_t i = 0; i < nb; ++i)
            if (k < order[i].size()) add_list(i, k, true);
    }
    const uint32_t final_start = (uint32_t)us.size();
    launch_start.push_back(final_start);
    for (uint32_t i = 0; i < nb; ++i)
        for (uint32_t k = 0; k < order[i].size(); ++k) add_list(i, k, false);
    const uint32_t nfinal = (uint32_t)us.size() - final_start;
    DevBuf dus(ctx), dpo(ctx), dpars(ctx);
    rc = dev_alloc(ctx, &dus.p, sizeof(uint32_t) * us.size());
    if (!rc) rc = dev_alloc(ctx, &dpo.p, sizeof(uint32_t) * par_off.size());
    if (!rc) rc = dev_alloc(ctx, &dpars.p, sizeof(Par) * std::max<size_t>(pars.size(), 1));
    if (!rc) rc = upload(ctx, dus.p, us.data(), sizeof(uint32_t) * us.size());
    if (!rc) rc = upload(ctx, dpo.p, par_off.data(), sizeof(uint32_t) * par_off.size());
    if (!rc && !pars.empty()) rc = upload(ctx, dpars.p, pars.data(), sizeof(Par) * pars.size());
    if (rc) return rc;
    const uint32_t lab0 = n ? g->lab[0] : 0;
    {
        KTimer kt(ctx, "filter_gen");
        for (uint32_t k = 0; k < maxlen; ++k) {
            const uint32_t a = launch_start[k], cnt = launch_start[k + 1] - a;
            hipLaunchKernelGGL(k_prune, dim3((W + 255) / 256, cnt), dim3(256), 0, ctx->stream, g->dev.nbr,
                               g->dev.loff, n, L, lab0, dbase.as<uint64_t>(), dgen.as<uint64_t>(), dgen.as<uint64_t>(),
                               cnt_gen, W, dus.as<uint32_t>() + a, dpo.as<uint32_t>() + a, dpars.as<Par>(), cnt_gen);
        }
        kt.stop();
        if ((rc = check_hip(ctx, hipGetLastError(), "k_prune gen"))) return rc;
        KTimer kf(ctx, "filter_final");
        hipLaunchKernelGGL(k_prune, dim3((W + 255) / 256, nfinal), dim3(256), 0, ctx->stream, g->dev.nbr,
                           g->dev.loff, n, L, lab0, dgen.as<uint64_t>(), dfin.as<uint64_t>(), dgen.as<uint64_t>(),
                           cnt_gen, W, dus.as<uint32_t>() + final_start, dpo.as<uint32_t>() + final_start,
                           dpars.as<Par>(), cnt_fin);
        kf.stop();
        if ((rc = check_hip(ctx, hipGetLastError(), "k_prune final"))) return rc;
    }
    std::vector<uint32_t> fin(S);
    if ((rc = download(ctx, fin.data(), cnt_fin, sizeof(uint32_t) * S))) return rc;

    // non-empty queries keep their slots; the others get no candidate storage
    std::vector<char> ok(nb, 1);
    for (uint32_t i = 0; i < nb; ++i)
        for (uint32_t s2 = sb[i]; s2 < sb[i + 1]; ++s2) ok[i] &= fin[s2] != 0;
    std::vector<uint64_t> cand_off(S + 1, 0);
    std::vector<uint32_t> live;
    for (uint32_t i = 0; i < nb; ++i)
        for (uint32_t s2 = sb[i]; s2 < sb[i + 1]; ++s2) {
            cand_off[s2 + 1] = cand_off[s2] + (ok[i] ? fin[s2] : 0);
            if (ok[i]) live.push_back(s2);
        }
    auto store = new asdm_cs::Store();
    store->refs = 0;
    auto drop_store = [&]() {
        dev_free(ctx, store->d_cand);
        dev_free(ctx, store->d_bits);
        dev_free(ctx, store->d_rank);
        dev_free(ctx, store->d_ce_off);
        dev_free(ctx, store->d_ce_val);
        delete store;
    };
    if (live.empty()) {
        delete store;
        for (uint32_t i = 0; i < nb; ++i) status[i] = ASDM_E_EMPTY_CANDIDATES;
        return ASDM_OK;
    }
    DevBuf dcoff(ctx), dlive(ctx);
    rc = dev_alloc(ctx, (void **)&store->d_cand, sizeof(uint32_t) * std::max<uint64_t>(cand_off[S], 1));
    if (!rc) rc = dev_alloc(ctx, (void **)&store->d_rank, sizeof(uint32_t) * (size_t)W * S);
    if (!rc) rc = dev_alloc(ctx, &dcoff.p, sizeof(uint64_t) * (S + 1));
    if (!rc) rc = dev_alloc(ctx, &dlive.p, sizeof(uint32_t) * live.size());
    if (!rc) rc = upload(ctx, dcoff.p, cand_off.data(), sizeof(uint64_t) * (S + 1));
    if (!rc) rc = upload(ctx, dlive.p, live.data(), sizeof(uint32_t) * live.size());
    if (rc) {
        drop_store();
        return rc;
    }
    {
        KTimer kt(ctx, "filter_compact");
        hipLaunchKernelGGL(k_compact, dim3((uint32_t)live.size()), dim3(1024), 0, ctx->stream, dfin.as<uint64_t>(), W,
                           store->d_rank, dcoff.as<uint64_t>(), store->d_cand, dlive.as<uint32_t>());
        kt.stop();
        if ((rc = check_hip(ctx, hipGetLastError(), "k_compact"))) {
            drop_store();
            return rc;
        }
    }
    store->d_bits = dfin.as<uint64_t>();
    dfin.p = nullptr;  // ownership moves to the store

    // candidate edges over every arc of every non-empty query (slot ids, global item space)
    std::vector<ArcInfo> arcs;
    std::vector<uint64_t> arc_first;   // first item of each arc (global)
    std::vector<uint32_t> qarc0(nb + 1, 0);
    uint64_t nitems = 0;
    for (uint32_t i = 0; i < nb; ++i) {
        qarc0[i] = (uint32_t)arcs.size();
        if (!ok[i]) continue;
        const asdm_graph *q = qs[i];
        for (uint32_t u = 0; u < q->n; ++u)
            for (uint32_t a = q->off[u]; a < q->off[u + 1]; ++a) {
                const uint32_t u2 = q->nbr[a];
                arcs.push_back({sb[i] + u, sb[i] + u2, q->lab[u2], nitems, cand_off[sb[i] + u]});
                arc_first.push_back(nitems);
                nitems += fin[sb[i] + u];
            }
    }
    qarc0[nb] = (uint32_t)arcs.size();
    const uint32_t narcs = (uint32_t)arcs.size();
    arc_first.push_back(nitems);
    DevBuf darcs(ctx), dcounts(ctx), dfirst(ctx), dstart(ctx);
    rc = dev_alloc(ctx, &darcs.p, sizeof(ArcInfo) * std::max<uint32_t>(narcs, 1));
    if (!rc && narcs) rc = upload(ctx, darcs.p, arcs.data(), sizeof(ArcInfo) * narcs);
    if (!rc) rc = dev_alloc(ctx, &dcounts.p, sizeof(uint32_t) * std::max<uint64_t>(nitems, 1));
    if (!rc) rc = dev_alloc(ctx, (void **)&store->d_ce_off, sizeof(uint64_t) * (nitems + 1));
    if (!rc) rc = dev_alloc(ctx, &dfirst.p, sizeof(uint64_t) * (narcs + 1));
    if (!rc) rc = dev_alloc(ctx, &dstart.p, sizeof(uint64_t) * (narcs + 1));
    if (!rc) rc = upload(ctx, dfirst.p, arc_first.data(), sizeof(uint64_t) * (narcs + 1));
    if (rc) {
        drop_store();
        return rc;
    }
    const unsigned ce_grid = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((nitems * 16 + 255) / 256, 1u << 20));
    if (nitems) {
        KTimer kt(ctx, "cand_edges_count");
        hipLaunchKernelGGL((k_cand_edges<16, false>), dim3(ce_grid), dim3(256), 0, ctx->stream, g->dev.nbr, g->dev.loff, L,
                           store->d_cand, darcs.as<ArcInfo>(), narcs, nitems, store->d_bits, store->d_rank, W,
                           dcounts.as<uint32_t>(), nullptr, nullptr);
        kt.stop();
        if ((rc = check_hip(ctx, hipGetLastError(), "k_cand_edges count"))) {
            drop_store();
            return rc;
        }
    }
    if ((rc = scan_u32_to_u64(ctx, dcounts.as<uint32_t>(), store->d_ce_off, nitems))) {
        drop_store();
        return rc;
    }
    hipLaunchKernelGGL(k_gather_u64, dim3((narcs + 1 + 255) / 256), dim3(256), 0, ctx->stream, store->d_ce_off,
                       dfirst.as<uint64_t>(), narcs + 1, dstart.as<uint64_t>());
    std::vector<uint64_t> astart(narcs + 1);
    if ((rc = download(ctx, astart.data(), dstart.p, sizeof(uint64_t) * (narcs + 1)))) {
        drop_store();
        return rc;
    }
    const uint64_t ce_total = astart[narcs];
    if ((rc = dev_alloc(ctx, (void **)&store->d_ce_val, sizeof(uint32_t) * std::max<uint64_t>(ce_total, 1)))) {
        drop_store();
        return rc;
    }
    if (nitems) {
        KTimer kt(ctx, "cand_edges_fill");
        hipLaunchKernelGGL((k_cand_edges<16, true>), dim3(ce_grid), dim3(256), 0, ctx->stream, g->dev.nbr, g->dev.loff, L,
                           store->d_cand, darcs.as<ArcInfo>(), narcs, nitems, store->d_bits, store->d_rank, W, nullptr,
                           store->d_ce_off, store->d_ce_val);
        kt.stop();
        if ((rc = check_hip(ctx, hipGetLastError(), "k_cand_edges fill"))) {
            drop_store();
            return rc;
        }
    }
    if ((rc = check_hip(ctx, hipStreamSynchronize(ctx->stream), "filter sync"))) {
        drop_store();
        return rc;
    }
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    ctx->stats.filter_ms = ms;

    // per-query views into the shared store
    for (uint32_t i = 0; i < nb; ++i) {
        if (!ok[i]) {
            status[i] = ASDM_E_EMPTY_CANDIDATES;
            continue;
        }
        const asdm_graph *q = qs[i];
        auto cs = new asdm_cs();
        cs->ctx = ctx;
        cs->store = store;
        store->refs++;
        cs->nq = q->n;
        cs->nv = n;
        cs->loop_free = g->loop_free;
        cs->W = W;
        cs->root = roots[i];
        cs->qlab = q->lab;
        cs->qoff = q->off;
        cs->qnbr = q->nbr;
        const uint64_t c0 = cand_off[sb[i]];
        cs->csize.assign(fin.begin() + sb[i], fin.begin() + sb[i + 1]);
        cs->cand_off.resize(q->n + 1);
        for (uint32_t u = 0; u <= q->n; ++u) cs->cand_off[u] = cand_off[sb[i] + u] - c0;
        cs->d_cand = store->d_cand + c0;
        cs->d_bits = store->d_bits + (size_t)sb[i] * W;
        cs->d_rank = store->d_rank + (size_t)sb[i] * W;
        cs->narcs = (uint32_t)q->m;
        cs->arc_u.resize(cs->narcs);
        cs->arc_u2.resize(cs->narcs);
        cs->ce_base.assign(cs->narcs + 1, 0);
        cs->arc_weight.resize(cs->narcs);
        const uint32_t a0 = qarc0[i];
        const uint64_t item0 = arc_first[a0];
        for (uint32_t a = 0; a < cs->narcs; ++a) {
            cs->arc_u[a] = arcs[a0 + a].u - sb[i];
            cs->arc_u2[a] = arcs[a0 + a].u2 - sb[i];
            cs->ce_base[a] = arc_first[a0 + a] - item0;
            cs->arc_weight[a] = astart[a0 + a + 1] - astart[a0 + a];
        }
        cs->ce_base[cs->narcs] = arc_first[a0 + cs->narcs] - item0;
        cs->d_ce_off = store->d_ce_off + item0;
        cs->d_ce_val = store->d_ce_val;  // list offsets are global into the shared values
        cs->ce_total = astart[a0 + cs->narcs] - astart[a0];
        cs->filter_ms = ms;
        out[i] = cs;
        status[i] = ASDM_OK;
    }
    if (store->refs == 0) drop_store();
    return ASDM_OK;
}

}  // namespace asdm

extern "C" {

int asdm_filter_cfl(asdm_ctx *ctx, const asdm_graph *q, const asdm_graph *g, asdm_cs **out) try {
    if (!ctx || !q || !g || !out) return ASDM_E_INVALID;
    *out = nullptr;
    int status = ASDM_E_INVALID;
    int rc = filter_many(ctx, g, &q, 1, out, &status);
    if (rc) return rc;
    return status;
} ASDM_CTX_CATCH(ctx)

int asdm_filter_many(asdm_ctx *ctx, const asdm_graph *g, const asdm_graph *const *qs, uint32_t nq, asdm_cs **cs,
                     int *status) try {
    if (!ctx || !g || (nq && (!qs || !cs || !status))) return ASDM_E_INVALID;
    for (uint32_t i = 0; i < nq; ++i)
        if (!qs[i]) return ASDM_E_INVALID;
    return filter_many(ctx, g, qs, nq, cs, status);
} ASDM_CTX_CATCH(ctx)

uint32_t asdm_cs_root(const asdm_cs *cs) { return cs ? cs->root : 0; }

int asdm_cs_info(const asdm_cs *cs, uint32_t *nq, uint32_t *nv) try {
    if (!cs) return ASDM_E_INVALID;
    if (nq) *nq = cs->nq;
    if (nv) *nv = cs->nv;
    return ASDM_OK;
} ASDM_HOST_CATCH

int asdm_cs_sizes(const asdm_cs *cs, uint32_t *sizes) try {
    if (!cs || !sizes) return ASDM_E_INVALID;
    std::memcpy(sizes, cs->csize.data(), sizeof(uint32_t) * cs->nq);
    return ASDM_OK;
} ASDM_HOST_CATCH

int asdm_cs_candidates(const asdm_cs *cs, uint32_t u, uint32_t *out, uint32_t cap, uint32_t *n) try {
    if (!cs || u >= cs->nq || !n) return ASDM_E_INVALID;
    *n = cs->csize[u];
    if (!out) return ASDM_OK;
    if (cap < cs->csize[u]) return ASDM_E_INVALID;
    (void)hipSetDevice(cs->ctx->device);
    return download(cs->ctx, out, cs->d_cand + cs->cand_off[u], sizeof(uint32_t) * cs->csize[u]);
} ASDM_HOST_CATCH

int asdm_cs_arc_weight(const asdm_cs *cs, uint32_t a, uint64_t *w) try {
    if (!cs || !w || a >= cs->narcs) return ASDM_E_INVALID;
    *w = cs->arc_weight[a];
    return ASDM_OK;
} ASDM_HOST_CATCH

}  // extern "C"
