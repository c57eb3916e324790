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
    uint32_t p;        // parent query vertex
    uint32_t lab;      // L(p)
    uint32_t labcount; // |V_L(p)|
};

__global__ void k_base(const uint32_t *__restrict__ glab, const uint32_t *__restrict__ goff,
                       const uint32_t *__restrict__ loff, uint32_t n, uint32_t L, const QV *__restrict__ qv,
                       uint64_t *__restrict__ bits, uint32_t W, uint32_t *__restrict__ cnt) {
    const uint32_t u = blockIdx.y;
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    const QV &q = qv[u];
    bool pass = false;
    if (v < n && glab[v] == q.lab && goff[v + 1] - goff[v] >= q.deg) {
        pass = true;
        const uint32_t *lo = loff + (size_t)v * (L + 1);
        for (uint32_t k = 0; k < q.nnlc; ++k) {
            uint32_t l = q.nlc_lab[k];
            if (l >= L || lo[l + 1] - lo[l] < q.nlc_cnt[k]) {
                pass = false;
                break;
            }
        }
    }
    uint64_t m = __ballot(pass);
    if ((threadIdx.x & 63) == 0 && (v >> 6) < W) {
        bits[(size_t)u * W + (v >> 6)] = m;
        if (m) atomicAdd(&cnt[u], (uint32_t)__popcll(m));
    }
}

// dst(u) = src(u) ∩ every parent test listed for u.  grid.y indexes (u, parent-list) pairs.
__global__ void k_prune(const uint32_t *__restrict__ nbr, const uint32_t *__restrict__ loff, uint32_t n, uint32_t L,
                        uint32_t lab0, const uint64_t *__restrict__ src, uint64_t *__restrict__ dst,
                        const uint64_t *__restrict__ genbits, const uint32_t *__restrict__ gcnt, uint32_t W,
                        const uint32_t *__restrict__ us, const uint32_t *__restrict__ par_off,
                        const Par *__restrict__ pars, uint32_t *__restrict__ dcnt) {
    const uint32_t slot = blockIdx.y;
    const uint32_t u = us[slot];
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    bool pass = false;
    if (v < n) pass = (src[(size_t)u * W + (v >> 6)] >> (v & 63)) & 1ull;
    if (pass) {
        const uint32_t *lo = loff + (size_t)v * (L + 1);
        // v adjacent to data vertex 0: 0 would lead v's label-L(0) slice (slices ascend by id)
        const bool adj0 = lab0 < L && lo[lab0 + 1] > lo[lab0] && nbr[lo[lab0]] == 0;
        for (uint32_t k = par_off[slot]; k < par_off[slot + 1] && pass; ++k) {
            const Par pr = pars[k];
            const bool pad = gcnt[pr.p] < pr.labcount;
            bool hit = pad && adj0;
            if (!hit && pr.lab < L) {
                const uint64_t *gb = genbits + (size_t)pr.p * W;
                for (uint32_t i = lo[pr.lab], e = lo[pr.lab + 1]; i < e; ++i) {
                    uint32_t w = nbr[i];
                    if ((gb[w >> 6] >> (w & 63)) & 1ull) {
                        hit = true;
                        break;
                    }
                }
            }
            pass = hit;
        }
    }
    uint64_t m = __ballot(pass);
    if ((threadIdx.x & 63) == 0 && (v >> 6) < W) {
        dst[(size_t)u * W + (v >> 6)] = m;
        if (m) atomicAdd(&dcnt[u], (uint32_t)__popcll(m));
    }
}

// Per query vertex: rank[w] = set bits before word w; cand[cand_off[u] + rank] = ascending ids.
__global__ void k_compact(const uint64_t *__restrict__ bits, uint32_t W, uint32_t *__restrict__ rank,
                          const uint64_t *__restrict__ cand_off, uint32_t *__restrict__ cand) {
    __shared__ uint64_t lds[1024 / 64 + 1];
    const uint32_t u = blockIdx.x;
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

uint32_t cfl_dummy_anchor();

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
    dev_free(ctx, cs->d_cand);
    dev_free(ctx, cs->d_bits);
    dev_free(ctx, cs->d_rank);
    dev_free(ctx, cs->d_ce_off);
    dev_free(ctx, cs->d_ce_val);
    (void)hipStreamSynchronize(ctx->stream);
    delete cs;
}

int asdm_filter_cfl(asdm_ctx *ctx, const asdm_graph *q, const asdm_graph *g, asdm_cs **out) {
    if (!ctx || !q || !g || !out) return ASDM_E_INVALID;
    *out = nullptr;
    if (g->dev.device != ctx->device) return ctx->fail(ASDM_E_INVALID, "data graph not uploaded to this device");
    if (q->n == 0 || q->n > kMaxQ) return ctx->fail(ASDM_E_UNSUPPORTED, "query must have 1..32 vertices");
    (void)hipSetDevice(ctx->device);
    auto t0 = std::chrono::steady_clock::now();
    const uint32_t nq = q->n, n = g->n, L = g->nlab;
    const uint32_t W = (n + 63) / 64;
    auto cs = std::make_unique<asdm_cs>();
    cs->ctx = ctx;
    cs->nq = nq;
    cs->nv = n;
    cs->W = W;
    cs->qlab = q->lab;
    cs->qoff = q->off;
    cs->qnbr = q->nbr;

    // query vertex descriptors
    std::vector<QV> qv(nq);
    for (uint32_t u = 0; u < nq; ++u) {
        QV &d = qv[u];
        std::memset(&d, 0, sizeof(d));
        d.lab = q->lab[u];
        d.deg = q->deg(u);
        for (uint32_t i = q->off[u]; i < q->off[u + 1];) {
            uint32_t l = q->lab[q->nbr[i]], j = i;
            while (j < q->off[u + 1] && q->lab[q->nbr[j]] == l) ++j;
            d.nlc_lab[d.nnlc] = l;
            d.nlc_cnt[d.nnlc] = j - i;
            d.nnlc++;
            i = j;
        }
    }
    DevBuf dqv(ctx), dbase(ctx), dgen(ctx), dfin(ctx), dcnt(ctx);
    int rc = dev_alloc(ctx, &dqv.p, sizeof(QV) * nq);
    if (!rc) rc = dev_alloc(ctx, &dbase.p, sizeof(uint64_t) * (size_t)W * nq);
    if (!rc) rc = dev_alloc(ctx, &dgen.p, sizeof(uint64_t) * (size_t)W * nq);
    if (!rc) rc = dev_alloc(ctx, &dfin.p, sizeof(uint64_t) * (size_t)W * nq);
    if (!rc) rc = dev_alloc(ctx, &dcnt.p, sizeof(uint32_t) * 3 * nq);  // base, gen, final counts
    if (rc) return rc;
    uint32_t *cnt_base = dcnt.as<uint32_t>(), *cnt_gen = cnt_base + nq, *cnt_fin = cnt_gen + nq;
    if ((rc = upload(ctx, dqv.p, qv.data(), sizeof(QV) * nq))) return rc;
    if ((rc = fill_u32(ctx, cnt_base, 3 * nq, 0))) return rc;
    {
        dim3 grid((W * 64 + 255) / 256, nq);
        KTimer kt(ctx, "filter_base");
        hipLaunchKernelGGL(k_base, grid, dim3(256), 0, ctx->stream, g->dev.lab, g->dev.off, g->dev.loff, n, L,
                           dqv.as<QV>(), dbase.as<uint64_t>(), W, cnt_base);
        kt.stop((uint64_t)n * 4 * 2 * nq);
        if ((rc = check_hip(ctx, hipGetLastError(), "k_base"))) return rc;
    }
    std::vector<uint32_t> base_cnt(nq);
    if ((rc = download(ctx, base_cnt.data(), cnt_base, sizeof(uint32_t) * nq))) return rc;

    // CFL root and BFS order (host, tiny)
    const uint32_t root = cfl_root(q, g, base_cnt);
    cs->root = root;
    std::vector<uint32_t> order = bfs_order(q, root);
    std::vector<uint32_t> pos(nq, UINT32_MAX);
    for (uint32_t i = 0; i < order.size(); ++i) pos[order[i]] = i;

    // parent lists: before (Gen passes, one launch per BFS position) and after (one launch)
    std::vector<uint32_t> us, par_off{0};
    std::vector<Par> pars;
    auto add_list = [&](uint32_t u, bool before) {
        us.push_back(u);
        for (uint32_t a = q->off[u]; a < q->off[u + 1]; ++a) {
            uint32_t p = q->nbr[a];
            if (pos[p] == UINT32_MAX) continue;
            if (before ? pos[p] < pos[u] : pos[p] > pos[u]) pars.push_back({p, q->lab[p], g->label_count(q->lab[p])});
        }
        par_off.push_back((uint32_t)pars.size());
    };
    for (uint32_t i = 0; i < order.size(); ++i) add_list(order[i], true);
    const uint32_t nfinal_start = (uint32_t)us.size();
    for (uint32_t i = 0; i < order.size(); ++i) add_list(order[i], false);
    DevBuf dus(ctx), dpo(ctx), dpars(ctx);
    rc = dev_alloc(ctx, &dus.p, sizeof(uint32_t) * us.size());
    if (!rc) rc = dev_alloc(ctx, &dpo.p, sizeof(uint32_t) * par_off.size());
    if (!rc) rc = dev_alloc(ctx, &dpars.p, sizeof(Par) * std::max<size_t>(pars.size(), 1));
    if (!rc) rc = upload(ctx, dus.p, us.data(), sizeof(uint32_t) * us.size());
    if (!rc) rc = upload(ctx, dpo.p, par_off.data(), sizeof(uint32_t) * par_off.size());
    if (!rc && !pars.empty()) rc = upload(ctx, dpars.p, pars.data(), sizeof(Par) * pars.size());
    // unreached query vertices (disconnected query) keep empty Gen/final bitmaps
    if (!rc) rc = fill_u64(ctx, dgen.as<uint64_t>(), (uint64_t)W * nq, 0);
    if (!rc) rc = fill_u64(ctx, dfin.as<uint64_t>(), (uint64_t)W * nq, 0);
    if (rc) return rc;
    const uint32_t lab0 = n ? g->lab[0] : 0;
    {
        KTimer kt(ctx, "filter_gen");
        for (uint32_t i = 0; i < order.size(); ++i) {
            hipLaunchKernelGGL(k_prune, dim3((W * 64 + 255) / 256, 1), dim3(256), 0, ctx->stream, g->dev.nbr, g->dev.loff, n,
                               L, lab0, dbase.as<uint64_t>(), dgen.as<uint64_t>(), dgen.as<uint64_t>(), cnt_gen, W,
                               dus.as<uint32_t>() + i, dpo.as<uint32_t>() + i, dpars.as<Par>(), cnt_gen);
        }
        kt.stop();
        if ((rc = check_hip(ctx, hipGetLastError(), "k_prune gen"))) return rc;
        KTimer kf(ctx, "filter_final");
        hipLaunchKernelGGL(k_prune, dim3((W * 64 + 255) / 256, (uint32_t)order.size()), dim3(256), 0, ctx->stream,
                           g->dev.nbr, g->dev.loff, n, L, lab0, dgen.as<uint64_t>(), dfin.as<uint64_t>(),
                           dgen.as<uint64_t>(), cnt_gen, W, dus.as<uint32_t>() + nfinal_start,
                           dpo.as<uint32_t>() + nfinal_start, dpars.as<Par>(), cnt_fin);
        kf.stop();
        if ((rc = check_hip(ctx, hipGetLastError(), "k_prune final"))) return rc;
    }
    std::vector<uint32_t> fin(nq);
    if ((rc = download(ctx, fin.data(), cnt_fin, sizeof(uint32_t) * nq))) return rc;
    for (uint32_t u = 0; u < nq; ++u)
        if (fin[u] == 0) {
            cs->filter_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            return ASDM_E_EMPTY_CANDIDATES;
        }
    cs->csize = fin;
    cs->cand_off.assign(nq + 1, 0);
    for (uint32_t u = 0; u < nq; ++u) cs->cand_off[u + 1] = cs->cand_off[u] + fin[u];

    // compaction into sorted candidate lists + per-word ranks
    DevBuf dcoff(ctx);
    rc = dev_alloc(ctx, (void **)&cs->d_cand, sizeof(uint32_t) * cs->cand_off[nq]);
    if (!rc) rc = dev_alloc(ctx, (void **)&cs->d_rank, sizeof(uint32_t) * (size_t)W * nq);
    if (!rc) rc = dev_alloc(ctx, &dcoff.p, sizeof(uint64_t) * (nq + 1));
    if (!rc) rc = upload(ctx, dcoff.p, cs->cand_off.data(), sizeof(uint64_t) * (nq + 1));
    if (rc) return rc;
    {
        KTimer kt(ctx, "filter_compact");
        hipLaunchKernelGGL(k_compact, dim3(nq), dim3(1024), 0, ctx->stream, dfin.as<uint64_t>(), W, cs->d_rank,
                           dcoff.as<uint64_t>(), cs->d_cand);
        kt.stop();
        if ((rc = check_hip(ctx, hipGetLastError(), "k_compact"))) return rc;
    }
    cs->d_bits = dfin.as<uint64_t>();
    dfin.p = nullptr;  // ownership moves to cs

    // candidate edges
    const uint32_t narcs = (uint32_t)q->m;
    cs->narcs = narcs;
    cs->arc_u.resize(narcs);
    cs->arc_u2.resize(narcs);
    std::vector<ArcInfo> arcs(narcs);
    cs->ce_base.assign(narcs + 1, 0);
    for (uint32_t u = 0; u < nq; ++u)
        for (uint32_t a = q->off[u]; a < q->off[u + 1]; ++a) {
            uint32_t u2 = q->nbr[a];
            cs->arc_u[a] = u;
            cs->arc_u2[a] = u2;
            arcs[a] = {u, u2, q->lab[u2], cs->ce_base[a], cs->cand_off[u]};
            cs->ce_base[a + 1] = cs->ce_base[a] + fin[u];
        }
    const uint64_t nitems = cs->ce_base[narcs];
    DevBuf darcs(ctx), dcounts(ctx);
    rc = dev_alloc(ctx, &darcs.p, sizeof(ArcInfo) * std::max<uint32_t>(narcs, 1));
    if (!rc && narcs) rc = upload(ctx, darcs.p, arcs.data(), sizeof(ArcInfo) * narcs);
    if (!rc) rc = dev_alloc(ctx, &dcounts.p, sizeof(uint32_t) * std::max<uint64_t>(nitems, 1));
    if (!rc) rc = dev_alloc(ctx, (void **)&cs->d_ce_off, sizeof(uint64_t) * (nitems + 1));
    if (rc) return rc;
    const unsigned ce_grid = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((nitems * 16 + 255) / 256, 1u << 20));
    if (nitems) {
        KTimer kt(ctx, "cand_edges_count");
        hipLaunchKernelGGL((k_cand_edges<16, false>), dim3(ce_grid), dim3(256), 0, ctx->stream, g->dev.nbr, g->dev.loff, L,
                           cs->d_cand, darcs.as<ArcInfo>(), narcs, nitems, cs->d_bits, cs->d_rank, W,
                           dcounts.as<uint32_t>(), nullptr, nullptr);
        kt.stop();
        if ((rc = check_hip(ctx, hipGetLastError(), "k_cand_edges count"))) return rc;
    }
    if ((rc = scan_u32_to_u64(ctx, dcounts.as<uint32_t>(), cs->d_ce_off, nitems))) return rc;
    std::vector<uint64_t> arc_ends(narcs + 1);
    // per-arc totals: offsets at each arc's first item
    {
        std::vector<uint64_t> hoff(narcs + 1);
        for (uint32_t a = 0; a <= narcs; ++a) {
            if ((rc = check_hip(ctx, hipMemcpyAsync(&hoff[a], cs->d_ce_off + cs->ce_base[a], sizeof(uint64_t),
                                                    hipMemcpyDeviceToHost, ctx->stream), "D2H")))
                return rc;
        }
        if ((rc = check_hip(ctx, hipStreamSynchronize(ctx->stream), "sync"))) return rc;
        cs->arc_weight.resize(narcs);
        for (uint32_t a = 0; a < narcs; ++a) cs->arc_weight[a] = hoff[a + 1] - hoff[a];
        cs->ce_total = hoff[narcs];
    }
    if ((rc = dev_alloc(ctx, (void **)&cs->d_ce_val, sizeof(uint32_t) * std::max<uint64_t>(cs->ce_total, 1)))) return rc;
    if (nitems) {
        KTimer kt(ctx, "cand_edges_fill");
        hipLaunchKernelGGL((k_cand_edges<16, true>), dim3(ce_grid), dim3(256), 0, ctx->stream, g->dev.nbr, g->dev.loff, L,
                           cs->d_cand, darcs.as<ArcInfo>(), narcs, nitems, cs->d_bits, cs->d_rank, W, nullptr,
                           cs->d_ce_off, cs->d_ce_val);
        kt.stop();
        if ((rc = check_hip(ctx, hipGetLastError(), "k_cand_edges fill"))) return rc;
    }
    if ((rc = check_hip(ctx, hipStreamSynchronize(ctx->stream), "filter sync"))) return rc;
    cs->filter_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    ctx->stats.filter_ms = cs->filter_ms;
    *out = cs.release();
    return ASDM_OK;
}

uint32_t asdm_cs_root(const asdm_cs *cs) { return cs ? cs->root : 0; }

int asdm_cs_sizes(const asdm_cs *cs, uint32_t *sizes) {
    if (!cs || !sizes) return ASDM_E_INVALID;
    std::memcpy(sizes, cs->csize.data(), sizeof(uint32_t) * cs->nq);
    return ASDM_OK;
}

int asdm_cs_candidates(const asdm_cs *cs, uint32_t u, uint32_t *out, uint32_t cap, uint32_t *n) {
    if (!cs || u >= cs->nq || !n) return ASDM_E_INVALID;
    *n = cs->csize[u];
    if (!out) return ASDM_OK;
    if (cap < cs->csize[u]) return ASDM_E_INVALID;
    (void)hipSetDevice(cs->ctx->device);
    return download(cs->ctx, out, cs->d_cand + cs->cand_off[u], sizeof(uint32_t) * cs->csize[u]);
}

int asdm_cs_arc_weight(const asdm_cs *cs, uint32_t a, uint64_t *w) {
    if (!cs || !w || a >= cs->narcs) return ASDM_E_INVALID;
    *w = cs->arc_weight[a];
    return ASDM_OK;
}

}  // extern "C"
