/*
 * asdm_oracle.c -- CPU restatement of the reference (ASDMatch) matching path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the HIP engine is compared against.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the product
 * (asdmatch_amd/ and its libasdm.so) never links, loads or calls it.
 *
 * What each function restates (reference file:line):
 *   orc_graph_load / orc_graph_from_edges   relation/graph.cpp:68-148 (t/v/e loader; adjacency
 *                                           sorted by (label, id); vertices grouped by label)
 *   nlc_ok                                  utility/utils.cpp:30-40 (nlcValid)
 *   select_root                             relation/candidate_space.cpp:723-763 (selectCFLRoot,
 *                                           including std::priority_queue tie order)
 *   bfs_tree                                relation/candidate_space.cpp:765-816
 *   orc_cfl                                 relation/candidate_space.cpp:550-721 (buildCandCFL)
 *   build_cand_edges                        relation/candidate_space.cpp:458-487
 *   leapfrog                                utility/compute_set_intersection.cpp:177-233
 *   orc_count                               embedding count of matching/adaptive.cpp:1179-2279 +
 *                                           matching/join.cpp:910-1014 (the count is the number of
 *                                           injective, label- and edge-preserving maps, so it is
 *                                           restated as backtracking over candidate edges with a
 *                                           counted last level; wraps mod 2^64 like size_t)
 *   orc_local_count                         matching/join.cpp:1016-1041 (LOCAL_COUNT) + main.cpp:135-146
 *   orc_unique_combine / orc_bf_unique_combine  utility/utils.cpp:145-201 / :203-233
 *   orc_bruteforce_count                    definition of the count (no filter), for pinning
 *   orc_paths_from_weights / orc_query_paths relation/candidate_space.cpp:507-528 +
 *                                           relation/graph.cpp:341-372 (query arc weights in 32
 *                                           bits, Floyd-Warshall over products in size_t)
 *   orc_reconstruct_path                    relation/candidate_space.cpp:529-537
 *   orc_cartesian                           plan/estimator.cpp:41-102 (optimizedCartesianProduct)
 *
 * Pinning (tests/): distances, successor matrices, paths and cartesian candidate sets against the
 * reference's own CandidateSpace (oracle/_ref/ref_cs_dump --paths, tests/golden/paths.json),
 * including config-2 queries whose products wrap mod 2^64; counts against the reference binary's golden counts (tests/golden), against
 * orc_bruteforce_count and closed forms; candidate sets against the reference's own
 * CandidateSpace output (tests/golden, produced by oracle/_ref/ref_cs_dump); unique-combine
 * against its brute-force twin; local counts against the reference's LOCAL_COUNT build.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    uint32_t n;      /* vertices */
    uint32_t m;      /* arcs (2 x undirected edges) */
    uint32_t nlab;   /* max label + 1 */
    uint32_t *lab;   /* [n] */
    uint32_t *off;   /* [n+1] */
    uint32_t *nbr;   /* [m] sorted by (label, id) per vertex */
    uint32_t *vbl_off; /* [nlab+1] */
    uint32_t *vbl;   /* [n] vertices grouped by label, ascending id */
} orc_graph;

typedef struct {
    uint32_t nq;
    int empty;
    uint32_t root;
    uint32_t *csize;   /* [nq] */
    uint32_t **cand;   /* [nq][csize] ascending data ids */
    /* candidate edges per query arc a (query CSR order): offsets over positions in C(u) */
    uint32_t narcs;
    uint32_t *arc_u, *arc_u2; /* [narcs] */
    uint64_t **ce_off;        /* [narcs][csize[u]+1] */
    uint32_t **ce_val;        /* [narcs][...] data ids */
} orc_cs;

/* ------------------------------------------------------------------ graph */

static int cmp_u32(const void *a, const void *b) {
    uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
    return x < y ? -1 : x > y;
}

static const uint32_t *g_sort_lab; /* qsort context for (label, id) order */
static int cmp_label_id(const void *a, const void *b) {
    uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
    uint32_t lx = g_sort_lab[x], ly = g_sort_lab[y];
    if (lx != ly) return lx < ly ? -1 : 1;
    return x < y ? -1 : x > y;
}

static void finish_graph(orc_graph *g) {
    g_sort_lab = g->lab;
    for (uint32_t v = 0; v < g->n; ++v)
        qsort(g->nbr + g->off[v], g->off[v + 1] - g->off[v], sizeof(uint32_t), cmp_label_id);
    g->vbl_off = (uint32_t *)calloc(g->nlab + 1, sizeof(uint32_t));
    g->vbl = (uint32_t *)malloc(sizeof(uint32_t) * (g->n ? g->n : 1));
    for (uint32_t v = 0; v < g->n; ++v) g->vbl_off[g->lab[v] + 1]++;
    for (uint32_t l = 0; l < g->nlab; ++l) g->vbl_off[l + 1] += g->vbl_off[l];
    uint32_t *pos = (uint32_t *)malloc(sizeof(uint32_t) * (g->nlab + 1));
    memcpy(pos, g->vbl_off, sizeof(uint32_t) * (g->nlab + 1));
    for (uint32_t v = 0; v < g->n; ++v) g->vbl[pos[g->lab[v]]++] = v;
    free(pos);
}

/* Reads the next non-space char (the reference reads the record type with `>> char`). */
static int next_char(FILE *f) {
    int c;
    do { c = fgetc(f); } while (c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '\v' || c == '\f');
    return c;
}

static int next_u32(FILE *f, uint32_t *out) {
    int c = next_char(f);
    if (c == EOF || c < '0' || c > '9') return 0;
    uint64_t x = 0;
    while (c >= '0' && c <= '9') { x = x * 10 + (uint64_t)(c - '0'); c = fgetc(f); }
    if (c != EOF) ungetc(c, f);
    *out = (uint32_t)x;
    return 1;
}

void orc_graph_free(orc_graph *g) {
    if (!g) return;
    free(g->lab); free(g->off); free(g->nbr); free(g->vbl_off); free(g->vbl);
    free(g);
}

/* relation/graph.cpp:68-148.  Returns NULL if the file cannot be opened or if the declared
 * degrees disagree with the edge records (the reference silently corrupts memory there). */
orc_graph *orc_graph_load(const char *path) {
    FILE *f = fopen(path, "r");
    if (!f) return NULL;
    orc_graph *g = (orc_graph *)calloc(1, sizeof(orc_graph));
    uint32_t n = 0, me = 0;
    int t = next_char(f);
    if (t == EOF || !next_u32(f, &n) || !next_u32(f, &me)) { fclose(f); free(g); return NULL; }
    g->n = n;
    g->m = me * 2;
    g->lab = (uint32_t *)calloc(n ? n : 1, sizeof(uint32_t));
    g->off = (uint32_t *)calloc(n + 1, sizeof(uint32_t));
    g->nbr = (uint32_t *)calloc(g->m ? g->m : 1, sizeof(uint32_t));
    uint32_t *fill = (uint32_t *)calloc(n ? n : 1, sizeof(uint32_t));
    uint32_t maxlab = 0;
    int bad = 0;
    for (;;) {
        int c = next_char(f);
        if (c == EOF) break;
        if (c == 'v') {
            uint32_t id, l, d;
            if (!next_u32(f, &id) || !next_u32(f, &l) || !next_u32(f, &d) || id >= n) { bad = 1; break; }
            g->lab[id] = l;
            g->off[id + 1] = g->off[id] + d;
            if (l > maxlab) maxlab = l;
        } else if (c == 'e') {
            uint32_t a, b;
            if (!next_u32(f, &a) || !next_u32(f, &b) || a >= n || b >= n) { bad = 1; break; }
            if (g->off[a] + fill[a] >= g->off[a + 1] || g->off[b] + fill[b] >= g->off[b + 1]) { bad = 1; break; }
            g->nbr[g->off[a] + fill[a]++] = b;
            g->nbr[g->off[b] + fill[b]++] = a;
        }
        /* any other token char is skipped, like the reference's `>> type` loop */
    }
    fclose(f);
    for (uint32_t v = 0; v < n && !bad; ++v)
        if (fill[v] != g->off[v + 1] - g->off[v]) bad = 1;
    if (!bad && g->off[n] != g->m) bad = 1;
    free(fill);
    if (bad) { orc_graph_free(g); return NULL; }
    g->nlab = n ? maxlab + 1 : 0;
    finish_graph(g);
    return g;
}

/* Same graph as orc_graph_load would build from a file with these v/e records, edges in order. */
orc_graph *orc_graph_from_edges(uint32_t n, const uint32_t *labels, uint32_t ne, const uint32_t *edges) {
    orc_graph *g = (orc_graph *)calloc(1, sizeof(orc_graph));
    g->n = n;
    g->m = ne * 2;
    g->lab = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
    g->off = (uint32_t *)calloc(n + 1, sizeof(uint32_t));
    g->nbr = (uint32_t *)malloc(sizeof(uint32_t) * (g->m ? g->m : 1));
    uint32_t maxlab = 0;
    for (uint32_t v = 0; v < n; ++v) { g->lab[v] = labels[v]; if (labels[v] > maxlab) maxlab = labels[v]; }
    for (uint32_t e = 0; e < ne; ++e) { g->off[edges[2 * e] + 1]++; g->off[edges[2 * e + 1] + 1]++; }
    for (uint32_t v = 0; v < n; ++v) g->off[v + 1] += g->off[v];
    uint32_t *fill = (uint32_t *)calloc(n ? n : 1, sizeof(uint32_t));
    for (uint32_t e = 0; e < ne; ++e) {
        uint32_t a = edges[2 * e], b = edges[2 * e + 1];
        g->nbr[g->off[a] + fill[a]++] = b;
        g->nbr[g->off[b] + fill[b]++] = a;
    }
    free(fill);
    g->nlab = n ? maxlab + 1 : 0;
    finish_graph(g);
    return g;
}

void orc_graph_info(const orc_graph *g, uint32_t *n, uint32_t *m, uint32_t *nlab) {
    *n = g->n; *m = g->m; *nlab = g->nlab;
}

void orc_graph_csr(const orc_graph *g, uint32_t *off, uint32_t *nbr, uint32_t *lab) {
    memcpy(off, g->off, sizeof(uint32_t) * (g->n + 1));
    memcpy(nbr, g->nbr, sizeof(uint32_t) * g->m);
    memcpy(lab, g->lab, sizeof(uint32_t) * g->n);
}

static inline uint32_t deg(const orc_graph *g, uint32_t v) { return g->off[v + 1] - g->off[v]; }

/* Label-l slice of v's adjacency: [*b, *b + *cnt).  relation/graph.h:111-114. */
static void label_slice(const orc_graph *g, uint32_t v, uint32_t l, uint32_t *b, uint32_t *cnt) {
    uint32_t lo = g->off[v], hi = g->off[v + 1];
    uint32_t a = lo, z = hi;
    while (a < z) { uint32_t mid = a + (z - a) / 2; if (g->lab[g->nbr[mid]] < l) a = mid + 1; else z = mid; }
    uint32_t s = a;
    z = hi;
    while (a < z) { uint32_t mid = a + (z - a) / 2; if (g->lab[g->nbr[mid]] <= l) a = mid + 1; else z = mid; }
    *b = s;
    *cnt = a - s;
}

/* Edge test via binary search in the label slice (relation/graph.cpp:150-176). */
static int has_edge(const orc_graph *g, uint32_t v, uint32_t w) {
    uint32_t b, c;
    label_slice(g, v, g->lab[w], &b, &c);
    uint32_t a = b, z = b + c;
    while (a < z) { uint32_t mid = a + (z - a) / 2; if (g->nbr[mid] < w) a = mid + 1; else z = mid; }
    return a < b + c && g->nbr[a] == w;
}

/* utility/utils.cpp:30-40: every neighbour-label count of u is met by v. */
static int nlc_ok(const orc_graph *q, uint32_t u, const orc_graph *g, uint32_t v) {
    uint32_t i = q->off[u], e = q->off[u + 1];
    while (i < e) {
        uint32_t l = q->lab[q->nbr[i]], j = i;
        while (j < e && q->lab[q->nbr[j]] == l) ++j;
        uint32_t b, c;
        if (l >= g->nlab) return 0;
        label_slice(g, v, l, &b, &c);
        if (c < j - i) return 0;
        i = j;
    }
    return 1;
}

static uint32_t label_count(const orc_graph *g, uint32_t l) {
    return l < g->nlab ? g->vbl_off[l + 1] - g->vbl_off[l] : 0;
}

/* --------------------------------------------------- CFL root (heap order) */

typedef struct { uint32_t u; double r; } rk_t;

/* std::priority_queue<pair, vector, (a.second < b.second)> is a max-heap on r; push_heap /
 * pop_heap are the textbook sift-up and hole-descent used by libstdc++, restated so that ties
 * between equal ranks leave the heap in the same order as the reference's. */
static void heap_push(rk_t *h, int len /* incl. new element at len-1 */) {
    int hole = len - 1;
    rk_t val = h[hole];
    int parent = (hole - 1) / 2;
    while (hole > 0 && h[parent].r < val.r) {
        h[hole] = h[parent];
        hole = parent;
        parent = (hole - 1) / 2;
    }
    h[hole] = val;
}

static void heap_pop(rk_t *h, int len) {
    if (len <= 1) return;
    int last = len - 1;
    rk_t val = h[last];
    h[last] = h[0];
    /* adjust heap over [0, last) with hole at 0, then push val */
    int n = last, hole = 0, top = 0, child = 0;
    while (child < (n - 1) / 2) {
        child = 2 * (child + 1);
        if (h[child].r < h[child - 1].r) child--;
        h[hole] = h[child];
        hole = child;
    }
    if ((n & 1) == 0 && child == (n - 2) / 2) {
        child = 2 * (child + 1);
        h[hole] = h[child - 1];
        hole = child - 1;
    }
    int parent = (hole - 1) / 2;
    while (hole > top && h[parent].r < val.r) {
        h[hole] = h[parent];
        hole = parent;
        parent = (hole - 1) / 2;
    }
    h[hole] = val;
}

static uint32_t select_root(const orc_graph *q, const orc_graph *g) {
    int nq = (int)q->n;
    rk_t *h = (rk_t *)malloc(sizeof(rk_t) * (nq + 1));
    int len = 0;
    for (int u = 0; u < nq; ++u) {
        h[len].u = (uint32_t)u;
        h[len].r = (double)label_count(g, q->lab[u]) / (double)deg(q, (uint32_t)u);
        ++len;
        heap_push(h, len);
    }
    while (len > 3) { heap_pop(h, len); --len; }
    uint32_t root = 0;
    double best = (double)g->n;
    while (len > 0) {
        uint32_t u = h[0].u;
        uint32_t l = q->lab[u], du = deg(q, u), num = 0;
        if (l < g->nlab)
            for (uint32_t i = g->vbl_off[l]; i < g->vbl_off[l + 1]; ++i) {
                uint32_t v = g->vbl[i];
                if (deg(g, v) < du) continue;
                if (!nlc_ok(q, u, g, v)) continue;
                ++num;
            }
        double score = (double)num / (double)du;
        if (score < best) { root = u; best = score; }
        heap_pop(h, len);
        --len;
    }
    free(h);
    return root;
}

/* ------------------------------------------------------------------- CFL */

/* BFS order from the root; neighbours enqueued in query adjacency order.
 * relation/candidate_space.cpp:765-816.  Only the order is needed below: `before` = earlier
 * neighbours in this order, and the union of sameLevelAfter and lowerLevelAfter = later ones. */
static int bfs_order(const orc_graph *q, uint32_t root, uint32_t *order) {
    uint32_t nq = q->n;
    char *seen = (char *)calloc(nq, 1);
    uint32_t head = 0, tail = 0;
    order[tail++] = root;
    seen[root] = 1;
    while (head < tail) {
        uint32_t u = order[head++];
        for (uint32_t i = q->off[u]; i < q->off[u + 1]; ++i) {
            uint32_t w = q->nbr[i];
            if (!seen[w]) { seen[w] = 1; order[tail++] = w; }
        }
    }
    free(seen);
    return (int)tail;
}

static void build_cand_edges(const orc_graph *q, const orc_graph *g, orc_cs *cs, char **inC);

void orc_cs_free(orc_cs *cs) {
    if (!cs) return;
    for (uint32_t u = 0; u < cs->nq; ++u) free(cs->cand[u]);
    for (uint32_t a = 0; a < cs->narcs; ++a) { free(cs->ce_off[a]); free(cs->ce_val[a]); }
    free(cs->cand); free(cs->csize); free(cs->arc_u); free(cs->arc_u2); free(cs->ce_off); free(cs->ce_val);
    free(cs);
}

/*
 * buildCandCFL (relation/candidate_space.cpp:550-721) in set form.  The reference keeps, per
 * query vertex p, an array of length |V_{L(p)}| whose first entries are the top-down generated
 * candidates Gen(p) and whose tail is zero-filled; the pruning passes walk that whole array.
 * Hence a parent p contributes the label-slice neighbours of
 *     P*(p) = Gen(p)  plus data vertex 0 when |Gen(p)| < |V_{L(p)}|,
 * and pruning only clears the pruned vertex's own bit.  With B(u) = label, degree and NLF test:
 *     Gen(u) = B(u) ∩ ⋂_{p earlier neighbour} N_{L(u)}(P*(p))        (top-down, BFS order)
 *     C(u)   = Gen(u) ∩ ⋂_{p later neighbour}  N_{L(u)}(P*(p))        (same-level + bottom-up)
 * Returns cs with empty=1 when some C(u) is empty (the reference then returns false).
 */
orc_cs *orc_cfl(const orc_graph *q, const orc_graph *g) {
    uint32_t nq = q->n, nv = g->n;
    orc_cs *cs = (orc_cs *)calloc(1, sizeof(orc_cs));
    cs->nq = nq;
    cs->csize = (uint32_t *)calloc(nq, sizeof(uint32_t));
    cs->cand = (uint32_t **)calloc(nq, sizeof(uint32_t *));
    uint32_t root = select_root(q, g);
    cs->root = root;
    uint32_t *order = (uint32_t *)malloc(sizeof(uint32_t) * nq);
    uint32_t *pos = (uint32_t *)malloc(sizeof(uint32_t) * nq);
    int reached = bfs_order(q, root, order);
    for (uint32_t u = 0; u < nq; ++u) pos[u] = UINT32_MAX;
    for (int i = 0; i < reached; ++i) pos[order[i]] = (uint32_t)i;

    char **gen = (char **)calloc(nq, sizeof(char *));   /* Gen(u) bitmaps */
    uint32_t *gsize = (uint32_t *)calloc(nq, sizeof(uint32_t));
    char *mark = (char *)malloc(nv ? nv : 1);
    for (uint32_t u = 0; u < nq; ++u) gen[u] = (char *)calloc(nv ? nv : 1, 1);

    /* N_{l}(P*(p)) into mark[] */
#define MARK_FROM(p, l)                                                          \
    do {                                                                         \
        memset(mark, 0, nv ? nv : 1);                                            \
        int pad = gsize[p] < label_count(g, q->lab[p]);                          \
        for (uint32_t w = 0; w < nv; ++w) {                                      \
            if (!(gen[p][w] || (pad && w == 0))) continue;                       \
            uint32_t b, c;                                                       \
            label_slice(g, w, (l), &b, &c);                                      \
            for (uint32_t k = 0; k < c; ++k) mark[g->nbr[b + k]] = 1;           \
        }                                                                        \
    } while (0)

    for (int i = 0; i < reached; ++i) {
        uint32_t u = order[i], l = q->lab[u];
        if (l < g->nlab)
            for (uint32_t j = g->vbl_off[l]; j < g->vbl_off[l + 1]; ++j) {
                uint32_t v = g->vbl[j];
                if (deg(g, v) >= deg(q, u) && nlc_ok(q, u, g, v)) gen[u][v] = 1;
            }
        for (uint32_t k = q->off[u]; k < q->off[u + 1]; ++k) {
            uint32_t p = q->nbr[k];
            if (pos[p] >= (uint32_t)i) continue; /* only earlier neighbours */
            MARK_FROM(p, l);
            for (uint32_t v = 0; v < nv; ++v) gen[u][v] &= mark[v];
        }
        uint32_t s = 0;
        for (uint32_t v = 0; v < nv; ++v) s += gen[u][v];
        gsize[u] = s;
    }
    /* final sets: prune by later neighbours */
    char **fin = (char **)calloc(nq, sizeof(char *));
    for (uint32_t u = 0; u < nq; ++u) {
        fin[u] = (char *)malloc(nv ? nv : 1);
        memcpy(fin[u], gen[u], nv ? nv : 1);
    }
    for (int i = 0; i < reached; ++i) {
        uint32_t u = order[i], l = q->lab[u];
        for (uint32_t k = q->off[u]; k < q->off[u + 1]; ++k) {
            uint32_t p = q->nbr[k];
            if (pos[p] <= (uint32_t)i) continue; /* only later neighbours */
            MARK_FROM(p, l);
            for (uint32_t v = 0; v < nv; ++v) fin[u][v] &= mark[v];
        }
    }
#undef MARK_FROM
    int empty = 0;
    for (uint32_t u = 0; u < nq; ++u) {
        uint32_t s = 0;
        for (uint32_t v = 0; v < nv; ++v) s += fin[u][v];
        cs->csize[u] = s;
        cs->cand[u] = (uint32_t *)malloc(sizeof(uint32_t) * (s ? s : 1));
        s = 0;
        for (uint32_t v = 0; v < nv; ++v) if (fin[u][v]) cs->cand[u][s++] = v;
        if (cs->csize[u] == 0) empty = 1;
    }
    cs->empty = empty;
    if (!empty) build_cand_edges(q, g, cs, fin);
    for (uint32_t u = 0; u < nq; ++u) { free(gen[u]); free(fin[u]); }
    free(gen); free(fin); free(gsize); free(mark); free(order); free(pos);
    return cs;
}

/* relation/candidate_space.cpp:458-487 with checkEdge=false: for arc (u,u2) and v in C(u),
 * the label-L(u2) slice of N(v) restricted to C(u2), in adjacency (ascending id) order. */
static void build_cand_edges(const orc_graph *q, const orc_graph *g, orc_cs *cs, char **inC) {
    cs->narcs = q->m;
    cs->arc_u = (uint32_t *)malloc(sizeof(uint32_t) * (q->m ? q->m : 1));
    cs->arc_u2 = (uint32_t *)malloc(sizeof(uint32_t) * (q->m ? q->m : 1));
    cs->ce_off = (uint64_t **)calloc(q->m ? q->m : 1, sizeof(uint64_t *));
    cs->ce_val = (uint32_t **)calloc(q->m ? q->m : 1, sizeof(uint32_t *));
    for (uint32_t u = 0; u < q->n; ++u)
        for (uint32_t a = q->off[u]; a < q->off[u + 1]; ++a) {
            uint32_t u2 = q->nbr[a], l2 = q->lab[u2];
            cs->arc_u[a] = u;
            cs->arc_u2[a] = u2;
            uint32_t nc = cs->csize[u];
            uint64_t *off = (uint64_t *)malloc(sizeof(uint64_t) * (nc + 1));
            off[0] = 0;
            for (uint32_t i = 0; i < nc; ++i) {
                uint32_t b, c, k = 0;
                label_slice(g, cs->cand[u][i], l2, &b, &c);
                for (uint32_t j = 0; j < c; ++j) k += inC[u2][g->nbr[b + j]] ? 1u : 0u;
                off[i + 1] = off[i] + k;
            }
            uint32_t *val = (uint32_t *)malloc(sizeof(uint32_t) * (off[nc] ? off[nc] : 1));
            for (uint32_t i = 0; i < nc; ++i) {
                uint32_t b, c;
                uint64_t k = off[i];
                label_slice(g, cs->cand[u][i], l2, &b, &c);
                for (uint32_t j = 0; j < c; ++j)
                    if (inC[u2][g->nbr[b + j]]) val[k++] = g->nbr[b + j];
            }
            cs->ce_off[a] = off;
            cs->ce_val[a] = val;
        }
}

int orc_cs_empty(const orc_cs *cs) { return cs->empty; }
uint32_t orc_cs_root(const orc_cs *cs) { return cs->root; }
uint32_t orc_cs_size(const orc_cs *cs, uint32_t u) { return cs->csize[u]; }
void orc_cs_cand(const orc_cs *cs, uint32_t u, uint32_t *out) {
    memcpy(out, cs->cand[u], sizeof(uint32_t) * cs->csize[u]);
}
/* Sum over v in C(u) of |candidateEdge[u][v][u2]| for query arc index a (query CSR order). */
uint64_t orc_cs_arc_weight(const orc_cs *cs, uint32_t a) {
    if (cs->empty || a >= cs->narcs) return 0;
    return cs->ce_off[a][cs->csize[cs->arc_u[a]]];
}

/* -------------------------------------------------------- set intersection */

/* utility/compute_set_intersection.cpp:177-191: exponential then binary search for the first
 * position >= target in src[begin, end). */
static uint32_t seek(const uint32_t *src, uint32_t begin, uint32_t end, uint32_t target) {
    uint32_t bound = 1, off = begin + bound;
    while (off < end && src[off] < target) { bound *= 2; off = begin + bound; }
    uint32_t lo = begin + bound / 2, hi = end;
    if (off + bound < end) hi = off + bound;
    while (lo < hi) { uint32_t mid = lo + (hi - lo) / 2; if (src[mid] < target) lo = mid + 1; else hi = mid; }
    return lo;
}

/* k-way intersection of ascending unique lists (utility/compute_set_intersection.cpp:193-233,
 * leapfrog over the lists ordered by first element). */
static uint32_t leapfrog(const uint32_t **arr, const uint32_t *cnt, int k, uint32_t *out) {
    if (k == 1) { memcpy(out, arr[0], sizeof(uint32_t) * cnt[0]); return cnt[0]; }
    const uint32_t *a[64];
    uint32_t c[64], it[64];
    for (int i = 0; i < k; ++i) { if (cnt[i] == 0) return 0; a[i] = arr[i]; c[i] = cnt[i]; it[i] = 0; }
    for (int i = 0; i < k; ++i)
        for (int j = i + 1; j < k; ++j)
            if (a[j][0] < a[i][0]) {
                const uint32_t *t = a[i]; a[i] = a[j]; a[j] = t;
                uint32_t tc = c[i]; c[i] = c[j]; c[j] = tc;
            }
    uint32_t n = 0, xp = a[k - 1][0];
    int p = 0;
    for (;;) {
        uint32_t x = a[p][it[p]];
        if (x == xp) { out[n++] = x; ++it[p]; }
        else it[p] = seek(a[p], it[p], c[p], xp);
        if (it[p] == c[p]) break;
        xp = a[p][it[p]];
        p = (p + 1) % k;
    }
    return n;
}

/* ------------------------------------------------------------ enumeration */

typedef struct {
    const orc_graph *q, *g;
    const orc_cs *cs;
    uint32_t nq;
    uint32_t order[64];
    int nb[64];               /* number of backward neighbours of order[d] */
    uint32_t barc[64][64];    /* arc index (p -> order[d]) per backward neighbour */
    uint32_t bdep[64][64];    /* depth of that backward neighbour */
    uint32_t *pos_in_c[64];   /* data id -> position in C(u) (UINT32_MAX if absent), per u */
    uint32_t f[64];           /* data id matched at depth */
    char *used;               /* [|Vg|] */
    uint32_t *buf[64];
    uint64_t *local;          /* optional |Vg| x |Vq| matrix */
    uint64_t count;
} enum_t;

static uint32_t arc_of(const orc_graph *q, uint32_t u, uint32_t u2) {
    for (uint32_t a = q->off[u]; a < q->off[u + 1]; ++a) if (q->nbr[a] == u2) return a;
    return UINT32_MAX;
}

/* Candidates of depth d given f[0..d-1] -> E->buf[d]; returns count. */
static uint32_t extend(enum_t *E, int d) {
    const uint32_t *arr[64];
    uint32_t cnt[64];
    for (int j = 0; j < E->nb[d]; ++j) {
        uint32_t a = E->barc[d][j];
        uint32_t p = E->cs->arc_u[a];
        uint32_t pv = E->f[E->bdep[d][j]];
        uint32_t i = E->pos_in_c[p][pv];
        arr[j] = E->cs->ce_val[a] + E->cs->ce_off[a][i];
        cnt[j] = (uint32_t)(E->cs->ce_off[a][i + 1] - E->cs->ce_off[a][i]);
    }
    return leapfrog(arr, cnt, E->nb[d], E->buf[d]);
}

static void dfs(enum_t *E, int d) {
    uint32_t u = E->order[d];
    uint32_t n;
    const uint32_t *list;
    if (d == 0) { n = E->cs->csize[u]; list = E->cs->cand[u]; }
    else { n = extend(E, d); list = E->buf[d]; }
    if (d == (int)E->nq - 1) {
        uint64_t k = 0;
        for (uint32_t i = 0; i < n; ++i) {
            if (E->used[list[i]]) continue;
            ++k;
            if (E->local) E->local[(uint64_t)list[i] * E->nq + u] += 1;
        }
        E->count += k;
        if (E->local && k)
            for (int j = 0; j < d; ++j) E->local[(uint64_t)E->f[j] * E->nq + E->order[j]] += k;
        return;
    }
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t v = list[i];
        if (E->used[v]) continue;
        E->used[v] = 1;
        E->f[d] = v;
        dfs(E, d + 1);
        E->used[v] = 0;
    }
}

/* Greedy connected order: start at argmin |C(u)|, then repeatedly the unvisited neighbour of the
 * matched set with the most matched neighbours, ties to the smaller |C(u)|, then smaller id. */
static void default_order(const orc_graph *q, const orc_cs *cs, uint32_t *order) {
    uint32_t nq = q->n;
    char in[64] = {0};
    uint32_t best = 0;
    for (uint32_t u = 1; u < nq; ++u) if (cs->csize[u] < cs->csize[best]) best = u;
    order[0] = best;
    in[best] = 1;
    for (uint32_t d = 1; d < nq; ++d) {
        int bu = -1, bconn = -1;
        for (uint32_t u = 0; u < nq; ++u) {
            if (in[u]) continue;
            int conn = 0;
            for (uint32_t a = q->off[u]; a < q->off[u + 1]; ++a) conn += in[q->nbr[a]];
            if (conn == 0) continue;
            if (conn > bconn || (conn == bconn && cs->csize[u] < cs->csize[bu])) { bu = (int)u; bconn = conn; }
        }
        if (bu < 0) /* disconnected query: take any remaining vertex */
            for (uint32_t u = 0; u < nq; ++u) if (!in[u]) { bu = (int)u; break; }
        order[d] = (uint32_t)bu;
        in[bu] = 1;
    }
}

static int run_enum(const orc_graph *q, const orc_graph *g, const orc_cs *cs, const uint32_t *order,
                    uint64_t *local, uint64_t *count) {
    *count = 0;
    if (cs->empty) return 1;
    if (q->n > 64) return -1;
    enum_t *E = (enum_t *)calloc(1, sizeof(enum_t));
    E->q = q; E->g = g; E->cs = cs; E->nq = q->n; E->local = local;
    if (order) memcpy(E->order, order, sizeof(uint32_t) * q->n);
    else default_order(q, cs, E->order);
    uint32_t maxc = 1;
    for (uint32_t u = 0; u < q->n; ++u) if (cs->csize[u] > maxc) maxc = cs->csize[u];
    int ok = 1;
    for (uint32_t d = 0; d < q->n; ++d) {
        uint32_t u = E->order[d];
        E->nb[d] = 0;
        for (uint32_t j = 0; j < d; ++j) {
            uint32_t a = arc_of(q, E->order[j], u);
            if (a != UINT32_MAX) { E->barc[d][E->nb[d]] = a; E->bdep[d][E->nb[d]] = j; E->nb[d]++; }
        }
        if (d > 0 && E->nb[d] == 0) ok = 0; /* disconnected order: not supported by the oracle */
        E->buf[d] = (uint32_t *)malloc(sizeof(uint32_t) * maxc);
        E->pos_in_c[u] = (uint32_t *)malloc(sizeof(uint32_t) * (g->n ? g->n : 1));
        for (uint32_t v = 0; v < g->n; ++v) E->pos_in_c[u][v] = UINT32_MAX;
        for (uint32_t i = 0; i < cs->csize[u]; ++i) E->pos_in_c[u][cs->cand[u][i]] = i;
    }
    E->used = (char *)calloc(g->n ? g->n : 1, 1);
    if (ok) dfs(E, 0);
    *count = E->count;
    for (uint32_t d = 0; d < q->n; ++d) { free(E->buf[d]); free(E->pos_in_c[E->order[d]]); }
    free(E->used);
    free(E);
    return ok ? 0 : -2;
}

/* Number of embeddings (mod 2^64).  order may be NULL (default connected order).
 * Returns 0 on success, 1 if the candidate space is empty (count 0, the reference prints nothing),
 * negative on unsupported input. */
int orc_count(const orc_graph *q, const orc_graph *g, const orc_cs *cs, const uint32_t *order, uint64_t *count) {
    return run_enum(q, g, cs, order, NULL, count);
}

/* LOCAL_COUNT matrix, row-major |Vg| x |Vq| (main.cpp:105-107,135-146; matching/join.cpp:1035-1041):
 * M[v][u] = number of embeddings f with f(u) = v.  M must be zeroed by the caller. */
int orc_local_count(const orc_graph *q, const orc_graph *g, const orc_cs *cs, const uint32_t *order,
                    uint64_t *M, uint64_t *count) {
    return run_enum(q, g, cs, order, M, count);
}

/* --------------------------------------------------------- brute force */

typedef struct {
    const orc_graph *q, *g;
    uint32_t order[64], nb[64], bdep[64][64];
    uint32_t f[64];
    char *used;
    uint64_t count;
} bf_t;

static void bf_dfs(bf_t *B, uint32_t d) {
    const orc_graph *q = B->q, *g = B->g;
    uint32_t u = B->order[d], l = q->lab[u];
    if (d == q->n) { B->count++; return; }
    if (d == 0) {
        for (uint32_t v = 0; v < g->n; ++v) {
            if (g->lab[v] != l) continue;
            B->used[v] = 1; B->f[0] = v; bf_dfs(B, 1); B->used[v] = 0;
        }
        return;
    }
    uint32_t anchor = B->f[B->bdep[d][0]], b, c;
    if (l >= g->nlab) return;
    label_slice(g, anchor, l, &b, &c);
    for (uint32_t k = 0; k < c; ++k) {
        uint32_t v = g->nbr[b + k];
        if (B->used[v]) continue;
        int ok = 1;
        for (uint32_t j = 1; j < B->nb[d] && ok; ++j) ok = has_edge(g, B->f[B->bdep[d][j]], v);
        if (!ok) continue;
        B->used[v] = 1; B->f[d] = v; bf_dfs(B, d + 1); B->used[v] = 0;
    }
}

/* The count by definition: injective maps preserving labels and query edges, no filtering.
 * The query must be connected; the order is BFS from vertex 0. */
int orc_bruteforce_count(const orc_graph *q, const orc_graph *g, uint64_t *count) {
    if (q->n == 0 || q->n > 64) return -1;
    bf_t *B = (bf_t *)calloc(1, sizeof(bf_t));
    B->q = q; B->g = g;
    int reached = bfs_order(q, 0, B->order);
    if (reached != (int)q->n) { free(B); return -2; }
    for (uint32_t d = 0; d < q->n; ++d) {
        B->nb[d] = 0;
        for (uint32_t j = 0; j < d; ++j)
            if (arc_of(q, B->order[j], B->order[d]) != UINT32_MAX) B->bdep[d][B->nb[d]++] = j;
    }
    B->used = (char *)calloc(g->n ? g->n : 1, 1);
    bf_dfs(B, 0);
    *count = B->count;
    free(B->used);
    free(B);
    return 0;
}

/* ------------------------------------------------------ unique combine */

static uint32_t isect_size(const uint32_t *a, uint32_t na, const uint32_t *b, uint32_t nb, uint32_t *out) {
    uint32_t i = 0, j = 0, n = 0;
    while (i < na && j < nb) {
        if (a[i] < b[j]) ++i;
        else if (b[j] < a[i]) ++j;
        else { if (out) out[n] = a[i]; ++n; ++i; ++j; }
    }
    return n;
}

static int uf_find(int *p, int x) { while (p[x] != x) { p[x] = p[p[x]]; x = p[x]; } return x; }

/* utility/utils.cpp:145-201: number of tuples (x_0..x_{k-1}), x_i in S_i, pairwise distinct, as
 * the signed sum over subsets of the k(k-1)/2 equality pairs of the product, over the connected
 * components they induce, of the size of the intersection of the component's sets.  Arithmetic
 * wraps mod 2^64 like the reference's size_t.  sets: concatenated ascending lists, sizes[k].
 * k <= 6 (2^15 pair subsets). */
uint64_t orc_unique_combine(const uint32_t *sets, const uint32_t *sizes, int k) {
    if (k <= 0) return 1;
    if (k > 6) return 0;
    const uint32_t *s[8];
    uint64_t o = 0;
    uint32_t maxs = 1;
    for (int i = 0; i < k; ++i) { s[i] = sets + o; o += sizes[i]; if (sizes[i] > maxs) maxs = sizes[i]; }
    /* intersection size of every subset mask */
    uint64_t isz[64];
    uint32_t *tmp = (uint32_t *)malloc(sizeof(uint32_t) * maxs), *tmp2 = (uint32_t *)malloc(sizeof(uint32_t) * maxs);
    for (int mask = 1; mask < (1 << k); ++mask) {
        int first = __builtin_ctz(mask);
        uint32_t n = sizes[first];
        memcpy(tmp, s[first], sizeof(uint32_t) * n);
        for (int i = first + 1; i < k; ++i)
            if (mask >> i & 1) { n = isect_size(tmp, n, s[i], sizes[i], tmp2); memcpy(tmp, tmp2, sizeof(uint32_t) * n); }
        isz[mask] = n;
    }
    free(tmp); free(tmp2);
    int np = 0, pa[32], pb[32];
    for (int i = 0; i < k; ++i) for (int j = i + 1; j < k; ++j) { pa[np] = i; pb[np] = j; ++np; }
    uint64_t num = 0;
    for (uint32_t sub = 0; sub < (1u << np); ++sub) {
        int par[8];
        for (int i = 0; i < k; ++i) par[i] = i;
        for (int e = 0; e < np; ++e)
            if (sub >> e & 1) { int x = uf_find(par, pa[e]), y = uf_find(par, pb[e]); if (x != y) par[x] = y; }
        uint64_t term = 1;
        for (int r = 0; r < k; ++r) {
            if (uf_find(par, r) != r) continue;
            int mask = 0;
            for (int i = 0; i < k; ++i) if (uf_find(par, i) == r) mask |= 1 << i;
            term *= isz[mask];
        }
        if (__builtin_popcount(sub) & 1) num -= term; else num += term;
    }
    return num;
}

/* utility/utils.cpp:203-233: the same number by enumeration. */
uint64_t orc_bf_unique_combine(const uint32_t *sets, const uint32_t *sizes, int k) {
    const uint32_t *s[16];
    uint64_t o = 0;
    for (int i = 0; i < k; ++i) { s[i] = sets + o; o += sizes[i]; }
    uint32_t idx[16] = {0}, tup[16];
    uint64_t num = 0;
    int d = 0;
    if (k == 0) return 1;
    while (d >= 0) {
        if (idx[d] >= sizes[d]) { idx[d] = 0; --d; if (d >= 0) ++idx[d]; continue; }
        uint32_t v = s[d][idx[d]];
        int clash = 0;
        for (int i = 0; i < d; ++i) if (tup[i] == v) { clash = 1; break; }
        if (clash) { ++idx[d]; continue; }
        tup[d] = v;
        if (d == k - 1) { ++num; ++idx[d]; }
        else { ++d; idx[d] = 0; }
    }
    return num;
}

/* -------------------------------------------------------- query paths, cartesian candidates */

/* relation/candidate_space.cpp:507-528 (setQueryGraphWeights) + relation/graph.cpp:341-372
 * (allSourcesShortestPaths): each arc (u -> u2) of the query, in query CSR order, weighs the total
 * length of its candidate-edge lists, held in a 32-bit `ui` (so it wraps mod 2^32); Floyd-Warshall
 * then minimises the PRODUCT of weights in size_t (wrapping mod 2^64), with dist[i][i] starting at
 * the maximum and next[i][i] = i.  w32[arc] are the arc weights already reduced to 32 bits. */
void orc_paths_from_weights(uint32_t nq, const uint32_t *qoff, const uint32_t *qnbr, const uint32_t *w32,
                            uint64_t *dist, uint32_t *next) {
    const uint64_t INF = UINT64_MAX;
    for (uint32_t i = 0; i < nq * nq; ++i) {
        dist[i] = INF;
        next[i] = UINT32_MAX;
    }
    for (uint32_t i = 0; i < nq; ++i) next[i * nq + i] = i;
    for (uint32_t v = 0; v < nq; ++v)
        for (uint32_t e = qoff[v]; e < qoff[v + 1]; ++e) {
            dist[v * nq + qnbr[e]] = w32[e];
            next[v * nq + qnbr[e]] = qnbr[e];
        }
    for (uint32_t k = 0; k < nq; ++k)
        for (uint32_t i = 0; i < nq; ++i)
            for (uint32_t j = 0; j < nq; ++j) {
                const uint64_t a = dist[i * nq + k], b = dist[k * nq + j];
                if (a == INF || b == INF) continue;
                if (a * b < dist[i * nq + j]) {
                    dist[i * nq + j] = a * b;
                    next[i * nq + j] = next[i * nq + k];
                }
            }
}

void orc_query_paths(const orc_graph *q, const orc_cs *cs, uint64_t *dist, uint32_t *next) {
    uint32_t *w = (uint32_t *)calloc(q->m ? q->m : 1, sizeof(uint32_t));
    /* the oracle's arcs are the query CSR's arcs in the same order */
    for (uint32_t a = 0; a < q->m && a < cs->narcs; ++a) w[a] = (uint32_t)orc_cs_arc_weight(cs, a);
    orc_paths_from_weights(q->n, q->off, q->nbr, w, dist, next);
    free(w);
}

/* relation/candidate_space.cpp:529-537 (reconstructPath): i, next[i][j], ... up to j.  Returns
 * the number of vertices, 0 when next[i][j] is unset, and -1 when following next does not reach
 * j within nq steps (the reference's loop would not end). */
int orc_reconstruct_path(uint32_t nq, const uint32_t *next, uint32_t i, uint32_t j, uint32_t *path) {
    if (next[i * nq + j] == UINT32_MAX) return 0;
    int len = 0;
    path[len++] = i;
    while (i != j) {
        if (len > (int)nq) return -1;
        i = next[i * nq + j];
        if (i == UINT32_MAX) return -1;
        path[len++] = i;
    }
    return len;
}

/* plan/estimator.cpp:41-102 (optimizedCartesianProduct): the candidates of path[len-1] reachable
 * from data vertex v (a candidate of path[0]) along candidate-edge lists of the arcs path[t] ->
 * path[t+1], ascending, without v.  The reference walks every path of candidate edges depth first
 * and merges the last lists; the union of their ends is the set reached layer by layer, which is
 * how it is restated here.  Returns the number written to out (at most |C(path[len-1])|). */
uint32_t orc_cartesian(const orc_graph *q, const orc_cs *cs, uint32_t v, const uint32_t *path, uint32_t len,
                       uint32_t *out) {
    if (cs->empty || len < 2) return 0;
    /* layer: a membership flag over the positions of C(path[t]) */
    uint32_t u0 = path[0];
    char *cur = (char *)calloc(cs->csize[u0] + 1, 1);
    uint32_t p = 0, lo = 0, hi = cs->csize[u0];
    while (lo < hi) { uint32_t mid = (lo + hi) / 2; if (cs->cand[u0][mid] < v) lo = mid + 1; else hi = mid; }
    p = lo;
    if (p < cs->csize[u0] && cs->cand[u0][p] == v) cur[p] = 1;
    for (uint32_t t = 0; t + 1 < len; ++t) {
        const uint32_t a = path[t], b = path[t + 1];
        char *nxt = (char *)calloc(cs->csize[b] + 1, 1);
        uint32_t arc = UINT32_MAX;
        for (uint32_t e = q->off[a]; e < q->off[a + 1]; ++e)
            if (q->nbr[e] == b) arc = e;
        if (arc != UINT32_MAX)
            for (uint32_t x = 0; x < cs->csize[a]; ++x) {
                if (!cur[x]) continue;
                for (uint64_t e = cs->ce_off[arc][x]; e < cs->ce_off[arc][x + 1]; ++e) {
                    const uint32_t id = cs->ce_val[arc][e];
                    uint32_t l2 = 0, h2 = cs->csize[b];
                    while (l2 < h2) { uint32_t mid = (l2 + h2) / 2; if (cs->cand[b][mid] < id) l2 = mid + 1; else h2 = mid; }
                    nxt[l2] = 1;
                }
            }
        free(cur);
        cur = nxt;
    }
    const uint32_t ul = path[len - 1];
    uint32_t n = 0;
    for (uint32_t x = 0; x < cs->csize[ul]; ++x)
        if (cur[x] && cs->cand[ul][x] != v) out[n++] = cs->cand[ul][x];
    free(cur);
    return n;
}
