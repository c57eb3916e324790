/*
 * asdm.h -- C ABI of the MI355X (gfx950) ASDMatch matching engine (libasdm.so).
 *
 * Plain C types only: opaque handles, pointers and sizes.  Every entry point returns an int
 * status (ASDM_OK = 0, negative on error); the message of the last error on a context is
 * available from asdm_last_error().  A context owns one HIP stream and is used by one host thread
 * at a time; distinct contexts may be used concurrently.
 *
 * Reference interfaces replaced (file:line in the reference repository):
 *   asdm_graph_load_text        Graph::loadGraphFromTextFile        relation/graph.cpp:68-148
 *   asdm_graph_from_edges       same layout, built from in-memory v/e records
 *   asdm_graph_write_text       Graph::writeToTextFile              relation/graph.cpp:451-471
 *   asdm_filter_cfl             CandidateSpace::buildCandCFL        relation/candidate_space.cpp:550-721
 *                               (+ buildCandidateEdge(false)        relation/candidate_space.cpp:458-487)
 *   asdm_cs_candidates          CandidateSpace::candidateSet        relation/candidate_space.h:62
 *   asdm_cs_arc_weight          setQueryGraphWeights' edge weight   relation/candidate_space.cpp:507-524
 *   asdm_cs_query_paths         allSourcesShortestPaths             relation/graph.cpp:341-372
 *   asdm_cs_query_path          CandidateSpace::reconstructPath     relation/candidate_space.cpp:529-537
 *   asdm_cartesian_candidates   optimizedCartesianProduct           plan/estimator.cpp:41-102
 *   asdm_plan_create            simplePlan / optCostPlan            plan/optimizer.cpp:7-200, :380-548
 *   asdm_plan_create_with_cards subsetToCard consumed by the order  plan/estimator.cpp:7
 *   asdm_plan_create_ordered    plan from a given attribute order   plan/optimizer.cpp:7-200
 *   asdm_count                  adaptiveShareJoin (count build)     matching/adaptive.cpp:1179-2279
 *   asdm_local_count            adaptiveShareJoin + produceResult   matching/join.cpp:1016-1041,
 *                               (ALL_LEVEL + LOCAL_COUNT build)     main.cpp:105-107,135-146
 *   asdm_intersect_k            ComputeSetIntersection::LeapfrogJoin utility/compute_set_intersection.cpp:193-233
 *   asdm_batch_count            (no counterpart: many queries per launch; see asdm_batch_count)
 *   asdm_filter_many            buildCandCFL for many queries       relation/candidate_space.cpp:550-721
 *   asdm_mpnn_*                 cardinality estimates for subsetToCard, in place of the sampling
 *                               estimator cardEstimateLabeled       plan/estimator.cpp:7, :173-291
 *   asdm_plan_create_learned    optimizer ordering by subsetToCard  plan/optimizer.cpp:7-200
 *   asdm_batch_count_learned    (no counterpart: asdm_batch_count with learned plans)
 *   asdm_batch_autotune,        (no counterpart: launch shape of the batched engine)
 *   asdm_ctx_{set,get}_batch_tuning
 *   asdm_ctx_{set,get}_bag_workers (no counterpart: concurrent bag-engine queries of a batch)
 *   asdm_ctx_set_table_form     (no counterpart: form of the bag engine's relation products)
 *   asdm_dense_product          (no counterpart: the bag engine's dense x dense product, for tests)
 *   asdm_plain_estimate         (no counterpart: this planner's own independence-model estimate,
 *                               the learned estimator's fallback)
 */
#ifndef ASDM_H
#define ASDM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---- */
#define ASDM_OK 0
#define ASDM_E_EMPTY_CANDIDATES 1 /* some C(u) is empty: count is 0; the CLI prints nothing */
#define ASDM_E_INVALID (-1)       /* bad argument / shape */
#define ASDM_E_IO (-2)            /* file cannot be opened or is malformed */
#define ASDM_E_OOM (-3)           /* device or host allocation failed */
#define ASDM_E_HIP (-4)           /* HIP runtime error (message has the HIP error string) */
#define ASDM_E_NODEVICE (-5)      /* no usable gfx950 device */
#define ASDM_E_BUDGET (-6)        /* frontier budget too small for one expansion */
#define ASDM_E_UNSUPPORTED (-7)   /* query outside the supported envelope (e.g. > 32 vertices) */
#define ASDM_E_TIMEOUT (-8)       /* the per-query time limit elapsed (count unknown) */
/* No C++ exception leaves the library: an int-returning entry point whose host allocation fails
 * returns ASDM_E_OOM (with a context, also its last error). */

typedef struct asdm_ctx asdm_ctx;     /* device, stream, scratch arena */
typedef struct asdm_graph asdm_graph; /* host graph (label-sorted CSR), optionally on device */
typedef struct asdm_cs asdm_cs;       /* device candidate space of (query, data) */
typedef struct asdm_plan asdm_plan;   /* host matching plan consumed by the engine */
typedef struct asdm_mpnn asdm_mpnn;   /* learned cardinality estimator (device weights) */

/* ---- context ---- */
int asdm_ctx_create(int device, asdm_ctx **out);
void asdm_ctx_destroy(asdm_ctx *ctx);
const char *asdm_last_error(const asdm_ctx *ctx);
/* Per-query wall-clock limit for asdm_count / asdm_local_count (0 = none), checked between
 * kernel launches; an exceeded limit returns ASDM_E_TIMEOUT.  The paper's experiments use a
 * one-hour limit per query in the same way. */
int asdm_ctx_set_time_limit(asdm_ctx *ctx, double ms);
/* Limit on partial embeddings materialized per count (0 = none), checked before each launch;
 * exceeding it returns ASDM_E_TIMEOUT.  Independent of device speed. */
int asdm_ctx_set_row_limit(asdm_ctx *ctx, uint64_t rows);
/* Library version string, no device needed. */
const char *asdm_version(void);

/* ---- graphs (host side; no device needed) ---- */
/* Reproduces the reference loader: header 't N M' (M undirected), 'v id label deg' records
 * fixing offsets from declared degrees, 'e a b [extra tokens]' inserted in both directions,
 * adjacency sorted by (label(nbr), nbr).  Declared degrees must equal the edge records.
 * Replaces Graph::loadGraphFromTextFile (relation/graph.cpp:68-148).  Files of 32 MB and more are
 * parsed by host threads, with the same result.  A file written by asdm_graph_save_binary (it
 * starts with "ASDMCSR1") is read as that binary CSR instead.  Labels above 2^24 - 1 return
 * ASDM_E_UNSUPPORTED (per-label tables hold max label + 1 entries); a header whose edge count the
 * file cannot hold, or degrees that do not add up to it, return ASDM_E_IO before any adjacency
 * entry is placed out of range. */
int asdm_graph_load_text(const char *path, asdm_graph **out);
/* Binary CSR cache of a loaded graph (labels, offsets, sorted adjacency); asdm_graph_load_text
 * reads it back without parsing or sorting.  The reference has a binary graph format
 * (relation/graph.cpp:293-332) that its main never uses; this one is for callers that load the
 * same large graph many times. */
int asdm_graph_save_binary(const asdm_graph *g, const char *path);
/* Graph from n labels and ne undirected edges (edges[2*i], edges[2*i+1]), in record order.
 * Endpoints >= n return ASDM_E_INVALID, labels above 2^24 - 1 ASDM_E_UNSUPPORTED. */
int asdm_graph_from_edges(uint32_t n, const uint32_t *labels, uint64_t ne, const uint32_t *edges,
                          asdm_graph **out);
int asdm_graph_write_text(const asdm_graph *g, const char *path);
void asdm_graph_free(asdm_graph *g);
/* n vertices, m arcs (= 2 x undirected edges), nlabels = max label + 1. */
int asdm_graph_info(const asdm_graph *g, uint32_t *n, uint64_t *m, uint32_t *nlabels);
/* Copies the label-sorted CSR out: off[n+1], nbr[m], lab[n] (any may be NULL). */
int asdm_graph_csr(const asdm_graph *g, uint32_t *off, uint32_t *nbr, uint32_t *lab);
/* Undirected edge list (a < b per edge, ascending (a, b)): edges[2*ne]. */
int asdm_graph_edges(const asdm_graph *g, uint32_t *edges, uint64_t cap, uint64_t *ne);
/* Puts CSR, labels, per-vertex label offsets and the vertices-by-label index into HBM. */
int asdm_graph_upload(asdm_ctx *ctx, asdm_graph *g);

/* ---- seeded synthetic inputs (portable xoshiro256** streams; no device needed) ---- */
/* Chung-Lu power law: weights w_i ~ (i+1)^(-1/(beta-1)), exactly ne distinct undirected edges,
 * no self-loops, vertex ids randomly permuted; labels iid Zipf(zipf_s) over nlabels (s = 0 gives
 * uniform labels). */
int asdm_gen_chung_lu(uint32_t n, uint64_t ne, double beta, uint32_t nlabels, double zipf_s,
                      uint64_t seed, asdm_graph **out);
/* R-MAT (a, b, c, d = 1-a-b-c), 2^scale vertices, edge_factor * 2^scale draws, symmetrized,
 * deduplicated, self-loops dropped, ids permuted; labels as above. */
int asdm_gen_rmat(uint32_t scale, uint32_t edge_factor, double a, double b, double c, uint32_t nlabels,
                  double zipf_s, uint64_t seed, asdm_graph **out);
/* Query of k distinct vertices taken by a random walk on g (labels copied).  Kept edges: the
 * walk's discovery tree plus a fraction `density` in [0,1] of the remaining induced edges
 * (density 1 = all induced edges), but at most max_extra extra edges (UINT32_MAX = no cap).
 * Returns ASDM_E_INVALID if no walk reaches k vertices. */
int asdm_gen_query(const asdm_graph *g, uint32_t k, double density, uint32_t max_extra, uint64_t seed,
                   asdm_graph **out);

/* ---- filter (CFL) and candidate space ---- */
/* Builds C(u) for every query vertex with the CFL rules of the reference (same root, same
 * candidate sets) and the candidate-edge index, all in HBM.  Returns ASDM_E_EMPTY_CANDIDATES
 * (and *out = NULL) when some C(u) is empty.  g must have been uploaded. */
int asdm_filter_cfl(asdm_ctx *ctx, const asdm_graph *q, const asdm_graph *g, asdm_cs **out);
/* Must be called before the context that built cs is destroyed. */
void asdm_cs_free(asdm_cs *cs);
uint32_t asdm_cs_root(const asdm_cs *cs);
/* Query vertices and data-graph vertices of the candidate space: a local-count matrix for cs
 * holds nv rows of nq values. */
int asdm_cs_info(const asdm_cs *cs, uint32_t *nq, uint32_t *nv);
/* |C(u)| for every u (sizes[nq]). */
int asdm_cs_sizes(const asdm_cs *cs, uint32_t *sizes);
/* Copies C(u) (ascending data ids) into out[cap]; *n = |C(u)|. */
int asdm_cs_candidates(const asdm_cs *cs, uint32_t u, uint32_t *out, uint32_t cap, uint32_t *n);
/* Sum over v in C(u) of the candidate-edge list length for query arc a (query CSR order). */
int asdm_cs_arc_weight(const asdm_cs *cs, uint32_t a, uint64_t *w);
/* The query's all-pairs paths as setQueryGraphWeights leaves them (relation/candidate_space.cpp:507-528,
 * relation/graph.cpp:341-372): each arc weighs its candidate-edge total reduced to 32 bits (the
 * reference's ui), and Floyd-Warshall minimises the product of weights in 64 bits, wrapping, with
 * dist[i][i] starting at UINT64_MAX.  dist[nq * nq] and next[nq * nq] (either may be NULL) receive the
 * distances and the successor matrix (UINT32_MAX: no path), row-major by (i, j). */
int asdm_cs_query_paths(const asdm_cs *cs, uint64_t *dist, uint32_t *next);
/* CandidateSpace::reconstructPath (relation/candidate_space.cpp:529-537): path[0] = i, ..., path[*len - 1]
 * = j; *len = 0 when there is no path.  When wrapped products leave a cycle in the successor matrix the
 * reference's loop never ends; here the call returns ASDM_E_UNSUPPORTED.  ASDM_E_BUDGET if cap is too small. */
int asdm_cs_query_path(const asdm_cs *cs, uint32_t i, uint32_t j, uint32_t *path, uint32_t cap, uint32_t *len);
/* optimizedCartesianProduct (plan/estimator.cpp:41-102), the candidates of a cartesian step: the candidates
 * of path[len-1] reached from data vertex v (the image of path[0]) through the candidate-edge lists of the
 * query arcs path[0] -> path[1] -> ... -> path[len-1], as ascending data ids, without v.  A v outside
 * C(path[0]) or a consecutive pair that is not a query arc reaches nothing.  *n receives the number of
 * candidates; ASDM_E_BUDGET when it exceeds cap.  Runs on the device (a bitmap per path layer). */
int asdm_cartesian_candidates(asdm_ctx *ctx, const asdm_cs *cs, uint32_t v, const uint32_t *path, uint32_t len,
                              uint32_t *out, uint32_t cap, uint32_t *n);

/* ---- plan ---- */
/* order: NULL to let the planner choose, else a connected permutation of the query vertices.
 * flags: ASDM_PLAN_LOCAL restricts leaf factorization to what local counting supports. */
#define ASDM_PLAN_LOCAL 1u
/* Count only the last vertex in closed form (no multi-leaf factorization): a plain enumeration
 * used to cross-check the leaf factorization. */
#define ASDM_PLAN_NO_FACTOR 2u
/* Cost-based order search (independence-model cardinalities): the order of the enumerated vertices
 * with the smallest summed prefix estimates -- exact over all connected orders (subset DP,
 * plan/optimizer.cpp:380-548) for at most 12 enumerated vertices, else the best greedy order grown
 * from each of them. */
#define ASDM_PLAN_COST 4u
/* Count through a tree decomposition of the query (bag tables of at most two query vertices,
 * joined by variable elimination; injectivity by Moebius inversion over the partitions of each
 * label class into non-adjacent blocks) instead of enumerating partial embeddings: the
 * reference's bag materialization and global join (matching/adaptive.cpp:10-330,
 * matching/join.cpp:56-103,910-1014,1109-1124) with the bags' tuples replaced by per-bag count
 * tables.  Fails with ASDM_E_UNSUPPORTED when some quotient of the query has treewidth above two
 * or a table would be too large.  Without this flag or ASDM_PLAN_FRONTIER, asdm_plan_create picks
 * this engine by itself when the frontier plan's estimated rows exceed 2^24 or it enumerates eight
 * or more vertices, and the bag plan costs at most 1e14 table-element operations (about ten
 * seconds on one MI355X). */
#define ASDM_PLAN_BAGS 8u
/* Never use the bag-table engine (enumerate partial embeddings). */
#define ASDM_PLAN_FRONTIER 16u
/* Which engine a plan uses: 1 = bag tables, 0 = frontier enumeration. */
int asdm_plan_uses_bags(const asdm_plan *p);
/* Size of a bag-table plan: partition terms, estimated table-element operations and entries of
 * its largest table (all 0 for a frontier plan). */
int asdm_plan_bag_info(const asdm_plan *p, uint32_t *terms, double *cost, double *peak_entries);
int asdm_plan_create(const asdm_graph *q, const asdm_cs *cs, const uint32_t *order, uint32_t flags,
                     asdm_plan **out);
/* Plan with an explicit order whose first num_enumerated vertices are enumerated and the rest are
 * leaves (pairwise non-adjacent, within the label-group limit); the analogue of building a plan
 * from a given attribute order (plan/optimizer.cpp:7-200). */
int asdm_plan_create_ordered(const asdm_graph *q, const asdm_cs *cs, const uint32_t *order,
                             uint32_t num_enumerated, uint32_t flags, asdm_plan **out);
void asdm_plan_free(asdm_plan *p);
/* Planner's estimate (independence model over candidate-edge list sizes, or the supplied
 * cardinalities) of the partial embeddings materialized at the leaf depth and over all depths. */
int asdm_plan_estimate(const asdm_plan *p, double *rows_leaf, double *rows_total);
/* order[nq] and the number of enumerated (non-leaf) depths. */
int asdm_plan_order(const asdm_plan *p, uint32_t *order, uint32_t *num_enumerated);
/* Planner cardinalities keyed by query-vertex subset bitmask (bit u = query vertex u), the
 * subsetToCard analogue.  Used to pick the order when asdm_plan_create is given order == NULL
 * after this call on the same (q, cs) pair via asdm_plan_create_with_cards. */
int asdm_plan_create_with_cards(const asdm_graph *q, const asdm_cs *cs, const uint64_t *masks,
                                const double *cards, uint32_t ncards, uint32_t flags, asdm_plan **out);

/* ---- execution ---- */
/* Exact embedding count mod 2^64.  budget_bytes bounds the frontier memory per level (0 = a
 * default sized for the device); any budget gives the same count. */
int asdm_count(asdm_ctx *ctx, const asdm_plan *plan, const asdm_cs *cs, uint64_t budget_bytes,
               uint64_t *count);
/* LOCAL_COUNT: mat[v * nq + u] = number of embeddings f with f(u) = v, row-major |Vg| x |Vq|,
 * caller-owned host memory; *count = total.  The plan must be created with ASDM_PLAN_LOCAL.
 * mat must hold nv * nq values, nv and nq as asdm_cs_info reports them for cs (the vertex count of
 * the data graph cs was filtered against): the whole matrix is written. */
int asdm_local_count(asdm_ctx *ctx, const asdm_plan *plan, const asdm_cs *cs, uint64_t budget_bytes,
                     uint64_t *mat, uint64_t *count);
/* COLLECT_RESULT: every embedding, out[i * nq + u] = data vertex of query vertex u in the i-th one,
 * in no particular order; *count = their number.  Replaces the result vector the reference fills
 * when built with COLLECT_RESULT (matching/join.cpp:1032-1034,1076-1078; matching/adaptive.cpp:131).
 * The plan must be created with ASDM_PLAN_NO_FACTOR.  out == NULL counts only; more embeddings than
 * cap_rows returns ASDM_E_BUDGET. */
int asdm_enumerate(asdm_ctx *ctx, const asdm_plan *plan, const asdm_cs *cs, uint64_t budget_bytes,
                   uint32_t *out, uint64_t cap_rows, uint64_t *count);

/* Per-kernel device timing (HIP events on the context stream), collected while profiling is on.
 * Turning profiling on or off clears the table.  No reference counterpart (measurement only). */
int asdm_ctx_set_profiling(asdm_ctx *ctx, int on);
/* Entry i of the kernel table: name, summed device ms, launches, algorithmic bytes. */
int asdm_kernel_stat(const asdm_ctx *ctx, uint32_t i, char *name, uint32_t name_cap, double *ms,
                     uint64_t *launches, uint64_t *bytes);

/* Statistics of the last asdm_count / asdm_local_count on this context. */
typedef struct {
    double filter_ms, plan_ms, exec_ms;   /* host wall clock of each phase */
    double kernel_ms;                     /* sum of device time of the enumeration kernels */
    uint64_t alg_bytes;                   /* algorithmic bytes: list reads + frontier traffic */
    uint64_t rows_total;                  /* partial embeddings materialized */
    uint64_t launches;                    /* kernel launches in the enumeration */
} asdm_stats;
int asdm_last_stats(const asdm_ctx *ctx, asdm_stats *out);

/* ---- batched multi-query engine ---- */
/* Counts nq queries against one uploaded data graph.  counts[nq] receive the exact counts;
 * lat_ms[nq] (optional) the completion time of each query measured from the call start. */
int asdm_batch_count(asdm_ctx *ctx, const asdm_graph *g, const asdm_graph *const *qs, uint32_t nq,
                     uint64_t budget_bytes, uint64_t *counts, float *lat_ms);

/* Batched filter: candidate spaces of nq queries against one uploaded data graph in one pass.
 * cs[i] is NULL with status[i] = ASDM_E_EMPTY_CANDIDATES when a candidate set of query i is
 * empty; the views share device storage that the last asdm_cs_free releases. */
int asdm_filter_many(asdm_ctx *ctx, const asdm_graph *g, const asdm_graph *const *qs, uint32_t nq, asdm_cs **cs,
                     int *status);
/* asdm_batch_count with plans ordered by the learned estimator m (the plain planner when m is
 * NULL).  Counts are the same exact values whatever the plans.  With status[nq] (optional) a query
 * that fails on its own -- the context's row limit (ASDM_E_TIMEOUT) or an unsupported shape --
 * gets its code there and count 0 while the rest are counted; status[i] = ASDM_E_EMPTY_CANDIDATES
 * for a query with an empty candidate set (count 0).  Without it such a query fails the call. */
int asdm_batch_count_learned(asdm_ctx *ctx, const asdm_graph *g, const asdm_graph *const *qs, uint32_t nq,
                             uint64_t budget_bytes, asdm_mpnn *m, uint64_t *counts, int *status, float *lat_ms);

/* Launch shape of the batched engine: lanes per leaf-depth row (1, 4, 16 or 64), a factor (> 0) on
 * the expected-list-length thresholds that pick 4, 16 or 64 lanes per row for the other rows, and
 * rows per lane group in a wave tile (1..64).  Defaults 16, 1.0, 4.  Counts do not depend on it. */
int asdm_ctx_set_batch_tuning(asdm_ctx *ctx, uint32_t leaf_lanes, double class_scale, uint32_t tile_iters);
int asdm_ctx_get_batch_tuning(const asdm_ctx *ctx, uint32_t *leaf_lanes, double *class_scale,
                              uint32_t *tile_iters);
/* Worker contexts that count the bag-engine queries of a batched call concurrently, each with its
 * own stream (1..16; default 6: more streams than the device's 4 hardware queues per process still
 * hide the workers' host synchronizations; 1 counts them one after the other on ctx).  Counts do not
 * depend on it. */
int asdm_ctx_set_bag_workers(asdm_ctx *ctx, uint32_t workers);
int asdm_ctx_get_bag_workers(const asdm_ctx *ctx, uint32_t *workers);
/* Form of the bag engine's products of two edge relations into a dense table: 0 (default) picks per
 * product from its term count (few terms per table cell: zero-fill + atomic scatter; many: row
 * accumulation in LDS), 1 always the LDS form, 2 always the scatter form.  For tests and tuning;
 * counts do not depend on it.  ASDM_E_INVALID outside 0..2. */
int asdm_ctx_set_table_form(asdm_ctx *ctx, int form);
/* Autotuner: runs the batch once untimed, then `reps` timed runs (at least one) for each shape of the grid leaf
 * lanes {4, 16, 64} x class factor {0.25, 0.5, 1, 2} x tile rows {2, 4, 8}; keeps the shape with the
 * smallest median time on the context and returns it (ms_best: that median).  Fails with
 * ASDM_E_INVALID if any run's counts differ from the first run's. */
int asdm_batch_autotune(asdm_ctx *ctx, const asdm_graph *g, const asdm_graph *const *qs, uint32_t nq,
                        uint64_t budget_bytes, asdm_mpnn *m, uint32_t reps, uint32_t *leaf_lanes,
                        double *class_scale, uint32_t *tile_iters, float *ms_best);

/* ---- learned cardinality estimator ----
 * A 3-layer message-passing network (hidden 64, node features ASDM_MPNN_FEATURES, graphs of at
 * most 32 nodes) predicting y = log2(1 + #embeddings) of a (sub-)query.  Weights are bf16 with
 * fp32 masters; all products accumulate in fp32.  A batch of B graphs is given as node_off[B+1]
 * (node ranges), x[N * ASDM_MPNN_FEATURES] (row-major node features), adj_off[N+1] and adj (global
 * node ids of each node's neighbours; the adjacency must be symmetric). */
#define ASDM_MPNN_FEATURES 8
#define ASDM_MPNN_HIDDEN 64
int asdm_mpnn_create(asdm_ctx *ctx, uint64_t seed, asdm_mpnn **out);
void asdm_mpnn_free(asdm_mpnn *m); /* before asdm_ctx_destroy of its context */
uint32_t asdm_mpnn_num_params(void);
int asdm_mpnn_get_params(asdm_mpnn *m, float *params);       /* fp32 masters */
int asdm_mpnn_set_params(asdm_mpnn *m, const float *params); /* also refreshes the bf16 copy */
int asdm_mpnn_save(asdm_mpnn *m, const char *path);
int asdm_mpnn_load(asdm_mpnn *m, const char *path);
int asdm_mpnn_forward(asdm_mpnn *m, uint32_t B, const uint32_t *node_off, const float *x, const uint32_t *adj_off,
                      const uint32_t *adj, float *y);
/* Gradient (fp32, w.r.t. the bf16 weights) of 0.5 * mean_b (y_b - t_b)^2, without an update. */
int asdm_mpnn_grad(asdm_mpnn *m, uint32_t B, const uint32_t *node_off, const float *x, const uint32_t *adj_off,
                   const uint32_t *adj, const float *targets, float *grad, float *y);
/* One Adam step (beta 0.9 / 0.999, eps 1e-8) on that loss; *loss is the loss before the step. */
int asdm_mpnn_train_step(asdm_mpnn *m, uint32_t B, const uint32_t *node_off, const float *x,
                         const uint32_t *adj_off, const uint32_t *adj, const float *targets, float lr, float *loss);
/* Features of the sub-queries induced by masks[n] on a candidate space (data graph g gives the
 * label statistics).  With node_off/x/adj_off/adj NULL only *num_nodes / *num_adj are returned;
 * otherwise they are capacities on input and sizes on output. */
int asdm_mpnn_features(const asdm_graph *g, const asdm_cs *cs, const uint64_t *masks, uint32_t n,
                       uint32_t *node_off, float *x, uint32_t *adj_off, uint32_t *adj, uint32_t *num_nodes,
                       uint32_t *num_adj);
/* Estimated cardinalities 2^y - 1 of the sub-queries induced by masks[n]. */
int asdm_mpnn_estimate(asdm_mpnn *m, const asdm_graph *g, const asdm_cs *cs, const uint64_t *masks, uint32_t n,
                       double *cards);
/* The plain planner's independence-model cardinalities of the same sub-queries:
 * prod_u |C(u)| * prod over sub-query edges (u, w) of W(u->w) / (|C(u)| |C(w)|). */
int asdm_plain_estimate(const asdm_cs *cs, const uint64_t *masks, uint32_t n, double *cards);
/* Plan whose enumeration order is chosen greedily by the learned estimates (the plain planner's
 * leaf set and first vertex; with ASDM_PLAN_COST in flags, the order with the smallest summed
 * estimated prefix cardinalities, as for asdm_plan_create).  The plain planner is used when the
 * network cannot run. */
int asdm_plan_create_learned(asdm_mpnn *m, const asdm_graph *g, const asdm_graph *q, const asdm_cs *cs,
                             uint32_t flags, asdm_plan **out);

/* ---- .card planner files ---- */
/* Reads the subsetToCard map (and the stored planning time) of a .card file written by the
 * reference's prepare.out (executable/prepare.cpp:65-71, plan/estimator.cpp:981-1008) or by
 * asdm_card_write.  With masks == NULL only *n is returned. */
int asdm_card_read(const char *path, uint64_t *masks, double *cards, uint32_t cap, uint32_t *n,
                   double *dp_time);
/* Writes a .card file holding only the map and the time (no per-bag DP blobs, which this
 * planner does not use); values are truncated to two decimals like saveSubsetToCard. */
int asdm_card_write(const char *path, const uint64_t *masks, const double *cards, uint32_t n,
                    double dp_time);
/* Planner subset cardinalities of a plan: masks/cards of every prefix of its order. */
int asdm_plan_cards(const asdm_plan *p, uint64_t *masks, double *cards, uint32_t cap, uint32_t *n);

/* ---- primitives (tests, autotuner) ---- */
/* k-way intersection of ascending unique device-independent host lists, computed on the GPU:
 * out[cap] receives the ascending intersection, *n its size. */
int asdm_intersect_k(asdm_ctx *ctx, const uint32_t *const *lists, const uint32_t *sizes, uint32_t k,
                     uint32_t *out, uint32_t cap, uint32_t *n);
/* asdm_intersect_k without any allocation, device or host: the lists are staged, and the result
 * built, in a fixed scratch array that each GPU holds for the life of the process.  It has no
 * counterpart in the reference; it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Limits: at most 64 lists and
 * 2^21 list elements in total per call, else ASDM_E_BUDGET.  Every call on one GPU shares that
 * scratch, so calls that overlap on one GPU (from different host threads or contexts) race. */
int asdm_intersect_k_noalloc(asdm_ctx *ctx, const uint32_t *const *lists, const uint32_t *sizes, uint32_t k,
                             uint32_t *out, uint32_t cap, uint32_t *n);
/* The bag engine's dense x dense table product G = A diag(u) B (mod 2^64) on host arrays (no
 * counterpart in the reference; for tests): A is nx x nv row-major, or its transpose (nv x nx) when
 * a_kmajor; B is nv x ny row-major when b_kmajor, else its transpose (ny x nv); u (nv) may be NULL
 * (all ones); G is nx x ny row-major.  method 1: the LDS-tiled 64-bit kernel (a_kmajor 0, b_kmajor 1
 * only); 2: the int8-digit matrix-core form. */
int asdm_dense_product(asdm_ctx *ctx, const uint64_t *A, int a_kmajor, const uint64_t *u, const uint64_t *B,
                       int b_kmajor, uint32_t nx, uint32_t nv, uint32_t ny, uint64_t *G, int method);

#ifdef __cplusplus
}
#endif
#endif /* ASDM_H */
