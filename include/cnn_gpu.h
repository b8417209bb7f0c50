/*
 * cnn_gpu.h -- C-ABI of the MI355X CommonNN clustering engine (libcnn_hip.so).
 *
 * Plain C types only: device pointers are `void *` / typed pointers to HBM,
 * sizes are int64_t.  Every function returns 0 (CNN_OK) or a negative
 * cnn_status; cnn_last_error() gives the message.  No C++ exception crosses
 * this boundary.  All work is enqueued on the context's stream; functions
 * that return a host-side size synchronise that stream.
 *
 * Reference interfaces replaced (cnnclustering 0.5.0, janjoswig/CommonNNClustering):
 *   cnn_graph_build / cnn_radius_csr
 *       NeighboursGetterExtBruteForce._get          _types.pyx:1692-1710
 *       DistanceGetterExtMetric._get_single         _types.pyx:1979-1985
 *       MetricExtEuclideanReduced._calc_distance    _types.pyx:2253-2267
 *       MetricExtEuclideanReduced._adjust_radius    _types.pyx:2286-2287
 *   cnn_labels_from_csr / cnn_graph_labels
 *       FitterExtBFS._fit                           _fit.pyx:408-491
 *       NeighboursExtVectorCPPUnorderedSet._enough  _types.pyx:1561-1564
 *       SimilarityCheckerExtSwitchContains._check   _types.pyx:2580-2610
 *       SimilarityCheckerExtScreensorted._check     _types.pyx:2654-2706
 *   cnn_sort_by_size
 *       Labels.sort_by_size                         _types.pyx:212-286
 *   cnn_fit
 *       Clustering.fit (fit + sort_by_size part)    cluster.pyx:256-364
 *   cnn_graph_rethreshold
 *       NeighboursGetterRecomputeLookup radius cache _types.pyx:1867-1873
 *       (re-fit at a smaller radius from a resident graph)
 *
 * Radius argument: every entry point takes r2, the squared cutoff computed in
 * double as r*r -- the value the reference's make_parameters stores in
 * ClusterParameters.radius_cutoff (_fit.pyx:136-141) and the getter compares
 * against (_types.pyx:1708).  Passing r2 avoids a lossy sqrt/square round trip.
 *
 * Semantics (bit-exact with the reference):
 *   s_ij = sum_k (x_ik - x_jk)^2, k ascending, FP64, each op separately rounded
 *   N(i) = { j : s_ij <= r2 }, i in N(i); core(i) = consider(i) && |N(i)| > m_core
 *   edge(i,j) = core(i) && core(j) && j in N(i) && |N(i) & N(j)| >= c_sim
 *   labels = connected components of (core points, edges), numbered by
 *   ascending smallest point index from current_start.
 *   With cnn_cutoff c, the reference's make_parameters gives c_sim = c + 2 and
 *   m_core = c + 1 (_fit.pyx:143-159).
 */
#ifndef CNN_GPU_H
#define CNN_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ status */
typedef enum {
    CNN_OK = 0,
    CNN_ERR_INVALID = -1,   /* bad argument (null pointer, negative size, bad enum) */
    CNN_ERR_OOM = -2,       /* device allocation failed or the memory budget is exceeded */
    CNN_ERR_HIP = -3,       /* a HIP runtime call failed */
    CNN_ERR_CAPACITY = -4,  /* caller buffer too small; required size returned */
    CNN_ERR_UNSUPPORTED = -5
} cnn_status;

/* dtype of a coordinate buffer */
enum { CNN_F32 = 0, CNN_F64 = 1 };

/* neighbour-search precision */
enum {
    CNN_PLAIN_F64 = 0,  /* every distance in FP64, reference order */
    CNN_F32_BAND = 1    /* FP32 distances (VALU direct form for d < 16, MFMA Gram
                           form for d >= 16); pairs within the guard band of r*r
                           are recomputed in FP64 in reference order */
};

/* similarity kernel selection.  A similarity test |N(i) & N(j)| >= c_sim is
 * either a merge of the two sorted lists or an AND + popcount of the rows'
 * 64-bit words over the column tiles both rows reach.  CNN_SIM_AUTO takes the
 * words when the graph keeps them (spatial graphs whose words take no more
 * memory than their lists: at least two neighbours per word on average) or,
 * for lists, when a sample of their rows (every stride-th row, the stride at
 * least 16 and about nnz / 2^21) averages at least 8 entries per word
 * (dense rows; the threshold is tunable, cnn_tune_params); otherwise the
 * merge. */
enum { CNN_SIM_AUTO = 0, CNN_SIM_MERGE = 1, CNN_SIM_BITSET = 2 };

typedef struct cnn_ctx cnn_ctx;
typedef struct cnn_graph cnn_graph;

typedef struct {
    int64_t pairs_evaluated;   /* point pairs whose distance was computed (tile and row-group culling
                                  applied; a symmetric launch computes each unordered pair once) */
    int64_t band_hits;         /* FP32 pairs recomputed in FP64 */
    int64_t band_flips;        /* band pairs whose FP64 decision differed from the FP32 one */
    double max_abs_err_band;   /* max |d2_fp32 - d2_fp64| over band pairs */
    int64_t tile_pairs;        /* 64x64 tile pairs evaluated */
    int64_t nnz;               /* neighbour-list entries (self included) */
    float ms_prep;             /* ordering, tiling, candidate tile pairs */
    float ms_distance;         /* distance kernel */
    float ms_emit;             /* degree scan + CSR emit */
    int32_t distance_launches; /* launches of the distance kernel timed in ms_distance */
    int32_t layout;            /* 0 identity order, 1 spatial (Morton) order */
    int64_t tile_pairs_computed; /* 64x64 tile pairs the distance kernel evaluated (symmetric: C >= R;
                                    dense layout: those the pivot bounds left) */
    float ms_distance_kernel;  /* the main distance kernel alone (ms_distance also holds guard-band settling) */
} cnn_graph_stats;

typedef struct {
    int64_t n_core;
    int64_t n_components;      /* components among core points (before sort_by_size) */
    int64_t n_clusters;        /* clusters after sort_by_size */
    int64_t tests;             /* similarity tests executed */
    int64_t tests_passed;
    int64_t sampled_components;/* components among core points after the sampling round */
    int64_t rows_scanned;      /* rows whose candidate edges the full round scanned */
    float ms_similarity;       /* similarity + union-find kernels */
    float ms_labels;           /* compress, label assignment, sort_by_size */
    int32_t sim_form;          /* how similarities were tested: CNN_SIM_FORM_* */
    float ms_hook;             /* the hooking kernels (k_hook_tile + k_hook_queue, or k_hook) */
    float ms_round;            /* the full-round kernels (k_round_tile + k_round_queue, or k_round_b) */
} cnn_label_stats;

/* cnn_label_stats.sim_form */
enum {
    CNN_SIM_FORM_MERGE = 1,       /* wave-wide merge of the two sorted neighbour lists */
    CNN_SIM_FORM_TILE_WORDS = 2,  /* AND + popcount of the rows' words in the spatial graph's
                                     tile-pair masks (no neighbour lists needed) */
    CNN_SIM_FORM_ROW_WORDS = 3    /* AND + popcount of per-row word lists built from the lists */
};

/* ------------------------------------------------------------------ context */
/* Create a context on HIP device `device`.  The context owns a non-blocking
 * stream (default priority) and a workspace arena. */
int cnn_ctx_create(int device, cnn_ctx **out);
int cnn_ctx_destroy(cnn_ctx *ctx);
/* Message of the last failing call on this context (never NULL). */
const char *cnn_last_error(const cnn_ctx *ctx);
/* Enqueue subsequent work on `stream` (a hipStream_t; NULL is the default
 * (null) stream).  Until this is called the context uses its own stream. */
int cnn_ctx_set_stream(cnn_ctx *ctx, void *stream);
void *cnn_ctx_get_stream(cnn_ctx *ctx);
int cnn_ctx_synchronize(cnn_ctx *ctx);
/* Device bytes currently held by the context's workspace arena. */
int cnn_ctx_workspace_bytes(const cnn_ctx *ctx, int64_t *bytes);
/* Release the workspace arena (kept between calls otherwise). */
int cnn_ctx_trim(cnn_ctx *ctx);

/* ------------------------------------------------------------------ phase N */
/* Build the radius-r neighbour graph of X (device, row-major n x d, dtype
 * CNN_F32 or CNN_F64) and keep it resident in device memory.  The graph is
 * stored in an internal point order (spatial order for d <= 8); export with
 * cnn_graph_export_csr. */
int cnn_graph_build(cnn_ctx *ctx, const void *X, int64_t n, int32_t d, int32_t dtype,
                    double r2, int32_t precision, cnn_graph **out);
int cnn_graph_destroy(cnn_graph *g);
int cnn_graph_nnz(const cnn_graph *g, int64_t *nnz);
int cnn_graph_stats_get(const cnn_graph *g, cnn_graph_stats *out);
/* Device bytes held by the graph (working coordinates, tile-pair lists, row
 * words and/or neighbour lists; lists emitted on demand count once emitted). */
int cnn_graph_bytes(const cnn_graph *g, int64_t *bytes);
/* Write the neighbour lists in the caller's point order, each row ascending:
 * row_ptr (device int64[n+1]), col (device int32[nnz]).  col_cap < nnz returns
 * CNN_ERR_CAPACITY. */
int cnn_graph_export_csr(cnn_ctx *ctx, const cnn_graph *g, int64_t *row_ptr, int32_t *col,
                         int64_t col_cap);
/* One-call neighbour search (build + export + destroy).  When col_cap is too
 * small, *nnz_out holds the required size and CNN_ERR_CAPACITY is returned. */
int cnn_radius_csr(cnn_ctx *ctx, const void *X, int64_t n, int32_t d, int32_t dtype,
                   double r2, int32_t precision, int64_t *row_ptr, int32_t *col,
                   int64_t col_cap, int64_t *nnz_out, cnn_graph_stats *stats_out);
/* Derive the graph at a smaller radius from a resident graph (neighbour sets
 * at r2' <= r2 are subsets; distances are re-evaluated exactly in FP64). */
int cnn_graph_rethreshold(cnn_ctx *ctx, const cnn_graph *g, double r2,
                          int32_t precision, cnn_graph **out);

/* ------------------------------------------------------------------ phase S+U */
/* Labels from a resident graph.  consider (device uint8[n], caller order) may
 * be NULL (all considered); when given it is set to 0 for every point, as the
 * reference's BFS leaves it.  labels (device int64[n], caller order) is written
 * only for points that join a cluster (other entries are left untouched), with
 * clusters numbered from current_start by ascending smallest point index. */
int cnn_graph_labels(cnn_ctx *ctx, const cnn_graph *g, int32_t c_sim, int32_t m_core,
                     uint8_t *consider, int64_t current_start, int32_t sim_mode,
                     int64_t *labels, cnn_label_stats *stats_out);
/* The same from a caller CSR (device; rows sorted ascending, no duplicates). */
int cnn_labels_from_csr(cnn_ctx *ctx, const int64_t *row_ptr, const int32_t *col, int64_t n,
                        int32_t c_sim, int32_t m_core, uint8_t *consider, int64_t current_start,
                        int32_t sim_mode, int64_t *labels, cnn_label_stats *stats_out);
/* cnn_labels_from_csr without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * stream capture, serving loops with a fixed memory plan).  Same arguments,
 * semantics and results as cnn_labels_from_csr; every temporary comes from a
 * fixed device workspace that the library holds per device (32 MiB, from the
 * first context created on the device until the process ends), so
 * n <= CNN_NOALLOC_MAX_POINTS (larger n: CNN_ERR_INVALID) and similarities are
 * tested by the sorted-list merge (CNN_SIM_BITSET: CNN_ERR_INVALID).  All calls
 * on one device share the workspace: calls that overlap on one device (other
 * host threads, other streams) race and must be serialised by the caller. */
#define CNN_NOALLOC_MAX_POINTS (1 << 20)
int cnn_labels_from_csr_noalloc(cnn_ctx *ctx, const int64_t *row_ptr, const int32_t *col, int64_t n,
                                int32_t c_sim, int32_t m_core, uint8_t *consider, int64_t current_start,
                                int32_t sim_mode, int64_t *labels, cnn_label_stats *stats_out);
/* Labels.sort_by_size in place on device int64 labels: clusters ranked by size
 * descending, ties by first occurrence; ranks beyond max_clusters (< 0: no
 * limit) and clusters smaller than member_cutoff become 0; survivors are
 * renumbered 1..K.  *n_clusters_out = K (may be NULL). */
int cnn_sort_by_size(cnn_ctx *ctx, int64_t *labels, int64_t n, int64_t member_cutoff,
                     int64_t max_clusters, int64_t *n_clusters_out);

/* ------------------------------------------------------------------ fused fit */
/* Fresh fit: neighbour graph, similarity, union-find and (sort != 0)
 * sort_by_size, writing all n labels (device int64, caller order).  When
 * graph_out is not NULL the neighbour graph stays resident for re-fits. */
int cnn_fit(cnn_ctx *ctx, const void *X, int64_t n, int32_t d, int32_t dtype,
            double r2, int32_t c_sim, int32_t m_core, int32_t precision,
            int32_t sort, int64_t member_cutoff, int64_t max_clusters, int64_t *labels,
            cnn_graph **graph_out, cnn_graph_stats *gstats, cnn_label_stats *lstats);
/* Fresh fit on a resident graph (labels for all n points). */
int cnn_graph_fit(cnn_ctx *ctx, const cnn_graph *g, int32_t c_sim, int32_t m_core,
                  int32_t sort, int64_t member_cutoff, int64_t max_clusters, int64_t *labels,
                  cnn_label_stats *lstats);

/* ------------------------------------------------------------------ batched */
/* Many independent datasets, each fitted as by cnn_fit.  Dataset b has
 * offsets[b] .. offsets[b+1] rows of X (device, row-major, common d and dtype;
 * offsets is a HOST int64[n_sets+1]) and its own r2 / c_sim / m_core (HOST
 * arrays).  labels (device int64[offsets[n_sets]]) receives every dataset's
 * sorted labels (sort_by_size per dataset with member_cutoff / max_clusters).
 * Replaces a loop of Clustering.fit calls over datasets (cluster.pyx:256-392). */
int cnn_fit_batch(cnn_ctx *ctx, const void *X, const int64_t *offsets, int32_t n_sets, int32_t d,
                  int32_t dtype, const double *r2, const int32_t *c_sim,
                  const int32_t *m_core, int32_t precision, int64_t member_cutoff,
                  int64_t max_clusters, int64_t *labels, cnn_graph_stats *gstats,
                  cnn_label_stats *lstats);

/* Parameter scan on one dataset: fits for every (r2[ri], c_sim[ci] / m_core[ci])
 * pair, as a loop of Clustering.fit calls or HierarchicalFitterRepeat._fit
 * steps (_fit.pyx:1058-1187) would produce them.  The neighbour graph is built
 * once at the largest radius; smaller radii filter the lists of the next larger
 * one exactly in FP64 (the radius cache of NeighboursGetterRecomputeLookup,
 * _types.pyx:1867-1873, without recomputation).  labels: device
 * int64[n_r * n_c * n], block (ri, ci) at offset (ri * n_c + ci) * n, sorted by
 * size with member_cutoff / max_clusters.  gstats: the graph at the largest
 * radius; lstats (nullable): HOST array of n_r * n_c entries. */
int cnn_fit_scan(cnn_ctx *ctx, const void *X, int64_t n, int32_t d, int32_t dtype,
                 const double *r2, int32_t n_r, const int32_t *c_sim, const int32_t *m_core,
                 int32_t n_c, int32_t precision, int64_t member_cutoff, int64_t max_clusters,
                 int64_t *labels, cnn_graph_stats *gstats, cnn_label_stats *lstats);

/* A parameter scan of a resident graph recorded once into a hipGraph and
 * replayed with one launch: for every r2[ri] (<= the graph's radius) and
 * every cutoff pair ci, the labels of a fresh fit, sorted by size, in block
 * (ri * n_c + ci) * n of labels (device int64).  Per radius the graph's row
 * words are re-thresholded exactly in FP64; each labelling kernel runs once
 * per radius for all cutoff pairs.  Needs a graph that keeps its row words
 * (graphs whose words take no more memory than their lists would, in either
 * layout: CNN_ERR_CAPACITY otherwise; use cnn_fit_scan, which falls back to
 * lists).  The scan holds its buffers until
 * cnn_scan_destroy; the graph and labels must outlive it.  lstats (nullable):
 * HOST array of n_r * n_c entries.  Replaces, like cnn_fit_scan, a loop of
 * Clustering.fit calls or HierarchicalFitterRepeat._fit steps
 * (_fit.pyx:1058-1187). */
typedef struct cnn_scan cnn_scan;
int cnn_graph_scan_create(cnn_ctx *ctx, const cnn_graph *g, const double *r2, int32_t n_r, const int32_t *c_sim,
                          const int32_t *m_core, int32_t n_c, int64_t member_cutoff, int64_t max_clusters,
                          int64_t *labels, cnn_scan **out);
int cnn_scan_launch(cnn_ctx *ctx, cnn_scan *scan, cnn_label_stats *lstats);
int cnn_scan_destroy(cnn_scan *scan);

/* ------------------------------------------------------------------ predict */
/* Label prediction for a second point set, replacing PredictorFirstmatch.predict
 * (_fit.pyx:1241-1291) as called by Clustering.predict (cluster.pyx:560-653)
 * with brute-force neighbour getters (_get / _get_other, _types.pyx:1692-1731).
 * X: reference points (n_x, d) with labels_x (device int64[n_x], caller order);
 * Y: points to label (n_y, d), same dtype as X.  allowed: device uint8 table
 * indexed by label (n_allowed entries; labels outside it are not allowed), the
 * reference's consider_set.  For every y with consider_y != 0 and more than
 * m_core reference points within r, the reference neighbours are visited in
 * ascending index; the first with an allowed label, more than m_core reference
 * neighbours of its own and at least c_sim reference neighbours shared with y
 * gives y its label and clears consider_y.  Other entries are left unchanged.
 * r2 = r*r, c_sim = c + 2, m_core = c + 1 as make_parameters (_fit.pyx:186-216).
 * tests_out (nullable, host): similarity tests executed. */
int cnn_predict(cnn_ctx *ctx, const void *X, int64_t n_x, const void *Y, int64_t n_y, int32_t d,
                int32_t dtype, double r2, int32_t precision, int32_t c_sim, int32_t m_core,
                const int64_t *labels_x, const uint8_t *allowed, int64_t n_allowed,
                int64_t *labels_y, uint8_t *consider_y, int64_t *tests_out);

/* ------------------------------------------------------------------ other metrics */
/* Neighbour lists from a precomputed distance matrix, replacing
 * NeighboursGetterExtBruteForce._get (_types.pyx:1692-1710) over
 * MetricExtPrecomputed (_types.pyx:2105-2127): N(i) = { j : D[i, j] <= r },
 * ascending j, r used as given.  D: device (n, n) row-major, CNN_F32 / CNN_F64
 * (compared in FP64).  row_ptr: device int64[n+1]; col: device int32[col_cap].
 * *nnz_out receives the entry count; when col_cap < nnz, CNN_ERR_CAPACITY is
 * returned with row_ptr complete and col untouched.  *asymmetric_out (host)
 * counts pairs {i, j} whose decisions for (i, j) and (j, i) differ: the labelling
 * entry points treat lists as a symmetric relation, which the reference's BFS
 * matches only when this count is 0.  (The non-reduced Euclidean metric needs
 * no entry point of its own: fl(sqrt(s)) <= r is decided as s <= T with T the
 * largest double whose correctly rounded square root is <= r, so the reduced-
 * metric kernels take T in place of r*r.) */
int cnn_precomputed_csr(cnn_ctx *ctx, const void *D, int64_t n, int32_t dtype, double r,
                        int64_t *row_ptr, int32_t *col, int64_t col_cap, int64_t *nnz_out,
                        int64_t *asymmetric_out);

/* Radius graph under the periodic reduced metric, replacing
 * NeighboursGetterExtBruteForce._get (_types.pyx:1692-1710) over
 * MetricExtEuclideanPeriodicReduced._calc_distance (_types.pyx:2296-2317): per
 * dimension k the difference |a_k - b_k| is reduced modulo bounds[k] and
 * replaced by bounds[k] - x above bounds[k]/2 (dimensions with bounds[k] <= 0
 * are not periodic), squared and summed in ascending k in FP64.  bounds: HOST
 * double[d].  Plain FP64, all tile pairs (no spatial culling); use the graph
 * with cnn_graph_labels / cnn_graph_fit / cnn_graph_export_csr.  (The
 * reference's _calc_distance_other for this metric adds (a_k - b_k)^2 instead
 * of the reduced difference, _types.pyx:2332-2343; prediction with it is not
 * offered.) */
int cnn_graph_build_periodic(cnn_ctx *ctx, const void *X, int64_t n, int32_t d, int32_t dtype,
                             double r2, const double *bounds, cnn_graph **out);

/* ------------------------------------------------------------------ MST hierarchy */
/* Similarity weight of every neighbour-list entry: weight[t] = |N(a) ∩ N(b)|
 * for the entry t of row a holding b, the count SimilarityChecker._get returns
 * (_types.pyx:2612-2634), which HierarchicalFitterMSTPrim._fit uses as edge
 * weight (_fit.pyx:851-856, 888-893).  row_ptr: device int64[n+1] from 0;
 * col: device int32[nnz], every row strictly ascending within [0, n) (checked:
 * CNN_ERR_INVALID otherwise); weight: device int32[nnz]. */
int cnn_csr_edge_similarity(cnn_ctx *ctx, const int64_t *row_ptr, const int32_t *col, int64_t n,
                            int32_t *weight);

/* Maximum spanning forest of the similarity-weighted lists, replacing the
 * Prim loop of HierarchicalFitterMSTPrim._fit (_fit.pyx:832-915) with its
 * PriorityQueueMaxHeap queues (_types.pyx:2825-2851): weights as
 * cnn_csr_edge_similarity on the device, then, on the host, every unvisited
 * point in ascending index starts a tree and the queued edge first in
 * (-weight, a, b) order with b unvisited joins it.  The tree edges (a, b,
 * weight) are written sorted in that order, the order in which the reference
 * pops its tree queue (_fit.pyx:937-939).  row_ptr / col: device, as
 * cnn_csr_edge_similarity; tree_a / tree_b / tree_w: HOST int64[cap]
 * (cap >= n - 1 always suffices); *n_edges_out: the tree edge count
 * (CNN_ERR_CAPACITY with outputs untouched when it exceeds cap). */
int cnn_similarity_mst(cnn_ctx *ctx, const int64_t *row_ptr, const int32_t *col, int64_t n,
                       int64_t *tree_a, int64_t *tree_b, int64_t *tree_w, int64_t cap,
                       int64_t *n_edges_out);

/* ------------------------------------------------------------------ tuning */
/* Launch shapes and algorithm switches chosen per problem class by the
 * autotuner (commonnnclustering_amd/autotune.py).  No counterpart in the
 * reference.  Every value is exact-result neutral: it changes speed only. */
typedef struct {
    int32_t mfma_col_chunk;  /* column tiles per workgroup of the dense MFMA distance kernel (8..1024; 64) */
    int32_t wide_col_chunk;  /* column tiles per workgroup of the dense plain-FP64 kernel (1..1024; 32) */
    int32_t emit_waves;      /* waves per workgroup of the neighbour-list emission (4 or 8; 8) */
    int32_t root_passes;     /* root-hooking passes of the labelling before its full round (0..8; 2) */
    int32_t words_min_fill;  /* CNN_SIM_AUTO on lists: per-row word lists when the lists hold at least
                                this many entries per word (estimated on every 16th row), else the
                                merge; 0 or more than 2^24: always the merge (0..2^30; 8) */
} cnn_tune_params;

/* Process-wide table of tuned shapes.  cnn_tune_set stores p for the class
 * d_lo <= d <= d_hi, n_lo <= n < n_hi, precision (CNN_PLAIN_F64 / CNN_F32_BAND);
 * later entries take precedence.  Graph builds and fits look up their class
 * (lists given by the caller: d = 0, precision CNN_PLAIN_F64); without an
 * entry the defaults above apply.  cnn_tune_get returns the shapes a problem
 * of that class uses; cnn_tune_clear empties the table. */
int cnn_tune_set(int32_t d_lo, int32_t d_hi, int64_t n_lo, int64_t n_hi, int32_t precision,
                 const cnn_tune_params *p);
int cnn_tune_get(int32_t d, int64_t n, int32_t precision, cnn_tune_params *out);
int cnn_tune_clear(void);

/* ------------------------------------------------------------------ info */
const char *cnn_version(void);

#ifdef __cplusplus
}
#endif
#endif /* CNN_GPU_H */
