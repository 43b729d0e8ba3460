/*
 * fh.h -- C ABI of the MI355X (gfx950) Euclidean MST library for HDBSCAN.
 *
 * The library replaces the Euclidean Boruvka branch of
 *   fast_hdbscan.hdbscan.compute_minimum_spanning_tree   (hdbscan.py:339-494)
 * i.e. build_kdtree (numba_kdtree.py:478-521) followed by
 *   parallel_boruvka (boruvka.py:552-647)
 * and the standalone kNN service
 *   parallel_tree_query (numba_kdtree.py:774-792).
 *
 * Conventions
 *   - Plain pointers and sizes only; no exceptions cross the ABI.
 *   - Every entry point returns FH_OK (0) or a negative FH_E* code; the
 *     message of the last failure on a context is fh_last_error().
 *   - Points are float32, row-major n x d.  Distances on the ABI are the
 *     reference's *squared* float32 distances (rdist); the caller takes the
 *     square roots exactly where the reference does (boruvka.py:646-647).
 *   - Edge rows come back in the reference's row order: the kNN-seeded
 *     initial edges in point order, then each Boruvka round's edges in the
 *     order of their component's lowest point index (boruvka.py:43-78,
 *     486-524, 608, 640).
 *   - A context owns one HIP stream and its device work buffers; it is not
 *     thread-safe.  All calls are synchronous: outputs are complete on
 *     return.
 */
#ifndef FH_H
#define FH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FH_OK 0
#define FH_EINVAL (-1)       /* bad argument (shape, size, null pointer)     */
#define FH_ENOMEM (-2)       /* device or host allocation failed              */
#define FH_EDEVICE (-3)      /* HIP runtime / kernel launch error             */
#define FH_EUNSUPPORTED (-4) /* valid input outside what the kernels handle   */

/* Largest dimension the KD-tree kernels accept.  The reference routes
 * d > 30 away from its KD-tree Boruvka (hdbscan.py:447-463). */
#define FH_MAX_DIM 64

typedef struct fh_ctx fh_ctx;

/* Per-call statistics; times are milliseconds measured with HIP events on
 * the context's stream (host phases with a steady clock). */
typedef struct {
    int32_t rounds;           /* Boruvka rounds after the kNN-seeded forest  */
    int32_t n_levels;         /* KD-tree levels                               */
    int64_t n_nodes;          /* KD-tree nodes                                */
    int64_t initial_components; /* components after the kNN-seeded forest    */
    int64_t n_edges;          /* rows written                                  */
    double t_tree_ms;         /* tree build (incl. its transfers)             */
    double t_knn_ms;          /* kNN / core-distance kernel                   */
    double t_init_ms;         /* kNN-seeded forest                             */
    double t_rounds_ms;       /* all Boruvka rounds                            */
    double t_query_ms;        /* per-round nearest-edge search, summed         */
    double t_query_main_ms;   /* its main-pass kernel alone, summed            */
    double t_total_ms;        /* whole call                                    */
    int32_t high_dim;         /* 1: MFMA candidate scans + exact re-scoring   */
    int32_t pad0;
    int64_t knn_tree_rows;    /* high_dim: kNN rows answered by the tree search */
    int64_t exact_rows;       /* high_dim: rows re-scanned exactly, summed     */
    int32_t learned;          /* learned-radius mode used for the kNN (0: off) */
    int32_t pad1;
    int64_t ball_misses;      /* learned: rows queried again without the ball  */
    double t_learn_ms;        /* learned: features, sample kNN, training       */
    double t_predict_ms;      /* learned: inference + radii                    */
    int64_t scored_tile_pairs; /* high_dim: 128 x 128 tile pairs the Boruvka scans scored, summed */
    int64_t tie_cycle_nodes;   /* components whose cycle test needed pointer doubling, summed over rounds */
    int32_t scan_rounds;       /* high_dim: rounds that ran the MFMA round scan (the first
                                  round may come from the kNN candidate lists); their
                                  main-pass time is t_query_main_ms */
    int32_t pad2;
} fh_stats;

/* Create a context on HIP device `device`.  Returns NULL on failure (no
 * device, out of memory).  Reference counterpart: none (numba keeps no
 * state); the context holds what parallel_boruvka allocates per call. */
fh_ctx *fh_create(int device);

/* Release the context, its stream and device buffers. */
void fh_destroy(fh_ctx *ctx);

/* Message of the last failing call on ctx (never NULL). */
const char *fh_last_error(const fh_ctx *ctx);

/* Library version string, e.g. "fh 0.1.0 gfx950". */
const char *fh_version(void);

/*
 * Mutual-reachability minimum spanning tree, inputs resident in HBM.
 *   Replaces: hdbscan.py:466 build_kdtree(data) +
 *             hdbscan.py:482-494 parallel_boruvka(tree, n_threads,
 *             min_samples, sample_weights, reproducible)
 *             (boruvka.py:552-647, all three branches).
 *
 * All pointers are DEVICE pointers on the context's device:
 *   X        [n*d]   float32 points, row-major (boruvka casts with
 *                    astype(float32), numba_kdtree.py:481; pass float32)
 *   sample_weights [n] float32 or NULL.  With weights the kNN width is
 *                    k = int(2 * min_samples / mean(w)) (boruvka.py:564-569),
 *                    distances are square-rooted before sorting and the
 *                    core is sample_weight_core_distance (527-539); the
 *                    returned core_sq and weights then hold the reference's
 *                    unsquared values, to which the caller still applies
 *                    its square root (boruvka.py:646-647).
 *   core_sq  [n]     squared core distance of each point: the
 *                    (min_samples+1)-th smallest rdist including the point
 *                    itself (boruvka.py:581-584); zeros when min_samples==1
 *                    (boruvka.py:592)
 *   nbr      [n*kn]  kn = fh_neighbor_columns(); kNN columns 1..kn of the
 *                    sorted query (boruvka.py:647 neighbors[:, 1:]), -1
 *                    where fewer than kn+1 points exist
 *   src,dst  [n-1]   edge endpoints (original point indices; dst may be -1
 *                    only when min_samples >= n, as in the reference)
 *   w_sq     [n-1]   float32 squared mutual-reachability weight
 * core_sq, nbr and stats may be NULL (not written).
 */
int fh_mst_euclidean_device(fh_ctx *ctx, const float *X, int64_t n, int32_t d, int32_t min_samples,
                            const float *sample_weights, float *core_sq, int32_t *nbr, int32_t *src, int32_t *dst,
                            float *w_sq, fh_stats *stats);

/* Same contract with HOST pointers; the library copies in and out. */
int fh_mst_euclidean(fh_ctx *ctx, const float *X, int64_t n, int32_t d, int32_t min_samples,
                     const float *sample_weights, float *core_sq, int32_t *nbr, int32_t *src, int32_t *dst,
                     float *w_sq, fh_stats *stats);

/*
 * Mutual-reachability MSTs of one dataset for several min_samples values
 * (a PLSCAN min_samples sweep), inputs resident in HBM.
 *   Replaces: for each value, hdbscan.py:466 build_kdtree(data) +
 *             parallel_boruvka(tree, n_threads, min_samples)
 *             (boruvka.py:552-647; layer_clusters.py:381-392 calls it once
 *             per PLSCAN fit): the same outputs as count calls of
 *             fh_mst_euclidean_device, computed from one tree and one kNN
 *             query of width max(min_samples) + 2 (rows whose
 *             min_samples + 2 smallest distances tie are queried again at
 *             their own width, whose heap order decides ties).
 * min_samples [count] HOST values in [2, 126]; DEVICE outputs laid out per
 * value: core_sq [count * n] (may be NULL), src, dst, w_sq [count * (n-1)].
 * Low-dimensional (KD-tree) path only, no sample weights: else
 * FH_EUNSUPPORTED.  stats: the last value's, t_total_ms for the call.
 */
int fh_mst_sweep_device(fh_ctx *ctx, const float *X, int64_t n, int32_t d, int32_t count,
                        const int32_t *min_samples, float *core_sq, int32_t *src, int32_t *dst, float *w_sq,
                        fh_stats *stats);

/*
 * The same MST without any allocation, for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.
 *
 * Same outputs as fh_mst_euclidean_device without sample weights (DEVICE
 * pointers on `device`; core_sq, nbr and stats may be NULL; stats carry no
 * times), with every work buffer carved from one static device array of
 * FH_NOALLOC_SCRATCH_MIB MiB that each GPU the library runs on holds until
 * the process ends.  The call makes no hipMalloc, hipHostMalloc or hipFree
 * and creates no stream or event: its kernels run on `stream` (a
 * hipStream_t; NULL for the null stream) and it returns once they are done.
 * Inputs that need more scratch than that (about 400,000 points at d = 2,
 * min_samples = 10; fewer in higher dimensions) return FH_EINVAL with a
 * message in err (err_cap bytes, may be NULL).  Calls on one device share
 * the array and run one at a time (a lock per device); the array keeps the
 * last call's contents, which no call reads.
 */
#define FH_NOALLOC_SCRATCH_MIB 128
int fh_mst_euclidean_device_noalloc(int32_t device, void *stream, const float *X, int64_t n, int32_t d,
                                    int32_t min_samples, float *core_sq, int32_t *nbr, int32_t *src, int32_t *dst,
                                    float *w_sq, fh_stats *stats, char *err, int32_t err_cap);

/* Width of the neighbour output: min_samples (1 when min_samples == 1), or
 * int(2 * min_samples / mean(w)) - 1 with HOST sample weights
 * (boruvka.py:564-569, 647).  -1 for invalid arguments. */
int64_t fh_neighbor_columns(int64_t n, int32_t min_samples, const float *sample_weights);

/*
 * Exact k nearest neighbours of Q's rows among X's rows.
 *   Replaces: numba_kdtree.py:765-792 parallel_tree_query(build_kdtree(X),
 *             Q, k, output_rdist): per row ascending float32 distances and
 *             int32 indices, in the reference's order for ties (its bounded
 *             max-heap and deheap_sort); rows with fewer than k points are
 *             padded with (+inf, -1).  output_rdist != 0: squared distances
 *             (rdist); 0: their float32 square roots, taken before the sort
 *             as numba_kdtree.py:792 does (equal roots of distinct squares
 *             keep the heap's order).
 * HOST pointers: X [n*d], Q [nq*d], dist [nq*k], idx [nq*k].
 */
int fh_knn(fh_ctx *ctx, const float *X, int64_t n, int32_t d, const float *Q, int64_t nq, int32_t k,
           int32_t output_rdist, float *dist, int32_t *idx);

/*
 * The KD-tree alone, built exactly as numba_kdtree.build_kdtree(X, 40)
 * (numba_kdtree.py:478-521).  HOST pointers; host-only work, so ctx may be
 * NULL.  N = fh_tree_nnodes(n).
 *   idx_array [n] int64, idx_start [N] int64, idx_end [N] int64,
 *   radius [N] float32, is_leaf [N] uint8, node_bounds [2*N*d] float32
 *   laid out [2][N][d] (min then max).
 */
int64_t fh_tree_nnodes(int64_t n);
int fh_build_kdtree(fh_ctx *ctx, const float *X, int64_t n, int32_t d, int64_t *idx_array, int64_t *idx_start,
                    int64_t *idx_end, float *radius, uint8_t *is_leaf, float *node_bounds);

/*
 * Kruskal MST of the mutual-reachability kNN graph.
 *   Replaces: kruskal.py:950-1034 (_kruskal_mst_knn, reached through
 *             hdbscan.py:466-478 with algorithm="kruskal", knn_k=int) without
 *             cannot-link constraints: tree, kNN of k_query columns, float64
 *             weights max(core_i, core_j, d_ij), stable global sort, Kruskal
 *             with union by rank and full path compression, +inf bridges
 *             between the smallest ids of disconnected components.
 * HOST pointers.  k_query = max(knn_k, min_samples) + 1 (min_samples > 1),
 * knn_k + 1 (min_samples == 1) or max(knn_k, k_for_core) + 1 with sample
 * weights, where k_for_core = int(2 * min_samples / mean(w)) is computed by
 * the caller as the reference does (float(np.mean(w)), kruskal.py:987-989).
 *   mst [(n-1)*3] float64 rows (u, v, w); nbr [n*nbr_cols] int32 with
 *   nbr_cols = min_samples, 1, or k_query-1 (weighted); core [n] float64.
 */
int fh_mst_kruskal_knn(fh_ctx *ctx, const float *X, int64_t n, int32_t d, int32_t min_samples, int32_t knn_k,
                       const float *sample_weights, int32_t k_for_core, double *mst, int32_t *nbr, int32_t nbr_cols,
                       double *core, fh_stats *stats);

/* ------------------------------------------------------------------ */
/* Many independent requests                                           */
/* ------------------------------------------------------------------ */

/*
 * The exact MSTs of count independent datasets (fh_mst_euclidean per
 * dataset, no sample weights).
 *   Replaces: a loop of compute_minimum_spanning_tree (hdbscan.py:339-494)
 *             calls, one per dataset (no counterpart in one reference call).
 * Datasets of at most a few thousand points (the KD-tree of one
 * work-group's LDS: 4032 points at d = 2, 3392 at d = 3, 2496 at d = 5,
 * 1792 at d = 8; min_samples <= 15) run in one persistent launch per
 * dimension: every work-group takes the next dataset from a device-side
 * queue and computes its tree, kNN, forest and every Boruvka round before
 * taking another.  The others run one at a time.  Rows, weights and core
 * distances equal fh_mst_euclidean's bit for bit.  HOST pointers: X[i]
 * [n[i]*d[i]]; outputs src[i], dst[i], w_sq[i] [n[i]-1] (the reference's
 * row order, squared weights), core_sq[i] [n[i]] or core_sq / core_sq[i]
 * NULL; persistent [count] (1: served by the persistent launch) or NULL.
 * Calls on one device run one at a time.
 */
int fh_mst_batch(int32_t device, int32_t count, const float *const *X, const int64_t *n, const int32_t *d,
                 int32_t min_samples, int32_t *const *src, int32_t *const *dst, float *const *w_sq,
                 float *const *core_sq, int32_t *persistent);

/*
 * Labels and membership strengths for count independent datasets (the
 * reference's fast_hdbscan(X_i, min_samples, min_cluster_size) per
 * dataset: Euclidean Boruvka MST, lexsort of the rows, EOM labels;
 * hdbscan.py:199-336).  The datasets fh_mst_batch takes in its persistent
 * launches run there; workers (1..16) threads, each with its own context
 * on `device`, take the others largest first, then the label stages of
 * the persistent launches' datasets.  HOST pointers: X[i] [n[i]*d[i]],
 * labels[i] [n[i]] int64, probabilities[i] [n[i]] float32; ms_out [count]
 * per-dataset wall times (the label stage alone for the persistent
 * launches' datasets) or NULL.  No context argument: the contexts are
 * created on first use and kept for later calls (released at process
 * exit); calls on one device run one at a time.
 */
int fh_hdbscan_batch(int32_t device, int32_t count, const float *const *X, const int64_t *n, const int32_t *d,
                     int32_t min_samples, int64_t min_cluster_size, int32_t workers, int64_t *const *labels,
                     float *const *probabilities, double *ms_out);

/* ------------------------------------------------------------------ */
/* Core graphs (lens-weighted kNN + MST graphs)                        */
/* ------------------------------------------------------------------ */

/*
 * The core graph as CSR.
 *   Replaces: core_graph.py:13-104 (knn_mst_union, flatten_to_csr,
 *             sort_by_lens): each point's children in insertion order (its
 *             kNN row skipping -1, then MST rows whose parent it is and
 *             whose child is new), weight = max(lens_p, lens_c), distance =
 *             max(core_p, core_c) (the MST distance for MST-only children),
 *             stored float32, each row ordered by numba's argsort of its
 *             weights.
 * HOST pointers.  neighbors [n*k] int32, core_distances [n] float32, mst
 * [m*3] float64 rows (parent, child, distance), lens [n] float64; outputs
 * indptr [n+1], indices/weights/distances [capacity] (n*k + m suffices).
 * Returns the entry count, or a negative FH_E* code.  Host-only.
 */
int64_t fh_core_graph(int64_t n, int32_t k, const int32_t *neighbors, const float *core_distances, int64_t m,
                      const double *mst, const double *lens, int32_t *indptr, int32_t *indices, float *weights,
                      float *distances, int64_t capacity);

/*
 * Boruvka on a CSR graph whose rows are sorted by weight.
 *   Replaces: core_graph.py:111-208 (boruvka_mst with select_components,
 *             merge_components, update_point_components,
 *             update_graph_components): the same edge rows in the same
 *             order, the reference's union-by-rank roots as component
 *             labels, stopping when a round adds no edge.
 * HOST pointers.  indptr [n+1], indices / weights [indptr[n]] (weights are
 * what the reference calls the graph's weights, e.g. lens values);
 * edges [(n-1)*3] float64 rows (point, child, weight); n_edges,
 * point_components [n] int64, n_components out.
 */
int fh_csr_boruvka(fh_ctx *ctx, int64_t n, const int32_t *indptr, const int32_t *indices, const float *weights,
                   double *edges, int64_t *n_edges, int64_t *point_components, int64_t *n_components);

/* ------------------------------------------------------------------ */
/* Learned search radii (no counterpart in the reference)              */
/* ------------------------------------------------------------------ */

/*
 * Core distances through learned ball bounds: a 3-layer FP32 MLP
 * (F = 2 + 2d inputs from KD-leaf statistics, 64, 64, 1) trained on the
 * device on a 1% sample's exact log core distances predicts a radius per
 * point; the reference's kNN query runs inside that ball and is repeated
 * without it wherever the ball could change the answer, so the outputs of
 * fh_mst_euclidean* are unchanged.  mode 0: off (default); 1: train on
 * every call; 2: train once per context and reuse.  Applies to d <= 8,
 * min_samples > 1, no sample weights.  FH_LEARNED_RADIUS sets the default.
 */
int fh_set_learned_radius(fh_ctx *ctx, int32_t mode);

/*
 * One training step of that network on a given batch, for verification:
 * forward, backward (loss = mean squared error), gradient summed over the
 * batch and one Adam step (fresh moments, step 1, learning rate lr; lr = 0
 * skips the update).  HOST pointers: x [B*F] (used as given), t [B],
 * params [P] in/out with P = 64F + 64 + 64*64 + 64 + 64 + 1 laid out
 * W1[64][F], b1, W2[64][64], b2, W3[64], b3; grad [P] and pred [B] may be
 * NULL.
 */
int fh_mlp_step(fh_ctx *ctx, int32_t F, int32_t B, const float *x, const float *t, float *params, float *grad,
                float *pred, float lr);

/* ------------------------------------------------------------------ */
/* Host label stage (the MST's consumer, kept on the host)             */
/* ------------------------------------------------------------------ */

typedef struct {
    int64_t min_cluster_size;          /* >= 1                                   */
    double max_cluster_size;           /* INFINITY: no limit (EOM only)         */
    double cluster_selection_epsilon;  /* 0: no epsilon search                  */
    int32_t selection_method;          /* 0: "eom", 1: "leaf"                  */
    int32_t allow_single_cluster;      /* 0 / 1                                  */
} fh_label_params;

/* Optional condensed-tree output; rows <= 2*(n-1).  `rows` is always set;
 * the arrays are filled when capacity >= rows. */
typedef struct {
    int64_t *parent;
    int64_t *child;
    float *lambda_val;
    float *child_size;
    int64_t capacity;
    int64_t rows;
} fh_condensed_tree;

/*
 * How the label stage may walk the rows; outputs never depend on it.
 *   point_order [n] or NULL: src/dst hold positions p of a point order and
 *     point_order[p] is the point id in [0, n) (fh_edges_to_tree_order;
 *     else FH_EINVAL); the union-find then runs over positions.
 *   prefix_rows k > 0: rows [0, k) form a forest whose connected
 *     components are given by row_order [k] (row indices grouped by
 *     component, increasing within a group) and group_sizes [n_groups]
 *     (fh_linkage_plan).  A row's linkage entry depends only on the earlier
 *     rows of its component, so the groups run on host threads (threads,
 *     0 = all cores); rows [k, n-1) that do not join the largest group run
 *     beside it, and those that do follow from running sums.
 */
typedef struct {
    const int32_t *point_order;
    int64_t prefix_rows;
    int64_t n_groups;
    const int32_t *row_order;
    const int32_t *group_sizes;
    int32_t threads;
    int32_t pad;
} fh_label_order;

/*
 * Labels and membership strengths from an MST whose rows are already
 * sorted by (weight, source, target).
 *   Replaces: hdbscan.py:521-574 (clusters_from_spanning_tree after its
 *             lexsort at 511-519), not semi-supervised,
 *             cluster_selection_persistence = 0: mst_to_linkage_tree (or
 *             mst_to_linkage_tree_w_sample_weights), condense_tree,
 *             extract_eom_clusters or cluster_tree_leaves,
 *             cluster_epsilon_search, get_cluster_label_vector,
 *             get_point_membership_strength_vector (cluster_trees.py).
 * HOST pointers.  src/dst [n-1]: sorted MST rows, endpoints in [-1, n) (-1
 * only as the MST gives it with min_samples >= n; with an order, positions),
 * else FH_EINVAL; their distances either
 * as w_sq [n-1] float32 squared (the float64 square root is taken inside,
 * as boruvka.py:646 does) or, when dist is not NULL, as dist [n-1] float64.
 * sample_weights [n] float32 or NULL (float32 weighted sizes in the
 * linkage and condensed trees, cluster_trees.py:29-35, 98-127, 240-250).
 * labels [n] int64, probabilities [n] float32.  linkage [(n-1)*4] float64
 * (rows left, right, distance, size, as mst_to_linkage_tree), condensed
 * and phase_ms [4] (linkage, condense, selection, labels; milliseconds)
 * may be NULL.  order may be NULL (see fh_label_order).  No context:
 * host-only.  n < 2^30 (cluster ids below 2n are held as int32), else
 * FH_EUNSUPPORTED.
 */
int fh_labels_from_sorted_mst(int64_t n, const int32_t *src, const int32_t *dst, const float *w_sq,
                              const double *dist, const float *sample_weights, const fh_label_params *params,
                              int64_t *labels, float *probabilities, double *linkage, fh_condensed_tree *condensed,
                              double *phase_ms, const fh_label_order *order);

/*
 * Sort MST rows by (weight, source, target) on the device -- the lexsort
 * of hdbscan.py:511-519 (the float64 square root preserves the order and
 * ties of the float32 squared weights).  DEVICE pointers, m rows; outputs
 * must not alias inputs.
 */
int fh_sort_mst_device(fh_ctx *ctx, int64_t m, const int32_t *src, const int32_t *dst, const float *w_sq,
                       int32_t *src_out, int32_t *dst_out, float *w_out);

/*
 * MST endpoints as positions in the KD-tree order of the context's last
 * MST call (the tree's idx_array, numba_kdtree.py:478-521): pos_src[i] =
 * position of src[i], likewise dst; order [n] = idx_array (point id at
 * each position).  Feeds fh_labels_from_sorted_mst's point_order, whose
 * union-find is faster over spatially ordered positions.  DEVICE pointers;
 * n must be the point count of that call.
 */
int fh_edges_to_tree_order(fh_ctx *ctx, int64_t n, int64_t m, const int32_t *src, const int32_t *dst,
                           int32_t *pos_src, int32_t *pos_dst, int32_t *order);

/*
 * Connected components of the forest of sorted MST rows [0, k) -- the
 * groups of an fh_label_order; no reference counterpart: it partitions
 * mst_to_linkage_tree's union-find, cluster_trees.py:66-95, into
 * independent parts.  DEVICE src/dst (point ids or positions, >= 0),
 * HOST outputs: row_order [k], group_sizes [k] (first *n_groups used).
 */
int fh_linkage_plan(fh_ctx *ctx, int64_t n, int64_t k, const int32_t *src, const int32_t *dst, int32_t *row_order,
                    int32_t *group_sizes, int64_t *n_groups);

/* ------------------------------------------------------------------ */
/* PLSCAN layer stage (host)                                           */
/* ------------------------------------------------------------------ */

typedef struct {
    int64_t base_min_cluster_size;      /* >= 1                                  */
    int64_t base_n_clusters;            /* 0: not given                          */
    double layer_similarity_threshold;  /* jaccard limit between layers         */
    int32_t max_layers;                 /* >= 1                                  */
} fh_layer_params;

/* Caller-allocated outputs of fh_cluster_layers. */
typedef struct {
    int64_t *labels;            /* [max_layers * n], layer-major              */
    float *strengths;           /* [max_layers * n]                           */
    float *persistence;         /* [max_layers]; 0 for the base layer         */
    float *sizes;               /* [capacity] min-cluster-size axis (may be NULL) */
    float *total_persistence;   /* [capacity] (may be NULL)                   */
    int64_t capacity;           /* entries of sizes / total_persistence      */
    int32_t n_layers;           /* out: base layer + selected peaks          */
    int64_t n_sizes;            /* out: 0 when there is no cluster split     */
    int64_t min_cluster_size;   /* out: the base layer's min cluster size    */
} fh_layers;

/*
 * Cluster layers from the MST rows already normalised and sorted the way
 * build_cluster_layers does before its linkage tree.
 *   Replaces: layer_clusters.py:399-481 (build_cluster_layers after
 *             np.argsort of the normalised weights): mst_to_linkage_tree,
 *             _binary_search_for_n_clusters, condense_tree(lambda "plscan"),
 *             extract_leaves, min_cluster_size_barcode,
 *             compute_total_persistence, find_peaks, select_diverse_peaks,
 *             extract_clusters_by_id.
 * HOST pointers.  src/dst [n-1] int32, dist [n-1] float64 (ascending).
 * Layers are returned unsorted: the base layer first, then one per
 * selected peak in selection order (the caller sorts them by cluster
 * count, layer_clusters.py:484-490).  No context: host-only.  n < 2^30,
 * else FH_EUNSUPPORTED.
 */
int fh_cluster_layers(int64_t n, const int32_t *src, const int32_t *dst, const double *dist,
                      const fh_layer_params *params, fh_layers *out);

/*
 * Parent/child links between clusters of stacked layers.
 *   Replaces: layer_clusters.py:498-525 (_build_cluster_tree).
 * HOST pointers.  labels [n_layers * n] int64, layer-major.  mapping
 * [capacity * 4] int64 rows (parent layer, parent cluster, child layer,
 * child cluster) in the reference's order; capacity = sum over layers of
 * (max label + 1) always suffices.  Returns the row count (rows beyond
 * capacity are counted, not written) or a negative FH_E* code.
 */
int64_t fh_layer_cluster_tree(int32_t n_layers, int64_t n, const int64_t *labels, int64_t *mapping,
                              int64_t capacity);

#ifdef __cplusplus
}
#endif

#endif /* FH_H */
