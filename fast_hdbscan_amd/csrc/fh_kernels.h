// Launch interface of the gfx950 kernels (fh_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fh {

struct KnnArgs {
    const float *xp;        // tree-order rows
    const int2 *range;      // per node (start, end)
    const float *box;       // per node min[d], max[d]
    const int *perm;        // tree position -> original index
    const float *queries;   // nq rows
    const int *query_perm;  // self-query: original index of each query row; else nullptr
    int64_t nq;
    int d;
    int k;
    int levels;
    int block;
    float *out_d;           // [rows, k] squared distances, ascending (nullptr: not written)
    int *out_i;             // [rows, k] original indices
    float *core_o;          // optional: core by original index
    float *core_p;          // optional: core by query (tree) position
    int core_from_knn;      // 1: core = last column; 0: zeros (min_samples == 1)
    float *heap_p;          // global heap slices for large k (nullptr: LDS)
    int *heap_i;
    int64_t heap_stride;
    int *nb_pos;            // optional: [nq, k-1] neighbour tree positions (columns 1..k-1)
    float *nb_d;            // optional: [nq, k-1] their squared distances
    int sqrt_dist;          // 1: float32 square roots before the sort (output_rdist=False)
    const int *query_pos;   // optional: the nq queries are these tree positions of xp
    // optional ball bound per query (learned radius, squared): rows whose
    // result could depend on it (heap not filled inside the ball, or a tie
    // at an eviction, a rejection or in the output) are listed instead
    const float *bound_sq;
    int *miss_list;
    int *miss_count;
};

struct QueryArgs {
    const float *xp;
    const int2 *range;
    const float *box;
    const int *perm;
    const float *core_p;
    const int *comp_p;
    const int *node_comp;
    float *bound;                // per component label
    unsigned long long *best;    // per component label
    int *partner;                // per tree position
    int64_t n;
    int d;
    int levels;
    int pass;                    // 0: every point; 1: sampled points; 2: the rest, minus certified subtrees
    int sample_mask;             // 2^s - 1: position p is sampled iff (p & mask) == mask >> 1
    const int *cert_of_pos;      // certification node (in-level index) of each position
    const unsigned char *cert;   // 1: the node's points cannot beat their component's bound
    int wave;                    // passes 0 and 2 for d <= 8: one wavefront per 64 positions walks the tree
};

// Subtree certification: nodes of one level whose points provably have no
// out-of-component candidate at or below the component bound.
struct CertArgs {
    const float *xp;
    const int2 *range;
    const float *box;
    const float *core_p;
    const int *comp_p;
    const int *node_comp;
    const float *bound;
    const float *mincore;        // per certification node
    unsigned char *cert;
    int level;                   // tree level of the certification nodes
    int levels;
    int d;
};

// The post-search work of a round with few components, in one work-group.
struct SmallRoundArgs {
    int c;                               // components (<= small_round_limit())
    int n;
    int base;                            // first output row of the round
    const int *roots;                    // [c] component label per slot
    const unsigned long long *best;      // packed best edge per label
    const int *partner;                  // per position
    const int *posof, *perm, *comp_p;
    int *rootslot;                       // label -> slot (read, then rewritten for the next round)
    int *parent;                         // union-find over labels
    int *roots_out;                      // [c] surviving labels, compacted
    int *src_out, *dst_out;
    float *w_out;
    int *small;                          // [0] rows emitted, [1] components left
};

struct CycleBufs {
    int *S0, *M0, *S1, *M1;
};

hipError_t launch_knn(const KnnArgs &a, hipStream_t st);
// core_pt [n]: each point's core distance, by point index
hipError_t launch_init_edges_pos(int n, int k, const int *nb_pos, const float *nb_d, const float *core_p,
                                 const int *perm, const int *posof, int *succ, int *dst_raw, float *w, float *core_pt,
                                 hipStream_t st);
hipError_t launch_sweep_slice(int64_t n, int K, int ms, const int *sw_pos, const float *sw_d, int *nb_pos,
                              float *nb_d, float *core_p, int *tie_list, int *tie_count, hipStream_t st);
hipError_t launch_outputs_by_point(int64_t n, int k, const int *posof, const int *perm, const float *core_p,
                                   const int *nb_pos, float *core_out, int *nbr, hipStream_t st);
hipError_t launch_cert(const CertArgs &a, hipStream_t st);
hipError_t launch_cert_setup(int level, const int2 *range, const float *core_p, int *cert_of_pos, float *mincore,
                             hipStream_t st);
hipError_t launch_seed_bounds(int64_t n, int kn, const int *nb_pos, const float *nb_d, const float *core_p,
                              const int *comp_p, float *bound, unsigned char *interior, hipStream_t st);
hipError_t launch_query(const QueryArgs &a, hipStream_t st);
hipError_t launch_posof(int n, const int *perm, int *posof, hipStream_t st);
hipError_t launch_weighted_core(int n, int k, const int *knn_i, const float *knn_d, const float *weights, int ms,
                                const int *posof, float *core_o, float *core_p, hipStream_t st);
hipError_t launch_neighbors(int64_t n, int k, const int *knn_i, int *nbr, hipStream_t st);

hipError_t launch_node_ranges(int64_t N, int n, int2 *range, hipStream_t st);

int scan_tiles(int m);
hipError_t launch_excl_scan(int m, const int *in, int *out, int *tile_sum, int *total, hipStream_t st);
hipError_t launch_cycle_keep(int m, const int *succ, CycleBufs &b, int *keep, hipStream_t st);
// Keyed form (w never increases along succ): keep[] for every node not
// flagged, tflag[] = 1 for the nodes that may lie on a cycle longer than 2;
// then, with tpos = exclusive scan of tflag and t flagged nodes, keep[] for
// those (pointer doubling over the t of them; tnode [m] scratch).
hipError_t launch_cycle_classify(int m, const int *succ, const float *w, int *keep, int *tflag, hipStream_t st);
hipError_t launch_cycle_ties(int m, int t, const int *succ, const int *tflag, const int *tpos, CycleBufs &b,
                             int *tnode, int *keep, hipStream_t st);

hipError_t launch_init_edges(int n, int k, const int *knn_i, const float *knn_d, const float *core_o, int *succ,
                             int *dst_raw, float *w, hipStream_t st);
hipError_t launch_init_emit_union(int n, const int *keep, const int *offs, const int *succ, const int *dst_raw,
                                  const float *w, int *parent, int *src_out, int *dst_out, float *w_out,
                                  hipStream_t st);
hipError_t launch_labels_and_root_flags(int n, const int *perm, const int *parent, int *comp_p, int *flag,
                                        hipStream_t st);
hipError_t launch_root_scatter(int n, const int *flag, const int *offs, int *roots, int *rootslot, hipStream_t st);

hipError_t launch_round_reset(int c, const int *roots, float *bound, unsigned long long *best, hipStream_t st);
hipError_t launch_purity(int levels, const int2 *range, const int *comp_p, int *node_comp, hipStream_t st);
hipError_t launch_select(int c, int n, const int *roots, const unsigned long long *best, const int *partner,
                         const int *posof, const int *perm, const int *comp_p, const int *rootslot, int *succ,
                         int *ei, int *ej, float *ew, hipStream_t st);
hipError_t launch_round_emit_union(int c, int base, const int *keep, const int *offs, const int *ei, const int *ej,
                                   const float *ew, const int *roots, const int *succ, int *parent, int *src_out,
                                   int *dst_out, float *w_out, hipStream_t st);
hipError_t launch_relabel_and_flags(int n, int c, const int *roots, const int *parent, int *comp_p, int *flag,
                                    hipStream_t st);
int small_round_limit();
hipError_t launch_round_small(const SmallRoundArgs &a, hipStream_t st);
hipError_t launch_relabel(int n, const int *parent, int *comp_p, hipStream_t st);
hipError_t launch_roots_compact(int c, const int *roots, const int *flag, const int *offs, int *roots_out,
                                int *rootslot, hipStream_t st);

}  // namespace fh
