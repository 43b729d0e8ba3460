// High-dimensional candidate search: MFMA (FP32 matrix core) distance tiles
// that find candidates under certified error bounds, followed by exact
// re-scoring in the reference's float32 arithmetic.  Used for the kNN core
// distances and the Boruvka rounds when the KD-tree stops pruning (d large).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fh {

// Padded dimension for the MFMA tiles (multiple of 4, one of the compiled
// sizes); 0 if d is above the largest.
int brute_padded_dim(int d);

struct BruteScanArgs {
    const float *xc;      // [npad * dp] centred points, tree order, in tiles of 128 rows
    const float *nrm;     // [npad] squared norms of the centred rows (approximate)
    int64_t n, npad;
    int dp;
    // split != 0: xc holds the FP16 hi / lo split of the centred rows scaled
    // by a power of two s (launch_center_pack with scale s; dp 32 or 64),
    // multiplied on the 16x16x32 FP16 MFMA, and dot_scale = -2 / s^2
    int split;
    float dot_scale;
    float pack_scale;  // s (0 without the split), for re-packing the rows
    // kNN candidate collection
    const float *thr;     // [n] row thresholds (approximate squared distance)
    int *cand;            // [n * cap] candidate positions
    int *cand_cnt;        // [n] candidates found (may exceed cap)
    int cap;
    // k-th distance bounds (MODE 2): the k-th smallest is taken with k = cap
    const unsigned int *nrm_max_bits;
    float margin_c;
    // MODE 2: scan only the column tiles t with (t - row tile) a multiple
    // of tile_step (the row tile itself and every tile_step-th from it;
    // <= 1: every tile).  The k-th smallest over that sample is then no
    // bound of the k-th nearest distance but a guess of a threshold, which
    // the re-scoring checks row by row (KnnRescoreArgs::thr)
    int tile_step;
    // Boruvka row minimum
    const float *core_p;  // [n]
    const int *comp_p;    // [n]
    float *row_min;       // [n] smallest approximate mutual-reachability value
    int *row_arg;         // [n] position that gave it
    // Boruvka: per 128-row tile the component of all its rows, -1 when
    // mixed (nullptr: none); tile pairs of one component are skipped and
    // the tile pairs scored are added to *scored_tiles
    const int *tile_comp;
    unsigned long long *scored_tiles;
    // Boruvka, symmetric form (row_key != nullptr): a block scans only the
    // column tiles from its own on, and each tile pair's column minima go
    // to the column rows too; per row the packed (value bits << 32 | arg)
    // minimum, set to ~0 before the scan and unpacked into row_min /
    // row_arg after it (launch_unpack_row_keys)
    unsigned long long *row_key;
    // kNN candidates (MODE 0), symmetric form (sym_knn): a block scans only
    // the column tiles from its own on and appends in both directions,
    // counting in cand_cnt (zeroed before the scan)
    int sym_knn;
    // Boruvka, symmetric form, tile-pair bounds (rows in component order):
    // tmin_out (non-null): per tile pair scanned, the smallest approximate
    // squared distance over its valid row-column pairs, at [A * ntiles + B]
    // (A the row tile, B >= A); tmin_in (non-null): a column tile B is
    // skipped when max(tmin_in[A * ntiles + B] - m, tminc[A], tminc[B]) >
    // max(tU[A], tU[B]), with m = 2 margin_c u max|y'|^2 the largest error
    // of any pair: a lower bound of every mutual-reachability value of the
    // pair above every upper bound of the components with rows in either
    // tile
    float *tmin_out;
    const float *tmin_in;
    const float *tminc;
    const float *tU;
};

// keys[p] = comp_p[p], vals[p] = p (the sort into component order)
hipError_t launch_order_keys(int64_t n, const int *comp_p, uint32_t *keys, int *vals, hipStream_t st);
// Rows in another order for the Boruvka scans (order[p2] = tree position,
// -1 for the padding rows p2 >= n): xc, nrm and core2 in that order.
hipError_t launch_order_pack(int64_t n, int d, int dp, int64_t npad, const int *order, const float *xp,
                             const float *root_box, const float *core_p, float *xc, float *nrm, float *core2,
                             float scale, hipStream_t st);
// comp2[p2] = comp_p[order[p2]] (-3 for the padding rows)
hipError_t launch_gather_comp(int64_t n, int64_t npad, const int *order, const int *comp_p, int *comp2,
                              hipStream_t st);
// per tile: smallest and largest core distance of its valid rows
hipError_t launch_tile_core_range(int64_t n, int64_t npad, const float *core2, float *tminc, float *tmaxc,
                                  hipStream_t st);
// Upper bounds of the components' minimum outgoing weight from the
// tile-pair minima: for tiles A < B whose rows all lie in components a != b,
// some pair (i in A, j in B) has exact squared distance <= tmin + m (m as
// above) and cores <= tmaxc, so max(tmin + m, tmaxc[A], tmaxc[B]) bounds
// both components' minimum (float-bits atomicMin into bound[a], bound[b]).
hipError_t launch_tile_bounds(int ntiles, const int *tile_comp, const float *tmin, const float *tmaxc, float margin_c,
                              const unsigned int *nrm_max_bits, float *bound, hipStream_t st);
// per tile: the largest bound[comp2[row]] over its valid rows
hipError_t launch_tile_umax(int64_t n, int64_t npad, const int *comp2, const float *bound, float *tU,
                            hipStream_t st);
// row_key (scan order) into row_min / row_arg by tree position
hipError_t launch_unpack_row_keys_ordered(int64_t n, const unsigned long long *row_key, const int *order,
                                          float *row_min, int *row_arg, hipStream_t st);

hipError_t launch_tile_comp(int64_t n, int64_t npad, const int *comp_p, int *tile_comp, hipStream_t st);

// scale 0: FP32 tiles; scale > 0 (a power of two with |x - centre| s <
// 2^14, dp 32 or 64): the FP16 hi / lo split of (x - centre) s
hipError_t launch_center_pack(int64_t n, int d, int dp, int64_t npad, const float *xp, const float *root_box,
                              float *xc, float *nrm, unsigned int *nrm_max_bits, float scale, hipStream_t st);
// row thresholds into row_min (k = cap), then candidate collection
hipError_t launch_scan_kth(const BruteScanArgs &a, hipStream_t st);
hipError_t launch_scan_knn(const BruteScanArgs &a, hipStream_t st);
hipError_t launch_scan_boruvka(const BruteScanArgs &a, hipStream_t st);
hipError_t launch_unpack_row_keys(int64_t n, const unsigned long long *row_key, float *row_min, int *row_arg,
                                  hipStream_t st);

struct KnnRescoreArgs {
    const float *xp;
    const int *perm;
    const int *cand;
    const int *cand_cnt;
    int cap;
    int64_t n;
    int d, k;
    int core_from_knn;
    float *out_d;
    int *out_i;
    float *core_o;
    float *core_p;
    int *flag_list;   // positions whose result must come from the tree search
    int *flag_count;
    // the candidate thresholds of the scan (thr) with the margin
    // margin_c u (nrm[p] + max nrm): a row whose k-th smallest exact value
    // plus that margin exceeds its threshold may have missed a neighbour
    // and is listed for the tree search as well
    const float *thr;
    const float *nrm;
    const unsigned int *nrm_max_bits;
    float margin_c;
    float *cand_d;  // optional [n * cap]: the candidates' exact values, in list order
};
hipError_t launch_knn_rescore(const KnnRescoreArgs &a, hipStream_t st);

// tree kNN (the reference's heap order) for listed positions, a wave each
struct KnnWaveArgs {
    const float *xp;
    const int2 *range;
    const float *box;
    const int *perm;
    const int *query_pos;
    int d, k, levels;
    int core_from_knn;
    float *out_d;
    int *out_i;
    float *core_o;
    float *core_p;
};
hipError_t launch_knn_wave(const KnnWaveArgs &a, int count, hipStream_t st);
// The same results, a work-group per listed position (the depth-first leaf
// order precomputed, leaves taken several at a time); trees of more than
// 19 levels or leaves above 256 points go to launch_knn_wave.
hipError_t launch_knn_block(const KnnWaveArgs &a, int count, int64_t n, hipStream_t st);

struct BoruvkaExactArgs {
    const float *xp;
    const int2 *range;
    const float *box;
    const int *perm;
    const float *core_p;
    const int *comp_p;
    const float *nrm;
    const unsigned int *nrm_max_bits;
    float margin_c;
    const float *row_min;
    const int *row_arg;
    float *bound;                 // [components] exact upper bound, float bits min
    int *rows;                    // candidate rows
    int *row_count;
    unsigned long long *best;     // [components] packed (weight bits, original index)
    int *partner;                 // [n] partner position of each scanned row
    int64_t n;
    int d, levels;
};
hipError_t launch_boruvka_upper(const BoruvkaExactArgs &a, hipStream_t st);
// The same upper bound from the approximate row minima plus the margin
// (no exact weights): bound[c] <= min over c's rows of row_min + margin.
hipError_t launch_boruvka_upper_approx(const BoruvkaExactArgs &a, hipStream_t st);
hipError_t launch_boruvka_rows(const BoruvkaExactArgs &a, hipStream_t st);
hipError_t launch_boruvka_row_scan(const BoruvkaExactArgs &a, hipStream_t st);
// First Boruvka round from the kNN candidate lists (all columns with
// approximate distance <= thr[p] per row): rows whose best candidate in
// another component weighs at most thr[p] - margin_p are settled exactly
// (best, partner, cert[p] = 1); every candidate weight lowers bound; the
// other rows whose lower bound reaches their component's bound are listed
// in rows / row_count for the exact re-scan.
struct CandRoundArgs {
    const float *xp;
    const int2 *range;
    const float *box;
    int levels, d;
    int64_t n;
    const int *perm;
    const float *core_p;
    const int *comp_p;
    const int *cand;
    const int *cand_cnt;
    const float *cand_d;  // optional: the candidates' exact values (knn_rescore), else recomputed
    int cap;
    const float *thr;
    const float *nrm;
    const unsigned int *nrm_max_bits;
    float margin_c;
    float *bound;
    unsigned long long *best;
    int *partner;
    unsigned char *cert;
    int *rows;
    int *row_count;  // zero before the launch
};
hipError_t launch_cand_round(const CandRoundArgs &a, hipStream_t st);
// The same re-scan of rows [row0, row0 + nrows) of the list, tiled (64 rows
// against 2,048 positions per work-group); slot_w / slot_j hold
// boruvka_rows_tiled_slots(n, nrows) entries.
int64_t boruvka_rows_tiled_slots(int64_t n, int max_rows);
hipError_t launch_boruvka_rows_tiled(const BoruvkaExactArgs &a, int row0, int nrows, float *slot_w, int *slot_j,
                                     hipStream_t st);

}  // namespace fh
