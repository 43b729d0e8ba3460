// High-dimensional candidate search on the matrix cores.
//
// In high dimension the KD-tree's box bounds stop pruning and every query
// degenerates into a scan of all leaves, so the kNN and the Boruvka rounds
// are done as blocked all-pairs scans instead.  The scans run on MFMA with
// the Gram expansion
//     ||x - y||^2 ~= ||x'||^2 + ||y'||^2 - 2 x'.y'     (x' = x - mu)
// which only *selects candidates*: every pair that could matter under a
// certified error bound is re-scored with the reference's own float32
// difference-sum (rdist, sequential FMA), so all outputs stay exact.
//
// Products.  FP32 form: v_mfma_f32_16x16x4f32 on x' itself.  Split form
// (the default; dp 32 or 64): x' is scaled by a power of two s with
// |x' s| < 2^14 (exact) and stored as the FP16 pair h = fp16(x' s),
// l = fp16(x' s - h) (x' s - h is exact in FP32); v_mfma_f32_16x16x32_f16
// sums h.l' and l.h' of every chunk first, then h.h' (each product exact in
// FP32), and -2 / s^2 (exact) scales the sum.  5.3x fewer matrix cycles
// than the FP32 form for the same margin.
//
// Error bound.  With u = 2^-24, |approx - rdist| <= (4d + 6) u (|x'|^2 +
// |y'|^2) covers the dot-product and norm sums (gamma_d each), the final
// combination, the centring roundings and rdist's own rounding (FP32
// form).  Split form: x' s = h + l + e with |e| <= 2^-22 |x' s| + 2^-25 (the
// second term from FP16 subnormals, or 2^-14 if the hardware flushes them:
// relative to s max|y'| either is below u).  The dropped l.l' and the e
// terms are at most 3 * 2^-22 = 12 u per |x'_i y'_i|; the small products'
// partial sums stay below 2^-9 sum|x'_i y'_i| s^2, so their roundings add at
// most 2 dp 2^-9 u; h.h' accumulates like the FP32 form (gamma_dp).  The
// dot product's error is thus gamma_dp + 13 u, and the total stays below
// (4 dp + 19) u (|x'|^2 + |y'|^2).  Both forms use margin_c * u * (|x'|^2 +
// max|y'|^2) with margin_c = 8 dp + 64, at least twice their bound.
//
// kNN (numba_kdtree.py:765-792): a first MFMA scan bounds each row's k-th
// distance from above (the k-th smallest of per-lane top-4 lists: k real
// columns); a second collects every column whose approximate distance is
// within two margins of that bound, a superset of the true k nearest;
// exact re-scoring sorts them.  When two
// of the k+1 smallest exact values tie, the reference's heap order decides
// the output, so such rows (and rows whose candidate list overflowed) are
// answered by the exact tree search instead.
//
// Boruvka round (boruvka.py:290-331 + 34-78): the MFMA scan gives each row
// the smallest approximate mutual-reachability value m_i to another
// component and the column j0 that gave it.  The exact weight of (i, j0) is
// a real candidate, so U_c = min_i exact(i, j0_i) bounds the component's
// minimum w*_c from above; any row with m_i - margin_i > U_c cannot reach
// w*_c.  The remaining rows (a few per component) are re-scanned exactly
// against every column, and ties in weight are broken by the reference's
// visiting order (first in the depth-first order of that row's KD-tree
// query: closer child first, left on equal bounds), which is the partner the
// tree search would report.  The packed atomicMin per component then picks
// the lowest original index at the minimum, as merge_components does.
#include "fh_brute.h"
#include "fh_device.h"

namespace fh {

namespace {

constexpr float kInfB = __builtin_huge_valf();
constexpr float kU = 5.9604644775390625e-08f;  // 2^-24

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef _Float16 half8 __attribute__((ext_vector_type(8)));

inline int64_t nblocks(int64_t n, int b) { return (n + b - 1) / b; }

#define FH_BRUTE_CHECK(x)                    \
    do {                                     \
        const hipError_t e_ = (x);           \
        if (e_ != hipSuccess) return e_;     \
    } while (0)

}  // namespace

int brute_padded_dim(int d) {
    static const int sizes[] = {16, 24, 32, 48, 64};
    for (int s : sizes)
        if (d <= s) return s;
    return 0;
}

// ------------------------------------------------------------- packing ----

// Rows are laid out in tiles of 128 (a tile is copied to LDS as it is):
// inside a tile, dimension k of row r sits at
//     (k / 4) * 512 + r * 4 + ((k % 4) ^ swz(r)),   swz(r) = 2 if r & 8,
// so one wave reads 64 consecutive floats per MFMA operand, and each half
// wave (lanes of g = k % 4 in {0, 1} or {2, 3}, rows r .. r + 15) hits 32
// distinct banks: without the swizzle rows r and r + 8 share a bank.
__device__ inline int tile_slot(int r, int k) { return (k >> 2) * 512 + r * 4 + ((k & 3) ^ ((r >> 2) & 2)); }

// The FP16 split form of a tile (SPLIT scans, dp = 32 or 64): dimension
// k = 32 c + 8 q + j of row r, plane p (0: hi, 1: lo), as a half at
//     ((c * 2 + p) * 4 + q) * 1024 + r * 8 + j
// (halves from the tile's start; a tile holds 128 dp floats' bytes).  The
// lane of the 16x16x32 MFMA that supplies row (or column) r, k = 8 q .. 8 q
// + 7 of chunk c reads one 16-byte vector; the 16 lanes of one q read 256
// consecutive bytes.
__device__ inline int split_slot(int r, int k) {
    return (((k >> 5) * 2) * 4 + ((k >> 3) & 3)) * 1024 + r * 8 + (k & 7);
}

// Store centred value v of row r, dimension k into a tile: FP32 (tile_slot)
// or, with scale > 0, the FP16 pair hi = fp16(v s), lo = fp16(v s - hi)
// (split_slot; s a power of two with |v s| < 2^14, so v s is exact).
__device__ inline void pack_store(float *tile, int r, int k, float v, float scale) {
    if (scale > 0.0f) {
        _Float16 *h = reinterpret_cast<_Float16 *>(tile);
        const float vs = v * scale;
        const _Float16 hi = (_Float16)vs;
        const _Float16 lo = (_Float16)(vs - (float)hi);
        const int slot = split_slot(r, k);
        h[slot] = hi;
        h[slot + 4096] = lo;  // plane 1: 4 blocks of 1024 halves further
    } else {
        tile[tile_slot(r, k)] = v;
    }
}

// Centre on the root box's midpoint, zero-pad to dp dimensions and pack the
// rows into tiles (tile_slot, or split_slot when scale > 0).
__global__ void center_pack_kernel(int64_t n, int d, int dp, int64_t npad, const float *__restrict__ xp,
                                   const float *__restrict__ root_box, float *__restrict__ xc, float *__restrict__ nrm,
                                   unsigned int *nrm_max_bits, float scale) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npad) return;
    // tile layout: 128 rows per tile, inside a tile [k / 4][row % 128][k % 4]
    float *tile = xc + (p / 128) * 128 * dp;
    const int r = (int)(p % 128);
    float s = 0.0f;
    for (int k = 0; k < dp; ++k) {
        float v = 0.0f;
        if (p < n && k < d) v = xp[p * d + k] - 0.5f * (root_box[k] + root_box[d + k]);
        pack_store(tile, r, k, v, scale);
        s = __builtin_fmaf(v, v, s);
    }
    nrm[p] = s;
    if (p < n) atomicMax(nrm_max_bits, __float_as_uint(s));
}

// --------------------------------------------------------- MFMA scans ----

// One block owns 128 rows (4 waves x 32) and streams every 128-column tile
// of the centred points through LDS, double-buffered (one barrier per
// tile).  Per tile each wave computes a 32 x 128 block of dot products:
// 2 row blocks x 8 column blocks of 16 x 16, K = 4 per instruction.
// MODE 0: collect kNN candidates under the row thresholds.
// MODE 1: smallest approximate mutual-reachability value to another
//         component per row.
// MODE 2: row thresholds for MODE 0.  Each lane keeps its 4 smallest
//         approximate distances per row; the k-th smallest of the union
//         (64 values from 16 lanes, all real columns) plus one margin
//         bounds the true k-th exact distance, plus a second margin bounds
//         the approximate value of every true k nearest neighbour.
template <int DP4, int MODE, bool SPLIT>
__global__ __launch_bounds__(256, MODE == 1 ? 2 : 3) void brute_scan_kernel(BruteScanArgs a) {
    constexpr int DP = 4 * DP4;
    // column tiles in LDS: double-buffered in the round scan (one barrier
    // per tile; its registers allow two wavefronts per SIMD anyway), single
    // in the kNN scans (two barriers per tile, three work-groups per CU)
    constexpr int NBUF = MODE == 1 ? 2 : 1;
    // SPLIT: FP16 hi / lo planes (split_slot), NC chunks of 32 dimensions
    constexpr int NC = SPLIT ? DP / 32 : 1;
    static_assert(!SPLIT || DP % 32 == 0, "split tiles hold whole 32-dimension chunks");
    constexpr int TILE = 128 * DP;  // floats per tile
    constexpr int NV = DP / 8;      // float4 loads per thread per tile
    __shared__ __align__(16) float bs[NBUF * TILE > 128 * 64 ? NBUF * TILE : 128 * 64];
    __shared__ float col_nrm[NBUF][128];
    __shared__ float col_core[NBUF][128];
    __shared__ int col_comp[NBUF][128];
    __shared__ float col_thr[NBUF][128];
    __shared__ int cnt_l[128];
    // MODE 1, symmetric: per wave and column, the column's minimum over the
    // wave's 32 rows (double-buffered like the tiles)
    __shared__ float cmin_s[MODE == 1 ? 2 : 1][4][128];
    __shared__ int carg_s[MODE == 1 ? 2 : 1][4][128];
    // MODE 1: the block's rows' core distances and components (read per
    // product from LDS rather than held in registers)
    __shared__ float row_core_s[MODE == 1 ? 128 : 1];
    __shared__ int row_comp_s[MODE == 1 ? 128 : 1];
    // MODE 1 with tmin_out: per wave the tile pair's smallest approximate
    // distance (double-buffered like the tiles)
    __shared__ float tmin_s[MODE == 1 ? 2 : 1][4];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int g = lane >> 4, c16 = lane & 15;
    const int64_t blk_row = (int64_t)blockIdx.x * 128;
    const int64_t rbase = blk_row + wave * 32;
    const int64_t n = a.n;

    float A[2][SPLIT ? 1 : DP4];
    half8 AH[2][NC], AL[2][NC];  // SPLIT: the row operands, hi and lo planes
#pragma unroll
    for (int rb = 0; rb < 2; ++rb) {
        if constexpr (SPLIT) {
            const _Float16 *src = reinterpret_cast<const _Float16 *>(a.xc + blk_row * DP) +
                                  (wave * 32 + rb * 16 + c16) * 8;
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                AH[rb][c] = *reinterpret_cast<const half8 *>(src + ((c * 2) * 4 + g) * 1024);
                AL[rb][c] = *reinterpret_cast<const half8 *>(src + ((c * 2 + 1) * 4 + g) * 1024);
            }
        } else {
            const float *src = a.xc + blk_row * DP + (wave * 32 + rb * 16 + c16) * 4 + (g ^ ((c16 >> 2) & 2));
#pragma unroll
            for (int s = 0; s < DP4; ++s) A[rb][s] = src[s * 512];
        }
    }
    // approx = |x'|^2 + |y'|^2 + c2 (x'.y') with the products' scale folded
    // in: -2 (FP32), -2 / s^2 (SPLIT: exact, s a power of two)
    const float c2 = SPLIT ? a.dot_scale : -2.0f;
    // the 8 output rows this lane holds: rbase + rb*16 + 4g + r
    float rn[2][4], rv[2][4], rcore[2][4];
    int ri[2][4], rcomp[2][4];  // MODE 1: rcomp -2 marks a padding row
    const bool sym = (MODE == 1 && a.row_key != nullptr) || (MODE == 0 && a.sym_knn != 0);
    // MODE 2: per row the lane's 4 smallest values, t[..][0] <= ... <= t[..][3]
    float t4[2][4][4];
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int q = 0; q < 4; ++q) t4[rb][r][q] = kInfB;
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int64_t row = rbase + rb * 16 + 4 * g + r;
            const bool ok = row < n;
            // padding rows and columns get +inf norms: every approximate
            // value of their pairs is +inf, which no comparison below takes
            rn[rb][r] = ok ? a.nrm[row] : kInfB;
            ri[rb][r] = -1;
            if constexpr (MODE == 0) {
                rv[rb][r] = ok ? a.thr[row] : -kInfB;
                rcore[rb][r] = 0.0f;
                rcomp[rb][r] = ok ? 0 : -2;
            } else if constexpr (MODE == 2) {
                rv[rb][r] = kInfB;
                rcore[rb][r] = 0.0f;
                rcomp[rb][r] = 0;
            } else {
                rv[rb][r] = kInfB;
                rcore[rb][r] = 0.0f;
                rcomp[rb][r] = 0;
            }
        }
    if (MODE == 0 && tid < 128) cnt_l[tid] = 0;
    if (MODE == 1 && tid < 128) {
        const int64_t row = blk_row + tid;
        row_core_s[tid] = row < n ? a.core_p[row] : 0.0f;
        row_comp_s[tid] = row < n ? a.comp_p[row] : -2;
    }

    const int ntiles = (int)(a.npad / 128);
    // the column tiles scanned: [tstart, ntiles)
    const int tstart = sym ? (int)blockIdx.x : 0;
    // MODE 1: column tiles whose rows all share this tile's single
    // component hold no candidate (boruvka.py:199-201) and are skipped
    const int my_tc = (MODE == 1 && a.tile_comp) ? a.tile_comp[blockIdx.x] : -1;
    // MODE 1 with tile-pair bounds: column tiles that no row of either tile
    // can use are skipped too
    const bool prune = MODE == 1 && a.tmin_in != nullptr;
    const float my_minc = prune ? a.tminc[blockIdx.x] : 0.0f;
    const float my_u = prune ? a.tU[blockIdx.x] : 0.0f;
    const float *tmin_row = prune ? a.tmin_in + (int64_t)blockIdx.x * ntiles : nullptr;
    const float tmargin = prune ? 2.0f * a.margin_c * kU * __uint_as_float(*a.nrm_max_bits) : 0.0f;
    auto next_tile = [&](int t) {
        if (MODE == 2 && a.tile_step > 1) {
            // the next sampled tile at or after t
            const int r = ((int)blockIdx.x - t) % a.tile_step;
            return t + (r < 0 ? r + a.tile_step : r);
        }
        for (; t < ntiles; ++t) {
            if (my_tc >= 0 && a.tile_comp[t] == my_tc) continue;
            if (prune) {
                float lb = tmin_row[t] - tmargin;
                const float mc = my_minc > a.tminc[t] ? my_minc : a.tminc[t];
                lb = lb > mc ? lb : mc;
                const float ub = my_u > a.tU[t] ? my_u : a.tU[t];
                if (lb > ub) continue;
            }
            break;
        }
        return t;
    };
    floatx4 pre[NV];
    float pn = 0.0f, pc = 0.0f, pt = 0.0f;
    int pk = -1;
    auto gload = [&](int t) {
        const floatx4 *src = reinterpret_cast<const floatx4 *>(a.xc + (int64_t)t * 128 * DP);
#pragma unroll
        for (int v = 0; v < NV; ++v) pre[v] = src[tid + v * 256];
        if (tid < 128) {
            const int64_t col = (int64_t)t * 128 + tid;
            pn = col < n ? a.nrm[col] : kInfB;
            if constexpr (MODE == 1) {
                pc = col < n ? a.core_p[col] : 0.0f;
                pk = col < n ? a.comp_p[col] : -3;
            }
            if constexpr (MODE == 0) pt = (sym && col < n) ? a.thr[col] : -kInfB;
            if constexpr (MODE == 1) {
                // symmetric: the column's best so far (any earlier value
                // of its packed key; one ulp above, so that equal values
                // still reach the packed minimum, which orders them)
                pt = -kInfB;
                if (sym && col < n) {
                    const unsigned long long key = a.row_key[col];
                    pt = key == ~0ull ? kInfB : __uint_as_float((unsigned int)(key >> 32) + 1u);
                }
            }
        }
    };
    auto lstore = [&](int buf) {
        floatx4 *dst = reinterpret_cast<floatx4 *>(bs + buf * TILE);
#pragma unroll
        for (int v = 0; v < NV; ++v) dst[tid + v * 256] = pre[v];
        if (tid < 128) {
            col_nrm[buf][tid] = pn;
            col_core[buf][tid] = pc;
            col_comp[buf][tid] = pk;
            col_thr[buf][tid] = pt;
        }
    };

    // the tiles scanned, in order: t (in buffer cur), then t1, then t2
    // (symmetric: from this block's own tile on)
    int t = next_tile(tstart);
    if (t < ntiles) {
        gload(t);
        lstore(0);
    }
    __syncthreads();
    int t1 = t < ntiles ? next_tile(t + 1) : ntiles;
    if (t1 < ntiles) gload(t1);
    int scored = 0;
    for (int cur = 0; t < ntiles; cur = NBUF == 2 ? cur ^ 1 : 0) {
        ++scored;
        const float *bt = bs + cur * TILE;
        floatx4 acc[2][8];
        const int64_t col0 = (int64_t)t * 128;
        float tmv = kInfB;  // MODE 1, tmin_out: the lane's smallest approximate distance of the pair
        // column block by column block: its 2 x DP4 products, then its
        // results, so that one block's compares overlap the next block's
        // products
#pragma unroll
        for (int cb = 0; cb < 8; ++cb) {
            acc[0][cb] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
            acc[1][cb] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
            if constexpr (SPLIT) {
                // hi.lo and lo.hi of every chunk first, hi.hi last: the
                // partial sums of the small products stay ~2^-10 of the
                // result, so their roundings add almost nothing
                const _Float16 *bh = reinterpret_cast<const _Float16 *>(bt) + (cb * 16 + c16) * 8;
                half8 BH[NC], BL[NC];
#pragma unroll
                for (int c = 0; c < NC; ++c) {
                    BH[c] = *reinterpret_cast<const half8 *>(bh + ((c * 2) * 4 + g) * 1024);
                    BL[c] = *reinterpret_cast<const half8 *>(bh + ((c * 2 + 1) * 4 + g) * 1024);
                }
#pragma unroll
                for (int c = 0; c < NC; ++c)
#pragma unroll
                    for (int rb = 0; rb < 2; ++rb) {
                        acc[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(AH[rb][c], BL[c], acc[rb][cb], 0, 0, 0);
                        acc[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(AL[rb][c], BH[c], acc[rb][cb], 0, 0, 0);
                    }
#pragma unroll
                for (int c = 0; c < NC; ++c)
#pragma unroll
                    for (int rb = 0; rb < 2; ++rb)
                        acc[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(AH[rb][c], BH[c], acc[rb][cb], 0, 0, 0);
            } else {
#pragma unroll
                for (int s = 0; s < DP4; ++s) {
                    const float b = bt[s * 512 + cb * 64 + c16 * 4 + (g ^ ((c16 >> 2) & 2))];
                    acc[0][cb] = __builtin_amdgcn_mfma_f32_16x16x4f32(A[0][s], b, acc[0][cb], 0, 0, 0);
                    acc[1][cb] = __builtin_amdgcn_mfma_f32_16x16x4f32(A[1][s], b, acc[1][cb], 0, 0, 0);
                }
            }
            const int cl = cb * 16 + c16;
            const int64_t col = col0 + cl;
            const bool valid = col < n;
            const float cn = col_nrm[cur][cl];
            if constexpr (MODE == 2) {
#pragma unroll
                for (int rb = 0; rb < 2; ++rb)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const float approx = __builtin_fmaf(c2, acc[rb][cb][r], rn[rb][r] + cn);
                        if (valid && approx < t4[rb][r][3]) {
                            float v = approx;
                            float *t = t4[rb][r];
                            // insert into the sorted 4-list
                            const float a3 = t[2] > v ? t[2] : v;
                            v = t[2] > v ? v : t[2];
                            const float a2 = t[1] > v ? t[1] : v;
                            v = t[1] > v ? v : t[1];
                            const float a1 = t[0] > v ? t[0] : v;
                            v = t[0] > v ? v : t[0];
                            t[3] = a3;
                            t[2] = a2;
                            t[1] = a1;
                            t[0] = v;
                        }
                    }
            } else if constexpr (MODE == 0) {
                // hits are rare: the compares first, the appends only when
                // some lane of the wave has one (the exact conditions are
                // checked again there)
                const bool other = t != (int)blockIdx.x;
                const float ct = (sym && other) ? col_thr[cur][cl] : -kInfB;
                float ap[2][4];
                bool hit = false;
#pragma unroll
                for (int rb = 0; rb < 2; ++rb)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        ap[rb][r] = __builtin_fmaf(c2, acc[rb][cb][r], rn[rb][r] + cn);
                        hit = hit || ap[rb][r] <= rv[rb][r] || ap[rb][r] <= ct;
                    }
                if (__builtin_amdgcn_ballot_w64(hit) != 0) {
#pragma unroll
                    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
                        for (int r = 0; r < 4; ++r) {
                            const float approx = ap[rb][r];
                            const int rl = wave * 32 + rb * 16 + 4 * g + r;
                            if (valid && approx <= rv[rb][r]) {
                                if (sym) {
                                    const int slot = atomicAdd(a.cand_cnt + blk_row + rl, 1);
                                    if (slot < a.cap) a.cand[(blk_row + rl) * a.cap + slot] = (int)col;
                                } else {
                                    const int slot = atomicAdd(&cnt_l[rl], 1);
                                    if (slot < a.cap) a.cand[(blk_row + rl) * a.cap + slot] = (int)col;
                                }
                            }
                            // the same pair for the column's row (tiles other
                            // than the block's own; the own tile gives both
                            // orders on the row side)
                            if (sym && other && valid && rcomp[rb][r] != -2 && approx <= ct) {
                                const int slot = atomicAdd(a.cand_cnt + col, 1);
                                if (slot < a.cap) a.cand[col * a.cap + slot] = (int)(blk_row + rl);
                            }
                        }
                }
            } else {
                // padding rows and columns give +inf values (their norms),
                // pairs inside one component +inf by selection.  A value
                // matters only below the row's best so far or the
                // column's (symmetric: seeded from its packed key), so the
                // minima are updated only when some lane of the wave has
                // such a value; then by compare and select, first index
                // kept on equal values
                const float cc = col_core[cur][cl];
                const int ck = col_comp[cur][cl];
                const float cv0 = col_thr[cur][cl];
                float mm[2][4];
                bool hit = false;
#pragma unroll
                for (int rb = 0; rb < 2; ++rb)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const float approx = __builtin_fmaf(c2, acc[rb][cb][r], rn[rb][r] + cn);
                        const int rl = wave * 32 + rb * 16 + 4 * g + r;
                        // no value here is NaN; the minimum only bounds
                        // values, so fminf (one instruction) serves
                        tmv = __builtin_fminf(approx, tmv);
                        // max(core_i, core_j, approx) on the signed bit
                        // patterns (one instruction): the cores are +0 or
                        // positive, so a negative or -0 approx loses and the
                        // result is never -0, whose sign bit would misorder
                        // the packed (value bits, index) keys
                        float m = __int_as_float(max(max(__float_as_int(row_core_s[rl]), __float_as_int(cc)),
                                                     __float_as_int(approx)));
                        m = ck != row_comp_s[rl] ? m : kInfB;
                        mm[rb][r] = m;
                        hit = hit || m < __builtin_fmaxf(rv[rb][r], cv0);
                    }
                float cv = cv0;  // this column's minimum over the lane's rows (below cv0)
                int ca = -1;
                if (__builtin_amdgcn_ballot_w64(hit) != 0) {
#pragma unroll
                    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
                        for (int r = 0; r < 4; ++r) {
                            const float m = mm[rb][r];
                            const bool lt = m < rv[rb][r];
                            rv[rb][r] = lt ? m : rv[rb][r];
                            ri[rb][r] = lt ? (int)col : ri[rb][r];
                            // rows in increasing order: the first equal one is kept
                            const bool lc = m < cv;
                            cv = lc ? m : cv;
                            ca = lc ? (int)(rbase + rb * 16 + 4 * g + r) : ca;
                        }
                }
                if (sym) {
                    if (__builtin_amdgcn_ballot_w64(ca >= 0) != 0) {
                        // the 4 lanes holding this column's other rows of the wave
#pragma unroll
                        for (int o = 16; o < 64; o <<= 1) {
                            const float ov = __shfl_xor(cv, o, 64);
                            const int oa = __shfl_xor(ca, o, 64);
                            if (oa >= 0 && (ca < 0 || ov < cv || (ov == cv && (unsigned)oa < (unsigned)ca))) {
                                cv = ov;
                                ca = oa;
                            }
                        }
                    }
                    if (g == 0) {
                        cmin_s[cur][wave][cl] = cv;
                        carg_s[cur][wave][cl] = ca;
                    }
                }
            }
        }
        if (MODE == 1 && a.tmin_out) {
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) {
                const float ov = __shfl_xor(tmv, o, 64);
                tmv = ov < tmv ? ov : tmv;
            }
            if (lane == 0) tmin_s[cur][wave] = tmv;
        }
        // tile t1 into the other buffer (last read in the previous
        // iteration, before the previous barrier), prefetch the one after,
        // then one barrier
        const int t2 = t1 < ntiles ? next_tile(t1 + 1) : ntiles;
        if constexpr (NBUF == 1) __syncthreads();  // every wave is done with the tile
        if (t1 < ntiles) {
            lstore(NBUF == 2 ? cur ^ 1 : 0);
            if (t2 < ntiles) gload(t2);
        }
        __syncthreads();
        if (MODE == 1 && a.tmin_out && tid == 0) {
            float v = tmin_s[cur][0];
#pragma unroll
            for (int w = 1; w < 4; ++w) v = tmin_s[cur][w] < v ? tmin_s[cur][w] : v;
            a.tmin_out[(int64_t)blockIdx.x * ntiles + t] = v;
        }
        if (sym && t != (int)blockIdx.x && tid < 128) {
            // tile t's column minima over this block's 128 rows, to the
            // column rows (the pair (t, this tile) is scanned here only);
            // buffer cur is written again only after the next barrier
            float v = cmin_s[cur][0][tid];
            int i = carg_s[cur][0][tid];
#pragma unroll
            for (int w = 1; w < 4; ++w) {
                const float ov = cmin_s[cur][w][tid];
                const int oi = carg_s[cur][w][tid];
                if (ov < v || (ov == v && (unsigned)oi < (unsigned)i)) {
                    v = ov;
                    i = oi;
                }
            }
            const int64_t col = (int64_t)t * 128 + tid;
            if (col < n && i >= 0)
                atomicMin(a.row_key + col, ((unsigned long long)__float_as_uint(v) << 32) | (unsigned int)i);
        }
        t = t1;
        t1 = t2;
    }
    if (MODE == 1 && a.scored_tiles && tid == 0) atomicAdd(a.scored_tiles, (unsigned long long)scored);

    if constexpr (MODE == 0) {
        __syncthreads();
        if (!sym && tid < 128 && blk_row + tid < n) a.cand_cnt[blk_row + tid] = cnt_l[tid];
    } else if constexpr (MODE == 2) {
        // gather the 16 lanes' lists per row, then one thread per row picks
        // the k-th smallest of its 64 values
        __syncthreads();
        float *un = bs;  // [128 rows][64]
#pragma unroll
        for (int rb = 0; rb < 2; ++rb)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int rl = wave * 32 + rb * 16 + 4 * g + r;
#pragma unroll
                for (int q = 0; q < 4; ++q) un[rl * 64 + c16 * 4 + q] = t4[rb][r][q];
            }
        __syncthreads();
        if (tid < 128 && blk_row + tid < n) {
            const int64_t row = blk_row + tid;
            const float *u = un + tid * 64;
            // k-th smallest (1-based k) by counting: the value v with
            // #(< v) < k <= #(<= v)
            float kth = kInfB;
            for (int i = 0; i < 64; ++i) {
                const float v = u[i];
                int lt = 0, le = 0;
                for (int j = 0; j < 64; ++j) {
                    lt += u[j] < v;
                    le += u[j] <= v;
                }
                if (lt < a.cap && a.cap <= le) kth = v;
            }
            const float nm = __uint_as_float(*a.nrm_max_bits);
            a.row_min[row] = kth + 2.0f * a.margin_c * kU * (a.nrm[row] + nm);
        }
    } else {
        // minimum over the 16 lanes that hold the same rows
#pragma unroll
        for (int rb = 0; rb < 2; ++rb)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                float v = rv[rb][r];
                int i = ri[rb][r];
#pragma unroll
                for (int m = 1; m < 16; m <<= 1) {
                    const float ov = __shfl_xor(v, m, 16);
                    const int oi = __shfl_xor(i, m, 16);
                    if (ov < v || (ov == v && (unsigned)oi < (unsigned)i)) {
                        v = ov;
                        i = oi;
                    }
                }
                const int64_t row = rbase + rb * 16 + 4 * g + r;
                if (c16 == 0 && row < n) {
                    if (sym) {
                        if (i >= 0)
                            atomicMin(a.row_key + row, ((unsigned long long)__float_as_uint(v) << 32) | (unsigned int)i);
                    } else {
                        a.row_min[row] = v;
                        a.row_arg[row] = i;
                    }
                }
            }
    }
}

// ------------------------------------------------- exact kNN re-score ----

// Exact rdist of every collected candidate, the k+1 smallest kept sorted in
// LDS; rows with a tie among them, an overflowed list or fewer than k
// candidates go to the tree search.
// A wavefront per row (four per work-group): the lanes compute the
// candidates' exact values in parallel (into LDS, and into cand_d for the
// first Boruvka round), then lane 0 inserts them in list order into the
// sorted k + 1 smallest, as one thread per row did.
constexpr int kRescoreCap = 512;  // candidates per row held in LDS

__global__ __launch_bounds__(256) void knn_rescore_kernel(KnnRescoreArgs a) {
    __shared__ float cd[4][kRescoreCap];
    __shared__ float sv[4][65];
    __shared__ int si[4][65];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t p = (int64_t)blockIdx.x * 4 + w;
    const bool live = p < a.n;
    const int K1 = a.k + 1;
    const int cnt = live ? a.cand_cnt[p] : 0;
    bool flag = cnt > a.cap;
    const int m = flag ? a.cap : cnt;  // the kept candidates (all of them, overflowed or not)
    if (live) {
        const float *x = a.xp + p * a.d;
        for (int t = lane; t < m; t += 64) {
            const int j = a.cand[p * a.cap + t];
            const float r = rdist<0>(x, a.xp + (int64_t)j * a.d, a.d);
            cd[w][t] = r;
            if (a.cand_d) a.cand_d[p * a.cap + t] = r;
        }
    }
    __syncthreads();
    if (!live || lane != 0) return;
    float *v = sv[w];
    int *id = si[w];
    for (int t = 0; t < K1; ++t) {
        v[t] = kInfB;
        id[t] = -1;
    }
    if (!flag) {
        for (int t = 0; t < cnt; ++t) {
            const float r = cd[w][t];
            if (r < v[K1 - 1]) {
                int i = K1 - 1;
                while (i > 0 && v[i - 1] > r) {
                    v[i] = v[i - 1];
                    id[i] = id[i - 1];
                    --i;
                }
                v[i] = r;
                id[i] = a.cand[p * a.cap + t];
            }
        }
        for (int t = 0; t < a.k; ++t)
            if (!(v[t] < v[t + 1])) flag = true;
        // every column whose exact value is at most the k-th smallest found
        // has an approximate value at most that plus the margin; below the
        // threshold it was collected, so the k + 1 smallest found are the
        // k + 1 smallest of all columns (ties at the k-th included)
        const float marg = a.margin_c * kU * (a.nrm[p] + __uint_as_float(*a.nrm_max_bits));
        if (!(v[a.k - 1] + marg <= a.thr[p])) flag = true;
    }
    if (flag) {
        const int slot = atomicAdd(a.flag_count, 1);
        a.flag_list[slot] = (int)p;
        return;
    }
    const int64_t row = a.perm[p];
    for (int t = 0; t < a.k; ++t) {
        a.out_d[row * a.k + t] = v[t];
        a.out_i[row * a.k + t] = a.perm[id[t]];
    }
    const float c = a.core_from_knn ? v[a.k - 1] : 0.0f;
    a.core_o[row] = c;
    a.core_p[p] = c;
}

// ---------------------------------------------- tree kNN, one wave/row ----

// The reference's kNN query for the rows the MFMA path hands back (ties):
// the same depth-first walk, pruning and heap pushes as knn_kernel, with a
// whole wave per row.  The walk is uniform across the wave; a leaf's rdist
// values are computed one point per lane and pushed by lane 0 in leaf
// order, child bounds by lanes 0 and 1.  In high dimension the walk visits
// almost every leaf, so this is what makes those rows affordable.
__global__ __launch_bounds__(64) void knn_wave_kernel(KnnWaveArgs a) {
    __shared__ float xs[64];
    __shared__ float hp[64];
    __shared__ int hi[64];
    __shared__ float pv[64];
    const int lane = threadIdx.x;
    const int pos = a.query_pos[blockIdx.x];
    const int d = a.d, k = a.k;
    if (lane < d) xs[lane] = a.xp[(int64_t)pos * d + lane];
    if (lane < k) {
        hp[lane] = kInfB;
        hi[lane] = -1;
    }
    __syncthreads();
    LdsHeap h{hp, hi, 1, k};
    const size_t stride = (size_t)2 * d;
    float v0 = 0.0f;
    if (lane == 0) v0 = box_lb<0>(xs, a.box, d);
    float lb = __shfl(v0, 0);
    int node = 0, depth = 0;
    uint32_t first_right = 0;
    bool entering = true;
    for (;;) {
        if (entering && !(lb > hp[0])) {
            if (depth == a.levels - 1) {
                const int2 r = a.range[node];
                for (int c0 = r.x; c0 < r.y; c0 += 64) {
                    const int i = c0 + lane;
                    if (i < r.y) pv[lane] = rdist<0>(xs, a.xp + (int64_t)i * d, d);
                    __syncthreads();
                    if (lane == 0) {
                        const int cnt = r.y - c0 < 64 ? r.y - c0 : 64;
                        for (int t = 0; t < cnt; ++t) {
                            const float dist = pv[t];
                            if (dist < hp[0]) h.push(dist, c0 + t);
                        }
                    }
                    __syncthreads();
                }
            } else {
                const int l = 2 * node + 1;
                float v = 0.0f;
                if (lane < 2) v = box_lb<0>(xs, a.box + (size_t)(l + lane) * stride, d);
                const float lbl = __shfl(v, 0), lbr = __shfl(v, 1);
                const bool rf = !(lbl <= lbr);
                ++depth;
                first_right = rf ? (first_right | (1u << depth)) : (first_right & ~(1u << depth));
                node = rf ? l + 1 : l;
                lb = rf ? lbr : lbl;
                continue;
            }
        }
        if (depth == 0) break;
        const bool rf = (first_right >> depth) & 1u;
        const bool is_right = (node & 1) == 0;
        if (is_right == rf) {
            node = is_right ? node - 1 : node + 1;
            float v = 0.0f;
            if (lane == 0) v = box_lb<0>(xs, a.box + (size_t)node * stride, d);
            lb = __shfl(v, 0);
            entering = true;
        } else {
            node = (node - 1) >> 1;
            --depth;
            entering = false;
        }
    }
    if (lane == 0) {
        h.sort();
        const int64_t row = a.perm[pos];
        for (int j = 0; j < k; ++j) {
            a.out_d[row * k + j] = hp[j];
            a.out_i[row * k + j] = hi[j] >= 0 ? a.perm[hi[j]] : -1;
        }
        const float c = a.core_from_knn ? hp[k - 1] : 0.0f;
        a.core_o[row] = c;
        a.core_p[pos] = c;
    }
}

// The same query, one work-group per row.  Pruning never changes which
// points the reference's heap receives: every point of a pruned subtree has
// rdist >= box_lb > top (both are FMA chains of the same shape over
// per-dimension gaps, the box's no larger, and rounding is monotone), and
// the top only falls.  So the heap receives, in the query's depth-first
// leaf order (closer child first, left on equal bounds) and in position
// order within a leaf, every point whose value is below its top at that
// moment.  The work-group first sets every internal node's child order (a
// bit in LDS), then takes the leaves in that order, 256 / leaf_cap at a
// time (leaf sizes differ by at most one, so leaf_cap = ceil(n / leaves)
// holds any leaf): a thread per point computes rdist, and one wavefront
// pushes the values in order (a ballot against the top, then lane 0
// re-checks and pushes each candidate in turn).
constexpr int kKnnBlockLevels = 19;  // child-order bits for up to 2^18 - 1 internal nodes

__global__ __launch_bounds__(256) void knn_block_kernel(KnnWaveArgs a, int leaf_cap) {
    __shared__ float xs[64];
    __shared__ float hp[64];
    __shared__ int hi[64];
    __shared__ float pv[256];
    __shared__ int pi[256];
    __shared__ unsigned int rfb[(1 << (kKnnBlockLevels - 1)) / 32];
    const int tid = threadIdx.x, lane = tid & 63;
    const int pos = a.query_pos[blockIdx.x];
    const int d = a.d, k = a.k, L = a.levels;
    if (tid < 64) xs[tid] = tid < d ? a.xp[(int64_t)pos * d + tid] : 0.0f;
    if (tid < k) {
        hp[tid] = kInfB;
        hi[tid] = -1;
    }
    const int ninternal = (1 << (L - 1)) - 1;
    for (int w = tid; w < (ninternal + 31) / 32; w += 256) rfb[w] = 0u;
    __syncthreads();
    const size_t stride = (size_t)2 * d;
    for (int node = tid; node < ninternal; node += 256) {
        const int l = 2 * node + 1;
        const float lbl = box_lb<0>(xs, a.box + (size_t)l * stride, d);
        const float lbr = box_lb<0>(xs, a.box + (size_t)(l + 1) * stride, d);
        if (!(lbl <= lbr)) atomicOr(&rfb[node >> 5], 1u << (node & 31));
    }
    __syncthreads();
    const int nleaves = 1 << (L - 1);
    const int lps = 256 / leaf_cap;  // leaves per step
    const int li = tid / leaf_cap, lo = tid % leaf_cap;
    LdsHeap h{hp, hi, 1, k};
    for (int r0 = 0; r0 < nleaves; r0 += lps) {
        // this thread's point: offset lo of the leaf of depth-first rank r0 + li
        float dist = kInfB;
        int pidx = -1;
        const int r = r0 + li;
        if (li < lps && r < nleaves) {
            int node = 0;
            for (int t = L - 2; t >= 0; --t) {
                const int rf = (rfb[node >> 5] >> (node & 31)) & 1;
                node = 2 * node + 1 + (((r >> t) & 1) ^ rf);
            }
            const int2 rg = a.range[node];
            if (rg.x + lo < rg.y) {
                pidx = rg.x + lo;
                dist = rdist<0>(xs, a.xp + (int64_t)pidx * d, d);
            }
        }
        pv[tid] = dist;
        pi[tid] = pidx;
        __syncthreads();
        if (tid < 64) {
            for (int c = 0; c < 256; c += 64) {
                uint64_t ball = __builtin_amdgcn_ballot_w64(pi[c + lane] >= 0 && pv[c + lane] < hp[0]);
                while (ball) {
                    const int l = __builtin_ctzll(ball);
                    ball &= ball - 1;
                    if (lane == 0) {
                        const float dv = pv[c + l];
                        if (dv < hp[0]) h.push(dv, pi[c + l]);
                    }
                }
            }
        }
        __syncthreads();
    }
    if (tid == 0) {
        h.sort();
        const int64_t row = a.perm[pos];
        for (int j = 0; j < k; ++j) {
            a.out_d[row * k + j] = hp[j];
            a.out_i[row * k + j] = hi[j] >= 0 ? a.perm[hi[j]] : -1;
        }
        const float c = a.core_from_knn ? hp[k - 1] : 0.0f;
        a.core_o[row] = c;
        a.core_p[pos] = c;
    }
}

// ------------------------------------------------ exact Boruvka steps ----

__global__ void boruvka_upper_kernel(BoruvkaExactArgs a) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.n) return;
    const int j = a.row_arg[p];
    if (j < 0) return;
    float w = rdist<0>(a.xp + p * a.d, a.xp + (int64_t)j * a.d, a.d);
    const float ci = a.core_p[p], cj = a.core_p[j];
    w = w > ci ? w : ci;
    w = w > cj ? w : cj;
    atomicMin(reinterpret_cast<unsigned int *>(a.bound + a.comp_p[p]), __float_as_uint(w));
}

// The same bound without the exact weights: a row's approximate value
// m_i = max(approx, core_i, core_j) is within the margin of the exact
// weight of (i, j0), so m_i + margin bounds it from above (the margin's
// factor two covers this sum's rounding).  No gather of j0's row.
// One atomic per run of equal components within a wavefront (rows are in
// tree order, where later rounds' components are long runs): a million
// atomics onto a few hundred slots serialise in L2.
__global__ void boruvka_upper_approx_kernel(BoruvkaExactArgs a) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    int key = -1;
    float u = kInfB;
    if (p < a.n && a.row_arg[p] >= 0) {
        key = a.comp_p[p];
        u = a.row_min[p] + a.margin_c * kU * (a.nrm[p] + __uint_as_float(*a.nrm_max_bits));
    }
    // segmented inclusive minimum over lanes of equal key: the last lane of
    // each run then holds its run's minimum
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const float ou = __shfl_up(u, o, 64);
        const int ok = __shfl_up(key, o, 64);
        if (lane >= o && ok == key) u = ou < u ? ou : u;
    }
    const int nk = __shfl_down(key, 1, 64);
    if (key >= 0 && (lane == 63 || nk != key))
        atomicMin(reinterpret_cast<unsigned int *>(a.bound + key), __float_as_uint(u));
}

__global__ void boruvka_rows_kernel(BoruvkaExactArgs a) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.n) return;
    const float nm = __uint_as_float(*a.nrm_max_bits);
    const float lo = a.row_min[p] - a.margin_c * kU * (a.nrm[p] + nm);
    if (lo <= a.bound[a.comp_p[p]]) {
        const int slot = atomicAdd(a.row_count, 1);
        a.rows[slot] = (int)p;
    }
}

// true when position j1 comes before j2 in the depth-first visiting order
// of the KD-tree query for point x (closer child first, left on equal
// bounds; within a leaf, position order)
__device__ bool dfs_before(const float *x, int j1, int j2, const int2 *__restrict__ range,
                           const float *__restrict__ box, int levels, int d) {
    int node = 0;
    for (int depth = 0; depth < levels - 1; ++depth) {
        const int l = 2 * node + 1;
        const int split = range[l].y;
        const bool a1 = j1 < split, a2 = j2 < split;
        if (a1 == a2) {
            node = a1 ? l : l + 1;
            continue;
        }
        const float lbl = box_lb<0>(x, box + (size_t)l * 2 * d, d);
        const float lbr = box_lb<0>(x, box + (size_t)(l + 1) * 2 * d, d);
        const bool left_first = lbl <= lbr;
        return a1 == left_first;
    }
    return j1 < j2;
}

__global__ __launch_bounds__(256) void boruvka_row_scan_kernel(BoruvkaExactArgs a) {
    __shared__ float xs[64];
    __shared__ float sw[256];
    __shared__ int sj[256];
    const int tid = threadIdx.x;
    const int count = *a.row_count;
    for (int it = blockIdx.x; it < count; it += gridDim.x) {
        const int p = a.rows[it];
        const int comp = a.comp_p[p];
        const float ci = a.core_p[p];
        const float B = a.bound[comp];
        if (tid < a.d) xs[tid] = a.xp[(int64_t)p * a.d + tid];
        __syncthreads();
        float bw = kInfB;
        int bj = -1;
        for (int64_t j = tid; j < a.n; j += 256) {
            if (a.comp_p[j] == comp) continue;
            const float cj = a.core_p[j];
            if (cj > B) continue;
            float w = rdist<0>(xs, a.xp + j * a.d, a.d);
            w = w > ci ? w : ci;
            w = w > cj ? w : cj;
            if (w > B) continue;
            if (w < bw || (w == bw && dfs_before(xs, (int)j, bj, a.range, a.box, a.levels, a.d))) {
                bw = w;
                bj = (int)j;
            }
        }
        sw[tid] = bw;
        sj[tid] = bj;
        __syncthreads();
        for (int off = 128; off > 0; off >>= 1) {
            if (tid < off) {
                const float ow = sw[tid + off];
                const int oj = sj[tid + off];
                const float mw = sw[tid];
                const int mj = sj[tid];
                bool take = false;
                if (oj >= 0) {
                    if (mj < 0 || ow < mw) take = true;
                    else if (ow == mw) take = dfs_before(xs, oj, mj, a.range, a.box, a.levels, a.d);
                }
                if (take) {
                    sw[tid] = ow;
                    sj[tid] = oj;
                }
            }
            __syncthreads();
        }
        if (tid == 0 && sj[0] >= 0) {
            const unsigned long long key =
                ((unsigned long long)__float_as_uint(sw[0]) << 32) | (unsigned int)a.perm[p];
            atomicMin(a.best + comp, key);
            a.partner[p] = sj[0];
        }
        __syncthreads();
    }
}

// The same re-scan, tiled: a work-group takes 64 listed rows (coordinates
// in LDS) against 2,048 positions (8 x 256, each thread's coordinates in
// registers, zero-padded to DP: the padded terms add +0 to rdist's sum), so
// a position is read once per 64 rows instead of once per row.  Per row the
// work-group's best (weight, position) under the same conditions and tie
// order goes to slot [row][blockIdx.x]; boruvka_slots_kernel reduces the
// slots.
template <int DP>
__global__ __launch_bounds__(256) void boruvka_rows_tiled_kernel(BoruvkaExactArgs a, int row0, int nrows, int ncb,
                                                                 float *slot_w, int *slot_j) {
    __shared__ float rs[64 * DP];
    __shared__ float rci[64], rB[64];
    __shared__ int rcomp[64];
    __shared__ float bw_s[4][64];
    __shared__ int bj_s[4][64];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int g0 = blockIdx.y * 64;
    if (g0 >= nrows) return;  // uniform over the work-group
    const int nr = nrows - g0 < 64 ? nrows - g0 : 64;
    const int d = a.d;
    for (int i = tid; i < 64 * DP; i += 256) {
        const int r = i / DP, k = i % DP;
        float v = 0.0f;
        if (r < nr && k < d) v = a.xp[(int64_t)a.rows[row0 + g0 + r] * d + k];
        rs[i] = v;
    }
    if (tid < 64) {
        if (tid < nr) {
            const int p = a.rows[row0 + g0 + tid];
            const int c = a.comp_p[p];
            rcomp[tid] = c;
            rci[tid] = a.core_p[p];
            rB[tid] = a.bound[c];
        }
#pragma unroll
        for (int w = 0; w < 4; ++w) {
            bw_s[w][tid] = kInfB;
            bj_s[w][tid] = -1;
        }
    }
    __syncthreads();
    const int64_t cbase = (int64_t)blockIdx.x * 2048;
    for (int sc = 0; sc < 8; ++sc) {
        const int64_t j = cbase + sc * 256 + tid;
        const bool ok = j < a.n;
        float y[DP];
#pragma unroll
        for (int k = 0; k < DP; ++k) y[k] = (ok && k < d) ? a.xp[j * d + k] : 0.0f;
        const float cj = ok ? a.core_p[j] : kInfB;
        const int kj = ok ? a.comp_p[j] : -1;
        for (int r = 0; r < nr; ++r) {
            float w = kInfB;
            const float B = rB[r];
            if (ok && kj != rcomp[r] && !(cj > B)) {
                const float *x = rs + r * DP;
                float acc = 0.0f;
#pragma unroll
                for (int k = 0; k < DP; ++k) {
                    const float df = x[k] - y[k];
                    acc = __builtin_fmaf(df, df, acc);
                }
                const float ci = rci[r];
                acc = acc > ci ? acc : ci;
                acc = acc > cj ? acc : cj;
                if (!(acc > B)) w = acc;
            }
            float wm = w;
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) {
                const float ow = __shfl_xor(wm, o, 64);
                wm = ow < wm ? ow : wm;
            }
            if (!(wm < kInfB)) continue;  // uniform: no candidate in this wave
            uint64_t ball = __builtin_amdgcn_ballot_w64(w == wm);
            int jb = __shfl((int)j, __builtin_ctzll(ball), 64);
            if (lane == 0) {
                // equal weights: the first in the row's depth-first order
                ball &= ball - 1;
                while (ball) {
                    const int l = __builtin_ctzll(ball);
                    ball &= ball - 1;
                    const int jl = (int)(cbase + sc * 256 + wave * 64 + l);
                    if (dfs_before(rs + r * DP, jl, jb, a.range, a.box, a.levels, d)) jb = jl;
                }
                const float ow = bw_s[wave][r];
                const int oj = bj_s[wave][r];
                if (wm < ow || oj < 0 || (wm == ow && dfs_before(rs + r * DP, jb, oj, a.range, a.box, a.levels, d))) {
                    bw_s[wave][r] = wm;
                    bj_s[wave][r] = jb;
                }
            }
        }
    }
    __syncthreads();
    if (tid < nr) {
        float bw = bw_s[0][tid];
        int bj = bj_s[0][tid];
#pragma unroll
        for (int w = 1; w < 4; ++w) {
            const float ow = bw_s[w][tid];
            const int oj = bj_s[w][tid];
            if (oj < 0) continue;
            if (bj < 0 || ow < bw || (ow == bw && dfs_before(rs + tid * DP, oj, bj, a.range, a.box, a.levels, d))) {
                bw = ow;
                bj = oj;
            }
        }
        slot_w[(int64_t)(g0 + tid) * ncb + blockIdx.x] = bw;
        slot_j[(int64_t)(g0 + tid) * ncb + blockIdx.x] = bj;
    }
}

// First Boruvka round from the kNN candidate lists.  Row p's list holds
// every column whose approximate distance is at most thr[p] (unless it
// overflowed), so every other column j has exact rdist(p, j) > L_p = thr[p]
// - margin_p, and so does its mutual-reachability weight.  A wavefront per
// row takes the exact weight of each candidate in another component (one
// lane per candidate; the same operations as the exact re-scans) and keeps
// the smallest, equal weights in the row's depth-first order.  When that
// weight w is at most L_p, no other column reaches it: (w, partner) is the
// row's exact answer over all columns, and the row enters its component's
// packed minimum and writes its partner, as the exact re-scan would.  Any
// candidate weight bounds its component's minimum from above (bound).
// cand_round_list_kernel then lists the rows not settled so whose lower
// bound L_p is at most their component's bound, and every row whose list
// overflowed (it has no such bound), for the exact re-scan; the other
// rows cannot reach the component's minimum.
__global__ __launch_bounds__(256) void cand_round_kernel(CandRoundArgs a) {
    const int lane = threadIdx.x & 63;
    const int64_t p = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (p >= a.n) return;  // uniform over the wavefront
    const int comp = a.comp_p[p];
    const float ci = a.core_p[p];
    const float *x = a.xp + p * a.d;
    const int cnt = a.cand_cnt[p];
    const int m = cnt < a.cap ? cnt : a.cap;
    float bw = kInfB;
    int bj = -1;
    for (int t0 = 0; t0 < m; t0 += 64) {
        const int t = t0 + lane;
        float w = kInfB;
        int j = -1;
        if (t < m) {
            j = a.cand[p * a.cap + t];
            if (a.comp_p[j] != comp) {
                float r = a.cand_d ? a.cand_d[p * a.cap + t] : rdist<0>(x, a.xp + (int64_t)j * a.d, a.d);
                const float cj = a.core_p[j];
                r = r > ci ? r : ci;
                w = r > cj ? r : cj;
            } else {
                j = -1;
            }
        }
        float wm = w;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            const float ow = __shfl_xor(wm, o, 64);
            wm = ow < wm ? ow : wm;
        }
        if (!(wm < kInfB) || wm > bw) continue;  // uniform
        uint64_t ball = __builtin_amdgcn_ballot_w64(j >= 0 && w == wm);
        int jb = __shfl(j, __builtin_ctzll(ball), 64);
        if (lane == 0) {
            ball &= ball - 1;
            while (ball) {
                const int l = __builtin_ctzll(ball);
                ball &= ball - 1;
                const int jl = a.cand[p * a.cap + t0 + l];
                if (dfs_before(x, jl, jb, a.range, a.box, a.levels, a.d)) jb = jl;
            }
            if (wm < bw || bj < 0 || dfs_before(x, jb, bj, a.range, a.box, a.levels, a.d)) {
                bw = wm;
                bj = jb;
            }
        }
        bw = __shfl(bw, 0, 64);
        bj = __shfl(bj, 0, 64);
    }
    if (lane == 0) {
        const float marg = a.margin_c * kU * (a.nrm[p] + __uint_as_float(*a.nrm_max_bits));
        const float L = a.thr[p] - marg;
        const bool settled = cnt <= a.cap && bj >= 0 && bw <= L;
        a.cert[p] = settled ? 1 : 0;
        if (bj >= 0) atomicMin(reinterpret_cast<unsigned int *>(a.bound + comp), __float_as_uint(bw));
        if (settled) {
            atomicMin(a.best + comp, ((unsigned long long)__float_as_uint(bw) << 32) | (unsigned int)a.perm[p]);
            a.partner[p] = bj;
        }
    }
}

__global__ void cand_round_list_kernel(CandRoundArgs a) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.n || a.cert[p]) return;
    const float marg = a.margin_c * kU * (a.nrm[p] + __uint_as_float(*a.nrm_max_bits));
    const float L = a.thr[p] - marg;
    // an overflowed list lost columns under the threshold: no lower bound
    if (a.cand_cnt[p] > a.cap || L <= a.bound[a.comp_p[p]]) {
        const int slot = atomicAdd(a.row_count, 1);
        a.rows[slot] = (int)p;
    }
}

// Per listed row: the best of its slots (depth-first order on equal
// weights), then the component's packed minimum and the row's partner, as
// boruvka_row_scan_kernel writes them.
__global__ void boruvka_slots_kernel(BoruvkaExactArgs a, int row0, int nrows, int ncb, const float *slot_w,
                                     const int *slot_j) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nrows) return;
    const int p = a.rows[row0 + i];
    const float *x = a.xp + (int64_t)p * a.d;
    float bw = kInfB;
    int bj = -1;
    for (int b = 0; b < ncb; ++b) {
        const int oj = slot_j[(int64_t)i * ncb + b];
        if (oj < 0) continue;
        const float ow = slot_w[(int64_t)i * ncb + b];
        if (bj < 0 || ow < bw || (ow == bw && dfs_before(x, oj, bj, a.range, a.box, a.levels, a.d))) {
            bw = ow;
            bj = oj;
        }
    }
    if (bj >= 0) {
        const unsigned long long key = ((unsigned long long)__float_as_uint(bw) << 32) | (unsigned int)a.perm[p];
        atomicMin(a.best + a.comp_p[p], key);
        a.partner[p] = bj;
    }
}

// ------------------------------------------------------------ launches ----

hipError_t launch_center_pack(int64_t n, int d, int dp, int64_t npad, const float *xp, const float *root_box,
                              float *xc, float *nrm, unsigned int *nrm_max_bits, float scale, hipStream_t st) {
    if (scale > 0.0f && dp != 32 && dp != 64) return hipErrorInvalidValue;
    hipLaunchKernelGGL(center_pack_kernel, dim3(nblocks(npad, 256)), dim3(256), 0, st, n, d, dp, npad, xp, root_box,
                       xc, nrm, nrm_max_bits, scale);
    return hipGetLastError();
}

template <int MODE>
static hipError_t launch_scan(const BruteScanArgs &a, hipStream_t st) {
    const dim3 g(a.npad / 128), b(256);
    if (a.split) {
        switch (a.dp) {
            case 32: hipLaunchKernelGGL((brute_scan_kernel<8, MODE, true>), g, b, 0, st, a); break;
            case 64: hipLaunchKernelGGL((brute_scan_kernel<16, MODE, true>), g, b, 0, st, a); break;
            default: return hipErrorInvalidValue;
        }
        return hipGetLastError();
    }
    switch (a.dp) {
        case 16: hipLaunchKernelGGL((brute_scan_kernel<4, MODE, false>), g, b, 0, st, a); break;
        case 24: hipLaunchKernelGGL((brute_scan_kernel<6, MODE, false>), g, b, 0, st, a); break;
        case 32: hipLaunchKernelGGL((brute_scan_kernel<8, MODE, false>), g, b, 0, st, a); break;
        case 48: hipLaunchKernelGGL((brute_scan_kernel<12, MODE, false>), g, b, 0, st, a); break;
        case 64: hipLaunchKernelGGL((brute_scan_kernel<16, MODE, false>), g, b, 0, st, a); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_scan_knn(const BruteScanArgs &a, hipStream_t st) { return launch_scan<0>(a, st); }
hipError_t launch_scan_kth(const BruteScanArgs &a, hipStream_t st) { return launch_scan<2>(a, st); }
hipError_t launch_scan_boruvka(const BruteScanArgs &a, hipStream_t st) { return launch_scan<1>(a, st); }

__global__ void unpack_row_keys_kernel(int64_t n, const unsigned long long *__restrict__ row_key, float *row_min,
                                       int *row_arg) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const unsigned long long key = row_key[p];
    if (key == ~0ull) {
        row_min[p] = kInfB;
        row_arg[p] = -1;
    } else {
        row_min[p] = __uint_as_float((unsigned int)(key >> 32));
        row_arg[p] = (int)(unsigned int)(key & 0xffffffffull);
    }
}

hipError_t launch_unpack_row_keys(int64_t n, const unsigned long long *row_key, float *row_min, int *row_arg,
                                  hipStream_t st) {
    hipLaunchKernelGGL(unpack_row_keys_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, row_key,
                       row_min, row_arg);
    return hipGetLastError();
}

// Component of every 128-row tile whose rows (below n) share one, else -1.
__global__ __launch_bounds__(128) void tile_comp_kernel(int64_t n, const int *__restrict__ comp_p, int *tile_comp) {
    __shared__ int mixed;
    const int64_t row0 = (int64_t)blockIdx.x * 128;
    const int c0 = comp_p[row0];
    if (threadIdx.x == 0) mixed = 0;
    __syncthreads();
    const int64_t row = row0 + threadIdx.x;
    if (row < n && comp_p[row] != c0) mixed = 1;
    __syncthreads();
    if (threadIdx.x == 0) tile_comp[blockIdx.x] = mixed ? -1 : c0;
}

// ------------------------------------------ rows in component order ----

// center_pack_kernel for the rows in `order` (scan position p2 holds tree
// position order[p2]; padding rows p2 >= n are zero), plus their cores.
__global__ void order_pack_kernel(int64_t n, int d, int dp, int64_t npad, const int *__restrict__ order,
                                  const float *__restrict__ xp, const float *__restrict__ root_box,
                                  const float *__restrict__ core_p, float *__restrict__ xc, float *__restrict__ nrm,
                                  float *__restrict__ core2, float scale) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npad) return;
    const int64_t q = p < n ? order[p] : -1;
    float *tile = xc + (p / 128) * 128 * dp;
    const int r = (int)(p % 128);
    float s = 0.0f;
    for (int k = 0; k < dp; ++k) {
        float v = 0.0f;
        if (q >= 0 && k < d) v = xp[q * d + k] - 0.5f * (root_box[k] + root_box[d + k]);
        pack_store(tile, r, k, v, scale);
        s = __builtin_fmaf(v, v, s);
    }
    nrm[p] = s;
    core2[p] = q >= 0 ? core_p[q] : 0.0f;
}

__global__ void order_keys_kernel(int64_t n, const int *__restrict__ comp_p, uint32_t *keys, int *vals) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    keys[p] = (uint32_t)comp_p[p];
    vals[p] = (int)p;
}

__global__ void gather_comp_kernel(int64_t n, int64_t npad, const int *__restrict__ order,
                                   const int *__restrict__ comp_p, int *comp2) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npad) return;
    comp2[p] = p < n ? comp_p[order[p]] : -3;
}

__global__ __launch_bounds__(128) void tile_core_range_kernel(int64_t n, const float *__restrict__ core2,
                                                              float *tminc, float *tmaxc) {
    __shared__ float lo_s[2], hi_s[2];
    const int64_t row = (int64_t)blockIdx.x * 128 + threadIdx.x;
    float lo = row < n ? core2[row] : kInfB, hi = row < n ? core2[row] : 0.0f;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const float a = __shfl_xor(lo, o, 64), b = __shfl_xor(hi, o, 64);
        lo = a < lo ? a : lo;
        hi = b > hi ? b : hi;
    }
    if ((threadIdx.x & 63) == 0) {
        lo_s[threadIdx.x >> 6] = lo;
        hi_s[threadIdx.x >> 6] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        tminc[blockIdx.x] = lo_s[0] < lo_s[1] ? lo_s[0] : lo_s[1];
        tmaxc[blockIdx.x] = hi_s[0] > hi_s[1] ? hi_s[0] : hi_s[1];
    }
}

// One work-group per row tile A, over the tile pairs (A, B), B > A, whose
// two tiles lie in two different single components: the pair's bound goes
// to both components.  Row A's side is reduced in registers over the whole
// row (one atomic per tile); the column side is reduced per run of equal
// components within each wavefront (the tiles are in component order, so
// the runs are long) before its atomic -- one atomic per pair onto a few
// hundred component slots serialises in L2.
__global__ __launch_bounds__(256) void tile_bounds_kernel(int ntiles, const int *__restrict__ tile_comp,
                                                          const float *__restrict__ tmin,
                                                          const float *__restrict__ tmaxc, float margin_c,
                                                          const unsigned int *__restrict__ nrm_max_bits,
                                                          float *bound) {
    __shared__ float ua_s[4];
    const int A = blockIdx.x, tid = threadIdx.x, lane = tid & 63;
    const int ca = tile_comp[A];
    if (ca < 0) return;  // uniform over the work-group
    const float tmargin = 2.0f * margin_c * kU * __uint_as_float(*nrm_max_bits);
    const float ta = tmaxc[A];
    const float *row = tmin + (int64_t)A * ntiles;
    float ua = kInfB;
    for (int B0 = A + 1; B0 < ntiles; B0 += 256) {
        const int B = B0 + tid;
        int key = -1;
        float u = kInfB;
        if (B < ntiles) {
            const int cb = tile_comp[B];
            if (cb >= 0 && cb != ca) {
                u = row[B] + tmargin;
                u = u > ta ? u : ta;
                const float tb = tmaxc[B];
                u = u > tb ? u : tb;
                key = cb;
            }
        }
        ua = u < ua ? u : ua;
        // segmented inclusive min over lanes of equal key: the last lane of
        // each run then holds its run's minimum
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const float ov = __shfl_up(u, o, 64);
            const int ok = __shfl_up(key, o, 64);
            if (lane >= o && ok == key) u = ov < u ? ov : u;
        }
        const int nk = __shfl_down(key, 1, 64);
        if (key >= 0 && (lane == 63 || nk != key))
            atomicMin(reinterpret_cast<unsigned int *>(bound + key), __float_as_uint(u));
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const float ov = __shfl_xor(ua, o, 64);
        ua = ov < ua ? ov : ua;
    }
    if (lane == 0) ua_s[tid >> 6] = ua;
    __syncthreads();
    if (tid == 0) {
        float v = ua_s[0];
#pragma unroll
        for (int w = 1; w < 4; ++w) v = ua_s[w] < v ? ua_s[w] : v;
        if (v < kInfB) atomicMin(reinterpret_cast<unsigned int *>(bound + ca), __float_as_uint(v));
    }
}

__global__ __launch_bounds__(128) void tile_umax_kernel(int64_t n, const int *__restrict__ comp2,
                                                        const float *__restrict__ bound, float *tU) {
    __shared__ float hi_s[2];
    const int64_t row = (int64_t)blockIdx.x * 128 + threadIdx.x;
    float hi = row < n ? bound[comp2[row]] : 0.0f;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const float b = __shfl_xor(hi, o, 64);
        hi = b > hi ? b : hi;
    }
    if ((threadIdx.x & 63) == 0) hi_s[threadIdx.x >> 6] = hi;
    __syncthreads();
    if (threadIdx.x == 0) tU[blockIdx.x] = hi_s[0] > hi_s[1] ? hi_s[0] : hi_s[1];
}

__global__ void unpack_row_keys_ordered_kernel(int64_t n, const unsigned long long *__restrict__ row_key,
                                               const int *__restrict__ order, float *row_min, int *row_arg) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const unsigned long long key = row_key[p];
    const int q = order[p];
    if (key == ~0ull) {
        row_min[q] = kInfB;
        row_arg[q] = -1;
    } else {
        row_min[q] = __uint_as_float((unsigned int)(key >> 32));
        row_arg[q] = order[(int)(unsigned int)(key & 0xffffffffull)];
    }
}

hipError_t launch_order_pack(int64_t n, int d, int dp, int64_t npad, const int *order, const float *xp,
                             const float *root_box, const float *core_p, float *xc, float *nrm, float *core2,
                             float scale, hipStream_t st) {
    if (scale > 0.0f && dp != 32 && dp != 64) return hipErrorInvalidValue;
    hipLaunchKernelGGL(order_pack_kernel, dim3(nblocks(npad, 256)), dim3(256), 0, st, n, d, dp, npad, order, xp,
                       root_box, core_p, xc, nrm, core2, scale);
    return hipGetLastError();
}

hipError_t launch_order_keys(int64_t n, const int *comp_p, uint32_t *keys, int *vals, hipStream_t st) {
    hipLaunchKernelGGL(order_keys_kernel, dim3((unsigned)nblocks(n, 256)), dim3(256), 0, st, n, comp_p, keys, vals);
    return hipGetLastError();
}

hipError_t launch_gather_comp(int64_t n, int64_t npad, const int *order, const int *comp_p, int *comp2,
                              hipStream_t st) {
    hipLaunchKernelGGL(gather_comp_kernel, dim3(nblocks(npad, 256)), dim3(256), 0, st, n, npad, order, comp_p, comp2);
    return hipGetLastError();
}

hipError_t launch_tile_core_range(int64_t n, int64_t npad, const float *core2, float *tminc, float *tmaxc,
                                  hipStream_t st) {
    hipLaunchKernelGGL(tile_core_range_kernel, dim3((unsigned)(npad / 128)), dim3(128), 0, st, n, core2, tminc, tmaxc);
    return hipGetLastError();
}

hipError_t launch_tile_bounds(int ntiles, const int *tile_comp, const float *tmin, const float *tmaxc, float margin_c,
                              const unsigned int *nrm_max_bits, float *bound, hipStream_t st) {
    if (ntiles <= 0) return hipSuccess;
    hipLaunchKernelGGL(tile_bounds_kernel, dim3((unsigned)ntiles), dim3(256), 0, st, ntiles, tile_comp, tmin, tmaxc,
                       margin_c, nrm_max_bits, bound);
    return hipGetLastError();
}

hipError_t launch_tile_umax(int64_t n, int64_t npad, const int *comp2, const float *bound, float *tU,
                            hipStream_t st) {
    hipLaunchKernelGGL(tile_umax_kernel, dim3((unsigned)(npad / 128)), dim3(128), 0, st, n, comp2, bound, tU);
    return hipGetLastError();
}

hipError_t launch_unpack_row_keys_ordered(int64_t n, const unsigned long long *row_key, const int *order,
                                          float *row_min, int *row_arg, hipStream_t st) {
    hipLaunchKernelGGL(unpack_row_keys_ordered_kernel, dim3((unsigned)nblocks(n, 256)), dim3(256), 0, st, n, row_key,
                       order, row_min, row_arg);
    return hipGetLastError();
}

hipError_t launch_tile_comp(int64_t n, int64_t npad, const int *comp_p, int *tile_comp, hipStream_t st) {
    hipLaunchKernelGGL(tile_comp_kernel, dim3((unsigned)(npad / 128)), dim3(128), 0, st, n, comp_p, tile_comp);
    return hipGetLastError();
}

hipError_t launch_knn_rescore(const KnnRescoreArgs &a, hipStream_t st) {
    if (a.k + 1 > 65 || a.cap > kRescoreCap) return hipErrorInvalidValue;
    hipLaunchKernelGGL(knn_rescore_kernel, dim3(nblocks(a.n, 4)), dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_knn_wave(const KnnWaveArgs &a, int count, hipStream_t st) {
    if (a.k > 64 || a.d > 64) return hipErrorInvalidValue;
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(knn_wave_kernel, dim3(count), dim3(64), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_knn_block(const KnnWaveArgs &a, int count, int64_t n, hipStream_t st) {
    if (a.k > 64 || a.d > 64) return hipErrorInvalidValue;
    if (count <= 0) return hipSuccess;
    const int64_t nleaves = int64_t(1) << (a.levels - 1);
    const int64_t leaf_cap = (n + nleaves - 1) / nleaves;
    if (a.levels < 1 || a.levels > kKnnBlockLevels || leaf_cap > 256) return launch_knn_wave(a, count, st);
    hipLaunchKernelGGL(knn_block_kernel, dim3(count), dim3(256), 0, st, a, (int)(leaf_cap < 1 ? 1 : leaf_cap));
    return hipGetLastError();
}

hipError_t launch_boruvka_upper(const BoruvkaExactArgs &a, hipStream_t st) {
    hipLaunchKernelGGL(boruvka_upper_kernel, dim3(nblocks(a.n, 256)), dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_boruvka_upper_approx(const BoruvkaExactArgs &a, hipStream_t st) {
    hipLaunchKernelGGL(boruvka_upper_approx_kernel, dim3(nblocks(a.n, 256)), dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_boruvka_rows(const BoruvkaExactArgs &a, hipStream_t st) {
    hipLaunchKernelGGL(boruvka_rows_kernel, dim3(nblocks(a.n, 256)), dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_boruvka_row_scan(const BoruvkaExactArgs &a, hipStream_t st) {
    hipLaunchKernelGGL(boruvka_row_scan_kernel, dim3(2048), dim3(256), 0, st, a);
    return hipGetLastError();
}

int64_t boruvka_rows_tiled_slots(int64_t n, int max_rows) { return (int64_t)max_rows * ((n + 2047) / 2048); }

hipError_t launch_cand_round(const CandRoundArgs &a, hipStream_t st) {
    if (a.n <= 0) return hipSuccess;
    if (a.d > 64) return hipErrorInvalidValue;
    hipLaunchKernelGGL(cand_round_kernel, dim3((unsigned)nblocks(a.n, 4)), dim3(256), 0, st, a);
    FH_BRUTE_CHECK(hipGetLastError());
    hipLaunchKernelGGL(cand_round_list_kernel, dim3((unsigned)nblocks(a.n, 256)), dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_boruvka_rows_tiled(const BoruvkaExactArgs &a, int row0, int nrows, float *slot_w, int *slot_j,
                                     hipStream_t st) {
    if (nrows <= 0) return hipSuccess;
    const int ncb = (int)((a.n + 2047) / 2048);
    const dim3 g((unsigned)ncb, (unsigned)((nrows + 63) / 64)), b(256);
    switch (brute_padded_dim(a.d)) {
        case 16: hipLaunchKernelGGL(boruvka_rows_tiled_kernel<16>, g, b, 0, st, a, row0, nrows, ncb, slot_w, slot_j); break;
        case 24: hipLaunchKernelGGL(boruvka_rows_tiled_kernel<24>, g, b, 0, st, a, row0, nrows, ncb, slot_w, slot_j); break;
        case 32: hipLaunchKernelGGL(boruvka_rows_tiled_kernel<32>, g, b, 0, st, a, row0, nrows, ncb, slot_w, slot_j); break;
        case 48: hipLaunchKernelGGL(boruvka_rows_tiled_kernel<48>, g, b, 0, st, a, row0, nrows, ncb, slot_w, slot_j); break;
        case 64: hipLaunchKernelGGL(boruvka_rows_tiled_kernel<64>, g, b, 0, st, a, row0, nrows, ncb, slot_w, slot_j); break;
        default: return hipErrorInvalidValue;
    }
    FH_BRUTE_CHECK(hipGetLastError());
    hipLaunchKernelGGL(boruvka_slots_kernel, dim3((unsigned)((nrows + 63) / 64)), dim3(64), 0, st, a, row0, nrows, ncb,
                       slot_w, slot_j);
    return hipGetLastError();
}

}  // namespace fh
