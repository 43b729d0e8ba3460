// Exact KD-tree construction on gfx950.
//
// The tree must equal numba_kdtree.build_kdtree(X, 40) (numba_kdtree.py:
// 478-521) including the order of points inside every node, which is the
// by-product of the reference's introselect swap sequence
// (numba_kdtree.py:280-407).  The build is level-synchronous: at level l all
// 2^l nodes compute their bounds and split dimension, then run their
// introselect concurrently.  Each introselect step is replayed exactly:
//
//   * median of three (280-299) and the pivot move: one lane per node;
//   * the two-pointer partition (313-350) pairs the m-th element >= pivot
//     counted from the left with the m-th element < pivot counted from the
//     right, for m < M = #(>= pivot) left of P = left + #(< pivot); both ranks
//     come from a prefix count of "< pivot" flags, so every swap of the
//     reference's scan happens here as one parallel exchange; the pivot
//     then swaps with position P;
//   * ranges of <= 16 (insertion sort, 192-201) and the depth-limit
//     fallback (heapsort, 256-266) both leave the range sorted; keys are
//     unique, so any sort reproduces them.
//
// Keys are 64-bit: the order-preserving image of the float coordinate on
// the split axis (-0 folded into +0, which the reference compares equal) in
// the high word and the original point index in the low word, so one
// integer comparison is the reference's (value, index) comparison
// (numba_kdtree.py:164-179) and the low word doubles as the permutation.
//
// Nodes larger than kMedium elements run their steps as multi-block kernels
// over global memory; nodes of a level whose largest node is at most kMedium
// run them in one work-group per node over global memory (no launches or
// host round trips per step); either way a range of at most kSmall elements
// finishes inside one work-group in LDS.
#include <hip/hip_runtime.h>

#include "fh_tree.h"

#include "fh_tree_dev.h"

namespace fh {

namespace {

// Level geometry: node k of level l covers [start(k), end(k)); the host
// uploads start/end for all nodes (they depend on n only).
struct LevelRef {
    const int2 *range;  // all nodes
    int first;          // 2^l - 1
    int count;          // 2^l
    int cps;            // chunks per node
};

__device__ __forceinline__ int wave_incl(int v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        int t = __shfl_up(v, o, 64);
        if (lane >= o) v += t;
    }
    return v;
}

// Block-wide exclusive scan for NT threads; total returned by reference.
template <int NT>
__device__ __forceinline__ int block_excl_nt(int v, int &total, int *ws) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int inc = wave_incl(v);
    if (lane == 63) ws[w] = inc;
    __syncthreads();
    int off = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < NT / 64; ++i) {
        int s = ws[i];
        off += i < w ? s : 0;
        tot += s;
    }
    __syncthreads();
    total = tot;
    return off + inc - v;
}

// ---------------------------------------------------------- bounds ----

// Blocks over (node, chunk): per-dimension min / max of the node's points,
// merged with order-preserving integer atomics (numba_kdtree.py:96-105).
// Up to kRegDims dimensions come from one gather of each point's row.
constexpr int kRegDims = 8;
__global__ __launch_bounds__(kTB) void bounds_kernel(LevelRef lv, const float *__restrict__ X, int d,
                                                     const uint64_t *__restrict__ A, int identity,
                                                     uint32_t *obox) {
    __shared__ uint32_t smin[kRegDims * (kTB / 64)], smax[kRegDims * (kTB / 64)];
    const int k = blockIdx.x / lv.cps, c = blockIdx.x % lv.cps;
    if (k >= lv.count) return;
    const int node = lv.first + k;
    const int2 r = lv.range[node];
    const int s = r.x + c * kChunk;
    if (s >= r.y) return;
    const int e = min(r.y, s + kChunk);
    if (d <= kRegDims) {
        // every dimension from one gather of the point's row
        uint32_t mn[kRegDims], mx[kRegDims];
#pragma unroll
        for (int j = 0; j < kRegDims; ++j) {
            mn[j] = 0xffffffffu;
            mx[j] = 0u;
        }
        // kBatch positions per lane at a time: their keys, then their rows,
        // so that the dependent loads overlap
        constexpr int kBatch = 8;
        for (int p0 = s + threadIdx.x; p0 < e; p0 += kBatch * kTB) {
            uint32_t idx[kBatch];
#pragma unroll
            for (int u = 0; u < kBatch; ++u) {
                const int p = p0 + u * kTB;
                idx[u] = p < e ? (identity ? (uint32_t)p : key_idx(A[p])) : 0xffffffffu;
            }
            float v[kBatch][kRegDims];
#pragma unroll
            for (int u = 0; u < kBatch; ++u)
#pragma unroll
                for (int j = 0; j < kRegDims; ++j)
                    if (j < d && idx[u] != 0xffffffffu) v[u][j] = X[(size_t)idx[u] * d + j];
#pragma unroll
            for (int u = 0; u < kBatch; ++u) {
                if (idx[u] == 0xffffffffu) continue;
#pragma unroll
                for (int j = 0; j < kRegDims; ++j) {
                    if (j < d) {
                        const uint32_t o = f2o(v[u][j]);
                        mn[j] = min(mn[j], o);
                        mx[j] = max(mx[j], o);
                    }
                }
            }
        }
#pragma unroll
        for (int j = 0; j < kRegDims; ++j) {
            if (j >= d) break;
            uint32_t a = mn[j], b = mx[j];
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) {
                a = min(a, (uint32_t)__shfl_xor((int)a, o, 64));
                b = max(b, (uint32_t)__shfl_xor((int)b, o, 64));
            }
            if ((threadIdx.x & 63) == 0) {
                smin[j * (kTB / 64) + (threadIdx.x >> 6)] = a;
                smax[j * (kTB / 64) + (threadIdx.x >> 6)] = b;
            }
        }
        __syncthreads();
        if (threadIdx.x < d) {
            const int j = threadIdx.x;
            uint32_t a = smin[j * (kTB / 64)], b = smax[j * (kTB / 64)];
            for (int w = 1; w < kTB / 64; ++w) {
                a = min(a, smin[j * (kTB / 64) + w]);
                b = max(b, smax[j * (kTB / 64) + w]);
            }
            atomicMin(obox + (size_t)node * 2 * d + j, a);
            atomicMax(obox + (size_t)node * 2 * d + d + j, b);
        }
        return;
    }
    for (int j = 0; j < d; ++j) {
        uint32_t mn = 0xffffffffu, mx = 0u;
        for (int p = s + threadIdx.x; p < e; p += kTB) {
            const uint32_t idx = identity ? (uint32_t)p : key_idx(A[p]);
            const uint32_t o = f2o(X[(size_t)idx * d + j]);
            mn = min(mn, o);
            mx = max(mx, o);
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            mn = min(mn, (uint32_t)__shfl_xor((int)mn, o, 64));
            mx = max(mx, (uint32_t)__shfl_xor((int)mx, o, 64));
        }
        if ((threadIdx.x & 63) == 0) {
            smin[threadIdx.x >> 6] = mn;
            smax[threadIdx.x >> 6] = mx;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int w = 1; w < kTB / 64; ++w) {
                mn = min(mn, smin[w]);
                mx = max(mx, smax[w]);
            }
            atomicMin(obox + (size_t)node * 2 * d + j, mn);
            atomicMax(obox + (size_t)node * 2 * d + d + j, mx);
        }
        __syncthreads();
    }
}

__global__ void bounds_init_kernel(int count, int first, int d, uint32_t *obox) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count * 2 * d) return;
    const int k = t / (2 * d), j = t % (2 * d);
    obox[(size_t)(first + k) * 2 * d + j] = j < d ? 0xffffffffu : 0u;
}

// Per node: float bounds, split dimension (first strictly largest spread,
// numba_kdtree.py:133-152) and the introselect state.
__global__ void axis_kernel(LevelRef lv, int d, const uint32_t *__restrict__ obox, float *box, int *axis,
                            SegState *seg, int internal) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= lv.count) return;
    const int node = lv.first + k;
    float best = 0.0f;
    int ax = 0;
    for (int j = 0; j < d; ++j) {
        const float lo = o2f(obox[(size_t)node * 2 * d + j]);
        const float hi = o2f(obox[(size_t)node * 2 * d + d + j]);
        box[(size_t)node * 2 * d + j] = lo;
        box[(size_t)node * 2 * d + d + j] = hi;
        const float spread = hi - lo;
        if (spread > best) {
            best = spread;
            ax = j;
        }
    }
    if (!internal) return;
    axis[k] = ax;
    const int2 r = lv.range[node];
    const int size = r.y - r.x;
    SegState st;
    st.L = r.x;
    st.R = r.y;
    st.nth = r.x + size / 2;
    st.depth = size > 16 ? 2 * floor_log2(size) : 0;
    st.flag = size > kSmall ? SEG_LARGE : SEG_SMALL;
    st.pivot = 0;
    seg[k] = st;
}

// Blocks over (node, chunk): keys for the node's split axis.
__global__ __launch_bounds__(kTB) void keys_kernel(LevelRef lv, const float *__restrict__ X, int d,
                                                   const int *__restrict__ axis, uint64_t *A, int identity) {
    const int k = blockIdx.x / lv.cps, c = blockIdx.x % lv.cps;
    if (k >= lv.count) return;
    const int2 r = lv.range[lv.first + k];
    const int s = r.x + c * kChunk;
    if (s >= r.y) return;
    const int e = min(r.y, s + kChunk);
    const int ax = axis[k];
    // all of the lane's positions at once: keys, then coordinates, then stores
    uint32_t idx[kPer];
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
        const int p = s + threadIdx.x + u * kTB;
        idx[u] = p < e ? (identity ? (uint32_t)p : key_idx(A[p])) : 0u;
    }
    float v[kPer];
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
        const int p = s + threadIdx.x + u * kTB;
        v[u] = p < e ? X[(size_t)idx[u] * d + ax] : 0.0f;
    }
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
        const int p = s + threadIdx.x + u * kTB;
        if (p < e) A[p] = make_key(v[u], idx[u]);
    }
}

// ------------------------------------------- multi-block introselect ----


// numba_kdtree.py:280-299 and the pivot move of 313-318.
__device__ __forceinline__ void median3_and_move(uint64_t *A, int L, int R) {
    const int mid = (L + R - 1) / 2;
    if (A[L] > A[mid]) swap64(A, L, mid);
    if (A[mid] > A[R - 1]) {
        swap64(A, mid, R - 1);
        if (A[L] > A[mid]) swap64(A, L, mid);
    }
    swap64(A, mid, R - 1);
}

// Pivot of the next step of a large segment (numba_kdtree.py:301-318):
// ranges of at most kSmall go to the in-LDS finish, ranges at the depth
// limit to the sort (counted in *sorts).  A segment left SEG_LARGE has its
// pivot moved to R-1 and counts itself in *active.
__device__ __forceinline__ void pivot_step(SegState &st, uint64_t *A, int *active, int *sorts) {
    const int size = st.R - st.L;
    if (size <= kSmall) {
        st.flag = SEG_SMALL;
        return;
    }
    if (st.depth == 0) {
        st.flag = SEG_SORT;
        atomicAdd(sorts, 1);
        return;
    }
    st.depth -= 1;
    median3_and_move(A, st.L, st.R);
    st.pivot = A[st.R - 1];
    atomicAdd(active, 1);
}

// First pivot of every large segment of a level; the node's exchange count
// starts at zero.
__global__ void big_pivot_kernel(int count, SegState *seg, uint64_t *A, int *aux, int *active, int *sorts) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    aux[4 * k + 2] = 0;
    SegState st = seg[k];
    if (st.flag != SEG_LARGE) return;
    pivot_step(st, A, active, sorts);
    seg[k] = st;
}

constexpr int kWavesTB = kTB / 64;
constexpr int kWaveChunk = kChunk / kWavesTB;  // positions per wave of a chunk

// Blocks over (segment, chunk of [L, R - 1)): keys < pivot, counted per wave
// over a contiguous quarter of the chunk (coalesced, four loads in flight).
__global__ __launch_bounds__(kTB) void big_count_kernel(int count, int cps, const SegState *__restrict__ seg,
                                                        const uint64_t *__restrict__ A, int *cnt) {
    const int k = blockIdx.x / cps, c = blockIdx.x % cps;
    if (k >= count) return;
    const SegState st = seg[k];
    if (st.flag != SEG_LARGE) return;
    const int w = threadIdx.x >> 6;
    const int s = st.L + c * kChunk;
    const int e = min(st.R - 1, s + kChunk);
    const int b = min(e, s + w * kWaveChunk), f = min(e, b + kWaveChunk);
    const int v = sub_count_pass(A, st.pivot, b, f);
    if ((threadIdx.x & 63) == 0) cnt[((size_t)k * cps + c) * kWavesTB + w] = v;
}

// Sums of the per-wave counts of one segment: those before flat index t
// (the same t for the whole block; returned) and all nch of them (total).
__device__ __forceinline__ int wave_count_prefix(const int *__restrict__ cnt, int nch, int t, int &total, int *ws) {
    int pre = 0, all = 0;
    for (int i = threadIdx.x; i < nch; i += kTB) {
        const int v = cnt[i];
        all += v;
        pre += i < t ? v : 0;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        pre += __shfl_xor(pre, o, 64);
        all += __shfl_xor(all, o, 64);
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        ws[2 * w] = pre;
        ws[2 * w + 1] = all;
    }
    __syncthreads();
    pre = 0;
    all = 0;
#pragma unroll
    for (int i = 0; i < kTB / 64; ++i) {
        pre += ws[2 * i];
        all += ws[2 * i + 1];
    }
    __syncthreads();
    total = all;
    return pre;
}

// Blocks over (segment, chunk): S = keys < pivot, the pivot's final place
// P = L + S, and the exchange partners of the reference's two-pointer scan
// (numba_kdtree.py:320-348) in part[]: the m-th ">= pivot" key left of P at
// L + m, the m-th "< pivot" key counted from the right at R - 2 - m.  aux
// holds (S, P, M) per segment; M (the exchange count) is summed per block.
__global__ __launch_bounds__(kTB) void big_rank_kernel(int count, int cps, const SegState *__restrict__ seg,
                                                       const uint64_t *__restrict__ A, const int *__restrict__ cnt,
                                                       int *part, int *aux) {
    __shared__ int ws[2 * (kTB / 64)];
    __shared__ int wm[kTB / 64];
    const int k = blockIdx.x / cps, c = blockIdx.x % cps;
    if (k >= count) return;
    const SegState st = seg[k];
    if (st.flag != SEG_LARGE) return;
    const int s = st.L + c * kChunk;
    const int e = min(st.R - 1, s + kChunk);
    if (s >= e) return;
    const int w = threadIdx.x >> 6;
    const int nch = (st.R - 1 - st.L + kChunk - 1) / kChunk;
    int S;
    const int *cn = cnt + (size_t)k * cps * kWavesTB;
    int pre = wave_count_prefix(cn, nch * kWavesTB, c * kWavesTB, S, ws);
    for (int i = 0; i < w; ++i) pre += cn[c * kWavesTB + i];
    const int P = st.L + S;
    if (c == 0 && threadIdx.x == 0) {
        aux[4 * k] = S;
        aux[4 * k + 1] = P;
    }
    const int b = min(e, s + w * kWaveChunk), f = min(e, b + kWaveChunk);
    const int m = sub_rank_pass(A, part, st.pivot, st.L, st.R, S, b, f, pre);
    if ((threadIdx.x & 63) == 0) wm[w] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        int mb = 0;
#pragma unroll
        for (int i = 0; i < kTB / 64; ++i) mb += wm[i];
        if (mb) atomicAdd(aux + 4 * k + 2, mb);
    }
}

// Blocks over (segment, chunk of exchange pairs): the exchanges.
__global__ __launch_bounds__(kTB) void big_swap_kernel(int count, int cps, const SegState *__restrict__ seg,
                                                       uint64_t *A, const int *__restrict__ part,
                                                       const int *__restrict__ aux) {
    const int k = blockIdx.x / cps, c = blockIdx.x % cps;
    if (k >= count) return;
    const SegState st = seg[k];
    if (st.flag != SEG_LARGE) return;
    const int m0 = c * kChunk;
    const int m1 = min(aux[4 * k + 2], m0 + kChunk);
    for (int m = m0 + threadIdx.x; m < m1; m += kTB) swap64(A, part[st.L + m], part[st.R - 2 - m]);
}

// Per segment: pivot into place (numba_kdtree.py:349), narrowing (376-383),
// then the pivot of the segment's next step; the exchange count restarts
// at zero.
__global__ void big_finish_kernel(int count, SegState *seg, uint64_t *A, int *aux, int *active, int *sorts) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    SegState st = seg[k];
    if (st.flag != SEG_LARGE) return;
    const int P = aux[4 * k + 1];
    swap64(A, P, st.R - 1);
    aux[4 * k + 2] = 0;
    if (st.nth < P) {
        st.R = P;
    } else if (st.nth > P) {
        st.L = P + 1;
    } else {
        st.flag = SEG_DONE;
        seg[k] = st;
        return;
    }
    pivot_step(st, A, active, sorts);
    seg[k] = st;
}

// ------------------------------------------------- in-LDS introselect ----


// One block of NT threads per node of the level: finish the introselect of
// segments that are (or became) at most CAP <= kSmall long.  Deep levels,
// whose nodes hold a few dozen keys, use a small CAP and NT so that many
// blocks share a CU.
template <int CAP, int NT>
__global__ __launch_bounds__(NT) void small_select_kernel(int count, SegState *seg, uint64_t *A) {
    __shared__ uint64_t sk[CAP];
    __shared__ int gt_s[CAP / 2], lt_s[CAP / 2];
    __shared__ int ws[NT / 64];
    __shared__ int sh[4];
    const int k = blockIdx.x;
    if (k >= count) return;
    SegState st = seg[k];
    if (st.flag != SEG_SMALL) return;
    const int L0 = st.L, R0 = st.R;
    const int n0 = R0 - L0;
    for (int p = threadIdx.x; p < n0; p += NT) sk[p] = A[L0 + p];
    __syncthreads();
    // local coordinates
    int L = 0, R = n0, nth = st.nth - L0, depth = st.depth;
    bool sort_rest = false;
    for (;;) {
        if (R - L <= 16) {
            sort_rest = true;
            break;
        }
        if (depth == 0) {
            sort_rest = true;
            break;
        }
        --depth;
        if (threadIdx.x == 0) median3_and_move(sk, L, R);
        __syncthreads();
        const uint64_t pivot = sk[R - 1];
        const int len = R - 1 - L;
        const int per = (len + NT - 1) / NT;
        const int b = L + threadIdx.x * per;
        int v = 0;
        for (int j = 0; j < per; ++j) {
            const int p = b + j;
            v += (p < R - 1 && sk[p] < pivot) ? 1 : 0;
        }
        int S;
        int before = block_excl_nt<NT>(v, S, ws);
        const int P = L + S;
        if (threadIdx.x == 0) sh[0] = 0;
        __syncthreads();
        for (int j = 0; j < per; ++j) {
            const int p = b + j;
            if (p >= R - 1) break;
            const bool lt = sk[p] < pivot;
            if (p == P) sh[0] = (P - L) - before;
            if (p < P) {
                if (!lt) gt_s[(p - L) - before] = p;
            } else if (lt) {
                lt_s[S - 1 - before] = p;
            }
            before += lt ? 1 : 0;
        }
        __syncthreads();
        const int M = sh[0];
        for (int m = threadIdx.x; m < M; m += NT) {
            const int a = gt_s[m], c = lt_s[m];
            const uint64_t t = sk[a];
            sk[a] = sk[c];
            sk[c] = t;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            const uint64_t t = sk[P];
            sk[P] = sk[R - 1];
            sk[R - 1] = t;
        }
        __syncthreads();
        if (nth < P)
            R = P;
        else if (nth > P)
            L = P + 1;
        else
            break;
    }
    if (sort_rest) {
        // insertion sort (<= 16) or heapsort at the depth limit: both leave
        // [L, R) ascending
        lds_sort(sk + L, R - L);
    }
    for (int p = threadIdx.x; p < n0; p += NT) A[L0 + p] = sk[p];
    seg[k].flag = SEG_DONE;
}

// ------------------------------------- one work-group per node ----

constexpr int kMTB = 1024;  // threads of the one-work-group-per-node steps
constexpr int kMW = kMTB / 64;


// One block per node of a level whose nodes are at most kMedium long: the
// steps over global memory while the active range is longer than kSmall
// (a range reaching the depth limit first is sorted in place, as the
// reference heap-sorts it); small_select_kernel then finishes in LDS.
__global__ __launch_bounds__(kMTB) void medium_select_kernel(int count, SegState *seg, uint64_t *A, int *part) {
    __shared__ int ws[kMW], ws2[kMW];
    const int k = blockIdx.x;
    if (k >= count) return;
    SegState st = seg[k];
    if (st.flag != SEG_LARGE) return;
    int L = st.L, R = st.R, depth = st.depth;
    const int nth = st.nth;
    bool done = false;
    while (R - L > kSmall) {
        if (depth == 0) {
            lds_sort(A + L, R - L);  // the same network, here over global memory
            done = true;
            break;
        }
        --depth;
        const int P = wg_select_step<kMTB>(A, part, L, R, ws, ws2);
        if (nth < P) {
            R = P;
        } else if (nth > P) {
            L = P + 1;
        } else {
            done = true;
            break;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        if (done) {
            seg[k].flag = SEG_DONE;
        } else {
            seg[k].L = L;
            seg[k].R = R;
            seg[k].depth = depth;
            seg[k].flag = SEG_SMALL;
        }
    }
}

// ---------------------------------------- one work-group per subtree ----

// (the build itself: subtree_build in fh_tree_dev.h)
template <int D>
__global__ __launch_bounds__(kSubNT) void subtree_kernel(SubtreeArgs a) {
    __shared__ __align__(16) unsigned char lds[subtree_lds_bytes<D>()];
    subtree_build<D>(a, blockIdx.x, lds);
}

// ----------------------------------------------------------- output ----

__global__ void finalize_kernel(int n, int d, const float *__restrict__ X, const uint64_t *__restrict__ A, int *perm,
                                int *posof, float *xp) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int idx = (int)key_idx(A[p]);
    perm[p] = idx;
    posof[idx] = p;
    for (int j = 0; j < d; ++j) xp[(size_t)p * d + j] = X[(size_t)idx * d + j];
}

__global__ void identity_keys_kernel(int n, uint64_t *A) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) A[p] = (uint64_t)p;
}

// One compare-exchange stage of the same-direction bitonic network over
// A[L, L+n) in global memory (positions >= n act as +inf).  Used only for a
// range that reaches the introselect depth limit while still longer than
// kSmall (the reference then heap-sorts it).
__global__ void gsort_stage_kernel(uint64_t *A, int n, int m, int size, int stride) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= m / 2) return;
    int lo, hi;
    if (stride == 0) {
        const int half = size >> 1;
        lo = (t / half) * size + (t % half);
        hi = (t / half) * size + size - 1 - (t % half);
    } else {
        lo = 2 * stride * (t / stride) + (t % stride);
        hi = lo + stride;
    }
    if (hi < n && A[lo] > A[hi]) {
        const uint64_t x = A[lo];
        A[lo] = A[hi];
        A[hi] = x;
    }
}

inline unsigned nblk(int64_t m, int b) { return (unsigned)((m + b - 1) / b); }

}  // namespace

int tree_small_limit() { return kSmall; }
int tree_medium_limit() { return kMedium; }
int tree_chunk() { return kChunk; }

hipError_t tree_level_bounds(const TreeBuildArgs &a, int level, int max_size, int identity, hipStream_t st) {
    LevelRef lv{a.range, (1 << level) - 1, 1 << level, (max_size + kChunk - 1) / kChunk};
    if (lv.cps < 1) lv.cps = 1;
    hipLaunchKernelGGL(bounds_init_kernel, dim3(nblk((int64_t)lv.count * 2 * a.d, 256)), dim3(256), 0, st, lv.count,
                       lv.first, a.d, a.obox);
    hipLaunchKernelGGL(bounds_kernel, dim3((unsigned)(lv.count * lv.cps)), dim3(kTB), 0, st, lv, a.X, a.d, a.A,
                       identity, a.obox);
    return hipGetLastError();
}

hipError_t tree_level_axis(const TreeBuildArgs &a, int level, int internal, hipStream_t st) {
    LevelRef lv{a.range, (1 << level) - 1, 1 << level, 1};
    hipLaunchKernelGGL(axis_kernel, dim3(nblk(lv.count, 256)), dim3(256), 0, st, lv, a.d, a.obox, a.box, a.axis,
                       a.seg, internal);
    return hipGetLastError();
}

hipError_t tree_level_keys(const TreeBuildArgs &a, int level, int max_size, int identity, hipStream_t st) {
    LevelRef lv{a.range, (1 << level) - 1, 1 << level, (max_size + kChunk - 1) / kChunk};
    if (lv.cps < 1) lv.cps = 1;
    hipLaunchKernelGGL(keys_kernel, dim3((unsigned)(lv.count * lv.cps)), dim3(kTB), 0, st, lv, a.X, a.d, a.axis, a.A,
                       identity);
    return hipGetLastError();
}

hipError_t tree_big_begin(const TreeBuildArgs &a, int count, hipStream_t st) {
    hipLaunchKernelGGL(big_pivot_kernel, dim3(nblk(count, 256)), dim3(256), 0, st, count, a.seg, a.A, a.aux, a.active,
                       a.sorts);
    return hipGetLastError();
}

hipError_t tree_big_step(const TreeBuildArgs &a, int count, int max_size, hipStream_t st) {
    const int cps = (max_size + kChunk - 1) / kChunk;
    hipLaunchKernelGGL(big_count_kernel, dim3((unsigned)(count * cps)), dim3(kTB), 0, st, count, cps, a.seg, a.A,
                       a.cnt);
    hipLaunchKernelGGL(big_rank_kernel, dim3((unsigned)(count * cps)), dim3(kTB), 0, st, count, cps, a.seg, a.A,
                       a.cnt, a.gtpos, a.aux);
    hipLaunchKernelGGL(big_swap_kernel, dim3((unsigned)(count * cps)), dim3(kTB), 0, st, count, cps, a.seg, a.A,
                       a.gtpos, a.aux);
    hipLaunchKernelGGL(big_finish_kernel, dim3(nblk(count, 256)), dim3(256), 0, st, count, a.seg, a.A, a.aux, a.active,
                       a.sorts);
    return hipGetLastError();
}

hipError_t tree_medium_select(const TreeBuildArgs &a, int count, hipStream_t st) {
    hipLaunchKernelGGL(medium_select_kernel, dim3((unsigned)count), dim3(kMTB), 0, st, count, a.seg, a.A, a.gtpos);
    return hipGetLastError();
}

hipError_t tree_small_select(const TreeBuildArgs &a, int count, int max_size, hipStream_t st) {
    if (max_size <= 128)
        hipLaunchKernelGGL((small_select_kernel<128, 64>), dim3((unsigned)count), dim3(64), 0, st, count, a.seg, a.A);
    else if (max_size <= 512)
        hipLaunchKernelGGL((small_select_kernel<512, 128>), dim3((unsigned)count), dim3(128), 0, st, count, a.seg,
                           a.A);
    else
        hipLaunchKernelGGL((small_select_kernel<kSmall, kTB>), dim3((unsigned)count), dim3(kTB), 0, st, count, a.seg,
                           a.A);
    return hipGetLastError();
}

hipError_t tree_sort_range(uint64_t *A, int L, int R, hipStream_t st) {
    const int n = R - L;
    int m = 1;
    while (m < n) m <<= 1;
    for (int size = 2; size <= m; size <<= 1) {
        hipLaunchKernelGGL(gsort_stage_kernel, dim3(nblk(m / 2, 256)), dim3(256), 0, st, A + L, n, m, size, 0);
        for (int stride = size >> 2; stride > 0; stride >>= 1)
            hipLaunchKernelGGL(gsort_stage_kernel, dim3(nblk(m / 2, 256)), dim3(256), 0, st, A + L, n, m, size,
                               stride);
    }
    return hipGetLastError();
}

hipError_t tree_identity_keys(int n, uint64_t *A, hipStream_t st) {
    hipLaunchKernelGGL(identity_keys_kernel, dim3(nblk(n, 256)), dim3(256), 0, st, n, A);
    return hipGetLastError();
}

int tree_subtree_cap(int d) {
    switch (d) {
        case 1: return sub_cap<1>();
        case 2: return sub_cap<2>();
        case 3: return sub_cap<3>();
        case 4: return sub_cap<4>();
        case 5: return sub_cap<5>();
        case 6: return sub_cap<6>();
        case 7: return sub_cap<7>();
        case 8: return sub_cap<8>();
        default: return 0;
    }
}

hipError_t tree_subtrees(const TreeBuildArgs &a, int level0, int levels, int *perm, int *posof, float *xp,
                         hipStream_t st) {
    SubtreeArgs s{};
    s.X = a.X;
    s.range = a.range;
    s.A = a.A;
    s.identity = level0 == 0 ? 1 : 0;
    s.level0 = level0;
    s.levels = levels;
    s.box = a.box;
    s.perm = perm;
    s.posof = posof;
    s.xp = xp;
    const dim3 grid(1u << level0), block(kSubNT);
    switch (a.d) {
        case 1: hipLaunchKernelGGL(subtree_kernel<1>, grid, block, 0, st, s); break;
        case 2: hipLaunchKernelGGL(subtree_kernel<2>, grid, block, 0, st, s); break;
        case 3: hipLaunchKernelGGL(subtree_kernel<3>, grid, block, 0, st, s); break;
        case 4: hipLaunchKernelGGL(subtree_kernel<4>, grid, block, 0, st, s); break;
        case 5: hipLaunchKernelGGL(subtree_kernel<5>, grid, block, 0, st, s); break;
        case 6: hipLaunchKernelGGL(subtree_kernel<6>, grid, block, 0, st, s); break;
        case 7: hipLaunchKernelGGL(subtree_kernel<7>, grid, block, 0, st, s); break;
        case 8: hipLaunchKernelGGL(subtree_kernel<8>, grid, block, 0, st, s); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t tree_finalize(const TreeBuildArgs &a, int n, int *perm, int *posof, float *xp, hipStream_t st) {
    hipLaunchKernelGGL(finalize_kernel, dim3(nblk(n, 256)), dim3(256), 0, st, n, a.d, a.X, a.A, perm, posof, xp);
    return hipGetLastError();
}

}  // namespace fh
