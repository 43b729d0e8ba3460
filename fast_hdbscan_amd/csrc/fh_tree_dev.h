// Device code of the exact KD-tree build shared by the level-synchronous
// build (fh_tree.hip) and the persistent batch kernel (fh_batch.hip): the
// 64-bit (value, index) keys, the ballot-ranked partition passes of one
// introselect step, and the build of a whole subtree by one work-group in
// LDS.  Internal linkage: each translation unit gets its own copy.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fh_tree.h"

namespace fh {

namespace {

constexpr int kSmall = 4096;    // largest range finished in LDS
constexpr int kMedium = 1 << 16;  // largest node whose steps one work-group runs
constexpr int kChunk = 4096;    // elements per block in the multi-block steps
constexpr int kTB = 256;        // threads per block
constexpr int kPer = kChunk / kTB;

__device__ __forceinline__ uint32_t f2o(float f) {
    uint32_t b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float o2f(uint32_t o) {
    uint32_t b = (o & 0x80000000u) ? (o & 0x7fffffffu) : ~o;
    return __uint_as_float(b);
}
__device__ __forceinline__ uint64_t make_key(float v, uint32_t idx) {
    if (v == 0.0f) v = 0.0f;  // -0 == +0 in the reference's comparison
    return ((uint64_t)f2o(v) << 32) | idx;
}
__device__ __forceinline__ uint32_t key_idx(uint64_t k) { return (uint32_t)(k & 0xffffffffull); }

__device__ __forceinline__ int floor_log2(int v) { return 31 - __clz(v); }

// ------------------------------------------- partition passes ----
//
// Helpers shared by the one-work-group-per-node steps (keys in global
// memory) and the subtree kernel (keys in LDS): a wave owns a contiguous
// run of positions and reads it 64 keys at a time, four loads in flight;
// ranks come from ballots.

__device__ __forceinline__ void sub_wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ void sub_swap(uint64_t *sk, uint16_t *sl, int i, int j) {
    const uint64_t t = sk[i];
    sk[i] = sk[j];
    sk[j] = t;
    const uint16_t u = sl[i];
    sl[i] = sl[j];
    sl[j] = u;
}

// median3_and_move on [L, R) (size > 2), the three keys and their slots
// read once: every lane computes the same result in registers, the caller's
// leader writes them back (positions L, mid, R - 1); the pivot is k[2].
struct Med3 {
    uint64_t k[3];
    uint16_t s[3];
};
__device__ __forceinline__ void sub_cswap(Med3 &m, int i, int j) {
    const uint64_t t = m.k[i];
    m.k[i] = m.k[j];
    m.k[j] = t;
    const uint16_t u = m.s[i];
    m.s[i] = m.s[j];
    m.s[j] = u;
}
__device__ __forceinline__ Med3 sub_median3(const uint64_t *sk, const uint16_t *sl, int L, int R) {
    const int mid = (L + R - 1) / 2;
    Med3 m;
    m.k[0] = sk[L];
    m.k[1] = sk[mid];
    m.k[2] = sk[R - 1];
    m.s[0] = sl[L];
    m.s[1] = sl[mid];
    m.s[2] = sl[R - 1];
    if (m.k[0] > m.k[1]) sub_cswap(m, 0, 1);
    if (m.k[1] > m.k[2]) {
        sub_cswap(m, 1, 2);
        if (m.k[0] > m.k[1]) sub_cswap(m, 0, 1);
    }
    sub_cswap(m, 1, 2);  // the pivot moves to R - 1
    return m;
}
__device__ __forceinline__ void sub_median3_store(const Med3 &m, uint64_t *sk, uint16_t *sl, int L, int R) {
    const int mid = (L + R - 1) / 2;
    sk[L] = m.k[0];
    sk[mid] = m.k[1];
    sk[R - 1] = m.k[2];
    sl[L] = m.s[0];
    sl[mid] = m.s[1];
    sl[R - 1] = m.s[2];
}

// lds_sort's network over [L, L+n) by one wave: ascending, slots moving
// along.
__device__ void sub_sort(uint64_t *sk, uint16_t *sl, int L, int n) {
    const int lane = threadIdx.x & 63;
    int m = 1;
    while (m < n) m <<= 1;
    for (int size = 2; size <= m; size <<= 1) {
        const int half = size >> 1;
        for (int t = lane; t < m / 2; t += 64) {
            const int lo = (t / half) * size + (t % half);
            const int hi = (t / half) * size + size - 1 - (t % half);
            if (hi < n && sk[L + lo] > sk[L + hi]) sub_swap(sk, sl, L + lo, L + hi);
        }
        sub_wave_sync();
        for (int stride = size >> 2; stride > 0; stride >>= 1) {
            for (int t = lane; t < m / 2; t += 64) {
                const int lo = 2 * stride * (t / stride) + (t % stride);
                const int hi = lo + stride;
                if (hi < n && sk[L + lo] > sk[L + hi]) sub_swap(sk, sl, L + lo, L + hi);
            }
            sub_wave_sync();
        }
    }
}

// Ranks of one partition step over [b, e) (part of [L, R - 1)) by one wave,
// four loads in flight: "< pivot" keys left of P go nowhere, ">= pivot"
// keys left of P are written at part[L + rank], "< pivot" keys at or right
// of P at part[R - 2 - rank from the right].  lts enters as the "< pivot"
// count before b and the return value is the ">= pivot" count left of P in
// [b, e).
template <typename T>
__device__ __forceinline__ int sub_rank_pass(const uint64_t *sk, T *part, uint64_t pivot, int L, int R,
                                             int S, int b, int e, int lts) {
    const int lane = threadIdx.x & 63;
    const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const int P = L + S;
    int M = 0;
    for (int p0 = b; p0 < e; p0 += 256) {
        uint64_t kv[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int p = p0 + u * 64 + lane;
            kv[u] = p < e ? sk[p] : ~0ull;
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int p = p0 + u * 64 + lane;
            const bool in = p < e;
            const bool lt = in && kv[u] < pivot;
            const uint64_t ltm = __ballot(lt);
            const int before = lts + __popcll(ltm & below);
            if (in) {
                if (p < P) {
                    if (!lt) part[L + (p - L) - before] = (T)p;
                } else if (lt) {
                    part[R - 2 - (S - 1 - before)] = (T)p;
                }
            }
            lts += __popcll(ltm);
            M += __popcll(__ballot(in && !lt && p < P));
        }
    }
    return M;
}

// "< pivot" keys in [b, e) by one wave, four loads in flight
__device__ __forceinline__ int sub_count_pass(const uint64_t *sk, uint64_t pivot, int b, int e) {
    const int lane = threadIdx.x & 63;
    int S = 0;
    for (int p0 = b; p0 < e; p0 += 256) {
        uint64_t kv[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int p = p0 + u * 64 + lane;
            kv[u] = p < e ? sk[p] : ~0ull;
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) S += __popcll(__ballot(kv[u] < pivot));
    }
    return S;
}

__device__ __forceinline__ void swap64(uint64_t *A, int i, int j) {
    uint64_t t = A[i];
    A[i] = A[j];
    A[j] = t;
}

// Ascending sort of n keys in LDS with a bitonic network whose merge steps
// all point the same way (the first step of each merge compares mirrored
// positions), so positions >= n behave as +inf and need no storage.
__device__ void lds_sort(uint64_t *sk, int n) {
    int m = 1;
    while (m < n) m <<= 1;
    for (int size = 2; size <= m; size <<= 1) {
        for (int t = threadIdx.x; t < m / 2; t += blockDim.x) {
            const int half = size >> 1;
            const int blk = t / half, off = t % half;
            const int lo = blk * size + off;
            const int hi = blk * size + size - 1 - off;
            if (hi < n && sk[lo] > sk[hi]) {
                const uint64_t x = sk[lo];
                sk[lo] = sk[hi];
                sk[hi] = x;
            }
        }
        __syncthreads();
        for (int stride = size >> 2; stride > 0; stride >>= 1) {
            for (int t = threadIdx.x; t < m / 2; t += blockDim.x) {
                const int lo = 2 * stride * (t / stride) + (t % stride);
                const int hi = lo + stride;
                if (hi < n && sk[lo] > sk[hi]) {
                    const uint64_t x = sk[lo];
                    sk[lo] = sk[hi];
                    sk[hi] = x;
                }
            }
            __syncthreads();
        }
    }
}

// One partition step of introselect on keys[L, R) (pivot of numba_kdtree.py:
// 280-299, the two-pointer exchanges of 313-350 as one parallel exchange,
// the pivot move of 349) by one work-group of NT threads, keys in global
// memory: every thread reads the median-of-three candidates, each wave
// counts and ranks one contiguous (NT / 64)-th of [L, R - 1) with coalesced
// loads, and part[] (the scratch array of the multi-block steps) holds the
// exchange partners: the m-th ">= pivot" key left of P at L + m, the m-th
// "< pivot" key from the right at R - 2 - m.  Returns the pivot's final
// position P.
template <int NT>
__device__ int wg_select_step(uint64_t *keys, int *part, int L, int R, int *ws, int *ws2) {
    constexpr int kMW = NT / 64;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int mid = (L + R - 1) / 2;
    uint64_t k0 = keys[L], k1 = keys[mid], k2 = keys[R - 1];
    if (k0 > k1) {
        const uint64_t t = k0;
        k0 = k1;
        k1 = t;
    }
    if (k1 > k2) {
        const uint64_t t = k1;
        k1 = k2;
        k2 = t;
        if (k0 > k1) {
            const uint64_t u = k0;
            k0 = k1;
            k1 = u;
        }
    }
    // the median moves to R - 1 and becomes the pivot
    const uint64_t pivot = k1;
    __syncthreads();
    if (threadIdx.x == 0) {
        keys[L] = k0;
        keys[mid] = k2;
        keys[R - 1] = k1;
    }
    __syncthreads();
    const int len = R - 1 - L;
    const int b = L + (int)((int64_t)len * w / kMW), e = L + (int)((int64_t)len * (w + 1) / kMW);
    const int c = sub_count_pass(keys, pivot, b, e);
    if (lane == 0) ws[w] = c;
    __syncthreads();
    int S = 0, pre = 0;
#pragma unroll
    for (int i = 0; i < kMW; ++i) {
        const int v = ws[i];
        S += v;
        pre += i < w ? v : 0;
    }
    const int P = L + S;
    const int mw = sub_rank_pass(keys, part, pivot, L, R, S, b, e, pre);
    if (lane == 0) ws2[w] = mw;
    __syncthreads();
    int M = 0;
#pragma unroll
    for (int i = 0; i < kMW; ++i) M += ws2[i];
    for (int m = threadIdx.x; m < M; m += NT) swap64(keys, part[L + m], part[R - 2 - m]);
    __syncthreads();
    if (threadIdx.x == 0) swap64(keys, P, R - 1);
    __syncthreads();
    return P;
}

// ---------------------------------------- one work-group per subtree ----
//
// All levels from level0 down of a subtree whose root holds at most
// sub_cap<D>() points, in one work-group over LDS: the points' rows are
// gathered once (slot order), every key (ordered coordinate << 32 | point
// index) moves together with its slot through the selects, and each level
// runs what the level-synchronous kernels above run over global memory --
// per-node bounds (bounds_kernel), split axis and box (axis_kernel), keys
// (keys_kernel) and the exact introselect of every internal node
// (small_select_kernel), here one wave per node.  Levels with fewer nodes
// than waves share each node's bounds and keys among several waves (bounds
// merged with LDS atomics); from the level with one node per wave on, each
// wave builds the subtrees of its own nodes with no work-group barrier.
// The work-group ends with the tree-order outputs of its positions
// (finalize_kernel).  LDS holds 12 + 4D bytes per point, sized so that two
// work-groups share a CU.

constexpr int kSubNT = 512;
constexpr int kSubW = kSubNT / 64;
constexpr int kSubLogW = 3;
static_assert((1 << kSubLogW) == kSubW, "waves per subtree work-group");

template <int D>
constexpr int sub_cap() {
    return ((81920 - 512) / (12 + 4 * D)) / 64 * 64;
}

// The introselect of node [L, R) towards nth (small_select_kernel's loop)
// by one wave: lanes over consecutive positions, ranks from ballots.  The
// exchange partners go to part[]: the m-th ">= pivot" key left of P at
// L + m, the m-th "< pivot" key from the right at R - 2 - m (2M <= R-1-L,
// so the two runs never meet inside [L, R - 1)).
__device__ void sub_select_wave(uint64_t *sk, uint16_t *sl, uint16_t *part, int L, int R, int nth, int depth) {
    const int lane = threadIdx.x & 63;
    for (;;) {
        if (R - L <= 16 || depth == 0) {
            sub_sort(sk, sl, L, R - L);
            return;
        }
        --depth;
        const Med3 md = sub_median3(sk, sl, L, R);
        sub_wave_sync();
        if (lane == 0) sub_median3_store(md, sk, sl, L, R);
        sub_wave_sync();
        const uint64_t pivot = md.k[2];
        const int S = sub_count_pass(sk, pivot, L, R - 1);
        const int P = L + S;
        const int M = sub_rank_pass(sk, part, pivot, L, R, S, L, R - 1, 0);
        sub_wave_sync();
        for (int m = lane; m < M; m += 64) sub_swap(sk, sl, part[L + m], part[R - 2 - m]);
        sub_wave_sync();
        if (lane == 0) sub_swap(sk, sl, P, R - 1);
        sub_wave_sync();
        if (nth < P)
            R = P;
        else if (nth > P)
            L = P + 1;
        else
            return;
    }
}

// The same introselect by groups of G waves, one node per group, all groups
// of the work-group stepping in lockstep (work-group barriers; a group whose
// node is finished idles through the remaining steps).  Each wave of a
// group takes one contiguous G-th of [L, R - 1); a range left at most 16
// long, or at the depth limit, is sorted by the group's first wave.
template <int G>
__device__ void sub_select_groups(uint64_t *sk, uint16_t *sl, uint16_t *part, int *ws, int *ws2, int L, int R,
                                  int nth, int depth) {
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int g0 = w - w % G;  // first wave of the group
    const bool leader = (w == g0) && lane == 0;
    bool stepping = !(R - L <= 16 || depth == 0);
    bool sort_rest = !stepping;
    for (;;) {
        if (!__syncthreads_or(stepping ? 1 : 0)) break;
        Med3 md{};
        if (stepping) {
            --depth;
            md = sub_median3(sk, sl, L, R);
        }
        __syncthreads();
        if (stepping && leader) sub_median3_store(md, sk, sl, L, R);
        __syncthreads();
        const int len = R - 1 - L;
        const int b = L + (int)((int64_t)len * (w - g0) / G), e = L + (int)((int64_t)len * (w - g0 + 1) / G);
        const uint64_t pivot = md.k[2];
        if (stepping) {
            const int c = sub_count_pass(sk, pivot, b, e);
            if (lane == 0) ws[w] = c;
        }
        __syncthreads();
        int S = 0, pre = 0, P = 0;
        if (stepping) {
            for (int i = 0; i < G; ++i) {
                const int c = ws[g0 + i];
                S += c;
                pre += i < w - g0 ? c : 0;
            }
            P = L + S;
            const int m = sub_rank_pass(sk, part, pivot, L, R, S, b, e, pre);
            if (lane == 0) ws2[w] = m;
        }
        __syncthreads();
        if (stepping) {
            int M = 0;
            for (int i = 0; i < G; ++i) M += ws2[g0 + i];
            for (int m = (w - g0) * 64 + lane; m < M; m += G * 64) sub_swap(sk, sl, part[L + m], part[R - 2 - m]);
        }
        __syncthreads();
        if (stepping) {
            if (leader) sub_swap(sk, sl, P, R - 1);
            if (nth < P) {
                R = P;
            } else if (nth > P) {
                L = P + 1;
            } else {
                stepping = false;
                continue;
            }
            if (R - L <= 16 || depth == 0) {
                stepping = false;
                sort_rest = true;
            }
        }
    }
    if (sort_rest && w == g0) sub_sort(sk, sl, L, R - L);
    __syncthreads();
}

struct SubtreeArgs {
    const float *X;       // points, original order
    const int2 *range;    // (idx_start, idx_end) of every node
    const uint64_t *A;    // keys after the selects of level0 - 1 (unused when identity)
    int identity;         // level0 == 0: position p holds point p
    int level0, levels;
    float *box;
    int *perm, *posof;
    float *xp;
};

template <int D>
__device__ __forceinline__ int sub_axis(const uint32_t *mn, const uint32_t *mx) {
    float best = 0.0f;
    int ax = 0;
#pragma unroll
    for (int j = 0; j < D; ++j) {
        const float spread = o2f(mx[j]) - o2f(mn[j]);
        if (spread > best) {
            best = spread;
            ax = j;
        }
    }
    return ax;
}

// per-dimension ordered min / max over positions [b, e) of one wave's lanes,
// reduced across the wave
template <int D>
__device__ __forceinline__ void sub_bounds(const uint16_t *sl, const float *sx, int b, int e, uint32_t *mn,
                                           uint32_t *mx) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int q = 0; q < D; ++q) {
        mn[q] = 0xffffffffu;
        mx[q] = 0u;
    }
    for (int p = b + lane; p < e; p += 64) {
        const int s = sl[p];
#pragma unroll
        for (int q = 0; q < D; ++q) {
            const uint32_t o = f2o(sx[s * D + q]);
            mn[q] = min(mn[q], o);
            mx[q] = max(mx[q], o);
        }
    }
#pragma unroll
    for (int q = 0; q < D; ++q) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            mn[q] = min(mn[q], (uint32_t)__shfl_xor((int)mn[q], o, 64));
            mx[q] = max(mx[q], (uint32_t)__shfl_xor((int)mx[q], o, 64));
        }
    }
}

// Bytes of LDS the subtree build carves its arrays from: keys, rows, the
// per-wave boxes and counts, slots and exchange partners.
template <int D>
constexpr int subtree_lds_bytes() {
    return sub_cap<D>() * (8 + 4 * D + 2 + 2) + kSubW * 2 * D * 4 + 2 * kSubW * 4;
}

// The subtree rooted at node (2^level0 - 1 + k0), by the calling work-group
// of kSubNT threads, its arrays in `lds` (subtree_lds_bytes<D>() bytes,
// 16-byte aligned).
template <int D>
__device__ __forceinline__ void subtree_build(const SubtreeArgs &a, int k0, unsigned char *lds) {
    constexpr int CAP = sub_cap<D>();
    uint64_t *sk = reinterpret_cast<uint64_t *>(lds);
    float *sx = reinterpret_cast<float *>(sk + CAP);
    uint32_t *lbox = reinterpret_cast<uint32_t *>(sx + CAP * D);
    int *ws = reinterpret_cast<int *>(lbox + kSubW * 2 * D);
    int *ws2 = ws + kSubW;
    uint16_t *sl = reinterpret_cast<uint16_t *>(ws2 + kSubW);
    uint16_t *part = sl + CAP;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int root = (1 << a.level0) - 1 + k0;
    const int2 rr = a.range[root];
    const int s0 = rr.x, m = rr.y - rr.x;
    if (m > CAP) return;  // the host sizes level0 so that this never happens
    for (int p = tid; p < m; p += kSubNT) {
        const uint32_t idx = a.identity ? (uint32_t)(s0 + p) : key_idx(a.A[s0 + p]);
        sk[p] = idx;
        sl[p] = (uint16_t)p;
        const float *row = a.X + (size_t)idx * D;
#pragma unroll
        for (int j = 0; j < D; ++j) sx[p * D + j] = row[j];
    }
    __syncthreads();
    const int total = a.levels - a.level0;  // local levels; the last one holds the leaves
    // levels with fewer nodes than waves: G waves share each node's bounds
    // and keys, one wave runs its select
    for (int j = 0; j < total && j < kSubLogW; ++j) {
        const int count = 1 << j;
        const int G = kSubW >> j;
        const int k = w / G, pw = w % G;
        const int node = (1 << (a.level0 + j)) - 1 + (k0 << j) + k;
        const int2 nr = a.range[node];
        const int L = nr.x - s0, R = nr.y - s0;
        const int len = R - L;
        const int b = L + (int)((int64_t)len * pw / G), e = L + (int)((int64_t)len * (pw + 1) / G);
        for (int t = tid; t < count * 2 * D; t += kSubNT) lbox[t] = (t % (2 * D)) < D ? 0xffffffffu : 0u;
        __syncthreads();
        {
            uint32_t mn[D], mx[D];
            sub_bounds<D>(sl, sx, b, e, mn, mx);
            if (lane == 0) {
#pragma unroll
                for (int q = 0; q < D; ++q) {
                    atomicMin(lbox + k * 2 * D + q, mn[q]);
                    atomicMax(lbox + k * 2 * D + D + q, mx[q]);
                }
            }
        }
        __syncthreads();
        const bool internal = j < total - 1;
        {
            uint32_t mn[D], mx[D];
#pragma unroll
            for (int q = 0; q < D; ++q) {
                mn[q] = lbox[k * 2 * D + q];
                mx[q] = lbox[k * 2 * D + D + q];
            }
            if (pw == 0 && lane < 2 * D) a.box[(size_t)node * 2 * D + lane] = o2f(lbox[k * 2 * D + lane]);
            if (internal) {
                const int ax = sub_axis<D>(mn, mx);
                for (int p = b + lane; p < e; p += 64) sk[p] = make_key(sx[sl[p] * D + ax], key_idx(sk[p]));
            }
        }
        __syncthreads();
        if (!internal) break;
        const int depth = len > 16 ? 2 * floor_log2(len) : 0;
        if (G == 8)
            sub_select_groups<8>(sk, sl, part, ws, ws2, L, R, L + len / 2, depth);
        else if (G == 4)
            sub_select_groups<4>(sk, sl, part, ws, ws2, L, R, L + len / 2, depth);
        else
            sub_select_groups<2>(sk, sl, part, ws, ws2, L, R, L + len / 2, depth);
    }
    // one wave per node from level kSubLogW on: each wave builds its own
    // subtrees
    for (int j = kSubLogW; j < total; ++j) {
        const bool internal = j < total - 1;
        const int per_wave = 1 << (j - kSubLogW);
        for (int kk = 0; kk < per_wave; ++kk) {
            const int k = w * per_wave + kk;
            const int node = (1 << (a.level0 + j)) - 1 + (k0 << j) + k;
            const int2 nr = a.range[node];
            const int L = nr.x - s0, R = nr.y - s0;
            uint32_t mn[D], mx[D];
            sub_bounds<D>(sl, sx, L, R, mn, mx);
            if (lane == 0) {
#pragma unroll
                for (int q = 0; q < D; ++q) {
                    a.box[(size_t)node * 2 * D + q] = o2f(mn[q]);
                    a.box[(size_t)node * 2 * D + D + q] = o2f(mx[q]);
                }
            }
            if (!internal) continue;
            const int ax = sub_axis<D>(mn, mx);
            for (int p = L + lane; p < R; p += 64) sk[p] = make_key(sx[sl[p] * D + ax], key_idx(sk[p]));
            sub_wave_sync();
            const int size = R - L;
            const int depth = size > 16 ? 2 * floor_log2(size) : 0;
            sub_select_wave(sk, sl, part, L, R, L + size / 2, depth);
        }
    }
    __syncthreads();
    for (int p = tid; p < m; p += kSubNT) {
        const int idx = (int)key_idx(sk[p]);
        a.perm[s0 + p] = idx;
        a.posof[idx] = s0 + p;
        const int s = sl[p];
#pragma unroll
        for (int j = 0; j < D; ++j) a.xp[(size_t)(s0 + p) * D + j] = sx[s * D + j];
    }
}


}  // namespace

}  // namespace fh
