// Device-side building blocks shared by the gfx950 kernels: the reference's
// float32 distance arithmetic, the exact bounded max-heap, and the
// stackless depth-first walk of the implicit KD-tree.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fh {

// Squared Euclidean distance in the reference's canonical float32 form:
// sequential fused multiply-add accumulation over dimensions
// (numba_kdtree.py:539-546 compiled with fastmath on an FMA host).
// D > 0: compile-time dimension with the query held in registers;
// D == 0: runtime dimension d, query read from memory.
template <int D>
__device__ __forceinline__ float rdist(const float *__restrict__ x, const float *__restrict__ y, int d) {
    float r = 0.0f;
    if constexpr (D > 0) {
#pragma unroll
        for (int j = 0; j < D; ++j) {
            float df = x[j] - y[j];
            r = __builtin_fmaf(df, df, r);
        }
    } else {
        for (int j = 0; j < d; ++j) {
            float df = x[j] - y[j];
            r = __builtin_fmaf(df, df, r);
        }
    }
    return r;
}

// Squared distance from a point to an axis-aligned box (lo[d] then hi[d]),
// numba_kdtree.py:566-575.
template <int D>
__device__ __forceinline__ float box_lb(const float *__restrict__ x, const float *__restrict__ b, int d) {
    const int dd = D > 0 ? D : d;
    float r = 0.0f;
#pragma unroll
    for (int j = 0; j < dd; ++j) {
        float lo = b[j], hi = b[dd + j], p = x[j];
        float dlo = lo > p ? lo - p : 0.0f;
        float dhi = p > hi ? p - hi : 0.0f;
        float s = dlo + dhi;
        r = __builtin_fmaf(s, s, r);
    }
    return r;
}

// Relaxed device-scope atomic load (bypasses the per-CU vector cache so a
// spinning reader sees other work-groups' stores).
template <typename T>
__device__ __forceinline__ T ld_relaxed(const T *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Bounded max-heap of K (priority, index) pairs held in LDS with a stride so
// that lane t's slot j lives at j*stride + t (no bank conflicts).
// Push and sort are numba_kdtree.py:593-635 and 647-677 step for step, so
// equal priorities end up in the reference's order.
struct LdsHeap {
    float *p;
    int *i;
    int stride;
    int k;

    __device__ __forceinline__ float &P(int j) const { return p[j * stride]; }
    __device__ __forceinline__ int &I(int j) const { return i[j * stride]; }

    __device__ void init() const {
        for (int j = 0; j < k; ++j) {
            P(j) = __builtin_huge_valf();
            I(j) = -1;
        }
    }

    // Caller has checked pr < P(0).
    __device__ void push(float pr, int id) const {
        int at = 0;
        for (;;) {
            int c1 = 2 * at + 1, c2 = c1 + 1, sw;
            if (c1 >= k) break;
            if (c2 >= k) {
                if (P(c1) > pr) sw = c1;
                else break;
            } else {
                float a = P(c1), b = P(c2);
                if (a >= b) {
                    if (pr < a) sw = c1;
                    else break;
                } else {
                    if (pr < b) sw = c2;
                    else break;
                }
            }
            P(at) = P(sw);
            I(at) = I(sw);
            at = sw;
        }
        P(at) = pr;
        I(at) = id;
    }

    __device__ void sift(int len, int elt) const {
        while (elt * 2 + 1 < len) {
            int l = elt * 2 + 1, r = l + 1, s = elt;
            if (P(s) < P(l)) s = l;
            if (r < len && P(s) < P(r)) s = r;
            if (s == elt) break;
            float tp = P(elt);
            P(elt) = P(s);
            P(s) = tp;
            int ti = I(elt);
            I(elt) = I(s);
            I(s) = ti;
            elt = s;
        }
    }

    __device__ void sort() const {
        for (int j = k - 1; j > 0; --j) {
            float tp = P(0);
            P(0) = P(j);
            P(j) = tp;
            int ti = I(0);
            I(0) = I(j);
            I(j) = ti;
            sift(j, 0);
        }
    }
};

}  // namespace fh
