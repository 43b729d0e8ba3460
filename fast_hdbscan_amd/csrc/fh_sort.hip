// Device sort of MST rows by (weight, source, target): two stable radix
// passes (least significant key first) with hipCUB, then a gather.
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include "fh_sort.h"

namespace fh {

namespace {

__global__ void sort_keys1_kernel(int m, const int *__restrict__ dst, uint32_t *k1, int *v1) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    k1[i] = static_cast<uint32_t>(dst[i]) ^ 0x80000000u;  // signed order
    v1[i] = i;
}

__global__ void sort_keys2_kernel(int m, const int *__restrict__ perm1, const int *__restrict__ src,
                                  const float *__restrict__ w, uint64_t *k2) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= m) return;
    const int i = perm1[j];
    // weights are >= 0 (or +inf): their bit patterns order like the values
    k2[j] = (static_cast<uint64_t>(__float_as_uint(w[i])) << 32) |
            (static_cast<uint32_t>(src[i]) ^ 0x80000000u);
}

__global__ void sort_gather_kernel(int m, const int *__restrict__ perm, const int *__restrict__ src,
                                   const int *__restrict__ dst, const float *__restrict__ w, int *so, int *dso,
                                   float *wo) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= m) return;
    const int i = perm[j];
    so[j] = src[i];
    dso[j] = dst[i];
    wo[j] = w[i];
}

inline unsigned nb(int m) { return (unsigned)((m + 255) / 256); }

__global__ void knn_core64_kernel(int64_t n, int kq, int mode, int ms, const float *__restrict__ knn_d,
                                  const float *__restrict__ core32, double *core64) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double c = 0.0;
    if (mode == 0) c = sqrt(static_cast<double>(knn_d[i * kq + ms]));
    else if (mode == 2) c = static_cast<double>(core32[i]);
    core64[i] = c;
}

// kruskal.py:483-510: w = max(cd_i, cd_j, d) by the same comparisons; a -1
// neighbour becomes an infinite self-loop, which the caller's filter drops
// (kruskal.py:1017-1021).  Dropped edges get the largest key so the sorted
// valid edges come first, in stable order.
__global__ void knn_edge_keys_kernel(int64_t n, int kq, int mode, const float *__restrict__ knn_d,
                                     const int *__restrict__ knn_i, const double *__restrict__ core64,
                                     uint64_t *keys, int *vals, int *valid_count) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int kc = kq - 1;
    if (e >= n * kc) return;
    const int64_t i = e / kc;
    const int t = (int)(e % kc) + 1;
    const int j = knn_i[i * kq + t];
    uint64_t key = ~0ull;
    if (j >= 0 && j != i) {
        const float dv = knn_d[i * kq + t];
        const double d = mode == 2 ? static_cast<double>(dv) : sqrt(static_cast<double>(dv));
        double w = core64[i];
        if (core64[j] > w) w = core64[j];
        if (d > w) w = d;
        if (isfinite(w)) {
            key = static_cast<uint64_t>(__double_as_longlong(w));  // w >= 0: bits order like values
            atomicAdd(valid_count, 1);
        }
    }
    keys[e] = key;
    vals[e] = (int)e;
}

__global__ void knn_edge_gather_kernel(int64_t m, int kq, const int *__restrict__ knn_i,
                                       const uint64_t *__restrict__ keys_s, const int *__restrict__ vals_s, int *so,
                                       int *dso, double *wo) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= m) return;
    const int64_t e = vals_s[r];
    const int kc = kq - 1;
    const int64_t i = e / kc;
    const int t = (int)(e % kc) + 1;
    so[r] = (int)i;
    dso[r] = knn_i[i * kq + t];
    wo[r] = __longlong_as_double(static_cast<long long>(keys_s[r]));
}

}  // namespace

size_t sort_u32_pairs_temp_bytes(int m) {
    size_t b = 0;
    hipError_t e = hipcub::DeviceRadixSort::SortPairs(nullptr, b, static_cast<uint32_t *>(nullptr),
                                                      static_cast<uint32_t *>(nullptr), static_cast<int *>(nullptr),
                                                      static_cast<int *>(nullptr), m);
    return e == hipSuccess ? b : 0;
}

hipError_t sort_u32_pairs(int m, const uint32_t *keys, uint32_t *keys_s, const int *vals, int *vals_s, void *temp,
                          size_t temp_bytes, hipStream_t st) {
    if (m <= 0) return hipSuccess;
    size_t tb = temp_bytes;
    return hipcub::DeviceRadixSort::SortPairs(temp, tb, keys, keys_s, vals, vals_s, m, 0, 32, st);
}

size_t knn_edges_temp_bytes(int64_t m) {
    size_t b = 0;
    hipError_t e = hipcub::DeviceRadixSort::SortPairs(nullptr, b, static_cast<uint64_t *>(nullptr),
                                                      static_cast<uint64_t *>(nullptr), static_cast<int *>(nullptr),
                                                      static_cast<int *>(nullptr), (int)m);
    return e == hipSuccess ? b : 0;
}

hipError_t knn_edges_sorted(const KnnEdgesArgs &a, hipStream_t st) {
    const int64_t m = a.n * (a.kq - 1);
    hipLaunchKernelGGL(knn_core64_kernel, dim3((unsigned)((a.n + 255) / 256)), dim3(256), 0, st, a.n, a.kq, a.mode,
                       a.ms, a.knn_d, a.core32, a.core64);
    if (m == 0) return hipGetLastError();
    const unsigned g = (unsigned)((m + 255) / 256);
    hipLaunchKernelGGL(knn_edge_keys_kernel, dim3(g), dim3(256), 0, st, a.n, a.kq, a.mode, a.knn_d, a.knn_i,
                       a.core64, a.keys, a.vals, a.valid_count);
    size_t tb = a.temp_bytes;
    hipError_t e = hipcub::DeviceRadixSort::SortPairs(a.temp, tb, a.keys, a.keys_s, a.vals, a.vals_s, (int)m, 0, 64, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(knn_edge_gather_kernel, dim3(g), dim3(256), 0, st, m, a.kq, a.knn_i, a.keys_s, a.vals_s,
                       a.src_out, a.dst_out, a.w_out);
    return hipGetLastError();
}

size_t sort_mst_temp_bytes(int m) {
    size_t a = 0, b = 0;
    hipError_t e1 = hipcub::DeviceRadixSort::SortPairs(nullptr, a, static_cast<uint32_t *>(nullptr), static_cast<uint32_t *>(nullptr),
                                       static_cast<int *>(nullptr), static_cast<int *>(nullptr), m);
    hipError_t e2 = hipcub::DeviceRadixSort::SortPairs(nullptr, b, static_cast<uint64_t *>(nullptr), static_cast<uint64_t *>(nullptr),
                                       static_cast<int *>(nullptr), static_cast<int *>(nullptr), m);
    if (e1 != hipSuccess || e2 != hipSuccess) return 0;
    return a > b ? a : b;
}

hipError_t sort_mst(const SortMstArgs &a, hipStream_t st) {
    const int m = a.m;
    if (m == 0) return hipSuccess;
    size_t tb = a.temp_bytes;
    hipLaunchKernelGGL(sort_keys1_kernel, dim3(nb(m)), dim3(256), 0, st, m, a.dst, a.k1, a.v1);
    hipError_t e = hipcub::DeviceRadixSort::SortPairs(a.temp, tb, a.k1, a.k1s, a.v1, a.v1s, m, 0, 32, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(sort_keys2_kernel, dim3(nb(m)), dim3(256), 0, st, m, a.v1s, a.src, a.w, a.k2);
    tb = a.temp_bytes;
    e = hipcub::DeviceRadixSort::SortPairs(a.temp, tb, a.k2, a.k2s, a.v1s, a.v1, m, 0, 64, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(sort_gather_kernel, dim3(nb(m)), dim3(256), 0, st, m, a.v1, a.src, a.dst, a.w, a.src_out,
                       a.dst_out, a.w_out);
    return hipGetLastError();
}

namespace {
__global__ void tree_positions_kernel(int m, const int *__restrict__ src, const int *__restrict__ dst,
                                      const int *__restrict__ posof, int *pos_src, int *pos_dst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const int s = src[i], t = dst[i];
    pos_src[i] = s < 0 ? s : posof[s];
    pos_dst[i] = t < 0 ? t : posof[t];
}
}  // namespace

hipError_t tree_positions(int m, const int *src, const int *dst, const int *posof, int *pos_src, int *pos_dst,
                          hipStream_t st) {
    if (m <= 0) return hipSuccess;
    hipLaunchKernelGGL(tree_positions_kernel, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, st, m, src, dst, posof,
                       pos_src, pos_dst);
    return hipGetLastError();
}

namespace {
__device__ __forceinline__ int pc_find(int *parent, int x) {
    while (parent[x] != x) {
        const int g = parent[parent[x]];
        parent[x] = g;  // halving; races only shorten paths to the same root
        x = g;
    }
    return x;
}

__global__ void pc_init_kernel(int n, int *parent) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) parent[i] = i;
}

// lock-free union to the smaller label
__global__ void pc_unite_kernel(int k, const int *__restrict__ src, const int *__restrict__ dst, int *parent) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= k) return;
    int a = src[i], b = dst[i];
    for (;;) {
        a = pc_find(parent, a);
        b = pc_find(parent, b);
        if (a == b) return;
        if (a < b) {
            const int t = a;
            a = b;
            b = t;
        }
        if (atomicCAS(parent + a, a, b) == a) return;
    }
}

__global__ void pc_keys_kernel(int k, const int *__restrict__ src, int *parent, uint32_t *keys, int *vals) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= k) return;
    keys[i] = static_cast<uint32_t>(pc_find(parent, src[i]));
    vals[i] = i;
}
}  // namespace

size_t prefix_components_temp_bytes(int k) {
    size_t a = 0, b = 0;
    hipError_t e1 = hipcub::DeviceRadixSort::SortPairs(nullptr, a, static_cast<uint32_t *>(nullptr),
                                                       static_cast<uint32_t *>(nullptr), static_cast<int *>(nullptr),
                                                       static_cast<int *>(nullptr), k, 0, 32);
    hipError_t e2 = hipcub::DeviceRunLengthEncode::Encode(nullptr, b, static_cast<uint32_t *>(nullptr),
                                                          static_cast<uint32_t *>(nullptr), static_cast<int *>(nullptr),
                                                          static_cast<int *>(nullptr), k);
    if (e1 != hipSuccess || e2 != hipSuccess) return 0;
    return a > b ? a : b;
}

hipError_t prefix_components(const PrefixComponentsArgs &a, hipStream_t st) {
    if (a.k <= 0) return hipSuccess;
    hipLaunchKernelGGL(pc_init_kernel, dim3(nb(a.n)), dim3(256), 0, st, a.n, a.parent);
    hipLaunchKernelGGL(pc_unite_kernel, dim3(nb(a.k)), dim3(256), 0, st, a.k, a.src, a.dst, a.parent);
    hipLaunchKernelGGL(pc_keys_kernel, dim3(nb(a.k)), dim3(256), 0, st, a.k, a.src, a.parent, a.keys, a.vals);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    int bits = 1;
    while (bits < 32 && (1u << bits) < static_cast<uint32_t>(a.n)) ++bits;
    size_t tb = a.temp_bytes;
    e = hipcub::DeviceRadixSort::SortPairs(a.temp, tb, a.keys, a.keys_s, a.vals, a.vals_s, a.k, 0, bits, st);
    if (e != hipSuccess) return e;
    tb = a.temp_bytes;
    return hipcub::DeviceRunLengthEncode::Encode(a.temp, tb, a.keys_s, a.keys, a.counts, a.n_runs, a.k, st);
}

}  // namespace fh
