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

}  // namespace

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

}  // namespace fh
