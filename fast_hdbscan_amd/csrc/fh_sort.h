// Device sort of MST rows (fh_sort.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace fh {

struct SortMstArgs {
    int m;
    const int *src;
    const int *dst;
    const float *w;
    int *src_out;
    int *dst_out;
    float *w_out;
    // scratch
    uint32_t *k1, *k1s;
    uint64_t *k2, *k2s;
    int *v1, *v1s;
    void *temp;
    size_t temp_bytes;
};

size_t sort_mst_temp_bytes(int m);
hipError_t sort_mst(const SortMstArgs &a, hipStream_t st);

}  // namespace fh
