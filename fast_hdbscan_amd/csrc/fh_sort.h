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

// Mutual-reachability edge list of a kNN graph, globally sorted by weight
// (kruskal.py:455-512 + the stable argsort at 1023): kNN rows by original
// index, columns 1..kq-1.  mode 0: core = sqrt(float64(rdist column ms));
// 1: core = 0; 2: core = float64(core32), distances already square roots.
struct KnnEdgesArgs {
    int64_t n;
    int kq;
    int mode;
    int ms;
    const float *knn_d;
    const int *knn_i;
    const float *core32;
    double *core64;     // [n] out
    uint64_t *keys;     // [m] scratch, m = n * (kq - 1)
    uint64_t *keys_s;   // [m] sorted keys
    int *vals;          // [m] scratch
    int *vals_s;        // [m] sorted edge ids
    int *valid_count;   // [1] out
    int *src_out;       // [m] sorted rows (first valid_count meaningful)
    int *dst_out;
    double *w_out;
    void *temp;
    size_t temp_bytes;
};
size_t knn_edges_temp_bytes(int64_t m);

// Connected components of the forest of rows [0, k) (non-negative
// endpoints < n): vals_s = row indices grouped by component, increasing
// within a group (stable radix sort on the component's root); counts [runs]
// = group sizes in the same order, n_runs [1].  keys is reused as the run
// keys' output.
struct PrefixComponentsArgs {
    int n, k;
    const int *src, *dst;
    int *parent;        // [n] scratch
    uint32_t *keys;     // [k] scratch
    uint32_t *keys_s;   // [k] scratch
    int *vals;          // [k] scratch
    int *vals_s;        // [k] out
    int *counts;        // [k] out
    int *n_runs;        // [1] out
    void *temp;
    size_t temp_bytes;
};
size_t prefix_components_temp_bytes(int k);
hipError_t prefix_components(const PrefixComponentsArgs &a, hipStream_t st);

// Stable sort of m (32-bit key, int value) pairs by key.
size_t sort_u32_pairs_temp_bytes(int m);
hipError_t sort_u32_pairs(int m, const uint32_t *keys, uint32_t *keys_s, const int *vals, int *vals_s, void *temp,
                          size_t temp_bytes, hipStream_t st);

// pos_src[i] = posof[src[i]], pos_dst[i] = posof[dst[i]] (negative ids kept)
hipError_t tree_positions(int m, const int *src, const int *dst, const int *posof, int *pos_src, int *pos_dst,
                          hipStream_t st);
hipError_t knn_edges_sorted(const KnnEdgesArgs &a, hipStream_t st);

}  // namespace fh
