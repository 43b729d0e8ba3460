// Launch interface of the persistent batch kernel (fh_batch.hip): the exact
// mutual-reachability MSTs of many small datasets in one launch, one
// work-group per dataset taken from a device-side work queue.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fh {

// One dataset of a batch launch.
struct BatchJob {
    int64_t x_off;    // first float of its rows (row-major, d per row) in X
    int64_t row_off;  // first output row of its MST (the n - 1 of earlier jobs)
    int64_t pt_off;   // first per-point output slot (the n of earlier jobs)
    int32_t n;        // points (2 <= n <= batch_max_points(d))
    int32_t d;        // dimensions (1 .. 8)
    int32_t levels;   // levels of its KD-tree (tree_levels(n, 40))
    int32_t pad;
};

struct BatchArgs {
    const float *X;           // every job's rows
    const BatchJob *jobs;     // in queue order (the costliest first)
    int count;
    int ms;                   // min_samples
    int kq;                   // neighbours per query, self included: ms + 1 (2 when ms == 1)
    int *next;                // work-queue counter, zero at launch
    unsigned char *scratch;   // one slice per work-group
    int64_t slice_bytes;
    int max_n;                // largest n of the batch (sizes the slices)
    int max_nodes;            // largest tree of the batch
    int max_d;
    int *src, *dst;           // MST rows of every job in the reference's row order
    float *w;                 // their squared weights
    float *core;              // squared core distance by point
    int *status;              // per job: rows written (n - 1), or -1 when a round made no progress
    uint64_t *phase_ticks;    // optional [5]: wall-clock ticks of tree, kNN, forest, searches, rest of rounds
};

// Largest dataset (points) a work-group of the batch kernel takes in d
// dimensions; 0 when d is not handled.
int batch_max_points(int d);
// Largest dataset whose whole KD-tree one subtree build makes in LDS (the
// levels above are built node by node over global memory).
int batch_subtree_points(int d);
// Largest neighbour count per query (min_samples + 1) the kernel takes.
int batch_max_kq();
// Work-groups of the d-dimensional kernel resident on one CU (registers
// and LDS permitting); 0 when unknown.
int batch_resident_per_cu(int d);
// Scratch bytes of one work-group's slice.
int64_t batch_slice_bytes(int max_n, int max_nodes, int max_d, int kq);
// Work-groups of one launch: resident on every CU (two per CU), at most one
// per job.
int batch_grid(int count, int cus);
hipError_t launch_batch(const BatchArgs &a, int grid, hipStream_t st);

}  // namespace fh
