// Launch interface of the exact device KD-tree build (fh_tree.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fh {

enum SegFlag : int {
    SEG_LARGE = 0,  // still running multi-block introselect steps
    SEG_SMALL = 1,  // to be finished in LDS by one work-group
    SEG_SORT = 2,   // hit the depth limit while large: range must be sorted
    SEG_DONE = 3,
};

// Introselect state of one node (numba_kdtree.py:361-386 locals).
struct SegState {
    int L, R, nth, depth, flag;
    uint64_t pivot;
};

struct TreeBuildArgs {
    const float *X;      // points, original order, row-major
    int d;
    const int2 *range;   // (idx_start, idx_end) of every node
    uint64_t *A;         // n keys: (ordered coordinate << 32) | original index
    uint32_t *obox;      // N * 2d order-preserving bounds scratch
    float *box;          // N * 2d bounds (min[d], max[d]) per node
    int *axis;           // per node of the current level
    SegState *seg;       // per node of the current level
    int *cnt;            // per (node, chunk, wave) counts
    int *gtpos, *ltpos;  // n each: exchange partners (the multi-block and one-work-group steps use gtpos)
    int *aux;            // 4 per node of the level: (S, P, M, unused) of the current step
    int *active;         // step counters
    int *sorts;          // nodes of the level left to the sort (depth limit)
};

int tree_small_limit();
int tree_medium_limit();
int tree_chunk();
hipError_t tree_identity_keys(int n, uint64_t *A, hipStream_t st);
hipError_t tree_level_bounds(const TreeBuildArgs &a, int level, int max_size, int identity, hipStream_t st);
hipError_t tree_level_axis(const TreeBuildArgs &a, int level, int internal, hipStream_t st);
hipError_t tree_level_keys(const TreeBuildArgs &a, int level, int max_size, int identity, hipStream_t st);
// multi-block introselect steps of the SEG_LARGE nodes of a level: begin
// chooses each one's first pivot; every step partitions around the pending
// pivot, narrows, and chooses the next pivot (counted in *active), so the
// steps may stop once a step counts no pivot
hipError_t tree_big_begin(const TreeBuildArgs &a, int count, hipStream_t st);
hipError_t tree_big_step(const TreeBuildArgs &a, int count, int max_size, hipStream_t st);
// steps of every SEG_LARGE node by one work-group each; leaves SEG_SMALL or
// SEG_DONE (for levels whose nodes are at most tree_medium_limit() long)
hipError_t tree_medium_select(const TreeBuildArgs &a, int count, hipStream_t st);
// max_size: the level's largest node (the segments left are no longer)
hipError_t tree_small_select(const TreeBuildArgs &a, int count, int max_size, hipStream_t st);
hipError_t tree_sort_range(uint64_t *A, int L, int R, hipStream_t st);
// Largest subtree root (points) that one work-group builds in LDS for d
// dimensions; 0 when d has no such kernel.
int tree_subtree_cap(int d);
// Every level from level0 to levels - 1 of the 2^level0 subtrees rooted at
// level0 (each at most tree_subtree_cap(d) points), one work-group each,
// then the tree-order outputs (as tree_finalize); a.A holds the keys left
// by the selects of level0 - 1 (not read when level0 == 0).
hipError_t tree_subtrees(const TreeBuildArgs &a, int level0, int levels, int *perm, int *posof, float *xp,
                         hipStream_t st);
hipError_t tree_finalize(const TreeBuildArgs &a, int n, int *perm, int *posof, float *xp, hipStream_t st);

}  // namespace fh
