// Host KD-tree construction shared by the C-ABI entry points.
#pragma once

#include <cstdint>
#include <vector>

namespace fh {

struct HostTree {
    int64_t n = 0;
    int d = 0;
    int levels = 0;
    int64_t n_nodes = 0;
    std::vector<int32_t> perm;   // tree position -> original index (idx_array)
    std::vector<float> xp;       // rows in tree order, n * d
    std::vector<int32_t> start;  // idx_start per node
    std::vector<int32_t> end;    // idx_end per node
    std::vector<float> radius;   // per node (not used by queries)
    std::vector<uint8_t> leaf;   // is_leaf per node
    std::vector<float> box;      // per node: min[d] then max[d]
    bool leaves_last_level = true;
};

int tree_levels(int64_t n, int64_t leaf_size);
int build_host_tree(const float *X, int64_t n, int d, int64_t leaf_size, int nthreads, HostTree &t);

}  // namespace fh
