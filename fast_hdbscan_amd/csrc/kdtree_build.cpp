// Host construction of the flat KD-tree that the gfx950 kernels traverse.
//
// The tree must be exactly numba_kdtree.build_kdtree(X, leaf_size=40)
// (numba_kdtree.py:478-521): the same implicit complete binary tree, the
// same split dimension (first dimension of strictly largest spread,
// numba_kdtree.py:133-152) and the same median selection by introselect on
// (value, original index) keys (numba_kdtree.py:164-407).  The *order of
// points inside each leaf* is a by-product of the introselect swap sequence
// and decides tie-breaking downstream (kNN heap pushes, Boruvka partner
// choice), so the selection below performs the reference's swaps one for
// one.  What differs is the data layout: the rows are carried along with
// the permutation (tree order, row-major), so every node is a contiguous
// block of rows that can be copied to HBM as-is and every comparison reads
// memory sequentially instead of through the index indirection.
//
// Independent subtrees are built on separate host threads; the result does
// not depend on the thread count.
#include "kdtree_build.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

namespace fh {

int tree_levels(int64_t n, int64_t leaf_size) {
    // numba_kdtree.py:489: int(log2(max(1, (n-1)/leaf_size)) + 1), evaluated
    // in integers: 1 + max{m >= 0 : leaf_size * 2^m <= n - 1}.
    if (n - 1 < leaf_size) return 1;
    int m = 0;
    while ((leaf_size << (m + 1)) <= n - 1) ++m;
    return m + 1;
}

namespace {

struct Rows {
    float *x;       // tree-order rows, n * d
    int32_t *perm;  // original index of each row
    int d;

    // numba_kdtree.py:164-179: order by value on `axis`, ties by original index.
    inline bool less(int64_t p, int64_t q, int axis) const {
        float a = x[p * d + axis], b = x[q * d + axis];
        if (a != b) return a < b;
        return perm[p] < perm[q];
    }
    inline bool greater(int64_t p, int64_t q, int axis) const { return less(q, p, axis); }
    inline void swap(int64_t p, int64_t q) {
        if (p == q) return;
        std::swap(perm[p], perm[q]);
        float *a = x + p * d, *b = x + q * d;
        for (int j = 0; j < d; ++j) std::swap(a[j], b[j]);
    }
    // Comparison against a saved pivot key (value, original index).
    inline bool less_key(int64_t p, float v, int32_t id, int axis) const {
        float a = x[p * d + axis];
        if (a != v) return a < v;
        return perm[p] < id;
    }
};

// numba_kdtree.py:192-201, insertion sort of [left, right).
void insertion_sort(Rows &r, int axis, int64_t left, int64_t right) {
    std::vector<float> row(r.d);
    for (int64_t i = left + 1; i < right; ++i) {
        int32_t key_id = r.perm[i];
        std::memcpy(row.data(), r.x + i * r.d, sizeof(float) * r.d);
        float key_v = row[axis];
        int64_t j = i - 1;
        // shift while element j compares greater than the key
        while (j >= left) {
            float a = r.x[j * r.d + axis];
            bool gt = (a > key_v) || (a == key_v && r.perm[j] > key_id);
            if (!gt) break;
            r.perm[j + 1] = r.perm[j];
            std::memcpy(r.x + (j + 1) * r.d, r.x + j * r.d, sizeof(float) * r.d);
            --j;
        }
        r.perm[j + 1] = key_id;
        std::memcpy(r.x + (j + 1) * r.d, row.data(), sizeof(float) * r.d);
    }
}

// numba_kdtree.py:214-266, heapsort of [left, right) (depth-limit fallback).
void sift_down(Rows &r, int axis, int64_t off, int64_t start, int64_t end) {
    int64_t root = start;
    while (root * 2 + 1 < end) {
        int64_t child = root * 2 + 1, s = root;
        if (r.less(off + s, off + child, axis)) s = child;
        if (child + 1 < end && r.less(off + s, off + child + 1, axis)) s = child + 1;
        if (s == root) return;
        r.swap(off + root, off + s);
        root = s;
    }
}

void heap_sort(Rows &r, int axis, int64_t left, int64_t right) {
    int64_t size = right - left;
    for (int64_t i = size / 2 - 1; i >= 0; --i) sift_down(r, axis, left, i, size);
    for (int64_t i = size - 1; i > 0; --i) {
        r.swap(left, left + i);
        sift_down(r, axis, left, 0, i);
    }
}

// numba_kdtree.py:280-299
int64_t median_of_three(Rows &r, int axis, int64_t left, int64_t right) {
    int64_t mid = (left + right - 1) / 2;
    if (r.greater(left, mid, axis)) r.swap(left, mid);
    if (r.greater(mid, right - 1, axis)) {
        r.swap(mid, right - 1);
        if (r.greater(left, mid, axis)) r.swap(left, mid);
    }
    return mid;
}

// numba_kdtree.py:313-350 (pivot moved to right-1, two-pointer exchange).
int64_t partition(Rows &r, int axis, int64_t left, int64_t right, int64_t pivot) {
    r.swap(pivot, right - 1);
    const float pv = r.x[(right - 1) * r.d + axis];
    const int32_t pid = r.perm[right - 1];
    int64_t i = left, j = right - 2;
    for (;;) {
        while (i <= j && r.less_key(i, pv, pid, axis)) ++i;
        while (i <= j && !r.less_key(j, pv, pid, axis)) --j;
        if (i >= j) break;
        r.swap(i, j);
        ++i;
        --j;
    }
    r.swap(i, right - 1);
    return i;
}

inline int64_t floor_log2(int64_t v) { return 63 - __builtin_clzll(static_cast<unsigned long long>(v)); }

// numba_kdtree.py:361-407
void introselect(Rows &r, int axis, int64_t left, int64_t right, int64_t nth) {
    if (right - left <= 16) {
        insertion_sort(r, axis, left, right);
        return;
    }
    int64_t depth = 2 * floor_log2(right - left);
    while (right - left > 16) {
        if (depth == 0) {
            heap_sort(r, axis, left, right);
            return;
        }
        --depth;
        int64_t pp = partition(r, axis, left, right, median_of_three(r, axis, left, right));
        if (nth < pp)
            right = pp;
        else if (nth > pp)
            left = pp + 1;
        else
            return;
    }
    insertion_sort(r, axis, left, right);
}

struct Builder {
    Rows rows;
    HostTree *t;

    // numba_kdtree.py:79-115 (bounds, radius) and 133-152 (split dimension:
    // the spread max-min of the same rows, first strictly largest wins).
    int init_node(int64_t node, int64_t s, int64_t e) {
        const int d = rows.d;
        float *lo = t->box.data() + node * 2 * d;
        float *hi = lo + d;
        for (int j = 0; j < d; ++j) {
            lo[j] = INFINITY;
            hi[j] = -INFINITY;
        }
        for (int64_t i = s; i < e; ++i) {
            const float *row = rows.x + i * d;
            for (int j = 0; j < d; ++j) {
                lo[j] = std::min(lo[j], row[j]);
                hi[j] = std::max(hi[j], row[j]);
            }
        }
        float rad = 0.0f;
        int axis = 0;
        float best = 0.0f;
        for (int j = 0; j < d; ++j) {
            float half = std::fabs(hi[j] - lo[j]) * 0.5f;
            rad += half * half;
            float spread = hi[j] - lo[j];
            if (spread > best) {
                best = spread;
                axis = j;
            }
        }
        t->start[node] = static_cast<int32_t>(s);
        t->end[node] = static_cast<int32_t>(e);
        t->radius[node] = std::sqrt(rad);
        return axis;
    }

    // numba_kdtree.py:414-475; subtrees below `par_depth` levels of
    // remaining parallelism are handed to new threads.
    void build(int64_t s, int64_t e, int64_t node, int par_depth) {
        int axis = init_node(node, s, e);
        if (2 * node + 1 >= t->n_nodes || e - s < 2) {
            t->leaf[node] = 1;
            return;
        }
        t->leaf[node] = 0;
        int64_t mid = s + (e - s) / 2;
        introselect(rows, axis, s, e, mid);
        if (par_depth > 0) {
            std::thread th([&] { build(s, mid, 2 * node + 1, par_depth - 1); });
            build(mid, e, 2 * node + 2, par_depth - 1);
            th.join();
        } else {
            build(s, mid, 2 * node + 1, 0);
            build(mid, e, 2 * node + 2, 0);
        }
    }
};

}  // namespace

int build_host_tree(const float *X, int64_t n, int d, int64_t leaf_size, int nthreads, HostTree &t) {
    if (n < 1 || d < 1 || leaf_size < 1) return -1;
    t.n = n;
    t.d = d;
    t.levels = tree_levels(n, leaf_size);
    t.n_nodes = (int64_t(1) << t.levels) - 1;
    t.perm.resize(n);
    t.xp.resize(static_cast<size_t>(n) * d);
    t.start.assign(t.n_nodes, 0);
    t.end.assign(t.n_nodes, 0);
    t.radius.assign(t.n_nodes, 0.0f);
    t.leaf.assign(t.n_nodes, 0);
    t.box.assign(static_cast<size_t>(t.n_nodes) * 2 * d, 0.0f);
    for (int64_t i = 0; i < n; ++i) t.perm[i] = static_cast<int32_t>(i);
    std::memcpy(t.xp.data(), X, sizeof(float) * static_cast<size_t>(n) * d);

    int par = 0;
    if (nthreads < 1) nthreads = static_cast<int>(std::max(1u, std::thread::hardware_concurrency()));
    while ((1 << (par + 1)) <= nthreads && par + 1 < t.levels) ++par;
    if (n < 65536) par = 0;

    Builder b{Rows{t.xp.data(), t.perm.data(), d}, &t};
    b.build(0, n, 0, par);

    // Every leaf of this implicit tree sits on the last level whenever all
    // upper-level nodes hold >= 2 points (always true for leaf_size 40 and
    // n >= 2); the device kernels rely on it.
    t.leaves_last_level = true;
    int64_t first_last = (t.n_nodes - 1) / 2;
    for (int64_t i = 0; i < first_last; ++i)
        if (t.leaf[i]) t.leaves_last_level = false;
    return 0;
}

}  // namespace fh
