// gfx950 kernels of the Euclidean mutual-reachability MST.
//
// Data layout in HBM (one tree per call):
//   xp[n*d]        points in KD-tree order (row-major); a leaf is a
//                  contiguous run of rows, so one wavefront whose lanes sit
//                  in the same leaf reads the same cache lines
//   perm[n]        tree position -> original index   (numba idx_array)
//   posof[n]       original index -> tree position
//   range[N]       (idx_start, idx_end) per node
//   box[N*2d]      per node min[d] then max[d]       (numba node_bounds)
//   core_p[n]      squared core distance by tree position
//   comp_p[n]      component label by tree position; the label of a
//                  component is its lowest original point index
//   node_comp[N]   label of a node whose points are all in one component,
//                  else -1                          (boruvka.py:98-124)
//   bound[n], best[n]  per-label pruning bound and packed best edge
//
// All leaves of the implicit tree are on its last level (checked on the
// host), so "leaf" == "depth == levels-1" inside the walks.
#include "fh_kernels.h"

#include "fh_device.h"
#include "fh_search.h"

namespace fh {

namespace {

// (the walk and visitors: fh_search.h)

template <int D, bool GLOBAL_HEAP>
__global__ __launch_bounds__(256) void knn_kernel(KnnArgs a) {
    extern __shared__ unsigned char smem_raw[];
    const int tid = threadIdx.x;
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + tid;
    if (q >= a.nq) return;
    const int k = a.k;
    LdsHeap h;
    if constexpr (!GLOBAL_HEAP) {
        float *sp = reinterpret_cast<float *>(smem_raw);
        h.p = sp + tid;
        h.i = reinterpret_cast<int *>(sp + (size_t)k * blockDim.x) + tid;
        h.stride = blockDim.x;
    } else {
        h.p = a.heap_p + q;
        h.i = a.heap_i + q;
        h.stride = (int)a.heap_stride;
    }
    h.k = k;
    h.init();

    const int dd = D > 0 ? D : a.d;
    // a listed tree position, or query row q
    const int64_t qp = a.query_pos ? a.query_pos[q] : q;
    KnnVisitor<D> v;
    v.x.load(a.query_pos ? a.xp + (size_t)qp * dd : a.queries + (size_t)q * dd, a.d);
    v.xp = a.xp;
    v.range = a.range;
    v.h = h;
    v.bound = a.bound_sq ? a.bound_sq[qp] : kInf;
    v.hmax = v.bound;
    v.tie = 0;
    v.d = a.d;
    walk<D>(a.box, a.levels, a.d, v);
    if (a.bound_sq) {
        bool miss = !(h.P(0) < kInf) || v.tie;
        if (!miss) {
            h.sort();
            for (int j = 0; j + 1 < k; ++j)
                if (h.P(j) == h.P(j + 1)) miss = true;
        }
        if (miss) {
            const int slot = atomicAdd(a.miss_count, 1);
            a.miss_list[slot] = (int)qp;
            return;
        }
    } else if (a.sqrt_dist) {
        // numba_kdtree.py:792: np.sqrt of the rows, then deheap_sort
        for (int j = 0; j < k; ++j) h.P(j) = sqrtf(h.P(j));
    }
    if (!a.bound_sq) h.sort();

    // Output row: the original index of the query for the self-query, the
    // query's own row otherwise.
    const int64_t row = a.query_pos ? a.perm[qp] : (a.query_perm ? a.query_perm[q] : q);
    for (int j = 0; j < k; ++j) {
        const int pj = h.I(j);
        if (a.out_d != nullptr) a.out_d[row * k + j] = h.P(j);
        if (a.out_i != nullptr) a.out_i[row * k + j] = pj >= 0 ? a.perm[pj] : -1;
    }
    if (a.nb_pos != nullptr) {
        for (int j = 1; j < k; ++j) {
            a.nb_pos[qp * (k - 1) + j - 1] = h.I(j);
            a.nb_d[qp * (k - 1) + j - 1] = h.P(j);
        }
    }
    if (a.core_o != nullptr || a.core_p != nullptr) {
        const float c = a.core_from_knn ? h.P(k - 1) : 0.0f;
        if (a.core_o != nullptr) a.core_o[row] = c;
        if (a.core_p != nullptr) a.core_p[qp] = c;
    }
}

// PASS 0: every point; 1: one position per group of sample_mask+1;
// 2: the others, except those inside certified subtrees.
template <int D, int PASS>
__global__ __launch_bounds__(256) void boruvka_query_kernel(QueryArgs a) {
    const int64_t pos = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (pos >= a.n) return;
    if constexpr (PASS != 0) {
        const bool sampled = (pos & a.sample_mask) == (a.sample_mask >> 1);
        if ((PASS == 1) != sampled) return;
        if (PASS == 2 && a.cert && a.cert[a.cert_of_pos[pos]]) return;
    }
    const int dd = D > 0 ? D : a.d;
    const int comp = a.comp_p[pos];
    BoruvkaVisitor<D> v;
    v.x.load(a.xp + (size_t)pos * dd, a.d);
    v.xp = a.xp;
    v.range = a.range;
    v.core_p = a.core_p;
    v.comp_p = a.comp_p;
    v.node_comp = a.node_comp;
    v.bound_c = a.bound + comp;
    v.core_i = a.core_p[pos];
    v.comp = comp;
    v.d = a.d;
    v.best = kInf;
    v.best_j = -1;
    v.B = ld_relaxed(v.bound_c);
    walk<D>(a.box, a.levels, a.d, v);
    if (v.best_j >= 0) {
        // Positive floats order like their bit patterns: the packed key
        // gives the smallest weight, then the lowest original point index
        // (merge_components keeps the first index at the minimum,
        // boruvka.py:43-59).
        const unsigned long long key =
            ((unsigned long long)__float_as_uint(v.best) << 32) | (unsigned int)a.perm[pos];
        atomicMin(a.best + comp, key);
        a.partner[pos] = v.best_j;
    }
}

// ------------------------------------------------ bound tightening ----
//
// The per-component bound only has to stay >= the component's minimum
// outgoing weight for the searches to remain exact (every prune compares
// strictly against it), so any real cross-component edge may lower it.

// Cross-component edges among the kNN lists.  The distances are the kNN
// kernel's rdist of the same two rows, bit-identical to what the search
// computes for that pair.  Components only merge, so a point whose
// neighbours all share its component keeps that property for the rest of
// the call: it is marked interior (interior[p] = 1) and skipped from then
// on.
__global__ void seed_bounds_kernel(int64_t n, int kn, const int *__restrict__ nb_pos, const float *__restrict__ nb_d,
                                   const float *__restrict__ core_p, const int *__restrict__ comp_p, float *bound,
                                   unsigned char *interior) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    seed_bounds_at(p, p < n, kn, nb_pos, nb_d, core_p, comp_p, bound, interior);
}

__global__ __launch_bounds__(64) void cert_setup_kernel(int level, const int2 *__restrict__ range,
                                                        const float *__restrict__ core_p, int *cert_of_pos,
                                                        float *mincore) {
    cert_setup_node(level, range, core_p, cert_of_pos, mincore, blockIdx.x);
}

// (the certification itself: cert_node in fh_search.h)
template <int D>
__global__ __launch_bounds__(64) void cert_kernel(CertArgs a) {
    __shared__ int q[kCertQueue];
    cert_node<D>(a, blockIdx.x, q);
}

// ------------------------------------------ wave-cooperative search ----

// (the search itself: wave_search in fh_search.h)
template <int D, int PASS>
__global__ __launch_bounds__(256) void boruvka_wave_kernel(QueryArgs a) {
    __shared__ float s_x[4][kWaveStage * D];
    __shared__ float s_c[4][kWaveStage];
    __shared__ int s_k[4][kWaveStage];
    const int wv = threadIdx.x >> 6;
    wave_search<D, PASS>(a, (int64_t)blockIdx.x * 256 + threadIdx.x, s_x[wv], s_c[wv], s_k[wv]);
}

// sample_weight_core_distance (boruvka.py:527-539): the first sorted
// neighbour at which the float64 running weight reaches min_samples.  A -1
// neighbour (k > n) reads the last weight, as numba's indexing does.
__global__ void weighted_core_kernel(int n, int k, const int *__restrict__ knn_i, const float *__restrict__ knn_d,
                                     const float *__restrict__ w, int ms, const int *__restrict__ posof,
                                     float *core_o, float *core_p) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double total = 0.0;
    int j = 0;
    while (total < (double)ms && j < k) {
        const int nb = knn_i[(size_t)i * k + j];
        total += (double)w[nb < 0 ? nb + n : nb];
        ++j;
    }
    const float c = knn_d[(size_t)i * k + (j - 1)];
    core_o[i] = c;
    core_p[posof[i]] = c;
}

// ------------------------------------------------- kNN-seeded forest ----

// boruvka.py:492-503: the first neighbour column j >= 1 whose core distance
// does not exceed the point's own.  Negative neighbour indices (rows with
// fewer real neighbours than k) read core[n-1], as numba's wrap-around
// indexing does.
__global__ void init_edges_kernel(int n, int k, const int *__restrict__ knn_i, const float *__restrict__ knn_d,
                                  const float *__restrict__ core_o, int *succ, int *dst_raw, float *w) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float ci = core_o[i];
    int s = -1, draw = -1;
    float wi = 0.0f;
    for (int j = 1; j < k; ++j) {
        const int kk = knn_i[(size_t)i * k + j];
        const int kw = kk < 0 ? kk + n : kk;
        if (ci >= core_o[kw]) {
            // without distance rows the core is the k-th distance: d_ij <= ci
            const float dj = knn_d != nullptr ? knn_d[(size_t)i * k + j] : ci;
            s = kw;
            draw = kk;
            wi = ci > dj ? ci : dj;
            break;
        }
    }
    succ[i] = s;
    dst_raw[i] = draw;
    w[i] = wi;
}

// The same choice from tree-order kNN rows (KD-tree positions, -1 padding):
// position p of point i = perm[p] scans its columns 1..k-1; the edge fields
// are written by point index for the emission in point order.
__global__ void init_edges_pos_kernel(int n, int k, const int *__restrict__ nb_pos, const float *__restrict__ nb_d,
                                      const float *__restrict__ core_p, const int *__restrict__ perm,
                                      const int *__restrict__ posof, int *succ, int *dst_raw, float *w,
                                      float *core_pt) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const ForestEdge e = forest_edge_at(p, n, k, nb_pos, nb_d, core_p, perm, posof);
    const int i = perm[p];
    succ[i] = e.succ;
    dst_raw[i] = e.dst_raw;
    w[i] = e.w;
    core_pt[i] = core_p[p];
}

// One min_samples value of a sweep from shared tree-order kNN rows of width
// K (columns 1..K of a (K+1)-nearest query): the first ms columns and the
// core (column ms).  The (ms+1)-nearest query's output equals these columns
// whenever its ms+2 smallest values are distinct; positions with an equal
// pair among them (a zero first column means a duplicate point, equal to the
// query's own zero) are listed for the exact (ms+1)-nearest query, whose
// heap order decides ties.
__global__ void sweep_slice_kernel(int64_t n, int K, int ms, const int *__restrict__ sw_pos,
                                   const float *__restrict__ sw_d, int *nb_pos, float *nb_d, float *core_p,
                                   int *tie_list, int *tie_count) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int *rp = sw_pos + p * K;
    const float *rd = sw_d + p * K;
    bool tie = !(rd[0] > 0.0f);
    for (int c = 0; c < ms; ++c) {
        nb_pos[p * ms + c] = rp[c];
        nb_d[p * ms + c] = rd[c];
        tie |= rd[c] == rd[c + 1];
    }
    core_p[p] = rd[ms - 1];
    if (tie) tie_list[atomicAdd(tie_count, 1)] = (int)p;
}

// Outputs by point index from tree-order rows: the core, and the neighbour
// columns as point indices (-1 kept).
__global__ void core_by_point_kernel(int n, const int *__restrict__ posof, const float *__restrict__ core_p,
                                     float *core_out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) core_out[i] = core_p[posof[i]];
}

__global__ void nbr_by_point_kernel(int64_t n, int C, const int *__restrict__ posof, const int *__restrict__ perm,
                                    const int *__restrict__ nb_pos, int *nbr) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n * C) return;
    const int64_t i = e / C;
    const int c = (int)(e - i * C);
    const int q = nb_pos[(int64_t)posof[i] * C + c];
    nbr[e] = q < 0 ? -1 : perm[q];
}

// ----------------------------------------------- cycles by doubling ----
//
// In a functional graph (one out-edge per node) every weakly connected
// piece holds at most one cycle.  Processing edges in key order with a
// union-find (boruvka.py:509-522, 64-76) skips exactly the cycle edge
// processed last: the one whose source has the largest key on the cycle.
// Pointer doubling finds it: after K steps with 2^K >= m, S(v) lies on v's
// cycle and M(S(v)) is that cycle's largest key.

__global__ void dbl_init_kernel(int m, const int *__restrict__ succ, int *S, int *M) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= m) return;
    S[v] = succ[v];
    M[v] = v;
}

__global__ void dbl_step_kernel(int m, const int *__restrict__ S, const int *__restrict__ M, int *S2, int *M2) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= m) return;
    const int s = S[v];
    const int mv = M[v];
    if (s < 0) {
        S2[v] = -1;
        M2[v] = mv;
    } else {
        S2[v] = S[s];
        const int ms = M[s];
        M2[v] = mv > ms ? mv : ms;
    }
}

// keep[v] = node has an edge and it is not its cycle's skipped edge.
__global__ void dbl_keep_kernel(int m, const int *__restrict__ succ, const int *__restrict__ S,
                                const int *__restrict__ M, int *keep) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= m) return;
    const int s = S[v];
    const bool skip = s >= 0 && M[s] == v;
    keep[v] = (succ[v] >= 0 && !skip) ? 1 : 0;
}

// A functional graph whose key w never increases along succ: in a round, a
// component's successor never has a heavier best edge (its own best is at
// most the edge the two share); in the kNN-seeded forest, a point's
// successor never has a larger core distance (boruvka.py:492-503).  So a
// cycle longer than two has equal keys all round, and each node on it has a
// successor of equal key that does not point back.  cyc_classify settles
// every other node at once -- the two nodes of a 2-cycle keep the edge of
// the smaller one (the larger key is the one the reference skips), any
// other node is on no cycle -- and flags the rest (tflag) for pointer
// doubling over just those (cyc_tinit, dbl_step, cyc_tkeep).
__global__ void cyc_classify_kernel(int m, const int *__restrict__ succ, const float *__restrict__ w, int *keep,
                                    int *tflag) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= m) return;
    const int s = succ[v];
    int k = 0, t = 0;
    if (s >= 0) {
        if (succ[s] == v) {
            k = v < s ? 1 : 0;
        } else {
            k = 1;
            t = w[s] == w[v] ? 1 : 0;
        }
    }
    keep[v] = k;
    tflag[v] = t;
}

// Flagged node v becomes entry tpos[v]: its successor's entry when that is
// flagged too (else -1: the path leaves the flagged set, so v is on no
// cycle), and its own id as the running maximum.
__global__ void cyc_tinit_kernel(int m, const int *__restrict__ succ, const int *__restrict__ tflag,
                                 const int *__restrict__ tpos, int *S, int *M, int *tnode) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= m || !tflag[v]) return;
    const int i = tpos[v], s = succ[v];
    S[i] = tflag[s] ? tpos[s] : -1;
    M[i] = v;
    tnode[i] = v;
}

__global__ void cyc_tkeep_kernel(int t, const int *__restrict__ S, const int *__restrict__ M,
                                 const int *__restrict__ tnode, int *keep) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= t) return;
    const int v = tnode[i], s = S[i];
    keep[v] = (s >= 0 && M[s] == v) ? 0 : 1;
}

// ------------------------------------------------------------ scans ----

constexpr int kScanBlock = 256;
constexpr int kScanItems = 4;
constexpr int kScanTile = kScanBlock * kScanItems;

__device__ __forceinline__ int wave_incl_scan(int v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int t = __shfl_up(v, o, 64);
        if (lane >= o) v += t;
    }
    return v;
}

// Exclusive scan of one value per thread over the block; returns the block
// total in `total`.
__device__ __forceinline__ int block_excl_scan(int v, int &total) {
    __shared__ int wsum[kScanBlock / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int inc = wave_incl_scan(v);
    if (lane == 63) wsum[wave] = inc;
    __syncthreads();
    int off = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < kScanBlock / 64; ++w) {
        const int s = wsum[w];
        off += (w < wave) ? s : 0;
        tot += s;
    }
    __syncthreads();
    total = tot;
    return off + inc - v;
}

__global__ __launch_bounds__(kScanBlock) void scan_tile_kernel(int m, const int *__restrict__ in, int *out,
                                                                int *tile_sum) {
    const int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    int v[kScanItems];
    int s = 0;
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        v[j] = (base + j < m) ? in[base + j] : 0;
        s += v[j];
    }
    int total;
    int off = block_excl_scan(s, total);
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        if (base + j < m) out[base + j] = off;
        off += v[j];
    }
    if (threadIdx.x == 0) tile_sum[blockIdx.x] = total;
}

__global__ __launch_bounds__(kScanBlock) void scan_sums_kernel(int tiles, int *tile_sum, int *total_out) {
    int carry = 0;
    for (int b0 = 0; b0 < tiles; b0 += kScanBlock) {
        const int t = b0 + threadIdx.x;
        const int v = t < tiles ? tile_sum[t] : 0;
        int tot;
        const int ex = block_excl_scan(v, tot);
        if (t < tiles) tile_sum[t] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) *total_out = carry;
}

__global__ __launch_bounds__(kScanBlock) void scan_add_kernel(int m, int *out, const int *__restrict__ tile_sum) {
    const int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    const int add = tile_sum[blockIdx.x];
#pragma unroll
    for (int j = 0; j < kScanItems; ++j)
        if (base + j < m) out[base + j] += add;
}


__global__ void iota_kernel(int m, int *out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) out[i] = i;
}

// Initial forest: emit kept edges in point order, then union them.
__global__ void init_emit_kernel(int n, const int *__restrict__ keep, const int *__restrict__ offs,
                                 const int *__restrict__ dst_raw, const float *__restrict__ w, int *src_out,
                                 int *dst_out, float *w_out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !keep[i]) return;
    const int r = offs[i];
    src_out[r] = i;
    dst_out[r] = dst_raw[i];
    w_out[r] = w[i];
}

__global__ void init_union_kernel(int n, const int *__restrict__ keep, const int *__restrict__ succ, int *parent) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !keep[i]) return;
    uf_unite(parent, i, succ[i]);
}

__global__ void compress_all_kernel(int n, int *parent) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    parent[i] = uf_find(parent, i);
}

__global__ void label_points_kernel(int n, const int *__restrict__ perm, const int *__restrict__ parent,
                                    int *comp_p) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    comp_p[p] = parent[perm[p]];
}

__global__ void root_flags_kernel(int n, const int *__restrict__ parent, int *flag) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    flag[i] = parent[i] == i ? 1 : 0;
}

__global__ void root_scatter_kernel(int n, const int *__restrict__ flag, const int *__restrict__ offs, int *roots,
                                    int *rootslot) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !flag[i]) return;
    roots[offs[i]] = i;
    rootslot[i] = offs[i];
}

// --------------------------------------------------- Boruvka rounds ----

__global__ void round_reset_kernel(int c, const int *__restrict__ roots, float *bound,
                                   unsigned long long *best) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= c) return;
    const int r = roots[s];
    bound[r] = kInf;
    best[r] = ~0ull;
}

// Node purity, boruvka.py:98-124.  A block turns 256 nodes of one level into
// `steps` levels above them.  At the leaf level the values come from the
// points; otherwise from node_comp of the level below.
__global__ __launch_bounds__(256) void purity_kernel(int level, int steps, int leaf_level,
                                                     const int2 *__restrict__ range, const int *__restrict__ comp_p,
                                                     int *node_comp) {
    __shared__ int sv[256];
    const int count = 1 << level;
    const int idx = blockIdx.x * 256 + threadIdx.x;
    int v = -1;
    if (idx < count) {
        const int node = count - 1 + idx;
        if (leaf_level) {
            const int2 r = range[node];
            const int c0 = comp_p[r.x];
            bool pure = true;
            for (int i = r.x + 1; i < r.y; ++i)
                if (comp_p[i] != c0) {
                    pure = false;
                    break;
                }
            v = pure ? c0 : -1;
            node_comp[node] = v;
        } else {
            v = node_comp[node];
        }
    }
    sv[threadIdx.x] = v;
    __syncthreads();
    for (int h = 1; h <= steps; ++h) {
        const int lvl = level - h;
        const int cnt = 1 << lvl;
        const int width = 256 >> h;
        int nv = -1;
        const bool act = threadIdx.x < width;
        const int gidx = (blockIdx.x * 256 >> h) + threadIdx.x;
        if (act) {
            const int a = sv[2 * threadIdx.x], b = sv[2 * threadIdx.x + 1];
            nv = (a == b) ? a : -1;
        }
        __syncthreads();
        if (act) {
            sv[threadIdx.x] = nv;
            if (gidx < cnt) node_comp[cnt - 1 + gidx] = nv;
        }
        __syncthreads();
    }
}

// Per component slot: its round edge (round_edge_of), the target component
// mapped to its slot.
__global__ void select_kernel(int c, int n, const int *__restrict__ roots, const unsigned long long *__restrict__ best,
                              const int *__restrict__ partner, const int *__restrict__ posof,
                              const int *__restrict__ perm, const int *__restrict__ comp_p,
                              const int *__restrict__ rootslot, int *succ, int *ei, int *ej, float *ew) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= c) return;
    const RoundEdge e = round_edge_of(roots[s], n, best, partner, posof, perm, comp_p);
    succ[s] = rootslot[e.tgt];
    ei[s] = e.i;
    ej[s] = e.j;
    ew[s] = e.w;
}

__global__ void round_emit_kernel(int c, int base, const int *__restrict__ keep, const int *__restrict__ offs,
                                  const int *__restrict__ ei, const int *__restrict__ ej,
                                  const float *__restrict__ ew, int *src_out, int *dst_out, float *w_out) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= c || !keep[s]) return;
    const int r = base + offs[s];
    src_out[r] = ei[s];
    dst_out[r] = ej[s];
    w_out[r] = ew[s];
}

__global__ void round_union_kernel(int c, const int *__restrict__ keep, const int *__restrict__ roots,
                                   const int *__restrict__ succ, int *parent) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= c || !keep[s]) return;
    uf_unite(parent, roots[s], roots[succ[s]]);
}

__global__ void compress_roots_kernel(int c, const int *__restrict__ roots, int *parent) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= c) return;
    const int r = roots[s];
    parent[r] = uf_find(parent, r);
}

__global__ void relabel_kernel(int n, const int *__restrict__ parent, int *comp_p) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    comp_p[p] = parent[comp_p[p]];
}

__global__ void still_root_kernel(int c, const int *__restrict__ roots, const int *__restrict__ parent, int *flag) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= c) return;
    const int r = roots[s];
    flag[s] = parent[r] == r ? 1 : 0;
}

__global__ void roots_compact_kernel(int c, const int *__restrict__ roots, const int *__restrict__ flag,
                                     const int *__restrict__ offs, int *roots_out, int *rootslot) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= c || !flag[s]) return;
    const int r = roots[s];
    roots_out[offs[s]] = r;
    rootslot[r] = offs[s];
}

// A round's work after the search when at most kSmallRound components are
// left, in one work-group (one thread per component slot s): the selection
// of select_kernel, the cycle skip by pointer doubling in LDS (dbl_*), the
// emission in slot order (round_emit_kernel), the unions and root
// compression (round_union_kernel, compress_roots_kernel), the surviving
// roots and their compaction (still_root_kernel, roots_compact_kernel):
// the same arithmetic as those kernels, with work-group barriers in place
// of kernel boundaries.  small[0] = rows emitted, small[1] = components
// left.  The relabelling of the points stays a grid-wide kernel.
constexpr int kSmallRound = 1024;

__device__ __forceinline__ int block_excl_scan_1024(int v, int &total, int *wsum) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int inc = wave_incl_scan(v);
    if (lane == 63) wsum[wave] = inc;
    __syncthreads();
    int off = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < kSmallRound / 64; ++w) {
        const int x = wsum[w];
        off += (w < wave) ? x : 0;
        tot += x;
    }
    __syncthreads();
    total = tot;
    return off + inc - v;
}

__global__ __launch_bounds__(kSmallRound) void round_small_kernel(SmallRoundArgs a) {
    __shared__ int S[2][kSmallRound], M[2][kSmallRound];
    __shared__ int wsum[kSmallRound / 64];
    const int s = threadIdx.x;
    const int c = a.c;
    const bool on = s < c;
    // selection (select_kernel)
    int i_star = -1, j_star = -1, succ = -1, r = -1;
    float w = kInf;
    if (on) {
        r = a.roots[s];
        const RoundEdge e = round_edge_of(r, a.n, a.best, a.partner, a.posof, a.perm, a.comp_p);
        i_star = e.i;
        j_star = e.j;
        w = e.w;
        succ = a.rootslot[e.tgt];
    }
    // cycle skip (dbl_init, dbl_step, dbl_keep)
    S[0][s] = succ;
    M[0][s] = s;
    __syncthreads();
    int cur = 0;
    for (int span = 1; span < c; span <<= 1) {
        if (on) {
            const int sv = S[cur][s], mv = M[cur][s];
            if (sv < 0) {
                S[cur ^ 1][s] = -1;
                M[cur ^ 1][s] = mv;
            } else {
                S[cur ^ 1][s] = S[cur][sv];
                const int ms = M[cur][sv];
                M[cur ^ 1][s] = mv > ms ? mv : ms;
            }
        }
        __syncthreads();
        cur ^= 1;
    }
    int keep = 0;
    if (on) {
        const int sv = S[cur][s];
        const bool skip = sv >= 0 && M[cur][sv] == s;
        keep = (succ >= 0 && !skip) ? 1 : 0;
    }
    // emission in slot order (round_emit_kernel)
    int emitted;
    const int off = block_excl_scan_1024(keep, emitted, wsum);
    if (keep) {
        const int row = a.base + off;
        a.src_out[row] = i_star;
        a.dst_out[row] = j_star;
        a.w_out[row] = w;
        // S[cur] is no longer read: the slots' roots for the unions
    }
    __syncthreads();
    if (on) S[cur][s] = r;
    __syncthreads();
    if (keep) uf_unite(a.parent, r, S[cur][succ]);
    __syncthreads();
    if (on) a.parent[r] = uf_find(a.parent, r);
    __syncthreads();
    // surviving roots, compacted (still_root_kernel, roots_compact_kernel)
    const int flag = (on && ld_relaxed(a.parent + r) == r) ? 1 : 0;
    int left;
    const int off2 = block_excl_scan_1024(flag, left, wsum);
    if (flag) {
        a.roots_out[off2] = r;
        a.rootslot[r] = off2;
    }
    if (s == 0) {
        a.small[0] = emitted;
        a.small[1] = left;
    }
}

__global__ void posof_kernel(int n, const int *__restrict__ perm, int *posof) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    posof[perm[p]] = p;
}

// Columns 1..k-1 of the sorted kNN rows (boruvka.py:647 neighbors[:, 1:]).
__global__ void neighbors_kernel(int64_t n, int k, const int *__restrict__ knn_i, int *nbr) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int kn = k - 1;
    if (t >= n * kn) return;
    const int64_t row = t / kn;
    const int col = (int)(t - row * kn);
    nbr[t] = knn_i[row * k + 1 + col];
}

inline unsigned blocks(int64_t m, int b) { return (unsigned)((m + b - 1) / b); }

}  // namespace

// ------------------------------------------------------------ launch ----

template <int D>
static hipError_t launch_knn_t(const KnnArgs &a, hipStream_t st) {
    const int block = a.block;
    if (a.heap_p) {
        hipLaunchKernelGGL((knn_kernel<D, true>), dim3(blocks(a.nq, block)), dim3(block), 0, st, a);
    } else {
        const size_t shmem = (size_t)a.k * block * (sizeof(float) + sizeof(int));
        hipLaunchKernelGGL((knn_kernel<D, false>), dim3(blocks(a.nq, block)), dim3(block), shmem, st, a);
    }
    return hipGetLastError();
}

template <int D>
static hipError_t launch_query_t(const QueryArgs &a, hipStream_t st) {
    if constexpr (D > 0) {
        if (a.wave && a.pass != 1) {
            if (a.pass == 2)
                hipLaunchKernelGGL((boruvka_wave_kernel<D, 2>), dim3(blocks(a.n, 256)), dim3(256), 0, st, a);
            else
                hipLaunchKernelGGL((boruvka_wave_kernel<D, 0>), dim3(blocks(a.n, 256)), dim3(256), 0, st, a);
            return hipGetLastError();
        }
    }
    if (a.pass == 1)
        hipLaunchKernelGGL((boruvka_query_kernel<D, 1>), dim3(blocks(a.n, 256)), dim3(256), 0, st, a);
    else if (a.pass == 2)
        hipLaunchKernelGGL((boruvka_query_kernel<D, 2>), dim3(blocks(a.n, 256)), dim3(256), 0, st, a);
    else
        hipLaunchKernelGGL((boruvka_query_kernel<D, 0>), dim3(blocks(a.n, 256)), dim3(256), 0, st, a);
    return hipGetLastError();
}

#define FH_DISPATCH_D(d, fn, ...)                 \
    switch (d) {                                  \
        case 1: return fn<1>(__VA_ARGS__);        \
        case 2: return fn<2>(__VA_ARGS__);        \
        case 3: return fn<3>(__VA_ARGS__);        \
        case 4: return fn<4>(__VA_ARGS__);        \
        case 5: return fn<5>(__VA_ARGS__);        \
        case 6: return fn<6>(__VA_ARGS__);        \
        case 7: return fn<7>(__VA_ARGS__);        \
        case 8: return fn<8>(__VA_ARGS__);        \
        default: return fn<0>(__VA_ARGS__);       \
    }

hipError_t launch_knn(const KnnArgs &a, hipStream_t st) { FH_DISPATCH_D(a.d, launch_knn_t, a, st) }

template <int D>
static hipError_t launch_cert_t(const CertArgs &a, hipStream_t st) {
    hipLaunchKernelGGL(cert_kernel<D>, dim3((unsigned)(1u << a.level)), dim3(64), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_cert(const CertArgs &a, hipStream_t st) { FH_DISPATCH_D(a.d, launch_cert_t, a, st) }

hipError_t launch_cert_setup(int level, const int2 *range, const float *core_p, int *cert_of_pos, float *mincore,
                             hipStream_t st) {
    hipLaunchKernelGGL(cert_setup_kernel, dim3((unsigned)(1u << level)), dim3(64), 0, st, level, range, core_p,
                       cert_of_pos, mincore);
    return hipGetLastError();
}

hipError_t launch_seed_bounds(int64_t n, int kn, const int *nb_pos, const float *nb_d, const float *core_p,
                              const int *comp_p, float *bound, unsigned char *interior, hipStream_t st) {
    if (kn <= 0) return hipSuccess;
    hipLaunchKernelGGL(seed_bounds_kernel, dim3(blocks(n, 256)), dim3(256), 0, st, n, kn, nb_pos, nb_d, core_p,
                       comp_p, bound, interior);
    return hipGetLastError();
}

hipError_t launch_query(const QueryArgs &a, hipStream_t st) { FH_DISPATCH_D(a.d, launch_query_t, a, st) }

hipError_t launch_weighted_core(int n, int k, const int *knn_i, const float *knn_d, const float *weights, int ms,
                                const int *posof, float *core_o, float *core_p, hipStream_t st) {
    hipLaunchKernelGGL(weighted_core_kernel, dim3(blocks(n, 256)), dim3(256), 0, st, n, k, knn_i, knn_d, weights, ms,
                       posof, core_o, core_p);
    return hipGetLastError();
}

hipError_t launch_posof(int n, const int *perm, int *posof, hipStream_t st) {
    hipLaunchKernelGGL(posof_kernel, dim3(blocks(n, 256)), dim3(256), 0, st, n, perm, posof);
    return hipGetLastError();
}

hipError_t launch_neighbors(int64_t n, int k, const int *knn_i, int *nbr, hipStream_t st) {
    if (k <= 1 || n == 0) return hipSuccess;
    hipLaunchKernelGGL(neighbors_kernel, dim3(blocks(n * (k - 1), 256)), dim3(256), 0, st, n, k, knn_i, nbr);
    return hipGetLastError();
}

int scan_tiles(int m) { return (m + kScanTile - 1) / kScanTile; }

// Node ranges of the implicit tree (numba_kdtree.py:426-473: a node's left
// child takes n_points // 2 of its points): each node descends from the root
// along the bits of its index.
__global__ void node_range_kernel(int64_t N, int n, int2 *range) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    range[i] = node_range_of(i, n);
}

hipError_t launch_node_ranges(int64_t N, int n, int2 *range, hipStream_t st) {
    hipLaunchKernelGGL(node_range_kernel, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, st, N, n, range);
    return hipGetLastError();
}

hipError_t launch_excl_scan(int m, const int *in, int *out, int *tile_sum, int *total, hipStream_t st) {
    const int tiles = scan_tiles(m);
    if (m == 0) {
        return hipMemsetAsync(total, 0, sizeof(int), st);
    }
    hipLaunchKernelGGL(scan_tile_kernel, dim3(tiles), dim3(kScanBlock), 0, st, m, in, out, tile_sum);
    hipLaunchKernelGGL(scan_sums_kernel, dim3(1), dim3(kScanBlock), 0, st, tiles, tile_sum, total);
    hipLaunchKernelGGL(scan_add_kernel, dim3(tiles), dim3(kScanBlock), 0, st, m, out, tile_sum);
    return hipGetLastError();
}

hipError_t launch_cycle_keep(int m, const int *succ, CycleBufs &b, int *keep, hipStream_t st) {
    if (m == 0) return hipSuccess;
    hipLaunchKernelGGL(dbl_init_kernel, dim3(blocks(m, 256)), dim3(256), 0, st, m, succ, b.S0, b.M0);
    int steps = 0;
    while ((int64_t(1) << steps) < m) ++steps;
    int *S = b.S0, *M = b.M0, *S2 = b.S1, *M2 = b.M1;
    for (int t = 0; t < steps; ++t) {
        hipLaunchKernelGGL(dbl_step_kernel, dim3(blocks(m, 256)), dim3(256), 0, st, m, S, M, S2, M2);
        std::swap(S, S2);
        std::swap(M, M2);
    }
    hipLaunchKernelGGL(dbl_keep_kernel, dim3(blocks(m, 256)), dim3(256), 0, st, m, succ, S, M, keep);
    return hipGetLastError();
}

hipError_t launch_cycle_classify(int m, const int *succ, const float *w, int *keep, int *tflag, hipStream_t st) {
    if (m == 0) return hipSuccess;
    hipLaunchKernelGGL(cyc_classify_kernel, dim3(blocks(m, 256)), dim3(256), 0, st, m, succ, w, keep, tflag);
    return hipGetLastError();
}

hipError_t launch_cycle_ties(int m, int t, const int *succ, const int *tflag, const int *tpos, CycleBufs &b,
                             int *tnode, int *keep, hipStream_t st) {
    if (t == 0) return hipSuccess;
    hipLaunchKernelGGL(cyc_tinit_kernel, dim3(blocks(m, 256)), dim3(256), 0, st, m, succ, tflag, tpos, b.S0, b.M0,
                       tnode);
    int steps = 0;
    while ((int64_t(1) << steps) < t) ++steps;
    int *S = b.S0, *M = b.M0, *S2 = b.S1, *M2 = b.M1;
    for (int i = 0; i < steps; ++i) {
        hipLaunchKernelGGL(dbl_step_kernel, dim3(blocks(t, 256)), dim3(256), 0, st, t, S, M, S2, M2);
        std::swap(S, S2);
        std::swap(M, M2);
    }
    hipLaunchKernelGGL(cyc_tkeep_kernel, dim3(blocks(t, 256)), dim3(256), 0, st, t, S, M, tnode, keep);
    return hipGetLastError();
}

hipError_t launch_init_edges_pos(int n, int k, const int *nb_pos, const float *nb_d, const float *core_p,
                                 const int *perm, const int *posof, int *succ, int *dst_raw, float *w, float *core_pt,
                                 hipStream_t st) {
    hipLaunchKernelGGL(init_edges_pos_kernel, dim3(blocks(n, 256)), dim3(256), 0, st, n, k, nb_pos, nb_d, core_p, perm,
                       posof, succ, dst_raw, w, core_pt);
    return hipGetLastError();
}

hipError_t launch_sweep_slice(int64_t n, int K, int ms, const int *sw_pos, const float *sw_d, int *nb_pos,
                              float *nb_d, float *core_p, int *tie_list, int *tie_count, hipStream_t st) {
    hipLaunchKernelGGL(sweep_slice_kernel, dim3(blocks(n, 256)), dim3(256), 0, st, n, K, ms, sw_pos, sw_d, nb_pos,
                       nb_d, core_p, tie_list, tie_count);
    return hipGetLastError();
}

hipError_t launch_outputs_by_point(int64_t n, int k, const int *posof, const int *perm, const float *core_p,
                                   const int *nb_pos, float *core_out, int *nbr, hipStream_t st) {
    if (core_out)
        hipLaunchKernelGGL(core_by_point_kernel, dim3(blocks(n, 256)), dim3(256), 0, st, (int)n, posof, core_p,
                           core_out);
    if (nbr && k > 1)
        hipLaunchKernelGGL(nbr_by_point_kernel, dim3(blocks(n * (k - 1), 256)), dim3(256), 0, st, n, k - 1, posof, perm,
                           nb_pos, nbr);
    return hipGetLastError();
}

hipError_t launch_init_edges(int n, int k, const int *knn_i, const float *knn_d, const float *core_o, int *succ,
                             int *dst_raw, float *w, hipStream_t st) {
    hipLaunchKernelGGL(init_edges_kernel, dim3(blocks(n, 256)), dim3(256), 0, st, n, k, knn_i, knn_d, core_o, succ,
                       dst_raw, w);
    return hipGetLastError();
}

hipError_t launch_init_emit_union(int n, const int *keep, const int *offs, const int *succ, const int *dst_raw,
                                  const float *w, int *parent, int *src_out, int *dst_out, float *w_out,
                                  hipStream_t st) {
    hipLaunchKernelGGL(init_emit_kernel, dim3(blocks(n, 256)), dim3(256), 0, st, n, keep, offs, dst_raw, w,
                       src_out, dst_out, w_out);
    hipLaunchKernelGGL(iota_kernel, dim3(blocks(n, 256)), dim3(256), 0, st, n, parent);
    hipLaunchKernelGGL(init_union_kernel, dim3(blocks(n, 256)), dim3(256), 0, st, n, keep, succ, parent);
    hipLaunchKernelGGL(compress_all_kernel, dim3(blocks(n, 256)), dim3(256), 0, st, n, parent);
    return hipGetLastError();
}

hipError_t launch_labels_and_root_flags(int n, const int *perm, const int *parent, int *comp_p, int *flag,
                                        hipStream_t st) {
    hipLaunchKernelGGL(label_points_kernel, dim3(blocks(n, 256)), dim3(256), 0, st, n, perm, parent, comp_p);
    hipLaunchKernelGGL(root_flags_kernel, dim3(blocks(n, 256)), dim3(256), 0, st, n, parent, flag);
    return hipGetLastError();
}

hipError_t launch_root_scatter(int n, const int *flag, const int *offs, int *roots, int *rootslot, hipStream_t st) {
    hipLaunchKernelGGL(root_scatter_kernel, dim3(blocks(n, 256)), dim3(256), 0, st, n, flag, offs, roots, rootslot);
    return hipGetLastError();
}

hipError_t launch_round_reset(int c, const int *roots, float *bound, unsigned long long *best, hipStream_t st) {
    hipLaunchKernelGGL(round_reset_kernel, dim3(blocks(c, 256)), dim3(256), 0, st, c, roots, bound, best);
    return hipGetLastError();
}

hipError_t launch_purity(int levels, const int2 *range, const int *comp_p, int *node_comp, hipStream_t st) {
    int level = levels - 1;
    bool leaf = true;
    for (;;) {
        const int steps = level < 8 ? level : 8;
        const int count = 1 << level;
        hipLaunchKernelGGL(purity_kernel, dim3(blocks(count, 256)), dim3(256), 0, st, level, steps, leaf ? 1 : 0,
                           range, comp_p, node_comp);
        leaf = false;
        level -= steps;
        if (level <= 0) break;
    }
    return hipGetLastError();
}

hipError_t launch_select(int c, int n, const int *roots, const unsigned long long *best, const int *partner,
                         const int *posof, const int *perm, const int *comp_p, const int *rootslot, int *succ,
                         int *ei, int *ej, float *ew, hipStream_t st) {
    hipLaunchKernelGGL(select_kernel, dim3(blocks(c, 256)), dim3(256), 0, st, c, n, roots, best, partner, posof, perm,
                       comp_p, rootslot, succ, ei, ej, ew);
    return hipGetLastError();
}

hipError_t launch_round_emit_union(int c, int base, const int *keep, const int *offs, const int *ei, const int *ej,
                                   const float *ew, const int *roots, const int *succ, int *parent, int *src_out,
                                   int *dst_out, float *w_out, hipStream_t st) {
    hipLaunchKernelGGL(round_emit_kernel, dim3(blocks(c, 256)), dim3(256), 0, st, c, base, keep, offs, ei, ej, ew,
                       src_out, dst_out, w_out);
    hipLaunchKernelGGL(round_union_kernel, dim3(blocks(c, 256)), dim3(256), 0, st, c, keep, roots, succ, parent);
    hipLaunchKernelGGL(compress_roots_kernel, dim3(blocks(c, 256)), dim3(256), 0, st, c, roots, parent);
    return hipGetLastError();
}

hipError_t launch_relabel_and_flags(int n, int c, const int *roots, const int *parent, int *comp_p, int *flag,
                                    hipStream_t st) {
    hipLaunchKernelGGL(relabel_kernel, dim3(blocks(n, 256)), dim3(256), 0, st, n, parent, comp_p);
    hipLaunchKernelGGL(still_root_kernel, dim3(blocks(c, 256)), dim3(256), 0, st, c, roots, parent, flag);
    return hipGetLastError();
}

int small_round_limit() { return kSmallRound; }

hipError_t launch_round_small(const SmallRoundArgs &a, hipStream_t st) {
    if (a.c < 1 || a.c > kSmallRound) return hipErrorInvalidValue;
    hipLaunchKernelGGL(round_small_kernel, dim3(1), dim3(kSmallRound), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_relabel(int n, const int *parent, int *comp_p, hipStream_t st) {
    hipLaunchKernelGGL(relabel_kernel, dim3(blocks(n, 256)), dim3(256), 0, st, n, parent, comp_p);
    return hipGetLastError();
}

hipError_t launch_roots_compact(int c, const int *roots, const int *flag, const int *offs, int *roots_out,
                                int *rootslot, hipStream_t st) {
    hipLaunchKernelGGL(roots_compact_kernel, dim3(blocks(c, 256)), dim3(256), 0, st, c, roots, flag, offs, roots_out,
                       rootslot);
    return hipGetLastError();
}

}  // namespace fh
