// Persistent work-queue kernel for many small independent datasets.
//
// Each work-group of one launch takes the next dataset from a device-side
// counter and runs the whole exact MST path for it -- KD-tree, kNN core
// distances, the kNN-seeded forest and every Boruvka round -- with
// work-group barriers where the per-call path (fh_api.hip run_mst) has
// kernel boundaries, then takes the next dataset, until the queue is empty.
// A batch costs one launch per dimension present, with no host round trip
// per dataset or per round, and the chip stays filled however small the
// datasets are.
//
// The steps are the per-call kernels' own device code (fh_tree_dev.h:
// subtree_build; fh_search.h: walk, KnnVisitor, seed_bounds_at,
// wave_search, forest_edge_at, round_edge_of, the union-find) plus
// work-group forms of the grid-wide passes (scans, pointer doubling,
// emission, unions, relabelling), so each dataset's rows, weights and core
// distances equal fh_mst_euclidean_device's bit for bit.
//
// Per work-group: 512 threads (8 wavefronts), 80 KB of LDS (two work-groups
// per CU) that each phase carves anew -- the subtree build's keys and rows,
// then the kNN heaps (k slots per thread), then the wave searches' leaf
// staging -- and a slice of global scratch sized for the batch's largest
// dataset (a few hundred KB, mostly L2-resident).
#include "fh_batch.h"

#include "fh_device.h"
#include "fh_search.h"
#include "fh_tree_dev.h"

namespace fh {

namespace {

constexpr int kBT = kSubNT;  // threads per work-group: the subtree build's
constexpr int kBW = kBT / 64;
constexpr int kMaxKq = 16;
constexpr int kCertDepth = 4;  // certification nodes this many levels above the leaves
constexpr int kMaxBatchN = 1 << 16;  // largest dataset a work-group takes
// rows read ahead in the kNN leaves and the search's staged leaves (4 and 2
// remove the register spills but measured no faster)
constexpr int kBatchLeafBatch = 8;
constexpr int kBatchStageBatch = 4;
// Two work-groups per CU: 80 KB each, less the 256 bytes the work-group
// reductions of the subtree build (__syncthreads_or) take.
constexpr int kLds = 81920 - 256;
constexpr int kLdsWork = kLds;

template <int D>
constexpr bool lds_fits() {
    return subtree_lds_bytes<D>() <= kLdsWork && kBW * kWaveStage * (4 * D + 8) <= kLdsWork;
}
static_assert(lds_fits<1>() && lds_fits<2>() && lds_fits<3>() && lds_fits<4>() && lds_fits<5>() && lds_fits<6>() &&
                  lds_fits<7>() && lds_fits<8>(),
              "every phase fits the work-group's LDS");
static_assert(kMaxKq * kBT * 8 <= kLdsWork, "kNN heaps fit the work-group's LDS");
static_assert(kBW * kCertQueue * 4 <= kLdsWork, "certification queues fit the work-group's LDS");

// A work-group's scratch arrays (n: points, N: tree nodes, d: dimensions,
// kq: neighbours per query), each 256-byte aligned: type, name, count.
#define FH_SLICE_ARRAYS(X)                 \
    X(int2, range, N)                      \
    X(float, box, N * 2 * d)               \
    X(uint64_t, keys, n)                   \
    X(int, part, n)                        \
    X(float, xp, n * d)                    \
    X(int, perm, n)                        \
    X(int, posof, n)                       \
    X(float, core_p, n)                    \
    X(int, nb_pos, n * (kq - 1))           \
    X(float, nb_d, n * (kq - 1))           \
    X(int, comp_p, n)                      \
    X(int, node_comp, N)                   \
    X(int, parent, n)                      \
    X(int, succ, n)                        \
    X(int, dst_raw, n)                     \
    X(float, w_init, n)                    \
    X(int, keep, n)                        \
    X(int, offs, n)                        \
    X(int, flag, n)                        \
    X(int, S0, n)                          \
    X(int, M0, n)                          \
    X(int, S1, n)                          \
    X(int, M1, n)                          \
    X(int, roots, n)                       \
    X(int, roots2, n)                      \
    X(int, rootslot, n)                    \
    X(float, bound, n)                     \
    X(unsigned long long, best, n)         \
    X(int, partner, n)                     \
    X(int, ei, n)                          \
    X(int, ej, n)                          \
    X(float, ew, n)                        \
    X(unsigned char, interior, n)          \
    X(int, cert_of_pos, n)                 \
    X(float, mincore, N)                   \
    X(unsigned char, cert, N)

struct Slice {
#define FH_SLICE_FIELD(T, name, count) T *name;
    FH_SLICE_ARRAYS(FH_SLICE_FIELD)
#undef FH_SLICE_FIELD
};

#define FH_SLICE_STEP(T, count) ((int64_t)sizeof(T) * (count) + 255) & ~int64_t(255)

__host__ __device__ inline int64_t slice_size(int64_t n, int64_t N, int64_t d, int64_t kq) {
    int64_t off = 0;
#define FH_SLICE_SIZE(T, name, count) off += FH_SLICE_STEP(T, count);
    FH_SLICE_ARRAYS(FH_SLICE_SIZE)
#undef FH_SLICE_SIZE
    return off;
}

__device__ __forceinline__ Slice slice_at(unsigned char *base, int64_t n, int64_t N, int64_t d, int64_t kq) {
    Slice s;
    int64_t off = 0;
#define FH_SLICE_PTR(T, name, count)                  \
    s.name = reinterpret_cast<T *>(base + off);       \
    off += FH_SLICE_STEP(T, count);
    FH_SLICE_ARRAYS(FH_SLICE_PTR)
#undef FH_SLICE_PTR
    return s;
}

// Exclusive scan of in[0, m) into out by the whole work-group (m the same
// on every thread); returns the total.  ws: kBW ints of LDS.  The caller has
// made in[] visible with a barrier; out[] is visible on return.
__device__ int wg_excl_scan(const int *in, int *out, int m, int *ws) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int carry = 0;
    for (int b0 = 0; b0 < m; b0 += kBT * 4) {
        const int base = b0 + threadIdx.x * 4;
        int v[4], sum = 0;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            v[u] = base + u < m ? in[base + u] : 0;
            sum += v[u];
        }
        int inc = sum;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int t = __shfl_up(inc, o, 64);
            if (lane >= o) inc += t;
        }
        if (lane == 63) ws[w] = inc;
        __syncthreads();
        int off = 0, tot = 0;
#pragma unroll
        for (int i = 0; i < kBW; ++i) {
            const int x = ws[i];
            off += i < w ? x : 0;
            tot += x;
        }
        __syncthreads();
        int o = carry + off + inc - sum;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if (base + u < m) out[base + u] = o;
            o += v[u];
        }
        carry += tot;
    }
    __syncthreads();
    return carry;
}

// keep[v] for a functional graph over m nodes (succ, -1: no edge): the node
// has an edge and it is not the edge its cycle skips, the one of the
// cycle's largest node (pointer doubling as dbl_init/dbl_step/dbl_keep).
__device__ void wg_cycle_keep(int m, const int *succ, int *S0, int *M0, int *S1, int *M1, int *keep) {
    for (int v = threadIdx.x; v < m; v += kBT) {
        S0[v] = succ[v];
        M0[v] = v;
    }
    __syncthreads();
    int *S = S0, *M = M0, *S2 = S1, *M2 = M1;
    for (int span = 1; span < m; span <<= 1) {
        for (int v = threadIdx.x; v < m; v += kBT) {
            const int sv = S[v], mv = M[v];
            if (sv < 0) {
                S2[v] = -1;
                M2[v] = mv;
            } else {
                S2[v] = S[sv];
                const int ms = M[sv];
                M2[v] = mv > ms ? mv : ms;
            }
        }
        __syncthreads();
        int *t = S;
        S = S2;
        S2 = t;
        t = M;
        M = M2;
        M2 = t;
    }
    for (int v = threadIdx.x; v < m; v += kBT) {
        const int sv = S[v];
        keep[v] = (succ[v] >= 0 && !(sv >= 0 && M[sv] == v)) ? 1 : 0;
    }
    __syncthreads();
}

// Levels 0 .. l_sub - 1 of a dataset larger than one subtree, node after
// node by the whole work-group over global memory: the node's bounds
// (numba_kdtree.py:96-105, as ordered integers), split axis (first strictly
// largest spread, 133-152), keys, and the exact introselect of 353-407
// (wg_select_step's partition steps; a range of at most 16, or one at the
// depth limit, sorted).  Leaves s.keys as the subtree builds of level l_sub
// read them.  lred: kBW * 2 * D words of LDS; ws, ws2: kBW ints each.
template <int D>
__device__ void build_top_levels(const float *__restrict__ X, int l_sub, const Slice &s, uint32_t *lred, int *ws,
                                 int *ws2) {
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    for (int l = 0; l < l_sub; ++l) {
        for (int k = 0; k < (1 << l); ++k) {
            const int node = (1 << l) - 1 + k;
            const int2 r = s.range[node];
            const int size = r.y - r.x;
            uint32_t mn[D], mx[D];
#pragma unroll
            for (int j = 0; j < D; ++j) {
                mn[j] = 0xffffffffu;
                mx[j] = 0u;
            }
            for (int p = r.x + tid; p < r.y; p += kBT) {
                const uint32_t idx = l == 0 ? (uint32_t)p : key_idx(s.keys[p]);
#pragma unroll
                for (int j = 0; j < D; ++j) {
                    const uint32_t o = f2o(X[(size_t)idx * D + j]);
                    mn[j] = min(mn[j], o);
                    mx[j] = max(mx[j], o);
                }
            }
#pragma unroll
            for (int j = 0; j < D; ++j) {
#pragma unroll
                for (int o = 32; o > 0; o >>= 1) {
                    mn[j] = min(mn[j], (uint32_t)__shfl_xor((int)mn[j], o, 64));
                    mx[j] = max(mx[j], (uint32_t)__shfl_xor((int)mx[j], o, 64));
                }
                if (lane == 0) {
                    lred[wv * 2 * D + j] = mn[j];
                    lred[wv * 2 * D + D + j] = mx[j];
                }
            }
            __syncthreads();
#pragma unroll
            for (int j = 0; j < D; ++j) {
                mn[j] = lred[j];
                mx[j] = lred[D + j];
                for (int w = 1; w < kBW; ++w) {
                    mn[j] = min(mn[j], lred[w * 2 * D + j]);
                    mx[j] = max(mx[j], lred[w * 2 * D + D + j]);
                }
            }
            if (tid < 2 * D) {
                uint32_t v = lred[tid];
                for (int w = 1; w < kBW; ++w) {
                    const uint32_t x = lred[w * 2 * D + tid];
                    v = tid < D ? min(v, x) : max(v, x);
                }
                s.box[(size_t)node * 2 * D + tid] = o2f(v);
            }
            __syncthreads();
            const int ax = sub_axis<D>(mn, mx);
            for (int p = r.x + tid; p < r.y; p += kBT) {
                const uint32_t idx = l == 0 ? (uint32_t)p : key_idx(s.keys[p]);
                s.keys[p] = make_key(X[(size_t)idx * D + ax], idx);
            }
            __syncthreads();
            int lo = r.x, hi = r.y, depth = size > 16 ? 2 * floor_log2(size) : 0;
            const int nth = r.x + size / 2;
            for (;;) {
                if (hi - lo <= 16 || depth == 0) {
                    lds_sort(s.keys + lo, hi - lo);  // the same network, here over global memory
                    break;
                }
                --depth;
                const int P = wg_select_step<kBT>(s.keys, s.part, lo, hi, ws, ws2);
                if (nth < P)
                    hi = P;
                else if (nth > P)
                    lo = P + 1;
                else
                    break;
            }
            __syncthreads();
        }
    }
}

// The exact MST of one dataset by the calling work-group.
template <int D>
__device__ void run_dataset(const BatchArgs &a, const BatchJob &job, int ji, const Slice &s, unsigned char *lds,
                            int *ws) {
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int n = job.n, L = job.levels;
    const int N = (1 << L) - 1;
    const int kq = a.kq, kn = kq - 1, ms = a.ms;
    // optional per-phase time (a.phase_ticks: tree, kNN, forest, searches,
    // the rest of the rounds; wall-clock ticks summed over datasets)
    uint64_t t_mark = a.phase_ticks ? wall_clock64() : 0;
    auto phase = [&](int k) {
        if (a.phase_ticks && tid == 0) {
            const uint64_t t = wall_clock64();
            atomicAdd(reinterpret_cast<unsigned long long *>(a.phase_ticks + k), (unsigned long long)(t - t_mark));
            t_mark = t;
        }
    };
    int *src = a.src + job.row_off;
    int *dst = a.dst + job.row_off;
    float *wout = a.w + job.row_off;

    // -- KD-tree (numba_kdtree.py:478-521): node ranges, then every level
    // of the tree in LDS (the root holds at most sub_cap<D>() points)
    for (int i = tid; i < N; i += kBT) s.range[i] = node_range_of(i, n);
    __syncthreads();
    {
        // the first level whose nodes fit one subtree build
        int l_sub = 0;
        while (((n + (1 << l_sub) - 1) >> l_sub) > sub_cap<D>()) ++l_sub;
        if (l_sub > 0) {
            uint32_t *lred = reinterpret_cast<uint32_t *>(lds + 64);
            int *tws = reinterpret_cast<int *>(lred + kBW * 2 * D);
            build_top_levels<D>(a.X + job.x_off, l_sub, s, lred, tws, tws + kBW);
        }
        SubtreeArgs t{};
        t.X = a.X + job.x_off;
        t.range = s.range;
        t.A = s.keys;
        t.identity = l_sub == 0 ? 1 : 0;
        t.level0 = l_sub;
        t.levels = L;
        t.box = s.box;
        t.perm = s.perm;
        t.posof = s.posof;
        t.xp = s.xp;
        for (int k0 = 0; k0 < (1 << l_sub); ++k0) {
            subtree_build<D>(t, k0, lds);
            __syncthreads();
        }
    }

    phase(0);
    // -- kNN of every tree position (numba_kdtree.py:765-792), rows in tree
    // order: columns 1..k-1 as positions and distances, the core by position
    {
        LdsHeap h;
        h.p = reinterpret_cast<float *>(lds) + tid;
        h.i = reinterpret_cast<int *>(lds + (size_t)kq * kBT * sizeof(float)) + tid;
        h.stride = kBT;
        h.k = kq;
        for (int q = tid; q < n; q += kBT) {
            h.init();
            KnnVisitor<D, kBatchLeafBatch> v;
            v.x.load(s.xp + (size_t)q * D, D);
            v.xp = s.xp;
            v.range = s.range;
            v.h = h;
            v.bound = kInf;
            v.hmax = kInf;
            v.tie = 0;
            v.d = D;
            walk<D>(s.box, L, D, v);
            h.sort();
            for (int j = 1; j < kq; ++j) {
                s.nb_pos[(size_t)q * kn + j - 1] = h.I(j);
                s.nb_d[(size_t)q * kn + j - 1] = h.P(j);
            }
            s.core_p[q] = ms > 1 ? h.P(kq - 1) : 0.0f;
        }
    }
    __syncthreads();

    // certification nodes kCertDepth levels above the leaves: which node
    // holds each position and the smallest core inside each (a wave per node)
    const int cert_level = L - 1 - kCertDepth > 0 ? L - 1 - kCertDepth : 0;
    for (int k = wv; k < (1 << cert_level); k += kBW)
        cert_setup_node(cert_level, s.range, s.core_p, s.cert_of_pos, s.mincore, k);
    phase(1);
    // -- kNN-seeded forest (boruvka.py:474-524), edges by point index
    for (int p = tid; p < n; p += kBT) {
        const ForestEdge e = forest_edge_at(p, n, kq, s.nb_pos, ms <= 1 ? s.nb_d : nullptr, s.core_p, s.perm, s.posof);
        const int i = s.perm[p];
        s.succ[i] = e.succ;
        s.dst_raw[i] = e.dst_raw;
        s.w_init[i] = e.w;
        s.parent[i] = i;
        s.interior[p] = 0;
    }
    __syncthreads();
    wg_cycle_keep(n, s.succ, s.S0, s.M0, s.S1, s.M1, s.keep);
    int base = wg_excl_scan(s.keep, s.offs, n, ws);
    // rows in point order, then the unions
    for (int i = tid; i < n; i += kBT) {
        if (!s.keep[i]) continue;
        const int r = s.offs[i];
        src[r] = i;
        dst[r] = s.dst_raw[i];
        wout[r] = s.w_init[i];
        uf_unite(s.parent, i, s.succ[i]);
    }
    __syncthreads();
    for (int i = tid; i < n; i += kBT) s.parent[i] = uf_find(s.parent, i);
    __syncthreads();
    for (int p = tid; p < n; p += kBT) s.comp_p[p] = s.parent[s.perm[p]];
    for (int i = tid; i < n; i += kBT) s.flag[i] = s.parent[i] == i ? 1 : 0;
    __syncthreads();
    int C = wg_excl_scan(s.flag, s.offs, n, ws);
    for (int i = tid; i < n; i += kBT) {
        if (!s.flag[i]) continue;
        s.roots[s.offs[i]] = i;
        s.rootslot[i] = s.offs[i];
    }
    __syncthreads();

    phase(2);
    // -- Boruvka rounds (boruvka.py:604-640)
    QueryArgs qa{};
    qa.xp = s.xp;
    qa.range = s.range;
    qa.box = s.box;
    qa.perm = s.perm;
    qa.core_p = s.core_p;
    qa.comp_p = s.comp_p;
    qa.node_comp = s.node_comp;
    qa.bound = s.bound;
    qa.best = s.best;
    qa.partner = s.partner;
    qa.n = n;
    qa.d = D;
    qa.levels = L;
    qa.pass = 0;
    qa.sample_mask = -1;  // no position counts as sampled
    qa.cert_of_pos = s.cert_of_pos;
    qa.cert = s.cert;
    qa.wave = 1;
    CertArgs ca{};
    ca.xp = s.xp;
    ca.range = s.range;
    ca.box = s.box;
    ca.core_p = s.core_p;
    ca.comp_p = s.comp_p;
    ca.node_comp = s.node_comp;
    ca.bound = s.bound;
    ca.mincore = s.mincore;
    ca.cert = s.cert;
    ca.level = cert_level;
    ca.levels = L;
    ca.d = D;
    float *sx = reinterpret_cast<float *>(lds) + wv * kWaveStage * D;
    float *sc = reinterpret_cast<float *>(lds) + kBW * kWaveStage * D + wv * kWaveStage;
    int *sk = reinterpret_cast<int *>(lds) + kBW * kWaveStage * (D + 1) + wv * kWaveStage;
    int *roots = s.roots, *roots2 = s.roots2;
    bool ok = true;
    while (C > 1) {
        for (int t = tid; t < C; t += kBT) {
            const int r = roots[t];
            s.bound[r] = kInf;
            s.best[r] = ~0ull;
        }
        // node purity (boruvka.py:98-124): leaves, then each level above
        const int nl = 1 << (L - 1);
        for (int k = tid; k < nl; k += kBT) {
            const int node = nl - 1 + k;
            const int2 r = s.range[node];
            const int c0 = s.comp_p[r.x];
            bool pure = true;
            for (int i = r.x + 1; i < r.y; ++i)
                if (s.comp_p[i] != c0) {
                    pure = false;
                    break;
                }
            s.node_comp[node] = pure ? c0 : -1;
        }
        __syncthreads();
        for (int l = L - 2; l >= 0; --l) {
            for (int k = tid; k < (1 << l); k += kBT) {
                const int node = (1 << l) - 1 + k;
                const int x = s.node_comp[2 * node + 1], y = s.node_comp[2 * node + 2];
                s.node_comp[node] = x == y ? x : -1;
            }
            __syncthreads();
        }
        // component bounds from the crossing kNN edges (every lane of a wave
        // takes part in each step)
        for (int p0 = 0; p0 < n; p0 += kBT)
            seed_bounds_at(p0 + tid, p0 + tid < n, kn, s.nb_pos, s.nb_d, s.core_p, s.comp_p, s.bound, s.interior);
        __syncthreads();
        // subtrees whose points cannot reach their component's bound
        // (cert_node: a wave per certification node, its walk queue in LDS)
        for (int k = wv; k < (1 << cert_level); k += kBW)
            cert_node<D>(ca, k, reinterpret_cast<int *>(lds) + wv * kCertQueue);
        __syncthreads();
        // nearest out-of-component edge of every other point, one wavefront
        // per 64 consecutive tree positions
        phase(4);
        for (int p0 = wv * 64; p0 < n; p0 += kBT) wave_search<D, 2, kBatchStageBatch>(qa, p0 + lane, sx, sc, sk);
        __syncthreads();
        phase(3);
        // merge_components (boruvka.py:34-78): each component's edge, the
        // cycle skip, rows in slot (= lowest point) order, unions
        for (int t = tid; t < C; t += kBT) {
            const RoundEdge e = round_edge_of(roots[t], n, s.best, s.partner, s.posof, s.perm, s.comp_p);
            s.succ[t] = s.rootslot[e.tgt];
            s.ei[t] = e.i;
            s.ej[t] = e.j;
            s.ew[t] = e.w;
        }
        __syncthreads();
        wg_cycle_keep(C, s.succ, s.S0, s.M0, s.S1, s.M1, s.keep);
        const int emitted = wg_excl_scan(s.keep, s.offs, C, ws);
        for (int t = tid; t < C; t += kBT) {
            if (!s.keep[t]) continue;
            const int r = base + s.offs[t];
            src[r] = s.ei[t];
            dst[r] = s.ej[t];
            wout[r] = s.ew[t];
            uf_unite(s.parent, roots[t], roots[s.succ[t]]);
        }
        __syncthreads();
        for (int t = tid; t < C; t += kBT) {
            const int r = roots[t];
            s.parent[r] = uf_find(s.parent, r);
        }
        __syncthreads();
        for (int p = tid; p < n; p += kBT) s.comp_p[p] = s.parent[s.comp_p[p]];
        for (int t = tid; t < C; t += kBT) {
            const int r = roots[t];
            s.flag[t] = s.parent[r] == r ? 1 : 0;
        }
        __syncthreads();
        const int left = wg_excl_scan(s.flag, s.offs, C, ws);
        for (int t = tid; t < C; t += kBT) {
            if (!s.flag[t]) continue;
            const int r = roots[t];
            roots2[s.offs[t]] = r;
            s.rootslot[r] = s.offs[t];
        }
        __syncthreads();
        if (emitted != C - left || left >= C) {
            ok = false;  // (the same on every thread: both come from the scans)
            break;
        }
        base += emitted;
        C = left;
        int *t = roots;
        roots = roots2;
        roots2 = t;
    }
    phase(4);
    for (int i = tid; i < n; i += kBT) a.core[job.pt_off + i] = s.core_p[s.posof[i]];
    if (tid == 0) a.status[ji] = ok ? base : -1;
    __syncthreads();
}

template <int D>
__global__ __launch_bounds__(kBT) void batch_kernel(BatchArgs a) {
    __shared__ __align__(16) unsigned char lds[kLds];
    // the job index and the scans' partial sums share the buffer's first
    // bytes with the phases: each is read only between barriers that no
    // phase's use of them spans
    int *s_job = reinterpret_cast<int *>(lds);
    int *ws = s_job + 1;
    const Slice s = slice_at(a.scratch + (int64_t)blockIdx.x * a.slice_bytes, a.max_n, a.max_nodes, a.max_d, a.kq);
    for (;;) {
        if (threadIdx.x == 0) *s_job = atomicAdd(a.next, 1);
        __syncthreads();
        const int ji = *s_job;
        __syncthreads();
        if (ji >= a.count) break;  // every work-group ends once the queue is empty
        run_dataset<D>(a, a.jobs[ji], ji, s, lds, ws);
    }
}

}  // namespace

int batch_max_points(int d) { return d >= 1 && d <= 8 ? kMaxBatchN : 0; }

int batch_subtree_points(int d) {
    switch (d) {
        case 1: return sub_cap<1>();
        case 2: return sub_cap<2>();
        case 3: return sub_cap<3>();
        case 4: return sub_cap<4>();
        case 5: return sub_cap<5>();
        case 6: return sub_cap<6>();
        case 7: return sub_cap<7>();
        case 8: return sub_cap<8>();
        default: return 0;
    }
}

int batch_max_kq() { return kMaxKq; }

int batch_resident_per_cu(int d) {
    int nb = 0;
    hipError_t e = hipErrorInvalidValue;
    switch (d) {
        case 1: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, batch_kernel<1>, kBT, 0); break;
        case 2: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, batch_kernel<2>, kBT, 0); break;
        case 3: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, batch_kernel<3>, kBT, 0); break;
        case 4: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, batch_kernel<4>, kBT, 0); break;
        case 5: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, batch_kernel<5>, kBT, 0); break;
        case 6: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, batch_kernel<6>, kBT, 0); break;
        case 7: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, batch_kernel<7>, kBT, 0); break;
        case 8: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, batch_kernel<8>, kBT, 0); break;
        default: break;
    }
    return e == hipSuccess ? nb : 0;
}

int64_t batch_slice_bytes(int max_n, int max_nodes, int max_d, int kq) {
    return slice_size(max_n, max_nodes, max_d, kq);
}

int batch_grid(int count, int cus) {
    const int g = 2 * (cus > 0 ? cus : 256);
    return count < g ? count : g;
}

hipError_t launch_batch(const BatchArgs &a, int grid, hipStream_t st) {
    if (grid < 1 || a.count < 1) return hipSuccess;
    if (a.kq < 2 || a.kq > kMaxKq) return hipErrorInvalidValue;
    // one dimension per launch: the jobs' d all equal max_d
    switch (a.max_d) {
        case 1: hipLaunchKernelGGL(batch_kernel<1>, dim3(grid), dim3(kBT), 0, st, a); break;
        case 2: hipLaunchKernelGGL(batch_kernel<2>, dim3(grid), dim3(kBT), 0, st, a); break;
        case 3: hipLaunchKernelGGL(batch_kernel<3>, dim3(grid), dim3(kBT), 0, st, a); break;
        case 4: hipLaunchKernelGGL(batch_kernel<4>, dim3(grid), dim3(kBT), 0, st, a); break;
        case 5: hipLaunchKernelGGL(batch_kernel<5>, dim3(grid), dim3(kBT), 0, st, a); break;
        case 6: hipLaunchKernelGGL(batch_kernel<6>, dim3(grid), dim3(kBT), 0, st, a); break;
        case 7: hipLaunchKernelGGL(batch_kernel<7>, dim3(grid), dim3(kBT), 0, st, a); break;
        case 8: hipLaunchKernelGGL(batch_kernel<8>, dim3(grid), dim3(kBT), 0, st, a); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace fh
