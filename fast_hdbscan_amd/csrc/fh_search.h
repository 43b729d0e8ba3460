// Device code of the tree searches shared by the per-call kernels
// (fh_kernels.hip) and the persistent batch kernel (fh_batch.hip): the
// stackless depth-first walk, the kNN and nearest out-of-component visitors,
// the wave-cooperative search, wave-reduced atomics and the lock-free
// union-find.  Internal linkage: each translation unit gets its own copy.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fh_device.h"
#include "fh_kernels.h"

namespace fh {

namespace {

constexpr float kInf = __builtin_huge_valf();

// (idx_start, idx_end) of node i of the implicit tree over n points
// (numba_kdtree.py:426-473: a node's left child takes n_points // 2 of its
// points): the node descends from the root along the bits of its index.
__device__ __forceinline__ int2 node_range_of(int64_t i, int n) {
    int level = 0;
    while (((int64_t(1) << (level + 1)) - 1) <= i) ++level;
    const int64_t k = i - ((int64_t(1) << level) - 1);  // index within the level
    int s = 0, e = n;
    for (int b = level - 1; b >= 0; --b) {
        const int m = s + (e - s) / 2;
        if ((k >> b) & 1) s = m;
        else e = m;
    }
    return make_int2(s, e);
}

// Depth-first walk of the implicit tree without a stack, visiting children
// in the reference's order: the child with the smaller box lower bound
// first, left on ties (numba_kdtree.py:747-760, boruvka.py:228-285).  A bit
// per depth remembers which child went first; on the way back the sibling's
// bound is recomputed with the same arithmetic, so the walk makes exactly the
// reference recursion's sequence of enter/prune decisions.
//
// Each lane walks on its own until it reaches the next leaf it enters (or
// ends); the lanes of a wavefront then scan their leaves together.  Lanes
// whose walks drift apart would otherwise run the leaf scan in almost every
// step of the loop, a few lanes at a time.  A lane's own sequence of
// decisions is unchanged: its bound only changes in its leaf scans.
template <int D, class V>
__device__ __forceinline__ void walk(const float *__restrict__ box, int levels, int d, V &v) {
    const float *x = v.x.p();
    const int stride = 2 * (D > 0 ? D : d);
    int node = 0, depth = 0;
    uint32_t first_right = 0;
    float lb = box_lb<D>(x, box, d);
    bool entering = true;
    for (;;) {
        bool done = false;
        for (;;) {
            if (entering && v.enter(node, lb)) {
                if (depth == levels - 1) break;
                const int l = 2 * node + 1;
                const float lbl = box_lb<D>(x, box + (size_t)l * stride, d);
                const float lbr = box_lb<D>(x, box + (size_t)(l + 1) * stride, d);
                const bool rf = !(lbl <= lbr);
                ++depth;
                first_right = rf ? (first_right | (1u << depth)) : (first_right & ~(1u << depth));
                node = rf ? l + 1 : l;
                lb = rf ? lbr : lbl;
                continue;
            }
            if (depth == 0) {
                done = true;
                break;
            }
            const bool rf = (first_right >> depth) & 1u;
            const bool is_right = (node & 1) == 0;
            if (is_right == rf) {
                node = is_right ? node - 1 : node + 1;
                lb = box_lb<D>(x, box + (size_t)node * stride, d);
                entering = true;
            } else {
                node = (node - 1) >> 1;
                --depth;
                entering = false;
            }
        }
        if (done) break;
        v.leaf(node);
        entering = false;  // then on as from a finished subtree
    }
}

// The query point: held in registers for a compile-time dimension, read
// through a pointer otherwise.
template <int D>
struct QueryPoint {
    float v[D];
    __device__ __forceinline__ void load(const float *row, int) {
#pragma unroll
        for (int j = 0; j < D; ++j) v[j] = row[j];
    }
    __device__ __forceinline__ const float *p() const { return v; }
};

template <>
struct QueryPoint<0> {
    const float *ptr;
    __device__ __forceinline__ void load(const float *row, int) { ptr = row; }
    __device__ __forceinline__ const float *p() const { return ptr; }
};

// ---------------------------------------------------------------- kNN ----

// LB: leaf rows loaded ahead of their distances (registers: LB * D).
template <int D, int LB = 8>
struct KnnVisitor {
    QueryPoint<D> x;
    const float *__restrict__ xp;
    const int2 *__restrict__ range;
    LdsHeap h;
    float hmax;
    float bound;  // +inf: the reference's unbounded query
    int tie;      // bounded query only: an equality that the heap layout decides
    int d;

    __device__ __forceinline__ bool enter(int, float lb) const { return !(lb > hmax); }

    __device__ __forceinline__ void consider(float dist, int i, bool bounded) {
        if (dist < hmax) {
            if (bounded && h.P(0) < kInf && h.k > 1) {
                float c = h.P(1);
                if (h.k > 2) c = h.P(2) > c ? h.P(2) : c;
                if (c == h.P(0)) tie = 1;  // which equal element leaves depends on the layout
            }
            h.push(dist, i);
            hmax = h.P(0) < bound ? h.P(0) : bound;
        } else if (bounded && dist == hmax && hmax == h.P(0)) {
            tie = 1;
        }
    }

    // The leaf's points in position order; for a compile-time dimension
    // their rows are loaded LB at a time ahead of the distances, so
    // the loads overlap (the pushes still follow position order).
    __device__ __forceinline__ void leaf(int node) {
        const int2 r = range[node];
        const bool bounded = bound < kInf;
        if constexpr (D > 0) {
            constexpr int kLeafBatch = LB;
            int i0 = r.x;
            for (; i0 + kLeafBatch <= r.y; i0 += kLeafBatch) {
                float pt[kLeafBatch][D];
#pragma unroll
                for (int u = 0; u < kLeafBatch; ++u)
#pragma unroll
                    for (int j = 0; j < D; ++j) pt[u][j] = xp[(size_t)(i0 + u) * D + j];
#pragma unroll
                for (int u = 0; u < kLeafBatch; ++u) consider(rdist<D>(x.p(), pt[u], D), i0 + u, bounded);
            }
            if (LB > 4 && i0 + 4 <= r.y) {
                float pt[4][D];
#pragma unroll
                for (int u = 0; u < 4; ++u)
#pragma unroll
                    for (int j = 0; j < D; ++j) pt[u][j] = xp[(size_t)(i0 + u) * D + j];
#pragma unroll
                for (int u = 0; u < 4; ++u) consider(rdist<D>(x.p(), pt[u], D), i0 + u, bounded);
                i0 += 4;
            }
            for (int i = i0; i < r.y; ++i) consider(rdist<D>(x.p(), xp + (size_t)i * D, D), i, bounded);
        } else {
            for (int i = r.x; i < r.y; ++i) consider(rdist<D>(x.p(), xp + (size_t)i * d, d), i, bounded);
        }
    }
};

// One lane per query row (queries in tree order for the self-query so a
// wavefront's lanes walk nearby paths).  Heap in LDS, or (GLOBAL_HEAP) in a
// global scratch slice for large k: two instantiations, so that the LDS
// heap's accesses compile to LDS instructions rather than generic ones.
// ----------------------------------------------------- Boruvka query ----

template <int D>
struct BoruvkaVisitor {
    QueryPoint<D> x;
    const float *__restrict__ xp;
    const int2 *__restrict__ range;
    const float *__restrict__ core_p;
    const int *__restrict__ comp_p;
    const int *__restrict__ node_comp;
    float *bound_c;
    float core_i;
    int comp;
    int d;
    float best;
    int best_j;
    float B;

    // boruvka.py:176-192: prune on the point's own best, on the component
    // bound (strict comparisons) and on single-component nodes.
    // B is the component bound as last read (at the start and on entering
    // each leaf): it only ever decreases, so a stale value prunes less but
    // never wrongly.
    __device__ __forceinline__ bool enter(int node, float lb) {
        if (lb > best) return false;
        if (lb > B || core_i > B) return false;
        if (node_comp[node] == comp) return false;
        return true;
    }

    // boruvka.py:196-212.  The candidate filter uses core_j <= B (the
    // reference's sequential run has core_j < B because the bound it sees is
    // still above the component minimum when the winning point runs; with
    // concurrent lanes the bound may already equal it).
    __device__ __forceinline__ void leaf(int node) {
        const int2 r = range[node];
        const int dd = D > 0 ? D : d;
        B = ld_relaxed(bound_c);
        for (int i = r.x; i < r.y; ++i) {
            if (comp_p[i] == comp) continue;
            const float cj = core_p[i];
            if (!(cj <= B)) continue;
            float dist = rdist<D>(x.p(), xp + (size_t)i * dd, d);
            dist = dist > core_i ? dist : core_i;
            dist = dist > cj ? dist : cj;
            if (dist < best) {
                best = dist;
                best_j = i;
                if (dist < B) {
                    atomicMin(reinterpret_cast<unsigned int *>(bound_c), __float_as_uint(dist));
                    B = dist;
                }
            }
        }
    }
};

// atomicMin(base + c, v) for every lane with has set, reduced first over
// the wave's lanes that share c (lanes of consecutive tree positions mostly
// share their component): one atomic per distinct c in the wave.  Every lane
// of the wave must call it.
template <typename T>
__device__ __forceinline__ void wave_atomic_min_by_key(T *base, int c, T v, bool has) {
    const int lane = threadIdx.x & 63;
    unsigned long long pending = __ballot(has);
    while (pending) {
        const int leader = __ffsll((long long)pending) - 1;
        const int c0 = __shfl(c, leader, 64);
        const bool mine = has && c == c0;
        T m = mine ? v : static_cast<T>(~static_cast<T>(0));
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            const T other = __shfl_xor(m, o, 64);
            m = other < m ? other : m;
        }
        if (lane == leader) atomicMin(base + c0, m);
        pending &= ~__ballot(mine);
    }
}

// The kNN-list bound seeding of tree position p (in_range: p < n) by one
// lane; every lane of the wave must call it (wave_atomic_min_by_key).
__device__ __forceinline__ void seed_bounds_at(int64_t p, bool in_range, int kn, const int *__restrict__ nb_pos,
                                               const float *__restrict__ nb_d, const float *__restrict__ core_p,
                                               const int *__restrict__ comp_p, float *bound,
                                               unsigned char *interior) {
    const bool act = in_range && !interior[p];
    int c = -1;
    float w = kInf;
    if (act) {
        c = comp_p[p];
        const float ci = core_p[p];
        bool cross = false;
        // neighbours four at a time: positions first, then their labels and
        // core distances together
        for (int j0 = 0; j0 < kn; j0 += 4) {
            int q[4], cq[4];
            float dq[4], kq[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                q[u] = j0 + u < kn ? nb_pos[p * kn + j0 + u] : -1;
                dq[u] = j0 + u < kn ? nb_d[p * kn + j0 + u] : 0.0f;
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                cq[u] = q[u] >= 0 ? comp_p[q[u]] : c;
                kq[u] = q[u] >= 0 ? core_p[q[u]] : 0.0f;
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                if (q[u] < 0 || cq[u] == c) continue;
                cross = true;
                float e = dq[u];
                e = e > ci ? e : ci;
                e = e > kq[u] ? e : kq[u];
                w = e < w ? e : w;
            }
        }
        if (!cross) interior[p] = 1;
    }
    // non-negative floats order as their bit patterns
    wave_atomic_min_by_key(reinterpret_cast<unsigned int *>(bound), c, __float_as_uint(w), act && w < kInf);
}

// ------------------------------------------------ forest and rounds ----

// The kNN-seeded forest edge of tree position p (boruvka.py:492-503): the
// first neighbour column j >= 1 whose core distance does not exceed the
// point's own, from tree-order rows (k - 1 columns of positions, -1
// padding).  A -1 column (rows with fewer real neighbours than k) reads
// the core of point n - 1, as numba's wrap-around indexing does.  Without
// distance rows (nb_d == nullptr) the core is the k-th distance, so
// max(core_i, d_ij) == core_i.  succ and dst_raw are point indices.
struct ForestEdge {
    int succ, dst_raw;
    float w;
};
__device__ __forceinline__ ForestEdge forest_edge_at(int p, int n, int k, const int *__restrict__ nb_pos,
                                                     const float *__restrict__ nb_d, const float *__restrict__ core_p,
                                                     const int *__restrict__ perm, const int *__restrict__ posof) {
    const float ci = core_p[p];
    const int C = k - 1;
    ForestEdge e{-1, -1, 0.0f};
    for (int j = 0; j < C; ++j) {
        const int q = nb_pos[(int64_t)p * C + j];
        const int kw = q < 0 ? n - 1 : perm[q];
        const int qp = q < 0 ? posof[n - 1] : q;
        if (ci >= core_p[qp]) {
            const float dj = nb_d != nullptr ? nb_d[(int64_t)p * C + j] : ci;
            e.succ = kw;
            e.dst_raw = q < 0 ? -1 : kw;
            e.w = ci > dj ? ci : dj;
            break;
        }
    }
    return e;
}

// A component's edge of one round from its packed best (weight, lowest
// point): the point i, its recorded partner j, the weight, and the label of
// j's component.  A component without any candidate (only when every core
// distance is +inf, i.e. min_samples >= n) keeps the reference's entry
// (lowest point, -1, inf) whose target resolves like ds_find(-1) = the
// component of point n - 1 (boruvka.py:34-78).
struct RoundEdge {
    int i, j, tgt;
    float w;
};
__device__ __forceinline__ RoundEdge round_edge_of(int r, int n, const unsigned long long *__restrict__ best,
                                                   const int *__restrict__ partner, const int *__restrict__ posof,
                                                   const int *__restrict__ perm, const int *__restrict__ comp_p) {
    const unsigned long long key = best[r];
    RoundEdge e;
    if (key == ~0ull) {
        e.i = r;
        e.j = -1;
        e.w = __builtin_huge_valf();
        e.tgt = comp_p[posof[n - 1]];
    } else {
        e.i = (int)(unsigned int)(key & 0xffffffffull);
        e.w = __uint_as_float((unsigned int)(key >> 32));
        const int pj = partner[posof[e.i]];
        e.j = perm[pj];
        e.tgt = comp_p[pj];
    }
    return e;
}

// Lower bound of the squared distance between two boxes, in the same
// sequential fused form as rdist so it never exceeds rdist of any pair of
// points inside them.
template <int D>
__device__ __forceinline__ float box_box_lb(const float *__restrict__ a, const float *__restrict__ b, int d) {
    const int dd = D > 0 ? D : d;
    float r = 0.0f;
#pragma unroll
    for (int j = 0; j < dd; ++j) {
        const float alo = a[j], ahi = a[dd + j], blo = b[j], bhi = b[dd + j];
        const float g1 = blo > ahi ? blo - ahi : 0.0f;
        const float g2 = alo > bhi ? alo - bhi : 0.0f;
        const float g = g1 + g2;
        r = __builtin_fmaf(g, g, r);
    }
    return r;
}

// Orders the LDS accesses of one wavefront's lanes (a wave executes in
// lockstep; this keeps the compiler from moving them across).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Which certification node (in-level index at `level`) holds each position,
// and the smallest core distance inside node k, by one wave.
__device__ __forceinline__ void cert_setup_node(int level, const int2 *__restrict__ range,
                                                const float *__restrict__ core_p, int *cert_of_pos, float *mincore,
                                                int k) {
    const int lane = threadIdx.x & 63;
    const int2 r = range[(1 << level) - 1 + k];
    float m = kInf;
    for (int p = r.x + lane; p < r.y; p += 64) {
        cert_of_pos[p] = k;
        m = core_p[p] < m ? core_p[p] : m;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const float t = __shfl_xor(m, o, 64);
        m = t < m ? t : m;
    }
    if (lane == 0) mincore[k] = m;
}

constexpr int kCertQueue = 1024;  // queued tree nodes of one certification walk

// One wave per certification node P of a single component C: the node
// fails as soon as some point j outside C could give a weight
// max(d(box P, x_j), core_j, min core of P) <= B_C (an existence test, so
// the visiting order does not change the answer).  The wave shares the
// walk: every lane holds one tree node, descends into its left child when
// the box-to-box bound and the single-component test let it pass and
// queues the right child in LDS; lanes without a node take queued ones.  A
// queue overflow leaves the node uncertified (always allowed: the main
// search then simply visits its points).  A certified node's points cannot
// reach B_C (nor, therefore, the component minimum), so the main search
// skips them.
template <int D>
__device__ __forceinline__ void cert_node(const CertArgs &a, int k, int *q) {
    const int lane = threadIdx.x & 63;
    const unsigned long long below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    const int dd = D > 0 ? D : a.d;
    const int stride = 2 * dd;
    const int P = (1 << a.level) - 1 + k;
    const int C = a.node_comp[P];
    if (C < 0) {
        if (lane == 0) a.cert[k] = 0;
        return;
    }
    const float B = ld_relaxed(a.bound + C);
    const float mc = a.mincore[k];
    if (mc > B || !(B < kInf)) {
        if (lane == 0) a.cert[k] = mc > B ? 1 : 0;
        return;
    }
    float pb[2 * (D > 0 ? D : 1)];
    const float *boxP = a.box + (size_t)P * stride;
    if constexpr (D > 0) {
#pragma unroll
        for (int j = 0; j < 2 * D; ++j) pb[j] = boxP[j];
    }
    const float *bp = D > 0 ? pb : boxP;
    const int first_leaf = (1 << (a.levels - 1)) - 1;
    int node = lane == 0 ? 0 : -1;
    int qn = 0;  // queued nodes (the same on every lane)
    bool fail = false, failed = false, overflow = false;
    for (;;) {
        bool push = false;
        int child = -1;
        if (node >= 0) {
            const float lb = box_box_lb<D>(bp, a.box + (size_t)node * stride, a.d);
            if (!(lb > B) && a.node_comp[node] != C) {
                if (node >= first_leaf) {
                    const int2 r = a.range[node];
                    for (int j = r.x; j < r.y; ++j) {
                        if (a.comp_p[j] == C) continue;
                        const float cj = a.core_p[j];
                        if (!(cj <= B)) continue;
                        float w = box_lb<D>(a.xp + (size_t)j * dd, bp, a.d);
                        w = w > cj ? w : cj;
                        w = w > mc ? w : mc;
                        if (!(w > B)) {
                            fail = true;
                            break;
                        }
                    }
                    node = -1;
                } else {
                    child = 2 * node + 2;
                    push = true;
                    node = 2 * node + 1;
                }
            } else {
                node = -1;
            }
        }
        if (__ballot(fail) != 0) {
            failed = true;  // (on every lane: the ballot is the wave's)
            break;
        }
        const unsigned long long pm = __ballot(push);
        const int npush = __popcll(pm);
        if (qn + npush > kCertQueue) {
            overflow = true;
            break;
        }
        if (push) q[qn + __popcll(pm & below)] = child;
        qn += npush;
        wave_sync();
        const unsigned long long im = __ballot(node < 0);
        const int idle = __popcll(im);
        const int take = idle < qn ? idle : qn;
        const int ir = __popcll(im & below);
        if (node < 0 && ir < take) node = q[qn - 1 - ir];
        qn -= take;
        wave_sync();
        if (__ballot(node >= 0) == 0) break;
    }
    if (lane == 0) a.cert[k] = (failed || overflow) ? 0 : 1;
}


// ------------------------------------------ wave-cooperative search ----
//
// The main nearest-edge search with one wavefront per 64 consecutive tree
// positions (spatially adjacent query points).  The wave walks the tree once
// for all its lanes, depth first, the child nearer to the lanes' bounding
// box first; it enters a node when at least one lane passes the per-point
// tests of boruvka.py:176-192 there (lower bound against the lane's own best
// and its component bound, core against the bound, single-component node),
// stages each leaf it reaches in LDS with coalesced loads (points, core
// distances, component labels), and every lane passing the tests at that
// leaf scans the staged points (boruvka.py:196-212).
//
// Exactness (the selection rule of the one-thread reference run): only the
// lanes whose point achieves its component's minimum w* matter, and for such
// a lane every candidate j of weight w* lies in nodes whose lower bound is
// <= w* <= best, <= w* <= B, with core_j <= w* <= B: no test prunes it, so
// the lane scans every candidate of weight w*.  Equal weights are resolved
// by the reference's depth-first order for that point (closer child first,
// left on equal bounds, position order within a leaf), so the lane keeps the
// first such candidate in that order, as the recursion records it.
// Lanes of larger weight report values >= w* that cannot win the packed
// atomicMin unless they also achieve w*.

// True when position j1 comes before j2 in the depth-first visiting order of
// the query for point x (the recursion of boruvka.py:228-285).
template <int D>
__device__ bool dfs_first(const float *x, int j1, int j2, const int2 *__restrict__ range,
                          const float *__restrict__ box, int levels) {
    int node = 0;
    for (int depth = 0; depth < levels - 1; ++depth) {
        const int l = 2 * node + 1;
        const int split = range[l].y;
        const bool a1 = j1 < split, a2 = j2 < split;
        if (a1 == a2) {
            node = a1 ? l : l + 1;
            continue;
        }
        const float lbl = box_lb<D>(x, box + (size_t)l * 2 * D, D);
        const float lbr = box_lb<D>(x, box + (size_t)(l + 1) * 2 * D, D);
        return a1 == (lbl <= lbr);
    }
    return j1 < j2;
}

constexpr int kWaveStage = 128;  // staged points per leaf (leaves hold < 2 * leaf_size = 80)

// The search of one wavefront whose lane l holds tree position pos (its
// wave's positions are consecutive); sx, sc, sk: this wave's LDS staging
// (kWaveStage points).
// SB: staged points read ahead of their tests (registers: SB * (D + 2)).
template <int D, int PASS, int SB = 4>
__device__ __forceinline__ void wave_search(const QueryArgs &a, int64_t pos, float *sx, float *sc, int *sk) {
    const int lane = threadIdx.x & 63;
    bool act = pos < a.n;
    if constexpr (PASS == 2) {
        if (act) {
            const bool sampled = (pos & a.sample_mask) == (a.sample_mask >> 1);
            act = !sampled && !(a.cert && a.cert[a.cert_of_pos[pos]]);
        }
    }
    if (__ballot(act) == 0) return;  // the whole wave has nothing to search
    float x[D];
    float core_i = kInf, B = 0.0f;
    int comp = -1;
    if (act) {
#pragma unroll
        for (int j = 0; j < D; ++j) x[j] = a.xp[pos * D + j];
        comp = a.comp_p[pos];
        core_i = a.core_p[pos];
        B = ld_relaxed(a.bound + comp);
    } else {
#pragma unroll
        for (int j = 0; j < D; ++j) x[j] = 0.0f;
    }
    // the active lanes' bounding box
    float gb[2 * D];
#pragma unroll
    for (int j = 0; j < D; ++j) {
        float lo = act ? x[j] : kInf, hi = act ? x[j] : -kInf;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            const float l2 = __shfl_xor(lo, o, 64), h2 = __shfl_xor(hi, o, 64);
            lo = l2 < lo ? l2 : lo;
            hi = h2 > hi ? h2 : hi;
        }
        gb[j] = lo;
        gb[D + j] = hi;
    }
    float best = kInf;
    int best_j = -1;
    const int levels = a.levels;
    int node = 0, depth = 0;
    uint32_t first_right = 0;
    bool entering = true;
    for (;;) {
        // the walk is the wave's: its node index lives in a scalar register,
        // so the node's box, label and range come through scalar loads
        node = __builtin_amdgcn_readfirstlane(node);
        if (entering) {
            const float *nb = a.box + (size_t)node * 2 * D;
            const float lb = box_lb<D>(x, nb, D);
            const bool t = act && !(core_i > B) && !(lb > best) && !(lb > B) && a.node_comp[node] != comp;
            if (__ballot(t) != 0) {
                if (depth == levels - 1) {
                    const int2 r = a.range[node];
                    const int cnt = r.y - r.x;
                    for (int q = lane; q < cnt; q += 64) {
                        const int64_t g = r.x + q;
#pragma unroll
                        for (int j = 0; j < D; ++j) sx[q * D + j] = a.xp[g * D + j];
                        sc[q] = a.core_p[g];
                        sk[q] = a.comp_p[g];
                    }
                    wave_sync();
                    if (t) {
                        B = ld_relaxed(a.bound + comp);
                        // staged points read SB at a time ahead of the tests
                        // (still taken in position order)
                        auto scan = [&](int q, int kq, float cj, const float *y) {
                            if (kq == comp || !(cj <= B)) return;
                            float dist = rdist<D>(x, y, D);
                            dist = dist > core_i ? dist : core_i;
                            dist = dist > cj ? dist : cj;
                            if (dist < best) {
                                best = dist;
                                best_j = r.x + q;
                                if (dist < B) {
                                    atomicMin(reinterpret_cast<unsigned int *>(a.bound + comp), __float_as_uint(dist));
                                    B = dist;
                                }
                            } else if (dist == best && best_j >= 0 &&
                                       dfs_first<D>(x, r.x + q, best_j, a.range, a.box, levels)) {
                                best_j = r.x + q;
                            }
                        };
                        int q = 0;
                        for (; q + SB <= cnt; q += SB) {
                            int kq[SB];
                            float cq[SB], yq[SB][D];
#pragma unroll
                            for (int u = 0; u < SB; ++u) {
                                kq[u] = sk[q + u];
                                cq[u] = sc[q + u];
#pragma unroll
                                for (int j = 0; j < D; ++j) yq[u][j] = sx[(q + u) * D + j];
                            }
#pragma unroll
                            for (int u = 0; u < SB; ++u) scan(q + u, kq[u], cq[u], yq[u]);
                        }
                        for (; q < cnt; ++q) scan(q, sk[q], sc[q], sx + q * D);
                    }
                    wave_sync();
                } else {
                    const int l = 2 * node + 1;
                    const float lbl = box_box_lb<D>(gb, a.box + (size_t)l * 2 * D, D);
                    const float lbr = box_box_lb<D>(gb, a.box + (size_t)(l + 1) * 2 * D, D);
                    const bool rf = !(lbl <= lbr);
                    ++depth;
                    first_right = rf ? (first_right | (1u << depth)) : (first_right & ~(1u << depth));
                    node = rf ? l + 1 : l;
                    continue;
                }
            }
        }
        if (depth == 0) break;
        const bool rf = (first_right >> depth) & 1u;
        const bool is_right = (node & 1) == 0;
        if (is_right == rf) {
            node = is_right ? node - 1 : node + 1;
            entering = true;
        } else {
            node = (node - 1) >> 1;
            --depth;
            entering = false;
        }
    }
    const bool found = act && best_j >= 0;
    unsigned long long key = ~0ull;
    if (found) {
        key = ((unsigned long long)__float_as_uint(best) << 32) | (unsigned int)a.perm[pos];
        a.partner[pos] = best_j;
    }
    wave_atomic_min_by_key(a.best, comp, key, found);
}

// ------------------------------------------------------- union-find ----
//
// Lock-free union-find over component labels: roots are linked larger
// under smaller with a CAS, so every root is the lowest original point index
// of its set (the label convention above).

__device__ __forceinline__ int uf_find(int *parent, int x) {
    for (;;) {
        const int p = ld_relaxed(parent + x);
        if (p == x) return x;
        x = p;
    }
}

__device__ __forceinline__ void uf_unite(int *parent, int a, int b) {
    for (;;) {
        a = uf_find(parent, a);
        b = uf_find(parent, b);
        if (a == b) return;
        if (a < b) {
            const int t = a;
            a = b;
            b = t;
        }
        if (atomicCAS(parent + a, a, b) == a) return;
    }
}

}  // namespace

}  // namespace fh
