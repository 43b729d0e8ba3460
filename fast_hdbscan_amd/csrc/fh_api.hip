// C-ABI entry points (include/fh.h) and the host driver of the MST path.
//
// Per call (run_mst):
//   1. KD-tree built on the device exactly as numba_kdtree.build_kdtree
//      (fh_tree.hip), rows kept in tree order.
//   2. knn kernel: exact (min_samples+1)-NN of every point, core distances
//      (the MFMA candidate scans of fh_brute.hip in high dimension).
//   3. kNN-seeded forest (boruvka.py:486-524): edge choice, cycle skip by
//      pointer doubling, emit in point order, lock-free union.
//   4. Boruvka rounds (boruvka.py:610-640) until one component is left:
//      node purity, bound seeding, nearest out-of-component search with
//      packed atomicMin per component, per-component selection, cycle skip,
//      emit in order of the component's lowest point, union + relabel, root
//      compaction.  The host reads back two integers per round (edges
//      emitted, components left).
// Many datasets (fh_mst_batch, fh_hdbscan_batch): a measured router splits
// them between persistent work-queue launches (fh_batch.hip: every step of a
// dataset in one work-group) and run_mst on worker streams.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <thread>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <mutex>
#include <random>
#include <string>
#include <vector>

#include "../../include/fh.h"
#include "fh_batch.h"
#include "fh_brute.h"
#include "fh_csr.h"
#include "fh_kernels.h"
#include "fh_mlp.h"
#include "fh_sort.h"
#include "fh_tree.h"
#include "kdtree_build.h"

using fh::HostTree;

namespace {

constexpr int64_t kLeafSize = 40;  // numba_kdtree.py:478 default, used by hdbscan.py:466
constexpr int kCertDepth = 4;      // certification nodes sit this many levels above the leaves
constexpr int kSampleMask = 255;   // one sampled search per 256 tree positions
constexpr int kOrderRound = 1;     // MFMA rounds: rows in component order from this round on

template <class T>
struct DevBuf {
    T *p = nullptr;
    size_t cap = 0;
};

struct Timer {
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    double ms() const {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
};

}  // namespace

struct fh_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    int host_threads = 0;

    DevBuf<float> xp, box, core_p, core_o, knn_d, w_init, bound, ew, heap_p, queries, out_w, out_core;
    DevBuf<int> perm, posof, comp_p, node_comp, knn_i, succ, dst_raw, S0, M0, S1, M1, keep, offs, parent, roots,
        roots2, rootslot, partner, ei, ej, tiles, small, heap_i, flag, out_src, out_dst, out_nbr;
    DevBuf<int2> range;
    DevBuf<unsigned long long> best;
    // device tree build
    DevBuf<float> xdev, wdev;
    DevBuf<uint64_t> tkeys;
    DevBuf<uint32_t> obox;
    DevBuf<int> taxis, tcnt, gtpos, ltpos, tactive, taux;
    DevBuf<fh::SegState> tseg;
    // pruned nearest-edge search
    DevBuf<int> nb_pos, cert_of_pos, sw_pos, sw_ties;
    DevBuf<float> sw_d;
    DevBuf<float> nb_d, mincore;
    DevBuf<unsigned char> cert, seed_int;
    int search_mode = 1;  // 0: one plain pass; 1: seeded + sampled + certified
    // high-dimensional MFMA candidate scans
    DevBuf<float> bx_xc, bx_nrm, bx_thr, bx_rowmin;
    DevBuf<int> bx_cand, bx_cnt, bx_flags, bx_rowarg, bx_rows, bx_small;
    DevBuf<unsigned int> bx_nmax;
    DevBuf<int> bx_tilecomp;
    // rounds in component order with tile-pair bounds
    DevBuf<float> bx_nrm2, bx_core2, bx_tmin, bx_tminc, bx_tmaxc, bx_tU;
    DevBuf<int> bx_comp2, bx_order, bx_ovals;
    DevBuf<uint32_t> bx_okeys, bx_okeys_s;
    DevBuf<unsigned char> bx_otemp;
    DevBuf<unsigned long long> bx_scored, bx_rowkey;
    // tiled exact re-scan: per listed row and work-group, the best weight and position
    DevBuf<float> bx_slotw;
    DevBuf<int> bx_slotj;
    DevBuf<unsigned char> bx_cert;  // first round from the candidate lists: settled rows
    DevBuf<float> bx_cand_d;        // the kNN candidates' exact values, for the first round
    int brute_min_dim = 9;   // d from which the MFMA scans replace the tree searches (use_brute)
    // learned search radii
    DevBuf<float> ml_feat, ml_w, ml_g, ml_m, ml_v, ml_mu, ml_sd, ml_scratch, ml_tgt, ml_pred, ml_bound, ml_stats;
    DevBuf<int> ml_rows, ml_miss;
    int learned_mode = 0;  // 0 off; 1 train on every call; 2 train once, reuse
    int ml_F = 0;          // feature count of the trained model (0: none)
    // kNN-graph Kruskal
    DevBuf<uint64_t> kr_keys, kr_keys_s;
    DevBuf<int> kr_vals, kr_vals_s, kr_src, kr_dst;
    DevBuf<double> kr_w, kr_core;
    int brute_cap = 512;     // kNN candidates kept per row
    // edge sort
    DevBuf<uint32_t> sk1, sk1s;
    DevBuf<uint64_t> sk2, sk2s;
    DevBuf<int> sv1, sv1s;
    // row groups for the host linkage
    DevBuf<uint32_t> pc_keys, pc_keys_s;
    DevBuf<int> pc_parent, pc_vals, pc_vals_s, pc_counts, pc_runs;
    DevBuf<unsigned char> pc_temp;
    DevBuf<unsigned char> stemp;
    // persistent batch kernel: packed rows, jobs, outputs, per-work-group
    // scratch; a pinned host buffer stages the rows up and the outputs down
    DevBuf<float> bt_x, bt_w, bt_core;
    DevBuf<int> bt_src, bt_dst, bt_status, bt_next;
    DevBuf<uint64_t> bt_ticks;
    DevBuf<fh::BatchJob> bt_jobs;
    DevBuf<unsigned char> bt_scratch;
    unsigned char *h_bt = nullptr;
    size_t h_bt_cap = 0;
    int64_t geom_n = -1;            // n the uploaded node ranges belong to
    int64_t tree_n = -1;            // n of the last tree built (perm, posof)
    int geom_maxsize[32] = {};      // largest node per level
    int geom_levels = 0;
    int *h_small = nullptr;  // readback slots: pinned, or a static host array in allocation-free mode
    hipEvent_t ev[10] = {};
    bool timing = true;      // events recorded (allocation-free mode creates none)
    // Allocation-free mode: buffers are carved from this device range in
    // call order, from offset 0 on every call, instead of hipMalloc'd.
    unsigned char *arena = nullptr;
    size_t arena_cap = 0, arena_used = 0;
};

namespace {

// Every device buffer of a context, for release and for the arena reset.
template <class F>
void for_each_buf(fh_ctx *c, F &&f) {
    auto all = [&](auto &...b) { (f(b), ...); };
    all(
        c->xp, c->box, c->core_p, c->core_o, c->knn_d, c->w_init, c->bound, c->ew,
        c->heap_p, c->queries, c->out_w, c->out_core, c->perm, c->posof, c->comp_p, c->node_comp,
        c->knn_i, c->succ, c->dst_raw, c->S0, c->M0, c->S1, c->M1, c->keep,
        c->offs, c->parent, c->roots, c->roots2, c->rootslot, c->partner, c->ei, c->ej,
        c->tiles, c->small, c->heap_i, c->flag, c->out_src, c->out_dst, c->out_nbr, c->range,
        c->best, c->xdev, c->wdev, c->tkeys, c->obox, c->taxis, c->tcnt, c->gtpos,
        c->ltpos, c->tactive, c->taux, c->tseg, c->nb_pos, c->cert_of_pos, c->sw_pos, c->sw_ties, c->sw_d, c->nb_d,
        c->mincore,
        c->cert, c->seed_int, c->bx_xc, c->bx_nrm, c->bx_thr, c->bx_rowmin, c->bx_cand, c->bx_cnt, c->bx_flags,
        c->bx_rowarg, c->bx_rows, c->bx_small, c->bx_nmax, c->bx_tilecomp, c->bx_scored, c->bx_rowkey, c->bx_nrm2, c->bx_core2, c->bx_tmin, c->bx_tminc, c->bx_tmaxc, c->bx_tU, c->bx_comp2, c->bx_order, c->bx_ovals, c->bx_okeys, c->bx_okeys_s, c->bx_otemp, c->bx_slotw, c->bx_slotj, c->bx_cert, c->bx_cand_d, c->ml_feat, c->ml_w, c->ml_g, c->ml_m,
        c->ml_v, c->ml_mu, c->ml_sd, c->ml_scratch, c->ml_tgt, c->ml_pred, c->ml_bound, c->ml_stats,
        c->ml_rows, c->ml_miss, c->kr_keys, c->kr_keys_s, c->kr_vals, c->kr_vals_s, c->kr_src, c->kr_dst,
        c->kr_w, c->kr_core, c->sk1, c->sk1s, c->sk2, c->sk2s, c->sv1, c->sv1s,
        c->pc_keys, c->pc_keys_s, c->pc_parent, c->pc_vals, c->pc_vals_s, c->pc_counts, c->pc_runs, c->pc_temp,
        c->stemp, c->bt_x, c->bt_w, c->bt_core, c->bt_src, c->bt_dst, c->bt_status, c->bt_next, c->bt_jobs,
        c->bt_scratch, c->bt_ticks);
}

}  // namespace

namespace {

int fail(fh_ctx *c, int code, const std::string &msg) {
    c->err = msg;
    return code;
}

#define FH_HIP(call)                                                                                   \
    do {                                                                                               \
        hipError_t e_ = (call);                                                                        \
        if (e_ != hipSuccess)                                                                          \
            return fail(ctx, FH_EDEVICE, std::string(#call) + ": " + hipGetErrorString(e_));           \
    } while (0)

template <class T>
int ensure(fh_ctx *ctx, DevBuf<T> &b, size_t count) {
    if (count == 0) count = 1;
    if (b.cap >= count) return FH_OK;
    if (ctx->arena) {
        // allocation-free mode: the next 256-byte aligned piece of the arena
        const size_t bytes = (count * sizeof(T) + 255) / 256 * 256;
        if (ctx->arena_used + bytes > ctx->arena_cap)
            return fail(ctx, FH_EINVAL,
                        "input needs more than the " + std::to_string(ctx->arena_cap >> 20) +
                            " MiB scratch of the allocation-free entry point (" +
                            std::to_string((ctx->arena_used + bytes) >> 20) + " MiB reached)");
        b.p = reinterpret_cast<T *>(ctx->arena + ctx->arena_used);
        b.cap = count;
        ctx->arena_used += bytes;
        return FH_OK;
    }
    if (b.p) {
        hipError_t e = hipFree(b.p);
        (void)e;
        b.p = nullptr;
        b.cap = 0;
    }
    // Grow with headroom so a sequence of similar sizes does not reallocate.
    size_t want = count + count / 8;
    if (hipMalloc(reinterpret_cast<void **>(&b.p), want * sizeof(T)) != hipSuccess) {
        b.p = nullptr;
        return fail(ctx, FH_ENOMEM, "hipMalloc of " + std::to_string(want * sizeof(T)) + " bytes failed");
    }
    b.cap = want;
    return FH_OK;
}

template <class T>
void release(DevBuf<T> &b) {
    if (b.p) {
        hipError_t e = hipFree(b.p);
        (void)e;
    }
    b.p = nullptr;
    b.cap = 0;
}

#define FH_TRY(x)               \
    do {                        \
        int r_ = (x);           \
        if (r_ != FH_OK) return r_; \
    } while (0)

// Phase marks of run_mst on the context's stream (none without events).
hipError_t mark(fh_ctx *c, int i) { return c->timing ? hipEventRecord(c->ev[i], c->stream) : hipSuccess; }

float event_ms(hipEvent_t a, hipEvent_t b);

float span(fh_ctx *c, int i, int j) { return c->timing ? event_ms(c->ev[i], c->ev[j]) : 0.0f; }

float event_ms(hipEvent_t a, hipEvent_t b) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0.0f;
    return ms;
}

// Node ranges of the implicit tree for n points (numba_kdtree.py:426-473:
// the left child takes n//2 points), computed on the device once per n.  The
// nodes of level l hold floor(n / 2^l) or ceil(n / 2^l) points.
int tree_geometry(fh_ctx *ctx, int64_t n) {
    if (ctx->geom_n == n) return FH_OK;
    const int levels = fh::tree_levels(n, kLeafSize);
    if (levels > 31) return fail(ctx, FH_EUNSUPPORTED, "tree deeper than 31 levels");
    const int64_t N = (int64_t(1) << levels) - 1;
    // every node above the last level must hold >= 2 points so that all
    // leaves sit on the last level (numba_kdtree.py:442-445)
    if (levels >= 2 && (n >> (levels - 2)) < 2) return fail(ctx, FH_EUNSUPPORTED, "tree has a leaf above the last level");
    for (int l = 0; l < levels && l < 32; ++l) ctx->geom_maxsize[l] = (int)((n + (int64_t(1) << l) - 1) >> l);
    FH_TRY(ensure(ctx, ctx->range, N));
    FH_HIP(fh::launch_node_ranges(N, (int)n, ctx->range.p, ctx->stream));
    ctx->geom_levels = levels;
    ctx->geom_n = n;
    return FH_OK;
}

// Exact KD-tree built on the device from points already in HBM; fills xp,
// perm, posof, range, box.
int build_tree_device(fh_ctx *ctx, const float *Xd, int64_t n, int d, int *levels_out, int64_t *nodes_out) {
    FH_TRY(tree_geometry(ctx, n));
    const int L = ctx->geom_levels;
    const int64_t N = (int64_t(1) << L) - 1;
    const int ni = (int)n;
    const int64_t last_count = int64_t(1) << (L - 1);
    FH_TRY(ensure(ctx, ctx->xp, (size_t)n * d));
    FH_TRY(ensure(ctx, ctx->perm, n));
    FH_TRY(ensure(ctx, ctx->posof, n));
    FH_TRY(ensure(ctx, ctx->box, (size_t)N * 2 * d));
    FH_TRY(ensure(ctx, ctx->tkeys, n));
    FH_TRY(ensure(ctx, ctx->obox, (size_t)N * 2 * d));
    FH_TRY(ensure(ctx, ctx->taxis, last_count));
    FH_TRY(ensure(ctx, ctx->tseg, last_count));
    // per (node, chunk, wave of the chunk's work-group) counts
    FH_TRY(ensure(ctx, ctx->tcnt, 4 * ((size_t)n / fh::tree_chunk() + 2 * last_count + 1)));
    FH_TRY(ensure(ctx, ctx->gtpos, n));
    FH_TRY(ensure(ctx, ctx->ltpos, n));
    // step counters (a batch's steps use slots from 0) and the sort count
    constexpr int kSortSlot = 15;
    FH_TRY(ensure(ctx, ctx->tactive, kSortSlot + 1));
    FH_TRY(ensure(ctx, ctx->taux, (size_t)4 * last_count));
    fh::TreeBuildArgs a{};
    a.X = Xd;
    a.d = d;
    a.range = ctx->range.p;
    a.A = ctx->tkeys.p;
    a.obox = ctx->obox.p;
    a.box = ctx->box.p;
    a.axis = ctx->taxis.p;
    a.seg = ctx->tseg.p;
    a.cnt = ctx->tcnt.p;
    a.gtpos = ctx->gtpos.p;
    a.ltpos = ctx->ltpos.p;
    a.active = ctx->tactive.p;
    a.sorts = ctx->tactive.p + kSortSlot;
    a.aux = ctx->taux.p;
    hipStream_t st = ctx->stream;
    const int small = fh::tree_small_limit();
    // steps launched between host checks: a large node needs ~9-14 steps
    // to come down to the LDS range, so the first batch covers most levels
    // in one check; a step whose nodes are all finished launches no-ops
    constexpr int kFirstBatch = 12, kBatch = 4;
    // the first level whose nodes fit one work-group's LDS: from there on
    // one launch builds every level (FH_TREE_SUBTREE=0 keeps the
    // level-by-level kernels down to the leaves)
    int l_sub = L;
    const char *sub_env = std::getenv("FH_TREE_SUBTREE");
    const int sub_cap = (sub_env && std::atoi(sub_env) == 0) ? 0 : fh::tree_subtree_cap(d);
    for (int l = 0; sub_cap > 0 && l < L; ++l)
        if (ctx->geom_maxsize[l] <= sub_cap) {
            l_sub = l;
            break;
        }
    if (l_sub > 0) FH_HIP(fh::tree_identity_keys(ni, a.A, st));
    for (int l = 0; l < l_sub; ++l) {
        const int count = 1 << l;
        const int maxs = ctx->geom_maxsize[l];
        const bool internal = l < L - 1;
        FH_HIP(fh::tree_level_bounds(a, l, maxs, 0, st));
        FH_HIP(fh::tree_level_axis(a, l, internal ? 1 : 0, st));
        if (!internal) break;
        FH_HIP(fh::tree_level_keys(a, l, maxs, 0, st));
        if (maxs > small && maxs <= fh::tree_medium_limit()) {
            // one work-group per node runs the steps down to an LDS-sized range
            FH_HIP(fh::tree_medium_select(a, count, st));
        } else if (maxs > small) {
            // multi-block steps until every node's active range fits in LDS
            FH_HIP(hipMemsetAsync(a.active, 0, sizeof(int) * (kSortSlot + 1), st));
            FH_HIP(fh::tree_big_begin(a, count, st));
            for (int batch = kFirstBatch;; batch = kBatch) {
                FH_HIP(hipMemsetAsync(a.active, 0, sizeof(int) * batch, st));
                for (int b = 0; b < batch; ++b) {
                    fh::TreeBuildArgs ab = a;
                    ab.active = a.active + b;
                    FH_HIP(fh::tree_big_step(ab, count, maxs, st));
                }
                FH_HIP(hipMemcpyAsync(ctx->h_small, a.active, sizeof(int) * (kSortSlot + 1), hipMemcpyDeviceToHost,
                                      st));
                FH_HIP(hipStreamSynchronize(st));
                if (ctx->h_small[batch - 1] == 0) break;
            }
            // ranges that hit the depth limit while still large (rare): sort them
            if (ctx->h_small[kSortSlot] > 0) {
                std::vector<fh::SegState> segs(count);
                FH_HIP(hipMemcpyAsync(segs.data(), a.seg, sizeof(fh::SegState) * count, hipMemcpyDeviceToHost, st));
                FH_HIP(hipStreamSynchronize(st));
                for (int k = 0; k < count; ++k)
                    if (segs[k].flag == fh::SEG_SORT) FH_HIP(fh::tree_sort_range(a.A, segs[k].L, segs[k].R, st));
            }
        }
        FH_HIP(fh::tree_small_select(a, count, maxs, st));
    }
    if (l_sub < L)
        FH_HIP(fh::tree_subtrees(a, l_sub, L, ctx->perm.p, ctx->posof.p, ctx->xp.p, st));
    else
        FH_HIP(fh::tree_finalize(a, ni, ctx->perm.p, ctx->posof.p, ctx->xp.p, st));
    ctx->tree_n = n;
    if (levels_out) *levels_out = L;
    if (nodes_out) *nodes_out = N;
    return FH_OK;
}

int knn_block(int k) {
    if (k <= 32) return 256;
    if (k <= 64) return 128;
    return 64;
}

// kNN of `nq` query rows against the uploaded tree.  Self-query when
// query_perm != nullptr (queries = tree-order rows).
int run_knn(fh_ctx *ctx, int d, int levels, const float *queries, const int *query_perm, int64_t nq, int k,
            float *out_d, int *out_i, float *core_o, float *core_p, int core_from_knn, int *nb_pos = nullptr,
            float *nb_d = nullptr, int sqrt_dist = 0, const int *query_pos = nullptr) {
    fh::KnnArgs a{};
    a.query_pos = query_pos;
    a.sqrt_dist = sqrt_dist;
    a.nb_pos = nb_pos;
    a.nb_d = nb_d;
    a.xp = ctx->xp.p;
    a.range = ctx->range.p;
    a.box = ctx->box.p;
    a.perm = ctx->perm.p;
    a.queries = queries;
    a.query_perm = query_perm;
    a.nq = nq;
    a.d = d;
    a.k = k;
    a.levels = levels;
    a.out_d = out_d;
    a.out_i = out_i;
    a.core_o = core_o;
    a.core_p = core_p;
    a.core_from_knn = core_from_knn;
    if (k <= 128) {
        a.block = knn_block(k);
    } else {
        a.block = 256;
        FH_TRY(ensure(ctx, ctx->heap_p, (size_t)nq * k));
        FH_TRY(ensure(ctx, ctx->heap_i, (size_t)nq * k));
        a.heap_p = ctx->heap_p.p;
        a.heap_i = ctx->heap_i.p;
        a.heap_stride = nq;
    }
    FH_HIP(fh::launch_knn(a, ctx->stream));
    return FH_OK;
}

int read_small(fh_ctx *ctx, int count) {
    FH_HIP(hipMemcpyAsync(ctx->h_small, ctx->small.p, sizeof(int) * count, hipMemcpyDeviceToHost, ctx->stream));
    FH_HIP(hipStreamSynchronize(ctx->stream));
    return FH_OK;
}

// boruvka.py:564-568: k = int(2 * min_samples / mean(w)), numba's float32
// mean (index-order float32 sum, float64 division).
int64_t weighted_k(const float *w, int64_t n, int ms) {
    float acc = 0.0f;
    for (int64_t i = 0; i < n; ++i) acc += w[i];
    const double mean = static_cast<double>(acc) / static_cast<double>(n);
    return static_cast<int64_t>(2.0 * (static_cast<double>(ms) / mean));
}

// kNN of every point by the MFMA candidate scan (fh_brute.hip) with exact
// re-scoring; rows that need the reference's heap order for ties go to the
// tree search.  Leaves the packed points and norms in ctx for the rounds.
// Whether the MFMA candidate scans replace the tree searches (both give the
// same result).  Measured on one MI355X (Gaussian mixtures, min_samples
// 10): the tree kernels are compiled per dimension up to 8 and degrade
// quickly above, while the scans cost the same for any d <= 32 (1.31 s at
// 1M points, 5 ms at 20k).  So:
//   - d >= 16: the scans (16-D at 200k: 54 ms against 1,172 ms);
//   - brute_min_dim (9) <= d < 16: the scans up to n = 4e5 * 2.5^(d - 9)
//     (9-D: 76 against 148 ms at 200k, 1,312 against 878 ms at 1M; 10-D:
//     about even at 1M; 12-D at 100k: 22 against 214 ms);
//   - small inputs from 5 dimensions on (n <= 3000 d), where a tree
//     search visits a large share of the leaves anyway and the per-thread
//     traversal of a few thousand points leaves most of the chip idle.
// FH_HIGH_DIM_FROM, read per call, replaces the rules with one switch-over
// dimension.
bool use_brute(const fh_ctx *ctx, int64_t n, int d, int kq, bool weighted) {
    int from = ctx->brute_min_dim;
    bool small = d >= 5 && n <= 3000 * (int64_t)d;
    double limit = d >= 16 ? 1e300 : 4e5 * std::pow(2.5, d - from);
    if (const char *m = std::getenv("FH_HIGH_DIM_FROM")) {
        from = std::atoi(m);
        small = false;
        limit = 1e300;
    }
    const bool large = d >= from && (double)n <= limit;
    return !weighted && fh::brute_padded_dim(d) > 0 && (large || small) && kq <= 64 && n >= 256;
}

constexpr int kSampleOrder = 8;     // order statistic of the sampled thresholds
constexpr int kRescanBatch = 4096;  // listed rows per tiled exact re-scan launch
constexpr int kCandRoundRows = 65536;  // first round from the candidate lists: most rows left for the exact re-scan
constexpr int kSampleRetry = 2048;  // flagged rows above which the kNN runs again unsampled

int brute_knn(fh_ctx *ctx, int64_t n, int d, int levels, int kq, int core_from_knn, fh::BruteScanArgs &bsa,
          fh_stats &S) {
    hipStream_t st = ctx->stream;
    int dp = fh::brute_padded_dim(d);
    // Products on the FP16 matrix cores from a hi / lo split of the centred
    // rows scaled by s = 2^(14 - e), where 2^e exceeds half the root box's
    // largest extent (so |x' s| < 2^14 and x' s is exact); dp 32 or 64.
    // The split keeps 22 bits of x' s, the hi.lo and lo.hi products are
    // summed before hi.hi, and the scale is exact, so the products' error
    // stays within the FP32 scans' margin (header of fh_brute.hip).
    // FH_MFMA_SPLIT=0: FP32 products (v_mfma_f32_16x16x4f32) throughout.
    bool split = true;
    if (const char *m = std::getenv("FH_MFMA_SPLIT")) split = std::atoi(m) != 0;
    float pack_scale = 0.0f, dot_scale = -2.0f;
    if (split) {
        float hb[2 * FH_MAX_DIM];  // the root box (d <= 64 on this path)
        FH_HIP(hipMemcpyAsync(hb, ctx->box.p, sizeof(float) * 2 * d, hipMemcpyDeviceToHost, st));
        FH_HIP(hipStreamSynchronize(st));
        float half_ext = 0.0f;
        for (int k = 0; k < d; ++k) half_ext = std::max(half_ext, 0.5f * std::fabs(hb[d + k] - hb[k]));
        int e = 0;
        if (half_ext > 0.0f) std::frexp(half_ext, &e);  // half_ext < 2^e
        if (!std::isfinite(half_ext) || e < -40 || e > 60) {
            split = false;
        } else {
            pack_scale = std::ldexp(1.0f, 14 - e);
            dot_scale = -std::ldexp(1.0f, 2 * e - 27);  // -2 / s^2
            dp = d <= 32 ? 32 : 64;
        }
    }
    const float margin_c = 8.0f * dp + 64.0f;
    const int64_t npad = (n + 127) / 128 * 128;
    S.high_dim = 1;
    FH_TRY(ensure(ctx, ctx->bx_xc, (size_t)npad * dp));
    FH_TRY(ensure(ctx, ctx->bx_nrm, npad));
    FH_TRY(ensure(ctx, ctx->bx_nmax, 1));
    FH_TRY(ensure(ctx, ctx->bx_thr, n));
    FH_TRY(ensure(ctx, ctx->bx_cand, (size_t)n * ctx->brute_cap));
    FH_TRY(ensure(ctx, ctx->bx_cnt, n));
    FH_TRY(ensure(ctx, ctx->bx_flags, n));
    FH_TRY(ensure(ctx, ctx->bx_rowmin, n));
    FH_TRY(ensure(ctx, ctx->bx_rowarg, n));
    FH_TRY(ensure(ctx, ctx->bx_rows, n));
    FH_TRY(ensure(ctx, ctx->bx_small, 4));
    FH_HIP(hipMemsetAsync(ctx->bx_nmax.p, 0, sizeof(unsigned int), st));
    FH_HIP(hipMemsetAsync(ctx->bx_small.p, 0, 4 * sizeof(int), st));
    FH_HIP(fh::launch_center_pack(n, d, dp, npad, ctx->xp.p, ctx->box.p, ctx->bx_xc.p, ctx->bx_nrm.p,
                                  ctx->bx_nmax.p, pack_scale, st));
    bsa.split = split ? 1 : 0;
    bsa.dot_scale = dot_scale;
    bsa.pack_scale = pack_scale;
    bsa.xc = ctx->bx_xc.p;
    bsa.nrm = ctx->bx_nrm.p;
    bsa.n = n;
    bsa.npad = npad;
    bsa.dp = dp;
    bsa.thr = ctx->bx_thr.p;
    bsa.cand = ctx->bx_cand.p;
    bsa.cand_cnt = ctx->bx_cnt.p;
    bsa.cap = ctx->brute_cap;
    bsa.core_p = ctx->core_p.p;
    bsa.comp_p = ctx->comp_p.p;
    bsa.row_min = ctx->bx_rowmin.p;
    bsa.row_arg = ctx->bx_rowarg.p;
    bsa.nrm_max_bits = ctx->bx_nmax.p;
    bsa.margin_c = margin_c;
    FH_TRY(ensure(ctx, ctx->bx_tilecomp, npad / 128));
    FH_TRY(ensure(ctx, ctx->bx_scored, 1));
    FH_HIP(hipMemsetAsync(ctx->bx_scored.p, 0, sizeof(unsigned long long), st));
    bsa.tile_comp = ctx->bx_tilecomp.p;
    bsa.scored_tiles = ctx->bx_scored.p;
    // each tile pair once in the rounds' scans (FH_SYM_SCAN=0: every row
    // tile against every column tile)
    const char *sym_env = std::getenv("FH_SYM_SCAN");
    if (!(sym_env && std::atoi(sym_env) == 0)) {
        FH_TRY(ensure(ctx, ctx->bx_rowkey, n));
        bsa.row_key = ctx->bx_rowkey.p;
    }
    // thresholds: MODE 2 writes them where MODE 0 reads them; the
    // candidates are collected once per tile pair, in both directions
    // (FH_SYM_SCAN=0: every pair from both sides).  From 256 tiles on (and
    // k + 1 <= 16) the thresholds come from a sample: the row tile and
    // every 16th tile from it, 8th smallest value (about the 128th nearest
    // neighbour in expectation), one sixteenth of the full scan; the
    // re-scoring checks each row against its threshold and a row that may
    // have missed a neighbour goes to the tree search, or, when more than
    // kSampleRetry rows do, the whole kNN runs again with full thresholds
    // (FH_KNN_SAMPLE: the tile step, 1 = full thresholds; FH_KNN_SAMPLE_RETRY:
    // that row count).
    const int64_t ntl = npad / 128;
    int step = (kq <= 16 && ntl >= 256) ? 16 : 1;
    if (const char *m = std::getenv("FH_KNN_SAMPLE")) step = std::max(1, std::atoi(m));
    fh::BruteScanArgs kth = bsa;
    kth.row_min = ctx->bx_thr.p;
    fh::BruteScanArgs kc = bsa;
    kc.sym_knn = bsa.row_key != nullptr ? 1 : 0;
    fh::KnnRescoreArgs ra{};
    ra.xp = ctx->xp.p;
    ra.perm = ctx->perm.p;
    ra.cand = ctx->bx_cand.p;
    ra.cand_cnt = ctx->bx_cnt.p;
    ra.cap = ctx->brute_cap;
    FH_TRY(ensure(ctx, ctx->bx_cand_d, (size_t)n * ctx->brute_cap));
    ra.cand_d = ctx->bx_cand_d.p;
    ra.n = n;
    ra.d = d;
    ra.k = kq;
    ra.core_from_knn = core_from_knn;
    ra.out_d = ctx->knn_d.p;
    ra.out_i = ctx->knn_i.p;
    ra.core_o = ctx->core_o.p;
    ra.core_p = ctx->core_p.p;
    ra.flag_list = ctx->bx_flags.p;
    ra.flag_count = ctx->bx_small.p;
    ra.thr = ctx->bx_thr.p;
    ra.nrm = ctx->bx_nrm.p;
    ra.nrm_max_bits = ctx->bx_nmax.p;
    ra.margin_c = margin_c;
    int retry = kSampleRetry;
    if (const char *m = std::getenv("FH_KNN_SAMPLE_RETRY")) retry = std::atoi(m);
    int nflag = 0;
    for (;;) {
        kth.tile_step = step;
        kth.cap = step > 1 ? kSampleOrder : kq;
        FH_HIP(fh::launch_scan_kth(kth, st));
        if (kc.sym_knn) FH_HIP(hipMemsetAsync(ctx->bx_cnt.p, 0, sizeof(int) * n, st));
        FH_HIP(fh::launch_scan_knn(kc, st));
        FH_HIP(hipMemsetAsync(ctx->bx_small.p, 0, sizeof(int), st));
        FH_HIP(fh::launch_knn_rescore(ra, st));
        FH_HIP(hipMemcpyAsync(ctx->h_small, ctx->bx_small.p, sizeof(int), hipMemcpyDeviceToHost, st));
        FH_HIP(hipStreamSynchronize(st));
        nflag = ctx->h_small[0];
        if (step == 1 || nflag <= retry) break;
        step = 1;
    }
    S.knn_tree_rows = nflag;
    if (nflag > 0) {
        fh::KnnWaveArgs wa{};
        wa.xp = ctx->xp.p;
        wa.range = ctx->range.p;
        wa.box = ctx->box.p;
        wa.perm = ctx->perm.p;
        wa.query_pos = ctx->bx_flags.p;
        wa.d = d;
        wa.k = kq;
        wa.levels = levels;
        wa.core_from_knn = core_from_knn;
        wa.out_d = ctx->knn_d.p;
        wa.out_i = ctx->knn_i.p;
        wa.core_o = ctx->core_o.p;
        wa.core_p = ctx->core_p.p;
        // a work-group per row (FH_KNN_TREE_FORM=0: a wavefront walking the tree)
        const char *form = std::getenv("FH_KNN_TREE_FORM");
        if (form && std::atoi(form) == 0)
            FH_HIP(fh::launch_knn_wave(wa, nflag, st));
        else
            FH_HIP(fh::launch_knn_block(wa, nflag, n, st));
    }
    return FH_OK;
}

// Learned-radius kNN.  A 1% subsample of tree positions gets the plain
// exact query; its log core distances train the MLP (fh_mlp.hip) on KD-leaf
// features; every position then runs the reference query inside the ball
// the network predicts (plus a residual margin), and positions where the
// ball could change the answer (not filled, or an equality the heap layout
// decides) are queried again without it.  Output identical to the plain
// kNN by construction.
constexpr int kLearnStride = 100;  // training sample: one position in 100
constexpr int kLearnSteps = 120;   // Adam steps per training
constexpr int kLearnBatch = 1024;
constexpr float kLearnRate = 3e-3f;
constexpr float kLearnZ = 2.0f;    // margin in residual standard deviations

int learned_knn(fh_ctx *ctx, int64_t n, int d, int levels, int kq, int *nb_pos, float *nb_d, fh_stats &S) {
    hipStream_t st = ctx->stream;
    const int F = 2 + 2 * d;
    const int P = fh::mlp_param_count(F);
    hipEvent_t e0, e1, e2;
    FH_HIP(hipEventCreate(&e0));
    FH_HIP(hipEventCreate(&e1));
    FH_HIP(hipEventCreate(&e2));
    FH_HIP(hipEventRecord(e0, st));
    FH_TRY(ensure(ctx, ctx->ml_feat, (size_t)n * F));
    FH_TRY(ensure(ctx, ctx->ml_w, P));
    FH_TRY(ensure(ctx, ctx->ml_g, P));
    FH_TRY(ensure(ctx, ctx->ml_m, P));
    FH_TRY(ensure(ctx, ctx->ml_v, P));
    FH_TRY(ensure(ctx, ctx->ml_mu, F));
    FH_TRY(ensure(ctx, ctx->ml_sd, F));
    FH_TRY(ensure(ctx, ctx->ml_pred, n));
    FH_TRY(ensure(ctx, ctx->ml_bound, n));
    FH_TRY(ensure(ctx, ctx->ml_stats, 2));
    FH_TRY(ensure(ctx, ctx->ml_miss, n));
    FH_TRY(ensure(ctx, ctx->bx_small, 4));
    fh::MlpParams mp{ctx->ml_w.p, ctx->ml_g.p, ctx->ml_m.p, ctx->ml_v.p, ctx->ml_mu.p, ctx->ml_sd.p, F};
    FH_HIP(fh::launch_mlp_features(n, d, levels, ctx->xp.p, ctx->range.p, ctx->box.p, ctx->ml_feat.p, st));
    const bool train = ctx->learned_mode == 1 || ctx->ml_F != F;
    // training sample: exact answers for one position in kLearnStride
    const int ntrain = (int)std::max<int64_t>(1, n / kLearnStride);
    std::vector<int> rows(ntrain);
    for (int i = 0; i < ntrain; ++i) rows[i] = (int)std::min<int64_t>(n - 1, (int64_t)i * kLearnStride + kLearnStride / 2);
    FH_TRY(ensure(ctx, ctx->ml_rows, ntrain));
    FH_TRY(ensure(ctx, ctx->ml_tgt, ntrain));
    FH_HIP(hipMemcpyAsync(ctx->ml_rows.p, rows.data(), sizeof(int) * ntrain, hipMemcpyHostToDevice, st));
    FH_TRY(run_knn(ctx, d, levels, nullptr, nullptr, ntrain, kq, ctx->knn_d.p, ctx->knn_i.p, ctx->core_o.p,
                   ctx->core_p.p, 1, nb_pos, nb_d, 0, ctx->ml_rows.p));
    if (train) {
        const int B = std::min(ntrain, kLearnBatch);
        FH_TRY(ensure(ctx, ctx->ml_scratch, (size_t)B * (F + 4 * fh::kMlpHidden + 2)));
        FH_HIP(fh::launch_mlp_log_targets(ntrain, ctx->ml_rows.p, ctx->core_p.p, ctx->ml_tgt.p, st));
        FH_HIP(fh::launch_mlp_standardise(ntrain, F, ctx->ml_feat.p, ctx->ml_rows.p, ctx->ml_mu.p, ctx->ml_sd.p, st));
        FH_HIP(fh::launch_mlp_init(mp, 12345u, st));
        for (int step = 1; step <= kLearnSteps; ++step)
            FH_HIP(fh::launch_mlp_train_step(mp, ntrain, B, (int)(((int64_t)(step - 1) * B) % ntrain),
                                             ctx->ml_feat.p, ctx->ml_rows.p, ctx->ml_tgt.p, ctx->ml_scratch.p,
                                             kLearnRate, step, st));
        // residual margin from the training rows
        FH_HIP(fh::launch_mlp_predict(mp, ntrain, ctx->ml_feat.p, ctx->ml_rows.p, ctx->ml_pred.p, st));
        FH_HIP(fh::launch_mlp_residual(ntrain, ctx->ml_pred.p, ctx->ml_tgt.p, ctx->ml_stats.p, st));
        ctx->ml_F = F;
    }
    FH_HIP(hipEventRecord(e1, st));
    FH_HIP(fh::launch_mlp_predict(mp, n, ctx->ml_feat.p, nullptr, ctx->ml_pred.p, st));
    FH_HIP(fh::launch_mlp_bounds(n, ctx->ml_pred.p, ctx->ml_stats.p, kLearnZ, ctx->ml_bound.p, st));
    FH_HIP(hipEventRecord(e2, st));
    // ball-bounded queries, then the plain query where the ball may matter
    FH_HIP(hipMemsetAsync(ctx->bx_small.p, 0, sizeof(int), st));
    {
        fh::KnnArgs a{};
        a.xp = ctx->xp.p;
        a.range = ctx->range.p;
        a.box = ctx->box.p;
        a.perm = ctx->perm.p;
        a.queries = ctx->xp.p;
        a.query_perm = ctx->perm.p;
        a.nq = n;
        a.d = d;
        a.k = kq;
        a.levels = levels;
        a.out_d = ctx->knn_d.p;
        a.out_i = ctx->knn_i.p;
        a.core_o = ctx->core_o.p;
        a.core_p = ctx->core_p.p;
        a.core_from_knn = 1;
        a.block = knn_block(kq);
        a.nb_pos = nb_pos;
        a.nb_d = nb_d;
        a.bound_sq = ctx->ml_bound.p;
        a.miss_list = ctx->ml_miss.p;
        a.miss_count = ctx->bx_small.p;
        if (kq > 128) return fail(ctx, FH_EUNSUPPORTED, "learned radii need k_query <= 128");
        FH_HIP(fh::launch_knn(a, st));
    }
    FH_HIP(hipMemcpyAsync(ctx->h_small, ctx->bx_small.p, sizeof(int), hipMemcpyDeviceToHost, st));
    FH_HIP(hipStreamSynchronize(st));
    const int misses = ctx->h_small[0];
    if (misses > 0)
        FH_TRY(run_knn(ctx, d, levels, nullptr, nullptr, misses, kq, ctx->knn_d.p, ctx->knn_i.p, ctx->core_o.p,
                       ctx->core_p.p, 1, nb_pos, nb_d, 0, ctx->ml_miss.p));
    S.learned = ctx->learned_mode;
    S.ball_misses = misses;
    S.t_learn_ms = event_ms(e0, e1);
    S.t_predict_ms = event_ms(e1, e2);
    FH_HIP(hipEventDestroy(e0));
    FH_HIP(hipEventDestroy(e1));
    FH_HIP(hipEventDestroy(e2));
    return FH_OK;
}

// The MST proper.  Xd: the points in HBM (original order).  weights_d:
// per-point sample weights in HBM or nullptr; kq_w: the kNN width for the
// weighted branch.  Outputs are device pointers.
// Shared kNN rows of a min_samples sweep: tree-order columns 1..K of the
// (K+1)-nearest query of the context's current tree.
struct SweepRows {
    const int *pos;
    const float *dist;
    int K;
};

int run_mst(fh_ctx *ctx, const float *Xd, int64_t n, int d, int ms, const float *weights_d, int kq_w, float *core_out,
            int *nbr_out, int *src_out, int *dst_out, float *w_out, fh_stats *stats, const SweepRows *sweep = nullptr) {
    Timer total;
    hipStream_t st = ctx->stream;
    fh_stats S{};
    FH_HIP(mark(ctx, 0));

    int levels = 0;
    int64_t n_nodes = 0;
    if (sweep) {
        // the tree of the sweep's points is already built
        if (ctx->tree_n != n) return fail(ctx, FH_EINVAL, "sweep rows belong to another tree");
        levels = ctx->geom_levels;
        n_nodes = (int64_t(1) << levels) - 1;
    } else {
        FH_TRY(build_tree_device(ctx, Xd, n, d, &levels, &n_nodes));
    }
    FH_HIP(mark(ctx, 1));
    S.n_levels = levels;
    S.n_nodes = n_nodes;

    const int ni = (int)n;
    const bool weighted = weights_d != nullptr;
    // boruvka.py:567-569 (weighted), 581-583, 593-595
    const int kq = weighted ? kq_w : (ms > 1 ? ms + 1 : 2);
    FH_TRY(ensure(ctx, ctx->core_p, n));
    FH_TRY(ensure(ctx, ctx->comp_p, n));
    FH_TRY(ensure(ctx, ctx->node_comp, n_nodes));
    FH_TRY(ensure(ctx, ctx->succ, n));
    FH_TRY(ensure(ctx, ctx->dst_raw, n));
    FH_TRY(ensure(ctx, ctx->w_init, n));
    FH_TRY(ensure(ctx, ctx->S0, n));
    FH_TRY(ensure(ctx, ctx->M0, n));
    FH_TRY(ensure(ctx, ctx->S1, n));
    FH_TRY(ensure(ctx, ctx->M1, n));
    FH_TRY(ensure(ctx, ctx->keep, n));
    FH_TRY(ensure(ctx, ctx->offs, n));
    FH_TRY(ensure(ctx, ctx->flag, n));
    FH_TRY(ensure(ctx, ctx->parent, n));
    FH_TRY(ensure(ctx, ctx->roots, n));
    FH_TRY(ensure(ctx, ctx->roots2, n));
    FH_TRY(ensure(ctx, ctx->rootslot, n));
    FH_TRY(ensure(ctx, ctx->partner, n));
    FH_TRY(ensure(ctx, ctx->bound, n));
    FH_TRY(ensure(ctx, ctx->best, n));
    FH_TRY(ensure(ctx, ctx->ei, n));
    FH_TRY(ensure(ctx, ctx->ej, n));
    FH_TRY(ensure(ctx, ctx->ew, n));
    FH_TRY(ensure(ctx, ctx->tiles, fh::scan_tiles(ni) + 1));
    FH_TRY(ensure(ctx, ctx->small, 8));

    // -- core distances
    const bool pruned = ctx->search_mode != 0;
    // Bound seeding reuses the kNN distances as candidate weights; in the
    // weighted branch those are square roots while the search compares
    // squared distances, so seeding is left out there.
    // high dimension: MFMA candidate scans instead of the tree searches
    // (FH_HIGH_DIM_FROM, read per call, moves the switch-over dimension)
    const int dp = fh::brute_padded_dim(d);
    const bool brute = use_brute(ctx, n, d, kq, weighted);
    const bool seed = pruned && !weighted && kq > 1 && !brute;
    if (seed) {
        FH_TRY(ensure(ctx, ctx->nb_pos, (size_t)n * (kq - 1)));
        FH_TRY(ensure(ctx, ctx->nb_d, (size_t)n * (kq - 1)));
    }
    const float margin_c = 8.0f * dp + 64.0f;
    fh::BruteScanArgs bsa{};
    bool knn_d_rows = true;  // the kNN distance rows were written
    const bool learned = !sweep && ctx->learned_mode != 0 && !weighted && ms > 1 && d <= 8 && kq <= 128 && n >= 1000;
    // Tree-order kNN rows (the self-query writes each tree position's
    // columns 1..k-1 as positions and its core by position, all coalesced
    // across a wavefront): the forest and the outputs are then taken from
    // those rows, and no point-order rows are written
    const bool tree_rows = !brute && !learned && !weighted && std::getenv("FH_KNN_ROWS") == nullptr;
    if (!tree_rows) {
        // point-order kNN rows
        FH_TRY(ensure(ctx, ctx->knn_d, (size_t)n * kq));
        FH_TRY(ensure(ctx, ctx->knn_i, (size_t)n * kq));
        FH_TRY(ensure(ctx, ctx->core_o, n));
    }
    if (tree_rows && sweep) {
        // this min_samples' columns of the shared rows; rows whose ms+2
        // smallest values hold an equal pair get the exact (ms+1)-query
        FH_TRY(ensure(ctx, ctx->nb_pos, (size_t)n * (kq - 1)));
        FH_TRY(ensure(ctx, ctx->nb_d, (size_t)n * (kq - 1)));
        FH_TRY(ensure(ctx, ctx->sw_ties, n));
        FH_HIP(hipMemsetAsync(ctx->small.p + 2, 0, sizeof(int), st));
        FH_HIP(fh::launch_sweep_slice(n, sweep->K, ms, sweep->pos, sweep->dist, ctx->nb_pos.p, ctx->nb_d.p,
                                      ctx->core_p.p, ctx->sw_ties.p, ctx->small.p + 2, st));
        FH_HIP(hipMemcpyAsync(ctx->h_small + 2, ctx->small.p + 2, sizeof(int), hipMemcpyDeviceToHost, st));
        FH_HIP(hipStreamSynchronize(st));
        const int ties = ctx->h_small[2];
        S.knn_tree_rows = ties;
        if (ties > 0)
            FH_TRY(run_knn(ctx, d, levels, nullptr, nullptr, ties, kq, nullptr, nullptr, nullptr, ctx->core_p.p, 1,
                           ctx->nb_pos.p, ctx->nb_d.p, 0, ctx->sw_ties.p));
    } else if (tree_rows) {
        FH_TRY(ensure(ctx, ctx->nb_pos, (size_t)n * (kq - 1)));
        FH_TRY(ensure(ctx, ctx->nb_d, (size_t)n * (kq - 1)));
        FH_TRY(run_knn(ctx, d, levels, ctx->xp.p, ctx->perm.p, n, kq, nullptr, nullptr, nullptr, ctx->core_p.p,
                       ms > 1 ? 1 : 0, ctx->nb_pos.p, ctx->nb_d.p, 0));
    } else if (brute) {
        FH_TRY(brute_knn(ctx, n, d, levels, kq, ms > 1 ? 1 : 0, bsa, S));
    } else if (learned) {
        FH_TRY(learned_knn(ctx, n, d, levels, kq, seed ? ctx->nb_pos.p : nullptr, seed ? ctx->nb_d.p : nullptr, S));
    } else {
        // the distance rows are read only by the weighted core and, when
        // the core is not a kNN column (min_samples == 1), by the forest
        // (with the core the k-th distance, max(core_i, d_ij) == core_i)
        knn_d_rows = weighted || ms <= 1;
        FH_TRY(run_knn(ctx, d, levels, ctx->xp.p, ctx->perm.p, n, kq, knn_d_rows ? ctx->knn_d.p : nullptr, ctx->knn_i.p,
                       ctx->core_o.p, ctx->core_p.p, (ms > 1 && !weighted) ? 1 : 0, seed ? ctx->nb_pos.p : nullptr,
                       seed ? ctx->nb_d.p : nullptr, weighted ? 1 : 0));
    }
    if (weighted)
        FH_HIP(fh::launch_weighted_core(ni, kq, ctx->knn_i.p, ctx->knn_d.p, weights_d, ms, ctx->posof.p, ctx->core_o.p,
                                        ctx->core_p.p, st));
    FH_HIP(mark(ctx, 2));
    // certification nodes: kCertDepth levels above the leaves
    const int cert_level = std::max(0, levels - 1 - kCertDepth);
    if (pruned && !brute) {
        FH_TRY(ensure(ctx, ctx->cert_of_pos, n));
        FH_TRY(ensure(ctx, ctx->mincore, int64_t(1) << cert_level));
        FH_TRY(ensure(ctx, ctx->cert, int64_t(1) << cert_level));
        FH_HIP(fh::launch_cert_setup(cert_level, ctx->range.p, ctx->core_p.p, ctx->cert_of_pos.p, ctx->mincore.p, st));
    }

    // -- kNN-seeded forest
    fh::CycleBufs cb{ctx->S0.p, ctx->M0.p, ctx->S1.p, ctx->M1.p};
    if (tree_rows)
        FH_HIP(fh::launch_init_edges_pos(ni, kq, ctx->nb_pos.p, ms <= 1 ? ctx->nb_d.p : nullptr, ctx->core_p.p,
                                         ctx->perm.p, ctx->posof.p, ctx->succ.p, ctx->dst_raw.p, ctx->w_init.p,
                                         ctx->ew.p, st));
    else
        FH_HIP(fh::launch_init_edges(ni, kq, ctx->knn_i.p, knn_d_rows ? ctx->knn_d.p : nullptr, ctx->core_o.p,
                                     ctx->succ.p, ctx->dst_raw.p, ctx->w_init.p, st));
    // Cycles of the forest's functional graph: 2-cycles settled at once,
    // pointer doubling only over points whose successor has an equal core
    // distance (FH_TIE_CYCLES: from how many points; flag, offs and parent
    // are free until the emission)
    int tie_cycles = 4096;
    if (const char *m = std::getenv("FH_TIE_CYCLES")) tie_cycles = std::atoi(m);
    if (ni >= tie_cycles) {
        FH_HIP(fh::launch_cycle_classify(ni, ctx->succ.p, tree_rows ? ctx->ew.p : ctx->core_o.p, ctx->keep.p,
                                         ctx->flag.p, st));
        FH_HIP(fh::launch_excl_scan(ni, ctx->flag.p, ctx->offs.p, ctx->tiles.p, ctx->small.p + 2, st));
        FH_HIP(hipMemcpyAsync(ctx->h_small + 2, ctx->small.p + 2, sizeof(int), hipMemcpyDeviceToHost, st));
        FH_HIP(hipStreamSynchronize(st));
        S.tie_cycle_nodes += ctx->h_small[2];
        FH_HIP(fh::launch_cycle_ties(ni, ctx->h_small[2], ctx->succ.p, ctx->flag.p, ctx->offs.p, cb, ctx->parent.p,
                                     ctx->keep.p, st));
    } else {
        FH_HIP(fh::launch_cycle_keep(ni, ctx->succ.p, cb, ctx->keep.p, st));
    }
    FH_HIP(fh::launch_excl_scan(ni, ctx->keep.p, ctx->offs.p, ctx->tiles.p, ctx->small.p + 0, st));
    FH_HIP(fh::launch_init_emit_union(ni, ctx->keep.p, ctx->offs.p, ctx->succ.p, ctx->dst_raw.p, ctx->w_init.p,
                                      ctx->parent.p, src_out, dst_out, w_out, st));
    FH_HIP(fh::launch_labels_and_root_flags(ni, ctx->perm.p, ctx->parent.p, ctx->comp_p.p, ctx->flag.p, st));
    FH_HIP(fh::launch_excl_scan(ni, ctx->flag.p, ctx->offs.p, ctx->tiles.p, ctx->small.p + 1, st));
    FH_HIP(fh::launch_root_scatter(ni, ctx->flag.p, ctx->offs.p, ctx->roots.p, ctx->rootslot.p, st));
    FH_TRY(read_small(ctx, 2));
    int base = ctx->h_small[0];
    int C = ctx->h_small[1];
    S.initial_components = C;
    FH_HIP(mark(ctx, 3));

    // -- Boruvka rounds
    fh::QueryArgs qa{};
    qa.xp = ctx->xp.p;
    qa.range = ctx->range.p;
    qa.box = ctx->box.p;
    qa.perm = ctx->perm.p;
    qa.core_p = ctx->core_p.p;
    qa.comp_p = ctx->comp_p.p;
    qa.node_comp = ctx->node_comp.p;
    qa.bound = ctx->bound.p;
    qa.best = ctx->best.p;
    qa.partner = ctx->partner.p;
    qa.n = n;
    qa.d = d;
    qa.levels = levels;
    qa.pass = 0;
    qa.sample_mask = kSampleMask;
    // the wave-cooperative main search stages one leaf (< 2 * leaf_size
    // points) per wave in LDS; FH_QUERY_KERNEL=lane keeps one lane per point
    qa.wave = (d <= 8 && ctx->geom_maxsize[levels - 1] <= 128) ? 1 : 0;
    if (const char *m = std::getenv("FH_QUERY_KERNEL")) qa.wave = std::strcmp(m, "lane") == 0 ? 0 : qa.wave;
    if (const char *m = std::getenv("FH_SAMPLE_MASK")) qa.sample_mask = std::atoi(m);
    qa.cert_of_pos = pruned ? ctx->cert_of_pos.p : nullptr;
    qa.cert = pruned ? ctx->cert.p : nullptr;
    fh::CertArgs ca{};
    ca.xp = ctx->xp.p;
    ca.range = ctx->range.p;
    ca.box = ctx->box.p;
    ca.core_p = ctx->core_p.p;
    ca.comp_p = ctx->comp_p.p;
    ca.node_comp = ctx->node_comp.p;
    ca.bound = ctx->bound.p;
    ca.mincore = pruned ? ctx->mincore.p : nullptr;
    ca.cert = pruned ? ctx->cert.p : nullptr;
    ca.level = cert_level;
    ca.levels = levels;
    ca.d = d;
    double query_ms = 0.0, query_main_ms = 0.0;
    int rounds = 0;
    // Which bound-tightening passes run before the main search (each only
    // prunes; outputs never change): FH_SEED_ROUNDS kNN seeding for the first
    // rounds only, FH_SAMPLED=1 adds the sampled pass (one lane per sampled
    // point; it paid before the main search was wave-cooperative, and costs
    // more than it prunes since: 2-D 2M 14.8 -> 13.7 ms, 3-D 10M 127 -> 116
    // ms without it), FH_CERT=0 drops subtree certification.
    struct {
        int seed_rounds = 1 << 30;
        bool sampled = false, cert = true;
        int small_round = fh::small_round_limit();
        int tie_cycles = 1 << 17;  // below this a round's doubling costs less than the host read
    } tune;
    int64_t tie_nodes = 0;
    tune.sampled = !qa.wave;
    if (const char *m = std::getenv("FH_SEED_ROUNDS")) tune.seed_rounds = std::atoi(m);
    if (const char *m = std::getenv("FH_SAMPLED")) tune.sampled = std::atoi(m) != 0;
    if (const char *m = std::getenv("FH_CERT")) tune.cert = std::atoi(m) != 0;
    // FH_SMALL_ROUND caps the component count below which a round's
    // selection, cycle skip, emission and unions run in one work-group
    if (const char *m = std::getenv("FH_SMALL_ROUND"))
        tune.small_round = std::min(std::atoi(m), fh::small_round_limit());
    // FH_TIE_CYCLES: the component count from which a round settles its
    // 2-cycles directly and doubles only over equal-weight chains (a large
    // value keeps the doubling over every component)
    if (const char *m = std::getenv("FH_TIE_CYCLES")) tune.tie_cycles = std::atoi(m);
    if (seed) {
        FH_TRY(ensure(ctx, ctx->seed_int, n));
        FH_HIP(hipMemsetAsync(ctx->seed_int.p, 0, (size_t)n, st));
    }
    fh::BoruvkaExactArgs be{};
    be.xp = ctx->xp.p;
    be.range = ctx->range.p;
    be.box = ctx->box.p;
    be.perm = ctx->perm.p;
    be.core_p = ctx->core_p.p;
    be.comp_p = ctx->comp_p.p;
    be.nrm = ctx->bx_nrm.p;
    be.nrm_max_bits = ctx->bx_nmax.p;
    be.margin_c = margin_c;
    be.row_min = ctx->bx_rowmin.p;
    be.row_arg = ctx->bx_rowarg.p;
    be.bound = ctx->bound.p;
    be.rows = ctx->bx_rows.p;
    be.row_count = ctx->bx_small.p + 1;
    be.best = ctx->best.p;
    be.partner = ctx->partner.p;
    be.n = n;
    be.d = d;
    be.levels = levels;
    int64_t exact_rows = 0;
    int scan_rounds = 0;
    // exact re-scan of the listed rows: tiled (64 rows per pass over the
    // positions), or one work-group per row (FH_TILED_RESCAN=0)
    bool tiled_rescan = true;
    if (const char *m = std::getenv("FH_TILED_RESCAN")) tiled_rescan = std::atoi(m) != 0;
    // first round from the kNN candidate lists (FH_CAND_ROUND=0: the MFMA scan)
    bool cand_round = brute;
    if (const char *m = std::getenv("FH_CAND_ROUND")) cand_round = cand_round && std::atoi(m) != 0;
    // rows in component order with tile-pair bounds in the MFMA rounds
    // (symmetric scans only; FH_TILE_BOUNDS=0 keeps tree order)
    bool tile_bounds = brute && bsa.row_key != nullptr;
    if (const char *m = std::getenv("FH_TILE_BOUNDS")) tile_bounds = tile_bounds && std::atoi(m) != 0;
    while (C > 1) {
        FH_HIP(fh::launch_round_reset(C, ctx->roots.p, ctx->bound.p, ctx->best.p, st));
        if (!brute) FH_HIP(fh::launch_purity(levels, ctx->range.p, ctx->comp_p.p, ctx->node_comp.p, st));
        FH_HIP(mark(ctx, 5));
        if (brute) {
            // approximate row minima on MFMA, exact upper bound per
            // component, exact re-scan of the rows that can reach it.  From
            // round kOrderRound on (symmetric scans) the scans take the rows
            // in component order, fixed at that round (components only
            // merge, so its single-component tiles stay so), and that
            // round's scan records every tile pair's smallest approximate
            // distance; later rounds bound their components from those
            // minima and skip the tile pairs that no row of either tile can
            // use (FH_TILE_BOUNDS=0: every round in tree order, no skips).
            const int64_t npad = bsa.npad, ntl = npad / 128;
            const bool order_rows = tile_bounds && rounds >= kOrderRound;
            if (order_rows && rounds == kOrderRound) {
                FH_TRY(ensure(ctx, ctx->bx_order, npad));
                FH_TRY(ensure(ctx, ctx->bx_okeys, n));
                FH_TRY(ensure(ctx, ctx->bx_okeys_s, n));
                FH_TRY(ensure(ctx, ctx->bx_ovals, n));
                const size_t tb = fh::sort_u32_pairs_temp_bytes(ni);
                FH_TRY(ensure(ctx, ctx->bx_otemp, tb));
                FH_TRY(ensure(ctx, ctx->bx_nrm2, npad));
                FH_TRY(ensure(ctx, ctx->bx_core2, npad));
                FH_TRY(ensure(ctx, ctx->bx_comp2, npad));
                FH_TRY(ensure(ctx, ctx->bx_tminc, ntl));
                FH_TRY(ensure(ctx, ctx->bx_tmaxc, ntl));
                FH_TRY(ensure(ctx, ctx->bx_tU, ntl));
                FH_TRY(ensure(ctx, ctx->bx_tmin, (size_t)(ntl * ntl)));
                FH_HIP(fh::launch_order_keys(n, ctx->comp_p.p, ctx->bx_okeys.p, ctx->bx_ovals.p, st));
                FH_HIP(fh::sort_u32_pairs(ni, ctx->bx_okeys.p, ctx->bx_okeys_s.p, ctx->bx_ovals.p, ctx->bx_order.p,
                                          ctx->bx_otemp.p, tb, st));
                // the packed rows of the kNN scans are not read again
                FH_HIP(fh::launch_order_pack(n, d, bsa.dp, npad, ctx->bx_order.p, ctx->xp.p, ctx->box.p, ctx->core_p.p,
                                             ctx->bx_xc.p, ctx->bx_nrm2.p, ctx->bx_core2.p, bsa.pack_scale, st));
                FH_HIP(fh::launch_tile_core_range(n, npad, ctx->bx_core2.p, ctx->bx_tminc.p, ctx->bx_tmaxc.p, st));
                FH_HIP(hipMemsetAsync(ctx->bx_tmin.p, 0, sizeof(float) * ntl * ntl, st));
            }
            // the first round from the kNN candidate lists: rows settled
            // there enter their component's minimum at once, the others
            // that can reach its bound are listed for the exact re-scan;
            // with more than kCandRoundRows of those, the MFMA scan lists
            // the rows instead (the settled rows' entries stand)
            bool from_cands = false;
            if (rounds == 0 && cand_round) {
                FH_TRY(ensure(ctx, ctx->bx_cert, n));
                fh::CandRoundArgs cr{};
                cr.xp = ctx->xp.p;
                cr.range = ctx->range.p;
                cr.box = ctx->box.p;
                cr.levels = levels;
                cr.d = d;
                cr.n = n;
                cr.perm = ctx->perm.p;
                cr.core_p = ctx->core_p.p;
                cr.comp_p = ctx->comp_p.p;
                cr.cand = bsa.cand;
                cr.cand_cnt = bsa.cand_cnt;
                cr.cand_d = ctx->bx_cand_d.p;
                cr.cap = bsa.cap;
                cr.thr = bsa.thr;
                cr.nrm = bsa.nrm;
                cr.nrm_max_bits = bsa.nrm_max_bits;
                cr.margin_c = bsa.margin_c;
                cr.bound = ctx->bound.p;
                cr.best = ctx->best.p;
                cr.partner = ctx->partner.p;
                cr.cert = ctx->bx_cert.p;
                cr.rows = be.rows;
                cr.row_count = be.row_count;
                // the main-pass span times the MFMA round scans only
                FH_HIP(mark(ctx, 8));
                FH_HIP(mark(ctx, 9));
                FH_HIP(hipMemsetAsync(be.row_count, 0, sizeof(int), st));
                FH_HIP(fh::launch_cand_round(cr, st));
                FH_HIP(hipMemcpyAsync(ctx->h_small + 4, be.row_count, sizeof(int), hipMemcpyDeviceToHost, st));
                FH_HIP(hipStreamSynchronize(st));
                int cap_rows = kCandRoundRows;
                if (const char *m = std::getenv("FH_CAND_ROUND_ROWS")) cap_rows = std::atoi(m);
                from_cands = ctx->h_small[4] <= cap_rows;
            }
            if (!from_cands) {
                ++scan_rounds;
                fh::BruteScanArgs rs = bsa;
                if (order_rows) {
                    FH_HIP(fh::launch_gather_comp(n, npad, ctx->bx_order.p, ctx->comp_p.p, ctx->bx_comp2.p, st));
                    FH_HIP(fh::launch_tile_comp(n, npad, ctx->bx_comp2.p, ctx->bx_tilecomp.p, st));
                    rs.nrm = ctx->bx_nrm2.p;
                    rs.core_p = ctx->bx_core2.p;
                    rs.comp_p = ctx->bx_comp2.p;
                    if (rounds == kOrderRound) {
                        rs.tmin_out = ctx->bx_tmin.p;
                    } else {
                        FH_HIP(fh::launch_tile_bounds((int)ntl, ctx->bx_tilecomp.p, ctx->bx_tmin.p, ctx->bx_tmaxc.p,
                                                      bsa.margin_c, bsa.nrm_max_bits, ctx->bound.p, st));
                        FH_HIP(fh::launch_tile_umax(n, npad, ctx->bx_comp2.p, ctx->bound.p, ctx->bx_tU.p, st));
                        rs.tmin_in = ctx->bx_tmin.p;
                        rs.tminc = ctx->bx_tminc.p;
                        rs.tU = ctx->bx_tU.p;
                    }
                } else {
                    FH_HIP(fh::launch_tile_comp(n, npad, ctx->comp_p.p, ctx->bx_tilecomp.p, st));
                }
                FH_HIP(mark(ctx, 8));
                if (rs.row_key) FH_HIP(hipMemsetAsync(rs.row_key, 0xff, sizeof(unsigned long long) * n, st));
                FH_HIP(fh::launch_scan_boruvka(rs, st));
                if (rs.row_key && order_rows)
                    FH_HIP(fh::launch_unpack_row_keys_ordered(n, rs.row_key, ctx->bx_order.p, rs.row_min, rs.row_arg, st));
                else if (rs.row_key)
                    FH_HIP(fh::launch_unpack_row_keys(n, rs.row_key, rs.row_min, rs.row_arg, st));
                FH_HIP(mark(ctx, 9));
                // components' upper bounds from the approximate minima plus
                // the margin (FH_EXACT_UPPER=1: exact weights of the row minima)
                const char *eu = std::getenv("FH_EXACT_UPPER");
                if (eu && std::atoi(eu) != 0)
                    FH_HIP(fh::launch_boruvka_upper(be, st));
                else
                    FH_HIP(fh::launch_boruvka_upper_approx(be, st));
                FH_HIP(hipMemsetAsync(be.row_count, 0, sizeof(int), st));
                FH_HIP(fh::launch_boruvka_rows(be, st));
            }
            if (tiled_rescan) {
                // the listed rows 64 at a time against every position, in
                // batches of kRescanBatch rows (their slots reused)
                FH_HIP(hipMemcpyAsync(ctx->h_small + 4, be.row_count, sizeof(int), hipMemcpyDeviceToHost, st));
                FH_HIP(hipStreamSynchronize(st));
                const int R = ctx->h_small[4];
                const int RB = std::min(R, kRescanBatch);
                if (RB > 0) {
                    const size_t ns = (size_t)fh::boruvka_rows_tiled_slots(n, RB);
                    FH_TRY(ensure(ctx, ctx->bx_slotw, ns));
                    FH_TRY(ensure(ctx, ctx->bx_slotj, ns));
                }
                for (int r0 = 0; r0 < R; r0 += RB)
                    FH_HIP(fh::launch_boruvka_rows_tiled(be, r0, std::min(RB, R - r0), ctx->bx_slotw.p,
                                                         ctx->bx_slotj.p, st));
            } else {
                FH_HIP(fh::launch_boruvka_row_scan(be, st));
            }
        } else if (pruned) {
            // bounds from kNN cross edges, then a sampled search, then
            // certification of whole subtrees, then everyone else
            if (seed && rounds < tune.seed_rounds)
                FH_HIP(fh::launch_seed_bounds(n, kq - 1, ctx->nb_pos.p, ctx->nb_d.p, ctx->core_p.p, ctx->comp_p.p,
                                              ctx->bound.p, ctx->seed_int.p, st));
            if (tune.sampled) {
                qa.pass = 1;
                FH_HIP(fh::launch_query(qa, st));
            } else {
                qa.sample_mask = -1;  // no position counts as sampled
            }
            qa.cert = tune.cert ? ctx->cert.p : nullptr;
            if (tune.cert) FH_HIP(fh::launch_cert(ca, st));
            qa.pass = 2;
            FH_HIP(mark(ctx, 8));
            FH_HIP(fh::launch_query(qa, st));
            FH_HIP(mark(ctx, 9));
        } else {
            qa.pass = 0;
            FH_HIP(mark(ctx, 8));
            FH_HIP(fh::launch_query(qa, st));
            FH_HIP(mark(ctx, 9));
        }
        FH_HIP(mark(ctx, 6));
        if (C <= tune.small_round) {
            fh::SmallRoundArgs sr{};
            sr.c = C;
            sr.n = ni;
            sr.base = (int)base;
            sr.roots = ctx->roots.p;
            sr.best = ctx->best.p;
            sr.partner = ctx->partner.p;
            sr.posof = ctx->posof.p;
            sr.perm = ctx->perm.p;
            sr.comp_p = ctx->comp_p.p;
            sr.rootslot = ctx->rootslot.p;
            sr.parent = ctx->parent.p;
            sr.roots_out = ctx->roots2.p;
            sr.src_out = src_out;
            sr.dst_out = dst_out;
            sr.w_out = w_out;
            sr.small = ctx->small.p;
            FH_HIP(fh::launch_round_small(sr, st));
            FH_HIP(fh::launch_relabel(ni, ctx->parent.p, ctx->comp_p.p, st));
        } else {
            FH_HIP(fh::launch_select(C, ni, ctx->roots.p, ctx->best.p, ctx->partner.p, ctx->posof.p, ctx->perm.p,
                                     ctx->comp_p.p, ctx->rootslot.p, ctx->succ.p, ctx->ei.p, ctx->ej.p, ctx->ew.p, st));
            if (C >= tune.tie_cycles) {
                // 2-cycles settled at once; doubling only over the nodes with
                // an equal-weight successor (flag, offs and dst_raw are free
                // until the emission below)
                FH_HIP(fh::launch_cycle_classify(C, ctx->succ.p, ctx->ew.p, ctx->keep.p, ctx->flag.p, st));
                FH_HIP(fh::launch_excl_scan(C, ctx->flag.p, ctx->offs.p, ctx->tiles.p, ctx->small.p + 2, st));
                FH_HIP(hipMemcpyAsync(ctx->h_small + 2, ctx->small.p + 2, sizeof(int), hipMemcpyDeviceToHost, st));
                FH_HIP(hipStreamSynchronize(st));
                tie_nodes += ctx->h_small[2];
                FH_HIP(fh::launch_cycle_ties(C, ctx->h_small[2], ctx->succ.p, ctx->flag.p, ctx->offs.p, cb,
                                             ctx->dst_raw.p, ctx->keep.p, st));
            } else {
                FH_HIP(fh::launch_cycle_keep(C, ctx->succ.p, cb, ctx->keep.p, st));
            }
            FH_HIP(fh::launch_excl_scan(C, ctx->keep.p, ctx->offs.p, ctx->tiles.p, ctx->small.p + 0, st));
            FH_HIP(fh::launch_round_emit_union(C, base, ctx->keep.p, ctx->offs.p, ctx->ei.p, ctx->ej.p, ctx->ew.p,
                                               ctx->roots.p, ctx->succ.p, ctx->parent.p, src_out, dst_out, w_out, st));
            FH_HIP(fh::launch_relabel_and_flags(ni, C, ctx->roots.p, ctx->parent.p, ctx->comp_p.p, ctx->flag.p, st));
            FH_HIP(fh::launch_excl_scan(C, ctx->flag.p, ctx->offs.p, ctx->tiles.p, ctx->small.p + 1, st));
            FH_HIP(fh::launch_roots_compact(C, ctx->roots.p, ctx->flag.p, ctx->offs.p, ctx->roots2.p, ctx->rootslot.p,
                                            st));
        }
        FH_TRY(read_small(ctx, 2));
        if (brute) {
            FH_HIP(hipMemcpyAsync(ctx->h_small + 4, be.row_count, sizeof(int), hipMemcpyDeviceToHost, st));
            FH_HIP(hipStreamSynchronize(st));
            exact_rows += ctx->h_small[4];
        }
        query_ms += span(ctx, 5, 6);
        query_main_ms += span(ctx, 8, 9);
        const int emitted = ctx->h_small[0];
        const int C2 = ctx->h_small[1];
        if (emitted != C - C2 || C2 >= C)
            return fail(ctx, FH_EDEVICE,
                        "Boruvka round made no consistent progress (components " + std::to_string(C) + " -> " +
                            std::to_string(C2) + ", edges " + std::to_string(emitted) + ")");
        base += emitted;
        C = C2;
        std::swap(ctx->roots, ctx->roots2);
        ++rounds;
    }
    if (brute) {
        FH_HIP(hipMemcpyAsync(ctx->h_small + 6, ctx->bx_scored.p, sizeof(unsigned long long), hipMemcpyDeviceToHost,
                              st));
        FH_HIP(hipStreamSynchronize(st));
        unsigned long long sc = 0;
        std::memcpy(&sc, ctx->h_small + 6, sizeof(sc));
        S.scored_tile_pairs = (int64_t)sc;
    }
    FH_HIP(mark(ctx, 4));

    // -- outputs
    if (tree_rows) {
        FH_HIP(fh::launch_outputs_by_point(n, kq, ctx->posof.p, ctx->perm.p, ctx->core_p.p, ctx->nb_pos.p, core_out,
                                           nbr_out, st));
    } else {
        if (core_out)
            FH_HIP(hipMemcpyAsync(core_out, ctx->core_o.p, sizeof(float) * n, hipMemcpyDeviceToDevice, st));
        if (nbr_out) FH_HIP(fh::launch_neighbors(n, kq, ctx->knn_i.p, nbr_out, st));
    }
    FH_HIP(mark(ctx, 7));
    FH_HIP(ctx->timing ? hipEventSynchronize(ctx->ev[7]) : hipStreamSynchronize(st));

    S.rounds = rounds;
    S.tie_cycle_nodes += tie_nodes;
    S.exact_rows = exact_rows;
    S.scan_rounds = scan_rounds;
    S.n_edges = base;
    S.t_tree_ms = span(ctx, 0, 1);
    S.t_knn_ms = span(ctx, 1, 2);
    S.t_init_ms = span(ctx, 2, 3);
    S.t_rounds_ms = span(ctx, 3, 4);
    S.t_query_ms = query_ms;
    S.t_query_main_ms = query_main_ms;
    S.t_total_ms = total.ms();
    if (base != n - 1)
        return fail(ctx, FH_EDEVICE, "spanning tree has " + std::to_string(base) + " edges for " +
                                         std::to_string(n) + " points");
    if (stats) *stats = S;
    return FH_OK;
}

// ---------------------------------------------- persistent batch ----
//
// Datasets small enough for one work-group (fh_batch.hip) run in one
// persistent launch per dimension: their rows are packed into a pinned
// buffer, uploaded with one copy, and the launch's work-groups take them
// from a device-side queue, costliest first; the rows of every MST, the
// core distances and a status per dataset come back with one copy each.

// Whether the persistent kernel can take a dataset: within its size limit,
// min_samples + 1 within its heap size, and a tree whose leaves all sit on
// the last level.  Which of these it is given is the router's choice.
bool batch_eligible(int64_t n, int d, int ms) {
    if (d < 1 || n < 2 || n > fh::batch_max_points(d)) return false;
    const int kq = ms > 1 ? ms + 1 : 2;
    if (kq > fh::batch_max_kq()) return false;
    const int levels = fh::tree_levels(n, kLeafSize);
    return !(levels >= 2 && (n >> (levels - 2)) < 2);
}

// One persistent launch: the datasets idx (all of dimension d) and where
// their results land in the context's pinned buffer.
struct BatchGroup {
    fh_ctx *ctx = nullptr;
    int d = 0;
    std::vector<int> idx;             // dataset indices in queue order
    std::vector<fh::BatchJob> jobs;
    int64_t total_n = 0, total_rows = 0;
    int grid = 0, max_n = 0, max_nodes = 0;
    int64_t slice = 0;
    bool phases = false;
    const int *src = nullptr, *dst = nullptr, *status = nullptr;
    const float *w = nullptr, *core = nullptr;
};

// Lays the group out and sizes its buffers (pinned and device): done for
// every group before any launch, since freeing a buffer to grow it waits
// for the whole device.
int batch_group_prepare(BatchGroup &g, const int64_t *n, int ms) {
    fh_ctx *ctx = g.ctx;
    int cus = 0;
    FH_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
    const int d = g.d;
    const int kq = ms > 1 ? ms + 1 : 2;
    std::stable_sort(g.idx.begin(), g.idx.end(), [&](int a, int b) { return n[a] > n[b]; });
    const int m = (int)g.idx.size();
    g.jobs.resize(m);
    int64_t xo = 0, ro = 0, po = 0;
    int max_n = 0, max_nodes = 1;
    for (int j = 0; j < m; ++j) {
        const int64_t ni = n[g.idx[j]];
        fh::BatchJob &b = g.jobs[j];
        b.x_off = xo;
        b.row_off = ro;
        b.pt_off = po;
        b.n = (int32_t)ni;
        b.d = d;
        b.levels = fh::tree_levels(ni, kLeafSize);
        b.pad = 0;
        xo += ni * d;
        ro += ni - 1;
        po += ni;
        max_n = std::max(max_n, (int)ni);
        max_nodes = std::max(max_nodes, (1 << b.levels) - 1);
    }
    g.total_n = po;
    g.total_rows = ro;
    // pinned layout: rows up; src, dst, w, core, status down
    const size_t up = sizeof(float) * xo;
    const size_t down = sizeof(int) * 2 * ro + sizeof(float) * ro + sizeof(float) * po + sizeof(int) * m;
    const size_t need = up + down + 64;
    if (ctx->h_bt_cap < need) {
        if (ctx->h_bt) FH_HIP(hipHostFree(ctx->h_bt));
        ctx->h_bt = nullptr;
        ctx->h_bt_cap = 0;
        const size_t want = need + need / 4;
        FH_HIP(hipHostMalloc(reinterpret_cast<void **>(&ctx->h_bt), want));
        ctx->h_bt_cap = want;
    }
    g.grid = fh::batch_grid(m, cus);
    g.slice = fh::batch_slice_bytes(max_n, max_nodes, d, kq);
    g.max_n = max_n;
    g.max_nodes = max_nodes;
    FH_TRY(ensure(ctx, ctx->bt_x, (size_t)xo));
    FH_TRY(ensure(ctx, ctx->bt_jobs, (size_t)m));
    // (one row past the last dataset's: a round that emits more rows than
    // its components allow writes at most there before it is caught)
    FH_TRY(ensure(ctx, ctx->bt_src, (size_t)ro + 1));
    FH_TRY(ensure(ctx, ctx->bt_dst, (size_t)ro + 1));
    FH_TRY(ensure(ctx, ctx->bt_w, (size_t)ro + 1));
    FH_TRY(ensure(ctx, ctx->bt_core, (size_t)po));
    FH_TRY(ensure(ctx, ctx->bt_status, (size_t)m));
    FH_TRY(ensure(ctx, ctx->bt_next, 1));
    FH_TRY(ensure(ctx, ctx->bt_scratch, (size_t)(g.slice * g.grid)));
    return FH_OK;
}

// Packs, uploads and launches; the outputs' copies to the pinned buffer are
// queued behind the kernel on the context's stream (batch_group_wait).
int batch_group_launch(BatchGroup &g, const float *const *X, int ms) {
    fh_ctx *ctx = g.ctx;
    hipStream_t st = ctx->stream;
    const int d = g.d;
    const int kq = ms > 1 ? ms + 1 : 2;
    const int m = (int)g.idx.size();
    const int64_t xo = g.total_n * d, ro = g.total_rows, po = g.total_n;
    const size_t up = sizeof(float) * xo;
    float *hx = reinterpret_cast<float *>(ctx->h_bt);
    for (int j = 0; j < m; ++j)
        std::memcpy(hx + g.jobs[j].x_off, X[g.idx[j]], sizeof(float) * (size_t)g.jobs[j].n * d);
    unsigned char *hd = ctx->h_bt + ((up + 15) / 16) * 16;
    int *hsrc = reinterpret_cast<int *>(hd);
    int *hdst = hsrc + ro;
    float *hw = reinterpret_cast<float *>(hdst + ro);
    float *hcore = hw + ro;
    int *hstat = reinterpret_cast<int *>(hcore + po);
    g.src = hsrc;
    g.dst = hdst;
    g.w = hw;
    g.core = hcore;
    g.status = hstat;

    FH_HIP(hipMemcpyAsync(ctx->bt_x.p, hx, up, hipMemcpyHostToDevice, st));
    FH_HIP(hipMemcpyAsync(ctx->bt_jobs.p, g.jobs.data(), sizeof(fh::BatchJob) * m, hipMemcpyHostToDevice, st));
    FH_HIP(hipMemsetAsync(ctx->bt_next.p, 0, sizeof(int), st));
    fh::BatchArgs a{};
    a.X = ctx->bt_x.p;
    a.jobs = ctx->bt_jobs.p;
    a.count = m;
    a.ms = ms;
    a.kq = kq;
    a.next = ctx->bt_next.p;
    a.scratch = ctx->bt_scratch.p;
    a.slice_bytes = g.slice;
    a.max_n = g.max_n;
    a.max_nodes = g.max_nodes;
    a.max_d = d;
    a.src = ctx->bt_src.p;
    a.dst = ctx->bt_dst.p;
    a.w = ctx->bt_w.p;
    a.core = ctx->bt_core.p;
    a.status = ctx->bt_status.p;
    // FH_BATCH_PHASES=1: per-phase device time, printed once the launch ends
    g.phases = std::getenv("FH_BATCH_PHASES") && std::atoi(std::getenv("FH_BATCH_PHASES")) != 0;
    if (g.phases) {
        FH_TRY(ensure(ctx, ctx->bt_ticks, 5));
        FH_HIP(hipMemsetAsync(ctx->bt_ticks.p, 0, sizeof(uint64_t) * 5, st));
        a.phase_ticks = ctx->bt_ticks.p;
    }
    FH_HIP(fh::launch_batch(a, g.grid, st));
    FH_HIP(hipMemcpyAsync(hsrc, ctx->bt_src.p, sizeof(int) * ro, hipMemcpyDeviceToHost, st));
    FH_HIP(hipMemcpyAsync(hdst, ctx->bt_dst.p, sizeof(int) * ro, hipMemcpyDeviceToHost, st));
    FH_HIP(hipMemcpyAsync(hw, ctx->bt_w.p, sizeof(float) * ro, hipMemcpyDeviceToHost, st));
    FH_HIP(hipMemcpyAsync(hcore, ctx->bt_core.p, sizeof(float) * po, hipMemcpyDeviceToHost, st));
    FH_HIP(hipMemcpyAsync(hstat, ctx->bt_status.p, sizeof(int) * m, hipMemcpyDeviceToHost, st));
    return FH_OK;
}

// Waits for the launch and its copies; every dataset must report n - 1 rows.
int batch_group_wait(BatchGroup &g) {
    fh_ctx *ctx = g.ctx;
    FH_HIP(hipStreamSynchronize(ctx->stream));
    if (g.phases) {
        uint64_t t[5];
        FH_HIP(hipMemcpy(t, ctx->bt_ticks.p, sizeof(t), hipMemcpyDeviceToHost));
        int rate = 0;  // kHz of the wall clock
        FH_HIP(hipDeviceGetAttribute(&rate, hipDeviceAttributeWallClockRate, ctx->device));
        const double ms_per_tick = rate > 0 ? 1.0 / rate : 0.0;
        std::fprintf(stderr,
                     "batch phases d=%d datasets=%zu grid=%d (work-group ms summed): tree %.1f knn %.1f forest %.1f "
                     "searches %.1f rest %.1f\n",
                     g.d, g.jobs.size(), g.grid, t[0] * ms_per_tick, t[1] * ms_per_tick, t[2] * ms_per_tick,
                     t[3] * ms_per_tick, t[4] * ms_per_tick);
    }
    for (size_t j = 0; j < g.jobs.size(); ++j)
        if (g.status[j] != g.jobs[j].n - 1)
            return fail(ctx, FH_EDEVICE, "persistent batch: dataset " + std::to_string(g.idx[j]) + " gave " +
                                             std::to_string(g.status[j]) + " rows for " +
                                             std::to_string(g.jobs[j].n) + " points");
    return FH_OK;
}

int check_args(fh_ctx *ctx, const void *X, int64_t n, int32_t d, int32_t ms) {
    if (!ctx) return FH_EINVAL;
    if (!X) return fail(ctx, FH_EINVAL, "X is null");
    if (n < 1) return fail(ctx, FH_EINVAL, "n must be >= 1");
    if (n > (int64_t(1) << 30)) return fail(ctx, FH_EUNSUPPORTED, "n above 2^30");
    if (d < 1 || d > FH_MAX_DIM) return fail(ctx, FH_EUNSUPPORTED, "d outside [1, FH_MAX_DIM]");
    if (ms < 1) return fail(ctx, FH_EINVAL, "min_samples must be >= 1");
    return FH_OK;
}

}  // namespace

// ------------------------------------------- allocation-free entry ----
//
// fh_mst_euclidean_device_noalloc keeps every device buffer of its calls in
// this file-scope array: FH_NOALLOC_SCRATCH_MIB MiB that each GPU the
// library runs on holds until the process ends.  Host code reaches it only
// through the address hipGetSymbolAddress returns on that device (the
// array's name is a host-side stand-in there).
static __device__ __align__(256) unsigned char scratch[size_t(FH_NOALLOC_SCRATCH_MIB) << 20];

namespace {

constexpr int kNoAllocDevices = 64;

struct NoAllocSlot {
    std::mutex mu;                   // one call at a time per device
    bool ready = false;
    unsigned char *scratch_ptr = nullptr;  // device address of `scratch`
    int small[64] = {};              // readback slots (pageable host memory)
    fh_ctx ctx;                      // host-side state of the calls
};

NoAllocSlot g_noalloc[kNoAllocDevices];

int noalloc_run(NoAllocSlot &sl, int device, void *stream, const float *X, int64_t n, int32_t d, int32_t ms,
                float *core_sq, int32_t *nbr, int32_t *src, int32_t *dst, float *w_sq, fh_stats *stats) {
    fh_ctx *ctx = &sl.ctx;
    FH_TRY(check_args(ctx, X, n, d, ms));
    if (n > 1 && (!src || !dst || !w_sq)) return fail(ctx, FH_EINVAL, "edge outputs are null");
    FH_HIP(hipSetDevice(device));
    if (!sl.ready) {
        // once per device, with that device current
        void *p = nullptr;
        FH_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        sl.scratch_ptr = static_cast<unsigned char *>(p);
        ctx->device = device;
        ctx->timing = false;
        ctx->h_small = sl.small;
        ctx->arena = sl.scratch_ptr;
        ctx->arena_cap = size_t(FH_NOALLOC_SCRATCH_MIB) << 20;
        sl.ready = true;
    }
    ctx->stream = static_cast<hipStream_t>(stream);
    // every call carves its buffers from offset 0: nothing survives a call
    for_each_buf(ctx, [](auto &b) {
        b.p = nullptr;
        b.cap = 0;
    });
    ctx->arena_used = 0;
    ctx->geom_n = -1;
    ctx->tree_n = -1;
    // a lower bound of what the buffers take (points in tree order, kNN
    // rows, per-point round state): larger inputs are refused before any
    // launch; the carving itself refuses any input that does not fit
    const int64_t kq = ms > 1 ? ms + 1 : 2;
    const double need = double(n) * double(4 * d + 8 * kq + 140);
    if (need > double(ctx->arena_cap))
        return fail(ctx, FH_EINVAL,
                    "input needs more than the " + std::to_string(FH_NOALLOC_SCRATCH_MIB) +
                        " MiB scratch of the allocation-free entry point (at least " +
                        std::to_string(int64_t(need) >> 20) + " MiB)");
    return run_mst(ctx, X, n, d, ms, nullptr, 0, core_sq, nbr, src, dst, w_sq, stats);
}

}  // namespace

extern "C" {

const char *fh_version(void) { return "fh 0.1.0 gfx950"; }

fh_ctx *fh_create(int device) {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) return nullptr;
    fh_ctx *c = new fh_ctx();
    c->device = device;
    if (const char *m = std::getenv("FH_SEARCH_MODE")) c->search_mode = std::atoi(m);
    if (const char *m = std::getenv("FH_HIGH_DIM_CAP")) c->brute_cap = std::max(16, std::atoi(m));
    if (const char *m = std::getenv("FH_LEARNED_RADIUS")) c->learned_mode = std::atoi(m);
    if (hipSetDevice(device) != hipSuccess) {
        delete c;
        return nullptr;
    }
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return nullptr;
    }
    for (auto &e : c->ev) {
        if (hipEventCreate(&e) != hipSuccess) {
            fh_destroy(c);
            return nullptr;
        }
    }
    if (hipHostMalloc(reinterpret_cast<void **>(&c->h_small), 64 * sizeof(int)) != hipSuccess) {
        c->h_small = nullptr;
        fh_destroy(c);
        return nullptr;
    }
    return c;
}

void fh_destroy(fh_ctx *c) {
    if (!c) return;
    hipError_t e = hipSetDevice(c->device);
    (void)e;
    if (c->stream) e = hipStreamSynchronize(c->stream);
    if (!c->arena) for_each_buf(c, [](auto &b) { release(b); });
    if (c->h_small) e = hipHostFree(c->h_small);
    if (c->h_bt) e = hipHostFree(c->h_bt);
    for (auto &ev : c->ev)
        if (ev) e = hipEventDestroy(ev);
    if (c->stream) e = hipStreamDestroy(c->stream);
    delete c;
}

const char *fh_last_error(const fh_ctx *c) { return c ? c->err.c_str() : "null context"; }

int64_t fh_tree_nnodes(int64_t n) {
    if (n < 1) return 0;
    return (int64_t(1) << fh::tree_levels(n, kLeafSize)) - 1;
}

int64_t fh_neighbor_columns(int64_t n, int32_t min_samples, const float *sample_weights) {
    if (n < 1 || min_samples < 1) return -1;
    if (!sample_weights) return min_samples > 1 ? min_samples : 1;
    return weighted_k(sample_weights, n, min_samples) - 1;
}

int fh_mst_euclidean_device(fh_ctx *ctx, const float *X, int64_t n, int32_t d, int32_t min_samples,
                            const float *sample_weights, float *core_sq, int32_t *nbr, int32_t *src, int32_t *dst,
                            float *w_sq, fh_stats *stats) {
    FH_TRY(check_args(ctx, X, n, d, min_samples));
    if (n > 1 && (!src || !dst || !w_sq)) return fail(ctx, FH_EINVAL, "edge outputs are null");
    FH_HIP(hipSetDevice(ctx->device));
    int kq_w = 0;
    if (sample_weights) {
        std::vector<float> hw(n);
        FH_HIP(hipMemcpyAsync(hw.data(), sample_weights, sizeof(float) * n, hipMemcpyDeviceToHost, ctx->stream));
        FH_HIP(hipStreamSynchronize(ctx->stream));
        const int64_t k = weighted_k(hw.data(), n, min_samples);
        if (k < 1 || k > 4096) return fail(ctx, FH_EUNSUPPORTED, "sample weights give a kNN width outside [1, 4096]");
        kq_w = static_cast<int>(k);
    }
    return run_mst(ctx, X, n, d, min_samples, sample_weights, kq_w, core_sq, nbr, src, dst, w_sq, stats);
}

int fh_mst_sweep_device(fh_ctx *ctx, const float *X, int64_t n, int32_t d, int32_t count,
                        const int32_t *min_samples, float *core_sq, int32_t *src, int32_t *dst, float *w_sq,
                        fh_stats *stats) {
    FH_TRY(check_args(ctx, X, n, d, 2));
    if (count < 1 || !min_samples) return fail(ctx, FH_EINVAL, "empty min_samples list");
    if (n > 1 && (!src || !dst || !w_sq)) return fail(ctx, FH_EINVAL, "edge outputs are null");
    int mmax = 0;
    for (int32_t i = 0; i < count; ++i) {
        if (min_samples[i] < 2 || min_samples[i] > 126)
            return fail(ctx, FH_EUNSUPPORTED, "sweep min_samples outside [2, 126]");
        mmax = std::max(mmax, (int)min_samples[i]);
    }
    if (use_brute(ctx, n, d, mmax + 2, false)) return fail(ctx, FH_EUNSUPPORTED, "sweep of the high-dimensional path");
    FH_HIP(hipSetDevice(ctx->device));
    Timer total;
    int levels = 0;
    FH_TRY(build_tree_device(ctx, X, n, d, &levels, nullptr));
    // columns 1..K of the (K+1)-nearest query, K = largest min_samples + 1
    SweepRows sw{nullptr, nullptr, mmax + 1};
    FH_TRY(ensure(ctx, ctx->sw_pos, (size_t)n * sw.K));
    FH_TRY(ensure(ctx, ctx->sw_d, (size_t)n * sw.K));
    FH_TRY(ensure(ctx, ctx->core_p, n));
    FH_TRY(run_knn(ctx, d, levels, ctx->xp.p, ctx->perm.p, n, sw.K + 1, nullptr, nullptr, nullptr, ctx->core_p.p, 1,
                   ctx->sw_pos.p, ctx->sw_d.p, 0));
    sw.pos = ctx->sw_pos.p;
    sw.dist = ctx->sw_d.p;
    fh_stats S{};
    const int64_t m = n > 1 ? n - 1 : 0;
    for (int32_t i = 0; i < count; ++i)
        FH_TRY(run_mst(ctx, X, n, d, min_samples[i], nullptr, 0, core_sq ? core_sq + (size_t)i * n : nullptr, nullptr,
                       src + (size_t)i * m, dst + (size_t)i * m, w_sq + (size_t)i * m, &S, &sw));
    S.t_total_ms = total.ms();
    if (stats) *stats = S;
    return FH_OK;
}

int fh_mst_euclidean_device_noalloc(int32_t device, void *stream, const float *X, int64_t n, int32_t d,
                                    int32_t min_samples, float *core_sq, int32_t *nbr, int32_t *src, int32_t *dst,
                                    float *w_sq, fh_stats *stats, char *err, int32_t err_cap) {
    int count = 0;
    if (device < 0 || device >= kNoAllocDevices || hipGetDeviceCount(&count) != hipSuccess || device >= count) {
        if (err && err_cap > 0) std::snprintf(err, (size_t)err_cap, "no HIP device %d", (int)device);
        return FH_EINVAL;
    }
    NoAllocSlot &sl = g_noalloc[device];
    std::lock_guard<std::mutex> lock(sl.mu);
    const int code = noalloc_run(sl, device, stream, X, n, d, min_samples, core_sq, nbr, src, dst, w_sq, stats);
    if (err && err_cap > 0) std::snprintf(err, (size_t)err_cap, "%s", code == FH_OK ? "" : sl.ctx.err.c_str());
    return code;
}

int fh_mst_euclidean(fh_ctx *ctx, const float *X, int64_t n, int32_t d, int32_t min_samples,
                     const float *sample_weights, float *core_sq, int32_t *nbr, int32_t *src, int32_t *dst,
                     float *w_sq, fh_stats *stats) {
    FH_TRY(check_args(ctx, X, n, d, min_samples));
    if (n > 1 && (!src || !dst || !w_sq)) return fail(ctx, FH_EINVAL, "edge outputs are null");
    FH_HIP(hipSetDevice(ctx->device));
    int kq_w = 0;
    if (sample_weights) {
        const int64_t k = weighted_k(sample_weights, n, min_samples);
        if (k < 1 || k > 4096) return fail(ctx, FH_EUNSUPPORTED, "sample weights give a kNN width outside [1, 4096]");
        kq_w = static_cast<int>(k);
    }
    const int kn = sample_weights ? kq_w - 1 : (min_samples > 1 ? min_samples : 1);
    const int64_t m = n - 1;
    FH_TRY(ensure(ctx, ctx->out_core, n));
    FH_TRY(ensure(ctx, ctx->out_nbr, (size_t)n * (kn > 0 ? kn : 1)));
    FH_TRY(ensure(ctx, ctx->out_src, m));
    FH_TRY(ensure(ctx, ctx->out_dst, m));
    FH_TRY(ensure(ctx, ctx->out_w, m));
    FH_TRY(ensure(ctx, ctx->xdev, (size_t)n * d));
    hipStream_t st = ctx->stream;
    FH_HIP(hipMemcpyAsync(ctx->xdev.p, X, sizeof(float) * n * d, hipMemcpyHostToDevice, st));
    if (sample_weights) {
        FH_TRY(ensure(ctx, ctx->wdev, n));
        FH_HIP(hipMemcpyAsync(ctx->wdev.p, sample_weights, sizeof(float) * n, hipMemcpyHostToDevice, st));
    }
    FH_TRY(run_mst(ctx, ctx->xdev.p, n, d, min_samples, sample_weights ? ctx->wdev.p : nullptr, kq_w,
                   ctx->out_core.p, ctx->out_nbr.p, ctx->out_src.p, ctx->out_dst.p, ctx->out_w.p, stats));
    if (core_sq) FH_HIP(hipMemcpyAsync(core_sq, ctx->out_core.p, sizeof(float) * n, hipMemcpyDeviceToHost, st));
    if (nbr && kn > 0) FH_HIP(hipMemcpyAsync(nbr, ctx->out_nbr.p, sizeof(int) * n * kn, hipMemcpyDeviceToHost, st));
    if (m > 0) {
        FH_HIP(hipMemcpyAsync(src, ctx->out_src.p, sizeof(int) * m, hipMemcpyDeviceToHost, st));
        FH_HIP(hipMemcpyAsync(dst, ctx->out_dst.p, sizeof(int) * m, hipMemcpyDeviceToHost, st));
        FH_HIP(hipMemcpyAsync(w_sq, ctx->out_w.p, sizeof(float) * m, hipMemcpyDeviceToHost, st));
    }
    FH_HIP(hipStreamSynchronize(st));
    return FH_OK;
}

int fh_knn(fh_ctx *ctx, const float *X, int64_t n, int32_t d, const float *Q, int64_t nq, int32_t k,
           int32_t output_rdist, float *dist, int32_t *idx) {
    FH_TRY(check_args(ctx, X, n, d, 1));
    if (!Q || !dist || !idx || nq < 0 || k < 1) return fail(ctx, FH_EINVAL, "bad query arguments");
    if (nq == 0) return FH_OK;
    FH_HIP(hipSetDevice(ctx->device));
    FH_TRY(ensure(ctx, ctx->xdev, (size_t)n * d));
    FH_HIP(hipMemcpyAsync(ctx->xdev.p, X, sizeof(float) * n * d, hipMemcpyHostToDevice, ctx->stream));
    int levels = 0;
    FH_TRY(build_tree_device(ctx, ctx->xdev.p, n, d, &levels, nullptr));
    FH_TRY(ensure(ctx, ctx->queries, (size_t)nq * d));
    FH_TRY(ensure(ctx, ctx->knn_d, (size_t)nq * k));
    FH_TRY(ensure(ctx, ctx->knn_i, (size_t)nq * k));
    hipStream_t st = ctx->stream;
    FH_HIP(hipMemcpyAsync(ctx->queries.p, Q, sizeof(float) * nq * d, hipMemcpyHostToDevice, st));
    // numba_kdtree.py:789-792: square roots (when asked) before deheap_sort
    FH_TRY(run_knn(ctx, d, levels, ctx->queries.p, nullptr, nq, k, ctx->knn_d.p, ctx->knn_i.p, nullptr, nullptr, 0,
                   nullptr, nullptr, output_rdist ? 0 : 1));
    FH_HIP(hipMemcpyAsync(dist, ctx->knn_d.p, sizeof(float) * nq * k, hipMemcpyDeviceToHost, st));
    FH_HIP(hipMemcpyAsync(idx, ctx->knn_i.p, sizeof(int) * nq * k, hipMemcpyDeviceToHost, st));
    FH_HIP(hipStreamSynchronize(st));
    return FH_OK;
}

int fh_build_kdtree(fh_ctx *ctx, const float *X, int64_t n, int32_t d, int64_t *idx_array, int64_t *idx_start,
                    int64_t *idx_end, float *radius, uint8_t *is_leaf, float *node_bounds) {
    if (!X || !idx_array || !idx_start || !idx_end || !radius || !is_leaf || !node_bounds || n < 1 || d < 1 ||
        n > (int64_t(1) << 30))
        return ctx ? fail(ctx, FH_EINVAL, "bad fh_build_kdtree arguments") : FH_EINVAL;
    if (!ctx) {
        // No context: the host construction (same tree, no device needed).
        HostTree t;
        if (fh::build_host_tree(X, n, d, kLeafSize, 0, t) != 0) return FH_EINVAL;
        const int64_t N = t.n_nodes;
        for (int64_t i = 0; i < n; ++i) idx_array[i] = t.perm[i];
        for (int64_t i = 0; i < N; ++i) {
            idx_start[i] = t.start[i];
            idx_end[i] = t.end[i];
            radius[i] = t.radius[i];
            is_leaf[i] = t.leaf[i];
            for (int j = 0; j < d; ++j) {
                node_bounds[i * d + j] = t.box[(size_t)i * 2 * d + j];
                node_bounds[(N + i) * d + j] = t.box[(size_t)i * 2 * d + d + j];
            }
        }
        return FH_OK;
    }
    if (d > FH_MAX_DIM) return fail(ctx, FH_EUNSUPPORTED, "d above FH_MAX_DIM");
    FH_HIP(hipSetDevice(ctx->device));
    FH_TRY(ensure(ctx, ctx->xdev, (size_t)n * d));
    hipStream_t st = ctx->stream;
    FH_HIP(hipMemcpyAsync(ctx->xdev.p, X, sizeof(float) * n * d, hipMemcpyHostToDevice, st));
    int levels = 0;
    int64_t N = 0;
    FH_TRY(build_tree_device(ctx, ctx->xdev.p, n, d, &levels, &N));
    std::vector<int> perm(n);
    std::vector<float> box((size_t)N * 2 * d);
    std::vector<int2> range(N);
    FH_HIP(hipMemcpyAsync(perm.data(), ctx->perm.p, sizeof(int) * n, hipMemcpyDeviceToHost, st));
    FH_HIP(hipMemcpyAsync(box.data(), ctx->box.p, sizeof(float) * N * 2 * d, hipMemcpyDeviceToHost, st));
    FH_HIP(hipMemcpyAsync(range.data(), ctx->range.p, sizeof(int2) * N, hipMemcpyDeviceToHost, st));
    FH_HIP(hipStreamSynchronize(st));
    for (int64_t i = 0; i < n; ++i) idx_array[i] = perm[i];
    const int64_t first_leaf = (N - 1) / 2;
    for (int64_t i = 0; i < N; ++i) {
        idx_start[i] = range[i].x;
        idx_end[i] = range[i].y;
        is_leaf[i] = i >= first_leaf ? 1 : 0;
        float r = 0.0f;
        for (int j = 0; j < d; ++j) {
            const float lo = box[(size_t)i * 2 * d + j], hi = box[(size_t)i * 2 * d + d + j];
            node_bounds[i * d + j] = lo;
            node_bounds[(N + i) * d + j] = hi;
            const float half = std::fabs(hi - lo) * 0.5f;
            r += half * half;
        }
        radius[i] = std::sqrt(r);
    }
    return FH_OK;
}

int fh_edges_to_tree_order(fh_ctx *ctx, int64_t n, int64_t m, const int32_t *src, const int32_t *dst,
                           int32_t *pos_src, int32_t *pos_dst, int32_t *order) {
    if (!ctx) return FH_EINVAL;
    if (n < 1 || n != ctx->tree_n) return fail(ctx, FH_EINVAL, "n is not the point count of the context's last tree");
    if (m < 0 || m > n) return fail(ctx, FH_EINVAL, "bad row count");
    if (m > 0 && (!src || !dst || !pos_src || !pos_dst)) return fail(ctx, FH_EINVAL, "null pointer");
    FH_HIP(hipSetDevice(ctx->device));
    FH_HIP(fh::tree_positions(static_cast<int>(m), src, dst, ctx->posof.p, pos_src, pos_dst, ctx->stream));
    if (order)
        FH_HIP(hipMemcpyAsync(order, ctx->perm.p, sizeof(int32_t) * (size_t)n, hipMemcpyDeviceToDevice, ctx->stream));
    FH_HIP(hipStreamSynchronize(ctx->stream));
    return FH_OK;
}

int fh_linkage_plan(fh_ctx *ctx, int64_t n, int64_t k, const int32_t *src, const int32_t *dst, int32_t *row_order,
                    int32_t *group_sizes, int64_t *n_groups) {
    if (!ctx) return FH_EINVAL;
    if (n < 2 || k < 0 || k > n - 1 || n > (int64_t(1) << 31) - 1) return fail(ctx, FH_EINVAL, "bad sizes");
    if (!n_groups || (k > 0 && (!src || !dst || !row_order || !group_sizes)))
        return fail(ctx, FH_EINVAL, "null pointer");
    *n_groups = 0;
    if (k == 0) return FH_OK;
    FH_HIP(hipSetDevice(ctx->device));
    const int ki = static_cast<int>(k);
    FH_TRY(ensure(ctx, ctx->pc_parent, n));
    FH_TRY(ensure(ctx, ctx->pc_keys, k));
    FH_TRY(ensure(ctx, ctx->pc_keys_s, k));
    FH_TRY(ensure(ctx, ctx->pc_vals, k));
    FH_TRY(ensure(ctx, ctx->pc_vals_s, k));
    FH_TRY(ensure(ctx, ctx->pc_counts, k));
    FH_TRY(ensure(ctx, ctx->pc_runs, 1));
    const size_t tb = fh::prefix_components_temp_bytes(ki);
    if (tb == 0) return fail(ctx, FH_EDEVICE, "hipCUB temp size query failed");
    FH_TRY(ensure(ctx, ctx->pc_temp, tb));
    fh::PrefixComponentsArgs a{};
    a.n = static_cast<int>(n);
    a.k = ki;
    a.src = src;
    a.dst = dst;
    a.parent = ctx->pc_parent.p;
    a.keys = ctx->pc_keys.p;
    a.keys_s = ctx->pc_keys_s.p;
    a.vals = ctx->pc_vals.p;
    a.vals_s = ctx->pc_vals_s.p;
    a.counts = ctx->pc_counts.p;
    a.n_runs = ctx->pc_runs.p;
    a.temp = ctx->pc_temp.p;
    a.temp_bytes = ctx->pc_temp.cap;
    FH_HIP(fh::prefix_components(a, ctx->stream));
    int runs = 0;
    FH_HIP(hipMemcpyAsync(&runs, ctx->pc_runs.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    FH_HIP(hipMemcpyAsync(row_order, ctx->pc_vals_s.p, sizeof(int32_t) * (size_t)k, hipMemcpyDeviceToHost,
                          ctx->stream));
    FH_HIP(hipStreamSynchronize(ctx->stream));
    if (runs < 1 || runs > ki) return fail(ctx, FH_EDEVICE, "bad group count");
    FH_HIP(hipMemcpyAsync(group_sizes, ctx->pc_counts.p, sizeof(int32_t) * (size_t)runs, hipMemcpyDeviceToHost,
                          ctx->stream));
    FH_HIP(hipStreamSynchronize(ctx->stream));
    *n_groups = runs;
    return FH_OK;
}

int fh_sort_mst_device(fh_ctx *ctx, int64_t m, const int32_t *src, const int32_t *dst, const float *w_sq,
                       int32_t *src_out, int32_t *dst_out, float *w_out) {
    if (!ctx) return FH_EINVAL;
    if (m < 0 || m > (int64_t(1) << 30)) return fail(ctx, FH_EINVAL, "bad row count");
    if (m == 0) return FH_OK;
    if (!src || !dst || !w_sq || !src_out || !dst_out || !w_out) return fail(ctx, FH_EINVAL, "null pointer");
    FH_HIP(hipSetDevice(ctx->device));
    const int mi = static_cast<int>(m);
    FH_TRY(ensure(ctx, ctx->sk1, m));
    FH_TRY(ensure(ctx, ctx->sk1s, m));
    FH_TRY(ensure(ctx, ctx->sk2, m));
    FH_TRY(ensure(ctx, ctx->sk2s, m));
    FH_TRY(ensure(ctx, ctx->sv1, m));
    FH_TRY(ensure(ctx, ctx->sv1s, m));
    const size_t tb = fh::sort_mst_temp_bytes(mi);
    FH_TRY(ensure(ctx, ctx->stemp, tb));
    fh::SortMstArgs a{};
    a.m = mi;
    a.src = src;
    a.dst = dst;
    a.w = w_sq;
    a.src_out = src_out;
    a.dst_out = dst_out;
    a.w_out = w_out;
    a.k1 = ctx->sk1.p;
    a.k1s = ctx->sk1s.p;
    a.k2 = ctx->sk2.p;
    a.k2s = ctx->sk2s.p;
    a.v1 = ctx->sv1.p;
    a.v1s = ctx->sv1s.p;
    a.temp = ctx->stemp.p;
    a.temp_bytes = ctx->stemp.cap;
    FH_HIP(fh::sort_mst(a, ctx->stream));
    FH_HIP(hipStreamSynchronize(ctx->stream));
    return FH_OK;
}

// Many independent requests: a pool of worker threads, each with its own
// context (own HIP stream and buffers), takes datasets from a shared
// counter and runs the MST on the device, sorts the rows on the host and
// runs the host label stage, so different datasets' kernels overlap on the
// GPU and their label stages on the host cores.
// Per device: the batch workers' contexts and the persistent launches'
// contexts (one per dimension in flight), kept for the process's lifetime
// and reused by later calls, so no stream or buffer is created or destroyed
// while other kernels run and a call starts with buffers already sized.
// One batch call per device at a time.
namespace {
std::mutex batch_mu[kNoAllocDevices];
std::vector<fh_ctx *> batch_ctx[kNoAllocDevices];
std::vector<fh_ctx *> batch_group_ctx[kNoAllocDevices];

int grow_pool(std::vector<fh_ctx *> &pool, int device, int size) {
    while ((int)pool.size() < size) {
        fh_ctx *c = fh_create(device);
        if (!c) return FH_EDEVICE;
        pool.push_back(c);
    }
    return FH_OK;
}

// ---------------------------------------------------- the router ----
//
// Which datasets of a batch go to the persistent launches and which to the
// per-call path on the worker streams, from a cost model whose constants
// are measured on the device (once per device, dimension and min_samples):
//   wg[d]: seconds per point for one work-group of the persistent kernel,
//          from launches of one synthetic dataset per resident work-group
//          at 2000 and at 16000 points (a power law in between and beyond);
//   st[d]: seconds per point of one per-call MST alone (host arrays in and
//          out), from a 32000-point synthetic dataset.
// The persistent launches finish after max(sum of their work-group times /
// resident work-groups, their longest work-group time); the streams after
// the sum of their datasets' times over the streams that overlap (at most
// the 4 hardware queues).  Starting with every dataset persistent, the
// longest work-group jobs move to the streams while that shortens the
// larger of the two.  FH_BATCH_MAX_N (points) replaces the model with a
// fixed cap; FH_BATCH_PERSISTENT=0 sends everything to the streams.
struct BatchCost {
    int ms = -1;  // min_samples of the measurement (-1: none yet)
    double wg1 = 0.0, n1 = 1.0, p = 0.0;  // work-group seconds per point at n1 points, growing as (n / n1)^p
    double st = 0.0;
    int slots = 0;  // resident work-groups of the launch
    // one work-group's time for a dataset of n points
    double wg(int64_t n) const { return n * wg1 * std::pow(std::max(1.0, double(n) / n1), p); }
};
BatchCost batch_cost[kNoAllocDevices][FH_MAX_DIM + 1];

// count Gaussian-mixture datasets of n points in d dimensions (5 centres in
// [-10, 10]^d, unit spread), deterministic
std::vector<std::vector<float>> calibration_data(int count, int n, int d, uint64_t seed) {
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<float> cen(-10.0f, 10.0f);
    std::normal_distribution<float> nrm(0.0f, 1.0f);
    std::vector<std::vector<float>> out(count, std::vector<float>((size_t)n * d));
    for (auto &X : out) {
        float c[5][FH_MAX_DIM];
        for (auto &cc : c)
            for (int j = 0; j < d; ++j) cc[j] = cen(rng);
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < d; ++j) X[(size_t)i * d + j] = c[i % 5][j] + nrm(rng);
    }
    return out;
}

// Seconds per point of one work-group on datasets of n points: one launch
// of one dataset per resident work-group (the second of two runs).
int batch_time_wg(int d, int ms, int n, int count, fh_ctx *gctx, double &out) {
    fh_ctx *ctx = gctx;
    auto data = calibration_data(count, n, d, 0x5eed0000ull + 131 * d + n);
    std::vector<const float *> xp(count);
    std::vector<int64_t> ns(count, n);
    for (int i = 0; i < count; ++i) xp[i] = data[i].data();
    double t = 0.0;
    for (int rep = 0; rep < 2; ++rep) {
        BatchGroup g;
        g.ctx = gctx;
        g.d = d;
        for (int i = 0; i < count; ++i) g.idx.push_back(i);
        FH_TRY(batch_group_prepare(g, ns.data(), ms));
        const Timer tm;
        FH_TRY(batch_group_launch(g, xp.data(), ms));
        FH_TRY(batch_group_wait(g));
        t = tm.ms() * 1e-3;
    }
    out = t / n;
    (void)ctx;
    return FH_OK;
}

int batch_calibrate(int device, int d, int ms, fh_ctx *gctx, fh_ctx *wctx) {
    BatchCost &c = batch_cost[device][d];
    if (c.ms == ms) return FH_OK;
    fh_ctx *ctx = gctx;
    int cus = 0;
    FH_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    const int res = std::max(1, fh::batch_resident_per_cu(d));
    const int slots = std::max(1, std::min(res * cus, fh::batch_grid(1 << 30, cus)));
    const int n1 = 2000, n2 = 16000;
    double c1 = 0.0, c2 = 0.0;
    FH_TRY(batch_time_wg(d, ms, n1, slots, gctx, c1));
    FH_TRY(batch_time_wg(d, ms, n2, slots, gctx, c2));
    c.wg1 = c1;
    c.n1 = n1;
    c.p = std::min(1.0, std::max(0.0, std::log(c2 / c1) / std::log(double(n2) / n1)));
    c.slots = slots;
    const int n_st = 32000;
    auto one = calibration_data(1, n_st, d, 0x5eed1000ull + d);
    std::vector<int32_t> src(n_st), dst(n_st);
    std::vector<float> w(n_st);
    double t = 0.0;
    for (int rep = 0; rep < 2; ++rep) {
        const Timer tm;
        FH_TRY(fh_mst_euclidean(wctx, one[0].data(), n_st, d, ms, nullptr, nullptr, nullptr, src.data(), dst.data(),
                                w.data(), nullptr));
        t = tm.ms() * 1e-3;
    }
    c.st = t / n_st;
    c.ms = ms;
    if (std::getenv("FH_BATCH_PHASES") && std::atoi(std::getenv("FH_BATCH_PHASES")) != 0)
        std::fprintf(stderr, "batch cost d=%d: work-group %.3g s/pt at %d points (exponent %.2f), alone %.3g s/pt, %d slots\n",
                     d, c.wg1, n1, c.p, c.st, c.slots);
    return FH_OK;
}

// Splits the batch: by_d[dd] the persistent launches' datasets, rest the
// streams'.  streams: how many worker streams overlap.
int batch_route(int device, int32_t count, const float *const *X, const int64_t *n, const int32_t *d, int32_t ms,
                int streams, std::vector<int> *by_d, std::vector<int32_t> &rest) {
    std::vector<int32_t> cand;
    for (int32_t i = 0; i < count; ++i) {
        if (n[i] >= 2 && d[i] >= 1 && d[i] <= FH_MAX_DIM && X[i] && batch_eligible(n[i], d[i], ms))
            cand.push_back(i);
        else
            rest.push_back(i);
    }
    if (const char *m = std::getenv("FH_BATCH_PERSISTENT"))
        if (std::atoi(m) == 0) {
            rest.insert(rest.end(), cand.begin(), cand.end());
            return FH_OK;
        }
    if (const char *m = std::getenv("FH_BATCH_MAX_N")) {
        const int64_t cap = std::atoll(m);
        for (int32_t i : cand) (n[i] <= cap ? by_d[d[i]].push_back(i) : rest.push_back(i));
        return FH_OK;
    }
    // the model's constants for every dimension present
    bool present[FH_MAX_DIM + 1] = {};
    for (int32_t i : cand) present[d[i]] = true;
    FH_TRY(grow_pool(batch_group_ctx[device], device, 1));
    FH_TRY(grow_pool(batch_ctx[device], device, 1));
    for (int dd = 1; dd <= FH_MAX_DIM; ++dd)
        if (present[dd]) FH_TRY(batch_calibrate(device, dd, ms, batch_group_ctx[device][0], batch_ctx[device][0]));
    const BatchCost *cost = batch_cost[device];
    std::stable_sort(cand.begin(), cand.end(),
                     [&](int32_t a, int32_t b) { return cost[d[a]].wg(n[a]) > cost[d[b]].wg(n[b]); });
    const double q = std::max(1, std::min(streams, 4));
    double tp_sum = 0.0, ts = 0.0;
    for (int32_t i : cand) tp_sum += cost[d[i]].wg(n[i]) / cost[d[i]].slots;
    size_t moved = 0;
    for (; moved < cand.size(); ++moved) {
        const int32_t i = cand[moved];
        const double job = cost[d[i]].wg(n[i]);
        const double now = std::max({tp_sum, job, ts});
        const double next_job = moved + 1 < cand.size() ? cost[d[cand[moved + 1]]].wg(n[cand[moved + 1]]) : 0.0;
        const double tp2 = tp_sum - job / cost[d[i]].slots;
        const double ts2 = ts + (n[i] * cost[d[i]].st + 3e-4) / q;
        if (std::max({tp2, next_job, ts2}) >= now) break;
        tp_sum = tp2;
        ts = ts2;
    }
    for (size_t k = 0; k < cand.size(); ++k) (k < moved ? rest.push_back(cand[k]) : by_d[d[cand[k]]].push_back(cand[k]));
    return FH_OK;
}

// The datasets the persistent kernel takes, one group per dimension, each
// launched on its own context's stream; `rest` gets the others.
int batch_start(int device, int32_t count, const float *const *X, const int64_t *n, const int32_t *d,
                int32_t ms, int streams, std::vector<BatchGroup> &groups, std::vector<int32_t> &rest) {
    std::vector<int> by_d[FH_MAX_DIM + 1];
    FH_TRY(batch_route(device, count, X, n, d, ms, streams, by_d, rest));
    int ng = 0;
    for (int dd = 1; dd <= FH_MAX_DIM; ++dd) ng += by_d[dd].empty() ? 0 : 1;
    FH_TRY(grow_pool(batch_group_ctx[device], device, ng));
    int gi = 0;
    for (int dd = 1; dd <= FH_MAX_DIM; ++dd) {
        if (by_d[dd].empty()) continue;
        BatchGroup g;
        g.ctx = batch_group_ctx[device][gi++];
        g.d = dd;
        g.idx = by_d[dd];
        groups.push_back(std::move(g));
    }
    for (auto &g : groups) FH_TRY(batch_group_prepare(g, n, ms));
    for (size_t k = 0; k < groups.size(); ++k) {
        const int rc = batch_group_launch(groups[k], X, ms);
        if (rc != FH_OK) {
            // nothing of this call may still run when it returns
            for (size_t j = 0; j < k; ++j) (void)hipStreamSynchronize(groups[j].ctx->stream);
            return rc;
        }
    }
    return FH_OK;
}
}  // namespace

int fh_mst_batch(int32_t device, int32_t count, const float *const *X, const int64_t *n, const int32_t *d,
                 int32_t min_samples, int32_t *const *src, int32_t *const *dst, float *const *w_sq,
                 float *const *core_sq, int32_t *persistent) {
    if (count < 0 || (count > 0 && (!X || !n || !d || !src || !dst || !w_sq)) || min_samples < 1) return FH_EINVAL;
    if (device < 0 || device >= kNoAllocDevices) return FH_EINVAL;
    std::lock_guard<std::mutex> batch_lock(batch_mu[device]);
    if (hipSetDevice(device) != hipSuccess) return FH_EDEVICE;
    std::vector<BatchGroup> groups;
    std::vector<int32_t> rest;
    FH_TRY(batch_start(device, count, X, n, d, min_samples, 1, groups, rest));
    if (persistent)
        for (int32_t i = 0; i < count; ++i) persistent[i] = 0;
    // the others one at a time on a worker context, beside the launches
    int rc = grow_pool(batch_ctx[device], device, 1);
    for (size_t r = 0; rc == FH_OK && r < rest.size(); ++r) {
        const int32_t i = rest[r];
        rc = (!X[i] || n[i] < 1) ? FH_EINVAL
                                 : fh_mst_euclidean(batch_ctx[device][0], X[i], n[i], d[i], min_samples, nullptr,
                                                    core_sq ? core_sq[i] : nullptr, nullptr, src[i], dst[i], w_sq[i],
                                                    nullptr);
    }
    for (auto &g : groups) {
        const int wrc = batch_group_wait(g);
        if (rc == FH_OK) rc = wrc;
    }
    if (rc != FH_OK) return rc;
    for (auto &g : groups) {
        for (size_t j = 0; j < g.idx.size(); ++j) {
            const int i = g.idx[j];
            const fh::BatchJob &b = g.jobs[j];
            std::memcpy(src[i], g.src + b.row_off, sizeof(int32_t) * (b.n - 1));
            std::memcpy(dst[i], g.dst + b.row_off, sizeof(int32_t) * (b.n - 1));
            std::memcpy(w_sq[i], g.w + b.row_off, sizeof(float) * (b.n - 1));
            if (core_sq && core_sq[i]) std::memcpy(core_sq[i], g.core + b.pt_off, sizeof(float) * b.n);
            if (persistent) persistent[i] = 1;
        }
    }
    return FH_OK;
}

int fh_hdbscan_batch(int32_t device, int32_t count, const float *const *X, const int64_t *n, const int32_t *d,
                     int32_t min_samples, int64_t min_cluster_size, int32_t workers, int64_t *const *labels,
                     float *const *probabilities, double *ms_out) {
    if (count < 0 || !X || !n || !d || !labels || !probabilities || min_samples < 1 || min_cluster_size < 1)
        return FH_EINVAL;
    const int W = std::max(1, std::min<int>(workers, 16));
    if (device < 0 || device >= kNoAllocDevices) return FH_EINVAL;
    std::lock_guard<std::mutex> batch_lock(batch_mu[device]);
    if (hipSetDevice(device) != hipSuccess) return FH_EDEVICE;
    // small datasets: persistent launches, started before the workers
    std::vector<BatchGroup> groups;
    std::vector<int32_t> queue;
    FH_TRY(batch_start(device, count, X, n, d, min_samples, W, groups, queue));
    // The rest one at a time on the workers' streams, largest first: a
    // worker's buffers reach their final size on its first dataset instead
    // of growing (hipFree waits for the whole device) as larger ones arrive,
    // and the long jobs do not trail at the end.
    std::stable_sort(queue.begin(), queue.end(), [&](int32_t a, int32_t b) {
        return n[a] * std::max(d[a], 1) > n[b] * std::max(d[b], 1);
    });
    const int nq = (int)queue.size();
    // the persistent launches' datasets, for the label stage once they are back
    // (group, job), published group by group as the launches complete
    std::vector<std::pair<int, int>> small;
    size_t n_small = 0;
    for (const auto &g : groups) n_small += g.idx.size();
    small.reserve(n_small);
    std::vector<fh_ctx *> &pool = batch_ctx[device];
    FH_TRY(grow_pool(pool, device, W));
    std::atomic<int> next{0}, next_small{0};
    std::atomic<int> err{0};
    std::mutex ready_mu;
    std::condition_variable ready_cv;
    int published = 0;
    bool all_done = false;
    auto work = [&](int wi) {
        fh_ctx *ctx = pool[wi];
        if (hipSetDevice(device) != hipSuccess) {
            err = FH_EDEVICE;
            return;
        }
        std::vector<int32_t> src, dst, s2, d2;
        std::vector<float> w, w2;
        std::vector<int32_t> order;
        fh_label_params prm{min_cluster_size, std::numeric_limits<double>::infinity(), 0.0, 0, 0};
        // hdbscan.py:511-519: rows by (weight, source, target), then the labels
        auto label = [&](int i, int64_t ni, const int32_t *rs, const int32_t *rd, const float *rw) -> int {
            const int64_t m = ni - 1;
            order.resize(m);
            for (int64_t r = 0; r < m; ++r) order[r] = (int32_t)r;
            std::sort(order.begin(), order.end(), [&](int32_t a, int32_t b) {
                if (rw[a] != rw[b]) return rw[a] < rw[b];
                if (rs[a] != rs[b]) return rs[a] < rs[b];
                return rd[a] < rd[b];
            });
            s2.resize(m);
            d2.resize(m);
            w2.resize(m);
            for (int64_t r = 0; r < m; ++r) {
                s2[r] = rs[order[r]];
                d2[r] = rd[order[r]];
                w2[r] = rw[order[r]];
            }
            return fh_labels_from_sorted_mst(ni, s2.data(), d2.data(), w2.data(), nullptr, nullptr, &prm, labels[i],
                                             probabilities[i], nullptr, nullptr, nullptr, nullptr);
        };
        for (;;) {
            const int qi = next.fetch_add(1);
            if (qi >= nq || err.load() != 0) break;
            const int i = queue[qi];
            const auto t0 = std::chrono::steady_clock::now();
            const int64_t ni = n[i];
            if (ni < 1 || !X[i] || !labels[i] || !probabilities[i]) {
                err = FH_EINVAL;
                break;
            }
            if (ni < 2) {
                labels[i][0] = -1;
                probabilities[i][0] = 0.0f;
            } else {
                const int64_t m = ni - 1;
                src.resize(m);
                dst.resize(m);
                w.resize(m);
                int rc = fh_mst_euclidean(ctx, X[i], ni, d[i], min_samples, nullptr, nullptr, nullptr, src.data(),
                                          dst.data(), w.data(), nullptr);
                if (rc == FH_OK) rc = label(i, ni, src.data(), dst.data(), w.data());
                if (rc != FH_OK) {
                    err = rc;
                    break;
                }
            }
            if (ms_out)
                ms_out[i] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        }
        for (;;) {
            const int si = next_small.fetch_add(1);
            {
                std::unique_lock<std::mutex> lk(ready_mu);
                ready_cv.wait(lk, [&] { return si < published || all_done; });
                if (si >= published) break;  // every launch is back and its jobs taken
            }
            if (err.load() != 0) break;
            const BatchGroup &g = groups[small[si].first];
            const int j = small[si].second;
            const int i = g.idx[j];
            const fh::BatchJob &b = g.jobs[j];
            if (!labels[i] || !probabilities[i]) {
                err = FH_EINVAL;
                break;
            }
            const auto t0 = std::chrono::steady_clock::now();
            const int rc = label(i, b.n, g.src + b.row_off, g.dst + b.row_off, g.w + b.row_off);
            if (rc != FH_OK) {
                err = rc;
                break;
            }
            if (ms_out)
                ms_out[i] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        }
    };
    std::vector<std::thread> th;
    for (int t = 0; t < W; ++t) th.emplace_back(work, t);
    // each launch's label stages start as soon as it is back
    std::vector<int> pending;
    for (int gi = 0; gi < (int)groups.size(); ++gi) pending.push_back(gi);
    while (!pending.empty()) {
        bool any = false;
        for (size_t k = 0; k < pending.size();) {
            BatchGroup &g = groups[pending[k]];
            if (hipStreamQuery(g.ctx->stream) == hipErrorNotReady) {
                ++k;
                continue;
            }
            const int rc = batch_group_wait(g);
            if (rc != FH_OK) err = rc;
            {
                std::lock_guard<std::mutex> lk(ready_mu);
                for (int j = 0; j < (int)g.idx.size(); ++j) small.emplace_back(pending[k], j);
                published = (int)small.size();
            }
            ready_cv.notify_all();
            pending.erase(pending.begin() + k);
            any = true;
        }
        if (!any) std::this_thread::sleep_for(std::chrono::microseconds(200));
    }
    {
        std::lock_guard<std::mutex> lk(ready_mu);
        all_done = true;
    }
    ready_cv.notify_all();
    for (auto &x : th) x.join();
    return err.load();
}

int fh_csr_boruvka(fh_ctx *ctx, int64_t n, const int32_t *indptr, const int32_t *indices, const float *weights,
                   double *edges, int64_t *n_edges, int64_t *point_components, int64_t *n_components) {
    if (!ctx) return FH_EINVAL;
    FH_HIP(hipSetDevice(ctx->device));
    FH_HIP(hipStreamSynchronize(ctx->stream));
    const int code = fh::csr_boruvka(ctx->device, n, indptr, indices, weights, edges, n_edges, point_components,
                                     n_components);
    if (code != FH_OK) return fail(ctx, code, "CSR Boruvka failed");
    return FH_OK;
}

int fh_set_learned_radius(fh_ctx *ctx, int32_t mode) {
    if (!ctx) return FH_EINVAL;
    if (mode < 0 || mode > 2) return fail(ctx, FH_EINVAL, "learned-radius mode must be 0, 1 or 2");
    ctx->learned_mode = mode;
    if (mode != 2) ctx->ml_F = 0;
    return FH_OK;
}

int fh_mlp_step(fh_ctx *ctx, int32_t F, int32_t B, const float *x, const float *t, float *params, float *grad,
                float *pred, float lr) {
    if (!ctx) return FH_EINVAL;
    if (F < 1 || F > fh::kMlpMaxF || B < 1 || !x || !t || !params) return fail(ctx, FH_EINVAL, "bad MLP step arguments");
    FH_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const int P = fh::mlp_param_count(F);
    FH_TRY(ensure(ctx, ctx->ml_feat, (size_t)B * F));
    FH_TRY(ensure(ctx, ctx->ml_tgt, B));
    FH_TRY(ensure(ctx, ctx->ml_w, P));
    FH_TRY(ensure(ctx, ctx->ml_g, P));
    FH_TRY(ensure(ctx, ctx->ml_m, P));
    FH_TRY(ensure(ctx, ctx->ml_v, P));
    FH_TRY(ensure(ctx, ctx->ml_mu, F));
    FH_TRY(ensure(ctx, ctx->ml_sd, F));
    FH_TRY(ensure(ctx, ctx->ml_scratch, (size_t)B * (F + 4 * fh::kMlpHidden + 2)));
    std::vector<float> zero(F, 0.0f), one(F, 1.0f), zp(P, 0.0f);
    FH_HIP(hipMemcpyAsync(ctx->ml_feat.p, x, sizeof(float) * B * F, hipMemcpyHostToDevice, st));
    FH_HIP(hipMemcpyAsync(ctx->ml_tgt.p, t, sizeof(float) * B, hipMemcpyHostToDevice, st));
    FH_HIP(hipMemcpyAsync(ctx->ml_w.p, params, sizeof(float) * P, hipMemcpyHostToDevice, st));
    FH_HIP(hipMemcpyAsync(ctx->ml_m.p, zp.data(), sizeof(float) * P, hipMemcpyHostToDevice, st));
    FH_HIP(hipMemcpyAsync(ctx->ml_v.p, zp.data(), sizeof(float) * P, hipMemcpyHostToDevice, st));
    FH_HIP(hipMemcpyAsync(ctx->ml_mu.p, zero.data(), sizeof(float) * F, hipMemcpyHostToDevice, st));
    FH_HIP(hipMemcpyAsync(ctx->ml_sd.p, one.data(), sizeof(float) * F, hipMemcpyHostToDevice, st));
    fh::MlpParams mp{ctx->ml_w.p, ctx->ml_g.p, ctx->ml_m.p, ctx->ml_v.p, ctx->ml_mu.p, ctx->ml_sd.p, F};
    FH_HIP(fh::launch_mlp_train_step(mp, B, B, 0, ctx->ml_feat.p, nullptr, ctx->ml_tgt.p, ctx->ml_scratch.p, lr, 1, st));
    if (grad) FH_HIP(hipMemcpyAsync(grad, ctx->ml_g.p, sizeof(float) * P, hipMemcpyDeviceToHost, st));
    FH_HIP(hipMemcpyAsync(params, ctx->ml_w.p, sizeof(float) * P, hipMemcpyDeviceToHost, st));
    if (pred) {
        // Y follows X, H1, H2, D1, D2, D3 in the scratch block
        const size_t yoff = (size_t)B * (F + 4 * fh::kMlpHidden + 1);
        FH_HIP(hipMemcpyAsync(pred, ctx->ml_scratch.p + yoff, sizeof(float) * B, hipMemcpyDeviceToHost, st));
    }
    FH_HIP(hipStreamSynchronize(st));
    ctx->ml_F = 0;  // the learned-radius model, if any, was overwritten
    return FH_OK;
}

// kruskal.py:950-1034 (_kruskal_mst_knn) without cannot-link constraints.
// Device: tree, kNN (k_query columns), weighted core, mutual-reachability
// edges in float64 and their stable global sort.  Host: Kruskal's
// union-find (_kruskal_core, kruskal.py:31-108) and the +inf bridges of a
// disconnected forest (_finish_mst, precomputed.bridge_forest_with_inf).
int fh_mst_kruskal_knn(fh_ctx *ctx, const float *X, int64_t n, int32_t d, int32_t min_samples, int32_t knn_k,
                       const float *sample_weights, int32_t k_for_core, double *mst, int32_t *nbr, int32_t nbr_cols,
                       double *core, fh_stats *stats) {
    FH_TRY(check_args(ctx, X, n, d, min_samples));
    if (knn_k < 1) return fail(ctx, FH_EINVAL, "knn_k must be >= 1");
    if (n > 1 && !mst) return fail(ctx, FH_EINVAL, "mst output is null");
    const bool weighted = sample_weights != nullptr;
    // kruskal.py:987-1011
    int64_t kq64;
    int want_cols;
    if (weighted) {
        kq64 = (int64_t)std::max(knn_k, k_for_core) + 1;
        want_cols = (int)kq64 - 1;
    } else if (min_samples > 1) {
        kq64 = (int64_t)std::max(knn_k, min_samples) + 1;
        want_cols = min_samples;
    } else {
        kq64 = (int64_t)knn_k + 1;
        want_cols = 1;
    }
    if (kq64 > 4096) return fail(ctx, FH_EUNSUPPORTED, "k_query above 4096");
    if (nbr && nbr_cols != want_cols) return fail(ctx, FH_EINVAL, "nbr_cols must be " + std::to_string(want_cols));
    const int kq = (int)kq64;
    const int64_t m = n * (kq - 1);
    if (m >= (int64_t(1) << 31)) return fail(ctx, FH_EUNSUPPORTED, "n * (k_query - 1) must stay below 2^31");
    FH_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    Timer total;
    fh_stats S{};
    FH_TRY(ensure(ctx, ctx->xdev, (size_t)n * d));
    FH_HIP(hipMemcpyAsync(ctx->xdev.p, X, sizeof(float) * n * d, hipMemcpyHostToDevice, st));
    if (weighted) {
        FH_TRY(ensure(ctx, ctx->wdev, n));
        FH_HIP(hipMemcpyAsync(ctx->wdev.p, sample_weights, sizeof(float) * n, hipMemcpyHostToDevice, st));
    }
    int levels = 0;
    int64_t n_nodes = 0;
    FH_TRY(build_tree_device(ctx, ctx->xdev.p, n, d, &levels, &n_nodes));
    S.n_levels = levels;
    S.n_nodes = n_nodes;
    FH_TRY(ensure(ctx, ctx->knn_d, (size_t)n * kq));
    FH_TRY(ensure(ctx, ctx->knn_i, (size_t)n * kq));
    FH_TRY(ensure(ctx, ctx->core_o, n));
    FH_TRY(ensure(ctx, ctx->core_p, n));
    FH_TRY(ensure(ctx, ctx->comp_p, n));
    const bool brute = use_brute(ctx, n, d, kq, weighted);
    if (brute) {
        fh::BruteScanArgs bsa{};
        FH_TRY(brute_knn(ctx, n, d, levels, kq, 0, bsa, S));
    } else {
        FH_TRY(run_knn(ctx, d, levels, ctx->xp.p, ctx->perm.p, n, kq, ctx->knn_d.p, ctx->knn_i.p, ctx->core_o.p,
                       ctx->core_p.p, 0, nullptr, nullptr, weighted ? 1 : 0));
    }
    if (weighted)
        FH_HIP(fh::launch_weighted_core((int)n, kq, ctx->knn_i.p, ctx->knn_d.p, ctx->wdev.p, min_samples, ctx->posof.p,
                                        ctx->core_o.p, ctx->core_p.p, st));
    // edges
    FH_TRY(ensure(ctx, ctx->kr_keys, m));
    FH_TRY(ensure(ctx, ctx->kr_keys_s, m));
    FH_TRY(ensure(ctx, ctx->kr_vals, m));
    FH_TRY(ensure(ctx, ctx->kr_vals_s, m));
    FH_TRY(ensure(ctx, ctx->kr_src, m));
    FH_TRY(ensure(ctx, ctx->kr_dst, m));
    FH_TRY(ensure(ctx, ctx->kr_w, m));
    FH_TRY(ensure(ctx, ctx->kr_core, n));
    FH_TRY(ensure(ctx, ctx->small, 8));
    const size_t tb = fh::knn_edges_temp_bytes(m);
    FH_TRY(ensure(ctx, ctx->stemp, tb));
    FH_HIP(hipMemsetAsync(ctx->small.p, 0, sizeof(int), st));
    fh::KnnEdgesArgs ka{};
    ka.n = n;
    ka.kq = kq;
    ka.mode = weighted ? 2 : (min_samples > 1 ? 0 : 1);
    ka.ms = min_samples;
    ka.knn_d = ctx->knn_d.p;
    ka.knn_i = ctx->knn_i.p;
    ka.core32 = ctx->core_o.p;
    ka.core64 = ctx->kr_core.p;
    ka.keys = ctx->kr_keys.p;
    ka.keys_s = ctx->kr_keys_s.p;
    ka.vals = ctx->kr_vals.p;
    ka.vals_s = ctx->kr_vals_s.p;
    ka.valid_count = ctx->small.p;
    ka.src_out = ctx->kr_src.p;
    ka.dst_out = ctx->kr_dst.p;
    ka.w_out = ctx->kr_w.p;
    ka.temp = ctx->stemp.p;
    ka.temp_bytes = ctx->stemp.cap;
    FH_HIP(fh::knn_edges_sorted(ka, st));
    int nvalid = 0;
    FH_HIP(hipMemcpyAsync(ctx->h_small, ctx->small.p, sizeof(int), hipMemcpyDeviceToHost, st));
    FH_HIP(hipStreamSynchronize(st));
    nvalid = ctx->h_small[0];
    std::vector<int> hs(nvalid), hd(nvalid);
    std::vector<double> hw(nvalid);
    if (nvalid > 0) {
        FH_HIP(hipMemcpyAsync(hs.data(), ctx->kr_src.p, sizeof(int) * nvalid, hipMemcpyDeviceToHost, st));
        FH_HIP(hipMemcpyAsync(hd.data(), ctx->kr_dst.p, sizeof(int) * nvalid, hipMemcpyDeviceToHost, st));
        FH_HIP(hipMemcpyAsync(hw.data(), ctx->kr_w.p, sizeof(double) * nvalid, hipMemcpyDeviceToHost, st));
    }
    if (core) FH_HIP(hipMemcpyAsync(core, ctx->kr_core.p, sizeof(double) * n, hipMemcpyDeviceToHost, st));
    std::vector<int> hk;
    if (nbr) {
        hk.resize((size_t)n * kq);
        FH_HIP(hipMemcpyAsync(hk.data(), ctx->knn_i.p, sizeof(int) * n * kq, hipMemcpyDeviceToHost, st));
    }
    FH_HIP(hipStreamSynchronize(st));
    if (nbr)
        for (int64_t i = 0; i < n; ++i)
            for (int c = 0; c < want_cols; ++c) nbr[i * want_cols + c] = hk[(size_t)i * kq + 1 + c];
    // _kruskal_core: union by rank, full path compression, stop at n-1
    std::vector<int> pred(n), rank(n, 0);
    for (int64_t i = 0; i < n; ++i) pred[i] = (int)i;
    auto root_of = [&](int x) {
        int r = x;
        while (pred[r] != r) r = pred[r];
        while (x != r) {
            const int nx = pred[x];
            pred[x] = r;
            x = nx;
        }
        return r;
    };
    int64_t n_mst = 0;
    for (int64_t r = 0; r < nvalid && n_mst < n - 1; ++r) {
        const int u = hs[r], v = hd[r];
        if (u == v) continue;
        const int ru = root_of(u), rv = root_of(v);
        if (ru == rv) continue;
        mst[3 * n_mst + 0] = (double)u;
        mst[3 * n_mst + 1] = (double)v;
        mst[3 * n_mst + 2] = hw[r];
        ++n_mst;
        if (rank[ru] > rank[rv]) {
            pred[rv] = ru;
        } else if (rank[ru] < rank[rv]) {
            pred[ru] = rv;
        } else {
            pred[rv] = ru;
            rank[ru] += 1;
        }
    }
    // _finish_mst: one representative (smallest id) per component, bridged
    // in ascending order with +inf edges
    if (n_mst < n - 1) {
        std::vector<char> seen(n, 0);
        std::vector<int> reps;
        for (int64_t i = 0; i < n; ++i) {
            int r = (int)i;
            while (pred[r] != r) r = pred[r];
            if (!seen[r]) {
                seen[r] = 1;
                reps.push_back((int)i);  // first (smallest) member met
            }
        }
        std::sort(reps.begin(), reps.end());
        for (size_t q = 0; q + 1 < reps.size(); ++q) {
            mst[3 * n_mst + 0] = (double)reps[q];
            mst[3 * n_mst + 1] = (double)reps[q + 1];
            mst[3 * n_mst + 2] = std::numeric_limits<double>::infinity();
            ++n_mst;
        }
    }
    if (n_mst != n - 1)
        return fail(ctx, FH_EDEVICE, "Kruskal produced " + std::to_string(n_mst) + " rows for " + std::to_string(n) +
                                         " points");
    S.n_edges = n_mst;
    S.t_total_ms = total.ms();
    if (stats) *stats = S;
    return FH_OK;
}

}  // extern "C"
