// Host label stage: sorted MST -> single-linkage tree -> condensed tree ->
// cluster selection -> labels and membership strengths.
//
// Counterpart of hdbscan.clusters_from_spanning_tree (hdbscan.py:505-574)
// for unit sample weights without semi-supervision, built on arrays instead
// of numba typed dicts.  Row orders, float32 rounding points and the
// union-by-rank tie rules follow cluster_trees.py so that labels and
// strengths come out identical:
//   mst_to_linkage_tree          cluster_trees.py:17-95
//   condense_tree                cluster_trees.py:115-352
//   extract_eom_clusters         cluster_trees.py:656-771
//   cluster_tree_leaves          cluster_trees.py:369-374
//   cluster_epsilon_search       cluster_trees.py:834-903
//   get_cluster_label_vector     cluster_trees.py:944-996
//   get_point_membership_strength_vector  cluster_trees.py:999-1031
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <new>
#include <thread>
#include <type_traits>
#include <utility>
#include <vector>

#include "../../include/fh.h"
#include "fh_numba_sort.h"
#include "fh_pool.h"

// phase timing hooks: empty unless a profiling build defines them
#ifndef FH_TMARK
#define FH_TMARK(x)
#define FH_TSTART()
#endif

namespace {

constexpr double kInfD = std::numeric_limits<double>::infinity();
constexpr float kInfF = std::numeric_limits<float>::infinity();

// One merge of the single-linkage tree: children and their sizes (16 bytes:
// the linkage writes and condensing reads them at random; the distance is
// the sorted row's own, Hierarchy::delta).
struct Merge {
    int32_t left, right;            // child node ids (points < n, merges >= n)
    int32_t left_size, right_size;  // points under each child
};

// delta(i): the merge distance of sorted row i -- the float64 square root of
// a squared float32 weight (boruvka.py:646), or a given float64 distance.
inline double row_delta(const float *w_sq, const double *dist, int64_t i) {
    return dist ? dist[i] : std::sqrt(static_cast<double>(w_sq[i]));
}

struct Hierarchy {
    std::vector<Merge> rows;  // row i is merge node n + i
    // the caller's sorted distances (valid for the call): row i's delta
    const float *w_sq = nullptr;
    const double *dist = nullptr;
    double delta(int64_t i) const { return row_delta(w_sq, dist, i); }
    // with sample weights: float32 weighted sizes of each row's children
    // (create_linkage_merge_data_w_sample_weights, cluster_trees.py:29-35)
    bool weighted = false;
    std::vector<float> lw, rw;
    bool literal = false;  // built by linkage_tree_literal (may hold -1 children)
};

inline void prefetch(const void *p) { __builtin_prefetch(p, 0, 1); }

// cluster_trees.py:66-95, reference form: the merge-node forest with
// linkage_merge_find (38-52).  Used when an endpoint is -1 (only when
// min_samples >= n), which the reference resolves through wrap-around
// indexing into the merge slots.
void linkage_tree_literal(int64_t n, const int32_t *src, const int32_t *dst, const float *w_sq, const double *dist,
                          Hierarchy &h) {
    const int64_t m = n - 1;
    const int64_t total = 2 * n - 1;
    std::vector<int64_t> parent(total, -1);
    std::vector<int64_t> size(total, 0);
    for (int64_t i = 0; i < n; ++i) size[i] = 1;
    h.rows.resize(m);
    h.literal = true;
    h.w_sq = w_sq;
    h.dist = dist;
    int64_t next = n;
    // node values stay as given (-1 stays -1); only array accesses wrap
    auto P = [&](int64_t x) -> int64_t & { return parent[x < 0 ? x + total : x]; };
    auto S = [&](int64_t x) -> int64_t & { return size[x < 0 ? x + total : x]; };
    auto find = [&](int64_t node) {
        int64_t relabel = node;
        while (P(node) != -1 && P(node) != node) node = P(node);
        P(node) = node;
        while (P(relabel) != node) {
            const int64_t nr = P(relabel);
            P(relabel) = node;
            relabel = nr;
        }
        return node;
    };
    for (int64_t i = 0; i < m; ++i) {
        const int64_t lc = find(src[i]);
        const int64_t rc = find(dst[i]);
        Merge &r = h.rows[i];
        const bool lbig = lc > rc;
        r.left = static_cast<int32_t>(lbig ? lc : rc);
        r.right = static_cast<int32_t>(lbig ? rc : lc);
        r.left_size = static_cast<int32_t>(S(lbig ? lc : rc));
        r.right_size = static_cast<int32_t>(S(lbig ? rc : lc));
        S(next) = S(lc) + S(rc);
        P(lc) = next;
        P(rc) = next;
        ++next;
    }
}

// Same tree from a union-find over the points (union by size, path
// halving) that records, per set, the merge node currently representing
// it: the node linkage_merge_find would return.  Upcoming edges' entries
// and their parents' entries are prefetched.
// With weights (cluster_trees.py:98-127) each set also carries its float32
// weighted size, summed in the reference's order (size[l] + size[r]).
// ord (or nullptr) changes only how the rows are walked:
//   point_order: the endpoints are positions p of some point order and
//     point_order[p] is the point's id.  The union-find then runs over
//     positions (a spatial order makes an edge's two endpoints, and
//     consecutive finds, share cache lines); which set becomes the root is
//     invisible in the rows.
//   prefix_rows k: rows [0, k) grouped by the connected components of
//     their forest.  Row i's entry depends only on the sets of its
//     endpoints at step i, i.e. on the earlier rows of its component, and
//     components touch disjoint union-find entries, so the groups run on
//     separate threads (each in row order) before rows [k, m) run in order.
// Returns FH_OK, FH_EINVAL for an endpoint outside [-1, n) or a point order
// value outside [0, n), or FH_EUNSUPPORTED for a -1 endpoint with weights.
int linkage_tree(int64_t n, const int32_t *src, const int32_t *dst, const float *w_sq, const double *dist,
                  Hierarchy &h, const float *weights = nullptr, const fh_label_order *ord = nullptr) {
    const int64_t m = n - 1;
    const int32_t *order = ord ? ord->point_order : nullptr;
    h.weighted = weights != nullptr;
    h.literal = false;
    h.w_sq = w_sq;
    h.dist = dist;
    const int nt = m > (1 << 16) ? fh::HostPool::width() : 1;
    // any -1 endpoint, any endpoint outside [-1, n)? (row ranges on threads)
    std::vector<uint8_t> neg(nt, 0), bad(nt, 0);
    fh::HostPool::run(nt, [&](int t) {
        uint8_t any = 0, out = 0;
        for (int64_t i = m * t / nt, e = m * (t + 1) / nt; i < e; ++i) {
            any |= (src[i] < 0) | (dst[i] < 0);
            out |= (static_cast<uint64_t>(int64_t{src[i]} + 1) > static_cast<uint64_t>(n)) |
                   (static_cast<uint64_t>(int64_t{dst[i]} + 1) > static_cast<uint64_t>(n));
        }
        neg[t] = any;
        bad[t] = out;
    });
    if (std::find(bad.begin(), bad.end(), 1) != bad.end()) return FH_EINVAL;
    if (order) {
        std::vector<uint8_t> obad(nt, 0);
        fh::HostPool::run(nt, [&](int t) {
            uint8_t out = 0;
            for (int64_t i = n * t / nt, e = n * (t + 1) / nt; i < e; ++i)
                out |= static_cast<uint32_t>(order[i]) >= static_cast<uint32_t>(n);
            obad[t] = out;
        });
        if (std::find(obad.begin(), obad.end(), 1) != obad.end()) return FH_EINVAL;
    }
    if (std::find(neg.begin(), neg.end(), 1) != neg.end()) {
        if (weights) return FH_EUNSUPPORTED;
        if (!order) {
            linkage_tree_literal(n, src, dst, w_sq, dist, h);
            return FH_OK;
        }
        std::vector<int32_t> s2(m), d2(m);
        for (int64_t j = 0; j < m; ++j) {
            s2[j] = src[j] < 0 ? src[j] : order[src[j]];
            d2[j] = dst[j] < 0 ? dst[j] : order[dst[j]];
        }
        linkage_tree_literal(n, s2.data(), d2.data(), w_sq, dist, h);
        return FH_OK;
    }
    static thread_local std::vector<float> wsz_tl;
    if (weights) {
        wsz_tl.resize(n);
        h.lw.resize(m);
        h.rw.resize(m);
    }
    struct Set {
        int32_t parent, size, node, pad;
    };
    // kept between calls: repeated calls of similar size touch no new pages
    static thread_local std::vector<Set> uf_tl;
    if (uf_tl.size() < static_cast<size_t>(n)) uf_tl.resize(n);
    // (worker threads below must use the caller's instances)
    Set *uf = uf_tl.data();
    float *wsz = wsz_tl.data();
    auto init_sets = [&] {
        fh::HostPool::run(nt, [&](int t) {
            for (int64_t i = n * t / nt, e = n * (t + 1) / nt; i < e; ++i) {
                uf[i] = {static_cast<int32_t>(i), 1, order ? order[i] : static_cast<int32_t>(i), 0};
                if (weights) wsz[i] = weights[order ? order[i] : i];
            }
        });
    };
    init_sets();
    if (h.rows.size() < static_cast<size_t>(m)) h.rows.resize(m);
    Merge *rows = h.rows.data();
    float *lw = weights ? h.lw.data() : nullptr;
    float *rw = weights ? h.rw.data() : nullptr;
    auto find = [uf](int32_t x) {
        while (uf[x].parent != x) {
            const int32_t g = uf[uf[x].parent].parent;
            uf[x].parent = g;
            x = g;
        }
        return x;
    };
    // row i merging sets (node lc, size ls, weight wa) and (rc, rs, wb):
    // the larger node id goes left
    auto emit = [&](int64_t i, int32_t lc, int32_t rc, int32_t ls, int32_t rs, float wa, float wb) {
        Merge &r = rows[i];
        const bool lbig = lc > rc;
        r.left = lbig ? lc : rc;
        r.right = lbig ? rc : lc;
        r.left_size = lbig ? ls : rs;
        r.right_size = lbig ? rs : ls;
        if (weights) {
            lw[i] = lbig ? wa : wb;
            rw[i] = lbig ? wb : wa;
        }
    };
    auto step = [&](int64_t i) {
        const int32_t a = find(src[i]);
        const int32_t b = find(dst[i]);
        const int32_t ls = uf[a].size, rs = uf[b].size;
        const float wa = weights ? wsz[a] : 0.0f, wb = weights ? wsz[b] : 0.0f;
        emit(i, uf[a].node, uf[b].node, ls, rs, wa, wb);
        int32_t root = a, child = b;
        if (ls < rs) {
            root = b;
            child = a;
        }
        uf[child].parent = root;
        uf[root].size = ls + rs;
        uf[root].node = static_cast<int32_t>(n + i);
        if (weights) wsz[root] = wa + wb;
    };
    // two prefetch stages: the endpoints' entries kAhead rows ahead, then
    // (those entries having arrived) their parents' entries kAhead2 ahead;
    // walks over a row list (not in row order) first fetch the row's
    // endpoints and output slot 2 kAhead ahead (prefetch_list)
    constexpr int64_t kAhead = 32, kAhead2 = 16;
    auto prefetch_row = [&](int64_t i, int stage) {
        if (stage == 0) {
            prefetch(&uf[src[i]]);
            prefetch(&uf[dst[i]]);
        } else {
            prefetch(&uf[uf[src[i]].parent]);
            prefetch(&uf[uf[dst[i]].parent]);
        }
    };
    auto prefetch_list = [&](int64_t i) {
        prefetch(&src[i]);
        prefetch(&dst[i]);
        __builtin_prefetch(&rows[i], 1, 1);
    };
    FH_TSTART();
    int64_t first_seq = 0;
    if (ord && ord->prefix_rows > 0 && ord->row_order && ord->group_sizes && ord->n_groups > 0 &&
        ord->prefix_rows <= m) {
        const int64_t k = ord->prefix_rows, G = ord->n_groups;
        const int32_t *ro = ord->row_order;
        std::vector<int64_t> start(G + 1, 0);
        int64_t big = 0;
        for (int64_t g = 0; g < G; ++g) {
            start[g + 1] = start[g] + ord->group_sizes[g];
            if (ord->group_sizes[g] > ord->group_sizes[big]) big = g;
        }
        if (start[G] != k) return FH_EINVAL;
        auto run_group = [&](int64_t g) {
            const int64_t e = start[g + 1];
            for (int64_t j = start[g]; j < e; ++j) {
                if (j + 2 * kAhead < e) prefetch_list(ro[j + 2 * kAhead]);
                if (j + kAhead < e) prefetch_row(ro[j + kAhead], 0);
                if (j + kAhead2 < e) prefetch_row(ro[j + kAhead2], 1);
                step(ro[j]);
            }
        };
        int nt = ord->threads > 0 ? ord->threads : static_cast<int>(std::thread::hardware_concurrency());
        nt = std::max(1, std::min(nt, 16));
        if (nt == 1 || k < (1 << 15)) {
            for (int64_t g = 0; g < G; ++g) run_group(g);
            first_seq = k;
        } else {
            // Task 0 runs the largest group.  The other tasks first mark
            // that group's points (uf[p].pad = 1, a field no step touches)
            // in pieces, then run the other groups in chunks of about
            // k / (16 nt) rows, all taken from a shared counter.  Task 1
            // then waits for those items and walks rows [k, m) in order
            // while task 0 may still run.  A row whose component at its step
            // holds no point of the largest group depends on earlier rows of
            // that component only, and those are all done, so it runs now.
            // A row that joins the largest group's component (the giant)
            // adds one other set to it: that set's node, size and weight
            // are final when the row is reached, so the row is kept with
            // them and the set's root is flagged (pad = 1) as part of the
            // giant.  Once task 0 has finished, the kept rows follow from
            // the giant's state after its own rows: row j merges the giant
            // (node n + previous kept row, size and weight the running sums)
            // with its set.  Rows run early touch no set of the giant, so
            // every row sees the sets it sees in the sequential order.  A
            // row with both ends in the giant closes a cycle (the rows are
            // no forest): then every row runs again in order.
            std::vector<int64_t> cut{0};
            const int64_t target = std::max<int64_t>(1, k / (16 * static_cast<int64_t>(nt)));
            int64_t acc = 0;
            for (int64_t g = 0; g < G; ++g) {
                if (g != big) acc += ord->group_sizes[g];
                if (acc >= target) {
                    cut.push_back(g + 1);
                    acc = 0;
                }
            }
            if (cut.back() != G) cut.push_back(G);
            const int64_t nchunks = static_cast<int64_t>(cut.size()) - 1;
            const int64_t b0 = start[big], b1 = start[big + 1];
            const int64_t npieces = std::min<int64_t>(4 * nt, (b1 - b0 + 32767) / 32768);
            const int64_t items = npieces + nchunks;
            std::atomic<int64_t> next{0}, done{0};
            struct Kept {
                int64_t row;
                int32_t node, size;
                float w;
            };
            static thread_local std::vector<Kept> kept_tl;
            std::vector<Kept> &kept = kept_tl;
            kept.clear();
            kept.reserve(static_cast<size_t>(m - k));
            bool cycle = false;
            auto tail_early = [&] {
                while (done.load(std::memory_order_acquire) < items) std::this_thread::yield();
                for (int64_t i = k; i < m; ++i) {
                    if (i + kAhead < m) {
                        prefetch(&uf[src[i + kAhead]]);
                        prefetch(&uf[dst[i + kAhead]]);
                    }
                    const int32_t a = src[i], b = dst[i];
                    // points of the largest group are never followed: task
                    // 0 may be changing their sets
                    const int32_t ra = uf[a].pad ? -1 : find(a);
                    const int32_t rb = uf[b].pad ? -1 : find(b);
                    const bool ga = ra < 0 || uf[ra].pad, gb = rb < 0 || uf[rb].pad;
                    if (ga && gb) {
                        cycle = true;
                        return;
                    }
                    if (ga || gb) {
                        const int32_t r = ga ? rb : ra;
                        kept.push_back({i, uf[r].node, uf[r].size, weights ? wsz[r] : 0.0f});
                        uf[r].pad = 1;
                    } else {
                        step(i);
                    }
                }
            };
            auto worker = [&](int t) {
                if (t == 0) run_group(big);
                for (;;) {
                    const int64_t c = next.fetch_add(1);
                    if (c >= items) break;
                    if (c < npieces) {
                        const int64_t e = b0 + (b1 - b0) * (c + 1) / npieces;
                        // a point ends several rows, possibly in two pieces
                        for (int64_t j = b0 + (b1 - b0) * c / npieces; j < e; ++j) {
                            __atomic_store_n(&uf[src[ro[j]]].pad, 1, __ATOMIC_RELAXED);
                            __atomic_store_n(&uf[dst[ro[j]]].pad, 1, __ATOMIC_RELAXED);
                        }
                    } else {
                        const int64_t q = c - npieces;
                        for (int64_t g = cut[q]; g < cut[q + 1]; ++g)
                            if (g != big) run_group(g);
                    }
                    done.fetch_add(1, std::memory_order_release);
                }
                if (t == 1) tail_early();
            };
            fh::HostPool::run(nt, worker);
            if (cycle) {
                init_sets();
                first_seq = 0;
            } else {
                const int32_t g = find(src[ro[b0]]);
                int32_t gnode = uf[g].node, gsize = uf[g].size;
                float gw = weights ? wsz[g] : 0.0f;
                for (const Kept &e : kept) {
                    emit(e.row, gnode, e.node, gsize, e.size, gw, e.w);
                    gnode = static_cast<int32_t>(n + e.row);
                    gsize += e.size;
                    gw += e.w;
                }
                first_seq = m;
            }
        }
    }
    FH_TMARK("link-groups");
    for (int64_t i = first_seq; i < m; ++i) {
        if (i + kAhead < m) prefetch_row(i + kAhead, 0);
        if (i + kAhead2 < m) prefetch_row(i + kAhead2, 1);
        step(i);
    }
    FH_TMARK("link-tail");
    return FH_OK;
}

// ids are below 2n; the label stage takes n < 2^30, so they fit int32
struct Condensed {
    std::vector<int32_t> parent, child;
    std::vector<float> lambda, size;
    int64_t rows = 0;
    int64_t next_label = 0;
    int64_t root = 0;                   // the root cluster's id (n)
    std::vector<int64_t> cluster_rows;  // rows whose child is a cluster id, ascending
    // score_condensed_tree_nodes (cluster_trees.py:656-669) per id - root,
    // filled by the chain form when asked for
    bool has_scores = false;
    std::vector<float> score;
};

// cluster_trees.py:205-352 (unit weights, max_cluster_size = inf: the
// reference does not pass it here, hdbscan.py:525-527, layer_clusters.py:
// 416-424).  Rows come out in the reference's order: merge nodes in
// breadth-first order from the root, each eliminated branch's points in
// breadth-first order of that branch.  lambda: 1/d (hdbscan_lambda_value,
// cluster_trees.py:192-197) or exp(-d) (plscan_lambda_value, 200-202).
enum LambdaMethod { kLambdaHdbscan = 0, kLambdaPlscan = 1 };

void condense_walk(int64_t n, Hierarchy &h, int64_t mcs, Condensed &c, int lambda_method) {
    const int64_t m = n - 1;
    const int32_t root = static_cast<int32_t>(2 * m);
    const int32_t num_points = static_cast<int32_t>(n);
    int64_t next_label = num_points + 1;
    // relabel is written for a node before the node is processed; only the
    // root needs a value up front.  Points are never processed: their
    // relabel is not kept.
    static thread_local std::vector<int32_t> relabel_tl;
    std::vector<int32_t> &relabel = relabel_tl;
    if (relabel.size() < static_cast<size_t>(m)) relabel.resize(m);
    relabel[root - num_points] = num_points;
    // a -1 child (only when min_samples >= n) indexes like numba: from the
    // end, i.e. the root's entry
    auto RL = [&](int32_t x, int32_t v) {
        if (x >= num_points) relabel[x - num_points] = v;
        else if (x < 0) relabel[x + root + 1 - num_points] = v;
    };
    // at most one row per node below the root
    const size_t cap = static_cast<size_t>(root);
    if (c.parent.size() < cap) {
        c.parent.resize(cap);
        c.child.resize(cap);
        c.lambda.resize(cap);
        c.size.resize(cap);
    }
    int64_t nrows = 0;
    const Merge *rows = h.rows.data();
    auto row = [&](int64_t p, int64_t ch, double lam, float sz) {
        c.parent[nrows] = static_cast<int32_t>(p);
        c.child[nrows] = static_cast<int32_t>(ch);
        c.lambda[nrows] = static_cast<float>(lam);
        c.size[nrows] = sz;
        ++nrows;
    };
    constexpr size_t kAhead = 16;
    // eliminate_branch (cluster_trees.py:115-145): the branch's points in
    // breadth-first order, written where the walk stands.  (A -1 leaf adds
    // a row but no size, so rows cannot be reserved by size here.)
    static thread_local std::vector<int32_t> sub_tl;
    std::vector<int32_t> &sub = sub_tl;
    auto eliminate = [&](int32_t branch, int64_t parent_node, double lam) {
        if (branch < num_points) {
            row(parent_node, branch, lam, 1.0f);
            return;
        }
        sub.clear();
        sub.push_back(branch);
        for (size_t head = 0; head < sub.size(); ++head) {
            const int32_t x = sub[head];
            if (x < num_points) {
                row(parent_node, x, lam, 1.0f);
            } else {
                sub.push_back(rows[x - num_points].left);
                sub.push_back(rows[x - num_points].right);
            }
        }
    };
    // The reference walks the whole breadth-first node list and skips
    // points and nodes inside eliminated branches (cluster_trees.py:228-230).
    // The nodes it does process are closed upwards, so a queue fed only with
    // the children that continue visits them in the same order.
    c.cluster_rows.clear();
    c.has_scores = false;
    static thread_local std::vector<int32_t> q;
    q.clear();
    q.reserve(static_cast<size_t>(n));
    q.push_back(root);
    auto push = [&](int32_t x) {
        if (x >= num_points) q.push_back(x);
    };
    const double dm = static_cast<double>(mcs);
    for (size_t qi = 0; qi < q.size(); ++qi) {
        if (qi + kAhead < q.size()) {
            const int32_t ahead = q[qi + kAhead];
            prefetch(&rows[ahead - num_points]);
        }
        const int32_t node = q[qi];
        const Merge &r = rows[node - num_points];
        const int64_t parent_node = relabel[node - num_points];
        const int32_t left = r.left, right = r.right;
        const double d = h.delta(node - num_points);
        const double lam = lambda_method == kLambdaHdbscan ? (d > 0.0 ? 1.0 / d : kInfD) : std::exp(-d);
        // counts as float32 (weighted sizes with weights), compared in
        // float64 as numba does.  A point child counts its sample weight,
        // 1 without weights (cluster_trees.py:252-260), also a -1 child,
        // whose recorded size is the wrap-around slot's
        const float lc = h.weighted ? h.lw[node - num_points]
                                    : (left < num_points ? 1.0f : static_cast<float>(r.left_size));
        const float rc = h.weighted ? h.rw[node - num_points]
                                    : (right < num_points ? 1.0f : static_cast<float>(r.right_size));
        const double dl = lc, dr = rc;
        if (left < num_points && dr >= dm) {
            RL(right, static_cast<int32_t>(parent_node));
            row(parent_node, left, lam, 1.0f);
            push(right);
        } else if (dl < dm && dr >= dm) {
            RL(right, static_cast<int32_t>(parent_node));
            eliminate(left, parent_node, lam);
            push(right);
        } else if (dl >= dm && dr < dm) {
            RL(left, static_cast<int32_t>(parent_node));
            eliminate(right, parent_node, lam);
            push(left);
        } else if (dl < dm && dr < dm) {
            eliminate(left, parent_node, lam);
            eliminate(right, parent_node, lam);
        } else {
            c.cluster_rows.push_back(nrows);
            c.cluster_rows.push_back(nrows + 1);
            RL(left, static_cast<int32_t>(next_label));
            row(parent_node, next_label, lam, lc);
            ++next_label;
            RL(right, static_cast<int32_t>(next_label));
            row(parent_node, next_label, lam, rc);
            ++next_label;
            push(left);
            push(right);
        }
    }
    c.rows = nrows;
    c.next_label = next_label;
    c.root = num_points;
}

// The same rows as condense_walk, computed without walking the tree one
// level at a time (the walk waits on memory once per level; a 2M-point
// tree has ~300k levels of ~4 processed nodes).  Processed merge nodes form
// a tree in which most nodes continue into one child only; a chain is a
// maximal path of processed nodes from the root or a split's child down to
// a split or to a node with no continuing child.  The walk's order is
// breadth-first: by depth, then left to right.  Two chains present at one
// depth lie in disjoint subtrees, so their left-to-right order is the
// preorder (left child first) of the chains' tree.  Passes (all but the
// small per-chain steps on the worker pool):
//   1a. classify every merge row: rows it writes, float32 lambda, and for
//       each merge child its parent row and role (dropped, same chain, new
//       chain's head);
//   1b. chain of every row, top-down (a parent row has a larger index than
//       its children): row ranges resolved locally, the few rows whose
//       parent lies in another range resolved in order, then substituted;
//   2.  nodes grouped by chain, top-down within each;
//   4.  the walk's order over depth ranges: each range starts from its
//       chains in preorder and its first slot, row and cluster id (from
//       per-depth node counts and per-range row and split totals): each
//       node's first row, the cluster ids of the splits (next_label counts
//       splits in the walk's order);
//   5.  rows written from those offsets; eliminated branches expanded
//       afterwards, breadth-first as eliminate_branch does.
// With want_scores, score_condensed_tree_nodes (cluster_trees.py:656-669)
// is summed per chain: a cluster's rows are its chain's rows, top-down,
// i.e. in row order, so each float32 sum runs in the reference's order.
void condense_chains(int64_t n, const Hierarchy &h, int64_t mcs, Condensed &c, int lambda_method, bool want_scores) {
    FH_TSTART();
    const int64_t m = n - 1;
    const int32_t N = static_cast<int32_t>(n);
    const Merge *rows = h.rows.data();
    const double dm = static_cast<double>(mcs);
    struct Node {
        int32_t row;    // merge row (node N + row)
        int32_t chain;
        int32_t nr;     // rows this node writes; -2 for a split (two cluster rows)
        float lam;      // float32 lambda of its rows
    };
    struct Chain {
        int32_t parent;     // chain ending in the split above, -1 for the root's
        int32_t len;
        int32_t kid[2];     // chains headed by the split's merge children, -1: none
        int32_t split_row;  // merge row of the split ending this chain, -1: none
        int32_t pad;
        int64_t start;        // depth of its head
        int64_t label;        // the cluster id its rows hang from
        int64_t split_label;  // the first cluster id its split creates
        int64_t first;        // its nodes' offset in `grouped`
    };
    static thread_local std::vector<int32_t> st_tl;
    static thread_local std::vector<Node> grouped_tl, bfs_tl;
    static thread_local std::vector<Chain> chains_tl;
    static thread_local std::vector<int64_t> roff_tl;
    std::vector<int32_t> &st = st_tl;
    std::vector<Node> &grouped = grouped_tl, &bfs = bfs_tl;
    std::vector<Chain> &chains = chains_tl;
    std::vector<int64_t> &roff = roff_tl;
    auto side_size = [&](int64_t i, int s) -> float {
        if (h.weighted) return s == 0 ? h.lw[i] : h.rw[i];
        return static_cast<float>(s == 0 ? rows[i].left_size : rows[i].right_size);
    };

    // 1a. what each merge row does if processed (independent rows:
    //     threads).  Each merge child learns its parent row and its role
    //     there: dropped (eliminated), continuing the parent's chain, or
    //     heading a new chain (left or right child of a split).
    enum : int32_t { kDrop = 0, kSame = 1, kHeadL = 2, kHeadR = 3 };
    struct Up {
        int32_t parent;
        int32_t role;
    };
    static thread_local std::vector<Up> up_tl;
    static thread_local std::vector<int32_t> nr_tl;
    static thread_local std::vector<float> lamv_tl;
    std::vector<Up> &up = up_tl;
    std::vector<int32_t> &nrv = nr_tl;
    std::vector<float> &lamv = lamv_tl;
    if (up.size() < static_cast<size_t>(m)) {
        up.resize(m);
        nrv.resize(m);
        lamv.resize(m);
    }
    if (st.size() < static_cast<size_t>(m)) st.resize(m);
    const int nt = m > (1 << 16) ? fh::HostPool::width() : 1;
    auto classify = [&](int64_t i0, int64_t i1) {
        for (int64_t i = i0; i < i1; ++i) {
            const Merge &r = rows[i];
            const double dl = side_size(i, 0), dr = side_size(i, 1);
            const double d = h.delta(i);
            const double lam = lambda_method == kLambdaHdbscan ? (d > 0.0 ? 1.0 / d : kInfD) : std::exp(-d);
            int32_t nr, lrole = kDrop, rrole = kDrop;
            if (r.left < N && dr >= dm) {
                nr = 1;
                rrole = kSame;
            } else if (dl < dm && dr >= dm) {
                nr = r.left_size;
                rrole = kSame;
            } else if (dl >= dm && dr < dm) {
                nr = r.right_size;
                lrole = kSame;
            } else if (dl < dm && dr < dm) {
                nr = r.left_size + r.right_size;
            } else {
                nr = -2;
                lrole = kHeadL;
                rrole = kHeadR;
            }
            if (r.left >= N) up[r.left - N] = {static_cast<int32_t>(i), lrole};
            if (r.right >= N) up[r.right - N] = {static_cast<int32_t>(i), rrole};
            nrv[i] = nr;
            lamv[i] = static_cast<float>(lam);
        }
    };
    fh::HostPool::run(nt, [&](int k) { classify(m * k / nt, m * (k + 1) / nt); });
    FH_TMARK("1a");

    // 1b. chain of every row, top-down (a row's parent has a larger index);
    //     -1: not processed.  Threads own row ranges.  In its range a thread
    //     resolves dropped rows (-1) and rows continuing a parent of the same
    //     range (the parent's value); the rest -- heads of new chains and
    //     rows whose parent lies in another range -- are pending, held as
    //     -2 - row, and resolved in decreasing row order on one thread (a
    //     few thousand), after which the threads replace every reference.
    chains.clear();
    chains.push_back({-1, 0, {-1, -1}, -1, 0, 0, N, 0, 0});
    std::vector<std::vector<int32_t>> pend(nt);
    auto lo_of = [&](int k) { return m * k / nt; };
    auto local = [&](int k) {
        int32_t *stp = st.data();
        const Up *upp = up.data();
        const int64_t lo = lo_of(k), hi = lo_of(k + 1);
        std::vector<int32_t> &pk = pend[k];
        pk.clear();
        for (int64_t i = hi - 1; i >= lo; --i) {
            if (i == m - 1) {
                stp[i] = 0;
                continue;
            }
            const Up u = upp[i];
            if (u.role == kDrop) {
                stp[i] = -1;
            } else if (u.role == kSame && u.parent < hi) {
                stp[i] = stp[u.parent];
            } else {
                stp[i] = -2 - static_cast<int32_t>(i);
                pk.push_back(static_cast<int32_t>(i));
            }
        }
    };
    auto resolve = [&](int k) {
        int32_t *stp = st.data();
        for (int64_t i = lo_of(k), e = lo_of(k + 1); i < e; ++i)
            if (stp[i] <= -2) stp[i] = stp[-2 - stp[i]];
    };
    auto on_threads = [&](const std::function<void(int)> &fn) { fh::HostPool::run(nt, fn); };
    on_threads(local);
    {
        int32_t *stp = st.data();
        for (int k = nt - 1; k >= 0; --k)
            for (const int32_t i : pend[k]) {
                const Up u = up[i];
                int32_t cp = stp[u.parent];
                if (cp <= -2) cp = stp[-2 - cp];
                int32_t ch = u.role == kSame ? cp : -1;
                if (cp >= 0 && u.role >= kHeadL) {
                    ch = static_cast<int32_t>(chains.size());
                    chains[cp].split_row = u.parent;
                    chains[cp].kid[u.role - kHeadL] = ch;
                    chains.push_back({cp, 0, {-1, -1}, -1, 0, 0, 0, 0, 0});
                }
                stp[i] = ch;
            }
    }
    on_threads(resolve);
    FH_TMARK("1b");

    // 2. nodes grouped by chain, top-down within each.  Chain sizes are
    //    counted per range; a range's nodes of a chain follow those of the
    //    ranges above it.
    const size_t C = chains.size();
    std::vector<std::vector<int64_t>> part(nt, std::vector<int64_t>(C + 1, 0));
    on_threads([&](int k) {
        int64_t *pc = part[k].data();
        const int32_t *stp = st.data();
        for (int64_t i = lo_of(k), e = lo_of(k + 1); i < e; ++i) ++pc[stp[i] + 1];
    });
    int64_t P = 0;
    for (size_t x = 0; x < C; ++x) {
        chains[x].first = P;
        int64_t acc = P;
        for (int k = nt - 1; k >= 0; --k) {
            const int64_t c_k = part[k][x + 1];
            part[k][x + 1] = acc;  // this range's first slot for chain x
            acc += c_k;
        }
        chains[x].len = static_cast<int32_t>(acc - P);
        P = acc;
    }
    if (grouped.size() < static_cast<size_t>(P)) grouped.resize(P);
    on_threads([&](int k) {
        int64_t *cur = part[k].data() + 1;
        const int32_t *stp = st.data();
        Node *g = grouped.data();
        for (int64_t i = lo_of(k + 1) - 1, lo = lo_of(k); i >= lo; --i) {
            const int32_t ch = stp[i];
            if (ch < 0) continue;
            g[cur[ch]++] = {static_cast<int32_t>(i), ch, nrv[i], lamv[i]};
        }
    });
    FH_TMARK("2");

    // 4. the walk's order: depth by depth, the chains present in preorder
    //    (a chain ending in a split is followed by its kids, left first);
    //    each node's first row and the cluster ids of the splits.  The list
    //    of chains changes only where one ends: it is rebuilt there only.
    //    Threads take ranges of depths: per-depth node, row and split counts
    //    give each range its first output slot, row and cluster id, and the
    //    chains present at its first depth, in preorder, its starting list.
    for (size_t x = 1; x < C; ++x) chains[x].start = chains[chains[x].parent].start + chains[chains[x].parent].len;
    int64_t depth_end = 0;
    for (const Chain &cx : chains) depth_end = std::max(depth_end, cx.start + cx.len);
    std::vector<int32_t> rank(C);
    {
        std::vector<int32_t> stack{0};
        int32_t r = 0;
        while (!stack.empty()) {
            const int32_t x = stack.back();
            stack.pop_back();
            rank[x] = r++;
            for (int s2 = 1; s2 >= 0; --s2)
                if (chains[x].kid[s2] >= 0) stack.push_back(chains[x].kid[s2]);
        }
    }
    // nodes per depth (chain spans), as first output slot per depth
    static thread_local std::vector<int64_t> dn_tl;
    std::vector<int64_t> &dn = dn_tl;
    dn.assign(depth_end + 1, 0);
    for (const Chain &cx : chains) {
        ++dn[cx.start];
        --dn[cx.start + cx.len];
    }
    {
        int64_t active = 0, on = 0;
        for (int64_t k = 0; k <= depth_end; ++k) {
            active += dn[k];
            dn[k] = on;
            on += active;
        }
    }
    bfs.resize(P);
    roff.resize(P);
    // depth ranges of about equal node counts; each range's rows and splits
    // summed by its thread, then scanned: its first row and cluster id
    const int T = P > (1 << 16) ? nt : 1;
    std::vector<int64_t> kcut(T + 1, depth_end);
    kcut[0] = 0;
    for (int t = 1; t < T; ++t)
        kcut[t] = std::upper_bound(dn.begin(), dn.begin() + depth_end, P * t / T) - dn.begin() - 1;
    for (int t = 1; t <= T; ++t) kcut[t] = std::max(kcut[t], kcut[t - 1]);
    std::vector<int64_t> row0(T + 1, 0), split0(T + 1, 0);
    fh::HostPool::run(T, [&](int t) {
        const int64_t k0 = kcut[t], k1 = kcut[t + 1];
        int64_t rws = 0, sps = 0;
        for (size_t x = 0; x < C; ++x) {
            const Chain &cx = chains[x];
            const int64_t lo = std::max(cx.start, k0), hi = std::min(cx.start + cx.len, k1);
            if (lo >= hi) continue;
            const Node *g = grouped.data() + cx.first - cx.start;
            for (int64_t k = lo; k < hi; ++k) {
                const int32_t nr = g[k].nr;
                rws += nr == -2 ? 2 : nr;
                sps += nr == -2;
            }
        }
        row0[t + 1] = rws;
        split0[t + 1] = sps;
    });
    for (int t = 0; t < T; ++t) {
        row0[t + 1] += row0[t];
        split0[t + 1] += split0[t];
    }
    const int64_t nrows = row0[T];
    const int64_t next_label = N + 1 + 2 * split0[T];
    std::vector<std::vector<int64_t>> crows(T);
    fh::HostPool::run(T, [&](int t) {
        const int64_t k0 = kcut[t], k1 = kcut[t + 1];
        if (k0 >= k1) return;
        struct Live {
            int32_t x;
            int64_t pos;
        };
        std::vector<Live> cur, nxt;
        for (size_t x = 0; x < C; ++x)
            if (chains[x].start <= k0 && k0 < chains[x].start + chains[x].len)
                cur.push_back({static_cast<int32_t>(x), chains[x].first + (k0 - chains[x].start)});
        std::sort(cur.begin(), cur.end(), [&](const Live &a, const Live &b) { return rank[a.x] < rank[b.x]; });
        Node *bp = bfs.data();
        int64_t *rp = roff.data();
        int64_t out = dn[k0], rw = row0[t], lab = N + 1 + 2 * split0[t];
        std::vector<int64_t> &cr = crows[t];
        int64_t k = k0;
        while (!cur.empty() && k < k1) {
            int64_t change = std::numeric_limits<int64_t>::max();
            for (const Live &l : cur) change = std::min(change, chains[l.x].start + chains[l.x].len);
            const int64_t stop = std::min(change, k1);
            for (; k < stop; ++k) {
                for (Live &l : cur) {
                    const Node v = grouped[l.pos++];
                    bp[out] = v;
                    rp[out++] = rw;
                    if (v.nr == -2) {
                        Chain &cx = chains[l.x];
                        cx.split_label = lab;
                        for (int s2 = 0; s2 < 2; ++s2)
                            if (cx.kid[s2] >= 0) chains[cx.kid[s2]].label = lab + s2;
                        cr.push_back(rw);
                        cr.push_back(rw + 1);
                        lab += 2;
                        rw += 2;
                    } else {
                        rw += v.nr;
                    }
                }
            }
            if (k == k1) break;
            nxt.clear();
            for (const Live &l : cur) {
                const Chain &cx = chains[l.x];
                if (cx.start + cx.len > change) {
                    nxt.push_back(l);
                } else {
                    for (int s2 = 0; s2 < 2; ++s2)
                        if (cx.kid[s2] >= 0) nxt.push_back({cx.kid[s2], chains[cx.kid[s2]].first});
                }
            }
            cur.swap(nxt);
        }
    });
    c.cluster_rows.clear();
    for (const auto &cr : crows) c.cluster_rows.insert(c.cluster_rows.end(), cr.begin(), cr.end());
    FH_TMARK("4");
    // 5. rows
    const size_t cap = static_cast<size_t>(2 * m);
    if (c.parent.size() < cap) {
        c.parent.resize(cap);
        c.child.resize(cap);
        c.lambda.resize(cap);
        c.size.resize(cap);
    }
    int32_t *cpar = c.parent.data();
    int32_t *cchild = c.child.data();
    float *clam = c.lambda.data();
    float *csize = c.size.data();
    constexpr int64_t kAhead = 16;
    // An eliminated branch's points in breadth-first order (eliminate_branch,
    // cluster_trees.py:115-145).  Each walk waits on memory once per level,
    // so kWalks branches are walked side by side, one queue entry each in
    // turn, and a merge node's row is prefetched when it is queued.
    struct Branch {
        int32_t root;
        int32_t label;
        float lam;
        int64_t at;
    };
    constexpr int kWalks = 16;
    auto expand = [&](const std::vector<Branch> &todo) {
        struct Walk {
            std::vector<int32_t> q;
            size_t head = 0;
            int64_t at = 0;
            int32_t label = 0;
            float lam = 0.0f;
            bool live = false;
        };
        Walk w[kWalks];
        size_t next = 0;
        auto start = [&](Walk &x) {
            x.live = next < todo.size();
            if (!x.live) return;
            const Branch &b = todo[next++];
            x.q.clear();
            x.q.push_back(b.root);
            x.head = 0;
            x.at = b.at;
            x.label = b.label;
            x.lam = b.lam;
        };
        for (auto &x : w) start(x);
        for (int live = kWalks; live > 0;) {
            live = 0;
            for (auto &x : w) {
                if (!x.live) continue;
                ++live;
                const int32_t y = x.q[x.head++];
                if (y < N) {
                    cpar[x.at] = x.label;
                    cchild[x.at] = y;
                    clam[x.at] = x.lam;
                    csize[x.at] = 1.0f;
                    ++x.at;
                } else {
                    const int32_t l = rows[y - N].left, r = rows[y - N].right;
                    if (l >= N) prefetch(&rows[l - N]);
                    if (r >= N) prefetch(&rows[r - N]);
                    x.q.push_back(l);
                    x.q.push_back(r);
                }
                if (x.head == x.q.size()) start(x);
            }
        }
    };
    auto emit = [&](int64_t p0, int64_t p1) {
        std::vector<Branch> todo;
        for (int64_t p = p0; p < p1; ++p) {
            if (p + kAhead < p1) prefetch(&rows[bfs[p + kAhead].row]);
            const Node &v = bfs[p];
            const Merge &r = rows[v.row];
            const Chain &cv = chains[v.chain];
            const int32_t label = static_cast<int32_t>(cv.label);
            const float lam = v.lam;
            int64_t at = roff[p];
            auto put = [&](int64_t child, float sz) {
                cpar[at] = label;
                cchild[at] = static_cast<int32_t>(child);
                clam[at] = lam;
                csize[at] = sz;
                ++at;
            };
            // a point's row now; a merge node's rows reserved for expand()
            auto eliminate = [&](int32_t b) {
                if (b < N) {
                    put(b, 1.0f);
                    return;
                }
                if (b < N + m) prefetch(&rows[b - N]);
                todo.push_back({b, label, lam, at});
                at += rows[b - N].left_size + rows[b - N].right_size;
            };
            if (v.nr == -2) {
                put(cv.split_label, side_size(v.row, 0));
                put(cv.split_label + 1, side_size(v.row, 1));
                continue;
            }
            const double dl = side_size(v.row, 0), dr = side_size(v.row, 1);
            if (r.left < N && dr >= dm) {
                put(r.left, 1.0f);
            } else if (dl < dm && dr >= dm) {
                eliminate(r.left);
            } else if (dl >= dm && dr < dm) {
                eliminate(r.right);
            } else {
                eliminate(r.left);
                eliminate(r.right);
            }
        }
        expand(todo);
    };
    if (nt == 1) {
        emit(0, P);
    } else {
        // pieces of about equal row counts, taken from a shared counter
        const int pieces = 8 * nt;
        std::vector<int64_t> cut(pieces + 1, P);
        cut[0] = 0;
        for (int k = 1; k < pieces; ++k)
            cut[k] = std::lower_bound(roff.begin(), roff.begin() + P, nrows * k / pieces) - roff.begin();
        fh::HostPool::run(pieces, [&](int k) { emit(cut[k], cut[k + 1]); });
    }
    FH_TMARK("5");
    c.rows = nrows;
    c.next_label = next_label;
    c.root = N;

    // scores per cluster id, each chain's rows in order
    c.has_scores = want_scores;
    if (!want_scores) return;
    c.score.assign(static_cast<size_t>(next_label - N), 0.0f);
    float *score = c.score.data();
    auto chain_scores = [&](size_t x0, size_t x1) {
        for (size_t x = x0; x < x1; ++x) {
            const Chain &cx = chains[x];
            const Node *g = grouped.data() + cx.first;
            float s = score[cx.label - N];  // the split row's initial value (or 0)
            for (int32_t j = 0; j < cx.len; ++j) {
                const float lam = g[j].lam;
                if (g[j].nr == -2) {
                    s = s + lam * side_size(g[j].row, 0);
                    s = s + lam * side_size(g[j].row, 1);
                } else {
                    for (int32_t k = 0; k < g[j].nr; ++k) s = s + lam;
                }
            }
            score[cx.label - N] = s;
        }
    };
    // split rows set their children's initial scores (-lambda * size when
    // size > 1), before any of the children's own rows
    for (const Chain &cx : chains) {
        if (cx.split_row < 0) continue;
        const Node &sv = grouped[cx.first + cx.len - 1];
        for (int s = 0; s < 2; ++s) {
            const float sz = side_size(cx.split_row, s);
            if (sz > 1.0f) score[cx.split_label + s - N] = -sv.lam * sz;
        }
    }
    if (nt == 1 || C < 64) {
        chain_scores(0, C);
    } else {
        fh::HostPool::run(nt, [&](int k) { chain_scores(C * k / nt, C * (k + 1) / nt); });
    }
}

// condense_tree (cluster_trees.py:205-352): the chain form, or the
// level-by-level walk when a merge has a -1 child (min_samples >= n,
// whose wrap-around indexing the walk reproduces).
void condense(int64_t n, Hierarchy &h, int64_t mcs, Condensed &c, int lambda_method = kLambdaHdbscan,
              bool want_scores = false) {
    if (h.literal || n < 3)
        condense_walk(n, h, mcs, c, lambda_method);
    else
        condense_chains(n, h, mcs, c, lambda_method, want_scores);
}

struct ClusterTree {
    int64_t root = -1;                        // smallest parent id (== n)
    int64_t n_ids = 0;                        // ids are [root, root + n_ids)
    std::vector<int64_t> parent, child;       // cluster rows (child_size > 1)
    std::vector<float> lambda, size;
    std::vector<int64_t> first_kid, next_sib;  // children lists in row order
};

void cluster_tree(const Condensed &c, ClusterTree &t) {
    t.parent.clear();
    t.child.clear();
    t.lambda.clear();
    t.size.clear();
    for (const int64_t i : c.cluster_rows) {
        if (c.size[i] > 1.0f) {
            t.parent.push_back(c.parent[i]);
            t.child.push_back(c.child[i]);
            t.lambda.push_back(c.lambda[i]);
            t.size.push_back(c.size[i]);
        }
    }
}

// extract_eom_clusters (cluster_trees.py:725-771), score_condensed_tree_nodes
// (656-669: float32 accumulation in row order) and eom_recursion (689-722,
// float64 totals), iteratively.
std::vector<int64_t> eom(const Condensed &c, const ClusterTree &ct, int64_t n, double max_cluster_size,
                         bool allow_single) {
    if (ct.parent.empty()) return {};
    const int64_t root = n;
    const int64_t nid = c.next_label - root;
    std::vector<float> scanned;
    if (!c.has_scores) {
        scanned.assign(nid, 0.0f);
        for (size_t i = 0; i < static_cast<size_t>(c.rows); ++i) {
            if (c.size[i] > 1.0f) scanned[c.child[i] - root] = -c.lambda[i] * c.size[i];
            float &s = scanned[c.parent[i] - root];
            s = s + c.lambda[i] * c.size[i];
        }
    }
    const float *score = c.has_scores ? c.score.data() : scanned.data();
    std::vector<float> node_size(nid, 0.0f);
    for (size_t i = 0; i < ct.child.size(); ++i) node_size[ct.child[i] - root] = ct.size[i];
    const int64_t ct_root = *std::min_element(ct.parent.begin(), ct.parent.end());
    node_size[ct_root - root] = static_cast<float>(ct_root - 1);
    std::vector<int64_t> first(nid, -1), nxt(ct.child.size(), -1), last(nid, -1);
    for (size_t i = 0; i < ct.child.size(); ++i) {
        const int64_t p = ct.parent[i] - root;
        if (first[p] < 0)
            first[p] = static_cast<int64_t>(i);
        else
            nxt[last[p]] = static_cast<int64_t>(i);
        last[p] = static_cast<int64_t>(i);
    }
    std::vector<uint8_t> selected(nid, 0);
    std::vector<double> ret(nid, 0.0);
    auto unselect_below = [&](int64_t node) {
        std::vector<int64_t> st{node};
        while (!st.empty()) {
            const int64_t x = st.back();
            st.pop_back();
            for (int64_t e = first[x - root]; e >= 0; e = nxt[e]) {
                selected[ct.child[e] - root] = 0;
                st.push_back(ct.child[e]);
            }
        }
    };
    // post-order evaluation of eom_recursion from `start`
    auto run = [&](int64_t start) {
        std::vector<std::pair<int64_t, int64_t>> st;  // (node, next child edge)
        st.push_back({start, first[start - root]});
        while (!st.empty()) {
            auto &top = st.back();
            if (top.second >= 0) {
                const int64_t e = top.second;
                top.second = nxt[e];
                st.push_back({ct.child[e], first[ct.child[e] - root]});
                continue;
            }
            const int64_t node = top.first;
            st.pop_back();
            const double cur = std::max(static_cast<double>(score[node - root]), 0.0);
            double total = 0.0;
            for (int64_t e = first[node - root]; e >= 0; e = nxt[e]) total += ret[ct.child[e] - root];
            if (total > cur || static_cast<double>(node_size[node - root]) > max_cluster_size) {
                ret[node - root] = total;
            } else {
                selected[node - root] = 1;
                unselect_below(node);
                ret[node - root] = cur;
            }
        }
    };
    if (allow_single) {
        run(ct_root);
    } else {
        // len(node_scores) > 1 holds whenever the cluster tree is non-empty
        for (int64_t e = first[ct_root - root]; e >= 0; e = nxt[e]) run(ct.child[e]);
    }
    std::vector<int64_t> out;
    for (int64_t i = 0; i < nid; ++i)
        if (selected[i]) out.push_back(root + i);
    return out;
}

// cluster_tree_leaves (cluster_trees.py:369-374)
std::vector<int64_t> leaves(const ClusterTree &ct, int64_t n) {
    if (ct.parent.empty()) return {};
    const int64_t n_nodes = *std::max_element(ct.child.begin(), ct.child.end()) + 1;
    std::vector<uint8_t> leaf(n_nodes - n, 1);
    for (const int64_t p : ct.parent) leaf[p - n] = 0;
    std::vector<int64_t> out;
    for (int64_t i = 0; i < n_nodes - n; ++i)
        if (leaf[i]) out.push_back(i + n);
    return out;
}

// cluster_epsilon_search, traverse_upwards, segments_in_branch
// (cluster_trees.py:834-903).  segments_in_branch looks every segment up with
// searchsorted among the sorted parents, so a segment that is not a parent
// picks up the two children of the next larger parent id; kept as is.
std::vector<int64_t> epsilon_search(const std::vector<int64_t> &clusters, const ClusterTree &ct, double min_eps) {
    const size_t r = ct.parent.size();
    std::vector<size_t> order(r);
    for (size_t i = 0; i < r; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return ct.parent[a] < ct.parent[b]; });
    std::vector<int64_t> parents(r), children(r);
    for (size_t i = 0; i < r; ++i) {
        parents[i] = ct.parent[order[i]];
        children[i] = ct.child[order[i]];
    }
    const int64_t root = *std::min_element(ct.parent.begin(), ct.parent.end());
    auto row_of_child = [&](int64_t c) {
        return static_cast<size_t>(std::lower_bound(ct.child.begin(), ct.child.end(), c) - ct.child.begin());
    };
    auto death_eps = [&](float lam) { return lam > 0.0f ? 1.0 / static_cast<double>(lam) : kInfD; };
    std::vector<int64_t> selected;
    std::vector<int64_t> processed;  // kept sorted
    auto is_processed = [&](int64_t c) { return std::binary_search(processed.begin(), processed.end(), c); };
    for (const int64_t cluster : clusters) {
        const size_t idx = row_of_child(cluster);
        if (death_eps(ct.lambda[idx]) < min_eps) {
            if (!is_processed(cluster)) {
                // traverse_upwards
                int64_t seg = cluster, parent;
                for (;;) {
                    parent = ct.parent[row_of_child(seg)];
                    if (parent == root) break;
                    if (death_eps(ct.lambda[row_of_child(parent)]) >= min_eps) break;
                    seg = parent;
                }
                selected.push_back(parent);
                // segments_in_branch
                std::vector<int64_t> to_process{parent}, found;
                while (!to_process.empty()) {
                    found.insert(found.end(), to_process.begin(), to_process.end());
                    std::vector<int64_t> kids;
                    for (const int64_t s : to_process) {
                        const size_t i =
                            static_cast<size_t>(std::lower_bound(parents.begin(), parents.end(), s) - parents.begin());
                        if (i >= parents.size()) continue;
                        kids.push_back(children[i]);
                        kids.push_back(children[i + 1]);
                    }
                    std::sort(kids.begin(), kids.end());
                    kids.erase(std::unique(kids.begin(), kids.end()), kids.end());
                    to_process.swap(kids);
                }
                processed.insert(processed.end(), found.begin(), found.end());
                std::sort(processed.begin(), processed.end());
                processed.erase(std::unique(processed.begin(), processed.end()), processed.end());
            }
        } else {
            selected.push_back(cluster);
        }
    }
    return selected;
}

// get_cluster_label_vector (cluster_trees.py:944-996) and
// get_point_membership_strength_vector (999-1031).
void label_vectors(const Condensed &c, std::vector<int64_t> clusters, double eps, int64_t n, int64_t *labels,
                   float *prob) {
    // Every point is the child of exactly one row unless the tree came
    // from the -1 wrap-around linkage: then points without a row stay noise.
    if (c.rows == 0 || clusters.empty() || c.rows - static_cast<int64_t>(c.cluster_rows.size()) != n) {
        for (int64_t i = 0; i < n; ++i) {
            labels[i] = -1;
            prob[i] = 0.0f;
        }
    }
    if (c.rows == 0 || clusters.empty()) return;
    // ids: the root cluster (smallest parent) up to the last cluster id
    // (largest parent or child; point ids are smaller)
    const int64_t root_cluster = c.root;
    const int64_t size = c.next_label;
    std::sort(clusters.begin(), clusters.end());
    clusters.erase(std::unique(clusters.begin(), clusters.end()), clusters.end());
    // cluster-id tables cover ids [root_cluster, size); points are never
    // clusters
    const int64_t nc = size - root_cluster;
    std::vector<int64_t> label_of(nc, -1);  // cluster id -> label
    std::vector<uint8_t> is_cluster(nc, 0);
    for (size_t k = 0; k < clusters.size(); ++k) {
        label_of[clusters[k] - root_cluster] = static_cast<int64_t>(k);
        is_cluster[clusters[k] - root_cluster] = 1;
    }
    auto selected = [&](int64_t id) { return id >= root_cluster && is_cluster[id - root_cluster]; };
    // The reference unions every non-selected child into its parent with
    // union by rank (cluster_trees.py:964-978).  Rows list parents before
    // children, so each child joins as a fresh singleton and every set keeps
    // its topmost node as root: a selected cluster, the root cluster, or a
    // point whose row was skipped (lambda 0).  That root is computed
    // directly: set_root(c) = c if selected or c is the root, else
    // set_root(parent(c)).
    std::vector<int64_t> set_root(nc, -1);
    set_root[0] = root_cluster;
    for (const int64_t r : c.cluster_rows) {
        const int64_t child = c.child[r];
        set_root[child - root_cluster] = selected(child) ? child : set_root[c.parent[r] - root_cluster];
    }
    // single selected cluster at the root: the largest lambda among the
    // root's rows (cluster_trees.py:984)
    const bool single_root = clusters.size() == 1 && clusters[0] == root_cluster;
    float root_max_lambda = 0.0f;
    if (single_root) {
        bool any = false;
        for (size_t r = 0; r < static_cast<size_t>(c.rows); ++r)
            if (c.parent[r] == root_cluster) {
                root_max_lambda = any ? std::max(root_max_lambda, c.lambda[r]) : c.lambda[r];
                any = true;
            }
    }
    // deaths: largest lambda of a point row directly under each selected
    // cluster (max_lambdas, cluster_trees.py:999-1008)
    // (a maximum: any split of the rows over threads gives the same values)
    const int64_t rows = c.rows;
    const int nt = rows > (1 << 16) ? fh::HostPool::width() : 1;
    std::vector<double> deaths(nc, 0.0);
    {
        std::vector<std::vector<double>> part(nt, std::vector<double>(nc, 0.0));
        auto scan = [&](int t) {
            std::vector<double> &dt = part[t];
            const int64_t r0 = rows * t / nt, r1 = rows * (t + 1) / nt;
            for (int64_t r = r0; r < r1; ++r) {
                const int64_t p = c.parent[r];
                if (selected(p) && c.size[r] == 1.0f)
                    dt[p - root_cluster] = std::max(dt[p - root_cluster], static_cast<double>(c.lambda[r]));
            }
        };
        fh::HostPool::run(nt, scan);
        for (int t = 0; t < nt; ++t)
            for (int64_t k = 0; k < nc; ++k) deaths[k] = std::max(deaths[k], part[t][k]);
    }
    // Every point is the child of exactly one row: its label follows from
    // that row alone (a skipped lambda-0 row leaves the point on its own,
    // i.e. noise), and so does its strength, which the reference computes
    // from the same row.  Rows are independent: split over threads.
    auto work = [&](int64_t r0, int64_t r1) {
        for (int64_t r = r0; r < r1; ++r) {
            const int64_t child = c.child[r];
            if (child >= n || child < 0) continue;
            const float lv = c.lambda[r];
            const int64_t cl = lv == 0.0f ? child : set_root[c.parent[r] - root_cluster];
            int64_t lab;
            if (cl < root_cluster) {
                lab = -1;
            } else if (cl == root_cluster) {
                // the reference looks the root up in the label map here; it
                // is only there when the root is the one selected cluster
                if (single_root) {
                    if (eps > 0.0)
                        lab = static_cast<double>(lv) >= 1.0 / eps ? label_of[0] : -1;
                    else
                        lab = lv >= root_max_lambda ? label_of[0] : -1;
                } else {
                    lab = -1;
                }
            } else {
                lab = label_of[cl - root_cluster];
            }
            labels[child] = lab;
            float pr = 0.0f;
            if (lab >= 0) {
                const double ml = deaths[clusters[lab] - root_cluster];
                if (ml == 0.0 || !std::isfinite(lv)) {
                    pr = 1.0f;
                } else {
                    const double lam = std::min(static_cast<double>(lv), ml);
                    pr = static_cast<float>(lam / ml);
                }
            }
            prob[child] = pr;
        }
    };
    fh::HostPool::run(nt, [&](int t) { work(rows * t / nt, rows * (t + 1) / nt); });
    // A -1 child (only with min_samples >= n) is a row of an +inf edge:
    // the reference skips it for labels and reads labels[-1] (the last
    // point) for the strength, which is noise in that configuration.
}

// ---------------------------------------------------------------- PLSCAN ---

// extract_leaves (cluster_trees.py:356-366): ids in [min parent, max parent]
// that have no child row of size > 1.
std::vector<int64_t> extract_leaves(const Condensed &c) {
    if (c.rows == 0) return {};
    const int64_t lo = *std::min_element(c.parent.begin(), c.parent.begin() + c.rows);
    const int64_t hi = *std::max_element(c.parent.begin(), c.parent.begin() + c.rows);
    std::vector<uint8_t> leaf(static_cast<size_t>(hi - lo + 1), 1);
    for (int64_t r = 0; r < c.rows; ++r)
        if (c.size[r] > 1.0f) leaf[c.parent[r] - lo] = 0;
    std::vector<int64_t> out;
    for (int64_t i = 0; i <= hi - lo; ++i)
        if (leaf[i]) out.push_back(lo + i);
    return out;
}

// find_peaks (layer_clusters.py:26-68)
std::vector<int64_t> find_peaks(const std::vector<float> &x) {
    std::vector<int64_t> mids;
    const int64_t i_max = static_cast<int64_t>(x.size()) - 1;
    int64_t i = 1;
    while (i < i_max) {
        if (x[i - 1] < x[i]) {
            int64_t ia = i + 1;
            while (ia < i_max && x[ia] == x[i]) ++ia;
            if (x[ia] < x[i]) {
                mids.push_back((i + ia - 1) / 2);
                i = ia;
            }
        }
        ++i;
    }
    return mids;
}

struct Barcode {
    std::vector<float> births, deaths, lambda_deaths;
};

// min_cluster_size_barcode (layer_clusters.py:183-212) over the rows whose
// child is a cluster; those come in sibling pairs with consecutive ids.
Barcode barcode(const std::vector<int64_t> &parent, const std::vector<int64_t> &child, const std::vector<float> &lam,
                const std::vector<float> &size, int64_t n, int64_t min_size) {
    const int64_t rows = static_cast<int64_t>(child.size());
    const int64_t n_nodes = child.back() - n + 1;
    Barcode b;
    b.births.assign(n_nodes, static_cast<float>(min_size));
    b.deaths.assign(n_nodes, 0.0f);
    b.lambda_deaths.assign(n_nodes, 0.0f);
    b.deaths[0] = static_cast<float>(n);
    for (int64_t idx = rows - 1; idx > 0; idx -= 2) {
        const int64_t o = child[idx] - n;
        const float lv = lam[idx];
        const float ld = lv > 0.0f ? static_cast<float>(std::exp(-1.0 / static_cast<double>(lv))) : 0.0f;
        b.lambda_deaths[o - 1] = b.lambda_deaths[o] = ld;
        const float ds = std::min(size[idx - 1], size[idx]);
        b.deaths[o - 1] = b.deaths[o] = ds;
        b.births[parent[idx] - n] = std::max(std::max(b.births[o - 1], b.births[o]), ds);
    }
    return b;
}

// compute_total_persistence (layer_clusters.py:215-248): float32 sums in
// node order over the sizes in [birth, death).
void total_persistence(const Barcode &b, std::vector<float> &sizes, std::vector<float> &tp) {
    sizes = b.births;
    std::sort(sizes.begin(), sizes.end());
    sizes.erase(std::unique(sizes.begin(), sizes.end()), sizes.end());
    tp.assign(sizes.size(), 0.0f);
    for (size_t i = 1; i < b.births.size(); ++i) {
        const float birth = b.births[i], death = b.deaths[i];
        if (death <= birth) continue;
        const size_t bi = static_cast<size_t>(std::lower_bound(sizes.begin(), sizes.end(), birth) - sizes.begin());
        const size_t di = static_cast<size_t>(std::lower_bound(sizes.begin(), sizes.end(), death) - sizes.begin());
        // (a birth above every size leaves birth_idx at 0 in the reference)
        const size_t b0 = bi == sizes.size() ? 0 : bi;
        const float prod = birth * b.lambda_deaths[i];
        const float v = death - prod;
        for (size_t k = b0; k < di; ++k) tp[k] += v;
    }
}

// clusters alive at size s: birth <= s < death (estimate_cluster_similarity,
// layer_clusters.py:265-289)
template <class S>
std::vector<int64_t> active_at(const Barcode &b, S s) {
    std::vector<int64_t> out;
    for (size_t i = 0; i < b.births.size(); ++i)
        if (static_cast<S>(b.births[i]) <= s && static_cast<S>(b.deaths[i]) > s) out.push_back(static_cast<int64_t>(i));
    return out;
}

// jaccard_similarity (layer_clusters.py:251-262) of two ascending id lists
double jaccard(const std::vector<int64_t> &a, const std::vector<int64_t> &b) {
    size_t i = 0, j = 0, inter = 0;
    while (i < a.size() && j < b.size()) {
        if (a[i] == b[j]) {
            ++inter;
            ++i;
            ++j;
        } else if (a[i] < b[j]) {
            ++i;
        } else {
            ++j;
        }
    }
    const size_t uni = a.size() + b.size() - inter;
    return uni > 0 ? static_cast<double>(inter) / static_cast<double>(uni) : 0.0;
}

// select_diverse_peaks (layer_clusters.py:309-350)
std::vector<int64_t> select_diverse_peaks(const std::vector<int64_t> &peaks, const std::vector<float> &tp,
                                          const std::vector<float> &sizes, const Barcode &b, double threshold,
                                          int64_t max_layers) {
    if (peaks.empty()) return {};
    std::vector<float> pp(peaks.size());
    for (size_t i = 0; i < peaks.size(); ++i) pp[i] = tp[peaks[i]];
    const std::vector<int64_t> order = fh::numba_argsort(pp.data(), static_cast<int64_t>(pp.size()));
    std::vector<int64_t> sel;
    std::vector<double> sel_births;
    for (size_t r = order.size(); r-- > 0;) {
        if (static_cast<int64_t>(sel.size()) >= max_layers) break;
        const int64_t peak = peaks[order[r]];
        const float birth = sizes[peak];
        const std::vector<int64_t> a = active_at<float>(b, birth);
        bool diverse = true;
        for (const double sb : sel_births) {
            if (jaccard(a, active_at<double>(b, sb)) > threshold) {
                diverse = false;
                break;
            }
        }
        if (diverse) {
            sel.push_back(peak);
            sel_births.push_back(static_cast<double>(birth));
        }
    }
    return sel;
}

// _binary_search_for_n_clusters (layer_clusters.py:71-153), hdbscan
// lambdas; round() is round-half-even as in Python and numba.
void binary_search_for_n_clusters(Hierarchy &h, int64_t n, int64_t approx, int64_t *labels, float *prob) {
    int64_t lo = 2, hi = n / 2;
    Condensed t;
    auto count = [&](int64_t mcs) {
        condense(n, h, mcs, t);
        return static_cast<int64_t>(extract_leaves(t).size());
    };
    int64_t hi_n = count(hi);
    int64_t lo_n = count(lo);
    while (hi - lo > 1) {
        const int64_t mid = static_cast<int64_t>(std::nearbyint((lo + hi) / 2.0));
        if (mid == lo || mid == hi) break;
        const int64_t mid_n = count(mid);
        if (mid_n < approx) {
            hi = mid;
            hi_n = mid_n;
        } else {
            lo = mid;
            lo_n = mid_n;
        }
    }
    const int64_t dl = std::llabs(lo_n - approx), du = std::llabs(hi_n - approx);
    if (dl != du) {
        condense(n, h, dl < du ? lo : hi, t);
        label_vectors(t, extract_leaves(t), 0.0, n, labels, prob);
        return;
    }
    std::vector<int64_t> lab_lo(n), lab_hi(n);
    std::vector<float> pr_lo(n), pr_hi(n);
    Condensed tl;
    condense(n, h, lo, tl);
    label_vectors(tl, extract_leaves(tl), 0.0, n, lab_lo.data(), pr_lo.data());
    condense(n, h, hi, t);
    label_vectors(t, extract_leaves(t), 0.0, n, lab_hi.data(), pr_hi.data());
    int64_t cl = 0, cu = 0;
    for (int64_t i = 0; i < n; ++i) {
        cl += lab_lo[i] >= 0;
        cu += lab_hi[i] >= 0;
    }
    const bool use_lo = cl > cu;
    std::memcpy(labels, (use_lo ? lab_lo : lab_hi).data(), sizeof(int64_t) * n);
    std::memcpy(prob, (use_lo ? pr_lo : pr_hi).data(), sizeof(float) * n);
}

}  // namespace

extern "C" int fh_labels_from_sorted_mst(int64_t n, const int32_t *src, const int32_t *dst, const float *w_sq,
                                         const double *dist, const float *sample_weights,
                                         const fh_label_params *params, int64_t *labels, float *probabilities,
                                         double *linkage, fh_condensed_tree *condensed, double *phase_ms,
                                         const fh_label_order *order) {
    using clk = std::chrono::steady_clock;
    auto ms_since = [](clk::time_point a) {
        return std::chrono::duration<double, std::milli>(clk::now() - a).count();
    };
    if (n < 1 || !params || !labels || !probabilities) return FH_EINVAL;
    if (n >= (int64_t{1} << 30)) return FH_EUNSUPPORTED;
    if (n > 1 && (!src || !dst || (!w_sq && !dist))) return FH_EINVAL;
    if (params->min_cluster_size < 1) return FH_EINVAL;
    if (params->selection_method != 0 && params->selection_method != 1) return FH_EINVAL;
    try {
        if (n < 2) {
            labels[0] = -1;
            probabilities[0] = 0.0f;
            if (condensed) condensed->rows = 0;
            return FH_OK;
        }
        if (order && order->prefix_rows > 0) {
            // the groups must be non-empty and cover rows [0, k), and
            // row_order must be a permutation of [0, k): in range, each row
            // once (a repeated row would leave another row's entry unset)
            const int64_t k = order->prefix_rows;
            if (k > n - 1 || order->n_groups < 1 || !order->row_order || !order->group_sizes) return FH_EINVAL;
            int64_t tot = 0;
            for (int64_t g = 0; g < order->n_groups; ++g) {
                if (order->group_sizes[g] < 1) return FH_EINVAL;
                tot += order->group_sizes[g];
            }
            if (tot != k) return FH_EINVAL;
            // (parallel pieces; the second sight of a row reads its byte as 1)
            std::vector<uint8_t> seen(static_cast<size_t>(k), 0);
            const int nt = k > (1 << 18) ? fh::HostPool::width() : 1;
            std::vector<uint32_t> bad(nt, 0);
            fh::HostPool::run(nt, [&](int t) {
                uint32_t any = 0;
                for (int64_t j = k * t / nt, e = k * (t + 1) / nt; j < e; ++j) {
                    const uint32_t r = static_cast<uint32_t>(order->row_order[j]);
                    if (r >= static_cast<uint32_t>(k)) {
                        any = 1;
                        continue;
                    }
                    any |= __atomic_exchange_n(&seen[r], static_cast<uint8_t>(1), __ATOMIC_RELAXED);
                }
                bad[t] = any;
            });
            for (const uint32_t b : bad)
                if (b) return FH_EINVAL;
        }
        // reused between calls on this thread (no fresh pages per call)
        static thread_local Hierarchy h;
        clk::time_point t = clk::now();
        const int lc = linkage_tree(n, src, dst, dist ? nullptr : w_sq, dist, h, sample_weights, order);
        if (lc != FH_OK) return lc;
        if (phase_ms) phase_ms[0] = ms_since(t);
        if (linkage) {
            for (int64_t i = 0; i < n - 1; ++i) {
                const Merge &r = h.rows[i];
                linkage[4 * i + 0] = static_cast<double>(r.left);
                linkage[4 * i + 1] = static_cast<double>(r.right);
                linkage[4 * i + 2] = h.delta(i);
                linkage[4 * i + 3] = h.weighted ? static_cast<double>(h.lw[i] + h.rw[i])
                                                : static_cast<double>(r.left_size) + static_cast<double>(r.right_size);
            }
        }
        static thread_local Condensed c;
        t = clk::now();
        condense(n, h, params->min_cluster_size, c, kLambdaHdbscan, params->selection_method == 0);
        if (phase_ms) phase_ms[1] = ms_since(t);
        t = clk::now();
        ClusterTree ct;
        cluster_tree(c, ct);
        std::vector<int64_t> sel;
        if (params->selection_method == 0) {
            sel = eom(c, ct, n, params->max_cluster_size, params->allow_single_cluster != 0);
        } else {
            sel = leaves(ct, n);
        }
        if (sel.size() > 1 && params->cluster_selection_epsilon > 0.0)
            sel = epsilon_search(sel, ct, params->cluster_selection_epsilon);
        if (phase_ms) phase_ms[2] = ms_since(t);
        t = clk::now();
        label_vectors(c, sel, params->cluster_selection_epsilon, n, labels, probabilities);
        if (phase_ms) phase_ms[3] = ms_since(t);
        if (condensed) {
            const int64_t rows = c.rows;
            condensed->rows = rows;
            if (rows <= condensed->capacity && condensed->parent && condensed->child && condensed->lambda_val &&
                condensed->child_size) {
                for (int64_t r = 0; r < rows; ++r) {
                    condensed->parent[r] = c.parent[r];
                    condensed->child[r] = c.child[r];
                }
                std::memcpy(condensed->lambda_val, c.lambda.data(), sizeof(float) * rows);
                std::memcpy(condensed->child_size, c.size.data(), sizeof(float) * rows);
            }
        }
    } catch (const std::bad_alloc &) {
        return FH_ENOMEM;
    }
    return FH_OK;
}

// layer_clusters.build_cluster_layers (layer_clusters.py:386-481) after the
// MST: linkage tree of the normalised, sorted edges, base layer, barcode,
// persistence peaks and one layer per selected peak.  Layers come out
// unsorted (base layer first, then peaks in selection order).
extern "C" int fh_cluster_layers(int64_t n, const int32_t *src, const int32_t *dst, const double *dist,
                                 const fh_layer_params *p, fh_layers *out) {
    if (n < 1 || !p || !out || !out->labels || !out->strengths || !out->persistence) return FH_EINVAL;
    if (n >= (int64_t{1} << 30)) return FH_EUNSUPPORTED;
    if (n > 1 && (!src || !dst || !dist)) return FH_EINVAL;
    if (p->base_min_cluster_size < 1 || p->max_layers < 1 || p->base_n_clusters < 0) return FH_EINVAL;
    try {
        out->n_layers = 1;
        out->n_sizes = 0;
        out->min_cluster_size = p->base_min_cluster_size;
        out->persistence[0] = 0.0f;
        if (n < 2) {
            out->labels[0] = -1;
            out->strengths[0] = 0.0f;
            return FH_OK;
        }
        Hierarchy h;
        const int lc = linkage_tree(n, src, dst, nullptr, dist, h);
        if (lc != FH_OK) return lc;
        int64_t mcs = p->base_min_cluster_size;
        Condensed c;
        if (p->base_n_clusters > 0) {
            binary_search_for_n_clusters(h, n, p->base_n_clusters, out->labels, out->strengths);
            // min(bincount(labels[labels >= 0])), zero counts included
            int64_t top = -1;
            for (int64_t i = 0; i < n; ++i) top = std::max(top, out->labels[i]);
            if (top >= 0) {
                std::vector<int64_t> cnt(static_cast<size_t>(top + 1), 0);
                for (int64_t i = 0; i < n; ++i)
                    if (out->labels[i] >= 0) ++cnt[out->labels[i]];
                mcs = std::max<int64_t>(1, *std::min_element(cnt.begin(), cnt.end()));
            }
            condense(n, h, mcs, c, kLambdaPlscan);
        } else {
            condense(n, h, mcs, c, kLambdaPlscan);
            label_vectors(c, extract_leaves(c), 0.0, n, out->labels, out->strengths);
        }
        out->min_cluster_size = mcs;
        // rows whose child is a cluster (mask_condensed_tree, child >= n)
        std::vector<int64_t> cp, cc;
        std::vector<float> cl, cs;
        for (int64_t r = 0; r < c.rows; ++r) {
            if (c.child[r] >= n) {
                cp.push_back(c.parent[r]);
                cc.push_back(c.child[r]);
                cl.push_back(c.lambda[r]);
                cs.push_back(c.size[r]);
            }
        }
        if (cc.empty()) return FH_OK;
        const Barcode b = barcode(cp, cc, cl, cs, n, mcs);
        std::vector<float> sizes, tp;
        total_persistence(b, sizes, tp);
        out->n_sizes = static_cast<int64_t>(sizes.size());
        if (out->sizes && out->total_persistence && out->n_sizes <= out->capacity) {
            std::memcpy(out->sizes, sizes.data(), sizeof(float) * sizes.size());
            std::memcpy(out->total_persistence, tp.data(), sizeof(float) * tp.size());
        }
        const std::vector<int64_t> peaks = find_peaks(tp);
        const std::vector<int64_t> sel =
            select_diverse_peaks(peaks, tp, sizes, b, p->layer_similarity_threshold, p->max_layers - 1);
        const int64_t n_samples = c.parent[0];
        for (size_t k = 0; k < sel.size(); ++k) {
            const float best = sizes[sel[k]];
            std::vector<int64_t> ids;
            for (size_t i = 0; i < b.births.size(); ++i)
                if (b.births[i] <= best && b.deaths[i] > best) ids.push_back(static_cast<int64_t>(i) + n);
            const int64_t layer = static_cast<int64_t>(k) + 1;
            label_vectors(c, ids, 0.0, n_samples, out->labels + layer * n, out->strengths + layer * n);
            out->persistence[layer] = tp[sel[k]];
        }
        out->n_layers = static_cast<int32_t>(1 + sel.size());
    } catch (const std::bad_alloc &) {
        return FH_ENOMEM;
    }
    return FH_OK;
}

// layer_clusters._build_cluster_tree (layer_clusters.py:498-525): the
// (parent layer, parent cluster, child layer, child cluster) rows in the
// reference's order; the points of each upper layer are grouped by numba's
// argsort of its labels.
extern "C" int64_t fh_layer_cluster_tree(int32_t n_layers, int64_t n, const int64_t *labels, int64_t *mapping,
                                         int64_t capacity) {
    if (n_layers < 1 || n < 1 || !labels) return FH_EINVAL;
    try {
        std::vector<std::vector<uint8_t>> found(n_layers);
        std::vector<int64_t> top(n_layers, -1);
        for (int32_t l = 0; l < n_layers; ++l) {
            for (int64_t i = 0; i < n; ++i) top[l] = std::max(top[l], labels[l * n + i]);
            found[l].assign(static_cast<size_t>(top[l] + 1), 0);
        }
        int64_t count = 0;
        auto emit = [&](int64_t a, int64_t b, int64_t c, int64_t d) {
            if (mapping && count < capacity) {
                mapping[4 * count + 0] = a;
                mapping[4 * count + 1] = b;
                mapping[4 * count + 2] = c;
                mapping[4 * count + 3] = d;
            }
            ++count;
        };
        for (int32_t up = 1; up < n_layers; ++up) {
            const int64_t *U = labels + up * n;
            std::vector<int64_t> uniq(U, U + n);
            std::sort(uniq.begin(), uniq.end());
            uniq.erase(std::unique(uniq.begin(), uniq.end()), uniq.end());
            const std::vector<int64_t> order = fh::numba_argsort(U, n);
            // np.split at cumsum(bincount(U + 1))[:-1]: group g holds label g - 1
            const int64_t ng = top[up] + 2;
            std::vector<int64_t> start(static_cast<size_t>(ng + 1), 0);
            for (int64_t i = 0; i < n; ++i) ++start[U[i] + 2];
            for (int64_t g = 0; g < ng; ++g) start[g + 1] += start[g];
            const int64_t drop = ng > static_cast<int64_t>(uniq.size()) ? 1 : 0;
            for (int32_t lo = up - 1; lo >= 0; --lo) {
                const int64_t *Lw = labels + lo * n;
                for (size_t i = 0; i < uniq.size(); ++i) {
                    const int64_t label = uniq[i];
                    const int64_t g = static_cast<int64_t>(i) + drop;
                    if (label < 0 || g >= ng) continue;
                    for (int64_t r = start[g]; r < start[g + 1]; ++r) {
                        const int64_t child = Lw[order[r]];
                        if (child >= 0 && !found[lo][child]) {
                            emit(up, label, lo, child);
                            found[lo][child] = 1;
                        }
                    }
                }
            }
        }
        for (int32_t lo = n_layers - 1; lo >= 0; --lo)
            for (int64_t child = 0; child <= top[lo]; ++child)
                if (!found[lo][child]) emit(n_layers, 0, lo, child);
        return count;
    } catch (const std::bad_alloc &) {
        return FH_ENOMEM;
    }
}
