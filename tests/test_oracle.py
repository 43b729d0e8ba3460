"""Pins for the CPU oracle (oracle/fh_oracle.c) -- CPU only.

Each check is chosen so that a plausible restatement mistake fails it:
closed forms, an independent library (sklearn KDTree, scipy MST), brute
force, a hand-derived fixture, and an unpruned re-derivation of every
Boruvka round.
"""
import numpy as np
import pytest

from oracle import oracle as O
from fast_hdbscan_amd import workloads


def _blobs(n, d=2, seed=0, centers=5):
    from sklearn.datasets import make_blobs

    return make_blobs(n_samples=n, n_features=d, centers=centers, random_state=seed)[0].astype(np.float32)


# ----------------------------------------------------------------- tree ----

@pytest.mark.parametrize("n", [1, 2, 40, 41, 42, 80, 81, 82, 161, 1000, 12345, 2_000_000, 10_000_000])
def test_tree_levels_closed_form(n):
    # numba_kdtree.py:489-490 evaluated in float64 as the reference does
    expect = int(np.log2(max(1, (n - 1) / 40)) + 1)
    assert O.tree_nlevels(n) == expect
    assert O.tree_nnodes(n) == 2 ** expect - 1


def _sk_tree(X):
    from sklearn.neighbors import KDTree

    t = KDTree(X.astype(np.float64), leaf_size=40)
    data, idx, nd, nb = t.get_arrays()
    return idx, nd, nb


@pytest.mark.parametrize("case", ["blobs2", "blobs3", "d1", "d5", "dups", "grid"])
def test_tree_matches_sklearn_structure(case):
    rng = np.random.default_rng(0)
    if case == "blobs2":
        X = _blobs(5000)
    elif case == "blobs3":
        X = _blobs(7000, d=3, seed=1)
    elif case == "d1":
        X = rng.standard_normal((3000, 1)).astype(np.float32)
    elif case == "d5":
        X = _blobs(4000, d=5, seed=2)
    elif case == "dups":
        X = _blobs(3000, seed=3)
        X[1500:] = X[:1500]
    else:
        X = np.stack(np.meshgrid(np.arange(50), np.arange(40)), -1).reshape(-1, 2).astype(np.float32)
        X = X[rng.permutation(X.shape[0])]
    T = O.build_kdtree(X)
    idx, nd, nb = _sk_tree(X)
    np.testing.assert_array_equal(T.idx_start, nd["idx_start"])
    np.testing.assert_array_equal(T.idx_end, nd["idx_end"])
    np.testing.assert_array_equal(T.is_leaf, nd["is_leaf"].astype(bool))
    np.testing.assert_array_equal(T.node_bounds, nb.astype(np.float32))
    for i in np.nonzero(T.is_leaf)[0]:
        s, e = T.idx_start[i], T.idx_end[i]
        assert set(T.idx_array[s:e]) == set(idx[s:e])


def test_tree_median_invariant():
    X = _blobs(20000, d=3, seed=4)
    T = O.build_kdtree(X)
    for node in range(T.n_nodes):
        if T.is_leaf[node]:
            continue
        s, e = T.idx_start[node], T.idx_end[node]
        mid = s + (e - s) // 2
        lo, hi = T.node_bounds[0, node], T.node_bounds[1, node]
        spread = hi - lo
        axis = int(np.argmax(spread))  # first maximum, like the strict '>' scan
        keys = [(X[i, axis], i) for i in T.idx_array[s:e]]
        left, right = keys[: mid - s], keys[mid - s:]
        # the split leaves the n//2 smallest (value, index) keys on the left;
        # positions inside each side are rearranged again by the children
        assert max(left) < min(right)


def test_introselect_partition_then_insertion_by_hand():
    # 20 points in 1-D, value(i) = 7*i mod 20, leaf_size 9: n_levels =
    # int(log2(19/9) + 1) = 2 (numba_kdtree.py:489), so only the root is
    # split, by _introselect(left=0, right=20, nth=10) (numba_kdtree.py:
    # 393-407, 449): size 20 > 16, max_depth = 2*int(log2 20) = 8.
    # Trace of _introselect_impl (366-390), idx_array starting as 0..19:
    #  _median_of_three_pivot (280-299): mid = 19//2 = 9; idx 0 (value 0),
    #    idx 9 (value 3), idx 19 (value 13) are already ordered -> pivot
    #    at position 9 (value 3).
    #  _partition_indices (313-350): swap positions 9 and 19, pivot idx 9;
    #    i=0, j=18.  i stops at 1 (value 7), j walks down to 6 (value 2):
    #    swap -> positions 1,6 hold idx 6,1; i=2, j=5.  i stops at 2
    #    (value 14), j at 3 (value 1): swap -> positions 2,3 hold idx 3,2;
    #    i=3, j=2: break; the pivot goes to position 3 (idx 2 to the end).
    #    Returns 3: idx_array = [0,6,3,9,4,5,1,7,8,19,10,...,18,2].
    #  nth=10 > 3 -> left = 4; right-left = 16 is not > 16, so positions
    #    4..19 are insertion-sorted by value: idx 12,15,18,1,4,7,10,13,16,
    #    19,2,5,8,11,14,17 (values 4..19).
    # The leaves take positions 0..9 and 10..19 without reordering, so the
    # left leaf keeps the partition's order 0,6,3,9 (values 0,2,1,3).
    X = ((7 * np.arange(20)) % 20).astype(np.float32).reshape(-1, 1)
    T = O.build_kdtree(X, leaf_size=9)
    assert T.n_nodes == 3
    assert list(T.idx_array) == [0, 6, 3, 9, 12, 15, 18, 1, 4, 7, 10, 13, 16, 19, 2, 5, 8, 11, 14, 17]
    assert list(T.idx_start) == [0, 0, 10] and list(T.idx_end) == [20, 10, 20]


def test_introselect_pivot_at_nth_exits_by_hand():
    # 21 points in 1-D, leaf_size 10: n_levels = int(log2(20/10) + 1) = 2,
    # root split by _introselect(0, 21, nth=10).  Values by index:
    V = [20, 11, 1, 12, 2, 13, 3, 4, 14, 5, 0, 6, 15, 7, 16, 17, 8, 18, 9, 19, 10]
    # _median_of_three_pivot: mid = 20//2 = 10; candidates idx 0 (20),
    #   idx 10 (0), idx 20 (10).  20 > 0: swap positions 0,10; then
    #   20 > 10: swap positions 10,20; then 0 < 10: no swap.  Pivot idx 20
    #   (value 10, rank 10) at position 10.
    # _partition_indices: positions 10 and 20 swap (idx 0, value 20, goes
    #   to position 10); i/j scans then swap the pairs of positions
    #   (1,18) (3,16) (5,13) (8,11): idx 18,16,13,11 come left and idx
    #   1,3,5,8 go right.  Then i stops at 10 (value 20), j falls to 9,
    #   break, and the pivot returns to position 10 (idx 0 to position 20).
    # pivot_pos == nth: _introselect_impl returns at once (385-387), with
    # no insertion sort, so both halves keep the partition's order.
    X = np.asarray(V, np.float32).reshape(-1, 1)
    T = O.build_kdtree(X, leaf_size=10)
    assert T.n_nodes == 3
    assert list(T.idx_array) == [10, 18, 2, 16, 4, 13, 6, 7, 11, 9, 20, 8, 12, 5, 14, 15, 3, 17, 1, 19, 0]
    assert list(T.idx_start) == [0, 0, 10] and list(T.idx_end) == [21, 10, 21]


def test_tree_permutation():
    X = _blobs(10000)
    T = O.build_kdtree(X)
    assert np.array_equal(np.sort(T.idx_array), np.arange(10000))


# ------------------------------------------------------------------ kNN ----

@pytest.mark.parametrize("d,k", [(2, 11), (3, 16), (5, 6), (8, 3)])
def test_knn_matches_brute_force(d, k):
    X = _blobs(2500, d=d, seed=d)
    T = O.build_kdtree(X)
    td, ti = O.tree_query(T, X, k)
    bd, bi = O.knn_brute(X, k)
    np.testing.assert_array_equal(td, bd)
    # indices agree wherever the row has no exact distance tie
    ties = np.zeros(X.shape[0], bool)
    for r in range(X.shape[0]):
        ties[r] = len(np.unique(bd[r])) < k
    np.testing.assert_array_equal(ti[~ties], bi[~ties])


def test_core_distance_vs_float64_kdtree():
    from sklearn.neighbors import KDTree

    X = workloads.blobs_small(10_000, seed=0)
    T = O.build_kdtree(X)
    td, _ = O.tree_query(T, X, 11)
    ref, _ = KDTree(X.astype(np.float64)).query(X.astype(np.float64), k=11)
    np.testing.assert_allclose(np.sqrt(td[:, -1].astype(np.float64)), ref[:, -1], rtol=1e-5, atol=1e-6)


def test_rdist_fma_form():
    # The canonical float32 form: r = d0*d0 rounded, then fma(d_i, d_i, r).
    rng = np.random.default_rng(5)
    for _ in range(200):
        x = rng.standard_normal(2).astype(np.float32)
        y = rng.standard_normal(2).astype(np.float32)
        d0 = np.float32(x[0] - y[0])
        d1 = np.float32(x[1] - y[1])
        # fma in exact arithmetic: float64 holds d1*d1 + r exactly enough for
        # a single final rounding whenever the exponents are close; use
        # fractions for an exact reference instead.
        from fractions import Fraction

        r0 = np.float32(d0 * d0)
        exact = Fraction(float(d1)) * Fraction(float(d1)) + Fraction(float(r0))
        want = np.float32(float(exact)) if True else None
        got = O.lib().orc_rdist(x, y, 2)
        # float(exact) rounds to float64 first; confirm no double-rounding
        # ambiguity by checking the neighbours of `want`.
        cands = [np.nextafter(want, np.float32(-np.inf)), want, np.nextafter(want, np.float32(np.inf))]
        best = min(cands, key=lambda c: abs(Fraction(float(c)) - exact))
        assert np.float32(got) == best


def test_duplicate_points_heap_layout_by_hand():
    # test_edge_cases.py:20-24 data.  One leaf (n_levels = 1) holding points
    # 0..49; every distance is 0.  Tracing simple_heap_push
    # (numba_kdtree.py:593-635) for pushes 0..5 leaves indices
    # [5, 2, 4, 0, 1, 3] in heap order (all later pushes are rejected since
    # 0 >= 0); deheap_sort (668-677) turns that into [2, 4, 0, 1, 3, 5].
    # initialize_boruvka_from_knn then links every i to column 1 = point 4;
    # (4, 4) is a self-loop and is skipped.
    X = np.full((50, 3), 1.0, np.float32)
    r = O.boruvka(O.build_kdtree(X), 5)
    for row in r.knn_i:
        assert list(row) == [2, 4, 0, 1, 3, 5]
    assert np.all(r.core_sq == 0)
    expect_src = [i for i in range(50) if i != 4]
    assert list(r.src) == expect_src
    assert list(r.dst) == [4] * 49
    assert np.all(r.w_sq == 0)
    assert r.rounds == 0


def test_heap_partial_ties_by_hand():
    # One leaf (8 points < 41, n_levels = 1) scanned in index order; the
    # query at 0 gives rdist 4,1,4,9,1,0,1,4 for points 0..7.  Tracing
    # simple_heap_push (numba_kdtree.py:593-635) with k = 5 (priorities /
    # indices in heap order):
    #   push 0 (4): sinks 0 -> 1 -> 3          [inf,inf,inf,4,inf] [-,-,-,0,-]
    #   push 1 (1): 0 -> 1 -> 4 (p[3]=4 < inf)  [inf,inf,inf,4,1]   [-,-,-,0,1]
    #   push 2 (4): 0 -> 1, stops (4 < 4 false) [inf,4,inf,4,1]     [-,2,-,0,1]
    #   push 3 (9): 0 -> 2 (p[1]=4 < inf)       [inf,4,9,4,1]       [-,2,3,0,1]
    #   push 4 (1): 0 -> 2 (4 < 9)              [9,4,1,4,1]         [3,2,4,0,1]
    #   push 5 (0): 0 -> 1 -> 3                 [4,4,1,0,1]         [2,0,4,5,1]
    #   push 6 (1): 0 -> 1, stops (1 < 1 false) [4,1,1,0,1]         [0,6,4,5,1]
    #     -- of the two 4s, point 2 leaves and point 0 stays
    #   push 7 (4): rejected, 4 >= p[0] = 4.
    # deheap_sort (667-677) with siftdown's strict '<' (647-664):
    #   j=4: swap -> [1,1,1,0,4] [1,6,4,5,0], no sift (ties)
    #   j=3: swap -> [0,1,1,1,4] [5,6,4,1,0], sift 0<->1 -> [6,5,4,1,0]
    #   j=2: swap -> [1,0,1,...] [4,5,6,1,0], no sift
    #   j=1: swap -> [0,1,1,1,4] [5,4,6,1,0]
    # (a stable sort of the kept points would give 5,1,4,6,0).
    X = np.asarray([2, 1, -2, 3, -1, 0, 1, -2], np.float32).reshape(-1, 1)
    Q = np.zeros((1, 1), np.float32)
    d, i = O.tree_query(O.build_kdtree(X), Q, 5)
    assert list(i[0]) == [5, 4, 6, 1, 0]
    assert list(d[0]) == [0, 1, 1, 1, 4]


def test_sqrt_mode_rows():
    # output_rdist=False rows are the square roots of the rdist rows (same
    # multiset per row); the index sets agree wherever the k-th squared
    # distance is below the (k+1)-th (numba_kdtree.py:789-792).
    X = _blobs(3000, seed=8)
    T = O.build_kdtree(X)
    d1, i1 = O.tree_query(T, X, 9)
    d2, i2 = O.tree_query_sqrt(T, X, 9)
    np.testing.assert_array_equal(np.sqrt(d1), d2)
    d10, _ = O.tree_query(T, X, 10)
    clean = d10[:, 8] < d10[:, 9]
    assert clean.mean() > 0.9
    np.testing.assert_array_equal(np.sort(i1[clean], 1), np.sort(i2[clean], 1))


# -------------------------------------------------------------- Boruvka ----

def _mrd_matrix(X, core_sq):
    n = X.shape[0]
    bd, bi = O.knn_brute(X, n)
    D = np.empty((n, n), np.float32)
    rows = np.repeat(np.arange(n), n)
    D[rows, bi.ravel()] = bd.ravel()
    return np.maximum(D, np.maximum(core_sq[:, None], core_sq[None, :]))


@pytest.mark.parametrize("n,ms,seed", [(600, 5, 0), (900, 10, 1), (500, 1, 2), (700, 3, 3)])
def test_mst_weights_match_scipy(n, ms, seed):
    from scipy.sparse.csgraph import minimum_spanning_tree

    X = _blobs(n, seed=seed)
    r = O.boruvka(O.build_kdtree(X), ms)
    M = _mrd_matrix(X, r.core_sq).astype(np.float64)
    # scipy drops explicit zeros; shift every weight by one
    G = M + 1.0
    np.fill_diagonal(G, 0.0)
    ref = np.sort(minimum_spanning_tree(G).tocoo().data - 1.0)
    assert r.src.shape[0] == n - 1
    np.testing.assert_array_equal(np.sort(r.w_sq.astype(np.float64)), ref)


def _spanning(n, src, dst):
    parent = list(range(n))

    def find(x):
        while parent[x] != x:
            x = parent[x]
        return x

    for a, b in zip(src, dst):
        ra, rb = find(int(a)), find(int(b))
        assert ra != rb, "cycle in MST"
        parent[ra] = rb
    return len({find(i) for i in range(n)}) == 1


def test_mst_is_spanning_tree():
    X = workloads.scale_2d(30_000, seed=3)
    r = O.boruvka(O.build_kdtree(X), 10)
    assert r.src.shape[0] == X.shape[0] - 1
    assert _spanning(X.shape[0], r.src, r.dst)
    # each edge weight is the mutual reachability of its endpoints
    d = np.array([O.lib().orc_rdist(X[a], X[b], 2) for a, b in zip(r.src[:2000], r.dst[:2000])], np.float32)
    w = np.maximum(d, np.maximum(r.core_sq[r.src[:2000]], r.core_sq[r.dst[:2000]]))
    np.testing.assert_array_equal(w, r.w_sq[:2000])


def _dfs_order(T, X, i):
    """All leaf points in the order a full (unpruned) walk from point i visits
    them: closer child first by box lower bound, left on ties."""
    order = []
    lib = O.lib()

    def lb(node):
        return lib.orc_box_lb(np.ascontiguousarray(T.node_bounds[0, node]),
                              np.ascontiguousarray(T.node_bounds[1, node]), X[i], X.shape[1])

    def rec(node):
        if T.is_leaf[node]:
            order.extend(T.idx_array[T.idx_start[node]:T.idx_end[node]].tolist())
            return
        l, r = 2 * node + 1, 2 * node + 2
        if lb(l) <= lb(r):
            rec(l)
            rec(r)
        else:
            rec(r)
            rec(l)

    rec(0)
    return order


@pytest.mark.parametrize("case", ["blobs", "grid"])
def test_boruvka_rounds_rederived_without_pruning(case):
    """Re-derive every Boruvka round from its definition (lowest point index
    at the component minimum; its first partner in unpruned walk order; edges
    in order of the component's lowest point; the last-processed edge of
    each cycle skipped) and compare with the pruned restatement."""
    if case == "blobs":
        X = _blobs(300, seed=11)
        ms = 5
    else:
        X = np.stack(np.meshgrid(np.arange(15), np.arange(14)), -1).reshape(-1, 2).astype(np.float32)
        X = X[np.random.default_rng(2).permutation(X.shape[0])]
        ms = 3
    n = X.shape[0]
    T = O.build_kdtree(X)
    r = O.boruvka(T, ms)
    core = r.core_sq
    bd, bi = O.knn_brute(X, n)
    D = np.empty((n, n), np.float32)
    D[np.repeat(np.arange(n), n), bi.ravel()] = bd.ravel()
    M = np.maximum(D, np.maximum(core[:, None], core[None, :]))

    # initial forest from the oracle's own kNN rows (E2), processed in order
    parent = list(range(n))

    def find(x):
        while parent[x] != x:
            x = parent[x]
        return x

    rows = []
    for i in range(n):
        for j in range(1, r.knn_i.shape[1]):
            kk = int(r.knn_i[i, j])
            if core[i] >= core[kk]:
                a, b = find(i), find(kk)
                if a != b:
                    rows.append((i, kk, max(core[i], r.knn_d[i, j])))
                    parent[a] = b
                break
    n_init = len(rows)
    assert list(zip(r.src[:n_init], r.dst[:n_init])) == [(a, b) for a, b, _ in rows]

    orders = {}
    while len({find(i) for i in range(n)}) > 1:
        comp = np.array([find(i) for i in range(n)])
        chosen = {}
        for c in sorted(set(comp.tolist()), key=lambda c: np.min(np.nonzero(comp == c)[0])):
            members = np.nonzero(comp == c)[0]
            others = comp != c
            best = min(M[i][others].min() for i in members)
            i_star = min(i for i in members if M[i][others].min() == best)
            if i_star not in orders:
                orders[i_star] = _dfs_order(T, X, i_star)
            j_star = next(j for j in orders[i_star] if comp[j] != c and M[i_star, j] == best)
            chosen[c] = (i_star, j_star, best)
        for c, (a, b, w) in chosen.items():
            fa, fb = find(a), find(b)
            if fa != fb:
                rows.append((a, b, w))
                parent[fa] = fb
    assert len(rows) == n - 1
    np.testing.assert_array_equal(r.src, [a for a, _, _ in rows])
    np.testing.assert_array_equal(r.dst, [b for _, b, _ in rows])
    np.testing.assert_array_equal(r.w_sq, np.array([w for _, _, w in rows], np.float32))


# ------------------------------------------------- sample-weight branch ----

def test_weighted_branch_unit_weights_core():
    # boruvka.py:564-578 with unit weights: k = int(2 * ms / 1.0); the core is
    # the square root of the ms-th smallest rdist (cumulative weight reaches
    # ms at column ms-1, self included).
    X = _blobs(3000, seed=8)
    ms = 7
    T = O.build_kdtree(X)
    r = O.boruvka_weighted(T, ms, np.ones(3000, np.float32))
    assert r.knn_i.shape[1] == 2 * ms
    rd, _ = O.tree_query(T, X, 2 * ms)
    np.testing.assert_array_equal(r.core_sq, np.sqrt(rd[:, ms - 1]))
    assert r.src.shape[0] == X.shape[0] - 1
    assert _spanning(X.shape[0], r.src, r.dst)


def test_weighted_k_float32_mean():
    # numba's np.mean on float32 accumulates in float32 in index order
    w = np.full(1_000_001, 0.1, np.float32)
    acc = np.float32(0.0)
    for v in w[:1000]:
        acc = np.float32(acc + v)
    k = O.lib().orc_weighted_k(w[:1000], 1000, 5)
    assert k == int(2.0 * (5 / (float(acc) / 1000)))


def test_weighted_branch_label_weights_spanning():
    # test_hdbscan.py:113-117 uses the blob labels (0, 1, 2) as weights
    from sklearn.datasets import make_blobs
    from sklearn.preprocessing import StandardScaler
    from sklearn.utils import shuffle

    X, y = make_blobs(n_samples=200, random_state=10)
    X, y = shuffle(X, y, random_state=7)
    X = StandardScaler().fit_transform(X)
    r = O.boruvka_weighted(O.build_kdtree(X), 10, y.astype(np.float32))
    assert r.src.shape[0] == 199
    assert _spanning(200, r.src, r.dst)


def test_racy_boruvka_keeps_the_weights():
    # the multi-threaded timing variant may pick other edges among exact
    # ties, but core distances and the MST weight multiset are fixed
    X = workloads.blobs_small(20_000, seed=5)
    t = O.build_kdtree(X)
    a = O.boruvka(t, 10, nthreads=1)
    b = O.boruvka_racy(t, 10, nthreads=4)
    np.testing.assert_array_equal(a.core_sq, b.core_sq)
    np.testing.assert_array_equal(np.sort(a.w_sq), np.sort(b.w_sq))
    assert b.src.shape[0] == X.shape[0] - 1
