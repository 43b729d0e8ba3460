"""Pins for the label-stage oracle and the host label stage (CPU only).

Known answers are the reference's own tests; the host stage (C++ in
libfh.so, no device) must reproduce the oracle row for row.
"""
import sys

import numpy as np
import pytest

from oracle import labels as LB
from oracle import oracle as O
from fast_hdbscan_amd import workloads


def _ref_X():
    # test_hdbscan.py:38-40
    from sklearn.datasets import make_blobs
    from sklearn.preprocessing import StandardScaler
    from sklearn.utils import shuffle

    X, y = make_blobs(n_samples=200, random_state=10)
    X, y = shuffle(X, y, random_state=7)
    return StandardScaler().fit_transform(X)


def test_known_answer_uniform_epsilon():
    # test_hdbscan.py:169-186
    np.random.seed(0)
    X = np.random.rand(150, 2)
    edges, _, _ = O.compute_minimum_spanning_tree(X, 5)
    lab, *_ = LB.clusters_from_spanning_tree(edges, min_cluster_size=5)
    u, c = np.unique(lab, return_counts=True)
    assert len(u) == 8 and c[u == -1] == 69
    lab, *_ = LB.clusters_from_spanning_tree(edges, min_cluster_size=5, cluster_selection_epsilon=0.2)
    u, c = np.unique(lab, return_counts=True)
    assert len(u) == 2 and c[u == -1] == 2


def test_known_answer_three_blobs():
    # test_hdbscan.py:84-93 (defaults: min_cluster_size 10, min_samples 10)
    edges, _, _ = O.compute_minimum_spanning_tree(_ref_X(), 10)
    lab, *_ = LB.clusters_from_spanning_tree(edges, min_cluster_size=10)
    assert len(set(lab)) - int(-1 in lab) == 3


def test_known_answer_no_clusters():
    # test_hdbscan.py:103-110
    X = _ref_X()
    edges, _, _ = O.compute_minimum_spanning_tree(X, len(X) + 1)
    lab, *_ = LB.clusters_from_spanning_tree(edges, min_cluster_size=len(X) + 1)
    assert np.all(lab == -1)


def test_known_answer_max_cluster_size():
    # test_hdbscan.py:189-195 (HDBSCAN defaults: min_cluster_size 5)
    edges, _, _ = O.compute_minimum_spanning_tree(_ref_X(), 5)
    lab, *_ = LB.clusters_from_spanning_tree(edges, min_cluster_size=5, max_cluster_size=30)
    assert len(set(lab)) >= 3
    for label in set(lab) - {-1}:
        assert np.sum(lab == label) <= 30


def test_known_answer_issue_63():
    # test_hdbscan.py:742-762
    for seed in [700, 727]:
        rng = np.random.default_rng(seed)
        centers = rng.uniform(-10, 10, size=(3, 3))
        data = np.vstack([rng.normal(loc=c, scale=0.5, size=(200, 3)) for c in centers])
        edges, _, _ = O.compute_minimum_spanning_tree(data, 1)
        lab, *_ = LB.clusters_from_spanning_tree(edges, min_cluster_size=192, cluster_selection_epsilon=0.05957,
                                                 allow_single_cluster=True)
        assert len(set(lab) - {-1}) >= 1
        assert np.sum(lab == -1) / len(lab) < 0.9


def test_edge_cases_all_noise_when_mcs_exceeds_n():
    # test_edge_cases.py:26-31
    np.random.seed(42)
    X = np.random.rand(10, 2)
    edges, _, _ = O.compute_minimum_spanning_tree(X, 100)
    lab, *_ = LB.clusters_from_spanning_tree(edges, min_cluster_size=100)
    assert np.all(lab == -1)


def test_agrees_with_sklearn_hdbscan():
    # Independent implementation (sklearn computes in float64, the reference
    # path in float32, so a few boundary points may differ); sklearn's
    # min_samples counts the point itself.
    from sklearn.cluster import HDBSCAN as SK

    X = workloads.blobs_small(3000, seed=4).astype(np.float64)
    edges, _, _ = O.compute_minimum_spanning_tree(X, 10)
    lab, prob, *_ = LB.clusters_from_spanning_tree(edges, min_cluster_size=25)
    sk = SK(min_cluster_size=25, min_samples=11).fit(X)
    assert len(set(lab.tolist())) == len(set(sk.labels_.tolist()))
    # best one-to-one matching by overlap
    agree = 0
    for c in set(lab.tolist()):
        vals, cnt = np.unique(sk.labels_[lab == c], return_counts=True)
        agree += cnt.max()
    assert agree / len(lab) > 0.99
    assert np.median(np.abs(prob - sk.probabilities_)) < 1e-3


CASES = [
    ("uniform", 5, 5, "eom", 0.0, False),
    ("uniform", 5, 5, "eom", 0.2, False),
    ("uniform", 5, 5, "leaf", 0.0, False),
    ("uniform", 5, 5, "eom", 0.0, True),
    ("tiny_k_gt_n", 100, 100, "eom", 0.0, False),
    ("dups", 5, 5, "eom", 0.0, False),
    ("blobs", 10, 10, "eom", 0.0, False),
    ("blobs", 10, 10, "leaf", 0.0, False),
    ("blobs", 10, 10, "eom", 0.5, False),
    ("blobs", 10, 10, "eom", 0.0, True),
    ("aniso", 10, 100, "eom", 0.0, False),
    # small cluster sizes: many splits, points as split children (mcs 1)
    ("blobs", 10, 1, "eom", 0.0, False),
    ("blobs", 5, 2, "leaf", 0.0, False),
    ("uniform", 3, 2, "eom", 0.0, True),
]


def _data(name):
    if name == "uniform":
        np.random.seed(0)
        return np.random.rand(150, 2)
    if name == "tiny_k_gt_n":
        return np.random.RandomState(42).rand(10, 2)
    if name == "dups":
        return np.full((50, 3), 1.0)
    if name == "blobs":
        return workloads.blobs_small(10_000)
    return workloads.scale_2d(60_000, seed=3)


@pytest.mark.parametrize("case", CASES, ids=[f"{c[0]}-{c[3]}-{c[4]}-{c[5]}" for c in CASES])
def test_host_label_stage_matches_oracle(case):
    from fast_hdbscan_amd import labels as PL

    name, ms, mcs, meth, eps, single = case
    if mcs <= 2:
        # the oracle's eom_recursion follows the reference's recursion
        # (cluster_trees.py:689-722); small cluster sizes make it deep
        sys.setrecursionlimit(max(sys.getrecursionlimit(), 20000))
    edges, _, _ = O.compute_minimum_spanning_tree(_data(name), ms)
    kw = dict(min_cluster_size=mcs, cluster_selection_method=meth, cluster_selection_epsilon=eps,
              allow_single_cluster=single)
    a = PL.clusters_from_spanning_tree(edges, **kw)
    b = LB.clusters_from_spanning_tree(edges, **kw)
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1], b[1])
    np.testing.assert_array_equal(a[2], b[2])
    for f in ("parent", "child", "lambda_val", "child_size"):
        np.testing.assert_array_equal(getattr(a[3], f), getattr(b[3], f))
    np.testing.assert_array_equal(a[4], b[4])


# ------------------------------------------------------ sample weights ----

def _weighted_case(seed, n=1500, ms=6):
    from oracle import oracle as O
    from sklearn.datasets import make_blobs

    X, _ = make_blobs(n, centers=5, cluster_std=1.2, random_state=seed)
    w = np.random.default_rng(seed).uniform(0.2, 3.0, n).astype(np.float32)
    e, _, _ = O.compute_minimum_spanning_tree(X, ms, sample_weights=w)
    return e, w


@pytest.mark.parametrize("seed,mcs,method", [(0, 10, "eom"), (1, 25, "eom"), (2, 15, "leaf"), (3, 40, "eom")])
def test_weighted_label_stage_matches_oracle(seed, mcs, method):
    from fast_hdbscan_amd.labels import clusters_from_spanning_tree
    from oracle import labels as OL

    e, w = _weighted_case(seed)
    ref = OL.clusters_from_spanning_tree(e, min_cluster_size=mcs, cluster_selection_method=method, sample_weights=w)
    got = clusters_from_spanning_tree(e, min_cluster_size=mcs, cluster_selection_method=method, sample_weights=w)
    np.testing.assert_array_equal(got[0], ref[0])
    np.testing.assert_array_equal(got[1], ref[1])
    np.testing.assert_array_equal(got[2], ref[2])
    for a, b in zip(got[3], (ref[3].parent, ref[3].child, ref[3].lambda_val, ref[3].child_size)):
        np.testing.assert_array_equal(a, b)


def test_reference_sample_weight_known_answer():
    # test_hdbscan.py:113-117: the blob labels as weights -> 2 clusters
    from fast_hdbscan_amd.labels import clusters_from_spanning_tree
    from oracle import oracle as O
    from sklearn.datasets import make_blobs
    from sklearn.preprocessing import StandardScaler
    from sklearn.utils import shuffle

    X, y = make_blobs(n_samples=200, random_state=10)
    X, y = shuffle(X, y, random_state=7)
    X = StandardScaler().fit_transform(X)
    sw = y.astype(np.float32)
    e, _, _ = O.compute_minimum_spanning_tree(X, 10, sample_weights=sw)
    labels = clusters_from_spanning_tree(e, min_cluster_size=10, sample_weights=sw)[0]
    assert len(set(labels)) - int(-1 in labels) == 2


@pytest.mark.parametrize("case", ["plain", "weighted", "k_exceeds_n"])
def test_point_order_changes_nothing(case):
    # endpoints given as positions of a permutation (point_order[p] = id)
    # must give the same labels, strengths, linkage and condensed tree
    from fast_hdbscan_amd import labels as PL

    rng = np.random.default_rng(11)
    w = None
    if case == "weighted":
        e, w = _weighted_case(4, n=800)
    elif case == "k_exceeds_n":
        e, _, _ = O.compute_minimum_spanning_tree(rng.normal(size=(12, 2)), 15)  # -1 endpoints
    else:
        e, _, _ = O.compute_minimum_spanning_tree(workloads.blobs_small(3000, seed=2), 5)
    n = e.shape[0] + 1
    o = np.lexsort((e[:, 1], e[:, 0], e[:, 2]))
    src, dst = e[o, 0].astype(np.int32), e[o, 1].astype(np.int32)
    dist = e[o, 2].copy()
    order = rng.permutation(n).astype(np.int32)
    pos = np.empty(n, np.int32)
    pos[order] = np.arange(n, dtype=np.int32)
    ps = np.where(src < 0, src, pos[np.maximum(src, 0)]).astype(np.int32)
    pd = np.where(dst < 0, dst, pos[np.maximum(dst, 0)]).astype(np.int32)
    params = PL.label_params(5, "eom", np.inf, False, 0.0)
    a = PL.labels_from_sorted_arrays(n, src, dst, None, params, return_trees=True, dist=dist, sample_weights=w)
    b = PL.labels_from_sorted_arrays(n, ps, pd, None, params, return_trees=True, dist=dist, sample_weights=w,
                                     point_order=order)
    for x, y in zip(a[:3], b[:3]):
        np.testing.assert_array_equal(x, y)
    for f in ("parent", "child", "lambda_val", "child_size"):
        np.testing.assert_array_equal(getattr(a[3], f), getattr(b[3], f))


@pytest.mark.parametrize("frac,threads,weighted", [(0.9, 4, False), (0.5, 16, False), (1.0, 3, False),
                                                   (0.9, 2, False), (0.7, 8, True)])
def test_component_groups_change_nothing(frac, threads, weighted):
    # rows [0, k) grouped by the components of their forest (computed here
    # with scipy) and linked on several threads, rows [k, m) that do not join
    # the largest group run beside it: identical outputs
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import connected_components

    from fast_hdbscan_amd import labels as PL

    X = workloads.anisotropic_mixture(120_000, 2, 12, 0.1, seed=9)
    e, _, _ = O.compute_minimum_spanning_tree(X, 5)
    n = e.shape[0] + 1
    o = np.lexsort((e[:, 1], e[:, 0], e[:, 2]))
    src, dst = e[o, 0].astype(np.int32), e[o, 1].astype(np.int32)
    dist = e[o, 2].copy()
    k = int((n - 1) * frac)
    g = coo_matrix((np.ones(k), (src[:k], dst[:k])), shape=(n, n))
    _, lab = connected_components(g, directed=False)
    comp = lab[src[:k]]
    ro = np.argsort(comp, kind="stable").astype(np.int32)
    gs = np.bincount(comp)
    gs = gs[gs > 0].astype(np.int32)  # sorted component ids: groups in the same order as ro
    params = PL.label_params(20, "eom", np.inf, False, 0.0)
    w = np.random.default_rng(3).uniform(0.5, 2.0, n).astype(np.float32) if weighted else None
    a = PL.labels_from_sorted_arrays(n, src, dst, None, params, return_trees=True, dist=dist, sample_weights=w)
    b = PL.labels_from_sorted_arrays(n, src, dst, None, params, return_trees=True, dist=dist, groups=(k, ro, gs),
                                     threads=threads, sample_weights=w)
    for x, y in zip(a[:3], b[:3]):
        np.testing.assert_array_equal(x, y)
    for f in ("parent", "child", "lambda_val", "child_size"):
        np.testing.assert_array_equal(getattr(a[3], f), getattr(b[3], f))


def test_component_groups_cycle_runs_in_order():
    # a row after the grouped prefix that closes a cycle inside the largest
    # group (the rows are no forest): the grouped linkage gives up its
    # parallel tail and reproduces the sequential result
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import connected_components

    from fast_hdbscan_amd import labels as PL

    X = workloads.anisotropic_mixture(120_000, 2, 12, 0.1, seed=9)
    e, _, _ = O.compute_minimum_spanning_tree(X, 5)
    n = e.shape[0] + 1
    o = np.lexsort((e[:, 1], e[:, 0], e[:, 2]))
    src, dst = e[o, 0].astype(np.int32), e[o, 1].astype(np.int32)
    dist = e[o, 2].copy()
    k = int((n - 1) * 0.9)
    _, lab = connected_components(coo_matrix((np.ones(k), (src[:k], dst[:k])), shape=(n, n)), directed=False)
    comp = lab[src[:k]]
    big = np.bincount(comp).argmax()
    pts = np.unique(np.concatenate([src[:k][comp == big], dst[:k][comp == big]]))
    src[k + 100], dst[k + 100] = pts[0], pts[1]
    ro = np.argsort(comp, kind="stable").astype(np.int32)
    gs = np.bincount(comp)
    gs = gs[gs > 0].astype(np.int32)
    params = PL.label_params(20, "eom", np.inf, False, 0.0)
    a = PL.labels_from_sorted_arrays(n, src, dst, None, params, return_trees=True, dist=dist)
    b = PL.labels_from_sorted_arrays(n, src, dst, None, params, return_trees=True, dist=dist, groups=(k, ro, gs),
                                     threads=4)
    for x, y in zip(a[:3], b[:3]):
        np.testing.assert_array_equal(x, y)


def test_component_groups_are_checked():
    from fast_hdbscan_amd import labels as PL

    e, _, _ = O.compute_minimum_spanning_tree(workloads.blobs_small(500, seed=1), 5)
    n = e.shape[0] + 1
    o = np.lexsort((e[:, 1], e[:, 0], e[:, 2]))
    src, dst = e[o, 0].astype(np.int32), e[o, 1].astype(np.int32)
    params = PL.label_params(5, "eom", np.inf, False, 0.0)
    with pytest.raises(ValueError):
        PL.labels_from_sorted_arrays(n, src, dst, None, params, dist=e[o, 2].copy(),
                                     groups=(10, np.arange(10, dtype=np.int32), np.array([4, 4], np.int32)))
    with pytest.raises(ValueError):
        PL.labels_from_sorted_arrays(n, src, dst, None, params, dist=e[o, 2].copy(),
                                     groups=(3, np.array([0, 1, 7], np.int32), np.array([3], np.int32)))
    # a repeated row (row_order not a permutation of [0, k))
    with pytest.raises(ValueError):
        PL.labels_from_sorted_arrays(n, src, dst, None, params, dist=e[o, 2].copy(),
                                     groups=(4, np.array([0, 1, 1, 3], np.int32), np.array([2, 2], np.int32)))
    # a negative group size that the others make up for
    with pytest.raises(ValueError):
        PL.labels_from_sorted_arrays(n, src, dst, None, params, dist=e[o, 2].copy(),
                                     groups=(4, np.arange(4, dtype=np.int32), np.array([5, -1], np.int32)))


def test_endpoints_and_point_order_are_checked():
    # endpoints outside [-1, n) and point ids outside [0, n) are errors, not
    # reads or writes out of bounds
    from fast_hdbscan_amd import labels as PL

    e, _, _ = O.compute_minimum_spanning_tree(workloads.blobs_small(500, seed=1), 5)
    n = e.shape[0] + 1
    o = np.lexsort((e[:, 1], e[:, 0], e[:, 2]))
    src, dst = e[o, 0].astype(np.int32), e[o, 1].astype(np.int32)
    dist = e[o, 2].copy()
    params = PL.label_params(5, "eom", np.inf, False, 0.0)
    for bad in (n, -2, np.iinfo(np.int32).max, np.iinfo(np.int32).min):
        s2 = src.copy()
        s2[7] = bad
        with pytest.raises(ValueError):
            PL.labels_from_sorted_arrays(n, s2, dst, None, params, dist=dist)
    order = np.arange(n, dtype=np.int32)
    order[3] = n
    with pytest.raises(ValueError):
        PL.labels_from_sorted_arrays(n, src, dst, None, params, dist=dist, point_order=order)
    lab, *_ = PL.labels_from_sorted_arrays(n, src, dst, None, params, dist=dist, point_order=np.arange(n, dtype=np.int32))
    ref, *_ = PL.labels_from_sorted_arrays(n, src, dst, None, params, dist=dist)
    np.testing.assert_array_equal(lab, ref)


def test_concurrent_label_stages_match():
    # several host threads in the label stage at once (the batch scheduler
    # and the PLSCAN sweep do this): the worker pool serves one, the others
    # run their parallel passes on threads of their own; results unchanged
    import threading

    from fast_hdbscan_amd import labels as PL

    X = workloads.anisotropic_mixture(90_000, 2, 12, 0.1, seed=5)
    e, _, _ = O.compute_minimum_spanning_tree(X, 5)
    n = e.shape[0] + 1
    o = np.lexsort((e[:, 1], e[:, 0], e[:, 2]))
    src, dst = e[o, 0].astype(np.int32), e[o, 1].astype(np.int32)
    dist = e[o, 2].copy()
    cases = [PL.label_params(mcs, "eom", np.inf, False, 0.0) for mcs in (5, 20, 50, 200)]
    ref = [PL.labels_from_sorted_arrays(n, src, dst, None, p, return_trees=True, dist=dist) for p in cases]
    out = [None] * (2 * len(cases))
    errs = []

    def run(i):
        try:
            out[i] = PL.labels_from_sorted_arrays(n, src, dst, None, cases[i % len(cases)], return_trees=True,
                                                  dist=dist)
        except Exception as ex:  # pragma: no cover - reported below
            errs.append(ex)

    th = [threading.Thread(target=run, args=(i,)) for i in range(len(out))]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert not errs
    for i, got in enumerate(out):
        want = ref[i % len(cases)]
        for x, y in zip(got[:3], want[:3]):
            np.testing.assert_array_equal(x, y)
        for f in ("parent", "child", "lambda_val", "child_size"):
            np.testing.assert_array_equal(getattr(got[3], f), getattr(want[3], f))


def test_k_exceeds_n_condensed_tree_matches_oracle():
    # min_samples > n: +inf rows with -1 endpoints, whose leaf adds a
    # condensed row but no size; an eliminated branch holding it must list
    # all its leaves (regression: a size-based reservation dropped a row)
    from fast_hdbscan_amd import labels as PL

    X = np.array([[3, 0], [0, 4], [4, 3], [0, 4], [3, 0], [2, 1], [0, 1], [5, 5]], dtype=np.float64)
    e, _, _ = O.compute_minimum_spanning_tree(X, 11)
    assert (e[:, :2] < 0).any()
    for mcs, single in ((6, True), (6, False), (3, True), (2, False)):
        kw = dict(min_cluster_size=mcs, allow_single_cluster=single)
        a = PL.clusters_from_spanning_tree(e, **kw)
        b = LB.clusters_from_spanning_tree(e, **kw)
        np.testing.assert_array_equal(a[0], b[0])
        np.testing.assert_array_equal(a[1], b[1])
        for f in ("parent", "child", "lambda_val", "child_size"):
            np.testing.assert_array_equal(getattr(a[3], f), getattr(b[3], f))


@pytest.mark.parametrize("case", [c for c in CASES if c[4] == 0.0], ids=[f"{c[0]}-{c[3]}-{c[2]}-{c[5]}" for c in CASES
                                                                     if c[4] == 0.0])
def test_compiled_label_oracle_matches_python_oracle(case):
    # the compiled restatement (the CPU baseline's label stage) against the
    # Python one, row for row
    name, ms, mcs, meth, eps, single = case
    if mcs <= 2:
        sys.setrecursionlimit(max(sys.getrecursionlimit(), 20000))
    edges, _, _ = O.compute_minimum_spanning_tree(_data(name), ms)
    kw = dict(min_cluster_size=mcs, cluster_selection_method=meth, allow_single_cluster=single)
    a = LB.clusters_from_spanning_tree(edges, **kw)
    b = LB.clusters_from_spanning_tree_c(edges, **kw)
    for x, y in zip(a[:3], b[:3]):
        np.testing.assert_array_equal(x, y)
    for f in ("parent", "child", "lambda_val", "child_size"):
        np.testing.assert_array_equal(getattr(a[3], f), getattr(b[3], f))
