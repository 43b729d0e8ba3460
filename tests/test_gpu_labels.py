"""End-to-end on the device: MST kernels + device edge sort + host label
stage, checked against the oracles (C MST restatement + Python label
restatement) and the reference's known answers."""
import numpy as np
import pytest

from fast_hdbscan_amd import workloads

pytestmark = pytest.mark.gpu


def _oracle(X, ms, mcs, **kw):
    from oracle import labels as LB
    from oracle import oracle as O

    edges, _, _ = O.compute_minimum_spanning_tree(X, ms)
    return LB.clusters_from_spanning_tree(edges, min_cluster_size=mcs, **kw)


@pytest.mark.parametrize("name,ms,mcs", [("blobs10k", 10, 10), ("aniso200k", 10, 100), ("grid", 4, 20)])
def test_device_pipeline_matches_oracle(name, ms, mcs):
    import torch

    from fast_hdbscan_amd.estimator import hdbscan_device

    if name == "blobs10k":
        X = workloads.blobs_small(10_000, seed=0)
    elif name == "aniso200k":
        X = workloads.scale_2d(200_000, seed=6)
    else:
        g = np.stack(np.meshgrid(np.arange(80), np.arange(60)), -1).reshape(-1, 2).astype(np.float32)
        X = g[np.random.default_rng(3).permutation(g.shape[0])]
    print(f"pipeline case {name} n={X.shape[0]}", flush=True)
    lab, prob = hdbscan_device(torch.from_numpy(X).cuda(), ms, mcs)
    olab, oprob, *_ = _oracle(X, ms, mcs)
    np.testing.assert_array_equal(lab, olab)
    np.testing.assert_array_equal(prob, oprob)


def test_device_pipeline_reuses_work_buffers():
    # one `work` dict over calls whose sizes shrink and grow (buffers kept,
    # regrown when too small, outputs reused): each result equals a call
    # that keeps nothing
    import torch

    from fast_hdbscan_amd.estimator import hdbscan_device

    work = {}
    for n, seed in [(150_000, 11), (20_000, 12), (160_000, 13)]:
        X = torch.from_numpy(workloads.scale_2d(n, seed=seed)).cuda()
        print(f"work reuse n={n}", flush=True)
        lab, prob = hdbscan_device(X, 10, 50, work=work)
        lab, prob = lab.copy(), prob.copy()
        flab, fprob = hdbscan_device(X, 10, 50)
        np.testing.assert_array_equal(lab, flab)
        np.testing.assert_array_equal(prob, fprob)
        assert lab.shape == (n,)


def test_device_sort_matches_lexsort():
    import ctypes

    import torch

    from fast_hdbscan_amd import _lib

    rng = np.random.default_rng(0)
    m = 300_000
    src = rng.integers(0, 1000, m).astype(np.int32)
    dst = rng.integers(-1, 1000, m).astype(np.int32)
    w = rng.integers(0, 50, m).astype(np.float32)  # many ties
    w[::97] = np.inf
    t = [torch.from_numpy(a).cuda() for a in (src, dst, w)]
    o = [torch.empty_like(a) for a in t]
    ctx = _lib.context(0)
    code = _lib.load().fh_sort_mst_device(ctx, m, *[ctypes.c_void_p(a.data_ptr()) for a in t + o])
    _lib.check(ctx, code)
    order = np.lexsort((dst.astype(np.float64), src.astype(np.float64), np.sqrt(w.astype(np.float64))))
    np.testing.assert_array_equal(o[0].cpu().numpy(), src[order])
    np.testing.assert_array_equal(o[1].cpu().numpy(), dst[order])
    np.testing.assert_array_equal(o[2].cpu().numpy(), w[order])


def test_estimator_known_answers():
    # test_hdbscan.py:169-186 through the estimator mirror on the device path
    from fast_hdbscan_amd.estimator import HDBSCAN

    np.random.seed(0)
    X = np.random.rand(150, 2)
    c = HDBSCAN(min_cluster_size=5).fit(X)
    u, cnt = np.unique(c.labels_, return_counts=True)
    assert len(u) == 8 and cnt[u == -1] == 69
    c = HDBSCAN(min_cluster_size=5, cluster_selection_epsilon=0.2).fit(X)
    u, cnt = np.unique(c.labels_, return_counts=True)
    assert len(u) == 2 and cnt[u == -1] == 2


def test_estimator_edge_cases():
    # test_edge_cases.py
    from fast_hdbscan_amd.estimator import HDBSCAN

    assert HDBSCAN(min_cluster_size=2).fit_predict(np.array([[1.0, 2.0]]))[0] == -1
    lab = HDBSCAN(min_cluster_size=2).fit_predict(np.array([[0.0, 0.0], [0.1, 0.0]]))
    assert len(set(lab) - {-1}) <= 1
    assert len(HDBSCAN(min_cluster_size=5).fit_predict(np.full((50, 3), 1.0))) == 50
    np.random.seed(42)
    assert np.all(HDBSCAN(min_cluster_size=100).fit_predict(np.random.rand(10, 2)) == -1)
    np.random.seed(42)
    X = np.concatenate([np.random.normal(0, 0.1, 50), np.random.normal(5, 0.1, 50)]).reshape(-1, 1)
    assert len(set(HDBSCAN(min_cluster_size=10).fit_predict(X)) - {-1}) >= 1
    np.random.seed(42)
    assert len(HDBSCAN(min_cluster_size=5, min_samples=18).fit_predict(np.random.rand(20, 2))) == 20


def test_estimator_missing_data():
    # test_hdbscan.py:42-56
    from sklearn.datasets import make_blobs
    from sklearn.preprocessing import StandardScaler
    from sklearn.utils import shuffle

    from fast_hdbscan_amd.estimator import HDBSCAN

    X, y = make_blobs(n_samples=200, random_state=10)
    X, y = shuffle(X, y, random_state=7)
    X = StandardScaler().fit_transform(X)
    Xm = X.copy()
    Xm[0] = [np.nan, 1]
    Xm[5] = [np.nan, np.nan]
    model = HDBSCAN().fit(Xm)
    assert model.labels_[0] == -1 and model.labels_[5] == -1
    assert model.probabilities_[0] == 0 and model.probabilities_[5] == 0
    clean = list(range(1, 5)) + list(range(6, 200))
    clean_model = HDBSCAN().fit(Xm[clean])
    assert np.allclose(clean_model.labels_, model.labels_[clean])


def test_reference_sample_weight_estimator():
    # test_hdbscan.py:113-117 end to end: device weighted MST + weighted labels
    from sklearn.datasets import make_blobs
    from sklearn.preprocessing import StandardScaler
    from sklearn.utils import shuffle

    from fast_hdbscan_amd.estimator import HDBSCAN, fast_hdbscan
    from oracle import labels as OL
    from oracle import oracle as O

    X, y = make_blobs(n_samples=200, random_state=10)
    X, y = shuffle(X, y, random_state=7)
    X = StandardScaler().fit_transform(X)
    sw = y.astype(np.float32)
    labels, p = fast_hdbscan(X, sample_weights=sw)
    assert len(set(labels)) - int(-1 in labels) == 2
    e, _, _ = O.compute_minimum_spanning_tree(X, 10, sample_weights=sw)
    ref = OL.clusters_from_spanning_tree(e, min_cluster_size=10, sample_weights=sw)
    np.testing.assert_array_equal(labels, ref[0])
    np.testing.assert_array_equal(p, ref[1])
    np.testing.assert_array_equal(HDBSCAN(min_cluster_size=10).fit(X, sample_weight=sw).labels_, labels)


def test_linkage_plan_matches_scipy_components():
    import ctypes

    import torch
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import connected_components

    from fast_hdbscan_amd import _lib

    rng = np.random.default_rng(5)
    n = 150_000
    # a random spanning tree in random row order
    parent = np.array([rng.integers(0, i) for i in range(1, n)], np.int32)
    src = np.arange(1, n, dtype=np.int32)
    dst = parent
    p = rng.permutation(n - 1)
    src, dst = src[p].copy(), dst[p].copy()
    k = 100_000
    print(f"linkage plan n={n} k={k}", flush=True)
    ctx = _lib.context(0)
    ds, dd = torch.from_numpy(src).cuda(), torch.from_numpy(dst).cuda()
    ro = np.empty(k, np.int32)
    gs = np.empty(k, np.int32)
    ng = ctypes.c_int64(0)
    _lib.check(ctx, _lib.load().fh_linkage_plan(ctx, n, k, ctypes.c_void_p(ds.data_ptr()), ctypes.c_void_p(dd.data_ptr()),
                                                _lib.ptr(ro), _lib.ptr(gs), ctypes.byref(ng)))
    gs = gs[: ng.value]
    _, lab = connected_components(coo_matrix((np.ones(k), (src[:k], dst[:k])), shape=(n, n)), directed=False)
    comp = lab[src[:k]]
    assert np.array_equal(np.sort(ro), np.arange(k))
    starts = np.concatenate([[0], np.cumsum(gs)])
    assert starts[-1] == k
    seen = set()
    for g in range(gs.shape[0]):
        rows = ro[starts[g]:starts[g + 1]]
        assert np.all(np.diff(rows) > 0)
        c = np.unique(comp[rows])
        assert c.shape[0] == 1 and int(c[0]) not in seen
        seen.add(int(c[0]))
    assert len(seen) == np.unique(comp).shape[0]
