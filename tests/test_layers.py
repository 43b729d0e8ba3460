"""PLSCAN layer stage: the host library against the oracle, and the
reference's own PLSCAN tests (test_plscan.py) on this package's PLSCAN.

The MSTs fed to the layer stage here come from the oracle, so these tests
run without a GPU; tests/test_gpu_layers.py runs the device MST."""
import numpy as np
import pytest
from sklearn.datasets import make_blobs

from fast_hdbscan_amd import layers as FL
from oracle import layers as OL
from oracle import oracle as O

X3, _ = make_blobs(n_samples=300, random_state=42, centers=3, cluster_std=1.0)


def _compare(edges, core, n, **kw):
    ro = OL.cluster_layers_from_mst(edges, core, n, **kw)
    rf = FL.layers_from_mst(edges, core, n, **kw)
    assert len(ro[0]) == len(rf[0])
    for a, b in zip(ro[0], rf[0]):
        np.testing.assert_array_equal(a, b)
    for a, b in zip(ro[1], rf[1]):
        np.testing.assert_array_equal(a, b)
    assert [float(v) for v in ro[2]] == [float(v) for v in rf[2]]
    np.testing.assert_array_equal(ro[3], rf[3])
    np.testing.assert_array_equal(ro[4], rf[4])
    assert OL.build_layer_cluster_tree(ro[0]) == FL.build_layer_cluster_tree(rf[0])
    return rf


@pytest.mark.parametrize("kw", [
    dict(base_min_cluster_size=5),
    dict(base_min_cluster_size=5, max_layers=3, layer_similarity_threshold=0.1),
    dict(base_min_cluster_size=5, max_layers=5, layer_similarity_threshold=0.01),
    dict(base_min_cluster_size=5, max_layers=5, layer_similarity_threshold=0.9),
    dict(base_min_cluster_size=5, base_n_clusters=4),
    dict(base_min_cluster_size=5, base_n_clusters=1),
    dict(base_min_cluster_size=5, max_layers=1),
    dict(base_min_cluster_size=2, max_layers=10, layer_similarity_threshold=0.5),
])
def test_layers_match_oracle_blobs(kw):
    e, _, core = O.compute_minimum_spanning_tree(X3, 5)
    _compare(e, core, 300, **kw)


@pytest.mark.parametrize("seed,n,ms,mcs", [(0, 2000, 5, 5), (1, 3000, 10, 10), (2, 1500, 3, 2), (3, 800, 1, 5)])
def test_layers_match_oracle_mixtures(seed, n, ms, mcs):
    rng = np.random.default_rng(seed)
    X, _ = make_blobs(n_samples=n, centers=rng.integers(3, 12), cluster_std=rng.uniform(0.3, 2.0),
                      random_state=seed)
    X = np.vstack([X, rng.uniform(X.min(0), X.max(0), (n // 10, 2))])
    e, _, core = O.compute_minimum_spanning_tree(X, ms)
    _compare(e, core, X.shape[0], base_min_cluster_size=mcs, max_layers=6, layer_similarity_threshold=0.3)


def test_layers_integer_grid_ties():
    # many equal weights: the argsort of the weights and numba's quicksort
    # orders are exercised together
    g = np.stack(np.meshgrid(np.arange(20), np.arange(15)), -1).reshape(-1, 2).astype(np.float64)
    e, _, core = O.compute_minimum_spanning_tree(g, 4)
    _compare(e, core, g.shape[0], base_min_cluster_size=4, max_layers=4)


def test_layers_tiny():
    for n in (2, 3, 4, 7):
        X = np.random.default_rng(n).normal(size=(n, 2))
        e, _, core = O.compute_minimum_spanning_tree(X, 2)
        _compare(e, core, n, base_min_cluster_size=2, max_layers=3)


# ------------------------------------------------------------ oracle pins ---

def test_find_peaks_matches_scipy():
    from scipy.signal import find_peaks

    rng = np.random.default_rng(5)
    for _ in range(50):
        x = rng.integers(0, 6, rng.integers(3, 60)).astype(np.float32)  # plateaus included
        np.testing.assert_array_equal(OL.find_peaks(x), find_peaks(x)[0])


def test_numba_argsort_is_a_sort():
    rng = np.random.default_rng(6)
    for n in (0, 1, 2, 5, 14, 15, 16, 40, 300):
        x = rng.permutation(n).astype(np.float32)  # no ties: unique answer
        np.testing.assert_array_equal(OL.numba_argsort(x), np.argsort(x, kind="stable"))
        y = rng.integers(0, 4, n)
        o = OL.numba_argsort(y)
        assert np.all(np.diff(y[o]) >= 0) and sorted(o.tolist()) == list(range(n))
        if n < 15:  # insertion sort alone: stable
            np.testing.assert_array_equal(o, np.argsort(y, kind="stable"))


def test_numba_argsort_ties_by_hand():
    # numba's quicksort argsort (numba/misc/quicksort.py: run_quicksort,
    # partition, insertion_sort; SMALL_QUICKSORT = 15) on 16 keys with ties.
    A = [1, 0, 1, 2, 0, 1, 2, 0, 1, 1, 2, 0, 1, 2, 0, 1]
    # run_quicksort: (0, 15) has high - low = 15 >= 15 -> partition(0, 15):
    #   mid = 7; A[R7]=0 < A[R0]=1: swap R0,R7; A[R15]=1 < A[R7]=1 no;
    #   A[R7]=1 < A[R0]=0 no; pivot = 1 (key 0, now at 7), parked at 15.
    #   R = [7,1,2,3,4,5,6,15,8,...,14,0]; i/j scans (strict '<' both
    #   ways, so keys equal to the pivot stop both scans) swap positions
    #   (2,14) (3,12) (5,11) (6,9) (7,8), then i=8 > j=7: the pivot goes
    #   to 8.  R = [7,1,14,12,4,11,9,8,0,6,10,5,3,13,2,15].
    #   high - i = 7 is not > i - low = 8: push (0,7), low = 9;
    #   15 - 9 < 15 -> insertion_sort(9,15) over keys 6,10,5,3,13,2,15
    #   (values 2,2,1,2,2,1,1) -> 5,2,15,6,10,3,13 (stable within each value).
    # pop (0,7): insertion_sort over 7,1,14,12,4,11,9,8 (values 0,0,0,1,0,0,
    #   1,1) -> 7,1,14,4,11,12,9,8.
    # (np.argsort(kind="stable") would give 1,4,7,11,14,0,2,5,...)
    expect = [7, 1, 14, 4, 11, 12, 9, 8, 0, 5, 2, 15, 6, 10, 3, 13]
    assert OL.numba_argsort(np.asarray(A, np.int64)).tolist() == expect
    assert OL.numba_argsort(np.asarray(A, np.float32)).tolist() == expect


def test_total_persistence_definition():
    rng = np.random.default_rng(7)
    births = np.concatenate([[5.0], rng.integers(2, 30, 40)]).astype(np.float32)
    deaths = (births + rng.integers(-3, 20, 41)).astype(np.float32)
    lam = rng.uniform(0, 1, 41).astype(np.float32)
    sizes, tp = OL.compute_total_persistence(births, deaths, lam)
    want = np.zeros(len(sizes))
    for i in range(1, 41):
        if deaths[i] > births[i]:
            m = (sizes >= births[i]) & (sizes < deaths[i])
            want[m] += float(deaths[i]) - float(births[i]) * float(lam[i])
    np.testing.assert_allclose(tp, want, rtol=1e-5)


# ------------------------------------------- reference PLSCAN test suite ---

def _plscan_on_oracle_mst(monkeypatch):
    # route the estimator's MST through the oracle: CPU-only tests
    import fast_hdbscan_amd.layers as layers_mod

    def mst(data, min_samples=10, sample_weights=None, **_):
        return O.compute_minimum_spanning_tree(np.asarray(data), min_samples, sample_weights=sample_weights)

    monkeypatch.setattr(layers_mod, "compute_minimum_spanning_tree", mst)
    from fast_hdbscan_amd.estimator import PLSCAN

    return PLSCAN


def test_plscan_reference_suite(monkeypatch):
    run_reference_plscan_suite(_plscan_on_oracle_mst(monkeypatch))


def run_reference_plscan_suite(PLSCAN):
    X = X3
    # test_plscan.py:23-27
    labels = PLSCAN().fit(X).labels_
    assert len(set(labels)) - int(-1 in labels) == 3
    # 30-32
    PLSCAN().fit([[1.0, 2.0], [3.0, 4.0]])
    # 35-43
    with pytest.raises(ValueError):
        PLSCAN().fit("fail")
    with pytest.raises(ValueError):
        PLSCAN().fit(None)
    with pytest.raises(ValueError):
        PLSCAN(min_samples="fail").fit(X)
    with pytest.raises(ValueError):
        PLSCAN(min_samples=-1).fit(X)
    # 46-48
    assert len(PLSCAN(max_layers=1).fit(X).cluster_layers_) == 1
    # 51-55
    lo = PLSCAN(layer_similarity_threshold=0.01, max_layers=5).fit(X)
    hi = PLSCAN(layer_similarity_threshold=0.9, max_layers=5).fit(X)
    assert len(lo.cluster_layers_) <= len(hi.cluster_layers_)
    # 58-61
    np.testing.assert_array_equal(PLSCAN(reproducible=True).fit(X).labels_, PLSCAN(reproducible=True).fit(X).labels_)
    # 64-69
    first = PLSCAN(base_n_clusters=4).fit(X).cluster_layers_[0]
    assert abs(len(set(first)) - int(-1 in first) - 4) <= 1
    # 72-77
    m = PLSCAN(max_layers=3, layer_similarity_threshold=0.1).fit(X)
    assert len(m.cluster_layers_) > 1
    assert len(set(map(tuple, [set(layer) for layer in m.cluster_layers_]))) == len(m.cluster_layers_)
    # 80-86, 225-240, 243-263
    tree = PLSCAN(max_layers=3).fit(X).cluster_tree_
    assert isinstance(tree, dict)
    for parent, children in tree.items():
        assert isinstance(parent, tuple) and len(parent) == 2
        for child in children:
            assert isinstance(child, tuple) and len(child) == 2 and child[0] < parent[0]
    mm = PLSCAN(max_layers=3).fit(X)
    seen = set()
    for parent, children in mm.cluster_tree_.items():
        seen.add(parent)
        seen.update(children)
    actual = {(i, int(lab)) for i, layer in enumerate(mm.cluster_layers_) for lab in np.unique(layer) if lab >= 0}
    assert actual.issubset(seen)
    # 89-117
    for bad in (0, -1, "invalid"):
        with pytest.raises(ValueError):
            PLSCAN(max_layers=bad).fit(X)
    for bad in (0, -5, "invalid"):
        with pytest.raises(ValueError):
            PLSCAN(base_min_cluster_size=bad).fit(X)
    for bad in (0, -1, "invalid"):
        with pytest.raises(ValueError):
            PLSCAN(base_n_clusters=bad).fit(X)
    # 120-145
    m = PLSCAN().fit(X)
    assert m.membership_strengths_.shape == m.labels_.shape
    assert np.all(m.membership_strengths_ >= 0) and np.all(m.membership_strengths_ <= 1)
    assert np.all(m.membership_strengths_[m.labels_ == -1] == 0)
    m = PLSCAN(max_layers=3).fit(X)
    assert len(m.membership_strength_layers_) == len(m.cluster_layers_)
    # 148-152
    m = PLSCAN(max_layers=5).fit(X)
    assert len(m.layer_persistence_scores_) == len(m.cluster_layers_)
    assert all(s >= 0 for s in m.layer_persistence_scores_)
    # 155-164
    m = PLSCAN(max_layers=5, layer_similarity_threshold=0.1).fit(X)
    if len(m.cluster_layers_) > 1:
        best = np.argmax(m.layer_persistence_scores_)
        np.testing.assert_array_equal(m.labels_, m.cluster_layers_[best])
    # 177-182
    small = PLSCAN(base_min_cluster_size=2).fit(np.array([[0, 0], [1, 1], [0, 1], [1, 0]]))
    assert len(small.labels_) == 4
    # 185-190
    Xs, _ = make_blobs(n_samples=100, centers=1, cluster_std=0.1, random_state=42)
    lab = PLSCAN(base_min_cluster_size=5).fit(Xs).labels_
    assert len(set(lab)) - int(-1 in lab) >= 1
    # 193-197
    PLSCAN(base_min_cluster_size=10).fit(np.random.RandomState(42).uniform(0, 10, (100, 2)))
    # 200-204, 207-211
    model = PLSCAN()
    assert model.fit(X) is model
    np.testing.assert_array_equal(PLSCAN().fit_predict(X), PLSCAN().fit(X).labels_)
    # 266-272, 275-285
    m = PLSCAN(max_layers=5, layer_similarity_threshold=0.1).fit(X)
    counts = [len(set(layer)) - int(-1 in layer) for layer in m.cluster_layers_]
    assert counts == sorted(counts, reverse=True)
    # 288-298: sample weights
    w = np.ones(X.shape[0])
    w[: X.shape[0] // 2] = 2.0
    assert hasattr(PLSCAN().fit(X, sample_weight=w), "labels_")
    # 301-322
    p = PLSCAN(min_samples=10, max_layers=5, base_min_cluster_size=15, layer_similarity_threshold=0.3,
               reproducible=True).get_params()
    assert p["min_samples"] == 10 and p["max_layers"] == 5 and p["base_min_cluster_size"] == 15
    model = PLSCAN()
    model.set_params(min_samples=15, max_layers=7)
    assert model.min_samples == 15 and model.max_layers == 7
