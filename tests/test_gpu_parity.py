"""Device path vs the CPU oracle, bit for bit, through the C-ABI.

Every case compares the library's outputs with the oracle's sequential
restatement of parallel_boruvka on the same float32 points:
core distances, neighbour columns, and the MST rows (source, target,
squared weight) in row order.
"""
import numpy as np
import pytest

from fast_hdbscan_amd import workloads

pytestmark = pytest.mark.gpu


def _oracle(X, ms):
    from oracle import oracle as O

    return O.boruvka(O.build_kdtree(X), ms)


def _device(X, ms):
    from fast_hdbscan_amd.hdbscan import mst_arrays

    stats = {}
    out = mst_arrays(X, ms, stats=stats)
    return out, stats


def _check(X, ms):
    X = np.ascontiguousarray(X, dtype=np.float32)
    print(f"case n={X.shape[0]} d={X.shape[1]} ms={ms}", flush=True)
    (src, dst, w, core, nbr), stats = _device(X, ms)
    ref = _oracle(X, ms)
    np.testing.assert_array_equal(core, ref.core_sq)
    np.testing.assert_array_equal(nbr, ref.knn_i[:, 1:])
    assert src.shape == ref.src.shape
    np.testing.assert_array_equal(w, ref.w_sq)
    np.testing.assert_array_equal(src, ref.src)
    np.testing.assert_array_equal(dst, ref.dst)
    assert stats["rounds"] == ref.rounds
    return stats


def _tree_case(name):
    rng = np.random.default_rng(0)
    if name == "blobs":
        return workloads.blobs_small(20_000, seed=0)
    if name == "scale2d_2M":
        return workloads.scale_2d(2_000_000, seed=1)
    if name == "scale3d_300k":
        return workloads.anisotropic_mixture(300_000, 3, 20, 0.05, seed=2)
    if name == "d8":
        return workloads.anisotropic_mixture(50_000, 8, 10, 0.05, seed=4)
    if name == "dups":
        X = workloads.blobs_small(30_000, seed=3)
        X[15000:] = X[:15000]
        return X
    if name == "grid":
        g = np.stack(np.meshgrid(np.arange(300), np.arange(200)), -1).reshape(-1, 2).astype(np.float32)
        return g[rng.permutation(g.shape[0])]
    if name == "sorted":
        return np.sort(rng.standard_normal((100_000, 2)).astype(np.float32), axis=0)
    if name == "reversed":
        return np.sort(rng.standard_normal((100_000, 2)).astype(np.float32), axis=0)[::-1].copy()
    if name == "organ_pipe":
        v = np.concatenate([np.arange(40_000), np.arange(40_000)[::-1]]).astype(np.float32)
        return np.stack([v, np.zeros_like(v)], 1)
    if name.startswith("dim"):
        # d-dimensional mixtures large enough that the subtree kernel starts
        # below the root (its LDS capacity shrinks with d)
        d = int(name[3:])
        return workloads.anisotropic_mixture(60_000 if d > 1 else 200_000, d, 12, 0.05, seed=40 + d)
    if name == "constant_axis":
        X = rng.standard_normal((60_000, 2)).astype(np.float32)
        X[:, 1] = 3.0
        return X
    raise KeyError(name)


@pytest.mark.parametrize("name", ["blobs", "scale2d_2M", "scale3d_300k", "d8", "dups", "grid", "sorted",
                                  "reversed", "organ_pipe", "constant_axis", "dim1", "dim4", "dim5", "dim6",
                                  "dim7"])
def test_device_tree_matches_oracle(name):
    _tree_check(name)


@pytest.mark.parametrize("name", ["blobs", "scale3d_300k", "d8", "dups", "grid", "organ_pipe"])
def test_device_tree_levels_only(name, monkeypatch):
    """The level-by-level kernels down to the leaves (FH_TREE_SUBTREE=0)
    instead of one work-group per subtree once a node fits in LDS."""
    monkeypatch.setenv("FH_TREE_SUBTREE", "0")
    _tree_check(name)


def _tree_check(name):
    from fast_hdbscan_amd.hdbscan import build_kdtree
    from oracle import oracle as O

    X = _tree_case(name)
    print(f"tree case {name} n={X.shape[0]} d={X.shape[1]}", flush=True)
    t = build_kdtree(X, on_device=True)
    o = O.build_kdtree(X)
    np.testing.assert_array_equal(t["idx_array"], o.idx_array)
    np.testing.assert_array_equal(t["idx_start"], o.idx_start)
    np.testing.assert_array_equal(t["idx_end"], o.idx_end)
    np.testing.assert_array_equal(t["is_leaf"], o.is_leaf)
    np.testing.assert_array_equal(t["node_bounds"], o.node_bounds)


def test_config1_blobs_10k():
    _check(workloads.blobs_small(10_000, seed=0), 10)


@pytest.mark.parametrize("ms", [1, 2, 5, 10, 15])
def test_blobs_min_samples(ms):
    _check(workloads.blobs_small(4000, seed=ms), ms)


@pytest.mark.parametrize("d", [1, 3, 4, 5, 8, 10])
@pytest.mark.parametrize("path", ["auto", "tree"])
def test_dimensions(d, path, monkeypatch):
    # small inputs from 5 dimensions on take the MFMA scans unless the
    # switch-over dimension is moved past d
    from sklearn.datasets import make_blobs

    if path == "tree":
        monkeypatch.setenv("FH_HIGH_DIM_FROM", "1000")
    X = make_blobs(n_samples=3000, n_features=d, centers=6, random_state=d)[0]
    st = _check(X, 5)
    assert st["high_dim"] == (1 if path == "auto" and d >= 5 else 0)


def test_uniform_reference_fixture():
    # test_hdbscan.py:169-186 data: np.random.seed(0); rand(150, 2)
    rng = np.random.RandomState(0)
    _check(rng.rand(150, 2), 5)


def test_all_duplicates():
    # test_edge_cases.py:20-24: 50 identical 3-D points
    _check(np.full((50, 3), 1.0), 5)


def test_duplicates_injected():
    X = workloads.blobs_small(5000, seed=3)
    rng = np.random.default_rng(0)
    dup = rng.choice(5000, 400, replace=False)
    X[dup[:200]] = X[dup[200:]]
    _check(X, 10)


def test_integer_grid_ties():
    g = np.stack(np.meshgrid(np.arange(60), np.arange(50)), -1).reshape(-1, 2).astype(np.float32)
    g = g[np.random.default_rng(1).permutation(g.shape[0])]
    _check(g, 4)


def test_k_exceeds_n():
    # test_edge_cases.py:26-31: min_samples=100 with 10 points
    rng = np.random.RandomState(42)
    _check(rng.rand(10, 2), 100)


def test_high_min_samples():
    # test_edge_cases.py:46-51
    rng = np.random.RandomState(42)
    _check(rng.rand(20, 2), 18)


def test_large_k_global_heap():
    _check(workloads.blobs_small(1500, seed=7), 150)


@pytest.mark.parametrize("n", [1, 2, 3, 41, 81, 82])
def test_tiny(n):
    rng = np.random.default_rng(n)
    _check(rng.standard_normal((n, 2)), 3)


def test_one_dimensional():
    # test_edge_cases.py:33-44
    rng = np.random.RandomState(42)
    X = np.concatenate([rng.normal(0, 0.1, 50), rng.normal(5, 0.1, 50)]).reshape(-1, 1)
    _check(X, 10)


def test_anisotropic_200k():
    _check(workloads.scale_2d(200_000, seed=5), 10)


def test_knn_external_queries():
    from fast_hdbscan_amd.hdbscan import parallel_tree_query
    from oracle import oracle as O

    X = workloads.blobs_small(20_000, seed=1)
    Q = workloads.blobs_small(3000, seed=2)
    d, i = parallel_tree_query(X, Q, k=7, output_rdist=True)
    rd, ri = O.tree_query(O.build_kdtree(X), Q, 7)
    np.testing.assert_array_equal(d, rd)
    np.testing.assert_array_equal(i, ri)


def test_knn_sqrt_mode_ties():
    """output_rdist=False: numba_kdtree.py:792 takes the float32 square root
    of the heap rows before deheap_sort.  Queries 20,000 units from an
    integer grid give distinct squared distances whose roots collide, so the
    sort then meets equal keys that the squared rows did not have."""
    from fast_hdbscan_amd.hdbscan import parallel_tree_query
    from oracle import oracle as O

    rng = np.random.default_rng(0)
    g = np.stack(np.meshgrid(np.arange(60), np.arange(60)), -1).reshape(-1, 2).astype(np.float32)
    X = g[rng.permutation(g.shape[0])]
    Q = np.stack([np.full(2000, -20000.0), rng.integers(0, 60, 2000)], 1).astype(np.float32)
    Q[:, 0] -= rng.integers(0, 50, 2000)
    tree = O.build_kdtree(X)
    for k in (12, 32):
        print(f"case sqrt-mode k={k}", flush=True)
        d, i = parallel_tree_query(X, Q, k=k, output_rdist=False)
        rd, ri = O.tree_query_sqrt(tree, Q, k)
        sq_d, sq_i = O.tree_query(tree, Q, k)
        # the case really exercises the order: squaring first would differ
        assert (sq_i != ri).any()
        np.testing.assert_array_equal(d, rd)
        np.testing.assert_array_equal(i, ri)


def test_repeatable():
    X = workloads.scale_2d(100_000, seed=9)
    a = _device(X, 10)[0]
    b = _device(X, 10)[0]
    for u, v in zip(a, b):
        np.testing.assert_array_equal(u, v)


def test_reproducible_flag_gives_the_one_thread_result():
    # boruvka_tree_query_reproducible (boruvka.py:334-471) is the reference's
    # blocked form of the same search; with one thread both equal the
    # sequential run, which is what the device path computes either way
    from fast_hdbscan_amd.hdbscan import compute_minimum_spanning_tree
    from oracle import oracle as O

    X = workloads.blobs_small(4000, seed=3)
    e0, n0, c0 = compute_minimum_spanning_tree(X, min_samples=7, reproducible=False)
    e1, n1, c1 = compute_minimum_spanning_tree(X, min_samples=7, reproducible=True)
    np.testing.assert_array_equal(e0, e1)
    np.testing.assert_array_equal(n0, n1)
    np.testing.assert_array_equal(c0, c1)
    er, nr, cr = O.compute_minimum_spanning_tree(X, 7)
    np.testing.assert_array_equal(e1, er)
    np.testing.assert_array_equal(n1, nr)
    np.testing.assert_array_equal(c1, cr)


def test_threads_get_their_own_contexts():
    """Two Python threads calling the library at once (ctypes releases the
    GIL) each use their own context and stream; results equal the
    one-thread calls."""
    import threading

    from fast_hdbscan_amd import _lib

    xs = [workloads.scale_2d(150_000, seed=21), workloads.blobs_small(120_000, seed=22)]
    want = [_device(x, 10)[0] for x in xs]
    got = [None, None]
    handles = [None, None]
    errors = []
    both = threading.Barrier(2, timeout=60)

    def run(j):
        try:
            handles[j] = _lib.context(0)
            both.wait()  # both contexts exist at once
            for _ in range(3):
                got[j] = _device(xs[j], 10)[0]
        except Exception as e:  # noqa: BLE001
            errors.append(e)

    ts = [threading.Thread(target=run, args=(j,)) for j in range(2)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert not errors, errors
    assert handles[0] != handles[1] and _lib.context(0) not in handles
    for j in range(2):
        for u, v in zip(want[j], got[j]):
            np.testing.assert_array_equal(u, v)


@pytest.mark.parametrize("kind", ["unit", "uniform", "labels", "sparse_zeros"])
def test_sample_weight_branch(kind):
    from oracle import oracle as O
    from fast_hdbscan_amd.hdbscan import mst_arrays

    rng = np.random.default_rng(11)
    X = workloads.blobs_small(6000, seed=12)
    if kind == "unit":
        w = np.ones(6000, np.float32)
    elif kind == "uniform":
        w = rng.uniform(0.5, 2.0, 6000).astype(np.float32)
    elif kind == "labels":
        w = rng.integers(0, 3, 6000).astype(np.float32)
    else:
        w = (rng.uniform(0, 1, 6000) > 0.7).astype(np.float32) * 3.0
    print(f"weighted case {kind}", flush=True)
    src, dst, wsq, core, nbr = mst_arrays(X, 10, sample_weights=w)
    ref = O.boruvka_weighted(O.build_kdtree(X), 10, w)
    np.testing.assert_array_equal(core, ref.core_sq)
    np.testing.assert_array_equal(nbr, ref.knn_i[:, 1:])
    np.testing.assert_array_equal(src, ref.src)
    np.testing.assert_array_equal(dst, ref.dst)
    np.testing.assert_array_equal(wsq, ref.w_sq)


def test_sample_weight_seam_contract():
    from oracle import oracle as O
    from fast_hdbscan_amd import compute_minimum_spanning_tree

    X = workloads.blobs_small(2000, seed=2).astype(np.float64)
    w = np.random.default_rng(1).uniform(0.5, 1.5, 2000).astype(np.float32)
    e, nb, core = compute_minimum_spanning_tree(X, min_samples=5, sample_weights=w)
    re, rnb, rcore = O.compute_minimum_spanning_tree(X, 5, sample_weights=w)
    np.testing.assert_array_equal(e, re)
    np.testing.assert_array_equal(nb, rnb)
    np.testing.assert_array_equal(core, rcore)
    assert e.dtype == np.float64 and nb.dtype == np.int32 and core.dtype == np.float32


@pytest.mark.parametrize("form", ["lane", "unpruned", "knn_point_rows", "sampled", "grid_rounds", "tie_cycles",
                                  "doubling_only"])
def test_search_kernel_forms(form, monkeypatch):
    """The main search exists as one lane per point (FH_QUERY_KERNEL=lane)
    and as one wavefront per 64 positions (default for d <= 8), and the
    unpruned single pass (FH_SEARCH_MODE=0, read when a context is made:
    run in a thread of its own); the kNN rows kept in point order
    (FH_KNN_ROWS=point) or in tree order (default); the sampled pre-pass on
    or off (FH_SAMPLED); the post-search work of every round as grid-wide
    kernels (FH_SMALL_ROUND=0) instead of one work-group once at most 1024
    components are left; the cycle skip with 2-cycles settled directly and
    doubling over equal-weight chains only, from any component count
    (FH_TIE_CYCLES=0), or pointer doubling over every component
    (doubling_only).  Each equals the oracle."""
    import threading

    cases = [workloads.scale_2d(120_000, seed=41), workloads.anisotropic_mixture(60_000, 3, 20, 0.05, seed=42)]
    g = np.stack(np.meshgrid(np.arange(90), np.arange(70)), -1).reshape(-1, 2).astype(np.float32)
    cases.append(g[np.random.default_rng(43).permutation(g.shape[0])])
    if form == "lane":
        monkeypatch.setenv("FH_QUERY_KERNEL", "lane")
    elif form == "knn_point_rows":
        monkeypatch.setenv("FH_KNN_ROWS", "point")
    elif form == "sampled":
        monkeypatch.setenv("FH_SAMPLED", "1")
    elif form == "grid_rounds":
        monkeypatch.setenv("FH_SMALL_ROUND", "0")
    elif form == "tie_cycles":
        monkeypatch.setenv("FH_SMALL_ROUND", "0")
        monkeypatch.setenv("FH_TIE_CYCLES", "0")
    elif form == "doubling_only":
        monkeypatch.setenv("FH_TIE_CYCLES", str(1 << 30))
    else:
        monkeypatch.setenv("FH_SEARCH_MODE", "0")
    errors = []

    def run():
        try:
            for X in cases:
                _check(X, 10 if X.shape[1] == 2 else 15)
        except Exception as e:  # noqa: BLE001
            errors.append(e)

    t = threading.Thread(target=run)
    t.start()
    t.join()
    assert not errors, errors
