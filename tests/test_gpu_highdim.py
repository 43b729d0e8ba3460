"""High-dimensional path: MFMA candidate scans + exact re-scoring.

Every case is checked bit for bit against the oracle (the reference's
KD-tree semantics) or against the device's own exact KD-tree path, which
the oracle pins at smaller sizes.  FH_HIGH_DIM_FROM moves the dimension
from which the MFMA path is used, so low-dimensional inputs with many ties
exercise its tie handling too."""
import numpy as np
import pytest

from fast_hdbscan_amd import workloads
from fast_hdbscan_amd.hdbscan import mst_arrays

pytestmark = pytest.mark.gpu


def _oracle(X, ms):
    from oracle import oracle as O

    return O.boruvka(O.build_kdtree(X), ms)


def _check(X, ms, expect_high_dim=True):
    st = {}
    src, dst, wsq, core, nbr = mst_arrays(X, ms, stats=st)
    assert st["high_dim"] == (1 if expect_high_dim else 0)
    ref = _oracle(X, ms)
    np.testing.assert_array_equal(core, ref.core_sq)
    np.testing.assert_array_equal(nbr, ref.knn_i[:, 1:])
    np.testing.assert_array_equal(src, ref.src)
    np.testing.assert_array_equal(dst, ref.dst)
    np.testing.assert_array_equal(wsq, ref.w_sq)
    return st


@pytest.mark.parametrize("d,n,ms", [(64, 6000, 10), (48, 5000, 5), (32, 6000, 10), (24, 4000, 1), (17, 5000, 30),
                                    (64, 3000, 63)])
def test_highdim_matches_oracle(d, n, ms):
    print(f"highdim d={d} n={n} ms={ms}", flush=True)
    X = workloads.gaussian_highdim(n, d, n_clusters=20, seed=d + ms)
    st = _check(X, ms)
    print(st["knn_tree_rows"], st["exact_rows"], st["rounds"], flush=True)


@pytest.mark.parametrize("sym", ["1", "0"])
def test_highdim_tile_pairs_once_or_twice(monkeypatch, sym):
    """The candidate and round scans take each 128 x 128 tile pair once
    (default: row and column minima / candidates from one product) or from
    both sides (FH_SYM_SCAN=0): both equal the oracle."""
    monkeypatch.setenv("FH_SYM_SCAN", sym)
    X = workloads.gaussian_highdim(7000, 40, n_clusters=12, seed=17)
    _check(X, 7)


def test_highdim_duplicates():
    # exact ties: duplicated rows send their kNN rows to the tree search and
    # make mutual-reachability ties that the depth-first order resolves
    rng = np.random.default_rng(4)
    X = workloads.gaussian_highdim(3000, 32, n_clusters=10, seed=4)
    X = np.vstack([X, X[rng.choice(3000, 600, replace=False)]])
    X = X[rng.permutation(X.shape[0])]
    st = _check(X, 8)
    assert st["knn_tree_rows"] > 0


def test_highdim_integer_lattice_ties():
    rng = np.random.default_rng(5)
    X = rng.integers(0, 4, size=(4000, 20)).astype(np.float32)
    _check(X, 5)


@pytest.mark.parametrize("d", [2, 3])
def test_mfma_path_forced_low_dim(monkeypatch, d):
    # the MFMA path on a low-dimensional integer grid: ties everywhere
    monkeypatch.setenv("FH_HIGH_DIM_FROM", "1")
    rng = np.random.default_rng(d)
    X = rng.integers(0, 40, size=(5000, d)).astype(np.float32)
    _check(X, 6)
    Y = workloads.blobs_small(6000, seed=d).astype(np.float32)[:, :d] if d <= 2 else rng.normal(size=(6000, d))
    _check(np.ascontiguousarray(Y, dtype=np.float32), 10)


def test_highdim_matches_tree_path_50k(monkeypatch):
    # same exact answer from both device paths at a size the oracle would
    # take minutes on
    X = workloads.gaussian_highdim(50000, 64, n_clusters=100, seed=3)
    a = mst_arrays(X, 10)
    monkeypatch.setenv("FH_HIGH_DIM_FROM", "1000")
    st = {}
    b = mst_arrays(X, 10, stats=st)
    assert st["high_dim"] == 0
    for u, v in zip(a, b):
        np.testing.assert_array_equal(u, v)


def test_tile_bounds_prune_and_agree(monkeypatch):
    """From the second round the scans take the rows in component order and
    skip the tile pairs that the first round's tile-pair minima rule out:
    fewer tile pairs are scored and the MST is the one every round in tree
    order (FH_TILE_BOUNDS=0) gives, bit for bit."""
    X = workloads.gaussian_highdim(40000, 48, n_clusters=60, seed=9)
    st_on = {}
    a = mst_arrays(X, 10, stats=st_on)
    monkeypatch.setenv("FH_TILE_BOUNDS", "0")
    st_off = {}
    b = mst_arrays(X, 10, stats=st_off)
    print(st_on["rounds"], st_on["scored_tile_pairs"], st_off["scored_tile_pairs"], flush=True)
    assert st_on["rounds"] >= 3
    assert st_on["scored_tile_pairs"] < st_off["scored_tile_pairs"]
    for u, v in zip(a, b):
        np.testing.assert_array_equal(u, v)


@pytest.mark.parametrize("step,retry", [("1", None), ("4", None), ("16", None), ("16", "0")])
def test_knn_sampled_thresholds(monkeypatch, step, retry):
    """kNN thresholds from every step-th tile (the 8th smallest value there):
    rows whose threshold could have missed a neighbour go to the tree search
    (knn_tree_rows), or every row runs again with full thresholds when more
    than the retry count do; the cores and neighbours equal the oracle's."""
    monkeypatch.setenv("FH_KNN_SAMPLE", step)
    if retry is not None:
        monkeypatch.setenv("FH_KNN_SAMPLE_RETRY", retry)
    X = workloads.gaussian_highdim(6000, 48, n_clusters=15, seed=21)
    st = _check(X, 10)
    print(step, retry, st["knn_tree_rows"], flush=True)


def test_knn_sampled_thresholds_small_clusters(monkeypatch):
    # clusters of ~12 points: a sample of whole tiles misses most of a
    # row's neighbours, so many thresholds are too large (overflowing
    # lists) or too small; every such row is caught by the check
    monkeypatch.setenv("FH_KNN_SAMPLE", "8")
    X = workloads.gaussian_highdim(5000, 32, n_clusters=400, seed=22)
    st = _check(X, 6)
    print(st["knn_tree_rows"], flush=True)


def test_knn_sampled_thresholds_duplicates(monkeypatch):
    monkeypatch.setenv("FH_KNN_SAMPLE", "4")
    rng = np.random.default_rng(23)
    X = workloads.gaussian_highdim(3000, 24, n_clusters=10, seed=23)
    X = np.vstack([X, X[rng.choice(3000, 900, replace=False)], X[:40], X[:40]])
    X = X[rng.permutation(X.shape[0])]
    _check(X, 5)


@pytest.mark.parametrize("scale", [1e-12, 1e-3, 1e6, 1e15])
def test_split_products_any_scale(scale):
    """The FP16 split scales the centred rows by a power of two chosen from
    the root box; results equal the oracle's whatever the data's scale."""
    X = (workloads.gaussian_highdim(4000, 40, n_clusters=12, seed=31) * scale).astype(np.float32)
    _check(X, 8)


def test_split_products_mixed_magnitudes():
    # dimensions 1e8 apart in scale and a large offset: the small ones fall
    # to FP16 subnormals after scaling, which the margin covers
    X = workloads.gaussian_highdim(4000, 33, n_clusters=12, seed=32).astype(np.float64)
    X[:, :10] *= 1e-5
    X[:, 10:20] *= 1e3
    X += 1e4
    _check(X.astype(np.float32), 7)


@pytest.mark.parametrize("split", ["1", "0"])
def test_split_and_fp32_products_agree_40k(monkeypatch, split):
    """FP16 split products (default) and FP32 products
    (FH_MFMA_SPLIT=0) select candidates differently but give the same
    exact MST."""
    X = workloads.gaussian_highdim(40000, 64, n_clusters=50, seed=33)
    monkeypatch.setenv("FH_MFMA_SPLIT", "0")
    ref = mst_arrays(X, 10)
    monkeypatch.setenv("FH_MFMA_SPLIT", split)
    got = mst_arrays(X, 10)
    for u, v in zip(got, ref):
        np.testing.assert_array_equal(u, v)


def test_tiled_exact_rescan_agrees_40k(monkeypatch):
    """The exact re-scan of the rows near their component's bound, tiled
    (64 rows against 2,048 positions per work-group, the default) or one
    work-group per row (FH_TILED_RESCAN=0): the same MST."""
    X = workloads.gaussian_highdim(40000, 48, n_clusters=40, seed=34)
    st = {}
    a = mst_arrays(X, 10, stats=st)
    monkeypatch.setenv("FH_TILED_RESCAN", "0")
    b = mst_arrays(X, 10)
    assert st["exact_rows"] > 0
    for u, v in zip(a, b):
        np.testing.assert_array_equal(u, v)


def test_tree_knn_forms_agree_duplicates_30k(monkeypatch):
    """Rows whose k + 1 nearest include ties take the reference's tree
    query: a work-group per row over the precomputed depth-first leaf order
    (default) or a wavefront walking the tree (FH_KNN_TREE_FORM=0) give the
    same neighbours, cores and MST."""
    rng = np.random.default_rng(35)
    X = workloads.gaussian_highdim(24000, 40, n_clusters=30, seed=35)
    X = np.vstack([X, X[rng.choice(24000, 6000, replace=False)]])
    X = X[rng.permutation(X.shape[0])]
    st = {}
    a = mst_arrays(X, 8, stats=st)
    assert st["knn_tree_rows"] > 100
    monkeypatch.setenv("FH_KNN_TREE_FORM", "0")
    b = mst_arrays(X, 8)
    for u, v in zip(a, b):
        np.testing.assert_array_equal(u, v)


@pytest.mark.parametrize("d,n", [(9, 60000), (12, 120000), (16, 250000)])
def test_mid_dimension_scans_match_tree_path(monkeypatch, d, n):
    """From 9 dimensions on the MFMA scans take inputs the tree path took
    before (up to n = 4e5 * 2.5^(d - 9); every n from 16 on): same exact
    MST as the tree path."""
    X = workloads.gaussian_highdim(n, d, n_clusters=max(3, n // 4000), seed=d + 40)
    st = {}
    a = mst_arrays(X, 10, stats=st)
    assert st["high_dim"] == 1
    monkeypatch.setenv("FH_HIGH_DIM_FROM", "1000")
    b = mst_arrays(X, 10)
    for u, v in zip(a, b):
        np.testing.assert_array_equal(u, v)


@pytest.mark.parametrize("split", ["1", "0"])
def test_zero_cores_from_duplicates(monkeypatch, split):
    """Every point has min_samples + 2 exact copies: core distances are 0,
    the mutual-reachability weights of copies are 0 and their approximate
    values can come out as -0; the packed (value, index) minima must still
    order them first (weights equal the oracle's, bit for bit)."""
    monkeypatch.setenv("FH_MFMA_SPLIT", split)
    rng = np.random.default_rng(42)
    base = workloads.gaussian_highdim(600, 48, n_clusters=6, seed=42)
    X = np.repeat(base, 7, axis=0)
    X = X[rng.permutation(X.shape[0])]
    _check(X, 5)


@pytest.mark.parametrize("form", ["lists", "fallback", "scan"])
def test_first_round_from_candidate_lists(monkeypatch, form):
    """The first round from the kNN candidate lists (default), the same
    with its listed rows handed to the MFMA scan (FH_CAND_ROUND_ROWS=0),
    and the MFMA scan alone (FH_CAND_ROUND=0): one exact MST."""
    X = workloads.gaussian_highdim(30000, 40, n_clusters=40, seed=44)
    ref = mst_arrays(X, 10) if form == "lists" else None
    if form == "fallback":
        monkeypatch.setenv("FH_CAND_ROUND_ROWS", "0")
    if form == "scan":
        monkeypatch.setenv("FH_CAND_ROUND", "0")
    got = mst_arrays(X, 10)
    monkeypatch.delenv("FH_CAND_ROUND_ROWS", raising=False)
    monkeypatch.setenv("FH_CAND_ROUND", "0")
    want = mst_arrays(X, 10)
    for u, v in zip(got, want):
        np.testing.assert_array_equal(u, v)
    if ref is not None:
        for u, v in zip(ref, want):
            np.testing.assert_array_equal(u, v)


@pytest.mark.parametrize("n,ms", [(40000, 1), (40111, 15), (33333, 16)])
def test_sampled_thresholds_and_first_round_edges(monkeypatch, n, ms):
    """Sampled kNN thresholds (from 256 tiles on, k + 1 <= 16) with the
    first round from the candidate lists at min_samples 1, at the largest
    sampled width (k + 1 = 16), one above it (full thresholds) and sizes
    that are not multiples of 128: the tree path's MST, bit for bit."""
    X = workloads.gaussian_highdim(n, 24, n_clusters=30, seed=n % 1000 + ms)
    st = {}
    a = mst_arrays(X, ms, stats=st)
    assert st["high_dim"] == 1
    monkeypatch.setenv("FH_HIGH_DIM_FROM", "1000")
    b = mst_arrays(X, ms)
    for u, v in zip(a, b):
        np.testing.assert_array_equal(u, v)


@pytest.mark.parametrize("exact_upper", ["0", "1"])
def test_round_upper_bounds_approx_or_exact(monkeypatch, exact_upper):
    """Components' upper bounds from the approximate row minima plus the
    margin (default) or from the exact weights of the row minima
    (FH_EXACT_UPPER=1): the same MST as the tree path."""
    monkeypatch.setenv("FH_EXACT_UPPER", exact_upper)
    X = workloads.gaussian_highdim(20000, 48, n_clusters=25, seed=46)
    a = mst_arrays(X, 10)
    monkeypatch.setenv("FH_HIGH_DIM_FROM", "1000")
    b = mst_arrays(X, 10)
    for u, v in zip(a, b):
        np.testing.assert_array_equal(u, v)
