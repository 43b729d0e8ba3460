"""kNN-graph Kruskal on the device against the oracle, bit for bit."""
import numpy as np
import pytest
from sklearn.datasets import make_blobs

from fast_hdbscan_amd import workloads

pytestmark = pytest.mark.gpu


def _cmp(X, ms, kk, w=None):
    from fast_hdbscan_amd.hdbscan import kruskal_mst_knn
    from oracle import kruskal as K

    print(f"kruskal case n={X.shape[0]} d={X.shape[1]} ms={ms} knn_k={kk} weighted={w is not None}", flush=True)
    m, nb, core = kruskal_mst_knn(X, ms, kk, sample_weights=w)
    rm, rnb, rcore = K.kruskal_mst_knn(X, ms, kk, sample_weights=w)
    np.testing.assert_array_equal(core, rcore)
    np.testing.assert_array_equal(nb, rnb)
    np.testing.assert_array_equal(m, rm)


@pytest.mark.parametrize("ms,kk", [(10, 10), (5, 15), (1, 8), (20, 5)])
def test_kruskal_knn_blobs(ms, kk):
    X = workloads.anisotropic_mixture(20000, 2, 15, 0.1, seed=ms + kk)
    _cmp(X, ms, kk)


def test_kruskal_knn_weighted():
    X = workloads.blobs_small(8000, seed=3)
    w = np.random.default_rng(2).uniform(0.5, 2.0, 8000).astype(np.float32)
    _cmp(X, 6, 10, w)


def test_kruskal_knn_disconnected_and_ties():
    rng = np.random.default_rng(7)
    g = rng.integers(0, 30, size=(4000, 2)).astype(np.float32)
    far = g + 10000.0
    _cmp(np.vstack([g, far]), 4, 6)


def test_kruskal_knn_high_dim():
    X = workloads.gaussian_highdim(6000, 48, n_clusters=15, seed=9)
    _cmp(X, 10, 12)


def test_kruskal_seam_dispatch():
    from fast_hdbscan_amd import compute_minimum_spanning_tree
    from oracle import kruskal as K

    X, _ = make_blobs(3000, random_state=0)
    m, nb, core = compute_minimum_spanning_tree(X, min_samples=5, algorithm="kruskal", knn_k=10)
    rm, rnb, rcore = K.kruskal_mst_knn(X, 5, 10)
    np.testing.assert_array_equal(m, rm)
    assert core.dtype == np.float64 and nb.dtype == np.int32
    with pytest.raises(NotImplementedError):
        compute_minimum_spanning_tree(X, min_samples=5, algorithm="kruskal")
