"""kNN-graph Kruskal oracle pins (CPU)."""
import numpy as np
import pytest
from scipy.sparse import coo_matrix
from scipy.sparse.csgraph import connected_components, minimum_spanning_tree
from sklearn.datasets import make_blobs

from oracle import kruskal as K
from oracle import oracle as O


def test_complete_knn_graph_matches_exact_mst_weights():
    # knn_k >= n - 1: every pair is an edge, so the weight multiset is the
    # exact mutual-reachability MST's
    X, _ = make_blobs(300, centers=4, random_state=3)
    for ms in (1, 4, 10):
        mst, nb, core = K.kruskal_mst_knn(X, ms, 299)
        ref = O.compute_minimum_spanning_tree(X, ms)[0]
        np.testing.assert_array_equal(np.sort(mst[:, 2]), np.sort(ref[:, 2]))
        np.testing.assert_array_equal(core, np.sqrt(O.boruvka(O.build_kdtree(X), ms).core_sq.astype(np.float64)))


@pytest.mark.parametrize("ms,kk", [(5, 8), (10, 10), (1, 6), (3, 15)])
def test_weights_match_scipy_on_the_same_graph(ms, kk):
    X, _ = make_blobs(800, centers=6, cluster_std=1.5, random_state=ms)
    mst, nb, core = K.kruskal_mst_knn(X, ms, kk)
    knn_d, knn_i, core2, _ = K.knn_graph(X, ms, kk)
    w, r, c = K.mrd_edges(knn_i, knn_d, core2)
    keep = (r != c) & np.isfinite(w) & (w > 0)
    G = coo_matrix((w[keep], (r[keep], c[keep])), shape=(800, 800)).tocsr()
    G = G.maximum(G.T)
    T = minimum_spanning_tree(G)
    ncomp = connected_components(G, directed=False)[0]
    finite = mst[np.isfinite(mst[:, 2]), 2]
    np.testing.assert_array_equal(np.sort(finite), np.sort(T.data))
    assert np.isinf(mst[:, 2]).sum() == ncomp - 1
    assert mst.shape == (799, 3)


def test_disconnected_graph_bridges():
    # two far clusters, knn_k small: one +inf bridge between the smallest ids
    rng = np.random.default_rng(0)
    X = np.vstack([rng.normal(0, 1, (100, 2)), rng.normal(1000, 1, (100, 2))])
    mst, nb, core = K.kruskal_mst_knn(X, 3, 5)
    inf_rows = mst[np.isinf(mst[:, 2])]
    assert inf_rows.shape[0] == 1
    assert inf_rows[0, 0] == 0.0 and inf_rows[0, 1] == 100.0


def test_weighted_branch_shapes():
    # test_kruskal.py:120-145 style: shapes with sample weights
    X, _ = make_blobs(400, random_state=2)
    w = np.random.default_rng(1).uniform(0.5, 2.0, 400).astype(np.float32)
    mst, nb, core = K.kruskal_mst_knn(X, 5, 10, sample_weights=w)
    assert mst.shape == (399, 3) and nb.shape[0] == 400 and core.shape == (400,)
