"""CSR Boruvka on the device against the oracle, bit for bit."""
import numpy as np
import pytest

from oracle import core_graph as OC
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _cmp(w, idx, indptr):
    from fast_hdbscan_amd.core_graph import CoreGraph, boruvka_mst

    g = CoreGraph(w, np.zeros_like(w), idx, indptr)
    nc, comps, edges = boruvka_mst(g)
    rnc, rcomps, redges = OC.boruvka_mst(w, idx, indptr)
    assert nc == rnc
    np.testing.assert_array_equal(comps, rcomps)
    np.testing.assert_array_equal(edges, redges)
    return nc


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_symmetric_random_graphs(seed):
    from test_core_graph import random_symmetric_csr

    w, idx, indptr, _ = random_symmetric_csr(seed, n=3000, m=9000)
    print(f"csr case seed={seed}", flush=True)
    _cmp(w, idx, indptr)


def test_empty_row_quirk():
    indptr = np.array([0, 1, 1, 3, 4], np.int32)
    idx = np.array([2, 0, 3, 2], np.int32)
    w = np.array([1.0, 1.0, 2.0, 2.0], np.float32)
    _cmp(w, idx, indptr)


@pytest.mark.parametrize("seed,n,ms", [(10, 50, 5), (3, 20000, 5), (7, 8000, 10)])
def test_core_graph_spanning_tree(seed, n, ms):
    from sklearn.datasets import make_blobs

    from fast_hdbscan_amd.core_graph import build_core_graph, core_graph_spanning_tree

    X, y = make_blobs(n_samples=n, random_state=seed)
    lens = np.random.RandomState(seed).normal(size=n)
    mst, indices, core = O.compute_minimum_spanning_tree(X, ms)
    g = build_core_graph(indices, core, mst, lens)
    nc, comps, edges, graph = core_graph_spanning_tree(indices, core, mst, lens)
    rnc, rcomps, redges = OC.boruvka_mst(g.weights, g.indices, g.indptr)
    assert nc == rnc
    np.testing.assert_array_equal(comps, rcomps)
    np.testing.assert_array_equal(edges, redges)


def test_disconnected_pieces():
    from test_core_graph import random_symmetric_csr

    w1, i1, p1, _ = random_symmetric_csr(5, n=500, m=600)
    w2, i2, p2, _ = random_symmetric_csr(6, n=500, m=600)
    w = np.concatenate([w1, w2])
    idx = np.concatenate([i1, i2 + 500]).astype(np.int32)
    indptr = np.concatenate([p1, p2[1:] + p1[-1]]).astype(np.int32)
    assert _cmp(w, idx, indptr) >= 2
