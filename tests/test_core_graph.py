"""Core graphs: the reference's core-graph tests, the host builder against
the oracle, and oracle pins for CSR Boruvka (CPU)."""
import numpy as np
import pytest
from scipy.sparse import csr_matrix
from scipy.sparse.csgraph import connected_components, minimum_spanning_tree
from sklearn.datasets import make_blobs

from oracle import core_graph as OC
from oracle import oracle as O


def _inputs(seed=10, n=50):
    rs = np.random.RandomState(seed)
    X, y = make_blobs(n_samples=n, random_state=seed)
    lens = rs.normal(size=y.shape)
    mst, indices, core = O.compute_minimum_spanning_tree(X, 5)
    return mst, indices, core, lens


def test_reference_core_graph_properties():
    # test_core_graph.py:36-70 on this package's builder
    from fast_hdbscan_amd.core_graph import build_core_graph, core_graph_to_edge_list, core_graph_to_rec_array

    mst, indices, core, lens = _inputs()
    csr = build_core_graph(indices, core, mst, lens)
    assert np.all(csr.distances > 0)
    for idx in range(len(csr.indptr) - 1):
        s, e = csr.indptr[idx], csr.indptr[idx + 1]
        assert np.all(np.diff(csr.weights[s:e]) >= 0)
        assert np.allclose(np.maximum(lens[idx], lens[csr.indices[s:e]]), csr.weights[s:e])
    rec = core_graph_to_rec_array(csr)
    assert np.all(rec["parent"] == np.repeat(np.arange(len(csr.indptr) - 1), np.diff(csr.indptr)))
    assert np.all(rec["child"] == csr.indices) and np.all(rec["weight"] == csr.weights)
    el = core_graph_to_edge_list(csr)
    assert np.all(el[:, 1] == csr.indices) and np.all(el[:, 3] == csr.distances)


@pytest.mark.parametrize("seed,n", [(10, 50), (3, 2000), (4, 5000)])
def test_builder_matches_oracle(seed, n):
    from fast_hdbscan_amd.core_graph import build_core_graph

    mst, indices, core, lens = _inputs(seed, n)
    lens = np.round(lens, 1)  # ties in the row sorts
    g = build_core_graph(indices, core, mst, lens)
    w, d, idx, indptr = OC.flatten_and_sort(OC.knn_mst_union(indices, core, mst, lens))
    np.testing.assert_array_equal(g.indptr, indptr)
    np.testing.assert_array_equal(g.indices, idx)
    np.testing.assert_array_equal(g.weights, w)
    np.testing.assert_array_equal(g.distances, d)


def random_symmetric_csr(seed, n=600, m=2400, ties=True):
    """Undirected random graph as CSR with each edge in both rows, rows
    sorted by weight with numba's argsort (as sort_by_lens does)."""
    rng = np.random.default_rng(seed)
    u = rng.integers(0, n, m)
    v = rng.integers(0, n, m)
    keep = u != v
    u, v = u[keep], v[keep]
    w = rng.normal(size=u.shape[0])
    if ties:
        w = np.round(w, 1)
    und = {}
    for a, b, x in zip(u, v, w):
        key = (min(a, b), max(a, b))
        if key not in und:
            und[key] = np.float32(x)
    deg = np.zeros(n, int)
    for a, b in und:
        deg[a] += 1
        deg[b] += 1
    for a in np.flatnonzero(deg == 0):  # no empty rows (see the quirk test below)
        b = (a + 1) % n
        und[(min(a, b), max(a, b))] = np.float32(0.5)
    rows = [[] for _ in range(n)]
    for (a, b), x in und.items():
        rows[a].append((b, x))
        rows[b].append((a, x))
    indptr = np.zeros(n + 1, np.int32)
    idx, ws = [], []
    for a in range(n):
        r = rows[a]
        rw = np.array([x for _, x in r], np.float32)
        order = OC.numba_argsort(rw, is_float=True) if len(r) else []
        idx.extend(r[o][0] for o in order)
        ws.extend(rw[o] for o in order)
        indptr[a + 1] = indptr[a] + len(r)
    return np.array(ws, np.float32), np.array(idx, np.int32), indptr, und


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_oracle_boruvka_matches_scipy_on_symmetric_graphs(seed):
    w, idx, indptr, und = random_symmetric_csr(seed)
    n = len(indptr) - 1
    nc, comps, edges = OC.boruvka_mst(w, idx, indptr)
    shift = 10.0
    a = np.array([k[0] for k in und]); b = np.array([k[1] for k in und])
    x = np.array([float(v) for v in und.values()]) + shift
    G = csr_matrix((x, (a, b)), shape=(n, n))
    T = minimum_spanning_tree(G)
    np.testing.assert_allclose(np.sort(edges[:, 2] + shift), np.sort(T.data), rtol=0, atol=1e-9)
    assert nc == connected_components(G, directed=False)[0]
    # labels: one root per component, shared by exactly its points
    cc = connected_components(G, directed=False)[1]
    for c in np.unique(cc):
        assert len(np.unique(comps[cc == c])) == 1


def test_empty_row_reads_the_next_row():
    # select_components (core_graph.py:119-121) reads indices[indptr[p]] for
    # an empty row p, i.e. the next row's first child
    indptr = np.array([0, 1, 1, 3, 4], np.int32)  # row 1 is empty
    idx = np.array([2, 0, 3, 2], np.int32)
    w = np.array([1.0, 1.0, 2.0, 2.0], np.float32)
    nc, comps, edges = OC.boruvka_mst(w, idx, indptr)
    assert (1.0, 0.0, 1.0) in [tuple(e) for e in edges]
