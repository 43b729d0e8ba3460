"""Oracle for the kNN-graph Kruskal MST -- TEST INFRASTRUCTURE ONLY.

Restates kruskal._kruskal_mst_knn (kruskal.py:950-1034) in plain Python +
numpy with the reference's dtypes: the oracle's KD-tree kNN (k_query
columns), float64 square roots, mutual-reachability weights
max(core_i, core_j, d) (_build_mrd_edges_knn, 455-512), the filter and
stable argsort (1017-1023), _kruskal_core's union by rank with full path
compression (31-108) and _finish_mst's +inf bridges between the smallest
ids of the components (723-752, precomputed.py:308-356).

Pinned by (tests/test_kruskal.py): with knn_k >= n-1 the kNN graph is
complete, so the weight multiset must equal the exact Boruvka MST's
(sqrt is monotone, so max and sqrt commute); on smaller knn_k the weight
multiset equals scipy.sparse.csgraph.minimum_spanning_tree on the same
graph; disconnected graphs get exactly (components - 1) +inf rows; the
reference's own shape tests (test_kruskal.py:120-145).
"""
from __future__ import annotations

import numpy as np

from . import oracle as O


def knn_graph(data, min_samples, knn_k, sample_weights=None):
    """kruskal.py:983-1011: (knn_dists f64, knn_indices i32, core f64,
    neighbors_out)."""
    T = O.build_kdtree(data)
    X = T.data
    if sample_weights is not None:
        mean_sw = float(np.mean(sample_weights))
        k_for_core = int(2 * min_samples / mean_sw)
        kq = max(knn_k, k_for_core) + 1
        d_raw, nb_raw = O.tree_query_sqrt(T, X, kq)
        core = O.sample_weight_core_distance(d_raw, nb_raw, np.asarray(sample_weights), min_samples).astype(
            np.float64)
        return d_raw[:, 1:].astype(np.float64), nb_raw[:, 1:].astype(np.int32), core, nb_raw[:, 1:]
    if min_samples > 1:
        kq = max(knn_k, min_samples) + 1
        rd, nb_raw = O.tree_query(T, X, kq)
        dist = np.sqrt(rd.astype(np.float64))
        return dist[:, 1:], nb_raw[:, 1:].astype(np.int32), dist[:, min_samples].copy(), nb_raw[:, 1:min_samples + 1]
    kq = knn_k + 1
    rd, nb_raw = O.tree_query(T, X, kq)
    dist = np.sqrt(rd.astype(np.float64))
    return dist[:, 1:], nb_raw[:, 1:].astype(np.int32), np.zeros(X.shape[0]), nb_raw[:, 1:2]


def mrd_edges(knn_i, knn_d, core):
    """_build_mrd_edges_knn (kruskal.py:455-512), vectorised."""
    n, k = knn_i.shape
    rows = np.repeat(np.arange(n, dtype=np.int32), k)
    cols = knn_i.ravel().astype(np.int32)
    d = knn_d.ravel()
    bad = cols < 0
    safe = np.where(bad, rows, cols)
    w = np.maximum(np.maximum(core[rows], core[safe]), d)
    w[bad] = np.inf
    cols = np.where(bad, rows, cols)
    return w, rows, cols


def kruskal_core(weights, rows, cols, order, n):
    """_kruskal_core (kruskal.py:31-108)."""
    pred = list(range(n))
    rank = [0] * n
    out = []
    for j in order:
        if len(out) == n - 1:
            break
        u, v = int(rows[j]), int(cols[j])
        if u == v:
            continue
        ru = u
        while pred[ru] != ru:
            ru = pred[ru]
        c = u
        while c != ru:
            nx = pred[c]
            pred[c] = ru
            c = nx
        rv = v
        while pred[rv] != rv:
            rv = pred[rv]
        c = v
        while c != rv:
            nx = pred[c]
            pred[c] = rv
            c = nx
        if ru != rv:
            out.append((float(u), float(v), float(weights[j])))
            if rank[ru] > rank[rv]:
                pred[rv] = ru
            elif rank[ru] < rank[rv]:
                pred[ru] = rv
            else:
                pred[rv] = ru
                rank[ru] += 1
    return np.asarray(out, dtype=np.float64).reshape(-1, 3), pred


def finish(mst, pred, n):
    """_finish_mst + bridge_forest_with_inf."""
    labels = []
    for i in range(n):
        r = pred[i]
        while pred[r] != r:
            r = pred[r]
        labels.append(r)
    labels = np.asarray(labels)
    comps = np.unique(labels)
    if len(comps) == 1:
        return mst
    reps = sorted(int(np.where(labels == c)[0].min()) for c in comps)
    bridges = np.array([(float(reps[i]), float(reps[i + 1]), np.inf) for i in range(len(reps) - 1)])
    if mst.shape[0] == 0:
        return bridges
    return np.vstack([mst, bridges])


def kruskal_mst_knn(data, min_samples, knn_k, sample_weights=None):
    """kruskal.py:950-1034 without cannot-link constraints."""
    knn_d, knn_i, core, nb_out = knn_graph(data, min_samples, knn_k, sample_weights)
    n = knn_i.shape[0]
    w, rows, cols = mrd_edges(knn_i, knn_d, core)
    valid = (rows != cols) & np.isfinite(w)
    w, rows, cols = w[valid], rows[valid], cols[valid]
    order = np.argsort(w, kind="stable")
    mst, pred = kruskal_core(w, rows, cols, order, n)
    return finish(mst, pred, n), nb_out, core
