"""Oracle for core graphs and CSR Boruvka -- TEST INFRASTRUCTURE ONLY.

Restates core_graph.py in plain Python: knn_mst_union (13-45) with
insertion-ordered dicts, flatten_to_csr (48-71), sort_by_lens (74-91) with
numba's quicksort (oracle.layers.numba_argsort), and boruvka_mst (172-208)
with select_components / merge_components / update_point_components /
update_graph_components and the union-by-rank disjoint set
(disjoint_set.py:11-42) -- one thread, point order.

Pinned by (tests/test_core_graph.py): the reference's own core-graph tests
(test_core_graph.py:36-70: distances > 0, rows sorted by weight, weights =
max of the two lens values, rec-array / edge-list views); the MST weight
multiset equals scipy.sparse.csgraph.minimum_spanning_tree on the same
(symmetrised) graph; the component count equals scipy's
connected_components on disconnected graphs.
"""
from __future__ import annotations

import numpy as np

from .layers import numba_argsort


def knn_mst_union(neighbors, core_distances, mst, lens):
    graph = [dict() for _ in range(neighbors.shape[0])]
    for p in range(len(core_distances)):
        ch = graph[p]
        pl = lens[p]
        pd = core_distances[p]
        for c in neighbors[p]:
            c = int(c)
            if c < 0:
                continue
            ch[c] = (max(pl, lens[c]), max(pd, core_distances[c]))
    for pa, c, dist in mst:
        pa, c = int(np.int32(pa)), int(np.int32(c))
        if c in graph[pa]:
            continue
        graph[pa][c] = (max(lens[pa], lens[c]), dist)
    return graph


def flatten_and_sort(graph):
    n = len(graph)
    indptr = np.zeros(n + 1, np.int32)
    for i, ch in enumerate(graph):
        indptr[i + 1] = indptr[i] + len(ch)
    w = np.empty(indptr[-1], np.float32)
    d = np.empty(indptr[-1], np.float32)
    idx = np.empty(indptr[-1], np.int32)
    for p, ch in enumerate(graph):
        s = indptr[p]
        for j, (c, (ww, dd)) in enumerate(ch.items()):
            w[s + j] = ww
            d[s + j] = dd
            idx[s + j] = c
    for p in range(n):
        s, e = indptr[p], indptr[p + 1]
        order = numba_argsort(w[s:e], is_float=True)
        w[s:e], d[s:e], idx[s:e] = w[s:e][order], d[s:e][order], idx[s:e][order]
    return w, d, idx, indptr


def boruvka_mst(weights, indices, indptr):
    distances = weights.copy()
    indices = indices.copy()
    n = len(indptr) - 1
    parent = list(range(n))
    rank = [0] * n

    def find(x):
        while parent[x] != x:
            x = parent[x]
        return x

    def union(x, y):
        x, y = find(x), find(y)
        if x == y:
            return
        if rank[x] < rank[y]:
            x, y = y, x
        parent[y] = x
        if rank[x] == rank[y]:
            rank[x] += 1

    comps = np.arange(n)
    n_comp = n
    edges = []
    nnz = len(indices)
    while n_comp > 1:
        best = {}
        for p, fc in enumerate(comps):
            s = indptr[p]
            if s == nnz or indices[s] == -1:
                continue
            nb, dist = int(indices[s]), distances[s]
            if fc in best:
                if dist < best[fc][2]:
                    best[fc] = (p, nb, dist)
            else:
                best[fc] = (p, nb, dist)
        new = []
        for u, v, w in best.values():
            a, b = find(u), find(v)
            if a != b:
                new.append((float(u), float(v), float(w)))
                union(a, b)
        if not new:
            break
        edges.extend(new)
        comps = np.array([find(i) for i in range(n)])
        for p in range(n):
            cnt = 0
            s, e = indptr[p], indptr[p + 1]
            for i in range(s, e):
                c = indices[i]
                if c == -1:
                    break
                if comps[c] != comps[p]:
                    indices[s + cnt] = indices[i]
                    distances[s + cnt] = distances[i]
                    cnt += 1
            indices[s + cnt:e] = -1
            distances[s + cnt:e] = np.inf
        n_comp -= len(new)
    return n_comp, comps, np.asarray(edges, dtype=np.float64).reshape(-1, 3)
