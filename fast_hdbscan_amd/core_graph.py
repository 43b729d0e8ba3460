"""Core graphs and Boruvka on CSR graphs.

Mirrors core_graph.py of the reference (CoreGraph, knn_mst_union +
flatten_to_csr + sort_by_lens as ``core_graph_spanning_tree``'s graph,
``boruvka_mst``, ``core_graph_clusters``, the rec-array / edge-list views):
the graph is built by the host library (``fh_core_graph``), the Boruvka
rounds run on the device (``fh_csr_boruvka``).
"""
from __future__ import annotations

import ctypes
from collections import namedtuple

import numpy as np

from . import _lib

CoreGraph = namedtuple("CoreGraph", ["weights", "distances", "indices", "indptr"])


def build_core_graph(neighbors, core_distances, min_spanning_tree, lens_values) -> CoreGraph:
    """sort_by_lens(flatten_to_csr(knn_mst_union(...))) (core_graph.py:13-104)."""
    nbr = np.ascontiguousarray(neighbors, dtype=np.int32)
    n = nbr.shape[0]
    k = nbr.shape[1] if nbr.ndim == 2 else 0
    core = np.ascontiguousarray(core_distances, dtype=np.float32)
    mst = np.ascontiguousarray(min_spanning_tree, dtype=np.float64).reshape(-1, 3)
    lens = np.ascontiguousarray(lens_values, dtype=np.float64)
    cap = max(n * k + mst.shape[0], 1)
    indptr = np.empty(n + 1, np.int32)
    indices = np.empty(cap, np.int32)
    weights = np.empty(cap, np.float32)
    distances = np.empty(cap, np.float32)
    nnz = _lib.load().fh_core_graph(n, k, _lib.ptr(nbr), _lib.ptr(core), mst.shape[0], _lib.ptr(mst), _lib.ptr(lens),
                                    _lib.ptr(indptr), _lib.ptr(indices), _lib.ptr(weights), _lib.ptr(distances), cap)
    if nnz < 0:
        raise _lib.FHError(f"fh_core_graph failed with code {nnz}")
    return CoreGraph(weights[:nnz].copy(), distances[:nnz].copy(), indices[:nnz].copy(), indptr)


def boruvka_mst(graph: CoreGraph, overwrite=False, device: int = 0):
    """core_graph.boruvka_mst (core_graph.py:172-208): returns (n_components,
    point_components int64[n], edges float64[m, 3]).  The input graph is
    never modified (the device works on its own copy)."""
    indptr = np.ascontiguousarray(graph.indptr, dtype=np.int32)
    indices = np.ascontiguousarray(graph.indices, dtype=np.int32)
    weights = np.ascontiguousarray(graph.weights, dtype=np.float32)
    n = indptr.shape[0] - 1
    edges = np.empty((max(n - 1, 1), 3), np.float64)
    ne = ctypes.c_int64(0)
    nc = ctypes.c_int64(0)
    comps = np.empty(n, np.int64)
    ctx = _lib.context(device)
    code = _lib.load().fh_csr_boruvka(ctx, n, _lib.ptr(indptr), _lib.ptr(indices), _lib.ptr(weights),
                                      _lib.ptr(edges), ctypes.byref(ne), _lib.ptr(comps), ctypes.byref(nc))
    _lib.check(ctx, code)
    return int(nc.value), comps, edges[: ne.value].copy()


def core_graph_spanning_tree(neighbors, core_distances, min_spanning_tree, lens):
    """core_graph.py:211-218"""
    graph = build_core_graph(neighbors, core_distances, min_spanning_tree, lens)
    return (*boruvka_mst(graph), graph)


def core_graph_clusters(lens, neighbors, core_distances, min_spanning_tree, **kwargs):
    """core_graph.py:221-249"""
    from .labels import clusters_from_spanning_tree

    num_components, component_labels, lensed_mst, graph = core_graph_spanning_tree(
        neighbors, core_distances, min_spanning_tree, lens)
    if num_components > 1:
        for i, label in enumerate(np.unique(component_labels)):
            component_labels[component_labels == label] = i
        empty = np.empty(0, dtype=np.intp), np.empty(0, dtype=np.float32)
        return (component_labels, np.ones(len(component_labels), dtype=np.float32), np.empty((0, 4)),
                (empty[0], empty[0], empty[1], empty[1]), lensed_mst, graph)
    return (*clusters_from_spanning_tree(lensed_mst, **kwargs), graph)


def core_graph_to_rec_array(graph):
    """core_graph.py:252-266"""
    result = np.empty(graph.indptr[-1], dtype=[("parent", np.int32), ("child", np.int32), ("weight", np.float32),
                                               ("distance", np.float32)])
    result["parent"] = np.repeat(np.arange(len(graph.indptr) - 1), np.diff(graph.indptr))
    result["child"] = graph.indices
    result["weight"] = graph.weights
    result["distance"] = graph.distances
    return result


def core_graph_to_edge_list(graph):
    """core_graph.py:269-274"""
    result = np.empty((graph.indptr[-1], 4), dtype=np.float64)
    result[:, 0] = np.repeat(np.arange(len(graph.indptr) - 1), np.diff(graph.indptr))
    result[:, 1] = graph.indices
    result[:, 2] = graph.weights
    result[:, 3] = graph.distances
    return result
