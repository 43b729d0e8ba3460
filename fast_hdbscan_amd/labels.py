"""Host label stage: MST -> single linkage -> condensed tree -> clusters.

``clusters_from_spanning_tree`` mirrors hdbscan.clusters_from_spanning_tree
(hdbscan.py:497-574) for unsupervised selection, with or without sample
weights; the work runs in the library's host C++ (fh_labels_from_sorted_mst).
"""
from __future__ import annotations

import ctypes
from collections import namedtuple

import numpy as np

from . import _lib

CondensedTree = namedtuple("CondensedTree", ["parent", "child", "lambda_val", "child_size"])

_METHODS = {"eom": 0, "leaf": 1}


def label_params(min_cluster_size=10, cluster_selection_method="eom", max_cluster_size=np.inf,
                 allow_single_cluster=False, cluster_selection_epsilon=0.0) -> _lib.FHLabelParams:
    if cluster_selection_method not in _METHODS:
        raise ValueError(f"Invalid cluster_selection_method {cluster_selection_method}")
    return _lib.FHLabelParams(int(min_cluster_size), float(max_cluster_size), float(cluster_selection_epsilon),
                              _METHODS[cluster_selection_method], 1 if allow_single_cluster else 0)


def labels_from_sorted_arrays(n, src, dst, w_sq, params, return_trees=False, phases: dict | None = None,
                              dist=None, sample_weights=None, point_order=None, groups=None, threads=0, out=None):
    """Run the host label stage on MST rows already sorted by (weight,
    source, target): squared float32 weights w_sq, or float64 distances
    dist (then w_sq may be None); float32 sample weights optional.  With
    ``point_order`` (int32 [n]) src / dst are positions p in that order and
    point_order[p] is the point id.  ``groups`` = (k, row_order int32 [k],
    group_sizes int32 [G]) as fh_linkage_plan gives them lets the linkage
    of rows [0, k) run on ``threads`` host threads.  Neither changes the
    outputs.  ``out`` = (labels int64 [>= n], strengths float32 [>= n]) to
    fill instead of new arrays.  Returns (labels int64, strengths float32,
    linkage or None, CondensedTree or None)."""
    src = np.ascontiguousarray(src, dtype=np.int32)
    dst = np.ascontiguousarray(dst, dtype=np.int32)
    w_sq = None if w_sq is None else np.ascontiguousarray(w_sq, dtype=np.float32)
    dist = None if dist is None else np.ascontiguousarray(dist, dtype=np.float64)
    sw = None if sample_weights is None else np.ascontiguousarray(sample_weights, dtype=np.float32)
    po = None if point_order is None else np.ascontiguousarray(point_order, dtype=np.int32)
    order = None
    if po is not None or groups is not None:
        order = _lib.FHLabelOrder(_lib.ptr(po) if po is not None else None, 0, 0, None, None, int(threads), 0)
        if groups is not None:
            k, ro, gs = groups
            ro = np.ascontiguousarray(ro, dtype=np.int32)
            gs = np.ascontiguousarray(gs, dtype=np.int32)
            order.prefix_rows, order.n_groups = int(k), int(gs.shape[0])
            order.row_order, order.group_sizes = _lib.ptr(ro), _lib.ptr(gs)
    if out is not None:
        labels, strengths = out[0][:n], out[1][:n]
        if labels.dtype != np.int64 or strengths.dtype != np.float32 or labels.shape[0] < n or \
                strengths.shape[0] < n or not (labels.flags.c_contiguous and strengths.flags.c_contiguous):
            raise ValueError("out must be contiguous (int64 [n], float32 [n]) arrays")
    else:
        labels = np.empty(n, np.int64)
        strengths = np.empty(n, np.float32)
    linkage = np.empty((max(n - 1, 0), 4), np.float64) if return_trees else None
    ct = None
    ctree = None
    if return_trees:
        cap = max(2 * (n - 1), 1)
        ct = {"parent": np.empty(cap, np.int64), "child": np.empty(cap, np.int64),
              "lambda_val": np.empty(cap, np.float32), "child_size": np.empty(cap, np.float32)}
        ctree = _lib.FHCondensedTree(ct["parent"].ctypes.data, ct["child"].ctypes.data, ct["lambda_val"].ctypes.data,
                                     ct["child_size"].ctypes.data, cap, 0)
    ph = np.zeros(4, np.float64)
    code = _lib.load().fh_labels_from_sorted_mst(
        n, _lib.ptr(src), _lib.ptr(dst), _lib.ptr(w_sq) if w_sq is not None else None,
        _lib.ptr(dist) if dist is not None else None, _lib.ptr(sw) if sw is not None else None,
        ctypes.byref(params), _lib.ptr(labels), _lib.ptr(strengths),
        _lib.ptr(linkage) if linkage is not None else None, ctypes.byref(ctree) if ctree is not None else None,
        _lib.ptr(ph), ctypes.byref(order) if order is not None else None)
    if code == _lib.FH_EINVAL:
        raise ValueError("invalid label-stage arguments")
    if code == _lib.FH_EUNSUPPORTED:
        raise NotImplementedError("sample weights with a -1 MST endpoint are not supported")
    if code != _lib.FH_OK:
        raise _lib.FHError(f"label stage failed with code {code}")
    if phases is not None:
        phases.update(dict(zip(("t_linkage_ms", "t_condense_ms", "t_select_ms", "t_label_pass_ms"), ph.tolist())))
    condensed = None
    if ctree is not None:
        r = ctree.rows
        condensed = CondensedTree(ct["parent"][:r], ct["child"][:r], ct["lambda_val"][:r], ct["child_size"][:r])
    return labels, strengths, linkage, condensed


def clusters_from_spanning_tree(
    minimum_spanning_tree,
    data_labels=None,
    semi_supervised=False,
    ss_algorithm="bc",
    min_cluster_size=10,
    cluster_selection_method="eom",
    max_cluster_size=np.inf,
    allow_single_cluster=False,
    cluster_selection_epsilon=0.0,
    cluster_selection_persistence=0.0,
    sample_weights=None,
):
    """Drop-in for hdbscan.clusters_from_spanning_tree (hdbscan.py:497-574):
    returns (labels, membership_strengths, linkage_tree, condensed_tree,
    sorted_mst)."""
    if semi_supervised:
        raise NotImplementedError("semi-supervised selection is not in the host label stage")
    if cluster_selection_persistence > 0.0:
        raise NotImplementedError("cluster_selection_persistence is not in the host label stage")
    mst = np.asarray(minimum_spanning_tree, dtype=np.float64)
    n = mst.shape[0] + 1
    # hdbscan.py:511-519
    sorted_mst = mst[np.lexsort((mst.T[1], mst.T[0], mst.T[2]))]
    params = label_params(min_cluster_size, cluster_selection_method, max_cluster_size, allow_single_cluster,
                          cluster_selection_epsilon)
    sw = None
    if sample_weights is not None:
        sw = np.asarray(sample_weights)
        if sw.shape != (n,):
            raise ValueError("sample_weights must have one value per point")
        if sw.dtype != np.float32 and not np.array_equal(sw.astype(np.float32).astype(sw.dtype), sw):
            raise NotImplementedError("sample weights must be float32-representable")
    labels, strengths, linkage, condensed = labels_from_sorted_arrays(
        n, sorted_mst[:, 0].astype(np.int32), sorted_mst[:, 1].astype(np.int32), None, params, return_trees=True,
        dist=np.ascontiguousarray(sorted_mst[:, 2]), sample_weights=sw)
    return labels, strengths, linkage, condensed, sorted_mst
