"""Many independent clustering requests in one call.

``hdbscan_batch`` runs, for every dataset, what fast_hdbscan(X_i,
min_samples, min_cluster_size) returns for Euclidean data (labels and
membership strengths, hdbscan.py:199-336), through the native batch
scheduler (``fh_hdbscan_batch``): datasets of up to a few thousand points
run in one persistent launch per dimension (a work-group per dataset, taken
from a device-side queue), the others on worker threads with one context
and HIP stream each, and the label stages on the host threads.

``mst_batch`` returns the MSTs alone (``fh_mst_batch``): what
compute_minimum_spanning_tree (hdbscan.py:339-494) computes per dataset,
as squared-weight rows in the reference's order plus squared core
distances, and which datasets the persistent launch served.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib


def hdbscan_batch(datasets, min_samples=5, min_cluster_size=20, workers=8, device=0, times=None):
    """Returns a list of (labels int64[n_i], probabilities float32[n_i])."""
    Xs = [np.ascontiguousarray(X, dtype=np.float32) for X in datasets]
    count = len(Xs)
    for X in Xs:
        if X.ndim != 2:
            raise ValueError("every dataset must be a 2-D array")
    labels = [np.empty(X.shape[0], np.int64) for X in Xs]
    probs = [np.empty(X.shape[0], np.float32) for X in Xs]
    xp = (ctypes.c_void_p * max(count, 1))(*[X.ctypes.data for X in Xs])
    lp = (ctypes.c_void_p * max(count, 1))(*[a.ctypes.data for a in labels])
    pp = (ctypes.c_void_p * max(count, 1))(*[a.ctypes.data for a in probs])
    ns = np.array([X.shape[0] for X in Xs], np.int64)
    ds = np.array([X.shape[1] for X in Xs], np.int32)
    ms = np.zeros(max(count, 1), np.float64)
    code = _lib.load().fh_hdbscan_batch(device, count, ctypes.cast(xp, ctypes.c_void_p), _lib.ptr(ns),
                                        _lib.ptr(ds), int(min_samples), int(min_cluster_size), int(workers),
                                        ctypes.cast(lp, ctypes.c_void_p), ctypes.cast(pp, ctypes.c_void_p),
                                        _lib.ptr(ms))
    if code != 0:
        raise _lib.FHError(f"fh_hdbscan_batch failed with code {code}")
    if times is not None:
        times.extend(ms[:count].tolist())
    return list(zip(labels, probs))


def mst_batch(datasets, min_samples=5, device=0):
    """Returns a list of (src int32[n-1], dst int32[n-1], w_sq float32[n-1],
    core_sq float32[n], persistent bool) per dataset (n >= 1)."""
    Xs = [np.ascontiguousarray(X, dtype=np.float32) for X in datasets]
    count = len(Xs)
    for X in Xs:
        if X.ndim != 2 or X.shape[0] < 1:
            raise ValueError("every dataset must be a 2-D array with at least one row")
    m = [max(X.shape[0] - 1, 0) for X in Xs]
    src = [np.empty(max(k, 1), np.int32) for k in m]
    dst = [np.empty(max(k, 1), np.int32) for k in m]
    w = [np.empty(max(k, 1), np.float32) for k in m]
    core = [np.empty(X.shape[0], np.float32) for X in Xs]
    arr = lambda a: (ctypes.c_void_p * max(count, 1))(*[x.ctypes.data for x in a])
    ns = np.array([X.shape[0] for X in Xs], np.int64)
    ds = np.array([X.shape[1] for X in Xs], np.int32)
    pers = np.zeros(max(count, 1), np.int32)
    code = _lib.load().fh_mst_batch(device, count, ctypes.cast(arr(Xs), ctypes.c_void_p), _lib.ptr(ns),
                                    _lib.ptr(ds), int(min_samples), ctypes.cast(arr(src), ctypes.c_void_p),
                                    ctypes.cast(arr(dst), ctypes.c_void_p), ctypes.cast(arr(w), ctypes.c_void_p),
                                    ctypes.cast(arr(core), ctypes.c_void_p), _lib.ptr(pers))
    if code != 0:
        raise _lib.FHError(f"fh_mst_batch failed with code {code}")
    return [(src[i][:m[i]], dst[i][:m[i]], w[i][:m[i]], core[i], bool(pers[i])) for i in range(count)]
