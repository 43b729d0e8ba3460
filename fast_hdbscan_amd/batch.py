"""Many independent clustering requests in one call.

``hdbscan_batch`` runs, for every dataset, what fast_hdbscan(X_i,
min_samples, min_cluster_size) returns for Euclidean data (labels and
membership strengths, hdbscan.py:199-336), through the native batch
scheduler (``fh_hdbscan_batch``): worker threads with one context and HIP
stream each, so the datasets' MST kernels overlap on the GPU and their
label stages on the host.
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
