"""Host-side mirror of fast_hdbscan's MST seam, backed by the gfx950 library.

``compute_minimum_spanning_tree`` has the signature, argument checks and
return contract of ``fast_hdbscan.hdbscan.compute_minimum_spanning_tree``
(hdbscan.py:339-494) for the Euclidean Boruvka branch, so rebinding the
reference's module attribute routes both ``HDBSCAN`` and ``PLSCAN`` through
it (``layer_clusters.py:371,381`` imports it at call time).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib

__all__ = [
    "compute_minimum_spanning_tree",
    "mst_arrays",
    "mst_device",
    "parallel_tree_query",
    "build_kdtree",
    "kruskal_mst_knn",
]


def _as_points(data) -> np.ndarray:
    # numba_kdtree.py:481: the tree works on a float32 copy of the data.
    X = np.ascontiguousarray(np.asarray(data).astype(np.float32))
    if X.ndim != 2:
        raise ValueError("data must be a 2-D array")
    return X


def _neighbor_columns(n, ms, weights_host):
    k = _lib.load().fh_neighbor_columns(n, ms, _lib.ptr(weights_host) if weights_host is not None else None)
    if k < 0:
        raise ValueError("invalid min_samples / sample weights")
    return int(k)


def mst_arrays(X: np.ndarray, min_samples: int = 10, device: int = 0, stats: dict | None = None,
               sample_weights=None):
    """Raw library outputs for host float32 points.

    Returns (src int32[n-1], dst int32[n-1], w_sq float32[n-1],
    core_sq float32[n], neighbors int32[n, k]).
    """
    X = np.ascontiguousarray(X, dtype=np.float32)
    n, d = X.shape
    ms = int(min_samples)
    w = None if sample_weights is None else np.ascontiguousarray(sample_weights, dtype=np.float32)
    if w is not None and w.shape != (n,):
        raise ValueError("sample_weights must have one value per point")
    kn = _neighbor_columns(n, ms, w)
    m = max(n - 1, 0)
    src = np.empty(max(m, 1), np.int32)
    dst = np.empty(max(m, 1), np.int32)
    wsq = np.empty(max(m, 1), np.float32)
    core = np.empty(n, np.float32)
    nbr = np.empty((n, max(kn, 0)), np.int32)
    st = _lib.FHStats()
    ctx = _lib.context(device)
    code = _lib.load().fh_mst_euclidean(
        ctx, _lib.ptr(X), n, d, ms, _lib.ptr(w) if w is not None else None, _lib.ptr(core),
        _lib.ptr(nbr) if kn > 0 else None, _lib.ptr(src), _lib.ptr(dst), _lib.ptr(wsq), ctypes.byref(st),
    )
    _lib.check(ctx, code)
    if stats is not None:
        stats.update(st.as_dict())
    return src[:m], dst[:m], wsq[:m], core, nbr


def mst_device(X, min_samples: int = 10, stats: dict | None = None, out: dict | None = None,
               sample_weights=None, core_and_neighbors: bool = True):
    """MST of a float32 torch tensor resident on a HIP device.

    Returns torch tensors on the same device: (src, dst, w_sq, core_sq,
    neighbors).  ``out`` may hold preallocated tensors under those names;
    the tensors allocated here are stored in it (and reused by a later call
    when large enough).  ``sample_weights`` (float32 tensor on the same
    device) selects the weighted-core branch.  With ``core_and_neighbors``
    False the library writes neither and both come back as None.
    """
    import torch

    if not (isinstance(X, torch.Tensor) and X.is_cuda):
        raise TypeError("mst_device expects a torch tensor on a HIP device")
    if X.dtype != torch.float32 or X.dim() != 2:
        raise ValueError("mst_device expects a 2-D float32 tensor")
    X = X.contiguous()
    n, d = X.shape
    ms = int(min_samples)
    wt = None
    if sample_weights is not None:
        wt = torch.as_tensor(sample_weights, dtype=torch.float32, device=X.device).contiguous()
        if wt.shape != (n,):
            raise ValueError("sample_weights must have one value per point")
        kn = _neighbor_columns(n, ms, wt.cpu().numpy())
    else:
        kn = ms if ms > 1 else 1
    m = max(n - 1, 0)
    dev = X.device
    o = out if out is not None else {}

    def buf(name, shape, dtype):
        t = o.get(name)
        if t is None or t.dtype != dtype or t.device != dev or t.numel() < int(np.prod(shape)):
            t = torch.empty(shape, dtype=dtype, device=dev)
            o[name] = t
        return t.view(-1)[:int(np.prod(shape))].view(shape)

    src = buf("src", (max(m, 1),), torch.int32)
    dst = buf("dst", (max(m, 1),), torch.int32)
    w = buf("w_sq", (max(m, 1),), torch.float32)
    core = buf("core_sq", (n,), torch.float32) if core_and_neighbors else None
    nbr = buf("neighbors", (n, kn), torch.int32) if core_and_neighbors else None
    # The library works on its own stream: inputs must be complete.
    torch.cuda.current_stream(dev).synchronize()
    st = _lib.FHStats()
    ctx = _lib.context(dev.index or 0)
    code = _lib.load().fh_mst_euclidean_device(
        ctx, ctypes.c_void_p(X.data_ptr()), n, d, ms, ctypes.c_void_p(wt.data_ptr()) if wt is not None else None,
        ctypes.c_void_p(core.data_ptr()) if core is not None else None,
        ctypes.c_void_p(nbr.data_ptr()) if nbr is not None else None,
        ctypes.c_void_p(src.data_ptr()), ctypes.c_void_p(dst.data_ptr()),
        ctypes.c_void_p(w.data_ptr()), ctypes.byref(st),
    )
    _lib.check(ctx, code)
    if stats is not None:
        stats.update(st.as_dict())
    return src[:m], dst[:m], w[:m], core, nbr


def _edges_f64(src, dst, w_sq) -> np.ndarray:
    # boruvka.py:640,646: float64 rows (src, dst, sqrt(weight)), the square
    # root taken in float64 of the float32 squared weight.
    edges = np.empty((src.shape[0], 3), dtype=np.float64)
    edges[:, 0] = src
    edges[:, 1] = dst
    edges[:, 2] = np.sqrt(w_sq.astype(np.float64))
    return edges


def compute_minimum_spanning_tree(
    data,
    min_samples=10,
    sample_weights=None,
    reproducible=False,
    metric="euclidean",
    algorithm="boruvka",
    knn_k=None,
    cannot_link=None,
    validate_cannot_link=True,
    cannot_link_groups=None,
    metric_kwds=None,
):
    """Drop-in for hdbscan.compute_minimum_spanning_tree (hdbscan.py:339-494),
    Euclidean Boruvka branch.

    Returns ``(mst_edges f64[n-1, 3], neighbors int32[n, k], core_distances
    float32[n])`` exactly as ``parallel_boruvka`` does (boruvka.py:646-647).
    ``reproducible`` is accepted and has no effect: the device path is
    deterministic and equals the reference's single-threaded run.
    ``sample_weights`` selects the reference's weighted-core branch
    (boruvka.py:564-578), quirks included.  Data with
    more than 30 columns stays on the exact KD-tree path (the reference sends
    it to approximate pynndescent, hdbscan.py:447-463).
    """
    # hdbscan.py:409-425
    if algorithm not in ("boruvka", "kruskal"):
        raise ValueError("algorithm must be 'boruvka' or 'kruskal'. Got: %s" % algorithm)
    if cannot_link is not None and cannot_link_groups is not None:
        raise ValueError(
            "cannot_link and cannot_link_groups are mutually exclusive. "
            "Provide one or the other, not both."
        )
    _has_cl = cannot_link is not None or cannot_link_groups is not None
    if _has_cl and algorithm != "kruskal":
        raise ValueError(
            "cannot_link constraints are only supported with "
            "algorithm='kruskal'. Got algorithm=%r." % algorithm
        )
    if metric == "precomputed":
        if sample_weights is not None:
            raise NotImplementedError("sample_weights is not supported with metric='precomputed'.")
        raise NotImplementedError("metric='precomputed' is not on the device MST path")
    if metric != "euclidean":
        raise NotImplementedError(f"metric={metric!r} is not on the device MST path (euclidean only)")
    if algorithm == "kruskal":
        # hdbscan.py:469-478 -> kruskal.kruskal_mst_from_feature_matrix
        if _has_cl:
            raise NotImplementedError("cannot-link constraints are not on the device MST path")
        if knn_k is None:
            if sample_weights is not None:
                raise NotImplementedError(
                    "sample_weights with knn_k=None (full pairwise) is not supported. "
                    "Use knn_k=<int> or algorithm='boruvka'."
                )
            raise NotImplementedError("algorithm='kruskal' with knn_k=None (dense float64 pairwise) is not on the "
                                      "device MST path; pass knn_k")
        return kruskal_mst_knn(_as_points(data), min_samples, knn_k, sample_weights)
    try:
        import torch

        if isinstance(data, torch.Tensor) and data.is_cuda:
            src, dst, w, core_sq, nbr = mst_device(data.float(), min_samples, sample_weights=sample_weights)
            return (
                _edges_f64(src.cpu().numpy(), dst.cpu().numpy(), w.cpu().numpy()),
                nbr.cpu().numpy(),
                np.sqrt(core_sq.cpu().numpy()),
            )
    except ImportError:
        pass

    X = _as_points(data)
    # boruvka.py:564-578: weights switch to the weighted-core branch
    w = None if sample_weights is None else np.asarray(sample_weights, dtype=np.float32)
    src, dst, wsq, core_sq, nbr = mst_arrays(X, min_samples, sample_weights=w)
    return _edges_f64(src, dst, wsq), nbr, np.sqrt(core_sq)


def kruskal_mst_knn(X, min_samples=10, knn_k=10, sample_weights=None, device: int = 0, stats: dict | None = None):
    """kruskal._kruskal_mst_knn (kruskal.py:950-1034) on the device: returns
    (mst_edges float64[n-1, 3], neighbors int32[n, c], core_distances
    float64[n]) with the reference's row order and +inf bridges."""
    X = np.ascontiguousarray(X, dtype=np.float32)
    n, d = X.shape
    ms = int(min_samples)
    kk = int(knn_k)
    w = None
    kfc = 0
    if sample_weights is not None:
        mean_sw = float(np.mean(sample_weights))  # kruskal.py:987-988
        kfc = int(2 * ms / mean_sw)
        w = np.ascontiguousarray(sample_weights, dtype=np.float32)
        if w.shape != (n,):
            raise ValueError("sample_weights must have one value per point")
        cols = max(kk, kfc)
    elif ms > 1:
        cols = ms
    else:
        cols = 1
    mst = np.empty((max(n - 1, 0), 3), np.float64)
    nbr = np.empty((n, cols), np.int32)
    core = np.empty(n, np.float64)
    st = _lib.FHStats()
    ctx = _lib.context(device)
    code = _lib.load().fh_mst_kruskal_knn(ctx, _lib.ptr(X), n, d, ms, kk, _lib.ptr(w) if w is not None else None,
                                          kfc, _lib.ptr(mst) if n > 1 else None, _lib.ptr(nbr), cols, _lib.ptr(core),
                                          ctypes.byref(st))
    _lib.check(ctx, code)
    if stats is not None:
        stats.update(st.as_dict())
    return mst, nbr, core


def parallel_tree_query(data, queries, k=10, output_rdist=False, device: int = 0):
    """numba_kdtree.parallel_tree_query(build_kdtree(data), queries, k,
    output_rdist) (numba_kdtree.py:765-792) on the device: squared distances
    with ``output_rdist``, else float32 square roots taken before the sort
    (numba_kdtree.py:792)."""
    X = _as_points(data)
    Q = _as_points(queries)
    if Q.shape[1] != X.shape[1]:
        raise ValueError("queries and data differ in dimension")
    nq = Q.shape[0]
    dist = np.empty((nq, int(k)), np.float32)
    idx = np.empty((nq, int(k)), np.int32)
    ctx = _lib.context(device)
    code = _lib.load().fh_knn(ctx, _lib.ptr(X), X.shape[0], X.shape[1], _lib.ptr(Q), nq, int(k),
                              1 if output_rdist else 0, _lib.ptr(dist), _lib.ptr(idx))
    _lib.check(ctx, code)
    return dist, idx


def build_kdtree(data, on_device: bool = False, device: int = 0):
    """numba_kdtree.build_kdtree(data, leaf_size=40) (numba_kdtree.py:478-521)
    as a dict of arrays with the reference's dtypes.  ``on_device`` builds
    it with the gfx950 kernels (the MST path's builder); otherwise with the
    library's host builder."""
    X = _as_points(data)
    n, d = X.shape
    L = _lib.load()
    N = int(L.fh_tree_nnodes(n))
    out = {
        "data": X,
        "idx_array": np.empty(n, np.int64),
        "idx_start": np.empty(N, np.int64),
        "idx_end": np.empty(N, np.int64),
        "radius": np.empty(N, np.float32),
        "is_leaf": np.empty(N, np.uint8),
        "node_bounds": np.empty((2, N, d), np.float32),
    }
    ctx = _lib.context(device) if on_device else None
    code = L.fh_build_kdtree(ctx, _lib.ptr(X), n, d, _lib.ptr(out["idx_array"]), _lib.ptr(out["idx_start"]),
                             _lib.ptr(out["idx_end"]), _lib.ptr(out["radius"]), _lib.ptr(out["is_leaf"]),
                             _lib.ptr(out["node_bounds"]))
    if ctx is not None:
        _lib.check(ctx, code)
    elif code != _lib.FH_OK:
        raise ValueError(f"fh_build_kdtree failed with code {code}")
    out["is_leaf"] = out["is_leaf"].astype(bool)
    return out
