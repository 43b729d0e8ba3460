"""PLSCAN cluster layers on top of the device MST.

``build_cluster_layers`` and ``build_layer_cluster_tree`` mirror
layer_clusters.build_cluster_layers (layer_clusters.py:353-495) and
build_layer_cluster_tree (528-554): the MST comes from the device path, the
NumPy steps the reference runs in plain Python (core-distance
normalisation, ``np.argsort`` of the weights, the final layer ordering) are
the same NumPy calls, and the numba-compiled stages run in the host C++
library (``fh_cluster_layers``, ``fh_layer_cluster_tree``).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .hdbscan import compute_minimum_spanning_tree


def _normalize(edges, core_distances):
    # layer_clusters.py:387-397
    finite_core = core_distances[np.isfinite(core_distances)]
    if len(finite_core) > 0:
        mean_core_distance = np.mean(finite_core)
        min_core_distance = np.min(finite_core)
    else:
        mean_core_distance = 0.0
        min_core_distance = 0.0
    if mean_core_distance > min_core_distance and np.isfinite(mean_core_distance):
        edges[:, 2] -= min_core_distance
        edges[:, 2] /= mean_core_distance - min_core_distance
    return edges


def layers_from_mst(edges, core_distances, n_samples, base_min_cluster_size=10, base_n_clusters=None,
                    layer_similarity_threshold=0.2, max_layers=10):
    """layer_clusters.py:386-495 for a given MST (edges float64[n-1, 3] and
    core distances as compute_minimum_spanning_tree returns them)."""
    edges = _normalize(np.array(edges, dtype=np.float64, copy=True), core_distances)
    sorted_mst = edges[np.argsort(edges.T[2])]  # layer_clusters.py:398
    n = int(n_samples)
    src = np.ascontiguousarray(sorted_mst[:, 0].astype(np.int32))
    dst = np.ascontiguousarray(sorted_mst[:, 1].astype(np.int32))
    dist = np.ascontiguousarray(sorted_mst[:, 2])
    ml = int(max_layers)
    labels = np.empty((ml, n), dtype=np.int64)
    strengths = np.empty((ml, n), dtype=np.float32)
    persistence = np.zeros(ml, dtype=np.float32)
    sizes = np.empty(max(n, 1), dtype=np.float32)
    total = np.empty(max(n, 1), dtype=np.float32)
    params = _lib.FHLayerParams(int(base_min_cluster_size), 0 if base_n_clusters is None else int(base_n_clusters),
                                float(layer_similarity_threshold), ml)
    out = _lib.FHLayers(_lib.ptr(labels), _lib.ptr(strengths), _lib.ptr(persistence), _lib.ptr(sizes),
                        _lib.ptr(total), sizes.shape[0], 0, 0, 0)
    code = _lib.load().fh_cluster_layers(n, _lib.ptr(src), _lib.ptr(dst), _lib.ptr(dist), ctypes.byref(params),
                                         ctypes.byref(out))
    if code != 0:
        raise _lib.FHError(f"fh_cluster_layers failed with code {code}")
    k = int(out.n_layers)
    cluster_layers = [labels[i].astype(np.intp) for i in range(k)]
    strength_layers = [strengths[i].copy() for i in range(k)]
    scores = [0.0] + [persistence[i] for i in range(1, k)]
    if out.n_sizes > 0:
        sizes_out = sizes[: out.n_sizes].copy()
        total_out = total[: out.n_sizes].copy()
    else:
        sizes_out = np.array([])
        total_out = np.array([])
    # layer_clusters.py:484-490: most clusters first
    n_clusters_per_layer = [layer.max() + 1 for layer in cluster_layers]
    order = np.argsort(n_clusters_per_layer)[::-1]
    return ([cluster_layers[i] for i in order], [strength_layers[i] for i in order], [scores[i] for i in order],
            sizes_out, total_out)


def plscan_sweep(data, min_samples_values, base_min_cluster_size=10, base_n_clusters=None,
                 layer_similarity_threshold=0.2, max_layers=10, workers=8, device=0):
    """build_cluster_layers (layer_clusters.py:353-495) of one dataset for
    each value in ``min_samples_values``, as a pipeline: the points go to the
    device once, the MSTs run one after another on the device, and each
    MST's layer stage (host NumPy + C++, which release the GIL) runs in one
    of ``workers`` host threads while the device computes the next MST.

    Returns {min_samples: (cluster_layers, membership_strength_layers,
    persistence_scores, sizes, total_persistence)}, each entry what
    ``layers_from_mst`` returns for that MST.
    """
    import concurrent.futures as cf

    import torch

    from .hdbscan import _as_points, _edges_f64, mst_device

    if isinstance(data, torch.Tensor) and data.is_cuda:
        Xd = data.contiguous().float()
    else:
        Xd = torch.from_numpy(_as_points(data)).to(f"cuda:{device}")
    n = int(Xd.shape[0])
    values = [int(m) for m in min_samples_values]
    out = {}
    shared = n > 1 and all(2 <= m <= 126 for m in values)
    if shared:
        try:
            msts = mst_sweep_device(Xd, values)
        except NotImplementedError:
            # the shared sweep runs the tree kNN; inputs that the MFMA scans
            # take (high dimension, small inputs from 5 dimensions on) run
            # value by value
            shared = False
    with cf.ThreadPoolExecutor(max_workers=max(1, int(workers))) as pool:
        futures = {}
        for i, m in enumerate(values):
            if shared:
                src, dst, w_sq, core_sq = (t[i] for t in msts)
            else:
                src, dst, w_sq, core_sq, _ = mst_device(Xd, m)
            edges = _edges_f64(src.cpu().numpy(), dst.cpu().numpy(), w_sq.cpu().numpy())
            core = np.sqrt(core_sq.cpu().numpy())
            futures[m] = pool.submit(layers_from_mst, edges, core, n, base_min_cluster_size, base_n_clusters,
                                     layer_similarity_threshold, max_layers)
        for m in values:
            out[m] = futures[m].result()
    return out


def mst_sweep_device(X, min_samples_values, stats: dict | None = None):
    """MSTs of one device-resident float32 tensor for several min_samples
    values (each in [2, 126]) from one KD-tree and one shared kNN query
    (fh_mst_sweep_device); equal, value by value, to mst_device.  Returns
    torch tensors on X's device: src, dst [count, n-1] int32, w_sq
    [count, n-1] float32, core_sq [count, n] float32."""
    import ctypes

    import torch

    from . import _lib

    if not (isinstance(X, torch.Tensor) and X.is_cuda) or X.dtype != torch.float32 or X.dim() != 2:
        raise TypeError("mst_sweep_device expects a 2-D float32 tensor on a HIP device")
    X = X.contiguous()
    n, d = X.shape
    values = np.ascontiguousarray([int(m) for m in min_samples_values], dtype=np.int32)
    c = values.shape[0]
    m = max(n - 1, 0)
    dev = X.device
    src = torch.empty((c, max(m, 1)), dtype=torch.int32, device=dev)
    dst = torch.empty((c, max(m, 1)), dtype=torch.int32, device=dev)
    w = torch.empty((c, max(m, 1)), dtype=torch.float32, device=dev)
    core = torch.empty((c, n), dtype=torch.float32, device=dev)
    torch.cuda.current_stream(dev).synchronize()
    st = _lib.FHStats()
    ctx = _lib.context(dev.index or 0)
    code = _lib.load().fh_mst_sweep_device(ctx, ctypes.c_void_p(X.data_ptr()), n, d, c, _lib.ptr(values),
                                           ctypes.c_void_p(core.data_ptr()), ctypes.c_void_p(src.data_ptr()),
                                           ctypes.c_void_p(dst.data_ptr()), ctypes.c_void_p(w.data_ptr()),
                                           ctypes.byref(st))
    _lib.check(ctx, code)
    if stats is not None:
        stats.update(st.as_dict())
    if m == 0:
        src, dst, w = src[:, :0], dst[:, :0], w[:, :0]
    return src, dst, w, core


def build_cluster_layers(
    data,
    *,
    min_samples=5,
    base_min_cluster_size=10,
    base_n_clusters=None,
    reproducible=False,
    layer_similarity_threshold=0.2,
    max_layers=10,
    sample_weights=None,
    metric="euclidean",
    algorithm="boruvka",
    knn_k=None,
    cannot_link=None,
    validate_cannot_link=True,
    metric_kwds=None,
    verbose=False,
):
    """Drop-in for layer_clusters.build_cluster_layers (layer_clusters.py:
    353-495).  Returns (cluster_layers, membership_strength_layers,
    persistence_scores, sizes, total_persistence)."""
    n_samples = data.shape[0]
    if verbose:
        print("Computing MST ...")
    edges, _neighbors, core_distances = compute_minimum_spanning_tree(
        data,
        min_samples=base_min_cluster_size if min_samples is None else min_samples,
        sample_weights=sample_weights,
        reproducible=reproducible,
        metric=metric,
        algorithm=algorithm,
        knn_k=knn_k,
        cannot_link=cannot_link,
        validate_cannot_link=validate_cannot_link,
        metric_kwds=metric_kwds,
    )
    if verbose:
        print("Finding optimal resolution layers ...")
    return layers_from_mst(edges, np.asarray(core_distances), n_samples, base_min_cluster_size, base_n_clusters,
                           layer_similarity_threshold, max_layers)


def build_layer_cluster_tree(labels):
    """Drop-in for layer_clusters.build_layer_cluster_tree (layer_clusters.
    py:528-554): {(parent layer, parent cluster): [(child layer, child
    cluster), ...]}."""
    stacked = np.ascontiguousarray(np.vstack(labels), dtype=np.int64)
    n_layers, n = stacked.shape
    cap = int(sum(int(row.max()) + 1 for row in stacked))
    mapping = np.empty((max(cap, 1), 4), dtype=np.int64)
    count = _lib.load().fh_layer_cluster_tree(n_layers, n, _lib.ptr(stacked), _lib.ptr(mapping), cap)
    if count < 0 or count > cap:
        raise _lib.FHError(f"fh_layer_cluster_tree failed with code {count}")
    result = {}
    for parent_layer, parent_cluster, child_layer, child_cluster in mapping[:count].tolist():
        result.setdefault((parent_layer, parent_cluster), []).append((child_layer, child_cluster))
    return result
