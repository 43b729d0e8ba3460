"""fast_hdbscan-compatible entry points on the gfx950 MST path.

``fast_hdbscan`` and ``HDBSCAN`` mirror hdbscan.fast_hdbscan
(hdbscan.py:199-336) and hdbscan.HDBSCAN (hdbscan.py:577-822) for Euclidean,
unsupervised clustering (sample weights included): the MST comes from the
device kernels and the labels from the host label stage.  ``hdbscan_device``
is the fully device-resident pipeline used by the benchmark.
"""
from __future__ import annotations

import numpy as np

from . import _lib
from .hdbscan import compute_minimum_spanning_tree, mst_device
from .labels import clusters_from_spanning_tree, label_params, labels_from_sorted_arrays


def _check_params(min_samples, min_cluster_size, cluster_selection_epsilon, cluster_selection_persistence):
    # hdbscan.py:254-275
    if (
        (not (np.issubdtype(type(min_samples), np.integer) or min_samples is None))
        or not np.issubdtype(type(min_cluster_size), np.integer)
        or (min_samples is not None and min_samples <= 0)
        or min_cluster_size <= 0
    ):
        raise ValueError("Min samples and min cluster size must be positive integers!")
    if not np.issubdtype(type(cluster_selection_epsilon), np.floating) or cluster_selection_epsilon < 0.0:
        raise ValueError("Cluster selection epsilon must be a positive floating point number!")
    if not np.issubdtype(type(cluster_selection_persistence), np.floating) or cluster_selection_persistence < 0.0:
        raise ValueError("Cluster selection persistence must be a positive floating point number!")


def fast_hdbscan(
    data,
    data_labels=None,
    semi_supervised=False,
    ss_algorithm="bc",
    min_samples=10,
    min_cluster_size=10,
    cluster_selection_method="eom",
    max_cluster_size=np.inf,
    allow_single_cluster=False,
    cluster_selection_epsilon=0.0,
    cluster_selection_persistence=0.0,
    sample_weights=None,
    reproducible=False,
    return_trees=False,
    metric="euclidean",
    algorithm="boruvka",
    knn_k=None,
    cannot_link=None,
    validate_cannot_link=True,
    cannot_link_groups=None,
    metric_kwds=None,
):
    """Drop-in for hdbscan.fast_hdbscan (hdbscan.py:199-336), Euclidean."""
    from sklearn.utils import check_array

    if metric != "euclidean":
        raise NotImplementedError("only metric='euclidean' runs on the device MST path")
    data = check_array(data)
    if semi_supervised:
        if data_labels is None:
            raise ValueError("data_labels must not be None when semi_supervised is set to True!")
        if ss_algorithm not in ["bc", "bc_simple"]:
            raise ValueError(f"Invalid ss_algorithm {ss_algorithm}")
    _check_params(min_samples, min_cluster_size, cluster_selection_epsilon, cluster_selection_persistence)
    if cluster_selection_method not in ("eom", "leaf"):
        raise ValueError(f"Invalid cluster_selection_method {cluster_selection_method}")

    # hdbscan.py:278-304
    if data.shape[0] < 2:
        n = data.shape[0]
        labels = np.full(n, -1, dtype=np.intp)
        probabilities = np.zeros(n, dtype=np.float64)
        single_linkage_tree = np.zeros((0, 4), dtype=np.float64)
        condensed_tree = np.zeros(0, dtype=[("parent", np.intp), ("child", np.intp), ("lambda_val", np.float64),
                                            ("child_size", np.intp)])
        mst = np.zeros((0, 3), dtype=np.float64)
        neighbors = np.zeros((n, 0), dtype=np.intp)
        core_distances = np.full(n, np.inf, dtype=np.float64)
        result = (labels, probabilities, single_linkage_tree, condensed_tree, mst, neighbors, core_distances)
        return result[: (None if return_trees else 2)]

    mst, neighbors, core_distances = compute_minimum_spanning_tree(
        data,
        min_samples=min_cluster_size if min_samples is None else min_samples,
        sample_weights=sample_weights,
        reproducible=reproducible,
        metric=metric,
        algorithm=algorithm,
        knn_k=knn_k,
        cannot_link=cannot_link,
        validate_cannot_link=validate_cannot_link,
        cannot_link_groups=cannot_link_groups,
        metric_kwds=metric_kwds,
    )
    return (
        *clusters_from_spanning_tree(
            mst,
            data_labels=data_labels,
            semi_supervised=semi_supervised,
            ss_algorithm=ss_algorithm,
            min_cluster_size=min_cluster_size,
            cluster_selection_method=cluster_selection_method,
            max_cluster_size=max_cluster_size,
            allow_single_cluster=allow_single_cluster,
            cluster_selection_epsilon=cluster_selection_epsilon,
            cluster_selection_persistence=cluster_selection_persistence,
            sample_weights=sample_weights,
        ),
        neighbors,
        core_distances,
    )[: (None if return_trees else 2)]


def to_numpy_rec_array(named_tuple_tree):
    """hdbscan.py:47-66"""
    size = named_tuple_tree.parent.shape[0]
    result = np.empty(size, dtype=[("parent", np.intp), ("child", np.intp), ("lambda_val", float),
                                   ("child_size", np.intp)])
    result["parent"] = named_tuple_tree.parent
    result["child"] = named_tuple_tree.child
    result["lambda_val"] = named_tuple_tree.lambda_val
    result["child_size"] = named_tuple_tree.child_size
    return result


class HDBSCAN:
    """hdbscan.HDBSCAN (hdbscan.py:577-822) for Euclidean, unsupervised
    clustering, with or without sample weights: fit / fit_predict, labels_,
    probabilities_.  Rows with
    non-finite values are left out of the clustering and labelled noise with
    probability 0 (hdbscan.py:768-820)."""

    def __init__(self, min_cluster_size=5, min_samples=None, cluster_selection_method="eom",
                 allow_single_cluster=False, max_cluster_size=np.inf, cluster_selection_epsilon=0.0,
                 cluster_selection_persistence=0.0, semi_supervised=False, ss_algorithm="bc", reproducible=False,
                 metric="euclidean", algorithm="boruvka", knn_k=None, cannot_link=None, validate_cannot_link=True,
                 cannot_link_groups=None, metric_kwds=None):
        self.min_cluster_size = min_cluster_size
        self.min_samples = min_samples
        self.cluster_selection_method = cluster_selection_method
        self.allow_single_cluster = allow_single_cluster
        self.max_cluster_size = max_cluster_size
        self.cluster_selection_epsilon = cluster_selection_epsilon
        self.cluster_selection_persistence = cluster_selection_persistence
        self.semi_supervised = semi_supervised
        self.ss_algorithm = ss_algorithm
        self.reproducible = reproducible
        self.metric = metric
        self.algorithm = algorithm
        self.knn_k = knn_k
        self.cannot_link = cannot_link
        self.validate_cannot_link = validate_cannot_link
        self.cannot_link_groups = cannot_link_groups
        self.metric_kwds = metric_kwds

    def get_params(self, deep=True):
        return {k: getattr(self, k) for k in (
            "min_cluster_size", "min_samples", "cluster_selection_method", "allow_single_cluster",
            "max_cluster_size", "cluster_selection_epsilon", "cluster_selection_persistence", "semi_supervised",
            "ss_algorithm", "reproducible", "metric", "algorithm", "knn_k", "cannot_link", "validate_cannot_link",
            "cannot_link_groups", "metric_kwds")}

    def fit(self, X, y=None, sample_weight=None, **fit_params):
        from sklearn.utils import check_array

        X = check_array(X, ensure_all_finite=False)
        if sample_weight is not None:
            sample_weight = _check_sample_weight(sample_weight, X, dtype=np.float32)
        all_finite = bool(np.all(np.isfinite(X)))
        if not all_finite:
            finite_index = np.where(np.isfinite(X).sum(axis=1) == X.shape[1])[0]
            clean = X[finite_index]
            if sample_weight is not None:
                sample_weight = sample_weight[finite_index]
        else:
            clean = X
        (self.labels_, self.probabilities_, self._single_linkage_tree, self._condensed_tree,
         self._min_spanning_tree, self._neighbors, self._core_distances) = fast_hdbscan(
            clean, None, return_trees=True, sample_weights=sample_weight, **self.get_params())
        if not isinstance(self._condensed_tree, np.ndarray):
            self._condensed_tree = to_numpy_rec_array(self._condensed_tree)
        if not all_finite:
            labels = np.full(X.shape[0], -1)
            labels[finite_index] = self.labels_
            self.labels_ = labels
            probs = np.zeros(X.shape[0])
            probs[finite_index] = self.probabilities_
            self.probabilities_ = probs
        return self

    def fit_predict(self, X, y=None, **fit_params):
        self.fit(X, y, **fit_params)
        return self.labels_


def hdbscan_device(X, min_samples=10, min_cluster_size=10, cluster_selection_method="eom",
                   max_cluster_size=np.inf, allow_single_cluster=False, cluster_selection_epsilon=0.0,
                   stats: dict | None = None, work: dict | None = None):
    """Labels and membership strengths for a float32 torch tensor resident
    on a HIP device: MST on the device, rows sorted on the device, then the
    host label stage.  Returns (labels int64, strengths float32) on the
    host.  ``work`` (a dict, kept by the caller between calls) holds the
    device, pinned and output buffers; with it the returned arrays are
    reused by the next call."""
    w = work if work is not None else {}
    staged = device_stage(X, min_samples, stats=stats, work=w)
    return host_stage(staged, min_cluster_size, cluster_selection_method, max_cluster_size, allow_single_cluster,
                      cluster_selection_epsilon, stats=stats, work=w, reuse=work is not None)


def device_stage(X, min_samples=10, stats: dict | None = None, work: dict | None = None):
    """hdbscan_device's device part: MST, row sort and row groups on the
    device, the sorted rows in pinned host memory of ``work`` on return.
    The returned handle is host_stage's input; the device is free for the
    next request meanwhile (a caller with two ``work`` dicts can overlap one
    request's host stage with the next one's device stage)."""
    import ctypes
    import time

    import torch

    n = X.shape[0]
    w = work if work is not None else {}
    src, dst, w_sq, _, _ = mst_device(X, min_samples, stats=stats, out=w.setdefault("mst", {}),
                                      core_and_neighbors=False)
    m = n - 1
    dev = X.device
    if "sorted" not in w or w["sorted"][0].shape[0] < m or w["order"][0].shape[0] < n:
        w["sorted"] = tuple(torch.empty(max(m, 1), dtype=t, device=dev)
                            for t in (torch.int32, torch.int32, torch.float32, torch.int32, torch.int32))
        w["host"] = tuple(torch.empty(max(m, 1), dtype=t).pin_memory() for t in (torch.int32, torch.int32, torch.float32))
        w["order"] = (torch.empty(max(n, 1), dtype=torch.int32, device=dev),
                      torch.empty(max(n, 1), dtype=torch.int32).pin_memory())
    ss, sd, sw, ps, pd = w["sorted"]
    od, oh = w["order"]
    t0 = time.perf_counter()
    ctx = _lib.context(dev.index or 0)
    L = _lib.load()
    vp = ctypes.c_void_p
    if m > 0:
        # rows sorted on the device, then their endpoints as KD-tree
        # positions (the host union-find runs faster in that order)
        _lib.check(ctx, L.fh_sort_mst_device(ctx, m, vp(src.data_ptr()), vp(dst.data_ptr()), vp(w_sq.data_ptr()),
                                             vp(ss.data_ptr()), vp(sd.data_ptr()), vp(sw.data_ptr())))
        _lib.check(ctx, L.fh_edges_to_tree_order(ctx, n, m, vp(ss.data_ptr()), vp(sd.data_ptr()), vp(ps.data_ptr()),
                                                 vp(pd.data_ptr()), vp(od.data_ptr())))
    # the sorted rows (ready: the library calls above end synchronised) go
    # to pinned host memory on torch's stream while the library's stream
    # computes the row groups
    hs, hd, hw = w["host"]
    hs[:m].copy_(ps[:m], non_blocking=True)
    hd[:m].copy_(pd[:m], non_blocking=True)
    hw[:m].copy_(sw[:m], non_blocking=True)
    oh[:n].copy_(od[:n], non_blocking=True)
    # the lightest 90% of the rows, grouped by the components they form:
    # the host linkage runs those groups on parallel threads
    k = (9 * m) // 10 if m >= (1 << 16) else 0
    groups = None
    if k > 0:
        if w.get("plan") is None or w["plan"][0].shape[0] < k:
            # pinned host buffers (kept in `w`), used through numpy views
            w["plan_pinned"] = tuple(torch.empty(k, dtype=torch.int32).pin_memory() for _ in range(2))
            w["plan"] = tuple(t.numpy() for t in w["plan_pinned"])
        ro, gs = w["plan"]
        ng = ctypes.c_int64(0)
        _lib.check(ctx, L.fh_linkage_plan(ctx, n, k, vp(ps.data_ptr()), vp(pd.data_ptr()), _lib.ptr(ro), _lib.ptr(gs),
                                          ctypes.byref(ng)))
        groups = (k, ro[:k], gs[:ng.value])
    torch.cuda.current_stream(dev).synchronize()
    if stats is not None:
        stats["t_sort_copy_ms"] = 1e3 * (time.perf_counter() - t0)
    return n, m, groups


def host_stage(staged, min_cluster_size=10, cluster_selection_method="eom", max_cluster_size=np.inf,
               allow_single_cluster=False, cluster_selection_epsilon=0.0, stats: dict | None = None,
               work: dict | None = None, reuse: bool = True):
    """hdbscan_device's host part on device_stage's result (the same
    ``work``): the label stage.  Returns (labels int64, strengths float32)."""
    import time

    n, m, groups = staged
    w = work
    hs, hd, hw = w["host"]
    oh = w["order"][1]
    t1 = time.perf_counter()
    params = label_params(min_cluster_size, cluster_selection_method, max_cluster_size, allow_single_cluster,
                          cluster_selection_epsilon)
    phases = {}
    # label outputs kept in `w` too: large fresh arrays would cost their page
    # faults on every call (the results alias them until the next call)
    if w.get("labels") is None or w["labels"][0].shape[0] < n:
        w["labels"] = (np.empty(n, np.int64), np.empty(n, np.float32))
    labels, strengths, _, _ = labels_from_sorted_arrays(n, hs[:m].numpy(), hd[:m].numpy(), hw[:m].numpy(), params,
                                                        phases=phases, point_order=oh[:n].numpy(), groups=groups,
                                                        out=w["labels"] if reuse else None)
    t2 = time.perf_counter()
    if stats is not None:
        stats["t_labels_ms"] = 1e3 * (t2 - t1)
        stats.update(phases)
    return labels, strengths


def _check_sample_weight(sample_weight, X, dtype):
    """hdbscan.py:44-90 (scalar or 1-D array, one per row, not all zero)."""
    import numbers

    from sklearn.utils import check_array

    n_samples = X.shape[0]
    if isinstance(sample_weight, numbers.Number):
        sample_weight = np.full(n_samples, sample_weight, dtype=dtype)
    else:
        sample_weight = check_array(sample_weight, accept_sparse=False, ensure_2d=False, dtype=dtype, order="C",
                                    input_name="sample_weight")
        if sample_weight.ndim != 1:
            raise ValueError(
                f"Sample weights must be 1D array or scalar, got {sample_weight.ndim}D array. Expected either a "
                f"scalar value or a 1D array of length {n_samples}."
            )
        if sample_weight.shape != (n_samples,):
            raise ValueError("sample_weight.shape == {}, expected {}!".format(sample_weight.shape, (n_samples,)))
    if np.all(sample_weight == 0):
        raise ValueError("Sample weights must contain at least one non-zero number.")
    return sample_weight


try:
    from sklearn.base import BaseEstimator, ClusterMixin
except ImportError:  # pragma: no cover - scikit-learn is a dependency of the reference
    BaseEstimator = object
    ClusterMixin = object


class PLSCAN(ClusterMixin, BaseEstimator):
    """hdbscan.PLSCAN (hdbscan.py:890-1187) for Euclidean data: the MST on
    the device, the layer stage on the host.  Attributes as the reference:
    cluster_layers_, membership_strength_layers_, layer_persistence_scores_,
    min_cluster_sizes_, total_persistence_, cluster_tree_, labels_,
    membership_strengths_ (the latter only with more than one layer)."""

    def __init__(self, *, min_samples=5, max_layers=10, base_min_cluster_size=5, base_n_clusters=None,
                 layer_similarity_threshold=0.2, reproducible=False, metric="euclidean", algorithm="boruvka",
                 knn_k=None, cannot_link=None, validate_cannot_link=True, metric_kwds=None, verbose=False):
        self.min_samples = min_samples
        self.max_layers = max_layers
        self.base_min_cluster_size = base_min_cluster_size
        self.base_n_clusters = base_n_clusters
        self.layer_similarity_threshold = layer_similarity_threshold
        self.reproducible = reproducible
        self.metric = metric
        self.algorithm = algorithm
        self.knn_k = knn_k
        self.cannot_link = cannot_link
        self.validate_cannot_link = validate_cannot_link
        self.metric_kwds = metric_kwds
        self.verbose = verbose
        self._validate_params()

    def _validate_params(self):
        # hdbscan.py:1042-1073
        if (not np.issubdtype(type(self.min_samples), np.integer)) or self.min_samples <= 0:
            raise ValueError("Min samples must be a positive integer!")
        if (not np.issubdtype(type(self.max_layers), np.integer)) or self.max_layers <= 0:
            raise ValueError("Max layers must be a positive integer!")
        if (not np.issubdtype(type(self.base_min_cluster_size), np.integer)) or self.base_min_cluster_size <= 0:
            raise ValueError("Base min cluster size must be a positive integer!")
        if self.base_n_clusters is not None:
            if (not np.issubdtype(type(self.base_n_clusters), np.integer)) or self.base_n_clusters <= 0:
                raise ValueError("Base n clusters must be a positive integer!")
        if self.algorithm not in ("boruvka", "kruskal"):
            raise ValueError("algorithm must be 'boruvka' or 'kruskal'. Got: %s" % self.algorithm)
        if self.cannot_link is not None and self.algorithm != "kruskal":
            raise ValueError(
                "cannot_link constraints are only supported with algorithm='kruskal'. Got algorithm=%r."
                % self.algorithm
            )

    def fit_predict(self, X, y=None, sample_weight=None, **fit_params):
        """hdbscan.py:1075-1160, metric='euclidean'."""
        import scipy.sparse
        from sklearn.utils.validation import validate_data

        from .layers import build_cluster_layers, build_layer_cluster_tree

        self._validate_params()
        if self.metric != "euclidean":
            raise NotImplementedError(f"metric={self.metric!r} is not on the device MST path (euclidean only)")
        if scipy.sparse.issparse(X):
            raise ValueError(
                "PLSCAN requires a dense feature matrix for metric='euclidean'. Sparse matrices are only "
                "supported with metric='precomputed'."
            )
        X = validate_data(self, X, accept_sparse=False, ensure_all_finite=False)
        self._raw_data = X
        if sample_weight is not None:
            sample_weight = _check_sample_weight(sample_weight, X, dtype=np.float32)
        kwargs = self.get_params()
        (self.cluster_layers_, self.membership_strength_layers_, self.layer_persistence_scores_,
         self.min_cluster_sizes_, self.total_persistence_) = build_cluster_layers(X, sample_weights=sample_weight,
                                                                                  **kwargs)
        self.cluster_tree_ = build_layer_cluster_tree(self.cluster_layers_)
        if len(self.cluster_layers_) == 1:
            self.labels_ = self.cluster_layers_[0]
            self.membership_strengths_ = self.membership_strength_layers_[0]
        else:
            best_layer = np.argmax(self.layer_persistence_scores_)
            self.labels_ = self.cluster_layers_[best_layer]
            self.membership_strengths_ = self.membership_strength_layers_[best_layer]
        return self.labels_

    def fit(self, X, y=None, sample_weight=None, **fit_params):
        _ = self.fit_predict(X, y=y, sample_weight=sample_weight, **fit_params)
        return self
