"""ctypes binding of libfh.so (include/fh.h).

The library is the only compute path: when it is missing or no HIP device is
visible, every entry point raises -- there is no CPU fallback.
"""
from __future__ import annotations

import atexit
import ctypes
import os
import threading

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfh.so")

FH_OK = 0
FH_EINVAL = -1
FH_ENOMEM = -2
FH_EDEVICE = -3
FH_EUNSUPPORTED = -4
FH_MAX_DIM = 64

# Every symbol include/fh.h declares.
EXPORTED_SYMBOLS = (
    "fh_create",
    "fh_destroy",
    "fh_last_error",
    "fh_version",
    "fh_mst_euclidean_device",
    "fh_mst_euclidean",
    "fh_mst_euclidean_device_noalloc",
    "fh_mst_sweep_device",
    "fh_knn",
    "fh_tree_nnodes",
    "fh_build_kdtree",
    "fh_labels_from_sorted_mst",
    "fh_sort_mst_device",
    "fh_edges_to_tree_order",
    "fh_linkage_plan",
    "fh_neighbor_columns",
    "fh_cluster_layers",
    "fh_layer_cluster_tree",
    "fh_mst_kruskal_knn",
    "fh_set_learned_radius",
    "fh_mlp_step",
    "fh_core_graph",
    "fh_csr_boruvka",
    "fh_hdbscan_batch",
    "fh_mst_batch",
)


class FHStats(ctypes.Structure):
    _fields_ = [
        ("rounds", ctypes.c_int32),
        ("n_levels", ctypes.c_int32),
        ("n_nodes", ctypes.c_int64),
        ("initial_components", ctypes.c_int64),
        ("n_edges", ctypes.c_int64),
        ("t_tree_ms", ctypes.c_double),
        ("t_knn_ms", ctypes.c_double),
        ("t_init_ms", ctypes.c_double),
        ("t_rounds_ms", ctypes.c_double),
        ("t_query_ms", ctypes.c_double),
        ("t_query_main_ms", ctypes.c_double),
        ("t_total_ms", ctypes.c_double),
        ("high_dim", ctypes.c_int32),
        ("pad0", ctypes.c_int32),
        ("knn_tree_rows", ctypes.c_int64),
        ("exact_rows", ctypes.c_int64),
        ("learned", ctypes.c_int32),
        ("pad1", ctypes.c_int32),
        ("ball_misses", ctypes.c_int64),
        ("t_learn_ms", ctypes.c_double),
        ("t_predict_ms", ctypes.c_double),
        ("scored_tile_pairs", ctypes.c_int64),
        ("tie_cycle_nodes", ctypes.c_int64),
        ("scan_rounds", ctypes.c_int32),
        ("pad2", ctypes.c_int32),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class FHLabelParams(ctypes.Structure):
    _fields_ = [
        ("min_cluster_size", ctypes.c_int64),
        ("max_cluster_size", ctypes.c_double),
        ("cluster_selection_epsilon", ctypes.c_double),
        ("selection_method", ctypes.c_int32),
        ("allow_single_cluster", ctypes.c_int32),
    ]


class FHLabelOrder(ctypes.Structure):
    _fields_ = [
        ("point_order", ctypes.c_void_p),
        ("prefix_rows", ctypes.c_int64),
        ("n_groups", ctypes.c_int64),
        ("row_order", ctypes.c_void_p),
        ("group_sizes", ctypes.c_void_p),
        ("threads", ctypes.c_int32),
        ("pad", ctypes.c_int32),
    ]


class FHCondensedTree(ctypes.Structure):
    _fields_ = [
        ("parent", ctypes.c_void_p),
        ("child", ctypes.c_void_p),
        ("lambda_val", ctypes.c_void_p),
        ("child_size", ctypes.c_void_p),
        ("capacity", ctypes.c_int64),
        ("rows", ctypes.c_int64),
    ]


class FHLayerParams(ctypes.Structure):
    _fields_ = [
        ("base_min_cluster_size", ctypes.c_int64),
        ("base_n_clusters", ctypes.c_int64),
        ("layer_similarity_threshold", ctypes.c_double),
        ("max_layers", ctypes.c_int32),
    ]


class FHLayers(ctypes.Structure):
    _fields_ = [
        ("labels", ctypes.c_void_p),
        ("strengths", ctypes.c_void_p),
        ("persistence", ctypes.c_void_p),
        ("sizes", ctypes.c_void_p),
        ("total_persistence", ctypes.c_void_p),
        ("capacity", ctypes.c_int64),
        ("n_layers", ctypes.c_int32),
        ("n_sizes", ctypes.c_int64),
        ("min_cluster_size", ctypes.c_int64),
    ]


class FHError(RuntimeError):
    pass


_lib = None
_lock = threading.Lock()

_vp = ctypes.c_void_p
_i64 = ctypes.c_int64
_i32 = ctypes.c_int32


def load() -> ctypes.CDLL:
    """Load libfh.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `make -C fast_hdbscan_amd/csrc` "
                "(or __graft_entry__.build()); there is no CPU fallback"
            )
        # torch (when installed) ships its own HIP runtime under the same
        # soname: load it first so that libfh.so binds to that one and the
        # process holds a single runtime (torch cannot see the GPU once
        # another HIP runtime has opened it)
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        L.fh_create.restype = _vp
        L.fh_create.argtypes = [ctypes.c_int]
        L.fh_destroy.restype = None
        L.fh_destroy.argtypes = [_vp]
        L.fh_last_error.restype = ctypes.c_char_p
        L.fh_last_error.argtypes = [_vp]
        L.fh_version.restype = ctypes.c_char_p
        L.fh_version.argtypes = []
        mst_args = [_vp, _vp, _i64, _i32, _i32, _vp, _vp, _vp, _vp, _vp, _vp, ctypes.POINTER(FHStats)]
        L.fh_mst_euclidean_device.restype = ctypes.c_int
        L.fh_mst_euclidean_device.argtypes = mst_args
        L.fh_mst_euclidean.restype = ctypes.c_int
        L.fh_mst_euclidean.argtypes = mst_args
        L.fh_mst_euclidean_device_noalloc.restype = ctypes.c_int
        L.fh_mst_euclidean_device_noalloc.argtypes = [_i32, _vp, _vp, _i64, _i32, _i32, _vp, _vp, _vp, _vp, _vp,
                                                      ctypes.POINTER(FHStats), ctypes.c_char_p, _i32]
        L.fh_mst_sweep_device.restype = ctypes.c_int
        L.fh_mst_sweep_device.argtypes = [_vp, _vp, _i64, _i32, _i32, _vp, _vp, _vp, _vp, _vp, ctypes.POINTER(FHStats)]
        L.fh_knn.restype = ctypes.c_int
        L.fh_knn.argtypes = [_vp, _vp, _i64, _i32, _vp, _i64, _i32, _i32, _vp, _vp]
        L.fh_tree_nnodes.restype = _i64
        L.fh_tree_nnodes.argtypes = [_i64]
        L.fh_build_kdtree.restype = ctypes.c_int
        L.fh_build_kdtree.argtypes = [_vp, _vp, _i64, _i32, _vp, _vp, _vp, _vp, _vp, _vp]
        L.fh_labels_from_sorted_mst.restype = ctypes.c_int
        L.fh_labels_from_sorted_mst.argtypes = [_i64, _vp, _vp, _vp, _vp, _vp, ctypes.POINTER(FHLabelParams), _vp,
                                                _vp, _vp, _vp, _vp, _vp]
        L.fh_neighbor_columns.restype = _i64
        L.fh_neighbor_columns.argtypes = [_i64, _i32, _vp]
        L.fh_cluster_layers.restype = ctypes.c_int
        L.fh_cluster_layers.argtypes = [_i64, _vp, _vp, _vp, ctypes.POINTER(FHLayerParams), ctypes.POINTER(FHLayers)]
        L.fh_layer_cluster_tree.restype = _i64
        L.fh_layer_cluster_tree.argtypes = [_i32, _i64, _vp, _vp, _i64]
        L.fh_mst_kruskal_knn.restype = ctypes.c_int
        L.fh_mst_kruskal_knn.argtypes = [_vp, _vp, _i64, _i32, _i32, _i32, _vp, _i32, _vp, _vp, _i32, _vp,
                                         ctypes.POINTER(FHStats)]
        L.fh_set_learned_radius.restype = ctypes.c_int
        L.fh_set_learned_radius.argtypes = [_vp, _i32]
        L.fh_mlp_step.restype = ctypes.c_int
        L.fh_mlp_step.argtypes = [_vp, _i32, _i32, _vp, _vp, _vp, _vp, _vp, ctypes.c_float]
        L.fh_core_graph.restype = _i64
        L.fh_core_graph.argtypes = [_i64, _i32, _vp, _vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _i64]
        L.fh_csr_boruvka.restype = ctypes.c_int
        L.fh_csr_boruvka.argtypes = [_vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp]
        L.fh_hdbscan_batch.restype = ctypes.c_int
        L.fh_hdbscan_batch.argtypes = [_i32, _i32, _vp, _vp, _vp, _i32, _i64, _i32, _vp, _vp, _vp]
        L.fh_mst_batch.restype = ctypes.c_int
        L.fh_mst_batch.argtypes = [_i32, _i32, _vp, _vp, _vp, _i32, _vp, _vp, _vp, _vp, _vp]
        L.fh_sort_mst_device.restype = ctypes.c_int
        L.fh_sort_mst_device.argtypes = [_vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp]
        L.fh_edges_to_tree_order.restype = ctypes.c_int
        L.fh_edges_to_tree_order.argtypes = [_vp, _i64, _i64, _vp, _vp, _vp, _vp, _vp]
        L.fh_linkage_plan.restype = ctypes.c_int
        L.fh_linkage_plan.argtypes = [_vp, _i64, _i64, _vp, _vp, _vp, _vp, ctypes.POINTER(ctypes.c_int64)]
        _lib = L
    return _lib


class _Context:
    """One library context (a HIP stream and its cached device buffers),
    owned by one Python thread.  A context is not thread-safe and ctypes
    releases the GIL during each call, so threads never share one; a
    thread's contexts are released when the thread ends."""

    def __init__(self, L, device: int):
        self._L = L
        self.handle = L.fh_create(int(device))

    def __del__(self):
        if self.handle and not _exiting:
            self._L.fh_destroy(self.handle)
            self.handle = None


_tls = threading.local()
_exiting = False


@atexit.register
def _at_exit():
    # the HIP runtime tears itself down at exit; contexts die with it
    global _exiting
    _exiting = True


def context(device: int = 0):
    """The calling thread's context on `device` (created on first use)."""
    L = load()
    per_thread = getattr(_tls, "contexts", None)
    if per_thread is None:
        per_thread = _tls.contexts = {}
    c = per_thread.get(device)
    if c is None:
        c = _Context(L, device)
        if not c.handle:
            raise FHError(f"fh_create({device}) failed: no usable HIP device")
        per_thread[device] = c
    return c.handle


def check(ctx, code: int) -> None:
    if code == FH_OK:
        return
    msg = load().fh_last_error(ctx).decode()
    if code == FH_EINVAL:
        raise ValueError(msg)
    if code == FH_EUNSUPPORTED:
        raise NotImplementedError(msg)
    if code == FH_ENOMEM:
        raise MemoryError(msg)
    raise FHError(msg)


def ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)
