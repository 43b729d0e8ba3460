"""MI355X (gfx950) implementation of fast_hdbscan's Euclidean MST path.

The public functions mirror the reference's seam
``fast_hdbscan.hdbscan.compute_minimum_spanning_tree`` and its building
blocks; the compute runs in hand-written HIP kernels (libfh.so).
"""
from .hdbscan import (  # noqa: F401
    build_kdtree,
    compute_minimum_spanning_tree,
    mst_arrays,
    mst_device,
    parallel_tree_query,
)

__version__ = "0.1.0"


def __getattr__(name):
    # the estimators pull in scikit-learn; load them on first use
    if name in ("HDBSCAN", "PLSCAN", "fast_hdbscan"):
        from . import estimator

        return getattr(estimator, name)
    if name in ("build_cluster_layers", "build_layer_cluster_tree"):
        from . import layers

        return getattr(layers, name)
    raise AttributeError(name)
