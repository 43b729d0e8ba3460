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
