"""CPU oracle for the fast_hdbscan Euclidean MST path -- TEST INFRASTRUCTURE.

Only tests/, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` may import this module.  It is the checker the HIP product path
is compared against, never part of the product.

The numerics live in ``fh_oracle.c`` (compiled by ``oracle/Makefile`` into
``oracle/_build/libfh_oracle.so``); this file is a thin ctypes wrapper that
returns arrays with the reference's dtypes and layout.

Reference functions restated (fast_hdbscan 0.3.2) and what pins each one:

=====================================  =====================================
restated function                      pinned by (tests/test_oracle.py)
=====================================  =====================================
numba_kdtree.build_kdtree 478-521      closed-form level/node counts;
                                       leaf sets, idx_start/idx_end and node
                                       bounds equal sklearn.neighbors.KDTree
                                       (same level formula, split rule and
                                       (value, index) median order); median
                                       invariant at every split.  Within-leaf
                                       ORDER (introselect swap sequence):
                                       two hand traces of _introselect_impl
                                       (partition then insertion sort; pivot
                                       landing on nth).  The depth-limit
                                       heapsort branch is pinned only by its
                                       sortedness (unique keys: any sort).
numba_kdtree.parallel_tree_query       brute-force kNN (same float32 rdist) on
774-792 (+ heap push / deheap_sort)    tie-free data: identical distances and
                                       indices; core distances vs float64
                                       sklearn KDTree within 1e-5 relative.
                                       Heap tie layout: two hand traces of
                                       simple_heap_push / deheap_sort (all
                                       points equal; partial ties where a
                                       stable sort would differ).
boruvka.parallel_boruvka 552-647       MST weight multiset equal to
(+ merge_components, init, disjoint    scipy.sparse.csgraph MST on the dense
set)                                   float32 mutual-reachability matrix;
                                       spanning-tree validity; the per-round
                                       edge characterisation (lowest index at
                                       the minimum, first partner in full DFS
                                       order) checked by an unpruned search.
                                       Row order / orientation under ties:
                                       the duplicate-point fixture by hand
                                       (star to point 4, rows in point
                                       order); elsewhere the restatement.
=====================================  =====================================

The float32 accumulation form of rdist (fma vs separate multiply/add) is not
pinned by any reference test; ``set_rdist_mode`` switches it.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libfh_oracle.so")
_lib = None

RDIST_FMA_SEQ = 0
RDIST_MULADD_SEQ = 1

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")


def build_library() -> str:
    """Compile the oracle with its Makefile; returns the library path."""
    import subprocess

    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build_library()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_tree_nnodes.restype = ctypes.c_int64
        L.orc_tree_nnodes.argtypes = [ctypes.c_int64, ctypes.c_int64]
        L.orc_tree_nlevels.restype = ctypes.c_int64
        L.orc_tree_nlevels.argtypes = [ctypes.c_int64, ctypes.c_int64]
        L.orc_set_rdist_mode.argtypes = [ctypes.c_int]
        L.orc_get_rdist_mode.restype = ctypes.c_int
        L.orc_build_kdtree.restype = ctypes.c_int
        L.orc_build_kdtree.argtypes = [_f32p, ctypes.c_int64, ctypes.c_int, ctypes.c_int64,
                                       _i64p, _i64p, _i64p, _f32p, _u8p, _f32p]
        tree_args = [_f32p, ctypes.c_int64, ctypes.c_int, _i64p, _i64p, _i64p, _u8p, _f32p, ctypes.c_int64]
        L.orc_knn.restype = ctypes.c_int
        L.orc_knn.argtypes = tree_args + [_f32p, ctypes.c_int64, ctypes.c_int, _f32p, _i32p, ctypes.c_int]
        L.orc_knn_brute.restype = ctypes.c_int
        L.orc_knn_brute.argtypes = [_f32p, ctypes.c_int64, ctypes.c_int, ctypes.c_int, _f32p, _i32p, ctypes.c_int]
        L.orc_boruvka.restype = ctypes.c_int64
        L.orc_boruvka.argtypes = tree_args + [ctypes.c_int, _i64p, _i64p, _f32p, _f32p, _i32p, _f32p,
                                              ctypes.POINTER(ctypes.c_int32), ctypes.c_int]
        L.orc_boruvka_racy.restype = ctypes.c_int64
        L.orc_boruvka_racy.argtypes = L.orc_boruvka.argtypes
        L.orc_weighted_k.restype = ctypes.c_int64
        L.orc_weighted_k.argtypes = [_f32p, ctypes.c_int64, ctypes.c_int]
        L.orc_boruvka_weighted.restype = ctypes.c_int64
        L.orc_boruvka_weighted.argtypes = tree_args + [ctypes.c_int, _f32p, ctypes.c_int, _i64p, _i64p, _f32p, _f32p,
                                                       _i32p, _f32p, ctypes.POINTER(ctypes.c_int32), ctypes.c_int]
        L.orc_knn_sqrt.restype = ctypes.c_int
        L.orc_knn_sqrt.argtypes = tree_args + [_f32p, ctypes.c_int64, ctypes.c_int, _f32p, _i32p, ctypes.c_int]
        L.orc_rdist.restype = ctypes.c_float
        L.orc_rdist.argtypes = [_f32p, _f32p, ctypes.c_int]
        L.orc_rdist_rows.restype = None
        L.orc_rdist_rows.argtypes = [_f32p, ctypes.c_int64, _f32p, ctypes.c_int64, ctypes.c_int, _f32p, ctypes.c_int]
        L.orc_rdist_pairs.restype = None
        L.orc_rdist_pairs.argtypes = [_f32p, ctypes.c_int, _i32p, _i32p, ctypes.c_int64, _f32p]
        _f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
        L.orc_labels.restype = ctypes.c_int64
        L.orc_labels.argtypes = [ctypes.c_int64, _f64p, ctypes.c_int64, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                 _i64p, _f32p, _f64p, _i64p, _i64p, _f32p, _f32p, ctypes.POINTER(ctypes.c_int64)]
        L.orc_box_lb.restype = ctypes.c_float
        L.orc_box_lb.argtypes = [_f32p, _f32p, _f32p, ctypes.c_int]
        _lib = L
    return _lib


def set_rdist_mode(mode: int) -> None:
    lib().orc_set_rdist_mode(int(mode))


def rdist_rows(Q, X, nthreads=None) -> np.ndarray:
    """float32 rdist (numba_kdtree.py:539-546) from every row of Q to every
    row of X: [m, n]."""
    Q = np.ascontiguousarray(Q, dtype=np.float32)
    X = np.ascontiguousarray(X, dtype=np.float32)
    out = np.empty((Q.shape[0], X.shape[0]), np.float32)
    lib().orc_rdist_rows(Q, Q.shape[0], X, X.shape[0], X.shape[1], out, _threads(nthreads))
    return out


def rdist_pairs(X, a, b) -> np.ndarray:
    """float32 rdist of the row pairs (X[a[i]], X[b[i]])."""
    X = np.ascontiguousarray(X, dtype=np.float32)
    a = np.ascontiguousarray(a, dtype=np.int32)
    b = np.ascontiguousarray(b, dtype=np.int32)
    out = np.empty(a.shape[0], np.float32)
    lib().orc_rdist_pairs(X, X.shape[1], a, b, a.shape[0], out)
    return out


def _threads(nthreads):
    return int(nthreads) if nthreads else (os.cpu_count() or 1)


@dataclass
class Tree:
    """numba_kdtree.NumbaKDTree (numba_kdtree.py:7-10) as plain arrays."""

    data: np.ndarray  # float32 [n, d]
    idx_array: np.ndarray  # int64 [n]
    idx_start: np.ndarray  # int64 [N]
    idx_end: np.ndarray  # int64 [N]
    radius: np.ndarray  # float32 [N]
    is_leaf: np.ndarray  # bool [N]
    node_bounds: np.ndarray  # float32 [2, N, d]

    @property
    def n_nodes(self) -> int:
        return self.idx_start.shape[0]

    def _args(self):
        X = self.data
        return (X, X.shape[0], X.shape[1], self.idx_array, self.idx_start, self.idx_end,
                self.is_leaf.view(np.uint8), self.node_bounds, self.n_nodes)


def tree_nlevels(n: int, leaf_size: int = 40) -> int:
    return int(lib().orc_tree_nlevels(n, leaf_size))


def tree_nnodes(n: int, leaf_size: int = 40) -> int:
    return int(lib().orc_tree_nnodes(n, leaf_size))


def build_kdtree(data, leaf_size: int = 40) -> Tree:
    """numba_kdtree.build_kdtree (numba_kdtree.py:478-521)."""
    X = np.ascontiguousarray(np.asarray(data).astype(np.float32))
    if leaf_size < 1:
        raise ValueError("leaf_size must be greater than or equal to 1")
    n, d = X.shape
    N = tree_nnodes(n, leaf_size)
    idx = np.empty(n, np.int64)
    s = np.empty(N, np.int64)
    e = np.empty(N, np.int64)
    rad = np.empty(N, np.float32)
    leaf = np.empty(N, np.uint8)
    bounds = np.empty((2, N, d), np.float32)
    lib().orc_build_kdtree(X, n, d, leaf_size, idx, s, e, rad, leaf, bounds)
    return Tree(X, idx, s, e, rad, leaf.astype(bool), bounds)


def tree_query(tree: Tree, queries, k: int, nthreads=None):
    """numba_kdtree.parallel_tree_query(tree, queries, k, output_rdist=True)
    (numba_kdtree.py:774-792): squared float32 distances, ascending."""
    Q = np.ascontiguousarray(np.asarray(queries, dtype=np.float32))
    nq = Q.shape[0]
    out_d = np.empty((nq, k), np.float32)
    out_i = np.empty((nq, k), np.int32)
    lib().orc_knn(*tree._args(), Q, nq, k, out_d, out_i, _threads(nthreads))
    return out_d, out_i


def tree_query_sqrt(tree: Tree, queries, k: int, nthreads=None):
    """parallel_tree_query(..., output_rdist=False) (numba_kdtree.py:
    788-792): float32 square roots taken before deheap_sort."""
    Q = np.ascontiguousarray(np.asarray(queries, dtype=np.float32))
    nq = Q.shape[0]
    out_d = np.empty((nq, k), np.float32)
    out_i = np.empty((nq, k), np.int32)
    lib().orc_knn_sqrt(*tree._args(), Q, nq, k, out_d, out_i, _threads(nthreads))
    return out_d, out_i


def sample_weight_core_distance(distances, neighbors, sample_weights, min_samples):
    """boruvka.py:527-539: float64 running weight, first column reaching
    min_samples; a -1 neighbour reads the last weight (numba wraps)."""
    n, k = distances.shape
    out = np.zeros(n, np.float32)
    for i in range(n):
        total = 0.0
        j = 0
        while total < min_samples and j < k:
            total += float(sample_weights[neighbors[i, j]])
            j += 1
        out[i] = distances[i, j - 1]
    return out


def knn_brute(data, k: int, nthreads=None):
    """Brute-force kNN, same float32 rdist, rows sorted by (distance, index)."""
    X = np.ascontiguousarray(np.asarray(data, dtype=np.float32))
    n, d = X.shape
    out_d = np.empty((n, k), np.float32)
    out_i = np.empty((n, k), np.int32)
    lib().orc_knn_brute(X, n, d, k, out_d, out_i, _threads(nthreads))
    return out_d, out_i


@dataclass
class BoruvkaResult:
    src: np.ndarray  # int64 [m]
    dst: np.ndarray  # int64 [m]
    w_sq: np.ndarray  # float32 [m] squared MRD
    core_sq: np.ndarray  # float32 [n]
    knn_i: np.ndarray  # int32 [n, kq]
    knn_d: np.ndarray  # float32 [n, kq]
    rounds: int

    def edges(self) -> np.ndarray:
        """float64 [m, 3] rows (src, dst, sqrt(w)) as boruvka.py:646-647."""
        out = np.empty((self.src.shape[0], 3), np.float64)
        out[:, 0] = self.src
        out[:, 1] = self.dst
        out[:, 2] = np.sqrt(self.w_sq.astype(np.float64))
        return out


def boruvka(tree: Tree, min_samples: int = 10, nthreads=None) -> BoruvkaResult:
    """boruvka.parallel_boruvka without sample weights (boruvka.py:552-647),
    executed with sequential (one-thread) semantics.  The kNN step may use
    several threads: each query row is independent and deterministic."""
    n = tree.data.shape[0]
    kq = min_samples + 1 if min_samples > 1 else 2
    m = max(n - 1, 0) + 1
    src = np.empty(m, np.int64)
    dst = np.empty(m, np.int64)
    w = np.empty(m, np.float32)
    core = np.empty(n, np.float32)
    ki = np.empty((n, kq), np.int32)
    kd = np.empty((n, kq), np.float32)
    rounds = ctypes.c_int32(0)
    cnt = lib().orc_boruvka(*tree._args(), int(min_samples), src, dst, w, core, ki, kd,
                            ctypes.byref(rounds), _threads(nthreads))
    return BoruvkaResult(src[:cnt], dst[:cnt], w[:cnt], core, ki, kd, int(rounds.value))


def boruvka_racy(tree: Tree, min_samples: int = 10, nthreads=None) -> BoruvkaResult:
    """parallel_boruvka as the reference runs it on several cores: kNN rows
    and each round's per-point searches spread over nthreads, the
    per-component bounds shared without synchronisation (boruvka.py:300-331
    under prange).  Tie choices and row order can differ from ``boruvka``;
    for timing the CPU path, never as a parity reference."""
    n = tree.data.shape[0]
    kq = min_samples + 1 if min_samples > 1 else 2
    m = max(n - 1, 0) + 1
    src = np.empty(m, np.int64)
    dst = np.empty(m, np.int64)
    w = np.empty(m, np.float32)
    core = np.empty(n, np.float32)
    ki = np.empty((n, kq), np.int32)
    kd = np.empty((n, kq), np.float32)
    rounds = ctypes.c_int32(0)
    cnt = lib().orc_boruvka_racy(*tree._args(), int(min_samples), src, dst, w, core, ki, kd,
                                 ctypes.byref(rounds), _threads(nthreads))
    return BoruvkaResult(src[:cnt], dst[:cnt], w[:cnt], core, ki, kd, int(rounds.value))


def boruvka_weighted(tree: Tree, min_samples: int, sample_weights, nthreads=None) -> BoruvkaResult:
    """parallel_boruvka's sample-weight branch (boruvka.py:564-578, 527-539)
    with one-thread semantics."""
    n = tree.data.shape[0]
    w = np.ascontiguousarray(np.asarray(sample_weights, dtype=np.float32))
    kq = int(lib().orc_weighted_k(w, n, int(min_samples)))
    if kq < 1:
        raise ValueError("sample weights give a kNN width below 1")
    m = max(n - 1, 0) + 1
    src = np.empty(m, np.int64)
    dst = np.empty(m, np.int64)
    ww = np.empty(m, np.float32)
    core = np.empty(n, np.float32)
    ki = np.empty((n, kq), np.int32)
    kd = np.empty((n, kq), np.float32)
    rounds = ctypes.c_int32(0)
    cnt = lib().orc_boruvka_weighted(*tree._args(), int(min_samples), w, kq, src, dst, ww, core, ki, kd,
                                     ctypes.byref(rounds), _threads(nthreads))
    return BoruvkaResult(src[:cnt], dst[:cnt], ww[:cnt], core, ki, kd, int(rounds.value))


def compute_minimum_spanning_tree(data, min_samples: int = 10, nthreads=None, sample_weights=None):
    """The Euclidean Boruvka branch of hdbscan.compute_minimum_spanning_tree
    (hdbscan.py:466,482-494): returns (edges f64[n-1,3], neighbors i32[n,ms'],
    core_distances f32[n])."""
    tree = build_kdtree(data)
    if sample_weights is None:
        r = boruvka(tree, min_samples, nthreads)
    else:
        r = boruvka_weighted(tree, min_samples, sample_weights, nthreads)
    core = np.sqrt(r.core_sq)
    return r.edges(), r.knn_i[:, 1:], core
