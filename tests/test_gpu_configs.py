"""The benchmark configurations on the device, checked against the oracle.

- 2-D scale workload at its full size (n = 2,000,000, min_samples 10):
  core distances, neighbour columns and MST rows bit for bit against the
  oracle's one-thread parallel_boruvka (boruvka.py:542-647), then labels and
  membership strengths at min_cluster_size 100 against the oracle's label
  stage (hdbscan.py:497-574).
- 3-D mixture workload (200 anisotropic Gaussians + 5% noise, seed 2,
  min_samples 15) at 300,000 points: the same comparisons, plain and with
  learned search radii (the network trained once, then reused).
- Full sizes beyond what the oracle finishes in seconds (10M x 3-D and
  1M x 64): size-independent properties -- n - 1 rows forming a spanning
  tree, each sampled row's weight equal to max(rdist, core_src, core_dst)
  in the oracle's float32 arithmetic, sampled core distances equal to the
  brute-force (min_samples + 1)-th rdist, sampled neighbour sets equal to
  brute force, and the cycle property of the MST for sampled kNN pairs.
"""
import numpy as np
import pytest

from fast_hdbscan_amd import workloads

pytestmark = pytest.mark.gpu


def _device_arrays(X, ms):
    from fast_hdbscan_amd.hdbscan import mst_arrays

    stats = {}
    out = mst_arrays(X, ms, stats=stats)
    return out, stats


def _compare_rows(out, ref):
    src, dst, w, core, nbr = out
    np.testing.assert_array_equal(core, ref.core_sq)
    np.testing.assert_array_equal(nbr, ref.knn_i[:, 1:])
    np.testing.assert_array_equal(w, ref.w_sq)
    np.testing.assert_array_equal(src, ref.src)
    np.testing.assert_array_equal(dst, ref.dst)


def _oracle_labels(ref, mcs):
    # the compiled label oracle (pinned row for row against the Python one
    # in tests/test_labels_oracle.py); the Python one takes ~20 s at 2M
    from oracle import labels as LB

    lab, prob, *_ = LB.clusters_from_spanning_tree_c(ref.edges(), min_cluster_size=mcs)
    return lab, prob


def test_scale_2d_full_size_vs_oracle():
    import torch

    from fast_hdbscan_amd.estimator import hdbscan_device
    from oracle import oracle as O

    X = workloads.scale_2d(2_000_000, seed=1)
    print("case 2-D scale workload n=2000000 ms=10 mcs=100", flush=True)
    out, stats = _device_arrays(X, 10)
    ref = O.boruvka(O.build_kdtree(X), 10)
    _compare_rows(out, ref)
    assert stats["rounds"] == ref.rounds
    # the bench's path: device MST, device sort, host label stage
    lab, prob = hdbscan_device(torch.from_numpy(X).cuda(), 10, 100)
    olab, oprob = _oracle_labels(ref, 100)
    np.testing.assert_array_equal(lab, olab)
    np.testing.assert_array_equal(prob, oprob)


@pytest.mark.parametrize("mode", [0, 2])
def test_mixture_3d_vs_oracle(mode):
    import torch

    from fast_hdbscan_amd import _lib
    from fast_hdbscan_amd.estimator import hdbscan_device
    from oracle import oracle as O

    X = workloads.anisotropic_mixture(300_000, 3, 200, 0.05, seed=2)
    print(f"case 3-D mixture n=300000 ms=15 learned mode {mode}", flush=True)
    ctx = _lib.context(0)
    L = _lib.load()
    _lib.check(ctx, L.fh_set_learned_radius(ctx, mode))
    try:
        if mode == 2:
            _device_arrays(X, 15)  # the one training
        out, stats = _device_arrays(X, 15)
        assert stats["learned"] == mode
        ref = O.boruvka(O.build_kdtree(X), 15)
        _compare_rows(out, ref)
        assert stats["rounds"] == ref.rounds
        lab, prob = hdbscan_device(torch.from_numpy(X).cuda(), 15, 100)
    finally:
        _lib.check(ctx, L.fh_set_learned_radius(ctx, 0))
    olab, oprob = _oracle_labels(ref, 100)
    np.testing.assert_array_equal(lab, olab)
    np.testing.assert_array_equal(prob, oprob)


def _spanning(n, src, dst):
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import connected_components

    g = coo_matrix((np.ones(src.shape[0], np.int8), (src, dst)), shape=(n, n))
    return connected_components(g, directed=False)[0] == 1


def _properties(X, ms, n_core_samples, rng):
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import connected_components

    from oracle import oracle as O

    n = X.shape[0]
    (src, dst, w, core, nbr), stats = _device_arrays(X, ms)
    print(f"  rounds={stats['rounds']} high_dim={stats['high_dim']}", flush=True)
    assert src.shape == (n - 1,)
    assert src.min() >= 0 and dst.min() >= 0 and src.max() < n and dst.max() < n
    assert not np.any(src == dst)
    assert _spanning(n, src, dst)
    # every row's weight is its mutual-reachability weight (sample of rows)
    rows = rng.choice(n - 1, 200_000, replace=False)
    rd = O.rdist_pairs(X, src[rows], dst[rows])
    want = np.maximum(rd, np.maximum(core[src[rows]], core[dst[rows]]))
    np.testing.assert_array_equal(w[rows], want)
    # core distances and neighbour sets by brute force
    pts = rng.choice(n, n_core_samples, replace=False)
    for chunk in np.array_split(pts, max(1, n_core_samples // 8)):
        D = O.rdist_rows(X[chunk], X)
        part = np.partition(D, ms, axis=1)[:, :ms + 1]
        srt = np.sort(part, axis=1)
        np.testing.assert_array_equal(core[chunk], srt[:, ms])
        for r, p in enumerate(chunk):
            kth, nxt = np.partition(D[r], [ms, ms + 1])[[ms, ms + 1]]
            if kth < nxt:  # no tie at the boundary: the k+1 nearest are unique
                brute = set(np.nonzero(D[r] <= kth)[0].tolist())
                got = set(nbr[p].tolist())  # columns 1..ms (column 0 not returned)
                assert len(got) == ms and got <= brute
        del D
    # cycle property: for a kNN pair (i, j) the rows of weight <= w_ij
    # already join i and j (otherwise (i, j) would beat a heavier row)
    order = np.argsort(w, kind="stable")
    ws = w[order]
    for p in rng.choice(n, 4, replace=False):
        j = int(nbr[p, ms - 1])
        wij = max(O.rdist_pairs(X, [p], [j])[0], core[p], core[j])
        m = np.searchsorted(ws, wij, side="right")
        g = coo_matrix((np.ones(m, np.int8), (src[order[:m]], dst[order[:m]])), shape=(n, n))
        lab = connected_components(g, directed=False)[1]
        assert lab[p] == lab[j]


def test_mixture_3d_full_size_properties():
    X = workloads.scale_3d()
    print(f"case 3-D mixture full size n={X.shape[0]} ms=15", flush=True)
    _properties(X, 15, 32, np.random.default_rng(0))


def test_highdim_full_size_properties():
    X = workloads.gaussian_highdim(1_000_000, 64, 100, seed=3)
    print("case 1M x 64 ms=10", flush=True)
    _properties(X, 10, 32, np.random.default_rng(1))
