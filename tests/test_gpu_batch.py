"""The batch scheduler against one-at-a-time calls and the oracle."""
import numpy as np
import pytest

from fast_hdbscan_amd import workloads

pytestmark = pytest.mark.gpu


def test_batch_matches_single_calls_and_oracle():
    from fast_hdbscan_amd.batch import hdbscan_batch
    from fast_hdbscan_amd.estimator import fast_hdbscan
    from oracle import labels as OL
    from oracle import oracle as O

    data = workloads.batch_datasets(24, n_hi=8000)
    data.append(np.zeros((1, 3), np.float32))
    data.append(np.random.default_rng(0).normal(size=(2, 5)).astype(np.float32))
    print("batch of", len(data), flush=True)
    out = hdbscan_batch(data, min_samples=5, min_cluster_size=20, workers=6)
    for i, (X, (lab, pr)) in enumerate(zip(data, out)):
        if X.shape[0] < 2:
            assert lab.tolist() == [-1] * X.shape[0]
            continue
        l1, p1 = fast_hdbscan(X, min_samples=5, min_cluster_size=20)
        np.testing.assert_array_equal(lab, l1)
        np.testing.assert_array_equal(pr, p1)
        if i < 6:
            e, _, _ = O.compute_minimum_spanning_tree(X, 5)
            ref = OL.clusters_from_spanning_tree(e, min_cluster_size=20)
            np.testing.assert_array_equal(lab, ref[0])
            np.testing.assert_array_equal(pr, ref[1])


def test_consecutive_batch_calls_reuse_workers():
    # the workers' contexts live across calls: later calls with fewer,
    # more and larger datasets give what one-at-a-time calls give
    from fast_hdbscan_amd.batch import hdbscan_batch
    from fast_hdbscan_amd.estimator import fast_hdbscan

    small = workloads.batch_datasets(10, n_hi=3000)
    large = workloads.batch_datasets(6, n_hi=30000)[::-1]
    for data, w in ((small, 4), (large, 8), (small[:3], 2), (large + small, 8)):
        print(f"batch call: {len(data)} datasets, {w} workers", flush=True)
        out = hdbscan_batch(data, min_samples=5, min_cluster_size=20, workers=w)
        for X, (lab, pr) in zip(data, out):
            l1, p1 = fast_hdbscan(X, min_samples=5, min_cluster_size=20)
            np.testing.assert_array_equal(lab, l1)
            np.testing.assert_array_equal(pr, p1)


def _batch_cases(rng):
    data = []
    for n, d in [(2, 2), (3, 3), (40, 2), (41, 2), (81, 2), (82, 3), (500, 1), (1000, 2), (4032, 2), (3392, 3),
                 (2496, 5), (1792, 8), (2000, 4), (700, 6), (900, 7), (5000, 2), (2600, 5)]:
        data.append(rng.standard_normal((n, d)).astype(np.float32))
    # exact ties: an integer lattice and repeated points
    g = np.stack(np.meshgrid(np.arange(30), np.arange(30)), -1).reshape(-1, 2).astype(np.float32)
    data.append(g)
    data.append(np.repeat(rng.standard_normal((60, 3)).astype(np.float32), 5, axis=0))
    data.append(np.repeat(rng.integers(0, 4, size=(50, 2)).astype(np.float32), 3, axis=0))
    data += workloads.batch_datasets(16, n_hi=6000)
    return data


@pytest.mark.parametrize("ms", [1, 5, 10, 15])
def test_mst_batch_matches_per_call_path(ms, monkeypatch):
    # the persistent launch's rows, weights and cores against the per-call
    # device path, bit for bit, over sizes up to one work-group's tree and
    # past it, every dimension 1-8, ties and duplicates (a fixed routing
    # cap: every dataset the kernel can take goes to it)
    from fast_hdbscan_amd.batch import mst_batch
    from fast_hdbscan_amd.hdbscan import mst_arrays

    monkeypatch.setenv("FH_BATCH_MAX_N", "65536")
    data = _batch_cases(np.random.default_rng(ms))
    print(f"mst_batch: {len(data)} datasets, min_samples={ms}", flush=True)
    out = mst_batch(data, min_samples=ms)
    served = 0
    for i, (X, (src, dst, w, core, pers)) in enumerate(zip(data, out)):
        n, d = X.shape
        kq = ms + 1 if ms > 1 else 2
        levels = int(np.log2(max(1, (n - 1) / 40)) + 1)
        geometry = not (levels >= 2 and (n >> (levels - 2)) < 2)
        assert pers == (n >= 2 and n <= 65536 and kq <= 16 and geometry), (i, n, d)
        served += pers
        s1, d1, w1, c1, _ = mst_arrays(X, ms)
        np.testing.assert_array_equal(core, c1, err_msg=f"dataset {i} n={n} d={d}")
        np.testing.assert_array_equal(w, w1, err_msg=f"dataset {i} n={n} d={d}")
        np.testing.assert_array_equal(src, s1, err_msg=f"dataset {i} n={n} d={d}")
        np.testing.assert_array_equal(dst, d1, err_msg=f"dataset {i} n={n} d={d}")
    assert served >= 10


def test_mst_batch_matches_oracle(monkeypatch):
    from fast_hdbscan_amd.batch import mst_batch
    from oracle import oracle as O

    monkeypatch.setenv("FH_BATCH_MAX_N", "4000")

    data = _batch_cases(np.random.default_rng(7))[:20:2]
    out = mst_batch(data, min_samples=5)
    for X, (src, dst, w, core, pers) in zip(data, out):
        print(f"oracle case n={X.shape[0]} d={X.shape[1]} persistent={pers}", flush=True)
        ref = O.boruvka(O.build_kdtree(X), 5)
        np.testing.assert_array_equal(core, ref.core_sq)
        np.testing.assert_array_equal(w, ref.w_sq)
        np.testing.assert_array_equal(src, ref.src)
        np.testing.assert_array_equal(dst, ref.dst)


def test_mst_batch_consecutive_calls():
    # the launches' buffers live across calls: smaller, larger and mixed
    # batches after each other give the per-call path's rows
    from fast_hdbscan_amd.batch import mst_batch
    from fast_hdbscan_amd.hdbscan import mst_arrays

    rng = np.random.default_rng(3)
    small = [rng.standard_normal((int(n), 2)).astype(np.float32) for n in rng.integers(50, 400, 40)]
    large = workloads.batch_datasets(12, n_hi=4000)
    for data in (small, large, small[:3], large + small):
        print(f"mst_batch call: {len(data)} datasets", flush=True)
        for X, (src, dst, w, core, _) in zip(data, mst_batch(data, min_samples=5)):
            s1, d1, w1, c1, _ = mst_arrays(X, 5)
            np.testing.assert_array_equal(core, c1)
            np.testing.assert_array_equal(w, w1)
            np.testing.assert_array_equal(src, s1)
            np.testing.assert_array_equal(dst, d1)


def test_mst_batch_large_datasets_in_the_persistent_kernel(monkeypatch):
    # datasets above one subtree build: the top tree levels built node by
    # node by the work-group, then the subtrees (FH_BATCH_MAX_N raises the
    # routing cap for this call)
    from fast_hdbscan_amd.batch import mst_batch
    from fast_hdbscan_amd.hdbscan import mst_arrays

    monkeypatch.setenv("FH_BATCH_MAX_N", "65536")
    rng = np.random.default_rng(11)
    data = [rng.standard_normal((n, d)).astype(np.float32)
            for n, d in [(4033, 2), (9000, 2), (20000, 3), (65536, 2), (12000, 5), (30000, 8), (6000, 1), (8193, 4)]]
    data += workloads.batch_datasets(8, n_lo=5000, n_hi=40000)
    data.append(np.repeat(rng.integers(0, 30, size=(3000, 2)).astype(np.float32), 3, axis=0))
    for ms in (5, 10):
        print(f"large mst_batch: {len(data)} datasets, min_samples={ms}", flush=True)
        out = mst_batch(data, min_samples=ms)
        for i, (X, (src, dst, w, core, pers)) in enumerate(zip(data, out)):
            assert pers, (i, X.shape)
            s1, d1, w1, c1, _ = mst_arrays(X, ms)
            np.testing.assert_array_equal(core, c1, err_msg=f"dataset {i} {X.shape}")
            np.testing.assert_array_equal(w, w1, err_msg=f"dataset {i} {X.shape}")
            np.testing.assert_array_equal(src, s1, err_msg=f"dataset {i} {X.shape}")
            np.testing.assert_array_equal(dst, d1, err_msg=f"dataset {i} {X.shape}")


def test_router_splits_change_nothing(monkeypatch):
    # the measured cost model's split, everything on the streams, and
    # everything the kernel can take in it: the same rows
    from fast_hdbscan_amd.batch import mst_batch

    data = workloads.batch_datasets(24, n_hi=30000) + [np.random.default_rng(5).standard_normal((60000, 2)).astype(np.float32)]
    base = mst_batch(data, min_samples=5)
    print("model split: persistent", sum(o[4] for o in base), "of", len(data), flush=True)
    for env in ({"FH_BATCH_PERSISTENT": "0"}, {"FH_BATCH_MAX_N": "65536"}):
        for k, v in env.items():
            monkeypatch.setenv(k, v)
        out = mst_batch(data, min_samples=5)
        for k in env:
            monkeypatch.delenv(k)
        if "FH_BATCH_PERSISTENT" in env:
            assert not any(o[4] for o in out)
        else:
            assert all(o[4] for o in out)
        for a, b in zip(base, out):
            for x, y in zip(a[:4], b[:4]):
                np.testing.assert_array_equal(x, y)


def test_mst_batch_edge_cases(monkeypatch):
    # an empty list; a single point; dimensions above the kernel's (the
    # per-call path, MFMA scans included); min_samples above its heap --
    # each dataset's rows as the per-call path gives them, and the kernel
    # takes only what it can
    from fast_hdbscan_amd.batch import mst_batch
    from fast_hdbscan_amd.hdbscan import mst_arrays

    assert mst_batch([], min_samples=5) == []
    monkeypatch.setenv("FH_BATCH_MAX_N", "65536")
    rng = np.random.default_rng(21)
    data = [rng.standard_normal((1, 2)).astype(np.float32), rng.standard_normal((500, 12)).astype(np.float32),
            rng.standard_normal((3000, 20)).astype(np.float32), rng.standard_normal((700, 3)).astype(np.float32)]
    for ms, expect in ((5, [False, False, False, True]), (20, [False, False, False, False])):
        out = mst_batch(data, min_samples=ms)
        for X, (src, dst, w, core, pers), e in zip(data, out, expect):
            print(f"edge case n={X.shape[0]} d={X.shape[1]} ms={ms} persistent={pers}", flush=True)
            assert pers == e
            if X.shape[0] < 2:
                assert src.shape == (0,)
                continue
            s1, d1, w1, c1, _ = mst_arrays(X, ms)
            np.testing.assert_array_equal(core, c1)
            np.testing.assert_array_equal(w, w1)
            np.testing.assert_array_equal(src, s1)
            np.testing.assert_array_equal(dst, d1)


def test_batch_full_workload_sampled():
    # the whole 4,096-dataset workload (sizes log-uniform 500-50,000, d in
    # {2, 3, 5, 8}) in one call; a sample of 48 datasets against
    # one-at-a-time calls, every dataset's labels in range
    from fast_hdbscan_amd.batch import hdbscan_batch
    from fast_hdbscan_amd.estimator import fast_hdbscan

    data = workloads.batch_datasets(4096)
    print("full batch workload: 4096 datasets", flush=True)
    out = hdbscan_batch(data, min_samples=5, min_cluster_size=20, workers=16)
    for X, (lab, pr) in zip(data, out):
        assert lab.shape == (X.shape[0],) and lab.min() >= -1
        assert np.all((pr >= 0) & (pr <= 1))
    rng = np.random.default_rng(0)
    for i in rng.choice(len(data), 48, replace=False):
        l1, p1 = fast_hdbscan(data[i], min_samples=5, min_cluster_size=20)
        np.testing.assert_array_equal(out[i][0], l1, err_msg=f"dataset {i} {data[i].shape}")
        np.testing.assert_array_equal(out[i][1], p1, err_msg=f"dataset {i} {data[i].shape}")
