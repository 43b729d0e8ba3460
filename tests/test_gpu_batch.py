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
