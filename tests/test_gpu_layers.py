"""PLSCAN layers on the device MST: the reference's PLSCAN test suite
(test_plscan.py) end to end, and the layer stage against the oracle on the
same device MST rows."""
import numpy as np
import pytest

from fast_hdbscan_amd import workloads

pytestmark = pytest.mark.gpu


def test_plscan_reference_suite_on_device():
    from fast_hdbscan_amd.estimator import PLSCAN
    from test_layers import run_reference_plscan_suite

    run_reference_plscan_suite(PLSCAN)


@pytest.mark.parametrize("ms,mcs", [(5, 5), (10, 20), (2, 10)])
def test_layers_device_mst_vs_oracle(ms, mcs):
    from fast_hdbscan_amd import compute_minimum_spanning_tree
    from fast_hdbscan_amd import layers as FL
    from oracle import layers as OL
    from oracle import oracle as O

    X = workloads.anisotropic_mixture(20000, 2, 12, 0.1, seed=ms).astype(np.float64)
    print(f"layers case ms={ms} mcs={mcs}", flush=True)
    e, nb, core = compute_minimum_spanning_tree(X, min_samples=ms)
    re, rnb, rcore = O.compute_minimum_spanning_tree(X, ms)
    np.testing.assert_array_equal(e, re)
    np.testing.assert_array_equal(core, rcore)
    ro = OL.cluster_layers_from_mst(re, rcore, X.shape[0], base_min_cluster_size=mcs, max_layers=6)
    rf = FL.layers_from_mst(e, core, X.shape[0], base_min_cluster_size=mcs, max_layers=6)
    assert len(ro[0]) == len(rf[0])
    for a, b in zip(ro[0] + ro[1], rf[0] + rf[1]):
        np.testing.assert_array_equal(a, b)
    assert [float(v) for v in ro[2]] == [float(v) for v in rf[2]]
    np.testing.assert_array_equal(ro[4], rf[4])


def test_plscan_sweep_matches_one_at_a_time():
    from fast_hdbscan_amd import compute_minimum_spanning_tree
    from fast_hdbscan_amd import layers as FL

    X = workloads.anisotropic_mixture(30000, 2, 15, 0.1, seed=4).astype(np.float32)
    values = [2, 3, 7, 16, 33]
    print("plscan sweep", values, flush=True)
    sweep = FL.plscan_sweep(X, values, base_min_cluster_size=5, workers=4)
    assert sorted(sweep) == values
    for m in values:
        e, _, core = compute_minimum_spanning_tree(X, min_samples=m)
        one = FL.layers_from_mst(e, core, X.shape[0], base_min_cluster_size=5)
        got = sweep[m]
        assert len(one[0]) == len(got[0])
        for a, b in zip(one[0] + one[1], got[0] + got[1]):
            np.testing.assert_array_equal(a, b)
        assert [float(v) for v in one[2]] == [float(v) for v in got[2]]
        np.testing.assert_array_equal(one[3], got[3])
        np.testing.assert_array_equal(one[4], got[4])


def test_plscan_sweep_high_dimension_runs_value_by_value():
    """Inputs that the MFMA scans take (here 12-D) are not served by the
    shared tree kNN of the sweep: plscan_sweep runs them value by value,
    with the same layers as one call per value."""
    from fast_hdbscan_amd import compute_minimum_spanning_tree
    from fast_hdbscan_amd import layers as FL

    X = workloads.gaussian_highdim(6000, 12, n_clusters=8, seed=41).astype(np.float32)
    values = [3, 9]
    sweep = FL.plscan_sweep(X, values, base_min_cluster_size=5, workers=2)
    for m in values:
        e, _, core = compute_minimum_spanning_tree(X, min_samples=m)
        one = FL.layers_from_mst(e, core, X.shape[0], base_min_cluster_size=5)
        got = sweep[m]
        assert len(one[0]) == len(got[0])
        for a, b in zip(one[0] + one[1], got[0] + got[1]):
            np.testing.assert_array_equal(a, b)
        np.testing.assert_array_equal(one[4], got[4])


@pytest.mark.parametrize("case", ["mixture", "grid_ties", "duplicates"])
def test_sweep_shared_knn_equals_single_mst(case):
    """fh_mst_sweep_device (one tree, one kNN of width max(min_samples)+2)
    gives, value by value, the single-MST outputs; ties in a row's shared
    columns send it back to the query at its own width (integer grids and
    duplicated points make many); two values against the oracle."""
    import torch

    from fast_hdbscan_amd.hdbscan import mst_device
    from fast_hdbscan_amd.layers import mst_sweep_device
    from oracle import oracle as O

    rng = np.random.default_rng(7)
    if case == "mixture":
        X = workloads.anisotropic_mixture(40_000, 2, 15, 0.1, seed=8)
    elif case == "grid_ties":
        g = np.stack(np.meshgrid(np.arange(150), np.arange(120)), -1).reshape(-1, 2).astype(np.float32)
        X = g[rng.permutation(g.shape[0])]
    else:
        X = workloads.anisotropic_mixture(20_000, 3, 10, 0.05, seed=9)
        X[10_000:12_000] = X[:2000]
    values = [2, 3, 5, 10, 17, 33, 64]
    print(f"sweep case {case} n={X.shape[0]} d={X.shape[1]}", flush=True)
    Xd = torch.from_numpy(np.ascontiguousarray(X, dtype=np.float32)).cuda()
    st = {}
    src, dst, w, core = mst_sweep_device(Xd, values, stats=st)
    for i, m in enumerate(values):
        s1, d1, w1, c1, _ = mst_device(Xd, m)
        np.testing.assert_array_equal(src[i].cpu().numpy(), s1.cpu().numpy())
        np.testing.assert_array_equal(dst[i].cpu().numpy(), d1.cpu().numpy())
        np.testing.assert_array_equal(w[i].cpu().numpy(), w1.cpu().numpy())
        np.testing.assert_array_equal(core[i].cpu().numpy(), c1.cpu().numpy())
    tree = O.build_kdtree(np.ascontiguousarray(X, dtype=np.float32))
    for i in (1, 4):
        ref = O.boruvka(tree, values[i])
        np.testing.assert_array_equal(src[i].cpu().numpy(), ref.src)
        np.testing.assert_array_equal(w[i].cpu().numpy(), ref.w_sq)
        np.testing.assert_array_equal(core[i].cpu().numpy(), ref.core_sq)
