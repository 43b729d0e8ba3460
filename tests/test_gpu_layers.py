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
