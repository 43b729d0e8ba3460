"""Learned search radii on the device: the network's gradients and Adam step
against the oracle, and MST outputs identical with and without the learned
ball bounds."""
import ctypes

import numpy as np
import pytest

from fast_hdbscan_amd import _lib, workloads
from fast_hdbscan_amd.hdbscan import mst_arrays

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("F,B", [(6, 256), (8, 1024), (18, 100)])
def test_mlp_step_matches_oracle(F, B):
    from oracle import mlp as M

    rng = np.random.default_rng(F)
    p = rng.normal(0, 0.2, M.param_count(F)).astype(np.float32)
    x = rng.normal(size=(B, F)).astype(np.float32)
    t = rng.normal(size=B).astype(np.float32)
    params = p.copy()
    grad = np.empty_like(p)
    pred = np.empty(B, np.float32)
    ctx = _lib.context(0)
    code = _lib.load().fh_mlp_step(ctx, F, B, _lib.ptr(x), _lib.ptr(t), _lib.ptr(params), _lib.ptr(grad),
                                   _lib.ptr(pred), ctypes.c_float(1e-3))
    _lib.check(ctx, code)
    loss, g, y = M.loss_and_grad(p.astype(np.float64), x.astype(np.float64), t.astype(np.float64))
    # float32 sums over at most 1024 samples / 64 units: relative 1e-4 of the
    # gradient's scale
    np.testing.assert_allclose(pred, y, rtol=1e-4, atol=1e-4 * np.abs(y).max())
    np.testing.assert_allclose(grad, g, rtol=1e-3, atol=1e-4 * np.abs(g).max())
    upd = M.adam_first_step(p.astype(np.float64), grad.astype(np.float64), 1e-3)
    np.testing.assert_allclose(params, upd, rtol=1e-5, atol=1e-6)


def _with_mode(mode):
    ctx = _lib.context(0)
    _lib.check(ctx, _lib.load().fh_set_learned_radius(ctx, mode))


@pytest.mark.parametrize("d,ms", [(2, 10), (3, 15), (5, 5)])
def test_learned_radius_mst_identical(d, ms):
    X = workloads.anisotropic_mixture(200000, d, 30, 0.1, seed=d)
    try:
        _with_mode(0)
        plain = mst_arrays(X, ms)
        _with_mode(1)
        st = {}
        learned = mst_arrays(X, ms, stats=st)
        _with_mode(2)
        st2 = {}
        reused = mst_arrays(X, ms, stats=st2)
    finally:
        _with_mode(0)
    for a, b, c in zip(plain, learned, reused):
        np.testing.assert_array_equal(a, b)
        np.testing.assert_array_equal(a, c)
    hit = 1.0 - st["ball_misses"] / X.shape[0]
    print(f"d={d} ms={ms} learned ball hit rate {hit:.3f}, train {st['t_learn_ms']:.2f} ms, "
          f"predict {st['t_predict_ms']:.2f} ms, kNN {st['t_knn_ms']:.2f} ms", flush=True)
    assert st["learned"] == 1 and st2["learned"] == 2
    assert hit > 0.5


def test_learned_radius_with_duplicates():
    rng = np.random.default_rng(9)
    X = rng.integers(0, 200, size=(50000, 2)).astype(np.float32)  # many exact ties
    try:
        _with_mode(0)
        plain = mst_arrays(X, 8)
        _with_mode(1)
        learned = mst_arrays(X, 8)
    finally:
        _with_mode(0)
    for a, b in zip(plain, learned):
        np.testing.assert_array_equal(a, b)
