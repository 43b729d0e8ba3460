"""Pins for the learned-radius network's oracle (CPU)."""
import numpy as np
import torch

from oracle import mlp as M


def _setup(F=6, B=32, seed=0):
    rng = np.random.default_rng(seed)
    p = rng.normal(0, 0.3, M.param_count(F))
    x = rng.normal(size=(B, F))
    t = rng.normal(size=B)
    return p, x, t


def test_gradient_matches_finite_differences():
    p, x, t = _setup()
    _, g, _ = M.loss_and_grad(p, x, t)
    rng = np.random.default_rng(1)
    idx = rng.choice(p.shape[0], 60, replace=False)
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (M.loss_and_grad(pp, x, t)[0] - M.loss_and_grad(pm, x, t)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd)


def test_matches_torch_autograd_and_adam():
    F = 8
    p, x, t = _setup(F=F, B=40, seed=3)
    loss, g, y = M.loss_and_grad(p, x, t)
    W1, b1, W2, b2, W3, b3 = (torch.tensor(a, dtype=torch.float64, requires_grad=True) for a in M.unpack(p, F))
    xt = torch.tensor(x)
    out = torch.relu(torch.relu(xt @ W1.T + b1) @ W2.T + b2) @ W3 + b3
    lt = ((out - torch.tensor(t)) ** 2).mean()
    lt.backward()
    tg = torch.cat([W1.grad.ravel(), b1.grad, W2.grad.ravel(), b2.grad, W3.grad, b3.grad.reshape(1)]).numpy()
    np.testing.assert_allclose(g, tg, rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(loss, lt.item(), rtol=1e-12)
    # Adam, first step
    params = [torch.tensor(a, dtype=torch.float64, requires_grad=True) for a in M.unpack(p, F)]
    opt = torch.optim.Adam(params, lr=1e-3, betas=(0.9, 0.999), eps=1e-8)
    W1, b1, W2, b2, W3, b3 = params
    out = torch.relu(torch.relu(xt @ W1.T + b1) @ W2.T + b2) @ W3 + b3
    ((out - torch.tensor(t)) ** 2).mean().backward()
    opt.step()
    tp = torch.cat([W1.ravel(), b1, W2.ravel(), b2, W3, b3.reshape(1)]).detach().numpy()
    np.testing.assert_allclose(M.adam_first_step(p, g, 1e-3), tp, rtol=1e-12, atol=1e-14)
