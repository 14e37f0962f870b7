"""Pins the surrogate oracle (oracle/surrogate.py) against independent routines, and checks the
library's host-side rank correlation.  CPU only."""
import numpy as np
import pytest
import scipy.stats
import torch

import surrogate as sgo


def _params(rng, scale=0.2):
    return rng.standard_normal(sgo.N_PARAMS) * scale


def _torch_forward(params, x):
    p = torch.tensor(params, dtype=torch.float64, requires_grad=True)
    w1 = p[sgo.OFF_W1:sgo.OFF_B1].reshape(sgo.IN, sgo.H1)
    b1 = p[sgo.OFF_B1:sgo.OFF_W2]
    w2 = p[sgo.OFF_W2:sgo.OFF_B2].reshape(sgo.H1, sgo.H2)
    b2 = p[sgo.OFF_B2:sgo.OFF_W3]
    w3 = p[sgo.OFF_W3:sgo.OFF_B3]
    b3 = p[sgo.OFF_B3]
    xt = torch.tensor(x, dtype=torch.float64)
    h1 = torch.nn.functional.linear(xt, w1.T, b1).relu()
    h2 = torch.nn.functional.linear(h1, w2.T, b2).relu()
    return p, h2 @ w3 + b3


def test_forward_matches_torch():
    rng = np.random.default_rng(1)
    params, x = _params(rng), rng.standard_normal((37, 192))
    _, yt = _torch_forward(params, x)
    y, _, _ = sgo.forward(params, x)
    np.testing.assert_allclose(y, yt.detach().numpy(), rtol=1e-12, atol=1e-12)


def test_gradient_matches_torch_autograd():
    rng = np.random.default_rng(2)
    params, x, t = _params(rng), rng.standard_normal((50, 192)), rng.standard_normal(50)
    p, yt = _torch_forward(params, x)
    (0.5 * ((yt - torch.tensor(t)) ** 2).mean()).backward()
    np.testing.assert_allclose(sgo.gradient(params, x, t), p.grad.numpy(), rtol=1e-10, atol=1e-12)


def test_gradient_matches_central_differences():
    rng = np.random.default_rng(3)
    params, x, t = _params(rng), rng.standard_normal((20, 192)), rng.standard_normal(20)
    g = sgo.gradient(params, x, t)
    # a sample from every block (weights and biases of all three layers)
    idx = np.concatenate([rng.integers(lo, hi, 12) for lo, hi in
                          [(sgo.OFF_W1, sgo.OFF_B1), (sgo.OFF_B1, sgo.OFF_W2), (sgo.OFF_W2, sgo.OFF_B2),
                           (sgo.OFF_B2, sgo.OFF_W3), (sgo.OFF_W3, sgo.OFF_B3)]] + [np.array([sgo.OFF_B3])])
    h = 1e-6
    for i in idx:
        e = np.zeros(sgo.N_PARAMS)
        e[i] = h
        fd = (sgo.loss(params + e, x, t) - sgo.loss(params - e, x, t)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])), (i, fd, g[i])


def test_adam_matches_torch_optimizer():
    rng = np.random.default_rng(4)
    params = _params(rng)
    p = torch.tensor(params, dtype=torch.float64, requires_grad=True)
    opt = torch.optim.Adam([p], lr=1e-3, betas=(0.9, 0.999), eps=1e-8)
    m = np.zeros(sgo.N_PARAMS)
    v = np.zeros(sgo.N_PARAMS)
    ours = params.copy()
    for step in range(1, 4):
        x, t = rng.standard_normal((64, 192)), rng.standard_normal(64)
        opt.zero_grad()
        # gradient through the oracle's own autograd-checked formula, fed to torch's optimizer
        p.grad = torch.tensor(sgo.gradient(p.detach().numpy(), x, t))
        opt.step()
        ours, m, v = sgo.adam_step(ours, m, v, step, x, t, 1e-3)
        np.testing.assert_allclose(ours, p.detach().numpy(), rtol=0, atol=1e-12)


@pytest.mark.parametrize("n", [3, 10, 257])
def test_rank_correlation_matches_scipy(pkg, n):
    rng = np.random.default_rng(n)
    a = rng.integers(0, 7, n).astype(np.float32)  # many ties
    b = (a + rng.standard_normal(n) * 2).astype(np.float32)
    b[: n // 3] = b[0]  # ties on the other side too
    expect = scipy.stats.spearmanr(a, b).statistic
    assert sgo.rank_correlation(a, b) == pytest.approx(expect, abs=1e-12)
    assert pkg.runtime.rank_correlation(a, b) == pytest.approx(expect, abs=1e-12)


def test_rank_correlation_degenerate_cases(pkg):
    rc = pkg.runtime.rank_correlation
    assert np.isnan(rc(np.array([1.0, 2.0]), np.array([2.0, 1.0])))  # fewer than 3 points
    assert np.isnan(rc(np.ones(5), np.arange(5.0)))                   # constant side
    assert np.isnan(rc(np.array([1.0, np.nan, 3.0]), np.arange(3.0)))
    assert rc(np.arange(6.0), np.arange(6.0) ** 3) == pytest.approx(1.0)
    assert rc(np.arange(6.0), -np.arange(6.0)) == pytest.approx(-1.0)
