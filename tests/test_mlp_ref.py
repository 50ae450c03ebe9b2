"""Pins the FP64 MLP checker by central finite differences (CPU), for both the
cross-entropy and the expected-cost loss."""
import numpy as np

import mlp_ref


def _fd_check(f, p, rng):
    loss, g = f(p)
    idx = rng.choice(p.size, 200, replace=False)
    h = 1e-6
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        num = (f(p + e)[0] - f(p - e)[0]) / (2 * h)
        assert abs(num - g[i]) <= 1e-6 + 1e-5 * abs(num), (i, num, g[i])
    return loss


def test_mlp_ref_gradient_finite_differences():
    rng = np.random.default_rng(0)
    p = rng.normal(0, 0.3, mlp_ref.nparams())
    X = rng.normal(0, 1, (17, 8))
    y = rng.integers(0, mlp_ref.OUT, 17)
    w = rng.uniform(0.1, 2.0, 17)
    assert _fd_check(lambda q: mlp_ref.loss_and_grad(q, X, y, w), p, rng) > 0


def test_mlp_ref_cost_gradient_finite_differences():
    rng = np.random.default_rng(1)
    p = rng.normal(0, 0.3, mlp_ref.nparams())
    X = rng.normal(0, 1, (13, 8))
    cost = rng.uniform(0.0, 3.0, (13, mlp_ref.OUT))
    loss = _fd_check(lambda q: mlp_ref.cost_loss_and_grad(q, X, cost, 7.0), p, rng)
    # the expected cost lies between the per-sample minima and maxima
    assert cost.min(1).sum() / 7.0 <= loss <= cost.max(1).sum() / 7.0


def test_mlp_ref_cost_loss_closed_forms():
    """one-hot policy: the loss is the chosen actions' cost; uniform policy: the mean"""
    rng = np.random.default_rng(2)
    cost = rng.uniform(0, 1, (5, mlp_ref.OUT))
    p = np.zeros(mlp_ref.nparams())  # all weights zero: uniform softmax
    loss, _ = mlp_ref.cost_loss_and_grad(p, rng.normal(0, 1, (5, 8)), cost, 1.0)
    assert abs(loss - cost.mean(1).sum()) < 1e-12
    W1, b1, W2, b2, W3, b3, W4, b4 = mlp_ref.unpack(p)
    b4[3] = 60.0  # views into p: action 3 dominates
    loss, _ = mlp_ref.cost_loss_and_grad(p, rng.normal(0, 1, (5, 8)), cost, 1.0)
    assert abs(loss - cost[:, 3].sum()) < 1e-9


def test_mlp_ref_layout_count():
    assert mlp_ref.nparams() == 9416
