"""Pins the FP64 MLP checker by central finite differences (CPU)."""
import numpy as np

import mlp_ref


def test_mlp_ref_gradient_finite_differences():
    rng = np.random.default_rng(0)
    p = rng.normal(0, 0.3, mlp_ref.nparams())
    X = rng.normal(0, 1, (17, 8))
    y = rng.integers(0, 2, 17)
    w = rng.uniform(0.1, 2.0, 17)
    loss, g = mlp_ref.loss_and_grad(p, X, y, w)
    idx = rng.choice(p.size, 200, replace=False)
    h = 1e-6
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        num = (mlp_ref.loss_and_grad(p + e, X, y, w)[0] - mlp_ref.loss_and_grad(p - e, X, y, w)[0]) / (2 * h)
        assert abs(num - g[i]) <= 1e-6 + 1e-5 * abs(num), (i, num, g[i])
    assert loss > 0


def test_mlp_ref_layout_count():
    assert mlp_ref.nparams() == 9026
