"""Pins of the proposal network's CPU oracle (oracle/mlp.py): the analytic
gradient against central finite differences in float64, and the loss gradient
identity d loss / d f = coef."""
import numpy as np

from oracle import mlp


def _params(seed):
    rng = np.random.default_rng(seed)
    p = np.zeros(mlp.NPARAM)
    p[mlp.W1:mlp.W2] = rng.uniform(-3, 3, mlp.W2)
    p[mlp.W2:mlp.B2] = rng.uniform(-0.2, 0.2, mlp.H * mlp.H)
    p[mlp.B2:mlp.W3] = rng.uniform(-0.1, 0.1, mlp.H)
    p[mlp.W3:mlp.B3] = rng.uniform(-0.3, 0.3, mlp.H)
    p[mlp.B3] = 0.2
    return p


def test_gradient_vs_finite_differences():
    rng = np.random.default_rng(1)
    p = _params(2)
    xy = rng.uniform(-1, 1, (40, 2))
    coef = rng.normal(size=40)
    g = mlp.gradient(p, xy, coef, bf16_ops=False)
    idx = np.concatenate([rng.choice(np.arange(mlp.W1, mlp.W2), 12, replace=False),
                          rng.choice(np.arange(mlp.W2, mlp.B2), 20, replace=False),
                          rng.choice(np.arange(mlp.B2, mlp.NPARAM), 12, replace=False), [mlp.B3]])
    h = 1e-6
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        fd = (coef @ mlp.forward(p + e, xy, False) - coef @ mlp.forward(p - e, xy, False)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * (1 + abs(g[i])), (i, fd, g[i])


def test_loss_coefficients_are_df_gradient():
    """d loss / d f_i: -1/nd on data, softmax weights on the uniform part (sum 1)."""
    rng = np.random.default_rng(3)
    p = _params(4)
    xd, xu = rng.uniform(-1, 1, (30, 2)), rng.uniform(-1, 1, (50, 2))
    loss, coef = mlp.mle_loss(p, xd, xu, bf16_ops=False)
    assert np.isclose(coef[30:].sum(), 1.0) and np.allclose(coef[:30], -1 / 30)
    # the parameter gradient of the loss is the coef-weighted gradient of f
    g = mlp.gradient(p, np.concatenate([xd, xu]), coef, bf16_ops=False)
    for i in (5, mlp.W2 + 77, mlp.W3 + 3, mlp.B3):
        e = np.zeros_like(p)
        e[i] = 1e-6
        fd = (mlp.mle_loss(p + e, xd, xu, False)[0] - mlp.mle_loss(p - e, xd, xu, False)[0]) / 2e-6
        assert abs(fd - g[i]) <= 1e-6 * (1 + abs(g[i]))


def test_bf16_rounding():
    x = np.array([1.0, 1.00390625, 1.005859375, -3.1415926, 65504.0], np.float32)
    r = mlp.bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0  # tie to even
    assert r[2] == 1.0078125            # above the tie rounds up
    assert abs(r[3] - (-3.140625)) < 1e-12
    assert r[4] == 65536.0
