"""The Minuit2-compatible minimiser (goofit_amd/host/Minuit2.cpp) on the CPU,
through the Minuit2 Python module: closed-form minima and covariances, limits,
MINOS, scans, and a maximum-likelihood fit whose estimates and errors are known
analytically.  (The reference's Minuit2 submodule is absent; these pin the
minimiser by the mathematics.)"""
import math

import numpy as np
import pytest

from goofit_amd import _goofit

M2 = _goofit.Minuit2


class Quadratic(M2.FCNBase):
    def __init__(self, A, a):
        super().__init__()
        self.A, self.a = np.asarray(A, float), np.asarray(a, float)

    def __call__(self, x):
        d = np.asarray(x) - self.a
        return 0.5 * float(d @ self.A @ d)

    def Up(self):
        return 0.5


def test_quadratic_minimum_and_covariance():
    """f = (x-a)^T A (x-a) / 2 with Up = 0.5: minimum a, covariance 2 Up A^-1 = A^-1."""
    A = np.array([[4.0, 1.0, 0.3], [1.0, 2.0, -0.4], [0.3, -0.4, 1.5]])
    a = np.array([1.0, -2.0, 0.5])
    f = Quadratic(A, a)
    up = M2.MnUserParameters()
    for i, s in enumerate((0.0, 0.0, 0.0)):
        up.Add(f"p{i}", s, 0.1)
    m = M2.MnMigrad(f, up, 1)(0, 0.1)
    assert m.IsValid() and m.HasValidCovariance()
    # EDM < 0.002 * tolerance * Up leaves the minimum within ~0.01 sigma
    np.testing.assert_allclose(m.UserParameters().Params(), a, atol=5e-3 * np.sqrt(np.diag(np.linalg.inv(A))).min())
    cov = m.UserCovariance().to_matrix()
    np.testing.assert_allclose(cov, np.linalg.inv(A), rtol=2e-3, atol=1e-5)
    np.testing.assert_allclose(m.UserParameters().Errors(), np.sqrt(np.diag(np.linalg.inv(A))), rtol=2e-3)
    # MINOS on a quadratic: symmetric errors equal to the parabolic ones
    lo, hi = M2.MnMinos(f, m, 1)(1)
    e = math.sqrt(np.linalg.inv(A)[1, 1])
    assert lo == pytest.approx(-e, rel=2e-3) and hi == pytest.approx(e, rel=2e-3)


class Rosenbrock(M2.FCNBase):
    def __call__(self, x):
        return 100.0 * (x[1] - x[0] ** 2) ** 2 + (1 - x[0]) ** 2

    def Up(self):
        return 0.5


def test_rosenbrock():
    up = M2.MnUserParameters()
    up.Add("x", -1.2, 0.1)
    up.Add("y", 1.0, 0.1)
    m = M2.MnMigrad(Rosenbrock(), up, 1)(0, 0.01)
    assert m.IsValid()
    np.testing.assert_allclose(m.UserParameters().Params(), [1.0, 1.0], atol=2e-3)


class GaussNLL(M2.FCNBase):
    def __init__(self, x):
        super().__init__()
        self.x = np.asarray(x, float)

    def __call__(self, p):
        mu, sig = p
        return float(np.sum(0.5 * ((self.x - mu) / sig) ** 2 + np.log(sig)))

    def Up(self):
        return 0.5


def test_gaussian_mle_estimates_and_errors():
    """ML fit of a Gaussian: mu = mean, sigma = rms; errors sigma/sqrt(N), sigma/sqrt(2N);
    sigma limited to (0.01, 10) (the sin transformation of limited parameters)."""
    x = np.random.default_rng(4).normal(1.3, 0.7, 5000)
    up = M2.MnUserParameters()
    up.Add("mu", 0.0, 0.1)
    up.Add("sig", 1.0, 0.1, 0.01, 10.0)
    m = M2.MnMigrad(GaussNLL(x), up, 1)(0, 0.1)
    assert m.IsValid()
    mu, sig = m.UserParameters().Params()
    emu, esig = m.UserParameters().Errors()
    s = math.sqrt(np.mean((x - x.mean()) ** 2))
    assert mu == pytest.approx(x.mean(), abs=1e-5) and sig == pytest.approx(s, rel=1e-5)
    assert emu == pytest.approx(s / math.sqrt(x.size), rel=1e-2)
    assert esig == pytest.approx(s / math.sqrt(2 * x.size), rel=1e-2)


def test_fixed_parameter_and_scan():
    A = np.diag([2.0, 3.0])
    f = Quadratic(A, [0.5, -0.5])
    up = M2.MnUserParameters()
    up.Add("a", 0.0, 0.1)
    up.Add("b", 1.0)  # constant
    m = M2.MnMigrad(f, up, 1)(0, 0.1)
    assert m.IsValid()
    a, b = m.UserParameters().Params()
    assert a == pytest.approx(0.5, abs=1e-5) and b == 1.0
    assert m.UserCovariance().Nrow() == 1
    # as Minuit2's MnParameterScan: the starting point first, then maxsteps points over [low, high]
    pts = M2.MnScan(f, up, 1).Scan(0, 11, -1.0, 2.0)
    assert len(pts) == 12 and pts[0][0] == 0.0
    assert pts[1][0] == pytest.approx(-1.0) and pts[-1][0] == pytest.approx(2.0)
    vals = [f([p, 1.0]) for p, _ in pts]
    np.testing.assert_allclose([v for _, v in pts], vals, rtol=1e-12)
    M2.MnPrint.SetGlobalLevel(0)
    assert M2.MnPrint.GlobalLevel() == 0
