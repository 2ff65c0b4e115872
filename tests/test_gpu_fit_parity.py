"""Fit-level parity: MIGRAD on the GPU likelihood against an independent minimisation of
the oracle's likelihood on the same events.

FitManagerMinuit2::fit (src/goofit/FitManagerMinuit2.cpp:31-119) with FCN::operator()
(src/goofit/FCN.cpp:18-28) runs on the GPU NLL; the independent minimum comes from scipy's
BFGS on the oracle NLL (the C restatement of the reference's OpenMP build) in coordinates
scaled by the fitted errors, polished by Newton steps with a central finite-difference
Hessian.  Every fitted parameter must lie within 0.01 of its error of that minimum.
"""
import numpy as np
import pytest

import goofit_amd as G
import oracle as O
from goofit_amd import workloads as W

from _models import dalitz, prod3d

pytestmark = pytest.mark.gpu


def _oracle_vars(model: O.Model):
    out = {}
    for node in model.op.prog.nodes:
        for p in node.params:
            out[p.name] = p
    return out


def _fd_grad_hess(f, u, h):
    """Central finite-difference gradient and Hessian of f at u with step h."""
    n = u.size
    f0 = f(u)
    g = np.empty(n)
    H = np.empty((n, n))
    fp, fm = np.empty(n), np.empty(n)
    for i in range(n):
        e = np.zeros(n)
        e[i] = h
        fp[i], fm[i] = f(u + e), f(u - e)
        g[i] = (fp[i] - fm[i]) / (2 * h)
        H[i, i] = (fp[i] - 2 * f0 + fm[i]) / (h * h)
    for i in range(n):
        for j in range(i + 1, n):
            ei, ej = np.zeros(n), np.zeros(n)
            ei[i], ej[j] = h, h
            v = (f(u + ei + ej) - f(u + ei - ej) - f(u - ei + ej) + f(u - ei - ej)) / (4 * h * h)
            H[i, j] = H[j, i] = v
    return f0, g, H


def independent_minimum(model: O.Model, names, start, scale, newton_steps=4):
    """Minimum of the oracle NLL over `names`: BFGS in u = (x - start) / scale from u = 0, then
    Newton steps u <- u - H^-1 g until the step is below 1e-6 (in units of scale)."""
    from scipy.optimize import minimize

    vars_ = _oracle_vars(model)
    start = np.asarray(start, float)
    scale = np.asarray(scale, float)

    def f(u):
        for n, v in zip(names, start + scale * u):
            vars_[n].value = float(v)
        return model.calculate_nll()

    r = minimize(f, np.zeros(len(names)), method="BFGS", options=dict(gtol=1e-6, maxiter=2000))
    u = r.x
    for _ in range(newton_steps):
        _, g, H = _fd_grad_hess(f, u, 0.02)
        step = np.linalg.solve(H, g)
        u = u - step
        if np.max(np.abs(step)) < 1e-6:
            break
    return start + scale * u, f(u)


def _gpu_fit(m, tolerance=0.001):
    fm = G.FitManager(m.pdf)
    fm.setVerbosity(0)
    fm.setTolerance(tolerance)
    res = fm.fit()
    assert res.IsValid()
    return res


def test_prod3d_fit_parity():
    """The 3-D ProdPdf fit (Gaussian x exponential x Gaussian-convolved exponential, free mu, sigma,
    alpha and the decay rate; the ExpGaus grid integral redone at every call) on 2e5 events."""
    n = 200_000
    cols = W.sample_prod3d(n, seed=31)
    m = W.build_prod3d(start=1.2)
    d = G.UnbinnedDataSet(m.observables)
    d.from_columns(list(cols))
    m.pdf.setData(d)
    res = _gpu_fit(m)
    names = ["mu", "sigma", "alpha", "rate"]
    gpu = np.array([m.params[k].value for k in names])
    err = np.array([m.params[k].error for k in names])
    assert np.all(err > 0)
    top, _ = prod3d(nbins=200)
    om = O.Model(top, np.stack(cols, 1))
    # independent start: the reference start values (+20 %), not the GPU result
    start = np.array([0.05 * 0.2, 1.2, -1.2 * 1.2, 1.2 / 1.5])
    xs, fs = independent_minimum(om, names, start, err)
    assert np.all(np.abs(gpu - xs) < 0.01 * err), (gpu, xs, err, (gpu - xs) / err)
    assert res.Fval() == pytest.approx(fs, abs=1e-5)


def test_dalitz8_fit_parity():
    """The Dalitz fit of an 8-resonance cut of the benchmark model (14 free coupling parameters, the
    rho+ coupling fixed), 5e4 events generated from the model, 200 x 200 grid normalisation."""
    table = W.DALITZ_RESONANCES[:8]
    nbins = 200
    m = W.build_dalitz(nbins=nbins, resonances=table)
    g12, g13 = W.generate_dalitz(m, 50_000, seed=52)
    m.pdf.setData(W.dalitz_dataset(m, g12, g13))
    m.pdf.setDataSize(len(g12))
    names = sorted(m.params)
    truth = np.array([m.params[k].value for k in names])
    rng = np.random.default_rng(9)
    for k, t in zip(names, truth):
        m.params[k].value = t + 0.05 * rng.standard_normal()
    res = _gpu_fit(m)
    gpu = np.array([m.params[k].value for k in names])
    err = np.array([m.params[k].error for k in names])
    assert np.all(err > 0)
    top, _ = dalitz(nbins=nbins, resonances=table)
    aos = np.stack([g12, g13, np.arange(len(g12), dtype=np.float64)], 1)
    om = O.Model(top, aos, decay=W.DALITZ_DECAY)
    xs, fs = independent_minimum(om, names, truth, err)
    assert np.all(np.abs(gpu - xs) < 0.01 * err), (gpu, xs, err, (gpu - xs) / err)
    assert res.Fval() == pytest.approx(fs, abs=1e-5)
