"""PdfBase::fillMCDataSimple (src/goofit/PdfBase.cpp:154-203) on the GPU, and the
call-limit path of FitManagerMinuit2::fit (src/goofit/FitManagerMinuit2.cpp:90-108).

fillMCDataSimple: the events a seed gives are predicted independently (the Gaussian
density in numpy at the bin centres, the std::mt19937 draws restated in Python and
pinned against libstdc++ on the CPU), and a large sample is compared with the
tabulated density by chi^2, cell by cell and inside the cells."""
import math

import numpy as np
import pytest
from scipy import stats

import goofit_amd as G
from goofit_amd import workloads as W

from _stdrandom import MT19937, uniform

pytestmark = pytest.mark.gpu

LO, HI, NB, MU, SIG = -5.0, 5.0, 100, 0.4, 1.3


def _gauss_model():
    x = G.Observable("x", LO, HI)
    x.numbins = NB
    g = G.GaussianPdf("g", x, G.Variable("m", MU), G.Variable("s", SIG))
    return x, g


def _centres():
    # ((hi - lo) / n) * (b + 0.5) + lo, one rounding per operation
    w = (HI - LO) / NB
    return np.array([w * (b + 0.5) + LO for b in range(NB)])


def _expected(nev, seed):
    dens = np.exp(-0.5 * (_centres() - MU) ** 2 / (SIG * SIG))
    cdf = np.cumsum(dens)
    cdf /= cdf[-1]
    g = MT19937(seed)
    w = (HI - LO) / NB
    cells = np.empty(nev, dtype=np.int64)
    vals = np.empty(nev)
    c = _centres()
    for i in range(nev):
        u = uniform(g, 0.0, 1.0)
        k = int(np.searchsorted(cdf, u, side="left"))
        cells[i] = k
        vals[i] = c[k] + w * uniform(g, -0.5, 0.5)
    return cells, vals


def test_fill_mc_simple_reproduces_the_seeded_events():
    x, g = _gauss_model()
    data = G.UnbinnedDataSet(x)
    g.setData(data)
    g.fillMCDataSimple(4000, seed=20231)
    got = np.asarray(data.column(x))
    assert got.size == 4000
    cells, want = _expected(4000, 20231)
    w = (HI - LO) / NB
    got_cells = np.floor((got - LO) / w).astype(np.int64)
    assert np.array_equal(got_cells, cells)
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-14)
    # the events are appended to the data set already there
    g.fillMCDataSimple(1000, seed=7)
    assert len(data) == 5000
    np.testing.assert_array_equal(np.asarray(data.column(x))[:4000], got)


def test_fill_mc_simple_distribution_chi2():
    x, g = _gauss_model()
    data = G.UnbinnedDataSet(x)
    g.setData(data)
    n = 200_000
    g.fillMCDataSimple(n, seed=99)
    v = np.asarray(data.column(x))
    w = (HI - LO) / NB
    cell = np.floor((v - LO) / w).astype(np.int64)
    assert cell.min() >= 0 and cell.max() < NB
    dens = np.exp(-0.5 * (_centres() - MU) ** 2 / (SIG * SIG))
    p = dens / dens.sum()
    obs = np.bincount(cell, minlength=NB)
    keep = n * p > 5
    exp = n * p[keep]
    chi2 = float(np.sum((obs[keep] - exp) ** 2 / exp))
    assert stats.chi2.sf(chi2, keep.sum() - 1) > 1e-4, chi2
    # uniform inside the cells
    frac = (v - LO) / w - cell
    h = np.histogram(frac, bins=20, range=(0, 1))[0]
    chi2u = float(np.sum((h - n / 20) ** 2 / (n / 20)))
    assert stats.chi2.sf(chi2u, 19) > 1e-4, chi2u


def test_fill_mc_simple_refusals():
    x, g = _gauss_model()
    with pytest.raises(G.GeneralError):
        g.fillMCDataSimple(10, seed=1)  # no data set
    y = G.Observable("y", 0.0, 1.0)
    p = G.ProdPdf("p", [g, G.ExpPdf("e", y, G.Variable("a", -1.0))])
    p.setData(G.UnbinnedDataSet([x, y]))
    with pytest.raises(G.GeneralError):
        p.fillMCDataSimple(10, seed=1)  # two observables


def test_fit_call_limit_makes_the_fit_invalid():
    """setMaxCalls -> MIGRAD stops at the limit -> HasReachedCallLimit -> FitErrors::InValid;
    the same fit without the limit is valid."""
    mdl = W.build_add1d(start=1.2)
    d = G.UnbinnedDataSet(mdl.observables)
    d.from_columns([W.sample_add1d(20_000, seed=5)])
    mdl.pdf.setData(d)
    fm = G.FitManager(mdl.pdf)
    fm.setVerbosity(0)
    assert int(fm) == 50 and not fm  # NotRun
    fm.setMaxCalls(12)
    m = fm.fit()
    assert m.HasReachedCallLimit() and not m.IsValid()
    assert m.NFcn() >= 12
    assert not fm and int(fm) == 51  # InValid
    for k, v in mdl.params.items():
        v.value = {"mu": 0.36, "sigma": 0.96, "alpha": -0.48, "frac": 0.36}.get(k, v.value)
    fm2 = G.FitManager(mdl.pdf)
    fm2.setVerbosity(0)
    m2 = fm2.fit()
    assert m2.IsValid() and not m2.HasReachedCallLimit()
    assert fm2 and int(fm2) == 0
