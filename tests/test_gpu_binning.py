"""Bit-exact binning on the device.

The cell centres the grid kernels compute (the midpoint-grid normalisation of
GooPdf::normalize / reduce_with_bins, src/PDFs/GooPdf.cu:83-108,340-396 with
MetricTaker.cu:73-86, and the Dalitz integration cells of
SpecialResonanceIntegrator.cu:59-76 / SpecialDalitzIntegrator.cu:75-82) must equal
the oracle's exactly: no tolerance.  Per-bin values go through the same cells.
"""
import numpy as np
import pytest

import oracle as O
from goofit_amd import workloads as W

from _models import add1d, dalitz, prod3d

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("ranges", [[-5.0, 5.0, 1000], [0.0, 10.0, 200], [0.1, 0.3, 3]])
def test_1d_cells_exact(ctx, ranges):
    assert np.array_equal(ctx.grid_cells(ranges), O.grid_cells(ranges))


def test_200_cubed_cells_exact(ctx):
    ranges = [-6.0, 6.0, 200, 0.0, 8.0, 200, 0.0, 10.0, 200]
    got = ctx.grid_cells(ranges)
    want = O.grid_cells(ranges)
    assert got.shape == (3, 200 ** 3)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("n12,n13", [(1414, 1414), (300, 200), (50, 80)])
@pytest.mark.parametrize("td", [False, True])
def test_dalitz_cells_exact(ctx, n12, n13, td):
    grid6 = [0.0, 3.0, n12, 0.0, 3.0, n13]
    g12, g13 = ctx.dalitz_grid_cells(grid6, time_dependent=td)
    o12, o13 = O.dalitz_grid_cells(grid6, time_dependent=td)
    assert np.array_equal(g12, o12)
    assert np.array_equal(g13, o13)


def _install(ctx, top, decay=None):
    om = O.Model(top, decay=decay)
    f, p, c, o, n = om.op.prog.arrays()
    ctx.set_program(f, p, c, o, n)
    if decay is not None:
        ctx.set_decay(decay)
    return om


def test_eval_bins_1d_matches_oracle_and_reduce(ctx):
    """Per-bin values of the Gaussian component of the AddPdf model on its 1000-bin grid: within
    rounding of the oracle's (exp of the same argument), and their sum is the grid reduction."""
    top, _ = add1d()
    om = _install(ctx, top)
    g = top.components[0]
    ranges = [-5.0, 5.0, 1000]
    got = ctx.eval_bins(g.function_idx, ranges)
    want = om.op.eval_bins(g.function_idx, ranges)
    np.testing.assert_allclose(got, want, rtol=2e-16 * 8, atol=0)
    assert float(np.sum(got)) == pytest.approx(ctx.reduce_bins(g.function_idx, ranges), rel=1e-14)


def test_eval_bins_3d_product(ctx):
    """The 3-D ProdPdf on a 40 x 30 x 20 grid: every cell is the product of the three component
    values at that cell's centres (first observable fastest)."""
    top, _ = prod3d()
    om = _install(ctx, top)
    ranges = [-6.0, 6.0, 40, 0.0, 8.0, 30, 0.0, 10.0, 20]
    got = ctx.eval_bins(top.function_idx, ranges)
    want = om.op.eval_bins(top.function_idx, ranges)
    np.testing.assert_allclose(got, want, rtol=1e-14, atol=0)


def test_eval_bins_rejects_wrong_size(ctx):
    from goofit_amd import capi
    top, _ = add1d()
    _install(ctx, top)
    import ctypes
    out = np.empty(999)
    r = np.array([-5.0, 5.0, 1000.0])
    st = capi.lib().gf_eval_bins(ctx.h, 1, r.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), 1,
                                 out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), out.size)
    assert st == 1  # GF_ERR_INVALID: the output does not hold one value per bin


def test_dalitz_gram_uses_the_reported_cells(ctx):
    """The grid Gram of a flat wave (NonRes, coupling 1) counts the cells inside the Dalitz region:
    its single entry equals the number of reported cells for which inDalitz holds."""
    table = [("nonr", "NONRES", 0.0, 0.0, 0, 0)]
    top, _ = dalitz(nbins=300, resonances=table)
    _install(ctx, top, decay=W.DALITZ_DECAY)
    grid6 = [0.0, 3.0, 300, 0.0, 3.0, 300]
    m12, m13 = ctx.dalitz_grid_cells(grid6)
    inside = sum(O.in_dalitz(a, b, *W.DALITZ_DECAY[:4]) for a, b in zip(m12, m13))
    om = O.Model(top, decay=W.DALITZ_DECAY)
    _, eff = om.op.dalitz_layout(top.function_idx)
    g = ctx.dalitz_gram(top.function_idx, grid6, eff, 1)
    assert g[0, 0].real == pytest.approx(float(inside), rel=1e-14)
    assert g[0, 0].imag == 0.0
