"""Grid-cell arithmetic of the oracle on the CPU.

The centres of every grid the reference integrates on are
((upper - lower) / n) * (local bin + 0.5) + lower with the first observable fastest
(src/PDFs/MetricTaker.cu:73-86), and the Dalitz integration cells take the m13 bin
count from the m12 entry (src/PDFs/physics/detail/SpecialResonanceIntegrator.cu:59-76,
hazard A3) unless the node is time dependent (detail/SpecialDalitzIntegrator.cu:75-82).
The oracle's cells are checked bit for bit against numpy evaluating the same
expression in the same order (numpy rounds every operation; nothing is fused), and
against the exact midpoints to within rounding.
"""
import math

import numpy as np
import pytest

import oracle as O


def numpy_centres(lo, hi, n, b):
    x = np.float64(hi) - np.float64(lo)
    x = x / np.float64(n)
    x = x * (b.astype(np.float64) + 0.5)
    return x + np.float64(lo)


@pytest.mark.parametrize("lo,hi,n", [(-5.0, 5.0, 1000), (0.0, 10.0, 200), (-6.0, 6.0, 7), (0.1, 0.3, 3)])
def test_1d_cells_bit_exact(lo, hi, n):
    got = O.grid_cells([lo, hi, n])[0]
    want = numpy_centres(lo, hi, n, np.arange(n))
    assert np.array_equal(got, want)
    mid = 0.5 * (np.linspace(lo, hi, n + 1)[:-1] + np.linspace(lo, hi, n + 1)[1:])
    np.testing.assert_allclose(got, mid, rtol=0, atol=4 * np.finfo(float).eps * max(abs(lo), abs(hi)))


def test_3d_cells_order_and_bits():
    """First observable fastest: bin b has local bins (b % n0, (b // n0) % n1, b // (n0 n1))."""
    ranges = [-6.0, 6.0, 20, 0.0, 8.0, 15, 0.0, 10.0, 12]
    got = O.grid_cells(ranges)
    b = np.arange(20 * 15 * 12)
    i0, i1, i2 = b % 20, (b // 20) % 15, b // 300
    assert np.array_equal(got[0], numpy_centres(-6.0, 6.0, 20, i0))
    assert np.array_equal(got[1], numpy_centres(0.0, 8.0, 15, i1))
    assert np.array_equal(got[2], numpy_centres(0.0, 10.0, 12, i2))


def test_200_cubed_cells_bit_exact():
    ranges = [-6.0, 6.0, 200, 0.0, 8.0, 200, 0.0, 10.0, 200]
    got = O.grid_cells(ranges)
    b = np.arange(200 ** 3)
    for d, (lo, hi) in enumerate([(-6.0, 6.0), (0.0, 8.0), (0.0, 10.0)]):
        assert np.array_equal(got[d], numpy_centres(lo, hi, 200, (b // 200 ** d) % 200))


@pytest.mark.parametrize("n12,n13", [(240, 240), (300, 200), (50, 80)])
def test_dalitz_cells_a3(n12, n13):
    """Time-integrated cells: m13 bin p // n12 with the m12 bin count in the width (A3);
    time-dependent cells: the node's own m13 count."""
    grid6 = [0.0, 3.0, n12, 0.1, 2.9, n13]
    p = np.arange(n12 * n13)
    m12, m13 = O.dalitz_grid_cells(grid6)
    assert np.array_equal(m12, numpy_centres(0.0, 3.0, n12, p % n12))
    assert np.array_equal(m13, numpy_centres(0.1, 2.9, n12, p // n12))
    t12, t13 = O.dalitz_grid_cells(grid6, time_dependent=True)
    assert np.array_equal(t12, m12)
    assert np.array_equal(t13, numpy_centres(0.1, 2.9, n13, p // n12))
    if n12 == n13:
        assert np.array_equal(t13, m13)


def test_dalitz_cells_1414_squared():
    n = 1414
    m12, m13 = O.dalitz_grid_cells([0.0, 3.0, n, 0.0, 3.0, n])
    assert m12.size == n * n
    p = np.arange(n * n)
    assert np.array_equal(m12, numpy_centres(0.0, 3.0, n, p % n))
    assert np.array_equal(m13, numpy_centres(0.0, 3.0, n, p // n))
    assert math.isclose(m12[0], 1.5 / n, rel_tol=1e-15) and math.isclose(m13[-1], 3.0 - 1.5 / n, rel_tol=1e-15)
