"""The boundary-fitted product Gauss-Legendre rule over the Dalitz plot
(goofit_amd.workloads.square_dalitz_rule) that judges the normalisation methods: against
adaptive one-dimensional integrals of closed-form inner integrals (scipy.integrate.quad), so
that a wrong Jacobian, map or weight fails; and its spectral convergence."""
import math

import numpy as np
from scipy import integrate

from goofit_amd import workloads as W


def _m13_range(m12, decay=W.DALITZ_DECAY):
    M, d1, d2, d3 = decay[:4]
    s = math.sqrt(m12)
    e1 = 0.5 * (m12 - d2 * d2 + d1 * d1) / s
    e3 = 0.5 * (M * M - m12 - d3 * d3) / s
    r1 = math.sqrt(max(e1 * e1 - d1 * d1, 0.0))
    r3 = math.sqrt(max(e3 * e3 - d3 * d3, 0.0))
    return (e1 + e3) ** 2 - (r1 + r3) ** 2, (e1 + e3) ** 2 - (r1 - r3) ** 2


def _quad_over_plot(inner):
    """Integral over m12 of inner(m12, lo, hi), adaptively, to 1e-13 relative."""
    M, d1, d2, d3 = W.DALITZ_DECAY[:4]
    lo, hi = (d1 + d2) ** 2, (M - d3) ** 2
    val, err = integrate.quad(lambda x: inner(x, *_m13_range(x)), lo, hi, epsabs=0, epsrel=1e-13, limit=400)
    return val


def test_points_inside_the_plot():
    a, b, w = W.square_dalitz_rule(64)
    assert a.size == 64 * 64 and np.all(w > 0)
    assert np.all(W.in_dalitz_np(a, b))


def test_area_and_moments_against_adaptive_quadrature():
    a, b, w = W.square_dalitz_rule(200)
    area = _quad_over_plot(lambda x, lo, hi: hi - lo)
    assert abs(w.sum() / area - 1) < 1e-12
    for p, q in ((1, 0), (0, 2), (2, 3), (3, 1)):
        # int int m12^p m13^q = int m12^p (hi^(q+1) - lo^(q+1)) / (q + 1) dm12
        ref = _quad_over_plot(lambda x, lo, hi: x ** p * (hi ** (q + 1) - lo ** (q + 1)) / (q + 1))
        got = float(np.sum(w * a ** p * b ** q))
        assert abs(got / ref - 1) < 1e-12, (p, q, got, ref)


def test_spectral_convergence_on_a_peaked_function():
    """A narrow Breit-Wigner band in m23 (width 0.05 GeV around 0.98 GeV): 1000 and 2000 nodes
    per axis agree to 1e-10, and 2000 nodes agree with adaptive quadrature to 1e-12."""
    M, d1, d2, d3 = W.DALITZ_DECAY[:4]
    tot = M * M + d1 * d1 + d2 * d2 + d3 * d3
    m0, g0 = 0.98, 0.05

    def f(x, y):
        s23 = tot - x - y
        return 1.0 / ((s23 - m0 * m0) ** 2 + (m0 * g0) ** 2)

    vals = []
    for nodes in (500, 1000, 2000):
        a, b, w = W.square_dalitz_rule(nodes)
        vals.append(float(np.sum(w * f(a, b))))
    assert abs(vals[2] / vals[1] - 1) < 1e-10, vals

    # inner integral over m13 in closed form: d(s23) = -d(m13)
    def inner(x, lo, hi):
        k = m0 * g0
        u_hi, u_lo = tot - x - lo - m0 * m0, tot - x - hi - m0 * m0
        return (math.atan(u_hi / k) - math.atan(u_lo / k)) / k

    ref = _quad_over_plot(inner)
    assert abs(vals[2] / ref - 1) < 1e-12, (vals[2], ref)
