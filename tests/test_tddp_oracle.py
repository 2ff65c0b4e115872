"""Pins of the oracle's time-dependent Dalitz restatement on the CPU.

Amp3Body_TD (src/PDFs/physics/Amp3Body_TD.cu:41-257, 552-748), TruthResolution
(src/PDFs/physics/TruthResolution.cu:6-71) and ThreeGaussResolution
(src/PDFs/physics/ThreeGaussResolution.cu:11-182) are fixed here by what the
mathematics gives independently of the oracle's formulas: decay-time integrals
in closed form, the narrow-resolution limit, the x = y = 0 reduction to the
time-integrated Dalitz density (pinned by its own tests), the charm-tag and
mistag identities, and the normalisation as the integral of the event density
over the Dalitz grid and the decay time.
"""
import math
import os
import sys

import numpy as np
import pytest
from scipy import integrate

import oracle as O
from goofit_amd import workloads as W

sys.path.insert(0, os.path.dirname(__file__))
import _models as M  # noqa: E402

TRUTH, TG = O.FN["TRUTH_RES"], O.FN["THREEGAUSS_RES"]
FIX = (0.4101, 0.0039, 0.0065)  # taufix, xfix, yfix of the mistag terms (Amp3Body_TD.cu:229-231)


def _tg_params(cf=1.0, tf=0.0, cb=0.0, cs=1.0, tb=0.0, ts=1.0, ob=0.0, os_=1.0, sb=0.0, cbo=0.0, cws=1.0,
               cbo2=0.0, cws2=1.0):
    return np.array([cf, tf, cb, cs, tb, ts, ob, os_, sb, cbo, cws, cbo2, cws2], np.float64)


# ------------------------------------------------------------ resolutions

@pytest.mark.parametrize("x,y", [(0.03, 0.05), (0.2, -0.3)])
def test_truth_time_integrals_closed_form(x, y):
    """int_0^inf e^{-t/tau} {cosh(y t/tau), cos(x t/tau), sinh, sin} dt in closed form."""
    tau = 0.41
    closed = {(1, 0, 0, 0): tau / (1 - y * y), (0, 1, 0, 0): tau / (1 + x * x),
              (0, 0, -0.5, 0): y * tau / (1 - y * y), (0, 0, 0, -0.5): x * tau / (1 + x * x)}
    for terms, expect in closed.items():
        # upper limit 120 tau: the tail beyond is below 1e-40 of the integral
        val = integrate.quad(lambda t: O.resolution_eval(TRUTH, None, *terms, tau, t, x, y, 0.0), 0, 120 * tau,
                             epsabs=0, epsrel=1e-13, limit=200)[0]
        assert val == pytest.approx(expect, rel=1e-10), terms


@pytest.mark.parametrize("kind", [TRUTH, TG])
def test_resolution_normalisation_is_time_integral(kind):
    """normalization(di1..di4) = integral over t of the resolution function fed
    term1 = di1 + di2, term2 = di1 - di2, Re/Im(AB*) = di3, di4 (Amp3Body_TD.cu:165-214)."""
    rng = np.random.default_rng(11)
    tau, x, y, sb = 0.41, 0.004, 0.006, 0.0
    p = M.threegauss().params
    par = np.array([q.value for q in p]) if kind == TG else None
    for _ in range(4):
        di = rng.uniform(-1, 1, size=4)
        di[0], di[1] = abs(di[0]) + 1.0, abs(di[1]) + 0.5
        t1, t2 = di[0] + di[1], di[0] - di[1]
        lo = 0.0 if kind == TRUTH else -3.0
        f = lambda t: O.resolution_eval(kind, par, t1, t2, di[2], di[3], tau, t, x, y, 0.05)
        val = integrate.quad(f, lo, 40 * tau, epsabs=0, epsrel=1e-13, limit=400, points=[0.0])[0]
        # ThreeGauss approximates cos/sin to second/third order in x t/tau: O(x^4) away
        assert val == pytest.approx(O.resolution_norm(kind, di, tau, x, y, sb), rel=1e-9)


@pytest.mark.parametrize("sb", [0.0, 0.3])
def test_threegauss_narrow_limit(sb):
    """sigma -> 0: P1..P4 approach e^{-(1/tau+s)t} times cosh(y t/tau), the second-order
    cos expansion, sinh(y t/tau) and the third-order sin expansion (ThreeGaussResolution.cu:11-61)."""
    tau, x, y, t, sig = 0.41, 0.3, 0.2, 0.37, 1e-4
    par = _tg_params(sb=sb)
    G = 1 / tau
    d = math.exp(-(G + sb) * t)
    u = x * G * t
    want = {(1, 0, 0, 0): d * math.cosh(y * G * t), (0, 1, 0, 0): d * (1 - u * u / 2),
            (0, 0, -0.5, 0): d * math.sinh(y * G * t), (0, 0, 0, -0.5): d * (u - u ** 3 / 6)}
    for terms, expect in want.items():
        got = O.resolution_eval(TG, par, *terms, tau, t, x, y, sig)
        assert got == pytest.approx(expect, rel=2e-6), terms


@pytest.mark.parametrize("sb", [0.0, 0.4])
def test_threegauss_time_integrals(sb):
    """Integrals over t of the three-Gaussian mixture (biases, scales, common offsets):
    cosh/sinh parts exactly those of e^{-(Gamma+s)t}; cos/sin parts those of their expansions."""
    tau, x, y = 0.41, 0.2, 0.15
    par = np.array([q.value for q in M.threegauss(sb=sb).params])
    G, a = 1 / tau, 1 / tau + sb
    want = {(1, 0, 0, 0): a / (a * a - (y * G) ** 2),
            (0, 1, 0, 0): 1 / a - (x * G) ** 2 / a ** 3,
            (0, 0, -0.5, 0): y * G / (a * a - (y * G) ** 2),
            (0, 0, 0, -0.5): x * G / a ** 2 - (x * G) ** 3 / a ** 4}
    for terms, expect in want.items():
        f = lambda t: O.resolution_eval(TG, par, *terms, tau, t, x, y, 0.05)
        val = integrate.quad(f, -3.0, 60 * tau, epsabs=0, epsrel=1e-13, limit=400, points=[0.0])[0]
        assert val == pytest.approx(expect, rel=1e-9), terms
    if sb == 0.0:
        # at s = 0 the cosh/sinh normalisation coincides with the truth resolution's
        assert O.resolution_norm(TG, [1, 0, 0, 0], tau, x, y, 0.0) == pytest.approx(
            O.resolution_norm(TRUTH, [1, 0, 0, 0], tau, x, y), rel=1e-14)


# ------------------------------------------------------------ event density

def _raw(top, aos):
    m = O.Model(top, aos=aos, decay=W.DALITZ_DECAY)
    m.normalize()
    m.op.sync_params()
    m.op.sync_norms()
    out = m.op.eval_metric(aos, "EVAL")
    assert m.op.error() == 0
    return out, m


def test_b_wave_is_swapped_lineshape():
    """B_i(m12, m13) = BW_i(m13, m12) (SpecialWaveCalculator.cu:55-58), checked with the
    Dalitz lineshape routine at the swapped point."""
    top, _ = M.tddp()
    aos = M.tddp_events(400)
    op = O.OracleProg(top, W.DALITZ_DECAY)
    res, _, _, _ = op.tddp_layout(0)
    for r in res[:4]:
        a, b = op.tddp_wave(0, r, aos)
        for k in range(0, 400, 37):
            m12, m13 = aos[k, 2], aos[k, 3]
            inside = O.in_dalitz(m12, m13, *W.DALITZ_DECAY[:4])
            wa = op.lineshape(r, m12, m13) if inside else 0
            wb = op.lineshape(r, m13, m12) if inside else 0
            assert a[k] == pytest.approx(wa, rel=1e-13, abs=1e-300)
            assert b[k] == pytest.approx(wb, rel=1e-12, abs=1e-300)


def test_no_mixing_reduces_to_dalitz_density():
    """x = y = 0: the density is 2 |A|^2 e^{-t/tau} eff, with |A|^2 eff the Amp3Body density."""
    table = W.DALITZ_RESONANCES[:6]
    eff_td = None
    top, (dt, sg, m12, m13, num, _, _) = M.tddp(resonances=table, mix=(0.41, 0.0, 0.0, 0.0, 0.0), eff=eff_td)
    aos = M.tddp_events(3000, seed=21)
    f_td, _ = _raw(top, aos)
    dz, _ = M.dalitz(resonances=table)
    f_dz, _ = _raw(dz, np.ascontiguousarray(aos[:, 2:5]))
    want = 2 * f_dz * np.exp(-aos[:, 0] / 0.41)
    np.testing.assert_allclose(f_td, want, rtol=1e-13, atol=0)
    assert np.count_nonzero(f_td == 0) > 0  # events outside the Dalitz region give 0


def test_charm_tag_selects_mixing_parameters():
    """tag +1 / -1 use (x + dx, y + dy) / (x - dx, y - dy); tag 0 uses (x, y) (Amp3Body_TD.cu:132-145)."""
    mix = (0.41, 0.02, 0.03, 0.005, -0.004)
    aos = M.tddp_events(800, seed=3)
    n = len(aos)
    for tag, (xs, ys) in ((1, (1, 1)), (-1, (-1, -1)), (0, (0, 0))):
        top, _ = M.tddp(mix=mix, charmtag=True, mistag=True)
        rows = np.concatenate([aos, np.zeros((n, 1)), np.full((n, 1), float(tag))], axis=1)
        f_tag, _ = _raw(top, rows)
        plain_mix = (mix[0], mix[1] + xs * mix[3], mix[2] + ys * mix[4], 0.0, 0.0)
        top2, _ = M.tddp(mix=plain_mix)
        f_plain, _ = _raw(top2, aos)
        np.testing.assert_allclose(f_tag, f_plain, rtol=1e-14, atol=0)


def test_mistag_mixture_identity():
    """mistag w: (1 - 2w) f(m12, m13) + w f_fix(m13, m12) + w f_fix(m12, m13), where f_fix has
    (tau, x, y) = (0.4101, 0.0039, 0.0065) and swapping m12, m13 exchanges A and B
    (equal daughter masses 2 and 3, constant efficiency)."""
    mix = (0.41, 0.02, 0.03, 0.0, 0.0)
    aos = M.tddp_events(1000, seed=9, outside=0.0)
    w = 0.13
    top, _ = M.tddp(mix=mix, mistag=True)
    f_w, _ = _raw(top, np.concatenate([aos, np.full((len(aos), 1), w)], axis=1))
    f0, _ = _raw(M.tddp(mix=mix)[0], aos)
    ffix, _ = _raw(M.tddp(mix=FIX + (0.0, 0.0))[0], aos)
    sw = aos.copy()
    sw[:, 2], sw[:, 3] = aos[:, 3], aos[:, 2]
    ffix_sw, _ = _raw(M.tddp(mix=FIX + (0.0, 0.0))[0], sw)
    want = (1 - 2 * w) * f0 + w * ffix_sw + w * ffix
    np.testing.assert_allclose(f_w, want, rtol=1e-11, atol=1e-14 * np.max(np.abs(want)))


def _gl_nodes(lo, hi, pieces, order, edges=None):
    x, w = np.polynomial.legendre.leggauss(order)
    edges = np.linspace(lo, hi, pieces + 1) if edges is None else edges
    ts, ws = [], []
    for a, b in zip(edges[:-1], edges[1:]):
        ts.append(0.5 * (b - a) * x + 0.5 * (b + a))
        ws.append(0.5 * (b - a) * w)
    return np.concatenate(ts), np.concatenate(ws)


@pytest.mark.parametrize("kind", ["truth", "threegauss"])
def test_normalisation_is_integral_of_density(kind):
    """Amp3Body_TD::normalize equals the integral of the event density over the m12 x m13
    midpoint grid (times the cell area) and the decay time, D0 and D0bar mixed 50:50 by
    the charm tag, with a mistag fraction and a non-constant efficiency."""
    nb = 24
    # ThreeGauss expands cos/sin to second/third order, so its normalisation is O(x^4) off the
    # integral: x + dx = 0.006 keeps that below 1e-9
    mix = (0.41, 0.012, -0.02, 0.003, 0.004) if kind == "truth" else (0.41, 0.004, -0.02, 0.002, 0.004)
    w = 0.08
    res = M.threegauss() if kind == "threegauss" else O.TruthResolution()
    tmp_eff = None
    top, obs = M.tddp(nbins=nb, mix=mix, resolution=res, mistag=True, charmtag=True, eff=tmp_eff)
    m12o, m13o, mt = obs[2], obs[3], obs[5]
    # linear efficiency in m12, m13 (exercised by both the fake grid event and the real events)
    eff = O.MultiPolynomial("eff", [m12o, m13o], [O.Var("c0", 1.0), O.Var("c1", 0.1), O.Var("c2", -0.05)],
                            [O.Var("z0", 0.0), O.Var("z1", 0.0)], 1)
    top.components[-1] = eff
    mt.value = w
    model = O.Model(top, aos=None, decay=W.DALITZ_DECAY)
    norm = model.normalize()
    c12 = M.np.array([(3.0 / nb) * (i + 0.5) for i in range(nb)])
    g12, g13 = np.meshgrid(c12, c12, indexing="ij")
    g12, g13 = g12.ravel(), g13.ravel()
    if kind == "truth":
        t, wt = _gl_nodes(0.0, 40 * mix[0], 10, 24)
    else:
        # the widest Gaussian (outlier scale 6 x 0.05) needs the range to start 13 sigma below 0
        # (pieces of 0.05 around t = 0, where the core Gaussian of width ~0.06 sits)
        edges = np.concatenate([np.linspace(-4.0, -1.0, 4)[:-1], np.linspace(-1.0, 1.0, 41),
                                np.linspace(1.0, 40 * mix[0], 16)[1:]])
        t, wt = _gl_nodes(0, 0, 0, 16, edges)
    ncell, nt = g12.size, t.size
    total = 0.0
    for tag in (1.0, -1.0):
        rows = np.column_stack([np.tile(t, ncell), np.full(ncell * nt, 0.05), np.repeat(g12, nt),
                                np.repeat(g13, nt), np.arange(ncell * nt, dtype=np.float64),
                                np.full(ncell * nt, w), np.full(ncell * nt, tag)])
        f, _ = _raw(top, rows)
        total += 0.5 * float(np.sum(f * np.tile(wt, ncell)))
    total *= (3.0 / nb) ** 2
    assert norm == pytest.approx(total, rel=1e-8)


def test_mistag_norm_terms_follow_observable_range():
    """The normalisation's mistag terms are applied when the mistag observable tests
    true (its value inside its limits, include/goofit/Variable.h:103-105)."""
    top, obs = M.tddp(mistag=True)
    mt = obs[5]
    mt.value = 0.1
    n_in = O.Model(top, decay=W.DALITZ_DECAY).normalize()
    mt.value = 0.7  # outside [0, 0.5]: no mistag terms
    top2, obs2 = M.tddp(mistag=True)
    obs2[5].value = 0.7
    n_out = O.Model(top2, decay=W.DALITZ_DECAY).normalize()
    top3, _ = M.tddp()
    n_none = O.Model(top3, decay=W.DALITZ_DECAY).normalize()
    assert n_out == pytest.approx(n_none, rel=1e-15)
    assert abs(n_in / n_none - 1) > 1e-4
