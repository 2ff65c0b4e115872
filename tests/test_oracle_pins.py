"""Pins of the oracle (oracle/goofit_oracle.c) on the CPU.

Each check fixes a result by something other than the oracle's own formula:
the reference's known-answer tests (tests/simple/NormalizeTest.cpp,
AddNormTest.cpp), closed forms from the physics or from scipy, invariants, and
brute force on small inputs.  A mistake such as a dropped term, a wrong sign or
index, a transposed operand or a wrong summation order fails one of them.
"""
import ctypes
import math

import numpy as np
import pytest
from scipy import special, stats

import oracle as O
from goofit_amd import workloads as W

FLT_EPS = float(np.finfo(np.float32).eps)


def catch_approx(actual, expected, eps=100 * FLT_EPS):
    """Catch2's default Approx: |a - e| <= 100 FLT_EPSILON * |e| (the reference tests' tolerance)."""
    return abs(actual - expected) <= eps * abs(expected)


def grid_centres(lo, hi, n):
    # midpoints of n equal bins, computed independently of the engine's formula
    edges = np.linspace(lo, hi, n + 1)
    return 0.5 * (edges[:-1] + edges[1:])


# ------------------------------------------------------------ reference KATs

GAUSS_KAT = {0: 0.00587129964482, 20: 0.0624318782242, 40: 0.157287605852, 60: 0.0938855055588}


def _gauss_model(mean=1.5, sigma=2.5):
    x = O.Obs("xvar", -5.0, 10.0, 100)
    g = O.Gaussian("gausspdf", x, O.Var("mean", mean), O.Var("sigma", sigma))
    return g, x


def test_gauss_1d_grid_kat():
    """NormalizeTest.cpp:56-100 ("1D Grid", "1D Eval"): normalised values at bin centres 0, 20, 40, 60."""
    g, x = _gauss_model()
    m = O.Model(g)
    norm_int = m.normalize()
    assert catch_approx(norm_int, 6.26657)  # NormalizeTest.cpp:189
    assert norm_int == pytest.approx(2.5 * math.sqrt(2 * math.pi), rel=1e-15)  # integral over R (A1)
    raw = m.op.eval_bins(g.function_idx, [-5.0, 10.0, 100])
    prob = raw / norm_int
    for i, v in GAUSS_KAT.items():
        assert catch_approx(prob[i], v), (i, prob[i], v)
    # every bin: scipy's normal density at the midpoints of 100 equal bins
    ref = stats.norm.pdf(grid_centres(-5.0, 10.0, 100), 1.5, 2.5)
    np.testing.assert_allclose(prob, ref, rtol=1e-14, atol=0)


def test_gauss_eval_kat():
    """NormalizeTest.cpp:191-216: Prob and Eval at x = 1.2 and 7.2.

    The reference's Eval KAT at 1.2 (0.99282585857903813) has two digits swapped
    relative to its own comment (0.9928258579038134); Catch's tolerance absorbs
    it, and the closed form pins the value to 1e-15."""
    g, x = _gauss_model()
    m = O.Model(g)
    m.normalize()
    m.op.sync_norms()
    for xv, p_kat, e_kat in ((1.2, 0.15843208471746245, 0.99282585857903813),
                             (7.2, 0.0118618339389365, 0.0743330208507896)):
        e = m.op.eval_at(g.function_idx, [xv])
        p = m.op.eval_metric(np.array([[xv]]), "PROB")[0]
        assert catch_approx(e, e_kat) and catch_approx(p, p_kat)
        assert e == pytest.approx(math.exp(-0.5 * ((xv - 1.5) / 2.5) ** 2), rel=1e-15)
        assert p == pytest.approx(stats.norm.pdf(xv, 1.5, 2.5), rel=1e-14)


def test_exp_norm_kat():
    """NormalizeTest.cpp:41-43: ExpPdf(alpha = -2) on [0, 10] integrates to 0.5."""
    x = O.Obs("xvar", 0.0, 10.0)
    e = O.Exp("exppdf", x, O.Var("alpha", -2.0))
    m = O.Model(e)
    v = m.normalize()
    assert catch_approx(v, 0.5)
    assert v == pytest.approx((1 - math.exp(-20.0)) / 2.0, rel=1e-15)
    # the midpoint grid over the same range converges to the same integral
    raw = m.op.eval_bins(e.function_idx, [0.0, 10.0, 20000])
    assert raw.sum() * 10.0 / 20000 == pytest.approx(v, rel=1e-7)


def test_add_norm_kat():
    """AddNormTest.cpp:16-44: f*A + (1-f)*B equals the weighted normalised components."""
    x = O.Obs("xvar", -5.0, 10.0, 100)
    mean = O.Var("mean", 1.5)
    ga = O.Gaussian("gaussA", x, mean, O.Var("sigmaA", 2.5))
    gb = O.Gaussian("gaussB", x, mean, O.Var("sigmaB", 0.8))
    add = O.Add("sum", [O.Var("frac", 0.6)], [ga, gb])
    pts = grid_centres(-5.0, 10.0, 100)[:, None]
    m = O.Model(add, aos=pts)
    m.normalize()
    m.op.sync_norms()
    comb = m.op.eval_metric(pts, "PROB")
    pa = stats.norm.pdf(pts[:, 0], 1.5, 2.5)
    pb = stats.norm.pdf(pts[:, 0], 1.5, 0.8)
    for i in range(0, 100, 7):
        assert catch_approx(comb[i], 0.6 * pa[i] + 0.4 * pb[i])
    np.testing.assert_allclose(comb, 0.6 * pa + 0.4 * pb, rtol=1e-14)


def test_2d_grid_order_kat():
    """NormalizeTest.cpp:105-174 ("2D Grid"): global bin b -> (b % 100, b / 100), first observable fastest."""
    x = O.Obs("xvar", -5.0, 10.0, 100)
    y = O.Obs("yvar", -5.0, 10.0, 100)
    g1 = O.Gaussian("g1", x, O.Var("mean1", 1.5), O.Var("sigma1", 2.5))
    g2 = O.Gaussian("g2", y, O.Var("mean2", -1.0), O.Var("sigma2", 0.5))
    prod = O.Prod("product", [g1, g2])
    m = O.Model(prod)
    m.normalize()
    m.op.sync_norms()
    raw = m.op.eval_bins(prod.function_idx, [-5.0, 10.0, 100, -5.0, 10.0, 100])
    c = grid_centres(-5.0, 10.0, 100)
    p2 = stats.norm.pdf(c, -1.0, 0.5)
    for i, v in GAUSS_KAT.items():
        # bin i: x-bin i, y-bin 0; the product carries the component normalisations
        assert catch_approx(raw[i] / p2[0], v)
        # bin 100*i: x-bin 0, y-bin i
        assert raw[100 * i] == pytest.approx(stats.norm.pdf(c[0], 1.5, 2.5) * p2[i], rel=1e-13)


# ------------------------------------------------------------ closed forms

def _metric(metric, raw, evt, norm, total):
    L = O.lib()
    L.gfo_metric.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.POINTER(ctypes.c_double), ctypes.c_double,
                             ctypes.c_double]
    L.gfo_metric.restype = ctypes.c_double
    e = (ctypes.c_double * 2)(*evt)
    return L.gfo_metric(O.METRIC[metric], raw, e, norm, total)


def test_metrics_closed_forms():
    """MetricPointer.cu:8-58 against the distributions they come from."""
    assert _metric("EVAL", 2.5, [0, 0], 0.1, 1) == 2.5
    assert _metric("PROB", 2.5, [0, 0], 0.1, 1) == pytest.approx(0.25, rel=1e-16)
    assert _metric("NLL", 2.5, [0, 0], 0.1, 1) == pytest.approx(-math.log(0.25), rel=1e-16)
    # A8: zero, not +inf, for a non-positive probability
    assert _metric("NLL", 0.0, [0, 0], 0.1, 1) == 0.0
    assert _metric("NLL", -1.0, [0, 0], 0.1, 1) == 0.0
    # BinAvg: Poisson NLL of n observed with expectation N*p*norm*volume, without the ln(n!) term
    n_obs, vol, total = 3.0, 0.5, 100.0
    mu = total * 2.0 * 0.25 * vol
    expect = -(stats.poisson.logpmf(3, mu) + math.lgamma(4.0))
    assert _metric("BINAVG", 2.0, [n_obs, vol], 0.25, total) == pytest.approx(expect, rel=1e-14)
    assert _metric("BINAVG", 0.0, [n_obs, vol], 0.25, total) == 0.0
    # BinWithError: ((f - y) / sigma)^2
    assert _metric("BINWITHERROR", 5.0, [3.0, 0.5], 7.0, total) == pytest.approx(16.0, rel=1e-16)
    # Chisq: Neyman chi^2 with the observed count floored at 1
    assert _metric("CHISQ", 2.0, [3.0, vol], 0.25, total) == pytest.approx((mu - 3.0) ** 2 / 3.0, rel=1e-15)
    assert _metric("CHISQ", 2.0, [0.0, vol], 0.25, total) == pytest.approx(mu ** 2, rel=1e-15)


def test_expgaus_vs_scipy():
    """device_ExpGaus (ExpGausPdf.cu:6-24) is the exponentially modified Gaussian density.

    scipy's exponnorm(K = 1/(sigma*lambda), loc = mu, scale = sigma) is the same
    distribution; the reference's truncated sqrt(2) (1.4142135623, A7) moves the
    erfc argument by 7e-11 relative, well inside rtol 1e-8."""
    z = O.Obs("z", 0.0, 10.0, 200)
    for mu, sig, lam in ((0.0, 0.1, 1 / 1.5), (0.3, 0.7, 2.0), (-0.5, 0.05, 0.3)):
        eg = O.ExpGaus("eg", z, O.Var("m", mu), O.Var("s", sig), O.Var("t", lam))
        op = O.OracleProg(eg)
        xs = np.linspace(-1.0, 9.0, 41)
        got = np.array([op.eval_at(0, [v]) for v in xs])
        ref = stats.exponnorm.pdf(xs, 1.0 / (sig * lam), loc=mu, scale=sig)
        np.testing.assert_allclose(got, ref, rtol=1e-8, atol=1e-300)


def test_two_body_cm_mom_vs_energy_form():
    """twoBodyCMmom (Resonance.cu:48-60, PDG kinematics): |p| = sqrt(E1^2 - m1^2), E1 = (s + m1^2 - m2^2)/(2 sqrt s)."""
    rng = np.random.default_rng(0)
    for _ in range(200):
        m1, m2 = rng.uniform(0.1, 0.6, 2)
        s = (m1 + m2) ** 2 * rng.uniform(1.001, 20.0)
        e1 = (s + m1 * m1 - m2 * m2) / (2 * math.sqrt(s))
        assert O.two_body_cm_mom(s, m1, m2) == pytest.approx(math.sqrt(e1 * e1 - m1 * m1), rel=1e-12)
    # below threshold the reference replaces each negative radicand by 1
    s = 0.5 * (0.3 + 0.4) ** 2
    k2 = 1 - (0.3 - 0.4) ** 2 / s
    assert O.two_body_cm_mom(s, 0.3, 0.4) == pytest.approx(0.5 * math.sqrt(s) * 1.0 * math.sqrt(k2), rel=1e-15)


def _boost(p4, beta):
    """Lorentz boost of four-vectors p4 (..., 4: E, px, py, pz) by velocity beta (3,)."""
    b2 = float(beta @ beta)
    if b2 == 0:
        return p4
    g = 1 / math.sqrt(1 - b2)
    bp = p4[..., 1:] @ beta
    e = g * (p4[..., 0] + bp)
    pv = p4[..., 1:] + ((g - 1) * bp / b2 + g * p4[..., 0])[..., None] * beta
    return np.concatenate([e[..., None], pv], axis=-1)


def _decay_sample(M, m1, m2, m3, n, rng):
    """(s12, s13) of n M -> 1 2 3 decays built from four-vectors: M -> (12) 3, then (12) -> 1 2, random angles."""
    out = []
    for _ in range(n):
        m12 = rng.uniform(m1 + m2, M - m3)
        # (12) and 3 back to back in the M frame
        q = math.sqrt((M * M - (m12 + m3) ** 2) * (M * M - (m12 - m3) ** 2)) / (2 * M)
        u = rng.normal(size=3); u /= np.linalg.norm(u)
        p12 = np.array([math.sqrt(m12 * m12 + q * q), *(q * u)])
        p3 = np.array([math.sqrt(m3 * m3 + q * q), *(-q * u)])
        # 1 and 2 back to back in the (12) frame, then boosted to the M frame
        k = math.sqrt((m12 * m12 - (m1 + m2) ** 2) * (m12 * m12 - (m1 - m2) ** 2)) / (2 * m12)
        v = rng.normal(size=3); v /= np.linalg.norm(v)
        p1 = _boost(np.array([math.sqrt(m1 * m1 + k * k), *(k * v)]), p12[1:] / p12[0])
        mink = lambda a: a[0] ** 2 - a[1:] @ a[1:]
        out.append((mink(p12), mink(p1 + p3), float(v @ u)))
    return np.array(out)


def test_in_dalitz_vs_four_vectors():
    """inDalitz (DalitzPlotHelpers.cu:14-59): every physical decay is inside, and the m13 boundary at
    fixed m12 is where the decay angle reaches 0 or pi."""
    M, m1, m2, m3 = W.DALITZ_DECAY[:4]
    rng = np.random.default_rng(1)
    s = _decay_sample(M, m1, m2, m3, 3000, rng)
    inner = np.abs(s[:, 2]) < 0.9999
    assert all(O.in_dalitz(a, b, M, m1, m2, m3) for a, b, _ in s[inner])
    # the extremes of s13 at fixed s12 over the angle, from the same four-vector construction
    for m12 in np.linspace(m1 + m2 + 1e-3, M - m3 - 1e-3, 25):
        s12 = m12 * m12
        vals = []
        for cos in (-1.0, 1.0):
            q = math.sqrt((M * M - (m12 + m3) ** 2) * (M * M - (m12 - m3) ** 2)) / (2 * M)
            k = math.sqrt((m12 * m12 - (m1 + m2) ** 2) * (m12 * m12 - (m1 - m2) ** 2)) / (2 * m12)
            p12 = np.array([math.sqrt(m12 * m12 + q * q), 0.0, 0.0, q])
            p3 = np.array([math.sqrt(m3 * m3 + q * q), 0.0, 0.0, -q])
            p1 = _boost(np.array([math.sqrt(m1 * m1 + k * k), 0.0, 0.0, k * cos]), p12[1:] / p12[0])
            a = p1 + p3
            vals.append(a[0] ** 2 - a[1:] @ a[1:])
        lo, hi = min(vals), max(vals)
        eps = 1e-7 * hi
        assert O.in_dalitz(s12, 0.5 * (lo + hi), M, m1, m2, m3)
        assert O.in_dalitz(s12, lo + eps, M, m1, m2, m3) and O.in_dalitz(s12, hi - eps, M, m1, m2, m3)
        assert not O.in_dalitz(s12, lo - eps, M, m1, m2, m3)
        assert not O.in_dalitz(s12, hi + eps, M, m1, m2, m3)
    # outside the m12 range altogether
    assert not O.in_dalitz((m1 + m2) ** 2 * 0.99, 1.0, M, m1, m2, m3)


def _lineshape_prog(kind, mass, width, spin, pair, norm=True):
    m12 = O.Obs("m12", 0.0, 3.0, 10)
    m13 = O.Obs("m13", 0.0, 3.0, 10)
    num = O.Obs("evt", 0.0, 1e9, 1, event_number=True)
    ar, ai = O.Var("re", 1.0), O.Var("im", 0.0)
    if kind == "RBW":
        r = O.RBW("r", ar, ai, O.Var("m", mass), O.Var("w", width), spin, pair, norm=norm)
    elif kind == "GS":
        r = O.GS("r", ar, ai, O.Var("m", mass), O.Var("w", width), spin, pair, norm=norm)
    else:
        r = O.NonRes("r", ar, ai)
    eff = O.MultiPolynomial("eff", [m12, m13], [O.Var("one", 1.0)], [O.Var("z0", 0.0), O.Var("z1", 0.0)], 0)
    top = O.Dalitz("dz", m12, m13, num, W.DALITZ_DECAY, [r], eff)
    op = O.OracleProg(top, W.DALITZ_DECAY)
    res, _ = op.dalitz_layout(0)
    return op, res[0]


def _p_cm(s, ma, mb):
    e = (s + ma * ma - mb * mb) / (2 * math.sqrt(s))
    return math.sqrt(e * e - ma * ma)


def test_rbw_spin0_textbook():
    """plainBW (RBW.cu:9-86) for spin 0: 1 / (m0^2 - s - i m0 Gamma(s)), Gamma(s) = Gamma0 (q/q0) (m0/sqrt s)."""
    M, mpi0, mpi, _ = W.DALITZ_DECAY[:4]
    m0, g0 = 1.2, 0.15
    # pair 12 = (daughter 1, daughter 2); bachelor 3
    op, rf = _lineshape_prog("RBW", m0, g0, 0, 0)
    M2 = M * M + mpi0 ** 2 + 2 * mpi ** 2
    for s12 in np.linspace(0.5, 2.5, 9):
        s13 = 0.5 * (M2 - s12)  # any m13: spin 0 has no angular factor
        q, q0 = _p_cm(s12, mpi0, mpi), _p_cm(m0 * m0, mpi0, mpi)
        gam = g0 * (q / q0) * (m0 / math.sqrt(s12))
        want = 1.0 / complex(m0 * m0 - s12, -m0 * gam)
        assert op.lineshape(rf, s12, s13) == pytest.approx(want, rel=1e-13)


# (A, B, C) of each pair index: resonance -> A B, bachelor C (daughters numbered 1, 2, 3)
PAIR_ABC = {0: (1, 2, 3), 1: (3, 1, 2), 2: (2, 3, 1)}


def _pole_config(pair, m0, cos_theta):
    """Invariants (s12, s13) of a decay whose pair has mass m0, built in the pair's rest frame from
    four-vectors: A and B back to back along z, the bachelor at angle theta.  Also returns the 3-vector
    products p_A.p_C and |p_A|^2 |p_C|^2 in that frame."""
    M = W.DALITZ_DECAY[0]
    mass = {1: W.DALITZ_DECAY[1], 2: W.DALITZ_DECAY[2], 3: W.DALITZ_DECAY[3]}
    a, b, c = PAIR_ABC[pair]
    ma, mb, mc = mass[a], mass[b], mass[c]
    k = math.sqrt((m0 * m0 - (ma + mb) ** 2) * (m0 * m0 - (ma - mb) ** 2)) / (2 * m0)
    ec = (M * M - m0 * m0 - mc * mc) / (2 * m0)
    pc = math.sqrt(ec * ec - mc * mc)
    sin_t = math.sqrt(1 - cos_theta ** 2)
    PA = np.array([math.sqrt(ma * ma + k * k), 0, 0, k])
    PB = np.array([math.sqrt(mb * mb + k * k), 0, 0, -k])
    PC = np.array([ec, pc * sin_t, 0, pc * cos_theta])
    mink = lambda v: v[0] ** 2 - v[1:] @ v[1:]
    four = {a: PA, b: PB, c: PC}
    s12 = mink(four[1] + four[2])
    s13 = mink(four[1] + four[3])
    return s12, s13, float(PA[1:] @ PC[1:]), k * k * pc * pc


@pytest.mark.parametrize("pair", [0, 1, 2])
@pytest.mark.parametrize("spin", [1, 2])
def test_rbw_pole_angular_factor(pair, spin):
    """At s_AB = m0^2 both Blatt-Weisskopf ratios are 1 and plainBW = i * Z / (m0 Gamma0), with the Zemach
    factor from four-vectors in the AB rest frame: Z = -4 pA.pC (spin 1),
    Z = 16 [(pA.pC)^2 - |pA|^2 |pC|^2 / 3] (spin 2)."""
    m0, g0 = 0.98, 0.07
    op, rf = _lineshape_prog("RBW", m0, g0, spin, pair)
    for cos in (-0.8, -0.3, 0.1, 0.55, 0.9):
        s12, s13, dot, mag2 = _pole_config(pair, m0, cos)
        z = -4 * dot if spin == 1 else 16 * (dot * dot - mag2 / 3)
        got = op.lineshape(rf, s12, s13)
        assert got == pytest.approx(1j * z / (m0 * g0), rel=1e-10, abs=1e-12 / (m0 * g0)), (cos, got)


def _barrier(L, x):
    """x^(2L+2) |h_L(x)|^2, h_L the spherical Hankel function of the first kind, from scipy's
    spherical Bessel functions: the Blatt-Weisskopf denominator of angular momentum L in
    x = radius * momentum."""
    return x ** (2 * L + 2) * (special.spherical_jn(L, x) ** 2 + special.spherical_yn(L, x) ** 2)


@pytest.mark.parametrize("pair", [0, 1, 2])
@pytest.mark.parametrize("spin,m0,g0", [(1, 0.98, 0.07), (1, 1.45, 0.3), (2, 1.27, 0.185)])
def test_rbw_off_pole_blatt_weisskopf(pair, spin, m0, g0):
    """plainBW with normalised barriers (RBW.cu:9-86) across the Dalitz plot:
    Z F_R F_D / (m0^2 - s - i m0 Gamma(s)), Gamma(s) = Gamma0 (q/q0)^(2L+1) (m0/sqrt s) F_R^2.
    F_R^2 = D_L(r q0) / D_L(r q) with D_L from the spherical Hankel functions and q the pair's
    breakup momentum; F_D^2 the same in the mother radius with the mother-side momentum of
    twoBodyCMMothermom (Resonance.cu:62-75), whose clamped square roots make it s / (2 M)
    inside the Dalitz plot; Z the Zemach factor from four-vectors in the pair's rest frame."""
    M, r, rm = W.DALITZ_DECAY[0], W.DALITZ_DECAY[4], W.DALITZ_DECAY[5]
    mass = {1: W.DALITZ_DECAY[1], 2: W.DALITZ_DECAY[2], 3: W.DALITZ_DECAY[3]}
    a, b, c = PAIR_ABC[pair]
    lo, hi = (mass[a] + mass[b]) ** 2, (M - mass[c]) ** 2
    op, rf = _lineshape_prog("RBW", m0, g0, spin, pair)
    q0 = _p_cm(m0 * m0, mass[a], mass[b])
    wants, gots = [], []
    for s in np.linspace(lo + 0.03 * (hi - lo), hi - 0.03 * (hi - lo), 7):
        for cos in (-0.7, 0.2, 0.85):
            s12, s13, dot, mag2 = _pole_config(pair, math.sqrt(s), cos)
            q = _p_cm(s, mass[a], mass[b])
            fr2 = _barrier(spin, r * q0) / _barrier(spin, r * q)
            fd2 = _barrier(spin, rm * m0 * m0 / (2 * M)) / _barrier(spin, rm * s / (2 * M))
            gam = g0 * (q / q0) ** (2 * spin + 1) * (m0 / math.sqrt(s)) * fr2
            z = -4 * dot if spin == 1 else 16 * (dot * dot - mag2 / 3)
            wants.append(z * math.sqrt(fr2 * fd2) / complex(m0 * m0 - s, -m0 * gam))
            gots.append(op.lineshape(rf, s12, s13))
    scale = max(abs(w) for w in wants)
    for w, g in zip(wants, gots):
        assert g == pytest.approx(w, rel=1e-10, abs=1e-12 * scale), (w, g)


def test_gs_pole_rho():
    """Gounaris-Sakurai (GS.cu:9-75) for the rho0 -> pi+ pi- (spin 1, pair 23) at s = m0^2:
    i (1 + d(m0^2) Gamma0/m0) Z / (m0 Gamma0), with the GS d-function for two equal-mass pions
    (Gounaris & Sakurai, PRL 21 (1968) 244) and the Zemach factor from four-vectors."""
    mpi = W.DALITZ_DECAY[2]
    m0, g0 = 0.77526, 0.1491
    op, rf = _lineshape_prog("GS", m0, g0, 1, 2)
    k0 = math.sqrt(m0 * m0 / 4 - mpi * mpi)
    d = (3 / math.pi) * mpi ** 2 / k0 ** 2 * math.log((m0 + 2 * k0) / (2 * mpi)) + m0 / (2 * math.pi * k0) \
        - mpi ** 2 * m0 / (math.pi * k0 ** 3)
    for cos in (-0.7, 0.2, 0.85):
        s12, s13, dot, _ = _pole_config(2, m0, cos)
        want = 1j * (1 + d * g0 / m0) * (-4 * dot) / (m0 * g0)
        assert op.lineshape(rf, s12, s13) == pytest.approx(want, rel=1e-10)


def test_gs_off_pole_rho():
    """Gounaris-Sakurai (GS.cu:9-75) for the rho0 -> pi+ pi- (spin 1, pair 23) across the plot:
    (1 + d Gamma0/m0) Z F_R F_D / (m0^2 - s + f(s) - i m0 Gamma0 (k/k0)^3 F_R^2) with
    f(s) = Gamma0 m0^2 / k0^3 [k^2 (h(s) - h(m0^2)) + (m0^2 - s) k0^2 h'(m0^2)] and
    h(s) = (2/pi) (k/sqrt s) ln((sqrt s + 2k) / (2 m_pi)) (Gounaris & Sakurai, PRL 21 (1968)
    244); h' by complex-step differentiation; the width term without the textbook m0/sqrt(s)
    factor, as GS.cu:64 writes it; barriers as in the RBW check."""
    import cmath
    M, r, rm, mpi = W.DALITZ_DECAY[0], W.DALITZ_DECAY[4], W.DALITZ_DECAY[5], W.DALITZ_DECAY[2]
    mc = W.DALITZ_DECAY[1]
    m0, g0 = 0.77526, 0.1491
    op, rf = _lineshape_prog("GS", m0, g0, 1, 2)

    def h(s):
        k = cmath.sqrt(s / 4 - mpi * mpi)
        return (2 / math.pi) * (k / cmath.sqrt(s)) * cmath.log((cmath.sqrt(s) + 2 * k) / (2 * mpi))

    s0 = m0 * m0
    k0 = math.sqrt(s0 / 4 - mpi * mpi)
    dh0 = h(complex(s0, 1e-30)).imag / 1e-30
    d = (3 / math.pi) * mpi ** 2 / k0 ** 2 * math.log((m0 + 2 * k0) / (2 * mpi)) + m0 / (2 * math.pi * k0) \
        - mpi ** 2 * m0 / (math.pi * k0 ** 3)
    lo, hi = 4 * mpi * mpi, (M - mc) ** 2
    wants, gots = [], []
    for s in np.linspace(lo + 0.03 * (hi - lo), hi - 0.03 * (hi - lo), 9):
        k = math.sqrt(s / 4 - mpi * mpi)
        f = g0 * s0 / k0 ** 3 * (k * k * (h(s).real - h(s0).real) + (s0 - s) * k0 * k0 * dh0)
        fr2 = _barrier(1, r * k0) / _barrier(1, r * k)
        fd2 = _barrier(1, rm * s0 / (2 * M)) / _barrier(1, rm * s / (2 * M))
        for cos in (-0.7, 0.2, 0.85):
            s12, s13, dot, _ = _pole_config(2, math.sqrt(s), cos)
            den = complex(s0 - s + f, -m0 * g0 * (k / k0) ** 3 * fr2)
            wants.append((1 + d * g0 / m0) * (-4 * dot) * math.sqrt(fr2 * fd2) / den)
            gots.append(op.lineshape(rf, s12, s13))
    scale = max(abs(w) for w in wants)
    for w, g in zip(wants, gots):
        assert g == pytest.approx(w, rel=1e-9, abs=1e-12 * scale), (w, g)


def test_nonres_is_one():
    op, rf = _lineshape_prog("NONRES", 0, 0, 0, 0)
    assert op.lineshape(rf, 1.0, 1.0) == 1.0


# ------------------------------------------------------------ sums and invariants

def test_chunked_sum_order():
    """A9: P contiguous chunks (the first n % P one longer), each summed from its first element, then
    the chunk sums in order, starting from 0."""
    t = np.array([1e16, 1.0, -1e16, 1.0, 3.0, 1e-3, 2.0])
    for P in (1, 2, 3, 4, 7):
        n, base, extra = t.size, t.size // P, t.size % P
        sums, i = [], 0
        for c in range(P):
            ln = base + (1 if c < extra else 0)
            acc = 0.0
            for v in t[i:i + ln]:
                acc += v
            sums.append(acc)
            i += ln
        want = 0.0
        for v in sums:
            want += v
        assert O.chunked_sum(t, P) == want, P
    # exact_sum is exact on a cancellation-heavy vector
    rng = np.random.default_rng(2)
    big = rng.normal(size=20000) * 10.0 ** rng.integers(-8, 8, size=20000)
    assert O.exact_sum(big) == math.fsum(big)


def test_dalitz_integrals_hermitian_psd_and_area():
    """Integrals I_ij (SpecialResonanceIntegrator.cu:52-121): Hermitian, positive semi-definite; a single
    flat wave integrates to the Dalitz-plot area, computed as the integral of 4 p1* p3* over s12."""
    top, obs = _dalitz_small()
    decay = W.DALITZ_DECAY
    op = O.OracleProg(top, decay)
    res, eff = op.dalitz_layout(0)
    I = op.dalitz_integrals(0, [0.0, 3.0, 200, 0.0, 3.0, 200], eff)
    np.testing.assert_allclose(I, I.conj().T, rtol=0, atol=1e-12 * abs(I).max())
    ev = np.linalg.eigvalsh(I)
    assert ev.min() > -1e-10 * ev.max()
    # flat amplitude
    opn, rf = _lineshape_prog("NONRES", 0, 0, 0, 0)
    _, eff1 = opn.dalitz_layout(0)
    n = 600
    I1 = opn.dalitz_integrals(0, [0.0, 3.0, n, 0.0, 3.0, n], eff1)
    area_grid = I1[0, 0].real * (3.0 / n) ** 2
    M, m1, m2, m3 = decay[:4]

    def width(s12):
        m12 = math.sqrt(s12)
        e1 = (s12 - m2 * m2 + m1 * m1) / (2 * m12)
        e3 = (M * M - s12 - m3 * m3) / (2 * m12)
        return 4 * math.sqrt(max(e1 * e1 - m1 * m1, 0)) * math.sqrt(max(e3 * e3 - m3 * m3, 0))

    from scipy.integrate import quad
    area = quad(width, (m1 + m2) ** 2, (M - m3) ** 2, limit=200)[0]
    assert area_grid == pytest.approx(area, rel=3e-3)  # midpoint grid of 600^2 cells on a curved boundary


def _dalitz_small():
    import sys, os
    sys.path.insert(0, os.path.dirname(__file__))
    import _models as M
    table = W.DALITZ_RESONANCES[:5]
    return M.dalitz(resonances=table, nbins=200)


def test_dalitz_norm_scales_quadratically():
    """NLL(c) invariant under c -> lambda c: the norm is a quadratic form in the couplings."""
    top, obs = _dalitz_small()
    op = O.OracleProg(top, W.DALITZ_DECAY)
    res, eff = op.dalitz_layout(0)
    I = op.dalitz_integrals(0, [0.0, 3.0, 120, 0.0, 3.0, 120], eff)
    c = W.dalitz_couplings(3, len(res))
    n1 = O.dalitz_norm_from_integrals(c, I, 0.025, 0.025)
    n2 = O.dalitz_norm_from_integrals(2.5 * np.asarray(c), I, 0.025, 0.025)
    assert n2 == pytest.approx(6.25 * n1, rel=1e-13)
    assert n1 == pytest.approx((np.conj(c) @ I.T @ c).real * 0.025 ** 2, rel=1e-12)


# ------------------------------------------------------------ whole likelihoods

def _models():
    import os
    import sys
    sys.path.insert(0, os.path.dirname(__file__))
    import _models as M
    return M


def test_add1d_nll_vs_scipy():
    """GooPdf::calculateNLL of f*Gauss + (1-f)*Exp (AddPdf.cu:16-51, GaussianPdf.cu, ExpPdf.cu):
    -sum log(f N(x; mu, sigma) + (1-f) alpha e^(alpha x) / (e^(alpha hi) - e^(alpha lo)))."""
    M = _models()
    top, x = M.add1d()
    xs = W.sample_add1d(3000, seed=7)
    got = O.Model(top, aos=xs[:, None]).calculate_nll()
    a, f = -0.4, 0.3
    pdf = f * stats.norm.pdf(xs, 0.3, 0.8) + (1 - f) * a * np.exp(a * xs) / (np.exp(5 * a) - np.exp(-5 * a))
    assert got == pytest.approx(math.fsum(-np.log(pdf)), rel=1e-12)


def test_prod3d_nll_vs_scipy():
    """ProdPdf of Gauss(x), Exp(y) and ExpGaus(z) (ProdPdf.cu:9-26); ExpGaus normalised on its 200-bin
    midpoint grid (GooPdf.cu:340-396)."""
    M = _models()
    top, obs = M.prod3d()
    x, y, z = W.sample_prod3d(3000, seed=8)
    got = O.Model(top, aos=np.stack([x, y, z], 1)).calculate_nll()
    zc = grid_centres(0.0, 10.0, 200)
    emg = lambda v: stats.exponnorm.pdf(v, 1.0 / (0.1 / 1.5), loc=0.0, scale=0.1)
    zint = emg(zc).sum() * 10.0 / 200
    a = -1.2
    pdf = stats.norm.pdf(x, 0, 1) * (a * np.exp(a * y) / (np.exp(8 * a) - 1.0)) * emg(z) / zint
    assert got == pytest.approx(math.fsum(-np.log(pdf)), rel=1e-9)


SPIN0_TABLE = [
    ("nonr", "NONRES", 0.0, 0.0, 0, 0),
    ("f0_980", "RBW", 0.980, 0.044, 0, 2),
    ("s12_1200", "RBW", 1.200, 0.300, 0, 0),
    ("s13_1400", "RBW", 1.400, 0.200, 0, 1),
]


def _inside(s12, s13, M, m1, m2, m3):
    m12 = np.sqrt(np.maximum(s12, 1e-300))
    e1 = (s12 - m2 * m2 + m1 * m1) / (2 * m12)
    e3 = (M * M - s12 - m3 * m3) / (2 * m12)
    p1 = np.sqrt(np.maximum(e1 * e1 - m1 * m1, 0))
    p3 = np.sqrt(np.maximum(e3 * e3 - m3 * m3, 0))
    lo = (e1 + e3) ** 2 - (p1 + p3) ** 2
    hi = (e1 + e3) ** 2 - (p1 - p3) ** 2
    return (s12 >= (m1 + m2) ** 2) & (s12 <= (M - m3) ** 2) & (s13 >= lo) & (s13 <= hi)


def test_dalitz_nll_spin0_model_vs_textbook():
    """Amp3Body NLL composed end to end (Amp3Body.cu:60-122, 238-359): coherent sum sum_i c_i W_i with
    spin-0 RBW waves from the textbook formula, |A|^2 on a 200 x 200 midpoint grid for the norm, times
    the bin areas; events outside the plot contribute 0 (A8)."""
    M_ = _models()
    top, obs = M_.dalitz(resonances=SPIN0_TABLE, nbins=200)
    aos = M_.dalitz_events(4000, seed=9)
    decay = W.DALITZ_DECAY
    got = O.Model(top, aos=aos, decay=decay).calculate_nll()

    M, m1, m2, m3 = decay[:4]
    msum = M * M + m1 * m1 + m2 * m2 + m3 * m3
    coups = np.asarray(W.dalitz_couplings(3, len(SPIN0_TABLE)))
    dm = {0: (m1, m2), 1: (m1, m3), 2: (m2, m3)}

    def amp(s12, s13):
        s = {0: s12, 1: s13, 2: msum - s12 - s13}
        A = np.zeros_like(s12, dtype=complex)
        for (name, kind, m0, g0, spin, pair), c in zip(SPIN0_TABLE, coups):
            if kind == "NONRES":
                A += c
                continue
            ss = s[pair]
            ma, mb = dm[pair]
            lam = lambda t: (t - (ma + mb) ** 2) * (t - (ma - mb) ** 2)
            q = np.sqrt(np.maximum(lam(ss), 0)) / (2 * np.sqrt(ss))
            q0 = math.sqrt(lam(m0 * m0)) / (2 * m0)
            gam = g0 * (q / q0) * (m0 / np.sqrt(ss))
            A += c / (m0 * m0 - ss - 1j * m0 * gam)
        return A

    c = grid_centres(0.0, 3.0, 200)
    S12, S13 = np.meshgrid(c, c, indexing="ij")
    ins = _inside(S12, S13, *decay[:4])
    norm = math.fsum((np.abs(amp(S12[ins], S13[ins])) ** 2).ravel()) * (3.0 / 200) ** 2
    e12, e13 = aos[:, 0], aos[:, 1]
    ev_in = _inside(e12, e13, *decay[:4])
    terms = -np.log(np.abs(amp(e12[ev_in], e13[ev_in])) ** 2 / norm)
    assert got == pytest.approx(math.fsum(terms), rel=1e-10)
