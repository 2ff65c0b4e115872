"""Pins of the oracle's remaining PDF functions on the CPU.

The polynomial and exponential-polynomial leaves, the extended AddPdf, the
two-channel RBW (RBW_Sym) and the lineshapes with unnormalised barriers
(norm = false) are fixed here by something other than the oracle's own
formula: numpy polynomial evaluation of the coefficient layout the reference's
constructors register, scipy densities, the coefficient order the reference
documents for PolynomialPdf, and lineshapes recomputed from four-vectors.
A dropped term, a shifted power, a swapped coefficient or a wrong channel
fails one of them.
"""
import cmath
import math

import numpy as np
import pytest
from numpy.polynomial import polynomial as npoly
from scipy import stats

import oracle as O
from goofit_amd import workloads as W


def _leaf(code, x, params, consts=()):
    p = O.Pdf("leaf", [x], params, consts)
    p.func = O.FN[code]
    return p


def _values(pdf, xs):
    op = O.OracleProg(pdf)
    return np.array([op.eval_event([float(v)]) for v in xs])


XS = np.array([-2.3, -1.0, -0.37, 0.0, 0.25, 0.9, 1.7, 3.1])


# ------------------------------------------------------------ 1-D leaves

@pytest.mark.parametrize("lowest", [0, 1, 3])
def test_polynomial_closed_form(lowest):
    """device_Polynomial (PolynomialPdf.cu:8-25) with the layout of the one-observable constructor
    (PolynomialPdf.cu:124-136: constant lowestDegree, then the weights): sum_i w_i x^(lowest + i)."""
    w = [0.7, -1.3, 0.45, 2.0]
    x = O.Obs("x", -5, 5)
    got = _values(_leaf("POLYNOMIAL", x, [O.Var(f"w{i}", v) for i, v in enumerate(w)], [lowest]), XS)
    want = npoly.polyval(XS, [0.0] * lowest + w)
    np.testing.assert_allclose(got, want, rtol=1e-13, atol=1e-14)


def test_offset_polynomial_closed_form():
    """device_OffsetPolynomial (PolynomialPdf.cu:28-45) with the layout of the offset constructor
    (PolynomialPdf.cu:138-153: the weights, then x0 as the last parameter).  As written there, the
    event value is not shifted, the sum starts at parameter 2 and the power is lowest + i - 2, so the
    first two weights drop out and x0 multiplies the highest power."""
    w = [0.7, -1.3, 0.45, 2.0]
    x0 = 0.6
    lowest = 1
    x = O.Obs("x", -5, 5)
    params = [O.Var(f"w{i}", v) for i, v in enumerate(w)] + [O.Var("x0", x0)]
    got = _values(_leaf("OFFSET_POLYNOMIAL", x, params, [lowest]), XS)
    coeffs = [0.0] * lowest + [0.45, 2.0, x0]  # parameters 2, 3, 4 at powers lowest, lowest+1, lowest+2
    np.testing.assert_allclose(got, npoly.polyval(XS, coeffs), rtol=1e-13, atol=1e-14)


def test_exp_offset_closed_form():
    """device_ExpOffset (ExpPdf.cu:19-30) with ExpPdf(name, x, alpha, offset)'s layout
    (ExpPdf.cu:82-88: offset first, then alpha): exp(alpha (x - x0))."""
    alpha, x0 = -0.8, 1.25
    x = O.Obs("x", -5, 5)
    got = _values(_leaf("EXP_OFFSET", x, [O.Var("x0", x0), O.Var("alpha", alpha)]), XS)
    np.testing.assert_allclose(got, np.exp(alpha * (XS - x0)), rtol=1e-14)


def test_exp_poly_closed_form():
    """device_ExpPoly (ExpPdf.cu:32-48), weights from the constant term up: exp(sum_i w_i x^i)."""
    w = [0.3, -0.5, 0.12, -0.04]
    x = O.Obs("x", -5, 5)
    got = _values(_leaf("EXP_POLY", x, [O.Var(f"w{i}", v) for i, v in enumerate(w)]), XS)
    np.testing.assert_allclose(got, np.exp(npoly.polyval(XS, w)), rtol=1e-13)


def test_exp_poly_offset_closed_form():
    """device_ExpPolyOffset (ExpPdf.cu:50-67) with ExpPdf(name, x, weights, offset)'s layout
    (ExpPdf.cu:105-118: offset first): the loop starts at parameter 1, so weight k multiplies
    (x - x0)^(k + 1) and there is no constant term."""
    w = [-0.5, 0.12, -0.04]
    x0 = -0.7
    x = O.Obs("x", -5, 5)
    params = [O.Var("x0", x0)] + [O.Var(f"w{i}", v) for i, v in enumerate(w)]
    got = _values(_leaf("EXP_POLY_OFFSET", x, params), XS)
    np.testing.assert_allclose(got, np.exp(npoly.polyval(XS - x0, [0.0] + w)), rtol=1e-13)


# ------------------------------------------------------------ MultiPolynomial

def _doc_order_2d(max_degree):
    """Coefficient order of the multivariate PolynomialPdf as its header documents it
    (include/goofit/PDFs/basic/PolynomialPdf.h:22-33): start at the constant term, raise the power of
    the leftmost observable until the total degree reaches the maximum, then reset it and raise the
    next one; for degree 3 in (x, y): a00 a10 a20 a30 a01 a11 a21 a02 a12 a03."""
    out = []
    for py in range(max_degree + 1):
        for px in range(max_degree + 1 - py):
            out.append((px, py))
    return out


def test_multipoly_documented_order_example():
    assert _doc_order_2d(3) == [(0, 0), (1, 0), (2, 0), (3, 0), (0, 1), (1, 1), (2, 1), (0, 2), (1, 2), (0, 3)]


@pytest.mark.parametrize("deg", [2, 3])
def test_multipoly_2d_unit_coefficients_follow_documented_order(deg):
    """device_MultiPolynomial (PolynomialPdf.cu:47-118): with coefficient k set to 1 and every other 0,
    the value is the k-th monomial of the documented order, in (x - x0, y - y0)."""
    order = _doc_order_2d(deg)
    x, y = O.Obs("x", -3, 3), O.Obs("y", -3, 3)
    x0, y0 = 0.4, -0.3
    rng = np.random.default_rng(5)
    pts = rng.uniform(-2, 2, size=(6, 2))
    for k, (px, py) in enumerate(order):
        coeffs = [O.Var(f"c{i}", 1.0 if i == k else 0.0) for i in range(len(order))]
        mp = O.MultiPolynomial("mp", [x, y], coeffs, [O.Var("x0", x0), O.Var("y0", y0)], deg)
        assert len(mp.params) == 2 + len(order)  # offsets, then exactly one coefficient per monomial
        op = O.OracleProg(mp)
        for a, b in pts:
            want = (a - x0) ** px * (b - y0) ** py
            assert op.eval_event([a, b]) == pytest.approx(want, rel=1e-13, abs=1e-15), (k, px, py)


def test_multipoly_3d_brute_force_monomial_sum():
    """Three observables, degree 2: the value equals a brute-force sum over every monomial of total
    degree <= 2 with the coefficients in the documented order (leftmost power fastest)."""
    deg = 2
    order = []
    for pz in range(deg + 1):
        for py in range(deg + 1 - pz):
            for px in range(deg + 1 - py - pz):
                order.append((px, py, pz))
    assert len(order) == 10
    rng = np.random.default_rng(6)
    c = rng.normal(size=len(order))
    off = rng.normal(scale=0.3, size=3)
    obs = [O.Obs(n, -3, 3) for n in "xyz"]
    mp = O.MultiPolynomial("mp", obs, [O.Var(f"c{i}", v) for i, v in enumerate(c)],
                           [O.Var(f"o{i}", v) for i, v in enumerate(off)], deg)
    op = O.OracleProg(mp)
    for p in rng.uniform(-2, 2, size=(20, 3)):
        d = p - off
        want = sum(ck * d[0] ** a * d[1] ** b * d[2] ** g for ck, (a, b, g) in zip(c, order))
        assert op.eval_event(list(p)) == pytest.approx(want, rel=1e-12, abs=1e-13)


# ------------------------------------------------------------ extended AddPdf

def test_add_extended_nll_vs_scipy():
    """AddPdf with one yield per component (AddPdf.cu:53-79 per event, normalize at :143-185, the
    extended term at :187-204): NLL = sum_e -log(sum_i N_i f_i(x_e) / sum_i N_i) + E - n log E with
    E = sum_i N_i and f_i the normalised component densities: scipy's normal density (the Gaussian is
    normalised over the real line, hazard A1) and the exponential truncated to the observable range."""
    lo, hi = -5.0, 5.0
    mu, sigma, alpha = 0.3, 0.8, -0.4
    n_sig, n_bkg = 3150.0, 6800.0
    x = O.Obs("x", lo, hi, 1000)
    g = O.Gaussian("g", x, O.Var("mu", mu), O.Var("sigma", sigma))
    e = O.Exp("e", x, O.Var("alpha", alpha))
    top = O.Add("ext", [O.Var("nsig", n_sig), O.Var("nbkg", n_bkg)], [g, e])
    assert top.extended
    data = W.sample_add1d(10_000, seed=21)
    got = O.Model(top, data.reshape(-1, 1)).calculate_nll()
    f_sig = stats.norm.pdf(data, mu, sigma)
    f_bkg = alpha * np.exp(alpha * data) / (math.exp(alpha * hi) - math.exp(alpha * lo))
    tot = n_sig + n_bkg
    want = math.fsum(-np.log((n_sig * f_sig + n_bkg * f_bkg) / tot)) + tot - data.size * math.log(tot)
    assert got == pytest.approx(want, rel=1e-12)


def test_add_extended_three_components_weights():
    """Extended sum of three components at single events: sum_i N_i f_i / sum_i N_i with every component
    normalised (the weight of the last component is its own yield, not 1 - sum of the others)."""
    lo, hi = 0.0, 4.0
    x = O.Obs("x", lo, hi, 1000)
    g1 = O.Gaussian("g1", x, O.Var("m1", 1.0), O.Var("s1", 0.3))
    g2 = O.Gaussian("g2", x, O.Var("m2", 2.5), O.Var("s2", 0.5))
    e = O.Exp("e", x, O.Var("a", -0.9))
    ns = [120.0, 300.0, 580.0]
    top = O.Add("ext3", [O.Var(f"n{i}", v) for i, v in enumerate(ns)], [g1, g2, e])
    m = O.Model(top, np.array([[0.5], [1.0], [2.2], [3.9]]))
    terms = m.nll_terms()
    xs = np.array([0.5, 1.0, 2.2, 3.9])
    f = [stats.norm.pdf(xs, 1.0, 0.3), stats.norm.pdf(xs, 2.5, 0.5),
         -0.9 * np.exp(-0.9 * xs) / (math.exp(-0.9 * hi) - math.exp(-0.9 * lo))]
    want = -np.log(sum(n * fi for n, fi in zip(ns, f)) / sum(ns))
    np.testing.assert_allclose(terms, want, rtol=1e-12)


# ------------------------------------------------------------ lineshapes

MASS = {1: W.DALITZ_DECAY[1], 2: W.DALITZ_DECAY[2], 3: W.DALITZ_DECAY[3]}
MOTHER, R_MESON, R_MOTHER = W.DALITZ_DECAY[0], W.DALITZ_DECAY[4], W.DALITZ_DECAY[5]
PAIRS = {0: (1, 2, 3), 1: (3, 1, 2), 2: (2, 3, 1)}  # pair index -> (A, B, bachelor C), the spin-factor roles


def _lineshape(kind, mass, width, spin, pair, norm=True, sym=False):
    m12, m13 = O.Obs("m12", 0.0, 3.0, 10), O.Obs("m13", 0.0, 3.0, 10)
    num = O.Obs("evt", 0.0, 1e9, 1, event_number=True)
    ar, ai = O.Var("re", 1.0), O.Var("im", 0.0)
    if kind == "RBW":
        r = O.RBW("r", ar, ai, O.Var("m", mass), O.Var("w", width), spin, pair, norm=norm, sym=sym)
    else:
        r = O.GS("r", ar, ai, O.Var("m", mass), O.Var("w", width), spin, pair, norm=norm)
    eff = O.MultiPolynomial("eff", [m12, m13], [O.Var("one", 1.0)], [O.Var("z0", 0.0), O.Var("z1", 0.0)], 0)
    top = O.Dalitz("dz", m12, m13, num, W.DALITZ_DECAY, [r], eff)
    op = O.OracleProg(top, W.DALITZ_DECAY)
    res, _ = op.dalitz_layout(0)
    return lambda s12, s13: op.lineshape(res[0], s12, s13)


def _mink(p):
    return p[0] ** 2 - p[1:] @ p[1:]


def _boost_to_rest(p, frame):
    """Boost four-vector p into the rest frame of four-vector `frame`."""
    beta = -frame[1:] / frame[0]
    b2 = beta @ beta
    if b2 == 0:
        return p
    g = 1 / math.sqrt(1 - b2)
    bp = p[1:] @ beta
    return np.array([g * (p[0] + bp), *(p[1:] + ((g - 1) * bp / b2 + g * p[0]) * beta)])


def _daughters(s12, s13):
    """Four-vectors of daughters 1, 2, 3 in the mother's rest frame for the invariants (s12, s13)."""
    M, m1, m2, m3 = MOTHER, MASS[1], MASS[2], MASS[3]
    s23 = M * M + m1 * m1 + m2 * m2 + m3 * m3 - s12 - s13
    e1, e2 = (M * M + m1 * m1 - s23) / (2 * M), (M * M + m2 * m2 - s13) / (2 * M)
    q1, q2 = math.sqrt(e1 * e1 - m1 * m1), math.sqrt(e2 * e2 - m2 * m2)
    cos12 = (e1 * e2 - 0.5 * (s12 - m1 * m1 - m2 * m2)) / (q1 * q2)
    p1 = np.array([e1, 0.0, 0.0, q1])
    p2 = np.array([e2, q2 * math.sqrt(max(0.0, 1 - cos12 * cos12)), 0.0, q2 * cos12])
    p3 = np.array([M, 0.0, 0.0, 0.0]) - p1 - p2
    return {1: p1, 2: p2, 3: p3}


def _pair_kinematics(s12, s13, pair):
    """Pair mass squared, breakup momentum q, the bachelor's momentum in the mother frame, and the
    Zemach factors of spin 1 and 2 (from pA, pC in the AB rest frame: -4 pA.pC and
    16 [(pA.pC)^2 - |pA|^2 |pC|^2 / 3]) at the decay with invariants (s12, s13)."""
    d = _daughters(s12, s13)
    a, b, c = PAIRS[pair]
    ab = d[a] + d[b]
    pa, pc = _boost_to_rest(d[a], ab)[1:], _boost_to_rest(d[c], ab)[1:]
    dot = pa @ pc
    z1, z2 = -4 * dot, 16 * (dot * dot - (pa @ pa) * (pc @ pc) / 3)
    return _mink(ab), math.sqrt(pa @ pa), math.sqrt(d[c][1:] @ d[c][1:]), {0: 1.0, 1: z1, 2: z2}


def _breakup(s, ma, mb):
    e = (s + ma * ma - mb * mb) / (2 * math.sqrt(s))
    return math.sqrt(e * e - ma * ma)


def _unnormalised_barrier(q, spin, radius):
    """dampingFactorSquare as Resonance.cu:77-86 writes it, z = (r q)^2: 2 z for spin 1 and, for spin 2,
    the nonstandard 13 z / (z - 3)^2 + 9 z^2 (hazard A5)."""
    z = radius * radius * q * q
    return 13 * z / (z - 3) ** 2 + 9 * z * z if spin == 2 else 2 * z


def _plot_points(pair, n_s=3):
    """Decays across the Dalitz plot: n_s pair masses between threshold and the kinematic limit,
    three bachelor directions each, returned as (s12, s13)."""
    M = MOTHER
    a, b, c = PAIRS[pair]
    lo, hi = (MASS[a] + MASS[b]) ** 2, (M - MASS[c]) ** 2
    out = []
    for s in np.linspace(lo + 0.1 * (hi - lo), hi - 0.1 * (hi - lo), n_s):
        m = math.sqrt(s)
        k = _breakup(s, MASS[a], MASS[b])
        ec = (M * M - s - MASS[c] ** 2) / (2 * m)
        pcm = math.sqrt(ec * ec - MASS[c] ** 2)
        for cos in (-0.6, 0.15, 0.8):
            sin = math.sqrt(1 - cos * cos)
            four = {a: np.array([math.sqrt(MASS[a] ** 2 + k * k), 0, 0, k]),
                    b: np.array([math.sqrt(MASS[b] ** 2 + k * k), 0, 0, -k]),
                    c: np.array([ec, pcm * sin, 0, pcm * cos])}
            out.append((_mink(four[1] + four[2]), _mink(four[1] + four[3])))
    return out


@pytest.mark.parametrize("pair", [0, 1, 2])
@pytest.mark.parametrize("spin,m0,g0", [(1, 0.98, 0.07), (2, 1.27, 0.185)])
def test_rbw_unnormalised_barriers(pair, spin, m0, g0):
    """plainBW with norm = false (RBW.cu:9-86): Z sqrt(B(q) B_D(p)) / (m0^2 - s - i m0^2 Gamma0 (q/q0)^(2L+1)
    B(q) / sqrt(s)), with B the unnormalised damping factors at the pair's breakup momentum q (meson
    radius) and at the bachelor's momentum p in the mother's rest frame (mother radius); q, p and the
    Zemach factor Z from four-vectors."""
    f = _lineshape("RBW", m0, g0, spin, pair, norm=False)
    a, b, _ = PAIRS[pair]
    q0 = _breakup(m0 * m0, MASS[a], MASS[b])
    wants, gots = [], []
    for s12, s13 in _plot_points(pair):
        s, q, p_bach, zem = _pair_kinematics(s12, s13, pair)
        br, bd = _unnormalised_barrier(q, spin, R_MESON), _unnormalised_barrier(p_bach, spin, R_MOTHER)
        den = complex(m0 * m0 - s, -m0 * m0 * g0 * (q / q0) ** (2 * spin + 1) * br / math.sqrt(s))
        wants.append(zem[spin] * math.sqrt(br * bd) / den)
        gots.append(f(s12, s13))
    scale = max(abs(w) for w in wants)
    for w, g in zip(wants, gots):
        assert g == pytest.approx(w, rel=1e-10, abs=1e-12 * scale), (w, g)


def test_unnormalised_spin2_barrier_quirk_values():
    """The helper's spin-2 form (hazard A5, which the spin-2 lineshape check above uses) at three
    momenta against hand values of 13 z/(z-3)^2 + 9 z^2: z = 0.5 gives 6.5/6.25 + 2.25, z = 2 gives
    26 + 36, z = 4 gives 52 + 144."""
    assert _unnormalised_barrier(math.sqrt(0.5) / R_MESON, 2, R_MESON) == pytest.approx(6.5 / 6.25 + 2.25, rel=1e-14)
    assert _unnormalised_barrier(math.sqrt(2.0) / R_MESON, 2, R_MESON) == pytest.approx(26 + 36, rel=1e-14)
    assert _unnormalised_barrier(2.0 / R_MESON, 2, R_MESON) == pytest.approx(52 + 144, rel=1e-14)


def test_gs_unnormalised_barriers_rho():
    """Gounaris-Sakurai with norm = false (GS.cu:9-75) for the rho0 -> pi+ pi- (spin 1, pair 23):
    (1 + d Gamma0/m0) Z sqrt(B(k) B_D(p)) / (m0^2 - s + f(s) - i m0 Gamma0 (k/k0)^3 B(k)), the
    unnormalised damping factors as in the RBW check, f and d the published Gounaris-Sakurai
    functions (PRL 21 (1968) 244), h' by complex-step differentiation."""
    mpi = MASS[2]
    m0, g0 = 0.77526, 0.1491
    f_gs = _lineshape("GS", m0, g0, 1, 2, norm=False)

    def h(s):
        k = cmath.sqrt(s / 4 - mpi * mpi)
        return (2 / math.pi) * (k / cmath.sqrt(s)) * cmath.log((cmath.sqrt(s) + 2 * k) / (2 * mpi))

    s0 = m0 * m0
    k0 = math.sqrt(s0 / 4 - mpi * mpi)
    dh0 = h(complex(s0, 1e-30)).imag / 1e-30
    d = (3 / math.pi) * mpi ** 2 / k0 ** 2 * math.log((m0 + 2 * k0) / (2 * mpi)) + m0 / (2 * math.pi * k0) \
        - mpi ** 2 * m0 / (math.pi * k0 ** 3)
    wants, gots = [], []
    for s12, s13 in _plot_points(2, n_s=4):
        s, k, p_bach, zem = _pair_kinematics(s12, s13, 2)
        fs = g0 * s0 / k0 ** 3 * (k * k * (h(s).real - h(s0).real) + (s0 - s) * k0 * k0 * dh0)
        br, bd = _unnormalised_barrier(k, 1, R_MESON), _unnormalised_barrier(p_bach, 1, R_MOTHER)
        den = complex(s0 - s + fs, -m0 * g0 * (k / k0) ** 3 * br)
        wants.append((1 + d * g0 / m0) * zem[1] * math.sqrt(br * bd) / den)
        gots.append(f_gs(s12, s13))
    scale = max(abs(w) for w in wants)
    for w, g in zip(wants, gots):
        assert g == pytest.approx(w, rel=1e-9, abs=1e-12 * scale), (w, g)


def _rbw_normalised_channel(s12, s13, pair, spin, m0, g0):
    """One channel of plainBW with normalised barriers, from four-vectors and the spherical-Hankel
    Blatt-Weisskopf factors (hazard A23: the mother-side momentum is s / (2 M) inside the plot)."""
    from scipy import special

    def dl(L, x):
        return x ** (2 * L + 2) * (special.spherical_jn(L, x) ** 2 + special.spherical_yn(L, x) ** 2)

    a, b, _ = PAIRS[pair]
    s, q, _, zem = _pair_kinematics(s12, s13, pair)
    q0 = _breakup(m0 * m0, MASS[a], MASS[b])
    fr2 = fd2 = 1.0
    if spin:
        fr2 = dl(spin, R_MESON * q0) / dl(spin, R_MESON * q)
        fd2 = dl(spin, R_MOTHER * m0 * m0 / (2 * MOTHER)) / dl(spin, R_MOTHER * s / (2 * MOTHER))
    gam = g0 * (q / q0) ** (2 * spin + 1) * (m0 / math.sqrt(s)) * fr2
    return zem[spin] * math.sqrt(fr2 * fd2) / complex(m0 * m0 - s, -m0 * gam)


@pytest.mark.parametrize("pair", [0, 1, 2])
@pytest.mark.parametrize("spin,m0,g0", [(0, 0.98, 0.07), (1, 0.7758, 0.1503)])
def test_rbw_sym_is_sum_of_two_channels(pair, spin, m0, g0):
    """plainBW<2> (RBW.cu:9-89, ptr_to_RBW_Sym): the same resonance in the channel of its pair index and
    in the next one (pair + 1 mod 3), each channel recomputed independently from four-vectors."""
    f = _lineshape("RBW", m0, g0, spin, pair, sym=True)
    nxt = (pair + 1) % 3
    pts = [p for p in _plot_points(pair) if O.in_dalitz(p[0], p[1], *W.DALITZ_DECAY[:4])]
    assert pts
    for s12, s13 in pts:
        want = _rbw_normalised_channel(s12, s13, pair, spin, m0, g0) + \
            _rbw_normalised_channel(s12, s13, nxt, spin, m0, g0)
        assert f(s12, s13) == pytest.approx(want, rel=1e-9, abs=1e-12 * abs(want)), (s12, s13)


@pytest.mark.parametrize("pair", [0, 1])
def test_gs_unequal_pair_masses_a6(pair):
    """Gounaris-Sakurai (GS.cu:9-75) in the pi0 pi+- channels (pairs 12 and 13): the width term and the
    barriers take the pair's own breakup momenta (pi0 and pi+-), while h, h' and d always take
    daughters 2 and 3 (both pi+-), whatever the channel (hazard A6, GS.cu:62-63); the published
    Gounaris-Sakurai functions of two pions of mass m_pi+-, barriers normalised as in the RBW checks."""
    from scipy import special

    def dl(L, x):
        return x ** (2 * L + 2) * (special.spherical_jn(L, x) ** 2 + special.spherical_yn(L, x) ** 2)

    mpi = MASS[2]
    m0, g0 = 0.7758, 0.1503
    f_gs = _lineshape("GS", m0, g0, 1, pair)

    def h(s):
        k = cmath.sqrt(s / 4 - mpi * mpi)
        return (2 / math.pi) * (k / cmath.sqrt(s)) * cmath.log((cmath.sqrt(s) + 2 * k) / (2 * mpi))

    s0 = m0 * m0
    kk0 = math.sqrt(s0 / 4 - mpi * mpi)  # pi+ pi- momenta of the GS functions
    dh0 = h(complex(s0, 1e-30)).imag / 1e-30
    d = (3 / math.pi) * mpi ** 2 / kk0 ** 2 * math.log((m0 + 2 * kk0) / (2 * mpi)) + m0 / (2 * math.pi * kk0) \
        - mpi ** 2 * m0 / (math.pi * kk0 ** 3)
    a, b, _ = PAIRS[pair]
    q0 = _breakup(s0, MASS[a], MASS[b])  # the pair's own momenta: width term and barriers
    wants, gots = [], []
    for s12, s13 in _plot_points(pair, n_s=4):
        s, q, _, zem = _pair_kinematics(s12, s13, pair)
        kk = math.sqrt(s / 4 - mpi * mpi)
        fs = g0 * s0 / kk0 ** 3 * (kk * kk * (h(s).real - h(s0).real) + (s0 - s) * kk0 * kk0 * dh0)
        fr2 = dl(1, R_MESON * q0) / dl(1, R_MESON * q)
        fd2 = dl(1, R_MOTHER * s0 / (2 * MOTHER)) / dl(1, R_MOTHER * s / (2 * MOTHER))
        den = complex(s0 - s + fs, -m0 * g0 * (q / q0) ** 3 * fr2)
        wants.append((1 + d * g0 / m0) * zem[1] * math.sqrt(fr2 * fd2) / den)
        gots.append(f_gs(s12, s13))
    scale = max(abs(w) for w in wants)
    for w, g in zip(wants, gots):
        assert g == pytest.approx(w, rel=1e-9, abs=1e-12 * scale), (w, g)
