"""Synthetic workloads of the benchmark configurations (seeded numpy generators).

* ``sample_add1d``   -- 1-D AddPdf fit: 0.3 Gaussian(0.3, 0.8) + 0.7 Exp(-0.4) on [-5, 5].
* ``sample_prod3d``  -- 3-D ProdPdf: Gaussian(0, 1) on [-6, 6] x Exp(-1.2) on [0, 8] x
  Gaussian-convolved exponential decay (tau 1.5, resolution 0.1) on [0, 10].
* ``DALITZ_*``       -- the D0 -> pi+ pi- pi0 model with 16 resonances (the resonance
  list of the reference's examples/dalitz/dalitz.cpp:124-296, with the rho0 as a
  Gounaris-Sakurai lineshape as in examples/pipipi0DPFit/pipipi0DPFit.cpp:563-570),
  couplings drawn from a seed with |c| in [0.1, 2] and a uniform phase.
* ``dalitz_uniform`` -- points uniform over the Dalitz plot (flat phase space).
"""
from __future__ import annotations

import math

import numpy as np

# --------------------------------------------------------------------- 1-D / 3-D


def _trunc_gauss(rng, n, mu, sigma, lo, hi):
    out = np.empty(0)
    while out.size < n:
        x = rng.normal(mu, sigma, size=max(2 * (n - out.size), 1024))
        out = np.concatenate([out, x[(x >= lo) & (x <= hi)]])
    return out[:n]


def _trunc_exp(rng, n, alpha, lo, hi):
    """Inverse CDF of exp(alpha x) on [lo, hi]."""
    u = rng.random(n)
    if alpha == 0:
        return lo + u * (hi - lo)
    a, b = math.exp(alpha * lo), math.exp(alpha * hi)
    return np.log(a + u * (b - a)) / alpha


def sample_add1d(n=100_000, seed=1, frac=0.3, mu=0.3, sigma=0.8, alpha=-0.4, lo=-5.0, hi=5.0):
    rng = np.random.default_rng(seed)
    is_sig = rng.random(n) < frac
    ns = int(is_sig.sum())
    x = np.empty(n)
    x[is_sig] = _trunc_gauss(rng, ns, mu, sigma, lo, hi)
    x[~is_sig] = _trunc_exp(rng, n - ns, alpha, lo, hi)
    return x


def sample_emg(rng, n, mu, sigma, tau, lo, hi):
    out = np.empty(0)
    while out.size < n:
        m = max(2 * (n - out.size), 1024)
        z = rng.normal(mu, sigma, size=m) + rng.exponential(tau, size=m)
        out = np.concatenate([out, z[(z >= lo) & (z <= hi)]])
    return out[:n]


def sample_prod3d(n=10_000_000, seed=2, mu=0.0, sigma=1.0, alpha=-1.2, tau=1.5, res=0.1):
    rng = np.random.default_rng(seed)
    x = _trunc_gauss(rng, n, mu, sigma, -6.0, 6.0)
    y = _trunc_exp(rng, n, alpha, 0.0, 8.0)
    z = sample_emg(rng, n, 0.0, res, tau, 0.0, 10.0)
    return x, y, z


# ------------------------------------------------------------------------ Dalitz

DALITZ_MOTHER = 1.8648
DALITZ_PI0 = 0.13498
DALITZ_PIPM = 0.13957
DALITZ_MESON_RADIUS = 1.5
DALITZ_MOTHER_RADIUS = 5.0  # set explicitly: the reference leaves it uninitialised
# (motherMass, daug1 = pi0, daug2 = pi+, daug3 = pi-, meson radius, mother meson radius)
DALITZ_DECAY = (DALITZ_MOTHER, DALITZ_PI0, DALITZ_PIPM, DALITZ_PIPM, DALITZ_MESON_RADIUS, DALITZ_MOTHER_RADIUS)

PAIR_12, PAIR_13, PAIR_23 = 0, 1, 2

# name, kind, mass, width, spin, pair
DALITZ_RESONANCES = [
    ("nonr", "NONRES", 0.0, 0.0, 0, 0),
    ("rhop", "RBW", 0.7758, 0.1503, 1, PAIR_12),
    ("rho0", "GS", 0.7758, 0.1503, 1, PAIR_23),
    ("rhom", "RBW", 0.7758, 0.1503, 1, PAIR_13),
    ("rhop_1450", "RBW", 1.465, 0.400, 1, PAIR_12),
    ("rho0_1450", "RBW", 1.465, 0.400, 1, PAIR_23),
    ("rhom_1450", "RBW", 1.465, 0.400, 1, PAIR_13),
    ("rhop_1700", "RBW", 1.720, 0.250, 1, PAIR_12),
    ("rho0_1700", "RBW", 1.720, 0.250, 1, PAIR_23),
    ("rhom_1700", "RBW", 1.720, 0.250, 1, PAIR_13),
    ("f0_980", "RBW", 0.980, 0.044, 0, PAIR_23),
    ("f0_1370", "RBW", 1.434, 0.173, 0, PAIR_23),
    ("f0_1500", "RBW", 1.507, 0.109, 0, PAIR_23),
    ("f0_1710", "RBW", 1.714, 0.140, 0, PAIR_23),
    ("f2_1270", "RBW", 1.2754, 0.1851, 2, PAIR_23),
    ("f0_600", "RBW", 0.500, 0.400, 0, PAIR_23),
]
DALITZ_FIXED = "rhop"  # coupling held at (1, 0): removes the overall complex scale


def dalitz_couplings(seed=3, n=len(DALITZ_RESONANCES), fixed_index=1):
    """|c| ~ U[0.1, 2], phase ~ U[0, 2 pi); the reference coupling is (1, 0)."""
    rng = np.random.default_rng(seed)
    mag = rng.uniform(0.1, 2.0, size=n)
    ph = rng.uniform(0.0, 2 * math.pi, size=n)
    c = mag * np.exp(1j * ph)
    if fixed_index is not None and fixed_index < n:
        c[fixed_index] = 1.0 + 0.0j
    return c


def dalitz_box(decay=DALITZ_DECAY):
    M, m1, m2, m3 = decay[:4]
    return ((m1 + m2) ** 2, (M - m3) ** 2, (m1 + m3) ** 2, (M - m2) ** 2)


def in_dalitz_np(m12, m13, decay=DALITZ_DECAY):
    """Vectorised kinematic boundary (same arithmetic as the engine's inDalitz)."""
    M, d1, d2, d3 = decay[:4]
    m23 = M * M + d1 * d1 + d2 * d2 + d3 * d3 - m12 - m13
    ok = (m23 >= 0) & (m23 <= (M - d1) ** 2)
    ok &= (m12 >= (d1 + d2) ** 2) & (m12 <= (M - d3) ** 2)
    with np.errstate(invalid="ignore", divide="ignore"):
        s = np.sqrt(m12)
        e1 = 0.5 * (m12 - d2 * d2 + d1 * d1) / s
        e3 = 0.5 * (M * M - m12 - d3 * d3) / s
        r1 = np.sqrt(e1 * e1 - d1 * d1)
        r3 = np.sqrt(e3 * e3 - d3 * d3)
        lo = (e1 + e3) ** 2 - (r1 + r3) ** 2
        hi = (e1 + e3) ** 2 - (r1 - r3) ** 2
        ok &= (m13 >= lo) & (m13 <= hi)
    return ok


def dalitz_uniform(n, seed=4, decay=DALITZ_DECAY):
    """n points uniform over the Dalitz plot (rejection from the bounding box)."""
    rng = np.random.default_rng(seed)
    a, b, c, d = dalitz_box(decay)
    m12 = np.empty(0)
    m13 = np.empty(0)
    while m12.size < n:
        k = max(int(2.5 * (n - m12.size)), 4096)
        x = rng.uniform(a, b, size=k)
        y = rng.uniform(c, d, size=k)
        ok = in_dalitz_np(x, y, decay)
        m12 = np.concatenate([m12, x[ok]])
        m13 = np.concatenate([m13, y[ok]])
    return m12[:n], m13[:n]


def dalitz_area(decay=DALITZ_DECAY, n=4000):
    """Area of the Dalitz region in (m12, m13), by the m13 range integrated over m12."""
    M, d1, d2, d3 = decay[:4]
    lo, hi = (d1 + d2) ** 2, (M - d3) ** 2
    # Gauss-Legendre over m12 of (m13max - m13min)
    x, w = np.polynomial.legendre.leggauss(n)
    m12 = 0.5 * (hi - lo) * x + 0.5 * (hi + lo)
    s = np.sqrt(m12)
    e1 = 0.5 * (m12 - d2 * d2 + d1 * d1) / s
    e3 = 0.5 * (M * M - m12 - d3 * d3) / s
    r1 = np.sqrt(np.maximum(e1 * e1 - d1 * d1, 0))
    r3 = np.sqrt(np.maximum(e3 * e3 - d3 * d3, 0))
    width = 4 * r1 * r3
    return float(0.5 * (hi - lo) * np.sum(w * width))


def square_dalitz_rule(nodes, decay=DALITZ_DECAY, order=8):
    """Points and weights of a product quadrature over the Dalitz region in (m12, m13):
    m12 = lo + (hi - lo) (1 - cos(pi u)) / 2 and, at each m12, m13 = centre + half-width * t,
    with u in [0, 1] and t in [-1, 1] each a composite Gauss-Legendre rule of nodes // order
    panels of `order` nodes; weight = wu * wt * dm12/du * half-width.  The kinematic boundary
    is an edge of the (u, t) rectangle, so unlike the midpoint grid no cell straddles it, and
    the cosine map absorbs the square-root behaviour of the half-width at both ends of m12;
    all nodes lie strictly inside the plot."""
    M, d1, d2, d3 = decay[:4]
    panels = max(1, nodes // order)
    x, w = np.polynomial.legendre.leggauss(order)

    def composite(lo, hi):
        edges = np.linspace(lo, hi, panels + 1)
        half = 0.5 * np.diff(edges)
        mid = 0.5 * (edges[1:] + edges[:-1])
        return (mid[:, None] + half[:, None] * x[None, :]).ravel(), (half[:, None] * w[None, :]).ravel()

    lo, hi = (d1 + d2) ** 2, (M - d3) ** 2
    u, wu = composite(0.0, 1.0)
    m12 = lo + (hi - lo) * 0.5 * (1.0 - np.cos(np.pi * u))
    w12 = wu * (hi - lo) * 0.5 * np.pi * np.sin(np.pi * u)
    t, wt = composite(-1.0, 1.0)
    s = np.sqrt(m12)
    e1 = 0.5 * (m12 - d2 * d2 + d1 * d1) / s
    e3 = 0.5 * (M * M - m12 - d3 * d3) / s
    r1 = np.sqrt(np.maximum(e1 * e1 - d1 * d1, 0))
    r3 = np.sqrt(np.maximum(e3 * e3 - d3 * d3, 0))
    centre = (e1 + e3) ** 2 - r1 * r1 - r3 * r3
    half = 2.0 * r1 * r3
    P12 = np.repeat(m12, t.size)
    P13 = (centre[:, None] + half[:, None] * t[None, :]).ravel()
    W = (w12[:, None] * half[:, None] * wt[None, :]).ravel()
    return P12, P13, W


# --------------------------------------------------------- models (C++ API)


class Model:
    """A PDF built with the GooFit API plus the objects it references."""

    def __init__(self, pdf, observables, params, keep=()):
        self.pdf = pdf
        self.observables = observables
        self.params = params          # name -> Variable
        self._keep = list(keep)      # component PDFs, decay info, datasets


def build_add1d(start=1.0, nbins=1000):
    """0.3 Gaussian(0.3, 0.8) + 0.7 Exp(-0.4) on [-5, 5]; free parameters start at start*truth."""
    import goofit_amd as G
    G.require_api()
    x = G.Observable("x", -5.0, 5.0)
    x.numbins = nbins
    mu = G.Variable("mu", 0.3 * start, 0.01, -2.0, 2.0)
    sigma = G.Variable("sigma", 0.8 * start, 0.01, 0.1, 3.0)
    alpha = G.Variable("alpha", -0.4 * start, 0.01, -2.0, -0.01)
    frac = G.Variable("frac", 0.3 * start, 0.01, 0.0, 1.0)
    g = G.GaussianPdf("gauss", x, mu, sigma)
    e = G.ExpPdf("exp", x, alpha)
    s = G.AddPdf("sum", frac, g, e)
    return Model(s, [x], dict(mu=mu, sigma=sigma, alpha=alpha, frac=frac), keep=[g, e])


def build_prod3d(start=1.0, nbins=200):
    """Gaussian(x) x Exp(y) x ExpGaus(z); free mu, sigma, alpha and the decay rate."""
    import goofit_amd as G
    G.require_api()
    x = G.Observable("x", -6.0, 6.0)
    y = G.Observable("y", 0.0, 8.0)
    z = G.Observable("z", 0.0, 10.0)
    for o in (x, y, z):
        o.numbins = nbins
    mu = G.Variable("mu", 0.0 + 0.05 * (start - 1.0), 0.01, -1.0, 1.0)
    sigma = G.Variable("sigma", 1.0 * start, 0.01, 0.2, 3.0)
    alpha = G.Variable("alpha", -1.2 * start, 0.01, -5.0, -0.05)
    emu = G.Variable("emg_mu", 0.0)
    esig = G.Variable("emg_sigma", 0.1)
    rate = G.Variable("rate", (1.0 / 1.5) * start, 0.01, 0.05, 5.0)
    gx = G.GaussianPdf("gx", x, mu, sigma)
    ey = G.ExpPdf("ey", y, alpha)
    ez = G.ExpGausPdf("ez", z, emu, esig, rate)
    p = G.ProdPdf("prod", [gx, ey, ez])
    return Model(p, [x, y, z], dict(mu=mu, sigma=sigma, alpha=alpha, rate=rate), keep=[gx, ey, ez])


def reseed_widths(table, seed, lo=0.75, hi=1.25):
    """The resonance table with every width scaled by a factor uniform in [lo, hi) drawn
    from seed, kept inside [0.05, 0.4] GeV (a table width below 0.05 bounds it from below)."""
    rng = np.random.default_rng(seed)
    out = []
    for (name, kind, mass, width, spin, pair) in table:
        f = rng.uniform(lo, hi)
        out.append((name, kind, mass, float(np.clip(width * f, min(0.05, width), 0.4)) if kind != "NONRES" else width, spin, pair))
    return out


def build_dalitz(nbins=1414, coup_seed=3, resonances=None, float_couplings=True, coup_scale=1.0, width_seed=None):
    """D0 -> pi+ pi- pi0 with the 16 resonances of DALITZ_RESONANCES, flat efficiency,
    m12 and m13 on [0, 3] with an nbins x nbins normalisation grid.  width_seed: widths
    reseeded by reseed_widths."""
    import goofit_amd as G
    G.require_api()
    table = resonances if resonances is not None else DALITZ_RESONANCES
    if width_seed is not None:
        table = reseed_widths(table, width_seed)
    m12 = G.Observable("m12", 0.0, 3.0)
    m13 = G.Observable("m13", 0.0, 3.0)
    m12.numbins = nbins
    m13.numbins = nbins
    evt = G.EventNumber("eventNumber")
    coups = dalitz_couplings(coup_seed, len(table))
    params, res = {}, []
    for (name, kind, mass, width, spin, pair), c in zip(table, coups):
        c = c * (coup_scale if name != DALITZ_FIXED else 1.0)
        if name == DALITZ_FIXED or not float_couplings:
            ar, ai = G.Variable(name + "_re", c.real), G.Variable(name + "_im", c.imag)
        else:
            ar = G.Variable(name + "_re", c.real, 0.01, 0.0, 0.0)
            ai = G.Variable(name + "_im", c.imag, 0.01, 0.0, 0.0)
            params[name + "_re"], params[name + "_im"] = ar, ai
        if kind == "NONRES":
            r = G.Resonances.NonRes(name, ar, ai)
        elif kind == "GS":
            r = G.Resonances.GS(name, ar, ai, G.Variable(name + "_mass", mass), G.Variable(name + "_width", width),
                                spin, pair)
        else:
            r = G.Resonances.RBW(name, ar, ai, G.Variable(name + "_mass", mass), G.Variable(name + "_width", width),
                                 spin, pair)
        res.append(r)
    decay = G.DecayInfo3()
    decay.motherMass, decay.daug1Mass, decay.daug2Mass, decay.daug3Mass = DALITZ_DECAY[:4]
    decay.meson_radius, decay.mother_meson_radius = DALITZ_DECAY[4:]
    decay.resonances = res
    eff = G.PolynomialPdf("constantEff", [m12, m13], [G.Variable("constantOne", 1.0)],
                          [G.Variable("constantZero0", 0.0), G.Variable("constantZero1", 0.0)], 0)
    amp = G.Amp3Body("signalPDF", m12, m13, evt, decay, eff)
    return Model(amp, [m12, m13, evt], params, keep=res + [decay, eff])


TD_MIX = (0.4101, 0.0016, 0.0055, 0.0, 0.0)  # tau [ps], x, y, dx, dy of the pipipi0 example


def build_tddp(nbins=240, coup_seed=3, resonances=None, mix=TD_MIX, resolution="truth", mistag=False,
               charmtag=False, float_mixing=True, tg=None):
    """Amp3Body_TD of D0 -> pi0 pi+ pi- (the first six resonances of DALITZ_RESONANCES by
    default, couplings fixed), observables (dtime, sigma_t, m12, m13, evt[, mistag][, tag]);
    tau, x, y free when float_mixing.  resolution: "truth" or "threegauss" (parameters tg)."""
    import goofit_amd as G
    G.require_api()
    table = resonances if resonances is not None else DALITZ_RESONANCES[:6]
    dtime = G.Observable("dtime", -2.0, 20.0)
    sigma = G.Observable("sigmat", 0.0, 0.8)
    m12 = G.Observable("m12", 0.0, 3.0)
    m13 = G.Observable("m13", 0.0, 3.0)
    m12.numbins = nbins
    m13.numbins = nbins
    evt = G.EventNumber("eventNumber")
    coups = dalitz_couplings(coup_seed, len(table))
    res = []
    for (name, kind, mass, width, spin, pair), c in zip(table, coups):
        ar, ai = G.Variable(name + "_re", c.real), G.Variable(name + "_im", c.imag)
        if kind == "NONRES":
            res.append(G.Resonances.NonRes(name, ar, ai))
        elif kind == "GS":
            res.append(G.Resonances.GS(name, ar, ai, G.Variable(name + "_mass", mass),
                                       G.Variable(name + "_width", width), spin, pair))
        else:
            res.append(G.Resonances.RBW(name, ar, ai, G.Variable(name + "_mass", mass),
                                        G.Variable(name + "_width", width), spin, pair))
    if float_mixing:
        tau = G.Variable("tau", mix[0], 0.001, 0.3, 0.5)
        x = G.Variable("xmixing", mix[1], 0.001, 0.0, 0.0)
        y = G.Variable("ymixing", mix[2], 0.001, 0.0, 0.0)
    else:
        tau, x, y = G.Variable("tau", mix[0]), G.Variable("xmixing", mix[1]), G.Variable("ymixing", mix[2])
    dx, dy = G.Variable("deltax", mix[3]), G.Variable("deltay", mix[4])
    decay = G.DecayInfo3t(tau, x, y, dx, dy)
    decay.motherMass, decay.daug1Mass, decay.daug2Mass, decay.daug3Mass = DALITZ_DECAY[:4]
    decay.meson_radius, decay.mother_meson_radius = DALITZ_DECAY[4:]
    decay.resonances = res
    if resolution == "threegauss":
        vals = tg if tg is not None else (0.7, 0.2, 0.05, 1.1, -0.1, 2.0, 0.0, 6.0, 0.0, 0.01, 1.05, 0.0, 1.0)
        names = ("cf", "tf", "cb", "cs", "tb", "ts", "ob", "os", "sb", "cbo", "cws", "cbo2", "cws2")
        rv = [G.Variable(n, v) for n, v in zip(names, vals)]
        r = G.ThreeGaussResolution(*rv)
    else:
        rv = []
        r = G.TruthResolution()
    eff = G.PolynomialPdf("constantEff", [m12, m13], [G.Variable("constantOne", 1.0)],
                          [G.Variable("constantZero0", 0.0), G.Variable("constantZero1", 0.0)], 0)
    obs = [dtime, sigma, m12, m13, evt]
    mt = G.Observable("mistag", 0.0, 0.5) if mistag else None
    tag = G.Observable("charmtag", -1.0, 1.0) if charmtag else None
    if mt is not None:
        obs.append(mt)
        if tag is not None:
            obs.append(tag)
        td = G.Amp3Body_TD("td", dtime, sigma, m12, m13, evt, decay, r, eff, mt, tag)
    else:
        td = G.Amp3Body_TD("td", dtime, sigma, m12, m13, evt, decay, r, eff)
    params = dict(tau=tau, xmixing=x, ymixing=y)
    return Model(td, obs, params, keep=res + rv + [decay, eff, r])


def dalitz_dataset(model, m12, m13):
    import goofit_amd as G
    d = G.UnbinnedDataSet(model.observables)
    d.from_columns([np.asarray(m12, float), np.asarray(m13, float), np.arange(len(m12), dtype=np.float64)])
    return d


def generate_dalitz(model, n, seed=6):
    """n events of the model (current parameters) by accept-reject on the GPU
    (Amp3Body.generateEvents).  Returns (m12, m13)."""
    import goofit_amd as G
    pdf = model.pdf
    pdf.SigGenSetIndices()
    d = G.UnbinnedDataSet(model.observables)
    pdf.generateEvents(d, int(n), int(seed))
    return d.column(model.observables[0]), d.column(model.observables[1])
