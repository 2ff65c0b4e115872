"""Test helpers: the benchmark models written with the oracle's Python model
classes, so that the same flattened arrays feed the oracle and the engine."""
import numpy as np

import oracle as O
from goofit_amd import workloads as W


def add1d(mu=0.3, sigma=0.8, alpha=-0.4, frac=0.3, lo=-5.0, hi=5.0, nbins=1000, extended=False):
    x = O.Obs("x", lo, hi, nbins)
    g = O.Gaussian("gauss", x, O.Var("mu", mu), O.Var("sigma", sigma))
    e = O.Exp("exp", x, O.Var("alpha", alpha))
    if extended:
        top = O.Add("sum", [O.Var("ns", frac), O.Var("nb", 1 - frac)], [g, e])
    else:
        top = O.Add("sum", [O.Var("frac", frac)], [g, e])
    return top, x


def prod3d(mu=0.0, sigma=1.0, alpha=-1.2, emg_mu=0.0, emg_sigma=0.1, rate=1 / 1.5, nbins=200):
    x = O.Obs("x", -6.0, 6.0, nbins)
    y = O.Obs("y", 0.0, 8.0, nbins)
    z = O.Obs("z", 0.0, 10.0, nbins)
    g = O.Gaussian("gx", x, O.Var("mu", mu), O.Var("sigma", sigma))
    e = O.Exp("ey", y, O.Var("alpha", alpha))
    eg = O.ExpGaus("egz", z, O.Var("emu", emg_mu), O.Var("esig", emg_sigma), O.Var("rate", rate))
    return O.Prod("prod", [g, e, eg]), (x, y, z)


def dalitz(coup_seed=3, nbins=240, resonances=None, eff=None):
    m12 = O.Obs("m12", 0.0, 3.0, nbins)
    m13 = O.Obs("m13", 0.0, 3.0, nbins)
    num = O.Obs("evt", 0.0, 1e9, 1, event_number=True)
    table = resonances if resonances is not None else W.DALITZ_RESONANCES
    coups = W.dalitz_couplings(coup_seed, len(table))
    res = []
    for (name, kind, mass, width, spin, pair), c in zip(table, coups):
        ar, ai = O.Var(name + "_re", c.real), O.Var(name + "_im", c.imag)
        if kind == "NONRES":
            res.append(O.NonRes(name, ar, ai))
        elif kind == "GS":
            res.append(O.GS(name, ar, ai, O.Var(name + "_m", mass), O.Var(name + "_w", width), spin, pair))
        elif kind == "RBW_SYM":
            res.append(O.RBW(name, ar, ai, O.Var(name + "_m", mass), O.Var(name + "_w", width), spin, pair,
                             sym=True))
        else:
            res.append(O.RBW(name, ar, ai, O.Var(name + "_m", mass), O.Var(name + "_w", width), spin, pair))
    if eff is None:
        eff = O.MultiPolynomial("eff", [m12, m13], [O.Var("one", 1.0)], [O.Var("z0", 0.0), O.Var("z1", 0.0)], 0)
    top = O.Dalitz("dalitz", m12, m13, num, W.DALITZ_DECAY, res, eff)
    return top, (m12, m13, num)


def dalitz_events(n, seed=5, outside=0.1):
    """n events: uniform in the Dalitz plot, a fraction outside it, event number = row."""
    m12, m13 = W.dalitz_uniform(n, seed)
    rng = np.random.default_rng(seed + 100)
    k = int(outside * n)
    if k:
        idx = rng.choice(n, size=k, replace=False)
        m12[idx] = rng.uniform(0.0, 3.0, size=k)
        m13[idx] = rng.uniform(0.0, 3.0, size=k)
    return np.stack([m12, m13, np.arange(n, dtype=np.float64)], axis=1)


def arrays(top):
    prog = O.flatten(top)
    return prog, prog.arrays()


TD_MIX = (0.4101, 0.0016, 0.0055, 0.0, 0.0)  # tau, x, y, dx, dy (examples/pipipi0DPFit/pipipi0DPFit.cpp:526-530)


def threegauss(cf=0.7, tf=0.2, cb=0.05, cs=1.1, tb=-0.1, ts=2.0, ob=0.0, os_=6.0, sb=0.0, cbo=0.01, cws=1.05,
               cbo2=0.0, cws2=1.0):
    v = [O.Var(n, x) for n, x in (("cf", cf), ("tf", tf), ("cb", cb), ("cs", cs), ("tb", tb), ("ts", ts),
                                  ("ob", ob), ("os", os_), ("sb", sb), ("cbo", cbo), ("cws", cws),
                                  ("cbo2", cbo2), ("cws2", cws2))]
    return O.ThreeGaussResolution("tg", *v)


def tddp(coup_seed=3, nbins=120, resonances=None, mix=TD_MIX, resolution=None, mistag=False, charmtag=False,
         eff=None):
    """Amp3Body_TD over (dtime, sigma_t, m12, m13, evt[, mistag][, tag]) with the config-2 resonances."""
    dtime = O.Obs("dtime", -2.0, 20.0, 100)
    sigma = O.Obs("sigmat", 0.0, 0.8, 100)
    m12 = O.Obs("m12", 0.0, 3.0, nbins)
    m13 = O.Obs("m13", 0.0, 3.0, nbins)
    num = O.Obs("evt", 0.0, 1e9, 1, event_number=True)
    mt = O.Obs("mistag", 0.0, 0.5, 10) if mistag else None
    tag = O.Obs("tag", -1.0, 1.0, 2) if charmtag else None
    table = resonances if resonances is not None else W.DALITZ_RESONANCES[:6]
    base, _ = dalitz(coup_seed=coup_seed, nbins=nbins, resonances=table)
    res = base.resonances
    if eff is None:
        eff = O.MultiPolynomial("eff", [m12, m13], [O.Var("one", 1.0)], [O.Var("z0", 0.0), O.Var("z1", 0.0)], 0)
    mixv = [O.Var(n, v) for n, v in zip(("tau", "xmix", "ymix", "dx", "dy"), mix)]
    resolution = resolution if resolution is not None else O.TruthResolution()
    top = O.Tddp("td", dtime, sigma, m12, m13, num, W.DALITZ_DECAY, mixv, res, resolution, eff, mistag=mt,
                 charmtag=tag)
    return top, (dtime, sigma, m12, m13, num, mt, tag)


def tddp_events(n, seed=7, tau=0.4101, sigma=0.05, mistag=None, tag=None, outside=0.05):
    """AoS rows (dtime, sigma_t, m12, m13, evt[, mistag][, tag]): Dalitz points uniform,
    decay times exponential smeared by sigma_t, event number = row."""
    d = dalitz_events(n, seed=seed, outside=outside)
    rng = np.random.default_rng(seed + 7)
    t = rng.exponential(tau, size=n) + sigma * rng.standard_normal(n)
    cols = [t, np.full(n, sigma), d[:, 0], d[:, 1], d[:, 2]]
    if mistag is not None:
        cols.append(np.full(n, mistag) if np.isscalar(mistag) else np.asarray(mistag, np.float64))
    if tag is not None:
        cols.append(np.asarray(tag, np.float64) if not np.isscalar(tag) else np.full(n, float(tag)))
    return np.stack(cols, axis=1)
