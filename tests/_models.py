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
