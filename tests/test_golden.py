"""The committed golden fixtures (tests/golden/, written by make_golden.py):
the oracle reproduces its recorded vectors, and the reference's known answers
hold for the oracle."""
import json
import math
import os

import numpy as np
import pytest

import oracle as O
from goofit_amd import workloads as W

import _models as M

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with open(os.path.join(GOLD, name)) as f:
        return json.load(f)


def test_oracle_reproduces_its_vectors():
    v = load("oracle_vectors.json")
    P = v["omp_chunks"]
    top, x = M.add1d()
    xs = np.array(v["add1d"]["events"])
    m = O.Model(top, xs[:, None], nchunks=P)
    assert m.calculate_nll() == v["add1d"]["nll"]
    np.testing.assert_array_equal(m.nll_terms()[:8], v["add1d"]["terms_first8"])
    top, obs = M.prod3d()
    m = O.Model(top, np.array(v["prod3d"]["events"]), nchunks=P)
    assert m.calculate_nll() == v["prod3d"]["nll"]
    top, obs = M.dalitz(nbins=v["dalitz16"]["grid"])
    m = O.Model(top, np.array(v["dalitz16"]["events"]), decay=W.DALITZ_DECAY, nchunks=P)
    assert m.calculate_nll() == v["dalitz16"]["nll"]
    I = m.integrals[top.function_idx]
    np.testing.assert_array_equal(I.real, v["dalitz16"]["integrals_re"])


def test_reference_kats_hold_for_the_oracle():
    k = load("reference_kats.json")
    tol = k["tolerance"]["relative"]
    g = k["gauss_1d_grid"]
    xo = O.Obs("x", *g["pdf"]["x"])
    pdf = O.Gaussian("g", xo, O.Var("mean", g["pdf"]["mean"]), O.Var("sigma", g["pdf"]["sigma"]))
    model = O.Model(pdf)
    norm = model.normalize()
    assert abs(norm - g["normalize"]) <= tol * g["normalize"]
    raw = model.op.eval_bins(pdf.function_idx, g["pdf"]["x"])
    for i, val in g["prob_at_bin"].items():
        assert abs(raw[int(i)] / norm - val) <= tol * val
    e = k["gauss_eval"]
    model.op.sync_norms()
    for xv, d in e["points"].items():
        ev = model.op.eval_at(pdf.function_idx, [float(xv)])
        pr = model.op.eval_metric(np.array([[float(xv)]]), "PROB")[0]
        assert abs(ev - d["eval"]) <= tol * d["eval"] and abs(pr - d["prob"]) <= tol * d["prob"]
    x = k["exp_normalize"]
    xo = O.Obs("x", *x["x"])
    ep = O.Exp("e", xo, O.Var("alpha", x["alpha"]))
    assert abs(O.Model(ep).normalize() - x["integral"]) <= tol * x["integral"]
