"""The drop-in C++ API from a C++ program: tests/cpp/goofit_client.cpp is compiled with
g++ against include/ and linked to libgoofit.so (tests/cpp/Makefile, run by the
build), builds the models the way the reference's examples do
(examples/dalitz/dalitz.cpp:281-335) and must reproduce the golden-fixture NLLs to
1e-11 and the fits of the Python binding."""
import json
import os
import subprocess

import numpy as np
import pytest

import goofit_amd as G
from goofit_amd import workloads as W

HERE = os.path.dirname(os.path.abspath(__file__))
CLIENT = os.path.join(HERE, "cpp", "goofit_client")
GOLDEN = os.path.join(HERE, "golden", "oracle_vectors.json")

pytestmark = pytest.mark.gpu


def _run(args):
    assert os.path.exists(CLIENT), "tests/cpp/goofit_client is not built (run the build first)"
    r = subprocess.run([CLIENT, *args], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = {"param": {}}
    for line in r.stdout.splitlines():
        key, *rest = line.split()
        if key == "param":
            out["param"][rest[0]] = (float(rest[1]), float(rest[2]))
        else:
            out[key] = float(rest[0])
    return out


def _write(path, rows):
    np.savetxt(path, np.asarray(rows, float), fmt="%.17g")
    return str(path)


def test_client_add1d_golden_nll(tmp_path):
    g = json.load(open(GOLDEN))["add1d"]
    out = _run(["add1d", _write(tmp_path / "ev.txt", np.asarray(g["events"])[:, None])])
    assert out["nll"] == pytest.approx(g["nll"], rel=1e-11)


def _model_file(path, table, coups):
    with open(path, "w") as f:
        for (name, kind, mass, width, spin, pair), c in zip(table, coups):
            # plain Python floats: numpy 2 scalars print as "np.float64(...)"
            f.write(f"{name} {kind} {float(mass)!r} {float(width)!r} {spin} {pair} {float(c.real)!r} "
                    f"{float(c.imag)!r}\n")
    return str(path)


def test_client_dalitz16_golden_nll(tmp_path):
    g = json.load(open(GOLDEN))["dalitz16"]
    ev = np.asarray(g["events"])
    assert np.array_equal(ev[:, 2], np.arange(len(ev)))  # event number = row, as the client numbers them
    model = _model_file(tmp_path / "model.txt", W.DALITZ_RESONANCES, W.dalitz_couplings(3, 16))
    out = _run(["dalitz", model, _write(tmp_path / "ev.txt", ev[:, :2]), str(g["grid"])])
    assert out["nll"] == pytest.approx(g["nll"], rel=1e-11)


def test_client_add1d_fit_matches_binding(tmp_path):
    """The same MIGRAD through the C++ program and through the Python binding of the same library:
    identical function minimum and parameters."""
    xs = W.sample_add1d(20_000, seed=41)
    out = _run(["add1d", _write(tmp_path / "ev.txt", xs[:, None]), "fit"])
    assert out["valid"] == 1.0
    m = W.build_add1d()
    d = G.UnbinnedDataSet(m.observables)
    d.from_columns([xs])
    m.pdf.setData(d)
    fm = G.FitManager(m.pdf)
    fm.setVerbosity(0)
    res = fm.fit()
    assert res.IsValid()
    assert out["fval"] == pytest.approx(res.Fval(), rel=1e-12)
    for cname, pname in (("mean", "mu"), ("sigma", "sigma"), ("alpha", "alpha"), ("frac", "frac")):
        v, e = out["param"][cname]
        assert v == pytest.approx(m.params[pname].value, abs=1e-9 * max(1.0, abs(v)))
        assert e == pytest.approx(m.params[pname].error, rel=1e-6)


def test_client_dalitz_fit_converges(tmp_path):
    """A 5-resonance Dalitz fit from the C++ program: valid, and its minimum is the binding's."""
    table = W.DALITZ_RESONANCES[:5]
    mdl = W.build_dalitz(nbins=150, resonances=table)
    g12, g13 = W.generate_dalitz(mdl, 30_000, seed=61)
    # the client fixes the first coupling of its table: put the fixed rho+ first
    order = [1, 0, 2, 3, 4]
    coups = W.dalitz_couplings(3, 5)
    model = _model_file(tmp_path / "model.txt", [table[i] for i in order], [coups[i] for i in order])
    out = _run(["dalitz", model, _write(tmp_path / "ev.txt", np.stack([g12, g13], 1)), "150", "fit"])
    assert out["valid"] == 1.0
    mdl.pdf.setData(W.dalitz_dataset(mdl, g12, g13))
    mdl.pdf.setDataSize(len(g12))
    fm = G.FitManager(mdl.pdf)
    fm.setVerbosity(0)
    res = fm.fit()
    assert res.IsValid()
    assert out["fval"] == pytest.approx(res.Fval(), abs=1e-4)
    for name in ("nonr", "rho0", "rhom", "rhop_1450"):
        for part, suffix in (("re", "real"), ("im", "imag")):
            v, e = out["param"][f"{name}_amp_{suffix}"]
            p = mdl.params[f"{name}_{part}"]
            assert abs(v - p.value) < 0.01 * p.error, (name, part, v, p.value, p.error)
