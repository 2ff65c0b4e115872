"""Write the golden fixtures of tests/golden/ (run from the repository root:
``python tests/golden/make_golden.py``).

* reference_kats.json -- the known answers the reference's own tests hold for
  this path (values copied from the test sources named per entry; data only).
* oracle_vectors.json -- the CPU oracle's outputs on small seeded inputs of the
  three benchmark models (per-event values, NLLs, grid integrals), so that the
  GPU tests can compare against committed numbers without rebuilding the
  inputs' reference values, and any later change of the oracle shows up.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

import oracle as O  # noqa: E402
from goofit_amd import workloads as W  # noqa: E402

import _models as M  # noqa: E402

KATS = {
    "gauss_1d_grid": {
        "source": "tests/simple/NormalizeTest.cpp:56-100",
        "pdf": {"kind": "Gaussian", "mean": 1.5, "sigma": 2.5, "x": [-5.0, 10.0, 100]},
        "prob_at_bin": {"0": 0.00587129964482, "20": 0.0624318782242, "40": 0.157287605852, "60": 0.0938855055588},
        "bin_centre": {"0": -4.925, "20": -1.925, "40": 1.075, "60": 4.075},
        "normalize": 6.26657,
    },
    "gauss_eval": {
        "source": "tests/simple/NormalizeTest.cpp:176-219",
        "points": {"1.2": {"prob": 0.15843208471746245, "eval": 0.99282585857903813},
                   "7.2": {"prob": 0.0118618339389365, "eval": 0.0743330208507896}},
    },
    "exp_normalize": {"source": "tests/simple/NormalizeTest.cpp:15-52", "alpha": -2.0, "x": [0.0, 10.0],
                      "integral": 0.5},
    "grid_order": {"source": "tests/simple/SimpleTest.cpp:81-127", "x": [0.0, 10.0, 100], "y": [0.0, 10.0, 100],
                   "rows": {"0": [0.05, 0.05], "1": [0.15, 0.05], "100": [0.05, 0.15]}},
    "add_norm": {"source": "tests/simple/AddNormTest.cpp:16-44", "frac": 0.6, "sigmaA": 2.5, "sigmaB": 0.8,
                 "mean": 1.5, "x": [-5.0, 10.0, 100]},
    "tolerance": {"source": "Catch2 Approx default", "relative": 100 * float(np.finfo(np.float32).eps)},
}


def vectors():
    out = {"omp_chunks": 8, "note": "oracle outputs with the reduction split into 8 chunks (hazard A9)"}
    # 1-D AddPdf
    top, x = M.add1d()
    xs = W.sample_add1d(2000, seed=101)
    m = O.Model(top, xs[:, None], nchunks=8)
    out["add1d"] = {"events": xs.tolist(), "nll": m.calculate_nll(), "terms_first8": m.nll_terms()[:8].tolist(),
                    "params": m.op.p.tolist(), "norms": m.op.n.tolist()}
    # 3-D ProdPdf
    top, obs = M.prod3d()
    a, b, c = W.sample_prod3d(1500, seed=102)
    aos = np.stack([a, b, c], 1)
    m = O.Model(top, aos, nchunks=8)
    out["prod3d"] = {"events": aos.tolist(), "nll": m.calculate_nll(), "params": m.op.p.tolist(),
                     "norms": m.op.n.tolist()}
    # 16-resonance Dalitz, 120 x 120 grid
    top, obs = M.dalitz(nbins=120)
    aos = M.dalitz_events(1200, seed=103)
    m = O.Model(top, aos, decay=W.DALITZ_DECAY, nchunks=8)
    nll = m.calculate_nll()
    I = m.integrals[top.function_idx]
    out["dalitz16"] = {"events": aos.tolist(), "nll": nll, "params": m.op.p.tolist(), "norms": m.op.n.tolist(),
                       "integrals_re": I.real.tolist(), "integrals_im": I.imag.tolist(), "grid": 120}
    return out


def main():
    with open(os.path.join(HERE, "reference_kats.json"), "w") as f:
        json.dump(KATS, f, indent=1)
    with open(os.path.join(HERE, "oracle_vectors.json"), "w") as f:
        json.dump(vectors(), f)
    print("wrote", os.listdir(HERE))


if __name__ == "__main__":
    main()
