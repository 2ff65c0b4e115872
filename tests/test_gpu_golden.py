"""The engine against the committed oracle vectors (tests/golden/oracle_vectors.json)."""
import json
import math
import os

import numpy as np
import pytest

import oracle as O
from goofit_amd import workloads as W

import _models as M

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "oracle_vectors.json")


def test_engine_matches_golden_vectors(ctx):
    v = json.load(open(GOLD))
    for name, top, cols in (("add1d", M.add1d()[0], 1), ("prod3d", M.prod3d()[0], 3)):
        f, p, c, o, n = O.flatten(top).arrays()
        ctx.set_program(f, v[name]["params"], c, o, v[name]["norms"])
        ev = np.array(v[name]["events"]).reshape(-1, cols)
        ctx.upload_columns([ev[:, k] for k in range(cols)])
        got = ctx.reduce_metric("NLL")[0]
        assert math.isclose(got, v[name]["nll"], rel_tol=1e-11), (name, got)
    d = v["dalitz16"]
    top, _ = M.dalitz(nbins=d["grid"])
    f, p, c, o, n = O.flatten(top).arrays()
    ctx.set_program(f, d["params"], c, o, d["norms"])
    ctx.set_decay(W.DALITZ_DECAY)
    ev = np.array(d["events"])
    ctx.upload_columns([ev[:, 0], ev[:, 1], ev[:, 2]])
    ctx.set_identity_column(2)
    ctx.dalitz_set_data_size(top.function_idx, ev.shape[0])
    ctx.dalitz_cache(top.function_idx)
    got = ctx.reduce_metric("NLL")[0]
    assert math.isclose(got, d["nll"], rel_tol=1e-11), got
    res, eff = O.OracleProg(top, W.DALITZ_DECAY).dalitz_layout(0)
    g = d["grid"]
    I = ctx.dalitz_gram(top.function_idx, [0.0, 3.0, g, 0.0, 3.0, g], eff, len(res))
    ref = np.array(d["integrals_re"]) + 1j * np.array(d["integrals_im"])
    np.testing.assert_allclose(I, ref, rtol=1e-11, atol=1e-13 * np.abs(ref).max())
