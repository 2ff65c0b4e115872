"""Toy batches on the GPU: per-toy NLLs of one batched launch against the oracle
(captured graph and plain launches, several replays with new parameters in
between), the autotuner, and batched MIGRAD fits against sequential ones."""
import math

import numpy as np
import pytest

import goofit_amd as G
import oracle as O
from goofit_amd import capi
from goofit_amd import workloads as W

from _models import add1d

pytestmark = pytest.mark.gpu


def _toys(counts, seed):
    return [W.sample_add1d(int(n), seed=seed + t) for t, n in enumerate(counts)]


def test_batch_reduce_matches_oracle(ctx):
    counts = [3000, 0, 777, 12_345, 64, 5000]
    data = _toys(counts, 40)
    first = np.concatenate([[0], np.cumsum(counts)[:-1]])
    top, x = add1d()
    prog = O.flatten(top)
    f, p, c, o, n = prog.arrays()
    ctx.set_program(f, p, c, o, n)
    ctx.upload_columns([np.concatenate(data)])
    b = capi.Batch(ctx, first, counts)
    rng = np.random.default_rng(3)
    try:
        for graph in (True, False):
            b.use_graph(graph)
            for rep in range(4):
                refs = []
                for t in range(len(counts)):
                    # a different parameter point per toy and per replay
                    mu, sig, alpha, frac = 0.3 + 0.1 * rng.normal(), 0.8 * (1 + 0.1 * rng.random()), -0.4, 0.3
                    tt, _ = add1d(mu=mu, sigma=sig, alpha=alpha, frac=frac)
                    m = O.Model(tt, data[t].reshape(-1, 1) if counts[t] else np.zeros((1, 1)))
                    m.normalize()
                    m.op.sync_params()
                    m.op.sync_norms()
                    b.set_toy(t, m.op.p, m.op.n)
                    refs.append(m.op.reduce_metric(data[t].reshape(-1, 1), "NLL") if counts[t] else 0.0)
                active = np.array([1, 1, 1, 1, 1, 1], np.uint8)
                if rep == 2:
                    active[3] = 0
                got = b.reduce(active)
                for t in range(len(counts)):
                    if not active[t]:
                        assert np.isnan(got[t])
                    elif counts[t] == 0:
                        assert got[t] == 0.0
                    else:
                        assert math.isclose(got[t], refs[t], rel_tol=1e-11), (graph, rep, t, got[t], refs[t])
                # the plain single-data-set path in between (eager work between replays)
                ctx.reduce_metric("NLL")
        chunk, us = b.autotune()
        assert chunk in (1024, 2048, 4096, 8192, 16384, 32768) and us > 0
    finally:
        b.close()


def test_batch_fitter_matches_sequential_fits():
    counts = [4000, 9000, 6000, 15_000, 7000, 5000, 12_000, 8000]
    data = _toys(counts, 70)
    m = W.build_add1d(start=1.1)
    start = {k: v.value for k, v in m.params.items()}
    d = G.UnbinnedDataSet(m.observables)
    d.from_columns([np.concatenate(data)])
    m.pdf.setData(d)
    first = list(np.concatenate([[0], np.cumsum(counts)[:-1]]).astype(int))
    bf = G.BatchFitter(m.pdf, first, counts)
    bf.fit(threads=4)
    vals, errs, valid = bf.values(), bf.errors(), bf.valid()
    st = bf.stats()
    assert st["launches"] > 0 and st["fcn_calls"] >= st["launches"]
    # sequential: one FitManager per toy from the same start
    for t, x in enumerate(data):
        for k, v in start.items():
            m.params[k].value = v
        dt = G.UnbinnedDataSet(m.observables)
        dt.from_columns([x])
        m.pdf.setData(dt)
        fm = G.FitManager(m.pdf)
        fm.setVerbosity(0)
        r = fm.fit()
        assert valid[t] and r.IsValid()
        sv, se = np.array(r.values()), np.array(r.errors())
        free = se > 0
        shift = np.abs(np.array(vals[t]) - sv)[free] / se[free]
        assert shift.max() < 0.01, (t, shift)
