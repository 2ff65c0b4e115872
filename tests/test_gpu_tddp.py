"""Time-dependent Dalitz (Amp3Body_TD) on the GPU: wave caches, ThreeComplex
integrals, likelihood parity with the oracle through the C ABI and the C++ API,
and a fit of the lifetime and mixing parameters.

Tolerances as for the time-integrated Dalitz PDF: waves within 1e-12, integrals
within 1e-11 relative to their largest entry, the NLL within 1e-11 relative.
"""
import math

import numpy as np
import pytest

import oracle as O
from goofit_amd import workloads as W

from _models import arrays, threegauss, tddp, tddp_events

pytestmark = pytest.mark.gpu

NLL_RTOL = 1e-11


def _install(ctx, top, aos, td=None):
    prog, (f, p, c, o, n) = arrays(top)
    ctx.set_program(f, p, c, o, n)
    ctx.set_decay(W.DALITZ_DECAY)
    ctx.upload_columns([np.ascontiguousarray(aos[:, k]) for k in range(aos.shape[1])])
    ctx.set_identity_column(4)
    ctx.dalitz_set_data_size((td or top).function_idx, len(aos))
    return prog


def _sync(ctx, model):
    model.op.sync_params()
    model.op.sync_norms()
    ctx.update_params(model.op.p)
    ctx.update_norms(model.op.n)


def test_tddp_waves(ctx):
    n = 6000
    top, _ = tddp()
    aos = tddp_events(n, seed=3)
    M = O.Model(top, aos, decay=W.DALITZ_DECAY)
    _install(ctx, top, aos)
    M.op.sync_params()
    ctx.update_params(M.op.p)
    ctx.dalitz_cache(top.function_idx)
    res_funcs = M.op.tddp_layout(top.function_idx)[0]
    for i, rf in enumerate(res_funcs):
        ra, rb = M.op.tddp_wave(top.function_idx, rf, aos)
        ga = ctx.dalitz_get_cache(top.function_idx, i, n)
        gb = ctx.dalitz_get_cache(top.function_idx, len(res_funcs) + i, n)
        np.testing.assert_allclose(ga, ra, rtol=1e-12, atol=1e-14 * np.abs(ra).max())
        np.testing.assert_allclose(gb, rb, rtol=1e-12, atol=1e-14 * np.abs(rb).max())


@pytest.mark.parametrize("nres", [1, 6, 8, 11, 16])
def test_tddp_integrals(ctx, nres):
    """One stacked Gram up to 8 resonances; pairs of 8-wave groups above."""
    top, _ = tddp(resonances=W.DALITZ_RESONANCES[:nres], nbins=90)
    aos = tddp_events(64, seed=4)
    M = O.Model(top, aos, decay=W.DALITZ_DECAY)
    _install(ctx, top, aos)
    M.op.sync_params()
    ctx.update_params(M.op.p)
    grid6 = [0.0, 3.0, 90, 0.0, 3.0, 70]  # unequal bin counts: the m13 count is read from its own entry
    ref = M.op.tddp_integrals(top.function_idx, grid6)
    got = ctx.tddp_gram(top.function_idx, grid6, nres)
    for g, r in zip(got, ref):
        scale = np.abs(r).max()
        np.testing.assert_allclose(g, r, rtol=1e-11, atol=1e-12 * scale)
    aa, ab, bb = got
    np.testing.assert_allclose(aa, aa.conj().T, rtol=0, atol=1e-13 * np.abs(aa).max())
    np.testing.assert_allclose(bb, bb.conj().T, rtol=0, atol=1e-13 * np.abs(bb).max())


@pytest.mark.parametrize("kind,tags", [("truth", False), ("truth", True), ("threegauss", False),
                                       ("threegauss", True)])
def test_tddp_nll(ctx, kind, tags):
    n = 30_000
    mix = (0.4101, 0.012, 0.02, 0.004, -0.003)
    res = threegauss() if kind == "threegauss" else O.TruthResolution()
    top, obs = tddp(mix=mix, resolution=res, mistag=tags, charmtag=tags, nbins=150)
    rng = np.random.default_rng(17)
    kw = dict(mistag=rng.uniform(0.0, 0.3, n), tag=rng.choice([-1.0, 1.0, 0.0], n)) if tags else {}
    aos = tddp_events(n, seed=8, **kw)
    if tags:
        obs[5].value = 0.1  # host value of the mistag observable (normalisation's mistag terms)
    M = O.Model(top, aos, decay=W.DALITZ_DECAY)
    ref = M.calculate_nll()
    _install(ctx, top, aos)
    M.op.sync_params()
    ctx.update_params(M.op.p)
    ctx.dalitz_cache(top.function_idx)
    _sync(ctx, M)
    got = ctx.reduce_metric("NLL")[0]
    assert math.isclose(got, ref, rel_tol=NLL_RTOL), (got, ref)
    # per-event values
    got_e = ctx.eval_metric("EVAL")
    ref_e = M.op.eval_metric(aos, "EVAL")
    np.testing.assert_allclose(got_e, ref_e, rtol=1e-12, atol=1e-14 * np.abs(ref_e).max())


def test_tddp_in_addpdf(ctx):
    """Signal Amp3Body_TD plus a flat Dalitz background (MultiPolynomial over m12, m13)."""
    n = 20_000
    top_td, obs = tddp(nbins=100)
    m12o, m13o = obs[2], obs[3]
    bkg = O.MultiPolynomial("bkg", [m12o, m13o], [O.Var("b0", 1.0), O.Var("b1", 0.2), O.Var("b2", 0.1)],
                            [O.Var("o0", 0.0), O.Var("o1", 0.0)], 1)
    top = O.Add("sum", [O.Var("fsig", 0.8)], [top_td, bkg])
    top.observables = list(top_td.observables)
    aos = tddp_events(n, seed=12)
    M = O.Model(top, aos, decay=W.DALITZ_DECAY)
    ref = M.calculate_nll()
    _install(ctx, top, aos, td=top_td)
    M.op.sync_params()
    ctx.update_params(M.op.p)
    ctx.dalitz_cache(top_td.function_idx)
    _sync(ctx, M)
    got = ctx.reduce_metric("NLL")[0]
    assert math.isclose(got, ref, rel_tol=NLL_RTOL), (got, ref)


# ------------------------------------------------------------ C++ API

def _api_events(model, n, seed=5, tau=0.4101, with_tags=False):
    import goofit_amd as G
    rng = np.random.default_rng(seed)
    m12, m13 = W.dalitz_uniform(n, seed)
    cols = [rng.exponential(tau, n), np.full(n, 0.05), m12, m13, np.arange(n, dtype=np.float64)]
    if with_tags:
        cols += [rng.uniform(0, 0.3, n), rng.choice([-1.0, 1.0], n)]
    d = G.UnbinnedDataSet(model.observables)
    d.from_columns(cols)
    return d, np.stack(cols, axis=1)


@pytest.mark.parametrize("kind", ["truth", "threegauss"])
def test_api_tddp_nll_parity(kind):
    import goofit_amd as G
    n = 20_000
    mix = (0.4101, 0.01, 0.015, 0.003, 0.002)
    model = W.build_tddp(nbins=120, mix=mix, resolution=kind, mistag=True, charmtag=True)
    d, aos = _api_events(model, n, with_tags=True)
    model.observables[5].value = 0.1
    pdf = model.pdf
    pdf.setData(d)
    pdf.setDataSize(n, 7)
    got = pdf.calculateNLL()
    res = threegauss() if kind == "threegauss" else O.TruthResolution()
    top, obs = tddp(nbins=120, mix=mix, resolution=res, mistag=True, charmtag=True)
    obs[5].value = 0.1
    ref = O.Model(top, aos, decay=W.DALITZ_DECAY).calculate_nll()
    assert math.isclose(got, ref, rel_tol=NLL_RTOL), (got, ref)
    # ThreeComplex integrals exposed by the API agree with the oracle's
    op = O.OracleProg(top, W.DALITZ_DECAY)
    op.sync_params()
    aa, ab, bb = op.tddp_integrals(0, [0.0, 3.0, 120, 0.0, 3.0, 120])
    np.testing.assert_allclose(np.array(pdf.getIntegralsAB()).reshape(aa.shape), ab, rtol=1e-11,
                               atol=1e-12 * np.abs(ab).max())
    fr = pdf.getFractions()
    assert len(fr) == aa.shape[0]


def test_api_tddp_fit_lifetime_and_mixing():
    """Events of the x = y = 0 model (time-integrated Dalitz density times an exponential
    in t, generated on the GPU) fitted with tau, x, y free: tau within 3 sigma of the
    truth, x and y within 3 sigma of 0."""
    import goofit_amd as G
    n = 60_000
    tau0 = 0.4101
    table = W.DALITZ_RESONANCES[:6]
    dz = W.build_dalitz(nbins=200, resonances=table, float_couplings=False)
    m12, m13 = W.generate_dalitz(dz, n, seed=21)
    model = W.build_tddp(nbins=200, resonances=table, mix=(0.42, 0.002, -0.002, 0.0, 0.0))
    rng = np.random.default_rng(22)
    cols = [rng.exponential(tau0, n), np.full(n, 0.05), np.asarray(m12), np.asarray(m13),
            np.arange(n, dtype=np.float64)]
    d = G.UnbinnedDataSet(model.observables)
    d.from_columns(cols)
    model.pdf.setData(d)
    model.pdf.setDataSize(n, 5)
    fm = G.FitManager(model.pdf)
    fm.verbosity = 0
    r = fm.fit()
    assert r.IsValid()
    tau, x, y = model.params["tau"], model.params["xmixing"], model.params["ymixing"]
    assert abs(tau.value - tau0) < 3 * tau.error, (tau.value, tau.error)
    assert abs(x.value) < 3 * x.error, (x.value, x.error)
    assert abs(y.value) < 3 * y.error, (y.value, y.error)
    assert tau.error < 0.01


def test_batch_fitter_tddp_toys_match_sequential_fits():
    """Amp3Body_TD toys over disjoint rows of one data set, fitted together by
    BatchFitter (tau, x, y free; the host computes each toy's decay-time integrals),
    against one FitManager fit per toy from the same start: every value within 0.01
    sigma."""
    import goofit_amd as G
    counts = [6000, 9000, 7000, 8000]
    n = sum(counts)
    table = W.DALITZ_RESONANCES[:6]
    dz = W.build_dalitz(nbins=120, resonances=table, float_couplings=False)
    m12, m13 = W.generate_dalitz(dz, n, seed=41)
    model = W.build_tddp(nbins=120, resonances=table, mix=(0.42, 0.002, -0.002, 0.0, 0.0))
    start = {k: v.value for k, v in model.params.items()}
    rng = np.random.default_rng(42)
    t_all, m12, m13 = rng.exponential(0.4101, n), np.asarray(m12), np.asarray(m13)

    def dataset(lo, hi):
        d = G.UnbinnedDataSet(model.observables)
        d.from_columns([t_all[lo:hi], np.full(hi - lo, 0.05), m12[lo:hi], m13[lo:hi],
                        np.arange(hi - lo, dtype=np.float64)])
        return d

    model.pdf.setData(dataset(0, n))
    model.pdf.setDataSize(n, 5)
    first = [int(v) for v in np.concatenate([[0], np.cumsum(counts)[:-1]])]
    bf = G.BatchFitter(model.pdf, first, counts)
    bf.fit(threads=4)
    vals, valid = bf.values(), bf.valid()
    for t in range(len(counts)):
        for k, v in start.items():
            model.params[k].value = v
        model.pdf.setData(dataset(first[t], first[t] + counts[t]))
        model.pdf.setDataSize(counts[t], 5)
        fm = G.FitManager(model.pdf)
        fm.verbosity = 0
        r = fm.fit()
        assert valid[t] and r.IsValid(), t
        sv, se = np.array(r.values()), np.array(r.errors())
        free = se > 0
        assert free.sum() == 3, se
        shift = np.abs(np.array(vals[t]) - sv)[free] / se[free]
        assert shift.max() < 0.01, (t, shift)
