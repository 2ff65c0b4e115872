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
        assert chunk in (2048, 4096, 8192, 16384) and us > 0
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


def _add1d_point(rng):
    """Parameter and normalisation arrays of the AddPdf model at a random point, and the model."""
    mu, sig = 0.3 + 0.1 * rng.normal(), 0.8 * (1 + 0.1 * rng.random())
    tt, _ = add1d(mu=mu, sigma=sig, alpha=-0.4 + 0.05 * rng.normal(), frac=0.3 + 0.05 * rng.random())
    return tt


def test_batch_points_equal_single_point_launches(ctx):
    """gf_batch_reduce_points on the interpreter path: every point's sum equals
    gf_batch_reduce's for the same arrays bit for bit (0, 1, 3, 5, 17 and 64 points per
    toy: one to four point groups), the first points of each toy agree with the oracle,
    and toys without points keep their output entries."""
    counts = [3000, 0, 777, 12_345, 64, 5000]
    npts = [1, 3, 0, 17, 64, 5]
    data = _toys(counts, 41)
    first = np.concatenate([[0], np.cumsum(counts)[:-1]])
    top, _ = add1d()
    f, p, c, o, n = O.flatten(top).arrays()
    ctx.set_program(f, p, c, o, n)
    ctx.upload_columns([np.concatenate(data)])
    b = capi.Batch(ctx, first, counts)
    rng = np.random.default_rng(8)
    try:
        pts = {}
        for t in range(len(counts)):
            for k in range(npts[t]):
                m = O.Model(_add1d_point(rng), data[t].reshape(-1, 1) if counts[t] else np.zeros((1, 1)))
                m.normalize()
                m.op.sync_params()
                m.op.sync_norms()
                ref = None
                if k < 2 and counts[t]:
                    ref = m.op.reduce_metric(data[t].reshape(-1, 1), "NLL")
                pts[t, k] = (m.op.p.copy(), m.op.n.copy(), ref)
                b.add_point(t, m.op.p, m.op.n)
        got = b.reduce_points()
        assert got.shape == (len(counts), capi.BATCH_MAX_POINTS)
        for t in range(len(counts)):
            assert np.isnan(got[t, npts[t]:]).all()
        for (t, k), (_, _, ref) in pts.items():
            if counts[t] == 0:
                assert got[t, k] == 0.0
            elif ref is not None:
                assert math.isclose(got[t, k], ref, rel_tol=1e-11), (t, k, got[t, k], ref)
        b.use_graph(False)
        for k in range(max(npts)):
            active = np.array([k < m for m in npts], np.uint8)
            for t in range(len(counts)):
                if active[t]:
                    b.set_toy(t, pts[t, k][0], pts[t, k][1])
            single = b.reduce(active)
            for t in range(len(counts)):
                if active[t]:
                    assert got[t, k] == single[t], (t, k, got[t, k], single[t])
        # a second request after the first: the points were cleared
        b.add_point(3, pts[3, 5][0], pts[3, 5][1])
        again = b.reduce_points()
        assert again[3, 0] == got[3, 5] and np.isnan(again[0]).all()
    finally:
        b.close()


def test_batch_points_rejects_bad_requests(ctx):
    top, _ = add1d()
    f, p, c, o, n = O.flatten(top).arrays()
    ctx.set_program(f, p, c, o, n)
    ctx.upload_columns([W.sample_add1d(2000, seed=3)])
    b = capi.Batch(ctx, [0, 1000], [1000, 1000])
    try:
        b.add_point(0, p, n)
        b.add_point(1, p, n)
        with pytest.raises(RuntimeError, match="one after the other"):
            b.add_point(0, p, n)
        b.clear_points()
        for _ in range(capi.BATCH_MAX_POINTS):
            b.add_point(1, p, n)
        with pytest.raises(RuntimeError, match="already has"):
            b.add_point(1, p, n)
        got = b.reduce_points()
        assert np.isfinite(got[1]).all() and np.isnan(got[0]).all()
        assert np.all(got[1] == got[1, 0])
    finally:
        b.close()


def _dalitz_batch_setup(ctx, n, eff=None, masses=None, nbins=64, seed=5):
    from _models import arrays, dalitz, dalitz_events
    top, _ = dalitz(nbins=nbins, eff=eff)
    if masses is not None:  # the efficiency's observables are the Dalitz PDF's
        top.m12, top.m13 = masses
        top.observables[0], top.observables[1] = masses
    aos = dalitz_events(n, seed)
    M = O.Model(top, aos, decay=W.DALITZ_DECAY)
    _, (f, p, c, o, nn) = arrays(top)
    ctx.set_program(f, p, c, o, nn)
    ctx.set_decay(W.DALITZ_DECAY)
    ctx.upload_columns([aos[:, 0], aos[:, 1], aos[:, 2]])
    ctx.set_identity_column(2)
    ctx.dalitz_set_data_size(top.function_idx, n)
    ref = M.calculate_nll()
    M.op.sync_params()
    M.op.sync_norms()
    ctx.update_params(M.op.p)
    ctx.update_norms(M.op.n)
    ctx.dalitz_cache(top.function_idx)
    return top, M, ref


def _coupling_sets(top, M, k, rng):
    sets = []
    for _ in range(k):
        p = M.op.p.copy()
        for i in range(len(top.resonances)):
            p[top.parameters_idx + 1 + 2 * i] *= 1 + 1e-3 * rng.standard_normal()
            p[top.parameters_idx + 2 + 2 * i] *= 1 + 1e-3 * rng.standard_normal()
        nn = M.op.n.copy()
        nn[top.normal_idx + 1] *= 1 + 1e-4 * rng.standard_normal()
        sets.append((p, nn))
    return sets


@pytest.mark.parametrize("const_eff", [True, False])
def test_batch_points_dalitz(ctx, const_eff):
    """The coherent-sum form of gf_batch_reduce_points (waves read once per 16 points):
    a toy over all rows reproduces the oracle's NLL at its first point; every point agrees
    with the interpreter's gf_batch_reduce within 1e-12 relative; with a constant
    efficiency the vector kernel that skips m12/m13 gives the same bits as the one that
    reads them, and the matrix-core groups agree with it within 1e-13."""
    n = 40_000
    eff = masses = None
    if not const_eff:
        masses = (O.Obs("m12", 0.0, 3.0, 64), O.Obs("m13", 0.0, 3.0, 64))
        eff = O.MultiPolynomial("eff", list(masses), [O.Var("k0", 0.8), O.Var("k1", 0.1), O.Var("k2", -0.05)],
                                [O.Var("o0", 0.5), O.Var("o1", 0.2)], 1)
    top, M, ref = _dalitz_batch_setup(ctx, n, eff=eff, masses=masses)
    first, counts = [0, 0, 10_000, 33_333], [n, 10_000, 23_333, n - 33_333]
    npts = [20, 1, 16, 40]
    rng = np.random.default_rng(12)
    sets = {t: _coupling_sets(top, M, npts[t], rng) for t in range(4)}
    sets[0][0] = (M.op.p.copy(), M.op.n.copy())
    b = capi.Batch(ctx, first, counts)
    try:
        def request():
            for t in range(4):
                for p, nn in sets[t]:
                    b.add_point(t, p, nn)
            return b.reduce_points()
        got = request()
        assert math.isclose(got[0, 0], ref, rel_tol=1e-11), (got[0, 0], ref)
        b.use_graph(False)
        for k in range(max(npts)):
            active = np.array([k < m for m in npts], np.uint8)
            for t in range(4):
                if active[t]:
                    b.set_toy(t, *sets[t][k])
            single = b.reduce(active)
            for t in range(4):
                if active[t]:
                    assert math.isclose(got[t, k], single[t], rel_tol=1e-12), (t, k, got[t, k], single[t])
        if const_eff:
            # with 16 resonances groups of 4 or more points sum on the matrix cores by default
            # (TUNE_BATCH_MX_PTS 0): against the vector kernel for every group (17: never),
            # and with every group of two or more points there (2)
            assert ctx.get_tuning(capi.TUNE_BATCH_MX_PTS) == 0
            ctx.set_tuning(capi.TUNE_BATCH_MX_PTS, 17)
            try:
                valu = request()
                ctx.set_tuning(capi.TUNE_ZERO_OUTSIDE, 0)
                plain = request()
            finally:
                ctx.set_tuning(capi.TUNE_ZERO_OUTSIDE, 1)
                ctx.set_tuning(capi.TUNE_BATCH_MX_PTS, 0)
            np.testing.assert_array_equal(plain, valu)
            np.testing.assert_allclose(got, valu, rtol=1e-13)
            assert not np.array_equal(got, valu, equal_nan=True)  # the matrix-core groups ran
            assert got[1, 0] == valu[1, 0]  # a lone point stays on the vector ALUs
            ctx.set_tuning(capi.TUNE_BATCH_MX_PTS, 2)
            try:
                mx2 = request()
            finally:
                ctx.set_tuning(capi.TUNE_BATCH_MX_PTS, 0)
            np.testing.assert_allclose(mx2, valu, rtol=1e-13)
            assert mx2[1, 0] == valu[1, 0]
            np.testing.assert_array_equal(mx2[2, :16], got[2, :16])  # full groups alike
            np.testing.assert_array_equal(mx2[0, 16:20], got[0, 16:20])  # a group of 4 in both
            with pytest.raises(capi.GooFitError):
                ctx.set_tuning(capi.TUNE_BATCH_MX_PTS, 1)
    finally:
        b.close()


def test_batch_fitter_multi_point_equals_single_point():
    """BatchFitter with derivative rounds in one request against one point per request:
    the same function values (interpreter path), hence the same fits and call counts."""
    counts = [4000, 9000, 6000, 15_000, 7000, 5000]
    data = _toys(counts, 90)
    m = W.build_add1d(start=1.1)
    d = G.UnbinnedDataSet(m.observables)
    d.from_columns([np.concatenate(data)])
    m.pdf.setData(d)
    first = list(np.concatenate([[0], np.cumsum(counts)[:-1]]).astype(int))
    out = {}
    for multi in (True, False):
        bf = G.BatchFitter(m.pdf, first, counts)
        bf.useMultiPoint(multi)
        bf.useDeviceNormalisation(False)  # host norms on both sides: the same function values
        assert bf.multiPoint() == multi
        bf.fit(threads=3, autotune=False)
        out[multi] = (bf.values(), bf.nfcn(), bf.valid(), bf.stats())
    assert out[True][0] == out[False][0]
    assert out[True][1] == out[False][1]
    assert all(out[True][2])
    assert out[True][3]["fcn_calls"] == out[False][3]["fcn_calls"]
    assert out[True][3]["launches"] < out[False][3]["launches"]


def test_batch_points_submit_and_wait(ctx):
    """gf_batch_submit_points / gf_batch_wait_points: two submissions in flight (the
    second staged while the first runs), a third refused until one is waited for, waits
    in either order, the same bits as gf_batch_reduce_points, stale tickets refused."""
    counts = [3000, 5000, 4000]
    data = _toys(counts, 61)
    first = np.concatenate([[0], np.cumsum(counts)[:-1]])
    top, _ = add1d()
    f, p, c, o, n = O.flatten(top).arrays()
    ctx.set_program(f, p, c, o, n)
    ctx.upload_columns([np.concatenate(data)])
    b = capi.Batch(ctx, first, counts)
    rng = np.random.default_rng(21)
    sets = []
    for _ in range(6):
        m = O.Model(_add1d_point(rng), np.zeros((1, 1)))
        m.normalize()
        m.op.sync_params()
        m.op.sync_norms()
        sets.append((m.op.p.copy(), m.op.n.copy()))
    try:
        plan1 = [(0, sets[0]), (0, sets[1]), (1, sets[2])]
        plan2 = [(2, sets[3]), (2, sets[4]), (2, sets[5])]
        for t, (pp, nn) in plan1:
            b.add_point(t, pp, nn)
        t1 = b.submit_points()
        for t, (pp, nn) in plan2:
            b.add_point(t, pp, nn)
        t2 = b.submit_points()
        with pytest.raises(RuntimeError, match="two submissions in flight"):
            b.add_point(1, *sets[0])
        with pytest.raises(RuntimeError, match="two submissions in flight"):
            b.submit_points()
        b.clear_points()  # no-op on a buffer in flight
        r2 = b.wait_points(t2)
        r1 = b.wait_points(t1)
        with pytest.raises(RuntimeError, match="no submission"):
            b.wait_points(t1)
        for plan, r in ((plan1, r1), (plan2, r2)):
            for t, (pp, nn) in plan:
                b.add_point(t, pp, nn)
            ref = b.reduce_points()
            np.testing.assert_array_equal(np.nan_to_num(r, nan=-1.0), np.nan_to_num(ref, nan=-1.0))
        assert np.isfinite(r1[0, :2]).all() and np.isfinite(r1[1, 0]) and np.isnan(r1[2]).all()
    finally:
        b.close()


def test_batch_fitter_refuses_a_floating_resonance_mass():
    """A resonance's mass sets the lineshape cache that all toys' rows share: a batched
    fit with it floating is refused before any toy starts; fixed again, the fit runs."""
    counts = [3000, 4000]
    n = sum(counts)
    dz = W.build_dalitz(nbins=64, resonances=W.DALITZ_RESONANCES[:4], float_couplings=True)
    m12, m13 = W.generate_dalitz(dz, n, seed=51)
    d = G.UnbinnedDataSet(dz.observables)
    d.from_columns([np.asarray(m12), np.asarray(m13), np.arange(n, dtype=np.float64)])
    dz.pdf.setData(d)
    dz.pdf.setDataSize(n)
    res = next(c for c in dz._keep if isinstance(c, G.Resonances.RBW))
    mass = next(v for v in res.getParameters() if v.name.endswith("_mass"))
    bf = G.BatchFitter(dz.pdf, [0, counts[0]], counts)
    mass.fixed = False
    try:
        with pytest.raises(G.GeneralError, match="must be fixed"):
            bf.fit(threads=2)
    finally:
        mass.fixed = True
    bf.fit(threads=2)
    assert min(bf.nfcn()) > 0


def _fit_both_ways(pdf, first, counts, threads=4):
    """The same batched fits with the points normalised in the launches and on the host."""
    out = {}
    for dev in (True, False):
        bf = G.BatchFitter(pdf, first, counts)
        bf.useDeviceNormalisation(dev)
        bf.fit(threads=threads, autotune=False)
        st = bf.stats()
        assert st["device_norms"] == dev
        out[dev] = (np.array(bf.values()), np.array(bf.errors()), list(bf.nfcn()), list(bf.fval()), list(bf.valid()),
                    st)
    return out


def test_device_normalisation_addpdf_matches_host():
    """Gaussian and Exp integrals of the AddPdf computed per point inside the launch: the
    Gaussian's is the host's product (same bits), the Exp's differs at most in the last bits
    of exp: fits within 1e-5 sigma of the host-normalised ones."""
    counts = [4000, 9000, 6000, 15_000, 7000, 5000]
    data = _toys(counts, 95)
    m = W.build_add1d(start=1.1)
    d = G.UnbinnedDataSet(m.observables)
    d.from_columns([np.concatenate(data)])
    m.pdf.setData(d)
    first = list(np.concatenate([[0], np.cumsum(counts)[:-1]]).astype(int))
    out = _fit_both_ways(m.pdf, first, counts, threads=3)
    vd, ed, _, fd, okd, _ = out[True]
    vh, eh, _, fh, okh, _ = out[False]
    assert all(okd) and all(okh)
    free = eh[0] > 0
    shift = np.abs(vd - vh)[:, free] / eh[:, free]
    assert shift.max() < 1e-5, shift.max()
    np.testing.assert_allclose(fd, fh, rtol=1e-12)


def test_device_normalisation_dalitz_is_the_host_quadratic_form():
    """An Amp3Body's coupling quadratic form over its cached integrals computed per point
    inside the launch, in the host's operation order: the same function values, hence the
    same fits and call counts as with the host's normalize() per point."""
    counts = [5000, 8000, 6000, 7000]
    n = sum(counts)
    dz = W.build_dalitz(nbins=200, resonances=W.DALITZ_RESONANCES[:6])
    m12, m13 = W.generate_dalitz(dz, n, seed=61)
    d = G.UnbinnedDataSet(dz.observables)
    d.from_columns([np.asarray(m12), np.asarray(m13), np.arange(n, dtype=np.float64)])
    dz.pdf.setData(d)
    dz.pdf.setDataSize(n)
    truth = {k: v.value for k, v in dz.params.items()}
    rng = np.random.default_rng(62)
    for k in dz.params:
        dz.params[k].value = truth[k] * (1 + 0.05 * rng.standard_normal())
    first = list(np.concatenate([[0], np.cumsum(counts)[:-1]]).astype(int))
    out = _fit_both_ways(dz.pdf, first, counts, threads=4)
    assert np.array_equal(out[True][0], out[False][0])
    assert out[True][2] == out[False][2]
    assert out[True][3] == out[False][3]
    assert all(out[True][4])


def test_extended_addpdf_toys_match_sequential_fits():
    """Extended AddPdf (signal and background yields): each toy's batched fit adds its own
    E - N log E (AddPdf.cu:187-204) with its own event count, so the yields come out within
    0.01 sigma of one FitManager fit per toy."""
    counts = [3000, 7000, 5000, 9000]
    data = _toys(counts, 120)
    x = G.Observable("x", -5.0, 5.0)
    mu = G.Variable("mu", 0.32, 0.01, -2.0, 2.0)
    sg = G.Variable("sigma", 0.85, 0.01, 0.1, 3.0)
    al = G.Variable("alpha", -0.38, 0.01, -2.0, -0.01)
    ns = G.Variable("nsig", 1500.0, 10.0, 0.0, 50000.0)
    nb = G.Variable("nbkg", 4000.0, 10.0, 0.0, 50000.0)
    g = G.GaussianPdf("g", x, mu, sg)
    e = G.ExpPdf("e", x, al)
    tot = G.AddPdf("ext", [ns, nb], [g, e])
    params = dict(mu=mu, sigma=sg, alpha=al, nsig=ns, nbkg=nb)
    start = {k: v.value for k, v in params.items()}
    d = G.UnbinnedDataSet([x])
    d.from_columns([np.concatenate(data)])
    tot.setData(d)
    first = list(np.concatenate([[0], np.cumsum(counts)[:-1]]).astype(int))
    bf = G.BatchFitter(tot, first, counts)
    bf.fit(threads=4, autotune=False)
    assert bf.stats()["device_norms"]
    vals, valid = np.array(bf.values()), bf.valid()
    for t, xs in enumerate(data):
        for k, v in start.items():
            params[k].value = v
        dt = G.UnbinnedDataSet([x])
        dt.from_columns([xs])
        tot.setData(dt)
        fm = G.FitManager(tot)
        fm.setVerbosity(0)
        r = fm.fit()
        assert valid[t] and r.IsValid()
        sv, se = np.array(r.values()), np.array(r.errors())
        free = se > 0
        shift = np.abs(vals[t] - sv)[free] / se[free]
        assert shift.max() < 0.01, (t, shift)
        # the yields add up to the toy's size (extended maximum likelihood)
        assert abs(params["nsig"].value + params["nbkg"].value - counts[t]) < 1e-3 * counts[t]
