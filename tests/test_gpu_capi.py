"""Parity of the gfx950 engine with the oracle, through the C ABI.

Tolerances: the NLL sum within 1e-11 relative; single
event values within 1e-13 relative (libm exp/log/erfc/pow differ by <= a few
ulp between glibc and the device math library); grid indices bit-exact.
"""
import math

import numpy as np
import pytest

import oracle as O
from goofit_amd import capi
from goofit_amd import workloads as W

from _models import add1d, arrays, dalitz, dalitz_events, prod3d

pytestmark = pytest.mark.gpu

NLL_RTOL = 1e-11
EVT_RTOL = 1e-13


def _install(ctx, top, decay=None):
    prog, (f, p, c, o, n) = arrays(top)
    ctx.set_program(f, p, c, o, n)
    if decay is not None:
        ctx.set_decay(decay)
    return prog


def _sync(ctx, model: O.Model):
    """Push the oracle model's current parameters and norms to the engine."""
    model.op.sync_params()
    model.op.sync_norms()
    ctx.update_params(model.op.p)
    ctx.update_norms(model.op.n)


@pytest.mark.parametrize("n", [1, 7, 513, 100_000])
def test_add1d_nll(ctx, n):
    top, x = add1d()
    data = W.sample_add1d(n, seed=11)
    aos = data.reshape(-1, 1)
    M = O.Model(top, aos)
    ref = M.calculate_nll()
    _install(ctx, top)
    ctx.upload_columns([data])
    _sync(ctx, M)
    got, comp = ctx.reduce_metric("NLL")
    assert math.isclose(got, ref, rel_tol=NLL_RTOL), (got, ref)
    exact = O.exact_sum(M.nll_terms())
    assert math.isclose(got, exact, rel_tol=1e-13), (got, exact)


def test_first_result_of_a_new_context_is_its_own():
    """Contexts created one after another (pinned result blocks recycled): each first
    reduction reports its own sum, not the last one a destroyed context left behind."""
    for k, n in enumerate((3000, 5000, 7000)):
        top, x = add1d()
        data = W.sample_add1d(n, seed=40 + k)
        M = O.Model(top, data.reshape(-1, 1))
        ref = M.calculate_nll()
        with capi.Context(0) as c:
            _install(c, top)
            c.upload_columns([data])
            _sync(c, M)
            got, _ = c.reduce_metric("NLL")
        assert math.isclose(got, ref, rel_tol=NLL_RTOL), (k, got, ref)


def test_add1d_per_event_prob(ctx):
    top, x = add1d()
    data = W.sample_add1d(4099, seed=12)
    M = O.Model(top, data.reshape(-1, 1))
    M.normalize()
    M.op.sync_params()
    M.op.sync_norms()
    ref = M.op.eval_metric(data.reshape(-1, 1), "PROB")
    _install(ctx, top)
    ctx.upload_columns([data])
    _sync(ctx, M)
    got = ctx.eval_metric("PROB")
    np.testing.assert_allclose(got, ref, rtol=EVT_RTOL, atol=0)


def test_add1d_extended(ctx):
    top, x = add1d(extended=True, frac=0.25)
    data = W.sample_add1d(20_000, seed=13)
    M = O.Model(top, data.reshape(-1, 1))
    M.normalize()
    _install(ctx, top)
    ctx.upload_columns([data])
    _sync(ctx, M)
    got = ctx.reduce_metric("NLL")[0]
    ref = M.op.reduce_metric(data.reshape(-1, 1), "NLL")
    assert math.isclose(got, ref, rel_tol=NLL_RTOL)


def test_gaussian_grid_kat(ctx):
    """tests/simple/NormalizeTest.cpp:57-100 -- Gaussian(1.5, 2.5) Prob on the 100-bin grid of [-5, 10]."""
    x = O.Obs("x", -5, 10, 100)
    g = O.Gaussian("g", x, O.Var("m", 1.5), O.Var("s", 2.5))
    grid = -5 + (np.arange(100) + 0.5) * 0.15
    M = O.Model(g, grid.reshape(-1, 1))
    M.normalize()
    _install(ctx, g)
    ctx.upload_columns([grid])
    _sync(ctx, M)
    vv = ctx.eval_metric("PROB")
    kat = {0: 0.00587129964482, 20: 0.0624318782242, 40: 0.157287605852, 60: 0.0938855055588}
    for i, v in kat.items():
        assert math.isclose(vv[i], v, rel_tol=1e-11), (i, vv[i], v)


def test_prod3d_nll_and_grid(ctx):
    top, (x, y, z) = prod3d()
    xs, ys, zs = W.sample_prod3d(200_000, seed=21)
    aos = np.stack([xs, ys, zs], axis=1)
    M = O.Model(top, aos)
    ref = M.calculate_nll()
    _install(ctx, top)
    ctx.upload_columns([xs, ys, zs])
    # the ExpGaus grid integral (GooPdf::normalize numeric branch)
    eg = top.components[2]
    ranges = [z.lo, z.hi, z.nbins]
    M.op.sync_params()
    ctx.update_params(M.op.p)
    got_bins = ctx.reduce_bins(eg.function_idx, ranges)
    ref_bins = M.op.reduce_bins(eg.function_idx, ranges)
    assert math.isclose(got_bins, ref_bins, rel_tol=1e-12), (got_bins, ref_bins)
    _sync(ctx, M)
    got = ctx.reduce_metric("NLL")[0]
    assert math.isclose(got, ref, rel_tol=NLL_RTOL), (got, ref)


def test_grid_nd_midpoint(ctx):
    """2-D grid over a ProdPdf of two Gaussians: the midpoint rule, first axis fastest."""
    x = O.Obs("x", -5, 10, 37)
    y = O.Obs("y", -5, 10, 23)
    g1 = O.Gaussian("g1", x, O.Var("m1", 1.5), O.Var("s1", 2.5))
    g2 = O.Gaussian("g2", y, O.Var("m2", -1.0), O.Var("s2", 0.5))
    top = O.Prod("p", [g1, g2])
    M = O.Model(top)
    M.normalize()
    _install(ctx, top)
    ctx.upload_columns([np.zeros(4), np.zeros(4)])
    _sync(ctx, M)
    r = [x.lo, x.hi, x.nbins, y.lo, y.hi, y.nbins]
    got = ctx.reduce_bins(0, r)
    ref = M.op.reduce_bins(0, r)
    assert math.isclose(got, ref, rel_tol=1e-12)
    # closed form of the same midpoint sum: product of the two 1-D sums
    cx = x.lo + (np.arange(x.nbins) + 0.5) * x.bin_size()
    cy = y.lo + (np.arange(y.nbins) + 0.5) * y.bin_size()
    sx = np.exp(-0.5 * (cx - 1.5) ** 2 / 2.5 ** 2).sum() / (2.5 * math.sqrt(2 * math.pi))
    sy = np.exp(-0.5 * (cy + 1.0) ** 2 / 0.5 ** 2).sum() / (0.5 * math.sqrt(2 * math.pi))
    assert math.isclose(got, sx * sy, rel_tol=1e-12)


def test_polynomials(ctx):
    x = O.Obs("x", -2, 3, 50)
    t = [O.Var("c0", 0.5), O.Var("c1", -0.25), O.Var("c2", 0.125)]
    pol = O.Pdf("poly", [x], t, [1])
    pol.func = O.FN["POLYNOMIAL"]
    ep = O.Pdf("expoly", [x], [O.Var("e0", 0.1), O.Var("e1", -0.3), O.Var("e2", 0.02)])
    ep.func = O.FN["EXP_POLY"]
    top = O.Add("s", [O.Var("f", 0.4)], [pol, ep])
    data = np.linspace(-1.9, 2.9, 301)
    M = O.Model(top, data.reshape(-1, 1))
    M.op.sync_params()
    ref = M.op.eval_metric(data.reshape(-1, 1), "EVAL")
    _install(ctx, top)
    ctx.upload_columns([data])
    _sync(ctx, M)
    got = ctx.eval_metric("EVAL")
    np.testing.assert_allclose(got, ref, rtol=EVT_RTOL, atol=1e-300)


def test_multipolynomial_2d(ctx):
    x = O.Obs("x", 0, 1, 10)
    y = O.Obs("y", 0, 1, 10)
    coeffs = [O.Var(f"k{i}", 0.3 + 0.1 * i) for i in range(6)]
    mp = O.MultiPolynomial("mp", [x, y], coeffs, [O.Var("ox", 0.1), O.Var("oy", -0.2)], 2)
    top = O.Prod("p", [mp])
    rng = np.random.default_rng(3)
    xs, ys = rng.random(1000), rng.random(1000)
    aos = np.stack([xs, ys], 1)
    M = O.Model(top, aos)
    M.op.sync_params()
    ref = M.op.eval_metric(aos, "EVAL")
    _install(ctx, top)
    ctx.upload_columns([xs, ys])
    _sync(ctx, M)
    got = ctx.eval_metric("EVAL")
    np.testing.assert_allclose(got, ref, rtol=EVT_RTOL)


# ------------------------------------------------------------------- Dalitz


def _dalitz_setup(ctx, n, nbins=64, seed=5):
    top, (m12, m13, num) = dalitz(nbins=nbins)
    aos = dalitz_events(n, seed)
    M = O.Model(top, aos, decay=W.DALITZ_DECAY)
    _install(ctx, top, decay=W.DALITZ_DECAY)
    ctx.upload_columns([aos[:, 0], aos[:, 1], aos[:, 2]])
    ctx.set_identity_column(2)
    ctx.dalitz_set_data_size(top.function_idx, n)
    return top, M, aos


def test_dalitz_wave_cache(ctx):
    n = 5000
    top, M, aos = _dalitz_setup(ctx, n)
    M.op.sync_params()
    ctx.update_params(M.op.p)
    ctx.dalitz_cache(top.function_idx)
    res_funcs, eff = M.op.dalitz_layout(top.function_idx)
    for i, rf in enumerate(res_funcs):
        ref = M.op.dalitz_wave(top.function_idx, rf, aos)
        got = ctx.dalitz_get_cache(top.function_idx, i, n)
        np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-14 * np.abs(ref).max())
    outside = ~W.in_dalitz_np(aos[:, 0], aos[:, 1])
    assert outside.any()
    assert np.all(ctx.dalitz_get_cache(top.function_idx, 1, n)[outside] == 0)


def test_dalitz_grid_gram(ctx):
    top, M, aos = _dalitz_setup(ctx, 100, nbins=97)
    M.op.sync_params()
    ctx.update_params(M.op.p)
    res_funcs, eff = M.op.dalitz_layout(top.function_idx)
    grid6 = [0.0, 3.0, 97, 0.0, 3.0, 97]
    ref = M.op.dalitz_integrals(top.function_idx, grid6, eff)
    got = ctx.dalitz_gram(top.function_idx, grid6, eff, len(res_funcs))
    scale = np.abs(ref).max()
    np.testing.assert_allclose(got, ref, rtol=1e-11, atol=1e-13 * scale)
    # Hermitian, real non-negative diagonal
    np.testing.assert_allclose(got, got.conj().T, rtol=0, atol=1e-13 * scale)
    assert np.all(np.diag(got).real >= 0)


def test_dalitz_nll(ctx):
    n = 20_000
    top, M, aos = _dalitz_setup(ctx, n, nbins=128)
    ref = M.calculate_nll()  # normalize (grid) + coherent-sum NLL
    M.op.sync_params()
    ctx.update_params(M.op.p)
    ctx.dalitz_cache(top.function_idx)
    _sync(ctx, M)
    got = ctx.reduce_metric("NLL")[0]
    assert math.isclose(got, ref, rel_tol=NLL_RTOL), (got, ref)


def test_dalitz_points_and_cache_gram(ctx):
    n = 3000
    top, M, aos = _dalitz_setup(ctx, n)
    M.op.sync_params()
    ctx.update_params(M.op.p)
    ctx.dalitz_cache(top.function_idx)
    res_funcs, eff = M.op.dalitz_layout(top.function_idx)
    waves = np.stack([M.op.dalitz_wave(top.function_idx, rf, aos) for rf in res_funcs])  # nres x n
    ref = waves @ waves.conj().T
    got = ctx.dalitz_cache_gram(top.function_idx, len(res_funcs))
    np.testing.assert_allclose(got, ref, rtol=1e-11, atol=1e-12 * np.abs(ref).max())
    w = np.random.default_rng(9).random(n)
    inside = W.in_dalitz_np(aos[:, 0], aos[:, 1]).astype(float)
    ref_p = (waves * (w * inside)) @ waves.conj().T
    got_p = ctx.dalitz_gram_points(top.function_idx, aos[:, 0], aos[:, 1], w, eff, len(res_funcs))
    np.testing.assert_allclose(got_p, ref_p, rtol=1e-11, atol=1e-12 * np.abs(ref_p).max())


def test_dalitz_event_number_out_of_cache(ctx):
    n = 256
    top, M, aos = _dalitz_setup(ctx, n)
    bad = aos.copy()
    bad[:, 2] = np.arange(n) + 10 * n  # event numbers past the cache
    ctx.upload_columns([bad[:, 0], bad[:, 1], bad[:, 2]])
    ctx.dalitz_set_data_size(top.function_idx, n)
    M.op.sync_params()
    ctx.update_params(M.op.p)
    ctx.dalitz_cache(top.function_idx)
    with pytest.raises(capi.GooFitError):
        ctx.reduce_metric("NLL")
    with pytest.raises(capi.GooFitError):
        ctx.set_identity_column(2)


def test_bad_program_rejected(ctx):
    top, x = add1d()
    prog, (f, p, c, o, n) = arrays(top)
    p2 = p.copy()
    p2[prog.nodes[1].parameters_idx] = 3  # Gaussian claims 3 parameters
    with pytest.raises(capi.GooFitError):
        ctx.set_program(f, p2, c, o, n)
    f2 = f.copy()
    f2[1] = 99
    with pytest.raises(capi.GooFitError):
        ctx.set_program(f2, p, c, o, n)


def test_dalitz_nll_paths(ctx):
    """Both Dalitz NLL paths against the oracle: the coherent-sum fast path with a
    degree-1 efficiency polynomial, and the interpreter when the event numbers are
    not the row index (a permutation)."""
    n = 30_000
    m12o = O.Obs("m12", 0.0, 3.0, 100)
    m13o = O.Obs("m13", 0.0, 3.0, 100)
    eff = O.MultiPolynomial("eff", [m12o, m13o], [O.Var("k0", 0.8), O.Var("k1", 0.1), O.Var("k2", -0.05)],
                            [O.Var("o0", 0.5), O.Var("o1", 0.2)], 1)
    top, _ = dalitz(nbins=100, eff=eff)
    top.m12, top.m13 = m12o, m13o
    top.observables[0], top.observables[1] = m12o, m13o
    aos = dalitz_events(n, 17)
    for permute in (False, True):
        a = aos.copy()
        if permute:
            perm = np.random.default_rng(3).permutation(n)
            a[:, 2] = perm.astype(float)
        M = O.Model(top, a, decay=W.DALITZ_DECAY)
        ref = M.calculate_nll()
        _install(ctx, top, decay=W.DALITZ_DECAY)
        ctx.upload_columns([a[:, 0], a[:, 1], a[:, 2]])
        if not permute:
            ctx.set_identity_column(2)
        ctx.dalitz_set_data_size(top.function_idx, n)
        M.op.sync_params()
        ctx.update_params(M.op.p)
        ctx.dalitz_cache(top.function_idx)
        _sync(ctx, M)
        got = ctx.reduce_metric("NLL")[0]
        assert math.isclose(got, ref, rel_tol=NLL_RTOL), (permute, got, ref)
        probs = ctx.eval_metric("PROB")
        np.testing.assert_allclose(probs, M.op.eval_metric(a, "PROB"), rtol=1e-12, atol=0)
    # the grid integrals with a non-constant efficiency (the general efficiency path)
    res_funcs, effi = M.op.dalitz_layout(top.function_idx)
    grid6 = [0.0, 3.0, 100, 0.0, 3.0, 100]
    ref_i = M.op.dalitz_integrals(top.function_idx, grid6, effi)
    got_i = ctx.dalitz_gram(top.function_idx, grid6, effi, len(res_funcs))
    np.testing.assert_allclose(got_i, ref_i, rtol=1e-11, atol=1e-13 * np.abs(ref_i).max())


def test_dalitz_zero_outside_cache(ctx):
    """The coherent-sum kernel that takes the cache's zeros outside the Dalitz region
    (constant efficiency, cache filled from these events) gives the same NLL and per-event
    values as the one that reads m12/m13; after new events are uploaded without refilling
    the cache it is not used."""
    n = 40_000
    top, M, aos = _dalitz_setup(ctx, n, nbins=80, seed=23)
    assert (~W.in_dalitz_np(aos[:, 0], aos[:, 1])).sum() > 1000
    ref = M.calculate_nll()
    M.op.sync_params()
    ctx.update_params(M.op.p)
    ctx.dalitz_cache(top.function_idx)
    _sync(ctx, M)
    out = {}
    for zo in (1, 0):
        ctx.set_tuning(capi.TUNE_ZERO_OUTSIDE, zo)
        out[zo] = (ctx.reduce_metric("NLL")[0], ctx.eval_metric("NLL"))
    ctx.set_tuning(capi.TUNE_ZERO_OUTSIDE, 1)
    assert out[1][0] == out[0][0]
    assert math.isclose(out[1][0], ref, rel_tol=NLL_RTOL), (out[1][0], ref)
    np.testing.assert_array_equal(out[1][1], out[0][1])
    # new m12/m13 values with the old cache: the kernel must read the masses again
    other = dalitz_events(n, 24)
    ctx.upload_columns([other[:, 0], other[:, 1], other[:, 2]])
    ctx.set_identity_column(2)
    stale = {}
    for zo in (1, 0):
        ctx.set_tuning(capi.TUNE_ZERO_OUTSIDE, zo)
        stale[zo] = ctx.reduce_metric("NLL")[0]
    ctx.set_tuning(capi.TUNE_ZERO_OUTSIDE, 1)
    assert stale[1] == stale[0]


def test_reduce_metric_multi_equals_single(ctx):
    """gf_reduce_metric_multi: each set's sum equals gf_reduce_metric's with that set
    installed, bit for bit, on the coherent-sum path (20 sets: one full sweep of 16
    and one of 4) and on the interpreter (one reduction per set)."""
    n = 50_000
    top, M, aos = _dalitz_setup(ctx, n, nbins=64, seed=31)
    M.calculate_nll()
    M.op.sync_params()
    ctx.update_params(M.op.p)
    ctx.dalitz_cache(top.function_idx)
    _sync(ctx, M)
    p0, n0 = M.op.p.copy(), M.op.n.copy()
    rng = np.random.default_rng(5)
    res = top.resonances
    sets_p, sets_n = [], []
    for k in range(20):
        p = p0.copy()
        for r in res:  # perturb every coupling (the waves do not depend on them)
            p[top.parameters_idx + 1 + 2 * res.index(r)] *= 1 + 1e-3 * rng.standard_normal()
        nn_ = n0.copy()
        nn_[top.normal_idx + 1] *= 1 + 1e-4 * rng.standard_normal()
        sets_p.append(p)
        sets_n.append(nn_)
    got = ctx.reduce_metric_multi("NLL", sets_p, sets_n)
    for k in range(20):
        ctx.update_params(sets_p[k])
        ctx.update_norms(sets_n[k])
        single = ctx.reduce_metric("NLL")[0]
        assert got[k] == single, (k, got[k], single)
    # the interpreter path: an AddPdf model
    top1, x = add1d()
    data = W.sample_add1d(20_000, seed=9)
    M1 = O.Model(top1, data.reshape(-1, 1))
    M1.calculate_nll()
    _install(ctx, top1)
    ctx.upload_columns([data])
    _sync(ctx, M1)
    q0, m0 = M1.op.p.copy(), M1.op.n.copy()
    sp = [q0 * (1 + 1e-3 * k) for k in range(3)]
    sn = [m0 for _ in range(3)]
    got1 = ctx.reduce_metric_multi("NLL", sp, sn)
    for k in range(3):
        ctx.update_params(sp[k])
        ctx.update_norms(sn[k])
        assert got1[k] == ctx.reduce_metric("NLL")[0]


def test_dalitz_generate(ctx):
    """Accept-reject generation on the GPU: reproducible for a seed, every event inside the plot,
    and the histogram over cells fully inside the plot follows |A|^2 (chi^2 per degree of freedom)."""
    top, M, aos = _dalitz_setup(ctx, 1000, nbins=64)
    M.op.sync_params()
    ctx.update_params(M.op.p)
    n = 200_000
    box = [0.0, 3.0, 0.0, 3.0]
    _, _, tried, fmax = ctx.dalitz_generate(top.function_idx, 1000, 5, box)  # envelope estimate
    assert tried >= 1000 and fmax > 0
    a12, a13, _, _ = ctx.dalitz_generate(top.function_idx, n, 5, box, fmax)
    b12, b13, _, _ = ctx.dalitz_generate(top.function_idx, n, 5, box, fmax)
    c12, _, _, _ = ctx.dalitz_generate(top.function_idx, n, 6, box, fmax)
    assert np.array_equal(a12, b12) and np.array_equal(a13, b13)
    assert not np.array_equal(a12, c12)
    assert W.in_dalitz_np(a12, a13).all()
    # expected content of 20 x 20 coarse cells from |A|^2 on a 400 x 400 midpoint grid
    F, C = 400, 20
    fc = (np.arange(F) + 0.5) * 3.0 / F
    g12, g13 = np.meshgrid(fc, fc, indexing="ij")
    pts = np.stack([g12.ravel(), g13.ravel()], 1)
    ctx.upload_columns([pts[:, 0], pts[:, 1], np.arange(len(pts), dtype=np.float64)])
    ctx.set_identity_column(2)
    ctx.dalitz_set_data_size(top.function_idx, len(pts))
    ctx.dalitz_cache(top.function_idx)
    val = ctx.eval_metric("EVAL").reshape(F, F)
    S = val.reshape(C, F // C, C, F // C).sum(axis=(1, 3))
    edges = np.linspace(0, 3, C + 1)
    inside = np.ones((C, C), bool)
    for i in range(C):
        for j in range(C):
            for x in (edges[i], edges[i + 1]):
                for y in (edges[j], edges[j + 1]):
                    inside[i, j] &= bool(W.in_dalitz_np(np.array([x]), np.array([y]))[0])
    H, _, _ = np.histogram2d(a12, a13, bins=[edges, edges])
    obs, exp = H[inside], S[inside]
    exp = exp / exp.sum() * obs.sum()
    chi2 = ((obs - exp) ** 2 / exp).sum()
    ndf = inside.sum() - 1
    assert ndf > 50
    assert chi2 / ndf < 1.35, (chi2, ndf)


@pytest.mark.parametrize("nres,n", [(16, 100_000), (12, 40_001), (8, 70_001), (5, 3_000), (3, 257)])
def test_multi_point_sums_on_the_matrix_cores_equal_single_sweeps(ctx, nres, n):
    """gf_reduce_metric_multi forms the coherent sums of its parameter sets on the FP64
    matrix cores (k_coh_mx) when the NLL of a constant efficiency allows it, and the single
    sweep then groups each lane's events the same way; v_mfma_f64_16x16x4f64, fed its K
    terms in the order of the vector path's fused complex multiply-adds, gives the same
    elements bit for bit.  So with the matrix-core path on, and with it off (vector
    kernels, their own grouping), every multi-point sum equals the single reduction of the
    same parameters exactly; the two paths agree with each other to rounding (1e-13) and
    with the oracle (1e-11).  5, 16, 23, 32 and 45 perturbed coupling sets: the matrix-core
    sweep takes up to 32 sets (two groups of 16 sharing each wave load), so these cover one
    group, two full and partial groups, and a request spread over two sweeps."""
    table = W.DALITZ_RESONANCES[:nres]
    top, _ = dalitz(nbins=64, resonances=table)
    aos = dalitz_events(n, 40 + nres)
    M = O.Model(top, aos, decay=W.DALITZ_DECAY)
    ref = M.calculate_nll()
    _install(ctx, top, decay=W.DALITZ_DECAY)
    ctx.upload_columns([aos[:, 0], aos[:, 1], aos[:, 2]])
    ctx.set_identity_column(2)
    ctx.dalitz_set_data_size(top.function_idx, n)
    M.op.sync_params()
    ctx.update_params(M.op.p)
    ctx.dalitz_cache(top.function_idx)
    _sync(ctx, M)
    p0, n0 = M.op.p.copy(), M.op.n.copy()
    rng = np.random.default_rng(nres)
    c0 = top.parameters_idx + 1
    try:
        for npts in (16, 5, 32, 23, 45):
            sets = []
            for _ in range(npts):
                p = p0.copy()
                p[c0:c0 + 2 * nres] *= 1 + 0.1 * rng.standard_normal(2 * nres)
                sets.append(p)
            got = {}
            for mx in (1, 0):
                ctx.set_tuning(capi.TUNE_COH_MFMA, mx)
                got[mx] = ctx.reduce_metric_multi("NLL", sets, [n0] * npts)
                for k in sorted({0, npts // 2, min(17, npts - 1), npts - 1}):
                    ctx.update_params(sets[k])
                    assert ctx.reduce_metric("NLL")[0] == got[mx][k], (mx, npts, k)
                ctx.update_params(p0)
            np.testing.assert_allclose(got[1], got[0], rtol=1e-13, atol=0)
        for mx in (1, 0):
            ctx.set_tuning(capi.TUNE_COH_MFMA, mx)
            assert math.isclose(ctx.reduce_metric_multi("NLL", [p0], [n0])[0], ref, rel_tol=1e-11)
    finally:
        ctx.set_tuning(capi.TUNE_COH_MFMA, 1)
        ctx.update_params(p0)
