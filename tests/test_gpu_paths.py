"""Every launch variant of the likelihood sweep against the oracle: parameters in
the kernel arguments or in device arrays, grid caps of 1, 4 and 16 workgroups per
CU, the interpreter compiled per node feature level (trees without a Dalitz node,
with the time-integrated one, with the time-dependent one), the coherent-sum Dalitz
kernel against the interpreter; an empty event set; and the two abort conditions of
GooPdf::calculateNLL (src/PDFs/GooPdf.cu:254-273).

Tolerance: the NLL within 1e-11 relative of the oracle; variants that differ only in
where the parameters are read give the same bits.
"""
import math

import numpy as np
import pytest

import goofit_amd as G
import oracle as O
from goofit_amd import capi
from goofit_amd import workloads as W

from _models import add1d, arrays, dalitz, dalitz_events, prod3d, tddp, tddp_events

pytestmark = pytest.mark.gpu

NLL_RTOL = 1e-11


def _install(ctx, top, cols, decay=None):
    prog, (f, p, c, o, n) = arrays(top)
    ctx.set_program(f, p, c, o, n)
    if decay is not None:
        ctx.set_decay(decay)
    ctx.upload_columns(cols)
    return prog


def _sync(ctx, M):
    M.op.sync_params()
    M.op.sync_norms()
    ctx.update_params(M.op.p)
    ctx.update_norms(M.op.n)


def _variants(ctx, ref):
    """NLL under every (inline, workgroups per CU) setting, each with the partials added
    inside the sweep and by a second launch (the same bits); the defaults restored after."""
    out = {}
    try:
        for blocks in (1, 4, 16):
            ctx.set_tuning(capi.TUNE_BLOCKS_PER_CU, blocks)
            ctx.set_tuning(capi.TUNE_FAST_BLOCKS_PER_CU, blocks)
            for inline in (1, 0):
                ctx.set_tuning(capi.TUNE_INLINE_ARGS, inline)
                got = {}
                for fused in (1, 0):
                    ctx.set_tuning(capi.TUNE_FUSED_FINISH, fused)
                    got[fused] = ctx.reduce_metric("NLL")
                assert got[1] == got[0], (blocks, inline, got)
                got = got[1][0]
                assert math.isclose(got, ref, rel_tol=NLL_RTOL), (blocks, inline, got, ref)
                out[blocks, inline] = got
            assert out[blocks, 1] == out[blocks, 0], blocks
    finally:
        ctx.set_tuning(capi.TUNE_BLOCKS_PER_CU, 4)
        ctx.set_tuning(capi.TUNE_FAST_BLOCKS_PER_CU, 3)
        ctx.set_tuning(capi.TUNE_INLINE_ARGS, 1)
        ctx.set_tuning(capi.TUNE_FUSED_FINISH, 1)
    return out


def test_variants_without_dalitz_nodes(ctx):
    """AddPdf(Gauss, Exp) and ProdPdf(Gauss, Exp, ExpGaus): the kernels without Dalitz nodes."""
    top, _ = add1d()
    data = W.sample_add1d(50_000, seed=31)
    M = O.Model(top, data.reshape(-1, 1))
    ref = M.calculate_nll()
    _install(ctx, top, [data])
    _sync(ctx, M)
    _variants(ctx, ref)
    top3, _ = prod3d()
    xs, ys, zs = W.sample_prod3d(100_000, seed=32)
    M3 = O.Model(top3, np.stack([xs, ys, zs], axis=1))
    ref3 = M3.calculate_nll()
    _install(ctx, top3, [xs, ys, zs])
    _sync(ctx, M3)
    _variants(ctx, ref3)


def test_variants_with_a_dalitz_node(ctx):
    """A lone Amp3Body: the coherent-sum kernel and the interpreter (fast path off), each
    under every inline / grid setting."""
    n = 30_000
    top, _ = dalitz(nbins=64)
    aos = dalitz_events(n, 33)
    M = O.Model(top, aos, decay=W.DALITZ_DECAY)
    ref = M.calculate_nll()
    _install(ctx, top, [aos[:, 0], aos[:, 1], aos[:, 2]], decay=W.DALITZ_DECAY)
    ctx.set_identity_column(2)
    ctx.dalitz_set_data_size(top.function_idx, n)
    M.op.sync_params()
    ctx.update_params(M.op.p)
    ctx.dalitz_cache(top.function_idx)
    _sync(ctx, M)
    fast = _variants(ctx, ref)
    try:
        ctx.set_tuning(capi.TUNE_FAST_DALITZ, 0)
        slow = _variants(ctx, ref)
    finally:
        ctx.set_tuning(capi.TUNE_FAST_DALITZ, 1)
    assert math.isclose(fast[4, 1], slow[4, 1], rel_tol=1e-13)


def test_variants_with_a_time_dependent_node(ctx):
    n = 20_000
    top, _ = tddp(nbins=80)
    aos = tddp_events(n, seed=34)
    M = O.Model(top, aos, decay=W.DALITZ_DECAY)
    ref = M.calculate_nll()
    prog, (f, p, c, o, nn) = arrays(top)
    ctx.set_program(f, p, c, o, nn)
    ctx.set_decay(W.DALITZ_DECAY)
    ctx.upload_columns([np.ascontiguousarray(aos[:, k]) for k in range(aos.shape[1])])
    ctx.set_identity_column(4)
    ctx.dalitz_set_data_size(top.function_idx, n)
    M.op.sync_params()
    ctx.update_params(M.op.p)
    ctx.dalitz_cache(top.function_idx)
    _sync(ctx, M)
    _variants(ctx, ref)


def test_empty_event_set(ctx):
    """No events: every sum is 0 (the reference then aborts in calculateNLL, below)."""
    top, _ = add1d()
    M = O.Model(top, np.zeros((1, 1)))
    M.normalize()
    _install(ctx, top, [np.zeros(0)])
    _sync(ctx, M)
    got, comp = ctx.reduce_metric("NLL")
    assert got == 0.0 and comp == 0.0
    assert ctx.eval_metric("NLL").size == 0


def _gauss_model():
    x = G.Observable("x", -5.0, 5.0)
    x.numbins = 100
    mu, sigma = G.Variable("mu", 0.0, 0.1, -1, 1), G.Variable("sigma", 1.0, 0.1, 0.1, 3)
    return x, mu, sigma, G.GaussianPdf("g", x, mu, sigma)


def test_calculate_nll_aborts_on_zero_nll():
    """GooPdf.cu:269-270: a likelihood sum of exactly 0 aborts -- here every event
    contributes -log(pdf * norm) = 0 because pdf * norm = 1."""
    x, mu, sigma, g = _gauss_model()
    # a 1-bin flat polynomial over a unit-width range: pdf * norm = 1 at every event
    y = G.Observable("y", 0.0, 1.0)
    y.numbins = 1
    c0 = G.Variable("c0", 1.0)
    flat = G.PolynomialPdf("flat", y, [c0], 0)
    d = G.UnbinnedDataSet([y])
    d.from_columns([np.linspace(0.1, 0.9, 50)])
    flat.setData(d)
    with pytest.raises(G.GeneralError, match="zero NLL"):
        flat.calculateNLL()
    # no events at all: the empty sum is 0 as well (GooPdf::reduce_with_metric over an
    # empty range)
    empty = G.UnbinnedDataSet([x])
    empty.from_columns([np.zeros(0)])
    g.setData(empty)
    with pytest.raises(G.GeneralError, match="zero NLL"):
        g.calculateNLL()
    # and the normal case on the same engine still works
    dg = G.UnbinnedDataSet([x])
    dg.from_columns([np.random.default_rng(1).normal(0, 1, 1000).clip(-4.9, 4.9)])
    g.setData(dg)
    assert g.calculateNLL() > 0


def test_calculate_nll_aborts_on_non_positive_normalisation():
    """A PDF whose grid integral is negative aborts before the sum: in
    GooPdf::normalize (GooPdf.cu:383-384) when the integral is numeric, as here, and
    otherwise in calculateNLL (GooPdf.cu:260-261)."""
    y = G.Observable("y", 0.0, 2.0)
    y.numbins = 50
    c0 = G.Variable("c0", -1.0)
    neg = G.PolynomialPdf("neg", y, [c0], 0)
    d = G.UnbinnedDataSet([y])
    d.from_columns([np.linspace(0.1, 1.9, 20)])
    neg.setData(d)
    with pytest.raises(G.GeneralError, match="(?i)non-positive normalization"):
        neg.calculateNLL()


def test_back_to_back_sweep_timing_leaves_the_context_as_it_was(ctx):
    """gf_time_sweep_back_to_back launches sweeps without their finishing kernel (or sweeps
    that add their partials themselves): the reductions before and after it give the same
    bits, with parameters inline and in device arrays."""
    n = 20_000
    top, _ = dalitz(nbins=64)
    aos = dalitz_events(n, 35)
    M = O.Model(top, aos, decay=W.DALITZ_DECAY)
    ref = M.calculate_nll()
    _install(ctx, top, [aos[:, 0], aos[:, 1], aos[:, 2]], decay=W.DALITZ_DECAY)
    ctx.set_identity_column(2)
    ctx.dalitz_set_data_size(top.function_idx, n)
    M.op.sync_params()
    ctx.update_params(M.op.p)
    ctx.dalitz_cache(top.function_idx)
    _sync(ctx, M)
    try:
        for inline in (1, 0):
            ctx.set_tuning(capi.TUNE_INLINE_ARGS, inline)
            before = ctx.reduce_metric("NLL")[0]
            assert ctx.time_sweep_back_to_back("NLL", 16) > 0
            after = ctx.reduce_metric("NLL")[0]
            assert before == after and math.isclose(after, ref, rel_tol=NLL_RTOL)
    finally:
        ctx.set_tuning(capi.TUNE_INLINE_ARGS, 1)


def test_in_kernel_finish_between_other_launches(ctx):
    """The coherent-sum sweep with its partials added inside the kernel leaves its slots empty
    for the next launch: 60 reductions with moving couplings, interleaved with back-to-back
    timing sweeps, multi-point sweeps, per-event evaluations and the interpreter, each equal
    bit for bit to the same reduction finished by a second launch."""
    n = 30_000
    top, _ = dalitz(nbins=64)
    aos = dalitz_events(n, 36)
    M = O.Model(top, aos, decay=W.DALITZ_DECAY)
    M.calculate_nll()
    _install(ctx, top, [aos[:, 0], aos[:, 1], aos[:, 2]], decay=W.DALITZ_DECAY)
    ctx.set_identity_column(2)
    ctx.dalitz_set_data_size(top.function_idx, n)
    M.op.sync_params()
    ctx.update_params(M.op.p)
    ctx.dalitz_cache(top.function_idx)
    _sync(ctx, M)
    base = np.array(M.op.p, dtype=np.float64)
    rng = np.random.default_rng(7)
    try:
        for it in range(60):
            p = base.copy()  # the couplings move (the waves do not depend on them)
            c0 = top.parameters_idx + 1
            p[c0:c0 + 8] *= 1 + 0.05 * rng.standard_normal(8)
            ctx.update_params(p)
            ctx.set_tuning(capi.TUNE_FUSED_FINISH, 1)
            fused = ctx.reduce_metric("NLL")
            kind = it % 4
            if kind == 0:
                assert ctx.time_sweep_back_to_back("NLL", 3) > 0
            elif kind == 1:
                multi = ctx.reduce_metric_multi("NLL", [p, p], [M.op.n, M.op.n])
                assert multi[0] == fused[0] and multi[1] == fused[0]
            elif kind == 2:
                assert ctx.eval_metric("NLL").size == n
            else:
                ctx.set_tuning(capi.TUNE_FAST_DALITZ, 0)
                slow = ctx.reduce_metric("NLL")[0]
                ctx.set_tuning(capi.TUNE_FAST_DALITZ, 1)
                assert math.isclose(slow, fused[0], rel_tol=1e-13)
            ctx.set_tuning(capi.TUNE_FUSED_FINISH, 0)
            second = ctx.reduce_metric("NLL")
            assert fused == second, (it, fused, second)
            ctx.set_tuning(capi.TUNE_FUSED_FINISH, 1)
            assert ctx.reduce_metric("NLL") == fused, it
    finally:
        ctx.set_tuning(capi.TUNE_FUSED_FINISH, 1)
        ctx.set_tuning(capi.TUNE_FAST_DALITZ, 1)


@pytest.mark.parametrize("n", [1, 255, 257, 4097, 1_000_003])
def test_in_kernel_finish_flat_sizes(ctx, n):
    """The flat ProdPdf sweep at sizes from one event (a single workgroup finishes itself) to
    1e6 + 3 (a grid capped by the device): the same bits with the partials added in the
    kernel or by a second launch, and the oracle's NLL."""
    top3, _ = prod3d()
    xs, ys, zs = W.sample_prod3d(n, seed=40 + n % 97)
    _install(ctx, top3, [xs, ys, zs])
    M3 = O.Model(top3, np.stack([xs, ys, zs], axis=1))
    ref = M3.calculate_nll()
    _sync(ctx, M3)
    try:
        got = {}
        for fused in (1, 0, 1):
            ctx.set_tuning(capi.TUNE_FUSED_FINISH, fused)
            got.setdefault(fused, []).append(ctx.reduce_metric("NLL"))
        assert got[1][0] == got[0][0] == got[1][1], got
        assert math.isclose(got[1][0][0], ref, rel_tol=NLL_RTOL)
    finally:
        ctx.set_tuning(capi.TUNE_FUSED_FINISH, 1)


def test_nll_with_subnormal_and_zero_pdf_values(ctx):
    """The NLL as one logarithm per lane over PDF values that reach the subnormal range and
    underflow to 0 (a Gaussian of width 0.05 over |x| < 2.2: exp(-0.5 (x/0.05)^2) passes
    1e-308 near |x| = 1.88 and is 0 beyond 1.93, so those events add 0, as the reference's
    metric gives them): the flat kernel, the interpreter, and both finishes give the same bits,
    within 1e-11 of the oracle's per-event -log sum."""
    n = 200_000
    top3, _ = prod3d(sigma=0.05)
    rng = np.random.default_rng(44)
    xs = rng.uniform(-2.2, 2.2, n)
    _, ys, zs = W.sample_prod3d(n, seed=45)
    _install(ctx, top3, [xs, ys, zs])
    M3 = O.Model(top3, np.stack([xs, ys, zs], axis=1))
    ref = M3.calculate_nll()
    _sync(ctx, M3)
    g = np.exp(-0.5 * (xs / 0.05) ** 2)
    assert np.any((g > 0) & (g < 2.3e-308)) and np.any(g == 0)
    got = {}
    try:
        for flat in (1, 0):
            ctx.set_tuning(capi.TUNE_FLAT, flat)
            for fused in (1, 0):
                ctx.set_tuning(capi.TUNE_FUSED_FINISH, fused)
                got[flat, fused] = ctx.reduce_metric("NLL")
    finally:
        ctx.set_tuning(capi.TUNE_FLAT, 1)
        ctx.set_tuning(capi.TUNE_FUSED_FINISH, 1)
    assert len(set(got.values())) == 1, got
    assert math.isclose(got[1, 1][0], ref, rel_tol=NLL_RTOL), (got[1, 1], ref)
