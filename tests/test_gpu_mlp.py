"""The learned importance-sampling proposal on the GPU, through the C ABI:
BF16-MFMA forward and backward against the CPU oracle (oracle/mlp.py), the
uniform and proposal samplers, and the gain of the trained proposal over
uniform phase-space Monte Carlo at equal point count."""
import math

import numpy as np
import pytest
import torch

import oracle as O
from goofit_amd import capi
from goofit_amd import workloads as W
from oracle import mlp

from _models import dalitz, dalitz_events

pytestmark = pytest.mark.gpu

BOX = [0.0, 3.0, 0.0, 3.0]


def _rand_params(seed):
    rng = np.random.default_rng(seed)
    p = np.zeros(mlp.NPARAM, np.float32)
    p[mlp.W1:mlp.W2] = rng.uniform(-3, 3, mlp.W2)
    p[mlp.W2:mlp.B2] = rng.uniform(-0.2, 0.2, mlp.H * mlp.H)
    p[mlp.B2:mlp.W3] = rng.uniform(-0.1, 0.1, mlp.H)
    p[mlp.W3:mlp.B3] = rng.uniform(-0.3, 0.3, mlp.H)
    p[mlp.B3] = 0.2
    return p


@pytest.fixture
def net(ctx):
    q = capi.Proposal(ctx, seed=5, batch_cap=4096)
    yield q
    q.close()


def test_forward_matches_oracle(net):
    """f within 2e-3 (1 + max|f|) of the oracle, which rounds the same operands to bf16 but
    accumulates in float64 (the device accumulates in float32)."""
    p = _rand_params(1)
    net.set_params(p)
    assert np.array_equal(net.get_params(), p)
    rng = np.random.default_rng(2)
    for n in (1000, 10_000):  # not a multiple of 64; more than one batch_cap chunk
        xy = rng.uniform(-1, 1, (n, 2)).astype(np.float32)
        got = net.forward(xy).astype(np.float64)
        ref = mlp.forward(p.astype(np.float64), xy.astype(np.float64), bf16_ops=True)
        tol = 2e-3 * (1 + np.abs(ref).max())
        assert np.abs(got - ref).max() <= tol, np.abs(got - ref).max()


def test_gradient_matches_oracle(net):
    """sum_i coef_i grad f(x_i): each parameter block within 2 % (relative L2) of the oracle."""
    p = _rand_params(3)
    net.set_params(p)
    rng = np.random.default_rng(4)
    n = 777
    xy = rng.uniform(-1, 1, (n, 2)).astype(np.float32)
    coef = rng.normal(size=n).astype(np.float32) / n
    got = net.gradient(xy, coef).astype(np.float64)
    ref = mlp.gradient(p.astype(np.float64), xy.astype(np.float64), coef.astype(np.float64), bf16_ops=True)
    for name, a, b in (("W1", mlp.W1, mlp.B1), ("b1", mlp.B1, mlp.W2), ("W2", mlp.W2, mlp.B2),
                       ("b2", mlp.B2, mlp.W3), ("w3", mlp.W3, mlp.B3), ("b3", mlp.B3, mlp.NPARAM)):
        err = np.linalg.norm(got[a:b] - ref[a:b]) / max(np.linalg.norm(ref[a:b]), 1e-30)
        assert err < 2e-2, (name, err)


def _dalitz_ctx(ctx, table=None, nbins=200):
    top, _ = dalitz(resonances=table, nbins=nbins)
    prog = O.flatten(top)
    f, p, c, o, n = prog.arrays()
    ctx.set_program(f, p, c, o, n)
    ctx.set_decay(W.DALITZ_DECAY)
    return top, prog


def _sq_amp(ctx, top, m12, m13):
    """|A|^2 * eff at points (0 outside the Dalitz plot) through the engine."""
    n = m12.size
    ctx.upload_columns([m12, m13, np.arange(n, dtype=np.float64)])
    ctx.set_identity_column(2)
    ctx.dalitz_set_data_size(top.function_idx, n)
    ctx.dalitz_cache(top.function_idx)
    return ctx.eval_metric("EVAL")


def test_uniform_points_integrate_the_dalitz_area(ctx):
    top, prog = _dalitz_ctx(ctx, table=[("nonr", "NONRES", 0.0, 0.0, 0, 0)])
    ctx.dalitz_set_data_size(top.function_idx, 1)
    n = 2_000_000
    ctx.dalitz_set_norm_points(top.function_idx, n, 11, box=BOX)
    res, eff = O.OracleProg(top, W.DALITZ_DECAY).dalitz_layout(0)
    I = ctx.dalitz_gram_norm_points(top.function_idx, eff, 0, n, 1)
    area = W.dalitz_area()
    p = area / 9.0
    sigma = 9.0 * math.sqrt(p * (1 - p) / n)
    # NonRes wave = 1 inside the plot, 0 outside
    val = I[0, 0].real
    assert abs(val - area) < 5 * sigma, (val, area, sigma)


def test_proposal_beats_uniform_at_equal_points(ctx):
    """Trained on events generated from the model: unbiased, and at N = 1e5 its norm-integral
    error is well below uniform Monte Carlo's."""
    top, prog = _dalitz_ctx(ctx)
    res, eff = O.OracleProg(top, W.DALITZ_DECAY).dalitz_layout(0)
    nres = len(res)
    coups = np.array([complex(r.amp_real.value, r.amp_imag.value) for r in top.resonances])
    # events from the model by accept-reject on uniform points
    rng = np.random.default_rng(21)
    m12 = rng.uniform(0, 3, 1_500_000)
    m13 = rng.uniform(0, 3, 1_500_000)
    # the model's |A|^2 needs the couplings: the oracle program holds them
    val = _sq_amp(ctx, top, m12, m13)
    keep = rng.uniform(0, val.max(), val.size) < val
    ev12, ev13 = m12[keep], m13[keep]
    assert ev12.size > 50_000
    # reference: fine midpoint grid
    G = 1500
    I_grid = ctx.dalitz_gram(top.function_idx, [0.0, 3.0, G, 0.0, 3.0, G], eff, nres) * (3.0 / G) ** 2
    ref_norm = O.dalitz_norm_from_integrals(coups, I_grid, 1.0, 1.0)

    q = capi.Proposal(ctx, seed=9, batch_cap=16384)
    try:
        l0, l1 = q.train_host(ev12, ev13, BOX, steps=600, batch=8192, lr=3e-3, seed=4)
        assert l1 < l0
        q.build(BOX, grid=256, alpha=0.1)
        N = 100_000
        ctx.dalitz_set_data_size(top.function_idx, 1)
        errs_u, errs_q = [], []
        for s in range(8):
            ctx.dalitz_set_norm_points(top.function_idx, N, 100 + s, box=BOX)
            Iu = ctx.dalitz_gram_norm_points(top.function_idx, eff, 0, N, nres)
            ctx.dalitz_set_norm_points(top.function_idx, N, 200 + s, proposal=q)
            Iq = ctx.dalitz_gram_norm_points(top.function_idx, eff, 0, N, nres)
            errs_u.append(O.dalitz_norm_from_integrals(coups, Iu, 1.0, 1.0) / ref_norm - 1)
            errs_q.append(O.dalitz_norm_from_integrals(coups, Iq, 1.0, 1.0) / ref_norm - 1)
        errs_u, errs_q = np.array(errs_u), np.array(errs_q)
        rms_u, rms_q = np.sqrt((errs_u ** 2).mean()), np.sqrt((errs_q ** 2).mean())
        print(f"norm rel. error at N=1e5: uniform {rms_u:.2e}, proposal {rms_q:.2e}; loss {l0:.3f} -> {l1:.3f}")
        assert rms_q < 0.5 * rms_u
        # unbiased: the mean proposal error is consistent with zero
        assert abs(errs_q.mean()) < 4 * errs_q.std(ddof=1) / math.sqrt(errs_q.size) + 2e-4
    finally:
        q.close()


def test_training_graph_replays_match_eager(ctx):
    """The captured training step replayed N times gives the same network as N eager steps
    (bitwise: deterministic kernels, same device-side step state)."""
    rng = np.random.default_rng(8)
    ev12, ev13 = rng.uniform(0, 3, 20_000), rng.uniform(0, 3, 20_000)
    nets = []
    for graph in (False, True):
        q = capi.Proposal(ctx, seed=12, batch_cap=4096)
        q.use_graph(graph)
        l0, l1 = q.train_host(ev12, ev13, BOX, steps=25, batch=2048, lr=3e-3, seed=6)
        nets.append((q.get_params(), l0, l1))
        # a second call continues the Adam step count
        q.train_host(ev12, ev13, BOX, steps=5, batch=2048, lr=1e-3, seed=7)
        nets.append((q.get_params(), 0.0, 0.0))
        q.close()
    assert np.array_equal(nets[0][0], nets[2][0]) and nets[0][1:] == nets[2][1:]
    assert np.array_equal(nets[1][0], nets[3][0])
