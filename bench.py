#!/usr/bin/env python3
"""Benchmark of the likelihood hot path on one MI355X.

Workload: the 16-resonance D0 -> pi+ pi- pi0 Dalitz model (Amp3Body with relativistic
Breit-Wigner and Gounaris-Sakurai lineshapes, spin 0/1/2, Blatt-Weisskopf radius 1.5,
couplings drawn from a seed), 1e6 events generated on the GPU by accept-reject on
|A|^2, normalisation on the 1414 x 1414 (m12, m13) midpoint grid, all FP64.

A step is one calculateNLL() of the GooFit API with new coupling values -- what
MIGRAD asks for at every function call: the normalisation (host quadratic form of
the cached Gram matrix, couplings only) and the coherent-sum NLL over the 1e6
events in HBM.  value = events * steps / wall time of the timed steps.

Also reported: the wall time of a full MIGRAD fit of the 30 free coupling
parameters, the lineshape-cache and Gram kernels, the 1-D AddPdf and 3-D ProdPdf
NLL rates, the roofline of the dominant kernel (coherent-sum NLL, HBM bound) and
a CPU baseline: the oracle (the C restatement of the reference's OMP algorithm)
on a bounded sample on the host cores.

Usage: python bench.py [--gpus 1] [--steps 50] [--warmup 5] [--events 1000000]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s
METRIC = "Dalitz fit wall-time & NLL events/s; norm-integral error at equal MC events"


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def measured_traffic(name="dalitz16_1e6_hbm_counters.md"):
    """HBM bytes per launch of the dominant kernel from the committed rocprofv3 counter summary
    (FETCH_SIZE + WRITE_SIZE, each scaled by the same counter on a 256 MiB copy)."""
    path = os.path.join(ROOT, "profiles", name)
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":")[1]), os.path.relpath(path, ROOT)
    except OSError:
        pass
    return None, None


def time_steps(pdf, params, steps, warmup, sync):
    """Time `steps` FCN calls (Params::from_minuit_vector + calculateNLL), each with
    slightly moved free parameters: exactly what MIGRAD requests per call, made one after
    the other from C++ (FitManager.time_calls) as the minimiser makes them; the vectors
    are built before the timed region."""
    import goofit_amd as G
    fm = G.FitManager(pdf)
    x0 = np.asarray(fm.minuit_vector(), float)
    free = np.asarray(fm.free_mask(), bool)
    idx = np.nonzero(free)[0]

    def vec(k):
        x = x0.copy()
        x[idx] = x0[idx] * (1.0 + 1e-3 * np.sin(0.37 * k + 1.3 * np.arange(idx.size)))
        return x.tolist()

    fm.time_calls([vec(k) for k in range(warmup)])
    timed = [vec(warmup + k) for k in range(steps)]
    sync()
    t0 = time.perf_counter()
    fm.time_calls(timed)
    sync()
    t1 = time.perf_counter()
    fm.evaluate(x0.tolist())
    return t1 - t0


# proposal training: 2e6 events of the model generated on the GPU and kept there, 300 Adam
# steps of 8192 (half events, half uniform points); ~0.05 s in all, relative variance per
# point 0.1-0.6 on reseeded models against 1.5-5 for uniform points
TRAIN = dict(events=2_000_000, steps=300, batch=8192, lr=6e-3)


def grid_integral(model, pdf, nb):
    """The model's normalisation integral on an nb x nb midpoint grid (numbins restored)."""
    o12, o13 = model.observables[0], model.observables[1]
    old = (o12.numbins, o13.numbins)
    o12.numbins, o13.numbins = nb, nb
    pdf.setNormalisationGrid()
    val = pdf.normalize()
    o12.numbins, o13.numbins = old
    pdf.setNormalisationGrid()
    return val


JUDGE_NODES = (1000, 2000, 4000)  # nodes per axis of the judge's product Gauss-Legendre rule
MIDPOINT_GRIDS = (1414, 2000, 2828, 4000, 5657, 8000)  # cells per axis of the midpoint grids beside it


def norm_error_study(model, pdf, sync, run_seed, target=1e-4, seeds=6):
    """Normalisation-integral error at equal Monte-Carlo points, uniform phase-space MC vs the
    learned proposal, judged against a converged deterministic integral: a product Gauss-
    Legendre rule over the Dalitz region in (m12, t) coordinates whose edges are the
    kinematic boundary (workloads.square_dalitz_rule), refined until successive rules agree
    far below the errors measured.  Reported beside it: the midpoint grids of the reference
    method (their boundary staircase leaves errors of order 1e-5 that do not fall smoothly
    with the cell size) and the 1e8-point uniform FP64 Monte-Carlo integral.  Then a MIGRAD
    fit with each method at the point count that reaches `target` relative error, and one on
    the midpoint grid.  The learned method is charged its training-event generation and
    training.  Leaves pdf on its midpoint grid."""
    import goofit_amd as G
    from goofit_amd import workloads as W
    out = {}
    truth = {n: v.value for n, v in model.params.items()}
    grid_norm = pdf.normalize()
    # judge: the boundary-fitted product rule at growing node counts
    ladder = []
    for nodes in JUDGE_NODES:
        t0 = time.perf_counter()
        a, b, w = W.square_dalitz_rule(nodes)
        val = pdf.integralAtPoints(a, b, w)
        sync()
        ladder.append(dict(nodes_per_axis=nodes, points=int(a.size), value=val,
                           seconds=round(time.perf_counter() - t0, 3)))
        del a, b, w
    for x, y in zip(ladder[:-1], ladder[1:]):
        y["vs_previous"] = y["value"] / x["value"] - 1
    ref_norm = ladder[-1]["value"]
    grids = []
    for nb in MIDPOINT_GRIDS:
        grids.append(dict(cells_per_axis=nb, vs_reference=grid_integral(model, pdf, nb) / ref_norm - 1))
    sync()
    out["reference"] = dict(kind="product Gauss-Legendre rule over (m12, t), boundary-fitted",
                            nodes_per_axis=JUDGE_NODES[-1], value=ref_norm, convergence=ladder,
                            est_rel_err=abs(ladder[-1]["vs_previous"]), midpoint_grids=grids,
                            fit_grid_vs_reference=grid_norm / ref_norm - 1)
    # the 1e8-point uniform FP64 Monte-Carlo integral, against the judge
    t0 = time.perf_counter()
    pdf.setNormalisationPoints(100_000_000, 991)
    val = pdf.normalize()
    sync()
    out["reference_1e8"] = dict(points=100_000_000, value=val, seconds=round(time.perf_counter() - t0, 3),
                                rel_stat_err=math.sqrt(pdf.normalisationRelativeVariance(16) / 1e8),
                                vs_judge=val / ref_norm - 1)
    log(f"reference integral: {JUDGE_NODES[-1]}^2-node boundary-fitted rule, last refinement "
        f"{ladder[-1]['vs_previous']:.1e}; midpoint grids vs it: "
        + ", ".join(f"{g['cells_per_axis']}: {g['vs_reference']:.1e}" for g in grids)
        + f"; 1e8 uniform points: {out['reference_1e8']['vs_judge']:.2e} (stat. {out['reference_1e8']['rel_stat_err']:.1e})")

    def errors(n, proposal, base):
        e, v = [], []
        for k in range(seeds):
            pdf.setNormalisationPoints(n, base + k, proposal)
            e.append(pdf.normalize() / ref_norm - 1)
            v.append(pdf.normalisationRelativeVariance(16))
        e = np.array(e)
        return dict(points=n, rms_rel_err=float(np.sqrt((e ** 2).mean())), mean_rel_err=float(e.mean()),
                    rel_var_per_point=float(np.mean(v)))

    uni = [errors(n, None, 1000 + 10 * i) for i, n in enumerate((100_000, 1_000_000))]
    # the learned method's whole cost: generating its training events, training, tabulating
    sync()
    t0 = time.perf_counter()
    q = G.DalitzProposal(model.observables[0], model.observables[1], seed=7)
    l0, l1 = q.trainOnModel(pdf, TRAIN["events"], run_seed + 3, steps=TRAIN["steps"], batch=TRAIN["batch"],
                            lr=TRAIN["lr"], seed=3)
    q.build(512, 0.1)
    sync()
    t_train = time.perf_counter() - t0
    # retraining trigger on: the first points drawn set the reference weight variance; a
    # later draw above 4x it (the model moved away from the training events) retrains
    q.setRetraining(events=TRAIN["events"], steps=TRAIN["steps"], batch=TRAIN["batch"], lr=TRAIN["lr"], grid=512,
                    alpha=0.1, factor=4.0)
    lrn = [errors(n, q, 2000 + 10 * i) for i, n in enumerate((100_000, 1_000_000))]
    out.update(uniform=uni, learned=lrn, train=dict(TRAIN, seconds_incl_generation=round(t_train, 3),
                                                    loss_first=l0, loss_last=l1))
    log(f"norm error at 1e6 points: uniform {uni[1]['rms_rel_err']:.2e}, learned {lrn[1]['rms_rel_err']:.2e} "
        f"(training {t_train:.2f} s)")

    # fits at equal target error
    fits = {}
    names = sorted(model.params)
    for label, var, prop in (("uniform", uni[1]["rel_var_per_point"], None),
                             ("learned", lrn[1]["rel_var_per_point"], q)):
        n = int(min(max(math.ceil(var / target ** 2), 100_000), 1_000_000_000))
        rng = np.random.default_rng(11)
        for k in names:
            model.params[k].value = truth[k] * (1.0 + 0.05 * rng.standard_normal())
        fm = G.FitManager(pdf)
        fm.setVerbosity(0)
        sync()
        t0 = time.perf_counter()
        pdf.setNormalisationPoints(n, 4242, prop)
        m = fm.fit()
        sync()
        wall = time.perf_counter() - t0
        fits[label] = dict(points=n, expected_rel_err=math.sqrt(var / n), wall_s=round(wall, 3), nfcn=m.NFcn(),
                           valid=bool(m.IsValid()), values=list(m.values()), errors=list(m.errors()))
        log(f"{label} fit at {n:.3g} points: {wall:.2f} s, {m.NFcn()} calls")
    # the reference's method from the same start: the midpoint-grid normalisation
    rng = np.random.default_rng(11)
    for k in names:
        model.params[k].value = truth[k] * (1.0 + 0.05 * rng.standard_normal())
    pdf.setNormalisationGrid()
    fm = G.FitManager(pdf)
    fm.setVerbosity(0)
    sync()
    t0 = time.perf_counter()
    m = fm.fit()
    sync()
    wall = time.perf_counter() - t0
    fits["grid"] = dict(cells=pdf_grid_cells(model), wall_s=round(wall, 3), nfcn=m.NFcn(), valid=bool(m.IsValid()),
                        values=list(m.values()), errors=list(m.errors()))
    log(f"grid fit: {wall:.2f} s, {m.NFcn()} calls")
    vg, eg = np.array(fits["grid"]["values"]), np.array(fits["grid"]["errors"])
    vu, vl = np.array(fits["uniform"]["values"]), np.array(fits["learned"]["values"])
    eu = np.array(fits["uniform"]["errors"])
    ok = eu > 0
    fits["max_param_shift_sigma"] = float(np.max(np.abs(vu - vl)[ok] / eu[ok])) if ok.any() else None
    okg = eg > 0
    for label in ("uniform", "learned"):
        shift = float(np.max(np.abs(np.array(fits[label]["values"]) - vg)[okg] / eg[okg])) if okg.any() else None
        fits[label]["max_shift_from_grid_fit_sigma"] = shift
        # the shift is the normalisation's Monte-Carlo fluctuation, which falls as 1/sqrt(points):
        # the point count at which it would reach 0.01 sigma
        if shift is not None and shift > 0:
            fits[label]["points_for_0p01_sigma"] = int(math.ceil(fits[label]["points"] * (shift / 0.01) ** 2))
        log(f"{label} fit: largest shift from the grid fit {shift:.3f} sigma")
    fits["learned_incl_training_and_generation_s"] = round(fits["learned"]["wall_s"] + t_train, 3)
    fits["learned_faster_than_uniform"] = fits["learned_incl_training_and_generation_s"] < fits["uniform"]["wall_s"]
    log(f"learned fit incl. generating its training events and training ({t_train:.3f} s): "
        f"{fits['learned_incl_training_and_generation_s']:.3f} s vs uniform {fits['uniform']['wall_s']:.3f} s")
    for k in names:
        model.params[k].value = truth[k]
    pdf.copyParams()
    pdf.setNormalisationGrid()
    out["fits"] = fits
    out["proposal_retrains"] = q.retrainCount()
    return out


def pdf_grid_cells(model):
    """Cells of the model's (m12, m13) normalisation grid."""
    return int(model.observables[0].numbins) * int(model.observables[1].numbins)


def tddp_workload(args, sync, n=1_000_000, grid=1414):
    """Time-dependent Dalitz NLL: the 16 resonances, three-Gaussian resolution, per-event
    mistag and charm tag, 1e6 events (uniform Dalitz points, exponential decay times);
    a step moves tau, x and y (no cache or integral redo)."""
    import goofit_amd as G
    from goofit_amd import workloads as W
    mdl = W.build_tddp(nbins=grid, resonances=W.DALITZ_RESONANCES, resolution="threegauss", mistag=True,
                       charmtag=True)
    rng = np.random.default_rng(31)
    m12, m13 = W.dalitz_uniform(n, seed=32)
    cols = [rng.exponential(W.TD_MIX[0], n), np.full(n, 0.05), m12, m13, np.arange(n, dtype=np.float64),
            rng.uniform(0.0, 0.3, n), rng.choice([-1.0, 1.0], n)]
    d = G.UnbinnedDataSet(mdl.observables)
    d.from_columns(cols)
    mdl.observables[5].value = 0.1
    pdf = mdl.pdf
    pdf.setData(d)
    pdf.setDataSize(n, 7)
    sync()
    t0 = time.perf_counter()
    pdf.calculateNLL()  # A/B wave caches of 16 resonances + ThreeComplex Gram (6 pair launches) + NLL
    sync()
    t_first = time.perf_counter() - t0
    steps = max(args.steps, 20)
    dtw = time_steps(pdf, mdl.params, steps, 3, sync)
    k_ms = G.engine.time_reduce_metric(1, 50) / 50
    bpe = 2 * 16 * 16 + 6 * 8  # A and B waves of 16 resonances + 6 columns (event number = row)
    out = dict(events=n, resonances=16, resolution="ThreeGaussResolution", mistag=True, charm_tag=True,
               nll_events_per_s=round(n * steps / dtw, 1), ms_per_step=round(dtw / steps * 1e3, 4),
               kernel_ms=round(k_ms, 5), bytes_per_event=bpe, hbm_gbs=round(bpe * n / (k_ms * 1e-3) / 1e9, 1),
               first_nll_ms=round(t_first * 1e3, 2), norm_grid=f"{grid}x{grid}")
    del d
    return out


TOY_RAW = {}  # per-toy pulls, sizes, values and errors of the last toy study (not printed)


def toy_counts(n, seed0):
    """Per-toy event counts: Poisson with mean uniform in [5e3, 8e4], each from its own seed."""
    out = []
    for t in range(n):
        rng = np.random.default_rng(seed0 + t)
        out.append(int(rng.poisson(rng.uniform(5e3, 8e4))))
    return np.array(out, np.int64)


def fit_toys(model, data, counts, truth, sync, seq_toys, threads, seq_seed=4321):
    """Batched MIGRAD fits of every toy, sequential fits of a random subset of seq_toys toys
    (for an Amp3Body the data-independent integrals are computed once and reused, as the
    batch does: only the per-event lineshapes are recomputed per toy); pulls."""
    import goofit_amd as G
    pdf = model.pdf
    names = sorted(model.params)
    for k in names:
        model.params[k].value = truth[k]
    first = np.concatenate([[0], np.cumsum(counts)[:-1]]).astype(np.int64)
    bf = G.BatchFitter(pdf, list(first), list(counts))
    sync()
    t0 = time.perf_counter()
    bf.fit(threads=threads)
    sync()
    t_batch = time.perf_counter() - t0
    st = bf.stats()
    vals, errs, valid = np.array(bf.values()), np.array(bf.errors()), np.array(bf.valid())
    fm0 = G.FitManager(pdf)
    mv = fm0.minuit_vector()
    free = np.array(fm0.free_mask(), bool)
    tv = np.array(mv)  # truth in Minuit order (the Variables hold the truth)
    ok = valid.astype(bool)
    pulls = (vals[ok][:, free] - tv[free]) / errs[ok][:, free]
    pm, pw = pulls.mean(0), pulls.std(0, ddof=1)
    # robust width (inter-quartile range / 1.349), tails, and the toys above the median size
    q75, q25 = np.percentile(pulls, [75, 25], axis=0)
    pw_robust = (q75 - q25) / 1.349
    big = counts[ok] >= np.median(counts)
    pw_big = pulls[big].std(0, ddof=1)
    pm_big, pm_small = pulls[big].mean(0), pulls[~big].mean(0)
    n_tail = int(np.sum(np.any(np.abs(pulls) > 5, axis=1)))
    # sequential fits of a random subset, one data set at a time
    seq_vals, t_seq = [], 0.0
    cols = [data.column(o) for o in model.observables]
    subset = np.sort(np.random.default_rng(seq_seed).choice(len(counts), size=min(seq_toys, len(counts)),
                                                            replace=False))
    for t in subset:
        for k in names:
            model.params[k].value = truth[k]
        d = G.UnbinnedDataSet(model.observables)
        sl = slice(int(first[t]), int(first[t] + counts[t]))
        parts = [c[sl] for c in cols]
        if len(parts) == 3:  # Dalitz: renumber the events of the toy
            parts[2] = np.arange(int(counts[t]), dtype=np.float64)
        d.from_columns(parts)
        sync()
        t0 = time.perf_counter()
        pdf.setData(d)
        if hasattr(pdf, "setDataSize"):
            pdf.setDataSize(int(counts[t]))
            if t != subset[0]:
                pdf.refreshCacheOnly()
        fm = G.FitManager(pdf)
        fm.setVerbosity(0)
        r = fm.fit()
        sync()
        t_seq += time.perf_counter() - t0
        seq_vals.append((np.array(r.values()), np.array(r.errors())))
    shifts = [np.max(np.abs(vals[t][free] - sv[free]) / se[free]) for t, (sv, se) in zip(subset, seq_vals)]
    for k in names:
        model.params[k].value = truth[k]
    pdf.setData(data)
    if hasattr(pdf, "setDataSize"):
        pdf.setDataSize(int(counts.sum()))
    n = len(counts)
    return dict(toys=n, events=int(counts.sum()), free_params=int(free.sum()), valid=int(ok.sum()),
                batched_s=round(t_batch, 3), batched_toys_per_hour=round(n / t_batch * 3600, 1),
                sequential_toys=len(subset), sequential_s=round(t_seq, 3),
                sequential_toys_per_hour=round(len(subset) / t_seq * 3600, 1) if t_seq > 0 else None,
                sequential_events=int(counts[subset].sum()),
                speedup=round((n / t_batch) / (len(subset) / t_seq), 3) if t_seq > 0 else None,
                speedup_per_event=round((counts.sum() / t_batch) / (counts[subset].sum() / t_seq), 3)
                if t_seq > 0 else None,
                two_stages=bool(st.get("two_stages", False)), device_norms=bool(st.get("device_norms", False)),
                launches=st["launches"], fcn_calls=st["fcn_calls"], gpu_wait_s=round(st["gpu_wait_s"], 3),
                host_prep_s=round(st["host_prep_s"], 3), gather_s=round(st["gather_s"], 3),
                chunk=st["chunk"], multi_point=bool(bf.multiPoint()), pull_mean=[round(float(v), 4) for v in pm],
                pull_width=[round(float(v), 4) for v in pw],
                max_abs_pull_mean=round(float(np.abs(pm).max()), 4),
                max_abs_pull_width_minus_1=round(float(np.abs(pw - 1).max()), 4),
                # the same in units of their statistical errors, 1/sqrt(N) and 1/sqrt(2(N-1))
                max_pull_mean_over_error=round(float(np.abs(pm).max() * math.sqrt(len(pulls))), 2),
                max_pull_width_dev_over_error=round(float(np.abs(pw - 1).max() * math.sqrt(2 * (len(pulls) - 1))), 2),
                pull_width_robust=[round(float(v), 4) for v in pw_robust],
                pull_width_large_toys=[round(float(v), 4) for v in pw_big],
                pull_mean_large_toys=[round(float(v), 4) for v in pm_big],
                pull_mean_small_toys=[round(float(v), 4) for v in pm_small],
                toys_with_pull_beyond_5=n_tail,
                batched_vs_sequential_max_shift_sigma=round(float(max(shifts)), 5) if shifts else None,
                _raw=dict(pulls=pulls, counts=counts[ok], vals=vals[ok], errs=errs[ok], truth=tv, free=free))


# Normalisation grid of the Dalitz toy fits.  The fits integrate the density with the
# midpoint rule over the bins whose centres lie inside the Dalitz region, while the toys
# are drawn from the exact density; the staircase along the kinematic boundary leaves a
# normalisation error of order one bin width, which at 400 x 400 shifted the pull means
# of 2e5-event toys by up to 0.1 sigma.  At 1600 x 1600 the shift is a quarter of that,
# and the integrals are computed once per toy batch (lineshapes are fixed), so the finer
# grid costs nothing per fit.
TOY_GRID = 1600


def toy_study(ntoys, sync, seq_toys=128, threads=256, add1d_threads=512, n_add1d=None, n_dalitz=None,
              rule_nodes=1000):
    """cfg: ntoys/2 toys of the 1-D AddPdf model and ntoys/2 of an 8-resonance cut of the Dalitz model
    (or n_add1d and n_dalitz), Poisson-varied sizes, fitted in batches (512 / 256 concurrent Minuit
    instances: the AddPdf fits are GPU-bound and gain from fuller launches) and a random subset of
    seq_toys of each model sequentially."""
    import goofit_amd as G
    from goofit_amd import workloads as W
    half = ntoys // 2
    n_add1d = half if n_add1d is None else n_add1d
    n_dalitz = half if n_dalitz is None else n_dalitz
    out = {}
    # 1-D AddPdf toys
    counts = toy_counts(n_add1d, 10_000)
    t0 = time.perf_counter()
    xs = np.concatenate([W.sample_add1d(int(n), seed=50_000 + t) for t, n in enumerate(counts)])
    t_gen = time.perf_counter() - t0
    m = W.build_add1d(start=1.0)
    truth = {k: v.value for k, v in m.params.items()}
    d = G.UnbinnedDataSet(m.observables)
    d.from_columns([xs])
    m.pdf.setData(d)
    out["add1d"] = fit_toys(m, d, counts, truth, sync, seq_toys, add1d_threads)
    TOY_RAW["add1d"] = out["add1d"].pop("_raw")
    out["add1d"]["generation_s"] = round(t_gen, 2)
    log("add1d toys", {k: out["add1d"][k] for k in ("batched_s", "batched_toys_per_hour",
                                                    "sequential_toys_per_hour", "max_abs_pull_mean",
                                                    "max_abs_pull_width_minus_1")})
    del d
    # 8-resonance Dalitz toys
    counts = toy_counts(n_dalitz, 20_000)
    md = W.build_dalitz(nbins=TOY_GRID, resonances=W.DALITZ_RESONANCES[:8])
    truth = {k: v.value for k, v in md.params.items()}
    t0 = time.perf_counter()
    g12, g13 = W.generate_dalitz(md, int(counts.sum()), seed=77)
    sync()
    t_gen = time.perf_counter() - t0
    dd = W.dalitz_dataset(md, g12, g13)
    md.pdf.setData(dd)
    md.pdf.setDataSize(int(counts.sum()))
    md.pdf.setForceIntegrals()
    md.pdf.calculateNLL()
    out["dalitz8"] = fit_toys(md, dd, counts, truth, sync, seq_toys, threads)
    TOY_RAW["dalitz8"] = out["dalitz8"].pop("_raw")
    out["dalitz8"]["generation_s"] = round(t_gen, 2)
    out["dalitz8"]["normalisation"] = f"midpoint grid {TOY_GRID}x{TOY_GRID}"
    log("dalitz8 toys", {k: out["dalitz8"][k] for k in ("batched_s", "batched_toys_per_hour",
                                                        "sequential_toys_per_hour", "max_abs_pull_mean",
                                                        "max_abs_pull_width_minus_1")})
    if rule_nodes:
        # the same toys normalised by the boundary-fitted product rule (no staircase along the
        # kinematic boundary): separates the grid's normalisation error from the estimator's bias
        md.pdf.setNormalisationRule(*W.square_dalitz_rule(rule_nodes))
        md.pdf.setData(dd)
        md.pdf.setDataSize(int(counts.sum()))
        md.pdf.calculateNLL()
        out["dalitz8_rule"] = fit_toys(md, dd, counts, truth, sync, 16, threads)
        TOY_RAW["dalitz8_rule"] = out["dalitz8_rule"].pop("_raw")
        out["dalitz8_rule"]["normalisation"] = f"product Gauss-Legendre rule, {rule_nodes}^2 nodes"
        log("dalitz8 toys, rule normalisation", {k: out["dalitz8_rule"][k] for k in (
            "batched_s", "max_abs_pull_mean", "max_abs_pull_width_minus_1", "max_pull_mean_over_error")})
        md.pdf.setNormalisationGrid()
    tot_t = out["add1d"]["batched_s"] + out["dalitz8"]["batched_s"]
    # sequential time of the same mix of toys: each model's toys at that model's sequential rate
    seq_t = sum(out[k]["toys"] / out[k]["sequential_toys_per_hour"] for k in ("add1d", "dalitz8")) * 3600
    n_all = n_add1d + n_dalitz
    out["all"] = dict(toys=n_all, batched_s=round(tot_t, 3), batched_toys_per_hour=round(n_all / tot_t * 3600, 1),
                      sequential_toys_per_hour=round(n_all / seq_t * 3600, 1),
                      speedup=round(seq_t / tot_t, 3))
    return out


def pull_study(ntoys, mean_events, sync, threads=512, nbins=1600):
    """Pulls of the 8-resonance Dalitz toys at a larger size: ntoys toys, Poisson sizes of mean
    mean_events, fitted in batches with an nbins x nbins normalisation grid (not part of the
    default run: --pull-study)."""
    import goofit_amd as G
    from goofit_amd import workloads as W
    rng = np.random.default_rng(90_000)
    counts = rng.poisson(mean_events, size=ntoys).astype(np.int64)
    md = W.build_dalitz(nbins=nbins, resonances=W.DALITZ_RESONANCES[:8])
    truth = {k: v.value for k, v in md.params.items()}
    t0 = time.perf_counter()
    g12, g13 = W.generate_dalitz(md, int(counts.sum()), seed=91)
    sync()
    t_gen = time.perf_counter() - t0
    dd = W.dalitz_dataset(md, g12, g13)
    md.pdf.setData(dd)
    md.pdf.setDataSize(int(counts.sum()))
    md.pdf.setForceIntegrals()
    md.pdf.calculateNLL()
    out = fit_toys(md, dd, counts, truth, sync, 4, threads)
    TOY_RAW["dalitz8_large"] = out.pop("_raw")
    out["generation_s"] = round(t_gen, 2)
    out["mean_events"] = mean_events
    out["norm_grid"] = f"{nbins}x{nbins}"
    return out


def host_cpu():
    """The host CPU the baseline ran on: model name, logical CPUs of the machine and of this process."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = None
    return dict(model=model, logical_cpus=os.cpu_count(), usable_cpus=usable,
                omp_num_threads=os.environ.get("OMP_NUM_THREADS"))


def cpu_baseline_dalitz(flat, decay, m12, m13, target_s=10.0):
    """Oracle (C/OpenMP restatement of the reference algorithm) NLL over the same events."""
    import oracle as O
    f, p, c, o, nrm = flat
    prog = O.RawProg(f, p, c, o, nrm, decay)
    n = len(m12)
    aos = np.stack([m12, m13, np.arange(n, dtype=np.float64)], axis=1)
    prog.set_total_events(n)
    res_funcs, _ = prog.dalitz_layout(0)
    slot0 = 16 * int(c[2])  # cResonances[16 * cacheToUse + i]
    t0 = time.perf_counter()
    for i, rf in enumerate(res_funcs):
        prog.set_cache(slot0 + i, prog.dalitz_wave(0, rf, aos))
    t_cache = time.perf_counter() - t0
    t0 = time.perf_counter()
    v = prog.reduce_metric(aos, "NLL")
    t1 = time.perf_counter() - t0
    reps = int(min(max(math.ceil(target_s / max(t1, 1e-6)), 2), 200))
    t0 = time.perf_counter()
    for _ in range(reps):
        v = prog.reduce_metric(aos, "NLL")
    dt = time.perf_counter() - t0
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return dict(value=n * reps / dt, unit="events/s", cores=threads, kind="port", host=host_cpu(),
                sample=f"{reps} NLL evaluations of the {n}-event 16-resonance Dalitz sample (cached lineshapes; "
                       f"lineshape cache build {t_cache:.2f} s not included)",
                nll=v, cache_build_s=t_cache)


def copy_256mib(sync):
    """Three device copies of 256 MiB: the reference dispatch for scaling the HBM counters."""
    import torch
    src = torch.ones(32 << 20, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    sync()


def profile_simple(name, sync):
    """Set up the 1-D AddPdf (1e5 events) or the 3-D ProdPdf (1e7 events) fit model and run 20
    NLL reductions plus the calibration copies, for rocprofv3."""
    import goofit_amd as G
    from goofit_amd import workloads as W
    if name == "prod3d":
        mdl, cols = W.build_prod3d(), list(W.sample_prod3d(10_000_000, seed=2))
    else:
        mdl, cols = W.build_add1d(), [W.sample_add1d(100_000, seed=1)]
    d = G.UnbinnedDataSet(mdl.observables)
    d.from_columns(cols)
    mdl.pdf.setData(d)
    for _ in range(20):
        mdl.pdf.calculateNLL()
    copy_256mib(sync)
    log(f"profile-only run done ({name}): 20 NLL reductions, 3 copies of 256 MiB")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--events", type=int, default=1_000_000)
    ap.add_argument("--grid", type=int, default=1414)
    ap.add_argument("--no-fit", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--no-learned", action="store_true", help="skip the normalisation-error study")
    ap.add_argument("--run-seed", type=int, default=None,
                    help="seed of the reseeded widths and couplings of the normalisation-error study (default: random)")
    ap.add_argument("--toys", type=int, default=4096, help="batched toy fits (0: skip)")
    ap.add_argument("--pull-study", type=int, default=0, metavar="NTOYS",
                    help="also fit NTOYS 8-resonance Dalitz toys of 2e5 events each for their pulls")
    ap.add_argument("--pull-grid", type=int, default=TOY_GRID, help="normalisation grid of the pull study's fits")
    ap.add_argument("--pulls", nargs=2, type=int, default=None, metavar=("N_ADD1D", "N_DALITZ8"),
                    help="pull study at resolution: the toy study's two models with these toy counts (e.g. 16384 8192)")
    ap.add_argument("--profile-only", nargs="?", const="dalitz16", default=None,
                    choices=("dalitz16", "prod3d", "add1d"),
                    help="set up one workload (default dalitz16: the 16-resonance Dalitz model, whose set-up "
                         "also runs the lineshape cache and the Gram once; prod3d: the 3-D ProdPdf of 1e7 "
                         "events; add1d: the 1-D AddPdf of 1e5 events), run 20 NLL reductions and a 256 MiB "
                         "device copy (HBM counter calibration), print nothing else; for rocprofv3 runs")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for {args.gpus}); the path shards by events across GPUs "
              f"but this benchmark runs on a single MI355X.", file=sys.stderr)
        sys.exit(2)

    import torch
    import goofit_amd as G
    from goofit_amd import workloads as W

    G.require_api()
    sync = torch.cuda.synchronize
    if args.profile_only in ("prod3d", "add1d"):
        profile_simple(args.profile_only, sync)
        return

    # ---------------------------------------------------------------- Dalitz
    log("building the 16-resonance Dalitz model")
    model = W.build_dalitz(nbins=args.grid)
    pdf = model.pdf
    t0 = time.perf_counter()
    m12, m13 = W.generate_dalitz(model, args.events, seed=6)
    t_gen = time.perf_counter() - t0
    log(f"generated {len(m12)} events on the GPU in {t_gen:.1f} s")
    data = W.dalitz_dataset(model, m12, m13)
    pdf.setData(data)
    pdf.setDataSize(len(m12))
    pdf.setForceIntegrals()
    sync()
    t0 = time.perf_counter()
    nll0 = pdf.calculateNLL()  # lineshape cache + Gram on the grid + NLL
    sync()
    t_first = time.perf_counter() - t0
    log(f"first NLL {nll0:.6f} ({t_first * 1e3:.1f} ms incl. cache and Gram)")
    if args.profile_only:
        for _ in range(20):
            pdf.calculateNLL()
        copy_256mib(sync)
        log("profile-only run done: 20 NLL reductions, 3 copies of 256 MiB")
        return

    # dominant kernel: the coherent-sum NLL sweep over the events, which adds its
    # workgroups' partials itself (the in-kernel finish, on by default), timed with HIP
    # events on the engine's stream: 200 sweeps launched back to back between two
    # events (the kernel's duration plus the dispatch gap between launches), and
    # 200 whole reductions with events around each sweep (which adds the idle gap
    # before each dispatch); measured first, so that the device is at its working
    # clocks for the timed steps.  The sweep alone (its partials left to a second
    # launch) is timed back to back beside it.
    reps = 200
    fused = G.engine.get_tuning(G.engine.TUNE_FUSED_FINISH) == 1
    ms_b2b = G.engine.time_sweep_back_to_back(1, reps)
    ms_sweep, ms_span = G.engine.time_reduce_sweep(1, reps)
    G.engine.set_tuning(G.engine.TUNE_FUSED_FINISH, 0)
    ms_sweep_only = G.engine.time_sweep_back_to_back(1, reps)
    G.engine.set_tuning(G.engine.TUNE_FUSED_FINISH, 1 if fused else 0)

    dt = time_steps(pdf, model.params, args.steps, args.warmup, sync)
    ms_per_step = dt / args.steps * 1e3
    value = args.events * args.steps / dt
    log(f"{args.steps} NLL steps: {ms_per_step:.4f} ms/step, {value:.4e} events/s")
    ms_kernel = ms_b2b
    # 16 complex waves; event number = row index; m12 and m13 are not read when the cache
    # (filled from these events, 0 outside the Dalitz region) settles inDalitz and the
    # efficiency is constant (GF_TUNE_ZERO_OUTSIDE, on by default)
    zero_outside = G.engine.get_tuning(G.engine.TUNE_ZERO_OUTSIDE) == 1
    bytes_per_event = 16 * 16 + (0 if zero_outside else 16)
    achieved = bytes_per_event * args.events / (ms_kernel * 1e-3) / 1e9
    traffic, traffic_src = measured_traffic() if args.events == 1_000_000 else (None, None)
    achieved_sweep_only = bytes_per_event * args.events / (ms_sweep_only * 1e-3) / 1e9
    roofline = dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(achieved / HBM_PEAK_GBS, 4), traffic=traffic, traffic_source=traffic_src,
                    kernel="k_dalitz_fast (coherent-sum NLL sweep"
                           + (", partials added in the kernel)" if fused else ")"),
                    kernel_ms=round(ms_kernel, 5),
                    kernel_timing="200 sweeps back to back between two HIP events on the engine stream",
                    kernel_ms_within_a_call=round(ms_sweep / reps, 5),
                    reduction_ms=round(ms_span / reps, 5), bytes_per_event=bytes_per_event,
                    bytes_per_launch=bytes_per_event * args.events,
                    kernel_ms_sweep_only=round(ms_sweep_only, 5),
                    frac_sweep_only=round(achieved_sweep_only / HBM_PEAK_GBS, 4))
    log(f"NLL sweep {ms_kernel * 1e3:.1f} us back to back -> {achieved:.0f} GB/s; {ms_sweep / reps * 1e3:.1f} us "
        f"within a call; whole reduction {ms_span / reps * 1e3:.1f} us; the sweep alone "
        f"{ms_sweep_only * 1e3:.1f} us ({achieved_sweep_only:.0f} GB/s)")
    # practical streaming-read rate of this device: torch's own reduction over 256 MiB,
    # timed the same way (events around each launch on the launching stream)
    buf = torch.ones(32 << 20, dtype=torch.float64, device="cuda")
    for _ in range(5):
        buf.sum()
    sync()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(50):  # back to back, as the sweep's launches are timed
        buf.sum()
    e1.record()
    sync()
    tot = e0.elapsed_time(e1)
    del buf
    ref_gbs = (32 << 20) * 8 / (tot / 50 * 1e-3) / 1e9
    roofline["device_read_gbs_torch_sum"] = round(ref_gbs, 1)
    roofline["frac_of_device_read"] = round(achieved / ref_gbs, 4)
    log(f"torch.sum over 256 MiB: {ref_gbs:.0f} GB/s (the sweep reaches {achieved / ref_gbs:.0%} of it)")

    extra = dict(events=len(m12), generation_s=round(t_gen, 2), first_nll_ms=round(t_first * 1e3, 2), nll=nll0)
    grid6 = [0.0, 3.0, float(args.grid), 0.0, 3.0, float(args.grid)]
    ms_gram = G.engine.time_dalitz_gram(pdf, grid6, 3) / 3
    cells = args.grid * args.grid
    extra["gram_ms"] = round(ms_gram, 3)
    extra["gram_cells"] = cells
    extra["gram_fp64_tflops_mfma"] = round(3 * 2 * 16 * 16 * 4 * (cells / 4) / (ms_gram * 1e-3) / 1e12, 3)
    ms_cache = G.engine.time_dalitz_cache(pdf, 5) / 5
    extra["cache_ms"] = round(ms_cache, 3)
    extra["cache_lineshapes_per_s"] = round(16 * args.events / (ms_cache * 1e-3), 1)

    nll_ref = pdf.calculateNLL()  # current parameters and normalisation for the CPU comparison
    flat = G.engine.flattened()

    fit = None
    if not args.no_fit:
        names = sorted(model.params)
        truth = {n: model.params[n].value for n in names}
        err0 = {n: model.params[n].error for n in names}
        rng = np.random.default_rng(11)
        for n in names:
            model.params[n].value = truth[n] * (1.0 + 0.05 * rng.standard_normal())
        fm = G.FitManager(pdf)
        fm.setVerbosity(0)
        sync()
        t0 = time.perf_counter()
        m = fm.fit()
        sync()
        t_fit = time.perf_counter() - t0
        fit = dict(wall_s=round(t_fit, 3), nfcn=m.NFcn(), valid=bool(m.IsValid()), fval=m.Fval(), edm=m.Edm(),
                   free_params=len(names), batched_derivatives=True)
        log(f"MIGRAD fit: {t_fit:.3f} s, {m.NFcn()} calls, valid={m.IsValid()}")
        vals_b = list(m.values())
        # the same fit with one request per function value (the reference's call pattern),
        # from the same start values and step sizes
        rng = np.random.default_rng(11)
        for n in names:
            model.params[n].value = truth[n] * (1.0 + 0.05 * rng.standard_normal())
            model.params[n].error = err0[n]
        fm1 = G.FitManager(pdf)
        fm1.setVerbosity(0)
        fm1.setBatchedDerivatives(False)
        sync()
        t0 = time.perf_counter()
        m1 = fm1.fit()
        sync()
        t_fit1 = time.perf_counter() - t0
        fit["one_point_per_request"] = dict(wall_s=round(t_fit1, 3), nfcn=m1.NFcn(),
                                            identical_values=list(m1.values()) == vals_b)
        log(f"MIGRAD fit with one request per function value: {t_fit1:.3f} s, {m1.NFcn()} calls "
            f"(identical result: {list(m1.values()) == vals_b})")
        for n in names:
            model.params[n].value = truth[n]
            model.params[n].error = err0[n]
        pdf.copyParams()

    norm_study = None
    if not args.no_learned:
        # the learned proposal on the 16-resonance model with widths and couplings reseeded per run
        run_seed = args.run_seed if args.run_seed is not None else int.from_bytes(os.urandom(3), "little")
        log(f"normalisation-integral error study: uniform MC vs learned proposal (model reseeded with {run_seed})")
        m3 = W.build_dalitz(nbins=args.grid, coup_seed=run_seed, width_seed=run_seed + 1)
        d12, d13 = W.generate_dalitz(m3, args.events, seed=run_seed + 2)
        data3 = W.dalitz_dataset(m3, d12, d13)
        m3.pdf.setData(data3)
        m3.pdf.setDataSize(len(d12))
        m3.pdf.setForceIntegrals()
        m3.pdf.calculateNLL()
        norm_study = norm_error_study(m3, m3.pdf, sync, run_seed)
        norm_study["run_seed"] = run_seed
        norm_study["widths"] = {r[0]: r[3] for r in W.reseed_widths(W.DALITZ_RESONANCES, run_seed + 1)}
        del m3, data3
        pdf.setData(data)
        pdf.setDataSize(len(m12))
        pdf.setForceIntegrals()
        pdf.calculateNLL()

    toys = None
    if args.toys > 0:
        log(f"toy study: {args.toys} toys")
        toys = toy_study(args.toys, sync)

    if args.pulls:
        log(f"pull study: {args.pulls[0]} AddPdf toys and {args.pulls[1]} 8-resonance Dalitz toys")
        pl = toy_study(0, sync, n_add1d=args.pulls[0], n_dalitz=args.pulls[1])
        toys = dict(toys or {}, pulls=pl)

    if args.pull_study > 0:
        log(f"pull study: {args.pull_study} toys of 2e5 events, {args.pull_grid}^2 normalisation grid")
        ps = pull_study(args.pull_study, 200_000, sync, nbins=args.pull_grid)
        toys = dict(toys or {}, dalitz8_2e5=ps)
        log("dalitz8 2e5-event toys", {k: ps[k] for k in ("batched_s", "max_abs_pull_mean",
                                                         "max_abs_pull_width_minus_1", "pull_width")})

    workloads = {}
    if not args.no_extra:
        for name, builder, sampler, nevt, bpe in (
                ("add1d_1e5", W.build_add1d, lambda n: [W.sample_add1d(n, seed=1)], 100_000, 8),
                ("prod3d_1e7", W.build_prod3d, lambda n: list(W.sample_prod3d(n, seed=2)), 10_000_000, 24)):
            mdl = builder()
            cols = sampler(nevt)
            d = G.UnbinnedDataSet(mdl.observables)
            d.from_columns(cols)
            mdl.pdf.setData(d)
            mdl.pdf.calculateNLL()
            dtw = time_steps(mdl.pdf, mdl.params, max(args.steps, 20), 3, sync)
            k_ms = G.engine.time_reduce_metric(1, 100) / 100
            workloads[name] = dict(events=nevt, nll_events_per_s=round(nevt * max(args.steps, 20) / dtw, 1),
                                   ms_per_step=round(dtw / max(args.steps, 20) * 1e3, 4), kernel_ms=round(k_ms, 5),
                                   hbm_gbs=round(bpe * nevt / (k_ms * 1e-3) / 1e9, 1))
            # the whole MIGRAD fit of the workload's free parameters from +20 % starting values
            fmdl = builder(start=1.2)
            fd = G.UnbinnedDataSet(fmdl.observables)
            fd.from_columns(cols)
            fmdl.pdf.setData(fd)
            fm = G.FitManager(fmdl.pdf)
            fm.setVerbosity(0)
            sync()
            t0 = time.perf_counter()
            m = fm.fit()
            sync()
            t_fit = time.perf_counter() - t0
            workloads[name]["fit"] = dict(
                wall_s=round(t_fit, 4), nfcn=m.NFcn(), valid=bool(m.IsValid()), start="truth x 1.2",
                values={k: v.value for k, v in fmdl.params.items()},
                errors={k: v.error for k, v in fmdl.params.items()})
            log(name, workloads[name])
            del d, fd
        workloads["tddp16_1e6_threegauss"] = tddp_workload(args, sync)
        log("tddp16_1e6_threegauss", workloads["tddp16_1e6_threegauss"])

    cpu = None
    if not args.no_cpu:
        log("CPU baseline (oracle, OpenMP) on the same events")
        decay = list(W.DALITZ_DECAY)
        cpu = cpu_baseline_dalitz(flat, decay, m12, m13)
        rel = abs(cpu["nll"] - nll_ref) / abs(nll_ref)
        cpu["nll_rel_diff_vs_gpu"] = rel
        log(f"CPU baseline {cpu['value']:.3e} events/s on {cpu['cores']} threads; NLL rel diff {rel:.2e}")
        cpu.pop("nll")

    line = dict(metric=METRIC, value=round(value, 1), unit="events/s", n_gpus=1, steps=args.steps,
                warmup=args.warmup, ms_per_step=round(ms_per_step, 5), higher_is_better=True, scaling="weak",
                vs_baseline=None, dtype="f64", data="synthetic (GPU accept-reject on the model)",
                config=dict(workload="dalitz16_d0_pipipi0", events=args.events, norm_grid=f"{args.grid}x{args.grid}",
                            resonances=16, free_params=30, step="calculateNLL with new couplings"),
                roofline=roofline, cpu_baseline=cpu, fit=fit, dalitz=extra, norm_error=norm_study,
                toys=toys, workloads=workloads)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
