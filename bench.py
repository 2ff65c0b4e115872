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
    slightly moved free parameters: exactly what MIGRAD requests per call."""
    import goofit_amd as G
    fm = G.FitManager(pdf)
    x0 = np.asarray(fm.minuit_vector(), float)
    free = np.asarray(fm.free_mask(), bool)
    idx = np.nonzero(free)[0]

    def vec(k):
        x = x0.copy()
        x[idx] = x0[idx] * (1.0 + 1e-3 * np.sin(0.37 * k + 1.3 * np.arange(idx.size)))
        return x.tolist()

    for k in range(warmup):
        fm.evaluate(vec(k))
    sync()
    t0 = time.perf_counter()
    for k in range(steps):
        fm.evaluate(vec(warmup + k))
    sync()
    t1 = time.perf_counter()
    fm.evaluate(x0.tolist())
    return t1 - t0


def norm_error_study(model, pdf, train_events, sync, target=1e-4, seeds=6):
    """Normalisation-integral error at equal Monte-Carlo points, uniform phase-space MC vs the
    learned proposal, against a 1e8-point uniform FP64 judge; then a MIGRAD fit with each method
    at the point count that reaches `target` relative error.  Leaves pdf on its midpoint grid."""
    import goofit_amd as G
    from goofit_amd import workloads as W
    out = {}
    truth = {n: v.value for n, v in model.params.items()}
    grid_norm = pdf.normalize()
    t0 = time.perf_counter()
    pdf.setNormalisationPoints(100_000_000, 991)
    judge = pdf.normalize()
    sync()
    out["judge"] = dict(points=100_000_000, seconds=round(time.perf_counter() - t0, 3),
                        rel_stat_err=math.sqrt(pdf.normalisationRelativeVariance(16) / 1e8),
                        grid_vs_judge=grid_norm / judge - 1)
    log(f"judge: 1e8 uniform points, grid/judge - 1 = {grid_norm / judge - 1:.2e}")

    def errors(n, proposal, base):
        e, v = [], []
        for k in range(seeds):
            pdf.setNormalisationPoints(n, base + k, proposal)
            e.append(pdf.normalize() / judge - 1)
            v.append(pdf.normalisationRelativeVariance(16))
        e = np.array(e)
        return dict(points=n, rms_rel_err=float(np.sqrt((e ** 2).mean())), mean_rel_err=float(e.mean()),
                    rel_var_per_point=float(np.mean(v)))

    uni = [errors(n, None, 1000 + 10 * i) for i, n in enumerate((100_000, 1_000_000))]
    t0 = time.perf_counter()
    q = G.DalitzProposal(model.observables[0], model.observables[1], seed=7)
    l0, l1 = q.train(train_events, steps=2000, batch=16384, lr=3e-3, seed=3)
    q.build(512, 0.1)
    sync()
    t_train = time.perf_counter() - t0
    lrn = [errors(n, q, 2000 + 10 * i) for i, n in enumerate((100_000, 1_000_000))]
    q.setReferenceVariance(lrn[-1]["rel_var_per_point"])
    out.update(uniform=uni, learned=lrn, train=dict(events=len(train_events), steps=2000, batch=16384,
                                                    seconds=round(t_train, 3), loss_first=l0, loss_last=l1))
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
    vu, vl = np.array(fits["uniform"]["values"]), np.array(fits["learned"]["values"])
    eu = np.array(fits["uniform"]["errors"])
    ok = eu > 0
    fits["max_param_shift_sigma"] = float(np.max(np.abs(vu - vl)[ok] / eu[ok])) if ok.any() else None
    fits["learned_incl_training_s"] = round(fits["learned"]["wall_s"] + t_train, 3)
    for k in names:
        model.params[k].value = truth[k]
    pdf.copyParams()
    pdf.setNormalisationGrid()
    out["fits"] = fits
    return out


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
    return dict(value=n * reps / dt, unit="events/s", cores=threads, kind="port",
                sample=f"{reps} NLL evaluations of the {n}-event 16-resonance Dalitz sample (cached lineshapes; "
                       f"lineshape cache build {t_cache:.2f} s not included)",
                nll=v, cache_build_s=t_cache)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--events", type=int, default=1_000_000)
    ap.add_argument("--grid", type=int, default=1414)
    ap.add_argument("--no-fit", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--no-learned", action="store_true", help="skip the normalisation-error study")
    ap.add_argument("--profile-only", action="store_true",
                    help="set up the Dalitz workload, run 20 NLL reductions and a 256 MiB device copy "
                         "(HBM counter calibration), print nothing else; for rocprofv3 runs")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for {args.gpus}); the path shards by events across GPUs "
              f"but this build runs on a single MI355X.", file=sys.stderr)
        sys.exit(2)

    import torch
    import goofit_amd as G
    from goofit_amd import workloads as W

    G.require_api()
    sync = torch.cuda.synchronize

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
        src = torch.ones(32 << 20, dtype=torch.float64, device="cuda")  # 256 MiB
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        sync()
        log("profile-only run done: 20 NLL reductions, 3 copies of 256 MiB")
        return

    dt = time_steps(pdf, model.params, args.steps, args.warmup, sync)
    ms_per_step = dt / args.steps * 1e3
    value = args.events * args.steps / dt
    log(f"{args.steps} NLL steps: {ms_per_step:.4f} ms/step, {value:.4e} events/s")

    # dominant kernel: the coherent-sum NLL sweep over the events, timed with HIP
    # events on the engine's stream around each sweep launch (the one-workgroup
    # kernel that adds the per-workgroup partials is timed separately)
    reps = 200
    ms_sweep, ms_span = G.engine.time_reduce_sweep(1, reps)
    ms_kernel = ms_sweep / reps
    bytes_per_event = 16 * 16 + 16  # 16 complex waves + (m12, m13); event number = row index
    achieved = bytes_per_event * args.events / (ms_kernel * 1e-3) / 1e9
    traffic, traffic_src = measured_traffic() if args.events == 1_000_000 else (None, None)
    roofline = dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(achieved / HBM_PEAK_GBS, 4), traffic=traffic, traffic_source=traffic_src,
                    kernel="k_dalitz_fast (coherent-sum NLL sweep)", kernel_ms=round(ms_kernel, 5),
                    reduction_ms=round(ms_span / reps, 5), bytes_per_event=bytes_per_event,
                    bytes_per_launch=bytes_per_event * args.events)
    log(f"NLL sweep {ms_kernel * 1e3:.1f} us -> {achieved:.0f} GB/s; whole reduction {ms_span / reps * 1e3:.1f} us")

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
                   free_params=len(names))
        log(f"MIGRAD fit: {t_fit:.2f} s, {m.NFcn()} calls, valid={m.IsValid()}")
        for n in names:
            model.params[n].value = truth[n]
        pdf.copyParams()

    norm_study = None
    if not args.no_learned:
        log("normalisation-integral error study: uniform MC vs learned proposal")
        t12, t13 = W.generate_dalitz(model, 2_000_000, seed=8)
        train_events = W.dalitz_dataset(model, t12, t13)
        pdf.setData(data)
        pdf.setDataSize(len(m12))
        pdf.setForceIntegrals()
        norm_study = norm_error_study(model, pdf, train_events, sync)
        pdf.setForceIntegrals()
        pdf.calculateNLL()

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
            log(name, workloads[name])
            del d

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
                workloads=workloads)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
