#!/usr/bin/env python3
"""Benchmark of the Dornic / Rietkerk step on one MI355X.

Headline (`value`): the batched FP32 DP quench over 256 mixed-size instances
(128 x 128^2 + 64 x 256^2 + 48 x 512^2 + 16 x 1024^2 = 35,651,584 cells/step;
a ramped 0.72 -> 0.70 over 10^4 steps; b = 2, D = 1, sigma = sqrt 2, dt = 0.1;
rho(t) every step in FP64), timed over the full 10^4-step run with every
rho(t) row copied back to the host.  Reported beside it, each also timed over
its full run: the same sweep in FP64, the plain FP64 single-instance Rietkerk
step (256^2, 2x10^4 steps), and the batched FP32 Rietkerk sweep (256 x 512^2,
R = linspace(0.5, 1.5, 16) x 16 seeds, 5x10^3 steps).  Every FP32 run is
compared with the FP64 run of the same seeds (rho(t) ensemble error).

Roofline: achieved HBM GB/s of the step kernel = algorithmic bytes per launch
(8 B/cell FP32 DP, 24 B/cell FP32 Rietkerk, 16 / 48 in FP64: one read and one
write of every field) / the kernel's average duration, measured with HIP
events on the context stream over a replay of the same trajectory.

Usage: python bench.py [--steps K] [--warmup W] [--gpus 1] [--c5-steps N] [--no-cpu]
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

METRIC = "cell-updates/s + % HBM peak: batched FP32 vs plain FP64, ρ(t) err ≤1e-3"
HBM_PEAK_GBS = 8000.0
SEED = 20251202
# HBM bytes per launch of the FP32 DP step kernels on the C4 batch, per kernel
# variant, from counter runs (profiles/dp_step_f32_c4_quench_hbm_traffic.md):
# FETCH_SIZE x 2.000 and WRITE_SIZE x 1.000 KiB, the factors a plain 256 MiB
# copy needs to read its known size.
FETCH_SCALE, WRITE_SCALE = 2.0, 1.0
# Multi-step variants: (steps per launch, (FETCH_SIZE, WRITE_SIZE) of a K-step
# launch, (FETCH_SIZE, WRITE_SIZE) of a single-step launch of the same tiles),
# means over the absorbing phase of the first 4000 steps.
C4_F32_BLOCKED_TRAFFIC_KIB = {
    "dp_f32_128x32_t256_x4a": (4, (87921.6, 141440.2), (75136.4, 139848.7)),
    "dp_f32_128x24_t256_x8a": (8, (175845.5, 145216.1), (114941.6, 140040.5)),
    "dp_f32_128x24_t256_x8aq": (8, (175843.5, 145216.1), (115013.0, 149042.5)),
}
C4_F32_TRAFFIC_KIB = {   # single-step variant: (FETCH_SIZE, WRITE_SIZE) per dispatch
    "dp_f32_128x32_t256_w6": (75198.8, 139824.1),
    "dp_f32_64x32_t256": (80202.1, 140385.4),
    "dp_f32_128x32_t256": (75199.2, 139824.1),
    "dp_f32_64x64_t256": (85723.7, 139837.5),
    "dp_f32_256x32_t256": (74841.1, 139556.2),
    "dp_f32_64x64_t256_w6": (85720.6, 139837.4),
    "dp_f32_128x16_t256_rm": (110913.2, 140372.5),
    "dp_f32_256x32_t256_rm": (89995.4, 139556.0),
}


def c4_batch(tm, precision):
    sizes = [128] * 128 + [256] * 64 + [512] * 48 + [1024] * 16
    p = tm.DPParams(a=0.72, a_end=0.70, ramp_steps=10_000, b=2.0, D=1.0, sigma=math.sqrt(2.0),
                    dt=0.1, dx=1.0, scheme=tm.DP_EULER, sampler=tm.SAMPLER_EXACT)
    inst = [tm.Instance(L, 0, SEED, k) for k, L in enumerate(sizes)]
    return dict(model=tm.MODEL_DP, precision=precision, params=[p], inst=inst, lo=[0.0], hi=[1.0],
                steps=10_000, nf=1, chunk=5000)


def c3_batch(tm, precision):
    Rs = np.linspace(0.5, 1.5, 16)
    ps = [tm.RietkerkParams(R=float(R)) for R in Rs]
    inst = [tm.Instance(512, k // 16, SEED, k) for k in range(256)]
    return dict(model=tm.MODEL_RIETKERK, precision=precision, params=ps, inst=inst,
                lo=[0.0, 0.0, 0.0], hi=[90.0, 1.0, 1.0], steps=5_000, nf=3, chunk=1000)


def c2_batch(tm, precision):
    return dict(model=tm.MODEL_RIETKERK, precision=precision, params=[tm.RietkerkParams(R=1.0)],
                inst=[tm.Instance(256, 0, SEED, 0)], lo=[0.0, 0.0, 0.0], hi=[90.0, 1.0, 1.0],
                steps=20_000, nf=3)


def c5_batch(tm, precision, steps=100_000):
    return dict(model=tm.MODEL_RIETKERK, precision=precision, params=[tm.RietkerkParams(R=0.75)],
                inst=[tm.Instance(8192, 0, SEED, 0)], lo=[0.0, 0.0, 0.0], hi=[90.0, 1.0, 1.0],
                steps=steps, nf=3, chunk=500)


def c6_batch(tm, model):
    """Vegetation + consumers (FP64): 16 a-values x 4 replicates of 256^2 with the
    grazer (and predator) values of order_2stocDP.cpp (dx = 0.1 km, dt = 0.1 hr,
    perception radius 0.521 km = 5 sites), vegetation from Uniform(0, 20), grazers
    Uniform(0, 2), predators Uniform(0, 0.5), perception fields 0."""
    ps = [tm.DPCParams(a=float(a), b=1e-6) for a in np.linspace(5e-6, 2e-5, 16)]
    inst = [tm.Instance(256, k // 4, SEED, k) for k in range(64)]
    if model == tm.MODEL_DP2:
        lo, hi, steps = [0.0, 0.0, 0.0], [20.0, 2.0, 0.0], 2000
    else:
        lo, hi, steps = [0.0, 0.0, 0.0, 0.0, 0.0], [20.0, 2.0, 0.5, 0.0, 0.0], 1000
    return dict(model=model, precision=tm.F64, params=ps, inst=inst, lo=lo, hi=hi, steps=steps,
                nf=tm.NFIELDS[model], chunk=500)


def movement_leg(tm, L=1024, reps=20):
    """The Rietkerk consumers' movement step (two species, 1024^2, the grazer and
    predator values of the reference's 3-species driver: D = 0.0298 and 0.05221
    km^2/hr, v = 0.45427 and 0.40587 km/hr, sigma = 1 and 3, dt = 0.1 hr, dx = 0.1
    km) through the host-array entry point: wall time per call, PCIe copies
    included; the kernel's own time is in the rocprof trace."""
    rng = np.random.default_rng(SEED)
    consts = [tm.fke_move_constants(0.0298, 0.45427, 1.0, 0.1, 0.1), tm.fke_move_constants(0.05221, 0.40587, 3.0, 0.1, 0.1)]
    rho = rng.uniform(0, 2, (2, L, L)) * (rng.random((2, L, L)) < 0.5)
    gam = rng.random((2, L, L))
    th = 2 * np.pi * rng.random((2, L, L))
    speed = np.array([0.45427, 0.40587])[:, None, None]
    v = np.stack([speed * np.cos(th), speed * np.sin(th)], axis=-1)
    args = ([c[0] for c in consts], [c[1] for c in consts], [c[2] for c in consts])
    tm.fke_movement(rho, gam, v, *args)       # warm-up
    t0 = time.perf_counter()
    for _ in range(reps):
        tm.fke_movement(rho, gam, v, *args)
    wall = (time.perf_counter() - t0) / reps
    return {"cells": 2 * L * L, "calls": reps, "ms_per_call": wall * 1e3, "species": [c[0] for c in consts],
            "cell_moves_per_s": 2 * L * L / wall}


def cells_of(b):
    return sum(i.L * i.L for i in b["inst"])


def bytes_per_cell(b):
    return b["nf"] * 2 * (8 if b["precision"] == 0 else 4)


def _kib_to_bytes(t):
    return (t[0] * FETCH_SCALE + t[1] * WRITE_SCALE) * 1024


def traffic_bytes(variant, single_steps=0, blocked_steps=0):
    """Measured HBM bytes per step of the timed run: per dispatch for a
    single-step variant; for a multi-step one, its two kinds of launch weighed by
    the steps each advanced."""
    if variant in C4_F32_BLOCKED_TRAFFIC_KIB:
        k, tk, t1 = C4_F32_BLOCKED_TRAFFIC_KIB[variant]
        n = single_steps + blocked_steps
        if n == 0:
            return None
        return int(round((single_steps * _kib_to_bytes(t1) + blocked_steps / k * _kib_to_bytes(tk)) / n))
    t = C4_F32_TRAFFIC_KIB.get(variant)
    return None if t is None else int(round(_kib_to_bytes(t)))


def run_workload(tm, b, steps, warmup, want_rows=True, tune=True, variant_name=None):
    """Warm up, then time `steps` steps of the real trajectory from t = 0: wall
    clock around the whole job, and the device time of every step launch (HIP
    events on the context stream inside each run() call, summed over chunks).
    The launch shape comes from the autotuner, or is the named variant."""
    with tm.Context(0) as ctx:
        ctx.configure(b["model"], b["precision"], b["params"], b["inst"])
        variant = None
        if variant_name:
            v = tm.variant_names(b["model"], b["precision"]).index(variant_name)
            ctx.set_launch(v, 100)
            variant = (v, 100)
        ctx.init_uniform(b["lo"], b["hi"])
        if not variant_name and tune:   # tuned on the initial state of the run
            variant = ctx.autotune(os.environ.get("TM_TUNE_CACHE"))
        if warmup:
            ctx.run(warmup, rows=want_rows)
        ctx.reset()
        ctx.init_uniform(b["lo"], b["hi"])
        chunk = max(1, min(steps, b.get("chunk", steps)))
        parts = []
        dev_ms = 0.0
        m0 = ctx.mode_steps()
        t0 = time.perf_counter()
        done = 0
        while done < steps:
            n = min(chunk, steps - done)
            parts.append(ctx.run(n, rows=want_rows))
            dev_ms += ctx.last_run_ms()
            done += n
            if done < steps:
                print(f"[bench]   ... {done}/{steps} steps", file=sys.stderr, flush=True)
        t1 = time.perf_counter()
        m1 = ctx.mode_steps()
        rows = np.concatenate(parts, axis=1) if want_rows else None
        final = [ctx.download(k) for k in range(min(len(b["inst"]), 4))]
        names = tm.variant_names(b["model"], b["precision"])
        vname = names[variant[0]] if variant else None
    cells = cells_of(b)
    wall = t1 - t0
    avg_step_s = dev_ms / 1e3 / steps
    achieved = cells * bytes_per_cell(b) / avg_step_s / 1e9
    return dict(value=cells * steps / wall, wall_s=wall, ms_per_step=wall / steps * 1e3,
                kernel_ms=avg_step_s * 1e3, achieved_gbs=achieved, frac=achieved / HBM_PEAK_GBS,
                cells=cells, steps=steps, variant=vname, graph_steps=variant[1] if variant else None,
                single_steps=m1[0] - m0[0], blocked_steps=m1[1] - m0[1], rows=rows, final=final)


def sweep_leg(tm, steps=None):
    """The sweep pieces on the C3 batch (FP32 Rietkerk, 16 R x 16 replicates of
    512^2): device frames, a measured run over the reference's log-spaced times,
    the replicate statistics, and the output layer on one instance."""
    import tempfile
    b = c3_batch(tm, tm.F32)
    dt = b["params"][0].dt
    t_meas = tm.sweep_times(25.0, dt)                       # 5000 steps of dt = 0.005
    meas = tm.measure_steps(t_meas, dt)
    if steps:   # a bounded run: the points before `steps`, then `steps` itself
        meas = np.append(meas[meas < steps], np.int64(steps))
    cells = cells_of(b)
    out = {"measurements": int(meas.size), "steps": int(meas[-1])}
    with tm.Context(0) as ctx:
        ctx.configure(b["model"], b["precision"], b["params"], b["inst"])
        t0 = time.perf_counter()
        ctx.init_gauss([5.0, 1.0, 1.0], 1, [8.0, 16.0])
        t1 = time.perf_counter()
        ctx.init_speckle(0.015, [90.0, 1.0, 1.0], [0.0, 0.2, 0.2])
        t2 = time.perf_counter()
        out["init_gauss_gbs"] = cells * 3 * 4 / (t1 - t0) / 1e9
        out["init_speckle_gbs"] = cells * 3 * 4 / (t2 - t1) / 1e9
        ctx.init_uniform(b["lo"], b["hi"])
        t0 = time.perf_counter()
        ctx.run_series(meas, copy=False)
        t1 = time.perf_counter()
        out["series_value"] = cells * int(meas[-1]) / (t1 - t0)
        out["series_step_device_ms"] = ctx.last_run_ms() / int(meas[-1])
        t0 = time.perf_counter()
        st = ctx.replicate_stats(tm.STATS_REFERENCE)
        t1 = time.perf_counter()
        out["replicate_stats_ms"] = (t1 - t0) * 1e3
        out["surviving_at_end"] = st[:, -1, 4].tolist()
        with tempfile.TemporaryDirectory() as d:
            t0 = time.perf_counter()
            ctx.save_snapshot(0, os.path.join(d, "s.tms"))
            t1 = time.perf_counter()
            ctx.write_frame_csv(0, os.path.join(d, "f.csv"), b["params"][0].R, tm.FRAME_HEADER[b["model"]])
            t2 = time.perf_counter()
            out["snapshot_mbs"] = 512 * 512 * 3 * 4 / (t1 - t0) / 1e6
            out["frame_csv_rows_per_s"] = 512 * 512 / (t2 - t1)
    return out


def _kernel_desc(variant):
    """The kernels an adaptive multi-step variant runs, from its name (dp_f32_TXxTY_tNT_xKa)."""
    v = variant or ""
    if v in C4_F32_BLOCKED_TRAFFIC_KIB:
        shape = v.split("_")[2]
        tx, ty = shape.split("x")
        k = C4_F32_BLOCKED_TRAFFIC_KIB[v][0]
        one = "two-pass (deferred Gamma rejections) " if v.endswith("q") else ""
        return (f"dp_stepk_kernel<float,{tx},{ty},256,{k}> ({k} steps per launch) once <= 2% of cells are "
                f"occupied, {one}dp_step_kernel<float,{tx},{ty},256> single steps before")
    return "dp_step_kernel<float>"


def roofline(h, b):
    """The headline's roofline: algorithmic bytes (one read and one write of the
    field per cell-update) per step over the mean device time per step, and the
    measured HBM bytes per step of the same run.  With temporal blocking (the
    _xK kernels read and write the field once per K steps) the algorithmic rate
    exceeds the HBM peak; traffic_frac is then the HBM utilisation."""
    traffic = traffic_bytes(h["variant"], h["single_steps"], h["blocked_steps"])
    step_s = h["kernel_ms"] / 1e3
    out = {"bound": "hbm", "achieved": h["achieved_gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": h["frac"], "traffic": traffic,
           "kernel": _kernel_desc(h["variant"]),
           "bytes_per_cell": bytes_per_cell(b), "cells_per_launch": cells_of(b),
           "step_device_ms": h["kernel_ms"], "single_steps": h["single_steps"],
           "blocked_steps": h["blocked_steps"],
           "timing": "HIP events on the context stream around every run() of the timed region; "
                     "the per-step share also holds the per-block constants, rho(t) fold and "
                     "counter bump launches"}
    if traffic is not None:
        out["traffic_gbs"] = traffic / step_s / 1e9
        out["traffic_frac"] = out["traffic_gbs"] / HBM_PEAK_GBS
        out["traffic_note"] = ("traffic: measured HBM bytes per step (FETCH_SIZE x 2 + WRITE_SIZE, per launch "
                               "kind, weighed by the steps each kind advanced in this run)")
    return out


def rho_error(r32, r64, nf):
    """Errors of FP32 rho(t) against the FP64 run of the same seeds, per field.
    ensemble: mean over instances of the spatial mean, per step -- relative error
    max / mean over all steps with a nonzero FP64 ensemble, and over the steps
    where the ensemble still holds >= 10% of its initial density (later, same-seed
    trajectories decorrelate cluster by cluster and the difference is sampling
    noise); bias_z_max: largest |paired mean difference| / standard error over
    steps (a systematic FP32 bias shows up here); survival: max difference of the
    fraction of instances with a live cell."""
    out = {}
    names = "rho" if nf == 1 else "PWO"
    for f in range(nf):
        c = 2 * f + 1
        m32, m64 = r32[:, :, c].mean(axis=0), r64[:, :, c].mean(axis=0)
        nz = np.abs(m64) > 0
        rel = np.abs(m32[nz] - m64[nz]) / np.abs(m64[nz])
        pop = m64 >= 0.1 * m64[0]
        relp = np.abs(m32[pop] - m64[pop]) / np.abs(m64[pop])
        d = r32[:, :, c] - r64[:, :, c]
        se = d.std(axis=0, ddof=1) / np.sqrt(max(d.shape[0], 1))
        zb = np.abs(d.mean(axis=0)) / np.where(se > 0, se, np.inf)
        s32 = (r32[:, :, 2 * f] > 0).mean(axis=0)
        s64 = (r64[:, :, 2 * f] > 0).mean(axis=0)
        out[names[f]] = dict(
            ensemble_rel_max=float(rel.max()) if rel.size else 0.0,
            ensemble_rel_mean=float(rel.mean()) if rel.size else 0.0,
            ensemble_rel_max_populated=float(relp.max()) if relp.size else 0.0,
            populated_steps=int(pop.sum()),
            bias_z_max=float(zb.max()) if zb.size else 0.0,
            survival_diff_max=float(np.abs(s32 - s64).max()))
    return out


def field_error(f32, f64):
    a = np.concatenate([x.ravel() for x in f32])
    b = np.concatenate([x.ravel() for x in f64])
    d = np.abs(a - b)
    return dict(max_abs=float(d.max()), mean_abs=float(d.mean()))


def cpu_baseline(max_threads=16):
    """The reference's own DP-Dornic integrator (percolationDornic_2D, built from
    its source into oracle/_ref) on the host cores: one independent call per
    thread (the reference's thread-per-parameter OpenMP pattern), L = 128,
    dt = 0.1, a = 0.72, b = 2, D = 1, sigma = sqrt 2, 2 replicates x 301 steps per
    thread.  Falls back to the oracle port when the reference library is absent."""
    from oracle import oracle
    threads = max(1, min(os.cpu_count() or 1, max_threads, 8))
    L, r, t_max, dt = 128, 2, 30.0, 0.1
    steps = int(round(t_max / dt)) + 1
    rho0 = oracle.init_uniform(L, [0.0], [1.0], SEED, 0).reshape(-1).copy()
    ref = oracle.ref_lib()
    if ref is not None:
        tmeas = np.array([2.0, 5.0, 10.0, 20.0, 30.0])
        fd = os.dup(1)
        devnull = os.open(os.devnull, os.O_WRONLY)
        os.dup2(devnull, 1)
        try:
            t0 = time.perf_counter()
            ref.ref_ssp_parallel(threads, L, 0.72, 2.0, 1.0, math.sqrt(2.0), dt, 1.0, r, rho0, tmeas,
                                 len(tmeas), t_max)
            wall = time.perf_counter() - t0
        finally:
            os.dup2(fd, 1)
            os.close(fd)
            os.close(devnull)
        work = threads * r * steps * L * L
        return dict(value=work / wall, unit="cell-updates/s", cores=threads, kind="reference",
                    sample=f"percolationDornic_2D (stochasticSPDP.cpp:940) x {threads} threads, "
                           f"L={L}, {r} replicates x {steps} steps each, a=0.72 b=2 D=1 sigma=sqrt2 dt=0.1",
                    wall_s=wall)
    p = oracle.dp_params(scheme=0)
    t0 = time.perf_counter()
    oracle.dp_run(p, rho0, L, SEED, 0, 0, steps, rows=True)
    wall = time.perf_counter() - t0
    return dict(value=L * L * steps / wall, unit="cell-updates/s", cores=1, kind="port",
                sample=f"CPU restatement, L={L}, {steps} steps, 1 thread", wall_s=wall)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed steps of the headline sweep (default: full 10^4)")
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--c5-steps", type=int, default=0,
                    help="also time the 8192^2 Rietkerk lattice for this many steps (its full run is 10^5)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-side", action="store_true", help="headline only")
    ap.add_argument("--variant", default=None,
                    help="kernel variant for the headline (or --only) run instead of the autotuner's choice")
    ap.add_argument("--only", choices=["c2", "c3", "c3-f64", "c4-f64", "c5", "sweep", "dp2", "dp3", "movement"],
                    default=None,
                    help="only this side workload, --steps steps (default: its full length), one JSON line")
    ap.add_argument("--headline-only", action="store_true",
                    help="only the FP32 headline run (no FP64 comparison, side workloads or CPU baseline)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run 5 plain device-to-device copies of MIB MiB (counter calibration)")
    args = ap.parse_args()
    if args.copy_probe > 0:
        import torch
        x = torch.empty(args.copy_probe * (1 << 20) // 4, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(5):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": x.numel() * 4, "copies": 5}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); instances are independent, "
              "so a batch shards across GPUs with no communication.", file=sys.stderr)
        sys.exit(2)

    import trophic_maelstorm_amd as tm

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    if args.only == "movement":
        print(json.dumps({"workload": "fke_movement", **movement_leg(tm)}), flush=True)
        return
    if args.only == "sweep":
        print(json.dumps({"workload": "c3_sweep", **sweep_leg(tm, args.steps)}), flush=True)
        return
    if args.only:
        b = {"c2": lambda: c2_batch(tm, tm.F64), "c3": lambda: c3_batch(tm, tm.F32),
             "c4-f64": lambda: c4_batch(tm, tm.F64),
             "c3-f64": lambda: c3_batch(tm, tm.F64), "c5": lambda: c5_batch(tm, tm.F32),
             "dp2": lambda: c6_batch(tm, tm.MODEL_DP2), "dp3": lambda: c6_batch(tm, tm.MODEL_DP3)}[args.only]()
        n = args.steps or b["steps"]
        log(f"[bench] {args.only}: {cells_of(b)} cells/step x {n} steps")
        r = run_workload(tm, b, n, args.warmup, variant_name=args.variant)
        print(json.dumps({"workload": args.only, "value": r["value"], "unit": "cell-updates/s", "steps": n,
                          "ms_per_step": r["ms_per_step"], "step_device_ms": r["kernel_ms"],
                          "achieved_gbs": r["achieved_gbs"], "frac": r["frac"], "variant": r["variant"]}),
              flush=True)
        return

    c4 = c4_batch(tm, tm.F32)
    K = args.steps or c4["steps"]
    log(f"[bench] C4 FP32 batched DP quench: {cells_of(c4)} cells/step x {K} steps")
    h = run_workload(tm, c4, K, args.warmup, variant_name=args.variant)
    log(f"[bench]   {h['value']:.4g} cell-updates/s, kernel {h['kernel_ms']:.4f} ms, "
        f"{h['achieved_gbs']:.0f} GB/s ({100 * h['frac']:.1f}% of peak)")
    if args.headline_only:
        print(json.dumps({"metric": METRIC, "value": h["value"], "unit": "cell-updates/s", "steps": K,
                          "ms_per_step": h["ms_per_step"], "step_device_ms": h["kernel_ms"],
                          "achieved_gbs": h["achieved_gbs"], "variant": h["variant"],
                          "roofline": roofline(h, c4)}), flush=True)
        return
    c4d = c4_batch(tm, tm.F64)
    log("[bench] C4 FP64 (same seeds)")
    h64 = run_workload(tm, c4d, K, args.warmup)
    log(f"[bench]   {h64['value']:.4g} cell-updates/s, {h64['achieved_gbs']:.0f} GB/s")
    errors = {"c4": dict(rho=rho_error(h["rows"], h64["rows"], 1), fields=field_error(h["final"], h64["final"]))}
    side = {"c4_fp64": {k: h64[k] for k in ("value", "ms_per_step", "kernel_ms", "achieved_gbs", "frac", "variant")}}

    if not args.no_side:
        c2 = c2_batch(tm, tm.F64)
        log("[bench] C2 plain FP64 Rietkerk 256^2")
        r2 = run_workload(tm, c2, c2["steps"], args.warmup)
        log(f"[bench]   {r2['value']:.4g} cell-updates/s, kernel {r2['kernel_ms']:.4f} ms")
        side["c2_plain_fp64"] = {k: r2[k] for k in ("value", "ms_per_step", "kernel_ms", "achieved_gbs", "frac", "variant")}
        c3 = c3_batch(tm, tm.F32)
        log("[bench] C3 FP32 batched Rietkerk 256 x 512^2")
        r3 = run_workload(tm, c3, c3["steps"], args.warmup)
        side["c3_batched_fp32"] = {k: r3[k] for k in ("value", "ms_per_step", "kernel_ms", "achieved_gbs", "frac", "variant")}
        log(f"[bench]   {r3['value']:.4g} cell-updates/s, {r3['achieved_gbs']:.0f} GB/s")
        c3d = c3_batch(tm, tm.F64)
        log("[bench] C3 FP64 (same seeds)")
        r3d = run_workload(tm, c3d, c3["steps"], args.warmup)
        log(f"[bench]   {r3d['value']:.4g} cell-updates/s, {r3d['achieved_gbs']:.0f} GB/s")
        side["c3_fp64"] = {k: r3d[k] for k in ("value", "ms_per_step", "kernel_ms", "achieved_gbs", "frac", "variant")}
        errors["c3"] = dict(rho=rho_error(r3["rows"], r3d["rows"], 3), fields=field_error(r3["final"], r3d["final"]))
        side["ratio_c4_fp32_over_c2_plain_fp64"] = h["value"] / r2["value"]
        side["ratio_c3_fp32_over_c2_plain_fp64"] = r3["value"] / r2["value"]
        log("[bench] sweep pieces on the C3 batch (frames, series, replicate stats, output)")
        side["c3_sweep"] = sweep_leg(tm)
        log(f"[bench]   series {side['c3_sweep']['series_value']:.4g} cell-updates/s, "
            f"stats {side['c3_sweep']['replicate_stats_ms']:.2f} ms")
        for name, model in (("c6_consumers_dp2_fp64", tm.MODEL_DP2), ("c6_consumers_dp3_fp64", tm.MODEL_DP3)):
            c6 = c6_batch(tm, model)
            log(f"[bench] {name}: 64 x 256^2, {c6['steps']} steps")
            r6 = run_workload(tm, c6, c6["steps"], args.warmup, want_rows=True)
            side[name] = {k: r6[k] for k in ("value", "ms_per_step", "kernel_ms", "achieved_gbs", "frac", "variant",
                                            "steps")}
            log(f"[bench]   {r6['value']:.4g} cell-updates/s, kernel {r6['kernel_ms']:.4f} ms")
        if args.c5_steps > 0:
            c5 = c5_batch(tm, tm.F32, args.c5_steps)
            log(f"[bench] C5 FP32 Rietkerk 8192^2, {args.c5_steps} of its 10^5 steps")
            r5 = run_workload(tm, c5, c5["steps"], args.warmup)
            side["c5_fp32"] = {k: r5[k] for k in ("value", "ms_per_step", "kernel_ms", "achieved_gbs", "frac",
                                                 "variant", "steps")}

    cpu = None if args.no_cpu else cpu_baseline()
    out = {
        "metric": METRIC,
        "value": h["value"],
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": h["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: i.i.d. Uniform initial fields from Philox(seed, instance)",
        "config": {"workload": "C4 batched FP32 DP quench: 128x128^2 + 64x256^2 + 48x512^2 + 16x1024^2, "
                               "a 0.72->0.70 over 10^4 steps, rho(t) FP64 every step",
                   "instances": len(c4["inst"]), "cells_per_step": cells_of(c4), "steps": K,
                   "kernel_variant": h["variant"], "graph_steps": h["graph_steps"]},
        "roofline": roofline(h, c4),
        "cpu_baseline": cpu,
        "side_by_side": side,
        "error_vs_fp64": errors,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
