#!/usr/bin/env python3
"""Benchmark of the Dornic / Rietkerk step on one MI355X.

Headline (`value`): the batched FP32 DP quench over 256 mixed-size instances
(128 x 128^2 + 64 x 256^2 + 48 x 512^2 + 16 x 1024^2 = 35,651,584 cells/step;
a ramped 0.72 -> 0.70 over 10^4 steps; b = 2, D = 1, sigma = sqrt 2, dt = 0.1;
rho(t) every step in FP64), timed over its first --steps steps from the
i.i.d. Uniform(0, 1) initial field, with every rho(t) row copied back to the
host.  Those steps are the live start of the quench: every cell is occupied and
takes a Poisson and a Gamma draw each step.

Reported beside it:
  * the whole 10^4-step C4 run in FP32 and FP64, timed per launch block, split
    into its live phase (single-step launches) and its absorbing phase (eight
    steps per launch), each with its own rate and HBM fraction;
  * the plain FP64 single-instance Rietkerk step (C2: 256^2, 2x10^4 steps);
  * the batched Rietkerk sweep (C3: 256 x 512^2, R = linspace(0.5, 1.5, 16) x 16
    seeds, 5x10^3 steps) in FP32 and FP64;
  * the 8192^2 Rietkerk lattice (C5, R = 0.75) for a bounded number of its 10^5
    steps in FP32 and FP64, with the cost of one snapshot (it takes one every
    1000 steps);
  * FP32-vs-FP64 error of every batch over its full run (C5: its bounded run);
  * the CPU baseline: the reference's own DP integrator (percolationDornic_2D,
    built from its source) on C1 (128^2, 8 replicates, up to 10^4 steps) on 8
    host threads and on 1, and the oracle port on C2 on 1 thread.

Roofline: achieved HBM GB/s of the dominant kernel of the timed region =
algorithmic bytes per launch (one read and one write of every field per
cell-update: 8 B/cell FP32 DP, 24 B/cell FP32 Rietkerk, 16 / 48 in FP64) / its
mean launch duration, from HIP events on the context stream.  A launch of K
steps moves the field once: its algorithmic bytes are one read and one write of
the field, and its cell-updates are K per cell.  `traffic` is the measured HBM
bytes per launch from rocprofv3 counter runs of the same kernel and workload.

Usage: python bench.py [--steps K] [--warmup W] [--gpus 1] [--c5-steps N] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "cell-updates/s + % HBM peak: batched FP32 vs plain FP64, ρ(t) err ≤1e-3"
HBM_PEAK_GBS = 8000.0
SEED = 20251202

# Measured HBM bytes per launch (read, write) of the kernels each workload runs,
# from rocprofv3 counter runs of the same workload: reads from the request-size
# counters (TCC_EA0_RDREQ_{32B,64B,128B} x 32/64/128 B), writes from WRITE_SIZE
# (KiB), each checked on a plain copy of known size.  Keyed by (variant, launch
# kind); "single" = one step per launch, "blocked" = K steps per launch.
TRAFFIC = {   # bytes per launch (read, write); profiles/dp_step_f32_c4_live_window.md
    ("dp_f32_128x24_t256_x8aq", "single"): (234_266_000, 143_368_000),    # two-pass single step, live start
    ("dp_f32_128x24_t256_x8aq", "blocked"): (359_357_000, 148_703_000),   # eight-step launch, absorbing field
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
                steps=20_000, nf=3, chunk=5000)


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


def c7_batch(tm, model):
    """Rietkerk vegetation + consumers as one step (FP64): 4 rainfall values x 4
    replicates of 256^2 with the constants of the reference's 3-species driver
    (order_3stoc_test_rietkerk.cpp: km, hr, kg; dt = 0.1 hr, Dt_Analytic = 0.118 hr),
    P ~ Uniform(0, 60), grazers Uniform(0, 3), predators Uniform(0, 0.5),
    W, O ~ Uniform(0, 2), perception fields 0."""
    ps = [tm.FKEParams(a=float(a) / 24.0) for a in np.linspace(0.8, 1.6, 4)]
    inst = [tm.Instance(256, k // 4, SEED, k) for k in range(16)]
    if model == tm.MODEL_FKE2:
        lo, hi = [0.0] * 5, [60.0, 3.0, 2.0, 2.0, 0.0]
    else:
        lo, hi = [0.0] * 7, [60.0, 3.0, 0.5, 2.0, 2.0, 0.0, 0.0]
    return dict(model=model, precision=tm.F64, params=ps, inst=inst, lo=lo, hi=hi, steps=200,
                nf=tm.NFIELDS[model], chunk=100)


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


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def _open(tm, b, variant_name=None, tune=True):
    ctx = tm.Context(0)
    ctx.configure(b["model"], b["precision"], b["params"], b["inst"])
    variant = None
    if variant_name:
        v = tm.variant_names(b["model"], b["precision"]).index(variant_name)
        ctx.set_launch(v, 100)
        variant = (v, 100)
    ctx.init_uniform(b["lo"], b["hi"])
    if not variant_name and tune:   # tuned on the initial state of the run
        variant = ctx.autotune(os.environ.get("TM_TUNE_CACHE"))
    if variant is None:
        variant = ctx.launch()
    names = tm.variant_names(b["model"], b["precision"])
    return ctx, names[variant[0]], variant[1]


def run_workload(tm, b, steps, warmup, want_rows=True, tune=True, variant_name=None, blocks=False,
                 keep_open=False):
    """Warm up, then time `steps` steps of the real trajectory from t = 0: wall
    clock around the whole job, and the device time of every step launch (HIP
    events on the context stream inside each run() call).  blocks: run one launch
    block per call, and split the device time by launch kind (single-step
    blocks, K-step blocks)."""
    ctx, vname, gsteps = _open(tm, b, variant_name, tune)
    try:
        if warmup:
            ctx.run(warmup, rows=want_rows)
        ctx.reset()
        ctx.init_uniform(b["lo"], b["hi"])
        chunk = max(1, min(steps, gsteps if blocks and gsteps else b.get("chunk", steps)))
        parts = []
        dev_ms = 0.0
        phase = {"single": [0, 0.0], "blocked": [0, 0.0]}
        m0 = ctx.mode_steps()
        t0 = time.perf_counter()
        done = 0
        next_log = time.perf_counter() + 10.0
        while done < steps:
            n = min(chunk, steps - done)
            a0 = ctx.mode_steps()
            parts.append(ctx.run(n, rows=want_rows))
            ms = ctx.last_run_ms()
            a1 = ctx.mode_steps()
            dev_ms += ms
            kind = "blocked" if a1[1] - a0[1] > a1[0] - a0[0] else "single"
            phase[kind][0] += n
            phase[kind][1] += ms
            done += n
            if done < steps and time.perf_counter() > next_log:
                log(f"[bench]   ... {done}/{steps} steps")
                next_log = time.perf_counter() + 10.0
        t1 = time.perf_counter()
        m1 = ctx.mode_steps()
        rows = np.concatenate(parts, axis=1) if want_rows else None
        final = [ctx.download(k) for k in range(min(len(b["inst"]), 4))]
    finally:
        if not keep_open:
            ctx.close()
    cells = cells_of(b)
    wall = t1 - t0
    avg_step_s = dev_ms / 1e3 / steps
    achieved = cells * bytes_per_cell(b) / avg_step_s / 1e9
    out = dict(value=cells * steps / wall, wall_s=wall, ms_per_step=wall / steps * 1e3,
               kernel_ms=avg_step_s * 1e3, achieved_gbs=achieved, frac=achieved / HBM_PEAK_GBS,
               cells=cells, steps=steps, variant=vname, graph_steps=gsteps,
               single_steps=m1[0] - m0[0], blocked_steps=m1[1] - m0[1], rows=rows, final=final,
               phase=phase)
    if keep_open:
        out["ctx"] = ctx
    return out


def phase_split(r, b, k_per_launch=8):
    """Live (single-step launches) and absorbing (K-step launches) phases of a
    run timed per launch block: steps, device ms per step, cell-updates/s and the
    HBM fraction of each kind's algorithmic bytes per launch."""
    cells, bpc = cells_of(b), bytes_per_cell(b)
    out = {}
    for kind, (n, ms) in r["phase"].items():
        if n == 0:
            continue
        step_ms = ms / n
        launch_ms = step_ms * (k_per_launch if kind == "blocked" else 1)
        gbs = cells * bpc / (launch_ms / 1e3) / 1e9
        out[kind] = dict(steps=n, device_ms_per_step=step_ms, cell_updates_per_s=cells / (step_ms / 1e3),
                         steps_per_launch=k_per_launch if kind == "blocked" else 1,
                         achieved_gbs=gbs, frac=gbs / HBM_PEAK_GBS)
    return out


def kernel_of(variant):
    """The single-step kernel a DP variant (dp_f32_TXxTY_tNT[_suffixes]) runs while
    the field is live: two-pass when a suffix ends in q, bounded to 80 VGPRs (six
    waves per SIMD) for _w6 and for the adaptive (_xKa...) variants."""
    parts = (variant or "").split("_")
    if len(parts) < 4 or not parts[0] == "dp":
        return variant
    tx, ty = parts[2].split("x")
    nt = parts[3][1:]
    suf = parts[4:]
    q = any(p.endswith("q") for p in suf)
    w6 = any(p.startswith("w6") for p in suf) or any(p.startswith("x") and "a" in p for p in suf)
    return (f"dp_step_kernel<{'float' if parts[1] == 'f32' else 'double'},{tx},{ty},{nt},0,{6 if w6 else 1},"
            f"{1 if q else 0}> ({'two-pass ' if q else ''}single DP step)")


def roofline(h, b, kind="single"):
    """The headline's roofline: its dominant kernel is the single-step (two-pass)
    DP kernel; algorithmic bytes per launch (one read and one write of the field)
    over its mean launch duration (device time per step of the timed region,
    HIP events on the context stream), against the measured HBM bytes per launch."""
    step_s = h["kernel_ms"] / 1e3
    achieved = cells_of(b) * bytes_per_cell(b) / step_s / 1e9
    t = TRAFFIC.get((h["variant"], kind))
    traffic = None if t is None else int(t[0] + t[1])
    out = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
           "kernel": kernel_of(h["variant"]),
           "bytes_per_cell": bytes_per_cell(b), "cells_per_launch": cells_of(b),
           "algorithmic_bytes_per_launch": cells_of(b) * bytes_per_cell(b),
           "step_device_ms": h["kernel_ms"], "single_steps": h["single_steps"],
           "blocked_steps": h["blocked_steps"],
           "timing": "HIP events on the context stream around every run() of the timed region; the per-step "
                     "share also holds the per-block constants, rho(t) fold and counter bump launches "
                     "(about 2% of it)"}
    if traffic is not None:
        out["traffic_gbs"] = traffic / step_s / 1e9
        out["traffic_frac"] = out["traffic_gbs"] / HBM_PEAK_GBS
        out["traffic_source"] = ("rocprofv3 counters of the same kernel on this workload: reads from "
                                 "TCC_EA0_RDREQ_{32B,64B,128B}, writes from WRITE_SIZE (profiles/)")
    return out


def side_fields(r, keys=("value", "ms_per_step", "kernel_ms", "achieved_gbs", "frac", "variant", "steps")):
    return {k: r[k] for k in keys}


def sweep_leg(tm, steps=None):
    """The sweep pieces on the C3 batch (FP32 Rietkerk, 16 R x 16 replicates of
    512^2): device frames, a measured run over the reference's log-spaced times,
    the replicate statistics, and the output layer on one instance."""
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


def c5_leg(tm, steps, warmup):
    """The 8192^2 Rietkerk lattice (R = 0.75, FP32) for `steps` of its 10^5 steps,
    the FP64 run of the same seeds for its error, and the cost of one snapshot (the
    reference writes a frame every 1000 steps, FKE:4090-4097), as a share of the
    1000 steps it covers."""
    from trophic_maelstorm_amd import compare
    b = c5_batch(tm, tm.F32, steps)
    r = run_workload(tm, b, steps, warmup, keep_open=True)
    ctx = r.pop("ctx")
    try:
        with tempfile.TemporaryDirectory() as d:
            t0 = time.perf_counter()
            ctx.save_snapshot(0, os.path.join(d, "c5.tms"))
            snap_s = time.perf_counter() - t0
    finally:
        ctx.close()
    bd = c5_batch(tm, tm.F64, steps)
    rd = run_workload(tm, bd, steps, warmup)
    out = side_fields(r)
    out["fp64"] = side_fields(rd)
    out["snapshot_s"] = snap_s
    out["snapshot_gbs"] = cells_of(b) * 3 * 4 / snap_s / 1e9
    out["snapshot_share_per_1000_steps"] = snap_s / (snap_s + 1000 * r["ms_per_step"] / 1e3)
    out["full_run_estimate_s"] = 1e5 * r["ms_per_step"] / 1e3 + 100 * snap_s
    err = dict(rho=compare.rows_report(r["rows"], rd["rows"], "PWO"),
               fields=compare.field_error(r["final"], rd["final"]))
    return out, err


def cpu_baseline(threads=8):
    """The reference's own DP-Dornic integrator (percolationDornic_2D,
    stochasticSPDP.cpp:940, built from its source into oracle/_ref) on C1: L = 128,
    a = 0.72, b = 2, D = 1, sigma = sqrt 2, dt = 0.1, 10^4 steps per replicate,
    logged every time unit.  It ends a replicate at the first logged time with no
    occupied cell (stochasticSPDP.cpp:1019-1020); the quench goes extinct after
    about 700 steps, so the steps it ran are counted from its rows.  On `threads`
    host threads (one replicate each, the reference's thread-per-parameter
    pattern) and on one thread (all replicates in one call).  C2 (Rietkerk 256^2)
    has no buildable reference here (exprtk is absent): the oracle port on one
    thread, 400 steps."""
    from oracle import oracle
    out = {}
    ref = oracle.ref_lib()
    L, dt = 128, 0.1
    rho0 = oracle.init_uniform(L, [0.0], [1.0], SEED, 0).reshape(-1).copy()
    t_meas = np.arange(2.0, 1001.0)          # logged every time unit to t = 1000 (10^4 steps)
    if ref is not None:
        for nthr, reps in ((threads, 1), (1, threads)):
            fd = os.dup(1)
            devnull = os.open(os.devnull, os.O_WRONLY)
            os.dup2(devnull, 1)     # the reference prints a line per replicate
            try:
                t0 = time.perf_counter()
                steps = ref.ref_ssp_parallel_steps(nthr, L, 0.72, 2.0, 1.0, math.sqrt(2.0), dt, 1.0, reps, rho0,
                                                   t_meas, len(t_meas), 1000.0)
                wall = time.perf_counter() - t0
            finally:
                os.dup2(fd, 1)
                os.close(fd)
                os.close(devnull)
            out[nthr] = dict(value=steps * L * L / wall, wall_s=wall, replicate_steps=steps, threads=nthr,
                             replicates=threads)
    c2 = oracle.rk_params(R=1.0, sigma=0.5)
    f0 = oracle.init_uniform(256, [0.0, 0.0, 0.0], [90.0, 1.0, 1.0], SEED, 0)
    n2 = 400
    t0 = time.perf_counter()
    oracle.rk_run(c2, f0, 256, SEED, 0, 0, n2)
    w2 = time.perf_counter() - t0
    c2_out = dict(value=256 * 256 * n2 / w2, unit="cell-updates/s", threads=1, kind="port",
                  sample=f"oracle port of the Rietkerk step (FKE:4154-4465), 256^2, {n2} of C2's 2x10^4 steps",
                  wall_s=w2)
    host = os.cpu_count()
    if threads in out:
        m = out[threads]
        return dict(value=m["value"], unit="cell-updates/s", cores=threads, kind="reference",
                    sample=f"C1: percolationDornic_2D (stochasticSPDP.cpp:940) built from source, L=128, a=0.72, "
                           f"b=2, D=1, sigma=sqrt2, dt=0.1, {threads} replicates (one per thread) of up to 10^4 "
                           f"steps; it stops each at extinction: {int(m['replicate_steps'])} replicate-steps run",
                    wall_s=m["wall_s"], threads=threads, host_logical_cpus=host,
                    one_thread=dict(value=out[1]["value"], wall_s=out[1]["wall_s"],
                                    replicate_steps=out[1]["replicate_steps"]),
                    c2_one_thread_port=c2_out)
    c2_out["host_logical_cpus"] = host
    c2_out["cores"] = 1
    return c2_out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20,
                    help="timed steps of the headline (the C4 quench from its initial field)")
    ap.add_argument("--warmup", type=int, default=40,
                    help="untimed steps first (the run then restarts from its initial field); the GPU "
                         "reaches its working clock within them")
    ap.add_argument("--c5-steps", type=int, default=200, help="steps of the 8192^2 lattice (its full run is 10^5)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-side", action="store_true", help="headline and the full C4 runs only")
    ap.add_argument("--variant", default=None,
                    help="kernel variant for the headline (or --only) run instead of the autotuner's choice")
    ap.add_argument("--only", choices=["c2", "c3", "c3-f64", "c4", "c4-f64", "c5", "sweep", "dp2", "dp3", "fke2", "fke3",
                                       "movement"],
                    default=None,
                    help="only this workload, --steps steps (default: its full length), one JSON line")
    ap.add_argument("--headline-only", action="store_true",
                    help="only the FP32 headline run (no FP64 comparison, side workloads or CPU baseline)")
    ap.add_argument("--tune", action="store_true",
                    help="with --only: run the autotuner's choice instead of the library's default launch shape")
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
    from trophic_maelstorm_amd import compare

    if args.only == "movement":
        print(json.dumps({"workload": "fke_movement", **movement_leg(tm)}), flush=True)
        return
    if args.only == "sweep":
        print(json.dumps({"workload": "c3_sweep", **sweep_leg(tm, args.steps)}), flush=True)
        return
    if args.only:
        b = {"c2": lambda: c2_batch(tm, tm.F64), "c3": lambda: c3_batch(tm, tm.F32),
             "c4": lambda: c4_batch(tm, tm.F32), "c4-f64": lambda: c4_batch(tm, tm.F64),
             "c3-f64": lambda: c3_batch(tm, tm.F64), "c5": lambda: c5_batch(tm, tm.F32),
             "dp2": lambda: c6_batch(tm, tm.MODEL_DP2), "dp3": lambda: c6_batch(tm, tm.MODEL_DP3),
             "fke2": lambda: c7_batch(tm, tm.MODEL_FKE2), "fke3": lambda: c7_batch(tm, tm.MODEL_FKE3)}[args.only]()
        n = args.steps if "--steps" in sys.argv else b["steps"]
        log(f"[bench] {args.only}: {cells_of(b)} cells/step x {n} steps")
        r = run_workload(tm, b, n, args.warmup, variant_name=args.variant, tune=args.tune,
                         blocks=args.only in ("c4", "c4-f64"))
        out = {"workload": args.only, "value": r["value"], "unit": "cell-updates/s", "steps": n,
               "ms_per_step": r["ms_per_step"], "step_device_ms": r["kernel_ms"], "variant": r["variant"]}
        if r["blocked_steps"]:
            # K-step launches advance several steps per field read: an HBM fraction
            # per step would count bytes never moved, so only the per-phase figures,
            # each over its own launches, are stated
            out["phases"] = phase_split(r, b)
        else:
            out.update(achieved_gbs=r["achieved_gbs"], frac=r["frac"])
        print(json.dumps(out), flush=True)
        return

    # ---- headline: the live start of the C4 quench, FP32 ----
    c4 = c4_batch(tm, tm.F32)
    K = args.steps
    log(f"[bench] C4 FP32 batched DP quench: {cells_of(c4)} cells/step x {K} steps from t = 0")
    # the C4 legs run the library's default launch shape (an adaptive variant built
    # for the whole quench); the autotuner, which times the initial live state only,
    # is not asked
    h = run_workload(tm, c4, K, args.warmup, variant_name=args.variant, tune=False)
    log(f"[bench]   {h['value']:.4g} cell-updates/s, kernel {h['kernel_ms']:.4f} ms, "
        f"{h['achieved_gbs']:.0f} GB/s ({100 * h['frac']:.1f}% of peak)")
    if args.headline_only:
        print(json.dumps({"metric": METRIC, "value": h["value"], "unit": "cell-updates/s", "steps": K,
                          "ms_per_step": h["ms_per_step"], "step_device_ms": h["kernel_ms"],
                          "achieved_gbs": h["achieved_gbs"], "variant": h["variant"],
                          "roofline": roofline(h, c4)}), flush=True)
        return
    side = {}
    errors = {}
    c4d = c4_batch(tm, tm.F64)
    log(f"[bench] C4 FP64, same seeds, {K} steps")
    h64 = run_workload(tm, c4d, K, args.warmup, tune=False)
    side["c4_fp64"] = side_fields(h64)

    # ---- the whole C4 quench, per phase, and its FP32 error ----
    full = {}
    for name, b in (("fp32", c4_batch(tm, tm.F32)), ("fp64", c4_batch(tm, tm.F64))):
        log(f"[bench] C4 {name} full run, {b['steps']} steps, timed per graph block")
        r = run_workload(tm, b, b["steps"], 0, blocks=True, tune=False)
        full[name] = r
        side[f"c4_full_run_{name}"] = dict(side_fields(r, ("value", "ms_per_step", "kernel_ms", "variant", "steps")),
                                           phases=phase_split(r, b))
    errors["c4"] = dict(rho=compare.rows_report(full["fp32"]["rows"], full["fp64"]["rows"], "r"),
                        fields=compare.field_error(full["fp32"]["final"], full["fp64"]["final"]),
                        steps=c4["steps"])
    del full

    if not args.no_side:
        c2 = c2_batch(tm, tm.F64)
        log("[bench] C2 plain FP64 Rietkerk 256^2")
        r2 = run_workload(tm, c2, c2["steps"], args.warmup, want_rows=False)
        side["c2_plain_fp64"] = side_fields(r2)
        c3 = c3_batch(tm, tm.F32)
        log("[bench] C3 FP32 batched Rietkerk 256 x 512^2")
        r3 = run_workload(tm, c3, c3["steps"], args.warmup)
        side["c3_batched_fp32"] = side_fields(r3)
        log(f"[bench]   {r3['value']:.4g} cell-updates/s, {r3['achieved_gbs']:.0f} GB/s")
        c3d = c3_batch(tm, tm.F64)
        log("[bench] C3 FP64 (same seeds)")
        r3d = run_workload(tm, c3d, c3["steps"], args.warmup)
        side["c3_fp64"] = side_fields(r3d)
        errors["c3"] = dict(rho=compare.rows_report(r3["rows"], r3d["rows"], "PWO"),
                            fields=compare.field_error(r3["final"], r3d["final"]), steps=c3["steps"])
        side["ratio_c4_fp32_over_c2_plain_fp64"] = h["value"] / r2["value"]
        side["ratio_c3_fp32_over_c2_plain_fp64"] = r3["value"] / r2["value"]
        del r3, r3d
        if args.c5_steps > 0:
            log(f"[bench] C5 Rietkerk 8192^2, {args.c5_steps} of its 10^5 steps, FP32 and FP64")
            side["c5_fp32"], errors["c5"] = c5_leg(tm, args.c5_steps, min(args.warmup, 5))
            errors["c5"]["steps"] = args.c5_steps
        log("[bench] sweep pieces on the C3 batch (frames, series, replicate stats, output)")
        side["c3_sweep"] = sweep_leg(tm)
        for name, model in (("c6_consumers_dp2_fp64", tm.MODEL_DP2), ("c6_consumers_dp3_fp64", tm.MODEL_DP3)):
            c6 = c6_batch(tm, model)
            log(f"[bench] {name}: 64 x 256^2, {c6['steps']} steps")
            r6 = run_workload(tm, c6, c6["steps"], args.warmup, want_rows=False)
            side[name] = side_fields(r6)
        for name, model in (("c7_rietkerk_consumers_fke2_fp64", tm.MODEL_FKE2),
                            ("c7_rietkerk_consumers_fke3_fp64", tm.MODEL_FKE3)):
            c7 = c7_batch(tm, model)
            log(f"[bench] {name}: 16 x 256^2, {c7['steps']} steps")
            r7 = run_workload(tm, c7, c7["steps"], args.warmup, want_rows=False)
            side[name] = side_fields(r7, ("value", "ms_per_step", "kernel_ms", "variant", "steps"))

    cpu = None
    if not args.no_cpu:
        log("[bench] CPU baseline (C1 reference on 8 and 1 host threads, C2 oracle port)")
        cpu = cpu_baseline()
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
        "config": {"workload": "C4 batched FP32 DP quench, live start: 128x128^2 + 64x256^2 + 48x512^2 + "
                               "16x1024^2, a 0.72->0.70 over 10^4 steps, steps from t = 0, rho(t) FP64 every step",
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
