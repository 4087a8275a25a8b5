"""FP32-storage runs against the FP64 run of the same seeds: the ensemble rho(t)
acceptance of the FP32 path, and the error figures stated for every batch.

rows: [n_inst, n_steps, 2 nf] as returned by Context.run (occupied count, spatial
mean per field).  The ensemble mean m(t) is the mean over instances of the
spatial mean.

Same-seed FP32 and FP64 trajectories share their Philox words, so they agree
until one integer draw differs (an FP32 rounding lands on the other side of a
sampler decision); from then on that instance is an independent realisation of
the same process.  While the ensemble is populated the difference is rounding
and must stay within 1e-3 relative.  Once most instances have thinned out, the
difference is dominated by the decorrelated ones, and the acceptance becomes
statistical: |m32 - m64| within 4 standard errors of the difference of two
independent ensembles, and the surviving fractions within the two-proportion
95% band (plus one instance).
"""
from __future__ import annotations

import math

import numpy as np

REL_TOL = 1e-3
POPULATED = 0.1      # "populated": the FP64 ensemble holds >= 10% of its initial mean
Z_TAIL = 4.0


def field_report(r32: np.ndarray, r64: np.ndarray, f: int = 0) -> dict:
    """Errors of field f: relative error while populated and over the run, the
    statistical tail check, the paired bias and the survival difference."""
    c = 2 * f + 1
    x32, x64 = r32[:, :, c], r64[:, :, c]
    n = x32.shape[0]
    m32, m64 = x32.mean(0), x64.mean(0)
    nz = m64 != 0
    rel = np.zeros_like(m64)
    rel[nz] = np.abs(m32[nz] - m64[nz]) / np.abs(m64[nz])
    pop = np.abs(m64) >= POPULATED * abs(m64[0])
    # independent-ensemble standard error of the difference (the tail criterion)
    se_ind = np.sqrt((x32.var(0, ddof=1) + x64.var(0, ddof=1)) / n) if n > 1 else np.zeros_like(m64)
    tail_ok = (rel <= REL_TOL) | (np.abs(m32 - m64) <= Z_TAIL * se_ind)
    # paired bias: a systematic FP32 offset shows as a large paired z
    d = x32 - x64
    se_p = d.std(0, ddof=1) / math.sqrt(n) if n > 1 else np.zeros_like(m64)
    zb = np.where(se_p > 0, np.abs(d.mean(0)) / np.where(se_p > 0, se_p, 1.0), 0.0)
    s32 = (r32[:, :, 2 * f] > 0).mean(0)
    s64 = (r64[:, :, 2 * f] > 0).mean(0)
    p = 0.5 * (s32 + s64)
    band = 1.96 * np.sqrt(2.0 * p * (1.0 - p) / n) + 1.0 / n
    surv_ok = np.abs(s32 - s64) <= band
    return dict(
        steps=int(m64.size),
        populated_steps=int(pop.sum()),
        ensemble_rel_max_populated=float(rel[pop].max()) if pop.any() else 0.0,
        ensemble_rel_mean_populated=float(rel[pop].mean()) if pop.any() else 0.0,
        ensemble_rel_max=float(rel.max()) if rel.size else 0.0,
        ensemble_rel_mean=float(rel[nz].mean()) if nz.any() else 0.0,
        tail_steps_outside_1e3=int((~pop & (rel > REL_TOL)).sum()),
        tail_ok=bool(tail_ok.all()),
        bias_z_max=float(zb.max()) if zb.size else 0.0,
        survival_diff_max=float(np.abs(s32 - s64).max()),
        survival_ok=bool(surv_ok.all()),
        accepted=bool(((rel <= REL_TOL) | ~pop).all() and tail_ok.all() and surv_ok.all()),
    )


def rows_report(r32: np.ndarray, r64: np.ndarray, names: str) -> dict:
    return {names[f]: field_report(r32, r64, f) for f in range(len(names))}


def field_error(f32, f64) -> dict:
    """Max and mean absolute difference of final fields (per-trajectory: meaningful
    only until the first divergent integer draw), and the same relative to the
    FP64 field's mean magnitude."""
    a = np.concatenate([np.asarray(x).ravel() for x in f32])
    b = np.concatenate([np.asarray(x).ravel() for x in f64])
    d = np.abs(a - b)
    scale = float(np.abs(b).mean()) or 1.0
    return dict(max_abs=float(d.max()), mean_abs=float(d.mean()), mean_abs_rel=float(d.mean() / scale),
                identical_fraction=float(np.mean(d == 0)))
