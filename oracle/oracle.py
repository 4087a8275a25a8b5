"""ctypes wrapper of the CPU oracle (oracle/liboracle.so) and of the reference's
own legacy DP integrator (oracle/_ref/libref_ssp.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product package never
imports this module.  See tm_oracle.h for what the oracle restates and how each
function is pinned.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libref_ssp.so")

_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


class DPParams(C.Structure):
    """tmo_dp_params (mirror of tm_dp_params)."""
    _fields_ = [("a", C.c_double), ("b", C.c_double), ("D", C.c_double), ("sigma", C.c_double),
                ("dt", C.c_double), ("dx", C.c_double), ("a_end", C.c_double),
                ("ramp_steps", C.c_int64), ("flush_eps", C.c_double), ("flush_neg", C.c_double),
                ("scheme", C.c_int32), ("sampler", C.c_int32)]


class RKParams(C.Structure):
    """tmo_rk_params (mirror of tm_rietkerk_params)."""
    _fields_ = [("R", C.c_double), ("c", C.c_double), ("gmax", C.c_double), ("alpha", C.c_double),
                ("rW", C.c_double), ("W0", C.c_double), ("d", C.c_double),
                ("D", C.c_double * 3), ("K", C.c_double * 3), ("sigma", C.c_double),
                ("dt", C.c_double), ("dx", C.c_double), ("flush_eps", C.c_double),
                ("sampler", C.c_int32), ("pad_", C.c_int32)]


class DPCParams(C.Structure):
    """tmo_dpc_params (mirror of tm_dpc_params)."""
    _fields_ = [("a", C.c_double), ("b", C.c_double), ("D", C.c_double * 3), ("sigma", C.c_double * 3),
                ("v", C.c_double * 3), ("M", C.c_double * 3), ("E", C.c_double * 3),
                ("A", (C.c_double * 3) * 3), ("H", (C.c_double * 3) * 3), ("pR", C.c_double * 3),
                ("dt", C.c_double), ("dx", C.c_double), ("a_end", C.c_double), ("ramp_steps", C.c_int64),
                ("flush_eps", C.c_double), ("flush_neg", C.c_double), ("sampler", C.c_int32),
                ("pad_", C.c_int32)]


FMT_PATH = os.path.join(HERE, "libfmt_check.so")
RADII_PATH = os.path.join(HERE, "_ref", "test_perception_radii")


def build(force: bool = False) -> str:
    """Compile liboracle.so and libfmt_check.so."""
    for path in (LIB_PATH, FMT_PATH):
        if force or not os.path.exists(path):
            subprocess.run(["make", "-C", HERE, os.path.basename(path)], check=True,
                           stdout=subprocess.DEVNULL)
    return LIB_PATH


def build_ref() -> str | None:
    """Compile oracle/_ref/libref_ssp.so from the reference sources, if present."""
    if os.path.isdir("/root/reference/simulations/Percolation_FD"):
        subprocess.run(["make", "-C", HERE, "ref"], check=True, stdout=subprocess.DEVNULL)
    return REF_PATH if os.path.exists(REF_PATH) else None


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.tmo_philox4x32_10.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        for name in ("tmo_exp", "tmo_log", "tmo_norm_icdf"):
            getattr(L, name).restype = C.c_double
            getattr(L, name).argtypes = [C.c_double]
        L.tmo_logfact.restype = C.c_double
        L.tmo_logfact.argtypes = [C.c_int64]
        L.tmo_sample.argtypes = [C.c_int, C.c_double, C.c_int, C.c_uint32, C.c_uint32, C.c_uint32,
                                 C.c_int64, C.c_int64, _f64p]
        L.tmo_dornic_constants.argtypes = [C.c_double] * 5 + [_f64p]
        L.tmo_dornic_cell.restype = C.c_double
        L.tmo_dornic_cell.argtypes = [C.c_double] * 5 + [C.c_int, C.c_uint32, C.c_uint32, C.c_uint32,
                                                         C.c_int64, C.POINTER(C.c_int64)]
        L.tmo_dornic_batch.argtypes = [C.c_double] * 5 + [C.c_int, C.c_uint32, C.c_uint32, C.c_int64,
                                                          C.c_int64, _f64p, C.c_void_p]
        L.tmo_init_uniform.argtypes = [C.c_int, C.c_int, C.c_uint32, C.c_uint32, _f64p, _f64p, _f64p]
        L.tmo_dp_run.argtypes = [C.POINTER(DPParams), C.c_int, C.c_uint32, C.c_uint32, C.c_int64,
                                 C.c_int64, _f64p, C.c_void_p, C.c_void_p, C.POINTER(C.c_int64)]
        L.tmo_rk_run.argtypes = [C.POINTER(RKParams), C.c_int, C.c_uint32, C.c_uint32, C.c_int64,
                                 C.c_int64, _f64p, C.c_void_p, C.c_void_p, C.POINTER(C.c_int64)]
        L.tmo_laplacian.argtypes = [C.c_int, _f64p, _f64p]
        L.tmo_rk_rhs.argtypes = [C.POINTER(RKParams), C.c_int, _f64p, _f64p]
        L.tmo_rk4_deterministic.argtypes = [C.POINTER(RKParams), C.c_int, C.c_int64, _f64p]
        L.tmo_dp_nonlinear.restype = C.c_double
        L.tmo_dp_nonlinear.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double]
        L.tmo_reduce.argtypes = [C.c_int, C.c_int, _f64p, _f64p]
        L.tmo_flush.restype = C.c_double
        L.tmo_flush.argtypes = [C.c_double] * 3
        L.tmo_init_speckle.argtypes = [C.c_int, C.c_int, C.c_uint32, C.c_uint32, C.c_double, _f64p, _f64p,
                                       _f64p]
        L.tmo_init_gauss.argtypes = [C.c_int, C.c_int, C.c_int, _f64p, C.c_int, _f64p, _f64p]
        L.tmo_replicate_stats.argtypes = [_f64p, C.c_int, C.c_int, C.c_int,
                                          np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS"),
                                          C.c_int, C.c_int, _f64p]
        L.tmo_species_stats.argtypes = [_f64p, C.c_int, C.c_int, C.c_int, C.c_int,
                                        np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS"),
                                        C.c_int, C.c_int, _f64p]
        L.tmo_fke_move_constants.restype = C.c_int
        L.tmo_fke_move_constants.argtypes = [C.c_double] * 5 + [C.c_int, C.c_double, C.c_int, C.POINTER(C.c_int32),
                                                                 C.POINTER(C.c_int32), C.POINTER(C.c_double), _f64p,
                                                                 C.c_int, _f64p, C.c_int]
        L.tmo_fke_movement.argtypes = [C.c_int, _f64p, _f64p, _f64p, _f64p, C.c_int, C.c_int, C.c_double, _f64p,
                                       _f64p, _f64p]
        L.tmo_fke_perception.argtypes = [C.c_int, C.c_int, _f64p, C.c_double, _f64p, C.c_void_p, C.c_int, C.c_uint32,
                                         C.c_uint32, C.c_int64, _f64p, _f64p, _f64p]
        L.tmo_sincos2pi.argtypes = [C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.tmo_fke_rk4.restype = C.c_int64
        L.tmo_fke_rk4.argtypes = [C.c_void_p, C.c_int, C.c_int, _f64p]
        L.tmo_fke_veg_dornic.argtypes = [C.POINTER(RKParams), C.c_int, C.c_uint32, C.c_uint32, C.c_int64, _f64p,
                                         _f64p]
        L.tmo_fke_movement_sep.argtypes = [C.c_int, _f64p, _f64p, _f64p, _f64p, C.c_int, C.c_int, C.c_int,
                                           C.c_double, _f64p, _f64p, _f64p]
        L.tmo_fke_mov_stats.argtypes = [C.c_int, _f64p, _f64p, _f64p, _f64p]
        L.tmo_fixed_sum.restype = C.c_double
        L.tmo_fixed_sum.argtypes = [_f64p, C.c_int64]
        L.tmo_dpc_rk4.argtypes = [C.POINTER(DPCParams), C.c_int, _f64p]
        L.tmo_perception_offsets.restype = C.c_int
        L.tmo_perception_offsets.argtypes = [C.c_int, C.c_void_p, C.c_int]
        L.tmo_dpc_run.restype = C.c_int
        L.tmo_dpc_run.argtypes = [C.POINTER(DPCParams), C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_uint32,
                                  C.c_uint32, C.c_int64, C.c_int64, _f64p, C.c_void_p, C.POINTER(C.c_int64)]
        _lib = L
    return _lib


def ref_lib():
    """The reference's legacy DP integrator, or None when it was not built."""
    global _ref
    if _ref is None:
        if not os.path.exists(REF_PATH):
            return None
        R = C.CDLL(REF_PATH)
        R.ref_ssp_percolation_dornic.restype = C.c_int
        R.ref_ssp_percolation_dornic.argtypes = [C.c_int] + [C.c_double] * 6 + [
            C.c_int, _f64p, _f64p, C.c_int, C.c_double, _f64p, C.c_int]
        R.ref_ssp_parallel.restype = C.c_int
        R.ref_ssp_parallel.argtypes = [C.c_int, C.c_int] + [C.c_double] * 6 + [
            C.c_int, _f64p, _f64p, C.c_int, C.c_double]
        R.ref_ssp_parallel_steps.restype = C.c_double
        R.ref_ssp_parallel_steps.argtypes = [C.c_int, C.c_int] + [C.c_double] * 6 + [
            C.c_int, _f64p, _f64p, C.c_int, C.c_double]
        R.ref_ssp_var_mean_surv.restype = None
        R.ref_ssp_var_mean_surv.argtypes = [_f64p, _f64p, C.c_int]
        R.ref_ssp_log10_time_bins.restype = C.c_int
        R.ref_ssp_log10_time_bins.argtypes = [C.c_double, C.c_double, C.c_void_p, C.c_int]
        _ref = R
    return _ref


# ---------------------------------------------------------------- helpers
def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().tmo_philox4x32_10(c, k, o)
    return [int(x) for x in o]


def sample(kind: str, n: int, param: float = 0.0, mode: int = 0, seed: int = 1, inst: int = 0,
           cell: int = 0, step: int = 0) -> np.ndarray:
    kinds = {"uniform": 0, "normal": 1, "poisson": 2, "gamma": 3}
    out = np.empty(n, np.float64)
    lib().tmo_sample(kinds[kind], float(param), int(mode), seed, inst, cell, step, n, out)
    return out


def dornic_constants(linear, D, sigma, dt, dx):
    out = np.empty(3, np.float64)
    lib().tmo_dornic_constants(linear, D, sigma, dt, dx, out)
    return out


def dornic_batch(alpha_i, rho, lam, lexp, sigma, n, sampler=0, seed=7, inst=0, step=0, draws=False):
    """n independent Dornic transitions of a cell (alpha_i, rho); returns rho* (and Poisson draws)."""
    out = np.empty(n, np.float64)
    k = np.empty(n, np.int64) if draws else None
    lib().tmo_dornic_batch(alpha_i, rho, lam, lexp, sigma, sampler, seed, inst, step, n, out,
                           k.ctypes.data if k is not None else None)
    return (out, k) if draws else out


def init_uniform(L, lo, hi, seed, inst):
    lo = np.ascontiguousarray(lo, np.float64)
    hi = np.ascontiguousarray(hi, np.float64)
    out = np.empty(len(lo) * L * L, np.float64)
    lib().tmo_init_uniform(L, len(lo), seed, inst, lo, hi, out)
    return out.reshape(len(lo), L, L)


def dp_params(a=0.72, b=2.0, D=1.0, sigma=2.0 ** 0.5, dt=0.1, dx=1.0, a_end=None, ramp_steps=0,
              flush_eps=1e-12, flush_neg=-1e-80, scheme=0, sampler=0) -> DPParams:
    return DPParams(a, b, D, sigma, dt, dx, a if a_end is None else a_end, ramp_steps, flush_eps,
                    flush_neg, scheme, sampler)


def rk_params(R=1.0, c=10.0, gmax=0.05, alpha=0.2, rW=0.2, W0=0.2, d=0.25, D=(0.1, 0.1, 100.0),
              K=(0.0, 5.0, 5.0), sigma=0.5, dt=0.005, dx=2.0, flush_eps=1e-10, sampler=0) -> RKParams:
    return RKParams(R, c, gmax, alpha, rW, W0, d, (C.c_double * 3)(*D), (C.c_double * 3)(*K), sigma,
                    dt, dx, flush_eps, sampler, 0)


def _run(fn, p, fields, L, seed, inst, step0, nsteps, nf, want_rows, want_draws):
    f = np.ascontiguousarray(fields, np.float64).copy().reshape(-1)
    rows = np.zeros((nsteps, 2 * nf), np.float64) if want_rows else None
    draws = np.zeros(L * L, np.int64) if want_draws else None
    fc = C.c_int64(-1)
    st = fn(C.byref(p), L, seed, inst, step0, nsteps, f,
            rows.ctypes.data if rows is not None else None,
            draws.ctypes.data if draws is not None else None, C.byref(fc))
    return {"fields": f.reshape(nf, L, L), "rows": rows,
            "draws": draws.reshape(L, L) if draws is not None else None,
            "status": st, "fault_cell": fc.value}


def dp_run(p: DPParams, rho, L, seed, inst, step0, nsteps, rows=True, draws=False):
    return _run(lib().tmo_dp_run, p, rho, L, seed, inst, step0, nsteps, 1, rows, draws)


def rk_run(p: RKParams, fields, L, seed, inst, step0, nsteps, rows=True, draws=False):
    return _run(lib().tmo_rk_run, p, fields, L, seed, inst, step0, nsteps, 3, rows, draws)


# ---------------------------------------------------------------- consumer species
def _v3(v):
    v = list(v) + [0.0] * (3 - len(v))
    return (C.c_double * 3)(*v)


def _m3(m):
    return ((C.c_double * 3) * 3)(*[(C.c_double * 3)(*row) for row in np.asarray(m, np.float64)])


def dpc_params(p) -> DPCParams:
    """DPCParams from any object with the tm_dpc_params field names (the product's
    DPCParams dataclass, or a SimpleNamespace)."""
    return DPCParams(p.a, p.b, _v3(p.D), _v3(p.sigma), _v3(p.v), _v3(p.M), _v3(p.E), _m3(p.A), _m3(p.H),
                     _v3(p.pR), p.dt, p.dx, p.a if getattr(p, "a_end", None) is None else p.a_end,
                     getattr(p, "ramp_steps", 0), getattr(p, "flush_eps", 1e-12),
                     getattr(p, "flush_neg", -1e-80), getattr(p, "sampler", 0), 0)


def sincos2pi(u):
    s, c = C.c_double(), C.c_double()
    lib().tmo_sincos2pi(float(u), C.byref(s), C.byref(c))
    return s.value, c.value


def fixed_sum(x):
    x = np.ascontiguousarray(x, np.float64).reshape(-1)
    return lib().tmo_fixed_sum(x, x.size)


def dpc_rk4(p: DPCParams, ns: int, y):
    y = np.ascontiguousarray(y, np.float64).copy()
    lib().tmo_dpc_rk4(C.byref(p), ns, y)
    return y


def perception_offsets(R: int) -> np.ndarray:
    buf = np.zeros(2 * (2 * R + 1) ** 2, np.int32)
    n = lib().tmo_perception_offsets(R, buf.ctypes.data, buf.size // 2)
    return buf[: 2 * n].reshape(n, 2)


def dpc_run(p: DPCParams, ns: int, fields, L, seed, inst, step0, nsteps, rows=True):
    """nsteps of the ns-species step; fields [2 ns - 1, L, L] (densities, then gammas)."""
    nf = 2 * ns - 1
    rmax = max(p.pR[s] for s in range(ns)) / p.dx
    disk = perception_offsets(int(rmax))
    f = np.ascontiguousarray(fields, np.float64).copy().reshape(-1)
    r = np.zeros((nsteps, 2 * nf), np.float64) if rows else None
    fc = C.c_int64(-1)
    st = lib().tmo_dpc_run(C.byref(p), ns, L, disk.ctypes.data, disk.shape[0], seed, inst, step0, nsteps, f,
                           r.ctypes.data if r is not None else None, C.byref(fc))
    return {"fields": f.reshape(nf, L, L), "rows": r, "status": st, "fault_index": fc.value}


def species_stats(series, params_index, n_params, nsp, flags):
    """series [n_inst, n_meas, 2 nf] -> [n_params, n_meas, 4 nsp] (multi-species form)."""
    x = np.ascontiguousarray(series, np.float64)
    n_inst, n_meas, w = x.shape
    pidx = np.ascontiguousarray(params_index, np.int32)
    out = np.zeros((n_params, n_meas, 4 * nsp), np.float64)
    lib().tmo_species_stats(x.reshape(-1), n_inst, n_meas, w // 2, nsp, pidx, n_params, flags, out.reshape(-1))
    return out


def fke_move_constants(D, v, sigma, dt_analytic, dx, dt_adv=0.0, mv_invariance=0, CI=4):
    rD, rV, pref = C.c_int32(), C.c_int32(), C.c_double()
    sd, sv = np.zeros(4096), np.zeros(4096)
    rc = lib().tmo_fke_move_constants(D, v, sigma, dt_analytic, dt_adv, mv_invariance, dx, CI, C.byref(rD), C.byref(rV),
                                      C.byref(pref), sd, sd.size, sv, sv.size)
    assert rc == 0
    return (rD.value, rV.value, pref.value), sd[: rD.value + 1].copy(), sv[: rV.value + 1].copy()


def fke_movement(rho, gamma, vx, vy, rD, rV, pref, SD, SV):
    """One species' movement step (advdiff_FKE_MultiSp restated): fields [L, L]."""
    r = np.ascontiguousarray(rho, np.float64)
    L = r.shape[-1]
    out = np.empty(L * L)
    lib().tmo_fke_movement(L, r.reshape(-1), np.ascontiguousarray(gamma, np.float64).reshape(-1),
                           np.ascontiguousarray(vx, np.float64).reshape(-1),
                           np.ascontiguousarray(vy, np.float64).reshape(-1), rD, rV, pref,
                           np.ascontiguousarray(SD, np.float64), np.ascontiguousarray(SV, np.float64), out)
    return out.reshape(L, L)


def fke_perception(rho, gamma, v_eff, pR, dx, v, seed=1, inst=0, step=0):
    """tmo_fke_perception: returns (gamma, v_eff) updated copies."""
    r = np.ascontiguousarray(rho, np.float64)
    ns, L = r.shape[0], r.shape[-1]
    rmax = max(pR[:ns]) / dx
    disk = perception_offsets(int(rmax))
    g = np.ascontiguousarray(gamma, np.float64).copy()
    ve = np.ascontiguousarray(v_eff, np.float64).copy()
    lib().tmo_fke_perception(ns, L, np.ascontiguousarray(list(pR)[:ns], np.float64), dx,
                             np.ascontiguousarray(list(v)[:ns], np.float64), disk.ctypes.data, disk.shape[0],
                             seed, inst, step, r.reshape(-1), g.reshape(-1), ve.reshape(-1))
    return g, ve


def ref_perception_sites():
    """The site list printed by the reference's Test/test_perception_radii.cpp (built
    from source into oracle/_ref): the sites within radius 8 of (2, 2) on a 20 x 20
    lattice, in its computeNeighboringSitesCentral order; None when not built."""
    import re
    if not os.path.exists(RADII_PATH):
        return None
    out = subprocess.run([RADII_PATH], capture_output=True, text=True, check=True).stdout
    body = out.split(":", 1)[1]        # drop the "Neighboring sites for (2, 2)" title
    return [(int(a), int(b)) for a, b in re.findall(r"\((\d+), (\d+)\)", body)]


def laplacian(x):
    x = np.ascontiguousarray(x, np.float64)
    out = np.empty_like(x)
    lib().tmo_laplacian(x.shape[0], x.reshape(-1), out.reshape(-1))
    return out


def rk_rhs(p: RKParams, y):
    y = np.ascontiguousarray(y, np.float64)
    f = np.empty_like(y)
    lib().tmo_rk_rhs(C.byref(p), y.shape[-1], y.reshape(-1), f.reshape(-1))
    return f


def rk4_deterministic(p: RKParams, y, nsteps):
    y = np.ascontiguousarray(y, np.float64).copy()
    lib().tmo_rk4_deterministic(C.byref(p), y.shape[-1], nsteps, y.reshape(-1))
    return y


def dp_nonlinear(scheme, b, dt, rho):
    return lib().tmo_dp_nonlinear(int(scheme), float(b), float(dt), float(rho))


def flush(x, eps, neg):
    return lib().tmo_flush(float(x), float(eps), float(neg))


def reduce(fields):
    f = np.ascontiguousarray(fields, np.float64)
    nf, L = f.shape[0], f.shape[-1]
    row = np.empty(2 * nf, np.float64)
    lib().tmo_reduce(L, nf, f.reshape(-1), row)
    return row


def replicate_stats(series, params_index, n_params, flags):
    """Across-replicate statistics (tmo_replicate_stats): [n_params, n_meas, nf+3]."""
    x = np.ascontiguousarray(series, np.float64)
    n_inst, n_meas, w = x.shape
    pi = np.ascontiguousarray(params_index, np.int32)
    out = np.empty((n_params, n_meas, w // 2 + 3), np.float64)
    lib().tmo_replicate_stats(x.reshape(-1), n_inst, n_meas, w // 2, pi, n_params, int(flags), out.reshape(-1))
    return out


def ref_var_mean_surv(tab, x):
    """The reference's var_mean_incremental_surv_runs (stochasticSPDP.cpp:173-207) on
    tab[size][5] (updated in place) with one replicate's x[size][2]."""
    x = np.ascontiguousarray(x, np.float64)
    ref_lib().ref_ssp_var_mean_surv(tab, x, tab.shape[0])
    return tab


def ref_log10_time_bins(t_max, dt):
    """The reference's logarithm10_time_bins (stochasticSPDP.cpp:336-377)."""
    R = ref_lib()
    n = R.ref_ssp_log10_time_bins(t_max, dt, None, 0)
    out = np.empty(n, np.float64)
    if n:
        R.ref_ssp_log10_time_bins(t_max, dt, out.ctypes.data, n)
    return out


def init_speckle(L, perc, high, low, seed, inst):
    """Speckle frame [nf, L, L] (tmo_init_speckle)."""
    hi = np.ascontiguousarray(high, np.float64)
    lo = np.ascontiguousarray(low, np.float64)
    out = np.empty(hi.size * L * L, np.float64)
    lib().tmo_init_speckle(L, hi.size, seed, inst, float(perc), hi, lo, out)
    return out.reshape(hi.size, L, L)


def init_gauss(L, n_gauss, sd_div, amp):
    """Gaussian frame [nf, L, L] (tmo_init_gauss)."""
    sd = np.ascontiguousarray(sd_div, np.float64)
    a = np.ascontiguousarray(amp, np.float64)
    out = np.empty(a.size * L * L, np.float64)
    lib().tmo_init_gauss(L, a.size, int(n_gauss), sd, sd.size, a, out)
    return out.reshape(a.size, L, L)


_fmt = None


def ostream_text(x: float, prec: int = 0) -> str:
    """x as std::ostream prints it with setprecision(prec) (0: the default, 6)."""
    global _fmt
    if _fmt is None:
        if not os.path.exists(FMT_PATH):
            build()
        _fmt = C.CDLL(FMT_PATH)
        _fmt.fmt_ostream.argtypes = [C.c_double, C.c_int, C.c_char_p, C.c_int]
        _fmt.fmt_ostream.restype = C.c_int
    buf = C.create_string_buffer(64)
    n = _fmt.fmt_ostream(float(x), int(prec), buf, 64)
    return buf.raw[:n].decode()


# ---------------------------------------------------------------- consumer RK4 (FKE)
class FkeRkParams(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("a", "c", "gmax", "alpha", "rW", "W0")] + [
        ("K", C.c_double * 3), ("A", (C.c_double * 3) * 3), ("H", (C.c_double * 3) * 3), ("E", C.c_double * 3),
        ("M", C.c_double * 3), ("D_W", C.c_double), ("D_O", C.c_double), ("dt", C.c_double), ("dt_rk", C.c_double),
        ("dx", C.c_double)]


def fke_rk_params(a=1.0, c=10.0, gmax=0.05, alpha=0.2, rW=0.2, W0=0.2, K=(0.0, 5.0, 5.0), A=None, H=None,
                  E=(1.0, 0.45, 0.3), M=(0.0, 0.01, 0.01), D_W=0.1, D_O=100.0, dt=0.005, dt_rk=None,
                  dx=2.0) -> FkeRkParams:
    A = np.zeros((3, 3)) if A is None else np.asarray(A, np.float64)
    H = np.zeros((3, 3)) if H is None else np.asarray(H, np.float64)
    m3 = lambda m: ((C.c_double * 3) * 3)(*[(C.c_double * 3)(*row) for row in m])
    return FkeRkParams(a, c, gmax, alpha, rW, W0, (C.c_double * 3)(*K), m3(A), m3(H), (C.c_double * 3)(*E),
                       (C.c_double * 3)(*M), D_W, D_O, dt, dt if dt_rk is None else dt_rk, dx)


def fke_rk4(p: FkeRkParams, y):
    """tmo_fke_rk4: one step of y [ns + 3, L, L] (P, G, (Pr,) W, O); returns (y, n_bad)."""
    y = np.ascontiguousarray(y, np.float64).copy()
    n = lib().tmo_fke_rk4(C.cast(C.byref(p), C.c_void_p), y.shape[0] - 3, y.shape[-1], y.reshape(-1))
    return y, int(n)


# ------------------------------------------- the Rietkerk consumer model as one step
def fke_veg_dornic(p: RKParams, P, seed, inst, step):
    """tmo_fke_veg_dornic: P* [L, L] from P (p.d = the vegetation's M[0])."""
    P = np.ascontiguousarray(P, np.float64)
    L = P.shape[-1]
    out = np.empty(L * L)
    lib().tmo_fke_veg_dornic(C.byref(p), L, seed, inst, step, P.reshape(-1), out)
    return out.reshape(L, L)


def fke_movement_sep(rho, gamma, vx, vy, rD, rV, S, pref, SD, SV):
    """tmo_fke_movement_sep: one species' movement, factored by rows and columns."""
    r = np.ascontiguousarray(rho, np.float64)
    L = r.shape[-1]
    out = np.empty(L * L)
    f = lambda x: np.ascontiguousarray(x, np.float64).reshape(-1)
    lib().tmo_fke_movement_sep(L, f(r), f(gamma), f(vx), f(vy), rD, rV, S, pref, f(SD), f(SV), out)
    return out.reshape(L, L)


def fke_mov_stats(gamma, vx, vy):
    """tmo_fke_mov_stats: (mean, population variance) of gamma, means of vx and vy."""
    g = np.ascontiguousarray(gamma, np.float64)
    out = np.zeros(4)
    f = lambda x: np.ascontiguousarray(x, np.float64).reshape(-1)
    lib().tmo_fke_mov_stats(g.shape[-1], f(g), f(vx), f(vy), out)
    return out


def fke_shift_reach(pref, vmax):
    """The largest shift |m - c| of a source's advected mean, plus one."""
    return int(math.floor(vmax * abs(pref) + 0.5)) + 1


def fke_window(rD, rV, pref, vmax):
    """The largest reach of a species' movement: 2 W + 1 must fit the lattice."""
    return max(rD, rV + fke_shift_reach(pref, vmax))


def fke_step(p, SP, fields, seed, inst, step):
    """One step of the Rietkerk consumer model (FKE:4100-4462) composed from the
    operators above: perception from the old densities, the vegetation's Dornic
    step, the RK4 of the trophic and water terms, each consumer's movement (separable
    order), checks and flush.  p: an object with tm_fke_params' fields; SP = 2 or 3
    species; fields [2 SP + 1, L, L] = P, G, (Pr,) W, O, gamma_1 (, gamma_2).
    Returns (fields, fault, v_eff [SP, L, L, 2])."""
    f = np.ascontiguousarray(fields, np.float64).copy()
    L, ND = f.shape[-1], SP + 2
    g, ve = fke_perception(f[:SP], f[ND:], np.zeros((SP, L, L, 2)), list(p.pR)[:SP], p.dx, list(p.v)[:SP],
                           seed, inst, step)
    veg = rk_params(c=p.c, gmax=p.gmax, alpha=p.alpha, rW=p.rW, W0=p.W0, d=p.M[0], D=(p.D_P, p.D_W, p.D_O),
                    K=p.K, sigma=p.sigma[0], dt=p.dt, dx=p.dx, flush_eps=p.flush_eps, sampler=p.sampler)
    y = f[:ND].copy()
    y[0] = fke_veg_dornic(veg, f[0], seed, inst, step)
    rk = fke_rk_params(a=p.a, c=p.c, gmax=p.gmax, alpha=p.alpha, rW=p.rW, W0=p.W0, K=p.K, A=p.A, H=p.H, E=p.E,
                       M=p.M, D_W=p.D_W, D_O=p.D_O, dt=p.dt, dt_rk=p.dt_analytic, dx=p.dx)
    y, _ = fke_rk4(rk, y)
    for s in range(1, SP):
        (rD, rV, pref), SD, SV = fke_move_constants(p.D_mov[s], p.v[s], p.sigma[s], p.dt_analytic, p.dx,
                                                    p.dt_adv[s], p.mv_invariance, p.CI)
        drawn = SP == 3
        vx = ve[s, ..., 0] if drawn else np.zeros((L, L))
        vy = ve[s, ..., 1] if drawn else np.zeros((L, L))
        S = fke_shift_reach(pref, abs(p.v[s]) if drawn else 0.0)
        y[s] = fke_movement_sep(y[s], g[s - 1], vx, vy, rD, rV, S, pref, SD, SV)
    fault = bool(np.any((y < 0) | ~np.isfinite(y)))
    y = np.where((y < p.flush_eps) & (y >= 0.0), 0.0, y)
    if SP == 2:
        ve = np.zeros_like(ve)
    return np.concatenate([y, g]), fault, ve
