"""trophic_maelstorm_amd -- MI355X (gfx950) integrator for the Dornic / Rietkerk
stochastic step of Trophic_Maelstorm.

The compute path is libtm.so (hand-written HIP for gfx950) behind the C ABI in
include/tm/tm.h.  This module is a thin ctypes binding of that ABI with the
reference's parameter packs as Python dataclasses; it performs no computation
of its own and raises if libtm.so is missing or no gfx950 device is present.

Reference call sites (relative to the reference's simulations/ directory):
  rietkerk_DorFPE_2D_MultiSp   Rietkerk_FPE/rietkerk_bjork_FKE.cpp:3697
  dP_Dornic_2D_MultiSp         Percolation_FD/multiSPDP.cpp:2421
"""
from __future__ import annotations

import ctypes as C
import math
import os
from dataclasses import dataclass, field

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libtm.so")

MODEL_DP, MODEL_DP2, MODEL_RIETKERK, MODEL_DP3, MODEL_FKE2, MODEL_FKE3 = 1, 2, 3, 4, 5, 6
# fields per instance; the consumer models hold the densities, then the perception fields
NFIELDS = {MODEL_DP: 1, MODEL_DP2: 3, MODEL_RIETKERK: 3, MODEL_DP3: 5, MODEL_FKE2: 5, MODEL_FKE3: 7}
NSPECIES = {MODEL_DP: 1, MODEL_DP2: 2, MODEL_RIETKERK: 3, MODEL_DP3: 3, MODEL_FKE2: 4, MODEL_FKE3: 5}
DPC_MAX_RADIUS = 12
F64, F32 = 0, 1
SAMPLER_EXACT, SAMPLER_COMPAT = 0, 1
DP_EULER, DP_RK4, DP_EXACT = 0, 1, 2

STATS_INTDIV_VARIANCE, STATS_STOP_AT_EXTINCTION, STATS_REFERENCE = 1, 2, 3

TM_OK, TM_FAULT_NUMERIC = 0, 1
_STATUS = {0: "TM_OK", -1: "TM_ERR_ARG", -2: "TM_ERR_HIP", -3: "TM_ERR_NOMEM", -4: "TM_ERR_STATE",
           -5: "TM_ERR_NODEV", -6: "TM_ERR_TOO_LARGE", 1: "TM_FAULT_NUMERIC"}


class TMError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"{_STATUS.get(status, status)}: {msg}")
        self.status = status


class NumericFault(TMError):
    """A negative or non-finite value appeared (the reference's exit(3) path,
    rietkerk_bjork_FKE.cpp:4489-4560)."""


class _DP(C.Structure):
    _fields_ = [("a", C.c_double), ("b", C.c_double), ("D", C.c_double), ("sigma", C.c_double),
                ("dt", C.c_double), ("dx", C.c_double), ("a_end", C.c_double),
                ("ramp_steps", C.c_int64), ("flush_eps", C.c_double), ("flush_neg", C.c_double),
                ("scheme", C.c_int32), ("sampler", C.c_int32)]


class _RK(C.Structure):
    _fields_ = [("R", C.c_double), ("c", C.c_double), ("gmax", C.c_double), ("alpha", C.c_double),
                ("rW", C.c_double), ("W0", C.c_double), ("d", C.c_double),
                ("D", C.c_double * 3), ("K", C.c_double * 3), ("sigma", C.c_double),
                ("dt", C.c_double), ("dx", C.c_double), ("flush_eps", C.c_double),
                ("sampler", C.c_int32), ("reserved", C.c_int32)]


class _DPC(C.Structure):
    _fields_ = [("a", C.c_double), ("b", C.c_double), ("D", C.c_double * 3), ("sigma", C.c_double * 3),
                ("v", C.c_double * 3), ("M", C.c_double * 3), ("E", C.c_double * 3),
                ("A", (C.c_double * 3) * 3), ("H", (C.c_double * 3) * 3), ("pR", C.c_double * 3),
                ("dt", C.c_double), ("dx", C.c_double), ("a_end", C.c_double), ("ramp_steps", C.c_int64),
                ("flush_eps", C.c_double), ("flush_neg", C.c_double), ("sampler", C.c_int32),
                ("reserved", C.c_int32)]


class _FKE(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("a", "c", "gmax", "alpha", "rW", "W0")] + [
        ("K", C.c_double * 3), ("A", (C.c_double * 3) * 3), ("H", (C.c_double * 3) * 3), ("E", C.c_double * 3),
        ("M", C.c_double * 3), ("D_P", C.c_double), ("D_W", C.c_double), ("D_O", C.c_double),
        ("D_mov", C.c_double * 3), ("v", C.c_double * 3), ("sigma", C.c_double * 3), ("pR", C.c_double * 3),
        ("dt", C.c_double), ("dt_analytic", C.c_double), ("dt_adv", C.c_double * 3), ("dx", C.c_double),
        ("flush_eps", C.c_double), ("mv_invariance", C.c_int32), ("CI", C.c_int32), ("sampler", C.c_int32),
        ("reserved", C.c_int32)]


class _Inst(C.Structure):
    _fields_ = [("L", C.c_int32), ("params_index", C.c_int32), ("seed", C.c_uint32),
                ("instance_id", C.c_uint32), ("step0", C.c_int64)]


@dataclass
class DPParams:
    """dP_Dornic_2D_MultiSp(..., a, b, c, D[Sp], ..., sigma[], ..., dt, dx, ...)
    (Percolation_FD/MultiSPDP.h:344-346) plus the quench ramp a -> a_end."""
    a: float = 0.72
    b: float = 2.0
    D: float = 1.0
    sigma: float = 2.0 ** 0.5
    dt: float = 0.1
    dx: float = 1.0
    a_end: float | None = None
    ramp_steps: int = 0
    flush_eps: float = 1e-12
    flush_neg: float = -1e-80
    scheme: int = DP_EULER
    sampler: int = SAMPLER_EXACT

    def _c(self) -> _DP:
        return _DP(self.a, self.b, self.D, self.sigma, self.dt, self.dx,
                   self.a if self.a_end is None else self.a_end, self.ramp_steps, self.flush_eps,
                   self.flush_neg, self.scheme, self.sampler)


@dataclass
class RietkerkParams:
    """rietkerk_DorFPE_2D_MultiSp(..., a(=R), c, gmax, alpha, rW, W0, D[Sp], ..., K[3],
    sigma[], ..., M[0](=d), ..., dt, dx, ...) (Rietkerk_FPE/rietkerk_bjork_basic.h:450-452)."""
    R: float = 1.0
    c: float = 10.0
    gmax: float = 0.05
    alpha: float = 0.2
    rW: float = 0.2
    W0: float = 0.2
    d: float = 0.25
    D: tuple = (0.1, 0.1, 100.0)
    K: tuple = (0.0, 5.0, 5.0)
    sigma: float = 0.5
    dt: float = 0.005
    dx: float = 2.0
    flush_eps: float = 1e-10
    sampler: int = SAMPLER_EXACT

    def _c(self) -> _RK:
        return _RK(self.R, self.c, self.gmax, self.alpha, self.rW, self.W0, self.d,
                   (C.c_double * 3)(*self.D), (C.c_double * 3)(*self.K), self.sigma, self.dt,
                   self.dx, self.flush_eps, self.sampler, 0)


def _mat3(m):
    m = np.zeros((3, 3)) if m is None else np.asarray(m, np.float64)
    return ((C.c_double * 3) * 3)(*[(C.c_double * 3)(*row) for row in m])


def _vec3(v):
    v = list(v) + [0.0] * (3 - len(v))
    return (C.c_double * 3)(*v)


@dataclass
class DPCParams:
    """Vegetation + consumers: dP_Dornic_2D_MultiSp(..., a, b, c, D[Sp], v[SpB], sigma[], ...,
    A[SpB][SpB], H[SpB][SpB], E[SpB], M[SpB], pR[], dtV[SpB], ..., dt, dx, ...)
    (Percolation_FD/MultiSPDP.h:344-346) for MODEL_DP2 (vegetation, grazer) and MODEL_DP3
    (and predator).  Defaults: the grazer and predator values of
    Percolation_FD/order_2stocDP.cpp:29-67 with dx = 0.1 and a perception radius
    of 5 sites.  b is the vegetation self-limitation of the trophic RK4 (the
    reference passes a there: b = a reproduces its runs)."""
    a: float = 0.03
    b: float = 1e-6
    D: tuple = (0.00025 / 24.0, 0.0298, 0.05221)
    sigma: tuple = ((0.00025 / 24.0) ** 0.5 / 0.1, 1.0, 3.0)
    v: tuple = (0.0, 0.45427, 0.40587)
    M: tuple = (0.0, 0.061609 * 20.0 ** -0.25 / 8760.0, 0.01)
    E: tuple = (1.0, 0.45, 0.3)
    A: tuple | None = None
    H: tuple | None = None
    pR: tuple = (0.0, 0.521425, 0.521425)
    dt: float = 0.1
    dx: float = 0.1
    a_end: float | None = None
    ramp_steps: int = 0
    flush_eps: float = 1e-12
    flush_neg: float = -1e-80
    sampler: int = SAMPLER_EXACT

    def __post_init__(self):
        if self.A is None:
            aij = 3.6 * 10.0 ** -6.08 * 20.0 ** -0.37
            self.A = ((0.0, aij, 0.0), (aij, 0.0, 0.05), (0.0, 0.05, 0.0))
        if self.H is None:
            self.H = ((0.0, 1.0, 0.0), (1.0, 0.0, 1.0), (0.0, 1.0, 0.0))

    def _c(self) -> _DPC:
        return _DPC(self.a, self.b, _vec3(self.D), _vec3(self.sigma), _vec3(self.v), _vec3(self.M),
                    _vec3(self.E), _mat3(self.A), _mat3(self.H), _vec3(self.pR), self.dt, self.dx,
                    self.a if self.a_end is None else self.a_end, self.ramp_steps, self.flush_eps,
                    self.flush_neg, self.sampler, 0)


def perception_offsets(R: int) -> np.ndarray:
    """Offsets (row, column) within radius R in computeNeighboringSitesCentral's order."""
    buf = np.zeros(2 * (2 * R + 1) ** 2, np.int32)
    n = lib().tm_perception_offsets(int(R), buf.ctypes.data_as(C.POINTER(C.c_int32)), buf.size // 2)
    if n < 0:
        raise ValueError(f"bad radius {R}")
    return buf[: 2 * n].reshape(n, 2)


class _MoveSpecies(C.Structure):
    _fields_ = [("rD", C.c_int32), ("rV", C.c_int32), ("mu_prefactor", C.c_double)]


def fke_move_constants(D, v, sigma, dt_analytic, dx, dt_adv=0.0, mv_invariance=0, CI=4):
    """One consumer species' movement constants and stencils (set_global_FKE_params,
    precompute_gaussian): ((rD, rV, mu_prefactor), stencil_D, stencil_V)."""
    sp = _MoveSpecies()
    sd, sv = np.zeros(4096), np.zeros(4096)
    rc = lib().tm_fke_move_constants(D, v, sigma, dt_analytic, dt_adv, mv_invariance, dx, CI, C.byref(sp), _dptr(sd),
                                     sd.size, _dptr(sv), sv.size)
    if rc != 0:
        raise ValueError("tm_fke_move_constants: bad arguments or stencil too wide")
    return (sp.rD, sp.rV, sp.mu_prefactor), sd[: sp.rD + 1].copy(), sv[: sp.rV + 1].copy()


def fke_movement(rho, gamma, v_eff, species, stencils_D, stencils_V, device: int = 0):
    """The Rietkerk consumers' movement step on the GPU: rho, gamma [ns, L, L],
    v_eff [ns, L, L, 2]; species [(rD, rV, mu_prefactor)] and their stencils."""
    r = np.ascontiguousarray(rho, np.float64)
    ns, L = r.shape[0], r.shape[-1]
    g = np.ascontiguousarray(gamma, np.float64)
    v = np.ascontiguousarray(v_eff, np.float64)
    sp = (_MoveSpecies * ns)(*[_MoveSpecies(int(a), int(b), float(c)) for a, b, c in species])
    sd = np.ascontiguousarray(np.concatenate(stencils_D), np.float64)
    sv = np.ascontiguousarray(np.concatenate(stencils_V), np.float64)
    out = np.empty_like(r)
    rc = lib().tm_fke_movement_host(device, ns, L, C.cast(sp, C.c_void_p), _dptr(sd), _dptr(sv), _dptr(r), _dptr(g),
                                    _dptr(v), _dptr(out))
    if rc != 0:
        raise TMError(rc, "tm_fke_movement_host failed")
    return out


def fke_perception(rho, gamma, v_eff, pR, dx, v, seed=1, inst=0, step=0, device: int = 0):
    """One step of the Rietkerk consumers' perception (and, with three species,
    velocity draws) on the GPU: returns (gamma, v_eff), updated copies."""
    r = np.ascontiguousarray(rho, np.float64)
    ns, L = r.shape[0], r.shape[-1]
    g = np.ascontiguousarray(gamma, np.float64).copy()
    ve = np.ascontiguousarray(v_eff, np.float64).copy()
    pr = np.ascontiguousarray(list(pR)[:ns], np.float64)
    vv = np.ascontiguousarray(list(v)[:ns], np.float64)
    rc = lib().tm_fke_perception_host(device, ns, L, _dptr(pr), dx, _dptr(vv), seed, inst, step, _dptr(r), _dptr(g),
                                      _dptr(ve))
    if rc != 0:
        raise TMError(rc, "tm_fke_perception_host failed")
    return g, ve


class _FKERK(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("a", "c", "gmax", "alpha", "rW", "W0")] + [
        ("K", C.c_double * 3), ("A", (C.c_double * 3) * 3), ("H", (C.c_double * 3) * 3), ("E", C.c_double * 3),
        ("M", C.c_double * 3), ("D_W", C.c_double), ("D_O", C.c_double), ("dt", C.c_double), ("dt_rk", C.c_double),
        ("dx", C.c_double)]


@dataclass
class FKERKParams:
    """The Rietkerk consumer model's RK4 (f_DorFKE_2Sp / f_DorFKE_3Sp arguments,
    Rietkerk_FPE/rietkerk_bjork_FKE.cpp:2449-2474, :2611-2650).  a is the rainfall;
    D_W and D_O the diffusion constants of W and O; dt the RK4's step (its third
    stage) and dt_rk Dt_Analytic, which sets its half and sixth steps (None: dt).
    Defaults: the Rietkerk constants of RietkerkParams and the trophic constants
    of DPCParams."""
    a: float = 1.0
    c: float = 10.0
    gmax: float = 0.05
    alpha: float = 0.2
    rW: float = 0.2
    W0: float = 0.2
    K: tuple = (0.0, 5.0, 5.0)
    A: tuple | None = None
    H: tuple | None = None
    E: tuple = (1.0, 0.45, 0.3)
    M: tuple = (0.0, 0.061609 * 20.0 ** -0.25 / 8760.0, 0.01)
    D_W: float = 0.1
    D_O: float = 100.0
    dt: float = 0.005
    dt_rk: float | None = None
    dx: float = 2.0

    def __post_init__(self):
        d = DPCParams()
        if self.A is None:
            self.A = d.A
        if self.H is None:
            self.H = d.H

    def _c(self) -> _FKERK:
        return _FKERK(self.a, self.c, self.gmax, self.alpha, self.rW, self.W0, _vec3(self.K), _mat3(self.A),
                      _mat3(self.H), _vec3(self.E), _vec3(self.M), self.D_W, self.D_O, self.dt,
                      self.dt if self.dt_rk is None else self.dt_rk, self.dx)


@dataclass
class FKEParams:
    """The Rietkerk model with consumers (MODEL_FKE2: vegetation and grazer; MODEL_FKE3:
    and predator) as one step: rietkerk_DorFPE_2D_MultiSp's arguments
    (Rietkerk_FPE/rietkerk_bjork_FKE.cpp:3697-3699) with set_global_FKE_params'
    (:195-250).  The reference's D[Sp] is split into D_P (vegetation), D_mov
    (consumers' movement) and D_W, D_O.  Defaults: the values of
    Rietkerk_FPE/order_3stoc_test_rietkerk.cpp:20-89 (km, hr, kg), with
    Dt_Analytic the smallest advection time (:47-48, FKE:197-203) and dt = 0.1 hr."""
    a: float = 1.0 / 24.0
    c: float = 10000.0
    gmax: float = 0.05 * 10.0 ** -3.0 / 24.0
    alpha: float = 0.2 / 24.0
    rW: float = 0.2 / 24.0
    W0: float = 0.2
    K: tuple = (0.0, 5.0, 5000.0)
    A: tuple | None = None
    H: tuple = ((0.0, 1.0, 0.0), (1.0, 0.0, 1.0), (0.0, 1.0, 0.0))
    E: tuple = (1.0, 0.45, 0.85)
    M: tuple = (0.25 / 24.0, 0.061609 * 20.0 ** -0.25 / 8760.0, 0.061609 * 100.0 ** -0.25 / 8760.0)
    D_P: float = 0.00025 / 24.0
    D_W: float = 0.00025 / 24.0
    D_O: float = 0.025 / 24.0
    D_mov: tuple = (0.0, 0.01028, 0.215965)
    v: tuple = (0.0, 0.45427, 0.40587)
    sigma: tuple = (math.sqrt((0.00025 / 24.0) / (0.1 * 0.1)), 1.0, 10.0)
    pR: tuple = (0.0, 1.04285 / 2.0, 1.25536 / 2.0)
    dt: float = 0.1
    dt_analytic: float = 0.11817455
    dt_adv: tuple = (0.0, 0.11817455, 0.765270868)
    dx: float = 0.1
    flush_eps: float = 1e-10
    mv_invariance: int = 0
    CI: int = 4
    sampler: int = SAMPLER_EXACT

    def __post_init__(self):
        if self.A is None:
            aij = 3.6 * 10.0 ** -6.08 * 20.0 ** -0.37
            ajm = 3.6 * 10.0 ** -6.08 * 100.0 ** -0.37
            self.A = ((0.0, aij, 0.0), (aij, 0.0, ajm), (0.0, ajm, 0.0))

    def _c(self) -> _FKE:
        return _FKE(self.a, self.c, self.gmax, self.alpha, self.rW, self.W0, _vec3(self.K), _mat3(self.A),
                    _mat3(self.H), _vec3(self.E), _vec3(self.M), self.D_P, self.D_W, self.D_O, _vec3(self.D_mov),
                    _vec3(self.v), _vec3(self.sigma), _vec3(self.pR), self.dt, self.dt_analytic,
                    _vec3(self.dt_adv), self.dx, self.flush_eps, self.mv_invariance, self.CI, self.sampler, 0)


def fke_rk4(fields, params: FKERKParams, device: int = 0):
    """One RK4 step of the Rietkerk consumer model on the GPU (tm_fke_rk4_host):
    fields [ns + 3, L, L] in the order P, G, (Pr,) W, O.  Returns (fields, n_bad):
    an updated copy and the number of values left negative or non-finite."""
    f = np.ascontiguousarray(fields, np.float64).copy()
    ns, L = f.shape[0] - 3, f.shape[-1]
    p = params._c()
    bad = C.c_int64(0)
    rc = lib().tm_fke_rk4_host(device, ns, L, C.cast(C.byref(p), C.c_void_p), _dptr(f), C.byref(bad))
    if rc != 0:
        raise TMError(rc, "tm_fke_rk4_host failed")
    return f, int(bad.value)


def fke_veg_dornic(P, params: RietkerkParams, seed: int, inst: int, step: int, device: int = 0):
    """The Rietkerk consumer model's vegetation Dornic sub-step on the GPU
    (tm_fke_veg_dornic_host): P* from P [L, L].  params.d is the vegetation's
    mortality M[0]; D[0], sigma, dt, dx and sampler enter."""
    p = np.ascontiguousarray(P, np.float64)
    out = np.empty_like(p)
    c = params._c()
    rc = lib().tm_fke_veg_dornic_host(device, p.shape[-1], C.cast(C.byref(c), C.c_void_p), seed, inst, step,
                                      _dptr(p), _dptr(out))
    if rc != 0:
        raise TMError(rc, "tm_fke_veg_dornic_host failed")
    return out


@dataclass
class Instance:
    L: int
    params_index: int = 0
    seed: int = 1
    instance_id: int = 0
    step0: int = 0


_lib = None


def lib() -> C.CDLL:
    """Load libtm.so; raises when it has not been built (no silent fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: run trophic_maelstorm_amd.build.build() "
                           "(hipcc, gfx950) before using the HIP integrator")
    L = C.CDLL(LIB_PATH)
    vp, i64, i32, dp = C.c_void_p, C.c_int64, C.c_int32, C.POINTER(C.c_double)
    sig = {
        "tm_create": ([C.POINTER(vp), C.c_int], C.c_int),
        "tm_destroy": ([vp], None),
        "tm_last_error": ([vp], C.c_char_p),
        "tm_configure": ([vp, C.c_int, C.c_int, vp, C.c_int, vp, C.c_int], C.c_int),
        "tm_upload_fields": ([vp, C.c_int, dp], C.c_int),
        "tm_download_fields": ([vp, C.c_int, dp], C.c_int),
        "tm_init_uniform": ([vp, dp, dp], C.c_int),
        "tm_run": ([vp, i64, vp], C.c_int),
        "tm_measure": ([vp, dp], C.c_int),
        "tm_mov_stats": ([vp, dp], C.c_int),
        "tm_set_record_draws": ([vp, C.c_int], C.c_int),
        "tm_download_draws": ([vp, C.c_int, C.POINTER(i64)], C.c_int),
        "tm_fault_info": ([vp, C.POINTER(i32), C.POINTER(i64), C.POINTER(i64)], C.c_int),
        "tm_get_step": ([vp, C.c_int, C.POINTER(i64)], C.c_int),
        "tm_reset": ([vp], C.c_int),
        "tm_set_launch": ([vp, C.c_int, C.c_int], C.c_int),
        "tm_get_launch": ([vp, C.POINTER(C.c_int), C.POINTER(C.c_int)], C.c_int),
        "tm_variant_count": ([C.c_int, C.c_int], C.c_int),
        "tm_variant_name": ([C.c_int, C.c_int, C.c_int], C.c_char_p),
        "tm_autotune": ([vp, C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int)], C.c_int),
        "tm_time_steps": ([vp, i64, dp, C.POINTER(i64)], C.c_int),
        "tm_last_run_time": ([vp, dp, C.POINTER(i64)], C.c_int),
        "tm_stream": ([vp], vp),
        "tm_mode_steps": ([vp, C.POINTER(i64), C.POINTER(i64)], C.c_int),
        "tm_log10_time_bins": ([C.c_double, C.c_double, dp, C.c_int], C.c_int),
        "tm_measure_steps": ([dp, C.c_int, C.c_double, C.POINTER(i64)], C.c_int),
        "tm_run_series": ([vp, C.POINTER(i64), C.c_int, vp], C.c_int),
        "tm_replicate_stats": ([vp, C.c_int, dp], C.c_int),
        "tm_init_speckle": ([vp, dp, dp, dp], C.c_int),
        "tm_init_gauss": ([vp, C.c_int, dp, C.c_int, dp], C.c_int),
        "tm_mft_frame_values": ([C.c_int, C.c_double, C.c_double, C.c_double, C.c_int, dp, dp, dp, dp],
                                C.c_int),
        "tm_write_frame_csv": ([vp, C.c_int, C.c_char_p, C.c_double, C.c_char_p, C.c_int], C.c_int),
        "tm_format_frame_csv": ([C.c_char_p, dp, C.c_int, C.c_int, C.c_double, C.c_char_p, C.c_int], C.c_int),
        "tm_write_prelim_csv": ([C.c_char_p, C.c_char_p, C.c_double, C.c_int, C.c_int, dp, dp, C.c_int,
                                 C.c_int], C.c_int),
        "tm_write_sweep_csv": ([C.c_char_p, C.c_char_p, dp, C.c_int, C.c_int], C.c_int),
        "tm_save_snapshot": ([vp, C.c_int, C.c_char_p], C.c_int),
        "tm_load_snapshot": ([vp, C.c_int, C.c_char_p, C.POINTER(i64)], C.c_int),
        "tm_snapshot_to_frame_csv": ([C.c_char_p, C.c_char_p, C.c_double, C.c_char_p, C.c_int], C.c_int),
        "tm_read_frame_csv": ([C.c_char_p, C.c_int, C.c_int, C.POINTER(C.c_char_p), C.c_int, dp], C.c_int),
        "tm_perception_offsets": ([C.c_int, C.POINTER(i32), C.c_int], C.c_int),
        "tm_fke_move_constants": ([C.c_double] * 5 + [C.c_int, C.c_double, C.c_int, vp, dp, C.c_int, dp, C.c_int],
                                  C.c_int),
        "tm_fke_movement_host": ([C.c_int, C.c_int, C.c_int, vp, dp, dp, dp, dp, dp, dp], C.c_int),
        "tm_fke_movement": ([C.c_int, C.c_int, vp, vp, vp, vp, vp, vp, vp, vp], C.c_int),
        "tm_fke_perception_host": ([C.c_int, C.c_int, C.c_int, dp, C.c_double, dp, C.c_uint32, C.c_uint32, i64, dp,
                                    dp, dp], C.c_int),
        "tm_fke_rk4_host": ([C.c_int, C.c_int, C.c_int, C.c_void_p, dp, C.POINTER(C.c_int64)], C.c_int),
        "tm_fke_rk4": ([C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.POINTER(C.c_int64), C.c_void_p], C.c_int),
        "tm_fke_rk4_noalloc": ([C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.POINTER(C.c_int64), C.c_void_p],
                               C.c_int),
        "tm_fke_veg_dornic_host": ([C.c_int, C.c_int, C.c_void_p, C.c_uint32, C.c_uint32, i64, dp, dp], C.c_int),
        "tm_fke_veg_dornic": ([C.c_int, C.c_void_p, C.c_uint32, C.c_uint32, i64, C.c_void_p, C.c_void_p, C.c_void_p],
                              C.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def variant_names(model: int, precision: int) -> list[str]:
    L = lib()
    return [L.tm_variant_name(model, precision, v).decode()
            for v in range(L.tm_variant_count(model, precision))]


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def log10_time_bins(t_max: float, dt: float) -> np.ndarray:
    """The reference's logarithm10_time_bins(t_max, dt) (rietkerk_bjork_FKE.cpp:1497-1536)."""
    n = lib().tm_log10_time_bins(t_max, dt, None, 0)
    out = np.empty(n, np.float64)
    if n:
        lib().tm_log10_time_bins(t_max, dt, _dptr(out), n)
    return out


def sweep_times(t_max: float, dt: float) -> np.ndarray:
    """The sweep's t_measure: the log-spaced bins with 0 in front, sorted, repeats
    removed (rietkerk_bjork_FKE.cpp:4676-4703, multiSPDP.cpp:3469-3490)."""
    return np.unique(np.concatenate([[0.0], log10_time_bins(t_max, dt)]))


def measure_steps(t_meas, dt: float) -> np.ndarray:
    """Steps after which the reference's time loop logs each t_meas[i]
    (rietkerk_bjork_FKE.cpp:4005-4017); raises if a point is never reached."""
    t = np.ascontiguousarray(t_meas, np.float64)
    out = np.empty(t.size, np.int64)
    n = lib().tm_measure_steps(_dptr(t), t.size, dt, out.ctypes.data_as(C.POINTER(C.c_int64)))
    if n != t.size:
        raise ValueError(f"measurement time {t[n]} is not reached on the dt = {dt} grid")
    return out


def sweep_values(start: float, end: float, num: int) -> np.ndarray:
    """a_space = linspace(a_start, a_end, div) (rietkerk_bjork_FKE.cpp:1374-1398,
    start + delta i with the last point = end, as numpy.linspace)."""
    return np.linspace(start, end, num)


def mft_frame_values(homogeneous: bool, R: float, R_c: float, dP: float, c_spread, mft):
    """(high, low) per field of the reference's mean-field initial frames at one
    parameter point (rietkerk_bjork_FKE.cpp:453-642); c_spread and mft have 2*nf
    entries (below R_c, then at or above)."""
    cs = np.ascontiguousarray(c_spread, np.float64)
    e = np.ascontiguousarray(mft, np.float64)
    if cs.size != e.size or cs.size % 2:
        raise ValueError("c_spread and mft need 2*nf entries each")
    nf = cs.size // 2
    hi, lo = np.empty(nf), np.empty(nf)
    if lib().tm_mft_frame_values(int(homogeneous), R, R_c, dP, nf, _dptr(cs), _dptr(e), _dptr(hi), _dptr(lo)):
        raise ValueError("tm_mft_frame_values: bad arguments")
    return hi, lo


# The reference's file headers (Rietkerk_FPE/rietkerk_bjork_constants_1Sp.h:6-9,
# Percolation_FD/MultiSPDP_constants_1Sp.h:6-9).
FRAME_HEADER = {
    MODEL_RIETKERK: "a_c,  x,  P(x; t), W(x; t), O(x; t) \n",
    MODEL_DP: "a_c,  x,  P(x; t) \n",
    MODEL_DP2: "a_c,  x,  P(x; t), G(x; t) \n",               # MultiSPDP_constants_2Sp.h:6
    MODEL_DP3: "a_c,  x,  P(x; t), G(x; t), Pr(x; t)\n",      # MultiSPDP_constants_3Sp.h:6
}
PRELIM_HEADER = {
    MODEL_RIETKERK: (" a , r, L, t , <<W(x; t)>_x>_r, <<O(x; t)>_x>_r,  <<P(x; t)>_x>_r, Var[<P(x; t)>_x]_r, "
                     "# Surviving Runs P(x; t), # Active Sites P(x; t), \n"),
    MODEL_DP: (" a , r, L, t , <<P(x; t)>_x>_r, Var[<P(x; t)>_x]_r, # Surviving Runs P(x; t),"
               " # Active Sites P(x; t), \n"),
    # MultiSPDP_constants_2Sp.h:10-11, MultiSPDP_constants_3Sp.h:10-12
    MODEL_DP2: (" a , r, L, t , <<P(x; t)>_x>_r, Var[<P(x; t)>_x]_r, # Surviving Runs P(x; t),"
                " # Active Sites P(x; t), <<G(x; t)>_x>_r, Var[<G(x; t)>_x]_r, # Surviving Runs G(x; t),"
                " # Active Sites G(x; t), \n"),
    MODEL_DP3: (" a , r, L, t , <<P(x; t)>_x>_r, Var[<P(x; t)>_x]_r, # Surviving Runs P(x; t),"
                " # Active Sites P(x; t), <<G(x; t)>_x>_r, Var[<G(x; t)>_x]_r, # Surviving Runs G(x; t),"
                " # Active Sites G(x; t),<<Pr(x; t)>_x>_r, Var[<Pr(x; t)>_x]_r, # Surviving Runs Pr(x; t),"
                " # Active Sites Pr(x; t), \n"),
}


def _b(x):
    return None if x is None else str(x).encode()


def _io_check(rc: int, what: str):
    if rc < 0:
        raise OSError(f"{what} failed ({rc})")
    return rc


def format_frame_csv(path, fields, a: float, header: str | None = None, digits: int = 6):
    """FRAME file from a host frame [nf, L, L] (save_frame's layout)."""
    f = np.ascontiguousarray(fields, np.float64)
    nf, L = f.shape[0], f.shape[-1]
    _io_check(lib().tm_format_frame_csv(_b(path), _dptr(f), nf, L, a, _b(header), digits), "tm_format_frame_csv")


def write_prelim_csv(path, header, a: float, r: int, g: int, t, stats):
    """PRELIM file: rows a, r+1, g, t[i], stats[i, :] (save_prelimframe's layout)."""
    tt = np.ascontiguousarray(t, np.float64)
    st = np.ascontiguousarray(stats, np.float64).reshape(tt.size, -1)
    _io_check(lib().tm_write_prelim_csv(_b(path), _b(header), a, r, g, _dptr(tt), _dptr(st), tt.size, st.shape[1]),
              "tm_write_prelim_csv")


def write_sweep_csv(path, header, rows):
    """The sweep's final file: header line, then rows with precision 16."""
    x = np.ascontiguousarray(rows, np.float64)
    _io_check(lib().tm_write_sweep_csv(_b(path), _b(header), _dptr(x), x.shape[0], x.shape[1]),
              "tm_write_sweep_csv")


def snapshot_to_frame_csv(snap, csv, a: float, header: str | None = None, digits: int = 6):
    _io_check(lib().tm_snapshot_to_frame_csv(_b(snap), _b(csv), a, _b(header), digits), "tm_snapshot_to_frame_csv")


def read_frame_csv(path, L: int, nf: int, cols, out=None):
    """Burn-in input: the named FRAME-file columns into fields [nf, L, L]; returns
    (fields, rows read).  Fields not named keep the values of `out` (zeros)."""
    f = np.zeros((nf, L, L)) if out is None else np.ascontiguousarray(out, np.float64).copy()
    names = (C.c_char_p * len(cols))(*[c.encode() for c in cols])
    n = lib().tm_read_frame_csv(_b(path), L, nf, names, len(cols), _dptr(f))
    if n == -2:
        raise KeyError(f"a column of {list(cols)} is missing in {path}")
    _io_check(n, "tm_read_frame_csv")
    return f, n


# Column names of Context.replicate_stats rows, per model (field 0 is P / rho).
STATS_COLUMNS = {
    MODEL_DP: ["mean", "var", "surviving", "active"],
    MODEL_RIETKERK: ["mean_W", "mean_O", "mean", "var", "surviving", "active"],
    MODEL_DP2: [f"{c}_{s}" for s in ("P", "G") for c in ("mean", "var", "surviving", "active")],
    MODEL_DP3: [f"{c}_{s}" for s in ("P", "G", "Pr") for c in ("mean", "var", "surviving", "active")],
}


class Context:
    """One tm_ctx: a batch of instances of one model at one precision on one GPU."""

    def __init__(self, device: int = 0):
        self._lib = lib()
        h = C.c_void_p()
        st = self._lib.tm_create(C.byref(h), device)
        if st != TM_OK:
            raise TMError(st, f"tm_create(device={device}) failed: a gfx950 GPU is required")
        self._h = h
        self.model = None
        self.instances: list[Instance] = []
        self._n_params = self._n_series = 0

    def close(self):
        if getattr(self, "_h", None):
            self._lib.tm_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, st: int):
        if st == TM_OK:
            return
        msg = self._lib.tm_last_error(self._h).decode()
        if st == TM_FAULT_NUMERIC:
            raise NumericFault(st, msg)
        raise TMError(st, msg)

    @property
    def nfields(self) -> int:
        return NFIELDS[self.model]

    @property
    def nspecies(self) -> int:
        return NSPECIES[self.model]

    def configure(self, model: int, precision: int, params, instances):
        plist = params if isinstance(params, (list, tuple)) else [params]
        ctype = {MODEL_DP: _DP, MODEL_DP2: _DPC, MODEL_DP3: _DPC, MODEL_FKE2: _FKE, MODEL_FKE3: _FKE}.get(model, _RK)
        parr = (ctype * len(plist))(*[p._c() for p in plist])
        iarr = (_Inst * len(instances))(*[_Inst(i.L, i.params_index, i.seed, i.instance_id, i.step0)
                                          for i in instances])
        self._check(self._lib.tm_configure(self._h, model, precision, C.cast(parr, C.c_void_p),
                                           len(plist), C.cast(iarr, C.c_void_p), len(instances)))
        self.model, self.precision, self.instances = model, precision, list(instances)
        self._n_params, self._n_series = len(plist), 0
        return self

    def upload(self, i: int, fields: np.ndarray):
        f = np.ascontiguousarray(fields, np.float64)
        L = self.instances[i].L
        if f.size != self.nfields * L * L:
            raise ValueError(f"instance {i}: expected {self.nfields}x{L}x{L} values, got {f.shape}")
        self._check(self._lib.tm_upload_fields(self._h, i, _dptr(f)))

    def download(self, i: int) -> np.ndarray:
        L = self.instances[i].L
        out = np.empty((self.nfields, L, L), np.float64)
        self._check(self._lib.tm_download_fields(self._h, i, _dptr(out)))
        return out

    def init_uniform(self, lo, hi):
        lo = np.ascontiguousarray(lo, np.float64)
        hi = np.ascontiguousarray(hi, np.float64)
        self._check(self._lib.tm_init_uniform(self._h, _dptr(lo), _dptr(hi)))

    def init_speckle(self, perc, high, low):
        """Speckle frames: per parameter point perc[p], high[p][f], low[p][f] (a single
        point's values are broadcast to every point)."""
        n_p, nf = self._n_params, self.nfields
        pc = np.ascontiguousarray(np.broadcast_to(np.asarray(perc, np.float64), (n_p,)))
        hi = np.ascontiguousarray(np.broadcast_to(np.asarray(high, np.float64), (n_p, nf)))
        lo = np.ascontiguousarray(np.broadcast_to(np.asarray(low, np.float64), (n_p, nf)))
        self._check(self._lib.tm_init_speckle(self._h, _dptr(pc), _dptr(hi), _dptr(lo)))

    def init_gauss(self, amp, n_gauss: int = 1, sd_div=(8.0, 16.0)):
        """Gaussian frames: fields < n_gauss are bumps of height amp[p][f] and width
        L / sd_div[f] at the corners, the others the constant amp[p][f]."""
        n_p, nf = self._n_params, self.nfields
        a = np.ascontiguousarray(np.broadcast_to(np.asarray(amp, np.float64), (n_p, nf)))
        sd = np.ascontiguousarray(sd_div, np.float64)
        self._check(self._lib.tm_init_gauss(self._h, int(n_gauss), _dptr(sd), sd.size, _dptr(a)))

    def run(self, n_steps: int, rows: bool = True):
        """Advance n_steps; returns rows [n_inst, n_steps, 2*nf] of (#>0, mean) per field."""
        out = np.empty((len(self.instances), n_steps, 2 * self.nfields), np.float64) if rows else None
        self._check(self._lib.tm_run(self._h, n_steps, out.ctypes.data if out is not None else None))
        return out

    def run_series(self, steps, copy: bool = True):
        """Advance to steps[-1] (from now), keeping every instance's rho(t) row at each
        steps[i] on the device; returns the series [n_inst, n_meas, 2*nf] (or None)."""
        st = np.ascontiguousarray(steps, np.int64)
        out = np.empty((len(self.instances), st.size, 2 * self.nfields), np.float64) if copy else None
        self._check(self._lib.tm_run_series(self._h, st.ctypes.data_as(C.POINTER(C.c_int64)), st.size,
                                            out.ctypes.data if out is not None else None))
        self._n_series = st.size
        return out

    def replicate_stats(self, flags: int = STATS_REFERENCE) -> np.ndarray:
        """Across-replicate statistics of the last run_series, computed on the device:
        [n_params, n_meas, nf+3] (consumer models: [.., 4 * species]), columns
        STATS_COLUMNS[model]."""
        ncol = 4 * self.nspecies if self.model in (MODEL_DP2, MODEL_DP3) else self.nfields + 3
        out = np.empty((self._n_params, self._n_series, ncol), np.float64)
        self._check(self._lib.tm_replicate_stats(self._h, flags, _dptr(out)))
        return out

    def write_frame_csv(self, i: int, path, a: float, header: str | None = None, digits: int = 6):
        self._check(self._lib.tm_write_frame_csv(self._h, i, _b(path), a, _b(header), digits))

    def save_snapshot(self, i: int, path):
        self._check(self._lib.tm_save_snapshot(self._h, i, _b(path)))

    def load_snapshot(self, i: int, path) -> int:
        """Load instance i's fields from a snapshot; returns the recorded step."""
        st = C.c_int64()
        self._check(self._lib.tm_load_snapshot(self._h, i, _b(path), C.byref(st)))
        return st.value

    def measure(self) -> np.ndarray:
        out = np.empty((len(self.instances), 2 * self.nfields), np.float64)
        self._check(self._lib.tm_measure(self._h, _dptr(out)))
        return out

    def mov_stats(self) -> np.ndarray:
        """MODEL_FKE2 / MODEL_FKE3: the Rho_Mov columns of the current state,
        [n_inst, 4 SpB] (tm_mov_stats)."""
        sp = 3 if self.model == MODEL_FKE3 else 2
        out = np.empty((len(self.instances), 4 * sp), np.float64)
        self._check(self._lib.tm_mov_stats(self._h, _dptr(out)))
        return out

    def record_draws(self, enable: bool = True):
        self._check(self._lib.tm_set_record_draws(self._h, int(enable)))

    def draws(self, i: int) -> np.ndarray:
        L = self.instances[i].L
        out = np.empty((L, L), np.int64)
        self._check(self._lib.tm_download_draws(self._h, i, out.ctypes.data_as(C.POINTER(C.c_int64))))
        return out

    def fault_info(self):
        inst, cell, step = C.c_int32(), C.c_int64(), C.c_int64()
        self._check(self._lib.tm_fault_info(self._h, C.byref(inst), C.byref(cell), C.byref(step)))
        return inst.value, cell.value, step.value

    def step(self, i: int = 0) -> int:
        s = C.c_int64()
        self._check(self._lib.tm_get_step(self._h, i, C.byref(s)))
        return s.value

    def reset(self):
        """Rewind every instance to its step0 and clear the fault record (fields untouched)."""
        self._check(self._lib.tm_reset(self._h))

    def set_launch(self, variant: int = -1, graph_steps: int = 100):
        self._check(self._lib.tm_set_launch(self._h, variant, graph_steps))

    def launch(self):
        """(variant index, steps per launch block) in use."""
        v, g = C.c_int(), C.c_int()
        self._check(self._lib.tm_get_launch(self._h, C.byref(v), C.byref(g)))
        return v.value, g.value

    def autotune(self, cache_path: str | None = None):
        v, g = C.c_int(), C.c_int()
        self._check(self._lib.tm_autotune(self._h, cache_path.encode() if cache_path else None,
                                          C.byref(v), C.byref(g)))
        return v.value, g.value

    def last_run_ms(self) -> float:
        """Device ms of the last run()'s launches (HIP events on the context stream
        around them; the rows copy to the host is outside)."""
        ms = C.c_double()
        self._check(self._lib.tm_last_run_time(self._h, C.byref(ms), None))
        return ms.value

    def mode_steps(self):
        """(single-step, multi-step) steps run() has advanced since configure."""
        a, b = C.c_int64(), C.c_int64()
        self._check(self._lib.tm_mode_steps(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def time_steps(self, n_steps: int) -> float:
        """Device ms of n_steps step-kernel launches (HIP events on the context stream)."""
        ms, n = C.c_double(), C.c_int64()
        self._check(self._lib.tm_time_steps(self._h, n_steps, C.byref(ms), C.byref(n)))
        return ms.value
