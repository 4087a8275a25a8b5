"""The Rietkerk consumer model's RK4 on the GPU (tm_fke_rk4, tm_fke_rk4_host) against the
oracle's restatement of RK4_Integrate_Stochastic_2Sp / _MultiSp
(Rietkerk_FPE/rietkerk_bjork_FKE.cpp:2476-2533, :2652-2720): bit for bit in FP64."""
import ctypes as C

import numpy as np
import pytest

from tests.test_fke_rk4 import A2, H2, fields

pytestmark = pytest.mark.gpu


def _params(tm, oracle, **kw):
    kw.setdefault("M", (0.0, 0.02, 0.03))     # the two wrappers' defaults differ; pass the same values
    return tm.FKERKParams(A=A2, H=H2, **kw), oracle.fke_rk_params(A=A2, H=H2, **kw)


@pytest.mark.parametrize("ns,L", [(1, 48), (2, 48), (1, 37), (2, 5)])
def test_rk4_matches_oracle_bit_for_bit(tm, oracle, ns, L):
    print(f"fke rk4 ns={ns} L={L}", flush=True)
    kw = dict(a=1.2, M=(0.0, 0.02, 0.03), D_W=0.1, D_O=0.5, dt=0.004, dt_rk=0.005, dx=2.0)
    pg, po = _params(tm, oracle, **kw)
    y = fields(ns, L, 21 + L)
    got, bad = tm.fke_rk4(y, pg)
    ref, rbad = oracle.fke_rk4(po, y)
    assert np.array_equal(got, ref)
    assert bad == rbad == 0
    # a few steps in a row stay equal
    for _ in range(3):
        got, _ = tm.fke_rk4(got, pg)
        ref, _ = oracle.fke_rk4(po, ref)
    assert np.array_equal(got, ref)


def test_negative_values_are_counted(tm, oracle):
    pg, po = _params(tm, oracle, a=-50.0, D_O=0.1, dt=0.05)
    y = fields(1, 16, 3)
    got, bad = tm.fke_rk4(y, pg)
    ref, rbad = oracle.fke_rk4(po, y)
    assert np.array_equal(got, ref)
    assert bad == rbad >= 256


def test_device_entry_equals_host_entry(tm):
    """tm_fke_rk4 on device arrays (hipMalloc'd here through libamdhip64) gives the host entry's result."""
    hip = C.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    H2D, D2H = 1, 2
    pg = tm.FKERKParams(A=A2, H=H2, M=(0.0, 0.02, 0.03), dt=0.004)
    y = np.ascontiguousarray(fields(2, 40, 8))
    want, wbad = tm.fke_rk4(y, pg)
    p = C.c_void_p()
    assert hip.hipMalloc(C.byref(p), y.nbytes) == 0
    try:
        assert hip.hipMemcpy(p, y.ctypes.data, y.nbytes, H2D) == 0
        par = pg._c()
        bad = C.c_int64(-1)
        rc = tm.lib().tm_fke_rk4(2, 40, C.cast(C.byref(par), C.c_void_p), p, C.byref(bad), None)
        assert rc == 0
        got = np.empty_like(y)
        assert hip.hipMemcpy(got.ctypes.data, p, got.nbytes, D2H) == 0
    finally:
        hip.hipFree(p)
    assert np.array_equal(got, want) and bad.value == wbad == 0


def test_rejects_bad_arguments(tm):
    pg = tm.FKERKParams()
    with pytest.raises(tm.TMError, match="TM_ERR_ARG"):
        tm.fke_rk4(np.ones((6, 8, 8)), pg)          # ns = 3 consumers is not a model
    with pytest.raises(tm.TMError, match="TM_ERR_ARG"):
        tm.fke_rk4(np.ones((4, 8, 8)), tm.FKERKParams(dx=0.0))


@pytest.mark.parametrize("sampler,L", [(0, 40), (1, 33)])
def test_vegetation_dornic_matches_oracle(tm, oracle, sampler, L):
    """tm_fke_veg_dornic against the oracle's Rietkerk step with no plant growth
    (c = 0): its RK4 then leaves P at P* exactly, so its P after one step is P* with
    the step's flush (below 1e-10 to 0), which the check applies to the operator's
    output.  Includes an absorbing patch, large densities (the compat sampler's
    Gaussian shortcuts above 100) and a step counter above 2^32."""
    rng = np.random.default_rng(L)
    f = np.stack([rng.uniform(0, 90, (L, L)), rng.random((L, L)), rng.random((L, L))])
    f[0, 3:9, 4:12] = 0.0
    f[0, 20:24, 2:6] = 400.0
    d, D, sigma, dt, dx = 0.25, (0.1, 0.1, 100.0), 0.5, 0.005, 2.0
    step = (1 << 32) + 17
    print(f"veg dornic sampler={sampler} L={L}", flush=True)
    po = oracle.rk_params(c=0.0, d=d, D=D, sigma=sigma, dt=dt, dx=dx, sampler=sampler)
    ref = oracle.rk_run(po, f, L, 7, 3, step, 1, rows=False)
    assert ref["status"] == 0
    pg = tm.RietkerkParams(c=0.0, d=d, D=D, sigma=sigma, dt=dt, dx=dx, sampler=sampler)
    got = tm.fke_veg_dornic(f[0], pg, 7, 3, step)
    got = np.where((got < 1e-10) & (got >= 0.0), 0.0, got)
    assert np.array_equal(got, ref["fields"][0])
    assert np.all(got[4:8, 5:11] == 0.0)
