"""The allocation-free entry point tm_fke_rk4_noalloc against tm_fke_rk4, through
the C ABI on device arrays: bit-identical fields and negative-value counts on two
calls in a row, the second on a smaller lattice (the library's fixed scratch keeps
the first call's contents), and a lattice beyond the scratch is refused."""
import ctypes as C

import numpy as np
import pytest

from tests.test_fke_rk4 import A2, H2, fields

pytestmark = pytest.mark.gpu
H2D, D2H = 1, 2


@pytest.fixture(scope="module")
def hip():
    h = C.CDLL("libamdhip64.so")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipFree.argtypes = [C.c_void_p]
    return h


def _call(tm, hip, name, y, pg):
    """One call of entry point `name` on a device copy of y; returns (fields, n_bad)."""
    ns, L = y.shape[0] - 3, y.shape[-1]
    y = np.ascontiguousarray(y, np.float64)
    p = C.c_void_p()
    assert hip.hipMalloc(C.byref(p), y.nbytes) == 0
    try:
        assert hip.hipMemcpy(p, y.ctypes.data, y.nbytes, H2D) == 0
        par = pg._c()
        bad = C.c_int64(-1)
        rc = getattr(tm.lib(), name)(ns, L, C.cast(C.byref(par), C.c_void_p), p, C.byref(bad), None)
        assert rc == 0, (name, rc)
        out = np.empty_like(y)
        assert hip.hipMemcpy(out.ctypes.data, p, out.nbytes, D2H) == 0
    finally:
        hip.hipFree(p)
    return out, bad.value


def test_noalloc_equals_allocating_entry(tm, hip):
    cases = [(2, 256, 5, dict(a=1.2, dt=0.004)),     # the largest lattice the scratch holds
             (1, 40, 6, dict(a=-50.0, D_O=0.1, dt=0.05))]   # then a smaller one, with negative values
    for ns, L, seed, kw in cases:
        print(f"noalloc rk4 ns={ns} L={L}", flush=True)
        pg = tm.FKERKParams(A=A2, H=H2, M=(0.0, 0.02, 0.03), **kw)
        y = fields(ns, L, seed)
        want, wbad = _call(tm, hip, "tm_fke_rk4", y, pg)
        got, gbad = _call(tm, hip, "tm_fke_rk4_noalloc", y, pg)
        assert np.array_equal(got, want)
        assert gbad == wbad
    assert wbad > 0                           # the second case counts negative values


def test_noalloc_refuses_larger_lattices(tm, hip):
    pg = tm.FKERKParams(A=A2, H=H2)
    par = pg._c()
    bad = C.c_int64(-1)
    p = C.c_void_p(1)                         # never dereferenced: the size check comes first
    rc = tm.lib().tm_fke_rk4_noalloc(1, 257, C.cast(C.byref(par), C.c_void_p), p, C.byref(bad), None)
    assert rc == -6                           # TM_ERR_TOO_LARGE
