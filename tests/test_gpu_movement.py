"""The Rietkerk consumers' movement step on the GPU (separable form) against the
CPU restatement of the reference's scatter, equal to summation-order rounding,
and bit for bit against the oracle's separable form (tmo_fke_movement_sep)."""
import numpy as np
import pytest

from tests.test_movement import _case

pytestmark = pytest.mark.gpu


def _species(oracle, D, v, sigma, dt, dx):
    return oracle.fke_move_constants(D, v, sigma, dt, dx)


def test_movement_matches_restatement_two_species(tm, oracle):
    L = 48
    consts = [_species(oracle, 0.5, 0.9, 0.8, 0.5, 0.25), _species(oracle, 0.2, 1.6, 1.3, 0.4, 0.2)]
    rho, gam, v = [], [], []
    for s in range(2):
        r, g, vx, vy = _case(10 + s, L)
        rho.append(r)
        gam.append(g)
        v.append(np.stack([vx * (1 + s), vy * (1 + s)], axis=-1))
    # exact half-site shifts exercise the rounding of the advected mean
    v[0][0, :, 0] = 0.5 / consts[0][0][2]
    v[0][1, :, 1] = -0.5 / consts[0][0][2]
    print("movement two species", flush=True)
    got = tm.fke_movement(np.array(rho), np.array(gam), np.array(v), [c[0] for c in consts],
                          [c[1] for c in consts], [c[2] for c in consts])
    for s in range(2):
        (rD, rV, pref), SD, SV = consts[s]
        vx, vy = v[s][..., 0], v[s][..., 1]
        want = oracle.fke_movement(rho[s], gam[s], vx, vy, rD, rV, pref, SD, SV)
        assert np.allclose(got[s], want, rtol=1e-12, atol=1e-15), np.max(np.abs(got[s] - want))
        assert got[s].sum() == pytest.approx(rho[s][rho[s] >= 1e-10].sum(), rel=1e-12)
        S = oracle.fke_shift_reach(pref, max(np.abs(vx).max(), np.abs(vy).max()))
        assert S >= 2 if s == 1 else S >= 1
        sep = oracle.fke_movement_sep(rho[s], gam[s], vx, vy, rD, rV, S, pref, SD, SV)
        assert np.array_equal(got[s], sep), np.max(np.abs(got[s] - sep))


def test_movement_refuses_reaches_beyond_the_lattice(tm, oracle):
    (rD, rV, pref), SD, SV = _species(oracle, 0.5, 0.9, 0.8, 0.5, 0.25)
    L = 2 * rD            # 2 rD + 1 > L
    z = np.zeros((1, L, L))
    with pytest.raises(tm.TMError):
        tm.fke_movement(z + 1.0, z, np.zeros((1, L, L, 2)), [(rD, rV, pref)], [SD], [SV])


def test_device_entry_point_equals_host_entry_point(tm, oracle):
    """tm_fke_movement on device arrays (the drop-in for launchAdvDiffKernel_MultiSp)
    gives the host entry point's result bit for bit.  Device memory comes from the
    HIP runtime the library itself links."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    H2D, D2H = 1, 2
    L = 40
    consts = [_species(oracle, 0.5, 0.9, 0.8, 0.5, 0.25)]
    r, g, vx, vy = _case(21, L)
    rho, gam = r[None].copy(), g[None].copy()
    v = np.ascontiguousarray(np.stack([vx, vy], axis=-1)[None])
    want = tm.fke_movement(rho, gam, v, [c[0] for c in consts], [c[1] for c in consts], [c[2] for c in consts])
    host = dict(rho=rho, gam=gam, v=v, sd=np.ascontiguousarray(consts[0][1]), sv=np.ascontiguousarray(consts[0][2]),
                out=np.zeros_like(rho))
    dev = {}
    try:
        for k, x in host.items():
            p = C.c_void_p()
            assert hip.hipMalloc(C.byref(p), x.nbytes) == 0
            dev[k] = p
            assert hip.hipMemcpy(p, x.ctypes.data, x.nbytes, H2D) == 0
        sp = (tm._MoveSpecies * 1)(tm._MoveSpecies(*consts[0][0]))
        print("movement device entry", flush=True)

        def run():
            return tm.lib().tm_fke_movement(1, L, C.cast(sp, C.c_void_p), dev["sd"], dev["sv"], dev["rho"],
                                            dev["gam"], dev["v"], dev["out"], None)
        assert run() == 0
        got = np.empty_like(rho)
        assert hip.hipMemcpy(got.ctypes.data, dev["out"], got.nbytes, D2H) == 0
        assert np.array_equal(got, want)
        v[0, 3, 4, 0] = float("nan")                 # a NaN velocity is refused
        assert hip.hipMemcpy(dev["v"], v.ctypes.data, v.nbytes, H2D) == 0
        assert run() < 0
    finally:
        for p in dev.values():
            hip.hipFree(p)


@pytest.mark.parametrize("ns", [2, 3])
def test_fke_perception_bit_exact(tm, oracle, ns):
    L, dx, pR, v = 37, 0.1, (0.0, 0.52, 0.31), (0.1, 0.45, 0.4)
    rng = np.random.default_rng(40 + ns)
    rho = np.zeros((ns, L, L))
    rho[0] = rng.uniform(0, 2, (L, L)) * (rng.random((L, L)) < 0.5)
    rho[1:] = rng.uniform(0, 0.4, (ns - 1, L, L))
    gam = rng.uniform(0, 0.1, (ns - 1, L, L))
    veff = rng.uniform(-1, 1, (ns, L, L, 2))
    print(f"fke perception ns={ns}", flush=True)
    g1, v1 = tm.fke_perception(rho, gam, veff, pR, dx, v, seed=9, inst=3, step=17)
    g2, v2 = oracle.fke_perception(rho, gam, veff, pR, dx, v, seed=9, inst=3, step=17)
    assert np.array_equal(g1, g2) and np.array_equal(v1, v2)
