"""The Rietkerk consumers' movement step on the GPU (gather form) against the
CPU restatement of the reference's scatter: equal to summation-order rounding."""
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
        want = oracle.fke_movement(rho[s], gam[s], v[s][..., 0], v[s][..., 1], rD, rV, pref, SD, SV)
        assert np.allclose(got[s], want, rtol=1e-12, atol=1e-15), np.max(np.abs(got[s] - want))
        assert got[s].sum() == pytest.approx(rho[s][rho[s] >= 1e-10].sum(), rel=1e-12)


def test_movement_refuses_reaches_beyond_the_lattice(tm, oracle):
    (rD, rV, pref), SD, SV = _species(oracle, 0.5, 0.9, 0.8, 0.5, 0.25)
    L = 2 * rD            # 2 rD + 1 > L
    z = np.zeros((1, L, L))
    with pytest.raises(tm.TMError):
        tm.fke_movement(z + 1.0, z, np.zeros((1, L, L, 2)), [(rD, rV, pref)], [SD], [SV])
