"""CPU pins of the oracle's Rietkerk consumer model as one step (the loop body of
Rietkerk_FPE/rietkerk_bjork_FKE.cpp:4100-4462), built from its operators:
  * the vegetation's Dornic step (tmo_fke_veg_dornic) equals the single-species
    Rietkerk step (tmo_rk_run) with no plant growth, after that step's flush;
  * the separable movement (tmo_fke_movement_sep) equals the scatter
    restatement of advdiff_FKE_MultiSp (tmo_fke_movement) up to summation order,
    conserves the moved mass, and does not depend on its shift reach S;
  * the movement statistics (tmo_fke_mov_stats) equal numpy's mean and
    population variance;
  * with no grazers, one step of the two-species model is the single-species
    Rietkerk step bit for bit on P, W and O (the composition's order and
    constants: Dornic, RK4, flush), and the extinct grazer perceives gamma 1;
  * the composed step equals the same composition with the scatter movement.
No GPU needed."""
import math

import numpy as np
import pytest

from tests.test_movement import _case


class P:
    """tm_fke_params' fields (the oracle reads attributes only)."""

    def __init__(self, **kw):
        aij = 3.6 * 10.0 ** -6.08 * 20.0 ** -0.37
        ajm = 3.6 * 10.0 ** -6.08 * 100.0 ** -0.37
        d = dict(a=1.0 / 24.0, c=10000.0, gmax=0.05 * 10.0 ** -3.0 / 24.0, alpha=0.2 / 24.0, rW=0.2 / 24.0, W0=0.2,
                 K=(0.0, 5.0, 5000.0), A=((0.0, aij, 0.0), (aij, 0.0, ajm), (0.0, ajm, 0.0)),
                 H=((0.0, 1.0, 0.0), (1.0, 0.0, 1.0), (0.0, 1.0, 0.0)), E=(1.0, 0.45, 0.85),
                 M=(0.25 / 24.0, 0.061609 * 20.0 ** -0.25 / 8760.0, 0.061609 * 100.0 ** -0.25 / 8760.0),
                 D_P=0.00025 / 24.0, D_W=0.00025 / 24.0, D_O=0.025 / 24.0, D_mov=(0.0, 0.01028, 0.215965),
                 v=(0.0, 0.45427, 0.40587), sigma=(math.sqrt((0.00025 / 24.0) / (0.1 * 0.1)), 1.0, 10.0),
                 pR=(0.0, 1.04285 / 2.0, 1.25536 / 2.0), dt=0.1, dt_analytic=0.11817455,
                 dt_adv=(0.0, 0.11817455, 0.765270868), dx=0.1, flush_eps=1e-10, mv_invariance=0, CI=4,
                 sampler=0)
        d.update(kw)
        self.__dict__.update(d)


def model_fields(SP, L, seed):
    """P, G, (Pr,) W, O, gamma_1 (, gamma_2) with empty patches."""
    rng = np.random.default_rng(seed)
    f = np.zeros((2 * SP + 1, L, L))
    f[0] = rng.uniform(0, 60, (L, L)) * (rng.random((L, L)) < 0.7)
    f[1] = rng.uniform(0, 3, (L, L)) * (rng.random((L, L)) < 0.6)
    if SP == 3:
        f[2] = rng.uniform(0, 0.5, (L, L)) * (rng.random((L, L)) < 0.5)
    f[SP] = rng.uniform(0, 2, (L, L))
    f[SP + 1] = rng.uniform(0, 2, (L, L))
    f[SP + 2:] = rng.uniform(0, 0.3, (SP - 1, L, L))
    return f


@pytest.mark.parametrize("sampler,L", [(0, 33), (1, 24)])
def test_vegetation_dornic_is_the_single_species_step_without_growth(oracle, sampler, L):
    rng = np.random.default_rng(L)
    f = np.stack([rng.uniform(0, 90, (L, L)), rng.random((L, L)), rng.random((L, L))])
    f[0, 2:7, 3:9] = 0.0
    f[0, 12:15, 1:4] = 400.0        # the compat sampler's Gaussian shortcuts
    step = (1 << 32) + 5
    rp = oracle.rk_params(c=0.0, d=0.3, D=(0.2, 0.1, 50.0), sigma=0.6, dt=0.01, dx=1.5, sampler=sampler)
    ref = oracle.rk_run(rp, f, L, 11, 4, step, 1, rows=False)
    got = oracle.fke_veg_dornic(rp, f[0], 11, 4, step)
    assert np.array_equal(np.where((got < 1e-10) & (got >= 0), 0.0, got), ref["fields"][0])
    assert np.all(got[3:6, 4:8] == 0.0)          # empty cells with empty neighbours are skipped
    assert not np.array_equal(got, f[0])


@pytest.mark.parametrize("D,v,sigma,dt,dx,vscale", [(0.5, 0.9, 0.8, 0.5, 0.25, 1.0), (0.215965, 0.40587, 10.0,
                                                                                      0.11817455, 0.1, 3.0)])
def test_separable_movement_is_the_scatter_in_another_order(oracle, D, v, sigma, dt, dx, vscale):
    L = 48
    (rD, rV, pref), SD, SV = oracle.fke_move_constants(D, v, sigma, dt, dx)
    rho, gam, vx, vy = _case(7, L)
    vx, vy = vx * vscale, vy * vscale
    vmax = max(np.abs(vx).max(), np.abs(vy).max())
    S = oracle.fke_shift_reach(pref, vmax)
    assert 2 * oracle.fke_window(rD, rV, pref, vmax) + 1 <= L and S >= 2
    got = oracle.fke_movement_sep(rho, gam, vx, vy, rD, rV, S, pref, SD, SV)
    want = oracle.fke_movement(rho, gam, vx, vy, rD, rV, pref, SD, SV)
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-15)
    assert got.sum() == pytest.approx(rho[rho >= 1e-10].sum(), rel=1e-12)
    # any reach at least the largest shift gives the same sums, bit for bit
    wider = oracle.fke_movement_sep(rho, gam, vx, vy, rD, rV, S + 2, pref, SD, SV)
    assert np.array_equal(got, wider)
    # a reach short of the shifts loses mass (the check above is not vacuous)
    short = oracle.fke_movement_sep(rho, gam, vx, vy, rD, rV, S - 2, pref, SD, SV)
    assert short.sum() < got.sum() * (1 - 1e-6)


def test_mov_stats_are_mean_and_population_variance(oracle):
    rng = np.random.default_rng(3)
    g, vx, vy = rng.random((50, 50)), rng.normal(size=(50, 50)), rng.normal(size=(50, 50))
    m, var, mx, my = oracle.fke_mov_stats(g, vx, vy)
    assert m == pytest.approx(g.mean(), rel=1e-14)
    assert var == pytest.approx(g.var(), rel=1e-12)          # ddof = 0, as variance_of_vector
    assert not var == pytest.approx(g.var(ddof=1), rel=1e-6)
    assert (mx, my) == (pytest.approx(vx.mean(), abs=1e-15), pytest.approx(vy.mean(), abs=1e-15))


@pytest.mark.parametrize("sampler", [0, 1])
def test_no_grazers_is_the_single_species_step(oracle, sampler):
    L = 28
    p = P(dt_analytic=0.1, sampler=sampler)      # the single-species RK4 takes dt/2 and dt/6
    f = model_fields(2, L, 5)
    f[1] = 0.0
    got, fault, _ = oracle.fke_step(p, 2, f, 9, 2, 77)
    rp = oracle.rk_params(R=p.a, c=p.c, gmax=p.gmax, alpha=p.alpha, rW=p.rW, W0=p.W0, d=p.M[0],
                          D=(p.D_P, p.D_W, p.D_O), K=p.K, sigma=p.sigma[0], dt=p.dt, dx=p.dx,
                          flush_eps=p.flush_eps, sampler=sampler)
    ref = oracle.rk_run(rp, f[[0, 2, 3]], L, 9, 2, 77, 1, rows=False)
    assert np.array_equal(got[[0, 2, 3]], ref["fields"])
    assert fault == (ref["status"] != 0)
    assert np.all(got[1] == 0.0)
    assert np.all(got[4] == 1.0)                  # an extinct consumer perceives gamma = 1


@pytest.mark.parametrize("SP", [2, 3])
def test_composed_step_equals_scatter_composition(oracle, SP):
    L, seed, inst, step = 30, 4, 1, 12
    p = P()
    f = model_fields(SP, L, 20 + SP)
    got, _, ve = oracle.fke_step(p, SP, f, seed, inst, step)
    # the same composition, spelled out, with the scatter movement
    g, ve2 = oracle.fke_perception(f[:SP], f[SP + 2:], np.zeros((SP, L, L, 2)), p.pR[:SP], p.dx, p.v[:SP],
                                   seed, inst, step)
    veg = oracle.rk_params(c=p.c, d=p.M[0], D=(p.D_P, p.D_W, p.D_O), sigma=p.sigma[0], dt=p.dt, dx=p.dx)
    y = f[:SP + 2].copy()
    y[0] = oracle.fke_veg_dornic(veg, f[0], seed, inst, step)
    y, _ = oracle.fke_rk4(oracle.fke_rk_params(a=p.a, c=p.c, gmax=p.gmax, alpha=p.alpha, rW=p.rW, W0=p.W0, K=p.K,
                                               A=p.A, H=p.H, E=p.E, M=p.M, D_W=p.D_W, D_O=p.D_O, dt=p.dt,
                                               dt_rk=p.dt_analytic, dx=p.dx), y)
    for s in range(1, SP):
        (rD, rV, pref), SD, SV = oracle.fke_move_constants(p.D_mov[s], p.v[s], p.sigma[s], p.dt_analytic, p.dx)
        vx, vy = (ve2[s, ..., 0], ve2[s, ..., 1]) if SP == 3 else (np.zeros((L, L)), np.zeros((L, L)))
        y[s] = oracle.fke_movement(y[s], g[s - 1], vx, vy, rD, rV, pref, SD, SV)
    y = np.where((y < p.flush_eps) & (y >= 0), 0.0, y)
    np.testing.assert_allclose(got[:SP + 2], y, rtol=1e-12, atol=1e-300)
    assert np.array_equal(got[SP + 2:], g)
    assert np.array_equal(ve, ve2 if SP == 3 else np.zeros_like(ve2))
    assert np.any(got[1] != f[1]) and np.any(got[SP + 2:] != f[SP + 2:])
