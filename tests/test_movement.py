"""Movement of the Rietkerk model's consumer species (advdiff_FKE_MultiSp,
rietkerk_bjork_FKE.cpp:3614-3695): the CPU restatement pinned by closed forms,
conservation, convolution identities and shifts; the product's host constants
against the restatement's.  No GPU needed."""
import math

import numpy as np
import pytest


def _conv_sep(x, st):
    """Circular separable convolution with the symmetric half-stencil st (numpy)."""
    R = len(st) - 1
    out = np.zeros_like(x)
    for dr in range(-R, R + 1):
        row = np.roll(x, dr, axis=0) * st[abs(dr)]
        for dc in range(-R, R + 1):
            out += np.roll(row, dc, axis=1) * st[abs(dc)]
    return out


@pytest.mark.parametrize("D,v,sigma,dt,dx", [(0.0298, 0.45427, 1.0, 0.1, 0.1), (0.05221, 0.40587, 3.0, 0.5, 0.1),
                                             (1.0, 2.0, 0.7, 1.0, 0.25)])
def test_constants_closed_forms(oracle, D, v, sigma, dt, dx):
    (rD, rV, pref), SD, SV = oracle.fke_move_constants(D, v, sigma, dt, dx)
    beta = 3.0 / dt
    assert rD == int(4 * math.sqrt(2 * D * dt) / dx)
    svx = math.sqrt(2 * beta * v) / beta * math.sqrt(dt + (1 - math.exp(-2 * beta * dt)) / (2 * beta)
                                                     - 2 / beta * (1 - math.exp(-beta * dt)))
    assert rV == int(4 * svx / dx)
    assert pref == pytest.approx(dt / dx * (1 - (1 - math.exp(-3.0)) / 3.0), rel=1e-15)
    for st, R in ((SD, rD), (SV, rV)):
        assert len(st) == R + 1
        assert st[0] + 2 * st[1:].sum() == pytest.approx(1.0, abs=1e-15)
        x = np.arange(R + 1)
        assert np.allclose(st / st[0], np.exp(-0.5 * x * x / sigma ** 2), rtol=1e-14)


def test_product_constants_equal_oracle(oracle):
    import os
    import trophic_maelstorm_amd as tm
    from trophic_maelstorm_amd import build as b
    if not os.path.exists(b.LIB):
        b.build()
    for args in [(0.0298, 0.45427, 1.0, 0.1, 0.1), (0.3, 1.7, 0.4, 0.2, 0.05)]:
        for mv in (0, 1):
            a = oracle.fke_move_constants(*args, dt_adv=0.3, mv_invariance=mv)
            b_ = tm.fke_move_constants(*args, dt_adv=0.3, mv_invariance=mv)
            assert a[0] == b_[0] and np.array_equal(a[1], b_[1]) and np.array_equal(a[2], b_[2])


def _case(seed, L=24):
    rng = np.random.default_rng(seed)
    rho = rng.uniform(0, 2, (L, L)) * (rng.random((L, L)) < 0.6)
    gam = rng.random((L, L))
    th = 2 * np.pi * rng.random((L, L))
    return rho, gam, 0.9 * np.cos(th), 0.9 * np.sin(th)


def test_mass_is_conserved(oracle):
    rho, gam, vx, vy = _case(1)
    rho[0, 0] = 5e-11                            # below the 1e-10 floor: not moved, lost
    (rD, rV, pref), SD, SV = oracle.fke_move_constants(0.5, 0.9, 0.8, 0.5, 0.25)
    out = oracle.fke_movement(rho, gam, vx, vy, rD, rV, pref, SD, SV)
    kept = rho[rho >= 1e-10].sum()
    assert out.sum() == pytest.approx(kept, rel=1e-13) and rD >= 2 and rV >= 1


def test_full_perception_is_diffusion(oracle):
    """gamma = 1: the separable convolution with the diffusion stencil, whatever the velocities."""
    rho, _, vx, vy = _case(2)
    (rD, rV, pref), SD, SV = oracle.fke_move_constants(0.5, 0.9, 0.8, 0.5, 0.25)
    out = oracle.fke_movement(rho, np.ones_like(rho), vx, vy, rD, rV, pref, SD, SV)
    assert np.allclose(out, _conv_sep(np.where(rho >= 1e-10, rho, 0.0), SD), rtol=1e-13, atol=1e-15)


@pytest.mark.parametrize("shift", [(0, 0), (2, 0), (0, -3), (1, 2)])
def test_no_perception_is_shifted_advection(oracle, shift):
    """gamma = 0 and one velocity for every cell, chosen so that v mu_prefactor is a
    whole number of sites: the advection stencil's convolution rolled by that many
    rows (vx) and columns (vy)."""
    rho, _, _, _ = _case(3)
    (rD, rV, pref), SD, SV = oracle.fke_move_constants(0.5, 0.9, 0.8, 0.5, 0.25)
    vx = np.full_like(rho, shift[0] / pref)
    vy = np.full_like(rho, shift[1] / pref)
    out = oracle.fke_movement(rho, np.zeros_like(rho), vx, vy, rD, rV, pref, SD, SV)
    want = np.roll(np.roll(_conv_sep(rho, SV), shift[0], axis=0), shift[1], axis=1)
    assert np.allclose(out, want, rtol=1e-13, atol=1e-15)


def test_mixed_perception_is_the_product_kernel(oracle):
    """One source cell: its spread is the outer product of gamma SD + (1 - gamma) SV
    (shifted) in rows and in columns."""
    L = 32
    (rD, rV, pref), SD, SV = oracle.fke_move_constants(0.5, 0.9, 0.8, 0.5, 0.25)
    rho = np.zeros((L, L))
    rho[10, 20] = 1.5
    g = 0.3
    vx, vy = np.full((L, L), 2.0 / pref), np.full((L, L), -1.0 / pref)
    out = oracle.fke_movement(rho, np.full((L, L), g), vx, vy, rD, rV, pref, SD, SV)

    def kern(center, mu):
        k = np.zeros(L)
        for d in range(-rD, rD + 1):
            k[(center + d) % L] += g * SD[abs(d)]
        for d in range(-rV, rV + 1):
            k[(mu + d) % L] += (1 - g) * SV[abs(d)]
        return k
    want = 1.5 * np.outer(kern(10, 12), kern(20, 19))
    assert np.allclose(out, want, rtol=1e-13, atol=1e-16)


# ------------------------------------------------------------------ perception and velocities (FKE)
def _fke_gamma_brute(pR, dx, ns, rho, gam_old):
    """calc_gamma_vel_2Sp/3Sp_NonRefugia's gamma, site by site from the reference's
    statements, with numpy means."""
    from oracle import oracle as o
    L = rho.shape[-1]
    eps = 1e-12
    rmax = max(pR[:ns]) / dx
    disk = o.perception_offsets(int(rmax))
    mean = [rho[s].mean() for s in range(ns)]
    nfac = max(1 - np.count_nonzero(rho[0] > 0) / (L * L), 0.35)
    new_len = int(nfac * nfac * len(disk))
    if ns == 3:
        new_len = max(new_len, 1)
    inv = [1 / m if m >= eps else 1 / eps for m in mean]
    out = gam_old.copy()
    for s in range(1, ns):
        fr = pR[s] / (dx * rmax)
        eff = int(fr * fr * new_len)
        if ns == 3 and eff < 1 and fr > 0:
            eff = 1
        for i in range(L):
            for j in range(L):
                if mean[s] < eps:
                    out[s - 1, i, j] = 1.0
                    continue
                if mean[s - 1] < eps:
                    out[s - 1, i, j] = 0.0
                    continue
                nb = [((i + a) % L, (j + b) % L) for a, b in disk[:eff]]
                g = gam_old[s - 1, i, j]
                if ns == 3 and s == 1 and mean[2] >= eps:
                    g += sum(rho[0][k] * (1 - rho[2][k] * inv[2]) for k in nb)
                else:
                    g += sum(rho[s - 1][k] for k in nb)
                g = g * inv[s - 1] / eff if (ns == 2 or s == 1) else g / eff * inv[s - 1]
                out[s - 1, i, j] = min(max(g, 0.0), 1.0)
    return out


@pytest.mark.parametrize("ns", [2, 3])
def test_fke_perception_against_brute_force(oracle, ns):
    L, dx, pR, v = 14, 0.1, (0.0, 0.35, 0.25), (0.1, 0.45, 0.4)
    rng = np.random.default_rng(30 + ns)
    rho = np.zeros((ns, L, L))
    rho[0] = rng.uniform(0, 2, (L, L)) * (rng.random((L, L)) < 0.4)
    rho[1:] = rng.uniform(0, 0.4, (ns - 1, L, L))
    gam = rng.uniform(0, 0.1, (ns - 1, L, L))
    veff = np.full((ns, L, L, 2), 7.0)
    g, ve = oracle.fke_perception(rho, gam, veff, pR, dx, v, seed=5, inst=2, step=11)
    assert np.allclose(g, _fke_gamma_brute(pR, dx, ns, rho, gam), rtol=1e-12, atol=1e-15)
    if ns == 2:
        assert np.all(ve == 7.0)                     # the two-species form draws no velocity
    else:
        for s in range(3):
            for (i, j) in [(0, 0), (3, 9), (13, 13)]:
                w = oracle.philox([i * L + j, 11, 0, (s << 24) | 0x800000], [5, 2])[0]
                ang = 2 * math.pi * ((int(w) + 0.5) * 2.0 ** -32)
                assert ve[s, i, j, 0] == pytest.approx(v[s] * math.cos(ang), abs=1e-15)
                assert ve[s, i, j, 1] == pytest.approx(v[s] * math.sin(ang), abs=1e-15)


def test_fke_extinct_consumer_has_full_perception(oracle):
    L = 12
    rho = np.zeros((2, L, L))
    rho[0] = 1.0
    g, _ = oracle.fke_perception(rho, np.full((1, L, L), 0.3), np.zeros((2, L, L, 2)), (0.0, 0.35), 0.1, (0.0, 0.45))
    assert np.all(g == 1.0)
    rho[0], rho[1] = 0.0, 0.5                        # no resource: gamma 0
    g, _ = oracle.fke_perception(rho, np.full((1, L, L), 0.3), np.zeros((2, L, L, 2)), (0.0, 0.35), 0.1, (0.0, 0.45))
    assert np.all(g == 0.0)
