"""CPU pins of the oracle's Rietkerk consumer RK4 (tmo_fke_rk4: RK4_Integrate_Stochastic_2Sp /
_MultiSp with f_DorFKE_2Sp / f_DorFKE_3Sp, Rietkerk_FPE/rietkerk_bjork_FKE.cpp:2449-2533,
:2611-2720).  Each pin ties it to something the restatement does not share:
  * no grazers: the single-species Rietkerk RK4 (tmo_rk4_deterministic, pinned by the
    mean-field fixed points and the P = 0 closed form), bit for bit;
  * no predators: the two-species form, bit for bit (the field order moves W and O);
  * no grazing: each consumer decays by RK4's own polynomial of e^{-M dt};
  * unit efficiency, no mortality, no plant growth: the biomass P + G (+ Pr) is
    conserved, which a wrong sign or factor in the coupling terms breaks;
  * dt_rk != dt: the half and sixth steps use dt_rk and the third stage dt."""
import numpy as np
import pytest

from oracle import oracle as o

A2 = [[0.0, 0.3, 0.0], [0.0, 0.0, 0.2], [0.0, 0.0, 0.0]]
H2 = [[0.0, 0.5, 0.0], [0.0, 0.0, 0.7], [0.0, 0.0, 0.0]]


def fields(ns, L, seed=0):
    rng = np.random.default_rng(seed)
    y = rng.random((ns + 3, L, L))
    y[0] *= 20.0
    return y


def test_no_grazers_is_the_single_species_rk4():
    L = 12
    y = fields(1, L)
    y[1] = 0.0
    p = o.fke_rk_params(a=1.1, A=A2, H=H2, E=(1.0, 0.45, 0.3), M=(0.0, 0.02, 0.01), D_W=0.1, D_O=100.0,
                        dt=0.005, dx=2.0)
    out, bad = o.fke_rk4(p, y)
    assert bad == 0
    q = o.rk_params(R=1.1, D=(0.1, 0.1, 100.0), dt=0.005, dx=2.0)
    ref = o.rk4_deterministic(q, y[[0, 2, 3]], 1)
    assert np.array_equal(out[[0, 2, 3]], ref)
    assert np.all(out[1] == 0.0)


def test_no_predators_is_the_two_species_form():
    L = 10
    y2 = fields(1, L, 3)
    y3 = np.stack([y2[0], y2[1], np.zeros((L, L)), y2[2], y2[3]])
    p = o.fke_rk_params(a=0.9, A=A2, H=H2, M=(0.0, 0.02, 0.03), dt=0.01)
    out2, _ = o.fke_rk4(p, y2)
    out3, _ = o.fke_rk4(p, y3)
    assert np.array_equal(out3[[0, 1, 3, 4]], out2)
    assert np.all(out3[2] == 0.0)


@pytest.mark.parametrize("ns", [1, 2])
def test_without_grazing_consumers_decay_by_the_rk4_polynomial(ns):
    L, M, dt = 8, (0.0, 0.7, 0.4), 0.05
    y = fields(ns, L, 5)
    p = o.fke_rk_params(A=np.zeros((3, 3)), M=M, dt=dt)
    out, _ = o.fke_rk4(p, y)
    for s in range(1, ns + 1):
        x = M[s] * dt
        np.testing.assert_allclose(out[s], y[s] * (1 - x + x * x / 2 - x ** 3 / 6 + x ** 4 / 24), rtol=1e-14)


@pytest.mark.parametrize("ns", [1, 2])
def test_biomass_is_conserved_without_growth_or_mortality(ns):
    L = 9
    y = fields(ns, L, 7)
    p = o.fke_rk_params(c=0.0, A=A2, H=H2, E=(1.0, 1.0, 1.0), M=(0.0, 0.0, 0.0), dt=0.05)
    out = y
    for _ in range(20):
        out, _ = o.fke_rk4(p, out)
    bio0, bio = y[: ns + 1].sum(0), out[: ns + 1].sum(0)
    np.testing.assert_allclose(bio, bio0, rtol=1e-13)
    assert np.all(out[0] < y[0])          # grazing moved biomass from P to G
    assert np.all(out[1] > y[1] * (1 - 1e-12)) or ns == 2


def test_dt_rk_sets_the_half_and_sixth_steps():
    L, M, dt, dtr = 6, (0.0, 0.9, 0.0), 0.04, 0.1
    y = fields(1, L, 9)
    p = o.fke_rk_params(A=np.zeros((3, 3)), M=M, dt=dt, dt_rk=dtr)
    out, _ = o.fke_rk4(p, y)
    m, h2, h6 = M[1], dtr / 2, dtr / 6
    k1 = -m
    k2 = -m * (1 + h2 * k1)
    k3 = -m * (1 + h2 * k2)
    k4 = -m * (1 + dt * k3)
    np.testing.assert_allclose(out[1], y[1] * (1 + h6 * (k1 + 2 * k2 + 2 * k3 + k4)), rtol=1e-14)
    same, _ = o.fke_rk4(o.fke_rk_params(A=np.zeros((3, 3)), M=M, dt=dt), y)
    assert not np.allclose(out[1], same[1], rtol=1e-6)


def test_negative_values_are_counted():
    y = fields(1, 6, 11)
    # a negative rainfall empties O within the step: O + dt a < 0 for O < 1
    p = o.fke_rk_params(a=-50.0, A=np.zeros((3, 3)), D_O=0.1, dt=0.05)
    out, bad = o.fke_rk4(p, y)
    assert np.all(out[3] < 0)
    assert bad == np.count_nonzero((out < 0) | ~np.isfinite(out)) >= 36
