"""The oracle's DP step against the reference's own integrator, in distribution.

percolationDornic_2D (Percolation_FD/stochasticSPDP.cpp:940-1143) is built from
the reference source into oracle/_ref by oracle/Makefile.  It seeds
std::mt19937_64 from std::random_device (stochasticSPDP.cpp:973-976) and uses
libstdc++'s Poisson and Gamma distributions, so only ensembles can agree.  Both
sides integrate the same initial frame with the Dornic step and the exact
nonlinear update rho / (1 + b rho dt) (stochasticSPDP.cpp:1064-1068; oracle
scheme 2, exact sampler: the legacy integrator has no Gaussian shortcuts, and
no flush of tiny values, so the oracle's flush is off here).

Compared per logged time (t = 0.1 ... 1.0 every step, then 2, 5, 10, 20): the
mean over surviving replicates of the spatial mean rho, and of the number of
occupied cells (the reference's columns 3 and 6).  Each difference must sit
within 4.5 standard errors, estimated from the oracle ensemble for both sides
(plus the Poisson variance of the empty-cell totals for the occupied count).
Over 28 comparisons the chance of a false failure is below 2e-4.
"""
import math
import os

import numpy as np
import pytest

A_VALUES = [0.72, 1.0]
L, R, DT, B, D, SIGMA = 32, 64, 0.1, 2.0, 1.0, math.sqrt(2.0)
T_MEAS = [2.0, 5.0, 10.0, 20.0]


def reference_rows(ref, a, rho0):
    out = np.zeros((64, 8), np.float64)
    tm = np.array(T_MEAS)
    fd = os.dup(1)
    devnull = os.open(os.devnull, os.O_WRONLY)
    os.dup2(devnull, 1)     # the reference prints a line per replicate
    try:
        n = ref.ref_ssp_percolation_dornic(L, a, B, D, SIGMA, DT, 1.0, R, rho0, tm, len(tm), T_MEAS[-1],
                                           out, out.shape[0])
    finally:
        os.dup2(fd, 1)
        os.close(fd)
        os.close(devnull)
    assert n > 0
    return out[:n]


@pytest.mark.parametrize("a", A_VALUES)
def test_dp_ensemble_matches_reference_integrator(oracle, a):
    ref = oracle.ref_lib()
    if ref is None:
        pytest.skip("reference library not built: /root/reference sources absent")
    rho0 = oracle.init_uniform(L, [0.0], [1.0], 77, 0).reshape(-1).copy()
    rows_ref = reference_rows(ref, a, rho0)
    nsteps = int(round(T_MEAS[-1] / DT))
    # the legacy integrator has no post-step flush (multiSPDP.cpp:3257-3258 added it later)
    p = oracle.dp_params(a=a, b=B, D=D, sigma=SIGMA, dt=DT, scheme=2, sampler=0, flush_eps=0.0, flush_neg=0.0)
    orc = np.stack([oracle.dp_run(p, rho0, L, 1234, k, 0, nsteps)["rows"] for k in range(R)])  # [R, steps, 2]
    checked = 0
    for row in rows_ref:
        t = row[2]
        n = int(round(t / DT))
        if n == 0:
            continue
        cnt, mean = orc[:, n - 1, 0], orc[:, n - 1, 1]
        alive = cnt > 0
        assert alive.sum() >= R // 2, "too few surviving oracle replicates to compare"
        m = alive.sum()
        x = mean[alive]
        se = x.std(ddof=1) / math.sqrt(m)
        zscore = abs(row[3] - x.mean()) / (math.sqrt(2.0) * se)
        assert zscore < 4.5, (a, t, "mean", row[3], x.mean(), zscore)
        # occupied cells: compare the empty-cell totals; while empty cells are
        # rare their count is Poisson-like and the oracle's sample variance can
        # be zero, so the Poisson variance of both totals is added
        empty_orc = float(np.sum(L * L - cnt[alive]))
        empty_ref = (L * L - row[6]) * m
        var = 2.0 * m * cnt[alive].var(ddof=1) + empty_orc + empty_ref + 1.0
        zscore = abs(empty_ref - empty_orc) / math.sqrt(var)
        assert zscore < 4.5, (a, t, "occupied", row[6], cnt[alive].mean(), zscore)
        checked += 2
    assert checked == 2 * (10 + len(T_MEAS))


# ---------------------------------------------------------------- survival
SURV_L, SURV_R = 8, 256
SURV_T = [2.0, 5.0, 8.0, 10.0, 15.0, 20.0, 30.0, 40.0]


@pytest.mark.parametrize("a", [0.5, 1.0])
def test_dp_survival_matches_reference_integrator(oracle, a):
    """Survival probability of a small lattice: the number of the reference's
    replicates still occupied at each logged time (its column 5, from
    var_mean_incremental_surv_runs) against the oracle's, 256 replicates each,
    as a two-proportion z-test (< 4.5 at each of 8 times).  At a = 0.5 the
    survival falls from 1 to ~0 by t = 40, at a = 1.0 from 1 to ~0.4.  The
    reference's second tally (column 7, surv_rep) sits in an array one element
    shorter than its times (stochasticSPDP.cpp:956), so its last entry is not
    a count; it is compared at the other times."""
    ref = oracle.ref_lib()
    if ref is None:
        pytest.skip("reference library not built: /root/reference sources absent")
    rho0 = oracle.init_uniform(SURV_L, [0.0], [1.0], 77, 0).reshape(-1).copy()
    out = np.zeros((64, 8), np.float64)
    tm = np.array(SURV_T)
    fd = os.dup(1)
    devnull = os.open(os.devnull, os.O_WRONLY)
    os.dup2(devnull, 1)
    try:
        n = ref.ref_ssp_percolation_dornic(SURV_L, a, B, D, SIGMA, DT, 1.0, SURV_R, rho0, tm, len(tm), SURV_T[-1],
                                           out, out.shape[0])
    finally:
        os.dup2(fd, 1)
        os.close(fd)
        os.close(devnull)
    rows = {round(r[2], 6): r for r in out[:n]}
    nsteps = int(round(SURV_T[-1] / DT))
    p = oracle.dp_params(a=a, b=B, D=D, sigma=SIGMA, dt=DT, scheme=2, sampler=0, flush_eps=0.0, flush_neg=0.0)
    cnt = np.stack([oracle.dp_run(p, rho0, SURV_L, 4321, k, 0, nsteps)["rows"][:, 0] for k in range(SURV_R)])
    # the oracle's survivors through the replicate statistics (break at the first empty time)
    series = np.zeros((SURV_R, len(SURV_T), 2))
    steps = [int(round(t / DT)) for t in SURV_T]
    series[:, :, 0] = cnt[:, [s - 1 for s in steps]]
    series[:, :, 1] = series[:, :, 0]
    surv_orc = oracle.replicate_stats(series, np.zeros(SURV_R, np.int32), 1, 3)[0, :, 2]
    for i, t in enumerate(SURV_T):
        row = rows[round(t, 6)]
        if i < len(SURV_T) - 1:
            assert row[5] == row[7]                # the reference's two tallies agree
        k_ref, k_orc = row[5], surv_orc[i]
        assert k_orc == np.count_nonzero(cnt[:, steps[i] - 1] > 0)
        pool = (k_ref + k_orc) / (2 * SURV_R)
        se = math.sqrt(max(pool * (1 - pool), 1.0 / SURV_R) * 2 / SURV_R)
        z = abs(k_ref - k_orc) / SURV_R / se
        assert z < 4.5, (a, t, k_ref, k_orc, z)


# ---------------------------------------------------------- steady state
SS_L, SS_R = 32, 32
SS_T = [30.0, 40.0, 50.0, 60.0, 70.0, 80.0]


def _ss_reference(oracle, ref, a, rho0):
    out = np.zeros((64, 8), np.float64)
    tm = np.array(SS_T)
    fd = os.dup(1)
    devnull = os.open(os.devnull, os.O_WRONLY)
    os.dup2(devnull, 1)
    try:
        n = ref.ref_ssp_percolation_dornic(SS_L, a, B, D, SIGMA, DT, 1.0, SS_R, rho0, tm, len(tm), SS_T[-1],
                                           out, out.shape[0])
    finally:
        os.dup2(fd, 1)
        os.close(fd)
        os.close(devnull)
    rows = {round(r[2], 6): r for r in out[:n]}
    return np.array([rows[round(t, 6)] for t in SS_T])


def _ss_oracle(oracle, a, rho0, sigma=SIGMA):
    p = oracle.dp_params(a=a, b=B, D=D, sigma=sigma, dt=DT, scheme=2, sampler=0, flush_eps=0.0, flush_neg=0.0)
    nsteps = int(round(SS_T[-1] / DT))
    idx = [int(round(t / DT)) - 1 for t in SS_T]
    rows = np.stack([oracle.dp_run(p, rho0, SS_L, 2468, k, 0, nsteps)["rows"][idx] for k in range(SS_R)])
    return rows                                                     # [R, times, (count, mean)]


def _ss_zscores(ref_rows, orc_rows):
    """Steady-state density and occupied fraction: the time average over t = 30 ... 80
    of the ensemble mean, reference against oracle.  The standard error comes from
    the spread of the oracle replicates' own time averages (the times within one
    replicate are correlated, the replicates are not), applied to both sides."""
    assert np.all(orc_rows[:, :, 0] > 0) and np.all(ref_rows[:, 5] == SS_R), "a replicate went extinct"
    rho_orc = orc_rows[:, :, 1].mean(1)
    occ_orc = orc_rows[:, :, 0].mean(1) / SS_L ** 2
    out = {}
    for name, x, r in (("density", rho_orc, ref_rows[:, 3].mean()),
                       ("occupied", occ_orc, ref_rows[:, 6].mean() / SS_L ** 2)):
        se = max(x.std(ddof=1), 1.0 / SS_L ** 2) / math.sqrt(SS_R)
        out[name] = abs(r - x.mean()) / (math.sqrt(2.0) * se)
    return out


@pytest.mark.parametrize("a", [1.2, 1.6])
def test_dp_steady_state_density_matches_reference_integrator(oracle, a):
    """Active phase (a = 1.2: rho ~ 0.20, a = 1.6: rho ~ 0.39 on 32^2): the steady-state
    density and occupied fraction of the oracle's ensemble (32 replicates, the
    exact nonlinear update as the reference) against percolationDornic_2D's
    (stochasticSPDP.cpp:940-1143), each within 4.5 standard errors of the
    difference.  At a = 1.2 a planted mistake in the oracle -- the noise amplitude
    sigma 10% too large, which lowers the density to ~0.11 -- must fail the same
    test."""
    ref = oracle.ref_lib()
    if ref is None:
        pytest.skip("reference library not built: /root/reference sources absent")
    rho0 = oracle.init_uniform(SS_L, [0.0], [1.0], 77, 0).reshape(-1).copy()
    ref_rows = _ss_reference(oracle, ref, a, rho0)
    z = _ss_zscores(ref_rows, _ss_oracle(oracle, a, rho0))
    assert z["density"] < 4.5 and z["occupied"] < 4.5, (a, z)
    if a == 1.2:
        z_bad = _ss_zscores(ref_rows, _ss_oracle(oracle, a, rho0, sigma=1.1 * SIGMA))
        assert z_bad["density"] > 10.0, z_bad
