"""The Dornic transition on the GPU against the closed-form CIR transition.

A uniform lattice rho0 (L = 1024, 2^20 cells) makes every cell an independent
draw of the same transition: alpha_i = 4 D rho0 / dx^2, and with b = 0 the
explicit-Euler sub-step leaves rho* unchanged.  Dornic's Poisson-Gamma
composition is exact for d rho = (alpha + beta rho) dt + sigma sqrt(rho) dW, so
    E[rho*]   = rho e^{b dt} + (alpha / beta)(e^{b dt} - 1)
    Var[rho*] = (sigma^2 / beta)[rho e^{b dt}(e^{b dt} - 1) + (alpha / (2 beta))(e^{b dt} - 1)^2]
(beta = a - 4 D / dx^2).  The recorded Poisson integers must follow Poisson(z),
z = lambda e^{beta dt} rho0 (chi-square on the pmf).  Regimes: upper-tail and
lower-tail inversion, PTRS, and PTRS at z ~ 6e4 (the Rietkerk regime); shapes
below and above 1.  Both storage precisions.  Tolerance: 5 standard errors.
"""
import math

import numpy as np
import pytest
from scipy import stats

pytestmark = pytest.mark.gpu

L = 1024
A, D, SIGMA, DT, DX = 0.72, 1.0, math.sqrt(2.0), 0.1, 1.0


def closed_form(rho):
    beta = A - 4 * D / DX ** 2
    e = math.exp(beta * DT)
    alpha = D / DX ** 2 * 4 * rho
    mean = rho * e + alpha / beta * (e - 1)
    var = SIGMA ** 2 / beta * (rho * e * (e - 1) + alpha / (2 * beta) * (e - 1) ** 2)
    lam = 2 * beta / (SIGMA ** 2 * (e - 1))
    return mean, var, lam * e * rho


def one_step(tm, prec, rho0, seed=99):
    with tm.Context() as ctx:
        ctx.configure(tm.MODEL_DP, prec, tm.DPParams(a=A, b=0.0, D=D, sigma=SIGMA, dt=DT, dx=DX,
                                                     flush_eps=0.0, flush_neg=0.0),
                      [tm.Instance(L, 0, seed, 5)])
        ctx.upload(0, np.full((1, L, L), rho0))
        ctx.record_draws(True)
        ctx.run(1)
        return ctx.download(0)[0].ravel(), ctx.draws(0).ravel()


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("rho0", [0.05, 0.5, 10.0, 7000.0])
def test_dornic_transition_moments(tm, prec, rho0):
    p = tm.F64 if prec == "f64" else tm.F32
    x, n = one_step(tm, p, rho0)
    mean, var, z = closed_form(rho0)
    N = x.size
    se_mean = math.sqrt(var / N)
    assert abs(x.mean() - mean) < 5 * se_mean, (x.mean(), mean, (x.mean() - mean) / se_mean)
    # variance: SE from the sample fourth central moment
    m4 = np.mean((x - x.mean()) ** 4)
    se_var = math.sqrt(max(m4 - var ** 2, 0.0) / N)
    assert abs(x.var() - var) < 5 * se_var, (x.var(), var, (x.var() - var) / se_var)
    # Poisson integers: chi-square against the pmf over bins with expected count >= 20
    assert n.min() >= 0
    lo, hi = int(stats.poisson.ppf(1e-6, z)), int(stats.poisson.ppf(1 - 1e-6, z)) + 1
    ks = np.arange(lo, hi + 1)
    inr = (n >= lo) & (n <= hi)
    obs = np.bincount(n[inr] - lo, minlength=hi - lo + 1).astype(float)
    exp = stats.poisson.pmf(ks, z) * N
    keep = exp >= 20
    o, e = obs[keep], exp[keep]
    o = np.append(o, N - o.sum())
    e = np.append(e, N - e.sum())
    chi2 = float(np.sum((o - e) ** 2 / np.maximum(e, 1e-300)))
    dof = len(o) - 1
    assert stats.chi2.sf(chi2, dof) > 1e-4, (chi2, dof)
