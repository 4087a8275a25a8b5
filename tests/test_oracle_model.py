"""Pins of the oracle's model functions by what the mathematics fixes.

Each test checks one oracle function against a closed form, an invariant or a
library routine, chosen so that a plausible slip fails it: a wrong stencil index
or sign (Fourier eigenvalues, wrap-around cells), a wrong Dornic constant or
shape (closed-form constants, CIR transition moments, Poisson chi-square, Gamma
KS), a wrong RK4 stage (fixed points, the P = 0 linear solution, RK4's
amplification factor on a diffusion mode, order-4 convergence), a wrong DP
nonlinear scheme (the closed form rho / (1 + b rho t)), a wrong flush boundary,
and a biased reduction.  Reference file:line for each rule is in the oracle's
own comments (oracle/tm_oracle.c).
"""
import math

import numpy as np
import pytest
from scipy import linalg, stats

SQ2 = math.sqrt(2.0)


# ------------------------------------------------------------------ Laplacian
@pytest.mark.parametrize("L,kx,ky", [(8, 1, 0), (8, 2, 3), (6, 1, 1), (5, 2, 4), (16, 7, 5)])
def test_laplacian_fourier_eigenvalue(oracle, L, kx, ky):
    y, x = np.meshgrid(np.arange(L), np.arange(L), indexing="ij")
    for phase in (np.cos, np.sin):
        f = phase(2 * np.pi * (kx * x + ky * y) / L)
        lam = 2 * math.cos(2 * math.pi * kx / L) + 2 * math.cos(2 * math.pi * ky / L) - 4
        np.testing.assert_allclose(oracle.laplacian(f), lam * f, atol=1e-13)


@pytest.mark.parametrize("L", [4, 5, 7, 8])
def test_laplacian_brute_force_and_wrap(oracle, L):
    rng = np.random.default_rng(L)
    f = rng.standard_normal((L, L))
    got = oracle.laplacian(f)
    for p in range(L):
        for q in range(L):
            want = f[(p - 1) % L, q] + f[(p + 1) % L, q] + f[p, (q - 1) % L] + f[p, (q + 1) % L] - 4 * f[p, q]
            assert abs(got[p, q] - want) < 1e-13
    # a unit spike at each corner reaches exactly its four wrapped neighbours
    for p, q in [(0, 0), (0, L - 1), (L - 1, 0), (L - 1, L - 1)]:
        e = np.zeros((L, L))
        e[p, q] = 1.0
        g = oracle.laplacian(e)
        want = np.zeros((L, L))
        want[p, q] = -4
        for dp, dq in [(-1, 0), (1, 0), (0, -1), (0, 1)]:
            want[(p + dp) % L, (q + dq) % L] += 1
        assert np.array_equal(g, want)


# ------------------------------------------------------------------ Dornic constants
def test_dornic_constants_closed_form(oracle):
    # DP, a = 0.72, D = 1, dx = 1, sigma = sqrt 2, dt = 0.1: beta = -3.28
    beta, e, lam = oracle.dornic_constants(0.72, 1.0, SQ2, 0.1, 1.0)
    assert beta == pytest.approx(-3.28, rel=1e-15)
    assert e == pytest.approx(math.exp(-0.328), rel=1e-15)
    assert lam == pytest.approx(2 * -3.28 / (2.0 * (math.exp(-0.328) - 1)), rel=1e-14)
    assert e == pytest.approx(0.720363, abs=5e-7)
    assert lam == pytest.approx(11.72949, abs=5e-6)
    assert lam * e == pytest.approx(8.44949, abs=5e-6)
    # Rietkerk P: linear = -d = -0.25, D_P = 0.1, dx = 2, sigma = 0.5, dt = 0.005: beta = -0.35
    beta, e, lam = oracle.dornic_constants(-0.25, 0.1, 0.5, 0.005, 2.0)
    assert beta == pytest.approx(-0.35, rel=1e-15)
    assert e == pytest.approx(0.998252, abs=5e-7)
    assert lam == pytest.approx(1601.400, abs=5e-4)
    assert lam * e == pytest.approx(1598.600, abs=5e-4)


# ------------------------------------------------------------------ samplers
def _chi2_pvalue(samples, pmf, lo, hi):
    ks = np.arange(lo, hi + 1)
    n = len(samples)
    inr = (samples >= lo) & (samples <= hi)
    obs = np.bincount((samples[inr] - lo).astype(np.int64), minlength=hi - lo + 1).astype(float)
    exp = pmf(ks) * n
    keep = exp >= 20
    o, e = obs[keep], exp[keep]
    o = np.append(o, n - o.sum())
    e = np.append(e, n - e.sum())
    e = np.maximum(e, 1e-300)
    chi2 = float(np.sum((o - e) ** 2 / e))
    return stats.chi2.sf(chi2, len(o) - 1)


@pytest.mark.parametrize("z", [0.01, 0.3, 2.5, 11.9, 12.5, 60.0, 1598.6])
def test_poisson_exact_chi_square(oracle, z):
    x = oracle.sample("poisson", 200_000, z, mode=0, seed=11, cell=int(z * 10))
    assert np.all(x == np.round(x)) and x.min() >= 0
    lo, hi = int(stats.poisson.ppf(1e-7, z)), int(stats.poisson.ppf(1 - 1e-7, z)) + 1
    assert _chi2_pvalue(x, lambda k: stats.poisson.pmf(k, z), lo, hi) > 1e-4
    assert abs(x.mean() - z) < 5 * math.sqrt(z / len(x))


@pytest.mark.parametrize("z", [150.0, 1598.6])
def test_poisson_compat_is_truncated_gaussian(oracle, z):
    """Reference shortcut above 100: long(N(0,1) sqrt z + z), truncation toward zero."""
    x = oracle.sample("poisson", 200_000, z, mode=1, seed=12)
    s = math.sqrt(z)
    pmf = lambda k: stats.norm.cdf((k + 1 - z) / s) - stats.norm.cdf((k - z) / s)
    lo, hi = int(z - 6 * s), int(z + 6 * s)
    assert _chi2_pvalue(x, pmf, lo, hi) > 1e-4


@pytest.mark.parametrize("shape", [0.02, 0.5, 1.0, 3.7, 150.0, 1.6e5])
def test_gamma_ks(oracle, shape):
    x = oracle.sample("gamma", 100_000, shape, seed=13, cell=int(shape * 100) % 1000)
    assert x.min() >= 0
    assert stats.kstest(x, stats.gamma(shape).cdf).pvalue > 1e-4
    se = math.sqrt(shape / len(x))
    assert abs(x.mean() - shape) < 5 * se


def test_normal_draws_ks(oracle):
    x = oracle.sample("normal", 200_000, seed=14)
    assert stats.kstest(x, stats.norm.cdf).pvalue > 1e-4


# ------------------------------------------------------------------ Dornic transition
def cir_moments(alpha_i, rho, beta, dt, sigma):
    e = math.exp(beta * dt)
    mean = rho * e + alpha_i / beta * (e - 1)
    var = sigma ** 2 / beta * (rho * e * (e - 1) + alpha_i / (2 * beta) * (e - 1) ** 2)
    return mean, var


CASES = [
    # (linear, D, sigma, dt, dx, neighbour sum, rho): DP regime and Rietkerk regime
    (0.72, 1.0, SQ2, 0.1, 1.0, 2.0, 0.5),
    (0.72, 1.0, SQ2, 0.1, 1.0, 0.3, 0.02),
    (0.72, 1.0, SQ2, 0.1, 1.0, 0.0, 0.7),
    (-0.25, 0.1, 0.5, 0.005, 2.0, 120.0, 40.0),
    (-0.25, 0.1, 0.5, 0.005, 2.0, 3.0, 0.001),
]


@pytest.mark.parametrize("case", CASES)
def test_dornic_transition_moments(oracle, case):
    lin, D, sigma, dt, dx, nsum, rho = case
    beta, e, lam = oracle.dornic_constants(lin, D, sigma, dt, dx)
    alpha_i = D / dx ** 2 * nsum
    x, n = oracle.dornic_batch(alpha_i, rho, lam, e, sigma, 1_000_000, sampler=0, seed=21, draws=True)
    mean, var = cir_moments(alpha_i, rho, beta, dt, sigma)
    N = len(x)
    assert abs(x.mean() - mean) < 5 * math.sqrt(var / N), ((x.mean() - mean) / math.sqrt(var / N))
    m4 = np.mean((x - x.mean()) ** 4)
    assert abs(x.var() - var) < 5 * math.sqrt(max(m4 - var ** 2, 0) / N)
    z = lam * e * rho
    lo, hi = int(stats.poisson.ppf(1e-7, z)), int(stats.poisson.ppf(1 - 1e-7, z)) + 1
    assert _chi2_pvalue(n, lambda k: stats.poisson.pmf(k, z), lo, hi) > 1e-4


def test_dornic_special_cases(oracle):
    beta, e, lam = oracle.dornic_constants(0.72, 1.0, SQ2, 0.1, 1.0)
    # alpha_i = 0 and rho = 0: absorbing, exactly 0 and no draw
    r = oracle.lib().tmo_dornic_cell
    import ctypes as C
    n = C.c_int64(99)
    assert r(0.0, 0.0, lam, e, SQ2, 0, 1, 0, 3, 0, C.byref(n)) == 0.0
    # rho = 0, alpha_i > 0: z = 0 so n = 0, and rho* = Gamma(2 alpha_i / sigma^2) / lambda
    alpha_i = 0.4
    x, k = oracle.dornic_batch(alpha_i, 0.0, lam, e, SQ2, 200_000, seed=5, draws=True)
    assert np.all(k == 0)
    shape = 2 * alpha_i / 2.0
    assert stats.kstest(x * lam, stats.gamma(shape).cdf).pvalue > 1e-4
    # sigma -> 0: the transition collapses onto the deterministic linear flow
    sig = 1e-5
    beta, e, lam = oracle.dornic_constants(0.72, 1.0, sig, 0.1, 1.0)
    x = oracle.dornic_batch(0.8, 0.5, lam, e, sig, 2000, seed=6)
    mean, _ = cir_moments(0.8, 0.5, beta, 0.1, sig)
    assert np.max(np.abs(x - mean)) < 1e-4 * mean


# ------------------------------------------------------------------ Rietkerk RHS and RK4
def uniform_state(L, P, W, O):
    return np.stack([np.full((L, L), P), np.full((L, L), W), np.full((L, L), O)]).astype(np.float64)


@pytest.mark.parametrize("R", [1.2, 1.5])
def test_rhs_vegetated_fixed_point(oracle, R):
    p = oracle.rk_params(R=R)
    c, gmax, d, K1, K2, W0, alpha, rW = 10.0, 0.05, 0.25, 5.0, 5.0, 0.2, 0.2, 0.2
    Ws = d * K1 / (c * gmax - d)
    Ps = c / d * (R - rW * Ws)
    Os = R / alpha * (Ps + K2) / (Ps + K2 * W0)
    f = oracle.rk_rhs(p, uniform_state(6, Ps, Ws, Os))
    # the P linear loss -d P lives in the Dornic step, so f_P balances it
    np.testing.assert_allclose(f[0], d * Ps, rtol=1e-13)
    np.testing.assert_allclose(f[1], 0.0, atol=1e-12)
    np.testing.assert_allclose(f[2], 0.0, atol=1e-12)


def test_rhs_bare_fixed_point(oracle):
    R = 0.8
    p = oracle.rk_params(R=R)
    f = oracle.rk_rhs(p, uniform_state(5, 0.0, R / 0.2, R / (0.2 * 0.2)))
    assert np.max(np.abs(f)) < 1e-13


def test_rk4_p_zero_linear_closed_form(oracle):
    """P = 0: O' = R - alpha W0 O, W' = alpha W0 O - rW W (uniform fields)."""
    R, alpha, W0, rW, dt, n = 1.0, 0.2, 0.2, 0.2, 0.005, 2000
    p = oracle.rk_params(R=R, dt=dt)
    y0 = np.array([0.3, 0.7])                       # (W, O)
    y = oracle.rk4_deterministic(p, uniform_state(4, 0.0, y0[0], y0[1]), n)
    assert np.all(y[0] == 0.0)
    # affine system y' = A y + b, solved with the matrix exponential
    A = np.array([[-rW, alpha * W0], [0.0, -alpha * W0]])
    b = np.array([0.0, R])
    ys = np.linalg.solve(A, -b)
    want = ys + linalg.expm(A * n * dt) @ (y0 - ys)
    np.testing.assert_allclose([y[1, 0, 0], y[2, 0, 0]], want, rtol=1e-11)


def test_rk4_amplification_on_a_diffusion_mode(oracle):
    """With the reaction switched off, W and O diffuse; a Fourier mode of the
    periodic Laplacian is multiplied each step by RK4's stability polynomial
    g(h) = 1 + h + h^2/2 + h^3/6 + h^4/24, h = dt D lam_k / dx^2.  This fixes the
    stage Laplacians of the stage values and the dt/2, dt/6 weights."""
    L, kx, ky, n, dt, dx = 8, 1, 2, 40, 0.005, 2.0
    Dw, Do = 0.1, 100.0
    p = oracle.rk_params(R=0.0, gmax=0.0, alpha=0.0, rW=0.0, D=(0.1, Dw, Do), dt=dt, dx=dx)
    yy, xx = np.meshgrid(np.arange(L), np.arange(L), indexing="ij")
    mode = np.cos(2 * np.pi * (kx * xx + ky * yy) / L)
    y0 = np.stack([np.zeros((L, L)), 1.0 + 0.5 * mode, 2.0 + 0.25 * mode])
    y = oracle.rk4_deterministic(p, y0, n)
    lam = (2 * math.cos(2 * math.pi * kx / L) + 2 * math.cos(2 * math.pi * ky / L) - 4) / dx ** 2
    g = lambda h: 1 + h + h * h / 2 + h ** 3 / 6 + h ** 4 / 24
    np.testing.assert_allclose(y[1], 1.0 + 0.5 * g(dt * Dw * lam) ** n * mode, atol=1e-13)
    np.testing.assert_allclose(y[2], 2.0 + 0.25 * g(dt * Do * lam) ** n * mode, atol=1e-13)


def test_rk4_converges_at_fourth_order(oracle):
    T = 8.0
    y0 = uniform_state(4, 20.0, 2.0, 3.0)

    def run(dt):
        p = oracle.rk_params(R=1.3, dt=dt)
        return oracle.rk4_deterministic(p, y0, int(round(T / dt)))[:, 0, 0]

    ref = run(T / 4096)
    errs = [np.max(np.abs(run(T / n) - ref)) for n in (64, 128, 256)]
    for e1, e2 in zip(errs, errs[1:]):
        assert 13.0 < e1 / e2 < 19.0, errs


# ------------------------------------------------------------------ DP nonlinear sub-steps
def test_dp_exact_scheme_is_the_closed_form(oracle):
    for rho in (0.0, 1e-9, 0.37, 5.0):
        assert oracle.dp_nonlinear(2, 2.0, 0.1, rho) == rho / (1 + 2.0 * rho * 0.1)


@pytest.mark.parametrize("scheme,order", [(0, 2), (1, 5)])
def test_dp_scheme_local_error_order(oracle, scheme, order):
    """Local error against rho / (1 + b rho dt): O(dt^2) for Euler, O(dt^5) for RK4."""
    b, rho = 2.0, 0.8
    exact = lambda dt: rho / (1 + b * rho * dt)
    dts = (0.02, 0.01, 0.005)
    errs = [abs(oracle.dp_nonlinear(scheme, b, dt, rho) - exact(dt)) for dt in dts]
    for e1, e2 in zip(errs, errs[1:]):
        assert 2 ** order * 0.8 < e1 / e2 < 2 ** order * 1.25, errs
    if scheme == 0:   # Euler's leading local error is b^2 rho^3 dt^2
        assert errs[-1] == pytest.approx(b * b * rho ** 3 * dts[-1] ** 2, rel=0.05)


# ------------------------------------------------------------------ flush and reductions
def test_flush_boundaries(oracle):
    f = oracle.flush
    # DP: x < 1e-12 and x >= -1e-80 -> 0
    assert f(np.nextafter(1e-12, 0), 1e-12, -1e-80) == 0.0
    assert f(1e-12, 1e-12, -1e-80) == 1e-12
    assert f(-1e-80, 1e-12, -1e-80) == 0.0
    assert f(np.nextafter(-1e-80, -1), 1e-12, -1e-80) < 0
    z = f(-0.0, 1e-12, -1e-80)
    assert z == 0.0 and math.copysign(1.0, z) == 1.0
    assert math.isnan(f(float("nan"), 1e-12, -1e-80))
    # Rietkerk: x < 1e-10 and x >= 0 -> 0; negatives stay (and are faults)
    assert f(np.nextafter(1e-10, 0), 1e-10, 0.0) == 0.0
    assert f(1e-10, 1e-10, 0.0) == 1e-10
    assert f(-1e-300, 1e-10, 0.0) == -1e-300


def test_reduce_counts_and_means(oracle):
    rng = np.random.default_rng(3)
    L = 33
    f = rng.integers(-3, 50, size=(3, L, L)).astype(np.float64) / 64.0   # exact dyadic sums
    row = oracle.reduce(f)
    for k in range(3):
        assert row[2 * k] == np.count_nonzero(f[k] > 0)
        assert row[2 * k + 1] == math.fsum(f[k].ravel()) / (L * L)


# ------------------------------------------------------------------ step composition
def linear_flow(P, beta, dt, Dc):
    """The sigma -> 0 limit of the Dornic transition: the mean of the CIR step with
    alpha_i = Dc * (sum of the four old neighbours)."""
    nsum = np.roll(P, 1, 0) + np.roll(P, -1, 0) + np.roll(P, 1, 1) + np.roll(P, -1, 1)
    e = math.exp(beta * dt)
    return P * e + Dc * nsum / beta * (e - 1)


def test_rietkerk_step_is_dornic_then_rk4(oracle):
    """One full step with sigma -> 0 equals RK4 started from the linearly evolved P
    and the OLD W, O (rietkerk_bjork_FKE.cpp:4154-4235 order).  A long dt (with
    slow O diffusion so RK4 stays stable) makes the other order visibly different."""
    L, sig, dt = 8, 1e-6, 0.5
    p = oracle.rk_params(R=1.1, sigma=sig, dt=dt, D=(0.1, 0.1, 0.1))
    y0 = oracle.init_uniform(L, [5.0, 0.5, 0.5], [90.0, 1.0, 1.0], 4, 0)
    got = oracle.rk_run(p, y0, L, 9, 0, 0, 1)["fields"]
    beta = -0.25 - 4 * 0.1 / 4.0
    mid = y0.copy()
    mid[0] = linear_flow(y0[0], beta, dt, 0.1 / 4.0)
    want = oracle.rk4_deterministic(p, mid, 1)
    np.testing.assert_allclose(got, want, rtol=1e-5)
    # and not RK4 first, then Dornic
    other = oracle.rk4_deterministic(p, y0, 1)
    other[0] = linear_flow(other[0], beta, dt, 0.1 / 4.0)
    assert np.max(np.abs(got[0] - other[0]) / other[0]) > 1e-3


@pytest.mark.parametrize("scheme", [0, 1, 2])
def test_dp_step_is_dornic_then_nonlinear(oracle, scheme):
    L, sig, a, b, dt = 8, 1e-6, 0.72, 2.0, 0.1
    p = oracle.dp_params(a=a, b=b, sigma=sig, dt=dt, scheme=scheme)
    rho0 = oracle.init_uniform(L, [0.2], [1.0], 5, 0)[0]
    got = oracle.dp_run(p, rho0, L, 9, 0, 0, 1)["fields"][0]
    lin = linear_flow(rho0, a - 4.0, dt, 1.0)
    want = np.vectorize(lambda r: oracle.dp_nonlinear(scheme, b, dt, r))(lin)
    np.testing.assert_allclose(got, want, rtol=1e-5)


def test_dornic_compat_moments(oracle):
    """Compat sampler at the Rietkerk C2 scale: n = trunc(N(0,1) sqrt z + z) and
    rho* = (N(0,1) sqrt s + s) / lambda with s = mu + 1 + n.  For z >> 1,
    E[trunc(X)] = z - 1/2 and Var = z + 1/12 (Sheppard), so
    E[lambda rho*] = mu + 1/2 + z and Var[lambda rho*] = E[s] + z + 1/12."""
    beta, e, lam = oracle.dornic_constants(-0.25, 0.1, 0.5, 0.005, 2.0)
    alpha_i, rho = 0.1 / 4.0 * 160.0, 40.0
    x = oracle.dornic_batch(alpha_i, rho, lam, e, 0.5, 1_000_000, sampler=1, seed=31) * lam
    z = lam * e * rho
    mu = -1 + 2 * alpha_i / 0.25
    m = mu + 0.5 + z
    v = (mu + 0.5 + z) + z + 1.0 / 12
    N = len(x)
    assert abs(x.mean() - m) < 5 * math.sqrt(v / N), (x.mean() - m) / math.sqrt(v / N)
    assert abs(x.var() - v) < 5 * v * math.sqrt(2.0 / N)
