"""The FP32 Gamma sampler's squeeze (gamma_f_try in csrc/tm_fast32.h).

A Marsaglia-Tsang trial with normal x, c = 1/sqrt(9d), y = c x, v = (1 + y)^3 > 0
accepts u when log u < Q(x) = x^2/2 + d(1 - v + log v).  The sampler first tries
the squeeze u < 1 - x^4 / (108 d m), m = min(1 + y, 1), and only where that fails
evaluates the logarithms.  The squeeze is exact only if it never accepts a u that
the logarithmic test rejects, that is 1 - x^4/(108 d m) <= e^Q for every x with
v > 0 and every d >= 2/3 (shape >= 1, or shape + 1 below that).  Checked here in
FP64 on a dense grid, with Q from log1p (or its series where |y| is small), and
shown to catch two plausible mistakes: the bound without m, and twice the
constant.  No GPU needed."""
import numpy as np


def q_of(x, d):
    """Q(x) = 3 d R(c x), R(y) = log(1 + y) - y + y^2/2 - y^3/3, free of cancellation."""
    c = 1.0 / np.sqrt(9.0 * d)
    y = c * x
    r = np.log1p(y) - y + y * y / 2 - y ** 3 / 3
    s = np.abs(y) < 0.05
    ys = y[s]
    r[s] = -ys ** 4 / 4 + ys ** 5 / 5 - ys ** 6 / 6 + ys ** 7 / 7 - ys ** 8 / 8 + ys ** 9 / 9
    return 3.0 * d * r, y


def violations(bound):
    """Grid points where the squeeze bound exceeds e^Q (would accept a rejected u)."""
    bad = 0
    for d in np.geomspace(2.0 / 3.0, 1e6, 120):
        c = 1.0 / np.sqrt(9.0 * d)
        x = np.linspace(-0.999999 / c, 12.0, 40001)
        q, y = q_of(x, d)
        m = np.minimum(1.0 + y, 1.0)
        bad += int(np.count_nonzero(bound(x, d, m) > np.exp(q) * (1 + 1e-12)))
    return bad


def test_squeeze_never_accepts_what_the_log_test_rejects():
    assert violations(lambda x, d, m: 1 - x ** 4 / (108 * d * m)) == 0


def test_check_catches_a_dropped_m_or_a_tighter_constant():
    assert violations(lambda x, d, m: 1 - x ** 4 / (108 * d)) > 0
    assert violations(lambda x, d, m: 1 - x ** 4 / (216 * d * m)) > 0


def test_squeeze_decides_more_trials_than_marsaglia_tsang():
    """Fraction of trials the squeeze settles, over x ~ N(0, 1): at least the
    classic 1 - 0.0331 x^4 everywhere, and nearly all of them at large shape."""
    for d, want in ((2.0 / 3.0, 0.93), (4.0, 0.99), (30.0, 0.998), (7.2e4, 0.99999)):
        c = 1.0 / np.sqrt(9.0 * d)
        x = np.linspace(max(-8.0, -0.999999 / c), 8.0, 200001)
        w = np.exp(-x * x / 2)
        w /= w.sum()
        m = np.minimum(1.0 + c * x, 1.0)
        ours = (np.clip(1 - x ** 4 / (108 * d * m), 0, 1) * w).sum()
        classic = (np.clip(1 - 0.0331 * x ** 4, 0, 1) * w).sum()
        assert ours >= classic - 1e-12
        assert ours > want, (d, ours)
