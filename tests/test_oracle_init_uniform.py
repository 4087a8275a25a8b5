"""Pins of the oracle's i.i.d. Uniform initial frames (tmo_init_uniform).

The frames are the starting state of every benchmark configuration: rho ~ U(0, 1)
for the DP model, P ~ U(0, 90) and W, O ~ U(0, 1) for the Rietkerk model, drawn
from Philox4x32-10 with key (seed, instance) and counter (cell, 0xFFFFFFFF,
0xFFFFFFFF, f / 4), field f taking word f % 4 of its block and mapping it to
lo + (hi - lo) (w + 1/2) 2^-32.

What pins it, independently of how the oracle writes it:
  * range: every value strictly inside (lo, hi) of its own field;
  * distribution: a Kolmogorov-Smirnov test against U(lo, hi) per field at
    2^18 cells, and the first two moments within 5 standard errors;
  * independence: near-zero correlation between fields of one cell, between
    instances and between seeds;
  * the counter layout: for sampled cells, the words come from the Philox block
    the layout names (Philox itself is pinned by Random123's known answers), and
    a fifth field comes from block 1, not a repeat of block 0.
The last test applies the same checks to deliberately broken frames (a dropped
(hi - lo) scale, one word shared by two fields, a wrong counter, a field taking
the word of another instance) and requires each one to be rejected.
"""
import math

import numpy as np
import pytest
from scipy import stats

L = 512                                   # 2^18 cells per field
LO = [0.0, 0.0, 0.0]
HI = [90.0, 1.0, 1.0]


def _u(w):
    return (np.asarray(w, np.float64) + 0.5) * 2.0 ** -32


@np.errstate(invalid="ignore", divide="ignore")
def frame_failures(f, lo, hi, other=None):
    """The checks, as a list of the ones that fail (empty: the frame passes).
    f: [nf, L, L]; other: a frame of another instance, same lo / hi."""
    bad = []
    nf = f.shape[0]
    x = f.reshape(nf, -1)
    n = x.shape[1]
    for k in range(nf):
        a, b = lo[k], hi[k]
        if not (np.all(x[k] > a) and np.all(x[k] < b)):
            bad.append(f"range[{k}]")
        u = (x[k] - a) / (b - a)
        if stats.kstest(u, "uniform").pvalue < 1e-4:
            bad.append(f"ks[{k}]")
        # mean (b + a) / 2 and variance (b - a)^2 / 12 of U(a, b)
        if abs(u.mean() - 0.5) > 5 * math.sqrt(1 / 12 / n):
            bad.append(f"mean[{k}]")
        if abs(u.var() - 1 / 12) > 5 * math.sqrt(1 / 180 / n):
            bad.append(f"var[{k}]")
    # independent fields: |r| <= 5 / sqrt(n) for each pair (a constant field has
    # no correlation coefficient: NaN, which fails the check)
    for i in range(nf):
        for j in range(i + 1, nf):
            r = np.corrcoef(x[i], x[j])[0, 1]
            if not abs(r) < 5 / math.sqrt(n):
                bad.append(f"corr[{i},{j}]")
    if other is not None:
        y = other.reshape(nf, -1)
        for k in range(nf):
            r = np.corrcoef(x[k], y[k])[0, 1]
            if not abs(r) < 5 / math.sqrt(n):
                bad.append(f"instance-corr[{k}]")
    return bad


def test_rietkerk_frame_range_distribution_independence(oracle):
    f = oracle.init_uniform(L, LO, HI, 20251202, 3)
    g = oracle.init_uniform(L, LO, HI, 20251202, 4)
    assert frame_failures(f, LO, HI, other=g) == []


def test_dp_frame_and_seeds(oracle):
    f = oracle.init_uniform(L, [0.0], [1.0], 5, 0)
    g = oracle.init_uniform(L, [0.0], [1.0], 6, 0)    # another seed, same instance
    assert frame_failures(f, [0.0], [1.0], other=g) == []


def test_shifted_interval(oracle):
    lo, hi = [-3.0, 2.5], [5.0, 2.75]
    f = oracle.init_uniform(256, lo, hi, 9, 1)
    assert frame_failures(f, lo, hi) == []


def test_counter_layout(oracle):
    """Field f of cell i is word f % 4 of Philox block (i, ~0, ~0, f / 4) under key
    (seed, instance): checked at sampled cells, including a fifth field."""
    n, seed, inst = 32, 11, 7
    lo = [0.0, 1.0, -2.0, 10.0, 0.0]
    hi = [1.0, 3.0, 2.0, 20.0, 1.0]
    f = oracle.init_uniform(n, lo, hi, seed, inst).reshape(5, -1)
    for i in (0, 1, 31, 32, 517, n * n - 1):
        b0 = oracle.philox([i, 0xFFFFFFFF, 0xFFFFFFFF, 0], [seed, inst])
        b1 = oracle.philox([i, 0xFFFFFFFF, 0xFFFFFFFF, 1], [seed, inst])
        words = b0 + b1[:1]
        for k in range(5):
            want = lo[k] + (hi[k] - lo[k]) * _u(words[k])
            assert f[k, i] == want, (i, k)
    # the fifth field is not a repeat of the first four words
    assert not np.any(np.isin(f[4], f[0]))


def _philox_words(oracle, cells, block, key, ctr_fn):
    return np.array([oracle.philox(ctr_fn(int(i), block), key) for i in cells], np.uint64)


def test_checks_reject_broken_frames(oracle):
    """Each plausible mistake in the frame generator fails at least one check."""
    n, seed, inst = 256, 13, 2
    good = oracle.init_uniform(n, LO, HI, seed, inst)
    other = oracle.init_uniform(n, LO, HI, seed, inst + 1)
    assert frame_failures(good, LO, HI, other=other) == []
    cells = np.arange(n * n)
    lay = lambda i, b: [i, 0xFFFFFFFF, 0xFFFFFFFF, b]              # noqa: E731
    w = _philox_words(oracle, cells, 0, [seed, inst], lay)          # [cells, 4]
    u = _u(w)
    sh = (3, n, n)

    def frame(cols, scale=True):
        return np.stack([(LO[k] + ((HI[k] - LO[k]) if scale else 1.0) * u[:, c]).reshape(n, n)
                         for k, c in enumerate(cols)])

    # the same words and map rebuild the good frame exactly
    assert np.array_equal(frame([0, 1, 2]), good)
    broken = {
        "dropped scale": frame([0, 1, 2], scale=False),
        "shared word W = O": frame([0, 1, 1]),
        "shared word P = W": frame([0, 0, 2]),
    }
    # a counter without the cell index: every cell draws the same block
    w_flat = np.array(oracle.philox([0, 0xFFFFFFFF, 0xFFFFFFFF, 0], [seed, inst]), np.uint64)
    broken["cell missing from the counter"] = np.stack(
        [np.full((n, n), LO[k] + (HI[k] - LO[k]) * _u(w_flat[k])) for k in range(3)])
    # the instance missing from the key: two instances draw the same frame
    g2 = oracle.init_uniform(n, LO, HI, seed, inst)
    for name, f in broken.items():
        assert f.shape == sh
        assert frame_failures(f, LO, HI, other=other) != [], name
    assert frame_failures(good, LO, HI, other=g2) != [], "instance missing from the key"
    # a counter that differs from the layout (step words 0 instead of ~0) keeps the
    # statistics but not the words: the layout check is what catches it
    w_bad = _philox_words(oracle, cells[:8], 0, [seed, inst], lambda i, b: [i, 0, 0, b])
    assert not np.array_equal(_u(w_bad[:, 0]), u[:8, 0])


@pytest.mark.parametrize("L_", [1, 3, 17])
def test_small_and_odd_sides(oracle, L_):
    f = oracle.init_uniform(L_, LO, HI, 1, 0)
    assert f.shape == (3, L_, L_)
    for k in range(3):
        assert np.all((f[k] > LO[k]) & (f[k] < HI[k]))
    # cell i of an L x L frame is cell i of any other frame size (the counter is the cell index)
    big = oracle.init_uniform(32, LO, HI, 1, 0).reshape(3, -1)
    assert np.array_equal(f.reshape(3, -1), big[:, : L_ * L_])
