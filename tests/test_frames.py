"""Initial frames on the CPU: the oracle's speckle and Gaussian frames, and the
product's host rule for the reference's mean-field frame values."""
import math
import os

import numpy as np
import pytest


@pytest.fixture(scope="module")
def host_tm():
    import trophic_maelstorm_amd as tm
    from trophic_maelstorm_amd import build as b
    if not os.path.exists(b.LIB):
        b.build()
    tm.lib()
    return tm


# ------------------------------------------------------------------ speckles
@pytest.mark.parametrize("perc", [0.015, 0.3, 0.0, 1.0])
def test_speckle_fraction_and_values(oracle, perc):
    L = 256
    hi, lo = [7.5, 0.25, 3.0], [0.5, 0.25, 1.0]
    f = oracle.init_speckle(L, perc, hi, lo, 11, 4).reshape(3, L * L)
    high = f[0] == hi[0]
    assert np.all(high | (f[0] == lo[0]))
    # every field switches with the same draw
    assert np.array_equal(f[2] == hi[2], high) and np.all(f[1] == 0.25)
    n, k = L * L, high.sum()
    sd = math.sqrt(n * perc * (1 - perc))
    assert abs(k - n * perc) <= 5 * sd + 1e-9
    if 0 < perc < 1:
        # independent of position: both halves of the lattice agree within 5 sd
        top, bot = high[: n // 2].sum(), high[n // 2:].sum()
        assert abs(top - bot) <= 5 * math.sqrt(2 * (n // 2) * perc * (1 - perc))


def test_speckle_stream_differs_from_uniform_frames(oracle):
    L = 128
    sp = oracle.init_speckle(L, 0.5, [1.0], [0.0], 3, 9)
    un = oracle.init_uniform(L, [0.0], [1.0], 3, 9)
    # the uniform frame's own u <= 0.5 pattern is not the speckle pattern
    agree = np.mean((un <= 0.5) == (sp == 1.0))
    assert abs(agree - 0.5) < 0.02


# ------------------------------------------------------------------ gaussians
def test_gauss_peaks_centres_and_widths(oracle):
    L = 64
    amp = [5.0, 2.0, 3.0]
    f = oracle.init_gauss(L, 3, [8.0, 16.0], amp).reshape(3, L, L)
    lo, hi = L // 8, 7 * L // 8
    centres = [(lo, lo), (hi, lo), (hi, hi)]          # counter-clockwise from the top left
    sds = [L / 8.0, L / 16.0, L / 16.0]               # the last width repeats
    for s in range(3):
        cx, cy = centres[s]
        assert np.unravel_index(np.argmax(f[s]), (L, L)) == (cx, cy)
        assert f[s, cx, cy] == amp[s]                  # exp(0) = 1 exactly
        d = int(sds[s])                                # one width along each axis
        for (i, j) in [(cx + d, cy), (cx - d, cy), (cx, cy + d), (cx, cy - d)]:
            if 0 <= i < L and 0 <= j < L:
                assert f[s, i, j] == pytest.approx(amp[s] * math.exp(-0.5), rel=1e-15)
        # not periodic: the bump is cut at the lattice edge, no wrap-around copy
        assert f[s].max() == f[s, cx, cy] and np.count_nonzero(f[s] > 0.5 * amp[s]) > 0


def test_gauss_matches_libm_exp(oracle):
    L = 48
    f = oracle.init_gauss(L, 2, [8.0], [4.0, 1.5, 0.75]).reshape(3, L, L)
    i, j = np.meshgrid(np.arange(L), np.arange(L), indexing="ij")
    for s, (cx, cy) in enumerate([(L // 8, L // 8), (7 * L // 8, L // 8)]):
        sd = L / 8.0
        want = [4.0, 1.5][s] * np.exp(-((i - cx) ** 2 + (j - cy) ** 2) / (2 * sd * sd))
        np.testing.assert_allclose(f[s], want, rtol=4e-16, atol=0)
    assert np.all(f[2] == 0.75)                        # beyond n_gauss: the constant


def test_gauss_reference_default_frame(oracle):
    """The reference's Rietkerk sweep frame: amp = clow[Sp..] = {5, 1, 1},
    sd = {g/8, g/16}, one Gaussian field (FKE:3956-3962, D1R:159)."""
    L = 32
    f = oracle.init_gauss(L, 1, [8.0, 16.0], [5.0, 1.0, 1.0]).reshape(3, L, L)
    assert f[0, 4, 4] == 5.0 and np.all(f[1] == 1.0) and np.all(f[2] == 1.0)
    assert f[0, 4, 8] == pytest.approx(5.0 * math.exp(-0.5), rel=1e-15)


# ------------------------------------------------------------ mean-field values
def test_mft_values_below_critical(host_tm):
    hi, lo = host_tm.mft_frame_values(False, 1.0, 1.5, 2.0, [0, 1, 1, 5, 1, 1], [0, 5, 3, 4, 0, 2])
    assert hi.tolist() == [2.0, 5.0, 3.0]     # max(dP, 0 * 0), c_spread * E
    assert lo.tolist() == [0.0, 5.0, 3.0]


def test_mft_values_above_critical(host_tm):
    hi, lo = host_tm.mft_frame_values(False, 2.0, 1.5, 2.0, [0, 1, 1, 5, 1, 1], [0, 5, 3, 4, 0, 2])
    assert hi.tolist() == [20.0, 1.0, 2.0]    # max(dP, 5 * 4); E = 0 -> bare c_spread
    assert lo.tolist() == [4.0, 1.0, 2.0]     # the background is the bare E_0


def test_mft_values_homogeneous_and_dp(host_tm):
    hi, lo = host_tm.mft_frame_values(True, 1.5, 1.5, 9.0, [0, 1, 1, 5, 1, 1], [0, 5, 3, 4, 0, 2])
    assert hi.tolist() == lo.tolist() == [20.0, 1.0, 2.0]
    hi, lo = host_tm.mft_frame_values(False, 0.1, 0.5, 0.25, [2.0, 3.0], [0.05, 0.4])
    assert hi.tolist() == [0.25] and lo.tolist() == [0.1]
    with pytest.raises(ValueError):
        host_tm.mft_frame_values(False, 0.1, 0.5, 0.25, [2.0, 3.0, 1.0], [0.05, 0.4, 1.0])
