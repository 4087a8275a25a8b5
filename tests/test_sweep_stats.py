"""Sweep measurement schedule and across-replicate statistics, on the CPU.

* The product's measurement times (tm_log10_time_bins, host code in libtm) equal
  the reference's logarithm10_time_bins, built from source, and the vectors it
  wrote to tests/golden/sweep_vectors.json.
* The measurement steps follow the reference's time-loop test.
* The oracle's replicate statistics equal the reference's own
  var_mean_incremental_surv_runs, folded replicate by replicate, bit for bit;
  and, without the reference's integer division, the numpy means and unbiased
  variances over the surviving replicates.
"""
import json
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "sweep_vectors.json")


@pytest.fixture(scope="module")
def golden():
    with open(GOLDEN) as f:
        return json.load(f)


@pytest.fixture(scope="module")
def host_tm():
    """libtm.so for its host-only functions (no device is touched)."""
    import trophic_maelstorm_amd as tm
    from trophic_maelstorm_amd import build as b
    if not os.path.exists(b.LIB):
        b.build()
    tm.lib()
    return tm


@pytest.fixture(scope="module")
def ref(oracle):
    if oracle.build_ref() is None:
        pytest.skip("reference sources not present")
    return oracle


# ------------------------------------------------------------------ time bins
def test_time_bins_equal_golden(host_tm, golden):
    for case in golden["time_bins"]:
        got = host_tm.log10_time_bins(case["t_max"], case["dt"])
        assert np.array_equal(got, np.array(case["t"], np.float64)), case["t_max"]


@pytest.mark.parametrize("t_max,dt", [(1e4, 0.1), (7.5e4, 0.01), (333.0, 0.3), (1e6, 0.5), (12.0, 0.7)])
def test_time_bins_equal_reference_build(host_tm, ref, t_max, dt):
    assert np.array_equal(host_tm.log10_time_bins(t_max, dt), ref.ref_log10_time_bins(t_max, dt))


def test_time_bins_properties(host_tm):
    dt = 0.1
    t = host_tm.log10_time_bins(1e4, dt)
    assert t[0] > 1.0 and t[-1] <= 1e4
    assert np.all(np.diff(t) >= dt / 2)
    k = t / dt
    assert np.allclose(k, np.round(k), atol=1e-6)                  # on the dt grid
    # about 25 points per decade over 4 decades, the first ones merged on the grid
    assert 90 <= t.size <= 101
    assert host_tm.log10_time_bins(1.0, dt).size == 0              # t_max <= 1: none


def test_sweep_times_insert_zero(host_tm):
    t = host_tm.sweep_times(1e3, 0.1)
    assert t[0] == 0.0 and np.all(np.diff(t) > 0)
    assert np.array_equal(t[1:], host_tm.log10_time_bins(1e3, 0.1))


# ------------------------------------------------------------ measurement steps
def test_measure_steps_on_grid(host_tm):
    for dt in (0.1, 0.005, 0.25):
        t = host_tm.sweep_times(2e3 if dt > 0.01 else 50.0, dt)
        steps = host_tm.measure_steps(t, dt)
        assert steps[0] == 0
        assert np.all(np.diff(steps) > 0)
        # the logged state is within dt/2 of the requested time
        assert np.all(np.abs(steps * dt - t) <= dt / 2 + 1e-9 * np.maximum(t, 1))
        assert np.array_equal(steps, np.round(t / dt).astype(np.int64))


def test_measure_steps_off_grid_point_stalls(host_tm):
    # 0.55 is 5.5 steps of 0.1: the window [0.5, 0.6) catches t = 0.5 (5 steps);
    # a point between two grid windows is reached by neither -> the rest stalls
    assert list(host_tm.measure_steps([0.0, 0.55, 1.0], 0.1)) == [0, 5, 10]
    with pytest.raises(ValueError):
        host_tm.measure_steps([0.0, 0.3, 0.3 + 1e-3, 1.0], 0.1)   # two points in one window


# ------------------------------------------------------- replicate statistics
def test_oracle_stats_equal_golden(oracle, golden):
    for run in golden["replicate_stats"]:
        x = np.array(run["series"], np.float64)
        got = oracle.replicate_stats(x, np.zeros(x.shape[0], np.int32), 1, 1)
        assert np.array_equal(got[0], np.array(run["table"], np.float64)), run["seed"]


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_oracle_stats_equal_reference_recurrence(oracle, ref, seed):
    """Reference function, fed the same replicates one after another, from the
    first replicate's table; both with the integer-division variance."""
    rng = np.random.default_rng(seed)
    n_rep, n_meas = 13, 33
    x = np.zeros((n_rep, n_meas, 2))
    for k in range(n_rep):
        end = int(rng.integers(0, n_meas + 3))
        x[k, :end, 0] = rng.integers(1, 10_000, size=min(end, n_meas))
        x[k, :end, 1] = rng.random(min(end, n_meas))
    pi = np.zeros(n_rep, np.int32)
    for flags in (1, 3):
        full = oracle.replicate_stats(x, pi, 1, flags)[0]
        tab = np.zeros((n_meas, 5))
        tab[:, 1:] = oracle.replicate_stats(x[:1], pi[:1], 1, flags)[0]
        for k in range(1, n_rep):
            ref.ref_var_mean_surv(tab, x[k])
        assert np.array_equal(tab[:, 1:], full)


def _sweep_series(rng, n_params, n_rep, n_meas, nf, revive=False):
    """Series of n_params x n_rep replicates (instance k: point k % n_params):
    field 0 alive up to a random measurement, then zero (or, with revive, some
    occupied cells again later); other fields always positive."""
    n = n_params * n_rep
    x = np.zeros((n, n_meas, 2 * nf))
    for k in range(n):
        end = int(rng.integers(0, n_meas + 3))
        for m in range(n_meas):
            if m < end or (revive and m > end + 2 and rng.random() < 0.5):
                x[k, m, 0] = rng.integers(1, 500)
                x[k, m, 1] = rng.random()
            for f in range(1, nf):
                x[k, m, 2 * f] = rng.integers(1, 500)
                x[k, m, 2 * f + 1] = 1.0 + rng.random()
    return x, (np.arange(n) % n_params).astype(np.int32)


@pytest.mark.parametrize("nf", [1, 3])
@pytest.mark.parametrize("flags", [0, 2])
def test_oracle_stats_closed_forms(oracle, nf, flags):
    """Without integer division: surviving = replicates occupied at every
    measurement so far, mean/var = numpy mean and unbiased variance of their
    spatial means, active = mean of their counts; other fields averaged over
    all replicates (zero after a replicate's end with flag 2)."""
    rng = np.random.default_rng(100 + nf + flags)
    n_params, n_rep, n_meas = 3, 7, 20
    x, pi = _sweep_series(rng, n_params, n_rep, n_meas, nf, revive=True)
    got = oracle.replicate_stats(x, pi, n_params, flags)
    alive = np.cumprod(x[:, :, 0] != 0, axis=1).astype(bool)        # occupied so far
    for g in range(n_params):
        ks = np.flatnonzero(pi == g)
        for m in range(n_meas):
            row = got[g, m]
            for f in range(1, nf):
                vals = x[ks, m, 2 * f + 1].copy()
                if flags & 2:
                    ended = np.array([m > np.argmin(alive[k]) and not alive[k].all() for k in ks])
                    vals[ended] = 0.0
                assert row[f - 1] == pytest.approx(vals.mean(), rel=1e-13)
            if m == 0 or not alive[ks[0], m]:
                continue   # the first replicate's own rule; covered below
            surv = ks[alive[ks, m]]
            assert row[nf + 1] == surv.size
            assert row[nf - 1] == pytest.approx(x[surv, m, 1].mean(), rel=1e-12)
            assert row[nf + 2] == pytest.approx(x[surv, m, 0].mean(), rel=1e-12)
            want_var = x[surv, m, 1].var(ddof=1) if surv.size > 1 else 0.0
            assert row[nf] == pytest.approx(want_var, rel=1e-9, abs=1e-15)


def test_oracle_stats_integer_division_closed_form(oracle):
    """With the reference's 1/(r-1) in integers, the variance after replicate r >= 3
    is (mean_r - mean_{r-1})^2 only; at r = 2 it is the true sample variance."""
    rng = np.random.default_rng(7)
    n_rep = 9
    x = np.zeros((n_rep, 1, 2))
    x[:, 0, 0] = 10.0
    x[:, 0, 1] = rng.random(n_rep)
    pi = np.zeros(n_rep, np.int32)
    for r in range(2, n_rep + 1):
        row = oracle.replicate_stats(x[:r], pi[:r], 1, 1)[0, 0]
        means = x[:r, 0, 1]
        if r == 2:
            assert row[1] == pytest.approx(means.var(ddof=1), rel=1e-12)
        else:
            assert row[1] == pytest.approx((means.mean() - means[:-1].mean()) ** 2, rel=1e-9, abs=1e-18)
        assert row[2] == r


def test_oracle_stats_first_replicate(oracle):
    x = np.array([[[5.0, 0.25, 3.0, 1.5, 4.0, 2.5], [0.0, 0.0, 7.0, 1.25, 1.0, 0.5]]])
    got = oracle.replicate_stats(x, np.zeros(1, np.int32), 1, 3)[0]
    assert got.tolist() == [[1.5, 2.5, 0.25, 0.0, 1.0, 5.0], [1.25, 0.5, 0.0, 0.0, 0.0, 0.0]]


def test_oracle_stats_empty_point_is_zero(oracle):
    x = np.ones((2, 4, 2))
    got = oracle.replicate_stats(x, np.array([0, 0], np.int32), 2, 3)
    assert np.all(got[1] == 0.0) and np.all(got[0, :, 2] == 2)
