"""A parameter sweep as one batch on the GPU: measurement series and
across-replicate statistics, against the oracle.

* The series recorded by tm_run_series equals the rows tm_run returns for the
  same steps (bit for bit) and, in FP64, the oracle's rows.
* The device statistics equal the oracle's tmo_replicate_stats on the same
  series bit for bit, for every flag combination, DP and Rietkerk, FP64 and FP32.
* Several advance chunks (a batch whose rows outgrow one chunk) give the same
  series as one tm_run.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

A_VALUES = (0.6, 0.9, 1.2)    # extinct early, mixed, surviving at L = 16 / 32 over ~900 steps
N_REP = 6


def dp_sweep(tm, n_rep=N_REP, sizes=(16, 32, 16)):
    ps = [tm.DPParams(a=float(a)) for a in A_VALUES]
    n_p = len(ps)
    insts = [tm.Instance(sizes[k % n_p], k % n_p, 9, 100 + k // n_p) for k in range(n_p * n_rep)]
    return ps, insts


def test_series_equals_run_rows_and_oracle(tm, oracle):
    ps, insts = dp_sweep(tm)
    t = tm.sweep_times(100.0, 0.1)
    steps = tm.measure_steps(t, 0.1)
    assert steps[0] == 0 and steps[-1] == 912     # the last bin is 10^1.96 = 91.2
    print(f"DP F64 sweep: {len(insts)} instances, {steps.size} measurements", flush=True)
    with tm.Context() as a, tm.Context() as b:
        for ctx in (a, b):
            ctx.configure(tm.MODEL_DP, tm.F64, ps, insts)
            ctx.init_uniform([0.0], [1.0])
        series = a.run_series(steps)
        row0 = b.measure()
        rows = b.run(int(steps[-1]))
        assert np.array_equal(series[:, 0], row0)
        assert np.array_equal(series[:, 1:], rows[:, steps[1:] - 1])
        for k in (0, 4, 17):
            assert np.array_equal(a.download(k), b.download(k))
    deaths = 0
    for k, d in enumerate(insts):
        rho0 = oracle.init_uniform(d.L, [0.0], [1.0], d.seed, d.instance_id)
        ref = oracle.dp_run(oracle.dp_params(a=ps[d.params_index].a), rho0, d.L, d.seed, d.instance_id, 0,
                            int(steps[-1]))["rows"]
        assert np.array_equal(series[k, 1:, 0], ref[steps[1:] - 1, 0]), k
        np.testing.assert_allclose(series[k, 1:, 1], ref[steps[1:] - 1, 1], rtol=1e-12, atol=0)
        deaths += int(np.any(series[k, :, 0] == 0))
    assert 0 < deaths < len(insts)    # the statistics below see both kinds of replicate


@pytest.mark.parametrize("precision", ["F64", "F32"])
def test_dp_replicate_stats_equal_oracle(tm, oracle, precision):
    ps, insts = dp_sweep(tm)
    steps = tm.measure_steps(tm.sweep_times(100.0, 0.1), 0.1)
    pidx = np.array([d.params_index for d in insts], np.int32)
    print(f"DP {precision} replicate stats, {len(insts)} instances", flush=True)
    with tm.Context() as ctx:
        ctx.configure(tm.MODEL_DP, getattr(tm, precision), ps, insts)
        ctx.init_uniform([0.0], [1.0])
        series = ctx.run_series(steps)
        for flags in range(4):
            got = ctx.replicate_stats(flags)
            want = oracle.replicate_stats(series, pidx, len(ps), flags)
            assert got.shape == (len(ps), steps.size, 4)
            assert np.array_equal(got, want), flags
    # the surviving count only falls, from N_REP at t = 0
    surv = got[:, :, 2]
    assert np.all(surv[:, 0] == N_REP) and np.all(np.diff(surv, axis=1) <= 0)


def test_rietkerk_replicate_stats_equal_oracle(tm, oracle):
    Rs = (0.6, 1.4)
    ps = [tm.RietkerkParams(R=R) for R in Rs]
    insts = [tm.Instance(32, k % 2, 5, 40 + k // 2) for k in range(8)]
    steps = tm.measure_steps(tm.sweep_times(2.0, 0.005), 0.005)
    pidx = np.array([d.params_index for d in insts], np.int32)
    print(f"Rietkerk F32 replicate stats, {steps.size} measurements", flush=True)
    with tm.Context() as ctx:
        ctx.configure(tm.MODEL_RIETKERK, tm.F32, ps, insts)
        ctx.init_uniform([0.0, 0.0, 0.0], [90.0, 1.0, 1.0])
        series = ctx.run_series(steps)
        for flags in (3, 0):
            got = ctx.replicate_stats(flags)
            assert got.shape == (2, steps.size, 6)
            assert np.array_equal(got, oracle.replicate_stats(series, pidx, 2, flags)), flags
    # W and O columns (flags 0): means over the four replicates of each point
    for g in range(2):
        np.testing.assert_allclose(got[g, :, 0], series[pidx == g, :, 3].mean(axis=0), rtol=1e-13)
        np.testing.assert_allclose(got[g, :, 1], series[pidx == g, :, 5].mean(axis=0), rtol=1e-13)


def test_series_over_several_chunks(tm):
    # 2048 instances x 2 row values: a rows chunk of ~32 MiB holds 1000 steps
    p = tm.DPParams(a=1.2)
    insts = [tm.Instance(16, 0, 3, k) for k in range(2048)]
    steps = np.array([0, 1, 999, 1000, 1001, 1999, 2000, 2001, 2500], np.int64)
    print("DP F32 series over 3 chunks, 2048 instances", flush=True)
    with tm.Context() as a, tm.Context() as b:
        for ctx in (a, b):
            ctx.configure(tm.MODEL_DP, tm.F32, p, insts)
            ctx.init_uniform([0.0], [1.0])
        series = a.run_series(steps)
        rows = b.run(int(steps[-1]))
        assert np.array_equal(series[:, 1:], rows[:, steps[1:] - 1])
        assert np.array_equal(a.download(2047), b.download(2047))
        assert a.step(0) == b.step(0) == 2500


def test_sweep_call_errors(tm):
    with tm.Context() as ctx:
        ctx.configure(tm.MODEL_DP, tm.F64, tm.DPParams(), [tm.Instance(16, 0, 1, 0)])
        with pytest.raises(tm.TMError, match="TM_ERR_STATE"):
            ctx.replicate_stats()
        with pytest.raises(tm.TMError, match="TM_ERR_ARG"):
            ctx.run_series([0, 5, 3])
        with pytest.raises(tm.TMError, match="TM_ERR_ARG"):
            ctx.run_series([-1, 2])
        ctx.run_series([0, 2])
        with pytest.raises(tm.TMError, match="TM_ERR_ARG"):
            ctx.replicate_stats(flags=4)
