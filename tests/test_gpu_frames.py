"""Initial frames generated on the GPU equal the oracle's, and the step runs
from them bit-exactly (FP64)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def mixed_batch(tm, n_params):
    sizes = [17, 64, 32, 128, 40]
    return [tm.Instance(sizes[k % len(sizes)], k % n_params, 21, 500 + k) for k in range(10)]


@pytest.mark.parametrize("precision", ["F64", "F32"])
def test_speckle_frames_equal_oracle(tm, oracle, precision):
    ps = [tm.RietkerkParams(R=R) for R in (0.5, 1.0, 1.5)]
    perc = np.array([0.015, 0.2, 0.9])
    high = np.array([[90.0, 1.0, 2.0], [10.0, 0.5, 0.25], [1e-3, 3.0, 4.0]])
    low = np.array([[0.0, 1.0, 1.0], [1.0, 0.5, 0.5], [0.0, 0.0, 0.0]])
    insts = mixed_batch(tm, 3)
    print(f"speckle frames {precision}", flush=True)
    with tm.Context() as ctx:
        ctx.configure(tm.MODEL_RIETKERK, getattr(tm, precision), ps, insts)
        ctx.init_speckle(perc, high, low)
        got = [ctx.download(k) for k in range(len(insts))]
    for k, d in enumerate(insts):
        p = d.params_index
        want = oracle.init_speckle(d.L, perc[p], high[p], low[p], d.seed, d.instance_id)
        if precision == "F32":
            want = want.astype(np.float32).astype(np.float64)
        assert np.array_equal(got[k], want), k


@pytest.mark.parametrize("precision", ["F64", "F32"])
def test_gauss_frames_equal_oracle(tm, oracle, precision):
    ps = [tm.RietkerkParams(R=R) for R in (0.5, 1.0)]
    amp = np.array([[5.0, 1.0, 1.0], [50.0, 2.0, 0.5]])
    insts = mixed_batch(tm, 2)
    print(f"gauss frames {precision}", flush=True)
    for n_gauss, sd_div in [(1, [8.0, 16.0]), (3, [8.0, 16.0]), (2, [5.0]), (0, [8.0])]:
        with tm.Context() as ctx:
            ctx.configure(tm.MODEL_RIETKERK, getattr(tm, precision), ps, insts)
            ctx.init_gauss(amp, n_gauss, sd_div)
            got = [ctx.download(k) for k in range(len(insts))]
        for k, d in enumerate(insts):
            want = oracle.init_gauss(d.L, n_gauss, sd_div, amp[d.params_index])
            if precision == "F32":
                want = want.astype(np.float32).astype(np.float64)
            assert np.array_equal(got[k], want), (n_gauss, k)


def test_rietkerk_from_reference_gauss_frame_bit_exact(tm, oracle):
    """The reference sweep's frame (amp {5, 1, 1}, sd {g/8, g/16}) then 3 steps."""
    L, steps = 48, 3
    print("Rietkerk F64 from the Gaussian frame", flush=True)
    with tm.Context() as ctx:
        ctx.configure(tm.MODEL_RIETKERK, tm.F64, tm.RietkerkParams(R=1.2), [tm.Instance(L, 0, 8, 3)])
        ctx.init_gauss([5.0, 1.0, 1.0], 1, [8.0, 16.0])
        ctx.record_draws(True)
        rows = ctx.run(steps)
        got, draws = ctx.download(0), ctx.draws(0)
    f0 = oracle.init_gauss(L, 1, [8.0, 16.0], [5.0, 1.0, 1.0])
    ref = oracle.rk_run(oracle.rk_params(R=1.2), f0, L, 8, 3, 0, steps, draws=True)
    assert np.array_equal(draws, ref["draws"])
    assert np.array_equal(got, ref["fields"])
    np.testing.assert_allclose(rows[0], ref["rows"], rtol=1e-12)


def test_dp_from_speckle_frame_bit_exact(tm, oracle):
    L, steps = 64, 5
    print("DP F64 from a speckle frame", flush=True)
    with tm.Context() as ctx:
        ctx.configure(tm.MODEL_DP, tm.F64, tm.DPParams(a=1.0), [tm.Instance(L, 0, 4, 77)])
        ctx.init_speckle([0.1], [[2.0]], [[0.0]])
        rows = ctx.run(steps)
        got = ctx.download(0)
    rho0 = oracle.init_speckle(L, 0.1, [2.0], [0.0], 4, 77)
    ref = oracle.dp_run(oracle.dp_params(a=1.0), rho0, L, 4, 77, 0, steps)
    assert np.array_equal(got, ref["fields"])
    np.testing.assert_allclose(rows[0], ref["rows"], rtol=1e-12)
