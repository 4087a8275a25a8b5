"""Parity at the benchmark sizes, through properties that do not depend on size.

* FP64 bit-exact against the oracle at the largest single instances of the
  batches: DP at L = 1024 (the uneven DP batch's largest lattice) and
  Rietkerk at L = 512 (the batched sweep's lattice), a few steps each.
* The whole uneven DP batch (256 instances, 35.7 M cells) and the whole
  Rietkerk sweep (256 x 512^2) in FP32: each instance's result does not depend
  on the batch it runs in (bit-identical to the instance run alone), and the
  rho(t) rows equal the spatial means and counts recomputed from the
  downloaded fields.
* The 8192^2 Rietkerk lattice in FP32: two runs are bit-identical, and the
  rows match the fields.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def dp_batch(tm):
    sizes = [128] * 128 + [256] * 64 + [512] * 48 + [1024] * 16
    p = tm.DPParams(a=0.72, a_end=0.70, ramp_steps=10_000, b=2.0, D=1.0, sigma=math.sqrt(2.0), dt=0.1)
    return p, [tm.Instance(L, 0, 20251202, k) for k, L in enumerate(sizes)]


def rk_batch(tm):
    Rs = np.linspace(0.5, 1.5, 16)
    ps = [tm.RietkerkParams(R=float(R)) for R in Rs]
    return ps, [tm.Instance(512, k % 16, 20251202 + k // 16, k) for k in range(256)]


def check_rows_against_fields(fields, row, nf):
    for f in range(nf):
        x = fields[f].astype(np.float64)
        assert row[2 * f] == np.count_nonzero(x > 0)
        assert row[2 * f + 1] == pytest.approx(math.fsum(x.ravel()) / x.size, rel=1e-12, abs=1e-300)


def test_dp_f64_bit_exact_at_1024(tm, oracle):
    L, steps = 1024, 3
    rho0 = oracle.init_uniform(L, [0.0], [1.0], 5, 250)
    p = tm.DPParams(a=0.72, a_end=0.70, ramp_steps=10_000)
    with tm.Context() as ctx:
        ctx.configure(tm.MODEL_DP, tm.F64, p, [tm.Instance(L, 0, 5, 250, 777)])
        ctx.upload(0, rho0)
        ctx.record_draws(True)
        rows = ctx.run(steps)
        got, draws = ctx.download(0), ctx.draws(0)
    ref = oracle.dp_run(oracle.dp_params(a=0.72, a_end=0.70, ramp_steps=10_000), rho0, L, 5, 250, 777, steps,
                        draws=True)
    assert np.array_equal(draws, ref["draws"])
    assert np.array_equal(got, ref["fields"])
    np.testing.assert_allclose(rows[0], ref["rows"], rtol=1e-12)


def test_rietkerk_f64_bit_exact_at_512(tm, oracle):
    L, steps = 512, 2
    f0 = oracle.init_uniform(L, [0.0, 0.0, 0.0], [90.0, 1.0, 1.0], 6, 17)
    with tm.Context() as ctx:
        ctx.configure(tm.MODEL_RIETKERK, tm.F64, tm.RietkerkParams(R=1.1), [tm.Instance(L, 0, 6, 17)])
        ctx.upload(0, f0)
        ctx.record_draws(True)
        rows = ctx.run(steps)
        got, draws = ctx.download(0), ctx.draws(0)
    ref = oracle.rk_run(oracle.rk_params(R=1.1), f0, L, 6, 17, 0, steps, draws=True)
    assert np.array_equal(draws, ref["draws"])
    assert np.array_equal(got, ref["fields"])
    np.testing.assert_allclose(rows[0], ref["rows"], rtol=1e-12)


@pytest.mark.parametrize("model", ["dp", "rietkerk"])
def test_full_batch_fp32_instances_independent_of_batch(tm, model):
    if model == "dp":
        params, insts = dp_batch(tm)
        mid, lo, hi, steps, pick = tm.MODEL_DP, [0.0], [1.0], 30, [0, 130, 200, 255]
        vname = "dp_f32_128x32_t256_w6"
    else:
        params, insts = rk_batch(tm)
        mid, lo, hi, steps, pick = tm.MODEL_RIETKERK, [0.0, 0.0, 0.0], [90.0, 1.0, 1.0], 6, [0, 77, 255]
        vname = "rk_f32_split_32x16_t256_q24"   # the FP32 batch default
    variant = tm.variant_names(mid, tm.F32).index(vname)   # the same kernels in both runs
    nf = 1 if model == "dp" else 3
    with tm.Context() as ctx:
        ctx.configure(mid, tm.F32, params, insts)
        ctx.set_launch(variant, 100)
        ctx.init_uniform(lo, hi)
        rows = ctx.run(steps)
        batch = {k: ctx.download(k) for k in pick}
    for k in pick:
        check_rows_against_fields(batch[k], rows[k, -1], nf)
        d = insts[k]
        plist = params if isinstance(params, list) else [params]
        with tm.Context() as ctx:
            ctx.configure(mid, tm.F32, plist[d.params_index], [tm.Instance(d.L, 0, d.seed, d.instance_id)])
            ctx.set_launch(variant, 100)
            ctx.init_uniform(lo, hi)
            alone_rows = ctx.run(steps)
            alone = ctx.download(0)
        assert np.array_equal(alone, batch[k]), k
        assert np.array_equal(alone_rows[0], rows[k]), k


def test_8192_lattice_fp32_deterministic(tm):
    p = tm.RietkerkParams(R=0.75)
    out = []
    for _ in range(2):
        with tm.Context() as ctx:
            ctx.configure(tm.MODEL_RIETKERK, tm.F32, p, [tm.Instance(8192, 0, 20251202, 0)])
            ctx.init_uniform([0.0, 0.0, 0.0], [90.0, 1.0, 1.0])
            rows = ctx.run(3)
            out.append((ctx.download(0), rows))
    (f1, r1), (f2, r2) = out
    assert np.array_equal(f1, f2)
    assert np.array_equal(r1, r2)
    check_rows_against_fields(f1, r1[0, -1], 3)


def test_default_dp_rows_do_not_depend_on_mode_or_batch(tm):
    """The shipped FP32 DP default switches from single-step to eight-step launches
    when fewer than 2% of the cells are occupied, on the batch's own occupancy, at
    a block boundary whose timing the host picks up asynchronously.  Every launch
    kind forms its rho(t) partials in the same thread order, so the rows do not
    depend on where the switch falls: two runs of the whole C4 batch past the
    switch give identical rows, and instances run alone (which switch at other
    steps) give their rows in the batch bit for bit."""
    params, insts = dp_batch(tm)
    steps, pick = 728, [3, 140, 230]
    runs = []
    for _ in range(2):
        with tm.Context() as ctx:
            ctx.configure(tm.MODEL_DP, tm.F32, params, insts)
            ctx.init_uniform([0.0], [1.0])
            runs.append(ctx.run(steps))
            single, blocked = ctx.mode_steps()
    assert blocked > 0 and single > 0, (single, blocked)   # the run crossed the switch
    assert np.array_equal(runs[0], runs[1])
    for k in pick:
        d = insts[k]
        with tm.Context() as ctx:
            ctx.configure(tm.MODEL_DP, tm.F32, params, [tm.Instance(d.L, 0, d.seed, d.instance_id)])
            ctx.init_uniform([0.0], [1.0])
            alone = ctx.run(steps)
        assert np.array_equal(alone[0], runs[0][k]), k


def test_c4_fp32_quench_through_extinction_matches_fp64(tm):
    """The whole C4 batch with the shipped defaults, FP32 against FP64 on the same
    seeds, for 1,120 steps: through the live start, the switch to eight-step
    launches and past extinction.  The acceptance of the FP32 path
    (trophic_maelstorm_amd.compare): ensemble rho(t) within 1e-3 relative while
    the FP64 ensemble holds >= 10% of its initial mean; afterwards within 1e-3 or
    within 4 standard errors of two independent ensembles (same-seed trajectories
    decorrelate once an integer draw differs); surviving fractions within the
    two-proportion 95% band (plus one instance) at every step."""
    from trophic_maelstorm_amd import compare
    params, insts = dp_batch(tm)
    steps = 1120
    rows = {}
    for prec in (tm.F32, tm.F64):
        with tm.Context() as ctx:
            ctx.configure(tm.MODEL_DP, prec, params, insts)
            ctx.init_uniform([0.0], [1.0])
            rows[prec] = ctx.run(steps)
    rep = compare.field_report(rows[tm.F32], rows[tm.F64])
    print(rep)
    assert rep["populated_steps"] > 20
    assert rep["ensemble_rel_max_populated"] < 1e-3, rep
    assert rep["tail_ok"] and rep["survival_ok"], rep
    assert rep["accepted"]
    # the run reaches extinction: no FP64 instance is occupied at the end, and at most
    # a straggler or two in FP32 (the survival band above bounds the difference)
    assert not np.any(rows[tm.F64][:, -1, 0] > 0)
    assert np.count_nonzero(rows[tm.F32][:, -1, 0] > 0) <= 2
