"""GPU parity: the HIP step (through the C ABI) against the CPU oracle.

FP64: bit-exact per-cell Poisson integers and fields / rho(t) within 1e-12
relative (the fields in fact match to the last bit; the rho(t) means differ only
by summation order).  Run on an MI355X with `pytest -m gpu`.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

REL = 1e-12   # FP64 tolerance on fields and rho(t) means


def assert_rows_close(got, want, rel=REL):
    got, want = np.asarray(got), np.asarray(want)
    assert got.shape == want.shape
    cnt = got[..., 0::2], want[..., 0::2]
    assert np.array_equal(cnt[0], cnt[1]), "occupied-site counts differ"
    m_g, m_w = got[..., 1::2], want[..., 1::2]
    assert np.all(np.abs(m_g - m_w) <= rel * np.maximum(np.abs(m_w), 1e-300)), \
        f"max rel {np.max(np.abs(m_g - m_w) / np.maximum(np.abs(m_w), 1e-300))}"


def dp_case(tm, oracle, L, steps, seed=11, inst=3, step0=0, sampler=0, scheme=0, ramp=0,
            variant=-1, graph=100, a=0.72, init_hi=1.0):
    p = tm.DPParams(a=a, a_end=0.70 if ramp else None, ramp_steps=ramp, sampler=sampler, scheme=scheme)
    po = oracle.dp_params(a=a, a_end=0.70 if ramp else None, ramp_steps=ramp, sampler=sampler,
                          scheme=scheme)
    rho0 = oracle.init_uniform(L, [0.0], [init_hi], seed, inst)
    with tm.Context() as ctx:
        ctx.configure(tm.MODEL_DP, tm.F64, p, [tm.Instance(L, 0, seed, inst, step0)])
        ctx.set_launch(variant, graph)
        ctx.upload(0, rho0)
        ctx.record_draws(True)
        rows = ctx.run(steps)
        got = ctx.download(0)
        draws = ctx.draws(0)
    ref = oracle.dp_run(po, rho0, L, seed, inst, step0, steps, rows=True, draws=True)
    return got, rows[0], draws, ref


@pytest.mark.parametrize("L,steps", [(16, 300), (40, 120), (64, 100), (128, 60)])
def test_dp_f64_bit_exact(tm, oracle, L, steps):
    got, rows, draws, ref = dp_case(tm, oracle, L, steps)
    assert ref["status"] == 0
    assert np.array_equal(draws, ref["draws"]), "Poisson integers differ"
    assert np.array_equal(got, ref["fields"]), \
        f"max rel {np.max(np.abs(got - ref['fields']) / np.maximum(np.abs(ref['fields']), 1e-300))}"
    assert_rows_close(rows, ref["rows"])


@pytest.mark.parametrize("sampler,scheme,ramp,a,hi", [
    (1, 0, 0, 0.72, 1.0),       # reference Gaussian shortcuts (z > 100 needs rho > 11.8)
    (1, 2, 0, 0.72, 40.0),      # ... with rho up to 40 (exact update: Euler would go negative)
    (0, 2, 0, 0.72, 40.0),      # PTRS Poisson branch (z >= 12)
    (0, 1, 0, 0.72, 1.0),       # RK4 nonlinear sub-step
    (0, 2, 0, 0.72, 1.0),       # exact nonlinear update
    (0, 0, 40, 0.72, 1.0),      # quench ramp a(t)
    (0, 0, 0, 1.6, 1.0),        # active phase
])
def test_dp_f64_modes(tm, oracle, sampler, scheme, ramp, a, hi):
    got, rows, draws, ref = dp_case(tm, oracle, 32, 60, sampler=sampler, scheme=scheme, ramp=ramp,
                                    a=a, init_hi=hi)
    assert np.array_equal(draws, ref["draws"])
    assert np.array_equal(got, ref["fields"])
    assert_rows_close(rows, ref["rows"])


def test_dp_f64_tiles_graphs_and_step0_do_not_change_results(tm, oracle):
    base = None
    for variant in range(len(tm.variant_names(tm.MODEL_DP, tm.F64))):
        for graph in (0, 7):
            got, rows, draws, ref = dp_case(tm, oracle, 48, 23, step0=1000003, variant=variant,
                                            graph=graph)
            assert np.array_equal(got, ref["fields"])
            if base is None:
                base = got
            assert np.array_equal(got, base)


def test_dp_odd_sizes_every_variant_bit_exact(tm, oracle):
    """Every DP kernel variant on a batch of lattice sides that are not multiples
    of 4 or of the tile width (per-cell stores, partial groups and tiles, wrapped
    halos, unaligned instance offsets), Poisson integers recorded, against the
    oracle."""
    names = tm.variant_names(tm.MODEL_DP, tm.F64)
    Ls = [17, 130, 30, 64, 50]
    p = tm.DPParams(a=1.1)
    po = oracle.dp_params(a=1.1)
    insts = [tm.Instance(L, 0, 70 + k, k, 11 * k) for k, L in enumerate(Ls)]
    for variant in range(len(names)):
        with tm.Context() as ctx:
            ctx.configure(tm.MODEL_DP, tm.F64, p, insts)
            ctx.set_launch(variant, 6)
            ctx.init_uniform([0.0], [1.0])
            init = [ctx.download(k) for k in range(len(Ls))]
            ctx.record_draws(True)
            rows = ctx.run(13)
            got = [ctx.download(k) for k in range(len(Ls))]
            draws = [ctx.draws(k) for k in range(len(Ls))]
        for k, d in enumerate(insts):
            ref = oracle.dp_run(po, init[k][0], d.L, d.seed, d.instance_id, d.step0, 13, draws=True)
            assert np.array_equal(draws[k], ref["draws"]), (names[variant], d.L)
            assert np.array_equal(got[k], ref["fields"]), (names[variant], d.L)
            assert_rows_close(rows[k], ref["rows"])


def rk_case(tm, oracle, L, steps, seed=5, inst=1, sampler=0, variant=-1, graph=100, R=1.0,
            lo=(0.0, 0.0, 0.0), hi=(90.0, 1.0, 1.0)):
    p = tm.RietkerkParams(R=R, sampler=sampler)
    po = oracle.rk_params(R=R, sampler=sampler)
    f0 = oracle.init_uniform(L, list(lo), list(hi), seed, inst)
    with tm.Context() as ctx:
        ctx.configure(tm.MODEL_RIETKERK, tm.F64, p, [tm.Instance(L, 0, seed, inst, 0)])
        ctx.set_launch(variant, graph)
        ctx.upload(0, f0)
        ctx.record_draws(True)
        rows = ctx.run(steps)
        got = ctx.download(0)
        draws = ctx.draws(0)
    ref = oracle.rk_run(po, f0, L, seed, inst, 0, steps, rows=True, draws=True)
    return got, rows[0], draws, ref


@pytest.mark.parametrize("L,steps", [(16, 40), (24, 30), (64, 20)])
def test_rietkerk_f64_bit_exact(tm, oracle, L, steps):
    got, rows, draws, ref = rk_case(tm, oracle, L, steps)
    assert ref["status"] == 0
    assert np.array_equal(draws, ref["draws"]), "Poisson integers differ"
    assert np.array_equal(got, ref["fields"]), \
        f"max rel {np.max(np.abs(got - ref['fields']) / np.maximum(np.abs(ref['fields']), 1e-300))}"
    assert_rows_close(rows, ref["rows"])


def test_rietkerk_f64_compat_and_tiles(tm, oracle):
    for variant in range(len(tm.variant_names(tm.MODEL_RIETKERK, tm.F64))):
        got, rows, draws, ref = rk_case(tm, oracle, 40, 12, sampler=1, variant=variant, graph=5)
        assert np.array_equal(draws, ref["draws"])
        assert np.array_equal(got, ref["fields"])
        assert_rows_close(rows, ref["rows"])


def test_rietkerk_split_step_mixed_batch_bit_exact(tm, oracle):
    """Split step (Dornic kernel on its own tiles, then RK4) on a batch of uneven
    L, none a multiple of the Dornic tile, exact sampler, against the oracle."""
    names = tm.variant_names(tm.MODEL_RIETKERK, tm.F64)
    split = [k for k, n in enumerate(names) if "split" in n]
    assert split
    Ls = [24, 40, 100, 70]
    ps = [tm.RietkerkParams(R=0.9), tm.RietkerkParams(R=1.3)]
    pos = [oracle.rk_params(R=0.9), oracle.rk_params(R=1.3)]
    insts = [tm.Instance(L, k % 2, 40 + k, 3 * k + 1, 5 * k) for k, L in enumerate(Ls)]
    for variant in split:
        with tm.Context() as ctx:
            ctx.configure(tm.MODEL_RIETKERK, tm.F64, ps, insts)
            ctx.set_launch(variant, 4)
            ctx.init_uniform([0.0, 0.0, 0.0], [90.0, 1.0, 1.0])
            init = [ctx.download(k) for k in range(len(Ls))]
            ctx.record_draws(True)
            rows = ctx.run(9)
            got = [ctx.download(k) for k in range(len(Ls))]
            draws = [ctx.draws(k) for k in range(len(Ls))]
        for k, d in enumerate(insts):
            ref = oracle.rk_run(pos[d.params_index], init[k], d.L, d.seed, d.instance_id, d.step0, 9,
                                rows=True, draws=True)
            assert np.array_equal(draws[k], ref["draws"]), (names[variant], d.L)
            assert np.array_equal(got[k], ref["fields"]), (names[variant], d.L)
            assert_rows_close(rows[k], ref["rows"])


def test_mixed_size_batch_matches_single_runs(tm, oracle):
    Ls = [16, 32, 24, 64, 16]
    ps = [tm.DPParams(a=0.72), tm.DPParams(a=1.1, a_end=1.0, ramp_steps=10)]
    pos = [oracle.dp_params(a=0.72), oracle.dp_params(a=1.1, a_end=1.0, ramp_steps=10)]
    insts = [tm.Instance(L, k % 2, 100 + k, 7 * k, 3 * k) for k, L in enumerate(Ls)]
    with tm.Context() as ctx:
        ctx.configure(tm.MODEL_DP, tm.F64, ps, insts)
        ctx.init_uniform([0.0], [1.0])
        init = [ctx.download(k) for k in range(len(Ls))]
        rows = ctx.run(37)
        got = [ctx.download(k) for k in range(len(Ls))]
    for k, d in enumerate(insts):
        want0 = oracle.init_uniform(d.L, [0.0], [1.0], d.seed, d.instance_id)
        assert np.array_equal(init[k], want0)
        ref = oracle.dp_run(pos[d.params_index], want0, d.L, d.seed, d.instance_id, d.step0, 37)
        assert np.array_equal(got[k], ref["fields"])
        assert_rows_close(rows[k], ref["rows"])


def test_init_uniform_and_measure(tm, oracle):
    with tm.Context() as ctx:
        ctx.configure(tm.MODEL_RIETKERK, tm.F64, tm.RietkerkParams(), [tm.Instance(33, 0, 9, 4)])
        ctx.init_uniform([0.0, 0.0, 0.0], [90.0, 1.0, 1.0])
        f = ctx.download(0)
        m = ctx.measure()
    want = oracle.init_uniform(33, [0, 0, 0], [90, 1, 1], 9, 4)
    assert np.array_equal(f, want)
    assert_rows_close(m[0], oracle.reduce(want))


def test_absorbing_state_is_exact(tm, oracle):
    with tm.Context() as ctx:
        ctx.configure(tm.MODEL_DP, tm.F64, tm.DPParams(), [tm.Instance(32)])
        ctx.record_draws(True)
        rows = ctx.run(10)
        assert np.all(ctx.download(0) == 0.0)
        assert np.all(ctx.draws(0) == -1)
        assert np.all(rows == 0.0)


@pytest.mark.parametrize("variant", [None, "dp_f64_64x32_t256_x2", "dp_f64_64x32_t256_x4", "dp_f64_64x32_t256_x8"])
def test_numeric_fault_is_reported(tm, oracle, variant):
    """The first fault (instance, cell, step) is reported.  With a multi-step
    variant the run covers every step of one launch: the faults of step 0 must
    win over the later faults their neighbourhoods produce at the next steps."""
    L = 160
    rho = np.zeros((1, L, L))
    rho[0, 5, 7] = 100.0     # explicit Euler drives it negative: 100 - 0.1*2*100^2 < 0
    rho[0, 130, 140] = 100.0
    po = oracle.dp_params()
    ref = oracle.dp_run(po, rho, L, 1, 0, 0, 1)
    assert ref["status"] == 1
    with tm.Context() as ctx:
        ctx.configure(tm.MODEL_DP, tm.F64, tm.DPParams(), [tm.Instance(L, 0, 1, 0)])
        if variant:
            ctx.set_launch(tm.variant_names(tm.MODEL_DP, tm.F64).index(variant), 0)
        ctx.upload(0, rho)
        with pytest.raises(tm.NumericFault):
            ctx.run(int(variant[-1]) if variant else 1)
        inst, cell, step = ctx.fault_info()
        assert (inst, step) == (0, 0)
        assert cell == ref["fault_cell"]


def test_autotune_does_not_change_state(tm, oracle, tmp_path):
    with tm.Context() as ctx:
        ctx.configure(tm.MODEL_DP, tm.F64, tm.DPParams(), [tm.Instance(64, 0, 1, 0), tm.Instance(32, 0, 1, 1)])
        ctx.init_uniform([0.0], [1.0])
        ctx.run(3)
        before = [ctx.download(k) for k in range(2)]
        v, g = ctx.autotune(str(tmp_path / "tune.txt"))
        assert 0 <= v < len(tm.variant_names(tm.MODEL_DP, tm.F64))
        assert [np.array_equal(ctx.download(k), before[k]) for k in range(2)] == [True, True]
        assert ctx.step(0) == 3
        # cached choice is reused
        assert ctx.autotune(str(tmp_path / "tune.txt")) == (v, g)
        # the timing aid: a fixed-mode variant, timed twice with its kept graphs
        names = tm.variant_names(tm.MODEL_DP, tm.F64)
        for want, steps in (("dp_f64_64x32_t256", 4), ("dp_f64_64x32_t256_x4", 9)):
            ctx.set_launch(names.index(want), 10)
            n0 = ctx.step(1)
            m0 = ctx.mode_steps()
            assert ctx.time_steps(steps) > 0 and ctx.time_steps(steps) > 0
            assert ctx.step(1) == n0 + 2 * steps
            m1 = ctx.mode_steps()
            assert (m1[0] - m0[0]) + (m1[1] - m0[1]) == 2 * steps
        # an adaptive variant chooses its mode from the rows fold, which the timing aid skips
        ctx.set_launch(names.index("dp_f64_64x32_t256_x8aq"), 10)
        with pytest.raises(tm.TMError, match="TM_ERR_STATE"):
            ctx.time_steps(4)


@pytest.mark.parametrize("variant", ["dp_f64_64x32_t256_x4a", "dp_f64_64x32_t256_x4", "dp_f64_64x32_t256_x8aq",
                                     "dp_f64_64x32_t256_x8aq_w6", "dp_f64_64x32_t256_x8a",
                                     "dp_f64_64x32_t256_x8", "dp_f64_64x8_t128_x16"])
def test_multistep_variant_through_extinction_bit_exact(tm, oracle, variant):
    """A DP batch that dies out (a = 0.3) over 300 steps in blocks of 10: the
    adaptive variant starts with single steps on the live field and turns to
    K-step launches once 2% or less of the cells are occupied (step 133); both kinds
    of launch, and the switch between them, reproduce the oracle bit for bit."""
    Ls = [32, 40, 17]
    p = tm.DPParams(a=0.3)
    po = oracle.dp_params(a=0.3)
    insts = [tm.Instance(L, 0, 5, 300 + k) for k, L in enumerate(Ls)]
    print(f"{variant} through extinction", flush=True)
    with tm.Context() as ctx:
        ctx.configure(tm.MODEL_DP, tm.F64, p, insts)
        ctx.set_launch(tm.variant_names(tm.MODEL_DP, tm.F64).index(variant), 10)
        ctx.init_uniform([0.0], [1.0])
        init = [ctx.download(k) for k in range(len(Ls))]
        rows = np.concatenate([ctx.run(150), ctx.run(150)], axis=1)
        got = [ctx.download(k) for k in range(len(Ls))]
    for k, d in enumerate(insts):
        ref = oracle.dp_run(po, init[k][0], d.L, d.seed, d.instance_id, 0, 300)
        assert np.array_equal(got[k], ref["fields"]), (variant, d.L)
        assert_rows_close(rows[k], ref["rows"])
    occupied = rows[:, :, 0].sum(axis=0) / sum(L * L for L in Ls)
    assert occupied[0] > 0.9 and occupied[-1] <= 0.02     # both modes were used
