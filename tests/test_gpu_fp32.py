"""FP32-storage path against the FP64 path (the FP64 GPU oracle) on the same seeds.

The FP32 step forms the Poisson mean lambda e^{beta dt} rho and draws the Poisson
integer in FP64 from the same Philox words as the FP64 step, so from identical
FP32-representable fields both paths draw the same integers except where the
FP32 rounding of rho moves the mean across a sampler decision.  Tolerances:
one step -- draws equal in >= 99.9% of live cells, rho* within 2e-6 relative
where they are equal; ensembles -- rho(t) ensemble mean within 1e-3 relative.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def one_step(tm, model, params, L, lo, hi, seed=3, inst=2):
    out = {}
    for prec in (tm.F32, tm.F64):
        with tm.Context() as ctx:
            ctx.configure(model, prec, params, [tm.Instance(L, 0, seed, inst)])
            ctx.init_uniform(lo, hi)
            f0 = ctx.download(0)
            ctx.record_draws(True)
            ctx.run(1)
            out[prec] = (f0, ctx.download(0), ctx.draws(0))
    return out


def run_batch(tm, model, prec, params, insts, lo, hi, steps):
    with tm.Context() as ctx:
        ctx.configure(model, prec, params, insts)
        ctx.init_uniform(lo, hi)
        return ctx.run(steps)


def test_fp32_init_is_rounded_fp64_init(tm):
    o = one_step(tm, tm.MODEL_RIETKERK, tm.RietkerkParams(), 64, [0, 0, 0], [90, 1, 1])
    assert np.array_equal(o[tm.F32][0], o[tm.F64][0].astype(np.float32).astype(np.float64))


@pytest.mark.parametrize("model", ["dp", "rietkerk"])
def test_fp32_one_step_tracks_fp64(tm, model):
    if model == "dp":
        # FP64 run from the FP32-rounded init so both start from identical fields
        L = 128
        p = tm.DPParams()
        res = {}
        for prec in (tm.F32, tm.F64):
            with tm.Context() as ctx:
                ctx.configure(tm.MODEL_DP, prec, p, [tm.Instance(L, 0, 3, 2)])
                ctx.init_uniform([0.0], [1.0])
                f0 = ctx.download(0).astype(np.float32).astype(np.float64)
                ctx.upload(0, f0)
                ctx.record_draws(True)
                ctx.run(1)
                res[prec] = (ctx.download(0), ctx.draws(0))
        f32, d32 = res[tm.F32]
        f64, d64 = res[tm.F64]
        field = 0
    else:
        L = 64
        res = {}
        for prec in (tm.F32, tm.F64):
            with tm.Context() as ctx:
                ctx.configure(tm.MODEL_RIETKERK, prec, tm.RietkerkParams(), [tm.Instance(L, 0, 3, 2)])
                ctx.init_uniform([0, 0, 0], [90, 1, 1])
                f0 = ctx.download(0).astype(np.float32).astype(np.float64)
                ctx.upload(0, f0)
                ctx.record_draws(True)
                ctx.run(1)
                res[prec] = (ctx.download(0), ctx.draws(0))
        f32, d32 = res[tm.F32]
        f64, d64 = res[tm.F64]
        field = 0
    live = d64 >= 0
    same = (d32 == d64) & live
    assert same.sum() >= 0.999 * live.sum()
    a, b = f32[field][same], f64[field][same]
    rel = np.abs(a - b) / np.maximum(np.abs(b), 1e-30)
    # FP32 rounding of the Gamma proposal; a Marsaglia-Tsang accept/reject decision
    # that falls within rounding of its threshold gives that cell a different
    # (equally valid) deviate, which must stay rare.
    assert np.median(rel) < 2e-7, np.median(rel)
    # the U^(1/shape) boost of small shapes multiplies the FP32 error of log2(U) by 1/shape
    assert np.quantile(rel, 0.99) < 1e-4, np.quantile(rel, 0.99)
    assert np.mean(rel > 1e-2) < 1e-3, np.mean(rel > 1e-2)


def ensemble_checks(r32, r64, field=0, floor=0.1, rel_tol=1e-3):
    """FP32 vs FP64 ensemble rho(t) on the same seeds: within rel_tol while the
    ensemble mean is at least `floor` of its initial value, and no bias over the
    whole run: paired per-instance differences within 4 standard errors, or below
    2e-5 of the initial mean.  Pairs on the same seeds share nearly all their
    draws, so their differences are mostly deterministic FP32 rounding rather than
    noise, and the standard error alone flags rounding-sized offsets: on the
    Rietkerk P field (values near 45 moved by small RK4 increments) the offset
    reaches 6e-6 of the mean at z above 4 for some seed sets, 170 times inside
    rel_tol."""
    c = 2 * field + 1
    m32, m64 = r32[:, :, c].mean(0), r64[:, :, c].mean(0)
    live = m64 >= floor * m64[0]
    rel = np.abs(m32 - m64)[live] / m64[live]
    assert rel.max() < rel_tol, rel.max()
    d = r32[:, :, c] - r64[:, :, c]
    se = d.std(0, ddof=1) / np.sqrt(d.shape[0])
    z = np.abs(d.mean(0)) / np.maximum(se, 1e-300)
    assert np.all((z < 4.0) | (np.abs(d.mean(0)) <= 2e-5 * np.abs(m64[0]))), z.max()


def test_fp32_dp_ensemble_matches_fp64(tm):
    insts = [tm.Instance(128, 0, 17, k) for k in range(128)]
    p = tm.DPParams(a=0.72, a_end=0.70, ramp_steps=400)
    r32 = run_batch(tm, tm.MODEL_DP, tm.F32, p, insts, [0.0], [1.0], 400)
    r64 = run_batch(tm, tm.MODEL_DP, tm.F64, p, insts, [0.0], [1.0], 400)
    ensemble_checks(r32, r64)


def test_fp32_rietkerk_ensemble_matches_fp64(tm):
    ps = [tm.RietkerkParams(R=R) for R in (0.8, 1.2)]
    insts = [tm.Instance(64, k % 2, 23, k) for k in range(32)]
    r32 = run_batch(tm, tm.MODEL_RIETKERK, tm.F32, ps, insts, [0, 0, 0], [90, 1, 1], 200)
    r64 = run_batch(tm, tm.MODEL_RIETKERK, tm.F64, ps, insts, [0, 0, 0], [90, 1, 1], 200)
    for f in range(3):
        ensemble_checks(r32, r64, field=f)


def test_fp32_rietkerk_split_matches_fused(tm):
    """The split FP32 step runs the same per-cell arithmetic as the fused kernel."""
    names = tm.variant_names(tm.MODEL_RIETKERK, tm.F32)
    out = {}
    for want in ("rk_f32_32x32_t256", "rk_f32_split_32x32_t256"):
        with tm.Context() as ctx:
            ctx.configure(tm.MODEL_RIETKERK, tm.F32, tm.RietkerkParams(R=1.1),
                          [tm.Instance(96, 0, 8, k) for k in range(3)])
            ctx.set_launch(names.index(want), 10)
            ctx.init_uniform([0, 0, 0], [90, 1, 1])
            ctx.record_draws(True)
            rows = ctx.run(20)
            out[want] = (np.stack([ctx.download(k) for k in range(3)]), ctx.draws(1), rows)
    (f_a, d_a, r_a), (f_b, d_b, r_b) = out.values()
    assert np.mean(d_a == d_b) >= 0.999
    rel = np.abs(f_a - f_b) / np.maximum(np.abs(f_b), 1e-30)
    assert np.quantile(rel, 0.99) < 1e-5, np.quantile(rel, 0.99)
    np.testing.assert_allclose(r_a[:, :, 1::2], r_b[:, :, 1::2], rtol=1e-5)


def test_fp32_dp_variants_agree(tm):
    """Every FP32 DP variant runs the same per-cell arithmetic from the same Philox
    words, so tile shape, register budget and tile order leave results unchanged."""
    names = tm.variant_names(tm.MODEL_DP, tm.F32)
    insts = [tm.Instance(L, 0, 9, k) for k, L in enumerate([96, 130, 64, 50])]
    out = []
    for v in range(len(names)):
        with tm.Context() as ctx:
            ctx.configure(tm.MODEL_DP, tm.F32, tm.DPParams(a=0.9), insts)
            ctx.set_launch(v, 10)
            ctx.init_uniform([0.0], [1.0])
            rows = ctx.run(30)
            out.append((np.concatenate([ctx.download(k).ravel() for k in range(len(insts))]), rows))
    f0, r0 = out[0]
    for v, (f, r) in enumerate(out[1:], 1):
        assert np.array_equal(f, f0), names[v]
        np.testing.assert_allclose(r, r0, rtol=1e-12, err_msg=names[v])


@pytest.mark.parametrize("hi,sampler", [(1.0, 0), (3.0, 0), (4.0, 0), (30.0, 0), (30.0, 1), (3.0, 1)])
def test_fp32_deferring_step_is_identical(tm, hi, sampler):
    """The two-pass single step (cells whose first Gamma or PTRS trial is rejected,
    or that need a fifth word, deferred to a queue in LDS) computes every cell as
    the one-pass step does: the same field and the same draws.  hi = 3 and 4 mix
    PTRS cells in; hi = 30 gives Poisson means up to 250 and Gamma shapes above
    100, where the reference's Gaussian shortcuts (sampler 1, compat) take over:
    the first pass defers compat means above 100 and draws the compat Gamma
    shortcut itself."""
    names = tm.variant_names(tm.MODEL_DP, tm.F32)
    insts = [tm.Instance(L, 0, 5, k) for k, L in enumerate([96, 130, 64, 50, 257])]
    out = {}
    for want in ("dp_f32_128x32_t256_w6", "dp_f32_128x24_t256_w6q"):
        with tm.Context() as ctx:
            # the exact nonlinear sub-step keeps large densities positive
            ctx.configure(tm.MODEL_DP, tm.F32, tm.DPParams(a=0.9, scheme=tm.DP_EXACT, sampler=sampler), insts)
            ctx.set_launch(names.index(want), 4)
            ctx.init_uniform([0.0], [hi])
            ctx.record_draws(True)
            rows = ctx.run(4)
            out[want] = (np.concatenate([ctx.download(k).ravel() for k in range(len(insts))]),
                         np.concatenate([ctx.draws(k).ravel() for k in range(len(insts))]), rows)
    (f_a, d_a, r_a), (f_b, d_b, r_b) = out.values()
    assert np.array_equal(f_a, f_b)
    assert np.array_equal(d_a, d_b)
    assert (d_a >= 0).sum() > 0.5 * d_a.size
    np.testing.assert_allclose(r_a, r_b, rtol=1e-12)


@pytest.mark.parametrize("prec, one_pass, two_pass, sampler", [
    ("F32", "dp_f32_128x32_t256_w6", "dp_f32_128x24_t256_w6q", 1),
    ("F64", "dp_f64_64x32_t256", "dp_f64_64x32_t256_x8aq_w6", 0),
])
def test_deferring_step_with_every_cell_deferred(tm, prec, one_pass, two_pass, sampler):
    """A full queue: with densities in [20, 30) every Poisson mean is above 100
    (about 8.5 per unit density here), so the first pass defers every cell -- FP64
    every PTRS mean, FP32 under the compat sampler every mean above 100 -- and the
    queue holds the whole tile.  The field, the draws and the rho(t) rows are
    still the one-pass step's."""
    P = getattr(tm, prec)
    names = tm.variant_names(tm.MODEL_DP, P)
    insts = [tm.Instance(L, 0, 9, k) for k, L in enumerate([96, 130, 64, 257])]
    out = {}
    for want in (one_pass, two_pass):
        with tm.Context() as ctx:
            ctx.configure(tm.MODEL_DP, P, tm.DPParams(a=0.9, scheme=tm.DP_EXACT, sampler=sampler), insts)
            ctx.set_launch(names.index(want), 1)
            ctx.init_uniform([20.0], [30.0])
            ctx.record_draws(True)
            rows = ctx.run(1)
            out[want] = (np.concatenate([ctx.download(k).ravel() for k in range(len(insts))]),
                         np.concatenate([ctx.draws(k).ravel() for k in range(len(insts))]), rows)
    (f_a, d_a, r_a), (f_b, d_b, r_b) = out.values()
    assert (d_a >= 100).mean() > 0.99            # every cell sampled, nearly all at means above 100
    assert np.array_equal(f_a, f_b)
    assert np.array_equal(d_a, d_b)
    assert np.array_equal(r_a, r_b)


@pytest.mark.parametrize("variant", ["dp_f32_128x32_t256_w6", "dp_f32_128x24_t256_w6q"])
def test_fp32_numeric_fault_is_reported(tm, variant):
    """The first fault (instance, cell, step) is reported by the FP32 step as by the
    FP64 step.  With b = 0.3 the explicit Euler sub-step y - 0.1 b y^2 turns negative
    only above y = 33: the planted densities (about 72 after the transition) fault,
    their neighbours (about 8.5) do not.  A density of 100 has a Poisson mean far
    above 12 (PTRS): the two-pass step records the fault from its first pass when
    the first PTRS and Gamma trials accept, from its second pass otherwise."""
    L = 160
    rho = np.zeros((1, L, L))
    rho[0, 5, 7] = 100.0
    rho[0, 130, 140] = 100.0
    info = {}
    for prec, name in ((tm.F32, variant), (tm.F64, None)):
        with tm.Context() as ctx:
            ctx.configure(tm.MODEL_DP, prec, tm.DPParams(b=0.3), [tm.Instance(L, 0, 1, 0)])
            if name:
                ctx.set_launch(tm.variant_names(tm.MODEL_DP, prec).index(name), 0)
            ctx.upload(0, rho)
            with pytest.raises(tm.NumericFault):
                ctx.run(1)
            info[prec] = ctx.fault_info()
    assert info[tm.F32] == info[tm.F64] == (0, 5 * L + 7, 0)


@pytest.mark.parametrize("sampler", [0, 1])
@pytest.mark.parametrize("want", ["rk_f32_split_32x16_t256_q", "rk_f32_split_32x16_t256_q24"])
def test_fp32_rietkerk_deferring_dornic_is_identical(tm, want, sampler):
    """The two-pass Dornic kernel of the split Rietkerk step (one PTRS trial and one
    Gamma trial first, the rest of the cells from a queue in LDS) gives the same
    P*, draws, fields and rho(t) rows as the one-pass kernel, on odd and even
    lattices, under both samplers."""
    names = tm.variant_names(tm.MODEL_RIETKERK, tm.F32)
    insts = [tm.Instance(L, k % 2, 8, k) for k, L in enumerate([96, 130, 64, 50, 257])]
    ps = [tm.RietkerkParams(R=0.9, sampler=sampler), tm.RietkerkParams(R=1.2, sampler=sampler)]
    out = {}
    for v in ("rk_f32_split_32x16_t256", want):
        with tm.Context() as ctx:
            ctx.configure(tm.MODEL_RIETKERK, tm.F32, ps, insts)
            ctx.set_launch(names.index(v), 5)
            ctx.init_uniform([0, 0, 0], [90, 1, 1])
            ctx.record_draws(True)
            rows = ctx.run(10)
            out[v] = (np.concatenate([ctx.download(k).ravel() for k in range(len(insts))]),
                      np.concatenate([ctx.draws(k).ravel() for k in range(len(insts))]), rows)
    (f_a, d_a, r_a), (f_b, d_b, r_b) = out.values()
    assert np.array_equal(d_a, d_b)
    assert (d_a >= 0).sum() > 0.9 * d_a.size
    assert np.array_equal(f_a, f_b)
    assert np.array_equal(r_a, r_b)


def test_fp32_one_step_dornic_bias(tm):
    """One Dornic transition of the Rietkerk P in FP32 against FP64 from identical
    FP32-representable fields (no plant growth, c = 0, so the RK4 leaves P at P*),
    paired per cell where the Poisson draws agree: the mean relative difference is
    below 1e-7 per step, a fifth of what 5000 steps within 1e-3 allow, and measured
    with a standard error small enough to resolve that bound."""
    L, n = 512, 8
    rng = np.random.default_rng(5)
    fields = [np.stack([rng.uniform(0, 90, (L, L)), rng.uniform(0, 1, (L, L)),
                        rng.uniform(0, 1, (L, L))]).astype(np.float32).astype(np.float64) for _ in range(n)]
    out = {}
    for prec in (tm.F32, tm.F64):
        with tm.Context() as ctx:
            ctx.configure(tm.MODEL_RIETKERK, prec, tm.RietkerkParams(c=0.0), [tm.Instance(L, 0, 7, k) for k in range(n)])
            for k in range(n):
                ctx.upload(k, fields[k])
            ctx.record_draws(True)
            ctx.run(1)
            out[prec] = (np.stack([ctx.download(k)[0] for k in range(n)]), np.stack([ctx.draws(k) for k in range(n)]))
    (p32, d32), (p64, d64) = out[tm.F32], out[tm.F64]
    same = d32 == d64
    assert same.mean() > 0.999
    rel = (p32 - p64)[same] / p64[same]
    se = rel.std() / np.sqrt(rel.size)
    print(f"one-step FP32 Dornic bias {rel.mean():.3e} (SE {se:.2e})", flush=True)
    assert se < 2e-8
    assert abs(rel.mean()) < 1e-7
