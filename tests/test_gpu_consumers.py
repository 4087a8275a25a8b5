"""The consumer-species step (vegetation + grazer, + predator) on the GPU against
the CPU oracle: FP64 fields bit for bit, rho(t) rows to compensated-sum rounding."""
import numpy as np
import pytest

from tests.test_consumers import params

pytestmark = pytest.mark.gpu


def _tm_params(tm, p):
    return tm.DPCParams(a=p.a, b=p.b, D=p.D, sigma=p.sigma, v=p.v, M=p.M, E=p.E, A=p.A, H=p.H, pR=p.pR,
                        dt=p.dt, dx=p.dx, a_end=p.a_end, ramp_steps=p.ramp_steps, flush_eps=p.flush_eps,
                        flush_neg=p.flush_neg, sampler=p.sampler)


def _assert_rows_close(got, want):
    assert got.shape == want.shape
    scale = np.maximum(np.abs(want), 1e-300)
    assert np.all((got == want) | (np.abs(got - want) <= 1e-12 * scale)), np.max(np.abs(got - want) / scale)


def _initial(rng, ns, L, sparse=1.0):
    f = np.zeros((2 * ns - 1, L, L))
    f[0] = rng.uniform(0, 2.0, (L, L)) * (rng.random((L, L)) < sparse)
    f[1] = rng.uniform(0, 0.8, (L, L))
    if ns == 3:
        f[2] = rng.uniform(0, 0.4, (L, L))
    f[ns:] = rng.uniform(0, 0.3, (ns - 1, L, L))
    return f


def _run_both(tm, oracle, model, plist, insts, init, steps, variant=None, graph=None):
    ns = tm.NSPECIES[model]
    with tm.Context() as ctx:
        ctx.configure(model, tm.F64, [_tm_params(tm, p) for p in plist], insts)
        if variant is not None or graph is not None:
            names = tm.variant_names(model, tm.F64)
            ctx.set_launch(names.index(variant) if variant else -1, 100 if graph is None else graph)
        for k, f in enumerate(init):
            ctx.upload(k, f)
        rows = ctx.run(steps)
        got = [ctx.download(k) for k in range(len(insts))]
    for k, d in enumerate(insts):
        ref = oracle.dpc_run(oracle.dpc_params(plist[d.params_index]), ns, init[k], d.L, d.seed, d.instance_id,
                             d.step0, steps)
        assert ref["status"] == 0
        assert np.array_equal(got[k], ref["fields"]), (model, d.L, np.argwhere(got[k] != ref["fields"])[:5])
        _assert_rows_close(rows[k], ref["rows"])
    return got, rows


@pytest.mark.parametrize("model", ["DP2", "DP3"])
def test_consumers_bit_exact_mixed_batch(tm, oracle, model):
    m = getattr(tm, "MODEL_" + model)
    ns = tm.NSPECIES[m]
    plist = [params(), params(pR=(0.0, 0.52, 0.31), v=(0.0, 0.6, 0.2), a=0.3)]   # R = 2 and R = 5
    Ls = [16, 37, 24, 61]
    insts = [tm.Instance(L, k % 2, 7, 40 + k, 3 * k) for k, L in enumerate(Ls)]
    rng = np.random.default_rng(10)
    init = [_initial(rng, ns, L, sparse=0.5 if k == 2 else 1.0) for k, L in enumerate(Ls)]
    print(f"consumers {model} mixed batch", flush=True)
    _, rows = _run_both(tm, oracle, m, plist, insts, init, 25)
    assert np.all(rows[:, -1, 2] > 0)                  # grazers alive to the end


@pytest.mark.parametrize("model", ["DP2", "DP3"])
def test_consumers_every_variant_eager_and_graph(tm, oracle, model):
    m = getattr(tm, "MODEL_" + model)
    ns = tm.NSPECIES[m]
    Ls = [29, 40]
    insts = [tm.Instance(L, 0, 3, k) for k, L in enumerate(Ls)]
    rng = np.random.default_rng(11)
    init = [_initial(rng, ns, L) for L in Ls]
    for k, name in enumerate(tm.variant_names(m, tm.F64)):
        print(f"consumers {model} {name}", flush=True)
        _run_both(tm, oracle, m, [params()], insts, init, 9, variant=name, graph=0 if k % 2 else 4)


@pytest.mark.parametrize("model", ["DP2", "DP3"])
def test_consumers_through_extinction(tm, oracle, model):
    """High consumer mortality and a vegetation ramp down: the consumers die out
    (their Dornic step is then skipped and gamma set to 0) and the vegetation
    thins below the 35% cover floor of the perception factor."""
    m = getattr(tm, "MODEL_" + model)
    ns = tm.NSPECIES[m]
    p = params(M=(0.0, 3.0, 4.0), a=0.45, a_end=-2.0, ramp_steps=40)
    insts = [tm.Instance(32, 0, 9, 1), tm.Instance(21, 0, 9, 2)]
    rng = np.random.default_rng(12)
    init = [_initial(rng, ns, d.L) for d in insts]
    print(f"consumers {model} extinction", flush=True)
    _, rows = _run_both(tm, oracle, m, [p], insts, init, 60, graph=16)
    assert np.all(rows[:, -1, 2] == 0.0)               # grazers gone
    assert np.all(rows[:, -1, 0] < 0.35 * np.array([32 * 32, 21 * 21]))


def test_consumer_fault_is_reported(tm, oracle):
    L = 24
    p = params(b=2.0)
    rng = np.random.default_rng(13)
    f = _initial(rng, 2, L)
    f[0, 5, 7] = 100.0            # the vegetation's RK4 of -b rho^2 overshoots below zero
    ref = oracle.dpc_run(oracle.dpc_params(p), 2, f, L, 1, 0, 0, 1)
    assert ref["status"] == 1 and ref["fault_index"] == 5 * L + 7
    print("consumers fault", flush=True)
    with tm.Context() as ctx:
        ctx.configure(tm.MODEL_DP2, tm.F64, _tm_params(tm, p), [tm.Instance(L, 0, 1, 0)])
        ctx.upload(0, f)
        with pytest.raises(tm.NumericFault):
            ctx.run(1)
        inst, cell, step = ctx.fault_info()
        assert (inst, cell, step) == (0, ref["fault_index"], 0)


def test_consumer_models_reject_bad_configurations(tm):
    with tm.Context() as ctx:
        with pytest.raises(tm.TMError):
            ctx.configure(tm.MODEL_DP2, tm.F32, _tm_params(tm, params()), [tm.Instance(16)])
        with pytest.raises(tm.TMError):          # R = (int)(0.3 / 0.1) = 2 is not below L = 2
            ctx.configure(tm.MODEL_DP2, tm.F64, _tm_params(tm, params()), [tm.Instance(2)])
        with pytest.raises(tm.TMError):          # R = 13 sites
            ctx.configure(tm.MODEL_DP3, tm.F64, _tm_params(tm, params(pR=(0.0, 1.35, 0.2))), [tm.Instance(64)])
        with pytest.raises(tm.TMError):          # no perception at all
            ctx.configure(tm.MODEL_DP2, tm.F64, _tm_params(tm, params(pR=(0.0, 0.0, 0.0))), [tm.Instance(64)])


@pytest.mark.parametrize("model", ["DP2", "DP3"])
def test_consumer_species_statistics(tm, oracle, model):
    """Per-species replicate statistics on the device equal the oracle's on the
    same device series, bit for bit, for every flag; the second parameter point's
    consumers die out within the run."""
    m = getattr(tm, "MODEL_" + model)
    ns = tm.NSPECIES[m]
    plist = [params(), params(M=(0.0, 3.0, 4.0))]
    insts = [tm.Instance(20 + 3 * (k % 3), k % 2, 5, k) for k in range(8)]
    pidx = np.array([d.params_index for d in insts])
    rng = np.random.default_rng(14)
    print(f"consumers {model} statistics", flush=True)
    with tm.Context() as ctx:
        ctx.configure(m, tm.F64, [_tm_params(tm, p) for p in plist], insts)
        for k, d in enumerate(insts):
            ctx.upload(k, _initial(rng, ns, d.L))
        series = ctx.run_series([1, 4, 9, 20, 45, 70])
        assert np.any(series[pidx == 1, -1, 2] == 0.0)
        for flags in (0, 1, 2, 3):
            got = ctx.replicate_stats(flags)
            assert got.shape == (2, 6, 4 * ns)
            assert np.array_equal(got, oracle.species_stats(series, pidx, 2, ns, flags)), flags
