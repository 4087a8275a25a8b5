"""The Rietkerk consumer model as one step on the GPU (MODEL_FKE2 / MODEL_FKE3: the
loop body of Rietkerk_FPE/rietkerk_bjork_FKE.cpp:4100-4462) against the oracle's
composition of its pinned operators (oracle.fke_step): bit for bit in FP64 over
several steps, on a mixed-size batch; graph and eager runs alike; the movement
statistics (Rho_Mov, FKE:4034-4060) against the oracle's; and, with no grazers,
the two-species model against the single-species Rietkerk model of the same
library."""
import numpy as np
import pytest

from tests.test_fke_model import model_fields

pytestmark = pytest.mark.gpu


def _oracle_run(oracle, p, SP, f, inst, nsteps):
    rows, ve = [], None
    for k in range(nsteps):
        f, fault, ve = oracle.fke_step(p, SP, f, inst.seed, inst.instance_id, inst.step0 + k)
        assert not fault
        rows.append(np.ravel([[np.count_nonzero(x > 0), x.mean()] for x in f]))
    return f, np.array(rows), ve


@pytest.mark.parametrize("SP", [2, 3])
def test_fke_model_matches_oracle_composition(tm, oracle, SP):
    model = tm.MODEL_FKE2 if SP == 2 else tm.MODEL_FKE3
    p = [tm.FKEParams(), tm.FKEParams(a=1.3 / 24.0, sampler=tm.SAMPLER_COMPAT)]
    insts = [tm.Instance(32, 0, 5, 0), tm.Instance(37, 1, 5, 1, step0=(1 << 32) + 3), tm.Instance(48, 0, 6, 2)]
    f0 = [model_fields(SP, d.L, 100 + d.L) for d in insts]
    nsteps = 4
    print(f"fke model SP={SP} L={[d.L for d in insts]}", flush=True)
    with tm.Context() as ctx:
        ctx.configure(model, tm.F64, p, insts)
        for k, f in enumerate(f0):
            ctx.upload(k, f)
        rows = ctx.run(nsteps)
        got = [ctx.download(k) for k in range(len(insts))]
        mov = ctx.mov_stats()
    for k, d in enumerate(insts):
        want, wrows, ve = _oracle_run(oracle, p[d.params_index], SP, f0[k], d, nsteps)
        assert np.array_equal(got[k], want), (k, np.abs(got[k] - want).max())
        np.testing.assert_allclose(rows[k], wrows, rtol=1e-12, atol=0)
        for s in range(1, SP):
            m = oracle.fke_mov_stats(want[SP + 1 + s], ve[s, ..., 0], ve[s, ..., 1])
            assert np.array_equal(mov[k, 4 * s:4 * s + 4], m), (k, s, mov[k], m)
        assert np.all(mov[k, :4] == 0.0)
        if SP == 2:
            assert np.all(mov[k, 6:8] == 0.0)      # the two-species step draws no velocity
        else:
            assert np.all(mov[k, [6, 7, 10, 11]] != 0.0)


@pytest.mark.parametrize("SP", [2, 3])
def test_fke_model_graph_equals_eager(tm, SP):
    model = tm.MODEL_FKE2 if SP == 2 else tm.MODEL_FKE3
    insts = [tm.Instance(40, 0, 9, k) for k in range(3)]
    out = []
    for gsteps in (0, 3):
        with tm.Context() as ctx:
            ctx.configure(model, tm.F64, tm.FKEParams(), insts)
            ctx.set_launch(0, gsteps)
            for k, d in enumerate(insts):
                ctx.upload(k, model_fields(SP, d.L, 60 + k))
            rows = ctx.run(7)
            out.append((np.concatenate([ctx.download(k).ravel() for k in range(3)]), rows, ctx.mov_stats()))
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])
    assert np.array_equal(out[0][2], out[1][2])


def test_no_grazers_equals_single_species_model(tm):
    """FKE2 with G = 0 and Dt_Analytic = dt runs the single-species Rietkerk step on
    P, W and O (the Rietkerk FP64 kernels, bit-exact with their own oracle)."""
    p = tm.FKEParams(dt_analytic=0.1)
    L = 36
    f = model_fields(2, L, 8)
    f[1] = 0.0
    rp = tm.RietkerkParams(R=p.a, c=p.c, gmax=p.gmax, alpha=p.alpha, rW=p.rW, W0=p.W0, d=p.M[0],
                           D=(p.D_P, p.D_W, p.D_O), K=p.K, sigma=p.sigma[0], dt=p.dt, dx=p.dx)
    inst = [tm.Instance(L, 0, 3, 4)]
    with tm.Context() as ctx:
        ctx.configure(tm.MODEL_FKE2, tm.F64, p, inst)
        ctx.upload(0, f)
        rows = ctx.run(5)
        got = ctx.download(0)
    with tm.Context() as ctx:
        ctx.configure(tm.MODEL_RIETKERK, tm.F64, rp, inst)
        ctx.upload(0, f[[0, 2, 3]])
        rrows = ctx.run(5)
        want = ctx.download(0)
    assert np.array_equal(got[[0, 2, 3]], want)
    assert np.all(got[1] == 0.0) and np.all(got[4] == 1.0)
    np.testing.assert_allclose(rows[0][:, [0, 1, 4, 5, 6, 7]], rrows[0], rtol=1e-12, atol=0)


def test_fke_model_rejects_movement_wider_than_lattice(tm):
    with tm.Context() as ctx:
        with pytest.raises(tm.TMError, match="TM_ERR_ARG"):
            ctx.configure(tm.MODEL_FKE3, tm.F64, tm.FKEParams(), [tm.Instance(16, 0, 1, 0)])   # predator W = 9
        with pytest.raises(tm.TMError, match="TM_ERR_ARG"):
            ctx.configure(tm.MODEL_FKE2, tm.F32, tm.FKEParams(), [tm.Instance(32, 0, 1, 0)])   # FP64 only
