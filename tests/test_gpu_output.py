"""Output layer with device fields: FRAME text straight from a context, and
snapshots that save, convert and resume a run exactly."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["F64", "F32"])
def test_frame_csv_and_snapshot_round_trip(tm, tmp_path, precision):
    prec = getattr(tm, precision)
    insts = [tm.Instance(48, 0, 2, 7), tm.Instance(33, 0, 2, 8)]
    header = tm.FRAME_HEADER[tm.MODEL_RIETKERK]
    print(f"frame csv + snapshot, {precision}", flush=True)
    with tm.Context() as ctx:
        ctx.configure(tm.MODEL_RIETKERK, prec, tm.RietkerkParams(R=1.1), insts)
        ctx.init_gauss([5.0, 1.0, 1.0], 1)
        ctx.run(4)
        f1 = ctx.download(1)
        ctx.write_frame_csv(1, tmp_path / "dev.csv", 1.1, header)
        ctx.save_snapshot(1, tmp_path / "s.tms")
    tm.format_frame_csv(tmp_path / "host.csv", f1, 1.1, header)
    assert (tmp_path / "dev.csv").read_bytes() == (tmp_path / "host.csv").read_bytes()
    tm.snapshot_to_frame_csv(tmp_path / "s.tms", tmp_path / "conv.csv", 1.1, header)
    assert (tmp_path / "conv.csv").read_bytes() == (tmp_path / "host.csv").read_bytes()
    with tm.Context() as ctx:
        ctx.configure(tm.MODEL_RIETKERK, prec, tm.RietkerkParams(R=1.1), insts)
        assert ctx.load_snapshot(1, tmp_path / "s.tms") == 4
        assert np.array_equal(ctx.download(1), f1)
        with pytest.raises(tm.TMError):
            ctx.load_snapshot(0, tmp_path / "s.tms")          # L = 48 != 33


def test_snapshot_resume_is_exact(tm, tmp_path):
    """Run 6 steps; or run 3, snapshot, load into a fresh context whose instance
    starts at the recorded step, run 3 more: the same fields (FP64)."""
    p = tm.DPParams(a=1.0)
    print("snapshot resume, DP F64", flush=True)
    with tm.Context() as ctx:
        ctx.configure(tm.MODEL_DP, tm.F64, p, [tm.Instance(64, 0, 12, 5)])
        ctx.init_uniform([0.0], [1.0])
        ctx.run(3)
        ctx.save_snapshot(0, tmp_path / "mid.tms")
        ctx.run(3)
        straight = ctx.download(0)
    with tm.Context() as ctx:
        ctx.configure(tm.MODEL_DP, tm.F64, p, [tm.Instance(64, 0, 12, 5, 3)])   # step0 = 3
        assert ctx.load_snapshot(0, tmp_path / "mid.tms") == 3
        ctx.run(3)
        resumed = ctx.download(0)
    assert np.array_equal(straight, resumed)
