"""The output layer on the CPU: CSV text byte-equal to what the reference's
std::ostream writers print, the snapshot converter, and the burn-in reader."""
import math
import os
import struct

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


SPECIALS = [0.0, -0.0, 1.0, 0.1, 1 / 3, 2.0 / 3, 1e-5, 1.5e-5, 123456.5, 1234567.0, 1e300, 1e-300, 5e-324,
            2.5, 0.000123456789, 99999.95, 999999.5, 1e16, -7.25, 89.99999999]


def expected_frame_text(oracle, header, a, fields, prec=0):
    nf, N = fields.shape[0], fields[0].size
    flat = fields.reshape(nf, N)
    lines = [header] if header else []
    astr = oracle.ostream_text(a, prec)
    for i in range(N):
        lines.append(",".join([astr, str(i)] + [oracle.ostream_text(flat[f, i], prec) for f in range(nf)]) + "\n")
    return "".join(lines)


def test_frame_text_equals_ostream(host_tm, oracle, tmp_path):
    rng = np.random.default_rng(3)
    L = 12
    f = np.empty((3, L * L))
    f[0] = rng.random(L * L) * 90.0
    f[1] = np.exp(rng.normal(0, 8, L * L))
    f[2, : len(SPECIALS)] = SPECIALS
    f[2, len(SPECIALS):] = rng.random(L * L - len(SPECIALS))
    f = f.reshape(3, L, L)
    header = host_tm.FRAME_HEADER[host_tm.MODEL_RIETKERK]
    p = tmp_path / "frame.csv"
    for a in (1.25, 0.000416666, 1.0 / 24.0):
        host_tm.format_frame_csv(p, f, a, header)
        assert p.read_text() == expected_frame_text(oracle, header, a, f)


def test_frame_text_many_threads(host_tm, tmp_path):
    """A frame large enough to be split over several formatting threads is the
    same text as one thread's (printf %.6g, which is the ostream default)."""
    L = 700
    rng = np.random.default_rng(4)
    f = (rng.random((1, L, L)) ** 3) * 50.0
    p = tmp_path / "big.csv"
    host_tm.format_frame_csv(p, f, 0.7, None)
    flat = f.reshape(-1)
    want = "".join(f"0.7,{i},{flat[i]:.6g}\n" for i in range(flat.size))
    assert p.read_text() == want


def test_frame_round_trip_at_17_digits(host_tm, tmp_path):
    rng = np.random.default_rng(5)
    L = 40
    f = rng.random((3, L, L)) * np.array([90.0, 1.0, 1.0])[:, None, None]
    p = tmp_path / "rt.csv"
    host_tm.format_frame_csv(p, f, 1.0, host_tm.FRAME_HEADER[host_tm.MODEL_RIETKERK], digits=17)
    g, n = host_tm.read_frame_csv(p, L, 3, ["P(x;t)", "W(x;t)", "O(x;t)"])
    assert n == L * L and np.array_equal(g, f)
    host_tm.format_frame_csv(p, f, 1.0, host_tm.FRAME_HEADER[host_tm.MODEL_RIETKERK])
    g, _ = host_tm.read_frame_csv(p, L, 3, ["P(x;t)", "W(x;t)", "O(x;t)"])
    assert np.array_equal(g, np.vectorize(lambda x: float(f"{x:.6g}"))(f))


def test_prelim_text(host_tm, oracle, tmp_path):
    t = np.array([0.0, 1.1, 91.2])
    stats = np.array([[0.2, 0.5, 12.5, 1e-7, 4.0, 1234.5], [0.21, 0.49, 0.0, 0.0, 0.0, 0.0],
                      [1 / 3, 2 / 3, 7.0, 0.25, 2.0, 999999.5]])
    header = host_tm.PRELIM_HEADER[host_tm.MODEL_RIETKERK]
    p = tmp_path / "prelim.csv"
    host_tm.write_prelim_csv(p, header, 1.0625, 7, 128, t, stats)
    want = header
    for i in range(3):
        cells = [oracle.ostream_text(1.0625), "8", "128", oracle.ostream_text(t[i])]
        cells += [oracle.ostream_text(x) for x in stats[i]]
        want += ",".join(cells) + ",\n"        # every statistic is followed by a comma
    assert p.read_text() == want


def test_sweep_text_precision_16(host_tm, oracle, tmp_path):
    rows = np.array([[0.5, 16, 512, 0.0, 0.2, 0.3, 12.345678901234567, 0.1, 16, 300.0],
                     [1.5, 16, 512, 91.2, 1 / 3, 1e-17, 5e-324, -2.5, 0, 1e22]])
    header = " a , r, L, t \n"
    p = tmp_path / "sweep.csv"
    host_tm.write_sweep_csv(p, header, rows)
    want = header + "\n" + "".join(",".join(oracle.ostream_text(x, 16) for x in r) + "\n" for r in rows)
    assert p.read_text() == want


def _write_snapshot(path, fields, precision, model, seed=1, inst=2, step=33, pidx=0):
    nf, L = fields.shape[0], fields.shape[-1]
    data = fields.astype(np.float64 if precision == 0 else np.float32).tobytes()
    head = b"TMSNAP1\0" + struct.pack("<iiiiIIqiiQQ", model, precision, nf, L, seed, inst, step, pidx, 0,
                                      len(data), 0)
    assert len(head) == 64
    with open(path, "wb") as fh:
        fh.write(head + data)


@pytest.mark.parametrize("precision", [0, 1])
def test_snapshot_converter(host_tm, tmp_path, precision):
    rng = np.random.default_rng(6 + precision)
    f = rng.random((3, 20, 20)) * 10
    snap, a, b = tmp_path / "s.tms", tmp_path / "a.csv", tmp_path / "b.csv"
    _write_snapshot(snap, f, precision, 3)
    host_tm.snapshot_to_frame_csv(snap, a, 0.75, "h\n")
    stored = f if precision == 0 else f.astype(np.float32).astype(np.float64)
    host_tm.format_frame_csv(b, stored, 0.75, "h\n")
    assert a.read_bytes() == b.read_bytes()
    # a damaged file is refused
    raw = bytearray(snap.read_bytes())
    raw[0] = ord("X")
    snap.write_bytes(bytes(raw))
    with pytest.raises(OSError):
        host_tm.snapshot_to_frame_csv(snap, a, 0.75, None)


def test_burnin_reader_columns_and_errors(host_tm, tmp_path):
    p = tmp_path / "burn.csv"
    # columns in another order, white space in the names, a bad cell
    p.write_text("a_c , x , O(x; t) , W(x; t) , P(x; t)\n"
                 "1,0,0.5,0.25,3.0\n"
                 "1,1,0.75,oops,4.5\n"
                 "1,2,1.0,0.5,6.0\n"
                 "1,3,1.25,0.75,7.5\n"
                 "1,4,9,9,9\n")
    base = np.full((3, 2, 2), -1.0)
    g, n = host_tm.read_frame_csv(p, 2, 3, ["W(x;t)", "O(x;t)"], out=base)
    assert n == 4                                             # at most L*L rows
    assert g[1].ravel().tolist() == [0.25, 0.0, 0.5, 0.75]     # W -> field nf-2, "oops" -> 0
    assert g[2].ravel().tolist() == [0.5, 0.75, 1.0, 1.25]     # O -> field nf-1
    assert np.all(g[0] == -1.0)                                # not named: untouched
    # any other column goes to field (its position - 2): P at position 4 -> field 2
    g, _ = host_tm.read_frame_csv(p, 2, 3, ["P(x;t)"], out=base)
    assert g[2].ravel().tolist() == [3.0, 4.5, 6.0, 7.5]
    with pytest.raises(KeyError):
        host_tm.read_frame_csv(p, 2, 3, ["G(x;t)"])
    with pytest.raises(OSError):
        host_tm.read_frame_csv(tmp_path / "missing.csv", 2, 3, ["W(x;t)"])


def test_ostream_default_is_printf_g(oracle):
    """The property the fast writer rests on: ostream's default float text is
    printf("%.6g") and setprecision(16) is printf("%.16g")."""
    rng = np.random.default_rng(8)
    xs = list(rng.normal(0, 1, 200) * 10.0 ** rng.integers(-30, 30, 200)) + SPECIALS + [math.inf, -math.inf]
    for x in xs:
        assert oracle.ostream_text(x) == "%.6g" % x
        assert oracle.ostream_text(x, 16) == "%.16g" % x


# ---------------------------------------------------------------- file names
def test_reference_file_names(oracle):
    from trophic_maelstorm_amd import names as nm
    f = nm.folders("PREFIX")
    assert f["frame_folder"] == "../Data/Rietkerk/Frames/Stochastic/1Sp/PREFIX_"
    assert nm.folders("X", 2, 1.5, 0.0)["prelim_folder"] == "../Data/Rietkerk/Prelims/Stochastic/2Sp_mG_1.5/X_"
    assert nm.folders("X", 3, 2.0, 0.125)["frame_folder"].startswith("../Data/Rietkerk/Frames/Stochastic/3Sp_mG_2_mP_0.13/")
    # FKE:2909-2912 with g = 128, t = 91.2, dt = 0.1, a = 1.0625 (precision 4 -> 1.062),
    # D1 = 0.00025/24 (precision 3), jID = 0, dx = 0.1, replicate 3
    assert nm.frame_path(f["frame_folder"], 0.5, 1.5, 1, 128, 91.2, 0.1, 1.0625, 0.00025 / 24.0, 0, 0.1, 3) == (
        "../Data/Rietkerk/Frames/Stochastic/1Sp/PREFIX_0.5-1.5_dP_1"
        "/FRAME_P_c_DP_G_128_T_91.2_dt_0.1_a_1.062_D1_1.04e-05_jID_0_dx_0.1_R_3.csv")
    assert nm.timeseries_path(f["prelim_folder"], 0.5, 1.5, 1, 64, 1000, 0.005, 1.23456789, 7, 0, Gstar=0.2) == (
        "../Data/Rietkerk/Prelims/Stochastic/1Sp/PREFIX_0.5-1.5_dP_1_Geq_0.2/TimeSeries"
        "/PRELIM_TSERIES_P_c_DP_G_64_T_1000_dt_0.005_a_1.2346_jID_7_R_0.csv")
    assert nm.aggregate_prelim_path(f["prelim_folder"], 0.5, 1.5, 1, 0.123, 64, 91201, 0.1, 1.0, 0.025 / 24, 5,
                                    Vstar=3.5) == (
        "../Data/Rietkerk/Prelims/Stochastic/1Sp/PREFIX_0.5-1.5_dP_1_Veq_3.5/TimeSeries"
        "/PRELIM_AGGRAND_P_c_ID_0.123_DP_G_64_T_91201_dt_0.1_a_1_D2_0.001042_R_5.csv")
    assert nm.sweep_path(f["stat_prefix"], 256, 91201, 0.1, 0.00025 / 24, 0.5, 1.5, 0.1, 16) == (
        "../Data/Rietkerk/Stochastic/1stOrderCC_Rietkerk_PREFIX_P_c_G_256_T_91201_dt_0.1_D1_1.04e-05"
        "_a1_0.5_a2_1.5_dx_0.1_R_16.csv")
    for x, p in [(1.0625, 4), (91201.0, 6), (0.00025 / 24, 3), (1.23456789, 5), (1e-7, 6)]:
        assert nm.ostream(x, p) == oracle.ostream_text(x, p)
