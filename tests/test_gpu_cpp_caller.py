"""The C ABI from C++: tests/cpp/rietkerk_loop.cpp, built with the host compiler
against include/tm/tm.h alone and linked to libtm.so, runs the Rietkerk loop body
of INTEGRATION.md (tm_create, tm_configure, tm_upload_fields, tm_log10_time_bins,
tm_measure_steps, tm_run between the reference's measurement times, tm_measure,
tm_fault_info, tm_download_fields) on a 3-replicate batch.  Its measured rows and
final fields must equal, bit for bit, those of the same run through the ctypes
binding.  Reference: rietkerk_DorFPE_2D_MultiSp (Rietkerk_FPE/rietkerk_bjork_basic.h:
450-452, loop body rietkerk_bjork_FKE.cpp:3887-4468)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_cpp_caller_matches_ctypes(tm, oracle, tmp_path):
    cxx = shutil.which("g++") or shutil.which("c++")
    assert cxx, "a host C++ compiler is needed"
    pkg = os.path.join(ROOT, "trophic_maelstorm_amd")
    exe = str(tmp_path / "rietkerk_loop")
    subprocess.run([cxx, "-std=c++17", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "rietkerk_loop.cpp"), "-o", exe, "-L", pkg, "-ltm",
                    f"-Wl,-rpath,{pkg}"], check=True)
    g, r, seed, t_max, dt = 48, 3, 41, 2.0, 0.005
    frames = np.stack([oracle.init_uniform(g, [0.0, 0.0, 0.0], [90.0, 1.0, 1.0], seed, j) for j in range(r)])
    fin = tmp_path / "fields.bin"
    frames.astype(np.float64).tofile(fin)
    fout = tmp_path / "out.bin"
    print(f"cpp caller: g={g} r={r} seed={seed} t_max={t_max}", flush=True)
    subprocess.run([exe, str(fin), str(fout), str(g), str(r), str(seed), str(t_max)], check=True, timeout=120)
    buf = np.fromfile(fout, np.float64)
    n_meas = int(buf[0])
    steps_cpp = buf[1:1 + n_meas].astype(np.int64)
    o = 1 + n_meas
    rho_cpp = buf[o:o + r * n_meas * 6].reshape(r, n_meas, 6)
    o += r * n_meas * 6
    fault_cpp = tuple(int(x) for x in buf[o:o + 3])
    o += 3
    fields_cpp = buf[o:].reshape(r, 3, g, g)

    # the same run through ctypes
    steps = tm.measure_steps(tm.log10_time_bins(t_max, dt), dt)
    assert np.array_equal(steps, steps_cpp) and n_meas > 3
    rho = np.zeros((r, n_meas, 6))
    with tm.Context() as ctx:
        ctx.configure(tm.MODEL_RIETKERK, tm.F64, tm.RietkerkParams(R=1.1), [tm.Instance(g, 0, seed, j)
                                                                            for j in range(r)])
        for j in range(r):
            ctx.upload(j, frames[j])
        done = 0
        for m, s in enumerate(steps):
            rows = ctx.run(int(s - done)) if s > done else None
            rho[:, m] = rows[:, -1] if rows is not None else ctx.measure()
            done = int(s)
        fault = ctx.fault_info()
        fields = np.stack([ctx.download(j) for j in range(r)])
    assert fault == fault_cpp == (-1, -1, -1)
    assert np.array_equal(rho, rho_cpp)
    assert np.array_equal(fields, fields_cpp)
    assert np.all(rho[:, -1, 0] > 0)          # the vegetation is alive: the run did work


def test_cpp_fke_caller_matches_ctypes(tm, tmp_path):
    """tests/cpp/fke_loop.cpp: the Rietkerk model with grazer and predator
    (TM_MODEL_FKE3) from C++, with the reference driver's constants spelled out as
    INTEGRATION.md maps them, stepped and read back with tm_mov_stats; bit for bit the
    run of tm.FKEParams() (the same constants) through the ctypes binding."""
    from tests.test_fke_model import model_fields
    cxx = shutil.which("g++") or shutil.which("c++")
    assert cxx, "a host C++ compiler is needed"
    pkg = os.path.join(ROOT, "trophic_maelstorm_amd")
    exe = str(tmp_path / "fke_loop")
    subprocess.run([cxx, "-std=c++17", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "fke_loop.cpp"), "-o", exe, "-L", pkg, "-ltm",
                    f"-Wl,-rpath,{pkg}"], check=True)
    g, r, seed, n = 40, 2, 17, 6
    frames = np.stack([model_fields(3, g, 70 + j) for j in range(r)])
    fin = tmp_path / "fields.bin"
    frames.astype(np.float64).tofile(fin)
    fout = tmp_path / "out.bin"
    print(f"cpp fke caller: g={g} r={r} seed={seed} n={n}", flush=True)
    subprocess.run([exe, str(fin), str(fout), str(g), str(r), str(seed), str(n)], check=True, timeout=120)
    buf = np.fromfile(fout, np.float64)
    o = r * n * 14
    rows_cpp = buf[:o].reshape(r, n, 14)
    mov_cpp = buf[o:o + r * 12].reshape(r, 12)
    o += r * 12
    fault_cpp = tuple(int(x) for x in buf[o:o + 3])
    fields_cpp = buf[o + 3:].reshape(r, 7, g, g)
    with tm.Context() as ctx:
        ctx.configure(tm.MODEL_FKE3, tm.F64, tm.FKEParams(), [tm.Instance(g, 0, seed, j) for j in range(r)])
        for j in range(r):
            ctx.upload(j, frames[j])
        rows = ctx.run(n)
        mov = ctx.mov_stats()
        fault = ctx.fault_info()
        fields = np.stack([ctx.download(j) for j in range(r)])
    assert fault == fault_cpp == (-1, -1, -1)
    assert np.array_equal(rows, rows_cpp)
    assert np.array_equal(mov, mov_cpp)
    assert np.array_equal(fields, fields_cpp)
    assert np.any(fields[:, 1] != frames[:, 1])         # the grazers moved
