"""The ctypes binding against the C header: a C++ program built from
include/tm/tm.h alone (tests/cpp/abi_layout.cpp, host compiler, no HIP) checks
the struct layouts with static_assert and prints every field's offset and size;
each must equal the binding's ctypes mirror in trophic_maelstorm_amd."""
import ctypes as C
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def layout(tmp_path_factory):
    cxx = shutil.which("g++") or shutil.which("c++")
    if cxx is None:
        pytest.skip("no host C++ compiler")
    exe = str(tmp_path_factory.mktemp("abi") / "abi_layout")
    subprocess.run([cxx, "-std=c++17", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "abi_layout.cpp"), "-o", exe], check=True)
    out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split("\n")
    lay = {}
    for line in out:
        if line.strip():
            s, f, off, size = line.split()
            lay.setdefault(s, {})[f] = (int(off), int(size))
    return lay


MIRRORS = {"tm_dp_params": "_DP", "tm_rietkerk_params": "_RK", "tm_dpc_params": "_DPC",
           "tm_instance_desc": "_Inst", "tm_fke_move_species": "_MoveSpecies", "tm_fke_rk_params": "_FKERK",
           "tm_fke_params": "_FKE"}


@pytest.mark.parametrize("cname", sorted(MIRRORS))
def test_ctypes_mirror_matches_header(layout, cname):
    import trophic_maelstorm_amd as tm
    cls = getattr(tm, MIRRORS[cname])
    c = layout[cname]
    assert C.sizeof(cls) == c["sizeof"][0], cname
    names = [f[0] for f in cls._fields_]
    assert sorted(names) == sorted(k for k in c if k != "sizeof"), cname
    for name in names:
        fd = getattr(cls, name)
        assert (fd.offset, fd.size) == c[name], (cname, name)
