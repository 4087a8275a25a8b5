"""Pins of the oracle's primitives against published vectors and library routines.

Philox4x32-10: Random123 known-answer vectors (Salmon et al. 2011, kat_vectors)
and rocRAND's host implementation; exp/log: numpy; the normal quantile:
scipy.special.ndtri; log k!: scipy.special.gammaln.
"""
import math
import os
import shutil
import subprocess

import numpy as np
import pytest
from scipy.special import gammaln, ndtri

KAT = [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
]


@pytest.mark.parametrize("ctr,key,want", KAT)
def test_philox_known_answers(oracle, ctr, key, want):
    assert oracle.philox(ctr, key) == want


ROCRAND_PROBE = r"""
#include <cstdio>
#include <rocrand/rocrand_philox4x32_10.h>
int main() {
    unsigned long long seeds[3] = {0ull, 0x1234567890abcdefull, 0xffffffffffffffffull};
    unsigned long long subs[3] = {0ull, 5ull, 0x100000007ull};
    for (int a = 0; a < 3; ++a) for (int b = 0; b < 3; ++b) {
        rocrand_device::philox4x32_10_engine e(seeds[a], subs[b], 0);
        for (int k = 0; k < 8; ++k) printf("%u ", e());
        printf("\n");
    }
}
"""


def test_philox_matches_rocrand_host(oracle, tmp_path):
    """rocRAND's engine(seed, subsequence): key = seed words, counter = (n, 0, sub lo, sub hi)."""
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc) or not os.path.exists("/opt/rocm/include/rocrand/rocrand_philox4x32_10.h"):
        pytest.skip("rocRAND headers / hipcc not available")
    src = tmp_path / "p.cpp"
    src.write_text(ROCRAND_PROBE)
    exe = tmp_path / "p"
    r = subprocess.run([hipcc, "-x", "c++", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-O1",
                        str(src), "-o", str(exe)], capture_output=True, text=True)
    if r.returncode != 0:
        pytest.skip("could not build the rocRAND host probe: " + r.stderr[-300:])
    lines = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    seeds = [0, 0x1234567890ABCDEF, 0xFFFFFFFFFFFFFFFF]
    subs = [0, 5, 0x100000007]
    row = 0
    for s in seeds:
        for sub in subs:
            got = [int(x) for x in lines[row].split()]
            row += 1
            key = [s & 0xFFFFFFFF, s >> 32]
            want = []
            for n in range(2):
                want += oracle.philox([n, 0, sub & 0xFFFFFFFF, sub >> 32], key)
            assert got == want, (hex(s), sub)


def test_exp_within_one_ulp(oracle):
    xs = np.concatenate([np.linspace(-745.0, 709.7, 40001), np.linspace(-3, 3, 20001),
                         [0.0, -0.0, 1e-300, -1e-300, 709.78, -708.4]])
    got = np.array([oracle.lib().tmo_exp(float(x)) for x in xs])
    want = np.exp(xs)
    normal = want > 2.3e-308
    ulps = np.abs(got[normal] - want[normal]) / np.spacing(want[normal])
    assert ulps.max() <= 1.0
    # subnormal results: within one subnormal spacing
    assert np.all(np.abs(got[~normal] - want[~normal]) <= 2 * 4.9e-324)
    assert oracle.lib().tmo_exp(710.0) == math.inf
    assert oracle.lib().tmo_exp(-746.0) == 0.0


def test_log_within_two_ulp(oracle):
    ys = np.concatenate([np.logspace(-320, 307, 40001), np.linspace(0.5, 2.0, 20001),
                         [1.0, 2.0, 0.5, 4.9e-324, 1.7976931348623157e308]])
    got = np.array([oracle.lib().tmo_log(float(y)) for y in ys])
    want = np.log(ys)
    nz = want != 0
    ulps = np.abs(got[nz] - want[nz]) / np.spacing(np.abs(want[nz]))
    assert ulps.max() <= 2.0
    assert oracle.lib().tmo_log(1.0) == 0.0
    assert oracle.lib().tmo_log(0.0) == -math.inf
    assert math.isnan(oracle.lib().tmo_log(-1.0))


def test_normal_quantile_matches_ndtri(oracle):
    ps = np.concatenate([np.linspace(1e-10, 1 - 1e-10, 200001), np.logspace(-10, -1, 2001),
                         1 - np.logspace(-10, -1, 2001), [0.02425, 1 - 0.02425, 0.5]])
    got = np.array([oracle.lib().tmo_norm_icdf(float(p)) for p in ps])
    want = ndtri(ps)
    rel = np.abs(got - want) / np.maximum(np.abs(want), 1e-300)
    assert rel[np.abs(want) > 1e-12].max() < 1.2e-9    # Acklam's published bound 1.15e-9
    assert got[-1] == 0.0
    # symmetry x(p) = -x(1-p) in the tails, at the sampler's uniforms (w + 1/2) 2^-32 for
    # which 1 - p is exact
    lo = (np.unique(np.logspace(0, 8, 50).astype(np.int64)) + 0.5) * 2.0 ** -32
    a = np.array([oracle.lib().tmo_norm_icdf(float(p)) for p in lo])
    b = np.array([oracle.lib().tmo_norm_icdf(float(1 - p)) for p in lo])
    assert np.allclose(a, -b, rtol=1e-9)


def test_logfact(oracle):
    ks = np.concatenate([np.arange(0, 2000), [10 ** 5, 10 ** 6, 123456789]])
    got = np.array([oracle.lib().tmo_logfact(int(k)) for k in ks])
    want = gammaln(ks + 1.0)
    assert np.all(np.abs(got - want) <= 4e-16 * np.maximum(want, 1.0) + 1e-15)
    assert oracle.lib().tmo_logfact(0) == 0.0 and oracle.lib().tmo_logfact(1) == 0.0


def test_uniforms_open_interval_and_exact(oracle):
    u = oracle.sample("uniform", 200000, seed=3, inst=9, cell=77, step=123456789012)
    assert u.min() > 0.0 and u.max() < 1.0
    w = u * 2.0 ** 32 - 0.5     # every uniform is (w + 1/2) 2^-32 for an integer w
    assert np.all(w == np.round(w))
    assert abs(u.mean() - 0.5) < 5 * math.sqrt(1 / 12 / len(u))


def test_streams_are_keyed_by_seed_instance_cell_step(oracle):
    base = oracle.sample("uniform", 8, seed=1, inst=2, cell=3, step=4)
    for kw in ({"seed": 2}, {"inst": 3}, {"cell": 4}, {"step": 5}, {"step": 4 + (1 << 32)}):
        args = dict(seed=1, inst=2, cell=3, step=4)
        args.update(kw)
        assert not np.array_equal(oracle.sample("uniform", 8, **args), base)
    assert np.array_equal(oracle.sample("uniform", 8, seed=1, inst=2, cell=3, step=4), base)
