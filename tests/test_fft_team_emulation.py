"""The GPU FFT's butterfly graph, run on the CPU.

tests/cpu_emul/fft_team_emul.cpp compiles the kernels' own fft_team<N>
(oceanlib_amd/csrc/fft_team.hpp) for the host, one thread per lane with a
std::barrier for the block barrier, and it must equal the oracle's Stockham
transform (reference src/core/fft.cpp:36-82) bit for bit at every legal N.
"""
import ctypes as C
import os
import shutil
import subprocess

import numpy as np
import pytest

from oracle_ffi import assert_bitexact, fptr, oracle

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "cpu_emul", "fft_team_emul.cpp")


@pytest.fixture(scope="module")
def emul(tmp_path_factory):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    out = str(tmp_path_factory.mktemp("emul") / "libfftemul.so")
    subprocess.run([hipcc, "-std=c++20", "-O2", "--offload-host-only", "-x", "hip",
                    "-ffp-contract=off", "-fPIC", "-shared", SRC, "-o", out], check=True,
                   capture_output=True)
    lib = C.CDLL(out)
    lib.emul_fft_team.argtypes = [C.c_uint32] + [C.POINTER(C.c_float)] * 4
    lib.emul_fft_team.restype = C.c_int
    lib.emul_fft_team_k.argtypes = [C.c_uint32, C.c_uint32] + [C.POINTER(C.c_float)] * 4
    lib.emul_fft_team_k.restype = C.c_int
    lib.emul_fft_team_r.argtypes = [C.c_uint32, C.c_uint32] + [C.POINTER(C.c_float)] * 4
    lib.emul_fft_team_r.restype = C.c_int
    return lib


@pytest.mark.parametrize("n", [16, 32, 64, 128, 256, 512, 1024, 2048, 4096])
def test_team_fft_equals_oracle(emul, n):
    twr = np.zeros(n // 2, np.float32)
    twi = np.zeros(n // 2, np.float32)
    oracle().orc_twiddles(n, fptr(twr), fptr(twi))
    rng = np.random.default_rng(n)
    for trial in range(3):
        re = rng.standard_normal(n).astype(np.float32)
        im = rng.standard_normal(n).astype(np.float32)
        if trial == 2:  # signed zeros and a lone impulse
            re[:] = 0.0
            im[:] = -0.0
            re[3] = -0.0
            im[n // 2 + 1] = 1.0
        a, b = re.copy(), im.copy()
        oracle().orc_fft_inverse(n, fptr(a), fptr(b))
        c, d = re.copy(), im.copy()
        assert emul.emul_fft_team(n, fptr(c), fptr(d), fptr(twr), fptr(twi)) == 0
        assert_bitexact(c, a, f"re n={n}")
        assert_bitexact(d, b, f"im n={n}")


@pytest.mark.parametrize("n", [16, 64, 512, 4096])
@pytest.mark.parametrize("k", [2, 3])
def test_team_fft_of_several_sequences_equals_oracle(emul, n, k):
    """fft_team_k: K sequences through the same exchanges, each as if alone."""
    twr = np.zeros(n // 2, np.float32)
    twi = np.zeros(n // 2, np.float32)
    oracle().orc_twiddles(n, fptr(twr), fptr(twi))
    rng = np.random.default_rng(100 * k + n)
    re = rng.standard_normal(k * n).astype(np.float32)
    im = rng.standard_normal(k * n).astype(np.float32)
    re[:n] = 0.0  # one sequence of signed zeros and an impulse
    im[:n] = -0.0
    im[n // 2 + 1] = 1.0
    want_re, want_im = re.copy(), im.copy()
    for j in range(k):
        a, b = want_re[j * n:(j + 1) * n].copy(), want_im[j * n:(j + 1) * n].copy()
        oracle().orc_fft_inverse(n, fptr(a), fptr(b))
        want_re[j * n:(j + 1) * n], want_im[j * n:(j + 1) * n] = a, b
    assert emul.emul_fft_team_k(n, k, fptr(re), fptr(im), fptr(twr), fptr(twi)) == 0
    assert_bitexact(re, want_re, f"re n={n} k={k}")
    assert_bitexact(im, want_im, f"im n={n} k={k}")


@pytest.mark.parametrize("n", [32, 64, 128, 256, 512, 1024, 2048])
@pytest.mark.parametrize("r", [4, 16])
def test_team_fft_other_radix_groupings_equal_oracle(emul, n, r):
    """fft_team with 4 or 16 points per lane (stages grouped in pairs or fours
    between the exchanges): the same butterflies, the oracle's bits."""
    twr = np.zeros(n // 2, np.float32)
    twi = np.zeros(n // 2, np.float32)
    oracle().orc_twiddles(n, fptr(twr), fptr(twi))
    rng = np.random.default_rng(7 * n + r)
    for trial in range(2):
        re = rng.standard_normal(n).astype(np.float32)
        im = rng.standard_normal(n).astype(np.float32)
        if trial == 1:
            re[:] = 0.0
            im[:] = -0.0
            im[n // 2 + 1] = 1.0
        a, b = re.copy(), im.copy()
        oracle().orc_fft_inverse(n, fptr(a), fptr(b))
        c, d = re.copy(), im.copy()
        assert emul.emul_fft_team_r(n, r, fptr(c), fptr(d), fptr(twr), fptr(twi)) == 0
        assert_bitexact(c, a, f"re n={n} r={r}")
        assert_bitexact(d, b, f"im n={n} r={r}")
