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
