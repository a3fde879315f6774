"""FoamField::coverage's ordered sum (oceanlib_amd/csrc/ordered_sum.hpp), run
on the CPU.

tests/cpu_emul/ordered_sum_emul.cpp runs the coverage kernel's steps (the
same binade rounding, run-end test and scan order) on the host and must give
the reference's sequential double sum (src/foam.cpp:307-313) bit for bit: on
foam fields of the heavy-chop patch as the oracle advects them, and on
adversarial float sequences - wide exponent ranges, leading zeros,
subnormals, terms that tie against the running value, negative and
non-finite terms, lengths that are not whole chunks.
"""
import ctypes as C
import os
import shutil
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "cpu_emul", "ordered_sum_emul.cpp")


@pytest.fixture(scope="module")
def emul(tmp_path_factory):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    out = str(tmp_path_factory.mktemp("emul") / "liborderedsum.so")
    subprocess.run([hipcc, "-std=c++20", "-O2", "--offload-host-only", "-x", "hip",
                    "-ffp-contract=off", "-fPIC", "-shared", SRC, "-o", out], check=True,
                   capture_output=True)
    lib = C.CDLL(out)
    fp = C.POINTER(C.c_float)
    lib.emul_sum_plain.argtypes = [fp, C.c_uint32]
    lib.emul_sum_plain.restype = C.c_double
    lib.emul_sum_ordered.argtypes = [fp, C.c_uint32, C.POINTER(C.c_uint32)]
    lib.emul_sum_ordered.restype = C.c_double
    return lib


def check(lib, v):
    v = np.ascontiguousarray(v, np.float32)
    p = v.ctypes.data_as(C.POINTER(C.c_float))
    want = lib.emul_sum_plain(p, v.size)
    steps = C.c_uint32(0)
    got = lib.emul_sum_ordered(p, v.size, C.byref(steps))
    assert np.float64(got).tobytes() == np.float64(want).tobytes(), (got, want, v.size)
    return steps.value


def test_foam_fields_of_the_heavy_chop_patch(emul):
    """The oracle's persistent foam on a 128^2 heavy-chop patch, every 10th
    substep: few workgroup steps per chunk (runs cross a binade rarely)."""
    import ctypes
    from oceanlib_amd.workloads import desc_from
    from oracle_ffi import OracleOcean, as_array, default_foam_params, oracle
    d = desc_from(dict(size=128, patch_length=120.0, wind_speed=22.0, choppiness=1.3,
                       compute_velocity=1, water_depth=50.0, seed=7))
    o = OracleOcean(d)
    lib = oracle()
    fp = default_foam_params()
    f = lib.orc_foam_create(o.h, ctypes.byref(fp))
    try:
        total_steps = 0
        for k in range(60):
            o.update(k / 60.0)
            lib.orc_foam_update(f, np.float32(1.0 / 60.0))
            if k % 10 == 9:
                field = as_array(lib.orc_foam_data(f), d.size * d.size)
                assert field.max() > 0.0
                total_steps += check(emul, field)
        # 4 chunks per field; a handful of binade crossings each
        assert total_steps < 6 * 4 * 20
    finally:
        lib.orc_foam_destroy(f)


@pytest.mark.parametrize("n", [1, 3, 4095, 4096, 4097, 65536, 262144 + 7])
def test_wide_exponents_and_lengths(emul, n):
    rng = np.random.default_rng(n)
    e = rng.integers(-40, 20, size=n)
    v = (rng.random(n) * np.exp2(e)).astype(np.float32)
    check(emul, v)


def test_leading_zeros_subnormals_and_signed_zeros(emul):
    rng = np.random.default_rng(3)
    v = np.zeros(20000, np.float32)
    v[9000:] = rng.random(11000).astype(np.float32) * 1e-3
    v[9000:9100] = np.float32(1e-44)  # float subnormals
    v[::7] = np.float32(-0.0)
    check(emul, v)
    check(emul, np.zeros(5000, np.float32))
    check(emul, np.full(5000, -0.0, np.float32))


def test_ties_against_the_running_value(emul):
    """s = 1 (binade [1, 2), u = 2^-52); terms 2^-53 are exact halfway cases,
    whose rounding goes to even and so depends on s's last bit; mixed with
    terms 3 * 2^-53 (also ties) and plain ones."""
    v = np.empty(9000, np.float32)
    v[0] = 1.0
    v[1:] = np.float32(2.0 ** -53)
    v[1::3] = np.float32(3 * 2.0 ** -53)
    v[2::5] = np.float32(0.25)
    check(emul, v)


def test_terms_that_end_runs_often(emul):
    """Growth by doubling: every term crosses into a new binade."""
    v = np.exp2(np.arange(-120, 120, dtype=np.float64)).astype(np.float32)
    check(emul, np.tile(v, 20))


def test_negative_and_non_finite_terms_fall_back(emul):
    rng = np.random.default_rng(5)
    v = rng.random(10000).astype(np.float32)
    w = v.copy()
    w[5000] = -3.0
    check(emul, w)
    w = v.copy()
    w[123] = -1e30
    check(emul, w)
    for bad in (np.inf, -np.inf, np.nan):
        w = v.copy()
        w[7000] = bad
        check(emul, w)
    check(emul, -v)
