"""The stamp mean's exact shortcut (oceanlib_amd/csrc/exact_sum.hpp), run on
the CPU.

tests/cpu_emul/exact_sum_emul.cpp compiles the kernel's own exactness test
and chunk statistics for the host and must give the plain sequential double
sum (the reference's shape_sum order, src/interaction.cpp:345-356) bit for
bit: on Ricker stamps of every size the interaction field draws (and many
sub-cell offsets), and on adversarial float sequences - wide exponent ranges,
subnormals, signed zeros, cancellation, terms that round against the running
sum - where the shortcut must fall back term by term.
"""
import ctypes as C
import os
import shutil
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "cpu_emul", "exact_sum_emul.cpp")


@pytest.fixture(scope="module")
def emul(tmp_path_factory):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    out = str(tmp_path_factory.mktemp("emul") / "libexactsum.so")
    subprocess.run([hipcc, "-std=c++20", "-O2", "--offload-host-only", "-x", "hip",
                    "-ffp-contract=off", "-fPIC", "-shared", SRC, "-o", out], check=True,
                   capture_output=True)
    lib = C.CDLL(out)
    fp = C.POINTER(C.c_float)
    lib.emul_sum_plain.argtypes = [fp, C.c_int]
    lib.emul_sum_plain.restype = C.c_double
    lib.emul_sum_chunked.argtypes = [fp, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int)]
    lib.emul_sum_chunked.restype = C.c_double
    return lib


def check(lib, t, chunk=256, group=4):
    """group: consecutive chunks whose merged statistics are tried first (the
    kernel's 1024-term groups of four 256-chunks)."""
    t = np.ascontiguousarray(t, np.float32)
    p = t.ctypes.data_as(C.POINTER(C.c_float))
    want = lib.emul_sum_plain(p, t.size)
    sk = C.c_int(0)
    got = lib.emul_sum_chunked(p, t.size, chunk, group, C.byref(sk))
    assert np.float64(got).tobytes() == np.float64(want).tobytes(), (got, want, t.size)
    return sk.value, (t.size + chunk - 1) // chunk


def ricker_box(sigma_cells, fx, fz):
    """The stamp's values over its clipped box in z-major order, as the field
    computes them (reference src/interaction.cpp:71-75, 322-356), with the
    cell size folded into sigma; float32 arithmetic, numpy's exp."""
    rad = int(np.ceil(4.0 * sigma_cells))
    g = np.float32(rad + fx), np.float32(rad + fz)
    x = np.arange(0, 2 * rad + 1, dtype=np.float32)
    px = (x + np.float32(0.5)) - g[0]
    pz = (x + np.float32(0.5)) - g[1]
    r2 = (pz[:, None] * pz[:, None] + px[None, :] * px[None, :]).astype(np.float32)
    inv = np.float32(1.0) / (np.float32(2.0) * np.float32(sigma_cells) * np.float32(sigma_cells))
    u = (r2 * inv).astype(np.float32)
    return ((np.float32(1.0) - u) * np.exp(-u).astype(np.float32)).astype(np.float32).reshape(-1)


def test_stamps_of_every_size(emul):
    rng = np.random.default_rng(3)
    skipped = chunks = 0
    for sigma in np.concatenate([np.linspace(2.5, 80.0, 40), rng.uniform(2.5, 80.0, 60)]):
        for fx, fz in rng.uniform(0.0, 1.0, size=(6, 2)):
            s, c = check(emul, ricker_box(float(sigma), float(fx), float(fz)))
            skipped += s
            chunks += c
    # the shortcut carries nearly every chunk of a real stamp
    assert skipped > 0.95 * chunks, (skipped, chunks)


@pytest.mark.parametrize("chunk,group", [(1, 1), (7, 3), (64, 4), (255, 1), (256, 1), (256, 4)])
def test_adversarial_sequences(emul, chunk, group):
    rng = np.random.default_rng(chunk * 10 + group)
    for trial in range(300):
        n = int(rng.integers(1, 3000))
        kind = trial % 6
        if kind == 0:    # wide exponents, random signs
            t = rng.standard_normal(n) * np.exp2(rng.integers(-140, 30, n))
        elif kind == 1:  # subnormals and zeros of both signs
            t = rng.integers(-(1 << 23), 1 << 23, n).astype(np.float64) * 2.0 ** -149
            t[rng.integers(0, n, n // 4)] = -0.0
        elif kind == 2:  # a large term and small ones that round against it
            t = rng.uniform(-1, 1, n) * 2.0 ** -30
            t[rng.integers(0, n, 3)] = rng.uniform(1e3, 1e6, 3)
        elif kind == 3:  # cancellation around a big running value
            t = np.repeat(rng.uniform(-1e7, 1e7, (n + 1) // 2), 2)[:n]
            t[1::2] *= -1
            t += rng.standard_normal(n) * 1e-3
        elif kind == 4:  # floats of full 24-bit mantissas near 1
            t = 1.0 + rng.integers(0, 1 << 23, n) * 2.0 ** -23
        else:            # the running value grows past 2^53 ulps of the small terms
            t = np.full(n, 3.0 * 2.0 ** 20) + rng.uniform(0, 1, n) * 2.0 ** -20
        with np.errstate(over="ignore"):
            check(emul, t.astype(np.float32), chunk, group)


@pytest.mark.parametrize("group", [1, 4])
def test_non_finite_terms_fall_back(emul, group):
    t = np.ones(1000, np.float32)
    t[517] = np.float32(np.inf)
    check(emul, t, group=group)
    t[300] = np.float32(-np.inf)
    check(emul, t, group=group)  # nan, as the sequential sum gives
    t = np.ones(2600, np.float32)
    t[5] = np.float32(np.nan)
    check(emul, t, group=group)
