"""The renderer handoff's oracle (oracle/pack_oracle.c), pinned on the CPU.

float_to_half restates the viewer's conversion (reference
examples/viewer/half.hpp:58-100): round to nearest even with half subnormals,
overflow to infinity, and every NaN to the quiet NaN sign | 0x7E00. Pinned
here on all 2^32 float32 bit patterns: against numpy's float16 cast (the
IEEE conversion, same rounding) for every non-NaN pattern and the closed form
for NaNs, and against the reference's own header compiled into oracle/_ref
(oracle/ref_shim.cpp). The pack (half.hpp:108-125) is checked against the
reference's on random buffers with and without the alpha substitution.
"""
import ctypes as C
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from oracle_ffi import oracle, ref

CHUNK = 1 << 26


def _bind(lib, prefix):
    f = getattr(lib, prefix + "half_range")
    f.argtypes = [C.c_uint32, C.c_uint32, C.c_void_p]
    f.restype = None
    p = getattr(lib, prefix + "pack_rgba16")
    p.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]
    p.restype = None
    return f, p


def _halves(fn, first, count):
    out = np.empty(count, np.uint16)
    fn(first, count, out.ctypes.data)
    return out


def _numpy_chunk(first):
    bits = np.arange(CHUNK, dtype=np.uint32) + np.uint32(first)
    f = bits.view(np.float32)
    with np.errstate(over="ignore", invalid="ignore"):
        want = f.astype(np.float16).view(np.uint16)
    nan = np.isnan(f)
    want = np.where(nan, ((bits >> 16) & 0x8000).astype(np.uint16) | np.uint16(0x7E00), want)
    return want


def test_float_to_half_every_pattern_against_numpy():
    orc_range, _ = _bind(oracle(), "orc_")

    def check(k):
        first = k * CHUNK
        got = _halves(orc_range, first, CHUNK)
        want = _numpy_chunk(first)
        bad = np.flatnonzero(got != want)
        return first, bad[:4], len(bad)

    with ThreadPoolExecutor(8) as ex:
        for first, bad, nbad in ex.map(check, range((1 << 32) // CHUNK)):
            assert nbad == 0, f"{nbad} mismatches from 0x{first:08x}, e.g. {[hex(first + b) for b in bad]}"


def test_float_to_half_every_pattern_against_reference():
    r = ref()
    if r is None:
        pytest.skip("oracle/_ref not built (make -C oracle ref)")
    orc_range, _ = _bind(oracle(), "orc_")
    ref_range, _ = _bind(r, "oref_")

    def check(k):
        first = k * CHUNK
        a = _halves(orc_range, first, CHUNK)
        b = _halves(ref_range, first, CHUNK)
        return first, int(np.count_nonzero(a != b))

    with ThreadPoolExecutor(8) as ex:
        for first, nbad in ex.map(check, range((1 << 32) // CHUNK)):
            assert nbad == 0, f"{nbad} mismatches from 0x{first:08x}"


def test_landmarks():
    orc_range, _ = _bind(oracle(), "orc_")
    cases = {
        0x00000000: 0x0000, 0x80000000: 0x8000,          # signed zeros
        0x3F800000: 0x3C00, 0xBF800000: 0xBC00,          # +-1
        0x477FE000: 0x7BFF,                              # 65504, half's largest
        0x477FF000: 0x7C00,                              # 65520 rounds to infinity
        0x477FEFFF: 0x7BFF,                              # just below the tie
        0x33800000: 0x0001,                              # 2^-24, smallest subnormal
        0x33000000: 0x0000,                              # 2^-25: a tie, to even (0)
        0x33000001: 0x0001,                              # just above the tie
        0x387FC000: 0x03FF,                              # largest subnormal
        0x7F800000: 0x7C00, 0xFF800000: 0xFC00,          # infinities
        0x7FC00000: 0x7E00, 0xFF800001: 0xFE00,          # NaNs
    }
    for bits, want in cases.items():
        assert _halves(orc_range, bits, 1)[0] == want, hex(bits)


def test_pack_matches_reference():
    r = ref()
    if r is None:
        pytest.skip("oracle/_ref not built (make -C oracle ref)")
    _, orc_pack = _bind(oracle(), "orc_")
    _, ref_pack = _bind(r, "oref_")
    rng = np.random.default_rng(9)
    cells = 4099
    src = rng.normal(0, 30, 4 * cells).astype(np.float32)
    src[::97] = np.float32(np.inf)
    src[5::101] = np.float32(1e-6)
    alpha = rng.uniform(0, 1, cells).astype(np.float32)
    for a in (alpha, None):
        got = np.zeros(4 * cells, np.uint16)
        want = np.ones(4 * cells, np.uint16)
        ap = None if a is None else a.ctypes.data
        orc_pack(got.ctypes.data, src.ctypes.data, ap, cells)
        ref_pack(want.ctypes.data, src.ctypes.data, ap, cells)
        np.testing.assert_array_equal(got, want)
