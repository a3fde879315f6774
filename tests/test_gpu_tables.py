"""The constant tables built on the device equal the host build bit for bit.

The device kernels (oceanlib_amd/csrc/spectrum_kernels.hip) restate the
spectrum, PCG32 rows and Box-Muller draws (reference src/core/spectrum.cpp,
src/core/rng.hpp, src/core/rng.cpp) with the device libm and flag every cell
whose float results a libm difference could change; the host recomputes
those. Compared here with ocean_amd_build_spectrum (the host chain, itself
equal to the oracle and the reference: tests/test_abi.py) over the workload
descriptors and 30 random ones covering depth, swell, wind direction, peak
enhancement, cutoff, fetch and calm seas.
"""
import ctypes as C

import numpy as np
import pytest

from oceanlib_amd import capi
from oceanlib_amd.workloads import (CONFIG1, CONFIG2, CONFIG3_LEVELS, SMALL_CASES, config4_levels,
                                    desc_from)
from oracle_ffi import assert_bitexact, fptr

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    return capi.load()


def random_descs(count=30, seed=21):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        out.append(dict(size=int(rng.choice([16, 32, 64, 128, 256, 512, 1024])),
                        patch_length=float(np.float32(10 ** rng.uniform(0.5, 3.3))),
                        wind_speed=float(np.float32(0.0 if i == 0 else rng.uniform(0.5, 40.0))),
                        fetch=float(np.float32(10 ** rng.uniform(3, 6))),
                        wind_direction=float(np.float32(rng.uniform(-7, 7))),
                        peak_enhancement=float(np.float32(rng.uniform(1.0, 7.0))),
                        swell=float(np.float32(rng.uniform(0, 1) if i % 3 == 0 else 0.0)),
                        small_wave_cutoff=float(np.float32(rng.uniform(0, 5))),
                        water_depth=float(np.float32(0.0 if i % 2 else rng.uniform(0.5, 500))),
                        seed=int(rng.integers(0, 2 ** 63)), choppiness=1.0))
    return out


CASES = ([("config1", CONFIG1), ("config2", CONFIG2)] +
         [(f"config3_{k}", kw) for k, kw in enumerate(CONFIG3_LEVELS)] +
         [(name, kw) for name, kw, _ in SMALL_CASES] +
         [(f"config4_{k}", kw) for k, kw in enumerate(config4_levels(20)[0])] +
         [(f"random_{k}", kw) for k, kw in enumerate(random_descs())])


@pytest.mark.parametrize("name,kw", CASES, ids=[c[0] for c in CASES])
def test_device_tables_equal_host_build(lib, name, kw):
    d = capi.default_desc(lib, **kw)
    n = d.size
    want_h0 = np.zeros(2 * n * n, np.float32)
    want_om = np.zeros(n * n, np.float32)
    assert lib.ocean_amd_build_spectrum(C.byref(d), fptr(want_h0), fptr(want_om)) == 0
    st = C.c_int(-1)
    sim = lib.ocean_create(C.byref(d), C.byref(st))
    assert sim, st.value
    try:
        h0 = np.zeros(2 * n * n, np.float32)
        om = np.zeros(n * n, np.float32)
        rep = C.c_uint32()
        assert lib.ocean_amd_get_tables(sim, fptr(h0), fptr(om), C.byref(rep)) == 0
        print(f"{name}: N={n} repaired {rep.value} of {n * n} cells")
        assert_bitexact(h0, want_h0, f"{name} h0")
        assert_bitexact(om, want_om, f"{name} omega")
        # cells on the axis opposite the wind (theta = pi exactly: the spreading
        # cosine is 6e-17, a branch no margin can certify) and the rare values
        # next to a float rounding midpoint: about one row's worth
        assert rep.value <= n + n * n // 200
    finally:
        lib.ocean_destroy(sim)


def extreme_descs(count=320, seed=77):
    """Descriptors at the parameter extremes, where the certification's error
    bound is stretched most: swell 1, depth down to 0.5 m, fetch 10^3,
    peak enhancement 7, winds 0.5 and 40 m/s, cutoffs and directions across
    their range, in every combination the draws give."""
    rng = np.random.default_rng(seed)
    pick = lambda *v: float(np.float32(v[int(rng.integers(0, len(v)))]))  # noqa: E731
    out = []
    for _ in range(count):
        out.append(dict(size=int(rng.choice([16, 32, 64, 128])),
                        patch_length=pick(3.0, 25.0, 250.0, 2000.0, float(10 ** rng.uniform(0.5, 3.3))),
                        wind_speed=pick(0.5, 40.0, 22.0, float(rng.uniform(0.5, 40.0))),
                        fetch=pick(1e3, 1e6, float(10 ** rng.uniform(3, 6))),
                        wind_direction=pick(0.0, 3.14159265, -1.5707964, float(rng.uniform(-7, 7))),
                        peak_enhancement=pick(1.0, 7.0, float(rng.uniform(1.0, 7.0))),
                        swell=pick(0.0, 1.0, float(rng.uniform(0, 1))),
                        small_wave_cutoff=pick(0.0, 5.0, float(rng.uniform(0, 5))),
                        water_depth=pick(0.0, 0.5, 2.0, float(rng.uniform(0.5, 500))),
                        seed=int(rng.integers(0, 2 ** 63)), choppiness=1.0))
    return out


def test_device_tables_equal_host_build_at_extremes(lib):
    """320 descriptors at the extremes, each table equal to the host build bit
    for bit (the repair count is reported, not bounded: extremes flag more)."""
    repaired = cells = 0
    for k, kw in enumerate(extreme_descs()):
        d = capi.default_desc(lib, **kw)
        n = d.size
        want_h0 = np.zeros(2 * n * n, np.float32)
        want_om = np.zeros(n * n, np.float32)
        assert lib.ocean_amd_build_spectrum(C.byref(d), fptr(want_h0), fptr(want_om)) == 0
        st = C.c_int(-1)
        sim = lib.ocean_create(C.byref(d), C.byref(st))
        assert sim, (k, kw, st.value)
        try:
            h0 = np.zeros(2 * n * n, np.float32)
            om = np.zeros(n * n, np.float32)
            rep = C.c_uint32()
            assert lib.ocean_amd_get_tables(sim, fptr(h0), fptr(om), C.byref(rep)) == 0
            assert_bitexact(h0, want_h0, f"extreme {k} h0 {kw}")
            assert_bitexact(om, want_om, f"extreme {k} omega {kw}")
            repaired += rep.value
            cells += n * n
        finally:
            lib.ocean_destroy(sim)
    print(f"extremes: repaired {repaired} of {cells} cells ({100.0 * repaired / cells:.3f}%)")
