"""Pins the CPU oracle (oracle/ocean_oracle.c) before anything is compared to it.

Each check is one that a plausible slip in the restatement (a dropped term, a
swapped operand, a wrong index or sign) would fail:
  - known answers the reference's own tests hold (PCG32 KAT);
  - the reference library itself, compiled from its sources (oracle/_ref),
    compared bit for bit on seeded inputs (skipped where it was not built);
  - golden fixtures written from that reference build (tests/golden/);
  - closed forms: sin/cos to 2e-7, FFT against a naive double DFT, the
    chop-0 and zero-wind special cases, Hermitian symmetry of h0.
"""
import ctypes as C
import hashlib
import json
import os

import numpy as np
import pytest

from oceanlib_amd.workloads import (CONFIG1, CONFIG2, CONFIG3_LEVELS, SMALL_CASES, desc_from,
                                    query_points)
from oracle_ffi import (OracleOcean, RefOcean, as_array, assert_bitexact, default_foam_params,
                        fptr, oracle, ref, spectrum_tables)

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

needs_ref = pytest.mark.skipif(ref() is None, reason="reference build oracle/_ref not present")


def digest(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()[:16]


@pytest.fixture(scope="module")
def golden():
    with open(os.path.join(GOLDEN, "digests.json")) as f:
        return json.load(f)


@pytest.fixture(scope="module")
def small():
    return np.load(os.path.join(GOLDEN, "small_buffers.npz"))


# ---------------------------------------------------------------- primitives

def test_pcg32_known_answers():
    """reference tests/test_rng.cpp:17-25 (seed 42, stream 54)."""
    want = [0xA15C02B7, 0x7B47F409, 0xBA1D3330, 0x83D2F293,
            0xBFA4784B, 0xCBED606E, 0xBFC6A3AD, 0x812FFF6D]
    out = (C.c_uint32 * 8)()
    oracle().orc_pcg32_draws(42, 54, 8, out)
    assert list(out) == want


@needs_ref
def test_pcg32_and_gaussians_match_reference():
    o, r = oracle(), ref()
    for seed, stream in ((0, 0), (7, 1), (2**63 + 5, 12345)):
        a, b = (C.c_uint32 * 1000)(), (C.c_uint32 * 1000)()
        o.orc_pcg32_draws(seed, stream, 1000, a)
        r.oref_pcg32_draws(seed, stream, 1000, b)
        assert list(a) == list(b)
    for seed, row in ((1337, 0), (7, 511), (2**40, 4095)):
        a = np.zeros(2000, np.float32)
        b = np.zeros(2000, np.float32)
        o.orc_row_gaussians(seed, row, 1000, fptr(a))
        r.oref_row_gaussians(seed, row, 1000, fptr(b))
        assert_bitexact(a, b, f"gaussians seed={seed} row={row}")


def test_gaussian_moments():
    """Box-Muller normals: mean 0, variance 1 (reference tests/test_rng.cpp:60-100)."""
    a = np.zeros(2 * 100000, np.float32)
    oracle().orc_row_gaussians(12345, 0, 100000, fptr(a))
    assert abs(a.mean()) < 0.01
    assert abs(a.var() - 1.0) < 0.01


def _sincos(lib, prefix, x):
    s = np.zeros_like(x)
    c = np.zeros_like(x)
    getattr(lib, prefix + "sincos")(fptr(x), len(x), fptr(s), fptr(c))
    return s, c


def test_sincos_accuracy_and_landmarks():
    """|err| <= 2e-7 vs double libm on [0, 2pi) (reference tests/test_trig.cpp:16-75)."""
    x = np.linspace(0.0, 2 * np.pi, 2_000_003, dtype=np.float64).astype(np.float32)
    x = x[x < np.float32(2 * np.pi)]
    s, c = _sincos(oracle(), "orc_", x)
    xd = x.astype(np.float64)
    assert np.max(np.abs(s - np.sin(xd))) <= 2e-7
    assert np.max(np.abs(c - np.cos(xd))) <= 2e-7
    land = np.array([0.0, np.pi / 2, np.pi, 3 * np.pi / 2], np.float32)
    s, c = _sincos(oracle(), "orc_", land)
    assert s[0] == 0.0 and c[0] == 1.0
    assert abs(s[1] - 1) < 1e-7 and abs(c[2] + 1) < 1e-7 and abs(s[3] + 1) < 1e-7


@needs_ref
def test_sincos_and_phase_bitexact_vs_reference():
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.uniform(0, 2 * np.pi, 1_000_000),
                        rng.uniform(-7, 7, 100_000)]).astype(np.float32)
    s1, c1 = _sincos(oracle(), "orc_", x)
    s2, c2 = _sincos(ref(), "oref_", x)
    assert_bitexact(s1, s2, "sin")
    assert_bitexact(c1, c2, "cos")
    w = rng.uniform(0, 40, 100_000).astype(np.float32)
    for t in (0.0, 1 / 60, 10.0, 3600.0, 36000.0):
        a = np.zeros_like(w)
        b = np.zeros_like(w)
        oracle().orc_reduced_phase(fptr(w), len(w), t, fptr(a))
        ref().oref_reduced_phase(fptr(w), len(w), t, fptr(b))
        assert_bitexact(a, b, f"phase t={t}")


# ---------------------------------------------------------------- spectrum

SPECTRUM_CASES = [CONFIG1, CONFIG2, CONFIG3_LEVELS[2],
                  dict(size=64, patch_length=30.0, wind_speed=9.0, swell=0.7, wind_direction=-2.0,
                       peak_enhancement=1.0, small_wave_cutoff=0.0, seed=2**63 + 11),
                  dict(size=32, patch_length=5.0, wind_speed=25.0, water_depth=0.5, fetch=2000.0)]


@needs_ref
@pytest.mark.parametrize("kw", SPECTRUM_CASES)
def test_spectrum_tables_vs_reference(kw):
    d = desc_from(kw)
    a = spectrum_tables(oracle(), "orc_", d)
    b = spectrum_tables(ref(), "oref_", d)
    for k in a:
        assert_bitexact(a[k], b[k], k)


def test_spectrum_tables_vs_golden_and_mirror(small):
    name, kw, _ = SMALL_CASES[0]
    d = desc_from(kw)
    a = spectrum_tables(oracle(), "orc_", d)
    for k in a:
        assert_bitexact(a[k], small[f"{name}_spectrum_{k}"], k)
    # conj(h0(-k)) mirror (reference tests/test_spectrum.cpp:373-394)
    n = d.size
    re, im = a["h0_re"].reshape(n, n), a["h0_im"].reshape(n, n)
    mz = (n - np.arange(n)) % n
    assert_bitexact(a["h0c_re"].reshape(n, n), re[mz][:, mz], "h0c_re mirror")
    assert_bitexact(a["h0c_im"].reshape(n, n), -im[mz][:, mz], "h0c_im mirror")
    assert a["k_inv"][0] == 0.0 and a["h0_re"][0] == 0.0 and a["omega"][0] == 0.0


def test_zero_wind_spectrum_is_flat():
    """Zero wind: no energy anywhere (reference tests/test_api.cpp:88-106)."""
    d = desc_from(dict(size=32, wind_speed=0.0))
    a = spectrum_tables(oracle(), "orc_", d)
    assert not np.any(a["h0_re"]) and not np.any(a["h0_im"])


# ---------------------------------------------------------------- FFT

def naive_idft(x):
    n = len(x)
    k = np.arange(n)
    return (np.exp(2j * np.pi * np.outer(k, k) / n) @ x.astype(np.complex128))


@pytest.mark.parametrize("n", [16, 64, 256])
def test_fft_vs_naive_dft(n):
    """Inverse, unnormalised (reference src/core/fft.hpp:16-26; tests/test_fft.cpp:102-169)."""
    rng = np.random.default_rng(n)
    re = rng.standard_normal(n).astype(np.float32)
    im = rng.standard_normal(n).astype(np.float32)
    want = naive_idft(re + 1j * im.astype(np.float64))
    r, i = re.copy(), im.copy()
    oracle().orc_fft_inverse(n, fptr(r), fptr(i))
    got = r + 1j * i.astype(np.float64)
    assert np.max(np.abs(got - want)) / np.max(np.abs(want)) < 1e-6
    # impulse at 0 -> constant (tests/test_fft.cpp:72-96)
    r = np.zeros(n, np.float32)
    i = np.zeros(n, np.float32)
    r[0] = 1.0
    oracle().orc_fft_inverse(n, fptr(r), fptr(i))
    assert np.all(r == 1.0) and np.all(i == 0.0)
    # +i sign convention: X[1] = 1 -> x[j] = exp(+2 pi i j / n) (tests/test_fft.cpp:46-70)
    r = np.zeros(n, np.float32)
    i = np.zeros(n, np.float32)
    r[1] = 1.0
    oracle().orc_fft_inverse(n, fptr(r), fptr(i))
    j = np.arange(n)
    assert np.max(np.abs(i - np.sin(2 * np.pi * j / n))) < 1e-6


@needs_ref
@pytest.mark.parametrize("n", [16, 32, 128, 512, 2048])
def test_fft_bitexact_vs_reference(n):
    rng = np.random.default_rng(n + 1)
    re = rng.standard_normal(n).astype(np.float32)
    im = rng.standard_normal(n).astype(np.float32)
    re[::7] = 0.0
    im[::5] = -0.0
    a, b = re.copy(), im.copy()
    c, d = re.copy(), im.copy()
    oracle().orc_fft_inverse(n, fptr(a), fptr(b))
    ref().oref_fft_inverse(n, fptr(c), fptr(d))
    assert_bitexact(a, c, "re")
    assert_bitexact(b, d, "im")
    if n <= 128:
        re2 = rng.standard_normal(n * n).astype(np.float32)
        im2 = rng.standard_normal(n * n).astype(np.float32)
        a, b, c, d = re2.copy(), im2.copy(), re2.copy(), im2.copy()
        oracle().orc_fft2d_inverse(n, fptr(a), fptr(b))
        ref().oref_fft2d_inverse(n, fptr(c), fptr(d))
        assert_bitexact(a, c, "2d re")
        assert_bitexact(b, d, "2d im")


# ---------------------------------------------------------------- full frames

def test_small_buffers_vs_golden(small):
    for name, kw, times in SMALL_CASES:
        o = OracleOcean(desc_from(kw))
        for t in times:
            o.update(t)
            key = f"{name}_t{t:g}"
            assert_bitexact(o.displacement(), small[key + "_disp"], key + " disp")
            assert_bitexact(o.normal(), small[key + "_normal"], key + " normal")
            if key + "_vel" in small:
                assert_bitexact(o.velocity(), small[key + "_vel"], key + " vel")


def test_config1_frames_vs_golden(golden):
    o = OracleOcean(desc_from(CONFIG1))
    for k in range(0, 601, 20):
        o.update(k / 60.0)
        assert [digest(o.displacement()), digest(o.normal())] == golden["config1"][k], k


def test_config2_frames_foam_queries_vs_golden(golden):
    """Heavy chop + velocity + advected foam + 4096 queries; the first 40 frames
    (foam is history-dependent, so every frame up to the last checked runs)."""
    lib = oracle()
    desc = desc_from(CONFIG2)
    o = OracleOcean(desc)
    fp = default_foam_params()
    foam = lib.orc_foam_create(o.h, C.byref(fp))
    n = desc.size
    frames = 40
    xz = query_points(600, 4096, desc.patch_length)[:frames]
    from oceanlib_amd.capi import OceanSurfaceEx
    try:
        for k in range(frames):
            o.update(k / 60.0)
            lib.orc_foam_update(foam, np.float32(1.0 / 60.0))
            fb = as_array(lib.orc_foam_data(foam), n * n)
            want = golden["config2"][k]
            assert digest(o.displacement()) == want[0], k
            assert digest(o.normal()) == want[1], k
            assert digest(o.velocity()) == want[2], k
            assert digest(fb) == want[3], k
            if k % 13 == 0:
                out = (OceanSurfaceEx * 4096)()
                pts = np.ascontiguousarray(xz[k].reshape(-1))
                lib.orc_sample_buffers(C.byref(desc), lib.orc_ocean_displacement(o.h),
                                       lib.orc_ocean_normal(o.h), lib.orc_ocean_velocity(o.h),
                                       fptr(pts), 4096, out)
                q = np.frombuffer(bytes(out), np.float32).reshape(4096, 10)
                assert digest(q) == want[4], k
    finally:
        lib.orc_foam_destroy(foam)


def test_config3_cascades_vs_golden(golden):
    levels = [OracleOcean(desc_from(kw)) for kw in CONFIG3_LEVELS]
    lib = oracle()
    arr = (C.c_void_p * 3)(*[lv.h for lv in levels])
    pts = query_points(1, 512, 800.0)[0]
    from oceanlib_amd.capi import OceanSurfaceEx
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                           "config3_frames.json")) as f:
        every = json.load(f)["cascade"]
    extra = [str(k) for k in np.random.default_rng(3).choice(1200, 5, replace=False)]
    for k in ("0", "1", "600", *extra):
        t = int(k) / 60.0
        for lv in levels:
            lv.update(t)
        g = golden["config3_cascade"].get(k)
        want = every[int(k)]
        if g is not None:
            assert [want[0:2], want[2:4], want[4:6]] == g["levels"] and want[6] == g["queries"]
        for li, lv in enumerate(levels):
            assert [digest(lv.displacement()), digest(lv.normal())] == want[2 * li:2 * li + 2], (k, li)
        q = np.zeros((len(pts), 10), np.float32)
        s = OceanSurfaceEx()
        for i, (x, z) in enumerate(pts):
            lib.orc_cascade_sample(arr, 3, float(x), float(z), C.byref(s))
            q[i] = np.frombuffer(bytes(s), np.float32)
        assert digest(q) == want[6], k


@needs_ref
def test_full_update_and_queries_vs_reference():
    for kw in (dict(size=128, wind_speed=30.0, choppiness=2.5, compute_velocity=1, seed=99),
               dict(size=64, patch_length=10.0, wind_speed=3.0, water_depth=0.2, seed=1)):
        d = desc_from(kw)
        o, r = OracleOcean(d), RefOcean(d)
        for t in (0.0, 0.37, 1e4):
            o.update(t)
            r.update(t)
            assert_bitexact(o.displacement(), r.displacement(), "disp")
            assert_bitexact(o.normal(), r.normal(), "normal")
            if d.compute_velocity:
                assert_bitexact(o.velocity(), r.velocity(), "vel")
            rng = np.random.default_rng(5)
            for x, z in rng.uniform(-300, 300, size=(200, 2)).astype(np.float32):
                assert bytes(o.sample(float(x), float(z))) == bytes(r.sample(float(x), float(z)))


def test_chop_zero_exact_special_case():
    """chop 0: J = 1, foam 0, zero horizontal displacement, exact lookup
    (reference tests/test_simulation.cpp:275-296, tests/test_query.cpp:98-118)."""
    d = desc_from(dict(size=64, choppiness=0.0, wind_speed=14.0))
    o = OracleOcean(d)
    o.update(2.0)
    disp = o.displacement().reshape(-1, 4)
    nrm = o.normal().reshape(-1, 4)
    assert np.all(disp[:, 0] == 0.0) and np.all(disp[:, 2] == 0.0)
    assert np.all(nrm[:, 3] == 1.0) and np.all(disp[:, 3] == 0.0)
    cell = d.patch_length / d.size
    s = o.sample(5 * cell, 7 * cell)
    assert s.height == disp[7 * d.size + 5, 1]


def test_parseval_energy():
    """Spatial variance of h equals the spectral energy (reference
    tests/test_simulation.cpp:118-185): sum |h(x)|^2 = N^2 sum |h(k,t)|^2."""
    d = desc_from(dict(size=64, wind_speed=12.0, seed=4))
    o = OracleOcean(d)
    n = d.size
    planes = np.zeros(8 * n * n, np.float32)
    oracle().orc_ocean_evolve(o.h, 3.0, fptr(planes))
    # field 0 = h + i*Dx; the real part after the transform is h.
    o.update(3.0)
    h = o.displacement().reshape(-1, 4)[:, 1].astype(np.float64)
    f0 = planes[0:n * n].astype(np.float64) + 1j * planes[n * n:2 * n * n]
    # Spectrum of the real field h alone: H(k) = (F0(k) + conj(F0(-k))) / 2
    F0 = f0.reshape(n, n)
    m = (n - np.arange(n)) % n
    H = 0.5 * (F0 + np.conj(F0[m][:, m]))
    lhs = np.sum(h * h)
    rhs = n * n * np.sum(np.abs(H) ** 2)
    assert abs(lhs - rhs) / rhs < 1e-4


@needs_ref
@pytest.mark.parametrize("kw,dts", [
    (dict(size=128, patch_length=120.0, wind_speed=22.0, choppiness=1.3, compute_velocity=1,
          water_depth=50.0, seed=7), [1 / 60.0]),
    (dict(size=64, patch_length=40.0, wind_speed=30.0, choppiness=2.5, compute_velocity=1,
          seed=99), [1 / 60.0, 0.05, 0.0, 0.3, 1 / 30.0]),
])
def test_foam_at_and_coverage_vs_reference(kw, dts):
    """orc_foam_at (FoamField::foam_at, reference src/foam.cpp:291-305) and
    orc_foam_coverage (src/foam.cpp:307-313) against the reference's own
    FoamField over a history of substeps (multi-substep and clamped dt
    included), at points on, off and far outside the patch."""
    d = desc_from(kw)
    o, r = OracleOcean(d), RefOcean(d)
    lib, rl = oracle(), ref()
    fp = default_foam_params()
    fo = lib.orc_foam_create(o.h, C.byref(fp))
    fr = rl.oref_foam_create(r.h, C.byref(fp))
    n = d.size
    rng = np.random.default_rng(11)
    pts = np.concatenate([
        rng.uniform(0.0, d.patch_length, size=(150, 2)),
        rng.uniform(-3.0 * d.patch_length, 4.0 * d.patch_length, size=(100, 2)),
        np.array([[0.0, 0.0], [d.patch_length, d.patch_length], [-1e4, 2e4], [-0.0, 1e-7]]),
    ]).astype(np.float32)
    try:
        t = 0.0
        for k in range(24):
            dt = dts[k % len(dts)]
            t += dt
            o.update(t)
            r.update(t)
            lib.orc_foam_update(fo, np.float32(dt))
            rl.oref_foam_update(fr, np.float32(dt))
            assert lib.orc_foam_last_substeps(fo) == rl.oref_foam_last_substeps(fr)
            assert_bitexact(as_array(lib.orc_foam_data(fo), n * n),
                            as_array(rl.oref_foam_data(fr), n * n), f"foam step {k}")
            assert np.float64(lib.orc_foam_coverage(fo)).tobytes() == \
                np.float64(rl.oref_foam_coverage(fr)).tobytes(), k
            if k % 6 == 5:
                for x, z in pts:
                    a = lib.orc_foam_at(fo, float(x), float(z))
                    b = rl.oref_foam_at(fr, float(x), float(z))
                    assert np.float32(a).tobytes() == np.float32(b).tobytes(), (k, x, z)
        assert lib.orc_foam_coverage(fo) > 0.0
    finally:
        lib.orc_foam_destroy(fo)
        rl.oref_foam_destroy(fr)
