"""The device tables' certification, checked on the CPU against libm errors.

The device builds h0, omega and the FFT twiddles with its own libm and flags
every cell whose floats a libm difference could change
(oceanlib_amd/csrc/spectrum_cell.hpp: every libm call is charged U = 8 double
ulps, device and host error together); the host recomputes flagged cells by
the reference's chain (reference src/core/spectrum.cpp:14-339,
src/core/rng.hpp:27-97, src/core/rng.cpp:7-17, src/core/fft.cpp:24-33).
tests/cpu_emul/spectrum_cert_emul.cpp runs that same per-cell code on the
host with a libm whose every result is moved by up to U ulps from glibc's:

- with no perturbation the cell code equals the host chain bit for bit on
  every cell (the device restatement is the chain, operation for operation);
- with perturbations of up to 8 ulps (half of the draws at +-8 exactly) every
  float that changes lies on a flagged cell, over the workload descriptors
  and descriptors at the parameter extremes, several perturbation streams each;
- perturbations far beyond the charged 8 ulps do change unflagged cells, so
  the check above is able to fail.
"""
import ctypes as C
import os
import shutil
import subprocess
import zlib

import numpy as np
import pytest

from oceanlib_amd.workloads import CONFIG1, CONFIG2, CONFIG3_LEVELS, desc_from

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = [os.path.join(HERE, "cpu_emul", "spectrum_cert_emul.cpp"),
       os.path.join(ROOT, "oceanlib_amd", "csrc", "spectrum_host.cpp")]

_F = C.POINTER(C.c_float)
_U8 = C.POINTER(C.c_uint8)


@pytest.fixture(scope="module")
def emul(tmp_path_factory):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    out = str(tmp_path_factory.mktemp("cert") / "libcertemul.so")
    subprocess.run([hipcc, "-std=c++20", "-O2", "--offload-host-only", "-x", "hip", "-ffp-contract=off",
                    "-fno-fast-math", "-fPIC", "-shared", "-I", os.path.join(ROOT, "include"),
                    "-I", os.path.join(ROOT, "oceanlib_amd", "csrc"), *SRC, "-o", out],
                   check=True, capture_output=True)
    lib = C.CDLL(out)
    lib.emul_tables.argtypes = [C.c_uint32, C.c_float, C.c_uint64, _F, C.c_int, C.c_uint64,
                                _F, _F, _U8, _F, _F]
    lib.emul_tables.restype = C.c_int
    lib.emul_twiddles.argtypes = [C.c_uint32, C.c_int, C.c_uint64, _F, _U8, _F]
    lib.emul_twiddles.restype = C.c_int
    return lib


def ptr(a, t=_F):
    return a.ctypes.data_as(t)


def tables(emul, kw, ulps, pseed):
    d = desc_from(kw)
    n = d.size
    s = d.spectrum
    spec = np.array([s.wind_speed, s.fetch, s.wind_direction, s.peak_enhancement, s.swell,
                     s.small_wave_cutoff, s.gravity, d.water_depth], np.float32)
    h0 = np.zeros(2 * n * n, np.float32)
    om = np.zeros(n * n, np.float32)
    fl = np.zeros(n * n, np.uint8)
    h0_host = np.zeros(2 * n * n, np.float32)
    om_host = np.zeros(n * n, np.float32)
    assert emul.emul_tables(n, d.patch_length, d.seed, ptr(spec), ulps, pseed, ptr(h0), ptr(om),
                            ptr(fl, _U8), ptr(h0_host), ptr(om_host)) == 0
    changed = ((h0.view(np.uint32) != h0_host.view(np.uint32)).reshape(-1, 2).any(1) |
               (om.view(np.uint32) != om_host.view(np.uint32)))
    return changed, fl.astype(bool)


def extreme_descs(count=48, seed=91):
    """Descriptors at the parameter extremes (swell 1, depth down to 0.5 m,
    fetch 10^3, peak enhancement 7, winds 0.5 and 40 m/s, cutoffs and
    directions across their range), as the device test draws them."""
    rng = np.random.default_rng(seed)
    pick = lambda *v: float(np.float32(v[int(rng.integers(0, len(v)))]))  # noqa: E731
    out = []
    for _ in range(count):
        out.append(dict(size=int(rng.choice([32, 64, 128])),
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


WORKLOADS = [CONFIG1, CONFIG2] + list(CONFIG3_LEVELS)
CASES = [(f"workload_{k}", kw) for k, kw in enumerate(WORKLOADS)] + \
        [(f"extreme_{k}", kw) for k, kw in enumerate(extreme_descs())]


@pytest.mark.parametrize("name,kw", CASES[:6], ids=[c[0] for c in CASES[:6]])
def test_unperturbed_cell_code_equals_the_host_chain(emul, name, kw):
    changed, _ = tables(emul, kw, 0, 1)
    assert not changed.any(), f"{name}: {int(changed.sum())} cells differ from the host chain"


def test_perturbed_libm_changes_only_flagged_cells(emul):
    total_changed = total_flagged = cells = 0
    for name, kw in CASES:
        for pseed in (11, 12, 13, 14, 15):
            changed, flags = tables(emul, kw, 8, pseed * 1000003 + zlib.crc32(name.encode()))
            escaped = changed & ~flags
            assert not escaped.any(), \
                f"{name} stream {pseed}: {int(escaped.sum())} changed cells not flagged, first " \
                f"{np.flatnonzero(escaped)[:5]}"
            total_changed += int(changed.sum())
            total_flagged += int(flags.sum())
            cells += changed.size
    print(f"8-ulp libm: {total_changed} changed cells, all flagged; {total_flagged} flagged "
          f"of {cells} ({100.0 * total_flagged / cells:.3f}%)")
    # the perturbations do reach the floats: the check is not vacuous
    assert total_changed > 0


def test_perturbations_beyond_the_charged_ulps_escape(emul):
    """2^24 ulps per call (far beyond the 8 the bound charges) must change
    unflagged cells somewhere: the containment check can fail."""
    escaped = 0
    for name, kw in CASES[:8]:
        changed, flags = tables(emul, kw, 1 << 24, 7)
        escaped += int((changed & ~flags).sum())
    assert escaped > 0


@pytest.mark.parametrize("n", [16, 64, 256, 1024, 4096])
def test_twiddles_perturbed_change_only_flagged(emul, n):
    tw = np.zeros(n, np.float32)
    tw_host = np.zeros(n, np.float32)
    fl = np.zeros(n // 2, np.uint8)
    assert emul.emul_twiddles(n, 0, 1, ptr(tw), ptr(fl, _U8), ptr(tw_host)) == 0
    assert np.array_equal(tw.view(np.uint32), tw_host.view(np.uint32))
    for pseed in (3, 4, 5, 6):
        assert emul.emul_twiddles(n, 8, pseed, ptr(tw), ptr(fl, _U8), ptr(tw_host)) == 0
        changed = (tw.view(np.uint32) != tw_host.view(np.uint32)).reshape(-1, 2).any(1)
        assert not (changed & ~fl.astype(bool)).any(), (n, pseed)
