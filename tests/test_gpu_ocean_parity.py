"""GPU parity: the HIP frame path through the C ABI vs the CPU oracle, bit for bit.

Mirrors the reference's determinism contract (every SIMD level memcmp-equal to
scalar, reference tests/test_simd.cpp:113-166; C == C++ bit for bit,
tests/test_c_api.cpp:51-95): here the HIP kernels must equal the oracle's
restatement of the reference's scalar path, word for word.
"""
import ctypes as C
import hashlib
import json
import os

import numpy as np
import pytest

from oceanlib_amd import capi
from oracle_ffi import OracleOcean, assert_bitexact, fptr

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    return capi.load()


class GpuOcean:
    def __init__(self, lib, desc):
        self.lib = lib
        st = C.c_int(-1)
        self.h = lib.ocean_create(C.byref(desc), C.byref(st))
        assert self.h, f"ocean_create status {st.value}"
        self.n = desc.size
        self.desc = desc

    def update(self, t):
        self.lib.ocean_update(self.h, t)

    def buffers(self):
        b = self.lib.ocean_get_buffers(self.h)
        n4 = 4 * self.n * self.n
        d = np.ctypeslib.as_array(b.displacement, shape=(n4,)).copy()
        nr = np.ctypeslib.as_array(b.normal, shape=(n4,)).copy()
        vp = self.lib.ocean_get_velocity(self.h)
        v = np.ctypeslib.as_array(vp, shape=(n4,)).copy() if vp else None
        return d, nr, v

    def sample_batch(self, xz):
        xz = np.ascontiguousarray(xz, dtype=np.float32)
        out = (capi.OceanSurfaceEx * len(xz))()
        rc = self.lib.ocean_amd_sample_batch(self.h, fptr(xz.reshape(-1)), len(xz), out)
        assert rc == 0
        return out

    def close(self):
        if self.h:
            assert self.lib.ocean_amd_last_error(self.h) == 0
            self.lib.ocean_destroy(self.h)
            self.h = None


def check_frame(g, o, t, what):
    g.update(t)
    o.update(t)
    d, nr, v = g.buffers()
    assert_bitexact(d, o.displacement(), f"{what} t={t} displacement")
    assert_bitexact(nr, o.normal(), f"{what} t={t} normal")
    ov = o.velocity()
    if ov is not None:
        assert v is not None
        assert_bitexact(v, ov, f"{what} t={t} velocity")


C1 = dict(size=256, patch_length=200.0, wind_speed=12.0, seed=1337, choppiness=1.0)
C2 = dict(size=512, patch_length=250.0, wind_speed=22.0, choppiness=1.3, compute_velocity=1,
          water_depth=50.0, seed=7)


def test_config1_plain_patch(lib):
    """N=256, 200 m, 12 m/s, seed 1337: t = 0..10 s at 1/60 s (sampled) plus far seeks."""
    d = capi.default_desc(lib, **C1)
    g, o = GpuOcean(lib, d), OracleOcean(d)
    try:
        d0, n0, _ = g.buffers()   # the state at creation is t = 0
        assert_bitexact(d0, o.displacement(), "creation displacement")
        assert_bitexact(n0, o.normal(), "creation normal")
        for k in list(range(0, 601, 37)) + [600]:
            check_frame(g, o, k / 60.0, "C1")
        for t in (3600.0, 36000.0, 1e-7, 123.456):
            check_frame(g, o, t, "C1 seek")
    finally:
        g.close()


def test_config1_every_frame_vs_reference_digests(lib):
    """The whole plain-patch config: every frame t = k/60, k = 0..600, of
    displacement and normal against the reference run's digests
    (tests/golden/digests.json, written from the reference build)."""
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                           "digests.json")) as f:
        want = json.load(f)["config1"]
    assert len(want) == 601
    d = capi.default_desc(lib, **C1)
    g = GpuOcean(lib, d)
    n4 = 4 * d.size * d.size
    try:
        for k in range(601):
            g.update(k / 60.0)
            b = lib.ocean_get_buffers(g.h)
            got = [hashlib.sha256(np.ctypeslib.as_array(p, shape=(n4,)).tobytes()).hexdigest()[:16]
                   for p in (b.displacement, b.normal)]
            assert got == want[k], f"frame {k}"
    finally:
        g.close()


def test_config2_heavy_chop_velocity_depth(lib):
    """N=512, 250 m, 22 m/s, chop 1.3, velocity on, depth 50 m, seed 7."""
    d = capi.default_desc(lib, **C2)
    g, o = GpuOcean(lib, d), OracleOcean(d)
    try:
        for k in (0, 1, 2, 59, 300, 599):
            check_frame(g, o, k / 60.0, "C2")
    finally:
        g.close()


@pytest.mark.parametrize("n", [16, 32, 64, 128, 1024])
def test_every_grid_size(lib, n):
    d = capi.default_desc(lib, size=n, patch_length=float(n) * 0.7, wind_speed=9.0,
                          seed=1000 + n, compute_velocity=int(n % 64 == 0))
    g, o = GpuOcean(lib, d), OracleOcean(d)
    try:
        for t in (0.5, 17.25):
            check_frame(g, o, t, f"N={n}")
    finally:
        g.close()


@pytest.mark.parametrize("n", [2048, 4096])
def test_largest_grid_sizes(lib, n):
    """The reference's upper bound (src/ocean.cpp:33-37)."""
    d = capi.default_desc(lib, size=n, patch_length=1000.0, wind_speed=14.0, seed=5)
    g, o = GpuOcean(lib, d), OracleOcean(d)
    try:
        check_frame(g, o, 2.0, f"N={n}")
    finally:
        g.close()


def test_special_cases(lib):
    """Zero wind -> bit-flat sea; chop 0 -> J = 1, zero horizontal displacement
    (reference tests/test_api.cpp:88-106, tests/test_simulation.cpp:275-296)."""
    d = capi.default_desc(lib, size=64, wind_speed=0.0)
    g = GpuOcean(lib, d)
    try:
        g.update(5.0)
        disp, nrm, _ = g.buffers()
        assert not np.any(disp.reshape(-1, 4)[:, :3])
        assert np.all(nrm.reshape(-1, 4)[:, 1] == 1.0)
    finally:
        g.close()
    d = capi.default_desc(lib, size=64, choppiness=0.0, wind_speed=15.0)
    g, o = GpuOcean(lib, d), OracleOcean(d)
    try:
        check_frame(g, o, 3.0, "chop0")
        disp, nrm, _ = g.buffers()
        assert np.all(disp.reshape(-1, 4)[:, 0] == 0.0)
        assert np.all(disp.reshape(-1, 4)[:, 2] == 0.0)
        assert np.all(nrm.reshape(-1, 4)[:, 3] == 1.0)
        assert np.all(disp.reshape(-1, 4)[:, 3] == 0.0)
    finally:
        g.close()


def test_queries_bitexact(lib):
    """Batched and single queries vs the oracle's sample_at (src/ocean.cpp:375-452)."""
    d = capi.default_desc(lib, **C2)
    g, o = GpuOcean(lib, d), OracleOcean(d)
    try:
        t = 7.5
        g.update(t)
        o.update(t)
        rng = np.random.default_rng(3)
        xz = rng.uniform(0.0, 250.0, size=(4096, 2)).astype(np.float32)
        xz[:8] = [[0, 0], [-0.0, 250.0], [-1e3, 3e3], [249.999, 0.001], [125, 125],
                  [-250, -250], [1e5, -1e5], [0.5, -0.5]]
        out = g.sample_batch(xz)
        for q in range(len(xz)):
            want = o.sample(float(xz[q, 0]), float(xz[q, 1]))
            assert bytes(out[q]) == bytes(want), (q, xz[q])
        # single-query C entry points
        for q in range(0, 64, 7):
            x, z = float(xz[q, 0]), float(xz[q, 1])
            want = o.sample(x, z)
            h = lib.ocean_height_at(g.h, x, z)
            assert np.float32(h).tobytes() == np.float32(want.height).tobytes()
            s = capi.OceanSurfaceEx()
            lib.ocean_sample_ex(g.h, x, z, C.byref(s))
            assert bytes(s) == bytes(want)
    finally:
        g.close()


@pytest.mark.parametrize("n", [32, 64, 128, 256, 512, 1024, 2048, 4096])
def test_every_launch_shape_is_bitexact(lib, n):
    """Each launch-shape variant of both passes (rows / columns per
    workgroup, field split, 4 / 8 / 16 FFT points per lane), in each layout
    of the fields between the passes (row-major; tiles of 2, 4, 8 columns
    from N = 256), gives the oracle's bits; the autotuner's pick does too."""
    vel = 1 if n in (256, 512, 1024) else 0
    d = capi.default_desc(lib, size=n, patch_length=float(n) * 0.7, wind_speed=17.0, seed=n + 3,
                          choppiness=1.2, compute_velocity=vel, water_depth=30.0 if vel else 0.0)
    g, o = GpuOcean(lib, d), OracleOcean(d)
    rows, cols = C.c_uint32(), C.c_uint32()
    try:
        t = 4.25
        o.update(t)
        want = (o.displacement(), o.normal(), o.velocity())
        layouts = [0] + ([1, 2, 3] if n >= 256 else [])
        assert lib.ocean_amd_set_frame_layout(g.h, 1) == (0 if n >= 256 else capi.OCEAN_ERROR_INVALID_ARG)
        assert lib.ocean_amd_set_frame_layout(g.h, 4) == capi.OCEAN_ERROR_INVALID_ARG
        for layout in layouts:
            assert lib.ocean_amd_set_frame_layout(g.h, layout) == 0
            assert lib.ocean_amd_get_frame_layout(g.h) == layout
            assert lib.ocean_amd_frame_shapes(g.h, C.byref(rows), C.byref(cols)) == 0
            assert rows.value >= 1 and cols.value >= 1
            shapes = [(r, 0) for r in range(rows.value)] + [(0, c) for c in range(1, cols.value)]
            for r, c in shapes:
                assert lib.ocean_amd_set_frame_shape(g.h, r, c) == 0
                g.update(t + 1.0)  # a different frame first, so stale buffers cannot pass
                g.update(t)
                dd, nr, v = g.buffers()
                what = f"n={n} layout={layout} shape=({r},{c})"
                assert_bitexact(dd, want[0], f"{what} displacement")
                assert_bitexact(nr, want[1], f"{what} normal")
                if vel:
                    assert_bitexact(v, want[2], f"{what} velocity")
            assert lib.ocean_amd_set_frame_shape(g.h, rows.value, 0) != 0
        rv, cv = C.c_uint32(), C.c_uint32()
        rms, cms = C.c_float(), C.c_float()
        assert lib.ocean_amd_autotune_frame(g.h, 3, C.byref(rv), C.byref(cv), C.byref(rms),
                                            C.byref(cms)) == 0
        assert lib.ocean_amd_get_frame_layout(g.h) in layouts
        assert lib.ocean_amd_frame_shapes(g.h, C.byref(rows), C.byref(cols)) == 0
        assert rv.value < rows.value and cv.value < cols.value and rms.value > 0 and cms.value > 0
        g.update(t)
        dd, nr, _ = g.buffers()
        assert_bitexact(dd, want[0], f"n={n} tuned displacement")
        assert_bitexact(nr, want[1], f"n={n} tuned normal")
    finally:
        g.close()


def test_frame_timing_inside_a_loop(lib):
    """ocean_amd_frame_timing: the timed frames' dispatches record their own
    events and compute the same bits; every k-th update is sampled, the ring
    keeps the last 64, reading clears it."""
    d = capi.default_desc(lib, size=256, patch_length=200.0, wind_speed=12.0, seed=1337)
    g, o = GpuOcean(lib, d), OracleOcean(d)
    try:
        r, c, k = C.c_float(), C.c_float(), C.c_uint32()
        assert lib.ocean_amd_frame_timing(g.h, 3) == 0
        for i in range(10):
            g.update(i / 60.0)
        assert lib.ocean_amd_frame_timing_read(g.h, C.byref(r), C.byref(c), C.byref(k)) == 0
        assert k.value == 4 and r.value > 0 and c.value > 0  # frames 0, 3, 6, 9
        o.update(9 / 60.0)
        dd, nr, _ = g.buffers()
        assert_bitexact(dd, o.displacement(), "timed frame displacement")
        assert_bitexact(nr, o.normal(), "timed frame normal")
        assert lib.ocean_amd_frame_timing_read(g.h, C.byref(r), C.byref(c), C.byref(k)) == 0
        assert k.value == 0
        lib.ocean_amd_frame_timing(g.h, 1)
        for i in range(70):
            g.update(i / 60.0)
        assert lib.ocean_amd_frame_timing_read(g.h, C.byref(r), C.byref(c), C.byref(k)) == 0
        assert k.value == 64
        assert lib.ocean_amd_frame_timing(g.h, 0) == 0
        g.update(1.0)
        assert lib.ocean_amd_frame_timing_read(g.h, C.byref(r), C.byref(c), C.byref(k)) == 0
        assert k.value == 0
    finally:
        g.close()
