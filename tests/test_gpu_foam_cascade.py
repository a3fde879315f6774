"""GPU parity for the persistent foam field and the cascade stack.

FoamField (reference src/foam.cpp:89-139,218-252,291-305) and CascadeStack
(reference src/cascade.cpp:42-153) through the additive C API, against the
oracle bit for bit and against the golden digests of the reference run.
"""
import ctypes as C
import hashlib
import json
import os

import numpy as np
import pytest

from oceanlib_amd import capi
from oceanlib_amd.workloads import CONFIG2, CONFIG3_LEVELS, desc_from, query_points
from oracle_ffi import OracleOcean, as_array, assert_bitexact, default_foam_params, fptr, oracle

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def digest(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()[:16]


@pytest.fixture(scope="module")
def lib():
    return capi.load()


@pytest.fixture(scope="module")
def golden():
    with open(os.path.join(GOLDEN, "digests.json")) as f:
        return json.load(f)


def test_config2_foam_and_queries_vs_golden_and_oracle(lib, golden):
    """The whole heavy-chop config: all 600 frames of surface, velocity,
    advected foam and the frame's 4096 queries against the reference run's
    digests, every frame; the first 120 frames also in lockstep with the
    oracle's foam field; after the last frame foam_at at 256 points and
    coverage against the reference's own values (tests/golden/config2_final.json)."""
    with open(os.path.join(GOLDEN, "config2_final.json")) as f:
        final = json.load(f)
    desc = desc_from(CONFIG2)
    n = desc.size
    st = C.c_int(-1)
    sim = lib.ocean_create(C.byref(desc), C.byref(st))
    assert sim, st.value
    fd = capi.FoamDesc()
    lib.ocean_amd_foam_desc_init(C.byref(fd))
    foam = lib.ocean_amd_foam_create(sim, C.byref(fd), C.byref(st))
    assert foam, st.value
    o = OracleOcean(desc)
    olib = oracle()
    fp = default_foam_params()
    ofoam = olib.orc_foam_create(o.h, C.byref(fp))
    frames = final["frames"]
    assert frames == len(golden["config2"]) == 600
    xz = query_points(frames, 4096, desc.patch_length)
    out = (capi.OceanSurfaceEx * 4096)()
    try:
        for k in range(frames):
            t = k / 60.0
            lib.ocean_update(sim, t)
            lib.ocean_amd_foam_update(foam, np.float32(1.0 / 60.0))
            want = golden["config2"][k]
            b = lib.ocean_get_buffers(sim)
            disp = np.ctypeslib.as_array(b.displacement, shape=(4 * n * n,))
            nrm = np.ctypeslib.as_array(b.normal, shape=(4 * n * n,))
            vel = np.ctypeslib.as_array(lib.ocean_get_velocity(sim), shape=(4 * n * n,))
            fb = np.ctypeslib.as_array(lib.ocean_amd_foam_data(foam), shape=(n * n,))
            pts = np.ascontiguousarray(xz[k].reshape(-1))
            assert lib.ocean_amd_sample_batch(sim, fptr(pts), 4096, out) == 0
            q = np.frombuffer(bytes(out), np.float32).reshape(4096, 10)
            assert [digest(disp), digest(nrm), digest(vel), digest(fb), digest(q)] == want, k
            if k < 120:
                o.update(t)
                olib.orc_foam_update(ofoam, np.float32(1.0 / 60.0))
                assert lib.ocean_amd_foam_last_substeps(foam) == olib.orc_foam_last_substeps(ofoam)
                if k % 10 == 0 or k == 119:
                    assert_bitexact(fb.copy(), as_array(olib.orc_foam_data(ofoam), n * n), "foam")
        pts = np.array(final["points"], np.uint32).view(np.float32).reshape(-1, 2)
        want_f = np.array(final["foam_at"], np.uint32)
        got_f = np.array([lib.ocean_amd_foam_at(foam, float(x), float(z)) for x, z in pts],
                         np.float32).view(np.uint32)
        assert np.array_equal(got_f, want_f), np.nonzero(got_f != want_f)[0][:8]
        # coverage: the reference's sequential double sum, bit for bit
        cov = np.float64(lib.ocean_amd_foam_coverage(foam))
        assert int(cov.view(np.uint64)) == final["coverage"] and cov > 0.0
        # clear -> zero field, accumulator reset (reference src/foam.cpp:254-260)
        lib.ocean_amd_foam_clear(foam)
        assert lib.ocean_amd_foam_coverage(foam) == 0.0
        assert lib.ocean_amd_last_error(sim) == 0
    finally:
        olib.orc_foam_destroy(ofoam)
        lib.ocean_amd_foam_destroy(foam)
        lib.ocean_destroy(sim)


def test_foam_requires_velocity_and_valid_desc(lib):
    d = desc_from(dict(size=32))
    st = C.c_int(-1)
    sim = lib.ocean_create(C.byref(d), C.byref(st))
    assert sim
    try:
        fd = capi.FoamDesc()
        lib.ocean_amd_foam_desc_init(C.byref(fd))
        assert not lib.ocean_amd_foam_create(sim, C.byref(fd), C.byref(st))
        assert st.value == capi.OCEAN_ERROR_INVALID_ARG
    finally:
        lib.ocean_destroy(sim)
    d = desc_from(dict(size=32, compute_velocity=1))
    sim = lib.ocean_create(C.byref(d), C.byref(st))
    try:
        for field, bad in (("decay", -1.0), ("fixed_dt", 0.0), ("max_substeps", 0)):
            fd = capi.FoamDesc()
            lib.ocean_amd_foam_desc_init(C.byref(fd))
            setattr(fd, field, bad)
            assert not lib.ocean_amd_foam_create(sim, C.byref(fd), C.byref(st)), field
            assert st.value == capi.OCEAN_ERROR_INVALID_ARG
    finally:
        lib.ocean_destroy(sim)


def test_foam_substep_clamp_and_whole_cell_shift(lib):
    """A huge dt runs max_substeps and discards the backlog (reference
    src/foam.cpp:234-240); the accumulator matches the oracle's step counts."""
    d = desc_from(dict(size=64, compute_velocity=1, wind_speed=20.0, choppiness=1.5))
    st = C.c_int(-1)
    sim = lib.ocean_create(C.byref(d), C.byref(st))
    fd = capi.FoamDesc()
    lib.ocean_amd_foam_desc_init(C.byref(fd))
    fd.wind_drift_x = 3.0
    foam = lib.ocean_amd_foam_create(sim, C.byref(fd), C.byref(st))
    o = OracleOcean(d)
    olib = oracle()
    fp = default_foam_params()
    fp.wind_drift_x = 3.0
    of = olib.orc_foam_create(o.h, C.byref(fp))
    try:
        for k, dt in enumerate((1 / 60, 1.0, 0.004, 0.013, 1 / 60, 1 / 30, 0.0, -1.0, 0.5)):
            t = 0.1 * k
            lib.ocean_update(sim, t)
            o.update(t)
            lib.ocean_amd_foam_update(foam, np.float32(dt))
            olib.orc_foam_update(of, np.float32(dt))
            assert lib.ocean_amd_foam_last_substeps(foam) == olib.orc_foam_last_substeps(of)
            fb = np.ctypeslib.as_array(lib.ocean_amd_foam_data(foam), shape=(64 * 64,)).copy()
            assert_bitexact(fb, as_array(olib.orc_foam_data(of), 64 * 64), f"foam k={k}")
        assert olib.orc_foam_last_substeps(of) == 4  # dt 0.5 s clamps at max_substeps
    finally:
        olib.orc_foam_destroy(of)
        lib.ocean_amd_foam_destroy(foam)
        lib.ocean_destroy(sim)


def test_config3_cascade_vs_golden_and_oracle(lib, golden):
    """All 1200 frames of the three cascades against the reference's digests
    (every level's displacement and normal and the 512 cascade queries, every
    frame: tests/golden/config3_frames.json), and the single-point cascade
    query against the oracle at the frames of the short fixture."""
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                           "config3_frames.json")) as f:
        every = json.load(f)["cascade"]
    assert len(every) == 1200
    levels = [desc_from(kw) for kw in CONFIG3_LEVELS]
    arr = (capi.OceanDesc * 3)(*levels)
    st = C.c_int(-1)
    h = lib.ocean_amd_cascade_create(arr, 3, C.byref(st))
    assert h, st.value
    assert lib.ocean_amd_cascade_level_count(h) == 3
    orcs = [OracleOcean(d) for d in levels]
    olib = oracle()
    harr = (C.c_void_p * 3)(*[o.h for o in orcs])
    pts = query_points(1, 512, 800.0)[0]
    try:
        flat = np.ascontiguousarray(pts.reshape(-1))
        for kf in range(1200):
            k = str(kf)
            t = kf / 60.0
            lib.ocean_amd_cascade_update(h, t)
            want = every[kf]
            if k in golden["config3_cascade"]:
                g = golden["config3_cascade"][k]
                assert [want[0:2], want[2:4], want[4:6]] == g["levels"] and want[6] == g["queries"]
            for li, d in enumerate(levels):
                b = lib.ocean_amd_cascade_get_buffers(h, li)
                nn = 4 * d.size * d.size
                disp = np.ctypeslib.as_array(b.displacement, shape=(nn,))
                nrm = np.ctypeslib.as_array(b.normal, shape=(nn,))
                assert [digest(disp), digest(nrm)] == want[2 * li:2 * li + 2], (k, li)
            out = (capi.OceanSurfaceEx * len(pts))()
            assert lib.ocean_amd_cascade_sample_batch(h, fptr(flat), len(pts), out) == 0
            q = np.frombuffer(bytes(out), np.float32).reshape(len(pts), 10)
            assert digest(q) == want[6], k
            if k not in golden["config3_cascade"]:
                continue
            # single-point form and the oracle
            for o in orcs:
                o.update(t)
            for i in range(0, len(pts), 97):
                s = capi.OceanSurfaceEx()
                lib.ocean_amd_cascade_sample(h, float(pts[i, 0]), float(pts[i, 1]), C.byref(s))
                w = capi.OceanSurfaceEx()
                olib.orc_cascade_sample(harr, 3, float(pts[i, 0]), float(pts[i, 1]), C.byref(w))
                assert bytes(s) == bytes(w)
    finally:
        lib.ocean_amd_cascade_destroy(h)


def test_cascade_rejects_empty_and_bad_levels(lib):
    st = C.c_int(-1)
    assert not lib.ocean_amd_cascade_create(None, 0, C.byref(st))
    assert st.value == capi.OCEAN_ERROR_INVALID_ARG
    levels = [desc_from(dict(size=64)), desc_from(dict(size=100))]
    arr = (capi.OceanDesc * 2)(*levels)
    assert not lib.ocean_amd_cascade_create(arr, 2, C.byref(st))
    assert st.value == capi.OCEAN_ERROR_INVALID_ARG


@pytest.mark.parametrize("q", [3000, 100_000])
def test_height_only_queries_equal_the_full_surfaces(lib, q):
    """Queries that ask for heights alone (no surface output) take the
    height taps without the normal and velocity loads; their heights equal,
    bit for bit, the full surface queries' heights at the same points,
    through ocean_amd_sample_batch_device and through the fused foam step
    with queries."""
    from oceanlib_amd.hipmem import DeviceArray
    desc = desc_from(CONFIG2)
    st = C.c_int(-1)
    sim = lib.ocean_create(C.byref(desc), C.byref(st))
    assert sim, st.value
    fd = capi.FoamDesc()
    lib.ocean_amd_foam_desc_init(C.byref(fd))
    foam = lib.ocean_amd_foam_create(sim, C.byref(fd), C.byref(st))
    assert foam, st.value
    try:
        xz = DeviceArray.from_numpy(query_points(1, q, desc.patch_length, seed=q).reshape(-1))
        out = DeviceArray((q, 10), np.float32)
        h_full, h_only, h_fused = (DeviceArray.from_numpy(np.full(q, np.nan, np.float32)) for _ in range(3))
        lib.ocean_update(sim, 3.25)
        assert lib.ocean_amd_sample_batch_device(sim, xz.ptr, q, out.ptr, h_full.ptr) == 0
        assert lib.ocean_amd_sample_batch_device(sim, xz.ptr, q, None, h_only.ptr) == 0
        assert lib.ocean_amd_foam_update_sample_device(foam, np.float32(1 / 60), xz.ptr, q, None,
                                                       h_fused.ptr) == 0
        lib.ocean_amd_synchronize(sim)
        assert lib.ocean_amd_last_error(sim) == 0
        full = h_full.numpy()
        assert not np.isnan(full).any()
        assert_bitexact(full, out.numpy()[:, 0].copy(), "heights vs surfaces")
        assert_bitexact(h_only.numpy(), full, "height-only batch query")
        assert_bitexact(h_fused.numpy(), full, "height-only fused foam step and queries")
    finally:
        lib.ocean_amd_foam_destroy(foam)
        lib.ocean_destroy(sim)


def test_fused_foam_step_and_queries_equal_the_two_calls(lib):
    """ocean_amd_foam_update_sample_device (the last foam substep and the
    frame's queries in one launch) gives the same foam field and the same
    query results, bit for bit, as ocean_amd_foam_update followed by
    ocean_amd_sample_batch_device, with 0, 1 and several substeps per call."""
    from oceanlib_amd.hipmem import DeviceArray
    desc = desc_from(CONFIG2)
    n = desc.size
    st = C.c_int(-1)
    sims = [lib.ocean_create(C.byref(desc), C.byref(st)) for _ in range(2)]
    assert all(sims)
    fd = capi.FoamDesc()
    lib.ocean_amd_foam_desc_init(C.byref(fd))
    foams = [lib.ocean_amd_foam_create(s, C.byref(fd), C.byref(st)) for s in sims]
    assert all(foams)
    q = 3000
    xz = DeviceArray.from_numpy(query_points(1, q, desc.patch_length).reshape(-1))
    outs = [DeviceArray((q, 10), np.float32) for _ in range(2)]
    hs = [DeviceArray((q,), np.float32) for _ in range(2)]
    dts = [1 / 60, 1 / 60, 0.0, 3 / 60, 1 / 120, 1 / 120, 1 / 60]
    try:
        t = 0.0
        for k in range(24):
            dt = np.float32(dts[k % len(dts)])
            t += float(dt)
            for s in sims:
                lib.ocean_update(s, t)
            lib.ocean_amd_foam_update(foams[0], dt)
            assert lib.ocean_amd_sample_batch_device(sims[0], xz.ptr, q, outs[0].ptr, hs[0].ptr) == 0
            assert lib.ocean_amd_foam_update_sample_device(foams[1], dt, xz.ptr, q, outs[1].ptr,
                                                           hs[1].ptr) == 0
            assert (lib.ocean_amd_foam_last_substeps(foams[0]) ==
                    lib.ocean_amd_foam_last_substeps(foams[1]))
            if k % 4 == 3:
                f0 = as_array(lib.ocean_amd_foam_data(foams[0]), n * n)
                f1 = as_array(lib.ocean_amd_foam_data(foams[1]), n * n)
                assert_bitexact(f1, f0, f"foam after frame {k}")
                assert_bitexact(outs[1].numpy().reshape(-1), outs[0].numpy().reshape(-1),
                                f"surfaces frame {k}")
                assert_bitexact(hs[1].numpy(), hs[0].numpy(), f"heights frame {k}")
        for s in sims:
            lib.ocean_amd_synchronize(s)
            assert lib.ocean_amd_last_error(s) == 0
        assert lib.ocean_amd_foam_update_sample_device(foams[1], np.float32(0.01), None, 5,
                                                       None, None) != 0
    finally:
        for f in foams:
            lib.ocean_amd_foam_destroy(f)
        for s in sims:
            lib.ocean_destroy(s)
