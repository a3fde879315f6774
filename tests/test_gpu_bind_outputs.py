"""Zero-copy outputs: ocean_amd_bind_outputs redirects a standalone ocean's
RGBA32F outputs into caller-owned device memory (what a renderer imports from
its graphics API), and every reader of the ocean follows them.

The bound buffers must hold the oracle's frame bit for bit (reference
src/core/evolve.cpp:186-327, the same bits ocean_get_buffers returns); the
foam step and the queries of a bound ocean must equal an unbound twin's; the
binding is refused, and left as it was, for pointers that are not device
memory of the right size and alignment, velocity without compute_velocity, and
the levels of a batch.
"""
import ctypes as C

import numpy as np
import pytest

from oceanlib_amd import capi
from oceanlib_amd.hipmem import DeviceArray
from oceanlib_amd.workloads import desc_from, query_points
from oracle_ffi import OracleOcean, as_array, assert_bitexact

pytestmark = pytest.mark.gpu

KW = dict(size=256, patch_length=180.0, wind_speed=17.0, choppiness=1.2, compute_velocity=1,
          water_depth=40.0, seed=21)


@pytest.fixture(scope="module")
def lib():
    return capi.load()


def create(lib, desc):
    st = C.c_int(-1)
    sim = lib.ocean_create(C.byref(desc), C.byref(st))
    assert sim, st.value
    return sim


def host_outputs(lib, sim, n):
    b = lib.ocean_get_buffers(sim)
    d = np.ctypeslib.as_array(b.displacement, shape=(4 * n * n,)).copy()
    nr = np.ctypeslib.as_array(b.normal, shape=(4 * n * n,)).copy()
    v = np.ctypeslib.as_array(lib.ocean_get_velocity(sim), shape=(4 * n * n,)).copy()
    return d, nr, v


def test_bound_buffers_hold_the_frame_and_readers_follow(lib):
    desc = desc_from(KW)
    n = desc.size
    sim, twin = create(lib, desc), create(lib, desc)
    o = OracleOcean(desc)
    bufs = [DeviceArray((4 * n * n,), np.float32) for _ in range(3)]
    fd = capi.FoamDesc()
    lib.ocean_amd_foam_desc_init(C.byref(fd))
    st = C.c_int(-1)
    foams = [lib.ocean_amd_foam_create(s, C.byref(fd), C.byref(st)) for s in (sim, twin)]
    assert all(foams)
    q = 2048
    xz = DeviceArray.from_numpy(query_points(1, q, desc.patch_length).reshape(-1))
    outs = [DeviceArray((q, 10), np.float32) for _ in range(2)]
    hs = [DeviceArray((q,), np.float32) for _ in range(2)]
    try:
        for s in (sim, twin):
            lib.ocean_update(s, 3.5)
        assert lib.ocean_amd_bind_outputs(sim, bufs[0].ptr, bufs[1].ptr, bufs[2].ptr) == 0
        o.update(3.5)
        # the current frame was re-evaluated into the bound memory
        for dev, want, what in zip(bufs, (o.displacement(), o.normal(), o.velocity()),
                                   ("displacement", "normal", "velocity")):
            assert_bitexact(dev.numpy(), want, f"bound {what} at bind")
        db = capi.DeviceBuffers()
        assert lib.ocean_amd_get_device_buffers(sim, C.byref(db)) == 0
        assert (db.displacement, db.normal, db.velocity) == tuple(b.ptr.value for b in bufs)

        dt = np.float32(1.0 / 60.0)
        for k in range(1, 13):
            t = 3.5 + k / 60.0
            for s in (sim, twin):
                lib.ocean_update(s, t)
            for i, f in enumerate(foams):
                assert lib.ocean_amd_foam_update_sample_device(f, dt, xz.ptr, q, outs[i].ptr,
                                                               hs[i].ptr) == 0
            if k % 4 == 0:
                o.update(t)
                for dev, want, what in zip(bufs, (o.displacement(), o.normal(), o.velocity()),
                                           ("displacement", "normal", "velocity")):
                    assert_bitexact(dev.numpy(), want, f"bound {what} frame {k}")
                for a, b, what in zip(host_outputs(lib, sim, n), host_outputs(lib, twin, n),
                                      ("displacement", "normal", "velocity")):
                    assert_bitexact(a, b, f"host mirror {what} frame {k}")
                assert_bitexact(outs[0].numpy().reshape(-1), outs[1].numpy().reshape(-1),
                                f"surfaces frame {k}")
                assert_bitexact(hs[0].numpy(), hs[1].numpy(), f"heights frame {k}")
                f0 = as_array(lib.ocean_amd_foam_data(foams[0]), n * n)
                f1 = as_array(lib.ocean_amd_foam_data(foams[1]), n * n)
                assert_bitexact(f0, f1, f"foam frame {k}")
        h16 = [np.zeros(4 * n * n, np.uint16) for _ in range(2)]
        for s, f, h in zip((sim, twin), foams, h16):
            assert lib.ocean_amd_pack_rgba16(s, f, h.ctypes.data, None) == 0
        np.testing.assert_array_equal(h16[0], h16[1])
        x, z = 17.25, -40.5
        assert lib.ocean_height_at(sim, x, z) == lib.ocean_height_at(twin, x, z)

        # unbinding: the ocean's own buffers again, the caller's memory left alone
        kept = bufs[0].numpy()
        assert lib.ocean_amd_bind_outputs(sim, None, None, None) == 0
        t = 9.0
        for s in (sim, twin):
            lib.ocean_update(s, t)
        o.update(t)
        d, nr, v = host_outputs(lib, sim, n)
        assert_bitexact(d, o.displacement(), "unbound displacement")
        assert_bitexact(nr, o.normal(), "unbound normal")
        assert_bitexact(v, o.velocity(), "unbound velocity")
        assert_bitexact(bufs[0].numpy(), kept, "caller memory after unbinding")
        for s in (sim, twin):
            lib.ocean_amd_synchronize(s)
            assert lib.ocean_amd_last_error(s) == 0
    finally:
        for f in foams:
            lib.ocean_amd_foam_destroy(f)
        lib.ocean_destroy(sim)
        lib.ocean_destroy(twin)


def test_partial_binding_and_refusals(lib):
    desc = desc_from(KW)
    n = desc.size
    sim = create(lib, desc)
    o = OracleOcean(desc)
    big = DeviceArray((4 * n * n + 8,), np.float32)
    small = DeviceArray((4 * n * n - 4,), np.float32)
    try:
        # normal only; the rest stays in the ocean's own buffers
        assert lib.ocean_amd_bind_outputs(sim, None, big.ptr, None) == 0
        lib.ocean_update(sim, 2.0)
        o.update(2.0)
        assert_bitexact(big.numpy()[:4 * n * n], o.normal(), "bound normal")
        d, nr, v = host_outputs(lib, sim, n)
        assert_bitexact(d, o.displacement(), "own displacement")
        assert_bitexact(nr, o.normal(), "mirror of the bound normal")
        db = capi.DeviceBuffers()
        assert lib.ocean_amd_get_device_buffers(sim, C.byref(db)) == 0
        before = (db.displacement, db.normal)
        assert before[1] == big.ptr.value

        offset = C.c_void_p(big.ptr.value + 4)       # not 16-byte aligned
        tail = C.c_void_p(big.ptr.value + 64)        # aligned, 8 floats short of N*N float4
        for args in ((small.ptr, None, None), (offset, None, None), (tail, None, None),
                     (C.c_void_p(0x1000), None, None)):
            assert lib.ocean_amd_bind_outputs(sim, *args) == capi.OCEAN_ERROR_INVALID_ARG, args
        assert lib.ocean_amd_bind_outputs(None, big.ptr, None, None) == capi.OCEAN_ERROR_INVALID_ARG
        # a refusal leaves the binding as it was
        assert lib.ocean_amd_get_device_buffers(sim, C.byref(db)) == 0
        assert (db.displacement, db.normal) == before
        assert lib.ocean_amd_last_error(sim) == 0

        plain = create(lib, desc_from(dict(KW, compute_velocity=0)))
        try:
            assert lib.ocean_amd_bind_outputs(plain, None, None, big.ptr) == \
                capi.OCEAN_ERROR_INVALID_ARG
        finally:
            lib.ocean_destroy(plain)
    finally:
        lib.ocean_destroy(sim)

    levels = (capi.OceanDesc * 2)(desc_from(dict(size=64, seed=1)), desc_from(dict(size=64, seed=2)))
    st = C.c_int(-1)
    b = lib.ocean_amd_batch_create(levels, 2, C.byref(st))
    assert b, st.value
    try:
        lvl = lib.ocean_amd_batch_level(b, 1)
        assert lib.ocean_amd_bind_outputs(lvl, big.ptr, None, None) == capi.OCEAN_ERROR_INVALID_ARG
    finally:
        lib.ocean_amd_batch_destroy(b)
