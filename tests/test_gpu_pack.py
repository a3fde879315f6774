"""GPU parity of the renderer handoff: RGBA16F packs made on the device.

The viewer converts each level's displacement (alpha replaced by the
persistent foam) and normal, and the interaction field, to RGBA16F before
upload (reference examples/viewer/ocean_view.cpp:818-870, half.hpp:58-130).
The device packs are compared bit for bit with the oracle's pack
(oracle/pack_oracle.c, pinned exhaustively in tests/test_pack_oracle.py) of
the same frame's buffers, and the device converter with the oracle on the
bit patterns where rounding decides: every exponent from half's underflow to
its overflow, ties, specials.
"""
import ctypes as C

import numpy as np
import pytest

from iwave_helpers import Field, idesc
from oceanlib_amd import capi
from oceanlib_amd.hipmem import DeviceArray
from oceanlib_amd.workloads import (CONFIG2, CONFIG3_INTERACTION, config3_disturbances,
                                    config3_sources, desc_from)
from oracle_ffi import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    return capi.load()


def orc_pack(src, alpha):
    o = oracle()
    o.orc_pack_rgba16.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]
    o.orc_pack_rgba16.restype = None
    cells = src.size // 4
    out = np.zeros(4 * cells, np.uint16)
    o.orc_pack_rgba16(out.ctypes.data, src.ctypes.data, None if alpha is None else alpha.ctypes.data,
                      cells)
    return out


def orc_halves(bits):
    o = oracle()
    f = np.ascontiguousarray(bits, np.uint32).view(np.float32)
    out = np.zeros(f.size, np.uint16)
    o.orc_pack_rgba16.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]
    o.orc_pack_rgba16(out.ctypes.data, f.ctypes.data, None, f.size // 4)
    return out


@pytest.fixture(scope="module")
def heavy_chop(lib):
    desc = desc_from(CONFIG2)
    st = C.c_int(-1)
    sim = lib.ocean_create(C.byref(desc), C.byref(st))
    assert sim, st.value
    fd = capi.FoamDesc()
    lib.ocean_amd_foam_desc_init(C.byref(fd))
    foam = lib.ocean_amd_foam_create(sim, C.byref(fd), C.byref(st))
    assert foam, st.value
    for k in range(40):
        lib.ocean_update(sim, k / 60.0)
        lib.ocean_amd_foam_update(foam, np.float32(1.0 / 60.0))
    lib.ocean_amd_synchronize(sim)
    assert lib.ocean_amd_last_error(sim) == 0
    yield desc, sim, foam
    lib.ocean_amd_foam_destroy(foam)
    lib.ocean_destroy(sim)


def host_buffers(lib, desc, sim, foam):
    n = desc.size
    b = lib.ocean_get_buffers(sim)
    disp = np.ctypeslib.as_array(b.displacement, shape=(4 * n * n,)).copy()
    nrm = np.ctypeslib.as_array(b.normal, shape=(4 * n * n,)).copy()
    fdata = np.ctypeslib.as_array(lib.ocean_amd_foam_data(foam), shape=(n * n,)).copy()
    return disp, nrm, fdata


def test_level_pack_host_and_device(lib, heavy_chop):
    desc, sim, foam = heavy_chop
    n = desc.size
    disp, nrm, fdata = host_buffers(lib, desc, sim, foam)
    want_d = orc_pack(disp, fdata)
    want_n = orc_pack(nrm, None)
    want_w = orc_pack(disp, None)
    assert np.any(want_d[3::4] != want_w[3::4])  # the foam alpha differs from disp.w

    got_d = np.zeros(4 * n * n, np.uint16)
    got_n = np.zeros(4 * n * n, np.uint16)
    assert lib.ocean_amd_pack_rgba16(sim, foam, got_d.ctypes.data, got_n.ctypes.data) == 0
    np.testing.assert_array_equal(got_d, want_d)
    np.testing.assert_array_equal(got_n, want_n)
    got_w = np.zeros(4 * n * n, np.uint16)
    assert lib.ocean_amd_pack_rgba16(sim, None, got_w.ctypes.data, None) == 0
    np.testing.assert_array_equal(got_w, want_w)

    dd = DeviceArray((4 * n * n,), np.uint16)
    dn = DeviceArray((4 * n * n,), np.uint16)
    assert lib.ocean_amd_pack_rgba16_device(sim, foam, dd.ptr, dn.ptr) == 0
    lib.ocean_amd_synchronize(sim)
    assert lib.ocean_amd_last_error(sim) == 0
    np.testing.assert_array_equal(dd.numpy(), want_d)
    np.testing.assert_array_equal(dn.numpy(), want_n)

    # a later frame: the pack follows the ocean's stream order
    lib.ocean_update(sim, 41 / 60.0)
    lib.ocean_amd_foam_update(foam, np.float32(1.0 / 60.0))
    assert lib.ocean_amd_pack_rgba16_device(sim, foam, dd.ptr, None) == 0
    lib.ocean_amd_synchronize(sim)
    assert lib.ocean_amd_last_error(sim) == 0
    disp2, _, fdata2 = host_buffers(lib, desc, sim, foam)
    np.testing.assert_array_equal(dd.numpy(), orc_pack(disp2, fdata2))


def test_pack_rejects_bad_arguments(lib, heavy_chop):
    desc, sim, foam = heavy_chop
    n = desc.size
    out = np.zeros(4 * n * n + 4, np.uint16)
    assert lib.ocean_amd_pack_rgba16(sim, foam, None, None) != 0
    assert lib.ocean_amd_pack_rgba16(sim, foam, out.ctypes.data + 2, None) != 0  # misaligned
    other = desc_from(dict(CONFIG2, seed=8))
    st = C.c_int(-1)
    sim2 = lib.ocean_create(C.byref(other), C.byref(st))
    assert sim2
    try:
        # the foam follows another ocean
        assert lib.ocean_amd_pack_rgba16(sim2, foam, out.ctypes.data, None) != 0
    finally:
        lib.ocean_destroy(sim2)


def test_converter_on_rounding_critical_patterns(lib, heavy_chop):
    _, sim, _ = heavy_chop
    rng = np.random.default_rng(12)
    parts = []
    for e32 in range(96, 148):  # half underflow .. overflow
        m = rng.integers(0, 1 << 23, 1 << 15, dtype=np.uint32)
        m[:16] = np.arange(16, dtype=np.uint32) << 9   # low mantissa bits zero
        m[16:32] = 0x1000 + np.arange(-8, 8, dtype=np.int64).astype(np.uint32)  # around the tie
        parts.append((np.uint32(e32) << 23) | m)
    parts.append(rng.integers(0, 1 << 32, 1 << 18, dtype=np.uint64).astype(np.uint32))
    parts.append(np.array([0, 0x7F800000, 0x7FC00000, 0x7F800001, 0x477FF000, 0x477FE000,
                           0x33000000, 0x33800000, 0x387FC000, 0x38800000], np.uint32))
    bits = np.concatenate(parts)
    bits = np.concatenate([bits, bits | np.uint32(0x80000000)])
    bits = np.concatenate([bits, np.zeros((-bits.size) % 4, np.uint32)])
    src = DeviceArray.from_numpy(bits.view(np.float32))
    dst = DeviceArray((bits.size,), np.uint16)
    assert lib.ocean_amd_pack_half_device(sim, src.ptr, dst.ptr, bits.size) == 0
    lib.ocean_amd_synchronize(sim)
    assert lib.ocean_amd_last_error(sim) == 0
    np.testing.assert_array_equal(dst.numpy(), orc_halves(bits))
    # a count that is not a multiple of four takes the tail path
    tail = DeviceArray((7,), np.uint16)
    assert lib.ocean_amd_pack_half_device(sim, src.ptr, tail.ptr, 7) == 0
    lib.ocean_amd_synchronize(sim)
    assert lib.ocean_amd_last_error(sim) == 0
    np.testing.assert_array_equal(tail.numpy(), orc_halves(bits[:8])[:7])


def test_interaction_field_pack(lib):
    d = idesc(lib.ocean_interaction_desc_init, **dict(CONFIG3_INTERACTION, size=128, extent=12.8))
    g = Field("gpu", lib, d)
    try:
        srcs = config3_sources()
        for k in range(30):
            for rec in config3_disturbances(srcs, k / 60.0, centre=(6.4, 6.4)):
                g.add(rec)
            g.update(1.0 / 60.0)
        field = g.field()
        want = orc_pack(field, None)
        got = np.zeros(field.size, np.uint16)
        assert lib.ocean_amd_interaction_pack_rgba16(g.h, got.ctypes.data) == 0
        np.testing.assert_array_equal(got, want)
        dev = DeviceArray((field.size,), np.uint16)
        assert lib.ocean_amd_interaction_pack_rgba16_device(g.h, dev.ptr) == 0
        np.testing.assert_array_equal(dev.numpy(), want)  # numpy() waits for the device
    finally:
        g.close()


def test_pack_of_batch_levels(lib):
    """The levels of a batch are oceans too: their packs equal the oracle's
    pack of their buffers."""
    descs = [desc_from(dict(size=n, patch_length=L, wind_speed=11.0, seed=s))
             for n, L, s in ((128, 90.0, 5), (256, 300.0, 6))]
    arr = (capi.OceanDesc * 2)(*descs)
    st = C.c_int(-1)
    b = lib.ocean_amd_batch_create(arr, 2, C.byref(st))
    assert b, st.value
    try:
        lib.ocean_amd_batch_update(b, 7.5)
        for i, d in enumerate(descs):
            lvl = lib.ocean_amd_batch_level(b, i)
            n = d.size
            bufs = lib.ocean_get_buffers(lvl)
            disp = np.ctypeslib.as_array(bufs.displacement, shape=(4 * n * n,)).copy()
            nrm = np.ctypeslib.as_array(bufs.normal, shape=(4 * n * n,)).copy()
            got_d = np.zeros(4 * n * n, np.uint16)
            got_n = np.zeros(4 * n * n, np.uint16)
            assert lib.ocean_amd_pack_rgba16(lvl, None, got_d.ctypes.data, got_n.ctypes.data) == 0
            np.testing.assert_array_equal(got_d, orc_pack(disp, None))
            np.testing.assert_array_equal(got_n, orc_pack(nrm, None))
    finally:
        assert lib.ocean_amd_batch_synchronize(b) == 0
        lib.ocean_amd_batch_destroy(b)
