"""GPU parity for the iWave interaction field and the combined query.

The field through the reference's C ABI (ocean_interaction_*) against the
oracle bit for bit, and against the reference run's digests for the
three-cascade scene (256^2 cells of 0.1 m, kernel radius 6, eight moving
continuous sources, a Neumann hull, 1200 frames at 1/60 s).
"""
import ctypes as C
import hashlib
import json
import os

import numpy as np
import pytest

from oceanlib_amd import capi
from oceanlib_amd.workloads import (CONFIG3_INTERACTION, config3_disturbances, config3_hull_mask,
                                    config3_sources, desc_from)
from iwave_helpers import Field, idesc
from oracle_ffi import OracleOcean, assert_bitexact, oracle

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def digest(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()[:16]


@pytest.fixture(scope="module")
def lib():
    return capi.load()


def test_config3_interaction_scene_vs_golden_and_oracle(lib):
    """The scene's field after every one of its 1200 substeps against the
    reference's digests (tests/golden/config3_frames.json), and bit for bit
    against the oracle, energy included, every 200th."""
    with open(os.path.join(GOLDEN, "digests.json")) as f:
        sparse = json.load(f)["config3_interaction"]
    with open(os.path.join(GOLDEN, "config3_frames.json")) as f:
        every = json.load(f)["interaction"]
    assert len(every) == 1200 and all(every[int(k)] == v for k, v in sparse.items())
    want = {str(k): v for k, v in enumerate(every)}
    d = idesc(lib.ocean_interaction_desc_init, **CONFIG3_INTERACTION)
    g = Field("gpu", lib, d)
    o = Field("orc", oracle(), d)
    try:
        assert g.fixed_dt() == o.fixed_dt() and g.dt_limit() == o.dt_limit()
        g.set_obstruction(config3_hull_mask())
        o.set_obstruction(config3_hull_mask())
        srcs = config3_sources()
        for k in range(1200):
            for rec in config3_disturbances(srcs, k / 60.0):
                g.add(rec)
                o.add(rec)
            g.update(1.0 / 60.0)
            o.update(1.0 / 60.0)
            if str(k) in want:
                fg = g.field()
                assert digest(fg) == want[str(k)], f"frame {k} vs reference digest"
                if k % 200 == 0:
                    assert_bitexact(fg, o.field(), f"frame {k}")
                    # energy(): the device's sequential sums in the reference's order
                    assert g.energy() == o.energy(), f"energy, frame {k}"
        rng = np.random.default_rng(9)
        for x, z in rng.uniform(-3, 29, size=(300, 2)).astype(np.float32):
            assert bytes(g.sample(x, z)) == bytes(o.sample(x, z)), (x, z)
            h = lib.ocean_interaction_height_at(g.h, np.float32(x), np.float32(z))
            assert np.float32(h).tobytes() == np.float32(o.sample(x, z).height).tobytes()
        assert lib.ocean_interaction_dispersion_error(g.h) > 0.0
    finally:
        g.close()
        o.close()


def test_dirichlet_recentre_impulses_clamp_vs_oracle(lib):
    imp = {5: [dict(world_x=np.float32(12.0), world_z=np.float32(9.0), radius=np.float32(0.3),
                    strength=np.float32(0.2), velocity_y=np.float32(-1.5), velocity_x=0.0,
                    velocity_z=0.0, kind=0)],
           31: [dict(world_x=np.float32(3.0), world_z=np.float32(20.0), radius=np.float32(0.01),
                     strength=np.float32(0.1), velocity_y=0.0, velocity_x=0.0, velocity_z=0.0,
                     kind=0)],
           44: [dict(world_x=np.float32(8.0), world_z=np.float32(8.0), radius=np.float32(6.0),
                     strength=np.float32(0.05), velocity_y=0.0, velocity_x=0.0, velocity_z=0.0,
                     kind=0)]}
    kw = dict(CONFIG3_INTERACTION, size=128, extent=16.0, kernel_radius=4, obstruction=1,
              max_sources=6, absorb_cells=8)
    d = idesc(lib.ocean_interaction_desc_init, **kw)
    g, o = Field("gpu", lib, d), Field("orc", oracle(), d)
    dts = [1 / 60, 0.0, 1 / 60, 0.5, 0.004, 1 / 30]
    recenter = {20: (9.0, 7.5), 50: (9.05, 7.5), 70: (200.0, 0.0), 80: (190.0, 3.0)}
    try:
        g.set_obstruction(config3_hull_mask(128))
        o.set_obstruction(config3_hull_mask(128))
        srcs = config3_sources()
        for k in range(90):
            for rec in config3_disturbances(srcs, k / 60.0) + imp.get(k, []):
                g.add(rec)
                o.add(rec)
            if k in recenter:
                assert g.recenter(*recenter[k]) == o.recenter(*recenter[k])
                assert_bitexact(g.field(), o.field(), f"after recentre {k}")
            dt = dts[k % len(dts)]
            g.update(dt)
            o.update(dt)
            assert lib.ocean_interaction_dropped(g.h) == oracle().orc_iw_dropped(o.h)
            if k % 3 == 0 or k == 89:
                assert_bitexact(g.field(), o.field(), f"frame {k}")
                assert g.energy() == o.energy(), f"energy, frame {k}"
        b = lib.ocean_interaction_get_buffers(g.h)
        assert b.origin_x == oracle().orc_iw_origin_x(o.h)
        assert b.origin_z == oracle().orc_iw_origin_z(o.h)
        g.set_obstruction(None)
        o.set_obstruction(None)
        for k in range(5):
            g.update(1 / 60)
            o.update(1 / 60)
        assert_bitexact(g.field(), o.field(), "mask cleared")
    finally:
        g.close()
        o.close()


def test_many_stamps_of_every_size_vs_oracle(lib):
    """Each stamp's mean is the reference's sequential double sum over its
    box (src/interaction.cpp:345-356); the device takes proven rounding-free
    chunks of it in one step (oceanlib_amd/csrc/exact_sum.hpp). Sixty-four
    sources a frame of radii 0.02-3 m at 0.1 m cells (boxes from 1 cell to
    clipped 241^2, the largest past the LDS staging), anywhere on and off the
    grid, impulses and continuous: the field equals the oracle's every frame."""
    rng = np.random.default_rng(17)
    kw = dict(CONFIG3_INTERACTION, max_sources=64)
    d = idesc(lib.ocean_interaction_desc_init, **kw)
    g, o = Field("gpu", lib, d), Field("orc", oracle(), d)
    try:
        for k in range(8):
            for _ in range(64):
                rec = dict(world_x=np.float32(rng.uniform(-4.0, 29.6)),
                           world_z=np.float32(rng.uniform(-4.0, 29.6)),
                           radius=np.float32(np.exp(rng.uniform(np.log(0.02), np.log(3.0)))),
                           strength=np.float32(rng.uniform(-0.3, 0.3)),
                           velocity_y=np.float32(rng.uniform(-2, 2)),
                           velocity_x=np.float32(rng.uniform(-3, 3)),
                           velocity_z=np.float32(rng.uniform(-3, 3)), kind=int(rng.integers(0, 2)))
                g.add(rec)
                o.add(rec)
            g.update(1 / 60)
            o.update(1 / 60)
            assert_bitexact(g.field(), o.field(), f"frame {k}")
    finally:
        g.close()
        o.close()


def test_large_stamps_vs_oracle(lib):
    """A 1024^2 field at 0.1 m: frames of 32 and of 8 sources of radius 5-6 m,
    away from the edges (unclipped boxes of 401^2-481^2 cells, five windows of
    the stamp mean's statistics each; the stamps ride in the kernel
    arguments), and of 40 (more stamps than the arguments carry, so
    uploaded): the field and energy() equal the oracle's every frame."""
    rng = np.random.default_rng(23)
    kw = dict(CONFIG3_INTERACTION, size=1024, extent=102.4, max_sources=40)
    d = idesc(lib.ocean_interaction_desc_init, **kw)
    g, o = Field("gpu", lib, d), Field("orc", oracle(), d)
    try:
        for k, count in enumerate((32, 32, 8, 40, 8)):
            for _ in range(count):
                rec = dict(world_x=np.float32(rng.uniform(30.0, 72.4)),
                           world_z=np.float32(rng.uniform(30.0, 72.4)),
                           radius=np.float32(rng.uniform(5.0, 6.0)),
                           strength=np.float32(rng.uniform(-0.05, 0.05)),
                           velocity_y=np.float32(rng.uniform(-1, 1)),
                           velocity_x=np.float32(rng.uniform(-3, 3)),
                           velocity_z=np.float32(rng.uniform(-3, 3)), kind=int(rng.integers(0, 2)))
                g.add(rec)
                o.add(rec)
            g.update(1 / 60)
            o.update(1 / 60)
            assert_bitexact(g.field(), o.field(), f"frame {k} ({count} sources)")
            assert g.energy() == o.energy(), f"energy, frame {k}"
    finally:
        g.close()
        o.close()


def test_random_blob_mask_neumann_vs_oracle(lib):
    """Neumann walls on a mask of many concave blobs (about a third of the
    cells solid, solid cells several cells from any fluid, nearest-fluid ties
    everywhere) and an all-solid border row: the nearest-fluid table
    (reference src/interaction.cpp:255-292) decides every solid cell's value,
    so the field equals the oracle's only if its BFS order is the
    reference's. Recentring shifts the mask; clearing it restores open
    water."""
    rng = np.random.default_rng(31)
    n = 256
    noise = rng.standard_normal((n // 8, n // 8))
    blob = np.kron(noise, np.ones((8, 8)))
    blob += 0.6 * rng.standard_normal((n, n))
    mask = (blob > 0.45).astype(np.uint8)
    mask[0, :] = 1
    d = idesc(lib.ocean_interaction_desc_init, **dict(CONFIG3_INTERACTION, max_sources=16))
    g, o = Field("gpu", lib, d), Field("orc", oracle(), d)
    try:
        g.set_obstruction(mask)
        o.set_obstruction(mask)
        srcs = config3_sources()
        for k in range(24):
            for rec in config3_disturbances(srcs, k / 60.0):
                g.add(rec)
                o.add(rec)
            if k == 10:
                assert g.recenter(14.1, 11.9) == o.recenter(14.1, 11.9)
            if k == 16:
                g.set_obstruction(mask[::-1].copy())
                o.set_obstruction(mask[::-1].copy())
            g.update(1 / 60)
            o.update(1 / 60)
            assert_bitexact(g.field(), o.field(), f"frame {k}")
        g.set_obstruction(None)
        o.set_obstruction(None)
        g.update(1 / 60)
        o.update(1 / 60)
        assert_bitexact(g.field(), o.field(), "mask cleared")
    finally:
        g.close()
        o.close()


def test_combined_query_vs_oracle(lib):
    """ocean_sample_combined (WaterSurface, reference src/interaction.cpp:863-901)."""
    od = desc_from(dict(size=128, patch_length=64.0, wind_speed=14.0, choppiness=1.2,
                        compute_velocity=1, seed=5))
    st = C.c_int(-1)
    sim = lib.ocean_create(C.byref(od), C.byref(st))
    assert sim
    oo = OracleOcean(od)
    d = idesc(lib.ocean_interaction_desc_init, **dict(CONFIG3_INTERACTION, size=128, extent=12.8))
    g, o = Field("gpu", lib, d), Field("orc", oracle(), d)
    try:
        srcs = config3_sources()
        for k in range(30):
            for rec in config3_disturbances(srcs, k / 60.0, centre=(6.4, 6.4)):
                g.add(rec)
                o.add(rec)
            g.update(1 / 60)
            o.update(1 / 60)
        lib.ocean_update(sim, 0.5)
        oo.update(0.5)
        rng = np.random.default_rng(3)
        for x, z in rng.uniform(-2, 15, size=(200, 2)).astype(np.float32):
            got = capi.OceanSurfaceEx()
            lib.ocean_sample_combined(sim, g.h, np.float32(x), np.float32(z), C.byref(got))
            want = oo.sample(float(x), float(z))
            e = o.sample(x, z)
            oracle().orc_compose(C.byref(want), C.byref(e))
            assert bytes(got) == bytes(want), (x, z)
        # NULL field -> plain ocean sample; NULL ocean -> field only
        got = capi.OceanSurfaceEx()
        lib.ocean_sample_combined(sim, None, np.float32(1.0), np.float32(2.0), C.byref(got))
        assert bytes(got) == bytes(oo.sample(1.0, 2.0))
        lib.ocean_sample_combined(None, g.h, np.float32(6.0), np.float32(6.0), C.byref(got))
        e = o.sample(6.0, 6.0)
        assert got.height == e.height and got.velocity_y == e.velocity_y and got.normal_y == 1.0
    finally:
        g.close()
        o.close()
        lib.ocean_destroy(sim)


def test_validation_and_null_handles(lib):
    for kw in (dict(size=48), dict(kernel_radius=13), dict(absorb_cells=200),
               dict(fixed_dt=10.0), dict(gravity=0.0)):
        d = idesc(lib.ocean_interaction_desc_init, **kw)
        st = C.c_int(-1)
        assert not lib.ocean_interaction_create(C.byref(d), C.byref(st)), kw
        assert st.value == capi.OCEAN_ERROR_INVALID_ARG
    lib.ocean_interaction_destroy(None)
    lib.ocean_interaction_update(None, 1.0)
    assert lib.ocean_interaction_recenter(None, 1.0, 1.0) == 0
    assert lib.ocean_interaction_dropped(None) == 0
    assert lib.ocean_interaction_height_at(None, 1.0, 1.0) == 0.0
