"""Pins the iWave part of the oracle (oracle/iwave_oracle.c) to the reference.

Bit-for-bit against the reference's own InteractionField compiled into
oracle/_ref (reference src/interaction.cpp, src/core/iwave_*.cpp): kernel
taps from the least-squares fit, fields after every checked frame of the
moving-source scene with a Neumann or Dirichlet hull, and the scheme's
energy() at each checked frame, recentring, impulses kept over zero-length
frames, the substep clamp and queue overflow. The
expf the Ricker stamp depends on is the host libm's in both.
"""
import ctypes as C

import numpy as np
import pytest

from oceanlib_amd.workloads import (CONFIG3_INTERACTION, config3_disturbances, config3_hull_mask,
                                    config3_sources)
from iwave_helpers import Field, idesc
from oracle_ffi import assert_bitexact, fptr, oracle, ref

needs_ref = pytest.mark.skipif(ref() is None, reason="reference build oracle/_ref not present")


@needs_ref
@pytest.mark.parametrize("radius,cell,depth", [(6, 0.1, 0.0), (1, 0.25, 0.0), (4, 0.25, 2.0),
                                               (12, 0.05, 0.0), (8, 0.5, 0.3)])
def test_kernel_taps_vs_reference(radius, cell, depth):
    w = 2 * radius + 1
    a = np.zeros(w * w, np.float32)
    b = np.zeros(w * w, np.float32)
    assert oracle().orc_iwave_taps(radius, cell, depth, fptr(a), None) == w * w
    assert ref().oref_iwave_taps(radius, cell, depth, fptr(b)) == w * w
    assert_bitexact(a, b, "taps")
    t = a.reshape(w, w)
    # D4 symmetry and zero sum (reference tests/test_iwave_kernel.cpp)
    assert np.array_equal(t, t.T) and np.array_equal(t, t[::-1]) and np.array_equal(t, t[:, ::-1])
    assert abs(float(t.astype(np.float64).sum())) < 1e-5


def run_pair(kind_a, lib_a, kind_b, lib_b, desc_kw, frames, mask=None, recenter_at=None,
             impulses=None, check_every=10, dts=None):
    init = ref().oref_interaction_desc_init
    d = idesc(init, **desc_kw)
    fa, fb = Field(kind_a, lib_a, d), Field(kind_b, lib_b, d)
    try:
        assert fa.fixed_dt() == fb.fixed_dt() and fa.dt_limit() == fb.dt_limit()
        if mask is not None:
            fa.set_obstruction(mask)
            fb.set_obstruction(mask)
        srcs = config3_sources()
        for k in range(frames):
            t = k / 60.0
            for rec in config3_disturbances(srcs, t):
                fa.add(rec)
                fb.add(rec)
            if impulses and k in impulses:
                for rec in impulses[k]:
                    fa.add(rec)
                    fb.add(rec)
            if recenter_at and k in recenter_at:
                x, z = recenter_at[k]
                assert fa.recenter(x, z) == fb.recenter(x, z)
            dt = dts[k % len(dts)] if dts else 1.0 / 60.0
            fa.update(dt)
            fb.update(dt)
            if k % check_every == 0 or k == frames - 1:
                assert_bitexact(fa.field(), fb.field(), f"frame {k}")
                assert fa.energy() == fb.energy(), f"energy, frame {k}"
        rng = np.random.default_rng(4)
        for x, z in rng.uniform(-2, 28, size=(200, 2)).astype(np.float32):
            assert bytes(fa.sample(x, z)) == bytes(fb.sample(x, z))
        assert fa.dropped() == fb.dropped()
    finally:
        fa.close()
        fb.close()


@needs_ref
def test_moving_sources_with_neumann_hull_vs_reference():
    run_pair("orc", oracle(), "oref", ref(), CONFIG3_INTERACTION, 240,
             mask=config3_hull_mask())


@needs_ref
def test_dirichlet_recentre_impulses_clamp_vs_reference():
    imp = {5: [dict(world_x=np.float32(12.0), world_z=np.float32(9.0), radius=np.float32(0.3),
                    strength=np.float32(0.2), velocity_y=np.float32(-1.5), velocity_x=0.0,
                    velocity_z=0.0, kind=0)],
           31: [dict(world_x=np.float32(3.0), world_z=np.float32(20.0), radius=np.float32(0.01),
                     strength=np.float32(0.1), velocity_y=0.0, velocity_x=0.0, velocity_z=0.0,
                     kind=0)]}
    kw = dict(CONFIG3_INTERACTION, size=128, extent=16.0, kernel_radius=4, obstruction=1,
              max_sources=6, absorb_cells=8)
    # dt pattern: normal, zero (impulses must survive), a hitch that clamps, short
    run_pair("orc", oracle(), "oref", ref(), kw, 90, mask=config3_hull_mask(128),
             recenter_at={20: (9.0, 7.5), 50: (9.05, 7.5), 70: (200.0, 0.0)}, impulses=imp,
             dts=[1 / 60, 0.0, 1 / 60, 0.5, 0.004, 1 / 30], check_every=3)


@needs_ref
def test_desc_validation_matches_reference():
    init = ref().oref_interaction_desc_init
    bad = [dict(size=48), dict(size=16), dict(size=2048), dict(extent=0.0),
           dict(kernel_radius=0), dict(kernel_radius=13), dict(size=32, kernel_radius=12),
           dict(damping=-1.0), dict(absorb_cells=200), dict(max_substeps=0),
           dict(max_sources=0), dict(gravity=0.0), dict(fixed_dt=-0.1), dict(fixed_dt=10.0)]
    import ctypes as C
    for kw in bad:
        d = idesc(init, **kw)
        st = C.c_int(-1)
        assert not ref().oref_interaction_create(C.byref(d), C.byref(st)), kw
        assert not oracle().orc_iw_create(C.byref(d)), kw


STENCIL_CASES = [(6, 0.1, 0.0), (1, 0.25, 0.0), (4, 0.25, 2.0), (12, 0.05, 0.0), (8, 0.5, 0.3),
                 (2, 1.0, 0.0), (3, 0.1, 1.5), (10, 0.02, 5.0), (5, 2.0, 0.5), (7, 0.125, 0.0)]


@needs_ref
@pytest.mark.parametrize("radius,cell,depth", STENCIL_CASES)
def test_stencil_summary_vs_reference(radius, cell, depth):
    """The fit's taps and its scan summary (max / min realised symbol, band
    edge, peak relative error: reference src/core/iwave_kernel.cpp:400-436),
    oracle against the reference build, bit for bit."""
    w = 2 * radius + 1
    a, b = np.zeros(w * w, np.float32), np.zeros(w * w, np.float32)
    sa, sb = np.zeros(4), np.zeros(4)
    dp = lambda v: v.ctypes.data_as(C.POINTER(C.c_double))
    assert oracle().orc_iwave_stencil(radius, cell, depth, fptr(a), dp(sa)) == w * w
    assert ref().oref_iwave_stencil(radius, cell, depth, fptr(b), dp(sb)) == w * w
    assert_bitexact(a, b, "taps")
    assert sa.tobytes() == sb.tobytes(), (sa, sb)
    assert sa[0] > 0 and sa[1] <= 0 and sa[3] > 0


def nearest_closed_form(mask):
    """The fluid cell at the smallest L1 distance, the smallest index among
    equals (brute force over all fluid cells); 0xFFFFFFFF without fluid."""
    n = mask.shape[0]
    fz, fx = np.nonzero(mask == 0)
    out = np.full(n * n, 0xFFFFFFFF, np.uint32)
    if fz.size == 0:
        return out
    fidx = fz.astype(np.int64) * n + fx
    for z in range(n):
        for x in range(n):
            d = np.abs(fx - x) + np.abs(fz - z)
            best = d.min()
            out[z * n + x] = fidx[d == best].min()
    return out


def test_nearest_fluid_bfs_equals_closed_form():
    """The reference's nearest-fluid BFS (src/interaction.cpp:255-292, restated
    as orc_nearest_fluid) assigns every cell the fluid cell at the smallest L1
    distance with the smallest index among equals: the closed form the device
    computes (iwave_init_kernels.hip). Checked by brute force on random masks
    of every density, sizes 1..40, blobs, lone fluid cells, no fluid."""
    rng = np.random.default_rng(11)
    cases = []
    for t in range(160):
        n = int(rng.integers(1, 41))
        p = float(rng.random())
        cases.append((rng.random((n, n)) < p).astype(np.uint8))
    for n in (7, 16, 33):
        m = np.ones((n, n), np.uint8)
        cases.append(m.copy())                       # no fluid at all
        m[n // 2, n // 3] = 0
        cases.append(m.copy())                       # one fluid cell
        m[n - 1, 0] = 0
        m[0, n - 1] = 0
        cases.append(m.copy())                       # equidistant fluid cells
        zz, xx = np.mgrid[0:n, 0:n]
        cases.append(((xx - n / 2) ** 2 + (zz - n / 2) ** 2 < (n / 3) ** 2).astype(np.uint8))
        cases.append(((xx + zz) % 3 == 0).astype(np.uint8))
    for mask in cases:
        n = mask.shape[0]
        got = np.zeros(n * n, np.uint32)
        mm = np.ascontiguousarray(mask)
        oracle().orc_nearest_fluid(mm.ctypes.data_as(C.POINTER(C.c_uint8)), n,
                                   got.ctypes.data_as(C.POINTER(C.c_uint32)))
        want = nearest_closed_form(mask)
        assert np.array_equal(got, want), (n, int(mask.sum()))


@needs_ref
def test_compose_vs_reference_combined_query():
    """orc_compose (WaterSurface::sample_at, reference src/interaction.cpp:863-901)
    applied to the oracle's ocean sample and field sample, against the
    reference's ocean_sample_combined on the same ocean and field history.
    The chop is heavy enough that some queried normals are folded
    (normal_y <= 1e-6), where the FFT normal must pass through unchanged."""
    import ctypes as C
    from oceanlib_amd.capi import InteractionSample, OceanSurfaceEx
    from oceanlib_amd.workloads import desc_from
    from oracle_ffi import OracleOcean, RefOcean

    od = desc_from(dict(size=64, patch_length=16.0, wind_speed=30.0, choppiness=2.8,
                        compute_velocity=1, seed=99))
    oo, ro = OracleOcean(od), RefOcean(od)
    init = ref().oref_interaction_desc_init
    d = idesc(init, **dict(CONFIG3_INTERACTION, size=128, extent=12.8))
    fa, fb = Field("orc", oracle(), d), Field("oref", ref(), d)
    folded = 0
    try:
        srcs = config3_sources()
        rng = np.random.default_rng(8)
        for k in range(40):
            for rec in config3_disturbances(srcs, k / 60.0, centre=(6.4, 6.4)):
                fa.add(rec)
                fb.add(rec)
            fa.update(1 / 60)
            fb.update(1 / 60)
            if k % 8 != 7:
                continue
            t = 0.37 + k / 60.0
            oo.update(t)
            ro.update(t)
            # plus the displaced positions of folded cells, where the query's
            # solved point lands in or next to the fold
            disp = oo.displacement().reshape(-1, 4)
            nrm = oo.normal().reshape(-1, 4)
            cell = od.patch_length / od.size
            idx = np.nonzero(nrm[:, 1] <= 1.0e-6)[0][:60]
            fx = (idx % od.size) * cell + disp[idx, 0]
            fz = (idx // od.size) * cell + disp[idx, 2]
            pts = np.concatenate([rng.uniform(-2.0, 15.0, size=(400, 2)),
                                  rng.uniform(-80.0, 80.0, size=(50, 2)),
                                  np.stack([fx, fz], axis=1)]).astype(np.float32)
            for x, z in pts:
                want = OceanSurfaceEx()
                ref().oref_sample_combined(ro.h, fb.h, float(x), float(z), C.byref(want))
                got = oo.sample(float(x), float(z))
                if got.normal_y <= 1.0e-6:
                    folded += 1
                e = fa.sample(x, z)
                oracle().orc_compose(C.byref(got), C.byref(e))
                assert bytes(got) == bytes(want), (k, x, z)
        assert folded > 0, "no folded crest was queried"
    finally:
        fa.close()
        fb.close()
