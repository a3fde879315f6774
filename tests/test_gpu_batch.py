"""GPU parity of batches of seas (ocean_amd_batch_*) through the C ABI.

A batch steps many levels of uneven N with one persistent launch per (N,
velocity) group and pass, replayed as a hipGraph. The reference has no batch;
its contract is per ocean (tests/test_determinism.cpp:29-76: the same desc and
time give the same buffers, bit for bit), so each level of a batch must equal
a standalone ocean with the same desc at the same time: the oracle's for mixed
sizes here, and the reference's own run level by level for the 256-sea
workload (tests/golden/config4_checksums.npz, made by make_config4.py).
"""
import ctypes as C
import json
import os

import numpy as np
import pytest

from oceanlib_amd import capi
from oceanlib_amd.workloads import checksum_words, config4_levels, desc_from
from oracle_ffi import (OracleOcean, as_array, assert_bitexact, default_foam_params, fptr,
                        oracle)

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "config4_checksums.npz")
MASK64 = (1 << 64) - 1


@pytest.fixture(scope="module")
def lib():
    return capi.load()


class Batch:
    def __init__(self, lib, descs):
        self.lib = lib
        arr = (capi.OceanDesc * len(descs))(*descs)
        st = C.c_int(-1)
        self.h = lib.ocean_amd_batch_create(arr, len(descs), C.byref(st))
        assert self.h, f"ocean_amd_batch_create status {st.value}"
        self.count = len(descs)
        self.descs = descs

    def update(self, t):
        self.lib.ocean_amd_batch_update(self.h, t)

    def level(self, i):
        p = self.lib.ocean_amd_batch_level(self.h, i)
        assert p
        return p

    def buffers(self, i):
        sim = self.level(i)
        n = self.descs[i].size
        b = self.lib.ocean_get_buffers(sim)
        n4 = 4 * n * n
        d = np.ctypeslib.as_array(b.displacement, shape=(n4,)).copy()
        nr = np.ctypeslib.as_array(b.normal, shape=(n4,)).copy()
        vp = self.lib.ocean_get_velocity(sim)
        v = np.ctypeslib.as_array(vp, shape=(n4,)).copy() if vp else None
        return d, nr, v

    def sample(self, first, count, xz):
        xz = np.ascontiguousarray(xz, dtype=np.float32)
        out = (capi.OceanSurfaceEx * len(xz))()
        rc = self.lib.ocean_amd_batch_sample(self.h, first, count, fptr(xz.reshape(-1)), len(xz), out)
        assert rc == 0
        return out

    def groups(self):
        out = []
        for g in range(self.lib.ocean_amd_batch_group_count(self.h)):
            info = capi.BatchGroup()
            assert self.lib.ocean_amd_batch_group_info(self.h, g, C.byref(info)) == 0
            out.append(info)
        return out

    def close(self):
        if self.h:
            assert self.lib.ocean_amd_batch_synchronize(self.h) == 0
            assert self.lib.ocean_amd_batch_last_error(self.h) == 0
            self.lib.ocean_amd_batch_destroy(self.h)
            self.h = None


def surf_fields(s):
    return np.array([getattr(s, f) for f, _ in capi.OceanSurfaceEx._fields_], np.float32)


MIXED = [
    dict(size=16, patch_length=20.0, wind_speed=10.0, seed=1),
    dict(size=64, patch_length=90.0, wind_speed=6.0, seed=2),
    dict(size=64, patch_length=33.0, wind_speed=14.0, seed=3, compute_velocity=1, water_depth=12.0),
    dict(size=128, patch_length=300.0, wind_speed=20.0, seed=4, choppiness=1.3),
    dict(size=32, patch_length=8.0, wind_speed=3.0, seed=5, compute_velocity=1),
    dict(size=256, patch_length=200.0, wind_speed=12.0, seed=1337),
    dict(size=256, patch_length=150.0, wind_speed=15.0, seed=42, compute_velocity=1, water_depth=80.0),
    dict(size=64, patch_length=400.0, wind_speed=25.0, seed=6, swell=0.6, wind_direction=1.1),
    dict(size=512, patch_length=250.0, wind_speed=22.0, seed=7, choppiness=1.3,
         compute_velocity=1, water_depth=50.0),
    dict(size=512, patch_length=900.0, wind_speed=9.0, seed=8),
    dict(size=1024, patch_length=1000.0, wind_speed=17.0, seed=9),
]


def test_mixed_batch_bitexact(lib):
    """Every level of a batch of 11 levels (N 16..1024, velocity on and off,
    finite depth, swell) equals the oracle's standalone ocean, bit for bit, at
    several times, before and after the autotuner re-picks grid sizes."""
    descs = [desc_from(kw) for kw in MIXED]
    b = Batch(lib, descs)
    orc = [OracleOcean(d) for d in descs]
    try:
        groups = b.groups()
        # N 16 and 32 are stepped level by level; the rest are grouped.
        assert sorted((g.size, g.velocity) for g in groups) == [
            (64, 0), (64, 1), (128, 0), (256, 0), (256, 1), (512, 0), (512, 1), (1024, 0)]
        assert sum(g.levels for g in groups) == 9
        for g in groups:
            assert g.items == g.levels * g.size * g.size // 2048
            assert 1 <= g.grid_row <= g.items and 1 <= g.grid_col <= g.items
        sums = (C.c_uint64 * len(descs))()
        for t in (0.0, 1.0 / 60.0, 37.5):
            b.update(t)
            assert lib.ocean_amd_batch_checksums(b.h, 0, len(descs), sums) == 0
            for i, o in enumerate(orc):
                o.update(t)
                d, nr, v = b.buffers(i)
                assert_bitexact(d, o.displacement(), f"level {i} t={t} displacement")
                assert_bitexact(nr, o.normal(), f"level {i} t={t} normal")
                if descs[i].compute_velocity:
                    assert_bitexact(v, o.velocity(), f"level {i} t={t} velocity")
                else:
                    assert v is None
                # the device checksum is the documented sum over these words
                words = [o.displacement(), o.normal()] + ([o.velocity()] if v is not None else [])
                assert sums[i] == checksum_words(np.concatenate(words)), f"level {i} checksum"
        part = (C.c_uint64 * 3)()
        assert lib.ocean_amd_batch_checksums(b.h, 4, 3, part) == 0
        assert list(part) == list(sums)[4:7]
        assert lib.ocean_amd_batch_checksums(b.h, 9, 3, part) != 0  # past the end
        lib.ocean_amd_batch_autotune(b.h, 2)
        tuned = b.groups()
        assert all(g.row_ms > 0 and g.col_ms > 0 for g in tuned)
        t = 3600.0
        b.update(t)
        for i, o in enumerate(orc):
            o.update(t)
            d, nr, _ = b.buffers(i)
            assert_bitexact(d, o.displacement(), f"tuned level {i} displacement")
            assert_bitexact(nr, o.normal(), f"tuned level {i} normal")
        assert abs(lib.ocean_get_time(b.level(3)) - t) == 0.0
    finally:
        b.close()


def test_every_batch_launch_choice_bitexact(lib):
    """Each group's launch choices set by hand - the layout of the fields
    between the passes (row-major, or tiles as wide as an item's columns for
    N = 256 .. 1024), 8 or 4 FFT points per lane for either pass (4: stages in
    pairs, every item in two blocks), a small persistent grid or one block per
    unit - give every level's frame the oracle's bits (checksums over the
    frame's output words)."""
    descs = [desc_from(kw) for kw in MIXED]
    b = Batch(lib, descs)
    orc = [OracleOcean(d) for d in descs]
    try:
        t = 21.75
        want = []
        for i, o in enumerate(orc):
            o.update(t)
            words = [o.displacement(), o.normal()] + ([o.velocity()] if descs[i].compute_velocity
                                                       else [])
            want.append(checksum_words(np.concatenate(words)))
        groups = b.groups()
        sums = (C.c_uint64 * len(descs))()
        tried = 0
        for g, info in enumerate(groups):
            radices = (8, 4) if info.size <= 1024 else (8,)
            layouts = (0, 1) if 256 <= info.size <= 1024 else (0,)
            for layout in layouts:
                for rr in radices:
                    for rc in radices:
                        for small in (True, False):
                            ur = info.items * (2 if rr == 4 else 1)
                            uc = info.items * (2 if rc == 4 else 1)
                            gr, gc = (min(8, ur), min(8, uc)) if small else (ur, uc)
                            assert lib.ocean_amd_batch_set_group_launch(b.h, g, layout, rr, gr, rc,
                                                                        gc) == 0
                            got = b.groups()[g]
                            assert (got.layout, got.radix_row, got.grid_row, got.radix_col,
                                    got.grid_col) == (layout, rr, gr, rc, gc)
                            b.update(t + 1.0)  # a different frame first
                            b.update(t)
                            assert lib.ocean_amd_batch_checksums(b.h, 0, len(descs), sums) == 0
                            assert list(sums) == want, (info.size, layout, rr, rc, small)
                            tried += 1
            # choices that do not exist are refused
            assert lib.ocean_amd_batch_set_group_launch(b.h, g, 0, 16, 1, 8, 1) != 0
            assert lib.ocean_amd_batch_set_group_launch(b.h, g, 0, 8, 0, 8, 1) != 0
            assert lib.ocean_amd_batch_set_group_launch(b.h, g, 0, 8, info.items + 1, 8, 1) != 0
            assert lib.ocean_amd_batch_set_group_launch(b.h, g, 2, 8, 1, 8, 1) != 0
            if not 256 <= info.size <= 1024:
                assert lib.ocean_amd_batch_set_group_launch(b.h, g, 1, 8, 1, 8, 1) != 0
        assert tried == sum(8 * (2 if 256 <= g.size <= 1024 else 1) for g in groups)
        lib.ocean_amd_batch_autotune(b.h, 2)
        b.update(t)
        assert lib.ocean_amd_batch_checksums(b.h, 0, len(descs), sums) == 0
        assert list(sums) == want
        print("tuned:", [(g.size, g.layout, g.radix_row, g.grid_row, g.radix_col, g.grid_col)
                         for g in b.groups()])
        # a fast mode has no tiled kernels: its groups fall back to row-major
        assert lib.ocean_amd_batch_set_precision(b.h, 1) == 0
        assert all(g.layout == 0 for g in b.groups())
        assert lib.ocean_amd_batch_set_precision(b.h, 0) == 0
        b.update(t)
        assert lib.ocean_amd_batch_checksums(b.h, 0, len(descs), sums) == 0
        assert list(sums) == want
    finally:
        b.close()


def test_batch_queries(lib):
    """Batch queries: one level answers as that ocean's sample_at, a range of
    levels as a cascade of those levels (reference src/cascade.cpp:82-153),
    both bit for bit against the oracle; the level handle serves ocean_sample_ex."""
    kws = [dict(size=256, patch_length=L, wind_speed=15.0, seed=42) for L in (800.0, 150.0, 25.0)]
    kws.append(dict(size=64, patch_length=50.0, wind_speed=8.0, seed=3, compute_velocity=1))
    descs = [desc_from(kw) for kw in kws]
    b = Batch(lib, descs)
    orc = [OracleOcean(d) for d in descs]
    ol = oracle()
    try:
        t = 12.25
        b.update(t)
        for o in orc:
            o.update(t)
        rng = np.random.default_rng(5)
        xz = rng.uniform(-900, 900, (300, 2)).astype(np.float32)
        single = b.sample(3, 1, xz)
        casc = b.sample(0, 3, xz)
        handles = (C.c_void_p * 3)(*[o.h for o in orc[:3]])
        for k, (x, z) in enumerate(xz):
            want = orc[3].sample(float(x), float(z))
            assert_bitexact(surf_fields(single[k]), surf_fields(want), f"single query {k}")
            wc = capi.OceanSurfaceEx()
            ol.orc_cascade_sample(handles, 3, float(x), float(z), C.byref(wc))
            assert_bitexact(surf_fields(casc[k]), surf_fields(wc), f"cascade query {k}")
            if k % 50 == 0:
                s = capi.OceanSurfaceEx()
                lib.ocean_sample_ex(b.level(3), float(x), float(z), C.byref(s))
                assert_bitexact(surf_fields(s), surf_fields(want), f"handle query {k}")
        bad = (capi.OceanSurfaceEx * 1)()
        one = np.zeros(2, np.float32)
        assert lib.ocean_amd_batch_sample(b.h, 3, 2, fptr(one), 1, bad) != 0  # past the end
        assert lib.ocean_amd_batch_sample(b.h, 0, 0, fptr(one), 1, bad) != 0  # no levels
    finally:
        b.close()


@pytest.mark.skipif(not os.path.exists(GOLDEN), reason="config4 fixture not generated")
def test_config4_batched_seas_vs_reference(lib):
    """256 seas, 516 levels of N 64..512, 300 frames at 1/60 s in one batch
    (autotuned), every level's frame checksums against the reference's own
    level-by-level run: frame 0, the last frame and the chain over all frames
    (sum of (2f + 1) * checksum_f mod 2^64)."""
    gold = np.load(GOLDEN)
    flat, labels = config4_levels()
    assert len(flat) == len(gold["size"])
    assert [kw["size"] for kw in flat] == gold["size"].tolist()
    frames = int(json.loads(str(gold["meta"]))["frames"])
    descs = [desc_from(kw) for kw in flat]
    b = Batch(lib, descs)
    count = len(descs)
    try:
        lib.ocean_amd_batch_autotune(b.h, 3)
        sums = (C.c_uint64 * count)()
        chain = [0] * count
        first = last = None
        for f in range(frames):
            b.update(f / 60.0)
            assert lib.ocean_amd_batch_checksums(b.h, 0, count, sums) == 0
            cur = list(sums)
            for i in range(count):
                chain[i] = (chain[i] + (2 * f + 1) * cur[i]) & MASK64
            if f == 0:
                first = cur
            last = cur
        for name, got in (("first", first), ("last", last), ("chain", chain)):
            want = gold[name].tolist()
            bad = [i for i in range(count) if got[i] != want[i]]
            assert not bad, (f"{name}: {len(bad)} of {count} levels differ, first "
                             f"(sea, level) {labels[bad[0]]}, N={flat[bad[0]]['size']}")
    finally:
        b.close()


def test_foam_on_batch_level_vs_oracle(lib):
    """A FoamField on a batch level (whose host query staging is created only
    on demand): the field after 40 substeps, foam_at at 128 points on and off
    the patch, and coverage, against the oracle's FoamField on the same ocean
    (reference src/foam.cpp:89-139,218-252,291-313)."""
    descs = [desc_from(dict(size=64, patch_length=90.0, wind_speed=20.0, choppiness=1.3,
                            compute_velocity=1, seed=31)),
             desc_from(dict(size=128, patch_length=35.0, wind_speed=20.0, choppiness=1.1,
                            compute_velocity=1, water_depth=12.0, seed=32))]
    b = Batch(lib, descs)
    olib = oracle()
    fp = default_foam_params()
    foams, ofoams, orcs = [], [], []
    try:
        fd = capi.FoamDesc()
        lib.ocean_amd_foam_desc_init(C.byref(fd))
        fd.source_gain = fp.source_gain
        st = C.c_int(-1)
        for i, d in enumerate(descs):
            f = lib.ocean_amd_foam_create(b.level(i), C.byref(fd), C.byref(st))
            assert f, st.value
            foams.append(f)
            o = OracleOcean(d)
            orcs.append(o)
            ofoams.append(olib.orc_foam_create(o.h, C.byref(fp)))
        for k in range(40):
            t = 2.0 + k / 60.0
            b.update(t)
            for f, o, of in zip(foams, orcs, ofoams):
                o.update(t)
                lib.ocean_amd_foam_update(f, np.float32(1.0 / 60.0))
                olib.orc_foam_update(of, np.float32(1.0 / 60.0))
        rng = np.random.default_rng(12)
        for i, (f, of) in enumerate(zip(foams, ofoams)):
            n = descs[i].size
            fb = np.ctypeslib.as_array(lib.ocean_amd_foam_data(f), shape=(n * n,)).copy()
            assert_bitexact(fb, as_array(olib.orc_foam_data(of), n * n), f"level {i} foam")
            for x, z in rng.uniform(-150.0, 250.0, size=(128, 2)).astype(np.float32):
                got = lib.ocean_amd_foam_at(f, float(x), float(z))
                want = olib.orc_foam_at(of, float(x), float(z))
                assert np.float32(got).tobytes() == np.float32(want).tobytes(), (i, x, z)
            cov, want = lib.ocean_amd_foam_coverage(f), olib.orc_foam_coverage(of)
            assert want > 0.0 and np.float64(cov).tobytes() == np.float64(want).tobytes()
            assert lib.ocean_amd_last_error(b.level(i)) == 0
    finally:
        for f in foams:
            lib.ocean_amd_foam_destroy(f)
        for of in ofoams:
            olib.orc_foam_destroy(of)
        b.close()


@pytest.mark.parametrize("n", [100_000, 5_000])
def test_cascade_heights_level_by_level_equal_the_full_query(lib, n):
    """Cascade heights alone are computed without the normal and velocity
    loads: for more than 64K points one level per launch (the running sum
    kept in the output between launches), for fewer all levels in one
    launch. They must equal, bit for bit, the heights of the full cascade
    query and the oracle's CascadeStack::sample_at heights (reference
    src/cascade.cpp:82-153)."""
    from oceanlib_amd.hipmem import DeviceArray
    kws = [dict(size=256, patch_length=L, wind_speed=15.0, seed=42) for L in (800.0, 150.0, 25.0)]
    descs = [desc_from(kw) for kw in kws]
    b = Batch(lib, descs)
    orc = [OracleOcean(d) for d in descs]
    try:
        b.update(7.5)
        for o in orc:
            o.update(7.5)
        rng = np.random.default_rng(6)
        xz = rng.uniform(-1500, 1500, (n, 2)).astype(np.float32)
        dxz = DeviceArray.from_numpy(xz)
        h = DeviceArray.from_numpy(np.full(n, np.nan, np.float32))
        h_full = DeviceArray.from_numpy(np.full(n, np.nan, np.float32))
        full = DeviceArray((n, 10), np.float32)
        assert lib.ocean_amd_batch_sample_device(b.h, 0, 3, dxz.ptr, n, None, h.ptr) == 0
        assert lib.ocean_amd_batch_sample_device(b.h, 0, 3, dxz.ptr, n, full.ptr, h_full.ptr) == 0
        assert lib.ocean_amd_batch_synchronize(b.h) == 0
        got, want_full = h.numpy(), h_full.numpy()
        assert_bitexact(got, want_full, "per-level heights vs full query")
        assert_bitexact(got, full.numpy()[:, 0].copy(), "per-level heights vs full surfaces")
        handles = (C.c_void_p * 3)(*[o.h for o in orc])
        wc = capi.OceanSurfaceEx()
        for k in range(0, n, 997 if n > 20_000 else 97):
            oracle().orc_cascade_sample(handles, 3, float(xz[k, 0]), float(xz[k, 1]), C.byref(wc))
            assert np.float32(got[k]).tobytes() == np.float32(wc.height).tobytes(), k
    finally:
        b.close()
