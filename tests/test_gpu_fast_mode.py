"""GPU tests of the opt-in fast mode (ocean_amd_set_precision).

No reference counterpart: the reference's frame is bit-exact only, and it
names a real-to-complex FFT as its speed-up path (ARCHITECTURE.md ADR-008).
The fast modes give up the reference's rounding on purpose, so they are
compared with the oracle (the reference's arithmetic) within a tolerance,
stated per mode, on each output channel as max |fast - oracle| over the
channel's max |oracle|:
  FAST       (hardware sin/cos, FMA butterflies, rsqrt)        <= 2e-5
  FAST_HALF  (FAST + fp16 fields between the passes)           <= 3e-3
The foam channel is a clamp of the Jacobian, (threshold - J) / threshold, and
is held to the Jacobian's own error over the threshold. Switching back to
EXACT gives the oracle's bits again; every launch shape of a fast mode gives
the same bits as its default shape.
"""
import ctypes as C

import numpy as np
import pytest

from oceanlib_amd import capi
from oceanlib_amd.workloads import desc_from
from oracle_ffi import OracleOcean, assert_bitexact

pytestmark = pytest.mark.gpu

EXACT, FAST, FAST_HALF = 0, 1, 2
TOL = {FAST: 2e-5, FAST_HALF: 3e-3}


@pytest.fixture(scope="module")
def lib():
    return capi.load()


def buffers(lib, sim, n):
    b = lib.ocean_get_buffers(sim)
    n4 = 4 * n * n
    d = np.ctypeslib.as_array(b.displacement, shape=(n4,)).copy().reshape(-1, 4)
    nr = np.ctypeslib.as_array(b.normal, shape=(n4,)).copy().reshape(-1, 4)
    vp = lib.ocean_get_velocity(sim)
    v = np.ctypeslib.as_array(vp, shape=(n4,)).copy().reshape(-1, 4) if vp else None
    return d, nr, v


def channel_errors(got, want):
    """max |got - want| / max |want| per channel (float64)."""
    g, w = got.astype(np.float64), want.astype(np.float64)
    scale = np.maximum(np.abs(w).max(axis=0), 1e-30)
    return np.abs(g - w).max(axis=0) / scale


def check_close(lib, sim, o, n, mode, what, foam_threshold=0.5):
    d, nr, v = buffers(lib, sim, n)
    od = o.displacement().reshape(-1, 4)
    on = o.normal().reshape(-1, 4)
    tol = TOL[mode]
    ed = channel_errors(d[:, :3], od[:, :3])
    en = channel_errors(nr, on)
    # foam = clamp((thr - J) / thr): its error is at most |dJ| / thr
    jerr = np.abs(nr[:, 3].astype(np.float64) - on[:, 3]).max()
    ferr = np.abs(d[:, 3].astype(np.float64) - od[:, 3]).max()
    assert ferr <= jerr / foam_threshold * (1 + 1e-6) + 1e-7, (what, ferr, jerr)
    errs = {"disp": ed.max(), "normal": en.max()}
    if v is not None:
        ev = channel_errors(v[:, :3], o.velocity().reshape(-1, 4)[:, :3])
        errs["velocity"] = ev.max()
    print(f"{what} mode {mode}: " + ", ".join(f"{k} {x:.2e}" for k, x in errs.items()))
    for k, x in errs.items():
        assert x <= tol, (what, mode, k, x)
    return errs


@pytest.mark.parametrize("kw", [
    dict(size=256, patch_length=200.0, wind_speed=12.0, seed=1337, choppiness=1.0),
    dict(size=512, patch_length=250.0, wind_speed=22.0, choppiness=1.3, compute_velocity=1,
         water_depth=50.0, seed=7),
    dict(size=64, patch_length=40.0, wind_speed=9.0, seed=3, compute_velocity=1),
    dict(size=2048, patch_length=1000.0, wind_speed=15.0, seed=5),
])
def test_fast_modes_within_tolerance_and_back_to_exact(lib, kw):
    d = desc_from(kw)
    st = C.c_int(-1)
    sim = lib.ocean_create(C.byref(d), C.byref(st))
    assert sim, st.value
    o = OracleOcean(d)
    n = d.size
    try:
        assert lib.ocean_amd_get_precision(sim) == EXACT
        for mode in (FAST, FAST_HALF):
            assert lib.ocean_amd_set_precision(sim, mode) == 0
            assert lib.ocean_amd_get_precision(sim) == mode
            for t in (0.0, 1.0 / 60.0, 9.75, 3600.0):
                lib.ocean_update(sim, t)
                o.update(t)
                check_close(lib, sim, o, n, mode, f"N={n} t={t}")
        assert lib.ocean_amd_set_precision(sim, EXACT) == 0
        lib.ocean_update(sim, 2.5)
        o.update(2.5)
        dd, nn, vv = buffers(lib, sim, n)
        assert_bitexact(dd.reshape(-1), o.displacement(), "exact again: displacement")
        assert_bitexact(nn.reshape(-1), o.normal(), "exact again: normal")
        assert lib.ocean_amd_set_precision(sim, 3) != 0
        assert lib.ocean_amd_get_precision(sim) == EXACT
        assert lib.ocean_amd_last_error(sim) == 0
    finally:
        lib.ocean_destroy(sim)


@pytest.mark.parametrize("mode", [FAST, FAST_HALF])
def test_every_fast_launch_shape_gives_the_same_bits(lib, mode):
    for kw in (dict(size=512, patch_length=250.0, wind_speed=22.0, choppiness=1.3,
                    compute_velocity=1, seed=7),
               dict(size=1024, patch_length=600.0, wind_speed=14.0, seed=2),
               dict(size=128, patch_length=90.0, wind_speed=14.0, seed=2)):
        d = desc_from(kw)
        st = C.c_int(-1)
        sim = lib.ocean_create(C.byref(d), C.byref(st))
        assert sim
        try:
            assert lib.ocean_amd_set_precision(sim, mode) == 0
            nr, nc = C.c_uint32(), C.c_uint32()
            assert lib.ocean_amd_frame_shapes(sim, C.byref(nr), C.byref(nc)) == 0
            assert nr.value >= 1 and nc.value >= 1
            lib.ocean_update(sim, 4.25)
            ref = buffers(lib, sim, d.size)
            for rv in range(nr.value):
                for cv in range(nc.value):
                    assert lib.ocean_amd_set_frame_shape(sim, rv, cv) == 0
                    lib.ocean_update(sim, 4.25)
                    got = buffers(lib, sim, d.size)
                    for a, b in zip(got, ref):
                        if a is not None:
                            assert_bitexact(a.reshape(-1), b.reshape(-1), f"shape {rv},{cv}")
            rv, cv, rms, cms = C.c_uint32(), C.c_uint32(), C.c_float(), C.c_float()
            assert lib.ocean_amd_autotune_frame(sim, 2, C.byref(rv), C.byref(cv), C.byref(rms),
                                                C.byref(cms)) == 0
            assert lib.ocean_amd_get_precision(sim) == mode  # the tuner keeps the mode
            lib.ocean_update(sim, 4.25)
            for a, b in zip(buffers(lib, sim, d.size), ref):
                if a is not None:
                    assert_bitexact(a.reshape(-1), b.reshape(-1), "autotuned shape")
        finally:
            lib.ocean_destroy(sim)


def test_batch_fast_modes(lib):
    kws = [dict(size=64, patch_length=50.0, wind_speed=8.0, seed=1),
           dict(size=128, patch_length=120.0, wind_speed=14.0, seed=2, compute_velocity=1),
           dict(size=256, patch_length=300.0, wind_speed=18.0, seed=3, water_depth=20.0),
           dict(size=512, patch_length=250.0, wind_speed=22.0, choppiness=1.3, seed=7,
                compute_velocity=1),
           dict(size=32, patch_length=10.0, wind_speed=5.0, seed=4)]
    descs = [desc_from(kw) for kw in kws]
    arr = (capi.OceanDesc * len(descs))(*descs)
    st = C.c_int(-1)
    b = lib.ocean_amd_batch_create(arr, len(descs), C.byref(st))
    assert b, st.value
    orcs = [OracleOcean(d) for d in descs]
    try:
        lib.ocean_amd_batch_autotune(b, 2)
        for mode in (FAST, FAST_HALF):
            assert lib.ocean_amd_batch_set_precision(b, mode) == 0
            for t in (0.5, 77.0):
                lib.ocean_amd_batch_update(b, t)
                for i, (d, o) in enumerate(zip(descs, orcs)):
                    o.update(t)
                    check_close(lib, lib.ocean_amd_batch_level(b, i), o, d.size, mode,
                                f"batch level {i} t={t}")
        assert lib.ocean_amd_batch_set_precision(b, EXACT) == 0
        lib.ocean_amd_batch_update(b, 1.25)
        for i, (d, o) in enumerate(zip(descs, orcs)):
            o.update(1.25)
            dd, nn, _ = buffers(lib, lib.ocean_amd_batch_level(b, i), d.size)
            assert_bitexact(dd.reshape(-1), o.displacement(), f"exact again: level {i}")
            assert_bitexact(nn.reshape(-1), o.normal(), f"exact again: level {i}")
        # a batch's level follows its batch's mode
        assert lib.ocean_amd_set_precision(lib.ocean_amd_batch_level(b, 0), FAST) != 0
        assert lib.ocean_amd_batch_set_precision(b, 7) != 0
        assert lib.ocean_amd_batch_synchronize(b) == 0
        assert lib.ocean_amd_batch_last_error(b) == 0
    finally:
        lib.ocean_amd_batch_destroy(b)
