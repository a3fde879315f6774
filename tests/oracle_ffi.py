"""ctypes access to the CPU checkers under oracle/ (test infrastructure only).

- `oracle()`  -> oracle/liboracle.so, the C restatement (ocean_oracle.c).
- `ref()`     -> oracle/_ref/liboceanref.so, the reference library's own
                 sources compiled here (oracle/Makefile), or None when it was
                 not built (it needs the reference tree at build time).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from oceanlib_amd.capi import (OceanDesc, OceanBuffers, OceanSurfaceEx, InteractionDesc,
                               Disturbance, InteractionBuffers, InteractionSample)

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "liboceanref.so")

_P = C.c_void_p
_F = C.POINTER(C.c_float)
_U32 = C.POINTER(C.c_uint32)
_D = C.POINTER(OceanDesc)
_DBL = C.POINTER(C.c_double)


class FoamParams(C.Structure):
    _fields_ = [(n, C.c_float) for n in ("decay", "source_gain", "advect_scale", "wind_drift_x",
                                         "wind_drift_z", "fixed_dt")] + \
               [("max_substeps", C.c_uint32)]


def default_foam_params() -> FoamParams:
    """Reference FoamDesc defaults (include/ocean/foam.hpp:38-74)."""
    return FoamParams(0.35, 4.0, 1.0, 0.0, 0.0, 1.0 / 60.0, 4)


_COMMON = {
    "pcg32_draws": (None, [C.c_uint64, C.c_uint64, C.c_uint32, _U32]),
    "row_gaussians": (None, [C.c_uint64, C.c_uint32, C.c_uint32, _F]),
    "sincos": (None, [_F, C.c_uint32, _F, _F]),
    "reduced_phase": (None, [_F, C.c_uint32, C.c_double, _F]),
    "spectrum": (None, [_D] + [_F] * 8),
    "wave_density": (C.c_double, [_D, C.c_double, C.c_double]),
    "fft_inverse": (None, [C.c_uint32, _F, _F]),
    "fft2d_inverse": (None, [C.c_uint32, _F, _F]),
}

_ORACLE_ONLY = {
    "orc_twiddles": (None, [C.c_uint32, _F, _F]),
    "orc_ocean_create": (_P, [_D]),
    "orc_ocean_destroy": (None, [_P]),
    "orc_ocean_update": (None, [_P, C.c_double]),
    "orc_ocean_displacement": (_F, [_P]),
    "orc_ocean_normal": (_F, [_P]),
    "orc_ocean_velocity": (_F, [_P]),
    "orc_ocean_planes": (_F, [_P]),
    "orc_ocean_evolve": (None, [_P, C.c_double, _F]),
    "orc_ocean_sample": (None, [_P, C.c_float, C.c_float, C.POINTER(OceanSurfaceEx)]),
    "orc_sample_buffers": (None, [_D, _F, _F, _F, _F, C.c_uint32, C.POINTER(OceanSurfaceEx)]),
    "orc_cascade_sample": (None, [C.POINTER(_P), C.c_uint32, C.c_float, C.c_float,
                                  C.POINTER(OceanSurfaceEx)]),
    "orc_foam_create": (_P, [_P, C.POINTER(FoamParams)]),
    "orc_foam_destroy": (None, [_P]),
    "orc_foam_update": (None, [_P, C.c_float]),
    "orc_foam_data": (_F, [_P]),
    "orc_foam_last_substeps": (C.c_uint32, [_P]),
    "orc_foam_at": (C.c_float, [_P, C.c_float, C.c_float]),
    "orc_foam_coverage": (C.c_double, [_P]),
    "orc_iwave_taps": (C.c_int, [C.c_uint32, C.c_double, C.c_double, _F,
                                 C.POINTER(C.c_double)]),
    "orc_iwave_stencil": (C.c_int, [C.c_uint32, C.c_double, C.c_double, _F, _DBL]),
    "orc_iwave_symbol_scan": (None, [_F, C.c_uint32, C.c_float, C.c_double, C.c_int, _DBL]),
    "orc_nearest_fluid": (None, [C.POINTER(C.c_uint8), C.c_uint32, C.POINTER(C.c_uint32)]),
    "orc_iw_mask_tables": (C.c_uint32, [_P, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    "orc_iw_create": (_P, [C.POINTER(InteractionDesc)]),
    "orc_iw_destroy": (None, [_P]),
    "orc_iw_add": (None, [_P, C.POINTER(Disturbance)]),
    "orc_iw_dropped": (C.c_uint64, [_P]),
    "orc_iw_update": (None, [_P, C.c_float]),
    "orc_iw_last_substeps": (C.c_uint32, [_P]),
    "orc_iw_recenter": (C.c_int32, [_P, C.c_float, C.c_float]),
    "orc_iw_set_obstruction": (None, [_P, C.POINTER(C.c_uint8)]),
    "orc_iw_sample_at": (None, [_P, C.c_float, C.c_float, C.POINTER(InteractionSample)]),
    "orc_iw_field": (_F, [_P]),
    "orc_iw_energy": (C.c_double, [_P]),
    "orc_iw_fixed_dt": (C.c_float, [_P]),
    "orc_iw_dt_limit": (C.c_float, [_P]),
    "orc_iw_origin_x": (C.c_float, [_P]),
    "orc_iw_origin_z": (C.c_float, [_P]),
    "orc_compose": (None, [C.POINTER(OceanSurfaceEx), C.POINTER(InteractionSample)]),
}

_REF_ONLY = {
    "oref_create": (_P, [_D, C.POINTER(C.c_int)]),
    "oref_destroy": (None, [_P]),
    "oref_update": (None, [_P, C.c_double]),
    "oref_get_buffers": (OceanBuffers, [_P]),
    "oref_get_velocity": (_F, [_P]),
    "oref_sample_ex": (None, [_P, C.c_float, C.c_float, C.POINTER(OceanSurfaceEx)]),
    "oref_desc_init": (None, [_D]),
    "oref_evolve": (None, [_D, C.c_double, _F]),
    "oref_cascade_create": (_P, [_D, C.c_uint32]),
    "oref_cascade_destroy": (None, [_P]),
    "oref_cascade_update": (None, [_P, C.c_double]),
    "oref_cascade_displacement": (_F, [_P, C.c_uint32]),
    "oref_cascade_normal": (_F, [_P, C.c_uint32]),
    "oref_cascade_sample": (None, [_P, C.c_float, C.c_float, C.POINTER(OceanSurfaceEx)]),
    "oref_foam_create": (_P, [_P, C.POINTER(FoamParams)]),
    "oref_foam_destroy": (None, [_P]),
    "oref_foam_update": (None, [_P, C.c_float]),
    "oref_foam_data": (_F, [_P]),
    "oref_foam_last_substeps": (C.c_uint32, [_P]),
    "oref_foam_at": (C.c_float, [_P, C.c_float, C.c_float]),
    "oref_foam_coverage": (C.c_double, [_P]),
    "oref_interaction_desc_init": (None, [C.POINTER(InteractionDesc)]),
    "oref_interaction_create": (_P, [C.POINTER(InteractionDesc), C.POINTER(C.c_int)]),
    "oref_interaction_destroy": (None, [_P]),
    "oref_interaction_add": (None, [_P, C.POINTER(Disturbance)]),
    "oref_interaction_update": (None, [_P, C.c_float]),
    "oref_interaction_recenter": (C.c_int32, [_P, C.c_float, C.c_float]),
    "oref_interaction_set_obstruction": (None, [_P, C.POINTER(C.c_uint8)]),
    "oref_interaction_get_buffers": (InteractionBuffers, [_P]),
    "oref_interaction_sample_at": (None, [_P, C.c_float, C.c_float,
                                          C.POINTER(InteractionSample)]),
    "oref_interaction_fixed_dt": (C.c_float, [_P]),
    "oref_interaction_dt_limit": (C.c_float, [_P]),
    "oref_interaction_dropped": (C.c_uint64, [_P]),
    "oref_interaction_energy": (C.c_double, [_P]),
    "oref_sample_combined": (None, [_P, _P, C.c_float, C.c_float, C.POINTER(OceanSurfaceEx)]),
    "oref_iwave_taps": (C.c_int, [C.c_uint32, C.c_double, C.c_double, _F]),
    "oref_iwave_stencil": (C.c_int, [C.c_uint32, C.c_double, C.c_double, _F, _DBL]),
}

_cache: dict = {}


def _bind(lib, table):
    for name, (res, args) in table.items():
        fn = getattr(lib, name, None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = args


def oracle():
    if "orc" not in _cache:
        if not os.path.exists(ORACLE_SO):
            raise OSError(f"{ORACLE_SO} missing: run `make -C oracle oracle`")
        lib = C.CDLL(ORACLE_SO)
        _bind(lib, {"orc_" + k: v for k, v in _COMMON.items()})
        _bind(lib, _ORACLE_ONLY)
        _cache["orc"] = lib
    return _cache["orc"]


def ref():
    if "ref" not in _cache:
        if not os.path.exists(REF_SO):
            _cache["ref"] = None
        else:
            lib = C.CDLL(REF_SO)
            _bind(lib, {"oref_" + k: v for k, v in _COMMON.items()})
            _bind(lib, _REF_ONLY)
            _cache["ref"] = lib
    return _cache["ref"]


def fptr(a: np.ndarray):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(_F)


def as_array(ptr, count) -> np.ndarray:
    """Copy `count` floats from a C pointer."""
    return np.ctypeslib.as_array(ptr, shape=(count,)).copy()


def spectrum_tables(lib, prefix: str, desc: OceanDesc) -> dict:
    n = desc.size
    names = ("h0_re", "h0_im", "h0c_re", "h0c_im", "omega", "kx", "kz", "k_inv")
    arrs = {k: np.zeros(n * n, np.float32) for k in names}
    getattr(lib, prefix + "spectrum")(C.byref(desc), *[fptr(arrs[k]) for k in names])
    return arrs


class OracleOcean:
    """The C oracle's ocean object (ocean_oracle.c: orc_ocean_*)."""

    def __init__(self, desc: OceanDesc):
        self.lib = oracle()
        self.desc = desc
        self.h = self.lib.orc_ocean_create(C.byref(desc))
        self.n = desc.size

    def update(self, t: float):
        self.lib.orc_ocean_update(self.h, t)

    def displacement(self) -> np.ndarray:
        return as_array(self.lib.orc_ocean_displacement(self.h), 4 * self.n * self.n)

    def normal(self) -> np.ndarray:
        return as_array(self.lib.orc_ocean_normal(self.h), 4 * self.n * self.n)

    def velocity(self):
        p = self.lib.orc_ocean_velocity(self.h)
        return as_array(p, 4 * self.n * self.n) if p else None

    def sample(self, x: float, z: float) -> OceanSurfaceEx:
        s = OceanSurfaceEx()
        self.lib.orc_ocean_sample(self.h, x, z, C.byref(s))
        return s

    def close(self):
        if self.h:
            self.lib.orc_ocean_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


class RefOcean:
    """The reference's own Ocean, via its C API (oracle/_ref)."""

    def __init__(self, desc: OceanDesc):
        self.lib = ref()
        assert self.lib is not None
        st = C.c_int(-1)
        self.h = self.lib.oref_create(C.byref(desc), C.byref(st))
        assert self.h, st.value
        self.n = desc.size
        self.desc = desc

    def update(self, t: float):
        self.lib.oref_update(self.h, t)

    def displacement(self) -> np.ndarray:
        return as_array(self.lib.oref_get_buffers(self.h).displacement, 4 * self.n * self.n)

    def normal(self) -> np.ndarray:
        return as_array(self.lib.oref_get_buffers(self.h).normal, 4 * self.n * self.n)

    def velocity(self):
        p = self.lib.oref_get_velocity(self.h)
        return as_array(p, 4 * self.n * self.n) if p else None

    def sample(self, x: float, z: float) -> OceanSurfaceEx:
        s = OceanSurfaceEx()
        self.lib.oref_sample_ex(self.h, x, z, C.byref(s))
        return s

    def close(self):
        if self.h:
            self.lib.oref_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


def bits(a: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


def assert_bitexact(got: np.ndarray, want: np.ndarray, what: str = ""):
    g, w = bits(got), bits(want)
    assert g.shape == w.shape, (what, g.shape, w.shape)
    bad = np.nonzero(g != w)[0]
    if bad.size:
        i = bad[0]
        raise AssertionError(f"{what}: {bad.size} of {g.size} words differ; first at {i}: "
                             f"got {got.ravel()[i]!r} want {want.ravel()[i]!r}")
