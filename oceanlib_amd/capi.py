"""ctypes mirror of the C ABI in include/ocean/ocean.h and include/ocean/ocean_amd.h.

The struct layouts are the reference's (reference include/ocean/ocean.h:35-415);
`OceanDesc` etc. are byte-compatible with what a C caller passes. This module
is the host-side Python binding a maintainer would use to call the library;
the parity tests and bench.py call the library through it.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboceanlib_amd.so")

OCEAN_OK = 0
OCEAN_ERROR_INVALID_ARG = 1
OCEAN_ERROR_OUT_OF_MEMORY = 2
OCEAN_ERROR_UNKNOWN = 3

TASK_FN = C.CFUNCTYPE(None, C.c_void_p, C.c_uint32)
PARALLEL_FOR_FN = C.CFUNCTYPE(None, C.c_void_p, TASK_FN, C.c_void_p, C.c_uint32)


class SpectrumDesc(C.Structure):
    _fields_ = [
        ("wind_speed", C.c_float),
        ("fetch", C.c_float),
        ("wind_direction", C.c_float),
        ("peak_enhancement", C.c_float),
        ("swell", C.c_float),
        ("small_wave_cutoff", C.c_float),
        ("gravity", C.c_float),
    ]


class OceanDesc(C.Structure):
    _fields_ = [
        ("struct_size", C.c_size_t),
        ("size", C.c_uint32),
        ("patch_length", C.c_float),
        ("spectrum", SpectrumDesc),
        ("choppiness", C.c_float),
        ("foam_threshold", C.c_float),
        ("seed", C.c_uint64),
        ("parallel_for", C.c_void_p),
        ("parallel_for_user", C.c_void_p),
        ("thread_count", C.c_uint32),
        ("water_depth", C.c_float),
        ("compute_velocity", C.c_int32),
    ]


class OceanBuffers(C.Structure):
    _fields_ = [
        ("displacement", C.POINTER(C.c_float)),
        ("normal", C.POINTER(C.c_float)),
        ("size", C.c_uint32),
        ("patch_length", C.c_float),
    ]


class OceanSurface(C.Structure):
    _fields_ = [(n, C.c_float) for n in
                ("height", "offset_x", "offset_z", "normal_x", "normal_y", "normal_z", "foam")]


class OceanSurfaceEx(C.Structure):
    _fields_ = [(n, C.c_float) for n in
                ("height", "offset_x", "offset_z", "normal_x", "normal_y", "normal_z", "foam",
                 "velocity_x", "velocity_y", "velocity_z")]


class InteractionDesc(C.Structure):
    _fields_ = [
        ("struct_size", C.c_size_t),
        ("size", C.c_uint32),
        ("extent", C.c_float),
        ("kernel_radius", C.c_uint32),
        ("damping", C.c_float),
        ("absorb_cells", C.c_uint32),
        ("fixed_dt", C.c_float),
        ("max_substeps", C.c_uint32),
        ("max_sources", C.c_uint32),
        ("gravity", C.c_float),
        ("water_depth", C.c_float),
        ("obstruction", C.c_int32),
        ("parallel_for", C.c_void_p),
        ("parallel_for_user", C.c_void_p),
        ("thread_count", C.c_uint32),
    ]


class Disturbance(C.Structure):
    _fields_ = [(n, C.c_float) for n in
                ("world_x", "world_z", "radius", "strength", "velocity_y", "velocity_x",
                 "velocity_z")] + [("kind", C.c_int32)]


class InteractionBuffers(C.Structure):
    _fields_ = [
        ("field", C.POINTER(C.c_float)),
        ("size", C.c_uint32),
        ("extent", C.c_float),
        ("origin_x", C.c_float),
        ("origin_z", C.c_float),
    ]


class InteractionSample(C.Structure):
    _fields_ = [(n, C.c_float) for n in ("height", "slope_x", "slope_z", "velocity_y")]


class FoamDesc(C.Structure):
    """ocean_amd_foam_desc (include/ocean/ocean_amd.h)."""
    _fields_ = [
        ("struct_size", C.c_size_t),
        ("decay", C.c_float),
        ("source_gain", C.c_float),
        ("advect_scale", C.c_float),
        ("wind_drift_x", C.c_float),
        ("wind_drift_z", C.c_float),
        ("fixed_dt", C.c_float),
        ("max_substeps", C.c_uint32),
    ]


class DeviceBuffers(C.Structure):
    """ocean_amd_device_buffers (include/ocean/ocean_amd.h)."""
    _fields_ = [("displacement", C.c_void_p), ("normal", C.c_void_p), ("velocity", C.c_void_p),
                ("size", C.c_uint32), ("patch_length", C.c_float), ("stream", C.c_void_p)]


class BatchGroup(C.Structure):
    """ocean_amd_batch_group (include/ocean/ocean_amd.h)."""
    _fields_ = [(n, C.c_uint32) for n in
                ("size", "velocity", "levels", "items", "grid_row", "grid_col")] + \
               [("row_ms", C.c_float), ("col_ms", C.c_float), ("radix_row", C.c_uint32),
                ("radix_col", C.c_uint32), ("layout", C.c_uint32)]


PROPOSAL_PARAMS = 9360
PROPOSAL_INPUTS = 64
PROPOSAL_OUTPUTS = 16


# (name, restype, argtypes) for every symbol the headers declare.
_P = C.c_void_p
_F = C.POINTER(C.c_float)
SIGNATURES = {
    # ---- include/ocean/ocean.h (reference ABI) ----
    "ocean_status_string": (C.c_char_p, [C.c_int]),
    "ocean_desc_init": (None, [C.POINTER(OceanDesc)]),
    "ocean_create": (_P, [C.POINTER(OceanDesc), C.POINTER(C.c_int)]),
    "ocean_destroy": (None, [_P]),
    "ocean_update": (None, [_P, C.c_double]),
    "ocean_get_buffers": (OceanBuffers, [_P]),
    "ocean_get_time": (C.c_double, [_P]),
    "ocean_height_at": (C.c_float, [_P, C.c_float, C.c_float]),
    "ocean_sample_at": (OceanSurface, [_P, C.c_float, C.c_float]),
    "ocean_sample_ex": (None, [_P, C.c_float, C.c_float, C.POINTER(OceanSurfaceEx)]),
    "ocean_get_velocity": (_F, [_P]),
    "ocean_simd_level": (C.c_char_p, []),
    "ocean_force_simd_level": (C.c_char_p, [C.c_char_p]),
    "ocean_version": (None, [C.POINTER(C.c_uint32)] * 3),
    "ocean_interaction_desc_init": (None, [C.POINTER(InteractionDesc)]),
    "ocean_interaction_create": (_P, [C.POINTER(InteractionDesc), C.POINTER(C.c_int)]),
    "ocean_interaction_destroy": (None, [_P]),
    "ocean_interaction_add": (None, [_P, C.POINTER(Disturbance)]),
    "ocean_interaction_dropped": (C.c_uint64, [_P]),
    "ocean_interaction_update": (None, [_P, C.c_float]),
    "ocean_interaction_recenter": (C.c_int32, [_P, C.c_float, C.c_float]),
    "ocean_interaction_set_obstruction": (None, [_P, C.POINTER(C.c_uint8)]),
    "ocean_interaction_get_buffers": (InteractionBuffers, [_P]),
    "ocean_interaction_sample_at": (None, [_P, C.c_float, C.c_float,
                                           C.POINTER(InteractionSample)]),
    "ocean_interaction_height_at": (C.c_float, [_P, C.c_float, C.c_float]),
    "ocean_interaction_fixed_dt": (C.c_float, [_P]),
    "ocean_interaction_dt_limit": (C.c_float, [_P]),
    "ocean_interaction_dispersion_error": (C.c_float, [_P]),
    "ocean_sample_combined": (None, [_P, _P, C.c_float, C.c_float,
                                     C.POINTER(OceanSurfaceEx)]),
}

# include/ocean/ocean_amd.h (additive entry points).
AMD_SIGNATURES = {
    "ocean_amd_get_device_buffers": (C.c_int, [_P, C.c_void_p]),
    "ocean_amd_bind_outputs": (C.c_int, [_P, C.c_void_p, C.c_void_p, C.c_void_p]),
    "ocean_amd_synchronize": (None, [_P]),
    "ocean_amd_last_error": (C.c_int, [_P]),
    "ocean_amd_profile_frame": (C.c_int, [_P, C.c_double, C.c_double, C.c_uint32, _F, _F]),
    "ocean_amd_get_tables": (C.c_int, [_P, _F, _F, C.POINTER(C.c_uint32)]),
    "ocean_amd_frame_shapes": (C.c_int, [_P, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    "ocean_amd_set_frame_layout": (C.c_int, [_P, C.c_uint32]),
    "ocean_amd_get_frame_layout": (C.c_uint32, [_P]),
    "ocean_amd_set_frame_shape": (C.c_int, [_P, C.c_uint32, C.c_uint32]),
    "ocean_amd_set_precision": (C.c_int, [_P, C.c_uint32]),
    "ocean_amd_get_precision": (C.c_uint32, [_P]),
    "ocean_amd_frame_timing": (C.c_int, [_P, C.c_uint32]),
    "ocean_amd_frame_timing_read": (C.c_int, [_P, C.POINTER(C.c_float), C.POINTER(C.c_float),
                                              C.POINTER(C.c_uint32)]),
    "ocean_amd_autotune_frame": (C.c_int, [_P, C.c_uint32, C.POINTER(C.c_uint32),
                                           C.POINTER(C.c_uint32), _F, _F]),
    "ocean_amd_sample_batch": (C.c_int, [_P, _F, C.c_uint32, C.POINTER(OceanSurfaceEx)]),
    "ocean_amd_sample_batch_device": (C.c_int, [_P, C.c_void_p, C.c_uint32, C.c_void_p,
                                                C.c_void_p]),
    "ocean_amd_build_spectrum": (C.c_int, [C.POINTER(OceanDesc), _F, _F]),
    "ocean_amd_build_twiddles": (C.c_int, [C.c_uint32, _F]),
    "ocean_amd_build_iwave_taps": (C.c_int, [C.c_uint32, C.c_double, C.c_double, _F,
                                             C.POINTER(C.c_double)]),
    "ocean_amd_build_iwave_stencil": (C.c_int, [C.c_uint32, C.c_double, C.c_double, _F,
                                                C.POINTER(C.c_double)]),
    "ocean_amd_foam_desc_init": (None, [C.POINTER(FoamDesc)]),
    "ocean_amd_foam_create": (_P, [_P, C.POINTER(FoamDesc), C.POINTER(C.c_int)]),
    "ocean_amd_foam_destroy": (None, [_P]),
    "ocean_amd_foam_update": (None, [_P, C.c_float]),
    "ocean_amd_foam_clear": (None, [_P]),
    "ocean_amd_foam_data": (_F, [_P]),
    "ocean_amd_foam_device_data": (C.c_void_p, [_P]),
    "ocean_amd_foam_update_sample_device": (C.c_int, [_P, C.c_float, _P, C.c_uint32, _P, _P]),
    # renderer handoff (RGBA16F); outputs as void* (host or device arrays of uint16)
    "ocean_amd_pack_rgba16": (C.c_int, [_P, _P, _P, _P]),
    "ocean_amd_pack_rgba16_device": (C.c_int, [_P, _P, _P, _P]),
    "ocean_amd_pack_half_device": (C.c_int, [_P, _P, _P, C.c_uint64]),
    "ocean_amd_interaction_pack_rgba16": (C.c_int, [_P, _P]),
    "ocean_amd_interaction_pack_rgba16_device": (C.c_int, [_P, _P]),
    "ocean_amd_interaction_energy": (C.c_double, [_P]),
    "ocean_amd_interaction_obstruction_tables": (C.c_uint32, [_P, C.POINTER(C.c_uint32),
                                                              C.POINTER(C.c_uint32), C.c_uint32]),
    "ocean_amd_foam_last_substeps": (C.c_uint32, [_P]),
    "ocean_amd_foam_at": (C.c_float, [_P, C.c_float, C.c_float]),
    "ocean_amd_foam_coverage": (C.c_double, [_P]),
    "ocean_amd_cascade_create": (_P, [C.POINTER(OceanDesc), C.c_uint32, C.POINTER(C.c_int)]),
    "ocean_amd_cascade_destroy": (None, [_P]),
    "ocean_amd_cascade_update": (None, [_P, C.c_double]),
    "ocean_amd_cascade_level_count": (C.c_uint32, [_P]),
    "ocean_amd_cascade_get_buffers": (OceanBuffers, [_P, C.c_uint32]),
    "ocean_amd_cascade_sample": (None, [_P, C.c_float, C.c_float, C.POINTER(OceanSurfaceEx)]),
    "ocean_amd_cascade_sample_batch": (C.c_int, [_P, _F, C.c_uint32,
                                                 C.POINTER(OceanSurfaceEx)]),
    "ocean_amd_batch_create": (_P, [C.POINTER(OceanDesc), C.c_uint32, C.POINTER(C.c_int)]),
    "ocean_amd_batch_destroy": (None, [_P]),
    "ocean_amd_batch_update": (None, [_P, C.c_double]),
    "ocean_amd_batch_level_count": (C.c_uint32, [_P]),
    "ocean_amd_batch_level": (_P, [_P, C.c_uint32]),
    "ocean_amd_batch_sample": (C.c_int, [_P, C.c_uint32, C.c_uint32, _F, C.c_uint32,
                                         C.POINTER(OceanSurfaceEx)]),
    "ocean_amd_batch_sample_device": (C.c_int, [_P, C.c_uint32, C.c_uint32, C.c_void_p,
                                                C.c_uint32, C.c_void_p, C.c_void_p]),
    "ocean_amd_batch_autotune": (None, [_P, C.c_uint32]),
    "ocean_amd_batch_set_precision": (C.c_int, [_P, C.c_uint32]),
    "ocean_amd_batch_group_count": (C.c_uint32, [_P]),
    "ocean_amd_batch_group_info": (C.c_int, [_P, C.c_uint32, C.POINTER(BatchGroup)]),
    "ocean_amd_batch_set_group_launch": (C.c_int, [_P] + [C.c_uint32] * 6),
    "ocean_amd_batch_checksums": (C.c_int, [_P, C.c_uint32, C.c_uint32,
                                            C.POINTER(C.c_uint64)]),
    "ocean_amd_batch_synchronize": (C.c_int, [_P]),
    "ocean_amd_proposal_create": (_P, [_F, C.c_uint32, C.POINTER(C.c_int)]),
    "ocean_amd_proposal_destroy": (None, [_P]),
    "ocean_amd_proposal_set_params": (C.c_int, [_P, _F]),
    "ocean_amd_proposal_get_params": (C.c_int, [_P, _F]),
    "ocean_amd_proposal_forward": (C.c_int, [_P, C.c_void_p, C.c_uint32, C.c_void_p]),
    "ocean_amd_proposal_train_step": (C.c_int, [_P, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32,
                                                C.c_float, _F]),
    "ocean_amd_proposal_gradient": (C.c_int, [_P, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32,
                                              _F, _F]),
    "ocean_amd_proposal_synchronize": (C.c_int, [_P]),
    "ocean_amd_batch_height_converged": (C.c_int, [_P, C.c_uint32, C.c_uint32, C.c_void_p,
                                                   C.c_uint32, C.c_float, _P, C.c_void_p,
                                                   C.c_void_p, C.c_void_p]),
    "ocean_amd_batch_height_converged_static": (C.c_int, [_P, C.c_uint32, C.c_uint32, C.c_void_p,
                                                          C.c_uint32, C.c_float, _P, C.c_void_p,
                                                          C.c_void_p, C.c_void_p]),
    "ocean_amd_cascade_height_converged": (C.c_int, [_P, C.c_void_p, C.c_uint32, C.c_float, _P,
                                                     C.c_void_p, C.c_void_p, C.c_void_p]),
    "ocean_amd_proposal_starts": (C.c_int, [_P, C.c_uint32, C.c_void_p, C.c_uint32, _P, C.c_void_p,
                                            C.c_void_p]),
    "ocean_amd_proposal_pairs": (C.c_int, [_P, C.c_uint32, C.c_void_p, C.c_uint32, C.c_float,
                                           C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "ocean_amd_batch_last_error": (C.c_int, [_P]),
}


def load(path: str | None = None) -> C.CDLL:
    """Load the HIP library and bind every declared symbol.

    Raises OSError when the library has not been built: there is no fallback.
    """
    path = path or LIB_PATH
    if not os.path.exists(path):
        raise OSError(f"{path} is missing: build it first (python -c "
                      "'import __graft_entry__ as g; g.build()')")
    lib = C.CDLL(path)
    for name, (res, args) in {**SIGNATURES, **AMD_SIGNATURES}.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def default_desc(lib=None, **kw) -> OceanDesc:
    """OceanDesc with the reference defaults (include/ocean/ocean.hpp:39-130)."""
    d = OceanDesc()
    if lib is not None:
        lib.ocean_desc_init(C.byref(d))
    else:
        d.struct_size = C.sizeof(OceanDesc)
        d.size = 256
        d.patch_length = 200.0
        d.spectrum = SpectrumDesc(10.0, 100000.0, 0.0, 3.3, 0.0, 0.5, 9.81)
        d.choppiness = 1.0
        d.foam_threshold = 0.5
        d.seed = 1337
        d.thread_count = 0
        d.water_depth = 0.0
        d.compute_velocity = 0
    spec_keys = {f[0] for f in SpectrumDesc._fields_}
    for k, v in kw.items():
        if k in spec_keys:
            setattr(d.spectrum, k, v)
        else:
            setattr(d, k, v)
    return d
