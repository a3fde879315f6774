"""The C boundary, checked without a GPU.

- every function include/ocean/*.h declares is exported by the built library;
- the headers compile as C (reference tests/c_header_check.c);
- struct layouts are byte-identical to the reference header's (offsetof and
  sizeof of every field, compiled against both headers when the reference
  tree is present; otherwise against the ctypes mirror in capi.py);
- the host-side table build inside the library equals the oracle's tables.
"""
import ctypes as C
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from oceanlib_amd import capi
from oceanlib_amd.workloads import CONFIG1, CONFIG2, CONFIG3_LEVELS, desc_from
from oracle_ffi import assert_bitexact, fptr, oracle, spectrum_tables

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")
REF_INCLUDE = "/root/reference/include"
HEADERS = [os.path.join(INCLUDE, "ocean", h) for h in ("ocean.h", "ocean_amd.h")]


def declared_functions(path):
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    names = set()
    for m in re.finditer(r"\b(ocean_[a-z0-9_]+)\s*\(", text):
        name = m.group(1)
        before = text[:m.start()].rstrip()
        if before.endswith("typedef") or before.endswith("(*"):
            continue
        names.add(name)
    # function-pointer typedef names are not symbols
    names -= {"ocean_task_fn", "ocean_parallel_for_fn"}
    return names


def test_every_declared_symbol_is_exported():
    lib = capi.load()
    declared = set()
    for h in HEADERS:
        declared |= declared_functions(h)
    assert len(declared_functions(HEADERS[0])) == 29, sorted(declared_functions(HEADERS[0]))
    missing = [n for n in sorted(declared) if not hasattr(lib, n)]
    assert not missing, missing
    # and the ctypes mirror binds exactly the declared set
    assert declared == set(capi.SIGNATURES) | set(capi.AMD_SIGNATURES)


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc missing")
def test_headers_compile_as_c(tmp_path):
    src = tmp_path / "check.c"
    src.write_text('#include "ocean/ocean.h"\n#include "ocean/ocean_amd.h"\n'
                   "int main(void) { ocean_desc d; ocean_desc_init(&d); return (int)d.size; }\n")
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-pedantic", "-fsyntax-only",
                    "-I", INCLUDE, str(src)], check=True)


LAYOUT_PROGRAM = r"""
#include <stdio.h>
#include <stddef.h>
#include "ocean/ocean.h"
#define F(T, f) printf(#T "." #f " %zu %zu\n", offsetof(T, f), sizeof(((T*)0)->f))
#define S(T) printf(#T " size %zu align %zu\n", sizeof(T), _Alignof(T))
int main(void) {
  S(ocean_spectrum_desc); S(ocean_desc); S(ocean_buffers); S(ocean_surface);
  S(ocean_surface_ex); S(ocean_interaction_desc); S(ocean_disturbance);
  S(ocean_interaction_buffers); S(ocean_interaction_sample);
  F(ocean_desc, struct_size); F(ocean_desc, size); F(ocean_desc, patch_length);
  F(ocean_desc, spectrum); F(ocean_desc, choppiness); F(ocean_desc, foam_threshold);
  F(ocean_desc, seed); F(ocean_desc, parallel_for); F(ocean_desc, parallel_for_user);
  F(ocean_desc, thread_count); F(ocean_desc, water_depth); F(ocean_desc, compute_velocity);
  F(ocean_spectrum_desc, wind_speed); F(ocean_spectrum_desc, gravity);
  F(ocean_buffers, displacement); F(ocean_buffers, normal); F(ocean_buffers, size);
  F(ocean_buffers, patch_length);
  F(ocean_surface_ex, foam); F(ocean_surface_ex, velocity_z);
  F(ocean_interaction_desc, struct_size); F(ocean_interaction_desc, size);
  F(ocean_interaction_desc, obstruction); F(ocean_interaction_desc, parallel_for);
  F(ocean_interaction_desc, thread_count);
  F(ocean_disturbance, kind); F(ocean_interaction_buffers, origin_z);
  F(ocean_interaction_sample, velocity_y);
  printf("enums %d %d %d %d %d\n", OCEAN_ERROR_UNKNOWN, OCEAN_OBSTRUCTION_DIRICHLET,
         OCEAN_SOURCE_CONTINUOUS, (int)sizeof(ocean_status), (int)sizeof(ocean_obstruction));
  return 0;
}
"""


def _layout(tmp_path, include_dir, tag):
    src = tmp_path / f"layout_{tag}.c"
    exe = tmp_path / f"layout_{tag}"
    src.write_text(LAYOUT_PROGRAM)
    subprocess.run(["gcc", "-std=c11", "-I", include_dir, str(src), "-o", str(exe)], check=True)
    return subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc missing")
def test_struct_layouts(tmp_path):
    ours = _layout(tmp_path, INCLUDE, "ours")
    if os.path.isdir(REF_INCLUDE):
        assert ours == _layout(tmp_path, REF_INCLUDE, "ref")
    # the ctypes mirror agrees with the C compiler
    sizes = dict(line.split()[0:3:2] for line in ours.splitlines() if " size " in line)
    assert int(sizes["ocean_desc"]) == C.sizeof(capi.OceanDesc)
    assert int(sizes["ocean_interaction_desc"]) == C.sizeof(capi.InteractionDesc)
    assert int(sizes["ocean_surface_ex"]) == C.sizeof(capi.OceanSurfaceEx)
    assert int(sizes["ocean_disturbance"]) == C.sizeof(capi.Disturbance)


def test_desc_defaults_match_reference():
    """ocean_desc_init fills the reference's defaults (include/ocean/ocean.hpp:39-130)."""
    lib = capi.load()
    d = capi.OceanDesc()
    lib.ocean_desc_init(C.byref(d))
    assert d.struct_size == C.sizeof(capi.OceanDesc)
    assert (d.size, d.patch_length, d.choppiness, d.foam_threshold, d.seed) == \
        (256, 200.0, 1.0, 0.5, 1337)
    s = d.spectrum
    assert (s.wind_speed, s.fetch, s.wind_direction, s.swell, s.small_wave_cutoff) == \
        (10.0, 100000.0, 0.0, 0.0, 0.5)
    assert abs(s.peak_enhancement - 3.3) < 1e-6 and abs(s.gravity - 9.81) < 1e-6
    assert d.water_depth == 0.0 and d.compute_velocity == 0


def test_create_rejects_bad_descriptors_without_device():
    """Validation happens before any device work (reference src/ocean.cpp:28-59,
    src/c_api.cpp:46-60)."""
    lib = capi.load()
    st = C.c_int(-1)
    assert not lib.ocean_create(None, C.byref(st)) and st.value == capi.OCEAN_ERROR_INVALID_ARG
    for kw in (dict(size=100), dict(size=8), dict(size=8192), dict(patch_length=0.0),
               dict(gravity=0.0), dict(fetch=-1.0), dict(wind_speed=-1.0),
               dict(peak_enhancement=0.5), dict(choppiness=-0.1)):
        d = capi.default_desc(lib, **kw)
        st = C.c_int(-1)
        assert not lib.ocean_create(C.byref(d), C.byref(st)), kw
        assert st.value == capi.OCEAN_ERROR_INVALID_ARG, kw
    d = capi.default_desc(lib)
    d.struct_size = 0
    assert not lib.ocean_create(C.byref(d), C.byref(st))
    d.struct_size = C.sizeof(capi.OceanDesc) + 8
    assert not lib.ocean_create(C.byref(d), C.byref(st))
    assert st.value == capi.OCEAN_ERROR_INVALID_ARG


def test_null_handles_are_tolerated():
    lib = capi.load()
    lib.ocean_destroy(None)
    lib.ocean_update(None, 1.0)
    b = lib.ocean_get_buffers(None)
    assert not b.displacement and b.size == 0
    assert lib.ocean_get_time(None) == 0.0
    assert lib.ocean_height_at(None, 1.0, 2.0) == 0.0
    s = lib.ocean_sample_at(None, 1.0, 2.0)
    assert s.normal_y == 1.0 and s.height == 0.0
    assert not lib.ocean_get_velocity(None)
    assert lib.ocean_status_string(1) == b"invalid argument"


@pytest.mark.parametrize("kw", [CONFIG1, CONFIG2, CONFIG3_LEVELS[0],
                                dict(size=16, wind_speed=40.0, swell=1.0, seed=3)])
def test_library_host_tables_equal_oracle(kw):
    """The tables the library uploads (built on the host, clang) equal the
    oracle's (gcc) and therefore the reference's."""
    lib = capi.load()
    d = desc_from(kw)
    n = d.size
    h0 = np.zeros(2 * n * n, np.float32)
    om = np.zeros(n * n, np.float32)
    assert lib.ocean_amd_build_spectrum(C.byref(d), fptr(h0), fptr(om)) == 0
    t = spectrum_tables(oracle(), "orc_", d)
    assert_bitexact(h0[0::2], t["h0_re"], "h0 re")
    assert_bitexact(h0[1::2], t["h0_im"], "h0 im")
    assert_bitexact(om, t["omega"], "omega")
    tw = np.zeros(n, np.float32)
    assert lib.ocean_amd_build_twiddles(n, fptr(tw)) == 0
    a = np.zeros(n // 2, np.float32)
    b = np.zeros(n // 2, np.float32)
    oracle().orc_twiddles(n, fptr(a), fptr(b))
    assert_bitexact(tw[0::2], a, "tw cos")
    assert_bitexact(tw[1::2], b, "tw sin")


def test_batch_create_rejects_bad_input_without_device():
    """A batch validates every level like ocean_create before any device work."""
    lib = capi.load()
    st = C.c_int(-1)
    assert not lib.ocean_amd_batch_create(None, 3, C.byref(st))
    assert st.value == capi.OCEAN_ERROR_INVALID_ARG
    good = capi.default_desc(lib, size=64)
    arr = (capi.OceanDesc * 2)(good, good)
    st = C.c_int(-1)
    assert not lib.ocean_amd_batch_create(arr, 0, C.byref(st))
    assert st.value == capi.OCEAN_ERROR_INVALID_ARG
    for kw in (dict(size=100), dict(patch_length=-1.0), dict(choppiness=-0.1)):
        arr = (capi.OceanDesc * 2)(good, capi.default_desc(lib, **kw))
        st = C.c_int(-1)
        assert not lib.ocean_amd_batch_create(arr, 2, C.byref(st)), kw
        assert st.value == capi.OCEAN_ERROR_INVALID_ARG, kw
    bad = capi.default_desc(lib)
    bad.struct_size = 0
    arr = (capi.OceanDesc * 2)(good, bad)
    assert not lib.ocean_amd_batch_create(arr, 2, C.byref(st))
    assert st.value == capi.OCEAN_ERROR_INVALID_ARG


def test_batch_null_handles_are_tolerated():
    lib = capi.load()
    lib.ocean_amd_batch_destroy(None)
    lib.ocean_amd_batch_update(None, 1.0)
    lib.ocean_amd_batch_autotune(None, 3)
    assert lib.ocean_amd_batch_level_count(None) == 0
    assert not lib.ocean_amd_batch_level(None, 0)
    assert lib.ocean_amd_batch_group_count(None) == 0
    g = capi.BatchGroup()
    assert lib.ocean_amd_batch_group_info(None, 0, C.byref(g)) == capi.OCEAN_ERROR_INVALID_ARG
    assert lib.ocean_amd_batch_synchronize(None) == capi.OCEAN_ERROR_INVALID_ARG
    assert lib.ocean_amd_batch_last_error(None) == 0
    out = (capi.OceanSurfaceEx * 1)()
    xz = np.zeros(2, np.float32)
    assert lib.ocean_amd_batch_sample(None, 0, 1, fptr(xz), 1, out) == capi.OCEAN_ERROR_INVALID_ARG
