"""Writes the golden fixtures in tests/golden/ from the reference library.

The reference is compiled from its own, unmodified sources by oracle/Makefile
(`make -C oracle ref`, needs the reference tree at /root/reference) into
oracle/_ref/liboceanref.so; this script drives it through ctypes and stores
only data: raw small buffers (.npz, no pickles) and 64-bit digests of larger
ones (JSON). The fixtures pin the oracle (tests/test_oracle_*.py) and, on a
GPU, the HIP path (tests/test_gpu_*.py) to the reference's own output.

    python tests/golden/make_golden.py                 (all fixtures)
    python tests/golden/make_golden.py config3_frames  (config3_frames.json only)
"""
from __future__ import annotations

import ctypes as C
import hashlib
import json
import os
import platform
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oceanlib_amd.capi import OceanSurfaceEx, default_desc  # noqa: E402
from oracle_ffi import RefOcean, FoamParams, default_foam_params, fptr, ref  # noqa: E402
from oceanlib_amd.workloads import (CONFIG1, CONFIG2, CONFIG3_LEVELS, SMALL_CASES,  # noqa: E402
                                    desc_from, query_points)


def digest(a) -> str:
    """First 16 hex digits of SHA-256 over the raw little-endian bytes."""
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()[:16]


def glibc() -> str:
    return " ".join(platform.libc_ver())


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def small_buffers(r):
    out = {}
    for name, kw, times in SMALL_CASES:
        o = RefOcean(desc_from(kw))
        for t in times:
            o.update(t)
            key = f"{name}_t{t:g}"
            out[key + "_disp"] = o.displacement()
            out[key + "_normal"] = o.normal()
            v = o.velocity()
            if v is not None:
                out[key + "_vel"] = v
        o.close()
    # spectrum tables of the first small case
    d = desc_from(SMALL_CASES[0][1])
    n = d.size
    names = ("h0_re", "h0_im", "h0c_re", "h0c_im", "omega", "kx", "kz", "k_inv")
    arrs = {k: np.zeros(n * n, np.float32) for k in names}
    r.oref_spectrum(C.byref(d), *[fptr(arrs[k]) for k in names])
    for k, v in arrs.items():
        out[f"{SMALL_CASES[0][0]}_spectrum_{k}"] = v
    return out


def run_config1():
    o = RefOcean(desc_from(CONFIG1))
    frames = []
    for k in range(0, 601):
        o.update(k / 60.0)
        frames.append([digest(o.displacement()), digest(o.normal())])
    o.close()
    return frames


def run_config2(r, frames=600):
    """Single heavy-chop patch with velocity + persistent foam + 4096 queries/frame."""
    desc = desc_from(CONFIG2)
    o = RefOcean(desc)
    fp = default_foam_params()
    foam = r.oref_foam_create(o.h, C.byref(fp))
    assert foam
    n = desc.size
    xz = query_points(frames, 4096, desc.patch_length)
    out = []
    surf = OceanSurfaceEx()
    for k in range(frames):
        t = k / 60.0
        o.update(t)
        r.oref_foam_update(foam, np.float32(1.0 / 60.0))
        fbuf = np.ctypeslib.as_array(r.oref_foam_data(foam), shape=(n * n,)).copy()
        q = np.zeros((4096, 10), np.float32)
        for i in range(4096):
            r.oref_sample_ex(o.h, float(xz[k, i, 0]), float(xz[k, i, 1]), C.byref(surf))
            q[i] = np.frombuffer(bytes(surf), np.float32)
        out.append([digest(o.displacement()), digest(o.normal()), digest(o.velocity()),
                    digest(fbuf), digest(q)])
    r.oref_foam_destroy(foam)
    o.close()
    return out


def run_config2_final(r, frames=600, points=256):
    """The heavy-chop patch's persistent foam after all 600 frames:
    FoamField::foam_at at points on and off the patch, and coverage(), as raw
    bits (the history-dependent state at the end of the config)."""
    desc = desc_from(CONFIG2)
    o = RefOcean(desc)
    fp = default_foam_params()
    foam = r.oref_foam_create(o.h, C.byref(fp))
    assert foam
    for k in range(frames):
        o.update(k / 60.0)
        r.oref_foam_update(foam, np.float32(1.0 / 60.0))
    rng = np.random.default_rng(0xF0A)
    xz = rng.uniform(-200.0, 450.0, size=(points, 2)).astype(np.float32)
    vals = np.array([r.oref_foam_at(foam, float(x), float(z)) for x, z in xz], np.float32)
    cov = np.float64(r.oref_foam_coverage(foam))
    r.oref_foam_destroy(foam)
    o.close()
    return {"frames": frames, "points": xz.view(np.uint32).tolist(),
            "foam_at": vals.view(np.uint32).tolist(), "coverage": int(cov.view(np.uint64))}


def run_config3_cascades(r, frames=(0, 1, 2, 60, 600, 1199)):
    levels = [desc_from(kw) for kw in CONFIG3_LEVELS]
    arr = (type(levels[0]) * len(levels))(*levels)
    h = r.oref_cascade_create(arr, len(levels))
    assert h
    out = {}
    surf = OceanSurfaceEx()
    pts = query_points(1, 512, 800.0)[0]
    for k in frames:
        t = k / 60.0
        r.oref_cascade_update(h, t)
        rows = []
        for li, d in enumerate(levels):
            n = d.size
            disp = np.ctypeslib.as_array(r.oref_cascade_displacement(h, li), shape=(4 * n * n,))
            nrm = np.ctypeslib.as_array(r.oref_cascade_normal(h, li), shape=(4 * n * n,))
            rows.append([digest(disp), digest(nrm)])
        q = np.zeros((len(pts), 10), np.float32)
        for i, (x, z) in enumerate(pts):
            r.oref_cascade_sample(h, float(x), float(z), C.byref(surf))
            q[i] = np.frombuffer(bytes(surf), np.float32)
        out[str(k)] = {"levels": rows, "queries": digest(q)}
    r.oref_cascade_destroy(h)
    return out


def run_config3_interaction(r, frames=1200, every=50):
    """The three-cascade scene's interaction field: 256^2 cells of 0.1 m,
    radius 6, eight moving continuous sources, one Neumann hull."""
    from iwave_helpers import Field, idesc
    from oceanlib_amd.workloads import (CONFIG3_INTERACTION, config3_disturbances,
                                        config3_hull_mask, config3_sources)
    f = Field("oref", r, idesc(r.oref_interaction_desc_init, **CONFIG3_INTERACTION))
    f.set_obstruction(config3_hull_mask())
    srcs = config3_sources()
    out = {}
    for k in range(frames):
        for rec in config3_disturbances(srcs, k / 60.0):
            f.add(rec)
        f.update(1.0 / 60.0)
        if k % every == 0 or k == frames - 1:
            out[str(k)] = digest(f.field())
    f.close()
    return out


def run_config3_every_frame(r, frames=1200):
    """The three-cascade scene at every one of its 1200 frames: each level's
    displacement and normal digests and the 512 cascade queries' digest, and
    the interaction field's digest after every substep."""
    levels = [desc_from(kw) for kw in CONFIG3_LEVELS]
    arr = (type(levels[0]) * len(levels))(*levels)
    h = r.oref_cascade_create(arr, len(levels))
    assert h
    surf = OceanSurfaceEx()
    pts = query_points(1, 512, 800.0)[0]
    cascade = []
    for k in range(frames):
        r.oref_cascade_update(h, k / 60.0)
        row = []
        for li, d in enumerate(levels):
            n = d.size
            row.append(digest(np.ctypeslib.as_array(r.oref_cascade_displacement(h, li), shape=(4 * n * n,))))
            row.append(digest(np.ctypeslib.as_array(r.oref_cascade_normal(h, li), shape=(4 * n * n,))))
        q = np.zeros((len(pts), 10), np.float32)
        for i, (x, z) in enumerate(pts):
            r.oref_cascade_sample(h, float(x), float(z), C.byref(surf))
            q[i] = np.frombuffer(bytes(surf), np.float32)
        row.append(digest(q))
        cascade.append(row)
    r.oref_cascade_destroy(h)
    every = run_config3_interaction(r, frames=frames, every=1)
    return {"cascade": cascade, "interaction": [every[str(k)] for k in range(frames)]}


def write_config3_frames(r):
    with open(os.path.join(HERE, "config3_frames.json"), "w") as f:
        json.dump(run_config3_every_frame(r), f, indent=0)


def write_config2_final(r):
    with open(os.path.join(HERE, "config2_final.json"), "w") as f:
        json.dump(run_config2_final(r), f)


def main():
    r = ref()
    if r is None:
        raise SystemExit("oracle/_ref/liboceanref.so missing: run `make -C oracle ref` first")
    small = small_buffers(r)
    np.savez_compressed(os.path.join(HERE, "small_buffers.npz"), **small)
    meta = {
        "generator": "tests/golden/make_golden.py",
        "reference_build": "oracle/Makefile target ref (g++ -O3 -ffp-contract=off)",
        "glibc": glibc(),
        "cpu": cpu_model(),
        "digest": "sha256 of raw float32 bytes, first 16 hex digits",
    }
    print("config1 ...", flush=True)
    c1 = run_config1()
    print("config2 ...", flush=True)
    c2 = run_config2(r)
    print("config3 cascades ...", flush=True)
    c3 = run_config3_cascades(r)
    print("config3 interaction ...", flush=True)
    c3i = run_config3_interaction(r)
    with open(os.path.join(HERE, "digests.json"), "w") as f:
        json.dump({"meta": meta, "config1": c1, "config2": c2, "config3_cascade": c3,
                   "config3_interaction": c3i}, f, indent=0)
    print("config2 final foam ...", flush=True)
    write_config2_final(r)
    print("wrote", os.path.join(HERE, "small_buffers.npz"), os.path.join(HERE, "digests.json"),
          os.path.join(HERE, "config2_final.json"))


if __name__ == "__main__":
    if sys.argv[1:] == ["config2_final"]:
        write_config2_final(ref())
    elif sys.argv[1:] == ["config3_frames"]:
        write_config3_frames(ref())
    else:
        main()
        write_config3_frames(ref())
