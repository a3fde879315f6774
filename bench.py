#!/usr/bin/env python3
"""Benchmark of the per-frame wave synthesis on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Workload (one "step" = one frame of it): the single heavy-chop patch -
N = 512, 250 m, 22 m/s, choppiness 1.3, orbital velocity on, finite depth
50 m, seed 7, t advancing 1/60 s per frame; each frame is
    ocean_update(t) -> FoamField step (dt = 1/60) -> 4096 height_at queries,
all enqueued through the C ABI on the ocean's stream with every input already
resident in device memory. The timed region is K frames bracketed by a
device-wide synchronize on both sides.

Prints one JSON line: value = frames (updates) per second, plus
  roofline     : the dominant kernel's algorithmic bytes per launch over its
                 mean device time (HIP events on the ocean's stream);
  cpu_baseline : the reference library itself (oracle/_ref, compiled from the
                 reference's sources, AVX2 kernels + its thread pool) running
                 a bounded sample of the same frames on this host's cores.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "updates/s at N=512 (1 and batched) vs CPU; height_at queries/s @0.01mm"
HBM_PEAK_GBS = 8000.0
TIMING_EVERY = 16
QUERIES_PER_FRAME = 4096


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def algorithmic_bytes(n: int, fields: int, velocity: bool):
    """Per-launch algorithmic HBM bytes of the two frame kernels.

    row pass : reads h0 (8 B) + omega (4 B) + 1/|k| (4 B) per cell once,
               writes the F complex fields (8F B per cell);
    col pass : reads the F fields (8F B), writes displacement + normal
               (+ velocity) RGBA32F (16 B each per cell).
    """
    cells = n * n
    row = cells * (16 + 8 * fields)
    col = cells * (8 * fields + 16 * (3 if velocity else 2))
    return row, col


def frame_bytes(n: int, fields: int, velocity: bool) -> int:
    """Algorithmic HBM bytes of one level's frame: h0 + omega + 1/|k| read
    once (16 B/cell), RGBA32F outputs written once (32 B/cell, +16 with
    velocity), and, for N >= 256 only, one write + read of the F complex
    fields between the passes (16F B/cell; a smaller level's fields could stay
    on chip)."""
    inter = 16 * fields if n >= 256 else 0
    return n * n * (16 + 32 + (16 if velocity else 0) + inter)


def traffic_from_profiles(kernels):
    """HBM bytes per launch (summed over `kernels`) measured with rocprofv3
    --pmc FETCH_SIZE and WRITE_SIZE, each scaled by what the counter reads for
    a 256 MiB copy (profiles/traffic_pmc.md); None if not recorded. An entry
    may be a tuple of names of the same pass (the first one recorded counts)."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            rec = json.load(f)["kernels"]
        total = 0.0
        for k in kernels:
            names = k if isinstance(k, tuple) else (k,)
            hit = next(n for n in names if n in rec)
            total += rec[hit]["traffic_bytes"]
        return total
    except (OSError, KeyError, ValueError, TypeError, StopIteration):
        return None


def batch_kernel_names(g, p):
    """The names (as the profiler prints them, arguments dropped) a batch
    group's row or column pass dispatches under: the looped or loop-free
    kernel of its radix, mode 0 (bit-exact) and field layout."""
    n, f = g["size"], (6 if g["velocity"] else 4)
    radix = g["radix_row"] if p == "row" else g["radix_col"]
    r4 = "_r4" if radix == 4 else ""
    w = 0 if g["layout"] == 0 else 2048 // n
    return tuple(f"batch_{p}{r4}{v}_kernel<{n}, {f}, 0, {w}>" for v in ("", "_item"))


def kernel_key(name: str) -> str:
    """A profiler kernel name without namespaces and arguments:
    'void ns::(anonymous namespace)::k<1, 2>(args)' -> 'k<1, 2>'."""
    head = name.split("(ocean")[0].split("(unsigned")[0].split("(HIP")[0].split("(float")[0]
    head = head.replace("(anonymous namespace)::", "")
    return head.split("::")[-1].strip()


def traffic_from_pmc(fetch_csv: str, write_csv: str, out_json: str):
    """profiles/traffic.json from two rocprofv3 --pmc runs (FETCH_SIZE, then
    WRITE_SIZE, KiB per dispatch) of `bench.py --profile-only`: each kernel's
    mean per dispatch, each counter scaled by what it reads for the run's
    256 MiB copy, and the single patch's frame kernels (the N = 512, F = 6
    row and column kernels of the fixed frame shape) also as row_pass /
    col_pass. Prints the table the traffic summary in profiles/ holds."""
    import csv
    from collections import defaultdict

    def load(path, counter):
        vals = defaultdict(list)
        for r in csv.DictReader(open(path)):
            if r["Counter_Name"] == counter:
                vals[kernel_key(r["Kernel_Name"])].append(float(r["Counter_Value"]))
        return vals
    fetch, write = load(fetch_csv, "FETCH_SIZE"), load(write_csv, "WRITE_SIZE")
    copy_kib = 262144.0
    cf = [v for v in fetch.get("__amd_rocclr_copyBuffer", []) if v > copy_kib / 4]
    cw = [v for v in write.get("__amd_rocclr_copyBuffer", []) if v > copy_kib / 2]
    f_fac = copy_kib / (sum(cf) / len(cf))
    w_fac = copy_kib / (sum(cw) / len(cw))
    rec = {"copy_factor_fetch": f_fac, "copy_factor_write": w_fac, "copy_kib": copy_kib,
           "copy_fetch_kib_measured": sum(cf) / len(cf), "copy_write_kib_measured": sum(cw) / len(cw),
           "kernels": {}}
    row_b, col_b = algorithmic_bytes(512, 6, True)
    for k in sorted(set(fetch) & set(write)):
        if not ("pass" in k or "split" in k or k.startswith("batch_")):
            continue
        fk = sum(fetch[k]) / len(fetch[k]) * f_fac
        wk = sum(write[k]) / len(write[k]) * w_fac
        e = {"fetch_kib": round(fk, 1), "write_kib": round(wk, 1),
             "traffic_bytes": round((fk + wk) * 1024), "dispatches": len(fetch[k])}
        rec["kernels"][k] = e
        # the frame loop's shape: the N = 512, F = 6 kernel with the most dispatches
        for alias, prefix, alg in (("row_pass", ("row_pass_kernel<512, 6,", "row_split_kernel<512, 6,"), row_b),
                                   ("col_pass", ("col_pass_kernel<512, 6,", "col_split_kernel<512, 6,"), col_b)):
            if k.startswith(prefix) and e["dispatches"] > rec["kernels"].get(alias, {}).get("dispatches", 0):
                rec["kernels"][alias] = dict(e, kernel=k, alg_bytes=alg)
    with open(out_json, "w") as f:
        json.dump(rec, f, indent=1)
    print(f"copy factors: fetch x{f_fac:.4f}, write x{w_fac:.4f}")
    print("| kernel | dispatches | fetch KiB | write KiB |")
    print("|---|---|---|---|")
    for k, e in rec["kernels"].items():
        print(f"| `{k}` | {e['dispatches']} | {e['fetch_kib']:,.0f} | {e['write_kib']:,.0f} |")


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def ref_lib():
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    try:
        from oracle_ffi import ref
        r = ref()
    except OSError:
        return None
    if r is not None:
        r.oref_hardware_concurrency.restype = C.c_uint32
        for name in ("oref_height_batch_mt", "oref_cascade_height_batch_mt"):
            getattr(r, name).argtypes = [C.c_void_p, C.POINTER(C.c_float), C.c_uint32,
                                         C.POINTER(C.c_float), C.c_uint32]
    return r


def host_threads() -> int:
    """std::thread::hardware_concurrency() of this host: the reference's
    thread count for thread_count = 0 (src/ocean.cpp:119-120)."""
    r = ref_lib()
    return int(r.oref_hardware_concurrency()) if r is not None else (os.cpu_count() or 1)


def cpu_baseline(desc, frames_budget_s: float = 12.0, threads: int = 0):
    """The reference library (oracle/_ref) on this host, same frames.
    threads = 0 is the reference's own default, one worker per hardware
    thread (std::thread::hardware_concurrency())."""
    import numpy as np
    from oracle_ffi import default_foam_params
    r = ref_lib()
    if r is None:
        return None
    cores = threads if threads > 0 else host_threads()
    kind = "reference"
    r.oref_sample_batch.argtypes = [C.c_void_p, C.POINTER(C.c_float), C.c_uint32, C.c_void_p]
    r.oref_height_batch.argtypes = [C.c_void_p, C.POINTER(C.c_float), C.c_uint32,
                                    C.POINTER(C.c_float)]
    d = type(desc).from_buffer_copy(desc)
    d.thread_count = threads
    st = C.c_int(-1)
    h = r.oref_create(C.byref(d), C.byref(st))
    fp = default_foam_params()
    foam = r.oref_foam_create(h, C.byref(fp))
    from oceanlib_amd.workloads import query_points
    xz = query_points(64, QUERIES_PER_FRAME, d.patch_length)
    heights = np.zeros(QUERIES_PER_FRAME, np.float32)
    fptr = C.POINTER(C.c_float)
    for k in range(5):  # warmup
        r.oref_update(h, k / 60.0)
        r.oref_foam_update(foam, np.float32(1 / 60))
        r.oref_height_batch(h, xz[k % 64].ctypes.data_as(fptr), QUERIES_PER_FRAME,
                            heights.ctypes.data_as(fptr))
    frames = 0
    t0 = time.perf_counter()
    while True:
        k = frames + 5
        r.oref_update(h, k / 60.0)
        r.oref_foam_update(foam, np.float32(1 / 60))
        r.oref_height_batch(h, xz[k % 64].ctypes.data_as(fptr), QUERIES_PER_FRAME,
                            heights.ctypes.data_as(fptr))
        frames += 1
        el = time.perf_counter() - t0
        if el >= frames_budget_s or frames >= 5000:
            break
    r.oref_foam_destroy(foam)
    r.oref_destroy(h)
    return {"value": frames / el, "unit": "updates/s", "cores": cores, "kind": kind,
            "sample": f"{frames} frames of the same workload ({el:.1f} s): reference "
                      f"Ocean::update with velocity, FoamField::update and "
                      f"{QUERIES_PER_FRAME} height_at calls per frame, thread_count={threads} "
                      f"({cores} workers)"}


def cpu_query_baseline(make, batch_fn, xz, budget_s: float = 2.0):
    """Queries/s of a reference query loop (oracle/_ref) over the points xz,
    at 1 thread and at one thread per hardware thread."""
    import numpy as np
    r = ref_lib()
    if r is None:
        return None
    h, destroy = make(r)
    n = len(xz)
    pts = np.ascontiguousarray(xz, np.float32).reshape(-1)
    out = np.empty(n, np.float32)
    fp = C.POINTER(C.c_float)
    res = {}
    for label, threads in (("1_thread", 1), ("all_threads", 0)):
        fn = getattr(r, batch_fn)
        fn(h, pts.ctypes.data_as(fp), min(n, 4096), out.ctypes.data_as(fp), threads)  # warm
        calls = 0
        t0 = time.perf_counter()
        while True:
            fn(h, pts.ctypes.data_as(fp), n, out.ctypes.data_as(fp), threads)
            calls += 1
            el = time.perf_counter() - t0
            if el >= budget_s:
                break
        res[label] = round(calls * n / el, 1)
    destroy(r, h)
    res["threads_all"] = host_threads()
    res["points_per_call"] = n
    return res


def batched_leg(lib, frames: int, warmup: int, tune: bool = True, precision: int = 0):
    """The batched uneven seas (256 seas, 1-3 levels of N 64..512 each) as one
    OceanBatch: autotuned, then `frames` frames timed between two
    synchronizes. Reports 512^2-equivalent patch updates per second (cells
    per second / 512^2) and the batch's HBM roofline fraction."""
    import numpy as np
    from oceanlib_amd.workloads import config4_levels, desc_from
    flat, _ = config4_levels()
    descs = [desc_from(kw) for kw in flat]
    arr = (capi_mod().OceanDesc * len(descs))(*descs)
    st = C.c_int(-1)
    t0 = time.perf_counter()
    b = lib.ocean_amd_batch_create(arr, len(descs), C.byref(st))
    if not b:
        raise SystemExit(f"ocean_amd_batch_create failed: {st.value}")
    build_s = time.perf_counter() - t0
    log(f"bench: batch of {len(descs)} levels built in {build_s:.1f} s")
    if precision:
        assert lib.ocean_amd_batch_set_precision(b, precision) == 0
    if tune:
        lib.ocean_amd_batch_autotune(b, 5)
    groups = []
    for g in range(lib.ocean_amd_batch_group_count(b)):
        info = capi_mod().BatchGroup()
        lib.ocean_amd_batch_group_info(b, g, C.byref(info))
        groups.append({k: getattr(info, k) for k, _ in info._fields_})
    for k in range(warmup):
        lib.ocean_amd_batch_update(b, k / 60.0)
    if lib.ocean_amd_batch_synchronize(b) != 0:
        raise SystemExit("batch warmup failed")
    log(f"bench: timing {frames} batched frames")
    t0 = time.perf_counter()
    for k in range(warmup, warmup + frames):
        lib.ocean_amd_batch_update(b, k / 60.0)
    if lib.ocean_amd_batch_synchronize(b) != 0:
        raise SystemExit(f"batch error {lib.ocean_amd_batch_last_error(b)}")
    el = time.perf_counter() - t0
    lib.ocean_amd_batch_destroy(b)
    cells = sum(kw["size"] ** 2 for kw in flat)
    alg = sum(frame_bytes(kw["size"], 4, False) for kw in flat)
    fps = frames / el
    gbs = alg * fps / 1e9
    return {
        "workload": "256 batched uneven seas, 1-3 levels each, N in {64,128,256,512}, "
                    "300-frame fixture workload (PCG32 seed 0xC4)",
        "levels": len(flat), "cells_per_frame": cells,
        "frames_per_s": round(fps, 2), "ms_per_frame": round(1000.0 / fps, 4),
        "patch_updates_512_per_s": round(cells * fps / (512 * 512), 1),
        "alg_bytes_per_frame": alg,
        "roofline": {"bound": "hbm", "achieved": round(gbs, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 5),
                     "traffic": traffic_from_profiles(
                         [batch_kernel_names(g, p) for g in groups for p in ("row", "col")])},
        "groups": groups, "build_s": round(build_s, 2), "frames": frames,
        "precision": ["exact", "fast", "fast_half"][precision],
    }


def batched_cpu_baseline(budget_s: float = 8.0):
    """The reference library on a bounded sample of the batched seas' levels
    (the reference steps levels one after another, src/cascade.cpp:42-55):
    cells per second over this host's cores, as 512^2-equivalent updates/s."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    try:
        from oracle_ffi import ref
        r = ref()
    except OSError:
        r = None
    if r is None:
        return None
    from oceanlib_amd.workloads import config4_levels, desc_from
    cores = host_threads()
    flat, _ = config4_levels(24)
    handles = []
    for kw in flat:
        d = desc_from(dict(kw, thread_count=0))
        st = C.c_int(-1)
        handles.append((r.oref_create(C.byref(d), C.byref(st)), kw["size"] ** 2))
    cells = 0
    frames = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < budget_s:
        for h, c in handles:
            r.oref_update(h, frames / 60.0)
            cells += c
        frames += 1
    el = time.perf_counter() - t0
    for h, _ in handles:
        r.oref_destroy(h)
    return {"value": round(cells / el / (512 * 512), 1), "unit": "512^2 patch-updates/s",
            "cores": cores, "kind": "reference",
            "sample": f"{frames} frames of the first 24 seas ({len(flat)} levels, "
                      f"{el:.1f} s): reference Ocean::update level by level, "
                      f"thread_count=0 ({cores} workers)"}


def learned_leg(lib, reps: int = 10):
    """Converged height queries, plain vs learned start point, on the learned
    proposal's sea: three 512^2 levels (800 / 150 / 25 m, 22 m/s, choppiness
    1.3, seeds 1337-1339) queried as a cascade at t = 5 s, 1M points uniform
    over a 2 km square, each level inverted until the step is <= 2.5e-6 m.
    Reports queries/s of both forms (HIP events around `reps` back-to-back
    calls on the levels' stream), taps per level (mean and per-wavefront
    maximum) and the largest height error against the same points solved
    plain to 1e-7 m."""
    import numpy as np
    import torch
    from oceanlib_amd import capi
    from oceanlib_amd.workloads import desc_from
    params_path = os.path.join(ROOT, "oceanlib_amd", "data", "height_proposal.npy")
    if not os.path.exists(params_path):
        return None
    params = np.load(params_path, allow_pickle=False)
    levels = [dict(size=512, patch_length=L, wind_speed=22.0, choppiness=1.3, seed=sd)
              for L, sd in ((800.0, 1337), (150.0, 1338), (25.0, 1339))]
    descs = (capi.OceanDesc * 3)(*[desc_from(kw) for kw in levels])
    st = C.c_int(-1)
    b = lib.ocean_amd_batch_create(descs, 3, C.byref(st))
    if not b:
        raise SystemExit(f"batch create failed: {st.value}")
    lib.ocean_amd_batch_update(b, 5.0)
    db = capi.DeviceBuffers()
    assert lib.ocean_amd_get_device_buffers(lib.ocean_amd_batch_level(b, 0), C.byref(db)) == 0
    stream = torch.cuda.ExternalStream(db.stream)
    nq = 1 << 20
    tol = 2.5e-6
    rng = np.random.default_rng(0xC5)
    xz = torch.from_numpy(rng.uniform(0.0, 2000.0, (nq, 2)).astype(np.float32)).cuda()
    model = lib.ocean_amd_proposal_create(params.ctypes.data_as(capi._F), nq, C.byref(st))
    if not model:
        raise SystemExit(f"proposal create failed: {st.value}")
    ref = torch.empty(nq, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    assert lib.ocean_amd_batch_height_converged(b, 0, 3, C.c_void_p(xz.data_ptr()), nq, 1e-7, None,
                                                C.c_void_p(ref.data_ptr()), None, None) == 0
    out = {}
    for name, m in (("plain", None), ("learned", model)):
        h = torch.empty(nq, dtype=torch.float32, device="cuda")
        it = torch.empty((nq, 3), dtype=torch.uint8, device="cuda")
        fl = torch.empty(nq, dtype=torch.uint8, device="cuda")

        def call(stats: bool):
            r = lib.ocean_amd_batch_height_converged(
                b, 0, 3, C.c_void_p(xz.data_ptr()), nq, tol, m, C.c_void_p(h.data_ptr()),
                C.c_void_p(it.data_ptr()) if stats else None,
                C.c_void_p(fl.data_ptr()) if stats else None)
            assert r == 0
        call(True)
        assert lib.ocean_amd_batch_synchronize(b) == 0
        itn = it.cpu().numpy().astype(np.int64)
        fln = fl.cpu().numpy()
        err = (h - ref).abs().max().item()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            call(False)
        e1.record(stream)
        e1.synchronize()
        ms = e0.elapsed_time(e1) / reps
        out[name] = {"queries_per_s": round(nq / (ms * 1e-3), 1), "ms_per_1M": round(ms, 4),
                     "taps_mean_per_level": [round(float(v), 3) for v in itn.mean(0)],
                     "taps_wave_max_per_level": [round(float(v), 3)
                                                 for v in itn.reshape(-1, 64, 3).max(1).mean(0)],
                     "max_err_mm": round(err * 1e3, 5),
                     "not_converged": int(np.count_nonzero(fln & 1)),
                     "fallbacks": int(np.count_nonzero(fln & 2))}
    # The reference's own cascade query (CascadeStack::height_at: four
    # fixed-point steps per level, src/cascade.cpp:118-130) on the same sea
    # and points: the device form over the 3-level range, and the reference
    # library on the host.
    h4 = torch.empty(nq, dtype=torch.float32, device="cuda")

    def call4():
        assert lib.ocean_amd_batch_sample_device(b, 0, 3, C.c_void_p(xz.data_ptr()), nq, None,
                                                 C.c_void_p(h4.data_ptr())) == 0
    call4()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        call4()
    e1.record(stream)
    e1.synchronize()
    ms4 = e0.elapsed_time(e1) / reps
    out["cascade_height_at_4_steps"] = {"queries_per_s": round(nq / (ms4 * 1e-3), 1),
                                        "ms_per_1M": round(ms4, 4)}
    descs_cpu = (capi.OceanDesc * 3)(*[desc_from(kw) for kw in levels])

    def make_cascade(r):
        c = r.oref_cascade_create(descs_cpu, 3)
        r.oref_cascade_update(c, 5.0)
        return c, lambda rr, hh: rr.oref_cascade_destroy(hh)
    out["cascade_height_at_4_steps"]["cpu_baseline"] = cpu_query_baseline(
        make_cascade, "oref_cascade_height_batch_mt", xz[: 1 << 18].cpu().numpy())
    lib.ocean_amd_proposal_destroy(model)
    lib.ocean_amd_batch_destroy(b)
    out["speedup"] = round(out["learned"]["queries_per_s"] / out["plain"]["queries_per_s"], 4)
    out["workload"] = ("3 x 512^2 levels as a cascade (800/150/25 m, 22 m/s, chop 1.3, seeds "
                       "1337-1339), 1M points over a 2 km square, step tolerance 2.5e-6 m")
    return out


def scene_inputs(lib, frames: int):
    """The three-cascade + interaction scene: cascade levels (256^2 of
    800 / 150 / 25 m, 15 m/s, seed 42), the interaction descriptor (256^2 at
    0.1 m, kernel radius 6), the static hull mask and, per frame, the eight
    moving sources' disturbance records as C structs."""
    import numpy as np
    from oceanlib_amd.workloads import (CONFIG3_INTERACTION, CONFIG3_LEVELS, config3_disturbances,
                                        config3_hull_mask, config3_sources, desc_from)
    cm = capi_mod()
    levels = (cm.OceanDesc * len(CONFIG3_LEVELS))(*[desc_from(kw) for kw in CONFIG3_LEVELS])
    idesc = cm.InteractionDesc()
    lib.ocean_interaction_desc_init(C.byref(idesc))
    for k, v in CONFIG3_INTERACTION.items():
        setattr(idesc, k, v)
    srcs = config3_sources()
    recs = []
    for k in range(frames):
        row = []
        for r in config3_disturbances(srcs, k / 60.0):
            d = cm.Disturbance()
            for f, v in r.items():
                setattr(d, f, v)
            row.append(d)
        recs.append(row)
    return levels, idesc, np.ascontiguousarray(config3_hull_mask(), np.uint8), recs


def scene_leg(lib, frames: int, warmup: int):
    """Three cascades plus interaction on the device: a step is the cascade
    update (three 256^2 levels in one batched launch per pass) and the
    interaction update with that frame's eight sources (stamps, hull, one
    leapfrog substep, RGBA finalize). Wall time of `frames` steps between two
    device synchronizes, and of each half alone."""
    import numpy as np
    from oceanlib_amd import hipmem
    levels, idesc, mask, recs = scene_inputs(lib, warmup + frames)
    st = C.c_int(-1)
    cas = lib.ocean_amd_cascade_create(levels, len(levels), C.byref(st))
    field = lib.ocean_interaction_create(C.byref(idesc), C.byref(st))
    if not cas or not field:
        raise SystemExit(f"scene create failed: {st.value}")
    lib.ocean_interaction_set_obstruction(field, mask.ctypes.data_as(C.POINTER(C.c_uint8)))
    dt = np.float32(1.0 / 60.0)

    def run(first, count, do_cascade, do_field):
        for k in range(first, first + count):
            if do_cascade:
                lib.ocean_amd_cascade_update(cas, k / 60.0)
            if do_field:
                for d in recs[k]:
                    lib.ocean_interaction_add(field, C.byref(d))
                lib.ocean_interaction_update(field, dt)

    out = {}
    for name, c, f in (("scene", True, True), ("cascade_only", True, False),
                       ("interaction_only", False, True)):
        run(0, warmup, c, f)
        hipmem.synchronize()
        t0 = time.perf_counter()
        run(warmup, frames, c, f)
        hipmem.synchronize()
        out[name] = frames / (time.perf_counter() - t0)
    lib.ocean_interaction_destroy(field)
    lib.ocean_amd_cascade_destroy(cas)
    return {
        "workload": "three cascades plus interaction: 3 x 256^2 levels (800/150/25 m, 15 m/s, "
                    "seed 42) and a 256^2 iWave grid at 0.1 m, kernel radius 6, 8 continuous "
                    "sources on circles, one static hull, 1/60 s steps",
        "frames": frames,
        "steps_per_s": round(out["scene"], 1),
        "ms_per_step": round(1000.0 / out["scene"], 4),
        "cascade_updates_per_s": round(out["cascade_only"], 1),
        "interaction_updates_per_s": round(out["interaction_only"], 1),
    }


def scene_cpu_baseline(budget_s: float = 8.0):
    """The reference library's CascadeStack and InteractionField on the same
    scene, on this host's cores, for a bounded sample of steps."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    try:
        from oracle_ffi import ref
        r = ref()
    except OSError:
        r = None
    if r is None:
        return None
    lib = capi_mod().load()
    cores = host_threads()
    levels, idesc, mask, recs = scene_inputs(lib, 1200)
    for d in levels:
        d.thread_count = 0
    idesc.thread_count = 0
    cas = r.oref_cascade_create(levels, len(levels))
    st = C.c_int(-1)
    field = r.oref_interaction_create(C.byref(idesc), C.byref(st))
    if not cas or not field:
        return None
    r.oref_interaction_set_obstruction(field, mask.ctypes.data_as(C.POINTER(C.c_uint8)))
    dt = np.float32(1.0 / 60.0)
    steps = 0
    t0 = time.perf_counter()
    while True:
        k = steps % len(recs)
        r.oref_cascade_update(cas, k / 60.0)
        for d in recs[k]:
            r.oref_interaction_add(field, C.byref(d))
        r.oref_interaction_update(field, dt)
        steps += 1
        el = time.perf_counter() - t0
        if el >= budget_s or steps >= 1200:
            break
    r.oref_interaction_destroy(field)
    r.oref_cascade_destroy(cas)
    return {"value": round(steps / el, 1), "unit": "steps/s", "cores": cores, "kind": "reference",
            "sample": f"{steps} steps of the same scene ({el:.1f} s): reference "
                      f"CascadeStack::update and InteractionField add + update, "
                      f"thread_count=0 ({cores} workers)"}


def interaction_setup_leg(lib, with_cpu: bool):
    """The interaction field's set-up work: creation (the least-squares
    stencil fit and its symbol scan, on the device) at the scene's 256^2 /
    radius 6 and at 1024^2 / radius 12, and set_obstruction of a moving hull
    (the scene's hull shifted one cell per call, so the mask tables are
    rebuilt on every call, as for a hull that moves each frame), per call
    including the device work. The reference library's times for the same
    calls on this host beside them."""
    import numpy as np
    from oceanlib_amd import hipmem
    from oceanlib_amd.workloads import config3_hull_mask
    cm = capi_mod()
    r = ref_lib() if with_cpu else None
    out = {}
    for n, radius in ((256, 6), (1024, 12)):
        d = cm.InteractionDesc()
        lib.ocean_interaction_desc_init(C.byref(d))
        d.size, d.extent, d.kernel_radius = n, 0.1 * n, radius
        st = C.c_int(-1)
        times = []
        for _ in range(3):
            t0 = time.perf_counter()
            f = lib.ocean_interaction_create(C.byref(d), C.byref(st))
            times.append(time.perf_counter() - t0)
            if not f:
                raise SystemExit(f"interaction create failed: {st.value}")
            if _ < 2:
                lib.ocean_interaction_destroy(f)
        hull = config3_hull_mask(n)
        masks = [np.ascontiguousarray(np.roll(hull, k, axis=1)) for k in range(16)]
        ptrs = [m.ctypes.data_as(C.POINTER(C.c_uint8)) for m in masks]
        calls = 200
        for k in range(20):
            lib.ocean_interaction_set_obstruction(f, ptrs[k % 16])
        hipmem.synchronize()
        t0 = time.perf_counter()
        for k in range(calls):
            lib.ocean_interaction_set_obstruction(f, ptrs[k % 16])
        hipmem.synchronize()
        obst_ms = (time.perf_counter() - t0) * 1000.0 / calls
        lib.ocean_interaction_destroy(f)
        row = {"create_ms": round(min(times) * 1000.0, 2),
               "set_obstruction_ms": round(obst_ms, 4),
               "solid_cells": int(hull.sum())}
        if r is not None:
            d.thread_count = 0
            t0 = time.perf_counter()
            rf = r.oref_interaction_create(C.byref(d), C.byref(st))
            row["reference_create_ms"] = round((time.perf_counter() - t0) * 1000.0, 2)
            if rf:
                rcalls = 50
                t0 = time.perf_counter()
                for k in range(rcalls):
                    r.oref_interaction_set_obstruction(rf, ptrs[k % 16])
                row["reference_set_obstruction_ms"] = round(
                    (time.perf_counter() - t0) * 1000.0 / rcalls, 4)
                r.oref_interaction_destroy(rf)
        out[f"{n}^2 radius {radius}"] = row
    return out


def handoff_leg(lib, sim, foam, n: int, reps: int = 200):
    """The renderer handoff of the same patch: the displacement (alpha = the
    persistent foam) and the normal as RGBA16F, into device arrays and into
    host memory (the D2H copy included). Wall time per call over back-to-back
    calls; the kernels' own durations are in the rocprofv3 trace."""
    import numpy as np
    import torch

    cells = n * n
    dd = torch.empty(4 * cells, dtype=torch.int16, device="cuda")
    dn = torch.empty(4 * cells, dtype=torch.int16, device="cuda")
    pd, pn = C.c_void_p(dd.data_ptr()), C.c_void_p(dn.data_ptr())
    for _ in range(10):
        assert lib.ocean_amd_pack_rgba16_device(sim, foam, pd, pn) == 0
    lib.ocean_amd_synchronize(sim)
    t0 = time.perf_counter()
    for _ in range(reps):
        lib.ocean_amd_pack_rgba16_device(sim, foam, pd, pn)
    lib.ocean_amd_synchronize(sim)
    dev_ms = (time.perf_counter() - t0) * 1e3 / reps
    hd = np.empty(4 * cells, np.uint16)
    hn = np.empty(4 * cells, np.uint16)
    lib.ocean_amd_pack_rgba16(sim, foam, hd.ctypes.data, hn.ctypes.data)
    t0 = time.perf_counter()
    for _ in range(20):
        lib.ocean_amd_pack_rgba16(sim, foam, hd.ctypes.data, hn.ctypes.data)
    host_ms = (time.perf_counter() - t0) * 1e3 / 20
    alg = cells * (16 + 4 + 16 + 16)  # disp + foam + normal read, 2 x 8 B written
    return {"what": "displacement (alpha = persistent foam) + normal of the 512^2 patch to "
                    "RGBA16F, the viewer's upload format",
            "device_wall_ms": round(dev_ms, 5), "alg_bytes": alg,
            "host_wall_ms": round(host_ms, 4), "host_bytes": cells * 16}


def capi_mod():
    from oceanlib_amd import capi
    return capi


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=600)
    ap.add_argument("--warmup", type=int, default=30)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--batch-frames", type=int, default=120,
                    help="timed frames of the batched-seas leg (0 skips it)")
    ap.add_argument("--no-learned", action="store_true",
                    help="skip the converged / learned height query leg")
    ap.add_argument("--no-autotune", action="store_true",
                    help="batched leg: keep the default grid sizes")
    ap.add_argument("--frame-shape", default=None, metavar="[LAYOUT:]ROW,COL",
                    help="fix the frame kernels' field layout and launch shapes instead of "
                         "autotuning (counter runs, where timing is perturbed)")
    ap.add_argument("--queries-only", action="store_true",
                    help="run only the converged / learned height query leg (for rocprofv3)")
    ap.add_argument("--scene-frames", type=int, default=600,
                    help="timed steps of the three-cascade + interaction scene (0 skips it)")
    ap.add_argument("--scene-only", action="store_true",
                    help="run only the three-cascade + interaction scene leg (for rocprofv3)")
    ap.add_argument("--profile-only", action="store_true",
                    help="run only warmup + timed frames (for rocprofv3 runs)")
    ap.add_argument("--traffic-from", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="write profiles/traffic.json from two rocprofv3 --pmc counter_collection "
                         "CSVs (FETCH_SIZE, WRITE_SIZE) of a --profile-only run, and exit")
    args = ap.parse_args()
    if args.traffic_from:
        traffic_from_pmc(args.traffic_from[0], args.traffic_from[1],
                         os.path.join(ROOT, "profiles", "traffic.json"))
        return
    if args.gpus != 1:
        print(f"bench.py: measures one GPU only; --gpus {args.gpus} is not supported "
              "(independent oceans would shard by ocean across GPUs with no communication)",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from oceanlib_amd import capi
    from oceanlib_amd.workloads import CONFIG2, desc_from, query_points

    if not torch.cuda.is_available():
        raise SystemExit("bench.py needs a HIP device")
    lib = capi.load()
    if args.queries_only:
        print(json.dumps({"converged_height_at": learned_leg(lib)}), flush=True)
        return
    if args.scene_only:
        print(json.dumps({"scene": scene_leg(lib, max(args.scene_frames, 1), 30),
                          "setup": interaction_setup_leg(lib, not args.no_cpu)}), flush=True)
        return
    desc = desc_from(CONFIG2)
    n = desc.size
    st = C.c_int(-1)
    sim = lib.ocean_create(C.byref(desc), C.byref(st))
    if not sim:
        raise SystemExit(f"ocean_create failed: {st.value}")
    fd = capi.FoamDesc()
    lib.ocean_amd_foam_desc_init(C.byref(fd))
    foam = lib.ocean_amd_foam_create(sim, C.byref(fd), C.byref(st))
    if not foam:
        raise SystemExit(f"foam create failed: {st.value}")

    # Launch shapes of the two frame kernels, timed on the device (the
    # autotuner; every shape gives the same bits).
    shape_r, shape_c = C.c_uint32(), C.c_uint32()
    tune_r, tune_c = C.c_float(), C.c_float()
    if args.frame_shape:
        spec = args.frame_shape
        layout = 0
        if ":" in spec:
            layout, spec = int(spec.split(":")[0]), spec.split(":")[1]
        shape_r.value, shape_c.value = (int(v) for v in spec.split(","))
        if (lib.ocean_amd_set_frame_layout(sim, layout) != 0 or
                lib.ocean_amd_set_frame_shape(sim, shape_r.value, shape_c.value) != 0):
            raise SystemExit(f"bad --frame-shape {args.frame_shape}")
        log(f"bench: frame layout {layout}, shapes row {shape_r.value} col {shape_c.value} (fixed)")
    else:
        if lib.ocean_amd_autotune_frame(sim, 20, C.byref(shape_r), C.byref(shape_c),
                                        C.byref(tune_r), C.byref(tune_c)) != 0:
            raise SystemExit("frame autotune failed")
        log(f"bench: frame layout {lib.ocean_amd_get_frame_layout(sim)}, "
            f"shapes row {shape_r.value} col {shape_c.value} "
            f"({tune_r.value * 1e3:.1f} / {tune_c.value * 1e3:.1f} us)")

    frame_layout = int(lib.ocean_amd_get_frame_layout(sim))
    total = args.warmup + args.steps
    frames_xz = 64
    xz_host = query_points(frames_xz, QUERIES_PER_FRAME, desc.patch_length)
    xz_dev = torch.from_numpy(xz_host.reshape(frames_xz, -1)).cuda()
    heights = torch.empty((frames_xz, QUERIES_PER_FRAME), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()

    # Device pointers of each frame's query points and results, made once so
    # the loop below is only the three library calls.
    xz_ptr = [C.c_void_p(xz_dev[j].data_ptr()) for j in range(frames_xz)]
    h_ptr = [C.c_void_p(heights[j].data_ptr()) for j in range(frames_xz)]
    dt = np.float32(1.0 / 60.0)

    def frame(k):
        # the foam step and the frame's queries share one launch
        lib.ocean_update(sim, k / 60.0)
        j = k % frames_xz
        if lib.ocean_amd_foam_update_sample_device(foam, dt, xz_ptr[j], QUERIES_PER_FRAME, None,
                                                   h_ptr[j]) != 0:
            raise SystemExit("foam step / queries failed")

    def frame_separate(k):
        # the reference's call sequence: update, foam update, queries
        lib.ocean_update(sim, k / 60.0)
        lib.ocean_amd_foam_update(foam, dt)
        j = k % frames_xz
        if lib.ocean_amd_sample_batch_device(sim, xz_ptr[j], QUERIES_PER_FRAME, None,
                                             h_ptr[j]) != 0:
            raise SystemExit("query failed")

    log(f"bench: warmup {args.warmup} frames")
    for k in range(args.warmup):
        frame(k)
    lib.ocean_amd_synchronize(sim)
    torch.cuda.synchronize()
    log(f"bench: timing {args.steps} frames")
    # The two frame kernels of every 16th frame of the timed loop record HIP
    # start/end events in their own dispatches, on the ocean's stream: the
    # kernels' durations inside this loop.
    if lib.ocean_amd_frame_timing(sim, TIMING_EVERY) != 0:
        raise SystemExit("frame timing failed")
    t0 = time.perf_counter()
    for k in range(args.warmup, total):
        frame(k)
    lib.ocean_amd_synchronize(sim)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    row_ms, col_ms, samples = C.c_float(), C.c_float(), C.c_uint32()
    if lib.ocean_amd_frame_timing_read(sim, C.byref(row_ms), C.byref(col_ms),
                                       C.byref(samples)) != 0:
        raise SystemExit("frame timing read failed")
    lib.ocean_amd_frame_timing(sim, 0)
    if lib.ocean_amd_last_error(sim) != 0:
        raise SystemExit(f"device error {lib.ocean_amd_last_error(sim)}")
    ms_per_step = el * 1000.0 / args.steps
    value = args.steps / el
    if args.profile_only:
        # A plain copy of known size, for scaling the FETCH_SIZE / WRITE_SIZE
        # counters (256 MiB read + 256 MiB written per launch).
        src = torch.empty(1 << 26, dtype=torch.float32, device="cuda").fill_(1.0)
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        rec = {"profile_only": True, "ms_per_step": ms_per_step, "copy_bytes": 4 << 26}
        if args.batch_frames > 0:
            rec["batched"] = batched_leg(lib, args.batch_frames, 2, tune=not args.no_autotune)
        print(json.dumps(rec), flush=True)
        return

    # The same frames with the three separate calls, for comparison.
    for k in range(total, total + args.warmup):
        frame_separate(k)
    lib.ocean_amd_synchronize(sim)
    t0 = time.perf_counter()
    for k in range(total + args.warmup, total + args.warmup + args.steps):
        frame_separate(k)
    lib.ocean_amd_synchronize(sim)
    separate_per_s = args.steps / (time.perf_counter() - t0)

    # The opt-in fast modes on the same frame loop (not the headline: the
    # headline is the bit-exact frame).
    fast = {}
    for mode, name in ((1, "fast"), (2, "fast_half")):
        assert lib.ocean_amd_set_precision(sim, mode) == 0
        fr_, fc_ = C.c_uint32(), C.c_uint32()
        lib.ocean_amd_autotune_frame(sim, 20, C.byref(fr_), C.byref(fc_), None, None)
        base = total + args.warmup + args.steps
        for k in range(base, base + args.warmup):
            frame(k)
        lib.ocean_amd_synchronize(sim)
        lib.ocean_amd_frame_timing(sim, TIMING_EVERY)
        t0 = time.perf_counter()
        for k in range(base + args.warmup, base + args.warmup + args.steps):
            frame(k)
        lib.ocean_amd_synchronize(sim)
        fps_fast = args.steps / (time.perf_counter() - t0)
        fr_ms, fc_ms, fs_ = C.c_float(), C.c_float(), C.c_uint32()
        lib.ocean_amd_frame_timing_read(sim, C.byref(fr_ms), C.byref(fc_ms), C.byref(fs_))
        lib.ocean_amd_frame_timing(sim, 0)
        fast[name] = {"frames_per_s": round(fps_fast, 1),
                      "update_device_ms": round(fr_ms.value + fc_ms.value, 5),
                      "row_ms": round(fr_ms.value, 5), "col_ms": round(fc_ms.value, 5),
                      "frame_shape": [fr_.value, fc_.value]}
    assert lib.ocean_amd_set_precision(sim, 0) == 0
    assert lib.ocean_amd_set_frame_layout(sim, frame_layout) == 0
    assert lib.ocean_amd_set_frame_shape(sim, shape_r.value, shape_c.value) == 0
    if lib.ocean_amd_last_error(sim) != 0:
        raise SystemExit(f"device error {lib.ocean_amd_last_error(sim)}")

    # The same kernels launched back to back with nothing between them.
    iso_r, iso_c = C.c_float(), C.c_float()
    assert lib.ocean_amd_profile_frame(sim, 100.0, 1.0 / 60.0, 200, C.byref(iso_r),
                                       C.byref(iso_c)) == 0
    fields = 6
    row_b, col_b = algorithmic_bytes(n, fields, True)
    kernels = {
        "row_pass": {"ms": row_ms.value, "alg_bytes": row_b,
                     "gbs": row_b / (row_ms.value * 1e-3) / 1e9,
                     "ms_back_to_back": iso_r.value},
        "col_pass": {"ms": col_ms.value, "alg_bytes": col_b,
                     "gbs": col_b / (col_ms.value * 1e-3) / 1e9,
                     "ms_back_to_back": iso_c.value},
        "timed_frames_sampled": samples.value,
    }
    dom = max(("row_pass", "col_pass"), key=lambda k: kernels[k]["ms"])
    achieved = kernels[dom]["gbs"]
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                "traffic": traffic_from_profiles([dom]), "kernel": dom,
                "alg_bytes_per_launch": kernels[dom]["alg_bytes"]}

    # height_at throughput: 1M queries per launch on the same sea.
    nq = 1 << 20
    rng = np.random.default_rng(11)
    big = torch.from_numpy(rng.uniform(0, desc.patch_length, (nq, 2)).astype(np.float32)).cuda()
    hq = torch.empty(nq, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    for _ in range(3):
        lib.ocean_amd_sample_batch_device(sim, C.c_void_p(big.data_ptr()), nq, None,
                                          C.c_void_p(hq.data_ptr()))
    lib.ocean_amd_synchronize(sim)
    reps = 20
    tq = time.perf_counter()
    for _ in range(reps):
        lib.ocean_amd_sample_batch_device(sim, C.c_void_p(big.data_ptr()), nq, None,
                                          C.c_void_p(hq.data_ptr()))
    lib.ocean_amd_synchronize(sim)
    qps = reps * nq / (time.perf_counter() - tq)

    # The reference's Ocean::height_at on the host, same sea, 256K of the points.
    lib.ocean_update(sim, 1000.0 / 60.0)
    cpu_queries = None
    if not args.no_cpu:
        def make_ocean(r):
            d = desc_from(CONFIG2)
            st2 = C.c_int(-1)
            h = r.oref_create(C.byref(d), C.byref(st2))
            r.oref_update(h, 1000.0 / 60.0)
            return h, lambda rr, hh: rr.oref_destroy(hh)
        cpu_queries = cpu_query_baseline(make_ocean, "oref_height_batch_mt",
                                         big[: 1 << 18].cpu().numpy())

    handoff = handoff_leg(lib, sim, foam, n)
    batched = batched_leg(lib, args.batch_frames, 5) if args.batch_frames > 0 else None
    if batched is not None:
        for mode, name in ((1, "fast"), (2, "fast_half")):
            bf = batched_leg(lib, args.batch_frames, 5, precision=mode)
            batched[name] = {"patch_updates_512_per_s": bf["patch_updates_512_per_s"],
                             "ms_per_frame": bf["ms_per_frame"]}
    learned = learned_leg(lib) if not args.no_learned else None
    scene = scene_leg(lib, args.scene_frames, 30) if args.scene_frames > 0 else None
    if scene is not None and not args.no_cpu:
        scene["cpu_baseline"] = scene_cpu_baseline()
    if scene is not None:
        scene["setup"] = interaction_setup_leg(lib, not args.no_cpu)
    cpu = None if args.no_cpu else cpu_baseline(desc)
    if cpu is not None:
        one = cpu_baseline(desc, 4.0, threads=1)
        cpu["value_1_thread"] = one["value"] if one else None
        sixteen = cpu_baseline(desc, 4.0, threads=16)
        cpu["value_16_threads"] = sixteen["value"] if sixteen else None
        cpu["cpu_model"] = cpu_model()
        cpu["hardware_threads"] = host_threads()
    if batched is not None and not args.no_cpu:
        batched["cpu_baseline"] = batched_cpu_baseline()

    lib.ocean_amd_foam_destroy(foam)
    lib.ocean_destroy(sim)

    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "single heavy-chop patch: N=512, 250 m, 22 m/s, chop 1.3, "
                               "velocity + persistent advected foam, depth 50 m, seed 7, "
                               "dt 1/60 s, 4096 height_at queries per frame (foam step and "
                               "queries in one launch)",
                   "size": n, "patch_length": 250.0, "queries_per_frame": QUERIES_PER_FRAME,
                   "parallelism": "single GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "kernels": kernels,
        "update_device_ms": round(row_ms.value + col_ms.value, 5),
        "frames_per_s_separate_calls": round(separate_per_s, 1),
        "frame_shape": {"layout": frame_layout, "row_variant": shape_r.value,
                        "col_variant": shape_c.value},
        "height_at_queries_per_s": round(qps, 1),
        "height_at_cpu_baseline": cpu_queries,
        "fast_mode": fast,
        "rgba16_handoff": handoff,
        "batched": batched,
        "converged_height_at": learned,
        "cascades_and_interaction": scene,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
