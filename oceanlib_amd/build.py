"""Builds liboceanlib_amd.so in-tree: HIP kernels for gfx950 + C++20 host code.

    python -m oceanlib_amd.build
    python -m oceanlib_amd.build --host-sanitize OUT_DIR   (ASan + UBSan on host code)

Every .hip file is compiled by hipcc for gfx950 with the bit-exactness flags
(no FP contraction, correctly rounded fp32 divide/sqrt, no packed-FP32 ops);
every .cpp file is host C++20. Objects go to oceanlib_amd/_build/, the shared
library next to this file. No JIT, no cache outside the tree.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "liboceanlib_amd.so")
ARCH = os.environ.get("OCEAN_AMD_ARCH", "gfx950")

COMMON = ["-std=c++20", "-O3", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
          "-I" + os.path.join(ROOT, "include"), "-I" + CSRC, "-Wall"]
HIP_FLAGS = ["-x", "hip", f"--offload-arch={ARCH}", "-fhip-fp32-correctly-rounded-divide-sqrt",
             "-fno-gpu-flush-denormals-to-zero",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (ROCm is required to build)")


def sources():
    out = []
    for name in sorted(os.listdir(CSRC)):
        if name.endswith(".hip") or name.endswith(".cpp"):
            out.append(os.path.join(CSRC, name))
    return out


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    flags = COMMON + (HIP_FLAGS if src.endswith(".hip") else [])
    cmd = [hipcc()] + flags + ["-c", src, "-o", obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    return obj


def build(verbose: bool = True) -> str:
    os.makedirs(BUILD, exist_ok=True)
    srcs = sources()
    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 4)) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs + \
        ["-lpthread"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} from {len(srcs)} sources", file=sys.stderr)
    return LIB


# Host-code sanitizers for a checking build (device code is built as usual;
# each flag reaches only the host compiler).
HOST_SANITIZE = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
                 "-Xarch_host", "-fno-omit-frame-pointer",
                 "-Xarch_host", "-fno-sanitize-recover=undefined"]


def build_host_sanitized(out_dir: str) -> str:
    """The library with AddressSanitizer and UBSan on its host code, into
    out_dir (objects and liboceanlib_amd.so), for a program that links the
    sanitizer runtime itself (e.g. tests/cpp_api built with
    -fsanitize=address,undefined -shared-libsan); run it with
    ASAN_OPTIONS=protect_shadow_gap=0:alloc_dealloc_mismatch=0 (the ROCm
    runtime's code object loader frees some operator-new memory with free)."""
    os.makedirs(out_dir, exist_ok=True)

    def comp(src):
        obj = os.path.join(out_dir, os.path.basename(src) + ".o")
        flags = COMMON + (HIP_FLAGS if src.endswith(".hip") else []) + HOST_SANITIZE
        res = subprocess.run([hipcc()] + flags + ["-c", src, "-o", obj], capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError(f"compile failed: {src}\n{res.stderr}")
        return obj
    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 4)) as ex:
        objs = list(ex.map(comp, sources()))
    lib = os.path.join(out_dir, "liboceanlib_amd.so")
    subprocess.run([hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", lib] + objs +
                   ["-lpthread", "-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
                    "-shared-libsan"], check=True)
    return lib


if __name__ == "__main__":
    if len(sys.argv) == 3 and sys.argv[1] == "--host-sanitize":
        print(build_host_sanitized(sys.argv[2]))
    else:
        build()
