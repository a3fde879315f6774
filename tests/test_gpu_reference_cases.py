"""The reference's own behavioural tests of the per-frame path, through this
library's C++ API and C ABI on the GPU.

tests/cpp_api/reference_cases.cpp and interaction_cases.cpp restate, case by
case, the reference's tests/test_api.cpp, test_query.cpp, test_simulation.cpp,
test_velocity.cpp, test_foam.cpp, test_cascade.cpp, test_c_api.cpp,
test_threading.cpp, test_interaction.cpp and test_c_interaction.cpp (the
cases on this path; each cites its file:lines) against ocean::Ocean,
CascadeStack, FoamField, InteractionField, WaterSurface and the ocean_*
functions, with the reference's tolerances: what a program written against
the reference's headers checks when it is linked against this library
instead. The case program is compiled with
g++ from the public headers, the way such a program would be, and run once;
each case is one test here.
"""
import os
import re
import shutil
import subprocess

import pytest

from oceanlib_amd import capi

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SOURCES = [os.path.join(HERE, "cpp_api", f) for f in ("reference_cases.cpp", "interaction_cases.cpp")]
LIB_DIR = os.path.dirname(capi.LIB_PATH)


def case_names():
    names = []
    for path in SOURCES:  # the program runs the files' case tables in this order
        with open(path) as f:
            names += re.findall(r'\{"([a-z_.0-9]+)", [a-z_0-9]+\}', f.read())
    return names


def build_program(out_dir) -> str:
    if shutil.which("g++") is None:
        pytest.skip("g++ missing")
    if not os.path.exists(capi.LIB_PATH):
        pytest.fail(f"{capi.LIB_PATH} not built (python -c 'import __graft_entry__ as g; g.build()')")
    exe = os.path.join(str(out_dir), "reference_cases")
    subprocess.run(["g++", "-std=c++20", "-O1", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                    *SOURCES, "-L", LIB_DIR, "-loceanlib_amd", f"-Wl,-rpath,{LIB_DIR}", "-o", exe],
                   check=True)
    return exe


def test_case_program_builds_from_the_public_headers(tmp_path):
    """No device needed: the program compiles against include/ocean as shipped,
    lists every case, and the validation case (no device work) passes."""
    exe = build_program(tmp_path)
    listed = subprocess.run([exe, "--list"], check=True, capture_output=True, text=True).stdout.split()
    assert listed == case_names() and len(listed) >= 50
    r = subprocess.run([exe, "api.bad_desc_throws"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "PASS api.bad_desc_throws" in r.stdout, r.stdout + r.stderr


@pytest.fixture(scope="module")
def case_results(tmp_path_factory):
    exe = build_program(tmp_path_factory.mktemp("reference_cases"))
    print("running the tests/cpp_api reference cases on the GPU", flush=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=600)
    results, last_run = {}, None
    for line in r.stdout.splitlines():
        kind, _, rest = line.partition(" ")
        if kind == "RUN":
            last_run = rest
        elif kind in ("PASS", "FAIL"):
            name, _, why = rest.partition(": ")
            results[name] = (kind, why)
    info = "\n".join(x for x in r.stdout.splitlines() if x.startswith("INFO"))
    print(info, flush=True)
    return results, r.returncode, last_run, r.stderr[-2000:]


@pytest.mark.gpu
@pytest.mark.parametrize("name", case_names())
def test_reference_case(case_results, name):
    results, rc, last_run, err = case_results
    assert name in results, f"no result: the program exited {rc} while running {last_run}\n{err}"
    kind, why = results[name]
    assert kind == "PASS", why
