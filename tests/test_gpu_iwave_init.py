"""GPU tests of the interaction field's set-up on the device.

- The stencil fit (reference src/core/iwave_kernel.cpp:177-437): normal
  equations and realised-symbol scan on the device, the scan's extremes
  certified and recomputed on the host; taps and the scan summary (max / min
  symbol, band edge, peak relative error) bit for bit against the oracle,
  whose summary is pinned to the reference build on the CPU
  (tests/test_iwave_oracle.py::test_stencil_summary_vs_reference).
- The certification's premise: the device scan (device double cosine) is
  within B / 50 of the exact scan at every point, B = 1e-12 * sum |taps| the
  bound the host certifies with.
- The obstruction tables (reference src/interaction.cpp:255-292): solid
  cells in index order and their nearest fluid cells, built on the device,
  against the oracle's BFS on every field size and through recentre shifts.
"""
import ctypes as C
import math

import numpy as np
import pytest

from oceanlib_amd import capi
from oceanlib_amd.hipmem import DeviceArray
from iwave_helpers import Field, idesc
from oracle_ffi import assert_bitexact, fptr, oracle

pytestmark = pytest.mark.gpu

KNONE = 0xFFFFFFFF


@pytest.fixture(scope="module")
def lib():
    lib = capi.load()
    lib.ocean_hip_iwave_symbol_scan.restype = C.c_int
    lib.ocean_hip_iwave_symbol_scan.argtypes = [C.c_void_p, C.c_uint32, C.c_double, C.c_double,
                                                C.c_double, C.c_uint32, C.c_void_p, C.c_void_p]
    return lib


def dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def u32ptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


STENCILS = [(r, 0.1, 0.0) for r in range(1, 13)] + [
    (6, 0.25, 0.0), (4, 0.25, 2.0), (12, 0.05, 0.0), (8, 0.5, 0.3), (2, 1.0, 0.0),
    (3, 0.1, 1.5), (10, 0.02, 5.0), (5, 2.0, 0.5), (7, 0.125, 0.0), (6, 0.0625, 0.05),
    (9, 64.0 / 256.0, 3.0), (6, 1000.0 / 1024.0, 0.0), (11, 0.3, 0.01), (1, 5.0, 100.0)]


@pytest.mark.parametrize("radius,cell,depth", STENCILS)
def test_device_stencil_equals_oracle(lib, radius, cell, depth):
    w = 2 * radius + 1
    a, b = np.zeros(w * w, np.float32), np.zeros(w * w, np.float32)
    sa, sb = np.zeros(4), np.zeros(4)
    assert lib.ocean_amd_build_iwave_stencil(radius, cell, depth, fptr(a), dptr(sa)) == w * w
    assert oracle().orc_iwave_stencil(radius, cell, depth, fptr(b), dptr(sb)) == w * w
    assert_bitexact(a, b, f"taps r={radius} cell={cell} depth={depth}")
    assert sa.tobytes() == sb.tobytes(), (sa, sb)
    m = C.c_double()
    assert lib.ocean_amd_build_iwave_taps(radius, cell, depth, fptr(a), C.byref(m)) == w * w
    assert m.value == sb[0]


@pytest.mark.parametrize("radius,cell,depth", [(1, 0.1, 0.0), (6, 0.1, 0.0), (12, 0.05, 0.0),
                                               (8, 0.5, 0.3), (3, 2.0, 1.5)])
def test_symbol_scan_within_bound(lib, radius, cell, depth):
    w = 2 * radius + 1
    taps, sym = np.zeros(w * w, np.float32), np.zeros(4)
    assert oracle().orc_iwave_stencil(radius, cell, depth, fptr(taps), dptr(sym)) == w * w
    scan = 192
    k_max = math.pi / cell
    cell_f = float(np.float32(cell))
    dtaps = DeviceArray.from_numpy(taps)
    out = DeviceArray(((scan + 1) ** 2, 2), np.float64)
    try:
        assert lib.ocean_hip_iwave_symbol_scan(dtaps.ptr, radius, cell_f, k_max, depth, scan,
                                               out.ptr, None) == 0
        dev = out.numpy()
    finally:
        dtaps.free()
        out.free()
    exact = np.zeros((scan + 1) ** 2)
    oracle().orc_iwave_symbol_scan(fptr(taps), radius, np.float32(cell), k_max, scan, dptr(exact))
    abs_taps = float(np.abs(taps.astype(np.float64)).sum())
    gap = float(np.abs(dev[:, 0] - exact).max())
    print(f"r={radius} cell={cell}: max |device - exact| = {gap / abs_taps:.2e} * sum|taps|")
    assert gap <= 1e-12 * abs_taps / 50
    # the scan's extremes are the oracle's
    assert max(0.0, exact.max()) == sym[0] and min(0.0, exact.min()) == sym[1]
    # the true symbol column: k tanh(k depth) (k when deep) at |k|
    zi, xi = np.divmod(np.arange((scan + 1) ** 2), scan + 1)
    kx, kz = k_max * xi / scan, k_max * zi / scan
    kk = np.sqrt(kx * kx + kz * kz)
    tr = kk if depth <= 0 else kk * np.tanh(kk * depth)
    assert np.allclose(dev[:, 1], tr, rtol=1e-13, atol=0)


def make_masks(n, rng):
    zz, xx = np.mgrid[0:n, 0:n]
    out = {
        "noise 5%": (rng.random((n, n)) < 0.05),
        "noise 40%": (rng.random((n, n)) < 0.4),
        "noise 95%": (rng.random((n, n)) < 0.95),
        "disc": ((xx - n / 2) ** 2 + (zz - n / 3) ** 2 < (n / 3.5) ** 2),
        "stripes": ((xx + 2 * zz) % 7 < 3),
        "all solid": np.ones((n, n), bool),
        "all fluid": np.zeros((n, n), bool),
    }
    lone = np.ones((n, n), bool)
    lone[n // 3, (2 * n) // 3] = False
    out["one fluid cell"] = lone
    two = np.ones((n, n), bool)
    two[0, n - 1] = False
    two[n - 1, 0] = False
    out["two far fluid cells"] = two
    blobs = np.zeros((n, n), bool)
    for _ in range(12):
        cx, cz, r = rng.uniform(0, n), rng.uniform(0, n), rng.uniform(1, n / 6)
        blobs |= (xx - cx) ** 2 + (zz - cz) ** 2 < r * r
    out["blobs"] = blobs
    return {k: v.astype(np.uint8) for k, v in out.items()}


def reference_tables(mask):
    n = mask.shape[0]
    near = np.zeros(n * n, np.uint32)
    m = np.ascontiguousarray(mask)
    oracle().orc_nearest_fluid(m.ctypes.data_as(C.POINTER(C.c_uint8)), n, u32ptr(near))
    solid = np.nonzero(m.ravel())[0].astype(np.uint32)
    return solid, near[solid]


def device_tables(lib, h, cells):
    solid = np.zeros(cells, np.uint32)
    source = np.zeros(cells, np.uint32)
    k = lib.ocean_amd_interaction_obstruction_tables(h, u32ptr(solid), u32ptr(source), cells)
    return solid[:k], source[:k]


@pytest.mark.parametrize("n", [32, 64, 128, 256, 512, 1024])
def test_mask_tables_equal_reference_bfs(lib, n):
    d = idesc(lib.ocean_interaction_desc_init, size=n, extent=0.1 * n, kernel_radius=3,
              absorb_cells=4)
    f = Field("gpu", lib, d)
    rng = np.random.default_rng(n)
    try:
        assert lib.ocean_amd_interaction_obstruction_tables(f.h, None, None, 0) == 0
        for name, mask in make_masks(n, rng).items():
            f.set_obstruction(mask)
            solid, source = device_tables(lib, f.h, n * n)
            want_solid, want_source = reference_tables(mask)
            assert np.array_equal(solid, want_solid), (n, name)
            assert np.array_equal(source, want_source), (n, name)
            if name == "all solid":
                assert (source == KNONE).all()
        f.set_obstruction(None)
        assert lib.ocean_amd_interaction_obstruction_tables(f.h, None, None, 0) == 0
    finally:
        f.close()


@pytest.mark.parametrize("n", [64, 256])
def test_mask_tables_through_recentre_vs_oracle_field(lib, n):
    """set_obstruction, then recentre shifts (partial, none, beyond the grid):
    the device's tables against the oracle field's (orc_iw_mask_tables)."""
    d = idesc(lib.ocean_interaction_desc_init, size=n, extent=0.1 * n, kernel_radius=3,
              absorb_cells=4)
    g = Field("gpu", lib, d)
    o = Field("orc", oracle(), d)
    rng = np.random.default_rng(7 + n)
    cells = n * n
    half = 0.05 * n
    try:
        masks = make_masks(n, rng)
        for name in ("blobs", "noise 40%", "disc"):
            g.set_obstruction(masks[name])
            o.set_obstruction(masks[name])
            for wx, wz in ((half + 0.1 * 5, half - 0.1 * 3), (half + 0.1 * 5, half - 0.1 * 3),
                           (half - 0.1 * 11, half + 0.1 * 17), (half + 0.1 * (n + 3), half)):
                assert g.recenter(wx, wz) == o.recenter(wx, wz)
                solid, source = device_tables(lib, g.h, cells)
                ws, wsrc = np.zeros(cells, np.uint32), np.zeros(cells, np.uint32)
                k = oracle().orc_iw_mask_tables(o.h, u32ptr(ws), u32ptr(wsrc))
                assert np.array_equal(solid, ws[:k]), (name, wx, wz)
                assert np.array_equal(source, wsrc[:k]), (name, wx, wz)
    finally:
        g.close()
        o.close()
