"""CPU oracle of the converged height query and of the proposal's training
pairs (test infrastructure only: the library never imports it).

The map is the reference's fixed-point inversion (src/ocean.cpp:375-400):
u <- X - Dx(u, v), v <- Z - Dz(u, v), with Dx, Dz the periodic bilinear tap
of the displacement buffer (src/ocean.cpp:322-358, weights (1-tu)(1-tv),
tu(1-tv), (1-tu)tv, tu tv). Here it runs in float64 in (integer base cell +
offset) coordinates, to a tolerance far below the GPU's, so its fixed point is
the exact solution the GPU's converged query approximates; a cascade's height
is the sum of the levels' heights at their own solved points
(src/cascade.cpp:118-130).

Features and spacing restate oceanlib_amd/csrc/proposal_kernels.hip
(level_features, spacing_kernel) in float32 numpy, operation for operation.
"""
from __future__ import annotations

import numpy as np


def base_cell(X, Z, n, L):
    g = np.float64(n) / np.float64(np.float32(L))
    gx = np.asarray(X, np.float32).astype(np.float64) * g
    gz = np.asarray(Z, np.float32).astype(np.float64) * g
    fx, fz = np.floor(gx), np.floor(gz)
    ix = (fx.astype(np.int64) & (n - 1))
    iz = (fz.astype(np.int64) & (n - 1))
    return ix, iz, gx - fx, gz - fz


def tap(disp, n, ix, iz, px, pz, channels=(0, 2)):
    """Bilinear channels of an (n*n, 4) float32 buffer at cell (ix+px, iz+pz), float64."""
    d = np.asarray(disp, np.float32).reshape(n * n, 4).astype(np.float64)
    flx, flz = np.floor(px), np.floor(pz)
    tu, tv = px - flx, pz - flz
    x0 = (ix + flx.astype(np.int64)) & (n - 1)
    z0 = (iz + flz.astype(np.int64)) & (n - 1)
    x1, z1 = (x0 + 1) & (n - 1), (z0 + 1) & (n - 1)
    w = ((1 - tu) * (1 - tv), tu * (1 - tv), (1 - tu) * tv, tu * tv)
    idx = (z0 * n + x0, z0 * n + x1, z1 * n + x0, z1 * n + x1)
    return [sum(d[i, c] * wi for i, wi in zip(idx, w)) for c in channels]


def solve_level(disp, n, L, X, Z, tol=1e-9, max_iter=500, du0=None, dv0=None):
    """Solved offsets (du, dv) in cells, iterations, converged mask."""
    ix, iz, fx, fz = base_cell(X, Z, n, L)
    inv_cell = n / np.float64(np.float32(L))
    du = np.zeros_like(fx) if du0 is None else np.asarray(du0, np.float64).copy()
    dv = np.zeros_like(fz) if dv0 is None else np.asarray(dv0, np.float64).copy()
    it = np.zeros(fx.shape, np.int64)
    done = np.zeros(fx.shape, bool)
    for k in range(1, max_iter + 1):
        dx, dz = tap(disp, n, ix, iz, fx + du, fz + dv)
        nu, nv = -dx * inv_cell, -dz * inv_cell
        step = np.maximum(np.abs(nu - du), np.abs(nv - dv)) / inv_cell
        act = ~done
        du = np.where(act, nu, du)
        dv = np.where(act, nv, dv)
        newly = act & (step <= tol)
        it[newly] = k
        done |= newly
        if done.all():
            break
    it[~done] = max_iter
    return du, dv, it, done


def height_converged(levels, X, Z, tol=1e-9):
    """levels: list of (disp (n*n*4 float32), n, L). Sum of solved heights (float64)."""
    h = np.zeros(np.shape(X))
    ok = np.ones(np.shape(X), bool)
    for disp, n, L in levels:
        du, dv, _, done = solve_level(disp, n, L, X, Z, tol)
        ix, iz, fx, fz = base_cell(X, Z, n, L)
        (hl,) = tap(disp, n, ix, iz, fx + du, fz + dv, channels=(1,))
        h += hl
        ok &= done
    return h, ok


def spacing(disp, n, L):
    """Stencil spacing of a level: the power of two nearest 2 rms / cell in
    log scale (s doubles while 2 rms / cell >= s sqrt 2), 1 .. 16."""
    d = np.asarray(disp, np.float32).reshape(n * n, 4).astype(np.float64)
    rms = np.sqrt(np.sum(d[:, 0] ** 2 + d[:, 2] ** 2) / (2.0 * n * n))
    cell = np.float64(np.float32(L)) / n
    r = 2.0 * rms / cell
    s = 1
    while s < 16 and r >= s * 1.4142135623730951:
        s *= 2
    return s


def level_features(disp, n, L, s, X, Z):
    """The 20 float32 features of one level (rows x 20): the (dx, dz) of the
    3 x 3 nodes of the level's coarse grid (every s-th cell) around the node
    nearest the query, over s * cell, then the query's offset from that node
    in units of s cells."""
    d = np.asarray(disp, np.float32).reshape(n * n, 4)
    ix, iz, fx, fz = base_cell(X, Z, n, L)
    fs = np.float32(s)
    ux = (ix.astype(np.float32) + fx.astype(np.float32)) / fs
    uz = (iz.astype(np.float32) + fz.astype(np.float32)) / fs
    rx = np.floor(ux + np.float32(0.5)).astype(np.float32)
    rz = np.floor(uz + np.float32(0.5)).astype(np.float32)
    cx, cz = rx.astype(np.int64), rz.astype(np.int64)
    nc = n // s
    inv_h = np.float32(n) / (np.float32(L) * fs)
    f = []
    for bz in (-1, 0, 1):
        for ax in (-1, 0, 1):
            xc = (cx + ax) & (nc - 1)
            zc = (cz + bz) & (nc - 1)
            v = d[(zc * s) * n + xc * s]
            f.append(v[:, 0] * inv_h)
            f.append(v[:, 2] * inv_h)
    f.append((ux - rx).astype(np.float32))
    f.append((uz - rz).astype(np.float32))
    return np.stack(f, 1).astype(np.float32)
