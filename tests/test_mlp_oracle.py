"""Pins of the CPU oracles for the learned proposal (no GPU).

- the network's backward formulas against central finite differences of the
  plain float64 network (a dropped term, a wrong transpose or a missing ReLU
  mask fails it);
- bf16 rounding against the definition (round to nearest even on the top 16
  bits) on ties, carries and specials;
- Adam against the textbook update written out for one scalar;
- the inversion oracle against closed forms: a constant displacement field
  (solution X - D exactly), zero choppiness (the identity), and the
  reference's 4-step result on the oracle's own ocean buffers.
"""
import numpy as np
import pytest

import inversion_oracle as inv
import mlp_oracle as mo


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = rng.normal(0, 0.3, mo.PARAMS)
    x = rng.normal(0, 1, (8, mo.IN))
    t = rng.normal(0, 1, (8, mo.OUT))
    loss, g = mo.loss_and_grad(p, x, t, rounding=False)
    idx = np.concatenate([rng.choice(mo.IN * mo.HID, 40, replace=False),
                          mo.IN * mo.HID + np.arange(0, mo.HID, 7),
                          mo.PARAMS - mo.OUT - mo.HID * mo.OUT + rng.choice(mo.HID * mo.OUT, 30),
                          mo.PARAMS - mo.OUT + np.arange(mo.OUT)])
    h = 1e-6
    for i in idx:
        e = np.zeros(mo.PARAMS)
        e[i] = h
        lp, _ = mo.loss_and_grad(p + e, x, t, rounding=False)
        lm, _ = mo.loss_and_grad(p - e, x, t, rounding=False)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])
    # padded outputs carry no gradient
    W3g = g[mo.PARAMS - mo.OUT - mo.HID * mo.OUT:mo.PARAMS - mo.OUT].reshape(mo.HID, mo.OUT)
    assert np.all(W3g[:, mo.USED_OUT:] == 0)


def test_bf16_rounding_definition():
    # bf16 keeps 8 significant bits: one ulp at 1.0 is 2^-7
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7 + 2 ** -8, 1.0 + 2 ** -8 + 2 ** -20,
                     1.0 + 2 ** -9, -2.5, 3.0e38, np.inf, -np.inf, 0.0, -0.0], np.float32)
    got = mo.bf16(vals)
    assert got[0] == 1.0
    assert got[1] == 1.0                                   # exact tie -> even (1.0)
    assert got[2] == np.float32(1.0 + 2 ** -6)             # tie between odd and even -> even
    assert got[3] == np.float32(1.0 + 2 ** -7)             # just above the tie -> up
    assert got[4] == 1.0                                   # a quarter ulp -> down
    assert got[5] == -2.5 and got[7] == np.inf and got[8] == -np.inf
    assert got[9] == 0.0 and np.signbit(got[10])
    nan = mo.bf16(np.array([np.nan], np.float32))
    assert np.isnan(nan[0])
    # every result has zero low 16 bits and is the nearest such value
    rng = np.random.default_rng(1)
    x = rng.normal(0, 100, 10000).astype(np.float32)
    r = mo.bf16(x)
    assert np.all(r.view(np.uint32) & 0xFFFF == 0)
    lo = (x.view(np.uint32) & 0xFFFF0000).view(np.float32)
    hi = ((x.view(np.uint32) & 0xFFFF0000) + 0x10000).view(np.float32)
    best = np.where(np.abs(x - lo) <= np.abs(hi - x), lo, hi)
    ties = np.abs(x - lo) == np.abs(hi - x)
    assert np.all((r == best) | ties)


def test_adam_matches_textbook():
    g, m, v, w = 0.5, 0.1, 0.01, 2.0
    step, lr = 3, 0.01
    a = 0.9 * m + 0.1 * g
    b = 0.999 * v + 0.001 * g * g
    want = w - lr * (a / (1 - 0.9 ** step)) / (np.sqrt(b / (1 - 0.999 ** step)) + 1e-8)
    got, a2, b2 = mo.adam_step([w], [m], [v], [g], step, lr)
    assert abs(got[0] - want) < 1e-6 and abs(a2[0] - a) < 1e-7 and abs(b2[0] - b) < 1e-7


def test_inversion_constant_field_and_identity():
    n, L = 16, 32.0
    disp = np.zeros((n * n, 4), np.float32)
    disp[:, 0], disp[:, 2], disp[:, 1] = 0.37, -0.81, 1.25
    X = np.array([0.0, 5.3, 31.9, -7.2, 1000.25], np.float32)
    Z = np.array([0.0, 2.2, 17.0, -3.3, -999.5], np.float32)
    du, dv, it, ok = inv.solve_level(disp, n, L, X, Z, tol=1e-12)
    cell = L / n
    assert ok.all() and np.all(it <= 2)
    np.testing.assert_allclose(du * cell, -0.37, atol=1e-12)
    np.testing.assert_allclose(dv * cell, 0.81, atol=1e-12)
    h, ok = inv.height_converged([(disp, n, L)], X, Z)
    np.testing.assert_allclose(h, 1.25, atol=1e-12)
    disp[:, 0] = disp[:, 2] = 0.0
    du, dv, it, ok = inv.solve_level(disp, n, L, X, Z)
    assert np.all(du == 0) and np.all(dv == 0) and np.all(it == 1)


@pytest.mark.parametrize("steps", [4])
def test_inversion_matches_reference_four_steps_on_oracle_buffers(steps):
    """Four iterations of the float64 oracle map from the query point land
    within float rounding of the reference's own four-step sample_at
    (here the C oracle's restatement of it, pinned to the reference)."""
    from oracle_ffi import OracleOcean
    from oceanlib_amd.workloads import desc_from
    o = OracleOcean(desc_from(dict(size=64, patch_length=40.0, wind_speed=14.0, choppiness=1.0,
                                   seed=11)))
    o.update(3.0)
    disp = o.displacement()
    rng = np.random.default_rng(2)
    X = rng.uniform(0, 40, 200).astype(np.float32)
    Z = rng.uniform(0, 40, 200).astype(np.float32)
    du, dv, _, _ = inv.solve_level(disp, 64, 40.0, X, Z, tol=0.0, max_iter=steps)
    ix, iz, fx, fz = inv.base_cell(X, Z, 64, 40.0)
    (h,) = inv.tap(disp, 64, ix, iz, fx + du, fz + dv, channels=(1,))
    want = np.array([o.sample(float(x), float(z)).height for x, z in zip(X, Z)])
    np.testing.assert_allclose(h, want, atol=2e-5)


def test_features_layout_and_spacing():
    """Spacing: the power of two nearest 2 rms / cell (log scale), closed form
    from the RMS; features: the coarse grid's nodes at multiples of s cells,
    the centre node nearest the query, offsets in units of s cells."""
    n, L = 32, 64.0
    rng = np.random.default_rng(3)
    disp = rng.normal(0, 0.5, (n * n, 4)).astype(np.float32)
    s = inv.spacing(disp, n, L)
    rms = np.sqrt(np.mean(disp[:, 0].astype(np.float64) ** 2 + disp[:, 2] ** 2) / 2)
    r = 2 * rms / (L / n)
    assert s == int(np.clip(2 ** np.rint(np.log2(r)), 1, 16))
    for scale, want in ((0.01, 1), (1.0, 1), (1.5, 2), (2.7, 2), (3.0, 4), (40.0, 16)):
        d = np.zeros((n * n, 4), np.float32)
        d[:, 0] = scale * (L / n) / 2 * np.sqrt(2.0)  # rms = scale * cell / 2 -> r = scale
        assert inv.spacing(d, n, L) == want, (scale, want)
    X = np.array([10.3], np.float32)
    Z = np.array([50.9], np.float32)
    for sp in (1, 4):
        f = inv.level_features(disp, n, L, sp, X, Z)
        assert f.shape == (1, 20)
        # centre node: the coarse node nearest the query (cell = 2 m, node = sp cells)
        ux, uz = 10.3 / 2 / sp, 50.9 / 2 / sp
        cx, cz = (int(np.floor(ux + 0.5)) % (n // sp)) * sp, (int(np.floor(uz + 0.5)) % (n // sp)) * sp
        inv_h = np.float32(n) / (np.float32(L) * np.float32(sp))
        assert f[0, 8] == np.float32(disp[cz * n + cx, 0] * inv_h)
        assert f[0, 9] == np.float32(disp[cz * n + cx, 2] * inv_h)
        # the +x neighbour is s cells further along the row (periodic)
        assert f[0, 10] == np.float32(disp[cz * n + (cx + sp) % n, 0] * inv_h)
        assert abs(f[0, 18] - (ux - round(ux))) < 1e-5 and abs(f[0, 19] - (uz - round(uz))) < 1e-5
