"""GPU tests of the converged height query and the learned start point.

The network's kernels (forward, backward, Adam; bf16 MFMA, fp32 accumulate)
are compared with the float64 numpy oracle that applies the same bf16
rounding points (oracle/mlp_oracle.py, itself pinned by finite differences in
tests/test_mlp_oracle.py). The converged queries are compared with the
float64 fixed point of the same bilinear map (oracle/inversion_oracle.py) on
the oracle's own ocean buffers, which the GPU's buffers equal bit for bit:
every height within 0.01 mm, plain and learned alike.
"""
import ctypes as C
import os

import numpy as np
import pytest

import inversion_oracle as inv
import mlp_oracle as mo
from oceanlib_amd import capi
from oceanlib_amd.hipmem import DeviceArray
from oceanlib_amd.workloads import desc_from
from oracle_ffi import OracleOcean

pytestmark = pytest.mark.gpu

PARAMS_FILE = os.path.join(os.path.dirname(os.path.dirname(__file__)), "oceanlib_amd", "data",
                           "height_proposal.npy")
TOL = 2.5e-6          # metres: the converged queries' step tolerance
MAX_ERR = 1.0e-5      # metres: 0.01 mm


@pytest.fixture(scope="module")
def lib():
    return capi.load()


def model(lib, params, max_queries=1 << 16):
    st = C.c_int(-1)
    p = None if params is None else np.ascontiguousarray(params, np.float32)
    h = lib.ocean_amd_proposal_create(None if p is None else p.ctypes.data_as(capi._F), max_queries,
                                      C.byref(st))
    assert h, st.value
    return h


def random_params(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    p = mo.init_params(seed).astype(np.float64)
    p += rng.normal(0, 0.05 * scale, p.size)
    return p.astype(np.float32)


def test_forward_matches_oracle(lib):
    rng = np.random.default_rng(4)
    p = random_params(1)
    rows = 1000  # not a multiple of the 64-row tiles: the tail rows are padding
    x = rng.normal(0, 1, (rows, mo.IN)).astype(np.float32)
    m = model(lib, p)
    try:
        fx = DeviceArray.from_numpy(x)
        out = DeviceArray((rows, mo.OUT), np.float32)
        assert lib.ocean_amd_proposal_forward(m, fx.ptr, rows, out.ptr) == 0
        assert lib.ocean_amd_proposal_synchronize(m) == 0
        got = out.numpy()
    finally:
        lib.ocean_amd_proposal_destroy(m)
    want, _ = mo.forward(p, x)
    err = np.abs(got - want)
    scale = np.abs(want).max()
    # fp32 accumulation order differs from float64; a hidden activation can
    # land on the other side of a bf16 rounding boundary (one bf16 ulp)
    assert np.median(err) <= 1e-5 * scale, (np.median(err), scale)
    assert err.max() <= 2e-2 * scale, (err.max(), scale)


def minibatch(seed, rows):
    rng = np.random.default_rng(seed)
    x = rng.normal(0, 1, (rows, mo.IN)).astype(np.float32)
    x[:, 60:] = 0.0
    t = np.zeros((rows, mo.OUT), np.float32)
    t[:, :6] = rng.normal(0, 1, (rows, 6))
    return x, t


def test_gradient_and_adam_step_match_oracle(lib):
    batch = 256
    x, t = minibatch(5, batch)
    p = random_params(2)
    m = model(lib, p)
    try:
        fx, ft = DeviceArray.from_numpy(x), DeviceArray.from_numpy(t)
        rows = DeviceArray.from_numpy(np.arange(batch, dtype=np.uint32))
        g = np.zeros(mo.PARAMS, np.float32)
        loss = C.c_float()
        assert lib.ocean_amd_proposal_gradient(m, fx.ptr, ft.ptr, rows.ptr, batch,
                                               g.ctypes.data_as(capi._F), C.byref(loss)) == 0
        want_loss, want_g = mo.loss_and_grad(p, x, t)
        assert abs(loss.value - want_loss) <= 1e-4 * want_loss
        rel = np.linalg.norm(g - want_g) / np.linalg.norm(want_g)
        assert rel <= 1e-2, rel
        # per-tensor check (a wrong tile layout would scramble one tensor)
        offs = np.cumsum([0, mo.IN * mo.HID, mo.HID, mo.HID * mo.HID, mo.HID, mo.HID * mo.OUT, mo.OUT])
        for a, b in zip(offs[:-1], offs[1:]):
            ga, wa = g[a:b], want_g[a:b]
            if np.linalg.norm(wa) > 0:
                assert np.linalg.norm(ga - wa) <= 2e-2 * np.linalg.norm(wa), (a, b)
            else:
                assert np.all(ga == 0)
        # one Adam step from zero moments uses exactly this gradient
        lr = 1e-3
        assert lib.ocean_amd_proposal_train_step(m, fx.ptr, ft.ptr, rows.ptr, batch, lr, None) == 0
        newp = np.zeros(mo.PARAMS, np.float32)
        assert lib.ocean_amd_proposal_get_params(m, newp.ctypes.data_as(capi._F)) == 0
        want_p, _, _ = mo.adam_step(p, np.zeros(mo.PARAMS), np.zeros(mo.PARAMS), g, 1, lr)
        np.testing.assert_allclose(newp, want_p, rtol=1e-6, atol=1e-9)
        # invalid minibatch sizes are refused
        assert lib.ocean_amd_proposal_train_step(m, fx.ptr, ft.ptr, rows.ptr, 100, lr, None) != 0
        assert lib.ocean_amd_proposal_train_step(m, fx.ptr, ft.ptr, rows.ptr, 16384, lr, None) != 0
    finally:
        lib.ocean_amd_proposal_destroy(m)


def test_training_fits_a_smooth_map(lib):
    """A few hundred steps of the device trainer fit a smooth 60 -> 6 map."""
    n = 8192
    rng = np.random.default_rng(6)
    x = rng.uniform(-1, 1, (n, mo.IN)).astype(np.float32)
    x[:, 60:] = 0.0
    A = rng.normal(0, 0.3, (60, 6))
    t = np.zeros((n, mo.OUT), np.float32)
    t[:, :6] = np.tanh(x[:, :60] @ A)
    m = model(lib, mo.init_params(3))
    try:
        fx, ft = DeviceArray.from_numpy(x), DeviceArray.from_numpy(t)
        losses = []
        for k in range(400):
            idx = DeviceArray.from_numpy(rng.permutation(n)[:1024].astype(np.uint32))
            loss = C.c_float()
            assert lib.ocean_amd_proposal_train_step(m, fx.ptr, ft.ptr, idx.ptr, 1024, 3e-3,
                                                     C.byref(loss)) == 0
            losses.append(loss.value)
        assert np.mean(losses[-20:]) < 0.1 * np.mean(losses[:5]), (losses[:5], losses[-20:])
    finally:
        lib.ocean_amd_proposal_destroy(m)


# ---- queries on a three-level sea ------------------------------------------------

SEA = [dict(size=512, patch_length=L, wind_speed=22.0, choppiness=1.3, seed=s)
       for L, s in ((800.0, 1337), (150.0, 1338), (25.0, 1339))]
T_SEA = 5.0


@pytest.fixture(scope="module")
def sea(lib):
    descs = [desc_from(kw) for kw in SEA]
    arr = (capi.OceanDesc * 3)(*descs)
    st = C.c_int(-1)
    b = lib.ocean_amd_batch_create(arr, 3, C.byref(st))
    assert b, st.value
    lib.ocean_amd_batch_update(b, T_SEA)
    assert lib.ocean_amd_batch_synchronize(b) == 0
    bufs = []
    for kw, d in zip(SEA, descs):
        o = OracleOcean(d)
        o.update(T_SEA)
        bufs.append((o.displacement(), kw["size"], kw["patch_length"]))
        o.close()
    yield b, bufs
    lib.ocean_amd_batch_destroy(b)


def points(n, seed):
    rng = np.random.default_rng(seed)
    return rng.uniform(0.0, 2000.0, (n, 2)).astype(np.float32)


def run_queries(lib, b, xz, m):
    n = len(xz)
    dxz = DeviceArray.from_numpy(xz)
    h = DeviceArray((n,), np.float32)
    it = DeviceArray((n, 3), np.uint8)
    fl = DeviceArray((n,), np.uint8)
    rc = lib.ocean_amd_batch_height_converged(b, 0, 3, dxz.ptr, n, TOL, m, h.ptr, it.ptr, fl.ptr)
    assert rc == 0
    assert lib.ocean_amd_batch_synchronize(b) == 0
    return h.numpy(), it.numpy().astype(np.int64), fl.numpy()


def test_pairs_match_oracle(lib, sea):
    b, bufs = sea
    n = 4096
    xz = points(n, 7)
    dxz = DeviceArray.from_numpy(xz)
    f = DeviceArray((n, mo.IN), np.float32)
    t = DeviceArray((n, mo.OUT), np.float32)
    ok = DeviceArray((n,), np.uint8)
    sp = DeviceArray((3,), np.int32)
    assert lib.ocean_amd_proposal_pairs(b, 0, dxz.ptr, n, 1e-6, f.ptr, t.ptr, ok.ptr, sp.ptr) == 0
    assert lib.ocean_amd_batch_synchronize(b) == 0
    f, t, ok, sp = f.numpy(), t.numpy(), ok.numpy(), sp.numpy()
    X, Z = xz[:, 0], xz[:, 1]
    assert ok.all()
    for k, (disp, nn, L) in enumerate(bufs):
        s = inv.spacing(disp, nn, L)
        assert sp[k] == s
        np.testing.assert_array_equal(f[:, 20 * k:20 * k + 20], inv.level_features(disp, nn, L, s, X, Z))
        du, dv, _, done = inv.solve_level(disp, nn, L, X, Z, tol=1e-10)
        assert done.all()
        cell = L / nn
        assert np.abs(t[:, 2 * k] * s - du).max() * cell < 2e-6
        assert np.abs(t[:, 2 * k + 1] * s - dv).max() * cell < 2e-6
    assert np.all(f[:, 60:] == 0) and np.all(t[:, 6:] == 0)


def check_starts_against_oracle_network(lib, b, n, seed):
    """Every level's proposed start offset from ocean_amd_proposal_starts, in
    cells, against the float64 oracle network on the same points' features
    from the pairs kernel, within the bf16 tolerance of the forward test.
    Returns the levels' spacings."""
    xz = points(n, seed)
    params = np.load(PARAMS_FILE, allow_pickle=False) if os.path.exists(PARAMS_FILE) \
        else random_params(5)
    m = model(lib, params, max_queries=n)
    try:
        dxz = DeviceArray.from_numpy(xz)
        f = DeviceArray((n, mo.IN), np.float32)
        t = DeviceArray((n, mo.OUT), np.float32)
        ok = DeviceArray((n,), np.uint8)
        sp = DeviceArray((3,), np.int32)
        assert lib.ocean_amd_proposal_pairs(b, 0, dxz.ptr, n, 1e-6, f.ptr, t.ptr, ok.ptr, sp.ptr) == 0
        st = DeviceArray((3, n, 2), np.float32)
        sp2 = DeviceArray((3,), np.int32)
        assert lib.ocean_amd_proposal_starts(b, 0, dxz.ptr, n, m, st.ptr, sp2.ptr) == 0
        assert lib.ocean_amd_batch_synchronize(b) == 0
        # more queries than the model's scratch, and no model, are refused
        assert lib.ocean_amd_proposal_starts(b, 0, dxz.ptr, n + 1, m, st.ptr, None) != 0
        assert lib.ocean_amd_proposal_starts(b, 0, dxz.ptr, n, None, st.ptr, None) != 0
    finally:
        lib.ocean_amd_proposal_destroy(m)
    f, sp, sp2, got = f.numpy(), sp.numpy(), sp2.numpy(), st.numpy()
    assert np.array_equal(sp, sp2)
    want, _ = mo.forward(params, f)
    for lv in range(3):
        w = want[:, 2 * lv:2 * lv + 2] * sp[lv]
        err = np.abs(got[lv] - w)
        scale = np.abs(w).max()
        if n >= 16:  # the median of a handful of points is one point's error
            assert np.median(err) <= 1e-5 * scale, (lv, np.median(err), scale)
        assert err.max() <= 2e-2 * scale, (lv, err.max(), scale)
    return sp


def test_proposal_starts_match_oracle_network(lib, sea):
    """The proposal kernel with every level on a coarse grid (its
    straight-line gather of 48-byte record spans) against the oracle network.
    The point count is not a multiple of a round's 64 queries."""
    b, _ = sea
    check_starts_against_oracle_network(lib, b, 4096 + 37, 12)


@pytest.mark.parametrize("n", [1, 63, 64, 65, 64 * 4 * 3 + 1])
def test_proposal_starts_at_round_boundaries(lib, sea, n):
    """Point counts at and around one wavefront's round of 64 queries (and
    one more than a few blocks' rounds): the lanes past the end gather at a
    stale point and store nothing; every stored proposal matches the oracle
    network."""
    b, _ = sea
    check_starts_against_oracle_network(lib, b, n, 100 + n)


def test_proposal_starts_with_a_level_past_the_coarse_grids(lib):
    """A 1024^2 level of small waves has stencil spacing 1, so its node grid
    (1024 a side) is larger than the coarse-grid scratch: the proposal kernel
    takes its general gather, reading that level's nodes from the
    displacement itself, and must still match the oracle network."""
    levels = [dict(size=1024, patch_length=1000.0, wind_speed=3.0, choppiness=1.0, seed=21),
              dict(size=512, patch_length=150.0, wind_speed=22.0, choppiness=1.3, seed=22),
              dict(size=512, patch_length=25.0, wind_speed=22.0, choppiness=1.3, seed=23)]
    arr = (capi.OceanDesc * 3)(*[desc_from(kw) for kw in levels])
    st = C.c_int(-1)
    b = lib.ocean_amd_batch_create(arr, 3, C.byref(st))
    assert b, st.value
    try:
        lib.ocean_amd_batch_update(b, T_SEA)
        assert lib.ocean_amd_batch_synchronize(b) == 0
        sp = check_starts_against_oracle_network(lib, b, 2048 + 5, 13)
    finally:
        lib.ocean_amd_batch_destroy(b)
    assert 1024 // sp[0] > 512, sp


def test_plain_converged_heights_within_0_01_mm(lib, sea):
    b, bufs = sea
    xz = points(16384, 8)
    h, it, fl = run_queries(lib, b, xz, None)
    want, ok = inv.height_converged(bufs, xz[:, 0], xz[:, 1])
    assert ok.all() and np.all(fl == 0)
    err = np.abs(h - want)
    print(f"plain: max err {err.max() * 1e3:.4f} mm, taps/level mean {it.mean(0)}, "
          f"wave max {it.reshape(-1, 64, 3).max(1).mean(0)}")
    assert err.max() <= MAX_ERR
    # the iteration count is the oracle's own count at the same tolerance, give or take one
    for k, (disp, nn, L) in enumerate(bufs):
        _, _, it_o, _ = inv.solve_level(disp, nn, L, xz[:, 0], xz[:, 1], tol=TOL)
        assert np.mean(np.abs(it[:, k] - it_o) <= 1) > 0.99


def test_learned_zero_and_bad_models_still_converge(lib, sea):
    """A zero network proposes the query point itself; a wild one is caught by
    the fallback; both end within 0.01 mm."""
    b, bufs = sea
    xz = points(8192, 9)
    want, _ = inv.height_converged(bufs, xz[:, 0], xz[:, 1])
    for params, name in ((None, "zero"), (random_params(4, scale=400.0), "wild")):
        m = model(lib, params)
        try:
            h, it, fl = run_queries(lib, b, xz, m)
        finally:
            lib.ocean_amd_proposal_destroy(m)
        assert np.all((fl & 1) == 0), name
        assert np.abs(h - want).max() <= MAX_ERR, name
        if name == "wild":
            assert np.mean((fl & 2) != 0) > 0.05  # the fallback was exercised
    # a model rejects more levels than it was built for, and more queries than its scratch
    m = model(lib, None, max_queries=1024)
    try:
        dxz = DeviceArray.from_numpy(xz[:2048])
        hh = DeviceArray((2048,), np.float32)
        assert lib.ocean_amd_batch_height_converged(b, 0, 3, dxz.ptr, 2048, TOL, m, hh.ptr,
                                                    None, None) != 0
    finally:
        lib.ocean_amd_proposal_destroy(m)


def test_learned_query_on_fewer_levels(lib, sea):
    """Learned queries over one or two levels (the proposal kernel's general
    gather, the missing levels' features zero) and over the stack's later
    levels only: every height within 0.01 mm of the oracle's fixed point over
    the same levels, every level converged (a poor proposal falls back)."""
    b, bufs = sea
    n = 4096 + 11
    xz = points(n, 14)
    params = np.load(PARAMS_FILE, allow_pickle=False) if os.path.exists(PARAMS_FILE) \
        else random_params(5)
    m = model(lib, params, max_queries=n)
    try:
        dxz = DeviceArray.from_numpy(xz)
        for first, count in ((0, 1), (0, 2), (1, 2)):
            h = DeviceArray((n,), np.float32)
            it = DeviceArray((n, count), np.uint8)
            fl = DeviceArray((n,), np.uint8)
            assert lib.ocean_amd_batch_height_converged(b, first, count, dxz.ptr, n, TOL, m, h.ptr,
                                                        it.ptr, fl.ptr) == 0
            assert lib.ocean_amd_batch_synchronize(b) == 0
            want, ok = inv.height_converged(bufs[first:first + count], xz[:, 0], xz[:, 1])
            assert ok.all()
            assert np.all((fl.numpy() & 1) == 0), (first, count)
            assert np.abs(h.numpy() - want).max() <= MAX_ERR, (first, count)
    finally:
        lib.ocean_amd_proposal_destroy(m)


@pytest.mark.skipif(not os.path.exists(PARAMS_FILE), reason="no trained proposal committed")
def test_trained_proposal_saves_iterations_at_equal_error(lib, sea):
    b, bufs = sea
    xz = points(65536, 10)
    want, _ = inv.height_converged(bufs, xz[:, 0], xz[:, 1])
    hp, itp, flp = run_queries(lib, b, xz, None)
    m = model(lib, np.load(PARAMS_FILE, allow_pickle=False), max_queries=len(xz))
    try:
        hl, itl, fll = run_queries(lib, b, xz, m)
    finally:
        lib.ocean_amd_proposal_destroy(m)
    ep, el = np.abs(hp - want).max(), np.abs(hl - want).max()
    wp = itp.reshape(-1, 64, 3).max(1).mean(0)
    wl = itl.reshape(-1, 64, 3).max(1).mean(0)
    print(f"plain   taps/level mean {itp.mean(0)} wave-max {wp} max err {ep * 1e3:.4f} mm")
    print(f"learned taps/level mean {itl.mean(0)} wave-max {wl} max err {el * 1e3:.4f} mm "
          f"fallbacks {np.mean((fll & 2) != 0):.4f}")
    assert ep <= MAX_ERR and el <= MAX_ERR
    assert np.all((fll & 1) == 0)
    assert itl.mean() < itp.mean()
    assert wl.sum() < wp.sum()


@pytest.mark.skipif(not os.path.exists(PARAMS_FILE), reason="no trained proposal committed")
def test_full_config_million_points_plain_and_learned(lib, sea):
    """The learned-proposal config at its full size: 1M query points uniform
    over the 2 km square on the three-level sea, the plain converged query and
    the learned start point both within 0.01 mm of the float64 fixed point on
    every point, no point left unconverged."""
    b, bufs = sea
    xz = points(1 << 20, 11)
    want, ok = inv.height_converged(bufs, xz[:, 0], xz[:, 1])
    assert ok.all()
    hp, itp, flp = run_queries(lib, b, xz, None)
    m = model(lib, np.load(PARAMS_FILE, allow_pickle=False), max_queries=len(xz))
    try:
        hl, itl, fll = run_queries(lib, b, xz, m)
    finally:
        lib.ocean_amd_proposal_destroy(m)
    ep, el = np.abs(hp - want).max(), np.abs(hl - want).max()
    print(f"1M points: plain max err {ep * 1e3:.4f} mm, taps/level {itp.mean(0)}; "
          f"learned max err {el * 1e3:.4f} mm, taps/level {itl.mean(0)}, "
          f"fallbacks {np.mean((fll & 2) != 0):.5f}")
    assert np.all((flp & 1) == 0) and np.all((fll & 1) == 0)
    assert ep <= MAX_ERR and el <= MAX_ERR
