"""Trains the learned height_at start-point proposal on the GPU.

    python -m oceanlib_amd.proposal_train [--out oceanlib_amd/data/height_proposal.npy]

Training set: 64 seas (three 512^2 levels of 800 / 150 / 25 m, choppiness
1.3, wind 5 + 20 s / 63 m/s for sea s, seed s for its three levels (seeds
0..63), each at its own time), 156,250 points per sea uniform over a 2 km
square: 10 M (query point ->
solved start offsets) pairs, solved by the plain fixed-point iteration to
1e-6 m on the device (ocean_amd_proposal_pairs). The network is then trained
with the library's own forward/backward/Adam kernels
(ocean_amd_proposal_train_step), minibatches of 4096, a few epochs with a
step-decayed learning rate; 1 % of the pairs is held out for validation.
Writes the fp32 parameters (.npy) and a .json of what was done.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from oceanlib_amd import capi  # noqa: E402
from oceanlib_amd.hipmem import DeviceArray, synchronize  # noqa: E402
from oceanlib_amd.workloads import desc_from  # noqa: E402

LEVEL_LENGTHS = (800.0, 150.0, 25.0)
DEFAULT_OUT = os.path.join(ROOT, "oceanlib_amd", "data", "height_proposal.npy")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def init_params(seed: int = 0) -> np.ndarray:
    """He-uniform weights on the used inputs/outputs; zero biases and padding."""
    rng = np.random.default_rng(seed)
    I, H, O = capi.PROPOSAL_INPUTS, 64, capi.PROPOSAL_OUTPUTS
    W1 = np.zeros((I, H))
    W1[:60] = rng.uniform(-1, 1, (60, H)) * np.sqrt(6.0 / 60)
    W2 = rng.uniform(-1, 1, (H, H)) * np.sqrt(6.0 / H)
    W3 = np.zeros((H, O))
    W3[:, :6] = rng.uniform(-1, 1, (H, 6)) * np.sqrt(6.0 / H)
    p = np.concatenate([W1.ravel(), np.zeros(H), W2.ravel(), np.zeros(H), W3.ravel(), np.zeros(O)])
    assert p.size == capi.PROPOSAL_PARAMS
    return p.astype(np.float32)


def sea_levels(s: int, seas: int):
    wind = 5.0 + 20.0 * s / max(1, seas - 1)
    return [dict(size=512, patch_length=L, wind_speed=float(np.float32(wind)), choppiness=1.3,
                 seed=s) for L in LEVEL_LENGTHS]


def make_pairs(lib, seas: int, per_sea: int, tol: float, seed: int):
    """Device features / targets for seas x per_sea points, and the ok mask."""
    rng = np.random.default_rng(seed)
    descs = []
    for s in range(seas):
        descs += [desc_from(kw) for kw in sea_levels(s, seas)]
    arr = (capi.OceanDesc * len(descs))(*descs)
    st = C.c_int(-1)
    b = lib.ocean_amd_batch_create(arr, len(descs), C.byref(st))
    if not b:
        raise SystemExit(f"batch create failed: {st.value}")
    times = rng.uniform(0.0, 600.0, seas)
    total = seas * per_sea
    feats = DeviceArray((total, capi.PROPOSAL_INPUTS), np.float32)
    targets = DeviceArray((total, capi.PROPOSAL_OUTPUTS), np.float32)
    ok = DeviceArray((total,), np.uint8)
    spacing = DeviceArray((3,), np.int32)
    spacings = []
    xz = DeviceArray((per_sea, 2), np.float32)
    for s in range(seas):
        # every sea at its own time: the batch steps all levels, only sea s is read
        lib.ocean_amd_batch_update(b, float(times[s]))
        pts = rng.uniform(0.0, 2000.0, (per_sea, 2)).astype(np.float32)
        xz.upload(pts)
        rc = lib.ocean_amd_proposal_pairs(b, 3 * s, xz.ptr, per_sea, tol,
                                          feats.offset(s * per_sea * capi.PROPOSAL_INPUTS * 4),
                                          targets.offset(s * per_sea * capi.PROPOSAL_OUTPUTS * 4),
                                          ok.offset(s * per_sea), spacing.ptr)
        if rc != 0:
            raise SystemExit(f"pairs failed rc={rc}")
        assert lib.ocean_amd_batch_synchronize(b) == 0
        spacings.append(spacing.numpy().tolist())
        if s % 16 == 0:
            log(f"pairs: sea {s}/{seas} spacing {spacings[-1]}")
    lib.ocean_amd_batch_destroy(b)
    xz.free()
    return feats, targets, ok.numpy().astype(bool), spacings, times


def train(lib, feats, targets, ok, epochs: int, batch: int, lr0: float, seed: int, params=None):
    rng = np.random.default_rng(seed + 1)
    rows = np.nonzero(ok)[0].astype(np.uint32)
    rng.shuffle(rows)
    n_val = max(batch, (len(rows) // 100) // batch * batch)
    val, tr = rows[:n_val], rows[n_val:]
    steps_per_epoch = len(tr) // batch
    st = C.c_int(-1)
    p0 = init_params(seed) if params is None else params
    model = lib.ocean_amd_proposal_create(p0.ctypes.data_as(capi._F), batch, C.byref(st))
    if not model:
        raise SystemExit(f"proposal create failed: {st.value}")
    idx = DeviceArray((steps_per_epoch * batch,), np.uint32)
    vidx = DeviceArray.from_numpy(val)
    grad = np.zeros(capi.PROPOSAL_PARAMS, np.float32)
    history = []

    def val_loss():
        tot, k = 0.0, 0
        for i in range(0, n_val, batch):
            loss = C.c_float()
            rc = lib.ocean_amd_proposal_gradient(model, feats.ptr, targets.ptr, vidx.offset(4 * i), batch,
                                                 grad.ctypes.data_as(capi._F), C.byref(loss))
            assert rc == 0
            tot += loss.value
            k += 1
        return tot / k

    t0 = time.time()
    for ep in range(epochs):
        lr = lr0 * (0.3 ** (ep * 3 // epochs))
        perm = tr.copy()
        rng.shuffle(perm)
        idx.upload(perm[:steps_per_epoch * batch])
        last = C.c_float()
        for k in range(steps_per_epoch):
            want = (k == steps_per_epoch - 1)
            rc = lib.ocean_amd_proposal_train_step(model, feats.ptr, targets.ptr, idx.offset(4 * k * batch),
                                                   batch, lr, C.byref(last) if want else None)
            if rc != 0:
                raise SystemExit(f"train step failed rc={rc}")
        vl = val_loss()
        history.append(dict(epoch=ep, lr=lr, train_loss=last.value, val_loss=vl,
                            seconds=round(time.time() - t0, 2)))
        log(f"epoch {ep}: lr {lr:.2e} train {last.value:.5f} val {vl:.5f} ({time.time() - t0:.1f} s)")
    assert lib.ocean_amd_proposal_synchronize(model) == 0
    out = np.zeros(capi.PROPOSAL_PARAMS, np.float32)
    assert lib.ocean_amd_proposal_get_params(model, out.ctypes.data_as(capi._F)) == 0
    lib.ocean_amd_proposal_destroy(model)
    return out, history, dict(train_rows=int(len(tr)), val_rows=int(n_val),
                              steps_per_epoch=int(steps_per_epoch))


def load_params(path: str = DEFAULT_OUT) -> np.ndarray:
    p = np.load(path, allow_pickle=False)
    if p.dtype != np.float32 or p.shape != (capi.PROPOSAL_PARAMS,):
        raise ValueError(f"{path}: expected {capi.PROPOSAL_PARAMS} float32 parameters")
    return p


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--seas", type=int, default=64)
    ap.add_argument("--per-sea", type=int, default=156250)
    ap.add_argument("--epochs", type=int, default=12)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default=DEFAULT_OUT)
    a = ap.parse_args()
    lib = capi.load()
    t0 = time.time()
    feats, targets, ok, spacings, times = make_pairs(lib, a.seas, a.per_sea, a.tol, a.seed)
    synchronize()
    t_pairs = time.time() - t0
    log(f"pairs: {ok.sum()} of {ok.size} converged, {t_pairs:.1f} s")
    params, history, info = train(lib, feats, targets, ok, a.epochs, a.batch, a.lr, a.seed)
    os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
    np.save(a.out, params, allow_pickle=False)
    meta = dict(seas=a.seas, per_sea=a.per_sea, pairs=int(ok.size), converged=int(ok.sum()),
                tol_m=a.tol, epochs=a.epochs, batch=a.batch, lr=a.lr, seed=a.seed,
                levels=[dict(patch_length=L, size=512, choppiness=1.3) for L in LEVEL_LENGTHS],
                winds="5 + 20 s / (seas - 1) m/s", level_seeds="s (seeds 0 .. seas - 1)",
                spacings=spacings, pair_seconds=round(t_pairs, 2), history=history, **info)
    with open(os.path.splitext(a.out)[0] + ".json", "w") as f:
        json.dump(meta, f, indent=1)
    print(json.dumps(dict(out=a.out, final=history[-1] if history else None)), flush=True)


if __name__ == "__main__":
    main()
