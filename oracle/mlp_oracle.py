"""CPU oracle of the learned height proposal's network (test infrastructure
only: the library never imports it; only tests/ and bench.py's CPU leg may).

No reference counterpart: the reference has no learned part. This restates
the network of oceanlib_amd/csrc/proposal_kernels.hip in float64 numpy with
the kernel's rounding points:

  X = bf16(features); W = bf16(master weights); b = fp32 biases
  Z1 = X W1 + b1;  H1 = bf16(relu(Z1))
  Z2 = H1 W2 + b2; H2 = bf16(relu(Z2))
  Y  = H2 W3 + b3                       (fp32 outputs; 6 of 16 used)
  loss = sum over rows and the 6 outputs of (Y - T)^2 / (B * 6)
  dY  = bf16(2 (Y - T) / (B * 6)) on the 6 outputs, 0 elsewhere
  dZ2 = bf16((dY W3^T) * [H2 > 0]);  dZ1 = bf16((dZ2 W2^T) * [H1 > 0])
  dW1 = X^T dZ1, dW2 = H1^T dZ2, dW3 = H2^T dY; db = column sums of dZ / dY
  Adam (beta1 0.9, beta2 0.999, eps 1e-8) on the fp32 master parameters.

With `rounding=False` every bf16() is the identity, which is the plain
float64 network that the finite-difference check in tests/test_mlp_oracle.py
pins the backward formulas against.
"""
from __future__ import annotations

import numpy as np

LEVELS = 3
IN, HID, OUT = 64, 64, 16
USED_OUT = 2 * LEVELS
PARAMS = IN * HID + HID + HID * HID + HID + HID * OUT + OUT


def bf16(a) -> np.ndarray:
    """fp32 -> bf16 -> fp32, round to nearest even (as the kernels' v_cvt_pk_bf16_f32)."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    nan = (u & 0x7F800000) == 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    r = np.where(nan, u & 0xFFFF0000, r)
    return (r & 0xFFFFFFFF).astype(np.uint32).view(np.float32)


def unpack(p: np.ndarray):
    p = np.asarray(p, np.float32)
    o = 0
    W1 = p[o:o + IN * HID].reshape(IN, HID); o += IN * HID
    b1 = p[o:o + HID]; o += HID
    W2 = p[o:o + HID * HID].reshape(HID, HID); o += HID * HID
    b2 = p[o:o + HID]; o += HID
    W3 = p[o:o + HID * OUT].reshape(HID, OUT); o += HID * OUT
    b3 = p[o:o + OUT]; o += OUT
    assert o == PARAMS
    return W1, b1, W2, b2, W3, b3


def pack(W1, b1, W2, b2, W3, b3) -> np.ndarray:
    return np.concatenate([np.ravel(x) for x in (W1, b1, W2, b2, W3, b3)]).astype(np.float64)


def _r(a, rounding):
    return bf16(a).astype(np.float64) if rounding else np.asarray(a, np.float64)


def forward(params, feats, rounding=True):
    """(Y, cache): Y rows x 16."""
    W1, b1, W2, b2, W3, b3 = [np.asarray(x, np.float64) for x in unpack(params)] if rounding \
        else [np.asarray(x, np.float64) for x in _unpack64(params)]
    X = _r(feats, rounding)
    W1r, W2r, W3r = _r(W1, rounding), _r(W2, rounding), _r(W3, rounding)
    Z1 = X @ W1r + b1
    H1 = _r(np.maximum(Z1, 0.0), rounding)
    Z2 = H1 @ W2r + b2
    H2 = _r(np.maximum(Z2, 0.0), rounding)
    Y = H2 @ W3r + b3
    return Y, dict(X=X, H1=H1, H2=H2, W1=W1r, W2=W2r, W3=W3r)


def _unpack64(p):
    p = np.asarray(p, np.float64)
    o = 0
    out = []
    for shape in ((IN, HID), (HID,), (HID, HID), (HID,), (HID, OUT), (OUT,)):
        n = int(np.prod(shape))
        out.append(p[o:o + n].reshape(shape))
        o += n
    return out


def loss_and_grad(params, feats, targets, rounding=True):
    """(loss, grad[PARAMS]) for the minibatch (all rows)."""
    Y, c = forward(params, feats, rounding)
    B = Y.shape[0]
    T = np.asarray(targets, np.float64)
    mask = np.zeros(OUT)
    mask[:USED_OUT] = 1.0
    d = (Y - T) * mask
    loss = float(np.sum(d * d) / (B * USED_OUT))
    dY = _r(2.0 * d / (B * USED_OUT), rounding)
    dZ2 = _r((dY @ c["W3"].T) * (c["H2"] > 0), rounding)
    dZ1 = _r((dZ2 @ c["W2"].T) * (c["H1"] > 0), rounding)
    g = pack(c["X"].T @ dZ1, dZ1.sum(0), c["H1"].T @ dZ2, dZ2.sum(0), c["H2"].T @ dY, dY.sum(0))
    return loss, g


def adam_step(master, m1, m2, grad, step, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam step in float32, operation for operation as the kernel
    (beta constants and 1 - beta in float32; c1, c2 from double)."""
    f = np.float32
    b1, b2, one = f(beta1), f(beta2), f(1.0)
    master = np.asarray(master, f).copy()
    g = np.asarray(grad, f)
    a = b1 * np.asarray(m1, f) + (one - b1) * g
    v = b2 * np.asarray(m2, f) + (one - b2) * g * g
    c1 = f(1.0 - 0.9 ** step)
    c2 = f(1.0 - 0.999 ** step)
    master = master - f(lr) * (a / c1) / (np.sqrt(v / c2) + f(eps))
    return master, a, v


def init_params(seed: int = 0) -> np.ndarray:
    """He-uniform weights for the used inputs/outputs, zero biases and padding."""
    rng = np.random.default_rng(seed)
    W1 = np.zeros((IN, HID)); W1[:60] = rng.uniform(-1, 1, (60, HID)) * np.sqrt(6.0 / 60)
    W2 = rng.uniform(-1, 1, (HID, HID)) * np.sqrt(6.0 / HID)
    W3 = np.zeros((HID, OUT)); W3[:, :USED_OUT] = rng.uniform(-1, 1, (HID, USED_OUT)) * np.sqrt(6.0 / HID)
    return pack(W1, np.zeros(HID), W2, np.zeros(HID), W3, np.zeros(OUT)).astype(np.float32)
