"""CPU oracle of the learned-FVS vertex scorer (test infrastructure only; the product never imports it).

The scorer has no counterpart in the reference: it is the 4 -> 64 -> 64 -> 1 MLP (ReLU, logistic
output, mean binary cross-entropy) whose device forward and backward kernels live in
``minimumcyclebasis_amd/csrc/hip/mcb_mlp.hip``.  This module restates the same function in numpy:

* ``emulate_bf16=True`` rounds to BF16 exactly where the device does (W1, W2, W3; the first hidden
  layer where it enters the 64 x 64 products; dH2 where it enters them) and accumulates in float32,
  so the device must match it to accumulation-order error;
* ``emulate_bf16=False`` is the exact float64 function, whose gradient the tests pin against central
  finite differences.

Parameter layout (MCB_MLP_PARAMS = 4545 floats): W1 [4][64], b1 [64], W2 [64][64] (row = input unit),
b2 [64], W3 [64], b3.
"""
from __future__ import annotations

import numpy as np

N_IN, N_HID = 4, 64
N_PARAMS = N_IN * N_HID + N_HID + N_HID * N_HID + N_HID + N_HID + 1


def bf16(x: np.ndarray) -> np.ndarray:
    """float32 -> nearest BF16 (ties to even), returned as float32."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def unpack(params: np.ndarray, dtype=np.float32):
    p = np.asarray(params, dtype=dtype)
    if p.size != N_PARAMS:
        raise ValueError(f"expected {N_PARAMS} parameters")
    o = 0
    W1 = p[o:o + N_IN * N_HID].reshape(N_IN, N_HID)
    o += N_IN * N_HID
    b1 = p[o:o + N_HID]
    o += N_HID
    W2 = p[o:o + N_HID * N_HID].reshape(N_HID, N_HID)
    o += N_HID * N_HID
    b2 = p[o:o + N_HID]
    o += N_HID
    W3 = p[o:o + N_HID]
    o += N_HID
    return W1, b1, W2, b2, W3, p[o]


def _forward(params, X, emulate_bf16: bool):
    dt = np.float32 if emulate_bf16 else np.float64
    W1, b1, W2, b2, W3, b3 = unpack(params, dt)
    X = np.asarray(X, dtype=dt)
    if emulate_bf16:
        W1, W2, W3 = bf16(W1), bf16(W2), bf16(W3)
    H1 = np.maximum(X @ W1 + b1, 0)
    H1q = bf16(H1) if emulate_bf16 else H1
    H2 = np.maximum(H1q @ W2 + b2, 0)
    z = H2 @ W3 + b3
    return (W1, W2, W3), H1, H1q, H2, z


def forward(params, X, emulate_bf16: bool = True) -> np.ndarray:
    """Logits of the rows of X (n x 4)."""
    return _forward(params, X, emulate_bf16)[-1]


def loss_and_grad(params, X, y, emulate_bf16: bool = True):
    """(mean binary cross-entropy, its gradient as a flat parameter vector)."""
    (W1, W2, W3), H1, H1q, H2, z = _forward(params, X, emulate_bf16)
    dt = z.dtype
    X = np.asarray(X, dtype=dt)
    y = np.asarray(y, dtype=dt)
    n = z.shape[0]
    loss = np.mean(np.maximum(z, 0) - z * y + np.log1p(np.exp(-np.abs(z))))
    dz = (1 / (1 + np.exp(-z)) - y) / n
    gW3 = H2.T @ dz
    gb3 = dz.sum()
    dH2 = dz[:, None] * W3[None, :] * (H2 > 0)
    gb2 = dH2.sum(axis=0)
    dH2q = bf16(dH2) if emulate_bf16 else dH2
    gW2 = H1q.T @ dH2q
    dH1 = (dH2q @ W2.T) * (H1 > 0)
    gW1 = X.T @ dH1
    gb1 = dH1.sum(axis=0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3, [gb3]]).astype(dt)
    return float(loss), g
