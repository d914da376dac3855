"""CPU ORACLE (test infrastructure only) of the learned middle state.

Used by tests/ to check the library's network kernels; nothing in the
product imports it.  The network is the library's own addition (the
reference's augmented Riemann solver iterates its middle state inside the
un-vendored solver submodule, called from SWE_DimensionalSplittingMpi.cpp:
403-410, 432-439), so there is no reference algorithm to restate: this file
states the network, its features, loss and gradients in numpy float64, and
emulates the BF16 inference of the matrix-core path.  Its gradients are
pinned against central finite differences (tests/test_mlp_oracle.py).

Parameter vector (float32, 4738): W1[6][64], b1[64], W2[64][64], b2[64],
W3[64][2], b3[2]; W[in][out].  Sample: x[6], y[2].
"""
from __future__ import annotations

import numpy as np

KIN, KHID, KOUT = 6, 64, 2
OFF_W1 = 0
OFF_B1 = OFF_W1 + KIN * KHID
OFF_W2 = OFF_B1 + KHID
OFF_B2 = OFF_W2 + KHID * KHID
OFF_W3 = OFF_B2 + KHID
OFF_B3 = OFF_W3 + KHID * KOUT
NPARAMS = OFF_B3 + KOUT
G = np.float32(9.81)


def unpack(p):
    p = np.asarray(p, np.float64)
    return (p[OFF_W1:OFF_B1].reshape(KIN, KHID), p[OFF_B1:OFF_W2], p[OFF_W2:OFF_B2].reshape(KHID, KHID),
            p[OFF_B2:OFF_W3], p[OFF_W3:OFF_B3].reshape(KHID, KOUT), p[OFF_B3:NPARAMS])


def pack(W1, b1, W2, b2, W3, b3):
    return np.concatenate([np.ravel(W1), b1, np.ravel(W2), b2, np.ravel(W3), b3]).astype(np.float32)


def features(hL, hR, uL, uR, bL, bR):
    """Float32 arithmetic in the device order: hs = (hL + hR)/2, cs = sqrt(g hs)."""
    f = np.float32
    hL, hR, uL, uR, bL, bR = (np.asarray(a, f) for a in (hL, hR, uL, uR, bL, bR))
    hs = f(0.5) * (hL + hR)
    inv_hs = f(1.0) / hs
    cs = np.sqrt(G * hs)
    inv_cs = f(1.0) / cs
    x = np.stack([hL * inv_hs, hR * inv_hs, uL * inv_cs, uR * inv_cs,
                  np.minimum(np.maximum((bR - bL) * inv_hs, f(-1.0)), f(1.0)), (uR - uL) * inv_cs], axis=-1)
    return x.astype(f), hs, cs


def linear_part(x):
    """Linearised (acoustic) middle state in network units: the network
    predicts the corrections to it."""
    f = np.float32
    x = np.asarray(x, f)
    return np.stack([f(1.0) - f(0.5) * x[..., 5], f(0.5) * (x[..., 2] + x[..., 3]) - f(0.5) * (x[..., 1] - x[..., 0])],
                    axis=-1)


def forward(p, x):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    x = np.asarray(x, np.float64)
    a1 = x @ W1 + b1
    h1 = np.maximum(a1, 0.0)
    a2 = h1 @ W2 + b2
    h2 = np.maximum(a2, 0.0)
    y = h2 @ W3 + b3
    return y, (x, h1, h2)


def loss_and_grad(p, samples):
    """Mean over samples of (y0 - t0)^2 + (y1 - t1)^2 and its gradient (float64)."""
    s = np.asarray(samples, np.float64).reshape(-1, KIN + KOUT)
    x, t = s[:, :KIN], s[:, KIN:]
    W1, b1, W2, b2, W3, b3 = unpack(p)
    y, (x, h1, h2) = forward(p, x)
    n = s.shape[0]
    e = y - t
    loss = float(np.mean(np.sum(e * e, axis=1)))
    dy = 2.0 * e / n
    dW3 = h2.T @ dy
    db3 = dy.sum(0)
    dh2 = (dy @ W3.T) * (h2 > 0)
    dW2 = h1.T @ dh2
    db2 = dh2.sum(0)
    dh1 = (dh2 @ W2.T) * (h1 > 0)
    dW1 = x.T @ dh1
    db1 = dh1.sum(0)
    return loss, np.concatenate([dW1.ravel(), db1, dW2.ravel(), db2, dW3.ravel(), db3])


def bf16(a):
    """Round float32 values to bfloat16 (round to nearest even), as float32."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def infer_bf16(p, x):
    """Emulation of the matrix-core inference: inputs as a BF16 high part plus
    a BF16 residual, BF16 weights (b1 and b2 folded in as BF16 weights), ReLU
    activations rounded to BF16 between layers, b3 in FP32; sums in float64."""
    W1, b1, W2, b2, W3, b3 = unpack(p)
    x = np.asarray(x, np.float32)
    hi = bf16(x)
    lo = bf16(x - hi)
    xs = hi.astype(np.float64) + lo.astype(np.float64)
    W1b, b1b = bf16(W1).astype(np.float64), bf16(b1).astype(np.float64)
    W2b, b2b = bf16(W2).astype(np.float64), bf16(b2).astype(np.float64)
    W3b = bf16(W3).astype(np.float64)
    h1 = bf16(np.maximum(xs @ W1b + b1b, 0.0).astype(np.float32)).astype(np.float64)
    h2 = bf16(np.maximum(h1 @ W2b + b2b, 0.0).astype(np.float32)).astype(np.float64)
    return h2 @ W3b + np.asarray(b3, np.float64)
