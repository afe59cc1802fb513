"""The refinement-indicator CNN in NumPy (oracle only): forward, loss, backward.

Three 3x3x3 convolutions over a periodic 3D grid, 8 -> 16 -> 16 -> 1 channels, ReLU
between them, a logit out; the operands are rounded to bfloat16 (round to nearest
even) and the products accumulated in float32, as the MFMA kernels do.  Layouts:
activations ``(nx, ny, nz, C)``; weights ``W[t, ci, co]`` with the tap
``t = 9 (dx+1) + 3 (dy+1) + (dz+1)`` reading the input at ``v + (dx, dy, dz)``
(a correlation, periodic wrap).  The loss is the mean binary cross-entropy of the
logits against the labels.  The GPU's summation order differs from NumPy's, so
parity is to a relative tolerance, stated in the tests.
"""

from __future__ import annotations

import numpy as np

TAPS = [(dx, dy, dz) for dx in (-1, 0, 1) for dy in (-1, 0, 1) for dz in (-1, 0, 1)]


def bf16(x: np.ndarray) -> np.ndarray:
    """Round float32 to bfloat16 (nearest, ties to even), returned as float32."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    out = (r & 0xFFFFFFFF).astype(np.uint32).view(np.float32)
    return np.where(np.isnan(a), a, out)


def shifted(x: np.ndarray, tap) -> np.ndarray:
    """x[v + tap] with periodic wrap (x is (nx, ny, nz, C))."""
    dx, dy, dz = tap
    return np.roll(x, shift=(-dx, -dy, -dz), axis=(0, 1, 2))


def conv(x: np.ndarray, w: np.ndarray, b: np.ndarray) -> np.ndarray:
    """sum_t x[v + t] @ w[t] + b in float32 (x, w already rounded)."""
    out = np.zeros(x.shape[:3] + (w.shape[2],), x.dtype)
    for t, tap in enumerate(TAPS):
        out += shifted(x, tap) @ w[t]
    return out + b


def conv_transpose(g: np.ndarray, w: np.ndarray) -> np.ndarray:
    """The input gradient of `conv`: sum_t g[v - t] @ w[t]^T."""
    out = np.zeros(g.shape[:3] + (w.shape[1],), g.dtype)
    for t, tap in enumerate(TAPS):
        out += shifted(g, tuple(-d for d in tap)) @ w[t].T
    return out


def weight_grad(x: np.ndarray, g: np.ndarray) -> np.ndarray:
    """dW[t] = sum_v x[v + t]^T g[v]."""
    cin, cout = x.shape[3], g.shape[3]
    dw = np.zeros((27, cin, cout), x.dtype)
    gf = g.reshape(-1, cout)
    for t, tap in enumerate(TAPS):
        dw[t] = shifted(x, tap).reshape(-1, cin).T @ gf
    return dw


def init(seed: int = 0):
    """He-initialised weights and zero biases (float32 masters)."""
    rng = np.random.default_rng(seed)
    shapes = [(8, 16), (16, 16), (16, 1)]
    params = []
    for cin, cout in shapes:
        w = rng.standard_normal((27, cin, cout)).astype(np.float32) * np.float32(np.sqrt(2.0 / (27 * cin)))
        params += [w, np.zeros(cout, np.float32)]
    return params


def same(x):
    """No rounding (the float64 gradient check)."""
    return x


def forward(x, params, rnd=bf16):
    """Activations of every layer (rounded as stored) and the logits.  ``rnd`` is the
    operand rounding: `bf16` as the kernels, or the identity (float64 checks)."""
    w1, b1, w2, b2, w3, b3 = params
    x = rnd(x)
    a1 = rnd(np.maximum(conv(x, rnd(w1), b1), 0.0))
    a2 = rnd(np.maximum(conv(a1, rnd(w2), b2), 0.0))
    logit = conv(a2, rnd(w3), b3)[..., 0]
    return x, a1, a2, logit


def bce(logit, y) -> float:
    z = np.asarray(logit, np.float64)
    return float(np.mean(np.maximum(z, 0) - z * y + np.log1p(np.exp(-np.abs(z)))))


def loss_and_grads(x, y, params, rnd=bf16):
    """Mean BCE-with-logits and the gradients of every parameter.  Every gradient
    is taken from the rounded (stored) gradients of the layer above, biases too."""
    w1, b1, w2, b2, w3, b3 = params
    x, a1, a2, logit = forward(x, params, rnd)
    n = logit.size
    loss = bce(logit, y)
    g3 = ((1.0 / (1.0 + np.exp(-logit))) - y).astype(logit.dtype) / logit.dtype.type(n)
    g3 = g3[..., None]
    g3b = rnd(g3)
    dw3 = weight_grad(a2, g3b)
    db3 = g3b.reshape(-1, 1).sum(axis=0)
    g2 = rnd(conv_transpose(g3b, rnd(w3)) * (a2 > 0))
    dw2 = weight_grad(a1, g2)
    db2 = g2.reshape(-1, 16).sum(axis=0)
    g1 = rnd(conv_transpose(g2, rnd(w2)) * (a1 > 0))
    dw1 = weight_grad(x, g1)
    db1 = g1.reshape(-1, 16).sum(axis=0)
    return loss, [dw1, db1, dw2, db2, dw3, db3], logit
