"""Numpy restatement of the learned warm start's network.

TEST INFRASTRUCTURE ONLY.  This module is the checker for the
pbf_warm_model_* functions of the HIP library in sph-cpp_amd/ (the MFMA
forward, backward and Adam kernels of csrc/pbf_mlp.hip).  Only tests/ may
import it; the product path never calls it.

The network has no counterpart in the reference (whose pressure solve starts
from lambda = 0 every step, Simulation/Source/particles.h:318-358).  What is
restated here is the library's own specification:
  8 features -> 64 (SiLU) -> 64 (SiLU) -> 1,  parameters W1[64][8], b1[64],
  W2[64][64] (row = output unit), b2[64], w3[64], b3 (4801 floats);
  loss = mean squared error; Adam (beta1 0.9, beta2 0.999, eps 1e-8).
With ``rounding=True`` every operand the device feeds to its BF16 matrix
cores is rounded to bfloat16 (round to nearest even) at the same points:
the features and W1 (layer 1), H1 and W2 (layer 2), dZ2 and H1 (dW2), W2 and
dZ2 (dH1), dZ1 and the features (dW1).  Biases, the last layer and all
elementwise work stay in float32 (here float64), as on the device.

Pins (tests/test_warm_oracle.py): without rounding the gradient matches
central finite differences of the loss; bf16() matches torch's bfloat16
conversion; one Adam step from zero moments moves each parameter by
learning_rate * sign(gradient) (closed form).
"""
from __future__ import annotations

import numpy as np

F, H = 8, 64
OFF_W1 = 0
OFF_B1 = OFF_W1 + H * F
OFF_W2 = OFF_B1 + H
OFF_B2 = OFF_W2 + H * H
OFF_W3 = OFF_B2 + H
OFF_B3 = OFF_W3 + H
NPAR = OFF_B3 + 1


def bf16(a) -> np.ndarray:
    """float32 -> nearest bfloat16 (ties to even), returned as float32 values."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return ((u & 0xFFFF) << 16).astype(np.uint32).view(np.float32)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return (p[OFF_W1:OFF_B1].reshape(H, F), p[OFF_B1:OFF_W2], p[OFF_W2:OFF_B2].reshape(H, H), p[OFF_B2:OFF_W3],
            p[OFF_W3:OFF_B3], p[OFF_B3])


def pack(W1, b1, W2, b2, w3, b3) -> np.ndarray:
    return np.concatenate([np.ravel(W1), b1, np.ravel(W2), b2, w3, np.atleast_1d(b3)])


def _sig(z):
    return 1.0 / (1.0 + np.exp(-z))


def silu(z):
    return z * _sig(z)


def dsilu(z):
    s = _sig(z)
    return s * (1.0 + z * (1.0 - s))


def _r(a, rounding):
    """The value the device feeds its matrix cores: bf16 of the float32 value."""
    if not rounding:
        return np.asarray(a, dtype=np.float64)
    return bf16(np.asarray(a, dtype=np.float32)).astype(np.float64)


def forward(p, X, rounding=True):
    """Network output for feature rows X (n x 8) and the activations."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    Xr = _r(X, rounding)
    Z1 = Xr @ _r(W1, rounding).T + b1
    H1 = silu(Z1)
    Z2 = _r(H1, rounding) @ _r(W2, rounding).T + b2
    H2 = silu(Z2)
    out = H2 @ w3 + b3
    return out, (Xr, Z1, H1, Z2, H2)


def loss(p, X, y, rounding=False):
    out, _ = forward(p, X, rounding)
    return float(np.mean((out - np.asarray(y, np.float64)) ** 2))


def gradient(p, X, y, rounding=True):
    """Gradient of the mean squared error and the loss."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    out, (Xr, Z1, H1, Z2, H2) = forward(p, X, rounding)
    y = np.asarray(y, dtype=np.float64)
    diff = out - y
    dout = 2.0 * diff / len(y)
    dw3 = H2.T @ dout
    db3 = dout.sum()
    dZ2 = dout[:, None] * w3[None, :] * dsilu(Z2)
    db2 = dZ2.sum(0)
    dZ2r = _r(dZ2, rounding)
    H1r = _r(H1, rounding)
    dW2 = dZ2r.T @ H1r
    dH1 = dZ2r @ _r(W2, rounding)
    dZ1 = dH1 * dsilu(Z1)
    db1 = dZ1.sum(0)
    dW1 = _r(dZ1, rounding).T @ Xr
    return pack(dW1, db1, dW2, db2, dw3, db3), float(np.mean(diff ** 2))


def adam(p, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (Kingma & Ba 2015, algorithm 1) in float32; returns (p, m, v)."""
    f = np.float32
    p, g, m, v = (np.asarray(a, dtype=np.float32) for a in (p, g, m, v))
    m = f(b1) * m + (f(1) - f(b1)) * g
    v = f(b2) * v + (f(1) - f(b2)) * g * g
    c1 = f(1) - f(b1) ** f(t)
    c2 = f(1) - f(b2) ** f(t)
    p = p - f(lr) * (m / c1) / (np.sqrt(v / c2) + f(eps))
    return p, m, v
