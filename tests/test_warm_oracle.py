"""CPU pins of the numpy restatement of the learned warm start's network."""
import os
import sys

import numpy as np
import pytest
import torch

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import warm_mlp as M  # noqa: E402


def _params(seed=0):
    rng = np.random.default_rng(seed)
    p = np.zeros(M.NPAR)
    p[M.OFF_W1:M.OFF_B1] = rng.uniform(-0.35, 0.35, M.H * M.F)
    p[M.OFF_B1:M.OFF_W2] = rng.uniform(-0.1, 0.1, M.H)
    p[M.OFF_W2:M.OFF_B2] = rng.uniform(-0.125, 0.125, M.H * M.H)
    p[M.OFF_B2:M.OFF_W3] = rng.uniform(-0.1, 0.1, M.H)
    p[M.OFF_W3:M.OFF_B3] = rng.uniform(-0.125, 0.125, M.H)
    p[M.OFF_B3] = 0.05
    return p


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    X = rng.normal(size=(37, 8))
    y = rng.normal(size=37)
    p = _params()
    g, _ = M.gradient(p, X, y, rounding=False)
    h = 1e-6
    idx = np.concatenate([rng.choice(np.arange(M.OFF_W1, M.OFF_B1), 20), rng.choice(np.arange(M.OFF_W2, M.OFF_B2), 40),
                          np.arange(M.OFF_B1, M.OFF_B1 + 5), np.arange(M.OFF_B2, M.OFF_B2 + 5),
                          np.arange(M.OFF_W3, M.OFF_W3 + 5), [M.OFF_B3]])
    for i in idx:
        a, b = p.copy(), p.copy()
        a[i] += h
        b[i] -= h
        fd = (M.loss(a, X, y) - M.loss(b, X, y)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])), (i, fd, g[i])


def test_bf16_matches_torch():
    rng = np.random.default_rng(2)
    x = np.concatenate([rng.normal(size=4000).astype(np.float32) * 10.0 ** rng.integers(-8, 8, 4000),
                        np.float32([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 0.0, 1e-30])]).astype(np.float32)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(M.bf16(x), ref)
    # ties go to even: 1 + 2^-8 lies halfway between 1 and 1 + 2^-7
    assert M.bf16(np.float32([1.0 + 2 ** -8]))[0] == np.float32(1.0)


def test_rounded_gradient_close_to_exact():
    rng = np.random.default_rng(3)
    X = rng.normal(size=(64, 8))
    y = rng.normal(size=64)
    p = _params(4)
    g0, _ = M.gradient(p, X, y, rounding=False)
    g1, _ = M.gradient(p, X, y, rounding=True)
    assert np.linalg.norm(g1 - g0) <= 0.05 * np.linalg.norm(g0)


def test_adam_first_step_is_signed_learning_rate():
    rng = np.random.default_rng(5)
    p = rng.normal(size=M.NPAR).astype(np.float32)
    g = rng.normal(size=M.NPAR).astype(np.float32)
    z = np.zeros(M.NPAR, np.float32)
    p1, m1, v1 = M.adam(p, g, z, z, 1, 1e-3)
    # m_hat = g and v_hat = g^2 after one step, so the move is lr * g / (|g| + eps)
    assert np.allclose(p - p1, 1e-3 * np.sign(g), rtol=1e-4, atol=1e-9)
    assert np.allclose(m1, 0.1 * g, rtol=1e-6) and np.allclose(v1, 0.001 * g * g, rtol=1e-4)  # 1 - 0.999 in float32


def test_output_gain_parabola():
    """The output-gain calibration (train_warm.gain_from_losses) returns the minimiser of the parabola
    through the losses at gains 0, 1/2, 1: exact for parabolas with a known minimum, clamped to [0, 1],
    0 for losses that do not curve upwards; the committed model's held-out losses give its gain 0.19."""
    import importlib
    tw = importlib.import_module("sph-cpp_amd.train_warm")
    for gstar, a, m in ((0.19, 30.0, 9.0), (0.7, 2.0, 1.0), (0.0, 5.0, 3.0)):
        L = lambda g: a * (g - gstar) ** 2 + m
        assert tw.gain_from_losses(L(0.0), L(0.5), L(1.0)) == pytest.approx(gstar, abs=1e-12)
    assert tw.gain_from_losses(1.96, 0.81, 0.16) == 1.0      # (g - 1.4)^2: clamped above
    assert tw.gain_from_losses(0.09, 0.64, 1.69) == 0.0      # (g + 0.3)^2: clamped below
    assert tw.gain_from_losses(1.0, 2.0, 1.0) == 0.0         # concave
    assert tw.gain_from_losses(10.17, 11.78, 27.87) == pytest.approx(0.194, abs=1e-3)
