"""Pins the guiding oracle (oracle/guide_oracle.py) by what the mathematics fixes.

- vMF lobes and the 8-lobe mixture integrate to 1 over the sphere (numerical
  quadrature, not the closed form being tested).
- The analytic float64 backward equals central finite differences of the loss,
  for every parameter block and the mix-weight logit; a dropped term, a wrong
  sign or a transposed weight fails this.
- The bf16-emulated gradient points the same way as the exact one.
- The frequency encoding has its closed-form values at p = 0 and p = 1/2, and
  its band recurrence tracks sin/cos(2^b π p).
"""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import guide_oracle as G  # noqa: E402

BLOCKS = {"W1": (G.W1, G.W2), "W2": (G.W2, G.W3), "W3": (G.W3, G.W4), "W4": (G.W4, G.B1), "b1": (G.B1, G.B2),
          "b2": (G.B2, G.B3), "b3": (G.B3, G.B4), "b4": (G.B4, G.NPARAM)}


@pytest.mark.parametrize("kappa", [1e-3, 0.05, 1.0, 30.0, float(np.exp(7.0))])
def test_vmf_integrates_to_one(kappa):
    # ∫ vmf dω = 2π ∫_{-1}^{1} vmf(κ, t) dt, integrated on a grid dense near t = 1
    s = np.linspace(0, 1, 400001)
    t = 1 - 2 * s ** 4  # concentrates nodes at t = 1 where the lobe lives
    f = G.vmf(np.full_like(t, kappa), t)
    integral = 2 * np.pi * np.trapezoid(f[::-1], t[::-1])
    assert abs(integral - 1) < 1e-6


def fibonacci_sphere(n):
    i = np.arange(n) + 0.5
    z = 1 - 2 * i / n
    r = np.sqrt(1 - z * z)
    phi = np.pi * (3 - np.sqrt(5)) * i
    return np.stack([r * np.cos(phi), r * np.sin(phi), z], axis=1)


def test_mixture_integrates_to_one():
    rng = np.random.default_rng(3)
    out = rng.normal(size=(4, 40))
    out[:, 3::5] = rng.uniform(-3, 3, size=(4, 8))  # kappa up to e^3
    dirs = fibonacci_sphere(400000)
    for o in out:
        g = G.mixture_pdf(np.repeat(o[None], len(dirs), 0), dirs)
        assert abs(g.mean() * 4 * np.pi - 1) < 2e-3


def test_features_closed_form():
    in9 = np.array([[0, 0.5, 0, 0.1, 0.2, 0.3, 0.4, 0.5, 0.6]], np.float32)
    f = G.features(in9)[0]
    assert np.allclose(f[0:12:2], 0) and np.allclose(f[1:12:2], 1)  # p_x = 0: sin 0, cos 1
    assert abs(f[12] - 1) < 1e-7 and abs(f[13]) < 1e-7            # p_y = 1/2, band 0: sin(π/2) = 1
    assert abs(f[14]) < 1e-6 and abs(f[15] + 1) < 1e-7             # band 1: sin π = 0, cos π = -1
    assert np.allclose(f[36:42], in9[0, 3:9]) and np.all(f[42:] == 0)


def fd_setup():
    p = G.init_params(7).astype(np.float64)
    rng = np.random.default_rng(8)
    p[G.W4:G.NPARAM] += rng.normal(0, 0.3, G.NPARAM - G.W4)  # sharper lobes, live log-κ and logits
    p[~G.live_mask()] = 0
    p[G.NPARAM] = 0.4
    rows = G.random_rows(24, seed=9).astype(np.float64)
    return p, rows


def test_gradient_matches_finite_differences():
    p, rows = fd_setup()
    g, _ = G.gradient(p, rows, emulate=False)
    rng = np.random.default_rng(10)
    live = G.live_mask()
    h = 1e-6
    for name, (lo, hi) in BLOCKS.items():
        idx = [i for i in rng.choice(np.arange(lo, hi), 40, replace=False) if live[i]][:12]
        for i in idx:
            pp, pm = p.copy(), p.copy()
            pp[i] += h
            pm[i] -= h
            fd = (G.loss_value(pp, rows) - G.loss_value(pm, rows)) / (2 * h)
            assert abs(fd - g[i]) <= 1e-5 * max(1.0, abs(fd)) + 1e-7, (name, i, fd, g[i])
    pp, pm = p.copy(), p.copy()
    pp[G.NPARAM] += h
    pm[G.NPARAM] -= h
    fd = (G.mix_loss_value(pp, rows) - G.mix_loss_value(pm, rows)) / (2 * h)
    assert abs(fd - g[G.NPARAM]) <= 1e-6 * max(1.0, abs(fd))


def test_padding_has_zero_gradient():
    p, rows = fd_setup()
    g, _ = G.gradient(p, rows, emulate=True)
    assert np.all(g[~G.live_mask()] == 0)


def test_emulated_gradient_close_to_exact():
    p = G.init_params(1)
    rows = G.random_rows(512, seed=2)
    ge, le = G.gradient(p, rows, emulate=True)
    gx, lx = G.gradient(p, rows, emulate=False)
    assert abs(le - lx) < 1e-4 * abs(lx)
    # Random rows give a gradient that is mostly cancellation between samples,
    # so the few ReLU gates that flip between the bf16 and the exact forward
    # (pre-activations within bf16 rounding of 0) dominate the difference.
    for name, (lo, hi) in BLOCKS.items():
        a, b = ge[lo:hi], gx[lo:hi]
        rel = np.linalg.norm(a - b) / np.linalg.norm(b)
        cos = a @ b / (np.linalg.norm(a) * np.linalg.norm(b))
        assert rel < 0.2 and cos > 0.98, (name, rel, cos)


def test_loss_ignores_rows_without_contribution():
    p = G.init_params(1)
    rows = G.random_rows(64, seed=4)
    rows[:, 12] = 0.0
    g, loss = G.gradient(p, rows)
    assert loss == 0 and np.all(g == 0)


def test_adam_first_step_is_signed_lr():
    p = G.init_params(3).astype(np.float64)
    grad = np.random.default_rng(5).normal(size=G.NPARAM + 1)
    q = G.adam_first_step(p, grad, 1e-3)
    live = G.live_mask()
    assert np.allclose((q - p)[live][:-1], -1e-3 * np.sign(grad[live][:-1]), atol=1e-6)


def test_band_recurrence_is_the_frequency_encoding():
    # the float32 doubling recurrence stays within 3e-5 of sin/cos(2^b π p)
    # (error grows about 3x per band), far below the bf16 step of the inputs
    p = np.random.default_rng(6).uniform(0, 1, (4096, 9)).astype(np.float32)
    f = G.features(p)
    for c in range(3):
        for b in range(6):
            x = (2.0 ** b) * np.pi * p[:, c].astype(np.float64)
            assert np.abs(f[:, 12 * c + 2 * b] - np.sin(x)).max() < 3e-5
            assert np.abs(f[:, 12 * c + 2 * b + 1] - np.cos(x)).max() < 3e-5
