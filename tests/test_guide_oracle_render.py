"""The oracle's guided estimator (oracle/oracle_guide.cc and the guided
vertex of oracle/oracle_render.cc's color()), on the CPU.  It restates the
device's guided shading, which has no reference counterpart; these tests pin
it: its network forward pass equals the numpy oracle's (itself pinned by
finite differences), its mixture sampler draws from the density it reports,
and the guided estimator keeps the plain estimator's expectation, whose
color() the reference's own pins (tests/test_oracle_vs_reference.py)."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import guide_oracle as G  # noqa: E402


def params(seed, mix_logit=0.5):
    p = G.init_params(seed)
    p[-1] = mix_logit
    return p


def test_forward_equals_numpy_oracle(oracle):
    """Same bf16 operands, float64 sums: equal to a few float ulps."""
    p = params(3)
    rows = G.random_rows(1 << 13, 1)
    a = oracle.guide_forward(p, rows[:, :9])
    b, _ = G.forward(p, rows[:, :9], emulate=True)
    assert np.all(np.abs(a - b) <= 1e-6 * np.abs(b) + 1e-7)


def fibonacci_sphere(n):
    k = np.arange(n) + 0.5
    z = 1 - 2 * k / n
    phi = np.pi * (1 + 5 ** 0.5) * k
    r = np.sqrt(1 - z * z)
    return np.stack([r * np.cos(phi), r * np.sin(phi), z], 1).astype(np.float32)


def test_mixture_sampler_draws_from_its_density(oracle):
    """For inputs of several random networks: the density integrates to 1 over
    the sphere (quadrature on 2^17 points), sampled directions are unit
    vectors, and E_g[1/g] = 4 pi (importance sampling of the constant 1): a
    sampler that does not draw from the density it reports fails it."""
    rng = np.random.default_rng(5)
    grid = fibonacci_sphere(1 << 17)
    for seed in range(4):
        p = params(seed)
        p[G.B4:G.NPARAM] += rng.normal(0, 1.0, G.NPARAM - G.B4).astype(np.float32) * G.live_mask()[G.B4:G.NPARAM]
        in9 = G.random_rows(1, 10 + seed)[0, :9]
        total = oracle.guide_pdf(p, in9, grid).astype(np.float64).mean() * 4 * np.pi
        assert abs(total - 1) < 2e-3, total
        u = rng.uniform(0, 1, (1 << 16, 3)).astype(np.float32)
        d, g = oracle.guide_sample(p, in9, u)
        assert np.abs(np.linalg.norm(d, axis=1) - 1).max() < 1e-5
        est = 1 / g.astype(np.float64)
        assert abs(est.mean() - 4 * np.pi) < 5 * est.std() / np.sqrt(len(est)), (est.mean() / (4 * np.pi))


def image_stats(sums):
    s, q, spp = sums
    m = s / spp
    v = np.maximum(q / spp - m * m, 0) / spp
    return m, v


def test_guided_estimator_keeps_the_plain_expectation(oracle):
    """Guided (one-sample MIS of the proposal with the reference sampler) and
    plain renders of the Cornell box on independent streams.  A random network
    at mix weight 0.62: per-pixel means within 3 sigma and image means within
    4 standard errors.  A sharp, misplaced proposal (kappa = e^6 in every lobe,
    mix weight 0.95, three guided vertices) makes the weights heavy-tailed (up
    to 20x the reference's per vertex), so per-pixel variance estimates run
    low; for it the image means are compared, at 4096 samples per pixel."""
    s = oracle.scene("cornell555")
    w, spp = 24, 256
    m_p, v_p = image_stats((*s.render(w, w, spp, seed=101), spp))
    m_g, v_g = image_stats((*s.render_guided(w, w, spp, params(3), seed=202), spp))
    z = (m_g - m_p) / np.sqrt(v_g + v_p + 1e-12)
    assert np.mean(np.abs(z) > 3) < 0.015
    se = np.sqrt((v_g + v_p).sum()) / v_p.size
    assert abs(m_g.mean() - m_p.mean()) <= 4 * se, (m_g.mean(), m_p.mean(), se)

    sharp = params(7, mix_logit=3.0)
    sharp[G.B4:G.NPARAM][3::5][:8] = 6.0  # log kappa of every lobe
    w, spp = 8, 4096
    m_p, v_p = image_stats((*s.render(w, w, spp, seed=103), spp))
    m_g, v_g = image_stats((*s.render_guided(w, w, spp, sharp, guide_verts=3, seed=204), spp))
    se = np.sqrt((v_g + v_p).sum()) / v_p.size
    assert abs(m_g.mean() - m_p.mean()) <= 4 * se, (m_g.mean(), m_p.mean(), se)
