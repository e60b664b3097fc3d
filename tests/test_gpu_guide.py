"""Guiding network kernels against the oracle (oracle/guide_oracle.py), through the C ABI.

- tpt_guide_infer: raw outputs equal the oracle's bf16-emulated forward to FP32
  accumulation rounding, for ragged n and for n beyond one grid of blocks.
- tpt_guide_gradient: every parameter block and the mix-weight logit equal the
  oracle's backward at the same bf16 rounding points.
- tpt_guide_train: the first Adam step is the textbook update of that gradient,
  and training on learnable data drives the loss toward the data's entropy.
- tpt_render_guided: on the same Philox streams it equals the oracle's guided
  color() (oracle/oracle_guide.cc) to rounding; the MIS estimator keeps the
  plain estimator's expectation (untrained and trained networks, and at the
  full 1200x1200 Cornell box); the recorded training rows are sound.
"""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import guide_oracle as G  # noqa: E402

pytestmark = pytest.mark.gpu

BLOCKS = {"W1": (G.W1, G.W2), "W2": (G.W2, G.W3), "W3": (G.W3, G.W4), "W4": (G.W4, G.B1), "b1": (G.B1, G.B2),
          "b2": (G.B2, G.B3), "b3": (G.B3, G.B4), "b4": (G.B4, G.NPARAM)}


def make_params(seed=5):
    p = G.init_params(seed).astype(np.float64)
    rng = np.random.default_rng(seed + 100)
    p[G.W4:G.NPARAM] += rng.normal(0, 0.3, G.NPARAM - G.W4)
    p[~G.live_mask()] = 0
    p[G.NPARAM] = 0.3
    return p.astype(np.float32)


def to_dev(a, device):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).to(device)


@pytest.fixture(scope="module")
def guide(tpt, device):
    return tpt.Guide(seed=1)


def midpoint_distance(v):
    """|v - m| for the midpoint m between the two bf16 values around each
    float32 value v (where rounding to bf16 changes direction)."""
    v32 = np.ascontiguousarray(v, np.float32)
    mid = ((v32.view(np.uint32) & np.uint32(0xFFFF0000)) | np.uint32(0x8000)).view(np.float32)
    return np.abs(v32.astype(np.float64) - mid.astype(np.float64))


def edge_rows(p, in9, act):
    """Rows with a bf16 operand closer to a rounding midpoint than the device's
    float32 result can be from the oracle's: a frequency feature of band b
    (sinpi/cospi, then b float angle doublings: within 2^b * 5e-7), a hidden
    activation (a float32 sum of K products: within K * 2^-24 times the sum of
    |terms|).  Only such a row may see an operand rounded one bf16 ulp the
    other way."""
    f = G.features(in9)
    band = np.zeros(48)
    band[:36] = (np.arange(36) % 12) // 2
    margin_f = np.where(np.arange(48) < 36, 2.0 ** band * 5e-7, 0.0)
    edge = (midpoint_distance(f) <= margin_f).any(1)
    w1, w2, w3, w4, b1, b2, b3, b4, _ = G.unpack(p)
    xin = G.bf16(f.astype(np.float32))
    for w, b, h, k in ((w1, b1, act["h1"], 48), (w2, b2, act["h2"], 64), (w3, b3, act["h3"], 64)):
        scale = np.abs(xin) @ np.abs(G.bf16(w)).T + np.abs(b)
        edge |= ((midpoint_distance(h) <= k * 2.0 ** -24 * scale) & (h > 0)).any(1)
        xin = G.bf16(h)
    return edge


@pytest.mark.parametrize("n", [1, 65, 1000, 200000])
def test_infer_matches_oracle(guide, device, n):
    p = make_params()
    guide.set_params(p)
    assert np.array_equal(guide.get_params(), p)
    rows = G.random_rows(n, seed=11 + n)
    out = guide.infer(to_dev(rows[:, :9], device)).cpu().numpy()
    ref, act = G.forward(p, rows[:, :9])
    err = np.abs(out - ref) / (1 + np.abs(ref))
    # Rows agree to FP32 accumulation rounding, except where an operand sits
    # at a bf16 rounding midpoint and the device's float32 value rounds it one
    # bf16 ulp the other way: every row that differs more must have such an
    # operand (a difference with no midpoint to explain it is a bug).
    off = (err > 1e-5).any(axis=1)
    unexplained = off & ~edge_rows(p, rows[:, :9], act)
    print(f"n={n}: rows off by more than 1e-5: {off.mean():.4f} (max error {err.max():.3g}), "
          f"without an operand at a bf16 midpoint: {unexplained.sum()}")
    assert not unexplained.any()
    assert off.mean() <= 0.01  # measured 0.0022-0.0040
    assert err.max() < 0.05


def test_infer_empty(guide, tpt):
    import ctypes as C

    assert tpt.lib().tpt_guide_infer(guide.h, None, 0, None, None) == 0
    assert tpt.lib().tpt_guide_gradient(guide.h, None, 0, None, C.byref(C.c_float()), None) != 0


@pytest.mark.parametrize("n", [64, 1000, 4096, 50000])
def test_gradient_matches_oracle(guide, device, n):
    p = make_params()
    guide.set_params(p)
    rows = G.random_rows(n, seed=23 + n)
    g_dev, loss = guide.gradient(to_dev(rows, device))
    g_dev = g_dev.cpu().numpy().astype(np.float64)
    g_ref, loss_ref = G.gradient(p, rows, emulate=True)
    assert abs(loss - loss_ref / n) <= 1e-4 * abs(loss_ref / n)
    for name, (lo, hi) in BLOCKS.items():
        a, b = g_dev[lo:hi], g_ref[lo:hi]
        rel = np.linalg.norm(a - b) / np.linalg.norm(b)
        print(f"n={n} {name}: relative error {rel:.2e}")
        assert rel < 1e-3, (name, rel)  # measured at most 5.0e-4 (W1, n = 50000)
    assert np.all(g_dev[~G.live_mask()] == 0)
    assert abs(g_dev[G.NPARAM] - g_ref[G.NPARAM]) <= 1e-3 * max(1e-3, abs(g_ref[G.NPARAM]))


def test_gradient_ignores_rows_without_contribution(guide, device):
    p = make_params()
    guide.set_params(p)
    rows = G.random_rows(3000, seed=31)
    keep = rows[:, 12] > 0
    rows[~keep, 12] = -1.0  # ignored rows, also negative or NaN contributions
    rows[np.flatnonzero(~keep)[::2], 12] = np.nan
    g_all, _ = guide.gradient(to_dev(rows, device))
    g_ref, _ = G.gradient(p, rows)
    rel = np.linalg.norm(g_all.cpu().numpy() - g_ref) / np.linalg.norm(g_ref)
    assert rel < 1e-2


def test_adam_first_step(guide, device):
    p = make_params(9)
    guide.set_params(p)
    rows = G.random_rows(4096, seed=41)
    rt = to_dev(rows, device)
    g_dev, _ = guide.gradient(rt)
    g_dev = g_dev.cpu().numpy().astype(np.float64)
    guide.train(rt, n_batch=4096, iters=1, lr=1e-3)
    q = guide.get_params().astype(np.float64)
    expect = G.adam_first_step(p, g_dev, 1e-3)
    err = np.abs(q - expect)
    print(f"max |param - Adam step| = {err.max():.3g}")
    assert err.max() < 2e-6


def vmf_samples(mu, kappa, rng):
    n = mu.shape[0]
    u = rng.uniform(size=n)
    w = 1 + np.log(u + (1 - u) * np.exp(-2 * kappa)) / kappa
    phi = rng.uniform(0, 2 * np.pi, n)
    helper = np.where(np.abs(mu[:, :1]) > 0.9, [[0, 1, 0]], [[1, 0, 0]])
    t1 = np.cross(mu, helper)
    t1 /= np.linalg.norm(t1, axis=1, keepdims=True)
    t2 = np.cross(mu, t1)
    s = np.sqrt(np.maximum(0, 1 - w * w))[:, None]
    return s * np.cos(phi)[:, None] * t1 + s * np.sin(phi)[:, None] * t2 + w[:, None] * mu


def learnable_rows(n, kappa, seed):
    """Directions drawn from one vMF lobe whose mean depends on the position."""
    rng = np.random.default_rng(seed)
    rows = G.random_rows(n, seed)
    d = rows[:, 0:3].astype(np.float64) - 0.5 + np.array([0.0, 0.0, 0.05])
    mu = d / np.linalg.norm(d, axis=1, keepdims=True)
    rows[:, 9:12] = vmf_samples(mu, kappa, rng)
    rows[:, 12] = 1.0
    rows[:, 13] = 1.0 / (4 * np.pi)
    return rows


def test_training_fits_a_position_dependent_lobe(tpt, device):
    kappa = 20.0
    rows = learnable_rows(1 << 16, kappa, seed=51)
    rt = to_dev(rows, device)
    g = tpt.Guide(seed=3)
    _, l0 = g.gradient(rt)
    losses = [g.train(rt, n_batch=1 << 14, iters=100, lr=5e-3) for _ in range(6)]
    _, l1 = g.gradient(rt)
    # entropy of a vMF lobe: -log C + 1 - 2κ e^{-2κ} / (1 - e^{-2κ}), C its normaliser
    e = np.exp(-2 * kappa)
    best = -np.log(kappa / (2 * np.pi * (1 - e))) + 1 - 2 * kappa * e / (1 - e)
    print(f"loss {l0:.3f} -> {l1:.3f} (batches {['%.3f' % x for x in losses]}); entropy of the data {best:.3f}")
    assert l1 < l0 - 1.5
    assert l1 < best + 0.6
    # the learned proposal beats the uniform one where it is sampled, so the mix weight rises
    assert g.get_params()[G.NPARAM] > 0.5


def cornell_pair(tpt, w, spp, seed, guide=None, guiding=0):
    sc = tpt.Scene.named("cornell555")
    s, q, st = sc.render(width=w, height=w, spp=spp, seed=seed, camera=sc.host.camera, guide=guide, guiding=guiding)
    return s.cpu().numpy(), q.cpu().numpy(), st, sc


def zfrac(s1, q1, s2, q2, spp):
    m1, m2 = s1 / spp, s2 / spp
    v1 = np.maximum(q1 / spp - m1 * m1, 0) / spp
    v2 = np.maximum(q2 / spp - m2 * m2, 0) / spp
    z = (m1 - m2) / np.sqrt(v1 + v2 + 1e-12)
    return np.mean(np.abs(z) > 3), z


def test_recorded_rows_are_sound(tpt, device):
    g = tpt.Guide(seed=4)
    w, spp, stride = 32, 64, 4
    _, _, st, _ = cornell_pair(tpt, w, spp, 5, guide=g, guiding=stride)
    rows = g.batch_rows().cpu().numpy()
    assert rows.shape == ((w * w * spp + stride - 1) // stride * 2, 16)
    used = rows[:, 12] > 0
    print(f"{rows.shape[0]} rows, {used.mean():.3f} with a contribution; paths {st['paths']}")
    assert used.mean() > 0.05
    r = rows[used]
    assert np.all(np.isfinite(r))
    assert np.all((r[:, 0:3] > -1e-3) & (r[:, 0:3] < 1 + 1e-3))
    for c in (3, 6, 9):
        assert np.allclose(np.linalg.norm(r[:, c:c + 3], axis=1), 1, atol=1e-4)
    assert np.all(r[:, 13] > 0)


@pytest.mark.parametrize("trained", [False, True])
def test_guided_render_keeps_the_expectation(tpt, device, trained):
    g = tpt.Guide(seed=6)
    w, spp = 32, 4096
    if trained:
        for k in range(4):
            cornell_pair(tpt, w, 64, 100 + k, guide=g, guiding=1)
            g.train(None, n_batch=1 << 14, iters=50, lr=5e-3)
    alpha = 1 / (1 + np.exp(-g.get_params()[G.NPARAM]))
    s_g, q_g, st, _ = cornell_pair(tpt, w, spp, 7, guide=g)
    s_p, q_p, _, _ = cornell_pair(tpt, w, spp, 8)
    assert st["paths"] == w * w * spp
    frac, _ = zfrac(s_g, q_g, s_p, q_p, spp)
    n = w * w * 3 * spp
    se = np.sqrt(q_g.mean() / spp / n + q_p.mean() / spp / n)
    dm = (s_g.mean() - s_p.mean()) / spp
    print(f"trained={trained} alpha={alpha:.3f}: |z|>3 in {frac:.4f}; image mean diff {dm:.5f} ({dm / se:.2f} SE)")
    assert frac < 0.015
    assert abs(dm) <= 4 * se


def test_trained_guide_lowers_the_error(tpt, device):
    """What the trained network gains: on a 96x96 Cornell box, after online
    training on its own rows, the guided estimator's squared error at equal
    spp (display values against a 16384-spp plain render) is clearly below the
    plain estimator's."""
    import torch

    sc = tpt.Scene.named("cornell555")
    w, spp = 96, 128
    cam = sc.host.camera
    ref_s, _, _ = sc.render(width=w, height=w, spp=16384, seed=900, camera=cam)
    ref = torch.clamp(torch.sqrt(ref_s / 16384), 0, 1)
    g = tpt.Guide(seed=12)
    for k in range(6):
        sc.render(width=w, height=w, spp=64, seed=500 + k, camera=cam, guide=g, guiding=1)
        g.train(None, n_batch=1 << 16, iters=64, lr=5e-3)

    def sq_err(guide, seed):
        s, _, _ = sc.render(width=w, height=w, spp=spp, seed=seed, camera=cam, guide=guide)
        return torch.mean((torch.clamp(torch.sqrt(s / spp), 0, 1) - ref) ** 2).item()

    e_plain = np.mean([sq_err(None, 40 + k) for k in range(3)])
    e_guided = np.mean([sq_err(g, 60 + k) for k in range(3)])
    print(f"squared display error at {spp} spp: plain {e_plain:.3g}, guided {e_guided:.3g} "
          f"({e_plain / e_guided:.2f}x lower)")
    assert e_guided * 1.3 < e_plain


GUIDED_SAME_STREAM = [("cornell555", 40, 16), ("random_spheres", 40, 8), ("ref_cornell_box", 32, 8)]


@pytest.mark.parametrize("name,w,spp", GUIDED_SAME_STREAM)
@pytest.mark.parametrize("verts", [1, 2])
def test_guided_same_streams_match_oracle(tpt, oracle, device, name, w, spp, verts):
    """tpt_render_guided against the oracle's guided color() on the same
    Philox streams and the same parameters: pixels equal to rounding, except
    where an fp32 last bit (the device's exp/log/sincos in the mixture, a
    network output on a bf16 rounding midpoint) moves a sampled direction.
    Scenes: a rect light (cornell555), no light list (the cosine-only q), and
    the reference driver's rect + sphere light list with metal (ref_cornell_box)."""
    p = make_params()
    g = tpt.Guide(seed=1)
    g.set_params(p)
    sc = tpt.Scene.named(name)
    s_g, q_g, st = sc.render(width=w, height=w, spp=spp, seed=17, camera=sc.host.camera, guide=g,
                             guide_vertices=verts)
    s_g, q_g = s_g.cpu().numpy(), q_g.cpu().numpy()
    s_c, q_c = oracle.scene(name).render_guided(w, w, spp, p, guide_verts=verts, seed=17)
    assert st["paths"] == w * w * spp
    close = (np.abs(s_g - s_c) <= 1e-3 * (np.abs(s_c) + 1e-2)).all(axis=2)
    frac = close.mean()
    z = (s_g - s_c) / spp / np.sqrt((np.maximum(q_g / spp - (s_g / spp) ** 2, 0) +
                                     np.maximum(q_c / spp - (s_c / spp) ** 2, 0)) / spp + 1e-12)
    print(f"{name} verts={verts}: {frac:.4f} of pixels equal to rounding; |z|>4 in {np.mean(np.abs(z) > 4):.4f}; "
          f"means {s_g.mean():.5f} {s_c.mean():.5f}")
    assert frac >= 0.997  # measured 0.9988-1.0 (at most 2 pixels of 1600)
    assert np.mean(np.abs(z) > 4) < 0.005


def test_guided_keeps_the_expectation_at_full_size(tpt, device):
    """The benchmark frame, Cornell 1200x1200: a guide trained by the online
    schedule (3 passes of 4 spp recording every 8th path, 64 Adam steps each),
    then 64 guided and 64 plain samples per pixel on independent streams:
    image means within 4 standard errors, |z| > 3 in under 1.5% of the pixel
    channels."""
    sc = tpt.Scene.named("cornell555")
    w, spp = 1200, 64
    cam = sc.host.camera
    g = tpt.Guide(seed=2)
    r = tpt.ProgressiveRender(sc, guide=g, seed=71, record_stride=8, iters=64, batch=1 << 18, lr=5e-3)
    for _ in range(3):
        r.add(4, train=True)
    s_g, q_g, st = sc.render(width=w, height=w, spp=spp, seed=72, camera=cam, guide=g)
    s_p, q_p, _ = sc.render(width=w, height=w, spp=spp, seed=73, camera=cam)
    s_g, q_g, s_p, q_p = (x.cpu().numpy().astype(np.float64) for x in (s_g, q_g, s_p, q_p))
    assert st["paths"] == w * w * spp
    frac, _ = zfrac(s_g, q_g, s_p, q_p, spp)
    n = w * w * 3 * spp
    se = np.sqrt(q_g.mean() / spp / n + q_p.mean() / spp / n)
    dm = (s_g.mean() - s_p.mean()) / spp
    alpha = 1 / (1 + np.exp(-g.get_params()[G.NPARAM]))
    print(f"1200^2: alpha={alpha:.3f}; |z|>3 in {frac:.4f}; image mean diff {dm:.6f} ({dm / se:.2f} SE)")
    assert abs(dm) <= 4 * se
    assert frac < 0.015



def test_overlapped_training_keeps_the_expectation(tpt, device):
    """ProgressiveRender(overlap=True): the Adam steps run on a second stream
    while the next passes render with whatever weights the guide holds when
    each block stages them.  The rows handed to training are the recorded
    ones, the guide trains (the parameters move), and the image keeps the
    plain estimator's expectation: 160x160 Cornell, 3 training passes of
    4 spp, then 500 more guided spp, against 512 plain spp on other streams."""
    import torch

    sc = tpt.Scene.named("cornell555")
    w = 160
    cam = sc.host.camera
    g = tpt.Guide(seed=3)
    p0 = g.get_params().copy()
    r = tpt.ProgressiveRender(sc, guide=g, seed=81, record_stride=2, iters=64, batch=1 << 16, lr=5e-3,
                              width=w, height=w, camera=cam, overlap=True)
    r.add(4, train=True)
    torch.cuda.synchronize()
    rows = g.copy_rows()
    assert rows.shape[0] > 0 and torch.equal(rows, g.batch_rows())
    for _ in range(2):
        r.add(4, train=True)
    r.add(500)
    r.finish()
    torch.cuda.synchronize()
    assert r.spp == 512
    assert np.abs(g.get_params() - p0).max() > 1e-3
    s_g, q_g = (x.cpu().numpy().astype(np.float64) for x in (r.sum, r.sumsq))
    s_p, q_p, _ = sc.render(width=w, height=w, spp=512, seed=82, camera=cam)
    s_p, q_p = (x.cpu().numpy().astype(np.float64) for x in (s_p, q_p))
    spp = 512
    frac, _ = zfrac(s_g, q_g, s_p, q_p, spp)
    n = w * w * 3 * spp
    se = np.sqrt(q_g.mean() / spp / n + q_p.mean() / spp / n)
    dm = (s_g.mean() - s_p.mean()) / spp
    print(f"overlap: |z|>3 in {frac:.4f}; image mean diff {dm:.6f} ({dm / se:.2f} SE)")
    assert abs(dm) <= 4 * se
    assert frac < 0.015
