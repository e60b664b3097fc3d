"""ORACLE — test infrastructure only (never imported by the product).

float64 numpy restatement of the guiding network, its vMF-mixture loss and the
backward pass, for checking the device kernels (tpt_guide_infer,
tpt_guide_gradient).  With `emulate=True` it rounds to bf16 at exactly the
points where the device feeds MFMA operands (weights, layer inputs, dout, dz,
and the activations/deltas of the weight-gradient products), so the device's
FP32-accumulated results agree to accumulation-order rounding.  With
`emulate=False` it is the exact float64 function, which tests pin by finite
differences.  Layout and formulas: include/tpt.h (learned guiding section).
"""
import numpy as np

G_IN, G_OUT, G_LOBES = 42, 40, 8
W1, W2, W3, W4 = 0, 64 * 48, 64 * 48 + 4096, 64 * 48 + 8192
B1 = W4 + 48 * 64
B2, B3, B4 = B1 + 64, B1 + 128, B1 + 192
NPARAM = B4 + 48
LOGK_MIN, LOGK_MAX = -3.0, 7.0


def bf16(x):
    """Round float32 values to bf16 (nearest even), returned as float64."""
    a = np.ascontiguousarray(x, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, np.float64)
    return (p[W1:W2].reshape(64, 48), p[W2:W3].reshape(64, 64), p[W3:W4].reshape(64, 64),
            p[W4:B1].reshape(48, 64), p[B1:B2], p[B2:B3], p[B3:B4], p[B4:NPARAM], p[NPARAM])


def live_mask():
    m = np.ones(NPARAM + 1, bool)
    m[W1:W2] = (np.arange(W2 - W1) % 48) < G_IN
    m[W4:B1] = (np.arange(B1 - W4) // 64) < G_OUT
    m[B4:NPARAM] = np.arange(48) < G_OUT
    return m


def features(in9):
    """(n, 9) -> (n, 48): sin/cos(2^b pi p_c) for c in xyz, b in 0..5, then n, wo, zeros.
    Band 0 is sin/cos(pi p) rounded to float32; band b+1 follows from band b by
    the float32 doublings sin 2x = (2 sin x) cos x, cos 2x = 1 - (2 sin x) sin x,
    evaluated in that order as the device does."""
    in9 = np.asarray(in9, np.float32)
    n = in9.shape[0]
    f = np.zeros((n, 48), np.float64)
    two, one = np.float32(2), np.float32(1)
    for c in range(3):
        p = in9[:, c].astype(np.float64)
        s = np.sin(np.pi * p).astype(np.float32)
        co = np.cos(np.pi * p).astype(np.float32)
        for b in range(6):
            f[:, 12 * c + 2 * b] = s
            f[:, 12 * c + 2 * b + 1] = co
            s2 = two * s
            s, co = s2 * co, one - s2 * s
    f[:, 36:42] = in9[:, 3:9]
    return f


def forward(p, in9, emulate=True):
    w1, w2, w3, w4, b1, b2, b3, b4, _ = unpack(p)
    r = bf16 if emulate else (lambda a: np.asarray(a, np.float64))
    x = features(in9)
    xs = r(x.astype(np.float32)) if emulate else x
    z1 = xs @ r(w1).T + b1
    h1 = np.maximum(z1, 0)
    z2 = r(h1) @ r(w2).T + b2 if emulate else h1 @ w2.T + b2
    h2 = np.maximum(z2, 0)
    z3 = r(h2) @ r(w3).T + b3 if emulate else h2 @ w3.T + b3
    h3 = np.maximum(z3, 0)
    out = (r(h3) @ r(w4).T if emulate else h3 @ w4.T) + b4
    return out[:, :G_OUT], dict(x=xs, h1=h1, h2=h2, h3=h3)


def decode(out):
    raw = out[:, :40].reshape(-1, 8, 5)
    mu_raw = raw[:, :, :3]
    nrm = np.linalg.norm(mu_raw, axis=2)
    mu = np.where(nrm[..., None] > 1e-10, mu_raw / np.maximum(nrm, 1e-30)[..., None], np.array([0, 0, 1.0]))
    lk = raw[:, :, 3]
    kappa = np.exp(np.clip(lk, LOGK_MIN, LOGK_MAX))
    z = raw[:, :, 4]
    w = np.exp(z - z.max(axis=1, keepdims=True))
    w /= w.sum(axis=1, keepdims=True)
    return mu, kappa, w, nrm, lk


def vmf(kappa, cosang):
    return kappa / (2 * np.pi * (-np.expm1(-2 * kappa))) * np.exp(kappa * (cosang - 1))


def mixture_pdf(out, omega):
    mu, kappa, w, _, _ = decode(out)
    cs = np.einsum("nlk,nk->nl", mu, omega)
    return (w * vmf(kappa, cs)).sum(axis=1)


def loss_and_dout(out, rows, n):
    """Loss sum of -c log g and d(loss/n)/d out; plus the mix-weight gradient."""
    omega, c, q = rows[:, 9:12], rows[:, 12], rows[:, 13]
    mu, kappa, w, nrm, lk = decode(out)
    cs = np.einsum("nlk,nk->nl", mu, omega)
    d = vmf(kappa, cs)
    g = np.maximum((w * d).sum(axis=1), 1e-30)
    valid = (c > 0) & np.isfinite(c)
    cc = np.where(valid, c, 0.0)
    loss = -(cc * np.log(g)).sum()
    f = -cc / (n * g)
    wd = w * d
    dout = np.zeros((out.shape[0], 8, 5))
    s = np.where(nrm > 1e-10, f[:, None] * wd * kappa / np.maximum(nrm, 1e-30), 0.0)
    dout[:, :, :3] = s[..., None] * (omega[:, None, :] - cs[..., None] * mu)
    live = (lk > LOGK_MIN) & (lk < LOGK_MAX)
    dout[:, :, 3] = np.where(live, f[:, None] * wd * (1 - 2 * kappa / np.expm1(2 * kappa) + kappa * (cs - 1)), 0.0)
    dout[:, :, 4] = f[:, None] * w * (d - g[:, None])
    return loss, dout.reshape(-1, 40), g, valid


def gradient(p, rows, emulate=True):
    """Gradient (NPARAM + 1) of -(1/n) sum c log g (+ the mix-weight term) and the loss sum."""
    rows = np.asarray(rows, np.float64)
    n = rows.shape[0]
    w1, w2, w3, w4, b1, b2, b3, b4, a = unpack(p)
    out, act = forward(p, rows[:, :9].astype(np.float32), emulate)
    loss, dout40, g, valid = loss_and_dout(out, rows, n)
    r = bf16 if emulate else (lambda v: np.asarray(v, np.float64))
    dout = np.zeros((n, 48))
    dout[:, :40] = dout40
    if emulate:
        dout = dout.astype(np.float32).astype(np.float64)
    gr = np.zeros(NPARAM + 1)
    h1, h2, h3, x = act["h1"], act["h2"], act["h3"], act["x"]
    d_b = r(dout) if emulate else dout
    # b4: the device sums the bf16 output deltas (the operand it already has
    # in the weight-gradient layout); b1..b3 sum the fp32 deltas
    gr[B4:NPARAM] = d_b.sum(axis=0)
    gr[W4:B1] = (d_b.T @ (r(h3) if emulate else h3)).reshape(-1)
    dz3 = (d_b @ r(w4)) * (h3 > 0) if emulate else (dout @ w4) * (h3 > 0)
    gr[B3:B4] = dz3.sum(axis=0)
    dz3b = r(dz3) if emulate else dz3
    gr[W3:W4] = (dz3b.T @ (r(h2) if emulate else h2)).reshape(-1)
    dz2 = (dz3b @ r(w3)) * (h2 > 0) if emulate else (dz3 @ w3) * (h2 > 0)
    gr[B2:B3] = dz2.sum(axis=0)
    dz2b = r(dz2) if emulate else dz2
    gr[W2:W3] = (dz2b.T @ (r(h1) if emulate else h1)).reshape(-1)
    dz1 = (dz2b @ r(w2)) * (h1 > 0) if emulate else (dz2 @ w2) * (h1 > 0)
    gr[B1:B2] = dz1.sum(axis=0)
    dz1b = r(dz1) if emulate else dz1
    gr[W1:W2] = (dz1b.T @ x).reshape(-1)
    alpha = 1 / (1 + np.exp(-a))
    q = rows[:, 13]
    m = alpha * g + (1 - alpha) * q
    cc = np.where(valid, rows[:, 12], 0.0)
    gr[NPARAM] = np.sum(np.where(m > 0, -cc / n * (g - q) * alpha * (1 - alpha) / np.maximum(m, 1e-300), 0.0))
    gr[~live_mask()] = 0.0
    return gr, loss


def loss_value(p, rows):
    """-(1/n) sum c log g in exact float64 (for finite differences)."""
    rows = np.asarray(rows, np.float64)
    out, _ = forward(p, rows[:, :9].astype(np.float32), emulate=False)
    loss, _, _, _ = loss_and_dout(out, rows, rows.shape[0])
    return loss / rows.shape[0]


def mix_loss_value(p, rows):
    """-(1/n) sum c log m, m = a g + (1 - a) q, a = sigmoid(mix logit), exact float64."""
    rows = np.asarray(rows, np.float64)
    out, _ = forward(p, rows[:, :9].astype(np.float32), emulate=False)
    g = np.maximum(mixture_pdf(out, rows[:, 9:12]), 1e-30)
    a = 1 / (1 + np.exp(-np.float64(p[NPARAM])))
    c = rows[:, 12]
    cc = np.where((c > 0) & np.isfinite(c), c, 0.0)
    return -(cc * np.log(a * g + (1 - a) * rows[:, 13])).sum() / rows.shape[0]


def adam_first_step(p, grad, lr):
    """Parameters after one Adam step (beta 0.9/0.999, eps 1e-8) from zero moments."""
    m = 0.1 * grad
    v = 0.001 * grad * grad
    mh, vh = m / (1 - 0.9), v / (1 - 0.999)
    q = np.asarray(p, np.float64) - lr * mh / (np.sqrt(vh) + 1e-8) * live_mask()
    q[NPARAM] = np.clip(q[NPARAM], -3, 3)
    return q


def init_params(seed=0):
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAM + 1)
    p[W1:W2] = rng.normal(0, np.sqrt(2 / 42), 64 * 48)
    p[W2:W3] = rng.normal(0, np.sqrt(2 / 64), 4096)
    p[W3:W4] = rng.normal(0, np.sqrt(2 / 64), 4096)
    p[W4:B1] = rng.normal(0, 0.1 * np.sqrt(2 / 64), 48 * 64)
    p[B1:NPARAM] = rng.normal(0, 0.1, NPARAM - B1)
    p[~live_mask()] = 0
    return p.astype(np.float32)


def random_rows(n, seed=0):
    rng = np.random.default_rng(seed)
    rows = np.zeros((n, 16), np.float32)
    rows[:, 0:3] = rng.uniform(0, 1, (n, 3))
    nv = rng.normal(size=(n, 3))
    rows[:, 3:6] = nv / np.linalg.norm(nv, axis=1, keepdims=True)
    wo = rng.normal(size=(n, 3))
    rows[:, 6:9] = wo / np.linalg.norm(wo, axis=1, keepdims=True)
    om = rng.normal(size=(n, 3))
    rows[:, 9:12] = om / np.linalg.norm(om, axis=1, keepdims=True)
    rows[:, 12] = rng.exponential(1.0, n) * (rng.uniform(size=n) < 0.8)
    rows[:, 13] = rng.uniform(0.05, 1.0, n)
    return rows
