"""Pins the error terms behind the sphere-only walk's box bound
(tiny-path-tracer_amd/csrc/traverse.h, sphere_walk_bound): a skipped box must
hold no sphere whose computed t can reach the held hit.  Two facts carry it:

  1. sphere_t's float coefficients move a root by at most
     8e-4 (|t| + r/|d|) from the exact root of the same (o, d, c, r)
     (measured worst 5.7e-4; at tangency the rounding of b^2 - 4ac passes
     through the square root);
  2. a sphere's box is c -+ r rounded to float, so the exact sphere stands out
     of it by at most half an ulp of the coordinates (2^-24 M).

The bound adds 2^-6 (|t| + |t_min|) + 2^-6 * 2 / |d| + 2^-22 M max|1/d_i| and
never bounds boxes that hold a sphere of radius above 2.  This test emulates
sphere_t on the CPU (float32 a, b, c and discriminant, each operation rounded;
float64 roots rounded to float32, as the device and the reference compute
them), takes the exact root in extended precision, and checks (1) on grazing,
tangent-adjacent and general rays for radii 0.2 to 1000, and that the bound's
terms hold it with the stated margin.  Fact 2 is checked directly on the
rounding of c -+ r."""
import numpy as np

EPS_TERM = 8e-4  # the root error term stated in traverse.h


def sphere_t_f32(o, d, c, r, tmin, tmax):
    """sphere_t of traverse.h (and the reference's sphere::hit), vectorised:
    float32 arithmetic, one rounding per operation, FP64 roots."""
    f = np.float32
    ox, oy, oz = (o[:, k] - c[:, k] for k in range(3))  # float32 - float32: rounded
    dx, dy, dz = d[:, 0], d[:, 1], d[:, 2]
    a = (dx * dx + dy * dy) + dz * dz
    b = f(2.0) * ((dx * ox + dy * oy) + dz * oz)
    cc = ((ox * ox + oy * oy) + oz * oz) - r * r
    disc = b * b - (f(4.0) * a) * cc
    hit = disc > 0
    sq = np.sqrt(np.maximum(disc, 0).astype(np.float64))
    den = (f(2.0) * a).astype(np.float64)
    nb = -b.astype(np.float64)
    t0 = ((nb - sq) / den).astype(np.float32)
    t1 = ((nb + sq) / den).astype(np.float32)
    use0 = (t0 < tmax) & (t0 > tmin)
    use1 = ~use0 & (t1 < tmax) & (t1 > tmin)
    t = np.where(use0, t0, t1)
    root0 = use0
    return hit & (use0 | use1), t, root0


def exact_roots(o, d, c, r):
    """Both roots of |o + t d - c| = r for the float32 inputs taken as exact,
    in extended precision (stable form)."""
    L = np.longdouble
    oc = o.astype(L) - c.astype(L)
    dd = d.astype(L)
    a = (dd * dd).sum(1)
    hb = (dd * oc).sum(1)
    cc = (oc * oc).sum(1) - r.astype(L) ** 2
    disc = hb * hb - a * cc
    sq = np.sqrt(np.maximum(disc, L(0)))
    q = -hb - np.copysign(sq, hb)
    with np.errstate(divide="ignore", invalid="ignore"):
        ra, rb = q / a, cc / q
    return np.minimum(ra, rb), np.maximum(ra, rb), disc


def rays(rng, n, radius):
    """Origins 1-40 radii out, directions aimed at points around the sphere
    (impact parameter near r: grazing and tangent-adjacent rays), lengths
    2^-3..2^3, centres up to 1000 from the origin of coordinates."""
    c = rng.uniform(-1000, 1000, (n, 3)).astype(np.float32)
    r = np.full(n, radius, np.float32)
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    dist = radius * rng.uniform(1.0, 40.0, (n, 1))
    o = (c + u * dist).astype(np.float32)
    # aim: the centre offset sideways by an impact parameter around r
    side = rng.normal(size=(n, 3))
    side -= (side * u).sum(1, keepdims=True) * u
    side /= np.linalg.norm(side, axis=1, keepdims=True)
    imp = radius * np.where(rng.random((n, 1)) < 0.7, 1.0 + rng.normal(0, 1e-3, (n, 1)), rng.uniform(0, 1.2, (n, 1)))
    target = c + side * imp
    d = target - o
    d *= 2.0 ** rng.uniform(-3, 3, (n, 1)) / np.linalg.norm(d, axis=1, keepdims=True)
    return o, d.astype(np.float32), c, r


def test_sphere_root_error_within_the_bound_term():
    rng = np.random.default_rng(5)
    worst = 0.0
    n_hits = 0
    for radius in (0.2, 1.0, 2.0, 90.0, 1000.0):
        o, d, c, r = rays(rng, 200_000, radius)
        hit, t, root0 = sphere_t_f32(o, d, c, r, np.float32(-3.4e38), np.float32(3.4e38))
        lo, hi, disc = exact_roots(o, d, c, r)
        # where the float walk reports a hit, compare with the nearer exact
        # root (the exact ray may graze or just miss: then the closest point)
        dl = np.linalg.norm(d.astype(np.float64), axis=1)
        L = np.longdouble
        tc = -(d.astype(L) * (o.astype(L) - c.astype(L))).sum(1) / (d.astype(L) ** 2).sum(1)
        exact = np.where(disc > 0, np.where(root0, lo, hi), tc).astype(np.float64)
        err = np.abs(t.astype(np.float64) - exact)[hit]
        scale = EPS_TERM * (np.abs(exact[hit]) + r[hit] / dl[hit])
        ratio = err / scale
        n_hits += hit.sum()
        worst = max(worst, float(ratio.max()))
        print(f"r={radius}: {hit.sum()} hits, worst err / (8e-4 (|t| + r/|d|)) = {ratio.max():.3f}")
    assert n_hits > 500_000
    assert worst < 1.0
    # the bound's terms are 2^-6 |t| and 2^-6 * 2/|d| (for r <= 2): at least
    # 19x the stated term
    assert (2.0 ** -6) / EPS_TERM >= 19.0


def test_float_box_misses_the_sphere_by_at_most_half_an_ulp():
    rng = np.random.default_rng(6)
    c = rng.uniform(-1000, 1000, (1_000_000, 3)).astype(np.float32)
    r = rng.uniform(0.01, 1000, (1_000_000, 1)).astype(np.float32)
    lo, hi = c - r, c + r  # the reference's bounding box: float32, rounded once
    M = np.maximum(np.abs(lo), np.abs(hi)).astype(np.float64)
    out_lo = (lo.astype(np.float64) - (c.astype(np.float64) - r.astype(np.float64))) / M
    out_hi = ((c.astype(np.float64) + r.astype(np.float64)) - hi.astype(np.float64)) / M
    assert max(out_lo.max(), out_hi.max()) <= 2.0 ** -24
    # the bound allows 2^-22 M per unit of 1/d: four times that
    assert 2.0 ** -22 / 2.0 ** -24 == 4.0
