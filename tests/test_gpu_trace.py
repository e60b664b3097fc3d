"""tpt_trace_closest against the oracle's hitable::hit: leaf id and t bit for
bit (t within 1e-6 relative is the stated bar; the kernel meets bit
equality), hit_record point/normal bit for bit, u/v within 1e-5 (device
atan2f/asinf), material index exact."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FULL = [("cornell555", 1 << 24), ("random_spheres", 1 << 24), ("random_spheres_n20000", 1 << 22)]
SMALL = ["ref_cornell_box", "box_list", "random_spheres_textured", "ref_light_spheres", "ref_two_perlin_spheres",
         "ref_sphere_cornell_box", "random_spheres_n20000", "cornell_smoke"]
CHUNK = 1 << 22


def compare(s_or, sc, rays, torch, device, record=True, time=None):
    want = s_or.trace(rays, record=record, time=time)
    got = sc.trace(torch.from_numpy(rays).to(device), record=record,
                   time=None if time is None else torch.from_numpy(time).to(device))
    got = [g.cpu().numpy() for g in got]
    leaf_ok = np.array_equal(got[0], want[0])
    hit = want[0] >= 0
    bad = np.nonzero(got[0] != want[0])[0]
    assert leaf_ok, f"{len(bad)} leaf mismatches, first at {bad[:5]}: rays {rays[:, bad[:2]].T} gpu {got[0][bad[:5]]} cpu {want[0][bad[:5]]}"
    assert np.array_equal(got[1][hit].view(np.uint32), want[1][hit].view(np.uint32)), "t differs"
    if record:
        g_rec, w_rec = got[2][:, hit], want[2][:, hit]
        assert np.array_equal(g_rec[2:].view(np.uint32), w_rec[2:].view(np.uint32)), "point/normal differ"
        assert np.nanmax(np.abs(g_rec[:2] - w_rec[:2]), initial=0) <= 1e-5, "u/v differ"
        assert np.array_equal(got[3], want[3]), "material differs"
    return hit.mean(), np.isnan(want[1][hit]).sum()


@pytest.mark.parametrize("name,n", FULL)
def test_trace_bit_exact_2e24(tpt, oracle, device, name, n):
    import torch

    s = oracle.scene(name)
    sc = tpt.Scene.named(name)
    # 25% camera primaries, 50% one-bounce secondaries, 25% uniform, + 2^14 adversarial
    plan = [(0, n // 4), (1, n // 2), (2, n // 4), (3, 1 << 14)]
    total_nan = 0
    for kind, count in plan:
        for off in range(0, count, CHUNK):
            m = min(CHUNK, count - off)
            rays = s.make_rays(kind, m, seed=1000 * kind + off // CHUNK + 1)
            frac, nans = compare(s, sc, rays, torch, device, record=(off == 0))
            total_nan += nans
            print(f"{name} kind {kind} chunk {off // CHUNK}: {m} rays, hit {frac:.3f}, nan-t hits {nans}",
                  flush=True)
            sys.stdout.flush()
    if name == "cornell555":
        assert total_nan > 0  # the in-plane NaN rule was exercised on the device


@pytest.mark.parametrize("name", SMALL)
def test_trace_bit_exact_scenes(tpt, oracle, device, name):
    import torch

    s = oracle.scene(name)
    sc = tpt.Scene.named(name)
    for kind in range(4):
        rays = s.make_rays(kind, 1 << 16, seed=77 + kind)
        compare(s, sc, rays, torch, device)


def test_trace_golden_reference_vectors(tpt, device, golden):
    """The device against the reference's own recorded hits."""
    import torch

    for fix, name in [("cornell555", "cornell555"), ("random_spheres", "random_spheres"),
                      ("native_cornell_box", "ref_cornell_box"), ("box_list", "box_list")]:
        sc = tpt.Scene.named(name)
        rays = golden[f"{fix}/rays"]
        leaf, t, rec, mat = [g.cpu().numpy() for g in sc.trace(torch.from_numpy(rays).to(device), record=True)]
        g_leaf = golden[f"{fix}/leaf"]
        hit = g_leaf >= 0
        assert np.array_equal(leaf, g_leaf)
        assert np.array_equal(t[hit].view(np.uint32), golden[f"{fix}/t"][hit].view(np.uint32))
        assert np.array_equal(rec[2:, hit].view(np.uint32), golden[f"{fix}/rec"][2:, hit].view(np.uint32))
        assert np.array_equal(mat, golden[f"{fix}/mat"])


def test_trace_tmin_tmax_and_empty(tpt, oracle, device):
    import torch

    s = oracle.scene("cornell555")
    sc = tpt.Scene.named("cornell555")
    rays = s.make_rays(1, 1 << 14, seed=5)
    for tmin, tmax in [(1e-4, 3e38), (0.001, 300.0), (10.0, 1000.0), (0.001, float("nan"))]:
        want = s.trace(rays, tmin=tmin, tmax=tmax)
        got = [g.cpu().numpy() for g in sc.trace(torch.from_numpy(rays).to(device), tmin=tmin, tmax=tmax)]
        assert np.array_equal(got[0], want[0])
    empty = torch.zeros((6, 0), dtype=torch.float32, device=device)
    leaf, t = sc.trace(empty)
    assert leaf.numel() == 0


def test_trace_moving_spheres(tpt, oracle, device):
    """Rays with times (tpt_trace_closest_timed) on the reference's
    random_scene: moving spheres placed at each ray's time."""
    import torch

    s = oracle.scene("ref_random_scene")
    sc = tpt.Scene.named("ref_random_scene")
    rng = np.random.default_rng(3)
    for kind in range(4):
        rays = s.make_rays(kind, 1 << 16, seed=177 + kind)
        time = rng.uniform(0, 1, rays.shape[1]).astype(np.float32)
        time[:8] = [0, 1, 0.5, 2, -1, 1e-7, 0.9999999, 3]
        compare(s, sc, rays, torch, device, time=time)
    # without times the rays see every sphere at time 0 (the ray default)
    rays = s.make_rays(0, 1 << 14, seed=5)
    compare(s, sc, rays, torch, device, record=False)


def test_trace_wrapped_bvh_media_and_earth(tpt, oracle, device, earth_cwd):
    """The reference's oneweek_final: a bvh_node under rotate_y and translate
    (a transformed segment), media inside the top-level list, a moving sphere
    and the earth texture; timed rays, bit for bit with the oracle."""
    import torch

    s = oracle.scene("ref_oneweek_final")
    sc = tpt.Scene.named("ref_oneweek_final")
    rng = np.random.default_rng(11)
    for kind in range(4):
        rays = s.make_rays(kind, 1 << 15, seed=401 + kind)
        time = rng.uniform(0, 1, rays.shape[1]).astype(np.float32)
        compare(s, sc, rays, torch, device, time=time)


def grazing_rays(host, n, seed):
    """Rays aimed where the sphere-only walk's box bound is tightest: grazing
    the largest sphere (the ground) from heights 1e-4..1 above it, tangent to
    the small spheres, leaving a sphere's surface nearly tangent to it, and
    nearly parallel to an axis; direction lengths 2^-3..2^3."""
    rng = np.random.default_rng(seed)
    sph = [(np.array(f[:3], np.float64), float(f[3])) for f, a, b in host.slots() if (b >> 28) == 1]
    big = max(sph, key=lambda s: abs(s[1]))
    small = [s for s in sph if s is not big]
    cs = np.array([c for c, _ in small])
    rs = np.array([r for _, r in small])
    lo, hi = cs.min(axis=0) - 2, cs.max(axis=0) + 2
    q = n // 4
    o = np.zeros((n, 3))
    d = np.zeros((n, 3))

    def perp(u):
        a = rng.normal(size=u.shape)
        a -= (a * u).sum(1, keepdims=True) * u
        return a / np.linalg.norm(a, axis=1, keepdims=True)

    def tangent_from(p, c, r, jitter):
        v = c - p
        dist = np.linalg.norm(v, axis=1, keepdims=True)
        u = v / dist
        beta = np.arcsin(np.clip(np.abs(r)[:, None] / dist, 0, 1)) * (1 + jitter[:, None])
        return np.cos(beta) * u + np.sin(beta) * perp(u)

    # (a) grazing the largest sphere from just above it
    k = rng.integers(0, len(small), q)
    p = cs[k] + rng.uniform(-1, 1, (q, 3)) * [1, 0, 1]
    up = p - big[0]
    up /= np.linalg.norm(up, axis=1, keepdims=True)
    p = big[0] + up * (abs(big[1]) + 10 ** rng.uniform(-4, 0, (q, 1)))
    o[:q], d[:q] = p, tangent_from(p, np.broadcast_to(big[0], (q, 3)), np.full(q, big[1]),
                                   rng.choice([-1, 1], q) * 10 ** rng.uniform(-7, -3, q))
    # (b) tangent to a small sphere from anywhere in the spheres' box
    k = rng.integers(0, len(small), q)
    p = rng.uniform(lo, hi, (q, 3))
    o[q:2 * q], d[q:2 * q] = p, tangent_from(p, cs[k], rs[k], rng.choice([-1, 1], q) * 10 ** rng.uniform(-7, -2, q))
    # (c) leaving a small sphere's surface, nearly tangent to it
    k = rng.integers(0, len(small), q)
    nrm = rng.normal(size=(q, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    o[2 * q:3 * q] = cs[k] + rs[k][:, None] * nrm
    d[2 * q:3 * q] = perp(nrm) + nrm * (rng.choice([-1, 1], q) * 10 ** rng.uniform(-6, 0, q))[:, None]
    # (d) nearly parallel to an axis
    m = n - 3 * q
    o[3 * q:] = rng.uniform(lo, hi, (m, 3))
    dd = rng.choice([-1, 1], (m, 3)) * 10 ** rng.uniform(-8, -1, (m, 3))
    ax = rng.integers(0, 3, m)
    dd[np.arange(m), ax] = rng.choice([-1.0, 1.0], m)
    d[3 * q:] = dd
    d *= 2.0 ** rng.uniform(-3, 3, (n, 1))
    return np.ascontiguousarray(np.concatenate([o, d], axis=1).T.astype(np.float32))


@pytest.mark.parametrize("name", ["random_spheres", "random_spheres_n20000"])
def test_trace_bounded_sphere_walk_adversarial(tpt, oracle, device, name):
    """The sphere-only walk skips boxes entered beyond a bound on the held hit
    (the reference tests them all).  Rays where float sphere roots stray
    furthest from their boxes (grazing the r = 1000 ground, tangents, rays
    leaving a surface, near-axis directions), at the default and at shifted
    (t_min, t_max): leaf and t bit for bit with the oracle's full walk."""
    import torch

    s = oracle.scene(name)
    sc = tpt.Scene.named(name)
    host = tpt.HostScene(name)
    rays = grazing_rays(host, 1 << 20, seed=23)
    frac, _ = compare(s, sc, rays, torch, device)
    print(f"{name}: grazing set hit {frac:.3f}", flush=True)
    for tmin, tmax in [(-50.0, 3e38), (0.5, 40.0)]:
        want = s.trace(rays, tmin=tmin, tmax=tmax)
        got = [g.cpu().numpy() for g in sc.trace(torch.from_numpy(rays).to(device), tmin=tmin, tmax=tmax)]
        assert np.array_equal(got[0], want[0])
        hit = want[0] >= 0
        assert np.array_equal(got[1][hit].view(np.uint32), want[1][hit].view(np.uint32))
