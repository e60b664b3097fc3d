"""Perlin noise of the oracle against the reference's own perlin_noise
(src/perlin_noise.cc:3-21, src/utils.cc:160-225, src/texture.cc:18-25),
compiled from its sources (oracle/_ref).

The reference draws its 256 gradients from drand_r (reproducible on a fresh
thread) and shuffles its permutations with a clock-seeded engine (not
reproducible). So the gradients are compared as built, and the evaluation of
noise, turb and the texture value is compared bit for bit after the oracle's
permutations are installed in the reference's static tables."""
import numpy as np
import pytest


def points(n, seed):
    rng = np.random.default_rng(seed)
    a = rng.uniform(-4, 4, (n // 4, 3))            # around the origin (negative cells: floor & 255)
    b = rng.uniform(-700, 700, (n // 4, 3))        # past the 256-cell wrap
    c = np.round(rng.uniform(-300, 300, (n // 8, 3)))  # lattice points
    d = c + rng.choice([-1, 1], c.shape) * rng.uniform(0, 1e-4, c.shape)  # just off the lattice
    e = rng.uniform(-2, 2, (n - len(a) - len(b) - len(c) - len(d), 3)).astype(np.float32) * 0.5 + 0.25
    return np.concatenate([a, b, c, d, e]).astype(np.float32)


def test_gradient_tables_equal_the_reference(oracle, reference):
    oracle.scene("ref_two_perlin_spheres")
    reference.scene("ref_two_perlin_spheres")
    g_o, p_o = oracle.perlin_tables()
    g_r, p_r = reference.perlin_tables()
    assert np.array_equal(g_o.view(np.uint32), g_r.view(np.uint32))
    for k in range(3):
        assert np.array_equal(np.sort(p_r[k]), np.arange(256)) and np.array_equal(np.sort(p_o[k]), np.arange(256))
    assert np.allclose(np.linalg.norm(g_o.astype(np.float64), axis=1), 1, atol=1e-6)


@pytest.mark.parametrize("scale", [2.0, 4.0, 0.5])
def test_noise_turb_and_texture_equal_the_reference(oracle, reference, scale):
    oracle.scene("ref_two_perlin_spheres")
    g, p = oracle.perlin_tables()
    pts = points(1 << 15, seed=int(scale * 10))
    n_o, t_o, x_o = oracle.perlin_eval(g, p, pts, scale)
    n_r, t_r, x_r = reference.perlin_eval(g, p, pts, scale)
    assert np.array_equal(n_o.view(np.uint32), n_r.view(np.uint32))
    assert np.array_equal(t_o.view(np.uint32), t_r.view(np.uint32))
    assert np.array_equal(x_o.view(np.uint32), x_r.view(np.uint32))
    # a different permutation changes the noise: the comparison sees the tables
    n_s, _, _ = reference.perlin_eval(g, p[:, ::-1], pts, scale)
    assert not np.array_equal(n_s, n_o)


def test_noise_closed_forms(oracle):
    """At lattice points every corner term vanishes (fade(0) = 0), so noise and
    turb are exactly 0 and the texture is 0.5 (1 + sin(scale z)); noise is
    periodic with period 256 in each axis."""
    oracle.scene("ref_two_perlin_spheres")
    g, p = oracle.perlin_tables()
    rng = np.random.default_rng(3)
    lat = np.round(rng.uniform(-500, 500, (4096, 3))).astype(np.float32)
    n, t, x = oracle.perlin_eval(g, p, lat, 2.0)
    assert not n.any() and not t.any()
    # sinf against float64 sin of the same float argument: within an ulp of 1
    want = 0.5 * (1 + np.sin((np.float32(2.0) * lat[:, 2]).astype(np.float64)))
    assert np.abs(x - want).max() <= 1.2e-7
    q = rng.uniform(0, 8, (4096, 3)).astype(np.float32)
    n1, _, _ = oracle.perlin_eval(g, p, q, 2.0)
    n2, _, _ = oracle.perlin_eval(g, p, q + np.float32(256), 2.0)
    assert np.allclose(n1, n2, atol=2e-4)
    # one octave of turb is |noise|; noise is bounded by the corner weights
    _, t1, _ = oracle.perlin_eval(g, p, q, 2.0, octaves=1)
    assert np.array_equal(t1, n1) and (n1 <= np.sqrt(3)).all()


def test_noise_equals_the_golden_vectors(oracle):
    """tests/golden/ref_perlin.npz (written by tests/golden/gen_golden.py from
    the reference build): the gradient table the scene API draws equals the
    reference's, and noise, turb and the texture value equal the reference's
    outputs bit for bit."""
    import os

    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_perlin.npz")
    gv = np.load(path)
    oracle.scene("ref_two_perlin_spheres")
    g_o, p_o = oracle.perlin_tables()
    assert np.array_equal(g_o.view(np.uint32), gv["grad"].view(np.uint32))
    assert np.array_equal(p_o, gv["perm"])
    n, t, x = oracle.perlin_eval(gv["grad"], gv["perm"], gv["pts"], float(gv["scale"]))
    for mine, ref in ((n, gv["noise"]), (t, gv["turb"]), (x, gv["tex"])):
        assert np.array_equal(mine.view(np.uint32), ref.view(np.uint32))
