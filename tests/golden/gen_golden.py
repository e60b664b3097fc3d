"""Writes the golden hit fixtures from the REFERENCE's own code.

Run in a container that has the reference tree (it needs
oracle/_ref/libtpt_ref.so, built by `make -C oracle ref` from the reference's
unmodified sources).  For each scene it stores seeded rays of the four kinds
(camera primaries, one-bounce secondaries, uniform, adversarial) with the
reference's hitable::hit result: leaf id, t, the hit_record (u, v, point,
normal) and the material index; plus camera fields and pre-order bvh_node
boxes; and the Perlin noise vectors of ref_perlin.npz.  The fixtures are data
only (inputs and expected outputs).

    python tests/golden/gen_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
from oracle_lib import Oracle, Reference  # noqa: E402

# (fixture name, scene name on the oracle side, scene name on the reference side)
SCENES = [
    ("cornell555", "cornell555", "cornell555"),
    ("random_spheres", "random_spheres", "random_spheres"),
    ("native_cornell_box", "ref_cornell_box", "native:cornell_box"),
    ("box_list", "box_list", "box_list"),
]
N_PER_KIND = 1024
CAMERAS = {
    # lookfrom, lookat, vup, vfov, aspect, aperture, focus, t0, t1
    "cornell555": [278, 278, -800, 278, 278, 0, 0, 1, 0, 40, 1, 0, 10, 0, 0],
    "random_spheres": [13, 2, 3, 0, 0, 0, 0, 1, 0, 20, 1, 0.1, 13.190906, 0, 0],
    "reference_main": [0, 0, 800, 0, 0, 0, 0, 1, 0, 90, 1, 0.1, 10, 0, 1],
    "odd": [1.5, -2.25, 7.125, 0.3, 0.7, -1.1, 0.2, 1, 0.1, 33.3, 1.7777778, 0.37, 4.2, 0.25, 0.75],
}


def main():
    O, R = Oracle(), Reference()
    out = {}
    for fix, mine, ref in SCENES:
        s = O.scene(mine)
        h = R.scene(ref)
        rays = np.concatenate([s.make_rays(k, N_PER_KIND, seed=2024 + k) for k in range(4)], axis=1)
        leaf, t, rec, mat = R.trace(h, rays, record=True)
        out[f"{fix}/rays"] = rays
        out[f"{fix}/leaf"] = leaf
        out[f"{fix}/t"] = t
        out[f"{fix}/rec"] = rec
        out[f"{fix}/mat"] = mat
        out[f"{fix}/bvh"] = R.bvh_dump(h)
        out[f"{fix}/n_leaves"] = np.array([R.leaf_count(h)], np.int32)
    for name, p in CAMERAS.items():
        out[f"camera/{name}/params"] = np.asarray(p, np.float32)
        out[f"camera/{name}/fields"] = R.camera_fields(p)
    path = os.path.join(HERE, "ref_hits.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path} ({os.path.getsize(path)} bytes)")
    perlin(O, R)


def perlin(O, R):
    """ref_perlin.npz: the reference's perlin_noise gradient table as its
    constructor draws it on a fresh thread, and its noise(p), turb(p) and
    perlin_noise_texture(2).value(0, 0, p).x() at 2048 points with that table
    and the scene API's permutations installed (the reference's own
    permutations come from the clock)."""
    O.scene("ref_two_perlin_spheres")
    R.scene("ref_two_perlin_spheres")
    grad, _ = R.perlin_tables()
    _, perm = O.perlin_tables()
    rng = np.random.default_rng(2025)
    pts = np.concatenate([rng.uniform(-4, 4, (1024, 3)), rng.uniform(-700, 700, (768, 3)),
                          np.round(rng.uniform(-300, 300, (256, 3)))]).astype(np.float32)
    noise, turb, tex = R.perlin_eval(grad, perm, pts, 2.0)
    path = os.path.join(HERE, "ref_perlin.npz")
    np.savez_compressed(path, grad=grad, perm=perm, pts=pts, scale=np.float32(2.0), noise=noise, turb=turb, tex=tex)
    print(f"wrote {path} ({os.path.getsize(path)} bytes)")


if __name__ == "__main__":
    main()
