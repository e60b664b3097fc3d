"""The oracle against golden vectors written by the reference's own code
(tests/golden/gen_golden.py): hit results must agree bit for bit."""
import numpy as np
import pytest

SCENES = [("cornell555", "cornell555"), ("random_spheres", "random_spheres"),
          ("native_cornell_box", "ref_cornell_box"), ("box_list", "box_list")]


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


@pytest.mark.parametrize("fix,scene", SCENES)
def test_hits_match_reference_golden(oracle, golden, fix, scene):
    s = oracle.scene(scene)
    rays = golden[f"{fix}/rays"]
    leaf, t, rec, mat = s.trace(rays, record=True)
    g_leaf = golden[f"{fix}/leaf"]
    assert s.n_leaves == int(golden[f"{fix}/n_leaves"][0])
    assert np.array_equal(leaf, g_leaf)
    hit = g_leaf >= 0
    assert hit.mean() > 0.2
    assert np.array_equal(bits(t[hit]), bits(golden[f"{fix}/t"][hit]))
    assert np.array_equal(bits(rec[:, hit]), bits(golden[f"{fix}/rec"][:, hit]))
    assert np.array_equal(mat, golden[f"{fix}/mat"])


@pytest.mark.parametrize("fix,scene", SCENES)
def test_bvh_topology_matches_reference_golden(oracle, golden, fix, scene):
    s = oracle.scene(scene)
    assert np.array_equal(bits(s.bvh_dump()), bits(golden[f"{fix}/bvh"]))


@pytest.mark.parametrize("cam", ["cornell555", "random_spheres", "reference_main", "odd"])
def test_camera_fields_match_reference_golden(oracle, golden, cam):
    fields = oracle.camera_from_params(golden[f"camera/{cam}/params"])
    assert np.array_equal(bits(fields), bits(golden[f"camera/{cam}/fields"]))


def test_nan_hits_are_covered(golden):
    # The adversarial set must exercise the in-plane (NaN t) rule.
    t = golden["cornell555/t"]
    assert np.isnan(t[golden["cornell555/leaf"] >= 0]).sum() > 10
