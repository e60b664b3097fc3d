"""tpt_render_noalloc (include/tpt.h): the render entry point that allocates
nothing, its state in the library's static device array.  Called through the
C ABI and compared with tpt_render of the same arguments: the same Philox
streams summed in sample order, so the sums must be equal bit for bit,
whatever the pass split (the array holds passes of at most
TPT_NOALLOC_PASS_ITEMS samples) and the queue capacity."""
import ctypes as C

import pytest

pytestmark = pytest.mark.gpu

NOALLOC_PASS_ITEMS = 3 << 19
NOALLOC_PATHS = 1 << 17


def noalloc(tpt, scene, cam, width, height, spp, seed, sample_offset=0, max_depth=50, queue_capacity=0, flags=0,
            guiding=0):
    import torch

    L = tpt.lib()
    L.tpt_render_noalloc.argtypes = [C.c_void_p, C.POINTER(tpt.Camera), C.POINTER(tpt.RenderDesc), C.c_void_p,
                                     C.c_void_p, C.POINTER(tpt.RenderStats), C.c_void_p]
    d = tpt.RenderDesc()
    d.width, d.height, d.spp, d.max_depth, d.seed = width, height, spp, max_depth, seed
    d.sample_offset, d.background = sample_offset, scene.host.info.background
    d.queue_capacity, d.flags, d.guiding = queue_capacity, flags, guiding
    s = torch.zeros((height, width, 3), dtype=torch.float32, device="cuda")
    q = torch.zeros_like(s)
    st = tpt.RenderStats()
    rc = L.tpt_render_noalloc(scene.h, C.byref(cam), C.byref(d), s.data_ptr(), q.data_ptr(), C.byref(st),
                              torch.cuda.current_stream().cuda_stream)
    if rc != 0:
        raise tpt.TptError(L.tpt_last_error().decode())
    torch.cuda.synchronize()
    return s, q, st.as_dict()


def camera_for(tpt, host, width, height):
    i = host.info
    return tpt.camera_make(tuple(i.lookfrom), tuple(i.lookat), (0.0, 1.0, 0.0), i.vfov, width / height,
                           i.aperture, i.focus, i.time0, i.time1)


CASES = [  # (scene, width, height, spp, seed, sample_offset, max_depth); each call smaller than the one before
    ("cornell555", 1024, 1024, 3, 5, 0, 50),       # 3 passes of one layer here, one pass in tpt_render
    ("cornell555", 200, 150, 9, 7, 2, 50),
    ("random_spheres_textured", 96, 64, 4, 3, 1, 50),  # textures: no fused tail
    ("random_spheres", 64, 40, 5, 11, 3, 50),      # sphere-only walk with the fused tail
    ("ref_cornell_box", 24, 24, 6, 2, 0, 12),
]


@pytest.fixture(scope="module")
def scenes(tpt, device):
    return {n: tpt.Scene.named(n) for n in {c[0] for c in CASES}}


def test_noalloc_equals_render_on_successive_calls(tpt, scenes):
    import torch

    for name, w, h, spp, seed, off, depth in CASES:
        sc = scenes[name]
        cam = camera_for(tpt, sc.host, w, h)
        print(f"noalloc {name} {w}x{h} spp={spp} seed={seed}", flush=True)
        s0, q0, st0 = noalloc(tpt, sc, cam, w, h, spp, seed, off, depth)
        s1, q1, st1 = sc.render(width=w, height=h, spp=spp, seed=seed, sample_offset=off, max_depth=depth,
                                camera=cam)
        assert st0["paths"] == w * h * spp == st1["paths"]
        assert st0["segments"] == st1["segments"] and st0["nan_samples"] == st1["nan_samples"]
        assert torch.equal(s0, s1), name
        assert torch.equal(q0, q1), name
        assert float(s0.sum()) > 0.0


def test_noalloc_small_capacity_and_repeat(tpt, scenes):
    """A capacity below the default, then the same call again: the array keeps
    the last call's contents, and nothing of it leaks into the next."""
    import torch

    sc = scenes["cornell555"]
    cam = camera_for(tpt, sc.host, 160, 120)
    ref, refq, _ = sc.render(width=160, height=120, spp=7, seed=9, camera=cam)
    for cap in (1 << 12, 0, 1 << 12):
        s, q, _ = noalloc(tpt, sc, cam, 160, 120, 7, 9, queue_capacity=cap)
        assert torch.equal(s, ref) and torch.equal(q, refq), cap


def test_noalloc_limits_are_refused(tpt, scenes):
    sc = scenes["cornell555"]
    big = 1300
    cam = camera_for(tpt, sc.host, big, big)
    assert big * big > NOALLOC_PASS_ITEMS
    with pytest.raises(tpt.TptError, match="TPT_NOALLOC_PASS_ITEMS"):
        noalloc(tpt, sc, cam, big, big, 1, 1)
    cam = camera_for(tpt, sc.host, 32, 32)
    with pytest.raises(tpt.TptError, match="TPT_NOALLOC_PATHS"):
        noalloc(tpt, sc, cam, 32, 32, 1, 1, queue_capacity=NOALLOC_PATHS * 2)
    with pytest.raises(tpt.TptError, match="flags"):
        noalloc(tpt, sc, cam, 32, 32, 1, 1, flags=1)
    with pytest.raises(tpt.TptError, match="guiding"):
        noalloc(tpt, sc, cam, 32, 32, 1, 1, guiding=8)
    s, _, st = noalloc(tpt, sc, cam, 32, 32, 0, 1)  # no samples: nothing to do
    assert st["paths"] == 0 and float(s.abs().sum()) == 0.0
