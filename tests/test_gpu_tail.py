"""The fused tail launch (k_tail) against the bounce loop alone.

Once a pass has handed out all its items and few paths are left, k_tail
carries them through their remaining bounces in one launch (the same trace
and shading functions, the same Philox streams).  Every sample, and so every
sum, must equal the bounce loop's bit for bit, and the render statistics
(segments, paths, light queries, visited slots) must be the same counts.
"""
import pytest

pytestmark = pytest.mark.gpu

COUNTS = ("paths", "segments", "light_queries", "nodes_visited", "nan_samples")


def same(tpt, a, b):
    import torch

    (s0, q0, st0), (s1, q1, st1) = a, b
    assert torch.equal(s0, s1) and torch.equal(q0, q1)
    for k in COUNTS:
        assert st0[k] == st1[k], k


@pytest.mark.parametrize("name,w,spp,cap", [
    ("cornell555", 96, 16, 0),            # small-scene sweep, every item in flight at once
    ("cornell555", 128, 24, 1 << 16),     # several capacity refills before the tail
    ("random_spheres", 80, 12, 0),        # sphere-only walk
    ("random_spheres", 80, 12, 1 << 15),
])
def test_tail_equals_bounce_loop(tpt, device, name, w, spp, cap):
    sc = tpt.Scene.named(name)
    h = w * sc.host.info.height // sc.host.info.width
    kw = dict(width=w, height=h, spp=spp, seed=7, camera=sc.host.camera, queue_capacity=cap)
    a = sc.render(tail=True, profile=True, **kw)
    b = sc.render(tail=False, profile=True, **kw)
    same(tpt, a, b)
    print(f"{name}: {a[2]['tail_segments']} of {a[2]['segments']} segments in {a[2]['tail_launches']} tail launches")
    assert a[2]["tail_launches"] >= 1 and a[2]["tail_segments"] > 0 and a[2]["tail_ms"] > 0
    assert b[2]["tail_launches"] == 0 and b[2]["tail_segments"] == 0


def test_tail_skipped_where_it_does_not_apply(tpt, device):
    # textured scene: no tail, same result either way
    sc = tpt.Scene.named("random_spheres_textured")
    h = 64 * sc.host.info.height // sc.host.info.width
    kw = dict(width=64, height=h, spp=4, seed=3, camera=sc.host.camera)
    a = sc.render(tail=True, **kw)
    assert a[2]["tail_launches"] == 0
    same(tpt, a, sc.render(tail=False, **kw))


@pytest.mark.parametrize("guide_vertices", [2, 50])
def test_guided_tail_hands_back_guided_vertices(tpt, device, guide_vertices):
    # guide_vertices=50: every lambertian vertex uses the network, so tail
    # paths are handed back to the guided shading at each diffuse hit
    import torch

    sc = tpt.Scene.named("cornell555")
    g = tpt.Guide(seed=9)
    kw = dict(width=96, height=96, spp=16, seed=5, camera=sc.host.camera, guide=g,
              guide_vertices=guide_vertices, guiding=4)
    a = sc.render(tail=True, **kw)
    rows_a = g.batch_rows()
    b = sc.render(tail=False, **kw)
    rows_b = g.batch_rows()
    same(tpt, a, b)
    assert torch.equal(rows_a, rows_b)
    assert a[2]["tail_launches"] >= 1
    if guide_vertices == 50:
        assert a[2]["tail_launches"] > 1  # handed-back paths came back to a later tail


@pytest.mark.parametrize("graph_iters", [0, 8])
@pytest.mark.parametrize("guided", [False, True])
def test_pooled_jobs_tail(tpt, device, graph_iters, guided):
    import torch

    hosts = [tpt.HostScene(f"random_spheres_n{n}", seed=k + 1) for k, n in enumerate((60, 400, 3000))]
    pool = tpt.Scene.pool(hosts)
    g = tpt.Guide(seed=2) if guided else None

    def run(flags):
        jobs, bufs = [], []
        for k, (res, spp) in enumerate(((48, 8), (96, 4), (64, 16))):
            j, s, q = pool.job(scene=k, width=res, height=res, spp=spp, seed=20 + k, max_depth=50)
            jobs.append(j)
            bufs.append((s, q))
        L = tpt.Launch()
        L.graph_iters = graph_iters
        L.flags = flags
        st = pool.render_jobs(jobs, launch=L, guide=g)
        return bufs, st

    (ba, sa), (bb, sb) = run(0), run(4)
    for (s0, q0), (s1, q1) in zip(ba, bb):
        assert torch.equal(s0, s1) and torch.equal(q0, q1)
    for k in COUNTS:
        assert sa[k] == sb[k], k
    assert sa["tail_launches"] >= 1 and sb["tail_launches"] == 0
