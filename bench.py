"""Benchmark: the plain GPU tracer on the 555-unit Cornell box, 1200x1200,
2048 samples per pixel, max depth 50 (one step = one full frame), against
the reference's own CPU render path timed on this host.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--spp S] [--cpu-spp C] [--no-cpu]

Prints one JSON line (see the keys below).  value = Msamples/s of the whole
job (camera paths per second); inputs (the flattened scene) are resident in
HBM before the timed region.
"""
import argparse
import importlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# Algorithmic HBM bytes per closest-hit segment, split by kernel: k_trace
# reads the ray (24 B) and its queue entry (4 B) and writes the hit (t, slot,
# entry, material: 16 B) and one material-queue entry (4 B); k_shade reads the
# hit (16 B), the ray (24 B), path state (throughput 12, radiance 12, pixel,
# sample, depth 12 B) and its queue entry (4 B), writes the next ray (24 B),
# throughput and radiance (24 B) and the next queue entry (4 B); per finished
# path it adds the accumulator read-modify-write (24 B).
TRACE_BYTES_PER_SEGMENT = 48
SHADE_BYTES_PER_SEGMENT = 116
ACCUM_BYTES_PER_PATH = 24
# Counted VALU work of the closest-hit walk, lane-operations per operation:
# a slab test 18, a sphere test 30 (its FP64 roots apart), a rect test 14, a
# wrapper chain (translate + rotate_y) 25.  Peak: the non-packed FP32 lane
# rate, 256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz.
LANE_OPS = {"box_tests": 18, "sphere_tests": 30, "rect_tests": 14, "xforms": 25}
VALU_PEAK_TLANE_OPS = 39.32
COUNT_SPP = 64


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=2)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--spp", type=int, default=None, help="samples per pixel per step (default: the scene's 2048)")
    p.add_argument("--cpu-spp", type=int, default=512,
                   help="spp of the bounded CPU baseline sample: the reference's render with snapshots every "
                        "--cpu-every spp (about 30 s at 1200x1200); its snapshots give the CPU time to RMSE 0.01")
    p.add_argument("--cpu-every", type=int, default=32)
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--scene", default="cornell555")
    p.add_argument("--no-rmse", action="store_true", help="skip the time-to-RMSE measurement")
    p.add_argument("--ref-spp", type=int, default=65536, help="spp of the converged plain render RMSE is taken against")
    p.add_argument("--workload", default="cornell", choices=["cornell", "batch64", "guide_train", "guided"],
                   help="cornell: the headline frame; batch64: 64 uneven random-sphere renders in one batch; "
                        "guide_train: Adam steps of the guiding network on 2^18 rows; guided: the guided "
                        "estimator's frames after the online training passes")
    p.add_argument("--launch", default=None, metavar="CAP,BLOCKS,GRAPH",
                   help="batch64: use this launch (queue capacity, blocks per CU, graph iterations) instead of "
                        "autotuning")
    p.add_argument("--queue-capacity", type=int, default=0,
                   help="frame workloads: path records in flight (0: the library's automatic choice)")
    p.add_argument("--no-compare", action="store_true",
                   help="batch64: skip the one-call-per-render and guided comparisons")
    p.add_argument("--copy-calibration", action="store_true",
                   help="run one 256 MiB device copy first (the known-size kernel for HBM counter scaling)")
    return p.parse_args()


RMSE_TARGET = 0.01
# Online guided schedule: 3 training passes of 4 spp, each recording every 8th
# path (2 diffuse vertices) and followed by 64 Adam steps of 2^18 rows at
# learning rate 5e-3; then passes of 16 spp with the trained guide.
GUIDED_SCHEDULE = dict(train_passes=3, pass_spp=4, record_stride=8, iters=64, batch=1 << 18, lr=5e-3)
SNAPSHOT_SPP = 16


def rmse_curve(tpt, torch, r, ref, train_passes=0, pass_spp=4, max_spp=8192):
    """Progressive snapshots at doubling spp (SNAPSHOT_SPP, 2x, 4x, ...): RMSE
    of the display image after each pass, with the GPU time of the passes
    (render + training; RMSE evaluation excluded).  Training passes (guided)
    come first.  Returns the points (spp, t, rmse) and the crossing of the
    target (crossing)."""
    t, pts, k = 0.0, [], 0
    target = SNAPSHOT_SPP
    while True:
        train = k < train_passes
        if not train:
            while target <= r.spp:
                target *= 2
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        r.add(pass_spp if train else target - r.spp, train=train)
        b.record()
        b.synchronize()  # this stream only: overlap-mode training may go on beside it
        t += a.elapsed_time(b) / 1e3
        k += 1
        e = tpt.display_rmse(r.sum, r.spp, ref)
        pts.append((r.spp, t, e))
        if (e <= RMSE_TARGET and k >= train_passes) or r.spp >= max_spp:
            break
    return pts, crossing(pts)


def wall_curve(tpt, torch, r, ref, train_passes, pass_spp, stop_spp):
    """As rmse_curve, for a renderer whose training runs on a stream of its
    own beside the passes (ProgressiveRender overlap mode): the time of a
    snapshot is the wall clock from the first pass to the end of its pass,
    host gaps and concurrent training included, and the RMSE of each snapshot
    (a copy of the sums taken on the render stream) is evaluated after the
    last pass, so that no evaluation runs between passes.  Renders to the
    first doubling target at or above stop_spp."""
    start = torch.cuda.Event(enable_timing=True)
    start.record()
    snaps, k, target = [], 0, SNAPSHOT_SPP
    while True:
        train = k < train_passes
        if not train:
            while target <= r.spp:
                target *= 2
        r.add(pass_spp if train else target - r.spp, train=train)
        e = torch.cuda.Event(enable_timing=True)
        e.record()
        snaps.append((r.spp, e, r.sum.clone()))
        k += 1
        if not train and r.spp >= stop_spp:
            break
    r.finish()
    torch.cuda.synchronize()
    pts = [(s, start.elapsed_time(e) / 1e3, tpt.display_rmse(x, s, ref)) for s, e, x in snaps]
    return pts, crossing(pts)


def crossing(pts):
    """Where RMSE crosses the target between bracketing snapshots (spp, t,
    rmse): the samples per pixel n* from RMSE^2 linear in 1/n (the variance of
    an unbiased estimate), then its time, linear in the samples within the
    pass that reaches it.  Returns (t*, n*), or (None, None)."""
    for (s0, t0, e0), (s1, t1, e1) in zip(pts, pts[1:]):
        if e0 > RMSE_TARGET >= e1:
            x0, x1, y0, y1 = 1 / s0, 1 / s1, e0 * e0, e1 * e1
            n = 1 / (x0 + (RMSE_TARGET ** 2 - y0) * (x1 - x0) / (y1 - y0))
            return t0 + (t1 - t0) * (n - s0) / (s1 - s0), n
    return None, None


def time_to_rmse(tpt, torch, sc, ref_spp, cpu):
    W, H = sc.host.info.width, sc.host.info.height
    ref_sum, ref_sq, _ = sc.render(spp=ref_spp, seed=987654321)
    ref = torch.clamp(torch.sqrt(ref_sum / ref_spp), 0.0, 1.0)
    # the reference image's own error (delta method on its per-pixel variance)
    m = ref_sum / ref_spp
    var = torch.clamp(ref_sq / ref_spp - m * m, min=0.0)
    live = (m > 0) & (torch.sqrt(m) < 1)
    floor = torch.sqrt(torch.where(live, var / (4 * m.clamp(min=1e-12)) / ref_spp, torch.zeros_like(m)).mean()).item()
    del ref_sq, var, m
    plain_pts, (t_plain, n_plain) = rmse_curve(tpt, torch, tpt.ProgressiveRender(sc, seed=1), ref)
    g = tpt.Guide(seed=1)
    sch = GUIDED_SCHEDULE
    rg = tpt.ProgressiveRender(sc, guide=g, seed=1, record_stride=sch["record_stride"], iters=sch["iters"],
                               batch=sch["batch"], lr=sch["lr"])
    guided_pts, (t_guided, n_guided) = rmse_curve(tpt, torch, rg, ref, train_passes=sch["train_passes"],
                                                  pass_spp=sch["pass_spp"])
    alpha = float(1 / (1 + np.exp(-g.get_params()[-1])))
    del rg
    # the same schedule at lr 1e-3 (the rate the workload description names;
    # the schedule above uses the tuned 5e-3)
    g3 = tpt.Guide(seed=1)
    r3 = tpt.ProgressiveRender(sc, guide=g3, seed=1, record_stride=sch["record_stride"], iters=sch["iters"],
                               batch=sch["batch"], lr=1e-3)
    _, (t_lr3, n_lr3) = rmse_curve(tpt, torch, r3, ref, train_passes=sch["train_passes"], pass_spp=sch["pass_spp"])
    del r3
    # the same schedule with training on a second stream beside the passes
    go = tpt.Guide(seed=1)
    ro = tpt.ProgressiveRender(sc, guide=go, seed=1, record_stride=sch["record_stride"], iters=sch["iters"],
                               batch=sch["batch"], lr=sch["lr"], overlap=True)
    over_pts, (t_over, n_over) = wall_curve(tpt, torch, ro, ref, sch["train_passes"], sch["pass_spp"],
                                            2 * (n_guided or 256))
    del ro
    r2 = lambda v: round(v, 1) if v else None
    out = {"target": RMSE_TARGET, "on": "clamp(sqrt(mean), 0, 1), all pixels and channels",
           "against": f"plain GPU render, {ref_spp} spp", "reference_rmse_floor": round(floor, 5),
           "plain_s": round(t_plain, 4) if t_plain else None, "plain_spp": r2(n_plain),
           "guided_s": round(t_guided, 4) if t_guided else None, "guided_spp": r2(n_guided),
           "guided_speedup": round(t_plain / t_guided, 3) if t_plain and t_guided else None,
           "guided_schedule": sch, "guided_mix_weight": round(alpha, 3),
           "guided_lr_1e-3_s": round(t_lr3, 4) if t_lr3 else None, "guided_lr_1e-3_spp": r2(n_lr3),
           "overlap_s": round(t_over, 4) if t_over else None, "overlap_spp": r2(n_over),
           "overlap_how": "training on a second stream beside the passes; wall clock, evaluation deferred",
           "snapshots": f"doubling spp from {SNAPSHOT_SPP} (guided: after the training passes)",
           "plain_curve": [[p[0], round(p[1], 4), round(p[2], 5)] for p in plain_pts],
           "guided_curve": [[p[0], round(p[1], 4), round(p[2], 5)] for p in guided_pts],
           "overlap_curve": [[p[0], round(p[1], 4), round(p[2], 5)] for p in over_pts]}
    if cpu is not None:
        # the reference's own render with its progressive snapshots: each
        # snapshot's RMSE, at the time the render took for its samples
        snaps, every, secs = cpu["snapshots"], cpu["every"], cpu["seconds"]
        n = snaps.shape[0]
        pts = []
        for k in range(n):
            s_k = torch.from_numpy(snaps[k]).cuda()
            pts.append(((k + 1) * every, secs * (k + 1) / n, tpt.display_rmse(s_k, (k + 1) * every, ref)))
        t_cpu, n_cpu = crossing(pts)
        out["cpu_s"] = round(t_cpu, 2) if t_cpu else None
        out["cpu_spp"] = r2(n_cpu)
        out["cpu_curve"] = [[p[0], round(p[1], 2), round(p[2], 5)] for p in pts]
        out["cpu_how"] = (f"the reference's own render (oracle/_ref) of {W}x{H} at {n * every} spp with its "
                          f"progressive snapshots every {every} spp (main.cpp:111-133), {secs:.1f} s on "
                          f"{cpu['cores']} hardware threads; a snapshot's time is the render's time for its samples")
        if t_cpu and t_plain:
            out["plain_gpu_x_cpu"] = round(t_cpu / t_plain, 1)
    return out


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(scene, w, h, spp, every, depth):
    """The bounded CPU sample: the reference's own color() and thread-per-row
    loop (oracle/_ref) on all host threads, with its progressive snapshots
    (kept for the CPU time to RMSE), and the same loop one row thread at a
    time for the single-thread rate.  Without the reference build: the CPU
    oracle's restatement (no snapshots)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import REF_SO, Oracle, Reference

    cores = os.cpu_count() or 1
    cam = {
        "cornell555": [278, 278, -800, 278, 278, 0, 0, 1, 0, 40, 1, 0, 10, 0, 0],
    }.get(scene)
    snaps = None
    rate1 = None
    if os.path.exists(REF_SO) and cam is not None:
        R = Reference()
        hnd = R.scene(scene)
        t0 = time.perf_counter()
        snaps = R.render_snapshots(hnd, cam, w, h, spp, every, depth)
        dt = time.perf_counter() - t0
        spp1 = 2
        t0 = time.perf_counter()
        R.render_snapshots(hnd, cam, w, h, spp1, spp1, depth, max_threads=1)
        rate1 = w * h * spp1 / (time.perf_counter() - t0) / 1e6
        kind = "reference"
        how = "the reference's own color() and thread-per-row loop (oracle/_ref, g++ -O3)"
    else:
        O = Oracle()
        s = O.scene(scene)
        t0 = time.perf_counter()
        s.render(w, h, spp, max_depth=depth, rng_mode=1, nthreads=cores)
        dt = time.perf_counter() - t0
        kind = "port"
        how = "the CPU oracle's color() restatement, per-row mt19937"
    out = {"value": round(w * h * spp / dt / 1e6, 4), "unit": "Msamples/s", "cores": cores, "kind": kind,
           "cpu_model": cpu_model(),
           "single_thread_value": round(rate1, 4) if rate1 else None,
           "sample": f"{scene} {w}x{h} x {spp} spp, depth {depth}: {how}; {dt:.1f} s on {cores} hardware threads "
                     f"(one std::async task per image row)" +
                     (f"; single-thread rate from {w}x{h} x 2 spp, one row thread at a time" if rate1 else "")}
    extra = None
    if snaps is not None:
        extra = {"snapshots": snaps, "every": every, "seconds": dt, "cores": cores}
    return out, extra


BATCH_SEED = 2026


def batch64_specs(n=64, seed=BATCH_SEED):
    """The uneven batch: sphere counts log-uniform in [50, 20000], resolution
    uniform in {128, 256, 512, 1024}^2, spp uniform in {4, 16, 64, 256}, depth 50."""
    rng = np.random.default_rng(seed)
    counts = np.round(np.exp(rng.uniform(np.log(50), np.log(20000), n))).astype(int)
    res = rng.choice([128, 256, 512, 1024], n)
    spp = rng.choice([4, 16, 64, 256], n)
    return [dict(spheres=int(c), res=int(r), spp=int(s)) for c, r, s in zip(counts, res, spp)]


def batch64(tpt, torch, args):
    """One step = the whole batch rendered by tpt_render_jobs with the
    autotuned launch; beside it, the same renders one tpt_render call each, and
    the batch with the guide on (one network for every scene, trained online
    on the batch's own rows first)."""
    specs = batch64_specs()
    t0 = time.perf_counter()
    hosts = [tpt.HostScene(f"random_spheres_n{sp['spheres']}", seed=k + 1) for k, sp in enumerate(specs)]
    pool = tpt.Scene.pool(hosts)
    t_build = time.perf_counter() - t0

    def make_jobs(seed):
        jobs, bufs = [], []
        for k, sp in enumerate(specs):
            j, s_, q_ = pool.job(scene=k, width=sp["res"], height=sp["res"], spp=sp["spp"], seed=seed + k,
                                 max_depth=50, sumsq_buf=False)
            jobs.append(j)
            bufs.append(s_)
        return jobs, bufs

    jobs, bufs = make_jobs(1)
    t0 = time.perf_counter()
    if args.launch:
        launch, calib_ms = tpt.Launch(), None
        launch.queue_capacity, launch.blocks_per_cu, launch.graph_iters = (int(x) for x in args.launch.split(","))
    else:
        launch, calib_ms = pool.autotune(jobs, calib_spp=1)
    t_tune = time.perf_counter() - t0
    samples = sum(sp["res"] ** 2 * sp["spp"] for sp in specs)
    for _ in range(args.warmup):
        pool.render_jobs(jobs, launch=launch)
    torch.cuda.synchronize()
    segs = 0
    t0 = time.perf_counter()
    for k in range(args.steps):
        jobs, bufs = make_jobs(100 + 64 * k)
        st = pool.render_jobs(jobs, launch=launch)
        segs += st["segments"]
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    detail = {"Mrays_per_s": round(segs / args.steps / dt / 1e6, 2),
              "launch": launch.as_dict(), "autotuned": not args.launch, "autotune_s": round(t_tune, 2),
              "calibration_ms": None if calib_ms is None else round(calib_ms, 2),
              "scene_build_s": round(t_build, 2)}
    if not args.no_compare:
        detail.update(batch64_compare(tpt, torch, pool, make_jobs, launch, dt, samples))
    value = samples / dt / 1e6
    return {
        "metric": "Msamples/s (batched renders: 64 random-sphere scenes, uneven resolutions and spp, one batch)",
        "value": round(value, 3),
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt * 1e3, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": f"synthetic (64 procedural random-sphere scenes, seed {BATCH_SEED})",
        "config": {"workload": "batch64: 64 random_spheres_n<N> scenes, N log-uniform in [50, 20000], "
                               "res uniform in {128,256,512,1024}^2, spp uniform in {4,16,64,256}, depth 50",
                   "samples": samples, "scenes": [sp["spheres"] for sp in specs]},
        "detail": detail,
    }


def batch64_compare(tpt, torch, pool, make_jobs, launch, dt, samples):
    """The batch's renders one tpt_render_jobs call each, and the batch with
    the guide on."""
    # the same renders one call each (one job per tpt_render_jobs call, the
    # tpt_render path; default launch settings)
    jobs, bufs = make_jobs(3000)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for j in jobs:
        pool.render_jobs([j])
    torch.cuda.synchronize()
    dt_single = time.perf_counter() - t0
    # guided batch: short online training on the batch (1-spp passes recording
    # every 16th path), then the timed batch with the trained guide
    g = tpt.Guide(seed=1)
    gl = tpt.Launch()
    gl.queue_capacity, gl.blocks_per_cu, gl.graph_iters = launch.queue_capacity, launch.blocks_per_cu, launch.graph_iters
    tl = tpt.Launch()
    tl.queue_capacity, tl.blocks_per_cu, tl.graph_iters, tl.guiding = gl.queue_capacity, gl.blocks_per_cu, 0, 16
    t0 = time.perf_counter()
    for r in range(4):
        tj, _ = make_jobs(5000 + 64 * r)
        for j in tj:
            j.spp = 1
        pool.render_jobs(tj, launch=tl, guide=g)
        g.train(None, n_batch=1 << 18, iters=32, lr=5e-3)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    jobs, bufs = make_jobs(7000)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    pool.render_jobs(jobs, launch=gl, guide=g)
    torch.cuda.synchronize()
    dt_guided = time.perf_counter() - t0
    return {"one_call_per_render_ms": round(dt_single * 1e3, 2),
            "batch_speedup_vs_one_call_per_render": round(dt_single / dt, 3),
            "guided_batch_ms": round(dt_guided * 1e3, 2),
            "guided_Msamples_per_s": round(samples / dt_guided / 1e6, 3),
            "guide_training_s": round(t_train, 3)}


MFMA_BF16_PEAK_TFLOPS = 2500.0
TRAIN_FLOP_PER_ROW = 80640  # forward 26,880 + backward data and weights (2x)


def guide_train(tpt, torch, args):
    """One step = one Adam step of the guiding network on 2^18 training rows
    (fused forward/backward, reduction, Adam, tile packing), on synthetic
    rows of the recorded layout, cycling through 2^20 of them."""
    rng = np.random.default_rng(7)
    n = 1 << 20
    rows = np.zeros((n, 16), np.float32)
    rows[:, 0:3] = rng.uniform(0, 1, (n, 3))
    for c in (3, 6, 9):
        v = rng.normal(size=(n, 3))
        rows[:, c:c + 3] = v / np.linalg.norm(v, axis=1, keepdims=True)
    rows[:, 12] = rng.exponential(1.0, n)
    rows[:, 13] = rng.uniform(0.05, 1.0, n)
    rt = torch.from_numpy(rows).cuda()
    g = tpt.Guide(seed=1)
    batch = 1 << 18
    steps = max(args.steps, 1) * 100
    g.train(rt, n_batch=batch, iters=max(args.warmup, 1) * 10, lr=1e-3)
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    loss = g.train(rt, n_batch=batch, iters=steps, lr=1e-3)
    b.record()
    torch.cuda.synchronize()
    ms = a.elapsed_time(b) / steps
    tflops = batch * TRAIN_FLOP_PER_ROW / (ms / 1e3) / 1e12
    return {
        "metric": "training rows/s of the guiding network (Adam steps of 2^18 rows)",
        "value": round(batch / (ms / 1e3) / 1e6, 3),
        "unit": "Mrows/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": max(args.warmup, 1) * 10,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16",
        "data": "synthetic training rows (uniform positions, random unit vectors, exponential contributions)",
        "config": {"workload": "guide_train: 42->3x64->40 vMF-mixture network, batch 2^18 rows, 2^20 rows cycled",
                   "batch": batch},
        "roofline": {"bound": "mfma", "kernel": "k_guide_train + reduce + Adam + pack", "achieved": round(tflops, 2),
                     "peak": MFMA_BF16_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": round(tflops / MFMA_BF16_PEAK_TFLOPS, 5),
                     "traffic": None, "flop_per_row": TRAIN_FLOP_PER_ROW},
        "detail": {"last_loss": round(loss, 5)},
    }


def guided(tpt, torch, args):
    """One step = one frame of `spp` (default 256) samples per pixel with the
    guided estimator, after the online schedule's training passes (not timed:
    the time-to-RMSE figure of the cornell workload includes them)."""
    sc = tpt.Scene.named(args.scene)
    info = sc.host.info
    W, H = info.width, info.height
    spp = args.spp or 256
    sch = GUIDED_SCHEDULE
    g = tpt.Guide(seed=1)
    rg = tpt.ProgressiveRender(sc, guide=g, seed=1, record_stride=sch["record_stride"], iters=sch["iters"],
                               batch=sch["batch"], lr=sch["lr"])
    t0 = time.perf_counter()
    for k in range(sch["train_passes"]):
        rg.add(sch["pass_spp"], train=True)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    sum_buf = torch.zeros((H, W, 3), dtype=torch.float32, device="cuda")
    sq_buf = torch.zeros_like(sum_buf)
    for i in range(args.warmup):
        sc.render(spp=spp, seed=1000 + i, sum_buf=sum_buf, sumsq_buf=sq_buf, guide=g,
                  queue_capacity=args.queue_capacity)
    torch.cuda.synchronize()
    segs = paths = 0
    t0 = time.perf_counter()
    for k in range(args.steps):
        sum_buf.zero_()
        sq_buf.zero_()
        _, _, st = sc.render(spp=spp, seed=k + 1, sum_buf=sum_buf, sumsq_buf=sq_buf, guide=g,
                             queue_capacity=args.queue_capacity)
        segs += st["segments"]
        paths += st["paths"]
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    value = args.steps * W * H * spp / dt / 1e6
    return {
        "metric": f"Msamples/s (guided GPU tracer, {args.scene} {W}x{H}, {spp} spp, depth {info.max_depth})",
        "value": round(value, 3), "unit": "Msamples/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(dt / args.steps * 1e3, 2), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 (network bf16)",
        "data": "synthetic (procedural scene built through the scene API; the guide trained online on its own rows)",
        "config": {"workload": f"{args.scene} {W}x{H}, {spp} spp, guided estimator (2 guided vertices)",
                   "schedule": sch},
        "detail": {"Mrays_per_s": round(segs / dt / 1e6, 2), "segments_per_path": round(segs / max(paths, 1), 3),
                   "training_s": round(t_train, 4),
                   "guided_mix_weight": round(float(1 / (1 + np.exp(-g.get_params()[-1]))), 3)},
    }


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU; --gpus above 1 is not supported (the path would shard by "
              "pixel tiles / sample ranges with one accumulator reduce)", file=sys.stderr)
        sys.exit(2)
    import torch

    tpt = importlib.import_module("tiny-path-tracer_amd")
    if args.workload == "batch64":
        print(json.dumps(batch64(tpt, torch, args)), flush=True)
        return
    if args.workload == "guide_train":
        print(json.dumps(guide_train(tpt, torch, args)), flush=True)
        return
    if args.workload == "guided":
        print(json.dumps(guided(tpt, torch, args)), flush=True)
        return
    if args.copy_calibration:
        src = torch.empty(256 << 20, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    hs = tpt.HostScene(args.scene)
    info = hs.info
    W, H, depth = info.width, info.height, info.max_depth
    spp = args.spp or info.spp
    sc = tpt.Scene(hs)
    sum_buf = torch.zeros((H, W, 3), dtype=torch.float32, device="cuda")
    sq_buf = torch.zeros_like(sum_buf)
    for i in range(args.warmup):
        sum_buf.zero_()
        sq_buf.zero_()
        sc.render(spp=spp, seed=1000 + i, sum_buf=sum_buf, sumsq_buf=sq_buf, queue_capacity=args.queue_capacity)
    torch.cuda.synchronize()
    totals = {"segments": 0, "paths": 0, "trace_ms": 0.0, "shade_ms": 0.0, "trace_launches": 0,
              "shade_launches": 0, "light_queries": 0, "nodes_visited": 0, "nan_samples": 0, "tail_segments": 0,
              "tail_ms": 0.0, "tail_launches": 0}
    t0 = time.perf_counter()
    for k in range(args.steps):
        sum_buf.zero_()
        sq_buf.zero_()
        _, _, st = sc.render(spp=spp, seed=k + 1, sum_buf=sum_buf, sumsq_buf=sq_buf, profile=True,
                             queue_capacity=args.queue_capacity)
        for key in totals:
            totals[key] += st[key]
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    samples = args.steps * W * H * spp
    value = samples / dt / 1e6
    segs = totals["segments"]
    paths = totals["paths"]
    # the bounce-loop kernels' share: the last paths of each pass are traced and
    # shaded by the fused tail launch (k_tail), timed apart
    loop_segs = segs - totals["tail_segments"]
    trace_bytes = TRACE_BYTES_PER_SEGMENT * loop_segs
    shade_bytes = SHADE_BYTES_PER_SEGMENT * loop_segs + ACCUM_BYTES_PER_PATH * paths
    if totals["trace_ms"] >= totals["shade_ms"]:
        kname, kbytes, kms, nl = "k_trace", trace_bytes, totals["trace_ms"], totals["trace_launches"]
    else:
        kname, kbytes, kms, nl = "k_shade", shade_bytes, totals["shade_ms"], totals["shade_launches"]
    achieved = kbytes / (kms / 1e3) / 1e9 if kms > 0 else 0.0
    # measured HBM bytes per segment of this kernel (two rocprofv3 counter passes,
    # scaled by a known-size copy): profiles/cornell555_1200_plain_traffic.md
    traffic = None
    tfile = os.path.join(ROOT, "profiles", f"{args.scene}_{W}_plain_traffic.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            tk = json.load(f)["kernels"].get(kname)
        if tk:
            traffic = round(tk["hbm_bytes_per_segment"] * loop_segs / max(nl, 1))
    out = {
        "metric": ("Msamples/s (plain GPU tracer, Cornell 1200^2, 2048 spp, depth 50)" if args.scene == "cornell555"
                   else f"Msamples/s (plain GPU tracer, {args.scene} {W}x{H}, {spp} spp, depth {depth})"),
        "value": round(value, 3),
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt / args.steps * 1e3, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": ("synthetic (procedural scene: the 555-unit Cornell box built through the scene API)"
                 if args.scene == "cornell555" else f"synthetic (procedural scene {args.scene}, built through the scene API)"),
        "config": {"workload": f"{args.scene} {W}x{H}, {spp} spp, max depth {depth}, plain estimator",
                   "width": W, "height": H, "spp": spp, "max_depth": depth},
        "roofline": {"bound": "hbm", "kernel": kname, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                     "avg_launch_ms": round(kms / max(nl, 1), 4),
                     "bytes_per_launch": round(kbytes / max(nl, 1))},
        "detail": {"Mrays_per_s": round(segs / dt / 1e6, 2), "segments_per_path": round(segs / max(paths, 1), 3),
                   "light_queries": totals["light_queries"], "nodes_per_segment":
                       round(totals["nodes_visited"] / max(segs, 1), 2), "trace_ms": round(totals["trace_ms"], 1),
                   "shade_ms": round(totals["shade_ms"], 1), "nan_samples": totals["nan_samples"],
                   "tail_segments": totals["tail_segments"], "tail_ms": round(totals["tail_ms"], 1),
                   "segments": segs, "tail_launches": totals["tail_launches"],
                   # BVH / primitive bytes the walk requests: one 32-byte slot per visit (boxes and
                   # leaves alike), served from L1/L2 (the Cornell box's slots are 576 bytes)
                   "bvh_slot_bytes_per_segment": round(32 * totals["nodes_visited"] / max(segs, 1), 1),
                   # (over the trace launches and the fused tail launches, which also walk)
                   "bvh_requested_GBps": round(32 * totals["nodes_visited"] /
                                               ((totals["trace_ms"] + totals["tail_ms"]) / 1e3) / 1e9, 1)
                   if totals["trace_ms"] else None},
    }
    # counted VALU work of the walk: one untimed render of the same frame at
    # COUNT_SPP samples per pixel with the counting trace kernel (every segment
    # through k_trace), per segment, times the timed run's bounce-loop segments
    _, _, cst = sc.render(spp=COUNT_SPP, seed=4242, sum_buf=sum_buf, sumsq_buf=sq_buf, count_ops=True)
    csegs = max(cst["segments"], 1)
    per_seg = {k: cst[k] / csegs for k in LANE_OPS}
    lane_ops_per_seg = sum(LANE_OPS[k] * v for k, v in per_seg.items())
    valu = {"bound": "valu", "kernel": "k_trace", "unit": "Tlane-ops/s", "peak": VALU_PEAK_TLANE_OPS,
            "counted_lane_ops_per_segment": round(lane_ops_per_seg, 2),
            "ops_per_segment": {k: round(v, 4) for k, v in per_seg.items()},
            "weights": LANE_OPS}
    if totals["trace_ms"] > 0:
        ach = lane_ops_per_seg * loop_segs / (totals["trace_ms"] / 1e3) / 1e12
        valu.update(achieved=round(ach, 3), frac=round(ach / VALU_PEAK_TLANE_OPS, 4))
    # executed VALU lane-operations per segment of the same kernel, from the
    # gfx950 counters (SQ_INSTS_VALU x active lanes) of a committed profile
    vfile = os.path.join(ROOT, "profiles", f"{args.scene}_{W}_plain_valu.json")
    if os.path.exists(vfile) and totals["trace_ms"] > 0:
        with open(vfile) as f:
            vk = json.load(f)["kernels"].get("k_trace")
        if vk:
            ex = vk["executed_lane_ops_per_segment"] * loop_segs / (totals["trace_ms"] / 1e3) / 1e12
            valu.update(executed_lane_ops_per_segment=vk["executed_lane_ops_per_segment"],
                        executed=round(ex, 3), executed_frac=round(ex / VALU_PEAK_TLANE_OPS, 4),
                        executed_source=f"profiles/{args.scene}_{W}_plain_valu.json")
    out["roofline"]["valu"] = valu
    cpu_extra = None
    if not args.no_cpu:
        out["cpu_baseline"], cpu_extra = cpu_baseline(args.scene, W, H, args.cpu_spp, args.cpu_every, depth)
        out["detail"]["x_host_cpu"] = round(value / out["cpu_baseline"]["value"], 1)
    if not args.no_rmse:
        del sum_buf, sq_buf
        out["time_to_rmse"] = time_to_rmse(tpt, torch, sc, args.ref_spp, cpu_extra)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
