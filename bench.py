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


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=2)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--spp", type=int, default=None, help="samples per pixel per step (default: the scene's 2048)")
    p.add_argument("--cpu-spp", type=int, default=256, help="spp of the bounded CPU baseline sample (about 15 s)")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--scene", default="cornell555")
    return p.parse_args()


def cpu_baseline(scene, w, h, spp, depth):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import REF_SO, Oracle, Reference

    cores = os.cpu_count() or 1
    cam = {
        "cornell555": [278, 278, -800, 278, 278, 0, 0, 1, 0, 40, 1, 0, 10, 0, 0],
    }.get(scene)
    if os.path.exists(REF_SO) and cam is not None:
        R = Reference()
        hnd = R.scene(scene)
        t0 = time.perf_counter()
        R.render(hnd, cam, w, h, spp, depth, decorrelate_rows=False)
        dt = time.perf_counter() - t0
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
    return {"value": round(w * h * spp / dt / 1e6, 4), "unit": "Msamples/s", "cores": cores, "kind": kind,
            "sample": f"{scene} {w}x{h} x {spp} spp, depth {depth}: {how}; {dt:.1f} s on {cores} hardware threads "
                      f"(one std::async task per image row)"}


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU; --gpus above 1 is not supported (the path would shard by "
              "pixel tiles / sample ranges with one accumulator reduce, see DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    import torch

    tpt = importlib.import_module("tiny-path-tracer_amd")
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
        sc.render(spp=spp, seed=1000 + i, sum_buf=sum_buf, sumsq_buf=sq_buf)
    torch.cuda.synchronize()
    totals = {"segments": 0, "paths": 0, "trace_ms": 0.0, "shade_ms": 0.0, "trace_launches": 0,
              "shade_launches": 0, "light_queries": 0, "nodes_visited": 0, "nan_samples": 0}
    t0 = time.perf_counter()
    for k in range(args.steps):
        sum_buf.zero_()
        sq_buf.zero_()
        _, _, st = sc.render(spp=spp, seed=k + 1, sum_buf=sum_buf, sumsq_buf=sq_buf, profile=True)
        for key in totals:
            totals[key] += st[key]
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    samples = args.steps * W * H * spp
    value = samples / dt / 1e6
    segs = totals["segments"]
    paths = totals["paths"]
    trace_bytes = TRACE_BYTES_PER_SEGMENT * segs
    shade_bytes = SHADE_BYTES_PER_SEGMENT * segs + ACCUM_BYTES_PER_PATH * paths
    if totals["trace_ms"] >= totals["shade_ms"]:
        kname, kbytes, kms, nl = "k_trace", trace_bytes, totals["trace_ms"], totals["trace_launches"]
    else:
        kname, kbytes, kms, nl = "k_shade", shade_bytes, totals["shade_ms"], totals["shade_launches"]
    achieved = kbytes / (kms / 1e3) / 1e9 if kms > 0 else 0.0
    out = {
        "metric": "Msamples/s (plain GPU tracer, Cornell 1200^2, 2048 spp, depth 50)",
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
        "data": "synthetic (procedural scene: the 555-unit Cornell box built through the scene API)",
        "config": {"workload": f"{args.scene} {W}x{H}, {spp} spp, max depth {depth}, plain estimator",
                   "width": W, "height": H, "spp": spp, "max_depth": depth},
        "roofline": {"bound": "hbm", "kernel": kname, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None,
                     "avg_launch_ms": round(kms / max(nl, 1), 4),
                     "bytes_per_launch": round(kbytes / max(nl, 1))},
        "detail": {"Mrays_per_s": round(segs / dt / 1e6, 2), "segments_per_path": round(segs / max(paths, 1), 3),
                   "light_queries": totals["light_queries"], "nodes_per_segment":
                       round(totals["nodes_visited"] / max(segs, 1), 2), "trace_ms": round(totals["trace_ms"], 1),
                   "shade_ms": round(totals["shade_ms"], 1), "nan_samples": totals["nan_samples"]},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.scene, W, H, args.cpu_spp, depth)
        out["detail"]["x_host_cpu"] = round(value / out["cpu_baseline"]["value"], 1)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
