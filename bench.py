#!/usr/bin/env python3
"""Benchmark: closest-hit traversal throughput on the 1M-triangle microbenchmark.

Workload (one GPU): 1,048,576 equilateral triangles (centres U[0,1]^3, random
rotation, edge U(0.002, 0.02)), binned-SAH BVH; 16,777,216 rays = 8,388,608
pinhole rays at 4096x2048 + 8,388,608 incoherent rays; closest hit (prim,
instance, t, u, v) per ray.  A step is one traversal launch over all rays,
resident in HBM.  Prints one JSON line.

Extra fields: any-hit throughput, coherent / incoherent split, traversal
statistics, the roofline of the traversal kernel (algorithmic bytes per ray =
32 + 24 + 64 N_int + 48 N_tri + 64 N_inst over the measured kernel time),
and the CPU baseline (the oracle's FP32 restatement of ClosestHit on a
bounded sample of the same rays, on all host cores).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
METRIC = "Mrays/s closest-hit; time-to-RMSE 1e-2 guided vs plain; p50/p99 job latency"


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--tris", type=int, default=1 << 20)
    ap.add_argument("--width", type=int, default=4096)
    ap.add_argument("--height", type=int, default=2048)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=6_000_000, help="rays in the CPU-baseline sample")
    ap.add_argument("--profile-only", action="store_true", help="skip stats / baseline (for rocprof runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); rays shard with no exchange "
              "across GPUs, so N GPUs run N independent copies of this workload.", file=sys.stderr)
        sys.exit(2)

    import torch
    from lavaframe_amd import lfh, scenes

    dev = torch.device("cuda:0")
    torch.cuda.init()
    t0 = time.time()
    tris = scenes.soup_triangles(args.tris, seed=0)
    s = lfh.Scene()
    s.add_material(lfh.MATERIAL_DEFAULT)
    s.add_instance(s.add_mesh(tris), scenes.affine(), 0)
    s.build()
    t_build = time.time() - t0
    log(f"scene: {args.tris} tris, BVH build {t_build:.1f}s")
    ctx = lfh.Context.from_scene(s, width=64, height=64, max_depth=1)
    # a dedicated (non-null) stream: the kernels and the timing events share it
    stream = torch.cuda.Stream()
    torch.cuda.set_stream(stream)
    ctx.set_stream(stream.cuda_stream)

    coh = scenes.pinhole_rays(args.width, args.height)
    inc = scenes.incoherent_rays(args.width * args.height, seed=1)
    rays = np.concatenate([coh, inc])
    n = len(rays)
    n_coh = len(coh)
    d_rays = torch.from_numpy(rays.view(np.uint8)).to(dev)
    d_hits = torch.empty(n * 24, dtype=torch.uint8, device=dev)
    d_occ = torch.empty(n, dtype=torch.uint8, device=dev)
    ptr_r, ptr_h, ptr_o = d_rays.data_ptr(), d_hits.data_ptr(), d_occ.data_ptr()
    log(f"rays: {n} ({n_coh} pinhole + {n - n_coh} incoherent) resident in HBM")

    def step():
        ctx.trace_closest_device(ptr_r, n, ptr_h)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t_wall = time.perf_counter()
    start.record(stream)
    for _ in range(args.steps):
        step()
    end.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_wall
    ms = start.elapsed_time(end)
    ms_per_step = ms / args.steps
    mrays = n / (ms_per_step * 1e-3) / 1e6
    log(f"closest hit: {ms_per_step:.3f} ms/step, {mrays:.1f} Mrays/s (wall {wall:.3f}s)")
    hits = d_hits.cpu().numpy().view(lfh.HIT_DTYPE)
    hit_frac = float((hits["prim"] >= 0).mean())

    extra = {}
    if args.profile_only:
        # a plain copy of known size (256 MiB read + 256 MiB write) calibrates
        # the HBM byte counters of a profiling run
        x = torch.empty(1 << 26, dtype=torch.float32, device=dev)
        y = x.clone()
        torch.cuda.synchronize()
        del x, y
    if not args.profile_only:
        # any hit over the same rays
        for _ in range(2):
            ctx.trace_any_device(ptr_r, n, ptr_o)
        a0 = torch.cuda.Event(enable_timing=True)
        a1 = torch.cuda.Event(enable_timing=True)
        a0.record(stream)
        for _ in range(args.steps):
            ctx.trace_any_device(ptr_r, n, ptr_o)
        a1.record(stream)
        torch.cuda.synchronize()
        any_ms = a0.elapsed_time(a1) / args.steps
        extra["anyhit_mrays_per_s"] = round(n / (any_ms * 1e-3) / 1e6, 1)
        # split by ray class
        for name, lo, hi in (("coherent", 0, n_coh), ("incoherent", n_coh, n)):
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            ctx.trace_closest_device(ptr_r + lo * 32, hi - lo, ptr_h + lo * 24)
            e0.record(stream)
            for _ in range(args.steps):
                ctx.trace_closest_device(ptr_r + lo * 32, hi - lo, ptr_h + lo * 24)
            e1.record(stream)
            torch.cuda.synchronize()
            extra[f"{name}_mrays_per_s"] = round((hi - lo) / (e0.elapsed_time(e1) / args.steps * 1e-3) / 1e6, 1)
        # traversal statistics of this kernel's visit order (for the roofline)
        d_cnt = torch.empty(n * 3, dtype=torch.int32, device=dev)
        ctx.trace_closest_stats_device(ptr_r, n, ptr_h, d_cnt.data_ptr())
        torch.cuda.synchronize()
        cnt = d_cnt.view(n, 3).double()
        n_int, n_tri, n_inst = (float(x) for x in cnt.mean(0).tolist())
        ci = cnt[:n_coh].mean(0).tolist()
        ii = cnt[n_coh:].mean(0).tolist()
        extra["traversal_stats"] = {"N_int": round(n_int, 2), "N_tri": round(n_tri, 2), "N_inst": round(n_inst, 3),
                                    "coherent": [round(x, 2) for x in ci], "incoherent": [round(x, 2) for x in ii]}
        bytes_per_ray = 32 + 24 + 64 * n_int + 48 * n_tri + 64 * n_inst
    else:
        bytes_per_ray = float("nan")
    achieved = bytes_per_ray * n / (ms_per_step * 1e-3) / 1e9
    traffic = read_traffic(ROOT / "profiles" / "traversal_c1_traffic.txt")
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "bytes_per_ray": round(bytes_per_ray, 1)}

    cpu = None
    if not args.no_cpu_baseline and not args.profile_only:
        cpu = cpu_baseline(s, rays, n_coh, args.cpu_sample)

    out = {
        "metric": METRIC,
        "value": round(mrays, 1),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "closest-hit traversal, 1M-triangle soup (binned SAH BVH), 16M rays "
                               f"({args.width}x{args.height} pinhole + {n - n_coh} incoherent)",
                   "triangles": args.tris, "rays": n, "hit_fraction": round(hit_frac, 4)},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    out.update(extra)
    print(json.dumps(out), flush=True)


def read_traffic(path):
    """HBM-side bytes per launch of the traversal kernel, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (counter values scaled by the
    same counters read on a copy of known size); None when absent."""
    if not path.exists():
        return None
    for line in path.read_text().splitlines():
        if line.startswith("hbm_bytes_per_launch:"):
            return int(line.split(":", 1)[1])
    return None


def cpu_baseline(scene, rays, n_coh, sample):
    """The oracle's FP32 restatement of ClosestHit (the reference visit order,
    no culling) on all host cores over an evenly strided sample of the rays."""
    try:
        from oracle import native
    except Exception as e:  # pragma: no cover
        log(f"cpu baseline unavailable: {e}")
        return None
    if not native.available():
        log("cpu baseline unavailable: oracle not built")
        return None
    half = sample // 2
    idx = np.concatenate([np.linspace(0, n_coh - 1, half).astype(np.int64),
                          np.linspace(n_coh, len(rays) - 1, sample - half).astype(np.int64)])
    sub = np.ascontiguousarray(rays[idx])
    orc = native.OracleScene(scene.arrays())
    threads = native.threads()
    t = time.perf_counter()
    orc.trace_closest(sub, tie_mode=0, nthreads=threads)
    dt = time.perf_counter() - t
    log(f"cpu baseline: {len(sub)} rays in {dt:.2f}s on {threads} threads")
    return {"value": round(len(sub) / dt / 1e6, 3), "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": f"{len(sub)} rays ({half} pinhole + {sample - half} incoherent, evenly strided) of the "
                      "same 16M-ray workload; FP32 restatement of ClosestHit in the reference visit order"}


if __name__ == "__main__":
    main()
