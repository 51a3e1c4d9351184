#!/usr/bin/env python3
"""Benchmark: closest-hit traversal throughput on the 1M-triangle microbenchmark.

Workload (one GPU): 1,048,576 equilateral triangles (centres U[0,1]^3, random
rotation, edge U(0.002, 0.02)), binned-SAH BVH; 16,777,216 rays = 8,388,608
pinhole rays at 4096x2048 + 8,388,608 incoherent rays; closest hit (prim,
instance, t, u, v) per ray.  A step is one traversal launch over all rays,
resident in HBM.  Prints one JSON line.

Extra fields: any-hit throughput, coherent / incoherent split, traversal
statistics, the roofline of the traversal kernel (algorithmic bytes per ray =
32 + 24 + 112 N_int + 48 N_tri + 64 N_inst over the measured kernel time:
ray in, hit out, the six child-box rows and the child references of each
BVH4 node visited, each triangle and instance record tested), against the
nominal 8 TB/s and a measured device copy, and the CPU baseline (the oracle's
FP32 restatement of ClosestHit on a bounded sample of the same rays, on all
host cores).  The path-tracing legs (Cornell, instanced scene) report
Msamples/s and Mrays/s, the RMSE of their timed frames against the FP64
oracle's 65,536-spp references in tests/golden, and their own CPU baselines
(the oracle's FP32 path tracer on a bounded pixel subset); the serving leg
reports p50 / p99 job latency against the same stream served one job at a
time.
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
NODE_BYTES = 112   # BVH4 node bytes read per visit: 6 float4 of child boxes + 1 of child refs
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
    ap.add_argument("--legs", default="c0,c2,c4",
                    help="extra path-tracing measurements reported beside the traversal line: c0 (Cornell "
                         "Msamples/s), c2 (instanced scene Msamples/s and Mrays/s), c4 (serving p50/p99)")
    ap.add_argument("--guiding", action="store_true",
                    help="instead: wall time to RMSE 1e-2, guided / plain, on the guiding room (own JSON line)")
    ap.add_argument("--guide-ref-spp", type=int, default=2048, help="spp of each reference half")
    ap.add_argument("--guide-max-spp", type=int, default=4096)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); rays shard with no exchange "
              "across GPUs, so N GPUs run N independent copies of this workload.", file=sys.stderr)
        sys.exit(2)

    import torch
    from lavaframe_amd import lfh, scenes

    dev = torch.device("cuda:0")
    torch.cuda.init()
    if args.guiding:
        print(json.dumps(guiding_leg(args)), flush=True)
        return
    tris = scenes.soup_triangles(args.tris, seed=0)
    s = lfh.Scene()
    s.add_material(lfh.MATERIAL_DEFAULT)
    s.add_instance(s.add_mesh(tris), scenes.affine(), 0)
    t0 = time.perf_counter()
    s.build()
    t_build = time.perf_counter() - t0
    log(f"scene: {args.tris} tris, host BVH build {t_build:.2f}s")
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
        bytes_per_ray = 32 + 24 + NODE_BYTES * n_int + 48 * n_tri + 64 * n_inst
        extra["device_bvh"] = device_bvh_leg(s, stream, ptr_r, n, d_hits, args.steps, t_build)
    else:
        bytes_per_ray = float("nan")
    achieved = bytes_per_ray * n / (ms_per_step * 1e-3) / 1e9
    traffic = read_traffic(ROOT / "profiles" / "traversal_c1_traffic.txt")
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "bytes_per_ray": round(bytes_per_ray, 1)}
    if not args.profile_only:
        cp = copy_peak(stream)
        roofline["copy_peak"] = round(cp, 1)
        roofline["frac_of_copy_peak"] = round(achieved / cp, 4)
        if traffic:
            roofline["traffic_GBps"] = round(traffic / (ms_per_step * 1e-3) / 1e9, 1)

    cpu = None
    if not args.no_cpu_baseline and not args.profile_only:
        cpu, extra["oracle_parity"] = cpu_baseline(s, rays, n_coh, args.cpu_sample, hits)
    if not args.profile_only and args.legs:
        del d_rays, d_hits, d_occ
        torch.cuda.empty_cache()
        legs = {"c0": cornell_leg, "c2": instanced_leg, "c4": serving_leg}
        for leg in args.legs.split(","):
            try:
                extra.update(legs[leg.strip()](stream, cpu=not args.no_cpu_baseline))
            except Exception as e:  # a failing extra leg must not hide the main measurement
                log(f"leg {leg} failed: {type(e).__name__}: {e}")
                extra[f"{leg}_error"] = f"{type(e).__name__}: {e}"

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


def device_bvh_leg(s, stream, ptr_r, n, d_hits, steps, host_build_s):
    """The same 1M-triangle scene with its BVH built on the GPU (PLOC + BVH4
    collapse): context creation wall time (host checks, upload, build) beside
    the host SAH build, closest-hit rate over the same rays, visit counts, and
    whether every hit equals the host-built tree's."""
    import torch
    from lavaframe_amd import lfh
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ctx = lfh.Context.from_scene(s, device_bvh=True, width=64, height=64, max_depth=1)
    create_s = time.perf_counter() - t0
    ctx.set_stream(stream.cuda_stream)
    d2 = torch.empty_like(d_hits)
    for _ in range(2):
        ctx.trace_closest_device(ptr_r, n, d2.data_ptr())
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(steps):
        ctx.trace_closest_device(ptr_r, n, d2.data_ptr())
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    same = bool(torch.equal(d2, d_hits))
    cnt = torch.empty(n * 3, dtype=torch.int32, device="cuda")
    ctx.trace_closest_stats_device(ptr_r, n, d2.data_ptr(), cnt.data_ptr())
    torch.cuda.synchronize()
    c = cnt.view(n, 3).double().mean(0).tolist()
    ctx.close()
    del d2, cnt
    torch.cuda.empty_cache()
    out = {"create_s": round(create_s, 3), "host_build_s": round(host_build_s, 3),
           "mrays_per_s": round(n / (ms * 1e-3) / 1e6, 1), "ms_per_launch": round(ms, 3),
           "N_int": round(c[0], 2), "N_tri": round(c[1], 2), "hits_identical_to_host_tree": same}
    log(f"device BVH: create {create_s:.3f}s (host build {host_build_s:.2f}s), {out['mrays_per_s']} Mrays/s, "
        f"N_int {c[0]:.2f} N_tri {c[1]:.2f}, identical hits {same}")
    return out


def copy_peak(stream):
    """Measured device-copy bandwidth (GB/s, read + write) of a 2 GiB buffer:
    the practical HBM ceiling beside the nominal 8 TB/s."""
    import torch
    n = 1 << 29
    x = torch.empty(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(2):
        y.copy_(x)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    reps = 10
    e0.record(stream)
    for _ in range(reps):
        y.copy_(x)
    e1.record(stream)
    torch.cuda.synchronize()
    gbs = 2 * 4 * n * reps / (e0.elapsed_time(e1) * 1e-3) / 1e9
    del x, y
    torch.cuda.empty_cache()
    log(f"device copy: {gbs:.1f} GB/s")
    return gbs


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


def cpu_baseline(scene, rays, n_coh, sample, gpu_hits):
    """The oracle's FP32 restatement of ClosestHit (the reference visit order,
    no culling) on all host cores over an evenly strided sample of the rays.
    Also compares its hits with the GPU's on that sample: ids and the bits of
    t, u, v; rays that differ under the reference's first-found tie rule are
    traced again with the lowest-index tie rule the GPU follows."""
    try:
        from oracle import native
    except Exception as e:  # pragma: no cover
        log(f"cpu baseline unavailable: {e}")
        return None, None
    if not native.available():
        log("cpu baseline unavailable: oracle not built")
        return None, None
    half = sample // 2
    idx = np.concatenate([np.linspace(0, n_coh - 1, half).astype(np.int64),
                          np.linspace(n_coh, len(rays) - 1, sample - half).astype(np.int64)])
    sub = np.ascontiguousarray(rays[idx])
    orc = native.OracleScene(scene.arrays())
    threads = native.threads()
    t = time.perf_counter()
    oh = orc.trace_closest(sub, tie_mode=0, nthreads=threads)
    dt = time.perf_counter() - t
    log(f"cpu baseline: {len(sub)} rays in {dt:.2f}s on {threads} threads")
    gh = gpu_hits[idx]

    def differs(a, b):
        d = np.zeros(len(a), bool)
        for k in ("prim", "inst", "mat"):
            d |= a[k] != b[k]
        for k in ("t", "u", "v"):
            d |= a[k].view(np.uint32) != b[k].view(np.uint32)
        return d

    bad0 = np.nonzero(differs(gh, oh))[0]
    bad1 = bad0
    if bad0.size:
        oh1 = orc.trace_closest(np.ascontiguousarray(sub[bad0]), tie_mode=1, nthreads=threads)
        bad1 = bad0[differs(gh[bad0], oh1)]
    parity = {"rays_compared": int(len(sub)), "differ_first_found_ties": int(bad0.size),
              "differ_lowest_index_ties": int(bad1.size), "hits_identical_to_oracle": bool(bad1.size == 0)}
    log(f"oracle parity on the sample: {parity}")
    cpu = {"value": round(len(sub) / dt / 1e6, 3), "unit": "Mrays/s", "cores": threads, "kind": "port",
           "sample": f"{len(sub)} rays ({half} pinhole + {sample - half} incoherent, evenly strided) of the "
                     "same 16M-ray workload; FP32 restatement of ClosestHit in the reference visit order"}
    return cpu, parity


# ---- path-tracing legs ------------------------------------------------------------------
def pt_cpu_baseline(sc, spp, seconds=8.0):
    """The oracle's FP32 restatement of the path tracer (pathtrace.glsl with the
    reference's RNG and BVH visit order) on all host cores, over an evenly
    strided subset of the frame's pixels, sized from a short calibration run to
    take about `seconds`.  Returns the rate in Msamples/s and the sample."""
    try:
        from oracle import native
    except Exception as e:  # pragma: no cover
        log(f"path-tracing cpu baseline unavailable: {e}")
        return None
    if not native.available():
        log("path-tracing cpu baseline unavailable: oracle not built")
        return None
    arr = sc.arrays()
    opts = sc.options()
    o = native.options_from(opts, sc.camera(), lights=len(arr["lights"]) > 0, envmap=bool(opts.use_env_map))
    orc = native.OracleScene(arr)
    threads = native.threads()
    w, h = opts.width, opts.height

    def run(n_px):
        px = np.linspace(0, w * h - 1, n_px).astype(np.int32)
        t = time.perf_counter()
        orc.render_pixels(o, px, 1, spp, precision=32, nthreads=threads)
        return time.perf_counter() - t

    n = 64 * threads
    dt = run(n)
    n = int(min(w * h, max(n, n * seconds / max(dt, 1e-3))))
    dt = run(n)
    samples = n * spp
    log(f"path-tracing cpu baseline: {n} pixels x {spp} spp in {dt:.2f}s on {threads} threads")
    return {"value": round(samples / dt / 1e6, 4), "unit": "Msamples/s", "cores": threads, "kind": "port",
            "sample": f"{n} of {w * h} pixels (evenly strided) x {spp} spp of the same frame; FP32 restatement "
                      "of the reference's path tracer"}


def frame_error(ctx, spp, ref_name):
    """RMSE of the context's accumulated frame (spp samples per pixel) against
    the FP64 oracle's 65,536-spp reference in tests/golden (every pixel, or the
    stored pixel subset), beside the RMSE the two sample counts' noise alone
    predicts from the reference's per-sample deviation."""
    f = ROOT / "tests" / "golden" / ref_name
    if not f.exists():
        return None
    z = np.load(f)
    mean = z["mean"].reshape(-1, 3).astype(np.float64)
    std = z["std"].reshape(-1, 3).astype(np.float64)
    s, c = ctx.accum()
    img = ((s.astype(np.float64) - c) / spp)[..., :3].reshape(-1, 3)
    if "pixels" in z.files:
        img = img[z["pixels"]]
    rmse = float(np.sqrt(np.mean((img - mean) ** 2)))
    noise = float(np.sqrt(np.mean(std ** 2) * (1.0 / spp + 1.0 / int(z["spp"]))))
    return {"pixels": len(mean), "spp": spp, "reference_spp": int(z["spp"]), "rmse": round(rmse, 6),
            "rmse_from_noise": round(noise, 6), "relative_rmse": round(rmse / float(mean.mean()), 5),
            "mean": round(float(img.mean()), 6), "reference_mean": round(float(mean.mean()), 6)}


def _timed_passes(ctx, stream, passes):
    """Renders `passes` full-frame passes; returns GPU milliseconds (HIP events
    on the context's stream)."""
    import torch
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    ctx.render_passes(passes)
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1)


def cornell_leg(stream, cpu=True):
    """Cornell box 256 x 256, 64 samples per pixel, max depth 4 (one tile)."""
    from lavaframe_amd import lfh, scenes
    sc = scenes.cornell(256, 256, max_depth=4)
    ctx = lfh.Context.from_scene(sc)
    ctx.set_stream(stream.cuda_stream)
    cam = sc.camera()
    spp = 64
    ctx.update(cam)
    _timed_passes(ctx, stream, spp)    # warm-up with the timed wavefront's size: its buffers exist before timing
    ctx.update(cam, reset=True)
    ctx.update(cam)
    r0 = ctx.ray_counts()
    ms = _timed_passes(ctx, stream, spp)
    r1 = ctx.ray_counts()
    samples = 256 * 256 * spp
    rays = (r1[0] - r0[0]) + (r1[1] - r0[1])
    log(f"cornell 256^2 x {spp} spp: {ms:.2f} ms, {samples / ms / 1e3:.1f} Msamples/s, {rays / ms / 1e3:.1f} Mrays/s")
    out = {"msamples_per_s": round(samples / (ms * 1e-3) / 1e6, 1),
           "mrays_per_s": round(rays / (ms * 1e-3) / 1e6, 1), "spp": spp,
           "resolution": [256, 256], "max_depth": 4, "ms": round(ms, 2)}
    out["error_vs_fp64_64k"] = frame_error(ctx, spp, "cornell_c0_fp64_64k.npz")
    if cpu:
        out["cpu_baseline"] = pt_cpu_baseline(sc, spp)
    return {"cornell_c0": out}


def instanced_leg(stream, cpu=True):
    """64 procedural meshes (1k-200k triangles), 1024 random affine instances,
    random Disney materials, 8 quad lights + 2048 x 1024 sky; 1920 x 1080,
    max depth 8; 16 samples per pixel timed."""
    from lavaframe_amd import lfh, scenes
    t = time.time()
    sc = scenes.instanced()
    ntri = int(sc.arrays()["vert_indices"].shape[0])
    log(f"instanced scene: {ntri} triangles, build {time.time() - t:.1f}s")
    ctx = lfh.Context.from_scene(sc)
    ctx.set_stream(stream.cuda_stream)
    cam = sc.camera()
    spp = 16
    ctx.update(cam)
    _timed_passes(ctx, stream, spp)    # warm-up: buffers of the timed wavefronts exist before timing
    ctx.update(cam, reset=True)
    r0 = ctx.ray_counts()
    ms = _timed_passes(ctx, stream, spp)
    r1 = ctx.ray_counts()
    samples = 1920 * 1080 * spp
    ext, sh = r1[0] - r0[0], r1[1] - r0[1]
    log(f"instanced 1920x1080 x {spp} spp: {ms:.1f} ms, {samples / ms / 1e3:.1f} Msamples/s, "
        f"{(ext + sh) / ms / 1e3:.1f} Mrays/s")
    out = {"msamples_per_s": round(samples / (ms * 1e-3) / 1e6, 1),
           "mrays_per_s": round((ext + sh) / (ms * 1e-3) / 1e6, 1),
           "extend_rays": ext, "shadow_rays": sh, "spp": spp, "triangles": ntri,
           "instances": 1024, "resolution": [1920, 1080], "max_depth": 8, "ms": round(ms, 1)}
    out["error_vs_fp64_64k"] = frame_error(ctx, spp, "instanced_c2_fp64_64k.npz")
    if cpu:
        del ctx
        out["cpu_baseline"] = pt_cpu_baseline(sc, 1)
    return {"instanced_c2": out}


def _job_stream(n_jobs, rate, seed):
    """Poisson arrivals (rate jobs/s) of jobs drawn from three scene families
    with random resolutions and sample targets."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n_jobs))
    res = [(256, 256), (512, 512), (640, 360), (960, 540)]
    spp = [16, 32, 64, 128]
    jobs = []
    for i in range(n_jobs):
        r = res[rng.choice(4, p=[0.4, 0.3, 0.2, 0.1])]
        jobs.append({"t": float(t[i]), "family": int(rng.integers(0, 3)), "res": r,
                     "spp": spp[rng.choice(4, p=[0.4, 0.3, 0.2, 0.1])]})
    return jobs


def _run_stream(scn, jobs, max_jobs):
    from lavaframe_amd import lfh
    import torch
    srv = lfh.Server(budget_bytes=0, max_jobs=max_jobs)
    ids = {}
    t0 = time.perf_counter()
    nxt = 0
    while True:
        now = time.perf_counter() - t0
        while nxt < len(jobs) and jobs[nxt]["t"] <= now:
            j = jobs[nxt]
            w, h = j["res"]
            sc = scn[j["family"]]
            ids[nxt] = (srv.submit_scene(sc, j["spp"], width=w, height=h, tile_width=w, tile_height=h),
                        time.perf_counter() - t0)
            nxt += 1
        active = srv.step()
        if active == 0:
            if nxt >= len(jobs):
                break
            time.sleep(max(0.0, jobs[nxt]["t"] - (time.perf_counter() - t0)))
    torch.cuda.synchronize()
    total = time.perf_counter() - t0
    lat, samples = [], 0
    for k, (jid, ts) in ids.items():
        st = srv.status(jid)
        lat.append(st.t_done - st.t_submit)
        w, h = jobs[k]["res"]
        samples += w * h * jobs[k]["spp"]
    srv.close()
    lat = np.array(lat)
    return {"p50_s": round(float(np.percentile(lat, 50)), 4), "p99_s": round(float(np.percentile(lat, 99)), 4),
            "aggregate_msamples_per_s": round(samples / total / 1e6, 1), "makespan_s": round(total, 3)}


def serving_leg(stream, cpu=True):
    """Render serving: 60 jobs, Poisson arrivals at 20 jobs/s, from the Cornell,
    instanced (small) and guiding-room generators at 256^2-960x540 and 16-128
    spp; latency = submit to target spp.  Baseline: the same stream with one
    job admitted at a time (so `cpu` is not used here)."""
    from lavaframe_amd import scenes
    scn = [scenes.cornell(256, 256), scenes.instanced(n_meshes=16, n_instances=128, max_tris=50000,
                                                      hdr_size=(512, 256)),
           scenes.aperture_room(target_tris=50000, max_depth=6)]
    jobs = _job_stream(60, 20.0, seed=0)
    batched = _run_stream(scn, jobs, 64)
    serial = _run_stream(scn, jobs, 1)
    log(f"serving: batched {batched}, serial {serial}")
    return {"serving_c4": {"jobs": len(jobs), "rate_jobs_per_s": 20.0, "batched": batched, "serial_baseline": serial}}


BLOCK = 16


def _blocks(a):
    """Means of BLOCK x BLOCK pixel blocks of an (H, W, 3) image."""
    H, W = a.shape[0] // BLOCK * BLOCK, a.shape[1] // BLOCK * BLOCK
    return a[:H, :W].reshape(H // BLOCK, BLOCK, W // BLOCK, BLOCK, 3).mean((1, 3))


def _rmse(img, ref, ref_var):
    """Relative RMSE of the block means against a noisy reference (block
    means), with the reference's own variance removed:
    E[(img - ref)^2] = MSE(img) + Var(ref); normalised by the mean radiance."""
    d = float(np.mean((_blocks(img) - ref) ** 2)) - ref_var
    return float(np.sqrt(max(d, 0.0))) / ref_mean_holder[0]


ref_mean_holder = [1.0]


def _converge(ctx, cam, ref, ref_var, stream, max_spp, target):
    """Renders 1, 2, 4, ... spp checkpoints; returns [(spp, seconds, RMSE)]
    with the wall clock paused while the image is read back."""
    import torch
    ctx.update(cam, reset=True)   # sample counter restarts; render_passes advances it per pass
    pts, spent, spp = [], 0.0, 0
    step = 1
    while spp < max_spp:
        torch.cuda.synchronize()
        t = time.perf_counter()
        ctx.render_passes(step)
        torch.cuda.synchronize()
        spent += time.perf_counter() - t
        spp += step
        img = ctx.output_hdr()
        e = _rmse(img, ref, ref_var)
        pts.append((spp, spent, e))
        if e < target:
            break
        step = spp
    return pts


def _time_to(pts, target):
    """Wall time at which rel-RMSE reaches target: log-log interpolation between
    checkpoints, or extrapolation with RMSE ~ t^-1/2 past the last one."""
    for (s0, t0, e0), (s1, t1, e1) in zip(pts, pts[1:]):
        if e1 <= target <= e0:
            a = (math.log(target) - math.log(e0)) / (math.log(e1) - math.log(e0))
            return math.exp(math.log(t0) + a * (math.log(t1) - math.log(t0))), False
    s, t, e = pts[-1]
    return t * (e / target) ** 2, e > target


GUIDE_FLOP_PER_QUERY = 2 * (3 * 64 * 64 + 64 * 40)   # 64-d encoding -> 3 x 64 ReLU -> 40 outputs


def _guide_mlp_rate(guide, sc, stream, n=1 << 20, reps=10):
    """MFMA rate of the guide's forward on n resident queries (points in the
    scene box, random unit normals and directions), HIP events on the stream."""
    import torch
    g = torch.Generator(device="cpu").manual_seed(3)
    v = sc.arrays()["vertices"][:, :3]
    lo, hi = torch.tensor(v.min(0), dtype=torch.float32), torch.tensor(v.max(0), dtype=torch.float32)
    pos = lo + (hi - lo) * torch.rand(n, 3, generator=g)
    nrm = torch.nn.functional.normalize(torch.randn(n, 3, generator=g), dim=1)
    wo = torch.nn.functional.normalize(torch.randn(n, 3, generator=g), dim=1)
    x = torch.cat([pos, nrm, wo], 1).contiguous().cuda()
    out = torch.empty(n, 40, dtype=torch.float32, device="cuda")
    for _ in range(2):
        guide.forward_device(x.data_ptr(), n, out.data_ptr(), stream.cuda_stream)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        guide.forward_device(x.data_ptr(), n, out.data_ptr(), stream.cuda_stream)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    tflops = GUIDE_FLOP_PER_QUERY * n / (ms * 1e-3) / 1e12
    log(f"guide MLP forward: {n} queries in {ms:.3f} ms, {tflops:.1f} TFLOP/s")
    return {"queries": n, "forward_ms": round(ms, 4), "flop_per_query": GUIDE_FLOP_PER_QUERY,
            "tflops": round(tflops, 1), "frac_of_bf16_dense_peak": round(tflops / 2500.0, 4)}


def guiding_leg(args):
    """Guiding room (about 200k triangles, light seen only through a door of 5 %
    of a wall), 1280 x 720, max depth 10.  Both runs use the consistent
    estimator (options.estimator = 1), which is unbiased; the reference is the
    plain estimator at 2 x --guide-ref-spp samples from a disjoint frame range,
    rendered as two halves whose difference gives the reference's own
    variance.  Error = RMSE of the 16 x 16-pixel block means (80 x 45 blocks)
    with that variance removed, relative to the mean radiance.  Guided: online
    Adam steps on the samples recorded from every path after each pass."""
    import torch
    from lavaframe_amd import lfh, scenes
    stream = torch.cuda.Stream()
    torch.cuda.set_stream(stream)
    sc = scenes.aperture_room()
    cam = sc.camera()
    target = 1e-2
    ref_ctx = lfh.Context.from_scene(sc, estimator=1)
    ref_ctx.set_stream(stream.cuda_stream)
    ref_ctx.set_frame_offset(1 << 24)
    ref_ctx.update(cam, reset=True)
    t = time.perf_counter()
    half = args.guide_ref_spp
    halves = []
    for h in range(2):
        done = 0
        while done < half:
            k = min(64, half - done)
            ref_ctx.render_passes(k)
            done += k
            torch.cuda.synchronize()
            log(f"reference half {h}: {done} spp ({time.perf_counter() - t:.1f}s)")
        s_, c_ = ref_ctx.accum()
        halves.append((s_ - c_).astype(np.float64))
    h1 = _blocks(halves[0] / half)
    h2 = _blocks((halves[1] - halves[0]) / half)
    ref = 0.5 * (h1 + h2)
    ref_var = float(np.mean((h1 - h2) ** 2)) / 4.0
    ref_mean_holder[0] = float(ref.mean())
    t_ref = time.perf_counter() - t
    ref_ctx.close()
    log(f"reference: mean radiance {ref.mean():.5f}, block noise RMS {math.sqrt(ref_var) / ref.mean():.4g} (relative)")
    plain_ctx = lfh.Context.from_scene(sc, estimator=1)
    plain_ctx.set_stream(stream.cuda_stream)
    plain = _converge(plain_ctx, cam, ref, ref_var, stream, args.guide_max_spp, target)
    log(f"plain: {plain}")
    npix = sc.options().width * sc.options().height
    plain_rays = [round(r / (npix * plain[-1][0]), 3) for r in plain_ctx.ray_counts()]
    plain_ctx.close()
    g_ctx = lfh.Context.from_scene(sc, estimator=1, guiding=1)
    g_ctx.set_stream(stream.cuda_stream)
    guide = lfh.Guide.for_scene(sc, seed=0)
    g_ctx.set_guide(guide, train_every=1)
    guided = _converge(g_ctx, cam, ref, ref_var, stream, args.guide_max_spp, target)
    log(f"guided: {guided} (Adam steps {guide.steps()}, samples {g_ctx.guide_samples_seen()})")
    guided_rays = [round(r / (npix * guided[-1][0]), 3) for r in g_ctx.ray_counts()]
    mlp = _guide_mlp_rate(guide, sc, stream)
    tp, xp = _time_to(plain, target)
    tg, xg = _time_to(guided, target)
    return {"metric": "wall time to relative RMSE 1e-2 (16x16 block means), guided / plain (consistent estimator)",
            "value": round(tg / tp, 4), "unit": "ratio", "higher_is_better": False, "n_gpus": 1,
            "guided_s": round(tg, 3), "plain_s": round(tp, 3), "guided_extrapolated": xg, "plain_extrapolated": xp,
            "plain_curve": [[s, round(t, 4), round(e, 6)] for s, t, e in plain],
            "guided_curve": [[s, round(t, 4), round(e, 6)] for s, t, e in guided],
            "reference": {"spp": 2 * half, "seconds": round(t_ref, 1), "mean_radiance": round(float(ref.mean()), 5),
                          "block_noise_rms": round(math.sqrt(ref_var), 6),
                          "block_noise_rms_rel": round(math.sqrt(ref_var) / float(ref.mean()), 6)},
            "rays_per_sample": {"plain": {"extended": plain_rays[0], "shadow": plain_rays[1]},
                                "guided": {"extended": guided_rays[0], "shadow": guided_rays[1]}},
            "ms_per_pass": {"plain": round(1e3 * plain[-1][1] / plain[-1][0], 3),
                            "guided": round(1e3 * guided[-1][1] / guided[-1][0], 3)},
            "guide_mlp": mlp,
            "adam_steps": guide.steps(), "training_samples": g_ctx.guide_samples_seen(),
            "dtype": "f32 (MLP bf16/f32-acc)", "data": "synthetic",
            "config": {"workload": "guiding room ~200k triangles, 1280x720, depth 10, door 5% of wall",
                       "triangles": int(sc.arrays()["vert_indices"].shape[0])}}

if __name__ == "__main__":
    main()
