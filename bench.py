#!/usr/bin/env python3
"""Benchmark: closest-hit traversal throughput on the 1M-triangle microbenchmark.

Workload (one GPU): 1,048,576 equilateral triangles (centres U[0,1]^3, random
rotation, edge U(0.002, 0.02)), binned-SAH BVH; 16,777,216 rays = 8,388,608
pinhole rays at 4096x2048 + 8,388,608 incoherent rays; closest hit (prim,
instance, t, u, v) per ray.  A step is one traversal launch over all rays,
resident in HBM.  Prints one JSON line.

Extra fields: any-hit throughput, coherent / incoherent split, traversal
statistics, the roofline of the traversal kernel (HBM bytes per launch from
the committed rocprofv3 FETCH_SIZE / WRITE_SIZE profile over this run's
launch time, against the nominal 8 TB/s and a measured device copy; beside it
the cache-inclusive algorithmic bytes per ray, 32 + 24 + 112 N_int + 48 N_tri
+ 64 N_inst: ray in, hit out, the six child-box rows and the child references
of each BVH4 node visited, each triangle and instance record tested), whether
the hits equal the oracle's on the CPU-baseline sample, and the CPU baseline
(the oracle's FP32 restatement of ClosestHit on a bounded sample of the same
rays, on all host cores).  The path-tracing legs report: Cornell and the
instanced scene, Msamples/s and Mrays/s, the RMSE of their timed frames
against the FP64 oracle's 65,536-spp references in tests/golden, and their
own CPU baselines (the oracle's FP32 path tracer on a bounded pixel subset);
the guiding room served by the render server, guided / plain wall time to
RMSE 1e-2 against its FP64 reference; the serving stream (60 jobs, Poisson
arrivals), p50 / p99 job latency against the same stream served one job at a
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
    ap.add_argument("--legs", default="c0,c2,c3,c4",
                    help="extra path-tracing measurements reported beside the traversal line: c0 (Cornell "
                         "Msamples/s), c2 (instanced scene Msamples/s and Mrays/s), c3 (guided / plain wall "
                         "time to RMSE 1e-2 against the FP64 reference, served), c4 (serving p50/p99)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); rays shard with no exchange "
              "across GPUs, so N GPUs run N independent copies of this workload.", file=sys.stderr)
        sys.exit(2)

    import torch
    from lavaframe_amd import lfh, scenes

    dev = torch.device("cuda:0")
    torch.cuda.init()
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
    # The traversal's bytes: the algorithmic model charges every BVH4 node
    # visit, triangle test and instance visit its record (cache-inclusive: L2
    # and the MALL serve the re-reads), so it is not an HBM figure; the HBM
    # figure is the PMC-measured traffic per launch (FETCH_SIZE + WRITE_SIZE,
    # calibrated on a copy) over this run's kernel time.
    algo_gbps = bytes_per_ray * n / (ms_per_step * 1e-3) / 1e9
    traffic = read_traffic(ROOT / "profiles" / "traversal_c1_traffic.txt")
    achieved = traffic / (ms_per_step * 1e-3) / 1e9 if traffic else None
    roofline = {"bound": "hbm", "achieved": round(achieved, 1) if achieved else None, "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None, "traffic": traffic,
                "achieved_from": "HBM bytes per launch from rocprofv3 FETCH_SIZE + WRITE_SIZE "
                                 "(profiles/traversal_c1_traffic.txt) over this run's average launch time",
                "algorithmic_bytes_per_ray": round(bytes_per_ray, 1),
                "algorithmic_GBps_cache_inclusive": round(algo_gbps, 1)}
    issue = read_issue(ROOT / "profiles" / "traversal_c1_sq.txt")
    if issue:
        roofline["issue"] = issue
        if issue.get("valu_busy_frac", 0.0) > 0.8:
            # the kernel is bound by VALU issue, not by bytes: HBM frac is the counter fraction
            roofline["limiter"] = (f"VALU issue: SIMD VALU {issue['valu_busy_frac']:.0%} busy "
                                   "(SQ_ACTIVE_INST_VALU, profiles/traversal_c1_sq.txt)")
    if not args.profile_only:
        cp = copy_peak(stream)
        roofline["copy_peak"] = round(cp, 1)
        if achieved:
            roofline["frac_of_copy_peak"] = round(achieved / cp, 4)

    cpu = None
    if not args.no_cpu_baseline and not args.profile_only:
        cpu, extra["oracle_parity"] = cpu_baseline(s, rays, n_coh, args.cpu_sample, hits)
    if not args.profile_only and args.legs:
        del d_rays, d_hits, d_occ
        torch.cuda.empty_cache()
        legs = {"c0": cornell_leg, "c2": instanced_leg, "c3": guided_served_leg, "c4": serving_leg}
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


def read_issue(path):
    """Instruction-issue figures of the traversal kernel from the committed
    rocprofv3 SQ-counter summary (lines `key: value`), or None."""
    if not path.exists():
        return None
    out = {}
    for line in path.read_text().splitlines():
        if line.startswith("issue_") and ":" in line:
            k, v = line.split(":", 1)
            try:
                out[k[len("issue_"):].strip()] = float(v)
            except ValueError:
                pass
    return out or None


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


# Job sizes: resolutions and sample targets over the configured ranges, small
# ones the most frequent (about 17M samples per job on average, so 20 jobs/s
# offer about 345 Msamples/s, some 70 % of what one GPU renders on this mix of
# scenes: a server sized for its load, where queues stay bounded).
C4_RES = [(256, 256), (512, 512), (640, 360), (960, 540), (1280, 720), (1920, 1080)]
C4_RES_P = [0.5, 0.25, 0.1, 0.08, 0.04, 0.03]
C4_SPP = [16, 32, 64, 128, 256, 512, 1024]
C4_SPP_P = [0.45, 0.25, 0.15, 0.08, 0.04, 0.02, 0.01]


def _job_stream(n_jobs, rate, seed):
    """Poisson arrivals (rate jobs/s) of jobs from three scene generators:
    family 0 the Cornell box, family 1 an instanced scene (16 meshes, 128
    instances, 512 x 256 sky) with a random seed, family 2 the guiding room
    (the 200k-triangle room of the guided leg, guided: the server trains one
    guide for the room across its jobs); resolutions 256^2 - 1920 x 1080 and
    targets 16 - 1024 spp drawn from C4_RES / C4_SPP.  Every scene is submitted without BVHs, so each job's
    admission includes its device BVH build (and, for the sky, the device
    environment tables)."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n_jobs))
    jobs = []
    for i in range(n_jobs):
        fam = int(rng.integers(0, 3))
        res = C4_RES[rng.choice(len(C4_RES), p=C4_RES_P)]
        jobs.append({"t": float(t[i]), "family": fam, "res": res, "seed": int(rng.integers(1, 1 << 30)),
                     "spp": C4_SPP[rng.choice(len(C4_SPP), p=C4_SPP_P)]})
    return jobs


def _c4_descs(jobs):
    """Flattened scene descs (no BVH nodes, no environment tables) of the
    stream's jobs; families 0 and 2 are one scene each, family 1 one scene
    per job seed."""
    from lavaframe_amd import scenes
    cache = {}
    out = []
    for j in jobs:
        key = (j["family"], j["seed"] if j["family"] == 1 else 0)
        if key not in cache:
            if j["family"] == 0:
                sc = scenes.cornell(256, 256, build=False)
            elif j["family"] == 1:
                sc = scenes.instanced(n_meshes=16, n_instances=128, max_tris=50000, hdr_size=(512, 256),
                                      seed=j["seed"], build=False)
            else:
                sc = scenes.aperture_room(build=False)
            d = sc.desc()
            d.hdr_marginal = None
            d.hdr_conditional = None
            cache[key] = (sc, d)
        out.append(cache[key])
    return out


def _run_stream(descs, jobs, max_jobs, policy=None):
    from lavaframe_amd import lfh
    import torch
    srv = lfh.Server(budget_bytes=0, max_jobs=max_jobs)
    if policy is not None:
        srv.set_policy(*policy)
    ids = {}
    t0 = time.perf_counter()
    nxt = 0
    while True:
        now = time.perf_counter() - t0
        while nxt < len(jobs) and jobs[nxt]["t"] <= now:
            j = jobs[nxt]
            w, h = j["res"]
            sc, d = descs[nxt]
            o = sc.options()
            o.width, o.height, o.tile_width, o.tile_height = w, h, w, h
            if j["family"] == 2:
                o.estimator, o.guiding = 1, 1   # the guided, served form
            ids[nxt] = srv.submit(d, o, sc.camera(), j["spp"])
            nxt += 1
        active = srv.step()
        if active == 0:
            if nxt >= len(jobs):
                break
            time.sleep(max(0.0, jobs[nxt]["t"] - (time.perf_counter() - t0)))
    torch.cuda.synchronize()
    total = time.perf_counter() - t0
    lat, samples, fam_lat = [], 0, {0: [], 1: [], 2: []}
    for k, jid in ids.items():
        st = srv.status(jid)
        assert st.state == 2, srv.job_error(jid)
        lat.append(st.t_done - st.t_submit)
        fam_lat[jobs[k]["family"]].append(st.t_done - st.t_submit)
        w, h = jobs[k]["res"]
        samples += w * h * jobs[k]["spp"]
    # the images of the jobs that render a reference's scene at its resolution
    results = {k: srv.result(jid, *jobs[k]["res"]) for k, jid in ids.items()
               if (jobs[k]["family"], jobs[k]["res"]) in ((0, (256, 256)), (2, (1280, 720)))}
    srv.close()
    lat = np.array(lat)
    out = {"p50_s": round(float(np.percentile(lat, 50)), 4), "p99_s": round(float(np.percentile(lat, 99)), 4),
           "aggregate_msamples_per_s": round(samples / total / 1e6, 1), "makespan_s": round(total, 3),
           "p50_by_family_s": {str(f): round(float(np.percentile(v, 50)), 4) for f, v in fam_lat.items() if v}}
    return out, results


def _c4_errors(jobs, results):
    """RMSE against the FP64 references of the jobs that render a reference's
    scene and resolution (Cornell 256^2; the guiding room at 1280 x 720, with
    the consistent estimator), over the noise-predicted RMSE of its samples."""
    out = {}
    for fam, name, ref in ((0, "cornell_256", "cornell_c0_fp64_64k.npz"), (2, "guiding_room_720p", "aperture_c3_fp64_64k.npz")):
        f = ROOT / "tests" / "golden" / ref
        if not f.exists():
            continue
        z = np.load(f)
        mean = z["mean"].reshape(-1, 3).astype(np.float64)
        std = z["std"].reshape(-1, 3).astype(np.float64)
        px = z["pixels"] if "pixels" in z.files else None
        ratios = []
        for k, img in results.items():
            if jobs[k]["family"] != fam:
                continue
            v = img.reshape(-1, 3).astype(np.float64)
            if px is not None:
                v = v[px]
            spp = jobs[k]["spp"]
            rmse = float(np.sqrt(np.mean((v - mean) ** 2)))
            pred = float(np.sqrt(np.mean(std ** 2) * (1.0 / spp + 1.0 / int(z["spp"]))))
            ratios.append(rmse / pred)
        if ratios:
            out[name] = {"jobs": len(ratios), "rmse_over_noise_median": round(float(np.median(ratios)), 3),
                         "rmse_over_noise_max": round(float(np.max(ratios)), 3)}
    out["instanced"] = "random-seed scenes: no FP64 reference"
    return out


C4_JOBS = 80


def serving_leg(stream, cpu=True):
    """Render serving: C4_JOBS jobs, Poisson arrivals at 20 jobs/s, from the
    Cornell, instanced (random seeds) and guiding-room generators at
    256^2 - 1920 x 1080 and 16 - 1024 spp; each scene's BVHs are built on the
    device at admission (inside the latency).  Latency = submit to target spp.
    Batched under the server's default step policy (earliest deadline first)
    and under fair share; baseline: the same stream with one job admitted at a
    time (first come, first served), so `cpu` is not used here."""
    jobs = _job_stream(C4_JOBS, 20.0, seed=0)
    descs = _c4_descs(jobs)
    batched, res_b = _run_stream(descs, jobs, 64)
    fair, _ = _run_stream(descs, jobs, 64, policy=(1, 4.5e8, 1e9))
    serial, _ = _run_stream(descs, jobs, 1)
    log(f"serving: batched (deadline order) {batched}")
    log(f"serving: batched (fair share) {fair}")
    log(f"serving: serial {serial}")
    return {"serving_c4": {"jobs": len(jobs), "rate_jobs_per_s": 20.0,
                           "samples": int(sum(j["res"][0] * j["res"][1] * j["spp"] for j in jobs)),
                           "policy": "earliest deadline first, deadline = submit + samples / 4.5e8 s",
                           "batched": batched, "batched_fair_share": fair, "serial_baseline": serial,
                           "error_vs_fp64": _c4_errors(jobs, res_b)}}


def _time_to(pts, target):
    """Wall time at which the error first reaches target: log-log
    interpolation between the checkpoints around the crossing (the first
    checkpoint's time if it is already below), or extrapolation with
    error ~ t^-1/2 past the last one."""
    if pts[0][2] <= target:
        return pts[0][1], False
    for (s0, t0, e0), (s1, t1, e1) in zip(pts, pts[1:]):
        if e1 <= target <= e0:
            e0, e1 = max(e0, 1e-12), max(e1, 1e-12)
            if e0 == e1:
                return t1, False
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


# ---- C3 served: time to RMSE against the FP64 reference ------------------------------
C3_TARGETS = (1e-2, 5e-3, 2.5e-3)


class _ServedRoom:
    """One render server for one estimator, warmed once before any timed run,
    as a serving loop holds it: the scene resident, and the step buffers sized
    by one full step of a job of the same kind on `stand_in` (a scene of the
    same size and depth).  Every timed run then serves the scene as one job
    (guided: the scene's guide restarted from scratch with reset_guides and
    trained online by the job's wavefronts; frame_offset: an independent
    random sequence), records after every server step the GPU wall time so
    far, the samples per pixel and the RMSE of the 1024 reference pixels
    against the FP64 reference (its own noise removed), and is cancelled once
    it reaches the target.  The clock pauses while the image is read back."""

    def __init__(self, sc, guided, stand_in=None):
        from lavaframe_amd import lfh
        self.sc, self.guided = sc, guided
        self.srv = lfh.Server(max_jobs=4)
        w = self.srv.submit_scene(sc, 1, estimator=1)
        self.srv.run_until_idle()
        self.srv.release(w)
        if stand_in is not None:   # after the 1-spp job, which leaves the step buffers small
            o = stand_in.options()
            step_passes = max(1, (1 << 24) // (o.width * o.height))
            w = self.srv.submit_scene(stand_in, step_passes, estimator=1, guiding=1 if guided else 0)
            self.srv.run_until_idle()
            self.srv.release(w)

    def curve(self, px, mean, var_ref, max_spp, target, frame_offset=0):
        import torch
        srv, o = self.srv, self.sc.options()
        if self.guided:
            srv.reset_guides()
        jid = srv.submit_scene(self.sc, max_spp, estimator=1, guiding=1 if self.guided else 0,
                               frame_offset=frame_offset)
        pts, spent = [], 0.0
        while True:
            torch.cuda.synchronize()
            t = time.perf_counter()
            active = srv.step()
            spent += time.perf_counter() - t
            st = srv.status(jid)
            if st.spp_done == 0:
                continue
            img = srv.result(jid, o.width, o.height).reshape(-1, 3)[px].astype(np.float64)
            e = float(np.sqrt(max(float(np.mean((img - mean) ** 2)) - var_ref, 0.0)))
            pts.append((int(st.spp_done), spent, e))
            if e < target or st.state == 2 or active == 0:
                break
        srv.release(jid)
        return pts

    def close(self):
        self.srv.close()


C3_REPEATS = 8    # independent random sequences (frame offsets) per estimator, run to every target
C3_REPEATS_1E2 = 32   # sequences, the first C3_REPEATS included, run to RMSE 1e-2 (the headline ratio)


def guided_served_leg(stream, cpu=True):
    """Guiding room (about 200k triangles, light seen only through a door of 5 %
    of a wall), 1280 x 720, max depth 10, served as one job of the render
    server, plain (BSDF sampling) and guided (the scene's MLP guide trained
    online from scratch, 50/50 one-sample MIS), both with the consistent
    estimator.  Error = RMSE (radiance units) of the 1024 pixels of the FP64
    oracle's 65,536-spp reference in tests/golden, the reference's own noise
    removed.  One run's error curve is noisy (the door's rare bright paths:
    single-run times to 1e-2 spread by +-20 %), so each estimator runs on
    C3_REPEATS_1E2 independent random sequences to RMSE 1e-2, the first
    C3_REPEATS of them on to 2.5e-3; reported: the median wall time to each
    target, the ratio guided / plain, and every run's times."""
    from lavaframe_amd import scenes
    f = ROOT / "tests" / "golden" / "aperture_c3_fp64_64k.npz"
    if not f.exists():
        return {"guided_c3_error": "tests/golden/aperture_c3_fp64_64k.npz missing"}
    z = np.load(f)
    px = z["pixels"]
    mean = z["mean"].astype(np.float64)
    std = z["std"].astype(np.float64)
    var_ref = float(np.mean(std ** 2)) / int(z["spp"])
    sc = scenes.aperture_room()
    stand_in = scenes.aperture_room(seed=1)   # same size and depth, other furniture: its own guide
    target = min(C3_TARGETS)
    runs = {"plain": [], "guided": []}
    rooms = {kind: _ServedRoom(sc, kind == "guided", stand_in) for kind in runs}
    for r in range(max(C3_REPEATS, C3_REPEATS_1E2)):
        for kind in ("plain", "guided"):
            runs[kind].append(rooms[kind].curve(px, mean, var_ref, 8192,
                                                target if r < C3_REPEATS else max(C3_TARGETS), frame_offset=r << 20))
    for room in rooms.values():
        room.close()
    out = {"targets": {}, "repeats": C3_REPEATS, "repeats_1e-2": C3_REPEATS_1E2,
           "metric": "median wall time to RMSE vs FP64 reference (1024 pixels), guided / plain, served",
           "reference": {"spp": int(z["spp"]), "pixels": len(px), "mean_radiance": round(float(mean.mean()), 6),
                         "noise_rms": round(math.sqrt(var_ref), 7)}}
    for tg in C3_TARGETS:
        n = C3_REPEATS_1E2 if tg == max(C3_TARGETS) else C3_REPEATS
        tp = [_time_to(c, tg) for c in runs["plain"][:n]]
        tgd = [_time_to(c, tg) for c in runs["guided"][:n]]
        mp = float(np.median([t for t, _ in tp]))
        mg = float(np.median([t for t, _ in tgd]))
        out["targets"][f"{tg:g}"] = {"plain_s": round(mp, 4), "guided_s": round(mg, 4), "ratio": round(mg / mp, 4),
                                     "plain_runs_s": [round(t, 4) for t, _ in tp],
                                     "guided_runs_s": [round(t, 4) for t, _ in tgd],
                                     "extrapolated": bool(any(x for _, x in tp + tgd))}
    out["ratio_at_1e-2"] = out["targets"]["0.01"]["ratio"]
    plain, guided = runs["plain"][0], runs["guided"][0]
    out["plain_curve"] = [[s_, round(t, 4), round(e, 6)] for s_, t, e in plain[:: max(1, len(plain) // 24)]]
    out["guided_curve"] = [[s_, round(t, 4), round(e, 6)] for s_, t, e in guided[:: max(1, len(guided) // 24)]]
    out["ms_per_pass"] = {k: round(1e3 * float(np.median([c[-1][1] / c[-1][0] for c in v[:C3_REPEATS]])), 3)
                          for k, v in runs.items()}
    from lavaframe_amd import lfh
    out["guide_mlp"] = _guide_mlp_rate(lfh.Guide.for_scene(sc, seed=0), sc, stream)
    log(f"c3 served: " + ", ".join(f"{k}: {v['plain_s']} / {v['guided_s']} s = {v['ratio']}"
                                   for k, v in out["targets"].items()))
    for k, v in out["targets"].items():
        log(f"c3 runs {k}: plain {v['plain_runs_s']}, guided {v['guided_runs_s']}")
    return {"guided_c3": out}


if __name__ == "__main__":
    main()
