"""lfh_path_trace_rays_device_noalloc (device arrays, path state in the
library's static scratch array, no allocation in the call) against
lfh_path_trace_rays (host arrays, path state allocated per call): radiance
bitwise equal on two calls in a row, the second with fewer queries, on a
scene with an environment map, quad lights and glass (every shading branch
that consumes scratch).  Queries above the maximum are rejected."""
import ctypes as C

import numpy as np
import pytest

from lavaframe_amd import lfh, scenes

pytestmark = pytest.mark.gpu
MAX_QUERIES = 1 << 16


def _scene():
    return scenes.instanced(n_meshes=6, n_instances=40, min_tris=100, max_tris=3000, width=64, height=48,
                            hdr_size=(256, 128), seed=9)


def _queries(sc, n, seed):
    rng = np.random.default_rng(seed)
    cam = sc.camera()
    d = rng.normal(size=(n, 3)) * 0.3 + np.array(cam.forward)
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    rays = np.concatenate([np.tile(np.array(cam.position, np.float32), (n, 1)), d], 1).astype(np.float32)
    seeds = rng.integers(0, 2**32, (n, 4), dtype=np.uint64).astype(np.uint32)
    return np.ascontiguousarray(rays), np.ascontiguousarray(seeds)


def test_noalloc_equals_path_trace_rays_on_two_calls():
    import torch
    print("[noalloc] two calls", flush=True)
    sc = _scene()
    ctx = lfh.Context.from_scene(sc)
    lib = ctx.lib
    for n, seed in ((MAX_QUERIES, 1), (20_011, 2)):
        rays, seeds = _queries(sc, n, seed)
        want = ctx.path_trace_rays(rays, seeds)
        d_rays = torch.from_numpy(rays).cuda()
        d_seeds = torch.from_numpy(seeds.view(np.int32)).cuda()
        d_out = torch.full((n, 3), -1.0, dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        code = lib.lfh_path_trace_rays_device_noalloc(ctx.h, C.c_void_p(d_rays.data_ptr()),
                                                      C.c_void_p(d_seeds.data_ptr()), n,
                                                      C.c_void_p(d_out.data_ptr()))
        assert code == 0, lib.lfh_last_error(ctx.h)
        assert lib.lfh_synchronize(ctx.h) == 0, lib.lfh_last_error(ctx.h)
        got = d_out.cpu().numpy()
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (n, np.abs(got - want).max())
        assert (got > 0).any(axis=1).mean() > 0.5
    # more queries than the scratch array holds: a clear error, nothing launched
    big = MAX_QUERIES + 1
    d_big = torch.zeros(big * 6, dtype=torch.float32, device="cuda")
    code = lib.lfh_path_trace_rays_device_noalloc(ctx.h, C.c_void_p(d_big.data_ptr()), C.c_void_p(d_big.data_ptr()),
                                                  big, C.c_void_p(d_big.data_ptr()))
    assert code == -1 and b"maximum" in lib.lfh_last_error(ctx.h)
