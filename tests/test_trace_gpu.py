"""GPU traversal parity: the HIP closest-hit / any-hit kernels against the CPU
restatement of ClosestHit / AnyHit (closest_hit.glsl, anyhit.glsl) on the
same flattened BVH and rays.  Bar: prim / instance / material ids and the
bits of t, u, v identical (ties to the lowest (instance, prim))."""
import numpy as np
import pytest

from lavaframe_amd import lfh, scenes
from oracle import native

pytestmark = pytest.mark.gpu


def aimed_rays(scene, n, seed=0):
    """Rays from random origins towards random triangle centroids."""
    v = scene.arrays()["vertices"][:, :3].reshape(-1, 3, 3)
    rng = np.random.default_rng(seed)
    tgt = v[rng.integers(0, len(v), n)].mean(1)
    r = scenes.incoherent_rays(n, seed=seed + 1)
    d = tgt - r["org"]
    r["dir"] = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    return r


def _ctx(scene):
    arr = scene.arrays()
    desc, keep = lfh.desc_from_arrays(arr)
    opts = scene.options()
    return lfh.Context(desc, opts), arr


def _compare(gh, oh):
    assert gh.shape == oh.shape
    for k in ("prim", "inst", "mat"):
        bad = np.nonzero(gh[k] != oh[k])[0]
        assert bad.size == 0, f"{k} differs on {bad.size} rays, first {bad[:5]}: {gh[bad[:3]]} vs {oh[bad[:3]]}"
    for k in ("t", "u", "v"):
        g = gh[k].view(np.uint32)
        o = oh[k].view(np.uint32)
        bad = np.nonzero(g != o)[0]
        assert bad.size == 0, f"{k} bits differ on {bad.size} rays: {gh[bad[:3]]} vs {oh[bad[:3]]}"


@pytest.mark.parametrize("ntri,nray", [(1, 4096), (10, 20000), (10000, 200000)])
def test_soup_closest_and_any(ntri, nray):
    print(f"[trace] soup ntri={ntri} nray={nray}", flush=True)
    sc = scenes.soup(ntri, seed=ntri)
    ctx, arr = _ctx(sc)
    orc = native.OracleScene(arr)
    rays = np.concatenate([scenes.incoherent_rays(nray // 2, seed=3),
                           scenes.pinhole_rays(64, nray // 4 // 64), aimed_rays(sc, nray // 4, seed=6)])
    gh = ctx.trace_closest(rays)
    oh = orc.trace_closest(rays, tie_mode=1)
    _compare(gh, oh)
    assert (gh["prim"] >= 0).sum() > nray // 20
    occ = ctx.trace_any(rays)
    oocc = orc.trace_any(rays)
    assert np.array_equal(occ, oocc)
    # any-hit must agree with closest-hit over the same interval
    assert np.array_equal(occ.astype(bool), gh["prim"] >= 0)


def test_cornell_instanced_closest():
    print("[trace] cornell instances", flush=True)
    sc = scenes.cornell()
    ctx, arr = _ctx(sc)
    orc = native.OracleScene(arr)
    rays = np.concatenate([scenes.incoherent_rays(50000, seed=5),
                           scenes.pinhole_rays(256, 256, eye=(0.5, 0.5, -1.35), lookat=(0.5, 0.5, 0.5))])
    gh = ctx.trace_closest(rays)
    oh = orc.trace_closest(rays, tie_mode=1)
    _compare(gh, oh)
    assert set(np.unique(gh["inst"])) >= {0, 1, 2, 3, 4}


def test_tmin_tmax_window():
    print("[trace] tmin/tmax", flush=True)
    sc = scenes.soup(2000, seed=7)
    ctx, arr = _ctx(sc)
    orc = native.OracleScene(arr)
    rays = scenes.incoherent_rays(50000, seed=9)
    rng = np.random.default_rng(0)
    rays["tmin"] = rng.uniform(0, 0.3, len(rays)).astype(np.float32)
    rays["tmax"] = rng.uniform(0.3, 1.0, len(rays)).astype(np.float32)
    _compare(ctx.trace_closest(rays), orc.trace_closest(rays, tie_mode=1))
    assert np.array_equal(ctx.trace_any(rays), orc.trace_any(rays))


def test_empty_batch_and_device_entry():
    import torch
    print("[trace] empty + device", flush=True)
    sc = scenes.soup(500, seed=11)
    ctx, arr = _ctx(sc)
    assert ctx.trace_closest(np.zeros(0, lfh.RAY_DTYPE)).shape == (0,)
    rays = scenes.incoherent_rays(10000, seed=12)
    d_rays = torch.from_numpy(rays.view(np.uint8)).cuda()
    d_hits = torch.empty(len(rays) * 24, dtype=torch.uint8, device="cuda")
    ctx.trace_closest_device(d_rays.data_ptr(), len(rays), d_hits.data_ptr())
    ctx.synchronize()
    gh = d_hits.cpu().numpy().view(lfh.HIT_DTYPE)
    _compare(gh, ctx.trace_closest(rays))


def test_small_and_odd_batches_trace_every_ray():
    """Ray ids are fetched from 8 contiguous slices of the batch, each with its
    own counter: batches smaller than the slice count (slices of 0 or 1 ray)
    and sizes just around a wave's width must still trace every ray once."""
    print("[trace] small batches", flush=True)
    sc = scenes.soup(2000, seed=21)
    ctx, arr = _ctx(sc)
    orc = native.OracleScene(arr)
    rays = aimed_rays(sc, 1100, seed=22)
    for n in (1, 2, 5, 7, 8, 9, 63, 64, 65, 127, 129, 1031):
        r = np.ascontiguousarray(rays[:n])
        _compare(ctx.trace_closest(r), orc.trace_closest(r, tie_mode=1))
        assert np.array_equal(ctx.trace_any(r), orc.trace_any(r)), n


def _leaving_rays(ctx, n, seed):
    """Continuation-like rays: from the hit points of random rays, offset
    1e-3 along a new random direction (pathtrace.glsl:291), many of them
    grazing the axis-aligned walls they start on."""
    rng = np.random.default_rng(seed)
    r = scenes.incoherent_rays(n, seed=seed)
    r["org"] = rng.random((n, 3)).astype(np.float32)
    h = ctx.trace_closest(r)
    hit = h["prim"] >= 0
    p = r["org"][hit] + r["dir"][hit] * h["t"][hit][:, None]
    m = int(hit.sum())
    d = rng.normal(size=(m, 3))
    d[: m // 2, 1] *= 1e-4          # half of them nearly parallel to the floor / ceiling
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    out = scenes.incoherent_rays(m, seed=seed + 1)
    out["dir"] = d
    out["org"] = (p + d * 1e-3).astype(np.float32)
    return out


@pytest.mark.parametrize("device_bvh", [False, True])
def test_rays_leaving_flat_walls_match_oracle(device_bvh):
    """A ray that leaves the plane of an axis-aligned wall at distance 0 does
    not enter that wall's flat box, so it cannot hit the wall at t = -0 (the
    reference's AABBIntersect > 0 rule): the BVH4 leaves split into
    one-triangle children and the device-built tree box triangles by their
    exact vertex bounds, and hit what the reference traversal hits."""
    print(f"[trace] leaving rays device_bvh={device_bvh}", flush=True)
    sc = scenes.cornell()
    ctx = lfh.Context.from_scene(sc, device_bvh=device_bvh)
    orc = native.OracleScene(sc.arrays())
    rays = _leaving_rays(ctx, 600_000, seed=11)
    gh = ctx.trace_closest(rays)
    _compare(gh, orc.trace_closest(rays, tie_mode=1))
    assert not (gh["t"][gh["prim"] >= 0] <= 0).any()
