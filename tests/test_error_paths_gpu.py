"""Error paths that must leave the library usable: a failed path-state
allocation, a scene that cannot be made resident at serving admission, and a
rejected instance update.  Failures are injected from the host with
lfh_debug_fail_next (the operation reports an out-of-memory error or a
rejected TLAS rebuild; no kernel is launched on bad buffers)."""
import numpy as np
import pytest

from lavaframe_amd import lfh, scenes

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.asarray(a).view(np.uint32)


def test_failed_path_allocation_is_retried_by_the_next_render():
    print("[err] path allocation", flush=True)
    sc = scenes.cornell(48, 40)
    cam = sc.camera()
    ref = lfh.Context.from_scene(sc)
    ref.update(cam, reset=True)
    ref.render_passes(6)
    want = ref.accum()[0]

    ctx = lfh.Context.from_scene(sc)
    ctx.update(cam, reset=True)
    lfh.debug_fail_next(1, 1)   # the 6-pass wavefront needs larger path buffers: that allocation fails
    try:
        with pytest.raises(lfh.LfhError) as e:
            ctx.render_passes(6)
        assert e.value.code == -2   # LFH_E_OOM
    finally:
        lfh.debug_fail_next(1, 0)
    # the failed call left no buffers: the next render allocates them again
    ctx.update(cam, reset=True)
    ctx.render_passes(6)
    assert np.array_equal(_bits(ctx.accum()[0]), _bits(want))
    # a single-pass render after a failure also reallocates
    lfh.debug_fail_next(1, 1)
    try:
        with pytest.raises(lfh.LfhError):
            ctx.render_passes(7)
    finally:
        lfh.debug_fail_next(1, 0)
    ctx.update(cam, reset=True)
    ctx.update(cam)
    ctx.render()
    ref.update(cam, reset=True)
    ref.update(cam)
    ref.render()
    assert np.array_equal(_bits(ctx.accum()[0]), _bits(ref.accum()[0]))


def test_server_rejects_a_job_whose_scene_upload_fails_and_serves_the_rest():
    print("[err] serving admission", flush=True)
    bad = scenes.cornell(40, 32, max_depth=3)
    good = [(scenes.cornell(48, 40), 5), (scenes.cornell(32, 32, max_depth=6), 4)]
    alone = []
    for sc, spp in good:
        srv = lfh.Server(max_jobs=4)
        j = srv.submit_scene(sc, spp)
        srv.run_until_idle()
        o = sc.options()
        alone.append(srv.result(j, o.width, o.height))
        srv.close()
    srv = lfh.Server(max_jobs=4)
    jb = srv.submit_scene(bad, 3)          # admitted first: its upload is the one that fails
    ids = [srv.submit_scene(sc, spp) for sc, spp in good]
    lfh.debug_fail_next(2, 1)
    try:
        srv.run_until_idle(max_steps=100)
    finally:
        lfh.debug_fail_next(2, 0)
    st = srv.status(jb)
    assert st.state == -1
    assert "scene upload" in srv.job_error(jb)
    for (sc, spp), j, a in zip(good, ids, alone):
        assert srv.status(j).state == 2 and srv.job_error(j) == ""
        o = sc.options()
        assert np.array_equal(_bits(srv.result(j, o.width, o.height)), _bits(a))
    srv.release(jb)
    # the scene of the rejected job can be served afterwards
    j = srv.submit_scene(bad, 3)
    srv.run_until_idle(max_steps=100)
    assert srv.status(j).state == 2


def _instanced():
    return scenes.instanced(n_meshes=5, n_instances=30, min_tris=50, max_tris=3000, width=32, height=32,
                            hdr_size=None, seed=6)


def test_rejected_host_tlas_update_leaves_the_scene_unchanged():
    print("[err] host TLAS update", flush=True)
    sc = _instanced()
    ctx = lfh.Context.from_scene(sc)
    rays = scenes.pinhole_rays(192, 96, eye=(0.0, 12.0, -55.0), lookat=(0.0, 3.0, 0.0), fov_deg=50.0)
    before = ctx.trace_closest(rays)
    a = sc.arrays()
    top = int(a["top_level_index"])
    tlas = a["nodes"][top:].copy()
    # corrupt one internal TLAS node: its left child now points into the BLAS
    # part, outside the update range, so its subtree is no longer reached
    internal = np.nonzero(tlas[:, 8] == 0)[0]
    assert len(internal) > 1
    bad = tlas.copy()
    bad[internal[1], 6] = 0.0
    red = a["materials"].copy()
    red[:, 0:3] = (1.0, 0.0, 0.0)
    with pytest.raises(lfh.LfhError):
        ctx.update_instances(transforms=a["transforms"], materials=red, tlas_nodes=bad, first_node=top)
    after = ctx.trace_closest(rays)
    for k in before.dtype.names:
        assert np.array_equal(before[k].view(np.uint32), after[k].view(np.uint32)), k
    # a valid update still goes through (the node slot map survived)
    ctx.update_instances(transforms=a["transforms"], tlas_nodes=tlas, first_node=top)
    again = ctx.trace_closest(rays)
    for k in before.dtype.names:
        assert np.array_equal(before[k].view(np.uint32), again[k].view(np.uint32)), k
    # and the materials of the rejected update were not applied: frames are those of a fresh context
    fresh = lfh.Context.from_scene(sc)
    out = []
    for c in (ctx, fresh):
        c.update(sc.camera(), reset=True)
        c.render_passes(2)
        out.append(c.accum()[0])
    assert np.array_equal(_bits(out[0]), _bits(out[1]))


def test_rejected_device_tlas_rebuild_leaves_the_scene_unchanged():
    print("[err] device TLAS rebuild", flush=True)
    sc = _instanced()
    dev = lfh.Context.from_scene(sc, device_bvh=True)
    rays = scenes.pinhole_rays(192, 96, eye=(0.0, 12.0, -55.0), lookat=(0.0, 3.0, 0.0), fov_deg=50.0)
    before = dev.trace_closest(rays)
    T = sc.arrays()["transforms"].copy()
    T2 = T.copy()
    T2[:, 12] += 5.0   # every instance moves along x
    lfh.debug_fail_next(3, 1)
    try:
        with pytest.raises(lfh.LfhError):
            dev.update_instances(transforms=T2)
    finally:
        lfh.debug_fail_next(3, 0)
    after = dev.trace_closest(rays)
    for k in before.dtype.names:
        assert np.array_equal(before[k].view(np.uint32), after[k].view(np.uint32)), k
    # the same update then succeeds and moves the hits
    dev.update_instances(transforms=T2)
    moved = dev.trace_closest(rays)
    assert not np.array_equal(moved["t"].view(np.uint32), before["t"].view(np.uint32))
