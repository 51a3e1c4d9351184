"""The preview engine (preview_flareon.glsl, drawn while the camera moves)
against the FP32 oracle pixel by pixel: the preview camera (frame-1 seeds of
the preview pixel, jitter over the screen size, lens samples used only with
USE_DOF), path depth 2, one sample, at preview scales below and above 1.  A
preview leaves the accumulation and the counters of the progressive render
untouched."""
import math

import numpy as np
import pytest

from lavaframe_amd import lfh, scenes
from oracle import native

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("scale,dof", [(1.0, False), (0.5, False), (1.5, False), (0.75, True)])
def test_preview_matches_oracle(scale, dof):
    print(f"[preview] scale {scale} dof {dof}", flush=True)
    sc = scenes.textured_room(96, 64, max_depth=6)
    if dof:
        sc.set_camera((0.5, 0.5, -1.35), (0.5, 0.5, 0.5), 40.0)
    ctx = lfh.Context.from_scene(sc)
    cam = sc.camera()
    if dof:
        cam.aperture = 0.002
        cam.focal_dist = 1.8
    ctx.update(cam, reset=True)
    g = ctx.render_preview(scale, dof=dof)
    ph, pw = g.shape[:2]
    assert (pw, ph) == (int(np.float32(96) * np.float32(scale)), int(np.float32(64) * np.float32(scale)))
    orc = native.OracleScene(sc.arrays())
    o = native.options_from(sc.options(), cam, lights=True, envmap=False)
    o.max_depth = 2
    c = orc.preview(o, pw, ph, dof=dof, precision=32)
    assert np.isfinite(g).all() and g.sum() > 0
    same = np.isclose(g, c, rtol=1e-4, atol=1e-5).all(-1)
    assert same.mean() > 0.99, same.mean()
    n = pw * ph
    se = c.reshape(n, 3).std(0) / math.sqrt(n)
    assert np.all(np.abs(g.reshape(n, 3).mean(0) - c.reshape(n, 3).mean(0)) < 4 * se + 1e-6)


def test_preview_leaves_the_progressive_render_untouched():
    print("[preview] untouched", flush=True)
    sc = scenes.cornell(48, 40)
    cam = sc.camera()
    a = lfh.Context.from_scene(sc)
    b = lfh.Context.from_scene(sc)
    for ctx, preview in ((a, False), (b, True)):
        ctx.update(cam, reset=True)
        ctx.render_passes(3)
        if preview:
            ctx.render_preview(0.5)
            ctx.render_preview(2.0)
        ctx.render_passes(2)
    assert a.sample_count() == b.sample_count() and a.frame() == b.frame()
    sa, sb = a.accum()[0], b.accum()[0]
    assert np.array_equal(sa.view(np.uint32), sb.view(np.uint32))
