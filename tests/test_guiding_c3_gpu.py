"""The guiding room (about 200k triangles, light seen only through a door of 5%
of a wall, 1280 x 720, depth 10) against the FP64 oracle's 65,536-spp
reference of the consistent estimator on 1024 pixels (one per cell of a
32 x 32 grid), written by tests/golden/make_fp64_references.py.

* Plain BSDF sampling (options.estimator = 1): per pixel and channel the
  difference over its standard error stays within 4.5 for 99% of the values,
  and the RMSE is what the two noises predict (at most 25% above it; the
  room's rare bright paths through the door put most of the variance in a few
  samples, so a typical frame falls below the prediction).
* Guided (the MLP trained online from the first pass, 50/50 one-sample MIS
  with the BSDF): no bias (the mean over the 1024 pixels within 4 standard
  errors of the reference's) and per-value z-scores within 4.5 for 99% of the
  values, against twice the plain estimator's variance (the defensive mixture
  at most doubles it).
* What the guide gains: at equal samples per pixel the guided frame's RMSE
  against the reference is below 0.8 of the plain frame's.
"""
from pathlib import Path

import numpy as np
import pytest

from lavaframe_amd import lfh, scenes

pytestmark = pytest.mark.gpu
GOLD = Path(__file__).resolve().parent / "golden" / "aperture_c3_fp64_64k.npz"
SPP = 256


@pytest.fixture(scope="module")
def room():
    if not GOLD.exists():
        pytest.skip("aperture_c3_fp64_64k.npz not generated (tests/golden/make_fp64_references.py c3)")
    z = np.load(GOLD)
    assert int(z["estimator"]) == 1
    ref = (z["pixels"], z["mean"].astype(np.float64), z["std"].astype(np.float64), int(z["spp"]))
    return scenes.aperture_room(), ref


def _frame(sc, px, guided, spp=SPP):
    ctx = lfh.Context.from_scene(sc, estimator=1, guiding=1 if guided else 0)
    g = None
    if guided:
        g = lfh.Guide.for_scene(sc, seed=0)
        ctx.set_guide(g, train_every=1)
    ctx.update(sc.camera(), reset=True)
    ctx.render_passes(spp)
    s, c = ctx.accum()
    img = ((s.astype(np.float64) - c) / spp)[..., :3].reshape(-1, 3)[px]
    if g is not None:
        assert g.steps() > 0
    return img


def _rmse(img, mean, std, ref_spp):
    """RMSE against the reference with the reference's own noise removed."""
    d = float(np.mean((img - mean) ** 2)) - float(np.mean(std ** 2)) / ref_spp
    return float(np.sqrt(max(d, 0.0)))


def test_plain_frame_matches_fp64_reference(room):
    print("[c3] plain 1280x720 x 256 spp", flush=True)
    sc, (px, mean, std, ref_spp) = room
    img = _frame(sc, px, guided=False)
    se = std * np.sqrt(1.0 / SPP + 1.0 / ref_spp)
    z = np.abs(img - mean) / np.maximum(se, 1e-7)
    assert np.mean(z < 4.5) > 0.99, np.mean(z < 4.5)
    rmse = np.sqrt(np.mean((img - mean) ** 2))
    expected = np.sqrt(np.mean(se ** 2))
    assert 0.4 < rmse / expected < 1.25, (rmse, expected)
    se_all = np.sqrt(np.sum(se ** 2)) / img.size
    assert abs(img.mean() - mean.mean()) < 4 * se_all, (img.mean(), mean.mean(), se_all)


def test_guided_frame_is_unbiased_and_gains_over_plain(room):
    print("[c3] guided vs plain 1280x720 x 256 spp", flush=True)
    sc, (px, mean, std, ref_spp) = room
    plain = _frame(sc, px, guided=False)
    guided = _frame(sc, px, guided=True)
    se2 = std * np.sqrt(2.0 / SPP + 1.0 / ref_spp)
    z = np.abs(guided - mean) / np.maximum(se2, 1e-7)
    assert np.mean(z < 4.5) > 0.99, np.mean(z < 4.5)
    e_p = _rmse(plain, mean, std, ref_spp)
    e_g = _rmse(guided, mean, std, ref_spp)
    # bias: the mean error over all values against its standard error (from
    # the guided frame's own scatter, which any bias would only inflate)
    se_mean = np.sqrt(np.mean((guided - mean) ** 2) / guided.size)
    assert abs(guided.mean() - mean.mean()) < 4 * se_mean, (guided.mean(), mean.mean(), se_mean)
    print(f"[c3] RMSE vs FP64 at {SPP} spp: plain {e_p:.5f}, guided {e_g:.5f} (ratio {e_g / e_p:.3f})", flush=True)
    assert e_g < 0.8 * e_p, (e_g, e_p)


def _served(sc, px, guided, spp, extra=()):
    """The room rendered as a job of the render server (guided: its scene's
    guide, trained by the job's own wavefronts), with `extra` plain jobs
    (scene, spp) submitted beside it.  Returns the room's pixels and the
    extra jobs' images."""
    srv = lfh.Server(max_jobs=8)
    o = sc.options()
    j = srv.submit_scene(sc, spp, estimator=1, guiding=1 if guided else 0)
    others = [srv.submit_scene(s2, n2) for s2, n2 in extra]
    srv.run_until_idle()
    st = srv.status(j)
    assert st.state == 2 and st.spp_done == spp, (st.state, srv.job_error(j))
    img = srv.result(j, o.width, o.height).reshape(-1, 3)   # (sum - compensation) / spp, float32
    outs = []
    for (s2, _), k in zip(extra, others):
        o2 = s2.options()
        outs.append(srv.result(k, o2.width, o2.height))
    srv.close()
    return img[px].astype(np.float64), outs


def test_served_plain_job_equals_the_context_render(room):
    print("[c3] served plain == context", flush=True)
    sc, (px, mean, std, ref_spp) = room
    spp = 32
    served, _ = _served(sc, px, guided=False, spp=spp)
    ctx = lfh.Context.from_scene(sc, estimator=1)
    ctx.update(sc.camera(), reset=True)
    ctx.render_passes(spp)
    s_, c_ = ctx.accum()
    # the server's read-back: (sum - compensation) * (1 / spp) in float32
    want = ((s_ - c_) * np.float32(1.0 / spp)).reshape(-1, 3)[px]
    assert np.array_equal(served.astype(np.float32).view(np.uint32), want.view(np.uint32))


def test_served_guided_job_is_unbiased_alone_and_in_a_mix(room):
    """A guided job of the render server (its scene's guide trained online by
    the job's wavefronts) is an unbiased estimate alone and when plain jobs of
    other scenes share its steps; the plain jobs' images are bitwise those of
    rendering them alone.  The guided image is not bitwise the same alone and
    in the mix (the steps group its passes differently, so the guide trains on
    different batches): both agree with the FP64 reference within the
    statistical tolerance, and their errors within 25% of each other."""
    print("[c3] served guided alone / in a mix", flush=True)
    sc, (px, mean, std, ref_spp) = room
    spp = 128
    others = [(scenes.cornell(128, 96), 200), (scenes.cornell(64, 64, max_depth=6), 90)]
    alone_plain = []
    for s2, n2 in others:
        srv = lfh.Server(max_jobs=4)
        k = srv.submit_scene(s2, n2)
        srv.run_until_idle()
        o2 = s2.options()
        alone_plain.append(srv.result(k, o2.width, o2.height))
        srv.close()
    g_alone, _ = _served(sc, px, guided=True, spp=spp)
    g_mix, outs = _served(sc, px, guided=True, spp=spp, extra=others)
    for a, b in zip(alone_plain, outs):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    se2 = std * np.sqrt(2.0 / spp + 1.0 / ref_spp)
    errs = []
    for img in (g_alone, g_mix):
        z = np.abs(img - mean) / np.maximum(se2, 1e-7)
        assert np.mean(z < 4.5) > 0.99, np.mean(z < 4.5)
        se_mean = np.sqrt(np.mean((img - mean) ** 2) / img.size)
        assert abs(img.mean() - mean.mean()) < 4 * se_mean, (img.mean(), mean.mean(), se_mean)
        errs.append(_rmse(img, mean, std, ref_spp))
    print(f"[c3] served guided RMSE alone {errs[0]:.5f}, in the mix {errs[1]:.5f}", flush=True)
    assert 0.8 < errs[1] / errs[0] < 1.25, errs
