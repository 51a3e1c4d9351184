"""GPU checks of the learned guide against the FP64 oracle (oracle/guide.py)
and of the estimators that use it.

* MLP forward (BF16 MFMA, FP32 accumulation) against the oracle evaluated on
  the same BF16-rounded operands: 99.9 % of outputs within 2e-3 of the
  output scale, all within 2e-2 (the residue is FP32-vs-FP64 accumulation and
  the occasional activation that rounds to the neighbouring BF16 value).
* Mixture pdf of the device lobes within 1e-3 relative of the oracle's; the
  device sampler's directions follow that pdf (histogram vs quadrature).
* Loss and gradient (BF16 activations and output gradients) within 5 % L2 of
  the FP64 analytic gradient, which is itself pinned by finite differences.
* Adam's first step moves every parameter by lr |g| / (|g| + eps).
* Training is deterministic (bitwise) and lowers the loss.
* Consistent-MIS estimator: paths equal the oracle's path by path.
* Guided rendering is unbiased: block means of a guided, online-trained render
  agree with the FP64 oracle of the consistent estimator.
"""
import math

import numpy as np
import pytest

from lavaframe_amd import lfh, scenes
from oracle import guide as G
from oracle import native

pytestmark = pytest.mark.gpu

LO = np.array([-1.0, -0.5, 0.0], np.float32)
HI = np.array([1.0, 1.5, 3.0], np.float32)


def _queries(n, seed):
    rng = np.random.default_rng(seed)
    q = np.zeros((n, 9), np.float32)
    q[:, 0:3] = LO + rng.random((n, 3)) * (HI - LO)
    for c in (3, 6):
        v = rng.normal(size=(n, 3))
        q[:, c:c + 3] = v / np.linalg.norm(v, axis=1, keepdims=True)
    return q


def _enc(q):
    return G.encode(G.normalise_pos(q[:, 0:3], LO, HI), q[:, 3:6], q[:, 6:9])


def _guide(seed=3, scale=None, kappa_bias=None):
    g = lfh.Guide(LO, HI, seed=seed)
    p = g.params().astype(np.float64)
    if scale is not None:
        p = np.random.default_rng(seed).normal(0, scale, p.size)
    if kappa_bias is not None:
        b4 = G.n_params() - 40
        p[b4 + 3:b4 + 40:5] = kappa_bias
    g.set_params(p.astype(np.float32))
    return g, g.params().astype(np.float64)


def test_forward_matches_bf16_oracle():
    print("[guide] forward", flush=True)
    g, p = _guide(scale=0.25)
    q = _queries(8192, 1)
    out = g.forward(q)
    ref, _ = G.forward(p, _enc(q), emulate_bf16=True)
    err = np.abs(out - ref)
    scale = np.abs(ref).max()
    assert scale > 1.0
    assert np.quantile(err, 0.999) < 2e-3 * scale, (np.quantile(err, 0.999), scale)
    assert err.max() < 2e-2 * scale, (err.max(), scale)


def test_pdf_and_sampler_match_oracle():
    print("[guide] pdf / sample", flush=True)
    g, p = _guide(seed=4, kappa_bias=2.0)
    q = _queries(20000, 2)
    seeds = np.random.default_rng(3).integers(0, 2**32, (len(q), 4), dtype=np.uint64).astype(np.uint32)
    d, pdf = g.sample(q, seeds)
    np.testing.assert_allclose(np.linalg.norm(d, axis=1), 1.0, atol=2e-6)
    mu, kappa, pi = G.lobes(g.forward(q))
    # the device stores each lobe's mu and kappa as binary16 (mu renormalised
    # after the rounding); the sampler and the pdf read the same stored lobes
    mu = mu.astype(np.float16).astype(np.float64)
    mu /= np.linalg.norm(mu, axis=2, keepdims=True)
    kappa = kappa.astype(np.float16).astype(np.float64)
    ref = G.mixture_pdf(mu, kappa, pi, d)
    rel = np.abs(pdf - ref) / np.maximum(ref, 1e-6)
    # a component that the device's float32 mu rounds to the other binary16
    # neighbour moves that lobe's term by up to kappa x 2^-11 (kappa about 7 here)
    assert np.quantile(rel, 0.999) < 1e-3, np.quantile(rel, 0.999)
    assert rel.max() < 2e-2, rel.max()
    # without the rounding the pdf would differ by more than that on some samples
    mu0, kappa0, _ = G.lobes(g.forward(q))
    assert np.abs(G.mixture_pdf(mu0, kappa0, pi, d) - ref).max() > 0
    np.testing.assert_array_equal(g.pdf(q, d), pdf)


def test_sampler_histogram_follows_pdf():
    print("[guide] sampler histogram", flush=True)
    g, p = _guide(seed=5, kappa_bias=1.5)
    n = 400_000
    q = np.repeat(_queries(1, 7), n, axis=0)
    seeds = np.random.default_rng(8).integers(0, 2**32, (n, 4), dtype=np.uint64).astype(np.uint32)
    d, _ = g.sample(q, seeds)
    mu, kappa, pi = G.lobes(g.forward(q[:1]))
    nb_t, nb_p = 16, 16
    ct = np.clip(((d[:, 2] + 1) * 0.5 * nb_t).astype(int), 0, nb_t - 1)
    ph = ((np.arctan2(d[:, 1], d[:, 0]) + math.pi) / (2 * math.pi) * nb_p).astype(int) % nb_p
    obs = np.zeros((nb_t, nb_p))
    np.add.at(obs, (ct, ph), 1)
    k = 24
    cz = -1 + (np.arange(nb_t * k) + 0.5) / (nb_t * k) * 2
    pp = (np.arange(nb_p * k) + 0.5) / (nb_p * k) * 2 * math.pi - math.pi
    CZ, PP = np.meshgrid(cz, pp, indexing="ij")
    s = np.sqrt(1 - CZ ** 2)
    L = np.stack([s * np.cos(PP), s * np.sin(PP), CZ], -1).reshape(-1, 3)
    pq = G.mixture_pdf(np.repeat(mu, len(L), 0), np.repeat(kappa, len(L), 0), np.repeat(pi, len(L), 0), L)
    dA = (2.0 / (nb_t * k)) * (2 * math.pi / (nb_p * k))
    expect = (pq.reshape(nb_t * k, nb_p * k) * dA).reshape(nb_t, k, nb_p, k).sum((1, 3)) * n
    assert abs(expect.sum() - n) < 0.01 * n
    m = expect > 50
    z = np.abs(obs[m] - expect[m]) / np.sqrt(expect[m])
    assert np.all(z < 6 + 0.02 * np.sqrt(expect[m])), z.max()


def _batch(n, seed):
    rng = np.random.default_rng(seed)
    q = _queries(n, seed)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    w = rng.random(n).astype(np.float32) * 2
    w[::7] = 0
    return q, d.astype(np.float32), w


def test_loss_and_gradient_match_oracle():
    print("[guide] loss / grad", flush=True)
    g, p = _guide(seed=6, scale=0.15)
    q, d, w = _batch(3000, 9)
    loss, grad = g.loss_grad(q, d, w)
    enc = _enc(q)
    ref = G.loss_grad(p, enc, d, w)
    rl = G.loss(p, enc, d, w)
    assert abs(loss - rl) < 1e-2 * abs(rl), (loss, rl)
    rel = np.linalg.norm(grad - ref) / np.linalg.norm(ref)
    cos = grad @ ref / (np.linalg.norm(grad) * np.linalg.norm(ref))
    assert rel < 0.05 and cos > 0.998, (rel, cos)
    # every layer's block, not only the largest, is right
    for l in range(4):
        o, n_ = G.layer_slice(l)
        r = ref[o:o + n_]
        assert np.linalg.norm(grad[o:o + n_] - r) < 0.08 * np.linalg.norm(r) + 1e-9, l


def test_adam_first_step_closed_form():
    print("[guide] adam", flush=True)
    g, p0 = _guide(seed=7)
    q, d, w = _batch(2048, 10)
    _, grad = g.loss_grad(q, d, w)
    g.train_step(q, d, w)
    delta = g.params().astype(np.float64) - p0
    gg = grad.astype(np.float64)
    expect = -1e-3 * gg / (np.abs(gg) + 1e-8)
    big = np.abs(gg) > 1e-6
    assert big.sum() > 1000
    np.testing.assert_allclose(delta[big], expect[big], rtol=2e-3, atol=2e-9)
    assert g.steps() == 1


def test_training_is_deterministic_and_learns():
    print("[guide] training", flush=True)
    rng = np.random.default_rng(11)
    n = 16384
    q = _queries(n, 12)
    # target: directions from a vMF about a direction that depends on position
    target = np.stack([np.cos(3 * q[:, 0]), np.sin(3 * q[:, 0]), 0.5 * np.ones(n)], 1)
    target /= np.linalg.norm(target, axis=1, keepdims=True)
    d = np.stack([G.vmf_sample(t, 20.0, rng.random(1), rng.random(1))[0] for t in target]).astype(np.float32)
    w = np.ones(n, np.float32)
    losses = []
    guides = [lfh.Guide(LO, HI, seed=21), lfh.Guide(LO, HI, seed=21)]
    for it in range(60):
        for k, gd in enumerate(guides):
            l = gd.train_step(q, d, w)
            if k == 0:
                losses.append(l)
    np.testing.assert_array_equal(guides[0].params(), guides[1].params())
    assert losses[-1] < losses[0] - 1.0, (losses[0], losses[-1])


def test_consistent_estimator_paths_match_oracle():
    """options.estimator = 1 (lobe-mixture pdf in the MIS weights after BSDF
    samples) on the GPU against the oracle's PathTrace with the same option."""
    print("[guide] estimator 1 paths", flush=True)
    sc = scenes.cornell(64, 64)
    ctx = lfh.Context.from_scene(sc, estimator=1)
    arr = sc.arrays()
    orc = native.OracleScene(arr)
    opts = sc.options()
    opts.estimator = 1
    o = native.options_from(opts, sc.camera(), lights=True, envmap=False)
    assert o.estimator == 1
    rng = np.random.default_rng(4)
    n = 100_000
    cam = sc.camera()
    rays = np.zeros((n, 6), np.float32)
    rays[:, 0:3] = cam.position
    dirs = rng.normal(size=(n, 3)) * 0.3 + np.asarray(cam.forward)
    rays[:, 3:6] = dirs / np.linalg.norm(dirs, axis=1, keepdims=True)
    seeds = rng.integers(0, 2**32, (n, 4), dtype=np.uint64).astype(np.uint32)
    g = ctx.path_trace_rays(rays, seeds)
    c = orc.path_trace_rays(o, rays, seeds, 32)
    same = np.isclose(g, c, rtol=1e-4, atol=1e-5).all(1)
    assert same.mean() > 0.99, same.mean()
    # the option changes the estimate on some paths (it is not a no-op)
    o.estimator = 0
    c0 = orc.path_trace_rays(o, rays, seeds, 32)
    assert (~np.isclose(c0, c, rtol=1e-4, atol=1e-5).all(1)).sum() > 10


def test_guided_render_is_unbiased():
    """Cornell 32 x 32, guided (online training every pass) vs the FP64
    oracle of the consistent estimator: 4 x 4-pixel block means agree."""
    print("[guide] guided render", flush=True)
    sc = scenes.cornell(32, 32)
    ctx = lfh.Context.from_scene(sc, estimator=1, guiding=1)
    guide = lfh.Guide.for_scene(sc, seed=1)
    ctx.set_guide(guide, train_every=1)
    cam = sc.camera()
    spp = 256
    ctx.update(cam)
    ctx.render_passes(spp - 1)
    ctx.render()
    img = ctx.output_hdr().reshape(32, 32, 3)
    assert guide.steps() > 0 and ctx.guide_samples_seen() > 1000
    assert np.isfinite(img).all()
    orc = native.OracleScene(sc.arrays())
    opts = sc.options()
    opts.estimator = 1
    o = native.options_from(opts, cam, lights=True, envmap=False)
    px = np.arange(32 * 32, dtype=np.int32)
    so = 1024
    s1, s2 = orc.render_pixels(o, px, 200_000, so, precision=64)
    mean = (s1 / so).reshape(32, 32, 3)
    var = np.maximum(s2 / so - (s1 / so) ** 2, 0).reshape(32, 32, 3)
    bm = lambda a: a.reshape(8, 4, 8, 4, 3).mean((1, 3))
    # the guided estimator's variance is at most ~2x the BSDF one (defensive 50 % mixture)
    se = np.sqrt(bm(var) / 16 * (2.0 / spp + 1.0 / so))
    z = np.abs(bm(img) - bm(mean)) / np.maximum(se, 1e-6)
    assert np.mean(z < 4.5) > 0.99 and z.max() < 7, (np.mean(z < 4.5), z.max())
    assert abs(img.mean() - mean.mean()) < 0.02 * mean.mean()


def _floor_scene(flipped):
    """A 4 x 4 floor (y = 0) under a quad light facing down; flipped: the
    floor's winding makes its normal point down, away from the light."""
    s = lfh.Scene()
    s.add_material(lfh.MATERIAL_DEFAULT)
    s.add_material(lfh.material((0.7, 0.7, 0.7), roughness=0.6))
    a, b, c, d = [np.array(p, np.float32) for p in ((-2, 0, -2), (-2, 0, 2), (2, 0, 2), (2, 0, -2))]
    tris = np.array([[a, b, c], [a, c, d]], np.float32)
    if flipped:
        tris = tris[:, ::-1].copy()
    s.add_instance(s.add_mesh(tris), None, 1)
    s.add_light(lfh.quad_light((-0.5, 2.0, -0.5), (0.5, 2.0, -0.5), (-0.5, 2.0, 0.5), (6, 6, 6)))
    s.set_camera((0, 1.5, -3), (0, 0, 0), 50)
    s.set_options(width=48, height=32, max_depth=3, enable_rr=0)
    return s.build()


def test_consistent_estimator_sees_lights_from_behind_one_sided_surfaces():
    """The reference offsets shadow-ray origins along the unflipped normal
    (pathtrace.glsl:130): seen from behind, a one-sided floor shadows its own
    light samples.  Estimator 1 offsets to the incident side, so flipping the
    floor's winding leaves its image unchanged; estimator 0 keeps the
    reference behaviour.  Both match the oracle path by path."""
    print("[guide] one-sided surfaces", flush=True)
    means = {}
    for est in (0, 1):
        for flipped in (False, True):
            sc = _floor_scene(flipped)
            ctx = lfh.Context.from_scene(sc, estimator=est)
            cam = sc.camera()
            ctx.update(cam, reset=True)
            ctx.render_passes(256)
            means[est, flipped] = ctx.output_hdr().mean()
            orc = native.OracleScene(sc.arrays())
            opts = sc.options()
            opts.estimator = est
            o = native.options_from(opts, cam, lights=True, envmap=False)
            rng = np.random.default_rng(7)
            n = 20000
            rays = np.zeros((n, 6), np.float32)
            rays[:, 0:3] = cam.position
            dd = rng.normal(size=(n, 3)) * 0.2 + np.asarray(cam.forward)
            rays[:, 3:6] = dd / np.linalg.norm(dd, axis=1, keepdims=True)
            seeds = rng.integers(0, 2**32, (n, 4), dtype=np.uint64).astype(np.uint32)
            g = ctx.path_trace_rays(rays, seeds)
            c = orc.path_trace_rays(o, rays, seeds, 32)
            assert np.isclose(g, c, rtol=1e-4, atol=1e-5).all(1).mean() > 0.99, (est, flipped)
    assert abs(means[1, True] - means[1, False]) < 0.03 * means[1, False], means
    assert means[0, True] < 0.5 * means[0, False], means


def test_forward_device_equals_host_forward():
    """lfh_guide_forward_device on resident queries gives the host entry's outputs bitwise."""
    import torch
    print("[guide] forward_device", flush=True)
    g, _ = _guide(seed=5, scale=0.3)
    n = 5000
    x = _queries(n, 11)
    host = g.forward(x)
    dx = torch.from_numpy(x).cuda()
    dout = torch.zeros(n, 40, dtype=torch.float32, device="cuda")
    g.forward_device(dx.data_ptr(), n, dout.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(dout.cpu().numpy(), host)
