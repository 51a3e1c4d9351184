"""Pins the CPU oracle (oracle/oracle.cpp) with checks that do not retype its
formulas: an independent vectorised RNG, brute-force traversal, closed forms
of the BSDF at normal incidence, sampling-vs-pdf histograms, light-sample
geometry, environment-pdf normalisation and simple scenes with known
radiance."""
import math

import numpy as np
import pytest

from lavaframe_amd import lfh, scenes
from oracle import native

pytestmark = pytest.mark.skipif(not native.available(), reason="oracle not built")


# ---------------------------------------------------------------- RNG
def pcg4d_np(v):
    """pcg4d on (n, 4) uint32 with numpy wrap-around arithmetic."""
    v = v.astype(np.uint32).copy()
    with np.errstate(over="ignore"):
        v = v * np.uint32(1664525) + np.uint32(1013904223)
        v[:, 0] += v[:, 1] * v[:, 3]
        v[:, 1] += v[:, 2] * v[:, 0]
        v[:, 2] += v[:, 0] * v[:, 1]
        v[:, 3] += v[:, 1] * v[:, 2]
        v ^= v >> np.uint32(16)
        v[:, 0] += v[:, 1] * v[:, 3]
        v[:, 1] += v[:, 2] * v[:, 0]
        v[:, 2] += v[:, 0] * v[:, 1]
        v[:, 3] += v[:, 1] * v[:, 2]
    return v


def test_pcg4d_and_rand_match_independent_formulation():
    L = native.lib()
    rng = np.random.default_rng(0)
    states = rng.integers(0, 2**32, (500, 4), dtype=np.uint64).astype(np.uint32)
    # seeds as renderer.glsl:InitRNG builds them
    states[:50] = np.array([[x, y, 7, x + y] for x, y in zip(range(50), range(100, 150))], np.uint32)
    ref = states.copy()
    for step in range(4):
        ref = pcg4d_np(ref)
        for i in range(len(states)):
            s = np.ascontiguousarray(states[i])
            r = L.orc_rand(s.ctypes.data_as(native.u32p))
            states[i] = s
            assert r == np.float32(np.float32(ref[i, 0]) / np.float32(4294967296.0))
        assert np.array_equal(states, ref)


def test_rand_reaches_one():
    # float(0xffffffff) rounds to 2^32, so x >= 2^32 - 128 gives exactly 1.0
    assert np.float32(np.float32(0xFFFFFFFF) / np.float32(0xFFFFFFFF)) == 1.0


# ---------------------------------------------------------------- traversal
def brute_force(tris, rays):
    """Closest hit over all triangles in float64 (no BVH)."""
    o = rays["org"].astype(np.float64)
    d = rays["dir"].astype(np.float64)
    v0 = tris[:, 0].astype(np.float64)
    e0 = tris[:, 1] - tris[:, 0]
    e1 = tris[:, 2] - tris[:, 0]
    best_t = np.full(len(rays), np.inf)
    best = np.full(len(rays), -1)
    second = np.full(len(rays), np.inf)
    for k in range(len(tris)):
        pv = np.cross(d, e1[k])
        det = pv @ e0[k]
        tv = o - v0[k]
        qv = np.cross(tv, e0[k])
        with np.errstate(divide="ignore", invalid="ignore"):
            u = np.einsum("ij,ij->i", tv, pv) / det
            v = (d * qv).sum(1) / det
            t = qv @ e1[k] / det
        ok = (u >= 0) & (v >= 0) & (u + v <= 1) & (t >= 0)
        t = np.where(ok, t, np.inf)
        closer = t < best_t
        second = np.where(closer, best_t, np.minimum(second, t))
        best = np.where(closer, k, best)
        best_t = np.where(closer, t, best_t)
    return best, best_t, second


@pytest.mark.parametrize("n", [1, 7, 300])
def test_closest_hit_matches_brute_force(n):
    sc = scenes.soup(n, seed=n + 100)
    arr = sc.arrays()
    orc = native.OracleScene(arr)
    rays = scenes.incoherent_rays(20000, seed=n)
    hits = orc.trace_closest(rays, tie_mode=1)
    # triangles in prim order: prim p = vertices 3p..3p+2
    tris = arr["vertices"][:, :3].reshape(-1, 3, 3)
    best, bt, second = brute_force(tris, rays)
    with np.errstate(invalid="ignore"):
        clear = np.isfinite(bt) & (second - bt > 1e-5 * np.maximum(bt, 1e-3))
    clear |= ~np.isfinite(bt)
    assert clear.mean() > 0.99
    assert np.array_equal(hits["prim"][clear], np.where(np.isfinite(bt[clear]), best[clear], -1))
    hit = clear & np.isfinite(bt)
    np.testing.assert_allclose(hits["t"][hit], bt[hit], rtol=2e-5)
    # barycentrics reproduce the hit point
    p = rays["org"][hit] + rays["dir"][hit] * hits["t"][hit, None]
    tri = tris[hits["prim"][hit]]
    q = tri[:, 0] * (1 - hits["u"][hit] - hits["v"][hit])[:, None] + tri[:, 1] * hits["u"][hit, None] + \
        tri[:, 2] * hits["v"][hit, None]
    np.testing.assert_allclose(p, q, atol=2e-5)


def test_any_hit_consistent_with_closest():
    sc = scenes.soup(500, seed=9)
    orc = native.OracleScene(sc.arrays())
    rays = scenes.incoherent_rays(20000, seed=4)
    rays["tmax"] = np.random.default_rng(1).uniform(0.01, 1.0, len(rays)).astype(np.float32)
    hits = orc.trace_closest(rays, tie_mode=1)
    occ = orc.trace_any(rays)
    assert np.array_equal(occ.astype(bool), hits["prim"] >= 0)


def test_tie_rules():
    """Duplicated triangles give exact-t ties: the reference's first-found
    rule and the lowest-index rule differ only there."""
    t = scenes.soup_triangles(64, seed=3)
    t = np.concatenate([t, t[::-1]])  # every triangle twice
    s = lfh.Scene()
    s.add_material(lfh.MATERIAL_DEFAULT)
    s.add_instance(s.add_mesh(t), scenes.affine(), 0)
    s.build()
    orc = native.OracleScene(s.arrays())
    rays = scenes.incoherent_rays(4000, seed=8)
    cen = t.mean(1)[np.arange(4000) % 64]
    d = cen - rays["org"]
    rays["dir"] = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    h1, _, ties = orc.trace_closest(rays, tie_mode=1, counts=True)
    h0 = orc.trace_closest(rays, tie_mode=0)
    hit = h1["prim"] >= 0
    assert ties > 0 and hit.sum() > 100
    # lowest index wins: the copy with the smaller prim id
    assert np.all(h1["prim"][hit] < 64 + 64)
    partner = np.where(h1["prim"][hit] < 64, 127 - h1["prim"][hit], h1["prim"][hit])
    assert np.all(h1["prim"][hit] <= partner)
    assert np.array_equal(h0["t"].view(np.uint32), h1["t"].view(np.uint32))
    # on a soup without duplicates both rules agree exactly
    sc = scenes.soup(2000, seed=5)
    o2 = native.OracleScene(sc.arrays())
    r2 = scenes.incoherent_rays(20000, seed=2)
    a, _, ties2 = o2.trace_closest(r2, tie_mode=1, counts=True)
    b = o2.trace_closest(r2, tie_mode=0)
    assert ties2 == 0 and np.array_equal(a, b)


# ---------------------------------------------------------------- BSDF
def frame_rows(V, mat_id, eta=1.0 / 1.45, n=1):
    fr = np.zeros((n, 16), np.float32)
    fr[:, 0:3] = (0, 0, 1)
    fr[:, 3:6] = (1, 0, 0)
    fr[:, 6:9] = (0, 1, 0)
    fr[:, 9:12] = V
    fr[:, 12] = eta
    fr[:, 13] = mat_id
    return fr


def _bsdf_scene(mats):
    s = lfh.Scene()
    s.add_material(lfh.MATERIAL_DEFAULT)
    for m in mats:
        s.add_material(m)
    s.add_instance(s.add_mesh(scenes.unit_cube_tris()), None, 0)
    return native.OracleScene(s.build().arrays())


@pytest.mark.parametrize("rough,alb", [(0.5, 0.8), (0.2, 0.3), (1.0, 1.0)])
def test_disney_eval_normal_incidence_closed_form(rough, alb):
    """V = L = N: f = 0.5 albedo / pi + 0.04 / (4 pi a^2), pdf = 0.5/pi + 1/(8 pi a^2)
    (default material: specular 0.5, metallic 0, clearcoat 0; disney.glsl:54-113)."""
    orc = _bsdf_scene([lfh.material((alb, alb, alb), roughness=rough)])
    for prec in (32, 64):
        f, pdf = orc.eval_bsdf(frame_rows((0, 0, 1), 1), np.array([[0, 0, 1]], np.float32), prec)
        a = rough
        np.testing.assert_allclose(f[0], 0.5 * alb / math.pi + 0.04 / (4 * math.pi * a * a), rtol=1e-5)
        np.testing.assert_allclose(pdf[0], 0.5 / math.pi + 1 / (8 * math.pi * a * a), rtol=1e-5)


def test_disney_eval_reciprocity():
    rng = np.random.default_rng(0)
    mats = [lfh.material((0.7, 0.4, 0.2), roughness=0.3), lfh.material((0.9, 0.9, 0.9), metallic=1.0, roughness=0.5),
            lfh.material((0.5, 0.5, 0.5), clearcoat=1.0, clearcoatRoughness=0.3, sheen=0.5, roughness=0.6)]
    orc = _bsdf_scene(mats)
    for m in range(1, 4):
        V = rng.normal(size=(2000, 3))
        L = rng.normal(size=(2000, 3))
        V[:, 2] = np.abs(V[:, 2]) + 0.05
        L[:, 2] = np.abs(L[:, 2]) + 0.05
        V /= np.linalg.norm(V, axis=1, keepdims=True)
        L /= np.linalg.norm(L, axis=1, keepdims=True)
        fr1 = frame_rows(V.astype(np.float32), m, n=len(V))
        fr1[:, 9:12] = V
        fr2 = frame_rows(L.astype(np.float32), m, n=len(L))
        fr2[:, 9:12] = L
        f1, _ = orc.eval_bsdf(fr1, L.astype(np.float32), 64)
        f2, _ = orc.eval_bsdf(fr2, V.astype(np.float32), 64)
        np.testing.assert_allclose(f1, f2, rtol=1e-5, atol=1e-7)


@pytest.mark.parametrize("mat", [dict(roughness=0.4), dict(roughness=0.15, metallic=1.0),
                                 dict(roughness=0.8, sheen=1.0, sheenTint=0.5),
                                 dict(roughness=0.05, metallic=0.5)])
def test_disney_sample_follows_eval_pdf(mat):
    """Histogram of DisneySample directions over upper-hemisphere bins against
    the integral of DisneyEval's pdf over each bin (disney.glsl:128-291)."""
    orc = _bsdf_scene([lfh.material((0.6, 0.5, 0.4), **mat)])
    V = np.array([0.5, 0.2, 0.0])
    V[2] = math.sqrt(1 - V[0] ** 2 - V[1] ** 2)
    N = 200_000
    fr = frame_rows(V.astype(np.float32), 1, n=N)
    seeds = np.random.default_rng(1).integers(0, 2**32, (N, 4), dtype=np.uint64).astype(np.uint32)
    L, f, pdf, _ = orc.sample_bsdf(fr, seeds, 64)
    nb_t, nb_p = 12, 12
    up = L[:, 2] > 0
    ct = np.clip((L[up, 2] * nb_t).astype(int), 0, nb_t - 1)
    ph = ((np.arctan2(L[up, 1], L[up, 0]) + math.pi) / (2 * math.pi) * nb_p).astype(int) % nb_p
    obs = np.zeros((nb_t, nb_p))
    np.add.at(obs, (ct, ph), 1)
    # quadrature of the pdf over bins (uniform in cos(theta) and phi)
    q = 12
    cz = (np.arange(nb_t * q) + 0.5) / (nb_t * q)
    pp = (np.arange(nb_p * q) + 0.5) / (nb_p * q) * 2 * math.pi - math.pi
    CZ, PP = np.meshgrid(cz, pp, indexing="ij")
    s = np.sqrt(1 - CZ ** 2)
    Ls = np.stack([s * np.cos(PP), s * np.sin(PP), CZ], -1).reshape(-1, 3).astype(np.float32)
    frq = frame_rows(V.astype(np.float32), 1, n=len(Ls))
    _, pq = orc.eval_bsdf(frq, Ls, 64)
    dA = (1.0 / (nb_t * q)) * (2 * math.pi / (nb_p * q))
    expect = (pq.reshape(nb_t * q, nb_p * q) * dA).reshape(nb_t, q, nb_p, q).sum((1, 3)) * N
    m = expect > 100
    assert m.sum() >= 10
    z = np.abs(obs[m] - expect[m]) / np.sqrt(expect[m])
    assert np.all(z < 6 + 0.03 * np.sqrt(expect[m])), (z.max(), obs[m][z.argmax()], expect[m][z.argmax()])


def test_clearcoat_sampling_quirk_reproduced():
    """ImportanceSampleGTR1 draws both the polar and the azimuthal angle from
    r1 (sampling.glsl:13-15), so sampled clearcoat half-vectors lie on one
    curve phi = 2 pi r1(cos theta) instead of filling the lobe.  With a mirror-
    like V = N, L = reflect(-V, H) keeps that dependence: phi(L) is a function
    of cos(theta_L)."""
    orc = _bsdf_scene([lfh.material((0.5, 0.5, 0.5), metallic=1.0, clearcoat=1e6, clearcoatRoughness=0.5,
                                    roughness=0.5)])
    N = 20000
    fr = frame_rows((0, 0, 1), 1, n=N)
    seeds = np.random.default_rng(7).integers(0, 2**32, (N, 4), dtype=np.uint64).astype(np.uint32)
    L, f, pdf, _ = orc.sample_bsdf(fr, seeds, 64)
    # with clearcoat huge, primarySpecRatio ~ 1e-6: almost every sample is clearcoat
    ok = L[:, 2] > 0.05
    ct = L[ok, 2]
    phi = np.arctan2(L[ok, 1], L[ok, 0]) % (2 * math.pi)
    order = np.argsort(ct)
    # azimuth is a deterministic (monotone modulo 2 pi) function of cos(theta):
    # neighbours in cos(theta) have neighbouring azimuths
    dphi = np.abs(np.diff(phi[order]))
    dphi = np.minimum(dphi, 2 * math.pi - dphi)
    assert np.median(dphi) < 0.05


# ---------------------------------------------------------------- lights, env
def test_quad_light_sample_geometry():
    sc = scenes.cornell()
    arr = sc.arrays()
    orc = native.OracleScene(arr)
    n = 5000
    rng = np.random.default_rng(0)
    pos = rng.uniform(0.1, 0.9, (n, 3)).astype(np.float32)
    seeds = rng.integers(0, 2**32, (n, 4), dtype=np.uint64).astype(np.uint32)
    d, nrm, em, dist, pdf, _ = orc.sample_light(pos, seeds, 64)
    L = arr["lights"][0]
    p0, u, v, area = L[0:3], L[6:9], L[9:12], L[13]
    pts = pos + d * dist[:, None]
    rel = pts - p0
    a = rel @ u / (u @ u)
    b = rel @ v / (v @ v)
    assert np.all((a > -1e-4) & (a < 1 + 1e-4) & (b > -1e-4) & (b < 1 + 1e-4))
    np.testing.assert_allclose(nrm, np.tile([0, -1, 0], (n, 1)), atol=1e-6)
    np.testing.assert_allclose(em, np.tile(L[3:6] * 1, (n, 1)), rtol=1e-6)
    np.testing.assert_allclose(pdf, dist ** 2 / (area * np.abs(d[:, 1])), rtol=1e-4)


def _env_scene(w=64, h=32):
    s = lfh.Scene()
    s.add_material(lfh.MATERIAL_DEFAULT)
    s.add_instance(s.add_mesh(scenes.unit_cube_tris()), None, 0)
    s.set_hdr(scenes.sky_hdr(w, h, sun_power=50.0))
    s.set_options(width=8, height=8, use_env_map=1, max_depth=2, hdr_multiplier=1.0)
    return s.build()


def test_env_pdf_normalised():
    orc = native.OracleScene(_env_scene().arrays())
    nt, nph = 400, 800
    th = (np.arange(nt) + 0.5) / nt * math.pi
    ph = (np.arange(nph) + 0.5) / nph * 2 * math.pi
    TH, PH = np.meshgrid(th, ph, indexing="ij")
    d = np.stack([np.sin(TH) * np.cos(PH), np.cos(TH), np.sin(TH) * np.sin(PH)], -1).reshape(-1, 3)
    pdf = orc.env_pdf(d.astype(np.float32), 64).reshape(nt, nph)
    integral = (pdf * np.sin(TH)).sum() * (math.pi / nt) * (2 * math.pi / nph)
    assert abs(integral - 1.0) < 0.02


def test_env_samples_are_texel_corners():
    """EnvSample returns directions at texel corners (u, v) = (i/W, j/H)
    (sampling.glsl:252-264)."""
    W, H = 64, 32
    orc = native.OracleScene(_env_scene(W, H).arrays())
    seeds = np.random.default_rng(3).integers(0, 2**32, (4000, 4), dtype=np.uint64).astype(np.uint32)
    d, pdf, col, _ = orc.sample_env(seeds, 1.0, 64)
    theta = np.arccos(np.clip(d[:, 1], -1, 1))
    v = theta / math.pi * H
    assert np.all(np.abs(v - np.round(v)) < 1e-3)
    ok = np.sin(theta) > 1e-3
    phi = np.arctan2(-d[ok, 2], -d[ok, 0]) % (2 * math.pi)
    u = phi / (2 * math.pi) * W
    assert np.all(np.minimum(np.abs(u - np.round(u)), W - np.abs(u - np.round(u))) < 2e-3)
    assert np.all(pdf[ok] > 0)


# ---------------------------------------------------------------- path tracer
def _opts(scene, **kw):
    o = native.options_from(scene.options(), scene.camera(), lights=len(scene.arrays()["lights"]) > 0,
                            envmap=bool(scene.options().use_env_map))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def test_constant_background_seen_directly():
    s = lfh.Scene()
    s.add_material(lfh.MATERIAL_DEFAULT)
    s.add_instance(s.add_mesh(scenes.unit_cube_tris()), scenes.affine(translate=(0, 0, 50)), 0)
    s.set_camera((0, 0, 0), (0, 0, -1), 30.0)
    s.set_options(width=16, height=16, max_depth=3, use_constant_bg=1, bg_color=(0.25, 0.5, 1.0))
    s.build()
    orc = native.OracleScene(s.arrays())
    acc = orc.render(_opts(s), 2, 3, precision=32)
    np.testing.assert_allclose(acc / 3.0, np.broadcast_to([0.25, 0.5, 1.0], acc.shape), rtol=1e-6)


def test_one_bounce_plane_matches_directional_albedo():
    """A camera looking straight down on a large plane under a constant
    background: the pixel mean converges to bg * integral of f cos over the
    hemisphere, which is evaluated here by quadrature of DisneyEval."""
    mat = lfh.material((0.6, 0.6, 0.6), roughness=0.7)
    s = lfh.Scene()
    s.add_material(lfh.MATERIAL_DEFAULT)
    mid = s.add_material(mat)
    quad = np.array([[[-1e3, 0, -1e3], [-1e3, 0, 1e3], [1e3, 0, 1e3]], [[-1e3, 0, -1e3], [1e3, 0, 1e3], [1e3, 0, -1e3]]],
                    np.float32)
    s.add_instance(s.add_mesh(quad), None, mid)
    s.set_camera((0, 1, 0), (0, 0, 0.0001), 1.0)
    s.set_options(width=4, height=4, max_depth=2, enable_rr=0, use_constant_bg=1, bg_color=(1.0, 1.0, 1.0))
    s.build()
    orc = native.OracleScene(s.arrays())
    opts = _opts(s)
    px = np.arange(16, dtype=np.int32)
    spp = 20000
    sm, sq = orc.render_pixels(opts, px, 2, spp, precision=64)
    mean = sm.sum(0) / (16 * spp)
    # quadrature of f cos with V = N
    nq = 400
    cz = (np.arange(nq) + 0.5) / nq
    ph = (np.arange(nq) + 0.5) / nq * 2 * math.pi
    CZ, PH = np.meshgrid(cz, ph, indexing="ij")
    sn = np.sqrt(1 - CZ ** 2)
    L = np.stack([sn * np.cos(PH), sn * np.sin(PH), CZ], -1).reshape(-1, 3).astype(np.float32)
    so = _bsdf_scene([mat])
    f, _ = so.eval_bsdf(frame_rows((0, 0, 1), 1, n=len(L)), L, 64)
    albedo = (f[:, 0] * L[:, 2]).sum() * (1.0 / nq) * (2 * math.pi / nq)
    assert abs(mean[0] - albedo) < 0.01 * albedo + 3e-3


def test_fp32_and_fp64_paths_agree_in_the_mean():
    s = scenes.cornell(32, 32)
    orc = native.OracleScene(s.arrays())
    o = _opts(s)
    a32 = orc.render(o, 2, 64, precision=32) / 64
    a64 = orc.render(o, 2, 64, precision=64) / 64
    assert abs(a32.mean() - a64.mean()) < 0.02 * a64.mean()


# ---------------------------------------------------------------- transmission, absorption, camera
def _fresnel_exact(cos_i, n1, n2):
    """Unpolarised Fresnel reflectance (textbook, Hecht's Optics 4.34/4.40)."""
    sin_t = n1 / n2 * math.sqrt(max(0.0, 1 - cos_i * cos_i))
    if sin_t >= 1:
        return 1.0
    cos_t = math.sqrt(1 - sin_t * sin_t)
    rs = (n1 * cos_i - n2 * cos_t) / (n1 * cos_i + n2 * cos_t)
    rp = (n2 * cos_i - n1 * cos_t) / (n2 * cos_i + n1 * cos_t)
    return 0.5 * (rs * rs + rp * rp)


@pytest.mark.parametrize("theta_deg,ior", [(30.0, 1.5), (60.0, 1.33), (10.0, 2.0)])
def test_smooth_dielectric_sampling_follows_snell_and_fresnel(theta_deg, ior):
    """specTrans = 1, metallic = 0, roughness at its floor: DisneySample reflects
    with the exact Fresnel probability and refracts along Snell's law
    (disney.glsl:128-173 against the textbook)."""
    orc = _bsdf_scene([lfh.material((1, 1, 1), transmission=1.0, ior=ior, roughness=0.0)])
    th = math.radians(theta_deg)
    V = np.array([math.sin(th), 0.0, math.cos(th)])
    N = 40000
    fr = frame_rows(V.astype(np.float32), 1, eta=1.0 / ior, n=N)
    seeds = np.random.default_rng(11).integers(0, 2**32, (N, 4), dtype=np.uint64).astype(np.uint32)
    L, f, pdf, _ = orc.sample_bsdf(fr, seeds, 64)
    refl = L[:, 2] > 0
    F = _fresnel_exact(math.cos(th), 1.0, ior)
    se = math.sqrt(F * (1 - F) / N)
    assert abs(refl.mean() - F) < 5 * se + 1e-4, (refl.mean(), F)
    # the roughness floor (GGX alpha 1e-3) spreads the microfacet normals with a
    # long tail: the median deviation stays at the alpha scale and 95 % within 1e-2
    def dev(Ls, target):
        return np.arccos(np.clip(Ls @ target, -1, 1))
    mirror = np.array([-V[0], -V[1], V[2]])
    dm = dev(L[refl], mirror)
    assert np.median(dm) < 3e-3 and np.quantile(dm, 0.95) < 1e-2, (np.median(dm), np.quantile(dm, 0.95))
    # Snell: sin(theta_t) = sin(theta_i) / ior, in the plane of incidence, below the surface
    st = math.sin(th) / ior
    t_dir = np.array([-st, 0.0, -math.sqrt(1 - st * st)])
    dt = dev(L[~refl], t_dir)
    assert np.median(dt) < 3e-3 and np.quantile(dt, 0.95) < 1e-2, (np.median(dt), np.quantile(dt, 0.95))


def test_beer_lambert_absorption_through_a_cube():
    """A nearly index-matched medium (ior 1.01, roughness at its floor) with
    extinction e and atDistance 1 inside the unit cube, lit by a constant white
    background: a ray crossing the cube over length l (Snell's law inside)
    arrives with e^l per channel (pathtrace.glsl:262-276: exp(-absorption t),
    absorption = -log(e) / atDistance), less the two Fresnel reflections.
    (ior exactly 1 makes the refraction term 0/0 in the reference formula.)"""
    ext = (0.5, 0.25, 0.8)
    ior = 1.01
    s = lfh.Scene()
    s.add_material(lfh.MATERIAL_DEFAULT)
    s.add_material(lfh.material((1, 1, 1), transmission=1.0, ior=ior, roughness=0.0, extinction=ext, atDistance=1.0))
    s.add_instance(s.add_mesh(scenes.unit_cube_tris()), None, 1)
    s.set_camera((0, 0, -3), (0, 0, 0), 40)
    s.set_options(width=8, height=8, max_depth=6, enable_rr=0, use_constant_bg=1, bg_color=(1.0, 1.0, 1.0))
    sc = s.build()
    orc = native.OracleScene(sc.arrays())
    o = native.options_from(sc.options(), sc.camera(), lights=False, envmap=False)
    rng = np.random.default_rng(2)
    n = 256
    for slope in (0.0, 0.25):
        d = np.array([1.0, slope, 0.0]) / math.hypot(1.0, slope)
        rays = np.zeros((n, 6), np.float32)
        y0 = rng.uniform(-0.4, 0.5 - slope - 0.1, n)
        z0 = rng.uniform(-0.4, 0.4, n)
        rays[:, 0] = -2.0
        rays[:, 1] = y0 - 1.5 * slope
        rays[:, 2] = z0
        rays[:, 3:6] = d
        seeds = rng.integers(0, 2**32, (n, 4), dtype=np.uint64).astype(np.uint32)
        rad = orc.path_trace_rays(o, rays, seeds, 64)
        th = math.atan(slope)
        tt = math.asin(math.sin(th) / ior)
        length = 1.0 / math.cos(tt)
        T = (1 - _fresnel_exact(math.cos(th), 1.0, ior)) * (1 - _fresnel_exact(math.cos(tt), ior, 1.0))
        # the continuation ray starts EPS = 1e-3 along L past the entry point
        # (pathtrace.glsl:291), so the distance absorbed over is l - EPS; a path
        # that reflects (probability ~1e-4 per face) misses the expected value
        expect = T * np.array(ext) ** (length - 1e-3)
        ok = np.isclose(rad, expect, rtol=5e-3).all(1)
        assert ok.mean() > 0.98, (ok.mean(), rad[~ok][:3], expect)
        # the microfacet normals (GGX alpha 1e-3) jitter single paths; the median is sharp
        np.testing.assert_allclose(np.median(rad[ok], 0), expect, rtol=2e-4)


def _cam_opts(W, H, fov_deg, aperture=0.0, focal=1.0):
    o = native.Options()
    o.width, o.height, o.tile_width, o.tile_height = W, H, W, H
    o.cam_pos[:] = [1.0, 2.0, 3.0]
    o.cam_right[:] = [1.0, 0.0, 0.0]
    o.cam_up[:] = [0.0, 1.0, 0.0]
    o.cam_fwd[:] = [0.0, 0.0, 1.0]
    o.fov = math.radians(fov_deg)
    o.aperture = aperture
    o.focal_dist = focal
    return o


def test_camera_pixel_mapping_and_tent_filter():
    """Pinhole camera (renderer.glsl:27-62): origins at the camera, the
    horizontal field of view maps pixel x + 0.5 + jitter to
    tan(fov/2) (2u/W - 1), rows scaled by H/W; the jitter of each axis follows
    the tent density 1 - |j| on [-1, 1] pixels."""
    W, H, fov = 64, 48, 60.0
    o = _cam_opts(W, H, fov)
    tan = math.tan(math.radians(fov) / 2)
    xs = np.arange(W) + 0.5
    ys = np.arange(H) + 0.5
    jx, jy = [], []
    for frame in range(1, 201):
        r = native.OracleScene.camera_rays(o, frame)
        np.testing.assert_array_equal(r[..., 0:3], np.broadcast_to([1.0, 2.0, 3.0], r[..., 0:3].shape))
        d = r[..., 3:6]
        u = (d[..., 0] / d[..., 2] / tan + 1) * W / 2
        v = (d[..., 1] / d[..., 2] / (tan * H / W) + 1) * H / 2
        jx.append(u - xs[None, :])
        jy.append(v - ys[:, None])
    for j in (np.concatenate(jx).ravel(), np.concatenate(jy).ravel()):
        assert np.abs(j).max() <= 1 + 1e-6
        assert abs(j.mean()) < 5 * math.sqrt(1 / 6 / j.size)
        edges = np.linspace(-1, 1, 21)
        obs = np.histogram(j, edges)[0]
        cdf = lambda x: np.where(x < 0, 0.5 * (1 + x) ** 2, 1 - 0.5 * (1 - x) ** 2)  # noqa: E731
        exp = np.diff(cdf(edges)) * j.size
        z = np.abs(obs - exp) / np.sqrt(exp)
        assert z.max() < 6, z.max()


def test_thin_lens_rays_meet_at_the_focal_point():
    """Thin lens: every lens sample of a pixel passes through the point at
    focalDist along the pinhole ray with the same jitter; lens positions fill
    a disk of radius sqrt(aperture) uniformly (E|a|^2 = aperture / 2)."""
    W, H = 32, 24
    ap, focal = 0.04, 2.5
    pin = _cam_opts(W, H, 50.0)
    lens = _cam_opts(W, H, 50.0, aperture=ap, focal=focal)
    pos = np.array([1.0, 2.0, 3.0])
    r2 = []
    for frame in range(1, 41):
        p = native.OracleScene.camera_rays(pin, frame)
        q = native.OracleScene.camera_rays(lens, frame)
        P = pos + focal * p[..., 3:6]
        off = q[..., 0:3] - pos
        assert np.all(np.abs(off[..., 2]) < 1e-12)
        r2.append((off ** 2).sum(-1))
        w = P - q[..., 0:3]
        dist = np.linalg.norm(w - (w * q[..., 3:6]).sum(-1, keepdims=True) * q[..., 3:6], axis=-1)
        assert dist.max() < 1e-9
    r2 = np.concatenate(r2).ravel()
    assert r2.max() <= ap + 1e-12
    assert abs(r2.mean() - ap / 2) < 5 * (ap / math.sqrt(12)) / math.sqrt(r2.size)
