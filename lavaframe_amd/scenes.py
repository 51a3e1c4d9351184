"""Synthetic scenes and ray sets of the benchmark configurations.

* cornell()          — Cornell box: 5 wall quads (10 tris), 2 rotated boxes
                       (24 tris), one quad light (17, 12, 4); diffuse albedos
                       0.73 / (0.65, 0.05, 0.05) / (0.12, 0.45, 0.15); 256x256,
                       depth 4.
* soup(n)            — n equilateral triangles, centres U[0,1]^3, random
                       rotation, edge U(0.002, 0.02); one mesh, identity
                       instance (the traversal microbenchmark, 1M triangles).
* traversal_rays()   — pinhole rays at 4096x2048 (pixel centres, eye
                       (0.5, 0.5, -1.5) -> (0.5, 0.5, 0.5), horizontal fov 40)
                       + incoherent rays (origins U[0,1]^3, directions
                       uniform on the sphere), tmax 1e6.
* instanced()        — procedural meshes (displaced spheres / tori) under random
                       affine instances with random Disney materials, quad
                       lights and a procedural HDR sky with a sun lobe.
* aperture_room()    — closed room lit only through a door aperture.

All are built through the Scene surface (lfh.Scene: AddMesh / AddMaterial /
AddLight / AddMeshInstance / CreateAccelerationStructures).
"""
from __future__ import annotations

import math

import numpy as np

from . import lfh

RAY_TMAX = 1.0e6


def _quad(p0, p1, p2, p3):
    """Two triangles (p0, p1, p2), (p0, p2, p3)."""
    return [p0, p1, p2, p0, p2, p3]


def _rot_y(deg):
    a = math.radians(deg)
    c, s = math.cos(a), math.sin(a)
    return np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]], np.float64)


def affine(rot3=None, scale=(1, 1, 1), translate=(0, 0, 0)) -> np.ndarray:
    """Column-major 4x4 (Mat4 data[col][row]) of T * R * S, flattened."""
    R = np.eye(3) if rot3 is None else np.asarray(rot3, np.float64)
    M = np.eye(4)
    M[:3, :3] = R @ np.diag(scale)
    M[:3, 3] = translate
    return np.ascontiguousarray(M.T, np.float32).reshape(16)


def unit_cube_tris() -> np.ndarray:
    """12 outward-wound triangles of [-0.5, 0.5]^3."""
    v = np.array([[x, y, z] for x in (-0.5, 0.5) for y in (-0.5, 0.5) for z in (-0.5, 0.5)], np.float32)
    idx = lambda x, y, z: (x > 0) * 4 + (y > 0) * 2 + (z > 0)  # noqa: E731
    faces = [  # outward normal, ccw seen from outside
        (idx(1, 0, 0), idx(1, 1, 0), idx(1, 1, 1), idx(1, 0, 1)),  # +x
        (idx(0, 0, 0), idx(0, 0, 1), idx(0, 1, 1), idx(0, 1, 0)),  # -x
        (idx(0, 1, 0), idx(0, 1, 1), idx(1, 1, 1), idx(1, 1, 0)),  # +y
        (idx(0, 0, 0), idx(1, 0, 0), idx(1, 0, 1), idx(0, 0, 1)),  # -y
        (idx(0, 0, 1), idx(1, 0, 1), idx(1, 1, 1), idx(0, 1, 1)),  # +z
        (idx(0, 0, 0), idx(0, 1, 0), idx(1, 1, 0), idx(1, 0, 0)),  # -z
    ]
    tris = []
    for a, b, c, d in faces:
        tris += _quad(v[a], v[b], v[c], v[d])
    return np.array(tris, np.float32).reshape(-1, 3, 3)


def cornell(width=256, height=256, max_depth=4, lib=None, build=True) -> lfh.Scene:
    """The Cornell configuration (unit box, open at z = 0, camera in front).
    build=False: flattened without host BVHs (device build by the context)."""
    s = lfh.Scene(lib)
    s.add_material(lfh.MATERIAL_DEFAULT)                         # 0 default
    white = s.add_material(lfh.material((0.73, 0.73, 0.73)))
    red = s.add_material(lfh.material((0.65, 0.05, 0.05)))
    green = s.add_material(lfh.material((0.12, 0.45, 0.15)))
    P = lambda x, y, z: np.array([x, y, z], np.float32)  # noqa: E731
    walls_white = (_quad(P(0, 0, 0), P(0, 0, 1), P(1, 0, 1), P(1, 0, 0))      # floor (+y)
                   + _quad(P(0, 1, 0), P(1, 1, 0), P(1, 1, 1), P(0, 1, 1))    # ceiling (-y)
                   + _quad(P(0, 0, 1), P(0, 1, 1), P(1, 1, 1), P(1, 0, 1)))   # back (-z)
    wall_red = _quad(P(0, 0, 0), P(0, 1, 0), P(0, 1, 1), P(0, 0, 1))          # left x = 0 (+x)
    wall_green = _quad(P(1, 0, 0), P(1, 0, 1), P(1, 1, 1), P(1, 1, 0))        # right x = 1 (-x)
    mw = s.add_mesh(np.array(walls_white, np.float32))
    mr = s.add_mesh(np.array(wall_red, np.float32))
    mg = s.add_mesh(np.array(wall_green, np.float32))
    mc = s.add_mesh(unit_cube_tris())
    eye = affine()
    s.add_instance(mw, eye, white)
    s.add_instance(mr, eye, red)
    s.add_instance(mg, eye, green)
    s.add_instance(mc, affine(_rot_y(-17), (0.3, 0.3, 0.3), (0.36, 0.15, 0.35)), white)
    s.add_instance(mc, affine(_rot_y(17), (0.3, 0.6, 0.3), (0.66, 0.30, 0.66)), white)
    s.add_light(lfh.quad_light((0.38, 0.999, 0.40), (0.62, 0.999, 0.40), (0.38, 0.999, 0.60), (17, 12, 4)))
    s.set_camera((0.5, 0.5, -1.35), (0.5, 0.5, 0.5), 40.0)
    s.set_options(width=width, height=height, tile_width=width, tile_height=height, max_depth=max_depth,
                  enable_rr=1, rr_depth=2, use_env_map=0, use_constant_bg=0, hdr_multiplier=1.0)
    return s.build() if build else s.flatten()


def soup_triangles(n, seed=0) -> np.ndarray:
    """n equilateral triangles, (n, 3, 3) float32."""
    rng = np.random.default_rng(seed)
    c = rng.random((n, 3))
    edge = rng.uniform(0.002, 0.02, n)
    # random rotations from unit quaternions
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    R = np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)], -1),
        np.stack([2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)], -1),
        np.stack([2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], -1),
    ], 1)
    ang = np.array([90.0, 210.0, 330.0]) * math.pi / 180
    base = np.stack([np.cos(ang), np.sin(ang), np.zeros(3)], -1) / math.sqrt(3)  # unit edge
    v = np.einsum("nij,kj->nki", R, base) * edge[:, None, None] + c[:, None, :]
    return v.astype(np.float32)


def soup(n, seed=0, lib=None, build=True) -> lfh.Scene:
    s = lfh.Scene(lib)
    s.add_material(lfh.MATERIAL_DEFAULT)
    m = s.add_mesh(soup_triangles(n, seed))
    s.add_instance(m, affine(), 0)
    return s.build() if build else s.flatten()


def camera_basis(eye, lookat):
    """position/right/up/forward of the orbit camera (Camera.cpp:101-169),
    evaluated in float64."""
    eye = np.asarray(eye, np.float64)
    d = np.asarray(lookat, np.float64) - eye
    d /= np.linalg.norm(d)
    pitch = math.asin(d[1])
    yaw = math.atan2(d[2], d[0])
    f = np.array([math.cos(yaw) * math.cos(pitch), math.sin(pitch), math.sin(yaw) * math.cos(pitch)])
    f /= np.linalg.norm(f)
    r = np.cross(f, [0, 1, 0])
    r /= np.linalg.norm(r)
    u = np.cross(r, f)
    u /= np.linalg.norm(u)
    return eye, r, u, f


def pinhole_rays(width, height, eye=(0.5, 0.5, -1.5), lookat=(0.5, 0.5, 0.5), fov_deg=40.0) -> np.ndarray:
    """Pixel-centre pinhole rays, horizontal fov (renderer.glsl:41-52 without jitter)."""
    pos, r, u, f = camera_basis(eye, lookat)
    scale = math.tan(math.radians(fov_deg) * 0.5)
    xs = (np.arange(width) + 0.5) / width * 2 - 1
    ys = (np.arange(height) + 0.5) / height * 2 - 1
    X, Y = np.meshgrid(xs * scale, ys * scale * height / width)
    d = X[..., None] * r + Y[..., None] * u + f
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    rays = np.zeros(width * height, lfh.RAY_DTYPE)
    rays["org"] = pos.astype(np.float32)
    rays["dir"] = d.reshape(-1, 3).astype(np.float32)
    rays["tmin"] = 0.0
    rays["tmax"] = RAY_TMAX
    return rays


def incoherent_rays(n, seed=1) -> np.ndarray:
    rng = np.random.default_rng(seed)
    o = rng.random((n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros(n, lfh.RAY_DTYPE)
    rays["org"] = o.astype(np.float32)
    rays["dir"] = d.astype(np.float32)
    rays["tmin"] = 0.0
    rays["tmax"] = RAY_TMAX
    return rays


def traversal_rays(width=4096, height=2048, n_incoherent=None, seed=1) -> tuple[np.ndarray, np.ndarray]:
    n_incoherent = width * height if n_incoherent is None else n_incoherent
    return pinhole_rays(width, height), incoherent_rays(n_incoherent, seed)


# ---------------------------------------------------------------------------
# instanced production-like scene
# ---------------------------------------------------------------------------
def _sphere_mesh(n_sub, rng, displace=0.15):
    """Displaced UV sphere with about 2 * n_sub^2 triangles."""
    nu, nv = 2 * n_sub, n_sub
    th = np.linspace(0, math.pi, nv + 1)
    ph = np.linspace(0, 2 * math.pi, nu + 1)
    TH, PH = np.meshgrid(th, ph, indexing="ij")
    k = rng.integers(2, 7, 3)
    rad = 1 + displace * np.sin(k[0] * TH) * np.cos(k[1] * PH) * np.sin(k[2] * (TH + PH))
    P = np.stack([rad * np.sin(TH) * np.cos(PH), rad * np.cos(TH), rad * np.sin(TH) * np.sin(PH)], -1)
    return _grid_tris(P)


def _torus_mesh(n_sub, rng, R=1.0, r=0.35, displace=0.08):
    nu, nv = 2 * n_sub, n_sub
    u = np.linspace(0, 2 * math.pi, nu + 1)
    v = np.linspace(0, 2 * math.pi, nv + 1)
    V, U = np.meshgrid(v, u, indexing="ij")
    k = rng.integers(2, 9, 2)
    rr = r * (1 + displace * np.sin(k[0] * U) * np.sin(k[1] * V))
    P = np.stack([(R + rr * np.cos(V)) * np.cos(U), rr * np.sin(V), (R + rr * np.cos(V)) * np.sin(U)], -1)
    return _grid_tris(P)


def _grid_tris(P):
    a = P[:-1, :-1]
    b = P[:-1, 1:]
    c = P[1:, 1:]
    d = P[1:, :-1]
    t1 = np.stack([a, b, c], -2).reshape(-1, 3, 3)
    t2 = np.stack([a, c, d], -2).reshape(-1, 3, 3)
    t = np.concatenate([t1, t2]).astype(np.float32)
    # drop degenerate triangles at the poles
    area = np.linalg.norm(np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 0]), axis=1)
    return t[area > 1e-12]


def random_material(rng) -> np.ndarray:
    kw = {"roughness": float(rng.uniform(0.05, 1.0))}
    alb = rng.uniform(0.05, 0.95, 3)
    r = rng.random()
    if r < 0.3:
        kw["metallic"] = 1.0
    elif r < 0.5:
        kw["transmission"] = 1.0
        kw["ior"] = 1.5
    if rng.random() < 0.1:
        kw["clearcoat"] = 1.0
        kw["clearcoatRoughness"] = float(rng.random())
    return lfh.material(tuple(alb), **kw)


def sky_hdr(width=2048, height=1024, sun_dir=(0.3, 0.8, 0.5), sun_power=2000.0) -> np.ndarray:
    """Procedural equirect sky (rows top to bottom, theta = pi * (row + 0.5) / H)
    with a sun lobe, as an (H, W, 3) float32 image."""
    th = (np.arange(height) + 0.5) / height * math.pi
    ph = (np.arange(width) + 0.5) / width * 2 * math.pi
    TH, PH = np.meshgrid(th, ph, indexing="ij")
    # direction convention of EnvSample: (-sin t cos p, cos t, -sin t sin p)
    d = np.stack([-np.sin(TH) * np.cos(PH), np.cos(TH), -np.sin(TH) * np.sin(PH)], -1)
    up = np.clip(d[..., 1], -1, 1)
    sky = np.stack([0.35 + 0.25 * up, 0.45 + 0.3 * up, 0.7 + 0.3 * up], -1).clip(0.02)
    ground = np.array([0.15, 0.13, 0.1])
    img = np.where((up > 0)[..., None], sky, ground)
    s = np.asarray(sun_dir, np.float64)
    s /= np.linalg.norm(s)
    cosang = d @ s
    img = img + sun_power * np.exp((cosang - 1) * 4000.0)[..., None] * np.array([1.0, 0.9, 0.75])
    return img.astype(np.float32)


def instanced(n_meshes=64, n_instances=1024, min_tris=1000, max_tris=200000, width=1920, height=1080,
              max_depth=8, hdr_size=(2048, 1024), seed=0, lib=None, light_count=8, build=True) -> lfh.Scene:
    rng = np.random.default_rng(seed)
    s = lfh.Scene(lib)
    s.add_material(lfh.MATERIAL_DEFAULT)
    # log-uniform triangle counts
    counts = np.exp(rng.uniform(math.log(min_tris), math.log(max_tris), n_meshes)).astype(int)
    meshes = []
    for i, nt in enumerate(counts):
        n_sub = max(4, int(math.sqrt(nt / 4.0)))
        tris = _sphere_mesh(n_sub, rng) if i % 2 == 0 else _torus_mesh(n_sub, rng)
        meshes.append(s.add_mesh(tris))
    mats = [s.add_material(random_material(rng)) for _ in range(n_instances)]
    floor = s.add_mesh(np.array(_quad(*[np.array(p, np.float32) for p in
                                        [(-40, 0, -40), (-40, 0, 40), (40, 0, 40), (40, 0, -40)]]), np.float32))
    for i in range(n_instances):
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        w, x, y, z = q
        R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                      [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                      [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
        sc = rng.uniform(0.3, 1.2) * np.array([1.0, rng.uniform(0.6, 1.4), 1.0])
        t = (rng.uniform(-30, 30), rng.uniform(0.8, 10), rng.uniform(-30, 30))
        s.add_instance(meshes[i % n_meshes], affine(R, sc, t), mats[i])
    s.add_instance(floor, affine(), s.add_material(lfh.material((0.5, 0.5, 0.5), roughness=0.8)))
    for k in range(light_count):
        c = np.array([rng.uniform(-25, 25), rng.uniform(12, 18), rng.uniform(-25, 25)], np.float32)
        a = rng.uniform(1.5, 3.0)
        s.add_light(lfh.quad_light(c, c + np.array([a, 0, 0], np.float32), c + np.array([0, 0, a], np.float32),
                                   tuple(rng.uniform(5, 25, 3))))
    if hdr_size:
        s.set_hdr(sky_hdr(*hdr_size))
    s.set_camera((0.0, 12.0, -55.0), (0.0, 3.0, 0.0), 50.0)
    s.set_options(width=width, height=height, tile_width=width, tile_height=height, max_depth=max_depth,
                  enable_rr=1, rr_depth=2, use_env_map=1 if hdr_size else 0, use_constant_bg=0, hdr_multiplier=1.0)
    return s.build() if build else s.flatten()


def aperture_room(width=1280, height=720, max_depth=10, target_tris=200000, seed=0, lib=None,
                  build=True) -> lfh.Scene:
    """Closed room (x, z in [0, 4], y in [0, 3]) whose only light sits in an
    annex behind the wall x = 4; the annex is seen through a door of 5% of
    that wall's area.  Furniture of displaced spheres fills the triangle
    budget."""
    rng = np.random.default_rng(seed)
    s = lfh.Scene(lib)
    s.add_material(lfh.MATERIAL_DEFAULT)
    wall = s.add_material(lfh.material((0.7, 0.7, 0.7), roughness=0.9))
    P = lambda x, y, z: np.array([x, y, z], np.float32)  # noqa: E731
    tris = []
    tris += _quad(P(0, 0, 0), P(0, 0, 4), P(4, 0, 4), P(4, 0, 0))          # floor
    tris += _quad(P(0, 3, 0), P(4, 3, 0), P(4, 3, 4), P(0, 3, 4))          # ceiling
    tris += _quad(P(0, 0, 0), P(4, 0, 0), P(4, 3, 0), P(0, 3, 0))          # z = 0
    tris += _quad(P(0, 0, 4), P(0, 3, 4), P(4, 3, 4), P(4, 0, 4))          # z = 4
    tris += _quad(P(0, 0, 0), P(0, 3, 0), P(0, 3, 4), P(0, 0, 4))          # x = 0
    # wall x = 4 with a door: door width 0.6 x height 1.0 = 0.6 = 5% of 12
    z0, z1, h = 1.7, 2.3, 1.0
    tris += _quad(P(4, 0, 0), P(4, 0, z0), P(4, 3, z0), P(4, 3, 0))
    tris += _quad(P(4, 0, z1), P(4, 0, 4), P(4, 3, 4), P(4, 3, z1))
    tris += _quad(P(4, h, z0), P(4, h, z1), P(4, 3, z1), P(4, 3, z0))
    # annex x in [4.02, 6], every face towards its inside; the dividing wall is
    # 2 cm thick, its annex face at x = 4.02 with the same door, and jambs line
    # the door opening
    xa = 4.02
    tris += _quad(P(xa, 0, 0), P(xa, 0, 4), P(6, 0, 4), P(6, 0, 0))          # annex floor (+y)
    tris += _quad(P(xa, 3, 0), P(6, 3, 0), P(6, 3, 4), P(xa, 3, 4))          # annex ceiling (-y)
    tris += _quad(P(6, 0, 0), P(6, 0, 4), P(6, 3, 4), P(6, 3, 0))            # x = 6 (-x)
    tris += _quad(P(xa, 0, 0), P(6, 0, 0), P(6, 3, 0), P(xa, 3, 0))          # z = 0 (+z)
    tris += _quad(P(xa, 0, 4), P(xa, 3, 4), P(6, 3, 4), P(6, 0, 4))          # z = 4 (-z)
    tris += _quad(P(xa, 0, 0), P(xa, 3, 0), P(xa, 3, z0), P(xa, 0, z0))      # annex face of the wall (+x)
    tris += _quad(P(xa, 0, z1), P(xa, 3, z1), P(xa, 3, 4), P(xa, 0, 4))
    tris += _quad(P(xa, h, z0), P(xa, 3, z0), P(xa, 3, z1), P(xa, h, z1))
    tris += _quad(P(4, 0, z0), P(xa, 0, z0), P(xa, h, z0), P(4, h, z0))      # jamb z = z0 (+z)
    tris += _quad(P(4, 0, z1), P(4, h, z1), P(xa, h, z1), P(xa, 0, z1))      # jamb z = z1 (-z)
    tris += _quad(P(4, h, z0), P(xa, h, z0), P(xa, h, z1), P(4, h, z1))      # lintel (-y)
    room = s.add_mesh(np.array(tris, np.float32))
    s.add_instance(room, affine(), wall)
    per = max(1000, target_tris // 8)
    n_sub = max(4, int(math.sqrt(per / 4.0)))
    for i in range(8):
        m = s.add_mesh(_sphere_mesh(n_sub, rng, 0.1))
        mat = s.add_material(lfh.material(tuple(rng.uniform(0.2, 0.8, 3)), roughness=float(rng.uniform(0.3, 1.0))))
        c = (rng.uniform(0.6, 3.4), rng.uniform(0.3, 0.6), rng.uniform(0.6, 3.4))
        s.add_instance(m, affine(None, (0.3, 0.3, 0.3), c), mat)
    # light in the annex, facing down
    s.add_light(lfh.quad_light((5.0, 2.95, 1.5), (5.6, 2.95, 1.5), (5.0, 2.95, 2.5), (40, 38, 34)))
    s.set_camera((0.4, 1.6, 0.4), (3.0, 1.0, 3.0), 70.0)
    s.set_options(width=width, height=height, tile_width=width, tile_height=height, max_depth=max_depth,
                  enable_rr=1, rr_depth=2, use_env_map=0, use_constant_bg=0, hdr_multiplier=1.0)
    return s.build() if build else s.flatten()


def _quad_uv(u0, v0, u1, v1):
    """Texture coordinates of _quad's six vertices for a quad spanning
    [u0, u1] x [v0, v1] (p0 at (u0, v0), p2 at (u1, v1))."""
    return [(u0, v0), (u1, v0), (u1, v1), (u0, v0), (u1, v1), (u0, v1)]


def texture_maps(size=32):
    """Four RGBA8 maps of one size: checker albedo, metallic (R) / roughness
    (G) ramps, a tangent-space normal map of sinusoidal bumps, an emission
    ring."""
    yy, xx = np.mgrid[0:size, 0:size]
    chk = ((xx // 4 + yy // 4) % 2)[..., None]
    albedo = np.where(chk, np.array([230, 200, 120, 255]), np.array([60, 90, 200, 255])).astype(np.uint8)
    mr = np.stack([xx * 255 // (size - 1), 40 + yy * 215 // (size - 1), np.zeros_like(xx),
                   np.full_like(xx, 255)], -1).astype(np.uint8)
    n = np.stack([0.45 * np.sin(xx * 2 * math.pi / 8), 0.45 * np.cos(yy * 2 * math.pi / 8), np.ones_like(xx, float)], -1)
    n /= np.linalg.norm(n, axis=-1, keepdims=True)
    nm = np.concatenate([np.round((n * 0.5 + 0.5) * 255), np.full(n.shape[:2] + (1,), 255)], -1).astype(np.uint8)
    r = np.hypot(xx - size / 2 + 0.5, yy - size / 2 + 0.5)
    ring = ((r > size * 0.2) & (r < size * 0.35))[..., None]
    em = np.where(ring, np.array([255, 180, 60, 255]), np.array([10, 10, 30, 255])).astype(np.uint8)
    return albedo, mr, nm, em


def textured_room(width=64, height=64, max_depth=4, lib=None, build=True) -> lfh.Scene:
    """Cornell-like box with texture maps: a checker albedo on the floor (its
    texture coordinates run over [0, 3], so the maps repeat) and on the back
    wall (coordinates from -0.5, negative ones wrap), a normal map on the
    floor, metallic-roughness maps on a box, an emission map on a panel."""
    s = lfh.Scene(lib)
    s.add_material(lfh.MATERIAL_DEFAULT)
    albedo, mr, nm, em = texture_maps()
    ta, tm, tn, te = (s.add_texture(a, n) for a, n in ((albedo, "albedo"), (mr, "mr"), (nm, "normal"),
                                                         (em, "emission")))
    def mat(texids, **kw):
        m = lfh.material(**kw)
        m[24:28] = texids
        return s.add_material(m)
    floor_m = mat((ta, -1, tn, -1), albedo=(0.9, 0.9, 0.9), roughness=0.4)
    back_m = mat((ta, -1, -1, -1), albedo=(0.8, 0.8, 0.8))
    red = s.add_material(lfh.material((0.65, 0.05, 0.05)))
    green = s.add_material(lfh.material((0.12, 0.45, 0.15)))
    white = s.add_material(lfh.material((0.73, 0.73, 0.73)))
    box_m = mat((-1, tm, -1, -1), albedo=(0.8, 0.6, 0.4))
    panel_m = mat((-1, -1, -1, te), albedo=(0.5, 0.5, 0.5))
    P = lambda x, y, z: np.array([x, y, z], np.float32)  # noqa: E731
    floor = s.add_mesh(np.array(_quad(P(0, 0, 0), P(0, 0, 1), P(1, 0, 1), P(1, 0, 0)), np.float32),
                       uvs=np.array(_quad_uv(0, 0, 3, 3), np.float32))
    back = s.add_mesh(np.array(_quad(P(0, 0, 1), P(0, 1, 1), P(1, 1, 1), P(1, 0, 1)), np.float32),
                      uvs=np.array(_quad_uv(-0.5, -0.5, 1.5, 1.5), np.float32))
    ceil = s.add_mesh(np.array(_quad(P(0, 1, 0), P(1, 1, 0), P(1, 1, 1), P(0, 1, 1)), np.float32))
    left = s.add_mesh(np.array(_quad(P(0, 0, 0), P(0, 1, 0), P(0, 1, 1), P(0, 0, 1)), np.float32))
    right = s.add_mesh(np.array(_quad(P(1, 0, 0), P(1, 0, 1), P(1, 1, 1), P(1, 1, 0)), np.float32))
    cube_uv = np.array([_quad_uv(0, 0, 1, 1)] * 6, np.float32).reshape(-1, 2)
    cube = s.add_mesh(unit_cube_tris(), uvs=cube_uv)
    panel = s.add_mesh(np.array(_quad(P(0.05, 0.45, 0.5), P(0.05, 0.85, 0.5), P(0.05, 0.85, 0.9),
                                      P(0.05, 0.45, 0.9)), np.float32), uvs=np.array(_quad_uv(0, 0, 1, 1), np.float32))
    eye = affine()
    s.add_instance(floor, eye, floor_m)
    s.add_instance(back, eye, back_m)
    s.add_instance(ceil, eye, white)
    s.add_instance(left, eye, red)
    s.add_instance(right, eye, green)
    s.add_instance(cube, affine(_rot_y(-17), (0.3, 0.3, 0.3), (0.36, 0.15, 0.35)), box_m)
    s.add_instance(panel, eye, panel_m)
    s.add_light(lfh.quad_light((0.38, 0.999, 0.40), (0.62, 0.999, 0.40), (0.38, 0.999, 0.60), (17, 12, 4)))
    s.set_camera((0.5, 0.5, -1.35), (0.5, 0.5, 0.5), 40.0)
    s.set_options(width=width, height=height, tile_width=width, tile_height=height, max_depth=max_depth,
                  enable_rr=1, rr_depth=2, use_env_map=0, use_constant_bg=0, hdr_multiplier=1.0)
    return s.build() if build else s.flatten()
