"""ctypes binding of the lfh C ABI (include/lfh.h).

The library is the in-tree ``lavaframe_amd/_lib/liblfh.so`` built by
``lavaframe_amd/build.py`` (hipcc, gfx950).  Loading fails loudly when it is
missing: there is no fallback path.

torch is imported before the library is loaded so that the process holds a
single HIP runtime (torch's), which the library then shares.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

PKG_DIR = Path(__file__).resolve().parent
LIB_PATH = PKG_DIR / "_lib" / "liblfh.so"

LFH_OK = 0
ERRORS = {
    -1: "LFH_E_INVALID", -2: "LFH_E_OOM", -3: "LFH_E_HIP", -4: "LFH_E_STATE",
    -5: "LFH_E_IO", -6: "LFH_E_UNSUPPORTED", -7: "LFH_E_BUDGET",
}

f32p = C.POINTER(C.c_float)
i32p = C.POINTER(C.c_int32)
u32p = C.POINTER(C.c_uint32)
u8p = C.POINTER(C.c_uint8)
i64p = C.POINTER(C.c_int64)


class Ray(C.Structure):
    _fields_ = [("org", C.c_float * 3), ("tmin", C.c_float), ("dir", C.c_float * 3), ("tmax", C.c_float)]


class Hit(C.Structure):
    _fields_ = [("prim", C.c_int32), ("inst", C.c_int32), ("mat", C.c_int32),
                ("t", C.c_float), ("u", C.c_float), ("v", C.c_float)]


RAY_DTYPE = np.dtype([("org", "<f4", 3), ("tmin", "<f4"), ("dir", "<f4", 3), ("tmax", "<f4")])
HIT_DTYPE = np.dtype([("prim", "<i4"), ("inst", "<i4"), ("mat", "<i4"), ("t", "<f4"), ("u", "<f4"), ("v", "<f4")])
assert RAY_DTYPE.itemsize == 32 and HIT_DTYPE.itemsize == 24


class SceneDesc(C.Structure):
    _fields_ = [
        ("nodes", f32p), ("num_nodes", C.c_int64), ("top_level_index", C.c_int32),
        ("vert_indices", i32p), ("num_tris", C.c_int64),
        ("vertices_uvx", f32p), ("normals_uvy", f32p), ("num_vertices", C.c_int64),
        ("materials", f32p), ("num_materials", C.c_int32),
        ("transforms", f32p), ("num_instances", C.c_int32),
        ("lights", f32p), ("num_lights", C.c_int32),
        ("hdr_width", C.c_int32), ("hdr_height", C.c_int32),
        ("hdr_cols", f32p), ("hdr_marginal", f32p), ("hdr_conditional", f32p),
        ("mesh_tri_offsets", i64p), ("num_meshes", C.c_int32),
        ("instance_meshes", i32p), ("instance_materials", i32p),
        ("textures", u8p), ("tex_width", C.c_int32), ("tex_height", C.c_int32), ("num_textures", C.c_int32),
    ]


class Options(C.Structure):
    _fields_ = [
        ("width", C.c_int32), ("height", C.c_int32), ("tile_width", C.c_int32), ("tile_height", C.c_int32),
        ("max_depth", C.c_int32), ("enable_rr", C.c_int32), ("rr_depth", C.c_int32),
        ("use_env_map", C.c_int32), ("use_constant_bg", C.c_int32), ("bg_color", C.c_float * 3),
        ("hdr_multiplier", C.c_float), ("kahan", C.c_int32), ("guiding", C.c_int32),
        ("estimator", C.c_int32), ("frame_offset", C.c_int32), ("reserved", C.c_int32 * 4),
    ]


class Camera(C.Structure):
    _fields_ = [("position", C.c_float * 3), ("right", C.c_float * 3), ("up", C.c_float * 3),
                ("forward", C.c_float * 3), ("fov", C.c_float), ("focal_dist", C.c_float),
                ("aperture", C.c_float)]


class Postprocess(C.Structure):
    _fields_ = [("tonemap_index", C.c_int32), ("use_ca", C.c_int32), ("use_ca_distortion", C.c_int32),
                ("ca_distance", C.c_float), ("ca_p1", C.c_float), ("ca_p2", C.c_float), ("ca_p3", C.c_float),
                ("use_vignette", C.c_int32), ("vignette_intensity", C.c_float), ("vignette_power", C.c_float)]

    @classmethod
    def make(cls, tonemap_index=0, use_ca=0, use_ca_distortion=0, ca_distance=0.05, ca_p1=5.0, ca_p2=-0.5,
             ca_p3=0.5, use_vignette=0, vignette_intensity=0.0, vignette_power=1.0):
        """Options with the reference defaults (Renderer.h:20-44) unless given."""
        return cls(tonemap_index, use_ca, use_ca_distortion, ca_distance, ca_p1, ca_p2, ca_p3, use_vignette,
                   vignette_intensity, vignette_power)

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class JobDesc(C.Structure):
    _fields_ = [("scene", C.POINTER(SceneDesc)), ("options", Options), ("camera", Camera),
                ("target_spp", C.c_int32), ("priority", C.c_int32), ("user_tag", C.c_int64)]


class JobStatus(C.Structure):
    _fields_ = [("job_id", C.c_int64), ("state", C.c_int32), ("spp_done", C.c_int32),
                ("t_submit", C.c_double), ("t_admit", C.c_double), ("t_done", C.c_double),
                ("device_bytes", C.c_int64)]


_P = C.c_void_p
_SIGS = {
    "lfh_abi_version": (C.c_int, []),
    "lfh_last_error": (C.c_char_p, [_P]),
    "lfh_device_memory": (C.c_int, [i64p, i64p]),
    "lfh_scene_create": (C.c_int, [C.POINTER(_P)]),
    "lfh_scene_destroy": (None, [_P]),
    "lfh_scene_load": (C.c_int, [_P, C.c_char_p]),
    "lfh_scene_add_mesh_file": (C.c_int, [_P, C.c_char_p, i32p]),
    "lfh_scene_add_mesh": (C.c_int, [_P, f32p, f32p, f32p, C.c_int64, i32p]),
    "lfh_scene_add_material": (C.c_int, [_P, f32p, i32p]),
    "lfh_scene_add_light": (C.c_int, [_P, f32p, i32p]),
    "lfh_scene_add_instance": (C.c_int, [_P, C.c_int32, f32p, C.c_int32, i32p]),
    "lfh_scene_set_camera": (C.c_int, [_P, f32p, f32p, C.c_float, C.c_float, C.c_float]),
    "lfh_scene_set_hdr": (C.c_int, [_P, f32p, C.c_int32, C.c_int32]),
    "lfh_scene_set_options": (C.c_int, [_P, C.POINTER(Options)]),
    "lfh_scene_get_options": (C.c_int, [_P, C.POINTER(Options)]),
    "lfh_scene_get_camera": (C.c_int, [_P, C.POINTER(Camera)]),
    "lfh_scene_get_postprocess": (C.c_int, [_P, C.POINTER(Postprocess)]),
    "lfh_get_output_post": (C.c_int, [_P, C.POINTER(Postprocess), C.c_float, f32p, C.c_int32, C.c_int32]),
    "lfh_get_output_ldr": (C.c_int, [_P, C.POINTER(Postprocess), C.c_float, u8p, C.c_int32, C.c_int32]),
    "lfh_scene_build": (C.c_int, [_P]),
    "lfh_scene_set_instance_transform": (C.c_int, [_P, C.c_int32, f32p]),
    "lfh_scene_rebuild_instances": (C.c_int, [_P]),
    "lfh_scene_get_desc": (C.c_int, [_P, C.POINTER(SceneDesc)]),
    "lfh_scene_flatten": (C.c_int, [_P]),
    "lfh_scene_add_texture": (C.c_int, [_P, u8p, C.c_int32, C.c_int32, C.c_char_p, i32p]),
    "lfh_scene_add_texture_file": (C.c_int, [_P, C.c_char_p, i32p]),
    "lfh_scene_counts": (C.c_int, [_P, i32p, i32p]),
    "lfh_create": (C.c_int, [C.POINTER(SceneDesc), C.POINTER(Options), C.POINTER(_P)]),
    "lfh_destroy": (None, [_P]),
    "lfh_set_stream": (C.c_int, [_P, _P]),
    "lfh_synchronize": (C.c_int, [_P]),
    "lfh_update_instances": (C.c_int, [_P, f32p, f32p, f32p, C.c_int64, C.c_int64]),
    "lfh_update": (C.c_int, [_P, C.POINTER(Camera), C.c_int]),
    "lfh_render": (C.c_int, [_P]),
    "lfh_get_sample_count": (C.c_int, [_P, i32p]),
    "lfh_get_frame": (C.c_int, [_P, i32p]),
    "lfh_get_progress": (C.c_int, [_P, f32p]),
    "lfh_get_output_hdr": (C.c_int, [_P, f32p, C.c_int32, C.c_int32]),
    "lfh_get_accum": (C.c_int, [_P, f32p, f32p, C.c_int32, C.c_int32]),
    "lfh_render_passes": (C.c_int, [_P, C.c_int32]),
    "lfh_render_preview": (C.c_int, [_P, C.c_float, C.c_int32, f32p, i32p, i32p]),
    "lfh_trace_closest": (C.c_int, [_P, _P, C.c_int64, _P]),
    "lfh_trace_any": (C.c_int, [_P, _P, C.c_int64, _P]),
    "lfh_trace_closest_device": (C.c_int, [_P, _P, C.c_int64, _P]),
    "lfh_trace_any_device": (C.c_int, [_P, _P, C.c_int64, _P]),
    "lfh_trace_closest_stats_device": (C.c_int, [_P, _P, C.c_int64, _P, _P]),
    "lfh_eval_bsdf": (C.c_int, [_P, f32p, f32p, C.c_int64, f32p, f32p]),
    "lfh_sample_bsdf": (C.c_int, [_P, f32p, u32p, C.c_int64, f32p, f32p, f32p]),
    "lfh_sample_light": (C.c_int, [_P, f32p, u32p, C.c_int64, f32p, f32p, f32p, f32p, f32p]),
    "lfh_sample_env": (C.c_int, [_P, u32p, C.c_int64, f32p, f32p, f32p]),
    "lfh_env_pdf": (C.c_int, [_P, f32p, C.c_int64, f32p]),
    "lfh_path_trace_rays": (C.c_int, [_P, f32p, u32p, C.c_int64, f32p]),
    "lfh_path_trace_rays_device_noalloc": (C.c_int, [_P, _P, _P, C.c_int64, _P]),
    "lfh_server_create": (C.c_int, [C.c_int64, C.c_int32, C.POINTER(_P)]),
    "lfh_server_destroy": (None, [_P]),
    "lfh_server_submit": (C.c_int, [_P, C.POINTER(JobDesc), i64p]),
    "lfh_server_step": (C.c_int, [_P, i32p]),
    "lfh_server_status": (C.c_int, [_P, C.c_int64, C.POINTER(JobStatus)]),
    "lfh_server_result": (C.c_int, [_P, C.c_int64, f32p, C.c_int32, C.c_int32]),
    "lfh_server_release": (C.c_int, [_P, C.c_int64]),
    "lfh_server_clock": (C.c_double, [_P]),
    "lfh_server_job_error": (C.c_char_p, [_P, C.c_int64]),
    "lfh_debug_fail_next": (C.c_int, [C.c_int32, C.c_int32]),
    "lfh_guide_create": (C.c_int, [f32p, f32p, C.c_uint64, C.c_float, C.POINTER(_P)]),
    "lfh_guide_destroy": (None, [_P]),
    "lfh_guide_get_params": (C.c_int, [_P, f32p, i64p]),
    "lfh_guide_set_params": (C.c_int, [_P, f32p, C.c_int64]),
    "lfh_guide_forward": (C.c_int, [_P, f32p, C.c_int64, f32p]),
    "lfh_guide_forward_device": (C.c_int, [_P, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p]),
    "lfh_guide_loss_grad": (C.c_int, [_P, f32p, f32p, f32p, C.c_int64, f32p, f32p]),
    "lfh_guide_train_step": (C.c_int, [_P, f32p, f32p, f32p, C.c_int64, f32p]),
    "lfh_set_guide": (C.c_int, [_P, _P, C.c_int32]),
    "lfh_guide_steps": (C.c_int, [_P, i64p]),
    "lfh_get_ray_counts": (C.c_int, [_P, i64p, i64p]),
    "lfh_set_frame_offset": (C.c_int, [_P, C.c_int32]),
    "lfh_server_set_policy": (C.c_int, [_P, C.c_int32, C.c_double, C.c_double]),
    "lfh_server_reset_guides": (C.c_int, [_P]),
    "lfh_guide_sample": (C.c_int, [_P, f32p, u32p, C.c_int64, f32p, f32p]),
    "lfh_guide_pdf": (C.c_int, [_P, f32p, f32p, C.c_int64, f32p]),
    "lfh_guide_samples_seen": (C.c_int, [_P, i64p]),
}

_lib = None


def header_symbols() -> list[str]:
    """Function names declared in include/lfh.h."""
    import re
    text = (PKG_DIR.parent / "include" / "lfh.h").read_text()
    return sorted(set(re.findall(r"\b(lfh_[a-z0-9_]+)\s*\(", text)))


def load(path: str | os.PathLike | None = None):
    """Load liblfh.so (raises if it is missing)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    try:  # one HIP runtime per process: torch's
        import torch  # noqa: F401
    except Exception:  # pragma: no cover - torch is part of the image
        pass
    p = Path(path) if path else Path(os.environ.get("LFH_LIBRARY", LIB_PATH))
    if not p.exists():
        raise RuntimeError(f"lfh native library not built: {p} (run lavaframe_amd/build.py)")
    lib = C.CDLL(str(p))
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name, None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def debug_fail_next(site: int, count: int = 1, lib=None):
    """Testing hook: the next `count` operations at `site` fail (1: render
    path-state allocation, 2: serving scene upload, 3: device TLAS rebuild)."""
    lib = lib or load()
    check(lib.lfh_debug_fail_next(int(site), int(count)), None, lib)


class LfhError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


def check(code, ctx=None, lib=None):
    lib = lib or load()
    if code != LFH_OK:
        msg = lib.lfh_last_error(ctx)
        raise LfhError(code, msg.decode() if msg else "")
    return code


def fptr(a: np.ndarray):
    assert a.dtype == np.float32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(f32p)


def iptr(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(i32p)


def uptr(a: np.ndarray):
    assert a.dtype == np.uint32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(u32p)


def vptr(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"]
    return C.c_void_p(a.ctypes.data)


MATERIAL_DEFAULT = np.array(
    [1, 1, 1, 0.5,  0, 0, 0, 0,  0, 0.5, 0, 0,  0, 0, 0, 0,  0, 1.45, 1, 0,  1, 1, 1, 0,  -1, -1, -1, -1],
    dtype=np.float32)


def material(albedo=(1, 1, 1), **kw) -> np.ndarray:
    """28-float material block (Material.h field order and defaults)."""
    m = MATERIAL_DEFAULT.copy()
    m[0:3] = albedo
    idx = {"specular": 3, "emission": (4, 7), "anisotropic": 7, "metallic": 8, "roughness": 9,
           "subsurface": 10, "specularTint": 11, "sheen": 12, "sheenTint": 13, "clearcoat": 14,
           "clearcoatRoughness": 15, "transmission": 16, "ior": 17, "atDistance": 18,
           "extinction": (20, 23)}
    for k, v in kw.items():
        i = idx[k]
        if isinstance(i, tuple):
            m[i[0]:i[1]] = v
        else:
            m[i] = v
    return m


def quad_light(position, v1, v2, emission) -> np.ndarray:
    """Light block of a `light { type Quad }` (Loader.cpp:175-181)."""
    p = np.asarray(position, np.float32)
    u = np.asarray(v1, np.float32) - p
    v = np.asarray(v2, np.float32) - p
    cr = np.cross(u.astype(np.float32), v.astype(np.float32)).astype(np.float32)
    area = np.float32(np.sqrt(np.float32(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2])))
    return np.concatenate([p, np.asarray(emission, np.float32), u, v,
                           np.array([0.0, area, 0.0], np.float32)]).astype(np.float32)


class Scene:
    """Host scene builder: the Scene surface (AddMesh/AddMaterial/AddLight/
    AddMeshInstance, the .scene loader, CreateAccelerationStructures)."""

    def __init__(self, lib=None):
        self.lib = lib or load()
        h = C.c_void_p()
        check(self.lib.lfh_scene_create(C.byref(h)), lib=self.lib)
        self.h = h
        self._keep = []

    def close(self):
        if self.h:
            self.lib.lfh_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, code):
        return check(code, None, self.lib)

    def load(self, path):
        self._chk(self.lib.lfh_scene_load(self.h, str(path).encode()))
        return self

    def add_mesh_file(self, path) -> int:
        i = C.c_int32()
        self._chk(self.lib.lfh_scene_add_mesh_file(self.h, str(path).encode(), C.byref(i)))
        return i.value

    def add_mesh(self, positions, normals=None, uvs=None) -> int:
        pos = np.ascontiguousarray(positions, np.float32).reshape(-1, 9)
        n = None if normals is None else np.ascontiguousarray(normals, np.float32).reshape(-1)
        uv = None if uvs is None else np.ascontiguousarray(uvs, np.float32).reshape(-1)
        i = C.c_int32()
        self._chk(self.lib.lfh_scene_add_mesh(self.h, fptr(pos), None if n is None else fptr(n),
                                              None if uv is None else fptr(uv), pos.shape[0], C.byref(i)))
        return i.value

    def add_material(self, m28) -> int:
        m = np.ascontiguousarray(m28, np.float32)
        assert m.size == 28
        i = C.c_int32()
        self._chk(self.lib.lfh_scene_add_material(self.h, fptr(m), C.byref(i)))
        return i.value

    def add_light(self, l15) -> int:
        m = np.ascontiguousarray(l15, np.float32)
        assert m.size == 15
        i = C.c_int32()
        self._chk(self.lib.lfh_scene_add_light(self.h, fptr(m), C.byref(i)))
        return i.value

    def add_instance(self, mesh_id, transform=None, material_id=0) -> int:
        t = np.eye(4, dtype=np.float32) if transform is None else np.ascontiguousarray(transform, np.float32)
        t = np.ascontiguousarray(t.reshape(16))
        i = C.c_int32()
        self._chk(self.lib.lfh_scene_add_instance(self.h, mesh_id, fptr(t), material_id, C.byref(i)))
        return i.value

    def set_camera(self, eye, lookat, fov_deg, aperture=0.0, focal_dist=1.0):
        e = np.asarray(eye, np.float32)
        l = np.asarray(lookat, np.float32)
        self._chk(self.lib.lfh_scene_set_camera(self.h, fptr(e), fptr(l), fov_deg, aperture, focal_dist))

    def set_hdr(self, rgb):
        a = np.ascontiguousarray(rgb, np.float32)
        h, w = a.shape[:2]
        self._chk(self.lib.lfh_scene_set_hdr(self.h, fptr(a), w, h))

    def set_options(self, **kw):
        o = self.options()
        for k, v in kw.items():
            if k == "bg_color":
                o.bg_color[:] = v
            else:
                setattr(o, k, v)
        self._chk(self.lib.lfh_scene_set_options(self.h, C.byref(o)))

    def options(self) -> Options:
        o = Options()
        self._chk(self.lib.lfh_scene_get_options(self.h, C.byref(o)))
        return o

    def camera(self) -> Camera:
        c = Camera()
        self._chk(self.lib.lfh_scene_get_camera(self.h, C.byref(c)))
        return c

    def postprocess(self) -> Postprocess:
        p = Postprocess()
        self._chk(self.lib.lfh_scene_get_postprocess(self.h, C.byref(p)))
        return p

    def build(self):
        self._chk(self.lib.lfh_scene_build(self.h))
        return self

    def add_texture(self, image, name: str = "") -> int:
        """Scene::AddTexture with decoded pixels: (H, W, 4) or (H, W, 3) uint8,
        rows top to bottom (alpha 255 when absent)."""
        a = np.asarray(image, np.uint8)
        if a.ndim != 3 or a.shape[2] not in (3, 4):
            raise ValueError("texture must be (H, W, 3|4) uint8")
        if a.shape[2] == 3:
            a = np.concatenate([a, np.full(a.shape[:2] + (1,), 255, np.uint8)], 2)
        a = np.ascontiguousarray(a)
        i = C.c_int32()
        self._chk(self.lib.lfh_scene_add_texture(self.h, a.ctypes.data_as(u8p), a.shape[1], a.shape[0],
                                                 name.encode(), C.byref(i)))
        return i.value

    def add_texture_file(self, path) -> int:
        i = C.c_int32()
        self._chk(self.lib.lfh_scene_add_texture_file(self.h, str(path).encode(), C.byref(i)))
        return i.value

    def set_instance_transform(self, inst: int, transform16):
        t = np.ascontiguousarray(transform16, np.float32).reshape(16)
        self._chk(self.lib.lfh_scene_set_instance_transform(self.h, int(inst), fptr(t)))

    def rebuild_instances(self):
        """RebuildInstances: the host TLAS and the flattened transforms."""
        self._chk(self.lib.lfh_scene_rebuild_instances(self.h))

    def flatten(self):
        """The flattened arrays without host BVH builds (lfh_scene_flatten):
        contexts made from this scene build their BVHs on the device."""
        self._chk(self.lib.lfh_scene_flatten(self.h))
        return self

    def desc(self) -> SceneDesc:
        d = SceneDesc()
        self._chk(self.lib.lfh_scene_get_desc(self.h, C.byref(d)))
        return d

    def arrays(self) -> dict:
        """Copies of the flattened reference-layout arrays."""
        d = self.desc()
        def a(ptr, n, dt):
            if n == 0 or not ptr:
                return np.zeros(0, dt)
            return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dt, copy=True)
        out = {
            "nodes": a(d.nodes, d.num_nodes * 9, np.float32).reshape(-1, 9),
            "top_level_index": d.top_level_index,
            "vert_indices": a(d.vert_indices, d.num_tris * 3, np.int32).reshape(-1, 3),
            "vertices": a(d.vertices_uvx, d.num_vertices * 4, np.float32).reshape(-1, 4),
            "normals": a(d.normals_uvy, d.num_vertices * 4, np.float32).reshape(-1, 4),
            "materials": a(d.materials, d.num_materials * 28, np.float32).reshape(-1, 28),
            "transforms": a(d.transforms, d.num_instances * 16, np.float32).reshape(-1, 16),
            "lights": a(d.lights, d.num_lights * 15, np.float32).reshape(-1, 15),
            "mesh_tri_offsets": a(d.mesh_tri_offsets, d.num_meshes + 1 if d.num_meshes else 0, np.int64),
            "instance_meshes": a(d.instance_meshes, d.num_instances, np.int32),
            "instance_materials": a(d.instance_materials, d.num_instances, np.int32),
        }
        if d.num_textures:
            out["textures"] = a(d.textures, d.num_textures * d.tex_height * d.tex_width * 4, np.uint8).reshape(
                d.num_textures, d.tex_height, d.tex_width, 4)
        if d.hdr_width:
            n = d.hdr_width * d.hdr_height
            out["hdr_cols"] = a(d.hdr_cols, n * 3, np.float32).reshape(d.hdr_height, d.hdr_width, 3)
            out["hdr_marginal"] = a(d.hdr_marginal, d.hdr_height * 2, np.float32).reshape(-1, 2)
            out["hdr_conditional"] = a(d.hdr_conditional, n * 2, np.float32).reshape(d.hdr_height, d.hdr_width, 2)
        return out


def desc_from_arrays(arr: dict) -> tuple[SceneDesc, list]:
    """SceneDesc pointing at numpy arrays (keep the returned list alive)."""
    keep = []
    def p(name, dt, ptype):
        if name not in arr or arr[name] is None or np.asarray(arr[name]).size == 0:
            return None, 0
        a = np.ascontiguousarray(arr[name], dt)
        keep.append(a)
        return a.ctypes.data_as(ptype), a.shape[0]
    d = SceneDesc()
    d.nodes, d.num_nodes = p("nodes", np.float32, f32p)
    d.top_level_index = int(arr.get("top_level_index", 0))
    d.vert_indices, d.num_tris = p("vert_indices", np.int32, i32p)
    d.vertices_uvx, d.num_vertices = p("vertices", np.float32, f32p)
    d.normals_uvy, _ = p("normals", np.float32, f32p)
    d.materials, d.num_materials = p("materials", np.float32, f32p)
    d.transforms, d.num_instances = p("transforms", np.float32, f32p)
    d.lights, d.num_lights = p("lights", np.float32, f32p)
    d.mesh_tri_offsets, n_off = p("mesh_tri_offsets", np.int64, i64p)
    d.num_meshes = max(n_off - 1, 0)
    d.instance_meshes, _ = p("instance_meshes", np.int32, i32p)
    d.instance_materials, _ = p("instance_materials", np.int32, i32p)
    if arr.get("textures") is not None and np.asarray(arr["textures"]).size:
        t = np.ascontiguousarray(arr["textures"], np.uint8)
        keep.append(t)
        d.textures = t.ctypes.data_as(u8p)
        d.num_textures, d.tex_height, d.tex_width = t.shape[:3]
    if "hdr_cols" in arr and arr["hdr_cols"] is not None:
        cols = np.ascontiguousarray(arr["hdr_cols"], np.float32)
        d.hdr_height, d.hdr_width = cols.shape[:2]
        keep.append(cols)
        d.hdr_cols = cols.ctypes.data_as(f32p)
        m = np.ascontiguousarray(arr["hdr_marginal"], np.float32); keep.append(m)
        c = np.ascontiguousarray(arr["hdr_conditional"], np.float32); keep.append(c)
        d.hdr_marginal = m.ctypes.data_as(f32p)
        d.hdr_conditional = c.ctypes.data_as(f32p)
    return d, keep


class Context:
    """A device context (lfh_create): the resident scene plus the Renderer
    surface (Update / Render / GetSampleCount / output) and direct traversal."""

    def __init__(self, desc: SceneDesc, options: Options, lib=None):
        self.lib = lib or load()
        h = C.c_void_p()
        check(self.lib.lfh_create(C.byref(desc), C.byref(options), C.byref(h)), None, self.lib)
        self.h = h
        self.options = options

    @classmethod
    def from_scene(cls, scene: "Scene", device_bvh: bool = False, **overrides):
        """Context over a built or flattened scene; device_bvh: ignore the
        host BVH and build on the device (always so for a flattened scene)."""
        o = scene.options()
        for k, v in overrides.items():
            setattr(o, k, v)
        d = scene.desc()
        if device_bvh:
            d.nodes = None
            d.num_nodes = 0
        return cls(d, o, scene.lib)

    def close(self):
        if getattr(self, "h", None):
            self.lib.lfh_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, code):
        return check(code, self.h, self.lib)

    def set_stream(self, stream_ptr: int | None):
        self._chk(self.lib.lfh_set_stream(self.h, C.c_void_p(stream_ptr) if stream_ptr else None))

    def synchronize(self):
        self._chk(self.lib.lfh_synchronize(self.h))

    # -- traversal ---------------------------------------------------------
    def trace_closest(self, rays: np.ndarray) -> np.ndarray:
        r = np.ascontiguousarray(rays)
        assert r.dtype == RAY_DTYPE
        hits = np.zeros(len(r), HIT_DTYPE)
        self._chk(self.lib.lfh_trace_closest(self.h, vptr(r), len(r), vptr(hits)))
        return hits

    def trace_any(self, rays: np.ndarray) -> np.ndarray:
        r = np.ascontiguousarray(rays)
        assert r.dtype == RAY_DTYPE
        occ = np.zeros(len(r), np.uint8)
        self._chk(self.lib.lfh_trace_any(self.h, vptr(r), len(r), vptr(occ)))
        return occ

    def trace_closest_device(self, d_rays: int, n: int, d_hits: int):
        self._chk(self.lib.lfh_trace_closest_device(self.h, C.c_void_p(d_rays), n, C.c_void_p(d_hits)))

    def trace_any_device(self, d_rays: int, n: int, d_occ: int):
        self._chk(self.lib.lfh_trace_any_device(self.h, C.c_void_p(d_rays), n, C.c_void_p(d_occ)))

    def trace_closest_stats_device(self, d_rays: int, n: int, d_hits: int, d_counts: int):
        self._chk(self.lib.lfh_trace_closest_stats_device(self.h, C.c_void_p(d_rays), n, C.c_void_p(d_hits),
                                                          C.c_void_p(d_counts)))

    # -- renderer surface --------------------------------------------------
    def update_instances(self, transforms=None, materials=None, tlas_nodes=None, first_node: int = 0):
        """Renderer::Update's re-upload after instance edits: transforms (n x 16),
        materials (m x 28), TLAS nodes from first_node (k x 9; None for a
        device-built context, whose TLAS is rebuilt from the transforms)."""
        def arr(x):
            return None if x is None else np.ascontiguousarray(x, np.float32)
        t, m, n = arr(transforms), arr(materials), arr(tlas_nodes)
        cnt = 0 if n is None else n.reshape(-1, 9).shape[0]
        self._chk(self.lib.lfh_update_instances(self.h, fptr(t) if t is not None else None,
                                                fptr(m) if m is not None else None,
                                                fptr(n) if n is not None else None, int(first_node), cnt))

    def update(self, cam: Camera, reset: bool = False):
        self._chk(self.lib.lfh_update(self.h, C.byref(cam), 1 if reset else 0))

    def render(self):
        self._chk(self.lib.lfh_render(self.h))

    def render_passes(self, passes: int):
        self._chk(self.lib.lfh_render_passes(self.h, passes))

    def render_preview(self, scale: float = 1.0, dof: bool = False) -> np.ndarray:
        """The preview engine's frame at (width * scale, height * scale):
        (h, w, 3) float32, rows bottom to top."""
        w = int(np.float32(self.options.width) * np.float32(scale))
        h = int(np.float32(self.options.height) * np.float32(scale))
        out = np.zeros((max(h, 1), max(w, 1), 3), np.float32)
        ow, oh = C.c_int32(), C.c_int32()
        self._chk(self.lib.lfh_render_preview(self.h, float(scale), 1 if dof else 0, fptr(out), C.byref(ow),
                                              C.byref(oh)))
        assert (ow.value, oh.value) == (w, h)
        return out

    def sample_count(self) -> int:
        v = C.c_int32()
        self._chk(self.lib.lfh_get_sample_count(self.h, C.byref(v)))
        return v.value

    def frame(self) -> int:
        v = C.c_int32()
        self._chk(self.lib.lfh_get_frame(self.h, C.byref(v)))
        return v.value

    def progress(self) -> float:
        v = C.c_float()
        self._chk(self.lib.lfh_get_progress(self.h, C.byref(v)))
        return v.value

    def output_hdr(self) -> np.ndarray:
        w, h = self.options.width, self.options.height
        out = np.zeros((h, w, 3), np.float32)
        self._chk(self.lib.lfh_get_output_hdr(self.h, fptr(out), w, h))
        return out

    def output_post(self, pp: Postprocess, inv_samples: float = 0.0) -> np.ndarray:
        """Post-processed RGB float image (bottom-up rows)."""
        w, h = self.options.width, self.options.height
        out = np.zeros((h, w, 3), np.float32)
        self._chk(self.lib.lfh_get_output_post(self.h, C.byref(pp), float(inv_samples), fptr(out), w, h))
        return out

    def output_ldr(self, pp: Postprocess, inv_samples: float = 0.0) -> np.ndarray:
        """RGB8 read-back of the post-processed image (bottom-up rows)."""
        w, h = self.options.width, self.options.height
        out = np.zeros((h, w, 3), np.uint8)
        self._chk(self.lib.lfh_get_output_ldr(self.h, C.byref(pp), float(inv_samples), out.ctypes.data_as(u8p),
                                              w, h))
        return out

    def accum(self) -> tuple[np.ndarray, np.ndarray]:
        w, h = self.options.width, self.options.height
        s = np.zeros((h, w, 3), np.float32)
        c = np.zeros((h, w, 3), np.float32)
        self._chk(self.lib.lfh_get_accum(self.h, fptr(s), fptr(c), w, h))
        return s, c

    # -- per-stage parity entry points ---------------------------------------
    def eval_bsdf(self, frames, L):
        fr = np.ascontiguousarray(frames, np.float32)
        l = np.ascontiguousarray(L, np.float32)
        n = fr.shape[0]
        f = np.zeros((n, 3), np.float32)
        pdf = np.zeros(n, np.float32)
        self._chk(self.lib.lfh_eval_bsdf(self.h, fptr(fr), fptr(l), n, fptr(f), fptr(pdf)))
        return f, pdf

    def sample_bsdf(self, frames, seeds):
        fr = np.ascontiguousarray(frames, np.float32)
        sd = np.ascontiguousarray(seeds, np.uint32).copy()
        n = fr.shape[0]
        L = np.zeros((n, 3), np.float32)
        f = np.zeros((n, 3), np.float32)
        pdf = np.zeros(n, np.float32)
        self._chk(self.lib.lfh_sample_bsdf(self.h, fptr(fr), uptr(sd), n, fptr(L), fptr(f), fptr(pdf)))
        return L, f, pdf, sd

    def sample_light(self, pos, seeds):
        p = np.ascontiguousarray(pos, np.float32)
        sd = np.ascontiguousarray(seeds, np.uint32).copy()
        n = p.shape[0]
        out = [np.zeros((n, 3), np.float32) for _ in range(3)] + [np.zeros(n, np.float32) for _ in range(2)]
        self._chk(self.lib.lfh_sample_light(self.h, fptr(p), uptr(sd), n, *[fptr(o) for o in out]))
        return (*out, sd)

    def sample_env(self, seeds):
        sd = np.ascontiguousarray(seeds, np.uint32).copy()
        n = sd.shape[0]
        d = np.zeros((n, 3), np.float32)
        pdf = np.zeros(n, np.float32)
        col = np.zeros((n, 3), np.float32)
        self._chk(self.lib.lfh_sample_env(self.h, uptr(sd), n, fptr(d), fptr(pdf), fptr(col)))
        return d, pdf, col, sd

    def env_pdf(self, dirs):
        d = np.ascontiguousarray(dirs, np.float32)
        pdf = np.zeros(d.shape[0], np.float32)
        self._chk(self.lib.lfh_env_pdf(self.h, fptr(d), d.shape[0], fptr(pdf)))
        return pdf

    def set_frame_offset(self, offset: int):
        self._chk(self.lib.lfh_set_frame_offset(self.h, int(offset)))

    def ray_counts(self):
        a, b = C.c_int64(), C.c_int64()
        self._chk(self.lib.lfh_get_ray_counts(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def set_guide(self, guide: "Guide | None", train_every: int = 1):
        self._guide = guide  # keep the guide alive while attached
        self._chk(self.lib.lfh_set_guide(self.h, guide.h if guide else None, int(train_every)))

    def guide_samples_seen(self):
        n = C.c_int64()
        self._chk(self.lib.lfh_guide_samples_seen(self.h, C.byref(n)))
        return n.value

    def path_trace_rays(self, rays6, seeds):
        r = np.ascontiguousarray(rays6, np.float32)
        sd = np.ascontiguousarray(seeds, np.uint32).copy()
        n = r.shape[0]
        rad = np.zeros((n, 3), np.float32)
        self._chk(self.lib.lfh_path_trace_rays(self.h, fptr(r), uptr(sd), n, fptr(rad)))
        return rad


class Guide:
    """Learned path guide (lfh_guide_*): MLP over a 64-feature encoding of
    (position, normal, wo) -> 8 vMF lobes; BF16 MFMA kernels, Adam."""

    N_PARAMS = 15080

    def __init__(self, scene_min, scene_max, seed: int = 0, lr: float = 1e-3, lib=None):
        self.lib = lib or load()
        h = C.c_void_p()
        lo = np.ascontiguousarray(scene_min, np.float32)
        hi = np.ascontiguousarray(scene_max, np.float32)
        check(self.lib.lfh_guide_create(fptr(lo), fptr(hi), int(seed), float(lr), C.byref(h)), None, self.lib)
        self.h = h

    @classmethod
    def for_scene(cls, scene: "Scene", **kw):
        v = scene.arrays()["vertices"][:, :3]
        return cls(v.min(0), v.max(0), **kw)

    def close(self):
        if getattr(self, "h", None):
            self.lib.lfh_guide_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, code):
        return check(code, None, self.lib)

    def params(self):
        p = np.zeros(self.N_PARAMS, np.float32)
        n = C.c_int64(self.N_PARAMS)
        self._chk(self.lib.lfh_guide_get_params(self.h, fptr(p), C.byref(n)))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        self._chk(self.lib.lfh_guide_set_params(self.h, fptr(p), p.size))

    def forward(self, inputs9):
        x = np.ascontiguousarray(inputs9, np.float32)
        out = np.zeros((x.shape[0], 40), np.float32)
        self._chk(self.lib.lfh_guide_forward(self.h, fptr(x), x.shape[0], fptr(out)))
        return out

    def forward_device(self, inputs9_ptr: int, n: int, out40_ptr: int, stream: int = 0):
        """Forward on device arrays (n x 9 in, n x 40 out), asynchronous on stream."""
        self._chk(self.lib.lfh_guide_forward_device(self.h, C.c_void_p(inputs9_ptr), n, C.c_void_p(out40_ptr),
                                                    C.c_void_p(stream)))

    def sample(self, inputs9, seeds):
        x = np.ascontiguousarray(inputs9, np.float32)
        sd = np.ascontiguousarray(seeds, np.uint32)
        n = x.shape[0]
        d = np.zeros((n, 3), np.float32)
        pdf = np.zeros(n, np.float32)
        self._chk(self.lib.lfh_guide_sample(self.h, fptr(x), uptr(sd), n, fptr(d), fptr(pdf)))
        return d, pdf

    def pdf(self, inputs9, dirs):
        x = np.ascontiguousarray(inputs9, np.float32)
        d = np.ascontiguousarray(dirs, np.float32)
        pdf = np.zeros(x.shape[0], np.float32)
        self._chk(self.lib.lfh_guide_pdf(self.h, fptr(x), fptr(d), x.shape[0], fptr(pdf)))
        return pdf

    def loss_grad(self, inputs9, dirs, weights):
        x = np.ascontiguousarray(inputs9, np.float32)
        d = np.ascontiguousarray(dirs, np.float32)
        w = np.ascontiguousarray(weights, np.float32)
        loss = np.zeros(1, np.float32)
        g = np.zeros(self.N_PARAMS, np.float32)
        self._chk(self.lib.lfh_guide_loss_grad(self.h, fptr(x), fptr(d), fptr(w), x.shape[0], fptr(loss), fptr(g)))
        return float(loss[0]), g

    def train_step(self, inputs9, dirs, weights):
        x = np.ascontiguousarray(inputs9, np.float32)
        d = np.ascontiguousarray(dirs, np.float32)
        w = np.ascontiguousarray(weights, np.float32)
        loss = np.zeros(1, np.float32)
        self._chk(self.lib.lfh_guide_train_step(self.h, fptr(x), fptr(d), fptr(w), x.shape[0], fptr(loss)))
        return float(loss[0])

    def steps(self):
        n = C.c_int64()
        self._chk(self.lib.lfh_guide_steps(self.h, C.byref(n)))
        return n.value


class Server:
    """Render service (lfh_server_*): continuous batching of path-tracing
    jobs against a device-memory budget."""

    def __init__(self, budget_bytes: int = 0, max_jobs: int = 64, lib=None):
        self.lib = lib or load()
        h = C.c_void_p()
        check(self.lib.lfh_server_create(int(budget_bytes), int(max_jobs), C.byref(h)), None, self.lib)
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.lfh_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, code):
        return check(code, None, self.lib)

    def submit(self, desc: SceneDesc, options: Options, camera: Camera, target_spp: int, priority: int = 0,
               tag: int = 0) -> int:
        jd = JobDesc()
        jd.scene = C.pointer(desc)
        jd.options = options
        jd.camera = camera
        jd.target_spp = target_spp
        jd.priority = priority
        jd.user_tag = tag
        jid = C.c_int64()
        code = self.lib.lfh_server_submit(self.h, C.byref(jd), C.byref(jid))
        if code != LFH_OK:
            raise LfhError(code, self.lib.lfh_last_error(None).decode())
        return jid.value

    def submit_scene(self, scene: "Scene", target_spp: int, **overrides) -> int:
        o = scene.options()
        for k, v in overrides.items():
            setattr(o, k, v)
        return self.submit(scene.desc(), o, scene.camera(), target_spp)

    def set_policy(self, policy: int = 0, nominal_rate: float = 4.5e8, max_offset: float = 1e9):
        """Step policy: 0 earliest deadline first, 1 fair share; deadline =
        submit + min(samples / nominal_rate, max_offset) seconds."""
        self._chk(self.lib.lfh_server_set_policy(self.h, int(policy), float(nominal_rate), float(max_offset)))

    def step(self) -> int:
        n = C.c_int32()
        self._chk(self.lib.lfh_server_step(self.h, C.byref(n)))
        return n.value

    def status(self, job: int) -> JobStatus:
        st = JobStatus()
        self._chk(self.lib.lfh_server_status(self.h, job, C.byref(st)))
        return st

    def result(self, job: int, width: int, height: int) -> np.ndarray:
        out = np.zeros((height, width, 3), np.float32)
        self._chk(self.lib.lfh_server_result(self.h, job, fptr(out), width, height))
        return out

    def release(self, job: int):
        """Forgets a job; a queued or running job is cancelled."""
        self._chk(self.lib.lfh_server_release(self.h, job))

    def reset_guides(self):
        """The next guided job of every resident scene starts a new guide."""
        self._chk(self.lib.lfh_server_reset_guides(self.h))

    def job_error(self, job: int) -> str:
        """Why a job was rejected ("" if it was not)."""
        return self.lib.lfh_server_job_error(self.h, job).decode()

    def clock(self) -> float:
        return self.lib.lfh_server_clock(self.h)

    def run_until_idle(self, max_steps: int = 1_000_000):
        for _ in range(max_steps):
            if self.step() == 0:
                return
        raise RuntimeError("server did not drain")
