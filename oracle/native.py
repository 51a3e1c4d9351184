"""ctypes binding of _build/liboracle.so, the CPU restatement of the reference
GLSL (TEST INFRASTRUCTURE ONLY: used by tests/, smoke() and bench.py's
cpu_baseline leg as the checker / CPU baseline, never by the product)."""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "_build" / "liboracle.so"

f32p = C.POINTER(C.c_float)
f64p = C.POINTER(C.c_double)
i32p = C.POINTER(C.c_int32)
u32p = C.POINTER(C.c_uint32)
u8p = C.POINTER(C.c_uint8)
P = C.c_void_p


class Options(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("tile_width", C.c_int32), ("tile_height", C.c_int32),
                ("max_depth", C.c_int32), ("enable_rr", C.c_int32), ("rr_depth", C.c_int32),
                ("envmap", C.c_int32), ("lights", C.c_int32), ("constant_bg", C.c_int32),
                ("bg", C.c_float * 3), ("hdr_multiplier", C.c_float),
                ("cam_pos", C.c_float * 3), ("cam_right", C.c_float * 3), ("cam_up", C.c_float * 3),
                ("cam_fwd", C.c_float * 3), ("fov", C.c_float), ("focal_dist", C.c_float), ("aperture", C.c_float),
                ("estimator", C.c_int32)]


_lib = None


def available() -> bool:
    return LIB.exists()


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            raise RuntimeError(f"{LIB} not built (make -C oracle)")
        L = C.CDLL(str(LIB))
        L.orc_pcg4d.argtypes = [u32p, C.c_int]
        L.orc_rand.argtypes = [u32p]
        L.orc_rand.restype = C.c_float
        L.orc_scene_create.restype = P
        L.orc_scene_create.argtypes = [f32p, C.c_int64, C.c_int32, i32p, C.c_int64, f32p, f32p, C.c_int64,
                                       f32p, C.c_int32, f32p, C.c_int32, f32p, C.c_int32, C.c_int32, C.c_int32,
                                       f32p, f32p, f32p]
        L.orc_scene_destroy.argtypes = [P]
        L.orc_scene_set_textures.argtypes = [P, u8p, C.c_int32, C.c_int32, C.c_int32]
        L.orc_texture_sample.argtypes = [P, f64p, i32p, C.c_int64, f64p]
        L.orc_preview.argtypes = [P, C.POINTER(Options), C.c_int32, C.c_int32, C.c_int32, f64p, C.c_int, C.c_int]
        L.orc_trace_closest.restype = C.c_int64
        L.orc_trace_closest.argtypes = [P, f32p, C.c_int64, P, C.c_int, i32p, C.c_int]
        L.orc_trace_any.argtypes = [P, f32p, C.c_int64, u8p, C.c_int]
        L.orc_eval_bsdf.argtypes = [P, f32p, f32p, C.c_int64, f32p, f32p, C.c_int]
        L.orc_sample_bsdf.argtypes = [P, f32p, u32p, C.c_int64, f32p, f32p, f32p, C.c_int]
        L.orc_sample_light.argtypes = [P, f32p, u32p, C.c_int64, f32p, f32p, f32p, f32p, f32p, C.c_int]
        L.orc_sample_env.argtypes = [P, u32p, C.c_int64, f32p, f32p, f32p, C.c_float, C.c_int]
        L.orc_env_pdf.argtypes = [P, f32p, C.c_int64, f32p, C.c_int]
        L.orc_path_trace_rays.argtypes = [P, C.POINTER(Options), f32p, u32p, C.c_int64, f32p, C.c_int, C.c_int]
        L.orc_render.argtypes = [P, C.POINTER(Options), C.c_int32, C.c_int32, f64p, C.c_int, C.c_int]
        L.orc_render_pixels.argtypes = [P, C.POINTER(Options), i32p, C.c_int64, C.c_int32, C.c_int32, f64p, f64p,
                                        C.c_int, C.c_int]
        L.orc_camera_rays.argtypes = [C.POINTER(Options), C.c_int32, C.c_int32, C.c_int32, f64p, C.c_int]
        _lib = L
    return _lib


def _f(a):
    a = np.ascontiguousarray(a, np.float32)
    return a, a.ctypes.data_as(f32p)


def threads() -> int:
    try:
        return max(1, len(os.sched_getaffinity(0)))
    except Exception:
        return os.cpu_count() or 1


class OracleScene:
    """The restatement's view of a scene given the reference flat arrays
    (the dict of lfh.Scene.arrays() / bind.RefScene.arrays())."""

    def __init__(self, arr: dict):
        L = lib()
        self.L = L
        keep = []

        def fp(name, dt=np.float32, ptype=f32p):
            if name not in arr or arr[name] is None or np.asarray(arr[name]).size == 0:
                return None
            a = np.ascontiguousarray(arr[name], dt)
            keep.append(a)
            return a.ctypes.data_as(ptype)
        nodes = np.asarray(arr["nodes"]).reshape(-1, 9)
        vi = np.asarray(arr["vert_indices"]).reshape(-1, 3)
        vt = np.asarray(arr["vertices"]).reshape(-1, 4)
        mats = np.asarray(arr["materials"]).reshape(-1, 28)
        xf = np.asarray(arr["transforms"]).reshape(-1, 16)
        lt = np.asarray(arr.get("lights", np.zeros((0, 15), np.float32))).reshape(-1, 15)
        hw = hh = 0
        if arr.get("hdr_cols") is not None:
            hh, hw = np.asarray(arr["hdr_cols"]).shape[:2]
        self.h = L.orc_scene_create(fp("nodes"), nodes.shape[0], int(arr["top_level_index"]),
                                    fp("vert_indices", np.int32, i32p), vi.shape[0], fp("vertices"), fp("normals"),
                                    vt.shape[0], fp("materials"), mats.shape[0], fp("transforms"), xf.shape[0],
                                    fp("lights"), lt.shape[0], hw, hh, fp("hdr_cols"), fp("hdr_marginal"),
                                    fp("hdr_conditional"))
        self.num_lights = lt.shape[0]
        self.has_env = hw > 0
        tex = arr.get("textures")
        if tex is not None and np.asarray(tex).size:
            t = np.ascontiguousarray(tex, np.uint8)   # (layers, H, W, 4)
            self._tex = t
            L.orc_scene_set_textures(self.h, t.ctypes.data_as(u8p), t.shape[2], t.shape[1], t.shape[0])

    def preview(self, opts: "Options", pw: int, ph: int, dof=False, precision=32, nthreads=None):
        """The preview engine's frame: (ph, pw, 3), one sample per pixel."""
        out = np.zeros((ph, pw, 3), np.float64)
        self.L.orc_preview(self.h, C.byref(opts), pw, ph, 1 if dof else 0, out.ctypes.data_as(f64p), precision,
                           nthreads or threads())
        return out

    def texture_sample(self, st, layer):
        """FP64 texture() reads: st (n, 2) texture coordinates, layer (n,)."""
        a = np.ascontiguousarray(st, np.float64).reshape(-1, 2)
        l = np.ascontiguousarray(np.broadcast_to(layer, (a.shape[0],)), np.int32)
        out = np.zeros((a.shape[0], 4))
        self.L.orc_texture_sample(self.h, a.ctypes.data_as(f64p), l.ctypes.data_as(i32p), a.shape[0],
                                  out.ctypes.data_as(f64p))
        return out

    def __del__(self):
        try:
            self.L.orc_scene_destroy(self.h)
        except Exception:
            pass

    def trace_closest(self, rays: np.ndarray, tie_mode=1, counts=False, nthreads=None):
        r = np.ascontiguousarray(rays)
        hits = np.zeros(len(r), [("prim", "<i4"), ("inst", "<i4"), ("mat", "<i4"), ("t", "<f4"), ("u", "<f4"),
                                 ("v", "<f4")])
        cnt = np.zeros((len(r), 3), np.int32) if counts else None
        ties = self.L.orc_trace_closest(self.h, r.ctypes.data_as(f32p), len(r), hits.ctypes.data_as(P), tie_mode,
                                        None if cnt is None else cnt.ctypes.data_as(i32p), nthreads or threads())
        return (hits, cnt, ties) if counts else hits

    def trace_any(self, rays: np.ndarray, nthreads=None):
        r = np.ascontiguousarray(rays)
        occ = np.zeros(len(r), np.uint8)
        self.L.orc_trace_any(self.h, r.ctypes.data_as(f32p), len(r), occ.ctypes.data_as(u8p), nthreads or threads())
        return occ

    def eval_bsdf(self, frames, L, precision=32):
        fr, frp = _f(frames)
        l, lp = _f(L)
        n = fr.shape[0]
        f = np.zeros((n, 3), np.float32)
        pdf = np.zeros(n, np.float32)
        self.L.orc_eval_bsdf(self.h, frp, lp, n, f.ctypes.data_as(f32p), pdf.ctypes.data_as(f32p), precision)
        return f, pdf

    def sample_bsdf(self, frames, seeds, precision=32):
        fr, frp = _f(frames)
        sd = np.ascontiguousarray(seeds, np.uint32).copy()
        n = fr.shape[0]
        L = np.zeros((n, 3), np.float32)
        f = np.zeros((n, 3), np.float32)
        pdf = np.zeros(n, np.float32)
        self.L.orc_sample_bsdf(self.h, frp, sd.ctypes.data_as(u32p), n, L.ctypes.data_as(f32p),
                               f.ctypes.data_as(f32p), pdf.ctypes.data_as(f32p), precision)
        return L, f, pdf, sd

    def sample_light(self, pos, seeds, precision=32):
        p, pp = _f(pos)
        sd = np.ascontiguousarray(seeds, np.uint32).copy()
        n = p.shape[0]
        out = [np.zeros((n, 3), np.float32) for _ in range(3)] + [np.zeros(n, np.float32) for _ in range(2)]
        self.L.orc_sample_light(self.h, pp, sd.ctypes.data_as(u32p), n, *[o.ctypes.data_as(f32p) for o in out],
                                precision)
        return (*out, sd)

    def sample_env(self, seeds, mult=1.0, precision=32):
        sd = np.ascontiguousarray(seeds, np.uint32).copy()
        n = sd.shape[0]
        d = np.zeros((n, 3), np.float32)
        pdf = np.zeros(n, np.float32)
        col = np.zeros((n, 3), np.float32)
        self.L.orc_sample_env(self.h, sd.ctypes.data_as(u32p), n, d.ctypes.data_as(f32p), pdf.ctypes.data_as(f32p),
                              col.ctypes.data_as(f32p), mult, precision)
        return d, pdf, col, sd

    def env_pdf(self, dirs, precision=32):
        d, dp = _f(dirs)
        pdf = np.zeros(d.shape[0], np.float32)
        self.L.orc_env_pdf(self.h, dp, d.shape[0], pdf.ctypes.data_as(f32p), precision)
        return pdf

    def path_trace_rays(self, opts: Options, rays6, seeds, precision=32, nthreads=None):
        r, rp = _f(rays6)
        sd = np.ascontiguousarray(seeds, np.uint32).copy()
        n = r.shape[0]
        rad = np.zeros((n, 3), np.float32)
        self.L.orc_path_trace_rays(self.h, C.byref(opts), rp, sd.ctypes.data_as(u32p), n, rad.ctypes.data_as(f32p),
                                   precision, nthreads or threads())
        return rad

    @staticmethod
    def camera_rays(opts: Options, frame: int, tile=(0, 0), precision=64):
        """(tile_h, tile_w, 6) origins and directions of one tile's camera rays."""
        out = np.zeros((opts.tile_height, opts.tile_width, 6), np.float64)
        lib().orc_camera_rays(C.byref(opts), tile[0], tile[1], frame, out.ctypes.data_as(f64p), precision)
        return out

    def render(self, opts: Options, frame0: int, passes: int, precision=32, nthreads=None):
        acc = np.zeros((opts.height, opts.width, 3), np.float64)
        self.L.orc_render(self.h, C.byref(opts), frame0, passes, acc.ctypes.data_as(f64p), precision,
                          nthreads or threads())
        return acc

    def render_pixels(self, opts: Options, pixels, frame0: int, spp: int, precision=64, nthreads=None):
        px = np.ascontiguousarray(pixels, np.int32)
        s = np.zeros((len(px), 3), np.float64)
        s2 = np.zeros((len(px), 3), np.float64)
        self.L.orc_render_pixels(self.h, C.byref(opts), px.ctypes.data_as(i32p), len(px), frame0, spp,
                                 s.ctypes.data_as(f64p), s2.ctypes.data_as(f64p), precision, nthreads or threads())
        return s, s2


def options_from(lfh_opts, cam, lights: bool, envmap: bool) -> Options:
    """Oracle options from lfh.Options + lfh.Camera."""
    o = Options()
    o.width, o.height = lfh_opts.width, lfh_opts.height
    o.tile_width = lfh_opts.tile_width or lfh_opts.width
    o.tile_height = lfh_opts.tile_height or lfh_opts.height
    o.max_depth, o.enable_rr, o.rr_depth = lfh_opts.max_depth, lfh_opts.enable_rr, lfh_opts.rr_depth
    o.envmap = 1 if (envmap and lfh_opts.use_env_map) else 0
    o.lights = 1 if lights else 0
    o.constant_bg = lfh_opts.use_constant_bg
    o.bg[:] = list(lfh_opts.bg_color)
    o.hdr_multiplier = lfh_opts.hdr_multiplier
    o.cam_pos[:] = list(cam.position)
    o.cam_right[:] = list(cam.right)
    o.cam_up[:] = list(cam.up)
    o.cam_fwd[:] = list(cam.forward)
    o.fov, o.focal_dist, o.aperture = cam.fov, cam.focal_dist, cam.aperture
    o.estimator = lfh_opts.estimator
    return o
