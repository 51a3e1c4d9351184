"""Python bindings of the oracle libraries (TEST INFRASTRUCTURE ONLY).

* ``_ref/libref.so``: the reference's own scene pipeline compiled from its
  sources (see ref_harness.cpp) — loader, OBJ ingest, BVH builders and
  flattening, HDR importance tables, camera basis.
* ``_build/liboracle.so``: the CPU restatement of the reference's GLSL path
  tracing core (oracle.cpp), FP32 and FP64.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
module, and only as the checker / CPU baseline; the product never does.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
REF_LIB = HERE / "_ref" / "libref.so"
ORACLE_LIB = HERE / "_build" / "liboracle.so"

f32p = C.POINTER(C.c_float)
i32p = C.POINTER(C.c_int32)
u32p = C.POINTER(C.c_uint32)
i64p = C.POINTER(C.c_int64)


def _fp(a):
    assert a.dtype == np.float32 and a.flags["C_CONTIGUOUS"], (a.dtype, a.flags)
    return a.ctypes.data_as(f32p)


def _ip(a):
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(i32p)


def _up(a):
    assert a.dtype == np.uint32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(u32p)


# ---------------------------------------------------------------------------
# reference scene pipeline
# ---------------------------------------------------------------------------
_ref = None


def ref_available() -> bool:
    return REF_LIB.exists()


def ref_lib():
    global _ref
    if _ref is None:
        if not REF_LIB.exists():
            raise RuntimeError(f"{REF_LIB} not built (oracle/Makefile, needs the reference sources)")
        lib = C.CDLL(str(REF_LIB))
        P = C.c_void_p
        lib.ref_scene_new.restype = P
        lib.ref_scene_free.argtypes = [P]
        lib.ref_scene_load.restype = P
        lib.ref_scene_load.argtypes = [C.c_char_p]
        lib.ref_scene_add_mesh.argtypes = [P, f32p, f32p, C.c_int]
        lib.ref_scene_add_material.argtypes = [P, f32p]
        lib.ref_scene_add_instance.argtypes = [P, C.c_int, f32p, C.c_int]
        lib.ref_scene_build.argtypes = [P]
        lib.ref_scene_arrays.argtypes = [P] + [C.POINTER(f32p)] + [C.POINTER(i32p)] + [C.POINTER(f32p)] * 5 + [i64p]
        lib.ref_scene_camera.argtypes = [P, f32p]
        lib.ref_make_camera.argtypes = [f32p, f32p, C.c_float, f32p]
        lib.ref_hdr_load.argtypes = [C.c_char_p, i32p, i32p, f32p, f32p, f32p, C.c_int]
        _ref = lib
    return _ref


class RefScene:
    """A reference Scene (LavaFrame/Scene.h) driven through its own API."""

    def __init__(self, handle):
        self.lib = ref_lib()
        self.h = handle

    @classmethod
    def load(cls, path):
        h = ref_lib().ref_scene_load(str(path).encode())
        if not h:
            raise RuntimeError(f"reference loader failed on {path}")
        return cls(h)

    @classmethod
    def new(cls):
        return cls(ref_lib().ref_scene_new())

    def __del__(self):
        try:
            if self.h:
                self.lib.ref_scene_free(self.h)
        except Exception:
            pass

    def add_mesh(self, positions, normals):
        p = np.ascontiguousarray(positions, np.float32).reshape(-1)
        n = np.ascontiguousarray(normals, np.float32).reshape(-1)
        return self.lib.ref_scene_add_mesh(self.h, _fp(p), _fp(n), p.size // 9)

    def add_material(self, m28):
        m = np.ascontiguousarray(m28, np.float32)
        return self.lib.ref_scene_add_material(self.h, _fp(m))

    def add_instance(self, mesh, transform16, mat):
        t = np.ascontiguousarray(np.asarray(transform16, np.float32).reshape(16))
        return self.lib.ref_scene_add_instance(self.h, mesh, _fp(t), mat)

    def build(self):
        self.lib.ref_scene_build(self.h)
        return self

    def arrays(self) -> dict:
        ptrs = [f32p(), i32p(), f32p(), f32p(), f32p(), f32p(), f32p()]
        counts = np.zeros(15, np.int64)
        self.lib.ref_scene_arrays(self.h, *[C.byref(p) for p in ptrs], counts.ctypes.data_as(i64p))
        nn, top, nt, nv, nm, ni, nl = (int(x) for x in counts[:7])

        def a(ptr, n, dt):
            if n == 0:
                return np.zeros(0, dt)
            return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dt, copy=True)
        return {
            "nodes": a(ptrs[0], nn * 9, np.float32).reshape(-1, 9),
            "top_level_index": top,
            "vert_indices": a(ptrs[1], nt * 3, np.int32).reshape(-1, 3),
            "vertices": a(ptrs[2], nv * 4, np.float32).reshape(-1, 4),
            "normals": a(ptrs[3], nv * 4, np.float32).reshape(-1, 4),
            "materials": a(ptrs[4], nm * 28, np.float32).reshape(-1, 28),
            "transforms": a(ptrs[5], ni * 16, np.float32).reshape(-1, 16),
            "lights": a(ptrs[6], nl * 15, np.float32).reshape(-1, 15),
            "options": dict(zip(["width", "height", "max_depth", "tile_width", "tile_height",
                                 "enable_rr", "rr_depth", "num_meshes"], [int(x) for x in counts[7:15]])),
        }

    def camera(self) -> np.ndarray:
        out = np.zeros(15, np.float32)
        if self.lib.ref_scene_camera(self.h, _fp(out)) != 0:
            raise RuntimeError("reference scene has no camera")
        return out


def ref_camera(eye, lookat, fov_deg) -> np.ndarray:
    out = np.zeros(15, np.float32)
    e = np.asarray(eye, np.float32)
    l = np.asarray(lookat, np.float32)
    ref_lib().ref_make_camera(_fp(e), _fp(l), fov_deg, _fp(out))
    return out


def write_rgbe(path, rgbe: np.ndarray):
    """Radiance file with flat (non run-length) RGBE scanlines, row 0 first."""
    h, w, _ = rgbe.shape
    with open(path, "wb") as f:
        f.write(b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n")
        f.write(f"-Y {h} +X {w}\n".encode())
        f.write(np.ascontiguousarray(rgbe, np.uint8).tobytes())


def ref_hdr_load(path, max_texels=1 << 22):
    lib = ref_lib()
    w = C.c_int32()
    h = C.c_int32()
    cols = np.zeros(max_texels * 3, np.float32)
    marg = np.zeros(max_texels * 2, np.float32)
    cond = np.zeros(max_texels * 2, np.float32)
    r = lib.ref_hdr_load(str(path).encode(), C.byref(w), C.byref(h), _fp(cols), _fp(marg), _fp(cond), max_texels)
    if r != 0:
        raise RuntimeError(f"reference HDR loader failed ({r})")
    W, H = w.value, h.value
    return (cols[: W * H * 3].reshape(H, W, 3).copy(), marg[: H * 2].reshape(H, 2).copy(),
            cond[: W * H * 2].reshape(H, W, 2).copy())
