// Device-side layouts and helpers shared by the gfx950 kernels.
//
// All device code is compiled with -ffp-contract=off: traversal and
// intersection arithmetic is evaluated in exactly the order written (one IEEE
// rounding per operation), which is what makes hit ids bit-exact against the
// CPU restatement of shaders/common/closest_hit.glsl.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define LFH_DEV __device__ __forceinline__
#define LFH_HD __host__ __device__ __forceinline__

namespace lfhd {

constexpr float kInfinity = 1000000.0f;   // globals.glsl:8
constexpr float kEps = 0.001f;            // globals.glsl:9
constexpr float kPi = 3.14159265358979323f;
constexpr float kTwoPi = 6.28318530717958648f;

// ---- child references (32 bit) ---------------------------------------------
//   internal node : index (< 2^30)
//   BLAS leaf     : 0x80000000 | (count << 28) | first triangle slot (< 2^28)
//   TLAS leaf     : 0x40000000 | instance index (< 2^30)
//   sentinel      : 0xFFFFFFFF (stack bottom / return from a BLAS)
constexpr uint32_t kRefSentinel = 0xFFFFFFFFu;
constexpr uint32_t kRefLeafBit = 0x80000000u;
constexpr uint32_t kRefInstBit = 0x40000000u;
LFH_HD bool ref_is_leaf(uint32_t r) { return (r & kRefLeafBit) != 0; }
LFH_HD bool ref_is_inst(uint32_t r) { return (r & 0xC0000000u) == kRefInstBit; }
LFH_HD uint32_t leaf_first(uint32_t r) { return r & 0x0FFFFFFFu; }
LFH_HD uint32_t leaf_count(uint32_t r) { return (r >> 28) & 0x7u; }
LFH_HD uint32_t make_leaf(uint32_t first, uint32_t count) { return kRefLeafBit | (count << 28) | first; }
LFH_HD uint32_t make_inst(uint32_t inst) { return kRefInstBit | inst; }

// ---- packed layouts ------------------------------------------------------------
// Internal BVH4 node, 128 B (one L2 line), the four child boxes in the parent
// as structure of arrays, child k in lane k of each float4:
//   f4[0..2] = min.x, min.y, min.z    f4[3..5] = max.x, max.y, max.z
//   f4[6]    = as_float(child refs)   f4[7]    = 0
//   an absent child has ref kRefSentinel.
// (Host intermediate: the BVH2 node of 64 B with both child boxes,
//  (lmin, lmax, rmin, rmax, left ref, right ref), collapsed into BVH4 nodes.)
// Triangle, 48 B, in vert_indices slot order:
//   f4[0] = (v0.xyz, as_float(prim)),  f4[1] = (e0 = v1 - v0, as_float(slot)),
//   f4[2] = (e1 = v2 - v0, 0)
// Instance, 64 B: rows 0..2 of inverse(M) as (inv[0][r], inv[1][r], inv[2][r],
//   inv[3][r]) and (as_float(blas ref), as_float(material), as_float(index), 0).

// Traversal stack: the newest kLdsStack entries of a lane live in LDS, older
// ones spill to a per-lane global area of kSpillStride entries.
#ifndef LFH_LDS_STACK
#define LFH_LDS_STACK 16
#endif
#ifndef LFH_TRACE_WAVES
// Waves per SIMD the ray-batch traversal kernel is register-budgeted for,
// closest hit and any hit.  With the octant slab (packed offsets), 1M-triangle
// soup: closest 6 waves 1641 Mrays/s (incoherent 1015, 28 B/lane spilled),
// 5 waves 1649 (incoherent 1038, coherent 3528 against 3672); any hit 6 waves
// 3876, 5 waves 3649.
#define LFH_TRACE_WAVES 5
#endif
#ifndef LFH_SLAB_OCTANT
#define LFH_SLAB_OCTANT 1   // node box tests with the near / far planes picked by the ray octant: C1 1633 / 1590 Mrays/s (incoherent 1011 / 984), any hit 3863 / 3701, C2 398 / 392 Msamples/s
#endif
#ifndef LFH_ANYHIT_WAVES
// any hit with 16-step slices: 5 / 6 / 7 / 8 waves per SIMD 4084 / 4305 / 4203 / 3860 Mrays/s
// (7 and 8 spill 28 and 56 B per lane); closest hit at 6 waves 1642.5 against 1649 at 5
#define LFH_ANYHIT_WAVES 6
#endif
#ifndef LFH_OCTANT_PACKED
#define LFH_OCTANT_PACKED 1   // the three near-plane offsets and the flag in one register (0: three + a flag)
#endif
#ifndef LFH_POP_CULL
#define LFH_POP_CULL 0   // 1: keep box entry distances on the stack and drop stale entries at pop
#endif
// Work counters of the persistent traversal kernels: a queue's ray ids are
// split into kFetchSlices contiguous slices, each with its own counter
// kFetchStride 64-bit words (256 B) from the next.  The path tracer's work
// buffer holds the extend queue's counters, then the shadow queue's, then
// two running ray totals at kWorkTotals.
#ifndef LFH_FETCH_SLICES
#define LFH_FETCH_SLICES 8
#endif
constexpr int kFetchSlices = LFH_FETCH_SLICES;
constexpr int kFetchStride = 32;
constexpr int kFetchWords = kFetchSlices * kFetchStride;
constexpr int kWorkTotals = 2 * kFetchWords;
constexpr int kWorkWords = kWorkTotals + 8;
#ifndef LFH_TRACE_SLICE
#define LFH_TRACE_SLICE 32   // C1 / C0 / C2: 16: 1557 / 1195 / 326, 24: 1557 / 1226 / 329, 32: 1573 / 1235 / 328, 48: 1558 / 1224 / 317, 96: 1457 / 1233 / 295
#endif
constexpr int kTraceSlice = LFH_TRACE_SLICE;  // traversal steps between refills of a wave's idle lanes
#ifndef LFH_ANY_SLICE
// the same for any-hit queries (C1 any hit; shadow rays of the path tracer): 32 steps 3885-3893
// Mrays/s, 24: 4107-4131, 16: 4288-4296, 12: 4293, 8: 4257 (C2 421.5 against 428 at 12-24)
#define LFH_ANY_SLICE 16
#endif
constexpr int kAnySlice = LFH_ANY_SLICE;
constexpr int kLdsStack = LFH_LDS_STACK;  // LDS entries per lane (power of 2)
constexpr int kTraceBlock = 256;        // threads per traversal block (4 wave64)
constexpr uint32_t kSpillStride = 128;  // max stack entries per lane

struct TraceScene {
    const float4* nodes;
    const float4* tris;
    const float4* insts;
    uint32_t root;
    uint32_t spill_stride;   // entries per lane in the spill area
};

struct V3 {
    float x, y, z;
};
LFH_HD V3 v3(float x, float y, float z) { return V3{x, y, z}; }
LFH_HD V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
LFH_HD V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
LFH_HD V3 operator-(V3 a) { return v3(-a.x, -a.y, -a.z); }
LFH_HD V3 operator*(V3 a, V3 b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
LFH_HD V3 operator*(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
LFH_HD V3 operator*(float s, V3 a) { return v3(s * a.x, s * a.y, s * a.z); }
LFH_HD V3 operator/(V3 a, float s) { return v3(a.x / s, a.y / s, a.z / s); }
LFH_HD V3 operator/(V3 a, V3 b) { return v3(a.x / b.x, a.y / b.y, a.z / b.z); }
// dot(a, b) = (a.x*b.x + a.y*b.y) + a.z*b.z
LFH_HD float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
LFH_HD V3 cross(V3 a, V3 b) {
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
#if defined(__HIP__)
// normalize(v) = v / sqrt(dot(v, v))
LFH_DEV V3 normalize(V3 a) {
    float l = sqrtf(dot(a, a));
    return v3(a.x / l, a.y / l, a.z / l);
}
LFH_DEV float length(V3 a) { return sqrtf(dot(a, a)); }
#endif
LFH_HD V3 xyz(float4 f) { return v3(f.x, f.y, f.z); }

#if defined(__HIP__)
LFH_DEV float as_f(uint32_t u) { return __uint_as_float(u); }
LFH_DEV uint32_t as_u(float f) { return __float_as_uint(f); }

// ---- slab test (intersection.glsl:56-72) -------------------------------------
// Returns the reference's AABBIntersect value: t0 if the box is hit in front
// (t0 > 0), the exit t1 if the origin is inside, -1 on a miss.  `entry`
// receives max(t0, 0) for culling.  min/max are IEEE minNum/maxNum.
// slab() with the near and far plane of each axis already chosen by the
// ray's octant (finite inv): the same t0 / t1, bit for bit, without the
// per-axis min / max.
LFH_DEV float slab_octant(V3 bnear, V3 bfar, V3 o, V3 inv, float* entry) {
    const float nx = (bnear.x - o.x) * inv.x, ny = (bnear.y - o.y) * inv.y, nz = (bnear.z - o.z) * inv.z;
    const float fx = (bfar.x - o.x) * inv.x, fy = (bfar.y - o.y) * inv.y, fz = (bfar.z - o.z) * inv.z;
    const float t1 = fminf(fx, fminf(fy, fz));
    const float t0 = fmaxf(nx, fmaxf(ny, nz));
    *entry = t0;
    return (t1 >= t0) ? (t0 > 0.f ? t0 : t1) : -1.0f;
}
LFH_DEV float slab(V3 bmin, V3 bmax, V3 o, V3 inv, float* entry) {
    float fx = (bmax.x - o.x) * inv.x, fy = (bmax.y - o.y) * inv.y, fz = (bmax.z - o.z) * inv.z;
    float nx = (bmin.x - o.x) * inv.x, ny = (bmin.y - o.y) * inv.y, nz = (bmin.z - o.z) * inv.z;
    float tmaxx = fmaxf(fx, nx), tmaxy = fmaxf(fy, ny), tmaxz = fmaxf(fz, nz);
    float tminx = fminf(fx, nx), tminy = fminf(fy, ny), tminz = fminf(fz, nz);
    float t1 = fminf(tmaxx, fminf(tmaxy, tmaxz));
    float t0 = fmaxf(tminx, fmaxf(tminy, tminz));
    *entry = t0;
    return (t1 >= t0) ? (t0 > 0.f ? t0 : t1) : -1.0f;
}

// ---- Moller-Trumbore as written in closest_hit.glsl:119-134 ------------------
// No epsilons; uvt.xyz / det as three IEEE divisions; w = (1 - u) - v.
// Returns true and fills (t, u, v) when u, v, w, t >= 0.
LFH_DEV bool tri_test(V3 o, V3 d, V3 v0, V3 e0, V3 e1, float* t, float* u, float* v) {
    V3 pv = cross(d, e1);
    float det = dot(e0, pv);
    V3 tv = o - v0;
    V3 qv = cross(tv, e0);
    float nu = dot(tv, pv);
    float nv = dot(d, qv);
    float nt = dot(e1, qv);
    float uu = nu / det, vv = nv / det, tt = nt / det;
    float ww = 1.0f - uu - vv;
    *t = tt;
    *u = uu;
    *v = vv;
    return uu >= 0.f && vv >= 0.f && tt >= 0.f && ww >= 0.f;
}

// Ray into instance space: inverse(M) * vec4(p, w) with the product summed
// column by column: ((c0*x + c1*y) + c2*z) + c3*w.
LFH_DEV V3 xform_inv(const float4& r0, const float4& r1, const float4& r2, V3 p, float w) {
    return v3(r0.x * p.x + r0.y * p.y + r0.z * p.z + r0.w * w,
              r1.x * p.x + r1.y * p.y + r1.z * p.z + r1.w * w,
              r2.x * p.x + r2.y * p.y + r2.z * p.z + r2.w * w);
}

LFH_DEV V3 rcp3(V3 d) { return v3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z); }
#endif  // __HIP__

}  // namespace lfhd
