// Device BVH build for scenes given as meshes and instances (no reference
// BVH arrays).  Every mesh gets its own BLAS and the instances one TLAS, all
// in the BVH4 layout the traversal reads (dev_common.h):
//
//  1. primitive boxes (triangles, or transformed BLAS boxes for the TLAS) and
//     per-segment centroid bounds (a segment is one mesh, or the TLAS);
//  2. 64-bit keys: segment id in bits 48..63, 48-bit Morton code of the
//     centroid in its segment's centroid box below; one radix sort;
//  3. PLOC clustering (Meister and Bittner, "Parallel Locally-Ordered
//     Clustering for Bounding Volume Hierarchy Construction", 2018): every
//     cluster finds the cluster within +-16 places of the sorted order whose
//     union box has the least surface area, mutual nearest neighbours merge,
//     the survivors are compacted in order.  Clusters of different segments
//     never merge, so the loop ends with one root per segment;
//  4. a top-down collapse, one tree level per pass, that turns each binary
//     node into a BVH4 node with the widening rule of the host path (replace
//     the internal child of largest surface area by its two children until
//     there are four); every triangle is a leaf of its own, its record
//     written at its vert_indices slot.
//
// Allocation of merged and BVH4 nodes uses prefix sums, so a build is
// deterministic.  Closest hits do not depend on the tree: a hit is the nearest
// triangle, equal distances resolved by (instance, prim), and the triangle
// records carry the reference's prim ids and vertex-index slots.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "ctx.h"
#include "dev_common.h"

namespace lfhd {
namespace {

#ifndef LFH_PLOC_RADIUS
#define LFH_PLOC_RADIUS 8
#endif
// On the 1M-triangle soup, radius 8 with one triangle per leaf traces fastest
// of radius 4-32 and leaves of 1-4 triangles.  With one triangle per leaf the
// triangle records stay in vert_indices order (record k = triangle k, as the
// host-built layout has them), so hits' slots index the records too.
constexpr int kBlk = 256;
constexpr int kRadius = LFH_PLOC_RADIUS;   // PLOC search radius

LFH_DEV uint32_t f2ord(float f) {
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
LFH_HD float ord2f(uint32_t u) {
    uint32_t b = (u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u;
    float f;
    std::memcpy(&f, &b, 4);
    return f;
}

// 16 bits spread to every third bit of 48
LFH_DEV uint64_t spread16(uint32_t x) {
    uint64_t v = x & 0xFFFFu;
    v = (v | (v << 32)) & 0x1F00000000FFFFull;
    v = (v | (v << 16)) & 0x1F0000FF0000FFull;
    v = (v | (v << 8)) & 0x100F00F00F00F00Full;
    v = (v | (v << 4)) & 0x10C30C30C30C30C3ull;
    v = (v | (v << 2)) & 0x1249249249249249ull;
    return v;
}


LFH_DEV float half_area(float4 lo, float4 hi) {
    const float dx = hi.x - lo.x, dy = hi.y - lo.y, dz = hi.z - lo.z;
    return dx * dy + dy * dz + dz * dx;
}

LFH_DEV int segment_of(const int64_t* offs, int nseg, int64_t k) {
    int lo = 0, hi = nseg - 1;   // largest m with offs[m] <= k
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (offs[mid] <= k) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Per-segment centroid bounds: 6 ordered uints (min xyz, max xyz).  A wave
// whose lanes share one segment reduces first and issues one set of atomics.
LFH_DEV void bound_centroid(uint32_t* segb, int seg, float3 c, bool live) {
    const int lane = threadIdx.x & 63;
    const uint64_t act = __ballot(live);
    if (!act) return;
    const int first = __builtin_ctzll(act);
    const int s0 = __shfl(seg, first);
    const bool same = __all(!live || seg == s0);
    float v[6] = {c.x, c.y, c.z, c.x, c.y, c.z};
    if (!live) {
        v[0] = v[1] = v[2] = __builtin_inff();
        v[3] = v[4] = v[5] = -__builtin_inff();
    }
    if (same) {
        for (int off = 32; off > 0; off >>= 1)
            for (int a = 0; a < 6; ++a) {
                const float o = __shfl_xor(v[a], off);
                v[a] = a < 3 ? fminf(v[a], o) : fmaxf(v[a], o);
            }
        if (lane == first)
            for (int a = 0; a < 6; ++a)
                a < 3 ? atomicMin(&segb[s0 * 6 + a], f2ord(v[a])) : atomicMax(&segb[s0 * 6 + a], f2ord(v[a]));
    } else if (live) {
        for (int a = 0; a < 6; ++a)
            a < 3 ? atomicMin(&segb[seg * 6 + a], f2ord(v[a])) : atomicMax(&segb[seg * 6 + a], f2ord(v[a]));
    }
}

__global__ __launch_bounds__(kBlk) void k_seg_init(uint32_t* segb, int nseg) {
    const int i = blockIdx.x * kBlk + threadIdx.x;
    if (i < nseg * 6) segb[i] = (i % 6) < 3 ? 0xFFFFFFFFu : 0u;
}

// Triangle primitives: node k is triangle k (vert_indices order).
__global__ __launch_bounds__(kBlk) void k_tri_prims(const float4* __restrict__ verts, const int4* __restrict__ vidx,
                                                    const int64_t* __restrict__ offs, int nseg, int n,
                                                    float4* __restrict__ nbox, int* __restrict__ ncount,
                                                    int* __restrict__ seg, uint32_t* __restrict__ segb) {
    const int k = blockIdx.x * kBlk + threadIdx.x;
    const bool live = k < n;
    float3 c = make_float3(0.f, 0.f, 0.f);
    int m = 0;
    if (live) {
        const int4 vi = vidx[k];
        const float4 a = verts[vi.x], b = verts[vi.y], d = verts[vi.z];
        const float4 lo = make_float4(fminf(fminf(a.x, b.x), d.x), fminf(fminf(a.y, b.y), d.y),
                                      fminf(fminf(a.z, b.z), d.z), 0.f);
        const float4 hi = make_float4(fmaxf(fmaxf(a.x, b.x), d.x), fmaxf(fmaxf(a.y, b.y), d.y),
                                      fmaxf(fmaxf(a.z, b.z), d.z), 0.f);
        nbox[(size_t)k * 2] = lo;
        nbox[(size_t)k * 2 + 1] = hi;
        ncount[k] = 1;
        m = segment_of(offs, nseg, k);
        seg[k] = m;
        c = make_float3(0.5f * (lo.x + hi.x), 0.5f * (lo.y + hi.y), 0.5f * (lo.z + hi.z));
    }
    bound_centroid(segb, m, c, live);
}

// Box primitives (instance world boxes, 6 floats each), one segment.
__global__ __launch_bounds__(kBlk) void k_box_prims(const float* __restrict__ boxes, int n, float4* __restrict__ nbox,
                                                    int* __restrict__ ncount, int* __restrict__ seg,
                                                    uint32_t* __restrict__ segb) {
    const int k = blockIdx.x * kBlk + threadIdx.x;
    const bool live = k < n;
    float3 c = make_float3(0.f, 0.f, 0.f);
    if (live) {
        const float* b = boxes + (size_t)k * 6;
        nbox[(size_t)k * 2] = make_float4(b[0], b[1], b[2], 0.f);
        nbox[(size_t)k * 2 + 1] = make_float4(b[3], b[4], b[5], 0.f);
        ncount[k] = 1;
        seg[k] = 0;
        c = make_float3(0.5f * (b[0] + b[3]), 0.5f * (b[1] + b[4]), 0.5f * (b[2] + b[5]));
    }
    bound_centroid(segb, 0, c, live);
}

__global__ __launch_bounds__(kBlk) void k_keys(const float4* __restrict__ nbox, const int* __restrict__ seg,
                                               const uint32_t* __restrict__ segb, int n, uint64_t* __restrict__ keys,
                                               int* __restrict__ vals) {
    const int k = blockIdx.x * kBlk + threadIdx.x;
    if (k >= n) return;
    const float4 lo = nbox[(size_t)k * 2], hi = nbox[(size_t)k * 2 + 1];
    const float c[3] = {0.5f * (lo.x + hi.x), 0.5f * (lo.y + hi.y), 0.5f * (lo.z + hi.z)};
    const int m = seg[k];
    uint32_t q[3];
    for (int a = 0; a < 3; ++a) {
        const float bl = ord2f(segb[m * 6 + a]), bh = ord2f(segb[m * 6 + 3 + a]);
        const float ext = bh - bl;
        float t = ext > 0.f ? (c[a] - bl) / ext : 0.5f;
        t = fminf(fmaxf(t, 0.f), 1.f);
        q[a] = (uint32_t)fminf(t * 65536.0f, 65535.0f);
    }
    keys[k] = ((uint64_t)m << 48) | spread16(q[0]) | (spread16(q[1]) << 1) | (spread16(q[2]) << 2);
    vals[k] = k;
}

__global__ __launch_bounds__(kBlk) void k_ploc_init(const int* __restrict__ vals, const int* __restrict__ seg,
                                                    const float4* __restrict__ nbox, int n, int* __restrict__ cl,
                                                    int* __restrict__ cseg, float4* __restrict__ cbox) {
    const int i = blockIdx.x * kBlk + threadIdx.x;
    if (i >= n) return;
    const int p = vals[i];
    cl[i] = p;
    cseg[i] = seg[p];
    cbox[(size_t)i * 2] = nbox[(size_t)p * 2];
    cbox[(size_t)i * 2 + 1] = nbox[(size_t)p * 2 + 1];
}

// Nearest neighbour of every cluster among the clusters i-16 .. i+16 of its
// segment: least union surface area, ties to the lowest index.
__global__ __launch_bounds__(kBlk) void k_ploc_nn(const int* __restrict__ cseg, const float4* __restrict__ cbox, int C,
                                                  int* __restrict__ nn) {
    constexpr int W = kBlk + 2 * kRadius;
    __shared__ float4 slo[W], shi[W];
    __shared__ int ss[W];
    const int b0 = blockIdx.x * kBlk;
    for (int t = threadIdx.x; t < W; t += kBlk) {
        const int g = b0 - kRadius + t;
        if (g >= 0 && g < C) {
            slo[t] = cbox[(size_t)g * 2];
            shi[t] = cbox[(size_t)g * 2 + 1];
            ss[t] = cseg[g];
        } else {
            ss[t] = -1;
        }
    }
    __syncthreads();
    const int i = b0 + threadIdx.x;
    if (i >= C) return;
    const int li = threadIdx.x + kRadius;
    const int s = ss[li];
    const float4 lo = slo[li], hi = shi[li];
    float best = __builtin_inff();
    int bj = -1;
    for (int d = -kRadius; d <= kRadius; ++d) {
        if (d == 0 || ss[li + d] != s) continue;
        const float4 l2 = slo[li + d], h2 = shi[li + d];
        const float4 ul = make_float4(fminf(lo.x, l2.x), fminf(lo.y, l2.y), fminf(lo.z, l2.z), 0.f);
        const float4 uh = make_float4(fmaxf(hi.x, h2.x), fmaxf(hi.y, h2.y), fmaxf(hi.z, h2.z), 0.f);
        const float a = half_area(ul, uh);
        if (a < best) {
            best = a;
            bj = i + d;
        }
    }
    nn[i] = bj;
}

// Per cluster: (merge << 32) | keep.  The lower index of a mutual pair merges
// (and keeps the merged cluster), the higher one is dropped.
__global__ __launch_bounds__(kBlk) void k_ploc_flags(const int* __restrict__ nn, int C,
                                                     unsigned long long* __restrict__ flags) {
    const int i = blockIdx.x * kBlk + threadIdx.x;
    if (i >= C) return;
    const int j = nn[i];
    const bool mutual = j >= 0 && nn[j] == i;
    const unsigned long long merge = (mutual && i < j) ? 1ull : 0ull;
    const unsigned long long keep = (mutual && j < i) ? 0ull : 1ull;
    flags[i] = (merge << 32) | keep;
}

__global__ __launch_bounds__(kBlk) void k_ploc_apply(const int* __restrict__ cl, const int* __restrict__ cseg,
                                                     const float4* __restrict__ cbox, const int* __restrict__ nn,
                                                     const unsigned long long* __restrict__ flags,
                                                     const unsigned long long* __restrict__ off, int C, int node_base,
                                                     float4* __restrict__ nbox, int2* __restrict__ nkids,
                                                     int* __restrict__ ncount, int* __restrict__ cl2,
                                                     int* __restrict__ cseg2, float4* __restrict__ cbox2,
                                                     unsigned long long* __restrict__ totals) {
    const int i = blockIdx.x * kBlk + threadIdx.x;
    if (i >= C) return;
    const unsigned long long f = flags[i], o = off[i];
    if (i == C - 1) totals[0] = o + f;
    if (!(f & 1ull)) return;
    const uint32_t dst = (uint32_t)o;
    float4 lo = cbox[(size_t)i * 2], hi = cbox[(size_t)i * 2 + 1];
    int id = cl[i];
    if (f >> 32) {
        const int j = nn[i];
        const float4 l2 = cbox[(size_t)j * 2], h2 = cbox[(size_t)j * 2 + 1];
        lo = make_float4(fminf(lo.x, l2.x), fminf(lo.y, l2.y), fminf(lo.z, l2.z), 0.f);
        hi = make_float4(fmaxf(hi.x, h2.x), fmaxf(hi.y, h2.y), fmaxf(hi.z, h2.z), 0.f);
        const int node = node_base + (int)(o >> 32);
        const int a = cl[i], b = cl[j];
        nbox[(size_t)node * 2] = lo;
        nbox[(size_t)node * 2 + 1] = hi;
        nkids[node] = make_int2(a, b);
        ncount[node] = ncount[a] + ncount[b];
        id = node;
    }
    cl2[dst] = id;
    cseg2[dst] = cseg[i];
    cbox2[(size_t)dst * 2] = lo;
    cbox2[(size_t)dst * 2 + 1] = hi;
}

// ---- collapse ---------------------------------------------------------------------

struct CollapseArgs {
    const float4* nbox;   // 2 per binary node
    const int2* nkids;    // internal binary nodes (ids >= nprim)
    const int* ncount;    // primitives below each binary node
    int nprim;            // binary nodes < nprim are primitives
    int tlas;             // 1: primitives are instances (leaf refs to instances)
    // triangle records (BLAS)
    const float4* verts;
    const int4* vidx;
    float4* tris;
    // output
    float4* nodes4;       // 8 float4 per BVH4 node
    int* need_max;        // deepest stack need seen at a leaf
};

struct Level {
    int* n2;     // binary node of each entry
    int* n4;     // its BVH4 node index
    int* off;    // first triangle record slot of its subtree
    int* need;   // stack entries pushed above it
};

LFH_DEV bool is_internal(const CollapseArgs& a, int n) { return n >= a.nprim; }

// Leaf ref of primitive n (a triangle, or an instance in the TLAS): triangle
// n's record (v0, prim id), (v1 - v0, vert_indices slot), (v2 - v0, 0) at
// record slot n.
LFH_DEV uint32_t make_leaf_ref(const CollapseArgs& a, int n) {
    if (a.tlas) return make_inst((uint32_t)n);
    const int4 vi = a.vidx[n];
    const float4 v0 = a.verts[vi.x], v1 = a.verts[vi.y], v2 = a.verts[vi.z];
    float4* t = a.tris + (size_t)n * 3;
    t[0] = make_float4(v0.x, v0.y, v0.z, __uint_as_float((uint32_t)(vi.x / 3)));
    t[1] = make_float4(v1.x - v0.x, v1.y - v0.y, v1.z - v0.z, __uint_as_float((uint32_t)n));
    t[2] = make_float4(v2.x - v0.x, v2.y - v0.y, v2.z - v0.z, 0.f);
    return make_leaf((uint32_t)n, 1u);
}

LFH_DEV bool is_leaf_sized(const CollapseArgs& a, int n) { return !is_internal(a, n); }

// Children of a BVH4 node from binary node n (widening rule of the host path).
LFH_DEV int expand(const CollapseArgs& a, int n, int ch[4]) {
    const int2 k = a.nkids[n];
    ch[0] = k.x;
    ch[1] = k.y;
    int c = 2;
    while (c < 4) {
        int best = -1;
        float ba = -1.0f;
        for (int i = 0; i < c; ++i)
            if (is_internal(a, ch[i])) {
                const float ar = half_area(a.nbox[(size_t)ch[i] * 2], a.nbox[(size_t)ch[i] * 2 + 1]);
                if (ar > ba) {
                    ba = ar;
                    best = i;
                }
            }
        if (best < 0) break;
        const int2 s = a.nkids[ch[best]];
        for (int i = c; i > best + 1; --i) ch[i] = ch[i - 1];
        ch[best] = s.x;
        ch[best + 1] = s.y;
        ++c;
    }
    return c;
}

// Roots: one per segment.  Internal roots get BVH4 nodes at the scan offset.
__global__ __launch_bounds__(kBlk) void k_root_flags(CollapseArgs a, const int* __restrict__ roots, int nseg,
                                                     int* __restrict__ flag) {
    const int m = blockIdx.x * kBlk + threadIdx.x;
    if (m >= nseg) return;
    const int r = roots[m];
    flag[m] = (is_internal(a, r) && !is_leaf_sized(a, r)) ? 1 : 0;
}

__global__ __launch_bounds__(kBlk) void k_root_write(CollapseArgs a, const int* __restrict__ roots,
                                                     const int64_t* __restrict__ seg_off, int nseg,
                                                     const int* __restrict__ flag, const int* __restrict__ scan,
                                                     int base4, Level q, uint32_t* __restrict__ seg_ref,
                                                     int* __restrict__ total) {
    const int m = blockIdx.x * kBlk + threadIdx.x;
    if (m >= nseg) return;
    if (m == nseg - 1) *total = scan[m] + flag[m];
    const int r = roots[m];
    const int o = seg_off ? (int)seg_off[m] : 0;
    if (flag[m]) {
        const int e = scan[m];
        q.n2[e] = r;
        q.n4[e] = base4 + e;
        q.off[e] = o;
        q.need[e] = 0;
        seg_ref[m] = (uint32_t)(base4 + e);
    } else {
        seg_ref[m] = make_leaf_ref(a, r);
        atomicMax(a.need_max, 0);
    }
}

// Pass 1 of a level: each entry's children, and how many become BVH4 nodes.
__global__ __launch_bounds__(kBlk) void k_level_expand(CollapseArgs a, Level q, int nq, int4* __restrict__ ech,
                                                       int* __restrict__ flag) {
    const int e = blockIdx.x * kBlk + threadIdx.x;
    if (e >= nq) return;
    int ch[4] = {-1, -1, -1, -1};
    const int c = expand(a, q.n2[e], ch);
    int inner = 0;
    for (int i = 0; i < c; ++i) inner += (is_internal(a, ch[i]) && !is_leaf_sized(a, ch[i])) ? 1 : 0;
    ech[e] = make_int4(ch[0], ch[1], ch[2], ch[3]);
    flag[e] = inner;
}

// Pass 2: write the BVH4 node, the leaves' triangles, and the next level.
__global__ __launch_bounds__(kBlk) void k_level_write(CollapseArgs a, Level q, int nq, const int4* __restrict__ ech,
                                                      const int* __restrict__ flag, const int* __restrict__ scan,
                                                      int next_base4, Level nx, int* __restrict__ total) {
    const int e = blockIdx.x * kBlk + threadIdx.x;
    if (e >= nq) return;
    if (e == nq - 1) *total = scan[e] + flag[e];
    const int4 c4 = ech[e];
    const int ch[4] = {c4.x, c4.y, c4.z, c4.w};
    int c = 0;
    while (c < 4 && ch[c] >= 0) ++c;
    const int need = q.need[e] + (c - 1);
    int o = q.off[e];
    int slot = scan[e];
    float row[8][4];
    for (int k = 0; k < 4; ++k) {
        uint32_t ref = kRefSentinel;
        float4 lo = make_float4(1.f, 1.f, 1.f, 0.f), hi = make_float4(-1.f, -1.f, -1.f, 0.f);
        if (k < c) {
            const int n = ch[k];
            // exact vertex bounds, as the reference builder boxes primitives: a
            // flat box is not entered by a ray leaving its plane at distance 0
            lo = a.nbox[(size_t)n * 2];
            hi = a.nbox[(size_t)n * 2 + 1];
            if (is_internal(a, n) && !is_leaf_sized(a, n)) {
                nx.n2[slot] = n;
                nx.n4[slot] = next_base4 + slot;
                nx.off[slot] = o;
                nx.need[slot] = need;
                ref = (uint32_t)(next_base4 + slot);
                ++slot;
            } else {
                ref = make_leaf_ref(a, n);
                atomicMax(a.need_max, need);
            }
            o += a.ncount[n];
        }
        row[0][k] = lo.x;
        row[1][k] = lo.y;
        row[2][k] = lo.z;
        row[3][k] = hi.x;
        row[4][k] = hi.y;
        row[5][k] = hi.z;
        row[6][k] = __uint_as_float(ref);
        row[7][k] = 0.f;
    }
    float4* dst = a.nodes4 + (size_t)q.n4[e] * 8;
    for (int r = 0; r < 8; ++r) dst[r] = make_float4(row[r][0], row[r][1], row[r][2], row[r][3]);
}

__global__ __launch_bounds__(kBlk) void k_root_boxes(const int* __restrict__ roots, const float4* __restrict__ nbox,
                                                     int nseg, float* __restrict__ out6) {
    const int m = blockIdx.x * kBlk + threadIdx.x;
    if (m >= nseg) return;
    const float4 lo = nbox[(size_t)roots[m] * 2], hi = nbox[(size_t)roots[m] * 2 + 1];
    float* o = out6 + (size_t)m * 6;
    o[0] = lo.x; o[1] = lo.y; o[2] = lo.z;
    o[3] = hi.x; o[4] = hi.y; o[5] = hi.z;
}

inline int nblk(int64_t n) { return (int)((n + kBlk - 1) / kBlk); }

}  // namespace
}  // namespace lfhd

namespace lfh {
namespace {

#define BCHK(expr, what)                                                                    \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess) return std::string(what) + ": " + hipGetErrorString(e_);      \
    } while (0)

// Scratch of one build (freed when it goes out of scope).
struct BuildBufs {
    DevBuf nbox, nkids, ncount, seg, segb, keys, keys2, vals, vals2, temp;
    DevBuf cl[2], cseg[2], cbox[2], nn, flags, off, totals;
    DevBuf roots, seg_ref, qa[4], qb[4], ech, iflag, iscan, need;
};

hipError_t ensure(DevBuf& b, size_t bytes) { return b.bytes >= bytes ? hipSuccess : b.alloc(bytes); }

template <class T>
hipError_t get(T* dst, const void* src, size_t n, hipStream_t s) {
    hipError_t e = hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyDeviceToHost, s);
    return e == hipSuccess ? hipStreamSynchronize(s) : e;
}

// Steps 1-3 for n primitives whose boxes and centroid bounds are in place:
// sort, cluster; leaves the segment roots (binary node ids) in b.roots.
std::string cluster(BuildBufs& b, int n, int nseg, hipStream_t s) {
    using namespace lfhd;
    BCHK(ensure(b.keys, (size_t)n * 8), "build: alloc");
    BCHK(ensure(b.keys2, (size_t)n * 8), "build: alloc");
    BCHK(ensure(b.vals, (size_t)n * 4), "build: alloc");
    BCHK(ensure(b.vals2, (size_t)n * 4), "build: alloc");
    hipLaunchKernelGGL(k_keys, dim3(nblk(n)), dim3(kBlk), 0, s, b.nbox.as<const float4>(), b.seg.as<const int>(),
                       b.segb.as<const uint32_t>(), n, b.keys.as<uint64_t>(), b.vals.as<int>());
    BCHK(hipGetLastError(), "build: keys");
    int end_bit = 48;
    while (end_bit < 64 && (1ll << (end_bit - 48)) < nseg) ++end_bit;
    size_t tb = 0;
    BCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, b.keys.as<uint64_t>(), b.keys2.as<uint64_t>(),
                                            b.vals.as<int>(), b.vals2.as<int>(), n, 0, end_bit, s),
         "build: sort size");
    size_t tb2 = 0;
    BCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, tb2, b.flags.as<unsigned long long>(),
                                          b.off.as<unsigned long long>(), n, s),
         "build: scan size");
    BCHK(ensure(b.temp, std::max(tb, tb2) + 256), "build: alloc");
    tb = b.temp.bytes;
    BCHK(hipcub::DeviceRadixSort::SortPairs(b.temp.p, tb, b.keys.as<uint64_t>(), b.keys2.as<uint64_t>(),
                                            b.vals.as<int>(), b.vals2.as<int>(), n, 0, end_bit, s),
         "build: sort");
    for (int k = 0; k < 2; ++k) {
        BCHK(ensure(b.cl[k], (size_t)n * 4), "build: alloc");
        BCHK(ensure(b.cseg[k], (size_t)n * 4), "build: alloc");
        BCHK(ensure(b.cbox[k], (size_t)n * 32), "build: alloc");
    }
    BCHK(ensure(b.nn, (size_t)n * 4), "build: alloc");
    BCHK(ensure(b.flags, (size_t)n * 8), "build: alloc");
    BCHK(ensure(b.off, (size_t)n * 8), "build: alloc");
    BCHK(ensure(b.totals, 64), "build: alloc");
    hipLaunchKernelGGL(k_ploc_init, dim3(nblk(n)), dim3(kBlk), 0, s, b.vals2.as<const int>(), b.seg.as<const int>(),
                       b.nbox.as<const float4>(), n, b.cl[0].as<int>(), b.cseg[0].as<int>(), b.cbox[0].as<float4>());
    BCHK(hipGetLastError(), "build: init");
    int C = n, cur = 0, node_base = n;
    for (int it = 0; C > nseg; ++it) {
        if (it > 100000) return "build: clustering does not converge";
        hipLaunchKernelGGL(k_ploc_nn, dim3(nblk(C)), dim3(kBlk), 0, s, b.cseg[cur].as<const int>(),
                           b.cbox[cur].as<const float4>(), C, b.nn.as<int>());
        hipLaunchKernelGGL(k_ploc_flags, dim3(nblk(C)), dim3(kBlk), 0, s, b.nn.as<const int>(), C,
                           b.flags.as<unsigned long long>());
        tb2 = b.temp.bytes;
        BCHK(hipcub::DeviceScan::ExclusiveSum(b.temp.p, tb2, b.flags.as<unsigned long long>(),
                                              b.off.as<unsigned long long>(), C, s),
             "build: scan");
        hipLaunchKernelGGL(k_ploc_apply, dim3(nblk(C)), dim3(kBlk), 0, s, b.cl[cur].as<const int>(),
                           b.cseg[cur].as<const int>(), b.cbox[cur].as<const float4>(), b.nn.as<const int>(),
                           b.flags.as<const unsigned long long>(), b.off.as<const unsigned long long>(), C, node_base,
                           b.nbox.as<float4>(), b.nkids.as<int2>(), b.ncount.as<int>(), b.cl[cur ^ 1].as<int>(),
                           b.cseg[cur ^ 1].as<int>(), b.cbox[cur ^ 1].as<float4>(),
                           b.totals.as<unsigned long long>());
        BCHK(hipGetLastError(), "build: clustering");
        unsigned long long t = 0;
        BCHK(get(&t, b.totals.p, 1, s), "build: clustering count");
        const int merges = (int)(t >> 32), keep = (int)(uint32_t)t;
        if (merges == 0) return "build: clustering made no progress";
        node_base += merges;
        C = keep;
        cur ^= 1;
    }
    if (C != nseg) return "build: a segment has no primitives";
    BCHK(ensure(b.roots, (size_t)nseg * 4), "build: alloc");
    BCHK(hipMemcpyAsync(b.roots.p, b.cl[cur].p, (size_t)nseg * 4, hipMemcpyDeviceToDevice, s), "build: roots");
    return std::string();
}

// Step 4: BVH4 nodes from base4 on (the caller sized nodes4), leaves'
// triangles into tris (BLAS).  Returns the number of BVH4 nodes written;
// seg_ref (host) gets each segment's root ref, *need the deepest stack need.
std::string collapse(BuildBufs& b, lfhd::CollapseArgs a, int nseg, const int64_t* seg_off_dev, int base4,
                     int64_t cap4, std::vector<uint32_t>* seg_ref, int* need, int64_t* written, hipStream_t s) {
    using namespace lfhd;
    const int n = a.nprim;
    for (int k = 0; k < 4; ++k) {
        BCHK(ensure(b.qa[k], (size_t)std::max(n, nseg) * 4 + 4), "build: alloc");
        BCHK(ensure(b.qb[k], (size_t)std::max(n, nseg) * 4 + 4), "build: alloc");
    }
    BCHK(ensure(b.ech, (size_t)std::max(n, nseg) * 16 + 16), "build: alloc");
    BCHK(ensure(b.iflag, (size_t)std::max(n, nseg) * 4 + 4), "build: alloc");
    BCHK(ensure(b.iscan, (size_t)std::max(n, nseg) * 4 + 4), "build: alloc");
    BCHK(ensure(b.seg_ref, (size_t)nseg * 4), "build: alloc");
    BCHK(ensure(b.need, 64), "build: alloc");
    BCHK(hipMemsetAsync(b.need.p, 0, 64, s), "build: clear");
    a.need_max = b.need.as<int>();
    Level qa{b.qa[0].as<int>(), b.qa[1].as<int>(), b.qa[2].as<int>(), b.qa[3].as<int>()};
    Level qb{b.qb[0].as<int>(), b.qb[1].as<int>(), b.qb[2].as<int>(), b.qb[3].as<int>()};
    int* total = b.totals.as<int>();
    size_t tb = 0;
    BCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, b.iflag.as<int>(), b.iscan.as<int>(), std::max(n, nseg), s),
         "build: scan size");
    BCHK(ensure(b.temp, tb + 256), "build: alloc");
    hipLaunchKernelGGL(k_root_flags, dim3(nblk(nseg)), dim3(kBlk), 0, s, a, b.roots.as<const int>(), nseg,
                       b.iflag.as<int>());
    tb = b.temp.bytes;
    BCHK(hipcub::DeviceScan::ExclusiveSum(b.temp.p, tb, b.iflag.as<int>(), b.iscan.as<int>(), nseg, s), "build: scan");
    hipLaunchKernelGGL(k_root_write, dim3(nblk(nseg)), dim3(kBlk), 0, s, a, b.roots.as<const int>(), seg_off_dev, nseg,
                       b.iflag.as<const int>(), b.iscan.as<const int>(), base4, qa, b.seg_ref.as<uint32_t>(), total);
    BCHK(hipGetLastError(), "build: roots");
    int nq = 0;
    BCHK(get(&nq, total, 1, s), "build: roots count");
    int64_t count4 = nq;
    int next_base = base4 + nq;
    while (nq > 0) {
        if (count4 > cap4) return "build: BVH4 node capacity exceeded";
        hipLaunchKernelGGL(k_level_expand, dim3(nblk(nq)), dim3(kBlk), 0, s, a, qa, nq, b.ech.as<int4>(),
                           b.iflag.as<int>());
        tb = b.temp.bytes;
        BCHK(hipcub::DeviceScan::ExclusiveSum(b.temp.p, tb, b.iflag.as<int>(), b.iscan.as<int>(), nq, s),
             "build: scan");
        hipLaunchKernelGGL(k_level_write, dim3(nblk(nq)), dim3(kBlk), 0, s, a, qa, nq, b.ech.as<const int4>(),
                           b.iflag.as<const int>(), b.iscan.as<const int>(), next_base, qb, total);
        BCHK(hipGetLastError(), "build: collapse");
        int nn = 0;
        BCHK(get(&nn, total, 1, s), "build: level count");
        std::swap(qa, qb);
        nq = nn;
        count4 += nn;
        next_base += nn;
    }
    if (count4 > cap4) return "build: BVH4 node capacity exceeded";
    seg_ref->resize(nseg);
    BCHK(get(seg_ref->data(), b.seg_ref.p, (size_t)nseg, s), "build: refs");
    BCHK(get(need, b.need.p, 1, s), "build: need");
    *written = count4;
    return std::string();
}

// World box of an instance: the BLAS root box under its transform, per axis
// the extreme products of the basis columns summed in the order of
// Scene::createTLAS (Scene.cpp:115-142), then the translation.
void instance_box(const float* t16, const float* b6, float* out6) {
    for (int a = 0; a < 3; ++a) {
        float mn[3], mx[3];
        for (int c = 0; c < 3; ++c) {
            const float x = t16[c * 4 + a] * b6[c], y = t16[c * 4 + a] * b6[3 + c];
            mn[c] = std::min(x, y);
            mx[c] = std::max(x, y);
        }
        out6[a] = ((mn[0] + mn[1]) + mn[2]) + t16[12 + a];
        out6[3 + a] = ((mx[0] + mx[1]) + mx[2]) + t16[12 + a];
    }
}

}  // namespace

std::string device_build_tlas(PackedScene* ps, DeviceScene* ds, hipStream_t s) {
    using namespace lfhd;
    const int n = ps->num_instances;
    BuildBufs b;
    std::vector<float> boxes((size_t)n * 6);
    for (int i = 0; i < n; ++i)
        instance_box(&ps->xforms[(size_t)i * 16], &ps->mesh_box[(size_t)ps->inst_mesh[i] * 6], &boxes[(size_t)i * 6]);
    DevBuf dboxes;
    BCHK(dboxes.alloc(boxes.size() * 4), "tlas: alloc");
    BCHK(hipMemcpyAsync(dboxes.p, boxes.data(), boxes.size() * 4, hipMemcpyHostToDevice, s), "tlas: upload");
    const size_t nodes2 = (size_t)2 * n;
    BCHK(b.nbox.alloc(nodes2 * 32), "tlas: alloc");
    BCHK(b.nkids.alloc(nodes2 * 8), "tlas: alloc");
    BCHK(b.ncount.alloc(nodes2 * 4), "tlas: alloc");
    BCHK(b.seg.alloc((size_t)n * 4), "tlas: alloc");
    BCHK(b.segb.alloc(64), "tlas: alloc");
    hipLaunchKernelGGL(k_seg_init, dim3(1), dim3(kBlk), 0, s, b.segb.as<uint32_t>(), 1);
    hipLaunchKernelGGL(k_box_prims, dim3(nblk(n)), dim3(kBlk), 0, s, dboxes.as<const float>(), n, b.nbox.as<float4>(),
                       b.ncount.as<int>(), b.seg.as<int>(), b.segb.as<uint32_t>());
    BCHK(hipGetLastError(), "tlas: boxes");
    std::string err = cluster(b, n, 1, s);
    if (!err.empty()) return "tlas " + err;
    CollapseArgs a{};
    a.nbox = b.nbox.as<const float4>();
    a.nkids = b.nkids.as<const int2>();
    a.ncount = b.ncount.as<const int>();
    a.nprim = n;
    a.tlas = 1;
    a.nodes4 = ds->nodes.as<float4>();
    // The collapse writes the TLAS region of the resident node array in
    // place; a rebuild that is then rejected (stack too deep) puts the
    // previous region back, so the scene stays as it was.
    const size_t region = (size_t)ps->tlas_cap4 * 128;
    float4* tlas_dev = ds->nodes.as<float4>() + (size_t)ps->blas_nodes4 * 8;
    DevBuf saved;
    BCHK(saved.alloc(region), "tlas: alloc");
    BCHK(hipMemcpyAsync(saved.p, tlas_dev, region, hipMemcpyDeviceToDevice, s), "tlas: save");
    auto restore = [&]() { (void)hipMemcpyAsync(tlas_dev, saved.p, region, hipMemcpyDeviceToDevice, s);
                           (void)hipStreamSynchronize(s); };
    std::vector<uint32_t> root;
    int need = 0;
    int64_t written = 0;
    err = collapse(b, a, 1, nullptr, (int)ps->blas_nodes4, ps->tlas_cap4, &root, &need, &written, s);
    if (!err.empty()) {
        restore();
        return "tlas " + err;
    }
    const int stack_need = need + 1 + ps->blas_need;   // TLAS pushes, the BLAS-return marker, the deepest BLAS
    if (stack_need > (int)kSpillStride || lfh_internal::inject_fault(lfh_internal::kFaultTlasBuild)) {
        restore();
        return "BVH too deep for the traversal stack (" + std::to_string(stack_need) + " entries needed, " +
               std::to_string(kSpillStride) + " available)";
    }
    ps->root4 = root[0];
    ps->stack_need = stack_need;
    device_insts4(ps);
    BCHK(ensure(ds->insts, ps->insts4.size() * 4), "tlas: alloc");
    BCHK(hipMemcpyAsync(ds->insts.p, ps->insts4.data(), ps->insts4.size() * 4, hipMemcpyHostToDevice, s),
         "tlas: instances");
    BCHK(hipStreamSynchronize(s), "tlas");
    return std::string();
}

std::string device_build(PackedScene* ps, DeviceScene* ds, hipStream_t s) {
    using namespace lfhd;
    const int n = (int)ps->num_tris;
    const int nseg = (int)ps->mesh_offsets.size() - 1;
    BuildBufs b;
    DevBuf dverts, doffs;
    BCHK(dverts.alloc(ps->verts.size() * 4), "build: alloc");
    BCHK(hipMemcpyAsync(dverts.p, ps->verts.data(), ps->verts.size() * 4, hipMemcpyHostToDevice, s), "build: upload");
    BCHK(doffs.alloc(ps->mesh_offsets.size() * 8), "build: alloc");
    BCHK(hipMemcpyAsync(doffs.p, ps->mesh_offsets.data(), ps->mesh_offsets.size() * 8, hipMemcpyHostToDevice, s),
         "build: upload");
    const size_t nodes2 = (size_t)2 * n;
    BCHK(b.nbox.alloc(nodes2 * 32), "build: alloc");
    BCHK(b.nkids.alloc(nodes2 * 8), "build: alloc");
    BCHK(b.ncount.alloc(nodes2 * 4), "build: alloc");
    BCHK(b.seg.alloc((size_t)n * 4), "build: alloc");
    BCHK(b.segb.alloc((size_t)nseg * 24), "build: alloc");
    hipLaunchKernelGGL(k_seg_init, dim3(nblk((int64_t)nseg * 6)), dim3(kBlk), 0, s, b.segb.as<uint32_t>(), nseg);
    hipLaunchKernelGGL(k_tri_prims, dim3(nblk(n)), dim3(kBlk), 0, s, dverts.as<const float4>(), ds->vidx.as<const int4>(),
                       doffs.as<const int64_t>(), nseg, n, b.nbox.as<float4>(), b.ncount.as<int>(), b.seg.as<int>(),
                       b.segb.as<uint32_t>());
    BCHK(hipGetLastError(), "build: triangles");
    std::string err = cluster(b, n, nseg, s);
    if (!err.empty()) return err;
    // BLAS: BVH4 nodes into a buffer of the largest possible size, then the
    // exact BLAS part plus the TLAS region into the scene's node array
    const int64_t cap4 = std::max<int64_t>(1, n - nseg + 1);
    DevBuf blas4;
    BCHK(blas4.alloc((size_t)cap4 * 128), "build: alloc");
    BCHK(ds->tris.alloc((size_t)n * 48), "build: alloc");
    CollapseArgs a{};
    a.nbox = b.nbox.as<const float4>();
    a.nkids = b.nkids.as<const int2>();
    a.ncount = b.ncount.as<const int>();
    a.nprim = n;
    a.tlas = 0;
    a.verts = dverts.as<const float4>();
    a.vidx = ds->vidx.as<const int4>();
    a.tris = ds->tris.as<float4>();
    a.nodes4 = blas4.as<float4>();
    int need = 0;
    int64_t written = 0;
    err = collapse(b, a, nseg, doffs.as<const int64_t>(), 0, cap4, &ps->mesh_ref4, &need, &written, s);
    if (!err.empty()) return err;
    ps->blas_need = need;
    ps->blas_nodes4 = written;
    // root boxes of the meshes (for the instance boxes of the TLAS)
    DevBuf rbox;
    BCHK(rbox.alloc((size_t)nseg * 24), "build: alloc");
    hipLaunchKernelGGL(k_root_boxes, dim3(nblk(nseg)), dim3(kBlk), 0, s, b.roots.as<const int>(),
                       b.nbox.as<const float4>(), nseg, rbox.as<float>());
    BCHK(hipGetLastError(), "build: root boxes");
    ps->mesh_box.assign((size_t)nseg * 6, 0.f);
    BCHK(get(ps->mesh_box.data(), rbox.p, (size_t)nseg * 6, s), "build: root boxes");
    ps->tlas_cap4 = std::max(1, ps->num_instances);
    BCHK(ds->nodes.alloc((size_t)(written + ps->tlas_cap4) * 128), "build: alloc");
    if (written > 0)
        BCHK(hipMemcpyAsync(ds->nodes.p, blas4.p, (size_t)written * 128, hipMemcpyDeviceToDevice, s), "build: nodes");
    BCHK(hipStreamSynchronize(s), "build");
    return device_build_tlas(ps, ds, s);
}

}  // namespace lfh
