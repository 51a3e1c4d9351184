// Two-level BVH traversal for gfx950: closest hit and any hit.
//
// Visit order and acceptance follow ClosestHit / AnyHit
// (shaders/common/closest_hit.glsl:69-201, anyhit.glsl:48-170):
//   * start at the top-level node without testing its box;
//   * at an internal node test the child boxes, descend into the nearest one
//     and push the others (the nodes are a BVH4 collapsed from the
//     reference's binary tree with the same boxes, so the same triangles are
//     reached);
//   * at a TLAS leaf move the ray into instance space with the inverse
//     transform, push a return marker and continue at the BLAS root;
//   * at a BLAS leaf run Moller-Trumbore on each triangle.
// Differences, all result-preserving:
//   * child boxes whose entry distance exceeds the current best t (times a
//     safety factor `cull`, default 1 + 2^-10) are skipped; the reference
//     visits them and rejects all their triangles;
//   * exact-t ties between triangles go to the lowest (instance, prim) pair
//     instead of the first one found, so the answer does not depend on visit
//     order;
//   * the per-lane stack keeps its newest entries in LDS (kLdsStack of them)
//     with their box entry distances, and spills older ones to a per-lane
//     global area; a popped entry whose entry distance is beyond the current
//     best t (times `cull`) is skipped.
#pragma once
#include "dev_common.h"

namespace lfhd {


struct LaneStack {
    uint32_t* lds;    // this lane's column of refs: entry e at lds[e * 64]
    float* ldk;       // entry distances of the LDS entries, same indexing
    uint32_t* spill;  // this lane's global spill area (refs only)
    int sp, lo;
    bool overflow;
    LFH_DEV void reset() { sp = 0; lo = 0; overflow = false; }
    LFH_DEV void push(uint32_t v, float key) {
        if (sp - lo == kLdsStack) {
            spill[lo] = lds[(lo & (kLdsStack - 1)) * 64];
            lo++;
        }
        if (sp >= (int)kSpillStride) { overflow = true; return; }
        lds[(sp & (kLdsStack - 1)) * 64] = v;
        if (LFH_POP_CULL) ldk[(sp & (kLdsStack - 1)) * 64] = key;
        sp++;
    }
    // Pops the newest entry; entries that come back from the spill area have
    // lost their distance and are never culled.
    LFH_DEV uint32_t pop(float& key) {
        if (sp <= 0) {
            key = -kInfinity;
            return kRefSentinel;
        }
        sp--;
        if (sp < lo) {
            lo = sp;
            key = -kInfinity;
            return spill[sp];
        }
        key = LFH_POP_CULL ? ldk[(sp & (kLdsStack - 1)) * 64] : -kInfinity;
        return lds[(sp & (kLdsStack - 1)) * 64];
    }
    // Pops until an entry whose box entry distance is within lim (or a marker).
    LFH_DEV uint32_t pop_live(float lim) {
        for (;;) {
            float k;
            const uint32_t r = pop(k);
            if (r == kRefSentinel || k <= lim) return r;
        }
    }
};

// Work fetch of the persistent traversal kernels.  All fetches to one
// counter serialise across the XCDs, so a queue's ids are split into
// kFetchSlices contiguous slices with a counter each (kFetchStride words
// apart).  A wave starts on slice blockIdx.x % kFetchSlices (blocks go to the
// XCDs in turn, so each XCD starts on its own run of neighbouring rays),
// reserves kFetchChunk ids at a time and hands them to its idle lanes over
// several refills; when its slice runs out it moves to the next, and after
// finding every slice exhausted it has no more work.  Counters only grow,
// so a slice once exhausted stays so, and every id below n is handed out
// exactly once.
#ifndef LFH_FETCH_CHUNK
#define LFH_FETCH_CHUNK 64
#endif
constexpr int kFetchChunk = LFH_FETCH_CHUNK;
struct WaveFetch {
    unsigned long long base = 0;   // next id of the wave's reservation (same in every lane)
    int left = 0;                  // ids left in it
    int slice = -1;                // current slice
    int tried = 0;                 // slices found exhausted
    // Ids for the lanes in `mask` (nonzero; called by every lane of the wave),
    // in rank order within the mask; n (no work) for lanes left without one.
    LFH_DEV int64_t next(unsigned long long* ctr, uint64_t mask, int lane, int64_t n, int first_slice) {
        const int want = __popcll(mask);
        const int rank = __popcll(mask & ((1ull << lane) - 1ull));
        if (left >= want) {
            const int64_t id = (int64_t)base + rank;
            base += (unsigned long long)want;
            left -= want;
            return id;
        }
        const int old = left;
        int64_t id = rank < old ? (int64_t)base + rank : n;
        left = 0;
        if (slice < 0) slice = first_slice;
        while (tried < kFetchSlices) {
            const int64_t lo = n * slice / kFetchSlices, hi = n * (slice + 1) / kFetchSlices;
            unsigned long long nb = 0;
            if (lane == 0) nb = atomicAdd(ctr + (size_t)slice * kFetchStride, (unsigned long long)kFetchChunk);
            nb = __shfl(nb, 0);
            const int64_t start = lo + (int64_t)nb;
            if (start >= hi) {
                slice = slice + 1 == kFetchSlices ? 0 : slice + 1;
                ++tried;
                continue;
            }
            const int avail = hi - start < kFetchChunk ? (int)(hi - start) : kFetchChunk;
            const int used = want - old < avail ? want - old : avail;
            if (rank >= old && rank < old + used) id = start + (rank - old);
            base = (unsigned long long)(start + used);
            left = avail - used;
            break;
        }
        return id;
    }
};

struct TraceStats {
    int nint, ntri, ninst;
};

// Per-lane traversal state, resumable (the persistent kernels run it in
// bounded slices so idle lanes can be refilled).
struct TraceState {
    V3 O, D, invD;      // world ray
    V3 o, d, inv;       // current-space ray
    float tmin, tbest;
    uint32_t cur;
    bool inBlas;
    int curInst, curMat;
    int bestPrim, bestInst, bestMat, bestSlot;
    float bu, bv;
    bool occluded;
#if LFH_SLAB_OCTANT
#if LFH_OCTANT_PACKED
    uint32_t oct;   // near-plane byte offsets x | y << 8 | z << 16, bit 24: inv has an infinite component
#else
    uint32_t nearx, neary, nearz;   // byte offsets of the near-side box planes in a node (ray octant)
    bool inf_inv;                   // some component of inv is infinite (axis-parallel ray)
#endif
#endif
};

#if LFH_SLAB_OCTANT
// The near-side plane of each axis follows from the sign of inv: min at
// bytes 0/16/32 of a node, max at 48/64/80.  For finite inv the near plane's
// distance equals the per-axis min of the two (t = (b - o) inv is monotonic
// in b), bit for bit; an infinite component (d = +-0) can make one of them
// 0 * inf = NaN, which the min/max form resolves differently, so such rays
// take the min/max form.
LFH_DEV void set_octant(TraceState& s) {
    const uint32_t nx = signbit(s.inv.x) ? 48u : 0u, ny = signbit(s.inv.y) ? 64u : 16u,
                   nz = signbit(s.inv.z) ? 80u : 32u;
    const bool inf = !(isfinite(s.inv.x) && isfinite(s.inv.y) && isfinite(s.inv.z));
#if LFH_OCTANT_PACKED
    s.oct = nx | (ny << 8) | (nz << 16) | (inf ? 1u << 24 : 0u);
#else
    s.nearx = nx;
    s.neary = ny;
    s.nearz = nz;
    s.inf_inv = inf;
#endif
}
#if LFH_OCTANT_PACKED
LFH_DEV bool octant_inf(const TraceState& s) { return (s.oct >> 24) != 0; }
LFH_DEV uint32_t near_x(const TraceState& s) { return s.oct & 0xFFu; }
LFH_DEV uint32_t near_y(const TraceState& s) { return (s.oct >> 8) & 0xFFu; }
LFH_DEV uint32_t near_z(const TraceState& s) { return (s.oct >> 16) & 0xFFu; }
#else
LFH_DEV bool octant_inf(const TraceState& s) { return s.inf_inv; }
LFH_DEV uint32_t near_x(const TraceState& s) { return s.nearx; }
LFH_DEV uint32_t near_y(const TraceState& s) { return s.neary; }
LFH_DEV uint32_t near_z(const TraceState& s) { return s.nearz; }
#endif
#endif

LFH_DEV void trace_init(TraceState& s, LaneStack& st, V3 O, V3 D, float tmin, float tmax, uint32_t root) {
    s.O = O;
    s.D = D;
    s.invD = rcp3(D);
    s.o = O;
    s.d = D;
    s.inv = s.invD;
    s.tmin = tmin;
    s.tbest = tmax;
    s.cur = root;
    s.inBlas = false;
    s.curInst = -1;
    s.curMat = 0;
    s.bestPrim = -1;
    s.bestInst = -1;
    s.bestMat = -1;
    s.bestSlot = -1;
    s.bu = 0.f;
    s.bv = 0.f;
    s.occluded = false;
#if LFH_SLAB_OCTANT
    set_octant(s);
#endif
    st.reset();
}

// Triangle tests of BLAS leaf `ref` (in the current instance's space).
// Returns true when an any-hit query found its hit.
template <bool ANY, bool STATS>
LFH_DEV bool leaf_test(const TraceScene& sc, TraceState& s, uint32_t ref, TraceStats* stats) {
    const uint32_t first = leaf_first(ref), count = leaf_count(ref);
    for (uint32_t k = 0; k < count; ++k) {
        const float4* tp = sc.tris + (size_t)(first + k) * 3;
        float4 a = tp[0], b = tp[1], c = tp[2];
        float t, u, v;
        if (STATS) stats->ntri++;
        bool h = tri_test(s.o, s.d, xyz(a), xyz(b), xyz(c), &t, &u, &v);
        if (ANY) {
            if (h && t >= s.tmin && t < s.tbest) {
                s.occluded = true;
                return true;
            }
        } else if (h && t >= s.tmin) {
            int prim = (int)as_u(a.w);
            bool better = t < s.tbest ||
                          (t == s.tbest && (s.curInst < s.bestInst ||
                                            (s.curInst == s.bestInst && prim < s.bestPrim)));
            if (better) {
                s.tbest = t;
                s.bestPrim = prim;
                s.bestSlot = (int)as_u(b.w);
                s.bestInst = s.curInst;
                s.bestMat = s.curMat;
                s.bu = u;
                s.bv = v;
            }
        }
    }
    return false;
}

// TLAS leaf: move the ray into instance space and continue at the BLAS root.
template <bool STATS>
LFH_DEV void enter_instance(const TraceScene& sc, TraceState& s, LaneStack& st, uint32_t cur, TraceStats* stats) {
    const uint32_t inst = cur & 0x3FFFFFFFu;
    const float4* ip = sc.insts + (size_t)inst * 4;
    float4 r0 = ip[0], r1 = ip[1], r2 = ip[2], meta = ip[3];
    if (STATS) stats->ninst++;
    s.o = xform_inv(r0, r1, r2, s.O, 1.0f);
    s.d = xform_inv(r0, r1, r2, s.D, 0.0f);
    s.inv = rcp3(s.d);
#if LFH_SLAB_OCTANT
    set_octant(s);
#endif
    st.push(kRefSentinel, -kInfinity);
    s.inBlas = true;
    s.curInst = (int)inst;
    s.curMat = (int)as_u(meta.y);
    s.cur = as_u(meta.x);
}

// BVH4 node: slab-test the four child boxes, descend into the nearest hit
// child, push the others farthest first.
template <bool STATS>
LFH_DEV void visit_node(const TraceScene& sc, TraceState& s, LaneStack& st, uint32_t cur, float cull,
                        TraceStats* stats) {
    if (STATS) stats->nint++;
    const float lim = s.tbest * cull;
    float h[4], e[4];
    uint32_t r[4];
#if LFH_SLAB_OCTANT
    if (__ballot(octant_inf(s)) == 0) {   // wave-uniform
        // near and far planes read by per-lane offsets: the per-axis min/max drop out
        const char* nb = reinterpret_cast<const char*>(sc.nodes);
        const uint32_t off = cur * 128u;
        const uint32_t ox = near_x(s), oy = near_y(s), oz = near_z(s);
        const float4 nx = *reinterpret_cast<const float4*>(nb + (off + ox));
        const float4 ny = *reinterpret_cast<const float4*>(nb + (off + oy));
        const float4 nz = *reinterpret_cast<const float4*>(nb + (off + oz));
        const float4 fx = *reinterpret_cast<const float4*>(nb + (off + (48u - ox)));
        const float4 fy = *reinterpret_cast<const float4*>(nb + (off + (80u - oy)));
        const float4 fz = *reinterpret_cast<const float4*>(nb + (off + (112u - oz)));
        const float4 rf = *reinterpret_cast<const float4*>(nb + (off + 96u));
        r[0] = as_u(rf.x); r[1] = as_u(rf.y); r[2] = as_u(rf.z); r[3] = as_u(rf.w);
        h[0] = slab_octant(v3(nx.x, ny.x, nz.x), v3(fx.x, fy.x, fz.x), s.o, s.inv, &e[0]);
        h[1] = slab_octant(v3(nx.y, ny.y, nz.y), v3(fx.y, fy.y, fz.y), s.o, s.inv, &e[1]);
        h[2] = slab_octant(v3(nx.z, ny.z, nz.z), v3(fx.z, fy.z, fz.z), s.o, s.inv, &e[2]);
        h[3] = slab_octant(v3(nx.w, ny.w, nz.w), v3(fx.w, fy.w, fz.w), s.o, s.inv, &e[3]);
    } else
#endif
    {
        const float4* np = sc.nodes + (size_t)cur * 8;
        const float4 mnx = np[0], mny = np[1], mnz = np[2], mxx = np[3], mxy = np[4], mxz = np[5];
        const float4 rf = np[6];
        r[0] = as_u(rf.x); r[1] = as_u(rf.y); r[2] = as_u(rf.z); r[3] = as_u(rf.w);
        h[0] = slab(v3(mnx.x, mny.x, mnz.x), v3(mxx.x, mxy.x, mxz.x), s.o, s.inv, &e[0]);
        h[1] = slab(v3(mnx.y, mny.y, mnz.y), v3(mxx.y, mxy.y, mxz.y), s.o, s.inv, &e[1]);
        h[2] = slab(v3(mnx.z, mny.z, mnz.z), v3(mxx.z, mxy.z, mxz.z), s.o, s.inv, &e[2]);
        h[3] = slab(v3(mnx.w, mny.w, mnz.w), v3(mxx.w, mxy.w, mxz.w), s.o, s.inv, &e[3]);
    }
    int nv = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const bool ok = r[k] != kRefSentinel && h[k] > 0.f && e[k] <= lim;
        nv += ok ? 1 : 0;
        h[k] = ok ? fminf(h[k], 3.0e38f) : __builtin_inff();
    }
    // sort by hit distance (absent / culled children last)
#define LFH_CSWAP(a, b)                                              \
    if (h[b] < h[a]) {                                               \
        float th = h[a]; h[a] = h[b]; h[b] = th;                     \
        float te = e[a]; e[a] = e[b]; e[b] = te;                     \
        uint32_t tr = r[a]; r[a] = r[b]; r[b] = tr;                  \
    }
    LFH_CSWAP(0, 1) LFH_CSWAP(2, 3) LFH_CSWAP(0, 2) LFH_CSWAP(1, 3) LFH_CSWAP(1, 2)
#undef LFH_CSWAP
    if (nv == 0) {
        s.cur = st.pop_live(lim);
    } else {
        if (nv > 3) st.push(r[3], e[3]);
        if (nv > 2) st.push(r[2], e[2]);
        if (nv > 1) st.push(r[1], e[1]);
        s.cur = r[0];
    }
}

// Runs up to `budget` loop steps; returns true when the traversal finished.
template <bool ANY, bool STATS>
LFH_DEV bool trace_run(const TraceScene& sc, TraceState& s, LaneStack& st, float cull, int budget,
                       TraceStats* stats) {
    for (int it = 0; it < budget; ++it) {
        uint32_t cur = s.cur;
        if (cur == kRefSentinel) {
            if (s.inBlas) {
                s.inBlas = false;
                s.o = s.O;
                s.d = s.D;
                s.inv = s.invD;
#if LFH_SLAB_OCTANT
                set_octant(s);
#endif
                s.cur = st.pop_live(s.tbest * cull);
                continue;
            }
            return true;
        }
        if (ref_is_leaf(cur)) {
            if (leaf_test<ANY, STATS>(sc, s, cur, stats)) return true;
            s.cur = st.pop_live(s.tbest * cull);
        } else if (ref_is_inst(cur)) {
            enter_instance<STATS>(sc, s, st, cur, stats);
        } else {
            visit_node<STATS>(sc, s, st, cur, cull, stats);
        }
        if (st.overflow) return true;
    }
    return false;
}

}  // namespace lfhd
