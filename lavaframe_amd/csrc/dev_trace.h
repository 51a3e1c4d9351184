// Two-level BVH traversal for gfx950: closest hit and any hit.
//
// Visit order and acceptance follow ClosestHit / AnyHit
// (shaders/common/closest_hit.glsl:69-201, anyhit.glsl:48-170):
//   * start at the top-level node without testing its box;
//   * at an internal node test both child boxes, descend into the nearer one
//     (ties to the left child) and push the farther one;
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
//     and spills older ones to a per-lane global area.
#pragma once
#include "dev_common.h"

namespace lfhd {

constexpr int kLdsStack = 16;          // LDS entries per lane (power of 2)
constexpr int kTraceBlock = 256;       // threads per block (4 wave64)
constexpr uint32_t kSpillStride = 96;  // max stack depth per lane

struct LaneStack {
    uint32_t* lds;    // this lane's column: entry e at lds[e * 64]
    uint32_t* spill;  // this lane's global spill area
    int sp, lo;
    bool overflow;
    LFH_DEV void reset() { sp = 0; lo = 0; overflow = false; }
    LFH_DEV void push(uint32_t v) {
        if (sp - lo == kLdsStack) {
            spill[lo] = lds[(lo & (kLdsStack - 1)) * 64];
            lo++;
        }
        if (sp >= (int)kSpillStride) { overflow = true; return; }
        lds[(sp & (kLdsStack - 1)) * 64] = v;
        sp++;
    }
    LFH_DEV uint32_t pop() {
        if (sp <= 0) return kRefSentinel;
        sp--;
        if (sp < lo) {
            lo = sp;
            return spill[sp];
        }
        return lds[(sp & (kLdsStack - 1)) * 64];
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
};

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
    st.reset();
}

// Runs up to `budget` node visits; returns true when the traversal finished.
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
                s.cur = st.pop();
                continue;
            }
            return true;
        }
        if (ref_is_leaf(cur)) {
            const uint32_t first = leaf_first(cur), count = leaf_count(cur);
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
            s.cur = st.pop();
        } else if (ref_is_inst(cur)) {
            const uint32_t inst = cur & 0x3FFFFFFFu;
            const float4* ip = sc.insts + (size_t)inst * 4;
            float4 r0 = ip[0], r1 = ip[1], r2 = ip[2], meta = ip[3];
            if (STATS) stats->ninst++;
            s.o = xform_inv(r0, r1, r2, s.O, 1.0f);
            s.d = xform_inv(r0, r1, r2, s.D, 0.0f);
            s.inv = rcp3(s.d);
            st.push(kRefSentinel);
            s.inBlas = true;
            s.curInst = (int)inst;
            s.curMat = (int)as_u(meta.y);
            s.cur = as_u(meta.x);
        } else {
            const float4* np = sc.nodes + (size_t)cur * 4;
            float4 n0 = np[0], n1 = np[1], n2 = np[2], n3 = np[3];
            if (STATS) stats->nint++;
            float el, er;
            float hl = slab(v3(n0.x, n0.y, n0.z), v3(n0.w, n1.x, n1.y), s.o, s.inv, &el);
            float hr = slab(v3(n1.z, n1.w, n2.x), v3(n2.y, n2.z, n2.w), s.o, s.inv, &er);
            const float lim = s.tbest * cull;
            bool vl = hl > 0.f && el <= lim;
            bool vr = hr > 0.f && er <= lim;
            uint32_t L = as_u(n3.x), R = as_u(n3.y);
            if (vl && vr) {
                if (hl > hr) {
                    s.cur = R;
                    st.push(L);
                } else {
                    s.cur = L;
                    st.push(R);
                }
            } else if (vl) {
                s.cur = L;
            } else if (vr) {
                s.cur = R;
            } else {
                s.cur = st.pop();
            }
        }
        if (st.overflow) return true;
    }
    return false;
}

}  // namespace lfhd
