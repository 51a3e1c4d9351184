// Wavefront path tracer for gfx950.
//
// One TiledRenderer pass over a tile (renderer.glsl:main + PathTrace,
// pathtrace.glsl:208-295) is split into stages that pass SoA queues through
// HBM:
//   generate  camera rays for the tile pixels, path state init
//   extend    closest hit (analytic lights first, then the BVH) per queued ray
//   shade     miss / emitter / surface shading, next-event candidates, BSDF
//             sampling, Russian roulette; appends continuing rays and shadow
//             rays with one atomic per block (ballot + rank, block offsets in LDS)
//   shadow    any hit per shadow ray; clears the candidate when occluded
//   accumulate  pending next-event terms + Kahan-compensated accumulation
// The random-number stream of each path is the reference's: it lives in the
// path state and is consumed in the reference's order; next-event terms are
// evaluated before their shadow test (they consume no random numbers) and
// added after it, in the reference's summation order.
#include "dev_guide.h"
#include "dev_shade.h"
#include "guide_kernels.h"
#include "pt_kernels.h"

namespace lfhd {

LFH_DEV float luminance(V3 c) { return 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z; }

// Training record of path p at depth d (only for recorded paths).
LFH_DEV float* guide_rec(const GuideHook& g, int p, int depth) {
    return reinterpret_cast<float*>(g.rec + ((size_t)(p / g.stride) * g.max_depth + depth) * 4);
}

// ---- queue helpers ---------------------------------------------------------------
#ifndef LFH_BLOCK_APPEND
#define LFH_BLOCK_APPEND 1   // 1: queue appends take one atomic per block, 0: one per wave
#endif
// Wave-aggregated append: returns this lane's slot (valid only if `want`).
LFH_DEV uint32_t wave_append(uint32_t* counter, bool want) {
    const uint64_t mask = __ballot(want);
    if (mask == 0) return 0;
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll((long long)mask) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(mask));
    base = __shfl(base, leader);
    return base + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
}

// Block-aggregated appends: every thread of the block calls them at the same
// point (the kernels using them have no early exit).  One atomic per block
// instead of one per wave: all the device's appends to a queue hit the same
// counter, whose atomics serialise across the XCDs.  `lds` holds 16 words.
// Returns this thread's slot for `want` (and for `want_b`, appended after
// the block's `want` entries of the same wave).
LFH_DEV uint32_t block_append2(uint32_t* counter, bool want, bool want_b, uint32_t* slot_b, uint32_t* lds) {
    const uint64_t ma = __ballot(want), mb = __ballot(want_b);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = (int)(blockDim.x >> 6);
    const uint32_t ca = (uint32_t)__popcll(ma);
    if (lane == 0) lds[wave] = ca + (uint32_t)__popcll(mb);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t tot = 0;
        for (int w = 0; w < nw; ++w) tot += lds[w];
        uint32_t base = tot ? atomicAdd(counter, tot) : 0u;
        for (int w = 0; w < nw; ++w) {
            const uint32_t c = lds[w];
            lds[w] = base;
            base += c;
        }
    }
    __syncthreads();
    const uint64_t below = (1ull << lane) - 1ull;
    const uint32_t wb = lds[wave];
    *slot_b = wb + ca + (uint32_t)__popcll(mb & below);
    const uint32_t sa = wb + (uint32_t)__popcll(ma & below);
    __syncthreads();   // lds is free for the next append
    return sa;
}
LFH_DEV uint32_t block_append(uint32_t* counter, bool want, uint32_t* lds) {
    uint32_t unused;
    return block_append2(counter, want, false, &unused, lds);
}

LFH_DEV float4 f4(V3 v, float w) { return make_float4(v.x, v.y, v.z, w); }

__global__ void k_pt_reset(uint32_t* counters, unsigned long long* work) {
    int i = threadIdx.x;
    if (i < 8) counters[i] = 0u;
    if (i < 2 * kFetchSlices) work[i * kFetchStride] = 0ull;
}

// work: the extend and shadow queues' fetch counters (zeroed after every
// bounce), then at kWorkTotals the running totals of extended and shadow
// rays (never reset by a pass).
__global__ void k_pt_advance(uint32_t* consumed, uint32_t* shadow_count, unsigned long long* work) {
    if (threadIdx.x == 0) {
        work[kWorkTotals] += *consumed;
        work[kWorkTotals + 1] += *shadow_count;
        *consumed = 0u;
        *shadow_count = 0u;
    }
    if (threadIdx.x < 2 * kFetchSlices) work[threadIdx.x * kFetchStride] = 0ull;
}

// One queue's half of k_pt_advance, for a chain whose shadow traversal runs
// on its own stream: the queue's count goes to its running total, then the
// count and the queue's fetch counters are cleared.
__global__ void k_pt_advance_queue(uint32_t* count, unsigned long long* fetch, unsigned long long* total) {
    if (threadIdx.x == 0) {
        *total += *count;
        *count = 0u;
    }
    if (threadIdx.x < kFetchSlices) fetch[threadIdx.x * kFetchStride] = 0ull;
}

// generate: one thread per tile pixel (MULTI: per pixel of every job)
template <bool MULTI>
__global__ __launch_bounds__(256) void k_generate(CamParams cam1, const JobDev* __restrict__ jobs, int njobs,
                                                  int total, PathBufs pb, float4* __restrict__ rayq,
                                                  uint32_t* __restrict__ count) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    int lp = p, j = 0;
    CamParams cam = cam1;
    bool inrange = p < total;
    if (MULTI && inrange) {
        while (j + 1 < njobs && jobs[j + 1].path_offset <= p) ++j;
        cam = jobs[j].cam;
        lp = p - jobs[j].path_offset;
        pb.job[p] = j;
    }
    const int P = cam.tile_w * cam.tile_h;
    if (cam.batch > 1 && inrange) {
        const int k = lp / P;   // sample k of the batch: the frame after k passes
        lp -= k * P;
        cam.frame += k;
    }
    bool valid = false;
    int lx = 0, ly = 0;
    if (inrange && lp < P) {
        lx = lp % cam.tile_w;
        ly = lp / cam.tile_w;
        valid = cam.preview ||
                ((cam.tileX * cam.tile_w + lx < cam.width) && (cam.tileY * cam.tile_h + ly < cam.height));
    }
    V3 o = v3(0.f, 0.f, 0.f), d = v3(0.f, 0.f, 1.f);
    uint4 seed = make_uint4(0, 0, 0, 0);
    if (valid) {
        if (cam.preview)
            camera_ray_preview(cam, lx, ly, o, d, seed);
        else
            camera_ray(cam, lx, ly, o, d, seed);
        pb.rad[p] = make_float4(0.f, 0.f, 0.f, 0.f);
        pb.thr[p] = make_float4(1.f, 1.f, 1.f, 0.f);
        pb.absorb[p] = make_float4(0.f, 0.f, 0.f, 0.f);
        pb.rng[p] = seed;
        pb.misc[p] = make_float4(0.f, __int_as_float(0), 0.f, 0.f);
        pb.nee[p * 3 + 2] = make_float4(0.f, 0.f, 0.f, 0.f);
    } else if (inrange) {
        pb.rad[p] = make_float4(0.f, 0.f, 0.f, 0.f);
        pb.nee[p * 3 + 2] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
#if LFH_BLOCK_APPEND
    __shared__ uint32_t app[16];
    uint32_t slot = block_append(count, valid, app);
#else
    uint32_t slot = wave_append(count, valid);
#endif
    if (valid) {
        rayq[(size_t)slot * 2] = f4(o, __int_as_float(p));
        rayq[(size_t)slot * 2 + 1] = f4(d, 0.f);
    }
}

// generate from caller-given rays and seeds (lfh_path_trace_rays)
__global__ __launch_bounds__(256) void k_generate_rays(const float* __restrict__ rays6, const uint4* __restrict__ seeds,
                                                       int n, PathBufs pb, float4* __restrict__ rayq,
                                                       uint32_t* __restrict__ count) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    bool valid = p < n;
    if (valid) {
        pb.rad[p] = make_float4(0.f, 0.f, 0.f, 0.f);
        pb.thr[p] = make_float4(1.f, 1.f, 1.f, 0.f);
        pb.absorb[p] = make_float4(0.f, 0.f, 0.f, 0.f);
        pb.rng[p] = seeds[p];
        pb.misc[p] = make_float4(0.f, __int_as_float(0), 0.f, 0.f);
        pb.nee[p * 3 + 2] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
#if LFH_BLOCK_APPEND
    __shared__ uint32_t app[16];
    uint32_t slot = block_append(count, valid, app);
#else
    uint32_t slot = wave_append(count, valid);
#endif
    if (valid) {
        const float* r = rays6 + (size_t)p * 6;
        rayq[(size_t)slot * 2] = make_float4(r[0], r[1], r[2], __int_as_float(p));
        rayq[(size_t)slot * 2 + 1] = make_float4(r[3], r[4], r[5], 0.f);
    }
}

// ---- extend: closest hit, lights first (closest_hit.glsl:7-67) ---------------------
LFH_DEV void light_closest(const ShadeScene& sc, V3 o, V3 d, float& t, int& light, float& lpdf) {
    for (int i = 0; i < sc.num_lights; ++i) {
        const float4* lp = sc.lights + (size_t)i * kLightF4;
        float4 a = lp[0], b = lp[1], c = lp[2];
        float type = a.w, area = b.w, radius = c.w;
        if (type == 0.0f) {
            V3 normal = ld3(lp[4]);
            if (dot(normal, d) > 0.0f) continue;
            float dd = rect_hit(lp, o, d);
            if (dd < 0.0f) dd = kInfinity;
            if (dd < t) {
                t = dd;
                float cosTheta = dot(-d, normal);
                lpdf = (t * t) / (area * cosTheta);
                light = i;
            }
        }
        if (type == 1.0f) {
            float dd = sphere_hit(radius, ld3(a), o, d);
            if (dd < 0.0f) dd = kInfinity;
            if (dd < t) {
                t = dd;
                lpdf = (t * t) / area;
                light = i;
            }
        }
    }
}

LFH_DEV bool light_any(const ShadeScene& sc, V3 o, V3 d, float maxDist) {
    for (int i = 0; i < sc.num_lights; ++i) {
        const float4* lp = sc.lights + (size_t)i * kLightF4;
        float4 a = lp[0], c = lp[2];
        if (a.w == 0.0f) {
            float dd = rect_hit(lp, o, d);
            if (dd > 0.0f && dd < maxDist) return true;
        }
        if (a.w == 1.0f) {
            float dd = sphere_hit(c.w, ld3(a), o, d);
            if (dd > 0.0f && dd < maxDist) return true;
        }
    }
    return false;
}

template <bool SHADOW, bool MULTI>
#ifndef LFH_QUEUE_WAVES
#define LFH_QUEUE_WAVES 6   // queue traversal register budget, waves per SIMD; with the octant slab 6: C0 1206 / C2 396 Msamples/s, 5: 1223 / 379 (C3 pass cost equal)
#endif
__global__ __launch_bounds__(kTraceBlock, LFH_QUEUE_WAVES) void k_trace_queue(ShadeScene sc, const JobDev* __restrict__ jobs,
                                                             const float4* __restrict__ q,
                                                             const uint32_t* __restrict__ count,
                                                             float4* __restrict__ hitq, PathBufs pb,
                                                             uint32_t* __restrict__ spill,
                                                             unsigned long long* __restrict__ work, float cull,
                                                             int* __restrict__ err) {
    __shared__ uint32_t stk[kTraceBlock / 64][kLdsStack][64];
    __shared__ float stkd[kTraceBlock / 64][LFH_POP_CULL ? kLdsStack : 1][64];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const size_t gtid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    LaneStack st;
    st.lds = &stk[wave][0][lane];
    st.ldk = &stkd[wave][0][lane];
    st.spill = spill + gtid * kSpillStride;
    st.reset();
    const int64_t n = (int64_t)*count;
    TraceState s;
    bool active = false;
    int64_t rid = 0;
    int light = -1;
    float lpdf = 0.f;
    TraceScene ts = sc.tr;
    WaveFetch fetch;
    for (;;) {
        const bool need = !active;
        const uint64_t mask = __ballot(need);
        if (mask) {
            const int64_t id = fetch.next(work, mask, lane, n, (int)(blockIdx.x % kFetchSlices));
            if (need) {
                rid = id;
                if (rid < n) {
                    float4 a = q[rid * 2], b = q[rid * 2 + 1];
                    V3 o = v3(a.x, a.y, a.z), d = v3(b.x, b.y, b.z);
                    const uint32_t tag = __float_as_uint(a.w);
                    const int pth = SHADOW ? (int)(tag >> 1) : (int)tag;
                    const ShadeScene& S = MULTI ? jobs[pb.job[pth]].sc : sc;
                    ts = S.tr;
                    if (SHADOW) {
                        float maxd = b.w;
                        trace_init(s, st, o, d, 0.0f, maxd, ts.root);
                        if (light_any(S, o, d, maxd)) {
                            s.occluded = true;
                            s.cur = kRefSentinel;  // finishes at once
                        }
                    } else {
                        float t = kInfinity;
                        light = -1;
                        lpdf = 0.f;
                        light_closest(S, o, d, t, light, lpdf);
                        trace_init(s, st, o, d, 0.0f, t, ts.root);
                    }
                    active = true;
                }
            }
        }
        if (__ballot(active) == 0) break;
        if (active) {
            bool done = trace_run<SHADOW, false>(ts, s, st, cull, SHADOW ? kAnySlice : kTraceSlice, nullptr);
            if (done) {
                if (st.overflow) atomicOr(err, 1);
                if (SHADOW) {
                    if (s.occluded) {
                        float4 a = q[rid * 2];
                        uint32_t tag = __float_as_uint(a.w);
                        uint32_t p = tag >> 1, kind = tag & 1u;
                        pb.nee[(size_t)p * 3 + kind].w = 0.0f;
                    }
                } else {
                    float4 h0, h1;
                    if (s.bestPrim >= 0) {
                        h0 = make_float4(s.tbest, s.bu, s.bv, __int_as_float(s.bestSlot));
                        h1 = make_float4(__int_as_float(s.bestInst), __int_as_float(s.bestMat), 0.f, __int_as_float(-1));
                    } else {
                        h0 = make_float4(s.tbest, 0.f, 0.f, __int_as_float(-1));
                        h1 = make_float4(__int_as_float(-1), __int_as_float(-1), lpdf, __int_as_float(light));
                    }
                    hitq[rid * 2] = h0;
                    hitq[rid * 2 + 1] = h1;
                }
                active = false;
            }
        }
    }
}

// ---- shade ------------------------------------------------------------------------
LFH_DEV void combine_nee(PathBufs& pb, int p, V3& rad) {
    float4 tn = pb.nee[(size_t)p * 3 + 2];
    if (tn.w != 0.0f) {
        float4 e = pb.nee[(size_t)p * 3 + 0], l = pb.nee[(size_t)p * 3 + 1];
        V3 Li = v3(0.f, 0.f, 0.f);
        if (e.w != 0.0f) Li = Li + ld3(e);
        if (l.w != 0.0f) Li = Li + ld3(l);
        rad = rad + Li * ld3(tn);
        pb.nee[(size_t)p * 3 + 2] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
}

// Shading class of hit-queue entry i (miss, light, then the material's
// dominant lobe group), for binning a block's entries before shading.
#ifndef LFH_SHADE_BIN
#define LFH_SHADE_BIN 0   // 1: a block shades its entries grouped by class (C2: 388.1 binned vs 390.6 Msamples/s; C0 unchanged)
#endif
constexpr int kShadeClasses = 6;
template <bool MULTI>
LFH_DEV int shade_class(const ShadeScene& sc1, const JobDev* jobs, const PathBufs& pb, const float4* rayq,
                        const float4* hitq, uint32_t i, uint32_t n) {
    if (i >= n) return kShadeClasses - 1;
    const float4 h0 = hitq[(size_t)i * 2], h1 = hitq[(size_t)i * 2 + 1];
    if (h0.x == kInfinity) return 0;
    if (__float_as_int(h0.w) < 0) return 1;
    const int p = __float_as_int(rayq[(size_t)i * 2].w);
    const ShadeScene& sc = MULTI ? jobs[pb.job[p]].sc : sc1;
    const float4* m = sc.materials + (size_t)__float_as_int(h1.y) * kMaterialF4;
    const float metallic = m[2].x, trans = m[4].x;
    if (trans > 0.0f && metallic < 1.0f) return 2;
    return metallic >= 1.0f ? 3 : 4;
}

// Block-local counting sort of the entries by class: returns the entry this
// thread shades.  Every entry of the block is shaded exactly once; a path's
// result does not depend on which thread shades it (each entry updates only
// its own path), so the permutation changes only the order of the appended
// rays.  `lds` holds kShadeClasses * 4 + 256 words.
LFH_DEV uint32_t bin_by_class(int cls, uint32_t* lds) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint64_t below = (1ull << lane) - 1ull;
    uint32_t rank = 0;
#pragma unroll
    for (int c = 0; c < kShadeClasses; ++c) {
        const uint64_t m = __ballot(cls == c);
        if (lane == 0) lds[c * 4 + wave] = (uint32_t)__popcll(m);
        if (cls == c) rank = (uint32_t)__popcll(m & below);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (int c = 0; c < kShadeClasses; ++c)
            for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
                const uint32_t k = lds[c * 4 + w];
                lds[c * 4 + w] = run;
                run += k;
            }
    }
    __syncthreads();
    uint32_t* perm = lds + kShadeClasses * 4;
    perm[lds[cls * 4 + wave] + rank] = threadIdx.x;
    __syncthreads();
    const uint32_t j = perm[threadIdx.x];
    __syncthreads();
    return j;
}

// GUIDED: the variant for guided wavefronts (guide sampling, MIS with the
// guide's pdf, training records); unguided wavefronts run the variant without
// that code, which would otherwise hold registers they never use.
template <bool MULTI, bool GUIDED>
#ifndef LFH_SHADE_WAVES
// k_shade register budget in waves per SIMD.  1 (153-158 VGPRs, 3 waves): C0 1207 /
// C2 391 Msamples/s, C3 1.463 / 2.134 ms per plain / guided pass; 4 (128 VGPRs,
// 100-140 B/lane spilled): 1208 / 384, 1.506 / 2.216; 5 (96 VGPRs, 244-276 B):
// 1037 / 363, 1.607 / 2.358.  The spills cost more than the occupancy gains.
#define LFH_SHADE_WAVES 1
#endif
__global__ __launch_bounds__(256, LFH_SHADE_WAVES) void k_shade(ShadeScene sc1, const JobDev* __restrict__ jobs, int depth,
                                               const float4* __restrict__ rayq,
                                               const float4* __restrict__ hitq, const uint32_t* __restrict__ count,
                                               float4* __restrict__ nextq, uint32_t* __restrict__ next_count,
                                               float4* __restrict__ shadowq, uint32_t* __restrict__ shadow_count,
                                               PathBufs pb, GuideHook gh) {
    const uint32_t n = *count;
#if LFH_SHADE_BIN
    __shared__ uint32_t bins[kShadeClasses * 4 + 256];
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if ((uint32_t)blockIdx.x * blockDim.x < n)   // block-uniform: blocks past the queue keep their order
        i = blockIdx.x * blockDim.x +
            bin_by_class(shade_class<MULTI>(sc1, jobs, pb, rayq, hitq, i, n), bins);
#else
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
#endif
    const bool live = i < n;
    bool cont = false;
    bool shEnv = false, shLight = false;
    V3 nextO = v3(0.f, 0.f, 0.f), nextD = v3(0.f, 0.f, 0.f);
    V3 shO = v3(0.f, 0.f, 0.f), shEnvD = v3(0.f, 0.f, 0.f), shLightD = v3(0.f, 0.f, 0.f);
    float shLightMax = 0.f;
    int p = 0;
    if (live) {
        float4 r0 = rayq[(size_t)i * 2], r1 = rayq[(size_t)i * 2 + 1];
        p = __float_as_int(r0.w);
        const ShadeScene& sc = MULTI ? jobs[pb.job[p]].sc : sc1;
        V3 ro = ld3(r0), rd = ld3(r1);
        float4 h0 = hitq[(size_t)i * 2], h1 = hitq[(size_t)i * 2 + 1];
        V3 rad = ld3(pb.rad[p]);
        V3 thr = ld3(pb.thr[p]);
        combine_nee(pb, p, rad);
        const bool recorded = GUIDED && gh.train && (p % gh.stride) == gh.phase;
        if (recorded && depth > 0) guide_rec(gh, p, depth - 1)[11] = luminance(rad);
        float4 misc = pb.misc[p];
        float prevPdf = misc.x;
        int lastMat = __float_as_int(misc.y);
        float2 texc = make_float2(misc.z, misc.w);   // last surface's texture coordinate
        const float t = h0.x;
        const int light = __float_as_int(h1.w);
        const int slot = __float_as_int(h0.w);
        if (t == kInfinity) {
            // miss (pathtrace.glsl:223-244)
            if (sc.const_bg) {
                rad = rad + v3(sc.bg[0], sc.bg[1], sc.bg[2]) * thr;
            } else if (sc.use_env) {
                float misWeight = 1.0f;
                float u = (kPi + atan2f(rd.z, rd.x)) * (1.0f / kTwoPi);
                float v = acosf(rd.y) * (1.0f / kPi);
                if (depth > 0) {
                    float lightPdf = EnvPdf(sc, rd);
                    misWeight = powerHeuristic(prevPdf, lightPdf);
                }
                rad = rad + misWeight * tex_linear(sc.hdr_cols, sc.hdr_w, sc.hdr_h, u, v) * thr * sc.hdr_mult;
            }
        } else if (slot < 0 && light >= 0) {
            // analytic light is the closest hit (pathtrace.glsl:246-260): the
            // material fetched is the last surface's (state is not refreshed),
            // its emission map read at that surface's texture coordinate
            Mat m = load_material(sc.materials, lastMat);
            rad = rad + textured_emission(sc, lastMat, make_float2(misc.z, misc.w), m.emission) * thr;
            V3 em = ld3(sc.lights[(size_t)light * kLightF4 + 1]);
            V3 Le = (depth == 0) ? em : powerHeuristic(prevPdf, h1.z) * em;
            rad = rad + Le * thr;
        } else {
            const int inst = __float_as_int(h1.x);
            const int matId = __float_as_int(h1.y);
            float u = h0.y, v = h0.z;
            float w = 1.0f - u - v;
            // hit point: object-space ray point, then the instance transform
            // (closest_hit.glsl:140-143)
            const float4* ip = sc.tr.insts + (size_t)inst * 4;
            float4 i0 = ip[0], i1 = ip[1], i2 = ip[2];
            V3 oo = xform_inv(i0, i1, i2, ro, 1.0f);
            V3 od = xform_inv(i0, i1, i2, rd, 0.0f);
            V3 po = oo + od * t;
            const float4* xp = sc.xforms + (size_t)inst * 4;
            float4 c0 = xp[0], c1 = xp[1], c2 = xp[2], c3 = xp[3];
            V3 fhp = v3(c0.x * po.x + c1.x * po.y + c2.x * po.z + c3.x * 1.0f,
                        c0.y * po.x + c1.y * po.y + c2.y * po.z + c3.y * 1.0f,
                        c0.z * po.x + c1.z * po.y + c2.z * po.z + c3.z * 1.0f);
            // shading normal (pathtrace.glsl:16-37)
            int4 vi = sc.vidx[slot];
            V3 n1 = ld3(sc.normals[vi.x]), n2 = ld3(sc.normals[vi.y]), n3 = ld3(sc.normals[vi.z]);
            V3 nrm = normalize(n1 * w + n2 * u + n3 * v);
            const float4* np = sc.nmats + (size_t)inst * 3;
            float4 m0 = np[0], m1 = np[1], m2 = np[2];
            nrm = normalize(v3(m0.x * nrm.x + m0.y * nrm.y + m0.z * nrm.z, m1.x * nrm.x + m1.y * nrm.y + m1.z * nrm.z,
                               m2.x * nrm.x + m2.y * nrm.y + m2.z * nrm.z));
            Mat m = load_material(sc.materials, matId);
            if (sc.tex) {
                const float2 t1 = sc.uvs[vi.x], t2 = sc.uvs[vi.y], t3 = sc.uvs[vi.z];
                texc = make_float2(t1.x * w + t2.x * u + t3.x * v, t1.y * w + t2.y * u + t3.y * v);
                apply_textures(sc, matId, texc, m, nrm);
            }
            Frame fr;
            fr.N = dot(nrm, rd) <= 0.0f ? nrm : nrm * -1.0f;
            onb(nrm, fr.T, fr.B);
            fr.eta = dot(rd, nrm) < 0.0f ? (1.0f / m.ior) : m.ior;
            V3 absorb = ld3(pb.absorb[p]);
            if (dot(nrm, fr.N) > 0.0f) absorb = v3(0.f, 0.f, 0.f);
            rad = rad + m.emission * thr;
            lastMat = matId;
            thr = thr * v3(expf(-absorb.x * t), expf(-absorb.y * t), expf(-absorb.z * t));
            uint4 rng = pb.rng[p];
            const float4* lob = gh.lobes + i;
            const size_t lst = (size_t)gh.lobe_stride;
            // the strategy choice depends only on the lobes, never on a sample
            const bool guided = GUIDED && gh.enabled != 0 && gh.sample != 0 && guide_valid(lob, lst);
            // next-event estimation (pathtrace.glsl:126-204)
            const V3 V = -rd;
            // the reference offsets along the unflipped normal (pathtrace.glsl:130), so
            // from behind a one-sided surface its shadow rays start on the far side;
            // the consistent estimator offsets to the side the ray came from
            V3 surfacePos = fhp + (sc.estimator == 1 ? fr.N : nrm) * kEps;
            shO = surfacePos;
            float4 e4 = make_float4(0.f, 0.f, 0.f, 0.f), l4 = make_float4(0.f, 0.f, 0.f, 0.f);
            if (sc.use_env) {
                V3 color, ldir;
                float lightPdf;
                EnvSample(sc, rng, color, ldir, lightPdf);
                float bp;
                V3 f = DisneyEval(m, fr.eta, V, fr.N, ldir, bp);
                if (bp > 0.0f) {
                    const float qp = guided ? (1.0f - kGuideFrac) * bp + kGuideFrac * guide_pdf(lob, ldir, lst) : bp;
                    float mw = powerHeuristic(lightPdf, qp);
                    if (mw > 0.0f) {
                        V3 c = mw * f * fabsf(dot(ldir, fr.N)) * color / lightPdf;
                        e4 = f4(c, 1.0f);
                        shEnv = true;
                        shEnvD = ldir;
                    }
                }
            }
            if (sc.num_lights > 0) {
                int index = (int)(rnd(rng) * (float)sc.num_lights);
                if (index >= sc.num_lights) index = sc.num_lights - 1;
                LightRec lr;
                sampleOneLight(sc, index, surfacePos, lr, rng);
                if (dot(lr.direction, lr.normal) < 0.0f) {
                    float bp;
                    V3 f = DisneyEval(m, fr.eta, V, fr.N, lr.direction, bp);
                    float area = sc.lights[(size_t)index * kLightF4 + 1].w;
                    float weight = 1.0f;
                    if (area > 0.0f)
                        weight = powerHeuristic(
                            lr.pdf, guided ? (1.0f - kGuideFrac) * bp + kGuideFrac * guide_pdf(lob, lr.direction, lst) : bp);
                    if (bp > 0.0f) {
                        V3 c = weight * f * fabsf(dot(fr.N, lr.direction)) * lr.emission / lr.pdf;
                        l4 = f4(c, 1.0f);
                        shLight = true;
                        shLightD = lr.direction;
                        shLightMax = lr.dist - kEps;
                    }
                }
            }
            if (shEnv || shLight) {
                pb.nee[(size_t)p * 3 + 0] = e4;
                pb.nee[(size_t)p * 3 + 1] = l4;
                pb.nee[(size_t)p * 3 + 2] = f4(thr, 1.0f);
            }
            // BSDF sampling (pathtrace.glsl:268-291)
            V3 L;
            float pdf, misPdf;
            V3 f;
            if (guided) {
                // one-sample MIS of the guide and the BSDF: the direction comes
                // from either, weighted by the mixture density q = (1-a) p_bsdf + a p_guide
                bool ok = true;
                if (rnd(rng) < kGuideFrac) {
                    const float u0 = rnd(rng), u1 = rnd(rng), u2 = rnd(rng);
                    L = guide_sample(lob, u0, u1, u2, lst);
                } else {
                    float pl;
                    DisneySample(m, fr, V, L, pl, rng);
                    ok = pl > 0.0f;
                }
                float bp;
                f = DisneyEval(m, fr.eta, V, fr.N, L, bp);
                pdf = ok ? (1.0f - kGuideFrac) * bp + kGuideFrac * guide_pdf(lob, L, lst) : 0.0f;
                misPdf = pdf;
            } else {
                f = DisneySample(m, fr, V, L, pdf, rng);
                misPdf = pdf;
                if (sc.estimator == 1 && pdf > 0.0f) {
                    // consistent MIS: the BSDF strategy's density is the lobe
                    // mixture's (disney.glsl:289-290), as in the light-sample weights
                    float bp;
                    DisneyEval(m, fr.eta, V, fr.N, L, bp);
                    misPdf = bp;
                }
            }
            if (dot(fr.N, L) < 0.0f)
                absorb = v3(-logf(m.extinction.x) / m.atDistance, -logf(m.extinction.y) / m.atDistance,
                            -logf(m.extinction.z) / m.atDistance);
            bool alive = pdf > 0.0f;
            const V3 thrBefore = thr;
            if (alive) thr = thr * ((f * fabsf(dot(fr.N, L))) / pdf);
            if (alive && sc.rr && depth >= sc.rr_depth) {
                float q = fminf(fmaxf(thr.x, fmaxf(thr.y, thr.z)) + 0.001f, 0.95f);
                if (rnd(rng) > q) alive = false;
                else thr = thr / q;
            }
            if (alive && depth + 1 < sc.max_depth) {
                cont = true;
                nextD = L;
                nextO = fhp + L * kEps;
            }
            if (recorded) {
                float* r = guide_rec(gh, p, depth);
                r[3] = cont ? 1.0f : 0.0f;
                r[7] = luminance(thrBefore);
                reinterpret_cast<float4*>(r)[3] = make_float4(L.x, L.y, L.z, 0.0f);
            }
            pb.rng[p] = rng;
            pb.thr[p] = f4(thr, 0.f);
            pb.absorb[p] = f4(absorb, 0.f);
            prevPdf = misPdf;
        }
        pb.rad[p] = f4(rad, 0.f);
        pb.misc[p] = make_float4(prevPdf, __int_as_float(lastMat), texc.x, texc.y);
    }
#if LFH_BLOCK_APPEND
    __shared__ uint32_t app[16];
    uint32_t s = block_append(next_count, cont, app);
    uint32_t sl;
    uint32_t se = block_append2(shadow_count, shEnv, shLight, &sl, app);
#else
    uint32_t s = wave_append(next_count, cont);
    uint32_t se = wave_append(shadow_count, shEnv);
    uint32_t sl = wave_append(shadow_count, shLight);
#endif
    if (cont) {
        nextq[(size_t)s * 2] = f4(nextO, __int_as_float(p));
        nextq[(size_t)s * 2 + 1] = f4(nextD, 0.f);
    }
    if (shEnv) {
        shadowq[(size_t)se * 2] = f4(shO, __uint_as_float(((uint32_t)p << 1) | 0u));
        shadowq[(size_t)se * 2 + 1] = f4(shEnvD, kInfinity - kEps);
    }
    if (shLight) {
        shadowq[(size_t)sl * 2] = f4(shO, __uint_as_float(((uint32_t)p << 1) | 1u));
        shadowq[(size_t)sl * 2 + 1] = f4(shLightD, shLightMax);
    }
}

// ---- accumulate: pending NEE + Kahan sum into the frame -----------------------------
template <bool MULTI>
__global__ __launch_bounds__(256) void k_accumulate(PathBufs pb, int total, CamParams cam1, float4* accum1,
                                                    float4* comp1, int kahan1, const JobDev* __restrict__ jobs) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= total) return;
    CamParams cam = cam1;
    float4* accum = accum1;
    float4* comp = comp1;
    int kahan = kahan1, lp = p;
    if (MULTI) {
        const JobDev& J = jobs[pb.job[p]];
        cam = J.cam;
        accum = J.accum;
        comp = J.comp;
        kahan = J.kahan;
        lp = p - J.path_offset;
    }
    if (lp >= cam.tile_w * cam.tile_h) return;
    const int lx = lp % cam.tile_w, ly = lp / cam.tile_w;
    const int gx = cam.tileX * cam.tile_w + lx, gy = cam.tileY * cam.tile_h + ly;
    if (gx >= cam.width || gy >= cam.height) return;
    const size_t pix = (size_t)gy * cam.width + gx;
    float4 s = accum[pix];
    float4 c = kahan ? comp[pix] : make_float4(0.f, 0.f, 0.f, 0.f);
    // the batch's samples of this pixel, added in frame order as separate passes would
    const int nb = cam.batch > 1 ? cam.batch : 1;
    const int P = cam.tile_w * cam.tile_h;
    for (int k = 0; k < nb; ++k) {
        const int q = p + k * P;
        V3 rad = ld3(pb.rad[q]);
        combine_nee(pb, q, rad);
        pb.rad[q] = f4(rad, 0.f);   // final path radiance (guide training targets)
        if (kahan) {
            float yx = rad.x - c.x, yy = rad.y - c.y, yz = rad.z - c.z;
            float tx = s.x + yx, ty = s.y + yy, tz = s.z + yz;
            c.x = (tx - s.x) - yx;
            c.y = (ty - s.y) - yy;
            c.z = (tz - s.z) - yz;
            s = make_float4(tx, ty, tz, s.w + 1.0f);
        } else {
            s = make_float4(s.x + rad.x, s.y + rad.y, s.z + rad.z, s.w + 1.0f);
        }
    }
    if (kahan) comp[pix] = c;
    accum[pix] = s;
}

// radiance out for lfh_path_trace_rays
__global__ void k_collect(PathBufs pb, int n, float* __restrict__ out) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    V3 rad = ld3(pb.rad[p]);
    combine_nee(pb, p, rad);
    out[(size_t)p * 3] = rad.x;
    out[(size_t)p * 3 + 1] = rad.y;
    out[(size_t)p * 3 + 2] = rad.z;
}

// linear HDR output: (sum - compensation) / samples, bottom-up rows
__global__ void k_output(const float4* __restrict__ accum, const float4* __restrict__ comp, int64_t npix, float inv,
                         float* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npix) return;
    float4 s = accum[i], c = comp[i];
    out[i * 3] = (s.x - c.x) * inv;
    out[i * 3 + 1] = (s.y - c.y) * inv;
    out[i * 3 + 2] = (s.z - c.z) * inv;
}

__global__ void k_clear_flag(int* f) {
    if (threadIdx.x == 0) f[0] = 0;
}

__global__ void k_clear4(float4* a, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// ---- per-stage parity kernels ---------------------------------------------------------
LFH_DEV void load_frame(const float* fr, Frame& f, V3& V, int& mat) {
    f.N = v3(fr[0], fr[1], fr[2]);
    f.T = v3(fr[3], fr[4], fr[5]);
    f.B = v3(fr[6], fr[7], fr[8]);
    V = v3(fr[9], fr[10], fr[11]);
    f.eta = fr[12];
    mat = (int)fr[13];
}

__global__ void k_eval_bsdf(ShadeScene sc, const float* frames, const float* L, int n, float* f, float* pdf) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Frame fr;
    V3 V;
    int mat;
    load_frame(frames + (size_t)i * 16, fr, V, mat);
    Mat m = load_material(sc.materials, mat);
    float p;
    V3 r = DisneyEval(m, fr.eta, V, fr.N, v3(L[i * 3], L[i * 3 + 1], L[i * 3 + 2]), p);
    f[i * 3] = r.x;
    f[i * 3 + 1] = r.y;
    f[i * 3 + 2] = r.z;
    pdf[i] = p;
}

__global__ void k_sample_bsdf(ShadeScene sc, const float* frames, uint4* seeds, int n, float* Lo, float* f,
                              float* pdf) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Frame fr;
    V3 V;
    int mat;
    load_frame(frames + (size_t)i * 16, fr, V, mat);
    Mat m = load_material(sc.materials, mat);
    uint4 s = seeds[i];
    V3 L = v3(0.f, 0.f, 0.f);
    float p;
    V3 r = DisneySample(m, fr, V, L, p, s);
    seeds[i] = s;
    Lo[i * 3] = L.x;
    Lo[i * 3 + 1] = L.y;
    Lo[i * 3 + 2] = L.z;
    f[i * 3] = r.x;
    f[i * 3 + 1] = r.y;
    f[i * 3 + 2] = r.z;
    pdf[i] = p;
}

__global__ void k_sample_light(ShadeScene sc, const float* pos, uint4* seeds, int n, float* dir, float* nrm,
                               float* em, float* dist, float* pdf) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint4 s = seeds[i];
    int index = (int)(rnd(s) * (float)sc.num_lights);
    if (index >= sc.num_lights) index = sc.num_lights - 1;
    LightRec r;
    sampleOneLight(sc, index, v3(pos[i * 3], pos[i * 3 + 1], pos[i * 3 + 2]), r, s);
    seeds[i] = s;
    dir[i * 3] = r.direction.x; dir[i * 3 + 1] = r.direction.y; dir[i * 3 + 2] = r.direction.z;
    nrm[i * 3] = r.normal.x; nrm[i * 3 + 1] = r.normal.y; nrm[i * 3 + 2] = r.normal.z;
    em[i * 3] = r.emission.x; em[i * 3 + 1] = r.emission.y; em[i * 3 + 2] = r.emission.z;
    dist[i] = r.dist;
    pdf[i] = r.pdf;
}

__global__ void k_sample_env(ShadeScene sc, uint4* seeds, int n, float* dir, float* pdf, float* col) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint4 s = seeds[i];
    V3 c, d;
    float p;
    EnvSample(sc, s, c, d, p);
    seeds[i] = s;
    dir[i * 3] = d.x; dir[i * 3 + 1] = d.y; dir[i * 3 + 2] = d.z;
    col[i * 3] = c.x; col[i * 3 + 1] = c.y; col[i * 3 + 2] = c.z;
    pdf[i] = p;
}

__global__ void k_env_pdf(ShadeScene sc, const float* dir, int n, float* pdf) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    pdf[i] = EnvPdf(sc, v3(dir[i * 3], dir[i * 3 + 1], dir[i * 3 + 2]));
}

}  // namespace lfhd

// ===========================================================================
// host launchers
// ===========================================================================
using namespace lfhd;

static inline int nblk(int64_t n, int b = 256) { return (int)((n + b - 1) / b); }

hipError_t pt_reset(uint32_t* counters, unsigned long long* work, hipStream_t s) {
    hipLaunchKernelGGL(k_pt_reset, dim3(1), dim3(64), 0, s, counters, work);
    return hipGetLastError();
}

hipError_t pt_generate(const CamParams& cam, const PathBufs& pb, float4* rayq, uint32_t* count, hipStream_t s) {
    int P = cam.tile_w * cam.tile_h * (cam.batch > 1 ? cam.batch : 1);
    hipLaunchKernelGGL(k_generate<false>, dim3(nblk(P)), dim3(256), 0, s, cam, (const JobDev*)nullptr, 1, P, pb, rayq,
                       count);
    return hipGetLastError();
}

hipError_t pt_generate_multi(const JobDev* jobs, int njobs, int total, const PathBufs& pb, float4* rayq,
                             uint32_t* count, hipStream_t s) {
    CamParams dummy{};
    hipLaunchKernelGGL(k_generate<true>, dim3(nblk(total)), dim3(256), 0, s, dummy, jobs, njobs, total, pb, rayq,
                       count);
    return hipGetLastError();
}

hipError_t pt_generate_rays(const float* rays6, const uint4* seeds, int n, const PathBufs& pb, float4* rayq,
                            uint32_t* count, hipStream_t s) {
    hipLaunchKernelGGL(k_generate_rays, dim3(nblk(n)), dim3(256), 0, s, rays6, seeds, n, pb, rayq, count);
    return hipGetLastError();
}

#ifndef LFH_LANE_GRID_NUM
// the lane's shadow traversal launches this many quarters of the resident grid: 2 / 3 / 4 quarters
// measured C0 1287.8 / 1255.5 / 1228.6 Msamples/s, C2 429.0 / 438.9 / 440.0, guided C3 pass
// 2.217 / 2.113 / 2.072 ms
#define LFH_LANE_GRID_NUM 4
#endif
#ifndef LFH_INFER_AFTER_SHADOW
#define LFH_INFER_AFTER_SHADOW 0   // 1: guide inference waits for the shadow traversal too (median
                                   // guided C3 time to 1e-2 over 32 sequences 0.158 s against 0.156)
#endif
// Shading of `depth` waits for the lane's shadow traversal of depth - 1.
static hipError_t wait_shadowed(const ShadowLane* lane, int depth, hipStream_t s) {
    return lane && depth > 0 ? hipStreamWaitEvent(s, lane->shadowed, 0) : hipSuccess;
}

// After shading: the extension queue's counters are reset on the main stream,
// and the shadow traversal with its reset runs on the lane.
template <bool MULTI>
static hipError_t shadow_on_lane(const ShadeScene& sc, const JobDev* jobs, uint32_t* count, float4* hitq,
                                 float4* shadowq, uint32_t* shadow_count, const PathBufs& pb,
                                 unsigned long long* work, int* err, float cull, int grid_blocks,
                                 const ShadowLane& lane, hipStream_t s) {
    hipLaunchKernelGGL(k_pt_advance_queue, dim3(1), dim3(64), 0, s, count, work, work + kWorkTotals);
    hipError_t e = hipEventRecord(lane.shaded, s);
    if (e == hipSuccess) e = hipStreamWaitEvent(lane.s, lane.shaded, 0);
    if (e != hipSuccess) return e;
    const int lane_blocks = grid_blocks * LFH_LANE_GRID_NUM / 4 > 0 ? grid_blocks * LFH_LANE_GRID_NUM / 4 : 1;
    hipLaunchKernelGGL((k_trace_queue<true, MULTI>), dim3(lane_blocks), dim3(kTraceBlock), 0, lane.s, sc, jobs,
                       shadowq, shadow_count, hitq, pb, lane.spill, work + kFetchWords, cull, err);
    hipLaunchKernelGGL(k_pt_advance_queue, dim3(1), dim3(64), 0, lane.s, shadow_count, work + kFetchWords,
                       work + kWorkTotals + 1);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipEventRecord(lane.shadowed, lane.s);
    return e;
}

hipError_t pt_bounce_join(const ShadowLane& lane, hipStream_t s) { return hipStreamWaitEvent(s, lane.shadowed, 0); }

hipError_t pt_bounce(const ShadeScene& sc, int depth, int maxq, float4* rayq, uint32_t* count, float4* nextq,
                     uint32_t* next_count, float4* hitq, float4* shadowq, uint32_t* shadow_count, const PathBufs& pb,
                     uint32_t* spill, unsigned long long* work, int* err, float cull, int grid_blocks,
                     const GuideHook& gh, const lfhd::GuideDev* gd, hipStream_t s, const ShadowLane* lane) {
    const JobDev* nj = nullptr;
    hipLaunchKernelGGL((k_trace_queue<false, false>), dim3(grid_blocks), dim3(kTraceBlock), 0, s, sc, nj, rayq, count,
                       hitq, pb, spill, work, cull, err);
    if (LFH_INFER_AFTER_SHADOW && gh.enabled) {
        hipError_t e = wait_shadowed(lane, depth, s);
        if (e != hipSuccess) return e;
    }
    if (gh.enabled) {
        hipError_t e = guide_infer_queue(*gd, sc, rayq, hitq, count, maxq, gh, depth,
                                         const_cast<float4*>(gh.lobes), s);
        if (e != hipSuccess) return e;
    }
    hipError_t e = wait_shadowed(lane, depth, s);
    if (e != hipSuccess) return e;
    if (gh.enabled)
        hipLaunchKernelGGL((k_shade<false, true>), dim3(nblk(maxq)), dim3(256), 0, s, sc, nj, depth, rayq, hitq, count,
                           nextq, next_count, shadowq, shadow_count, pb, gh);
    else
        hipLaunchKernelGGL((k_shade<false, false>), dim3(nblk(maxq)), dim3(256), 0, s, sc, nj, depth, rayq, hitq, count,
                           nextq, next_count, shadowq, shadow_count, pb, gh);
    if (lane)
        return shadow_on_lane<false>(sc, nj, count, hitq, shadowq, shadow_count, pb, work, err, cull, grid_blocks,
                                     *lane, s);
    hipLaunchKernelGGL((k_trace_queue<true, false>), dim3(grid_blocks), dim3(kTraceBlock), 0, s, sc, nj, shadowq,
                       shadow_count, hitq, pb, spill, work + kFetchWords, cull, err);
    hipLaunchKernelGGL(k_pt_advance, dim3(1), dim3(64), 0, s, count, shadow_count, work);
    return hipGetLastError();
}

hipError_t pt_bounce_multi(const JobDev* jobs, int depth, int maxq, float4* rayq, uint32_t* count, float4* nextq,
                           uint32_t* next_count, float4* hitq, float4* shadowq, uint32_t* shadow_count,
                           const PathBufs& pb, uint32_t* spill, unsigned long long* work, int* err, float cull,
                           int grid_blocks, const GuideHook& gh, const lfhd::GuideDev* gd, const ShadeScene* guide_sc,
                           hipStream_t s, const ShadowLane* lane) {
    ShadeScene dummy{};
    hipLaunchKernelGGL((k_trace_queue<false, true>), dim3(grid_blocks), dim3(kTraceBlock), 0, s, dummy, jobs, rayq,
                       count, hitq, pb, spill, work, cull, err);
    if (LFH_INFER_AFTER_SHADOW && gh.enabled) {
        hipError_t e = wait_shadowed(lane, depth, s);
        if (e != hipSuccess) return e;
    }
    if (gh.enabled) {
        // every job of a guided wavefront draws on one scene (the guide's)
        hipError_t e = guide_infer_queue(*gd, *guide_sc, rayq, hitq, count, maxq, gh, depth,
                                         const_cast<float4*>(gh.lobes), s);
        if (e != hipSuccess) return e;
    }
    hipError_t e = wait_shadowed(lane, depth, s);
    if (e != hipSuccess) return e;
    if (gh.enabled)
        hipLaunchKernelGGL((k_shade<true, true>), dim3(nblk(maxq)), dim3(256), 0, s, dummy, jobs, depth, rayq, hitq,
                           count, nextq, next_count, shadowq, shadow_count, pb, gh);
    else
        hipLaunchKernelGGL((k_shade<true, false>), dim3(nblk(maxq)), dim3(256), 0, s, dummy, jobs, depth, rayq, hitq,
                           count, nextq, next_count, shadowq, shadow_count, pb, gh);
    if (lane)
        return shadow_on_lane<true>(dummy, jobs, count, hitq, shadowq, shadow_count, pb, work, err, cull,
                                    grid_blocks, *lane, s);
    hipLaunchKernelGGL((k_trace_queue<true, true>), dim3(grid_blocks), dim3(kTraceBlock), 0, s, dummy, jobs, shadowq,
                       shadow_count, hitq, pb, spill, work + kFetchWords, cull, err);
    hipLaunchKernelGGL(k_pt_advance, dim3(1), dim3(64), 0, s, count, shadow_count, work);
    return hipGetLastError();
}

hipError_t pt_accumulate_multi(const JobDev* jobs, int total, const PathBufs& pb, hipStream_t s) {
    CamParams dummy{};
    hipLaunchKernelGGL(k_accumulate<true>, dim3(nblk(total)), dim3(256), 0, s, pb, total, dummy, (float4*)nullptr,
                       (float4*)nullptr, 0, jobs);
    return hipGetLastError();
}

hipError_t pt_accumulate(const PathBufs& pb, const CamParams& cam, float4* accum, float4* comp, int kahan,
                         hipStream_t s) {
    int P = cam.tile_w * cam.tile_h;
    hipLaunchKernelGGL(k_accumulate<false>, dim3(nblk(P)), dim3(256), 0, s, pb, P, cam, accum, comp, kahan,
                       (const JobDev*)nullptr);
    return hipGetLastError();
}

hipError_t pt_collect(const PathBufs& pb, int n, float* out, hipStream_t s) {
    hipLaunchKernelGGL(k_collect, dim3(nblk(n)), dim3(256), 0, s, pb, n, out);
    return hipGetLastError();
}

hipError_t pt_output(const float4* accum, const float4* comp, int64_t npix, float inv, float* out, hipStream_t s) {
    hipLaunchKernelGGL(k_output, dim3(nblk(npix)), dim3(256), 0, s, accum, comp, npix, inv, out);
    return hipGetLastError();
}

hipError_t pt_clear4(float4* a, int64_t n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_clear4, dim3(nblk(n)), dim3(256), 0, s, a, n);
    return hipGetLastError();
}

hipError_t pt_clear_flag(int* flag, hipStream_t s) {
    hipLaunchKernelGGL(k_clear_flag, dim3(1), dim3(64), 0, s, flag);
    return hipGetLastError();
}

int pt_trace_grid_blocks() {
    static int blocks = 0;
    if (blocks == 0) {
        int dev = 0, a = 0, b = 0;
        hipGetDevice(&dev);
        hipDeviceProp_t prop;
        hipGetDeviceProperties(&prop, dev);
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, (const void*)k_trace_queue<false, true>, kTraceBlock, 0) !=
                hipSuccess || a <= 0)
            a = 4;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, (const void*)k_trace_queue<false, false>, kTraceBlock, 0) !=
                hipSuccess || b <= 0)
            b = 4;
        blocks = (a > b ? a : b) * prop.multiProcessorCount;
    }
    return blocks;
}

hipError_t stage_eval_bsdf(const ShadeScene& sc, const float* fr, const float* L, int n, float* f, float* pdf,
                           hipStream_t s) {
    hipLaunchKernelGGL(k_eval_bsdf, dim3(nblk(n)), dim3(256), 0, s, sc, fr, L, n, f, pdf);
    return hipGetLastError();
}
hipError_t stage_sample_bsdf(const ShadeScene& sc, const float* fr, uint4* seeds, int n, float* L, float* f,
                             float* pdf, hipStream_t s) {
    hipLaunchKernelGGL(k_sample_bsdf, dim3(nblk(n)), dim3(256), 0, s, sc, fr, seeds, n, L, f, pdf);
    return hipGetLastError();
}
hipError_t stage_sample_light(const ShadeScene& sc, const float* pos, uint4* seeds, int n, float* dir, float* nrm,
                              float* em, float* dist, float* pdf, hipStream_t s) {
    hipLaunchKernelGGL(k_sample_light, dim3(nblk(n)), dim3(256), 0, s, sc, pos, seeds, n, dir, nrm, em, dist, pdf);
    return hipGetLastError();
}
hipError_t stage_sample_env(const ShadeScene& sc, uint4* seeds, int n, float* dir, float* pdf, float* col,
                            hipStream_t s) {
    hipLaunchKernelGGL(k_sample_env, dim3(nblk(n)), dim3(256), 0, s, sc, seeds, n, dir, pdf, col);
    return hipGetLastError();
}
hipError_t stage_env_pdf(const ShadeScene& sc, const float* dir, int n, float* pdf, hipStream_t s) {
    hipLaunchKernelGGL(k_env_pdf, dim3(nblk(n)), dim3(256), 0, s, sc, dir, n, pdf);
    return hipGetLastError();
}
