// Host launch interface of the wavefront path-tracing stages.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dev_shade.h"

// Path state, indexed by path id (tile pixel).
struct PathBufs {
    float4* rad;     // radiance
    float4* thr;     // throughput
    float4* absorb;  // Beer-Lambert absorption coefficient
    uint4* rng;      // pcg4d state
    float4* misc;    // (previous BSDF pdf, last material id, 0, 0)
    float4* nee;     // 3 per path: env term, light term (w = 1 while unoccluded), (throughput, pending)
};

// Training-sample hook of the learned guide (disabled when enabled == 0):
// records per path vertex (position, normal, wo, sampled wi, BSDF pdf) and the
// throughput, for the radiance targets assembled after the pass.
struct GuideHook {
    int enabled;
    int max_depth;
    float4* rec;     // 4 float4 per (path, depth)
};

__device__ __forceinline__ void guide_record(const GuideHook& g, int p, int depth, lfhd::V3 x, lfhd::V3 n, lfhd::V3 wo,
                                             lfhd::V3 wi, float pdf, lfhd::V3 thr, bool alive) {
    float4* r = g.rec + ((size_t)p * g.max_depth + depth) * 4;
    r[0] = make_float4(x.x, x.y, x.z, alive ? 1.0f : 0.0f);
    r[1] = make_float4(n.x, n.y, n.z, pdf);
    r[2] = make_float4(wo.x, wo.y, wo.z, 0.0f);
    r[3] = make_float4(wi.x, wi.y, wi.z, (thr.x + thr.y + thr.z) * (1.0f / 3.0f));
}

hipError_t pt_reset(uint32_t* counters, unsigned long long* work, hipStream_t s);
hipError_t pt_generate(const lfhd::CamParams& cam, const PathBufs& pb, float4* rayq, uint32_t* count, hipStream_t s);
hipError_t pt_generate_rays(const float* rays6, const uint4* seeds, int n, const PathBufs& pb, float4* rayq,
                            uint32_t* count, hipStream_t s);
hipError_t pt_bounce(const lfhd::ShadeScene& sc, int depth, int maxq, float4* rayq, uint32_t* count, float4* nextq,
                     uint32_t* next_count, float4* hitq, float4* shadowq, uint32_t* shadow_count, const PathBufs& pb,
                     uint32_t* spill, unsigned long long* work, int* err, float cull, int grid_blocks,
                     const GuideHook& gh, hipStream_t s);
hipError_t pt_accumulate(const PathBufs& pb, const lfhd::CamParams& cam, float4* accum, float4* comp, int kahan,
                         hipStream_t s);
hipError_t pt_collect(const PathBufs& pb, int n, float* out, hipStream_t s);
hipError_t pt_output(const float4* accum, const float4* comp, int64_t npix, float inv, float* out, hipStream_t s);
hipError_t pt_clear4(float4* a, int64_t n, hipStream_t s);
int pt_trace_grid_blocks();

hipError_t stage_eval_bsdf(const lfhd::ShadeScene& sc, const float* fr, const float* L, int n, float* f, float* pdf,
                           hipStream_t s);
hipError_t stage_sample_bsdf(const lfhd::ShadeScene& sc, const float* fr, uint4* seeds, int n, float* L, float* f,
                             float* pdf, hipStream_t s);
hipError_t stage_sample_light(const lfhd::ShadeScene& sc, const float* pos, uint4* seeds, int n, float* dir,
                              float* nrm, float* em, float* dist, float* pdf, hipStream_t s);
hipError_t stage_sample_env(const lfhd::ShadeScene& sc, uint4* seeds, int n, float* dir, float* pdf, float* col,
                            hipStream_t s);
hipError_t stage_env_pdf(const lfhd::ShadeScene& sc, const float* dir, int n, float* pdf, hipStream_t s);
