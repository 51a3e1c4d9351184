// Host launch interface of the wavefront path-tracing stages.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dev_guide.h"
#include "pt_types.h"

// Path state, indexed by path id (tile pixel).
struct PathBufs {
    float4* rad;     // radiance
    float4* thr;     // throughput
    float4* absorb;  // Beer-Lambert absorption coefficient
    uint4* rng;      // pcg4d state
    float4* misc;    // (previous BSDF pdf, last material id, 0, 0)
    float4* nee;     // 3 per path: env term, light term (w = 1 while unoccluded), (throughput, pending)
    int* job;        // job id per path (multi-job passes only)
};

// Learned-guide hook of a pass.
//   enabled  the shade stage draws from the guide (lobes of the current hit
//            queue, kGuideLobeF4 float4 per entry) mixed with the BSDF
//   train    training records are written for paths p with p % stride == phase:
//            4 float4 per (recorded path, depth) = (pos, continued),
//            (n, lum throughput before the vertex), (wo, lum radiance before
//            the next vertex), (wi, 0)
namespace lfhd {
struct GuideHook {
    int enabled = 0;
    int sample = 1;   // 0: the guide is queried and trained but directions come from the BSDF alone
    int train = 0;
    int max_depth = 0;
    int stride = 1;
    int phase = 0;
    float4* rec = nullptr;
    // processed lobes of hit-queue entry i: float4 k at lobes[k * lobe_stride + i]
    // (structure of arrays, so each of a wave's ten loads is one contiguous run)
    const float4* lobes = nullptr;
    int lobe_stride = 0;
};
}  // namespace lfhd
using lfhd::GuideHook;


// Second stream of a wavefront's bounce chain.  The shadow-ray traversal of
// depth d and its counter reset run on it, concurrently with the extension
// traversal (and guide inference) of depth d + 1 on the main stream; shading
// of depth d + 1 waits for them (it reads the occlusion results and appends to
// the shadow queue again).  The two traversals use separate spill stacks.
struct ShadowLane {
    hipStream_t s = nullptr;
    uint32_t* spill = nullptr;     // as large as the main stream's
    hipEvent_t shaded = nullptr;   // recorded on the main stream after shading
    hipEvent_t shadowed = nullptr; // recorded on s after the shadow traversal
};

hipError_t pt_reset(uint32_t* counters, unsigned long long* work, hipStream_t s);
hipError_t pt_generate(const lfhd::CamParams& cam, const PathBufs& pb, float4* rayq, uint32_t* count, hipStream_t s);
hipError_t pt_generate_rays(const float* rays6, const uint4* seeds, int n, const PathBufs& pb, float4* rayq,
                            uint32_t* count, hipStream_t s);
hipError_t pt_bounce(const lfhd::ShadeScene& sc, int depth, int maxq, float4* rayq, uint32_t* count, float4* nextq,
                     uint32_t* next_count, float4* hitq, float4* shadowq, uint32_t* shadow_count, const PathBufs& pb,
                     uint32_t* spill, unsigned long long* work, int* err, float cull, int grid_blocks,
                     const GuideHook& gh, const lfhd::GuideDev* gd, hipStream_t s,
                     const ShadowLane* lane = nullptr);
// After the last depth of a chain that used `lane`: the main stream waits for
// the lane's last shadow traversal.
hipError_t pt_bounce_join(const ShadowLane& lane, hipStream_t s);
// Multi-job pass (serving): every job renders its full frame once; rays of all
// jobs share the queues.  jobs: device table of njobs entries.
hipError_t pt_generate_multi(const lfhd::JobDev* jobs, int njobs, int total, const PathBufs& pb, float4* rayq,
                             uint32_t* count, hipStream_t s);
// gh.enabled: a guided wavefront, all of whose jobs render guide_sc's scene.
hipError_t pt_bounce_multi(const lfhd::JobDev* jobs, int depth, int maxq, float4* rayq, uint32_t* count,
                           float4* nextq, uint32_t* next_count, float4* hitq, float4* shadowq, uint32_t* shadow_count,
                           const PathBufs& pb, uint32_t* spill, unsigned long long* work, int* err, float cull,
                           int grid_blocks, const GuideHook& gh, const lfhd::GuideDev* gd,
                           const lfhd::ShadeScene* guide_sc, hipStream_t s, const ShadowLane* lane = nullptr);
hipError_t pt_accumulate_multi(const lfhd::JobDev* jobs, int total, const PathBufs& pb, hipStream_t s);
hipError_t pt_accumulate(const PathBufs& pb, const lfhd::CamParams& cam, float4* accum, float4* comp, int kahan,
                         hipStream_t s);
hipError_t pt_collect(const PathBufs& pb, int n, float* out, hipStream_t s);
hipError_t pt_output(const float4* accum, const float4* comp, int64_t npix, float inv, float* out, hipStream_t s);
hipError_t pt_clear4(float4* a, int64_t n, hipStream_t s);
hipError_t pt_clear_flag(int* flag, hipStream_t s);
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
