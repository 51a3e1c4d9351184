// Launchers of the guide kernels (kernels_guide.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "dev_guide.h"
#include "pt_types.h"

namespace lfhd {

struct GuideHook;

// Adam state and hyper-parameters of one optimisation step; c1, c2 are the
// bias corrections 1 / (1 - b1^t), 1 / (1 - b2^t).
struct GuideAdam {
    float* master;
    float* m;
    float* v;
    uint16_t* wf;
    uint16_t* wb;
    float* bias;
    float lr, b1, b2, eps;
    float c1, c2;
};

hipError_t guide_forward_host(const GuideDev& g, const float* in9, int n, float* raw40, float4* lobes,
                              hipStream_t s);
hipError_t guide_infer_queue(const GuideDev& g, const ShadeScene& sc, const float4* rayq, const float4* hitq,
                             const uint32_t* count, int maxq, const GuideHook& gh, int depth, float4* lobes,
                             hipStream_t s);
hipError_t guide_sample(const float4* lobes, const uint4* seeds, int n, float* dirs, float* pdf, hipStream_t s);
hipError_t guide_pdf(const float4* lobes, const float* dirs, int n, float* pdf, hipStream_t s);
int guide_train_blocks(int batch);
size_t guide_partial_floats(int batch);
// One batch of at most `batch` samples of the device sample list (count
// valid): every stride-th sample, stride = ceil(count / batch), from
// start % stride.  grad_out != NULL: store the gradient only; else one Adam step.
// wsum != NULL: sample weights are normalised by their mean (*wsum / *count)
// and capped at wcap.
hipError_t guide_train(const GuideDev& g, const float4* samples, const int* count, int start, int batch,
                       float* partial, const GuideAdam& adam, float* grad_out, float* loss_out, hipStream_t s,
                       const float* wsum = nullptr, float wcap = 0.0f);
hipError_t guide_refresh(const float* master, uint16_t* wf, uint16_t* wb, float* bias, hipStream_t s);
hipError_t guide_collect(const float4* rec, int nslots, int max_depth, int stride, int phase, const float4* rad,
                         int npaths, int* block_counts, int* block_offsets, int* total, float4* samples,
                         int max_samples, float* block_wsum, float* wsum, unsigned long long* seen,
                         hipStream_t s);
int guide_collect_blocks(int nslots);

}  // namespace lfhd
