// Environment-map importance tables built on the device (kernels_hdr.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace lfhd {
// Floats of scratch memory hdr_tables needs for a W x H map.
size_t hdr_scratch_floats(int W, int H);
// marginal (H float2) and conditional (W * H float2) tables of the RGB map
// cols (float4 per texel, rows first), as HDRLoader::buildDistributions builds them.
hipError_t hdr_tables(const float4* cols, int W, int H, float* scratch, float2* marg, float2* cond, hipStream_t s);
}  // namespace lfhd
