// Host-side launch interface of the device kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dev_common.h"

struct TraceLaunch {
    lfhd::TraceScene scene;
    const float4* rays;       // 2 float4 per ray (org, tmin), (dir, tmax)
    int64_t n;
    int32_t* hits;            // 6 int32 per ray (lfh_hit)
    uint8_t* occluded;        // 1 byte per ray
    int32_t* counts;          // 3 int32 per ray or NULL
    uint32_t* spill;          // grid_blocks * 256 * spill_stride
    unsigned long long* work; // kFetchSlices counters, kFetchStride words apart
    int* err;                 // bit 0: stack overflow
    float cull;
    bool any;
    int grid_blocks;
};

int lfh_trace_grid_blocks();
int lfh_trace_threads_per_block();
uint32_t lfh_trace_spill_stride();
hipError_t lfh_launch_trace(const TraceLaunch& L, hipStream_t stream);
