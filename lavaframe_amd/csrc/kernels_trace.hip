// Persistent traversal kernels: closest hit, any hit, closest hit + counters.
//
// One launch fills the chip with exactly as many 256-thread blocks as can be
// resident.  Lanes fetch rays dynamically: every idle lane of a wave asks for
// work in the same step and takes the next ids of a contiguous chunk its wave
// reserved with one atomic (WaveFetch), in rank order among the idle lanes, so
// refills stay coalesced while busy lanes keep traversing.
#include "dev_trace.h"
#include "kernels.h"

namespace lfhd {

__global__ void k_zero_fetch(unsigned long long* work) {
    if (threadIdx.x < kFetchSlices) work[threadIdx.x * kFetchStride] = 0ull;
}

template <bool ANY, bool STATS>
__global__ __launch_bounds__(kTraceBlock, ANY ? LFH_ANYHIT_WAVES : LFH_TRACE_WAVES) void k_trace(TraceScene sc, const float4* __restrict__ rays, int64_t n,
                                                       int32_t* __restrict__ hits, uint8_t* __restrict__ occl,
                                                       int32_t* __restrict__ counts, uint32_t* __restrict__ spill,
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

    TraceState s;
    TraceStats stats{0, 0, 0};
    bool active = false;
    int64_t rid = 0;
    WaveFetch fetch;

    for (;;) {
        const bool need = !active;
        const uint64_t mask = __ballot(need);
        if (mask) {
            const int64_t id = fetch.next(work, mask, lane, n, (int)(blockIdx.x % kFetchSlices));
            if (need) {
                rid = id;
                if (rid < n) {
                    float4 a = rays[rid * 2], b = rays[rid * 2 + 1];
                    trace_init(s, st, v3(a.x, a.y, a.z), v3(b.x, b.y, b.z), a.w, b.w, sc.root);
                    if (STATS) stats = TraceStats{0, 0, 0};
                    active = true;
                }
            }
        }
        if (__ballot(active) == 0) break;
        if (active) {
            bool done = trace_run<ANY, STATS>(sc, s, st, cull, ANY ? kAnySlice : kTraceSlice, &stats);
            if (done) {
                if (st.overflow) atomicOr(err, 1);
                if (ANY) {
                    occl[rid] = s.occluded ? 1 : 0;
                } else {
                    int32_t* h = hits + rid * 6;
                    h[0] = s.bestPrim;
                    h[1] = s.bestInst;
                    h[2] = s.bestMat;
                    h[3] = (int32_t)as_u(s.tbest);
                    h[4] = (int32_t)as_u(s.bu);
                    h[5] = (int32_t)as_u(s.bv);
                }
                if (STATS) {
                    counts[rid * 3 + 0] = stats.nint;
                    counts[rid * 3 + 1] = stats.ntri;
                    counts[rid * 3 + 2] = stats.ninst;
                }
                active = false;
            }
        }
    }
}

}  // namespace lfhd

using namespace lfhd;

static int occupancy_blocks(const void* fn) {
    int dev = 0, nb = 0;
    hipGetDevice(&dev);
    hipDeviceProp_t prop;
    hipGetDeviceProperties(&prop, dev);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, kTraceBlock, 0) != hipSuccess || nb <= 0) nb = 4;
    return nb * prop.multiProcessorCount;
}

// Resident blocks of each variant on the whole device: closest hit, any hit, closest hit with counts.
// The variants are register-budgeted for different waves per SIMD, so each launches its own
// resident count; blocks beyond it would only start once the work queue is empty.
struct VariantBlocks {
    int n[3];
};

static const int* variant_blocks() {
    static const VariantBlocks blocks = {{occupancy_blocks((const void*)k_trace<false, false>),
                                          occupancy_blocks((const void*)k_trace<true, false>),
                                          occupancy_blocks((const void*)k_trace<false, true>)}};
    return blocks.n;
}

int lfh_trace_grid_blocks() {
    const int* v = variant_blocks();
    int blocks = v[0] > v[1] ? v[0] : v[1];
    return blocks > v[2] ? blocks : v[2];
}

int lfh_trace_threads_per_block() { return kTraceBlock; }
uint32_t lfh_trace_spill_stride() { return kSpillStride; }

hipError_t lfh_launch_trace(const TraceLaunch& L, hipStream_t stream) {
    // L.grid_blocks sizes the spill buffer; each variant launches at most its own resident count
    const int* v = variant_blocks();
    const int which = L.any ? 1 : (L.counts ? 2 : 0);
    const int blocks = v[which] < L.grid_blocks ? v[which] : L.grid_blocks;
    hipLaunchKernelGGL(k_zero_fetch, dim3(1), dim3(64), 0, stream, L.work);
    if (L.any) {
        hipLaunchKernelGGL((k_trace<true, false>), dim3(blocks), dim3(kTraceBlock), 0, stream, L.scene, L.rays, L.n,
                           L.hits, L.occluded, L.counts, L.spill, L.work, L.cull, L.err);
    } else if (L.counts) {
        hipLaunchKernelGGL((k_trace<false, true>), dim3(blocks), dim3(kTraceBlock), 0, stream, L.scene, L.rays, L.n,
                           L.hits, L.occluded, L.counts, L.spill, L.work, L.cull, L.err);
    } else {
        hipLaunchKernelGGL((k_trace<false, false>), dim3(blocks), dim3(kTraceBlock), 0, stream, L.scene, L.rays, L.n,
                           L.hits, L.occluded, L.counts, L.spill, L.work, L.cull, L.err);
    }
    return hipGetLastError();
}
