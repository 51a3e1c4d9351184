// Environment-map importance tables on the device (HDRLoader::buildDistributions,
// hdrloader.cpp:57-131): per-row luminance CDFs, the row marginal, and the
// inverse-CDF lookups at (i + 1) / N.  The sums run in the reference's
// order (one thread per row, one for the marginal) and every operation
// rounds once (-ffp-contract=off), so the tables equal the host's bit for bit.
#include "dev_common.h"
#include "hdr_kernels.h"

namespace lfhd {
namespace {

__global__ __launch_bounds__(256) void k_hdr_rows(const float4* __restrict__ cols, int W, int H,
                                                  float* __restrict__ pdf2, float* __restrict__ cdf2,
                                                  float* __restrict__ rowsum) {
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= H) return;
    const size_t r = (size_t)j * W;
    float sum = 0.f;
    for (int i = 0; i < W; ++i) {
        const float4 c = cols[r + i];
        const float w = c.x * 0.3f + c.y * 0.6f + c.z * 0.1f;
        sum += w;
        pdf2[r + i] = w;
        cdf2[r + i] = sum;
    }
    for (int i = 0; i < W; ++i) {
        pdf2[r + i] /= sum;
        cdf2[r + i] /= sum;
    }
    rowsum[j] = sum;
}

// one thread: the column sums in row order, then both normalised
__global__ void k_hdr_marginal(const float* __restrict__ rowsum, int H, float* __restrict__ pdf1,
                               float* __restrict__ cdf1) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    float col = 0.f;
    for (int j = 0; j < H; ++j) {
        col += rowsum[j];
        pdf1[j] = rowsum[j];
        cdf1[j] = col;
    }
    for (int j = 0; j < H; ++j) {
        cdf1[j] /= col;
        pdf1[j] /= col;
    }
}

LFH_DEV int lower_bound(const float* a, int lo, int hi, float value) {
    while (lo < hi) {
        const int mid = lo + (hi - lo) / 2;
        if (a[mid] < value) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

__global__ __launch_bounds__(256) void k_hdr_lookup(const float* __restrict__ pdf2, const float* __restrict__ cdf2,
                                                    const float* __restrict__ pdf1, const float* __restrict__ cdf1,
                                                    int W, int H, float2* __restrict__ marg, float2* __restrict__ cond) {
    const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t n = (int64_t)W * H;
    if (k < H) {
        const int i = (int)k;
        const float inv = (float)(i + 1) / (float)H;
        const int row = lower_bound(cdf1, 0, H, inv);
        marg[i] = make_float2((float)row / (float)H, pdf1[i]);
    }
    if (k < n) {
        const int j = (int)(k / W), i = (int)(k % W);
        const float inv = (float)(i + 1) / (float)W;
        const int col = lower_bound(cdf2, j * W, (j + 1) * W, inv) - j * W;
        cond[k] = make_float2((float)col / (float)W, pdf2[k]);
    }
}

}  // namespace

size_t hdr_scratch_floats(int W, int H) { return (size_t)W * H * 2 + (size_t)H * 3; }

hipError_t hdr_tables(const float4* cols, int W, int H, float* scratch, float2* marg, float2* cond, hipStream_t s) {
    if (W <= 0 || H <= 0) return hipSuccess;
    float* pdf2 = scratch;
    float* cdf2 = pdf2 + (size_t)W * H;
    float* rowsum = cdf2 + (size_t)W * H;
    float* pdf1 = rowsum + H;
    float* cdf1 = pdf1 + H;
    hipLaunchKernelGGL(k_hdr_rows, dim3((H + 255) / 256), dim3(256), 0, s, cols, W, H, pdf2, cdf2, rowsum);
    hipLaunchKernelGGL(k_hdr_marginal, dim3(1), dim3(64), 0, s, rowsum, H, pdf1, cdf1);
    const int64_t n = (int64_t)W * H;
    hipLaunchKernelGGL(k_hdr_lookup, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, pdf2, cdf2, pdf1, cdf1, W, H,
                       marg, cond);
    return hipGetLastError();
}

}  // namespace lfhd
