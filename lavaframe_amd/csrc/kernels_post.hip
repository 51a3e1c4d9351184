// Post-process pass (shaders/postprocess.glsl:25-171) as one kernel per
// output pixel: texture reads of the accumulation with GL_LINEAR filtering
// and GL_MIRRORED_REPEAT wrapping (the accumulation texture's parameters,
// TiledRenderer.cpp:163-171), chromatic aberration, the seven tonemaps,
// vignette, and GetOutputBuffer's RGB8 conversion (TiledRenderer.cpp:382-397).
#include "dev_common.h"
#include "post_kernels.h"

namespace lfhd {

LFH_DEV int mirror_index(int i, int n) {
    int m = i % (2 * n);
    if (m < 0) m += 2 * n;
    return m < n ? m : 2 * n - 1 - m;
}

LFH_DEV float texel(const float4* accum, const float4* comp, int W, int H, int i, int j, int ch) {
    const size_t k = (size_t)mirror_index(j, H) * W + mirror_index(i, W);
    const float4 s = accum[k], c = comp[k];
    return ch == 0 ? s.x - c.x : (ch == 1 ? s.y - c.y : s.z - c.z);
}

// GL_LINEAR read of one channel at normalised (u, v)
LFH_DEV float sample_linear(const float4* accum, const float4* comp, int W, int H, float u, float v, int ch) {
    const float x = u * (float)W - 0.5f, y = v * (float)H - 0.5f;
    const float fx = floorf(x), fy = floorf(y);
    const int i0 = (int)fx, j0 = (int)fy;
    const float a = x - fx, b = y - fy;
    const float t00 = texel(accum, comp, W, H, i0, j0, ch), t10 = texel(accum, comp, W, H, i0 + 1, j0, ch);
    const float t01 = texel(accum, comp, W, H, i0, j0 + 1, ch), t11 = texel(accum, comp, W, H, i0 + 1, j0 + 1, ch);
    return (1.0f - a) * (1.0f - b) * t00 + a * (1.0f - b) * t10 + (1.0f - a) * b * t01 + a * b * t11;
}

LFH_DEV float clamp01(float x) { return fminf(fmaxf(x, 0.0f), 1.0f); }
LFH_DEV float gamma22(float x) { return powf(x, 1.0f / 2.2f); }

LFH_DEV void tonemap(float c[3], int index) {
    switch (index) {
    case 1: {  // tonemap(c, 2): c / (1 + lum / 2)
        const float lum = 0.3f * c[0] + 0.6f * c[1] + 0.1f * c[2];
        const float k = 1.0f / (1.0f + lum / 2.0f);
        for (int i = 0; i < 3; ++i) c[i] = gamma22(c[i] * k);
        break;
    }
    case 2:  // fitted ACES
        for (int i = 0; i < 3; ++i)
            c[i] = gamma22(clamp01((c[i] * (2.51f * c[i] + 0.03f)) / (c[i] * (2.43f * c[i] + 0.59f) + 0.14f)));
        break;
    case 3:  // Reinhard
        for (int i = 0; i < 3; ++i) c[i] = gamma22(clamp01(c[i] / (c[i] + 1.0f)));
        break;
    case 4:  // Kanjero
        for (int i = 0; i < 3; ++i) {
            float x = c[i];
            x = powf((x * (x * (1.2295f * x + 0.3135f) + 1.1935f * 0.4655f) / (x * (1.1935f * x + 0.4655f) + 0.073f)),
                     1.7f);
            x = powf(x, 1.0f / 0.8f) * 0.8f;
            c[i] = gamma22(clamp01(x));
        }
        break;
    case 5:  // Hejl-Dawson (no gamma)
        for (int i = 0; i < 3; ++i) {
            const float x = fmaxf(0.0f, c[i] - 0.004f);
            c[i] = (x * (6.2f * x + 0.5f)) / (x * (6.2f * x + 1.7f) + 0.06f);
        }
        break;
    case 6: {  // Uncharted 2, scaled by 1.75 after the gamma
        const float A = 0.15f, B = 0.50f, C = 0.10f, D = 0.20f, E = 0.02f, F = 0.30f;
        for (int i = 0; i < 3; ++i) {
            const float x = c[i];
            c[i] = gamma22(((x * (A * x + C * B) + D * E) / (x * (A * x + B) + D * F)) - E / F) * 1.75f;
        }
        break;
    }
    default:
        break;
    }
}

__global__ void k_post(const float4* __restrict__ accum, const float4* __restrict__ comp, int W, int H, float inv,
                       PostParams p, float* __restrict__ out_rgb, uint8_t* __restrict__ out_ldr) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= W * H) return;
    const int x = k % W, y = k / W;
    const float u = ((float)x + 0.5f) / (float)W, v = ((float)y + 0.5f) / (float)H;
    float c[3];
    if (p.use_ca) {
        float s;
        if (p.use_ca_distortion) {
            const float du = u - p.ca_p3, dv = v - p.ca_p3;
            const float dist = powf(sqrtf(du * du + dv * dv), p.ca_p1) * p.ca_p2;
            s = p.ca_distance * dist;
        } else {
            s = p.ca_distance * 0.025f * p.ca_p2;
        }
        c[0] = sample_linear(accum, comp, W, H, u + s, v + s, 0) * inv;
        c[1] = sample_linear(accum, comp, W, H, u, v, 1) * inv;
        c[2] = sample_linear(accum, comp, W, H, u - s, v - s, 2) * inv;
    } else {
        // a GL_LINEAR read at a texel centre is the texel
        const float4 s4 = accum[k], c4 = comp[k];
        c[0] = (s4.x - c4.x) * inv;
        c[1] = (s4.y - c4.y) * inv;
        c[2] = (s4.z - c4.z) * inv;
    }
    tonemap(c, p.tonemap_index);
    if (p.use_vignette) {
        const float du = u - 0.5f, dv = v - 0.5f;
        const float d = 1.0f - powf(sqrtf(du * du + dv * dv), p.vignette_power) * p.vignette_intensity;
        for (int i = 0; i < 3; ++i) c[i] *= d;
    }
    if (out_rgb)
        for (int i = 0; i < 3; ++i) out_rgb[(size_t)k * 3 + i] = c[i];
    if (out_ldr)
        for (int i = 0; i < 3; ++i) out_ldr[(size_t)k * 3 + i] = (uint8_t)floorf(clamp01(c[i]) * 255.0f + 0.5f);
}

hipError_t post_process(const float4* accum, const float4* comp, int width, int height, float inv,
                        const PostParams& p, float* out_rgb, uint8_t* out_ldr, hipStream_t s) {
    const int n = width * height;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_post, dim3((n + 255) / 256), dim3(256), 0, s, accum, comp, width, height, inv, p, out_rgb,
                       out_ldr);
    return hipGetLastError();
}

}  // namespace lfhd
