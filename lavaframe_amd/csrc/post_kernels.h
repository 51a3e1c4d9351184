// Post-process pass of the output (shaders/postprocess.glsl) on the device.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lfhd {

// Uniforms of postprocess.glsl:11-22 (set at TiledRenderer.cpp:541-552).
struct PostParams {
    int tonemap_index;        // 0 linear, 1 "linear" curve, 2 ACES, 3 Reinhard, 4 Kanjero, 5 Hejl, 6 Uncharted 2
    int use_ca, use_ca_distortion;
    float ca_distance, ca_p1, ca_p2, ca_p3;
    int use_vignette;
    float vignette_intensity, vignette_power;
};

// Reads the accumulation (sum - compensation, W x H, row 0 at the bottom) as
// the pass's RGB32F GL_LINEAR / GL_MIRRORED_REPEAT texture, scales by inv and
// writes RGB float (out_rgb) and/or RGB8 (out_ldr, GL's round(255 clamp(c))).
hipError_t post_process(const float4* accum, const float4* comp, int width, int height, float inv,
                        const PostParams& p, float* out_rgb, uint8_t* out_ldr, hipStream_t s);

}  // namespace lfhd
