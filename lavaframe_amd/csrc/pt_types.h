// Device views used by the path-tracing stages.
#pragma once
#include "dev_common.h"

namespace lfhd {

// Light record, 7 float4 (112 B):
//   [0] (position, type)  [1] (emission, area)  [2] (u, radius)  [3] (v, 0)
//   [4] (n = normalize(cross(u, v)), dot(n, position))   — quad lights
//   [5] (u * (1 / dot(u, u)), 0)  [6] (v * (1 / dot(v, v)), 0)
// [4..6] are the per-ray quantities of ClosestHit (closest_hit.glsl:30-35)
// evaluated once with the same operations.
constexpr int kLightF4 = 7;
constexpr int kMaterialF4 = 7;

struct ShadeScene {
    TraceScene tr;
    const float4* xforms;     // 4 per instance, columns of M
    const float4* nmats;      // 3 per instance, rows of transpose(inverse(mat3(M)))
    const int4* vidx;         // per triangle slot
    const float4* normals;    // per vertex (n, uv.y)
    const float4* materials;  // 7 per material
    const float4* lights;     // 7 per light
    const float4* hdr_cols;   // W*H
    const float2* hdr_marg;   // H
    const float2* hdr_cond;   // W*H
    int hdr_w, hdr_h;
    float hdr_res;            // W*H (hdrResolution)
    float hdr_mult;
    int num_lights;
    int max_depth;
    int rr, rr_depth;
    int use_env;              // ENVMAP && !CONSTANT_BG
    int const_bg;             // CONSTANT_BG
    float bg[3];
    int estimator;            // 0: reference MIS weights; 1: mixture pdf after BSDF samples
    const uint32_t* tex;      // RGBA8 texture layers (NULL: no textures)
    const float2* uvs;        // per vertex (uv.x, uv.y) when textures exist
    int tex_w, tex_h, tex_layers;
};

// Camera uniforms + tile / frame of one pass (renderer.glsl inputs).
struct CamParams {
    float pos[3], right[3], up[3], fwd[3];
    float fov, focal_dist, aperture;
    int width, height, tile_w, tile_h, tileX, tileY;
    int frame;
    int batch;   // samples per pixel in this wavefront: path p is pixel p % P of frame + p / P
    int preview; // 1: the preview engine's camera (tile_w x tile_h = preview size, width x height = screen)
    int dof;     // preview only: thin-lens rays (USE_DOF)
};

// One job of a multi-job (serving) pass: its scene, camera and frame, its
// accumulation buffers, and the first global path id of its pixels.
struct JobDev {
    ShadeScene sc;
    CamParams cam;
    float4* accum;
    float4* comp;
    int path_offset;
    int kahan;
    int pad[2];
};

}  // namespace lfhd
