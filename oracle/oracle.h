/*
 * ORACLE — test infrastructure only.
 *
 * A CPU restatement of the reference's GLSL path-tracing core
 * (the shaders/common GLSL files and shaders/renderer.glsl) over the reference's own
 * flattened scene arrays (Scene.h:67-101), templated on the arithmetic type:
 * the FP32 instance is the parity checker and the CPU baseline, the FP64
 * instance the converged reference.  Each function cites the GLSL it follows.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * this library.  The product never links or calls it.
 *
 * Arithmetic conventions where GLSL leaves the evaluation open (these are
 * also what the product kernels implement, so parity is defined over them):
 *   - no FMA contraction; dot(a,b) = (a.x*b.x + a.y*b.y) + a.z*b.z;
 *   - a / b and 1.0 / x are IEEE divisions; normalize(v) = v / sqrt(dot(v,v));
 *   - min / max are IEEE minNum / maxNum (fminf / fmaxf);
 *   - inverse(mat4) = adjugate * (1 / det), cofactors by expansion along the
 *     first row (3x3 minors by the rule a00(a11a22 - a12a21) - a01(...) + ...);
 *   - mat * vec sums columns in order: ((c0 x + c1 y) + c2 z) + c3 w;
 *   - texture() filtering: NEAREST = texel floor(s * W) wrapped (GL_REPEAT),
 *     LINEAR = bilinear between floor(s * W - 0.5) and the next texel with
 *     weights frac(s * W - 0.5), wrapped.
 * Parity status of each function is listed in the product's design notes.
 */
#ifndef LFH_ORACLE_H
#define LFH_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_scene orc_scene;

/* pcg4d (globals.glsl:122-128) applied n times to one state. */
void orc_pcg4d(uint32_t* v4, int n);
/* rand() (globals.glsl:130-133): advances the state, returns float(x)/float(0xffffffff). */
float orc_rand(uint32_t* v4);

/* Scene from the reference flat arrays (copied).  hdr arrays may be NULL. */
orc_scene* orc_scene_create(const float* nodes, int64_t num_nodes, int32_t top_level_index,
                            const int32_t* vert_indices, int64_t num_tris,
                            const float* vertices_uvx, const float* normals_uvy, int64_t num_vertices,
                            const float* materials, int32_t num_materials,
                            const float* transforms, int32_t num_instances,
                            const float* lights, int32_t num_lights,
                            int32_t hdr_w, int32_t hdr_h, const float* hdr_cols,
                            const float* hdr_marginal, const float* hdr_conditional);
/* Texture maps: RGBA8 layers of width x height, first row first
 * (textureMapsArray, Scene.cpp:223-230). */
void orc_scene_set_textures(orc_scene* s, const uint8_t* rgba8, int32_t width, int32_t height, int32_t layers);
/* texture() reads (FP64) of layer[i] at texture coordinates st[2i], st[2i+1]. */
void orc_texture_sample(const orc_scene* s, const double* st, const int32_t* layer, int64_t n, double* out4);
void orc_scene_destroy(orc_scene* s);

/* ClosestHit over the BVH only (closest_hit.glsl:69-204, no analytic lights),
 * FP32.  rays: 8 floats (org, tmin, dir, tmax); hits: 6 x 4 B (prim, inst,
 * mat, t, u, v) like lfh_hit.  tie_mode 0: first found wins (the reference's
 * strict '<'); 1: lowest (instance, prim) wins exact-t ties.  counts (may be
 * NULL): 3 int32 per ray (internal visits, triangle tests, TLAS-leaf visits)
 * of the reference visit order.  Returns the number of exact-t ties seen. */
int64_t orc_trace_closest(const orc_scene* s, const float* rays, int64_t n, void* hits, int tie_mode,
                          int32_t* counts, int threads);
/* AnyHit over the BVH only (anyhit.glsl:48-172). */
void orc_trace_any(const orc_scene* s, const float* rays, int64_t n, uint8_t* occluded, int threads);

/* Options of the path tracer (the #defines of TiledRenderer.cpp:77-108 and
 * the per-frame uniforms). */
typedef struct orc_options {
    int32_t width, height, tile_width, tile_height;
    int32_t max_depth, enable_rr, rr_depth;
    int32_t envmap, lights, constant_bg;   /* ENVMAP / LIGHTS / CONSTANT_BG */
    float bg[3];
    float hdr_multiplier;
    float cam_pos[3], cam_right[3], cam_up[3], cam_fwd[3];
    float fov, focal_dist, aperture;
    int32_t estimator;   /* 0: the reference's MIS weights; 1: lobe-mixture pdf after BSDF samples */
} orc_options;

/* Shading frame of the per-stage entry points: 16 floats
 * {N(3), T(3), B(3), V(3), eta, mat_id, pad, pad}. */
/* DisneyEval (disney.glsl:228-291): f (3) + pdf. precision 32 or 64. */
void orc_eval_bsdf(const orc_scene* s, const float* frames, const float* L, int64_t n, float* f, float* pdf,
                   int precision);
/* DisneySample (disney.glsl:128-225) with the reference RNG (seeds advanced). */
void orc_sample_bsdf(const orc_scene* s, const float* frames, uint32_t* seeds, int64_t n, float* L, float* f,
                     float* pdf, int precision);
/* Light pick + sampleOneLight (pathtrace.glsl:168-181, sampling.glsl:172-230). */
void orc_sample_light(const orc_scene* s, const float* pos, uint32_t* seeds, int64_t n, float* dir,
                      float* normal, float* emission, float* dist, float* pdf, int precision);
/* EnvSample (sampling.glsl:246-265): dir (3), pdf, colour (3). */
void orc_sample_env(const orc_scene* s, uint32_t* seeds, int64_t n, float* dir, float* pdf, float* color,
                    float hdr_multiplier, int precision);
/* EnvPdf (sampling.glsl:236-243). */
void orc_env_pdf(const orc_scene* s, const float* dir, int64_t n, float* pdf, int precision);

/* PathTrace (pathtrace.glsl:208-295) for given rays (6 floats) and seeds. */
void orc_path_trace_rays(const orc_scene* s, const orc_options* o, const float* rays, uint32_t* seeds,
                         int64_t n, float* radiance, int precision, int threads);

/* One full pass of renderer.glsl:main over the frame (tile by tile, in the
 * TiledRenderer order) for frames [frame0, frame0 + passes): accum (w*h*3,
 * double) += each pass's samples.  Pixels are bottom-up rows like GL. */
/* renderer.glsl:27-62 camera rays of one tile (origin, direction per tile
 * pixel, row-major), for the given tile and frame. */
void orc_camera_rays(const orc_options* o, int32_t tileX, int32_t tileY, int32_t frame, double* rays6, int precision);
/* The preview engine (preview_flareon.glsl) over a pw x ph preview of the
 * o->width x o->height screen, one sample per pixel, o->max_depth as given. */
void orc_preview(const orc_scene* s, const orc_options* o, int32_t pw, int32_t ph, int32_t dof, double* out3,
                 int precision, int threads);
void orc_render(const orc_scene* s, const orc_options* o, int32_t frame0, int32_t passes, double* accum,
                int precision, int threads);
/* Same, for a subset of pixels only (pixel ids y*w + x), accumulating per
 * pixel sum and sum of squares (3 doubles each) over `spp` samples with
 * frames frame0 .. frame0+spp-1 (single tile). */
void orc_render_pixels(const orc_scene* s, const orc_options* o, const int32_t* pixels, int64_t npix,
                       int32_t frame0, int32_t spp, double* sum, double* sumsq, int precision, int threads);

#ifdef __cplusplus
}
#endif
#endif
