/*
 * lfh.h — C ABI of the MI355X (gfx950) path-tracing core.
 *
 * Every entry point takes plain pointers and sizes; no C++ or torch types cross
 * this boundary.  Status codes are returned as int (LFH_OK == 0, negative on
 * error) and the message of the last failure is available from
 * lfh_last_error().  No entry point throws or exits.
 *
 * Reference interfaces replaced (paths relative to the LavaFrame sources):
 *   Scene surface          LavaFrame/Scene.h:55-65, Scene.cpp:15-231
 *   .scene/.lfs loader     LavaFrame/Loader.cpp:45-386 (LoadSceneFromFile)
 *   Renderer / Tiled       LavaFrame/Renderer.h:71-117, TiledRenderer.h:57-71
 *   scene upload           LavaFrame/Renderer.cpp:76-188 (Renderer::Init)
 *   instance re-upload     LavaFrame/Renderer.cpp:190-204 (Renderer::Update)
 *   per-frame counters     LavaFrame/TiledRenderer.cpp:421-554 (Update)
 *   one pass / tile        LavaFrame/TiledRenderer.cpp:317-352 (Render)
 *   HDR readback           LavaFrame/TiledRenderer.cpp:399-414 (GetOutputBufferHDR)
 *   ClosestHit / AnyHit    shaders/common/closest_hit.glsl:7-205, anyhit.glsl:7-172
 *   DisneySample/Eval      shaders/common/disney.glsl:128-291
 *   light / env sampling   shaders/common/sampling.glsl:172-265
 *
 * Threading: one lfh_ctx per host thread, or external synchronisation.  A ctx
 * owns its device memory and launches on one HIP stream (its own, or the one
 * given to lfh_set_stream).  Host pointers passed in are borrowed only for the
 * duration of the call.
 */
#ifndef LFH_H
#define LFH_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LFH_ABI_VERSION 1

/* ---- status codes ------------------------------------------------------ */
#define LFH_OK            0
#define LFH_E_INVALID    -1   /* bad argument, size or layout               */
#define LFH_E_OOM        -2   /* device or host allocation failed           */
#define LFH_E_HIP        -3   /* HIP runtime error                          */
#define LFH_E_STATE      -4   /* call not valid in the current state        */
#define LFH_E_IO         -5   /* file could not be opened or parsed         */
#define LFH_E_UNSUPPORTED -6  /* feature outside this build's scope         */
#define LFH_E_BUDGET     -7   /* serving: job exceeds the device budget     */

typedef struct lfh_ctx lfh_ctx;       /* device-resident scene + renderer  */
typedef struct lfh_scene lfh_scene;   /* host-side scene builder           */

/* ---- record layouts ---------------------------------------------------- */

/* One ray, 32 B.  Hits are accepted for tmin <= t < tmax (the reference
 * accepts t >= 0 and t < INFINITY = 1e6, closest_hit.glsl:10,134). */
typedef struct lfh_ray {
    float org[3];
    float tmin;
    float dir[3];
    float tmax;
} lfh_ray;

/* One closest-hit result, 24 B.  prim = global triangle id (the first vertex
 * index of the hit triangle / 3, Scene.cpp:201-214), inst = instance index
 * (TLAS leaf -leaf-1, closest_hit.glsl:152), mat = instance material id,
 * (u, v) = barycentrics of vertices 1 and 2 (uvt.xy, closest_hit.glsl:131),
 * t = hit distance.  A miss has prim = inst = mat = -1 and t = tmax. */
typedef struct lfh_hit {
    int32_t prim;
    int32_t inst;
    int32_t mat;
    float t;
    float u;
    float v;
} lfh_hit;

/* Scene data contract: the reference's flattened arrays after
 * Scene::CreateAccelerationStructures (Scene.h:67-101, App. layouts):
 *   nodes        9 floats / node  {bboxmin, bboxmax, LRLeaf}   (bvh_translator.h:48-53)
 *   vert_indices 3 int32 / tri    (Scene.h:42-45)
 *   vertices_uvx 4 floats / vertex (x, y, z, u)
 *   normals_uvy  4 floats / vertex (nx, ny, nz, v)
 *   materials    28 floats / material (Material.h:48-75)
 *   transforms   16 floats / instance, Mat4 data[col][row] (Mat4.h:24)
 *   lights       15 floats / light {pos, emission, u, v, radius, area, type}
 *   hdr_*        env radiance RGB (3 floats/texel) + the two inverse-CDF tables
 *                (2 floats / entry), hdrloader.h:19-28.  hdr_width == 0: no env.
 *                With both tables NULL they are built on the device from the
 *                radiance (HDRLoader::buildDistributions, bit-identical).
 * All pointers are borrowed for the duration of lfh_create only.
 *
 * Device BVH build (no reference counterpart): with nodes == NULL the BVHs are
 * built on the GPU from the mesh and instance tables below, by lfh_create and
 * at serving admission.  Mesh m owns the triangles [mesh_tri_offsets[m],
 * mesh_tri_offsets[m + 1]) of vert_indices (ranges contiguous, in order, none
 * empty); instance i draws mesh instance_meshes[i] with material
 * instance_materials[i] and transforms[i].  Hits report the same prim ids as
 * for the reference tree.  lfh_scene_get_desc fills these tables for built and
 * flattened scenes alike. */
typedef struct lfh_scene_desc {
    const float*   nodes;
    int64_t        num_nodes;
    int32_t        top_level_index;
    const int32_t* vert_indices;
    int64_t        num_tris;
    const float*   vertices_uvx;
    const float*   normals_uvy;
    int64_t        num_vertices;
    const float*   materials;
    int32_t        num_materials;
    const float*   transforms;
    int32_t        num_instances;
    const float*   lights;
    int32_t        num_lights;
    int32_t        hdr_width;
    int32_t        hdr_height;
    const float*   hdr_cols;
    const float*   hdr_marginal;
    const float*   hdr_conditional;
    const int64_t* mesh_tri_offsets;   /* num_meshes + 1 */
    int32_t        num_meshes;
    const int32_t* instance_meshes;    /* num_instances */
    const int32_t* instance_materials; /* num_instances */
    /* Texture maps (Scene.h textureMapsArray / texWidth / texHeight): RGBA8,
     * num_textures layers of tex_width x tex_height texels, first row first.
     * A material's last four floats name its albedo, metallic-roughness,
     * normal and emission layers (-1: none). */
    const uint8_t* textures;
    int32_t        tex_width, tex_height, num_textures;
} lfh_scene_desc;

/* RenderOptions subset that reaches the path tracer (Renderer.h:18-67) and the
 * compile-time defines of TiledRenderer.cpp:77-108 (ENVMAP, LIGHTS, RR,
 * RR_DEPTH, CONSTANT_BG), which become kernel variants here. */
typedef struct lfh_options {
    int32_t width, height;           /* resolution                          */
    int32_t tile_width, tile_height; /* TiledRenderer tiles (0 = full frame) */
    int32_t max_depth;
    int32_t enable_rr;
    int32_t rr_depth;
    int32_t use_env_map;             /* ENVMAP iff use_env_map && hdr given  */
    int32_t use_constant_bg;         /* CONSTANT_BG                         */
    float   bg_color[3];
    float   hdr_multiplier;
    int32_t kahan;                   /* 1: compensated accumulation (default) */
    int32_t guiding;                 /* 1: reserve learned-guide buffers (lfh_set_guide) */
    int32_t estimator;               /* 0: the reference's MIS weights (drop-in);
                                      * 1: consistent MIS, the BSDF strategy weighted by
                                      *    its lobe-mixture pdf (disney.glsl:289-290) after
                                      *    BSDF samples too (unbiased; used with guiding) */
    int32_t frame_offset;            /* added to the frame number that seeds each pass's
                                      * random numbers (0 = the reference's sequence, >= 0);
                                      * see lfh_set_frame_offset.  No reference counterpart */
    int32_t reserved[4];
} lfh_options;

/* Post-process options: the RenderOptions fields behind postprocess.glsl's
 * uniforms (Renderer.h:18-67, set at TiledRenderer.cpp:541-552). */
typedef struct lfh_postprocess {
    int32_t tonemap_index;           /* 0 none, 1 linear curve, 2 ACES, 3 Reinhard, 4 Kanjero,
                                      * 5 Hejl-Dawson, 6 Uncharted 2 (postprocess.glsl:130-162) */
    int32_t use_ca;                  /* chromatic aberration (postprocess.glsl:96-116)      */
    int32_t use_ca_distortion;
    float   ca_distance, ca_p1, ca_p2, ca_p3;
    int32_t use_vignette;            /* postprocess.glsl:119-122, 164-167                    */
    float   vignette_intensity, vignette_power;
} lfh_postprocess;

/* Camera uniforms (TiledRenderer.cpp:507-513); fov in radians. */
typedef struct lfh_camera {
    float position[3];
    float right[3];
    float up[3];
    float forward[3];
    float fov;
    float focal_dist;
    float aperture;
} lfh_camera;

/* ---- library ----------------------------------------------------------- */
int         lfh_abi_version(void);
/* Message of the last failed call on this ctx (ctx == NULL: of the calling
 * thread's last failed ctx-less call).  Never NULL. */
const char* lfh_last_error(const lfh_ctx* ctx);
/* Device memory the library may use (bytes free / total on the current device). */
int         lfh_device_memory(int64_t* free_bytes, int64_t* total_bytes);

/* ---- host scene builder (Scene.h:55-65, Loader.h:15) -------------------- */
int  lfh_scene_create(lfh_scene** out);
void lfh_scene_destroy(lfh_scene* s);
/* LoadSceneFromFile: parses a .scene/.lfs file, adds its meshes, materials,
 * lights, camera and options, then builds the acceleration structures. */
int  lfh_scene_load(lfh_scene* s, const char* path);
/* AddMesh from an OBJ file; returns the mesh id in *id (-1 if it failed). */
int  lfh_scene_add_mesh_file(lfh_scene* s, const char* path, int32_t* id);
/* AddMesh from de-indexed triangle arrays: 3*num_tris vertices, xyz each;
 * normals may be NULL (face normals are used), uvs may be NULL (0). */
int  lfh_scene_add_mesh(lfh_scene* s, const float* positions, const float* normals,
                        const float* uvs, int64_t num_tris, int32_t* id);
int  lfh_scene_add_material(lfh_scene* s, const float* material28, int32_t* id);
/* Scene::AddTexture (Scene.cpp:45-66) with decoded RGBA8 pixels (rows top to
 * bottom, as stbi_load returns them); a name seen before returns its id. */
int  lfh_scene_add_texture(lfh_scene* s, const uint8_t* rgba8, int32_t width, int32_t height,
                           const char* name, int32_t* id);
/* Scene::AddTexture from a file: PNG or binary PPM/PGM. */
int  lfh_scene_add_texture_file(lfh_scene* s, const char* path, int32_t* id);
int  lfh_scene_add_light(lfh_scene* s, const float* light15, int32_t* id);
int  lfh_scene_add_instance(lfh_scene* s, int32_t mesh_id, const float* transform16,
                            int32_t material_id, int32_t* id);
/* AddCamera(eye, lookat, fov in degrees) followed by aperture / focal distance. */
int  lfh_scene_set_camera(lfh_scene* s, const float* eye3, const float* lookat3,
                          float fov_deg, float aperture, float focal_dist);
/* AddHDR from an RGB float image (3 floats / texel, row 0 first); builds the
 * importance tables of hdrloader.cpp:57-131. */
int  lfh_scene_set_hdr(lfh_scene* s, const float* rgb, int32_t width, int32_t height);
int  lfh_scene_set_options(lfh_scene* s, const lfh_options* opt);
int  lfh_scene_get_options(const lfh_scene* s, lfh_options* opt);
int  lfh_scene_get_camera(const lfh_scene* s, lfh_camera* cam);
/* Post-process options of a loaded scene (Renderer block of the .lfs,
 * Loader.cpp:262-282; defaults Renderer.h:20-44). */
int  lfh_scene_get_postprocess(const lfh_scene* s, lfh_postprocess* pp);
/* CreateAccelerationStructures (Scene.cpp:287-338). */
int  lfh_scene_build(lfh_scene* s);
/* The flattened arrays of CreateAccelerationStructures without its BVH builds
 * (triangles in mesh order): lfh_scene_get_desc then returns nodes == NULL,
 * for a device BVH build by lfh_create or lfh_server_submit. */
int  lfh_scene_flatten(lfh_scene* s);
/* RebuildInstances (Scene.cpp:272-285) after lfh_scene_set_instance_transform. */
int  lfh_scene_set_instance_transform(lfh_scene* s, int32_t inst, const float* transform16);
int  lfh_scene_rebuild_instances(lfh_scene* s);
/* Pointers into the scene's flattened arrays (valid until the scene changes). */
int  lfh_scene_get_desc(const lfh_scene* s, lfh_scene_desc* desc);
/* Number of meshes / instances and BLAS node count of mesh i. */
int  lfh_scene_counts(const lfh_scene* s, int32_t* num_meshes, int32_t* num_instances);

/* ---- device context: the Renderer surface ------------------------------- */
/* Renderer::Init: uploads every scene array (Renderer.cpp:76-188) and repacks
 * the BVH into the device layout. */
int  lfh_create(const lfh_scene_desc* desc, const lfh_options* opt, lfh_ctx** out);
/* Renderer::Finish. */
void lfh_destroy(lfh_ctx* ctx);
/* Launch on this HIP stream (hipStream_t as void*; NULL = the ctx's own). */
int  lfh_set_stream(lfh_ctx* ctx, void* hip_stream);
int  lfh_synchronize(lfh_ctx* ctx);
/* Renderer::Update's partial re-upload on instance edits (Renderer.cpp:192-203):
 * transforms (16 floats x num_instances), materials (28 floats x
 * num_materials) and TLAS nodes [first_node, first_node + count) (9 floats
 * each).  Any pointer may be NULL to leave that array unchanged.  A context
 * whose BVHs were built on the device takes no TLAS nodes (tlas_nodes NULL):
 * its TLAS is rebuilt on the device from the transforms. */
int  lfh_update_instances(lfh_ctx* ctx, const float* transforms, const float* materials,
                          const float* tlas_nodes, int64_t first_node, int64_t count);
/* TiledRenderer::Update: reset != 0 (camera moving / instances modified) clears
 * the accumulation and restarts the counters, else advances tile / sample /
 * frame counters exactly as TiledRenderer.cpp:471-501. */
int  lfh_update(lfh_ctx* ctx, const lfh_camera* cam, int reset);
/* TiledRenderer::Render: path-traces the current tile (one sample per pixel)
 * and accumulates it. */
int  lfh_render(lfh_ctx* ctx);
int  lfh_get_sample_count(const lfh_ctx* ctx, int32_t* out);
/* Offset added to the frame number that seeds each pass's random numbers
 * (InitRNG's frame, globals.glsl:116-120; default 0 = the reference's
 * sequence).  Contexts with disjoint frame ranges render independent samples
 * of the same image, e.g. one sample-index shard per GPU.  No reference
 * counterpart. */
int  lfh_set_frame_offset(lfh_ctx* ctx, int32_t offset);
/* Rays traced by this ctx's path tracer so far: closest-hit (camera and
 * bounce) rays and shadow rays.  No reference counterpart (the reference keeps
 * no counters); used for rays/s figures. */
int  lfh_get_ray_counts(lfh_ctx* ctx, int64_t* extended, int64_t* shadow);
int  lfh_get_frame(const lfh_ctx* ctx, int32_t* out);
int  lfh_get_progress(const lfh_ctx* ctx, float* out);
/* Linear HDR accum / samples, RGB32F, bottom-up rows like GL; dst is caller-
 * allocated w*h*3 floats and w, h must equal the resolution. */
int  lfh_get_output_hdr(lfh_ctx* ctx, float* dst, int32_t w, int32_t h);
/* The post-processed image (postprocess.glsl over the accumulation scaled by
 * inv_samples; inv_samples <= 0: 1 / sample counter), RGB float rows bottom-up:
 * what GetOutputBufferHDR reads (TiledRenderer.cpp:399-414); and its RGB8
 * read-back, round(255 clamp(c)), what GetOutputBuffer reads
 * (TiledRenderer.cpp:382-397).  Caller-allocated w*h*3 buffers. */
int  lfh_get_output_post(lfh_ctx* ctx, const lfh_postprocess* pp, float inv_samples, float* dst,
                         int32_t w, int32_t h);
int  lfh_get_output_ldr(lfh_ctx* ctx, const lfh_postprocess* pp, float inv_samples, uint8_t* dst,
                        int32_t w, int32_t h);
/* Raw accumulation sum (RGB, w*h*3) and its Kahan compensation (may be NULL). */
int  lfh_get_accum(lfh_ctx* ctx, float* sum, float* comp, int32_t w, int32_t h);
/* Render `passes` full passes back to back (each = Update(no reset) + Render
 * over all tiles); for benchmarks and serving. */
int  lfh_render_passes(lfh_ctx* ctx, int32_t passes);
/* The preview engine's frame (preview_flareon.glsl; TiledRenderer::Render
 * draws it while the camera moves or instances change, TiledRenderer.cpp:
 * 327-333): one sample per pixel of a (width * scale) x (height * scale)
 * frame with the current camera, RNG frame 1, path depth 2, no
 * accumulation; use_dof: thin-lens rays (USE_DOF, off by default in the
 * reference).  dst (may be NULL): w * h * 3 floats, rows bottom to top;
 * *w, *h receive the size.  The accumulation is left untouched. */
int  lfh_render_preview(lfh_ctx* ctx, float scale, int32_t use_dof, float* dst, int32_t* w, int32_t* h);

/* ---- direct traversal (microbenchmark + parity) -------------------------- */
/* Host arrays: copies in, traces, copies out, synchronises. */
int  lfh_trace_closest(lfh_ctx* ctx, const lfh_ray* rays, int64_t n, lfh_hit* hits);
int  lfh_trace_any(lfh_ctx* ctx, const lfh_ray* rays, int64_t n, uint8_t* occluded);
/* Device arrays: asynchronous on the ctx stream. */
int  lfh_trace_closest_device(lfh_ctx* ctx, const lfh_ray* d_rays, int64_t n, lfh_hit* d_hits);
int  lfh_trace_any_device(lfh_ctx* ctx, const lfh_ray* d_rays, int64_t n, uint8_t* d_occluded);
/* Traversal statistics per ray (internal-node visits, triangle tests, TLAS-leaf
 * visits), written to d_counts as 3 int32 per ray; for the roofline model. */
int  lfh_trace_closest_stats_device(lfh_ctx* ctx, const lfh_ray* d_rays, int64_t n,
                                    lfh_hit* d_hits, int32_t* d_counts);

/* ---- per-stage parity entry points (host arrays) -------------------------
 * Shading frame per query: 16 floats {N(3), T(3), B(3), V(3), eta, mat_id(as
 * float), pad, pad}.  Material ids index the ctx material table. */
/* DisneyEval: L (3 floats/query) -> f (3) + pdf (1). */
int  lfh_eval_bsdf(lfh_ctx* ctx, const float* frames, const float* L, int64_t n,
                   float* f, float* pdf);
/* DisneySample with the reference RNG: seeds (4 uint32/query, the pcg4d state,
 * advanced and written back) -> L (3) + f (3) + pdf (1). */
int  lfh_sample_bsdf(lfh_ctx* ctx, const float* frames, uint32_t* seeds, int64_t n,
                     float* L, float* f, float* pdf);
/* sampleOneLight after the light pick of DirectLight (pathtrace.glsl:168-181):
 * surface positions (3/query), seeds (4 uint32, advanced) -> direction (3),
 * normal (3), emission (3), dist (1), pdf (1). */
int  lfh_sample_light(lfh_ctx* ctx, const float* pos, uint32_t* seeds, int64_t n,
                      float* dir, float* normal, float* emission, float* dist, float* pdf);
/* EnvSample (sampling.glsl:246-265): seeds -> direction (3) + pdf (1) + colour (3). */
int  lfh_sample_env(lfh_ctx* ctx, uint32_t* seeds, int64_t n, float* dir, float* pdf,
                    float* color);
/* EnvPdf (sampling.glsl:236-243) for directions (3/query). */
int  lfh_env_pdf(lfh_ctx* ctx, const float* dir, int64_t n, float* pdf);
/* One path sample per query with the reference estimator (PathTrace,
 * pathtrace.glsl:208-295) from given rays and RNG seeds; radiance (3/query). */
int  lfh_path_trace_rays(lfh_ctx* ctx, const float* rays6, uint32_t* seeds, int64_t n,
                         float* radiance);

/* lfh_path_trace_rays on device arrays with no device allocation in the
 * call.  No counterpart in the reference: it is for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  d_rays6: n x 6 floats (origin, direction); d_seeds: n x 4
 * uint32 pcg4d states (read only); d_radiance: n x 3 floats; all device
 * pointers, asynchronous on the ctx stream.  The path state lives in a static
 * device array of about 19 MiB per GPU (held until the process ends), so n is
 * at most 65536 (larger n: LFH_E_INVALID); every call on one GPU shares that
 * array, so calls that overlap on one GPU (other streams or host threads)
 * race.  Radiance is bitwise that of lfh_path_trace_rays. */
int  lfh_path_trace_rays_device_noalloc(lfh_ctx* ctx, const float* d_rays6, const uint32_t* d_seeds, int64_t n,
                                        float* d_radiance);

/* ---- serving: continuous-batching render service ------------------------ */
typedef struct lfh_server lfh_server;
typedef struct lfh_job_desc {
    const lfh_scene_desc* scene;   /* borrowed until lfh_server_submit returns */
    lfh_options           options;
    lfh_camera            camera;
    int32_t               target_spp;
    int32_t               priority;
    int64_t               user_tag;
} lfh_job_desc;
typedef struct lfh_job_status {
    int64_t job_id;
    int32_t state;                 /* 0 queued, 1 running, 2 done, -1 rejected */
    int32_t spp_done;
    double  t_submit, t_admit, t_done;   /* seconds, server clock */
    int64_t device_bytes;
} lfh_job_status;

/* budget_bytes: device memory the server may hold across jobs (0 = 90% of free).
 * max_jobs: concurrent job slots (<= 64). */
int  lfh_server_create(int64_t budget_bytes, int32_t max_jobs, lfh_server** out);
void lfh_server_destroy(lfh_server* srv);
/* Queue a job; its scene arrays are copied.  Returns LFH_E_BUDGET (and
 * rejects) if the job can never fit the budget. */
int  lfh_server_submit(lfh_server* srv, const lfh_job_desc* job, int64_t* job_id);
/* Run one scheduling step: admit queued jobs that fit, render passes of the
 * running jobs in one wavefront (rays of all jobs packed into shared queues;
 * the step's path budget goes to the jobs by the server's policy, see
 * lfh_server_set_policy; the earliest deadline always advances), retire
 * finished jobs.  *active = number of jobs queued or running afterwards. */
int  lfh_server_step(lfh_server* srv, int32_t* active);
/* Step policy.  A job's deadline is its submit time + min(samples /
 * nominal_rate, max_offset) seconds.  policy 0: earliest deadline first, each
 * job in deadline order takes what the step's path budget still holds;
 * policy 1: fair share, every running job is offered an equal part of the
 * budget and what a job does not need goes to the others.  Default: policy 0,
 * nominal_rate 4.5e8 samples/s, no max offset.  No reference counterpart (the
 * reference renders one scene). */
int  lfh_server_set_policy(lfh_server* srv, int32_t policy, double nominal_rate, double max_offset);
int  lfh_server_status(lfh_server* srv, int64_t job_id, lfh_job_status* st);
/* Result of a finished job: linear HDR (w*h*3) like lfh_get_output_hdr. */
int  lfh_server_result(lfh_server* srv, int64_t job_id, float* dst, int32_t w, int32_t h);
/* Forgets a job and returns its memory.  A queued or running job is cancelled
 * (its image is discarded); the server goes on with its other jobs. */
int  lfh_server_release(lfh_server* srv, int64_t job_id);
/* Drops the guides of the resident scenes: the next guided job of a scene
 * starts a new guide from scratch (after the scene's lighting changed, or to
 * time training from scratch on a warm server).  LFH_E_STATE while a guided
 * job is queued or running.  No counterpart in the reference. */
int  lfh_server_reset_guides(lfh_server* srv);
/* Why a job was rejected (state -1): at submission (invalid scene, over the
 * budget) or at admission (scene upload or device BVH build failed, buffers
 * could not be allocated; the server goes on with its other jobs).  "" for
 * jobs that were not rejected.  Valid until the job is released. */
const char* lfh_server_job_error(lfh_server* srv, int64_t job_id);
double lfh_server_clock(const lfh_server* srv);

/* ---- learned path guiding ------------------------------------------------ */
typedef struct lfh_guide lfh_guide;
/* MLP: 64-dim frequency encoding of (position, normal, wo) -> 3 x 64 ReLU ->
 * 8 vMF lobes (mu 3, log kappa 1, weight logit 1 each = 40 outputs).  Weights
 * BF16, accumulation FP32, Adam at learning_rate, halved after 160 steps.  A
 * render server's guides use 1.5e-3. */
int  lfh_guide_create(const float* scene_min3, const float* scene_max3, uint64_t seed,
                      float learning_rate, lfh_guide** out);
void lfh_guide_destroy(lfh_guide* g);
/* Parameters as FP32 (master copy), for checks. n_params in/out. */
int  lfh_guide_get_params(lfh_guide* g, float* params, int64_t* n_params);
int  lfh_guide_set_params(lfh_guide* g, const float* params, int64_t n_params);
/* Forward on host arrays: inputs 9 floats/query (pos, normal, wo) -> 40 raw
 * outputs per query. */
int  lfh_guide_forward(lfh_guide* g, const float* inputs9, int64_t n, float* out40);
/* Forward on device arrays, asynchronous on `stream` (a hipStream_t; NULL:
 * the null stream): inputs9 and out40 are device pointers of n x 9 and n x 40
 * floats.  No reference counterpart: it serves callers whose queries are
 * already resident, and timing of the MLP alone. */
int  lfh_guide_forward_device(lfh_guide* g, const float* inputs9, int64_t n, float* out40, void* stream);
/* Loss and gradients for a batch of (inputs, target direction, weight) samples:
 * loss = -sum_i w_i log p(d_i | x_i) / n; grads written as FP32 (n_params). */
int  lfh_guide_loss_grad(lfh_guide* g, const float* inputs9, const float* dirs3,
                         const float* weights, int64_t n, float* loss, float* grads);
/* One Adam step (lr as created, beta 0.9 / 0.999, eps 1e-8) on a batch of at
 * most 65536 (inputs, direction, weight) samples; loss as in lfh_guide_loss_grad. */
int  lfh_guide_train_step(lfh_guide* g, const float* inputs9, const float* dirs3,
                          const float* weights, int64_t n, float* loss);
/* Number of Adam steps taken (explicit and online). */
int  lfh_guide_steps(const lfh_guide* g, int64_t* steps);
/* The mixture at each query: one direction drawn per query (seeds: 4 uint32
 * pcg4d states per query, read only) with its mixture pdf; or the pdf of given
 * directions. */
int  lfh_guide_sample(lfh_guide* g, const float* inputs9, const uint32_t* seeds, int64_t n,
                      float* dirs3, float* pdf);
int  lfh_guide_pdf(lfh_guide* g, const float* inputs9, const float* dirs3, int64_t n, float* pdf);
/* Attach (g != NULL) or detach a guide.  The ctx must have been created with
 * options.guiding = 1 and options.estimator = 1: the shading stage then draws
 * each BSDF direction from the guide or DisneySample with probability 1/2 and
 * weights it by the mixture density.  train_every > 0: every train_every-th
 * pass records training samples from a subset of its paths and takes one Adam
 * step per 65536 of them after the pass; 0 = frozen guide. */
int  lfh_set_guide(lfh_ctx* ctx, lfh_guide* g, int32_t train_every);
/* Training samples collected by a ctx's online training so far. */
int  lfh_guide_samples_seen(const lfh_ctx* ctx, int64_t* n);

/* ---- testing hook ------------------------------------------------------
 * No reference counterpart.  Makes the next `count` operations at `site` fail
 * as a device out-of-memory would (site 1: allocation of a render pass's
 * path state; site 2: scene upload at serving admission) or as a rejected
 * device TLAS rebuild would (site 3: stack need too deep).  count 0 disarms
 * the site.  For tests of the error paths; process-wide. */
int  lfh_debug_fail_next(int32_t site, int32_t count);

#ifdef __cplusplus
}
#endif
#endif /* LFH_H */
