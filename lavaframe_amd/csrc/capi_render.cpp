// C ABI: the Renderer / TiledRenderer surface (Update, Render, counters,
// output) and the per-stage parity entry points.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <string>
#include <vector>

#include "capi_common.h"
#include "ctx.h"
#include "kernels.h"
#include "post_kernels.h"
#include "pt_kernels.h"
#include "render.h"

using lfh::DevBuf;

static int rfail(lfh_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    lfh_internal::set_thread_error(msg);
    return code;
}
static int rhip(lfh_ctx* c, hipError_t e, const char* where) {
    return rfail(c, e == hipErrorOutOfMemory ? LFH_E_OOM : LFH_E_HIP, std::string(where) + ": " + hipGetErrorString(e));
}
#define RCHK(c, expr, where)                              \
    do {                                                  \
        hipError_t _e = (expr);                           \
        if (_e != hipSuccess) return rhip(c, _e, where);  \
    } while (0)

namespace lfh {

// Replays g's captured chain, capturing `launch` first when the key differs.
hipError_t run_graph(BounceGraph& g, const std::vector<uintptr_t>& key, hipStream_t s,
                     const std::function<hipError_t()>& launch) {
    if (!g.exec || g.key != key) {
        g.reset();
        hipGraph_t graph = nullptr;
        hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) return e;
        hipError_t le = launch();
        e = hipStreamEndCapture(s, &graph);
        if (le != hipSuccess) e = le;
        if (e == hipSuccess) e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
        if (graph) (void)hipGraphDestroy(graph);
        if (e != hipSuccess) {
            g.exec = nullptr;
            return e;
        }
        g.key = key;
    }
    return hipGraphLaunch(g.exec, s);
}

PathBufs view(const PassBufs& b) {
    PathBufs p;
    p.rad = b.rad.as<float4>();
    p.thr = b.thr.as<float4>();
    p.absorb = b.absorb.as<float4>();
    p.rng = b.rng.as<uint4>();
    p.misc = b.misc.as<float4>();
    p.nee = b.nee.as<float4>();
    p.job = b.job.as<int>();
    return p;
}

}  // namespace lfh

namespace lfh {

int render_init(lfh_ctx* c) {
    const lfh_options& o = c->opt;
    c->tileW = o.tile_width > 0 ? o.tile_width : o.width;
    c->tileH = o.tile_height > 0 ? o.tile_height : o.height;
    if (c->tileW > o.width) c->tileW = o.width;
    if (c->tileH > o.height) c->tileH = o.height;
    c->numTilesX = (int)std::ceil((float)o.width / c->tileW);
    c->numTilesY = (int)std::ceil((float)o.height / c->tileH);
    c->tileX = -1;
    c->tileY = c->numTilesY - 1;
    c->sampleCounter = 1;
    c->frameCounter = 1;
    int64_t npix = (int64_t)o.width * o.height;
    c->rs.npix = npix;
    c->rs.npaths = c->tileW * c->tileH;
    hipError_t e = c->rs.accum.alloc(npix * sizeof(float4));
    if (e == hipSuccess) e = c->rs.comp.alloc(npix * sizeof(float4));
    if (e == hipSuccess) e = pt_clear4(c->rs.accum.as<float4>(), npix, c->stream);
    if (e == hipSuccess) e = pt_clear4(c->rs.comp.as<float4>(), npix, c->stream);
    if (e == hipSuccess) e = c->pass.alloc(c->rs.npaths);
    if (e == hipSuccess) e = c->rflag.alloc(64);
    if (e == hipSuccess) e = hipMemsetAsync(c->rflag.p, 0, 64, c->stream);
    if (e == hipSuccess) {
        c->pt_grid = pt_trace_grid_blocks();
        size_t need = (size_t)c->pt_grid * lfh_trace_threads_per_block() * lfh_trace_spill_stride() * 4;
        if (need > c->spill.bytes) e = c->spill.alloc(need);
        if (e == hipSuccess && shadow_lane_enabled()) e = c->lane.init(need);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return rhip(c, e, "render_init");
    return LFH_OK;
}

void render_release(lfh_ctx* c) {
    c->bgraph.reset();
    c->lane.release();
}

bool bounce_graphs_enabled() {
    static const bool on = [] {
        const char* e = std::getenv("LFH_BOUNCE_GRAPHS");
        return !(e && e[0] == '0');
    }();
    return on;
}

bool shadow_lane_enabled() {
    static const bool on = [] {
        const char* e = std::getenv("LFH_SHADOW_LANE");
        return !(e && e[0] == '0');
    }();
    return on;
}

// Reads (and clears) the overflow flag the path-tracing passes set when a
// traversal stack overflowed; the accumulation is then incomplete.  The
// scene's stack need is checked at lfh_create, so this is a safety net.
int check_render_flag(lfh_ctx* c) {
    int flag = 0;
    RCHK(c, hipMemcpyAsync(&flag, c->rflag.p, sizeof(int), hipMemcpyDeviceToHost, c->stream), "render flag");
    RCHK(c, hipStreamSynchronize(c->stream), "render flag");
    if (flag) {
        RCHK(c, pt_clear_flag(c->rflag.as<int>(), c->stream), "render flag");
        return rfail(c, LFH_E_UNSUPPORTED, "traversal stack overflow while rendering (image incomplete)");
    }
    return LFH_OK;
}

// Paths per wavefront targeted by multi-sample passes over one-tile frames
// (292 B of path state each, 9.8 GB at 2^25).  The last bounces of a
// wavefront hold few paths and cost a traversal's memory latency per launch,
// so several passes share them.  2^23 -> 2^24: instanced 1080p scene 326 ->
// 366 Msamples/s; guided / plain time on the guiding room 1.028 / 1.654 s ->
// 0.967 / 1.470 s.  2^24 -> 2^25 (the 16 passes of the 1080p scene in one
// wavefront instead of two): 416.4 / 417.1 -> 439.5 / 440.7 Msamples/s.
#ifndef LFH_BATCH_PATHS_LOG2
#define LFH_BATCH_PATHS_LOG2 25
#endif
constexpr int64_t kBatchPaths = (int64_t)1 << LFH_BATCH_PATHS_LOG2;
constexpr int kMaxBatch = 64;

// One path-traced pass over the current tile (TiledRenderer::Render's
// pathTraceShader + outputShader draws, TiledRenderer.cpp:334-344); batch > 1
// renders that many consecutive passes of a one-tile frame in one wavefront
// (frames frameCounter .. frameCounter + batch - 1), accumulated per pixel in
// frame order.
int render_tile(lfh_ctx* c, int batch) {
    lfhd::CamParams cp{};
    std::memcpy(cp.pos, c->cam.position, 12);
    std::memcpy(cp.right, c->cam.right, 12);
    std::memcpy(cp.up, c->cam.up, 12);
    std::memcpy(cp.fwd, c->cam.forward, 12);
    cp.fov = c->cam.fov;
    cp.focal_dist = c->cam.focal_dist;
    cp.aperture = c->cam.aperture;
    cp.width = c->opt.width;
    cp.height = c->opt.height;
    cp.tile_w = c->tileW;
    cp.tile_h = c->tileH;
    cp.tileX = c->tileX;
    cp.tileY = c->tileY;
    cp.frame = c->frameCounter + c->frameOffset;
    cp.batch = batch;
    {
        // also after a failed allocation, which leaves cap 0
        const int64_t need = (int64_t)c->rs.npaths * (batch > 1 ? batch : 1);
        if (c->pass.cap < need) {
            hipError_t e = hipStreamSynchronize(c->stream);
            if (e == hipSuccess) e = c->pass.alloc((int)need);
            if (e != hipSuccess) return rhip(c, e, "render: path buffers");
        }
    }
    PassBufs& pb = c->pass;
    lfhd::ShadeScene sc = shade_view(c->ps, c->ds, c->opt);
    uint32_t* cnt = pb.counters.as<uint32_t>();
    unsigned long long* work = pb.work.as<unsigned long long>();
    RCHK(c, pt_reset(cnt, work, c->stream), "render: reset");
    RCHK(c, pt_generate(cp, lfh::view(pb), pb.rq0.as<float4>(), cnt + 0, c->stream), "render: generate");
    GuideHook gh{};
    lfhd::GuideDev gd{};
    if (guide_pass_begin(c, batch, &gh, &gd) < 0) return rfail(c, LFH_E_HIP, "render: guide buffers");
    const ShadowLane lane = c->lane.lane();
    const ShadowLane* lp = c->lane.ready() ? &lane : nullptr;
    auto chain = [&]() -> hipError_t {
        for (int depth = 0; depth < c->opt.max_depth; ++depth) {
            const int cur = depth & 1;
            float4* q[2] = {pb.rq0.as<float4>(), pb.rq1.as<float4>()};
            hipError_t e = pt_bounce(sc, depth, pb.cap, q[cur], cnt + cur, q[cur ^ 1], cnt + (cur ^ 1),
                                     pb.hq.as<float4>(), pb.sq.as<float4>(), cnt + 2, lfh::view(pb),
                                     c->spill.as<uint32_t>(), work, c->rflag.as<int>(), c->cull, c->pt_grid, gh, &gd,
                                     c->stream, lp);
            if (e != hipSuccess) return e;
        }
        return lp && c->opt.max_depth > 0 ? pt_bounce_join(*lp, c->stream) : hipSuccess;
    };
    if (!gh.enabled && bounce_graphs_enabled()) {
        // the guide hook changes every pass; unguided chains are replayed
        const std::vector<uintptr_t> key = {(uintptr_t)pb.rad.p,  (uintptr_t)pb.cap,     (uintptr_t)c->opt.max_depth,
                                            (uintptr_t)c->stream, (uintptr_t)c->spill.p, (uintptr_t)c->scene_version,
                                            (uintptr_t)c->lane.s, (uintptr_t)c->lane.spill.p};
        RCHK(c, run_graph(c->bgraph, key, c->stream, chain), "render: bounce graph");
    } else {
        RCHK(c, chain(), "render: bounce");
    }
    RCHK(c, pt_accumulate(lfh::view(pb), cp, c->rs.accum.as<float4>(), c->rs.comp.as<float4>(), c->opt.kahan, c->stream),
         "render: accumulate");
    int gr = guide_pass_end(c, gh, batch);
    if (gr != LFH_OK) return rfail(c, gr, std::string("render: ") + lfh_internal::thread_error());
    return LFH_OK;
}

// TiledRenderer::Update's counter logic (TiledRenderer.cpp:471-501).
void advance_counters(lfh_ctx* c, bool reset) {
    if (reset) {
        c->tileX = -1;
        c->tileY = c->numTilesY - 1;
        c->sampleCounter = 1;
        c->frameCounter = 1;
    } else {
        c->frameCounter++;
        c->tileX++;
        if (c->tileX >= c->numTilesX) {
            c->tileX = 0;
            c->tileY--;
            if (c->tileY < 0) {
                c->tileX = 0;
                c->tileY = c->numTilesY - 1;
                c->sampleCounter++;
            }
        }
    }
}

// The preview engine's frame (preview_flareon.glsl, drawn by
// TiledRenderer::Render while the camera moves, TiledRenderer.cpp:327-333):
// one sample per pixel of a (width x scale) x (height x scale) frame, RNG
// frame 1, path depth 2 (the maxDepth uniform while moving,
// TiledRenderer.cpp:533), no accumulation.
int render_preview(lfh_ctx* c, float scale, int dof, int* pw, int* ph) {
    const int W = (int)((float)c->opt.width * scale), H = (int)((float)c->opt.height * scale);
    if (W <= 0 || H <= 0 || (int64_t)W * H > (int64_t)1 << 26)
        return rfail(c, LFH_E_INVALID, "lfh_render_preview: preview size out of range");
    const int64_t n = (int64_t)W * H;
    if (c->pass.cap < n) {
        hipError_t e = hipStreamSynchronize(c->stream);
        if (e == hipSuccess) e = c->pass.alloc((int)n);
        if (e != hipSuccess) return rhip(c, e, "lfh_render_preview: buffers");
    }
    // the frame, then a zero compensation buffer for the readback kernel
    if (c->preview.bytes < (size_t)n * 32) RCHK(c, c->preview.alloc((size_t)n * 32), "lfh_render_preview: alloc");
    RCHK(c, pt_clear4(c->preview.as<float4>(), 2 * n, c->stream), "lfh_render_preview: clear");
    lfhd::CamParams cp{};
    std::memcpy(cp.pos, c->cam.position, 12);
    std::memcpy(cp.right, c->cam.right, 12);
    std::memcpy(cp.up, c->cam.up, 12);
    std::memcpy(cp.fwd, c->cam.forward, 12);
    cp.fov = c->cam.fov;
    cp.focal_dist = c->cam.focal_dist;
    cp.aperture = c->cam.aperture;
    cp.width = c->opt.width;     // screenResolution (jitter and aspect)
    cp.height = c->opt.height;
    cp.tile_w = W;
    cp.tile_h = H;
    cp.frame = 1;
    cp.batch = 1;
    cp.preview = 1;
    cp.dof = dof;
    PassBufs& pb = c->pass;
    lfh_options o = c->opt;
    o.max_depth = 2;
    lfhd::ShadeScene sc = shade_view(c->ps, c->ds, o);
    uint32_t* cnt = pb.counters.as<uint32_t>();
    unsigned long long* work = pb.work.as<unsigned long long>();
    RCHK(c, pt_reset(cnt, work, c->stream), "lfh_render_preview: reset");
    RCHK(c, pt_generate(cp, lfh::view(pb), pb.rq0.as<float4>(), cnt + 0, c->stream), "lfh_render_preview: generate");
    GuideHook gh{};
    for (int depth = 0; depth < o.max_depth; ++depth) {
        const int cur = depth & 1;
        float4* q[2] = {pb.rq0.as<float4>(), pb.rq1.as<float4>()};
        RCHK(c,
             pt_bounce(sc, depth, pb.cap, q[cur], cnt + cur, q[cur ^ 1], cnt + (cur ^ 1), pb.hq.as<float4>(),
                       pb.sq.as<float4>(), cnt + 2, lfh::view(pb), c->spill.as<uint32_t>(), work, c->rflag.as<int>(),
                       c->cull, c->pt_grid, gh, nullptr, c->stream),
             "lfh_render_preview: bounce");
    }
    lfhd::CamParams ca = cp;   // accumulate over the preview frame itself
    ca.width = W;
    ca.height = H;
    RCHK(c, pt_accumulate(lfh::view(pb), ca, c->preview.as<float4>(), c->preview.as<float4>(), 0, c->stream),
         "lfh_render_preview: output");
    *pw = W;
    *ph = H;
    return LFH_OK;
}

}  // namespace lfh

extern "C" {

int lfh_render_preview(lfh_ctx* c, float scale, int32_t use_dof, float* dst, int32_t* w, int32_t* h) {
    if (!c || !w || !h) return rfail(c, LFH_E_INVALID, "lfh_render_preview: NULL argument");
    if (!c->has_cam) return rfail(c, LFH_E_STATE, "lfh_render_preview: lfh_update has not been called");
    if (!(scale > 0.0f) || scale > 4.0f) return rfail(c, LFH_E_INVALID, "lfh_render_preview: scale must be in (0, 4]");
    int W = 0, H = 0;
    int r = lfh::render_preview(c, scale, use_dof ? 1 : 0, &W, &H);
    if (r != LFH_OK) return r;
    *w = W;
    *h = H;
    if (dst) {
        const int64_t n = (int64_t)W * H;
        const size_t bytes = (size_t)n * 3 * sizeof(float);
        if (c->out.bytes < bytes) RCHK(c, c->out.alloc(bytes), "lfh_render_preview: alloc");
        RCHK(c, pt_output(c->preview.as<float4>(), c->preview.as<float4>() + n, n, 1.0f, c->out.as<float>(),
                          c->stream),
             "lfh_render_preview: output");
        RCHK(c, hipMemcpyAsync(dst, c->out.p, bytes, hipMemcpyDeviceToHost, c->stream), "lfh_render_preview: copy");
    }
    RCHK(c, hipStreamSynchronize(c->stream), "lfh_render_preview: sync");
    return LFH_OK;
}

int lfh_update(lfh_ctx* c, const lfh_camera* cam, int reset) {
    if (!c || !cam) return rfail(c, LFH_E_INVALID, "lfh_update: NULL argument");
    c->cam = *cam;
    c->has_cam = true;
    lfh::advance_counters(c, reset != 0);
    if (reset) {
        RCHK(c, pt_clear4(c->rs.accum.as<float4>(), c->rs.npix, c->stream), "lfh_update: clear");
        RCHK(c, pt_clear4(c->rs.comp.as<float4>(), c->rs.npix, c->stream), "lfh_update: clear");
    }
    return LFH_OK;
}

int lfh_render(lfh_ctx* c) {
    if (!c) return rfail(c, LFH_E_INVALID, "lfh_render: NULL ctx");
    if (!c->has_cam) return rfail(c, LFH_E_STATE, "lfh_render: lfh_update has not been called");
    if (c->tileX < 0) return LFH_OK;  // after a reset the reference draws only the preview
    return lfh::render_tile(c, 1);
}

int lfh_render_passes(lfh_ctx* c, int32_t passes) {
    if (!c || passes < 0) return rfail(c, LFH_E_INVALID, "lfh_render_passes: bad argument");
    if (!c->has_cam) return rfail(c, LFH_E_STATE, "lfh_render_passes: lfh_update has not been called");
    const int tiles = c->numTilesX * c->numTilesY;
    if (tiles == 1) {
        // one-tile frames: consecutive passes share one wavefront (with a guide,
        // once it has finished warming up)
        const int kmax = (int)std::min<int64_t>(lfh::kMaxBatch, std::max<int64_t>(1, lfh::kBatchPaths / c->rs.npaths));
        int done = 0;
        while (done < passes) {
            const int k = std::min({kmax, passes - done, lfh::guide_batch_limit(c)});
            lfh::advance_counters(c, false);
            const int frame0 = c->frameCounter;
            for (int i = 1; i < k; ++i) lfh::advance_counters(c, false);  // frames frame0 .. frame0 + k - 1
            const int keep = c->frameCounter;
            c->frameCounter = frame0;
            int r = lfh::render_tile(c, k);
            c->frameCounter = keep;
            if (r != LFH_OK) return r;
            done += k;
        }
        return LFH_OK;
    }
    for (int p = 0; p < passes; ++p) {
        for (int t = 0; t < tiles; ++t) {
            lfh::advance_counters(c, false);
            int r = lfh::render_tile(c, 1);
            if (r != LFH_OK) return r;
        }
    }
    return LFH_OK;
}

int lfh_set_frame_offset(lfh_ctx* c, int32_t offset) {
    if (!c || offset < 0) return rfail(c, LFH_E_INVALID, "lfh_set_frame_offset: bad argument");
    c->frameOffset = offset;
    return LFH_OK;
}

int lfh_get_ray_counts(lfh_ctx* c, int64_t* extended, int64_t* shadow) {
    if (!c || !extended || !shadow) return rfail(c, LFH_E_INVALID, "lfh_get_ray_counts: NULL argument");
    unsigned long long t[2] = {0, 0};
    RCHK(c, hipMemcpyAsync(t, c->pass.work.as<unsigned long long>() + lfhd::kWorkTotals, sizeof t, hipMemcpyDeviceToHost, c->stream),
         "lfh_get_ray_counts");
    RCHK(c, hipStreamSynchronize(c->stream), "lfh_get_ray_counts");
    *extended = (int64_t)t[0];
    *shadow = (int64_t)t[1];
    return LFH_OK;
}

int lfh_get_sample_count(const lfh_ctx* c, int32_t* out) {
    if (!c || !out) return rfail(nullptr, LFH_E_INVALID, "lfh_get_sample_count: NULL argument");
    *out = c->sampleCounter;
    return LFH_OK;
}

int lfh_get_frame(const lfh_ctx* c, int32_t* out) {
    if (!c || !out) return rfail(nullptr, LFH_E_INVALID, "lfh_get_frame: NULL argument");
    *out = c->frameCounter;
    return LFH_OK;
}

// TiledRenderer::GetProgress (TiledRenderer.cpp:377-380)
int lfh_get_progress(const lfh_ctx* c, float* out) {
    if (!c || !out) return rfail(nullptr, LFH_E_INVALID, "lfh_get_progress: NULL argument");
    *out = float((c->numTilesY - c->tileY - 1) * c->numTilesX + c->tileX) / float(c->numTilesX * c->numTilesY);
    return LFH_OK;
}

int lfh_get_output_hdr(lfh_ctx* c, float* dst, int32_t w, int32_t h) {
    if (!c || !dst) return rfail(c, LFH_E_INVALID, "lfh_get_output_hdr: NULL argument");
    if (w != c->opt.width || h != c->opt.height) return rfail(c, LFH_E_INVALID, "lfh_get_output_hdr: size mismatch");
    size_t bytes = (size_t)c->rs.npix * 3 * sizeof(float);
    if (c->out.bytes < bytes) RCHK(c, c->out.alloc(bytes), "lfh_get_output_hdr: alloc");
    float inv = 1.0f / (float)c->sampleCounter;
    RCHK(c, pt_output(c->rs.accum.as<float4>(), c->rs.comp.as<float4>(), c->rs.npix, inv, c->out.as<float>(), c->stream),
         "lfh_get_output_hdr: kernel");
    RCHK(c, hipMemcpyAsync(dst, c->out.p, bytes, hipMemcpyDeviceToHost, c->stream), "lfh_get_output_hdr: copy");
    RCHK(c, hipStreamSynchronize(c->stream), "lfh_get_output_hdr: sync");
    return lfh::check_render_flag(c);
}

static int get_post(lfh_ctx* c, const lfh_postprocess* pp, float inv, float* rgb, uint8_t* ldr, int32_t w,
                    int32_t h, const char* fn) {
    if (!c || !pp || (!rgb && !ldr)) return rfail(c, LFH_E_INVALID, std::string(fn) + ": NULL argument");
    if (w != c->opt.width || h != c->opt.height) return rfail(c, LFH_E_INVALID, std::string(fn) + ": size mismatch");
    if (pp->tonemap_index < 0 || pp->tonemap_index > 6)
        return rfail(c, LFH_E_INVALID, std::string(fn) + ": tonemap_index must be 0..6");
    lfhd::PostParams p;
    p.tonemap_index = pp->tonemap_index;
    p.use_ca = pp->use_ca != 0;
    p.use_ca_distortion = pp->use_ca_distortion != 0;
    p.ca_distance = pp->ca_distance;
    p.ca_p1 = pp->ca_p1;
    p.ca_p2 = pp->ca_p2;
    p.ca_p3 = pp->ca_p3;
    p.use_vignette = pp->use_vignette != 0;
    p.vignette_intensity = pp->vignette_intensity;
    p.vignette_power = pp->vignette_power;
    if (!(inv > 0.0f)) inv = 1.0f / (float)c->sampleCounter;
    const size_t n = (size_t)c->rs.npix * 3;
    const size_t bytes = rgb ? n * sizeof(float) : n;
    if (c->out.bytes < bytes) RCHK(c, c->out.alloc(bytes), fn);
    RCHK(c, lfhd::post_process(c->rs.accum.as<float4>(), c->rs.comp.as<float4>(), w, h, inv, p,
                               rgb ? c->out.as<float>() : nullptr, rgb ? nullptr : c->out.as<uint8_t>(), c->stream),
         fn);
    RCHK(c, hipMemcpyAsync(rgb ? (void*)rgb : (void*)ldr, c->out.p, bytes, hipMemcpyDeviceToHost, c->stream), fn);
    RCHK(c, hipStreamSynchronize(c->stream), fn);
    return lfh::check_render_flag(c);
}

int lfh_get_output_post(lfh_ctx* c, const lfh_postprocess* pp, float inv, float* dst, int32_t w, int32_t h) {
    return get_post(c, pp, inv, dst, nullptr, w, h, "lfh_get_output_post");
}

int lfh_get_output_ldr(lfh_ctx* c, const lfh_postprocess* pp, float inv, uint8_t* dst, int32_t w, int32_t h) {
    return get_post(c, pp, inv, nullptr, dst, w, h, "lfh_get_output_ldr");
}

int lfh_get_accum(lfh_ctx* c, float* sum, float* comp, int32_t w, int32_t h) {
    if (!c || !sum) return rfail(c, LFH_E_INVALID, "lfh_get_accum: NULL argument");
    if (w != c->opt.width || h != c->opt.height) return rfail(c, LFH_E_INVALID, "lfh_get_accum: size mismatch");
    std::vector<float> tmp((size_t)c->rs.npix * 4);
    RCHK(c, hipMemcpyAsync(tmp.data(), c->rs.accum.p, tmp.size() * 4, hipMemcpyDeviceToHost, c->stream), "accum copy");
    RCHK(c, hipStreamSynchronize(c->stream), "accum sync");
    for (int64_t i = 0; i < c->rs.npix; ++i)
        for (int k = 0; k < 3; ++k) sum[i * 3 + k] = tmp[i * 4 + k];
    if (comp) {
        RCHK(c, hipMemcpyAsync(tmp.data(), c->rs.comp.p, tmp.size() * 4, hipMemcpyDeviceToHost, c->stream), "comp copy");
        RCHK(c, hipStreamSynchronize(c->stream), "comp sync");
        for (int64_t i = 0; i < c->rs.npix; ++i)
            for (int k = 0; k < 3; ++k) comp[i * 3 + k] = tmp[i * 4 + k];
    }
    return LFH_OK;
}

// ---- per-stage parity entry points ------------------------------------------------
namespace {
struct Tmp {
    std::vector<DevBuf*> bufs;
    ~Tmp() {
        for (auto* b : bufs) delete b;
    }
    void* get(size_t bytes, hipError_t* e) {
        DevBuf* b = new DevBuf;
        *e = b->alloc(bytes);
        bufs.push_back(b);
        return b->p;
    }
};
}  // namespace

#define TMP(var, T, count)                                              \
    T* var = static_cast<T*>(tmp.get(sizeof(T) * (size_t)(count), &e)); \
    if (e != hipSuccess) return rhip(c, e, fn);

static const char* stage_check(lfh_ctx* c, int64_t n) {
    if (!c) return "NULL ctx";
    if (n < 0 || n > (1LL << 30)) return "bad query count";
    return nullptr;
}

int lfh_eval_bsdf(lfh_ctx* c, const float* frames, const float* L, int64_t n, float* f, float* pdf) {
    const char* fn = "lfh_eval_bsdf";
    if (const char* m = stage_check(c, n)) return rfail(c, LFH_E_INVALID, std::string(fn) + ": " + m);
    if (n == 0) return LFH_OK;
    for (int64_t i = 0; i < n; ++i) {
        int mat = (int)frames[i * 16 + 13];
        if (mat < 0 || mat >= c->ps.num_materials) return rfail(c, LFH_E_INVALID, "lfh_eval_bsdf: material id out of range");
    }
    Tmp tmp;
    hipError_t e;
    TMP(dfr, float, n * 16) TMP(dl, float, n * 3) TMP(df, float, n * 3) TMP(dp, float, n)
    RCHK(c, hipMemcpyAsync(dfr, frames, n * 64, hipMemcpyHostToDevice, c->stream), fn);
    RCHK(c, hipMemcpyAsync(dl, L, n * 12, hipMemcpyHostToDevice, c->stream), fn);
    RCHK(c, stage_eval_bsdf(lfh::shade_view(c->ps, c->ds, c->opt), dfr, dl, (int)n, df, dp, c->stream), fn);
    RCHK(c, hipMemcpyAsync(f, df, n * 12, hipMemcpyDeviceToHost, c->stream), fn);
    RCHK(c, hipMemcpyAsync(pdf, dp, n * 4, hipMemcpyDeviceToHost, c->stream), fn);
    RCHK(c, hipStreamSynchronize(c->stream), fn);
    return LFH_OK;
}

int lfh_sample_bsdf(lfh_ctx* c, const float* frames, uint32_t* seeds, int64_t n, float* L, float* f, float* pdf) {
    const char* fn = "lfh_sample_bsdf";
    if (const char* m = stage_check(c, n)) return rfail(c, LFH_E_INVALID, std::string(fn) + ": " + m);
    if (n == 0) return LFH_OK;
    for (int64_t i = 0; i < n; ++i) {
        int mat = (int)frames[i * 16 + 13];
        if (mat < 0 || mat >= c->ps.num_materials) return rfail(c, LFH_E_INVALID, "lfh_sample_bsdf: material id out of range");
    }
    Tmp tmp;
    hipError_t e;
    TMP(dfr, float, n * 16) TMP(ds, uint4, n) TMP(dl, float, n * 3) TMP(df, float, n * 3) TMP(dp, float, n)
    RCHK(c, hipMemcpyAsync(dfr, frames, n * 64, hipMemcpyHostToDevice, c->stream), fn);
    RCHK(c, hipMemcpyAsync(ds, seeds, n * 16, hipMemcpyHostToDevice, c->stream), fn);
    RCHK(c, stage_sample_bsdf(lfh::shade_view(c->ps, c->ds, c->opt), dfr, ds, (int)n, dl, df, dp, c->stream), fn);
    RCHK(c, hipMemcpyAsync(seeds, ds, n * 16, hipMemcpyDeviceToHost, c->stream), fn);
    RCHK(c, hipMemcpyAsync(L, dl, n * 12, hipMemcpyDeviceToHost, c->stream), fn);
    RCHK(c, hipMemcpyAsync(f, df, n * 12, hipMemcpyDeviceToHost, c->stream), fn);
    RCHK(c, hipMemcpyAsync(pdf, dp, n * 4, hipMemcpyDeviceToHost, c->stream), fn);
    RCHK(c, hipStreamSynchronize(c->stream), fn);
    return LFH_OK;
}

int lfh_sample_light(lfh_ctx* c, const float* pos, uint32_t* seeds, int64_t n, float* dir, float* normal,
                     float* emission, float* dist, float* pdf) {
    const char* fn = "lfh_sample_light";
    if (const char* m = stage_check(c, n)) return rfail(c, LFH_E_INVALID, std::string(fn) + ": " + m);
    if (c->ps.num_lights <= 0) return rfail(c, LFH_E_STATE, "lfh_sample_light: scene has no lights");
    if (n == 0) return LFH_OK;
    Tmp tmp;
    hipError_t e;
    TMP(dpos, float, n * 3) TMP(ds, uint4, n) TMP(dd, float, n * 3) TMP(dn, float, n * 3) TMP(de, float, n * 3)
    TMP(dist_d, float, n) TMP(dp, float, n)
    RCHK(c, hipMemcpyAsync(dpos, pos, n * 12, hipMemcpyHostToDevice, c->stream), fn);
    RCHK(c, hipMemcpyAsync(ds, seeds, n * 16, hipMemcpyHostToDevice, c->stream), fn);
    RCHK(c, stage_sample_light(lfh::shade_view(c->ps, c->ds, c->opt), dpos, ds, (int)n, dd, dn, de, dist_d, dp, c->stream),
         fn);
    RCHK(c, hipMemcpyAsync(seeds, ds, n * 16, hipMemcpyDeviceToHost, c->stream), fn);
    RCHK(c, hipMemcpyAsync(dir, dd, n * 12, hipMemcpyDeviceToHost, c->stream), fn);
    RCHK(c, hipMemcpyAsync(normal, dn, n * 12, hipMemcpyDeviceToHost, c->stream), fn);
    RCHK(c, hipMemcpyAsync(emission, de, n * 12, hipMemcpyDeviceToHost, c->stream), fn);
    RCHK(c, hipMemcpyAsync(dist, dist_d, n * 4, hipMemcpyDeviceToHost, c->stream), fn);
    RCHK(c, hipMemcpyAsync(pdf, dp, n * 4, hipMemcpyDeviceToHost, c->stream), fn);
    RCHK(c, hipStreamSynchronize(c->stream), fn);
    return LFH_OK;
}

int lfh_sample_env(lfh_ctx* c, uint32_t* seeds, int64_t n, float* dir, float* pdf, float* color) {
    const char* fn = "lfh_sample_env";
    if (const char* m = stage_check(c, n)) return rfail(c, LFH_E_INVALID, std::string(fn) + ": " + m);
    if (c->ps.hdr_w <= 0) return rfail(c, LFH_E_STATE, "lfh_sample_env: scene has no environment map");
    if (n == 0) return LFH_OK;
    Tmp tmp;
    hipError_t e;
    TMP(ds, uint4, n) TMP(dd, float, n * 3) TMP(dp, float, n) TMP(dc, float, n * 3)
    RCHK(c, hipMemcpyAsync(ds, seeds, n * 16, hipMemcpyHostToDevice, c->stream), fn);
    RCHK(c, stage_sample_env(lfh::shade_view(c->ps, c->ds, c->opt), ds, (int)n, dd, dp, dc, c->stream), fn);
    RCHK(c, hipMemcpyAsync(seeds, ds, n * 16, hipMemcpyDeviceToHost, c->stream), fn);
    RCHK(c, hipMemcpyAsync(dir, dd, n * 12, hipMemcpyDeviceToHost, c->stream), fn);
    RCHK(c, hipMemcpyAsync(pdf, dp, n * 4, hipMemcpyDeviceToHost, c->stream), fn);
    RCHK(c, hipMemcpyAsync(color, dc, n * 12, hipMemcpyDeviceToHost, c->stream), fn);
    RCHK(c, hipStreamSynchronize(c->stream), fn);
    return LFH_OK;
}

int lfh_env_pdf(lfh_ctx* c, const float* dir, int64_t n, float* pdf) {
    const char* fn = "lfh_env_pdf";
    if (const char* m = stage_check(c, n)) return rfail(c, LFH_E_INVALID, std::string(fn) + ": " + m);
    if (c->ps.hdr_w <= 0) return rfail(c, LFH_E_STATE, "lfh_env_pdf: scene has no environment map");
    if (n == 0) return LFH_OK;
    Tmp tmp;
    hipError_t e;
    TMP(dd, float, n * 3) TMP(dp, float, n)
    RCHK(c, hipMemcpyAsync(dd, dir, n * 12, hipMemcpyHostToDevice, c->stream), fn);
    RCHK(c, stage_env_pdf(lfh::shade_view(c->ps, c->ds, c->opt), dd, (int)n, dp, c->stream), fn);
    RCHK(c, hipMemcpyAsync(pdf, dp, n * 4, hipMemcpyDeviceToHost, c->stream), fn);
    RCHK(c, hipStreamSynchronize(c->stream), fn);
    return LFH_OK;
}

int lfh_path_trace_rays(lfh_ctx* c, const float* rays6, uint32_t* seeds, int64_t n, float* radiance) {
    const char* fn = "lfh_path_trace_rays";
    if (const char* m = stage_check(c, n)) return rfail(c, LFH_E_INVALID, std::string(fn) + ": " + m);
    if (n == 0) return LFH_OK;
    lfh::PassBufs pb;
    RCHK(c, pb.alloc((int)n), fn);
    Tmp tmp;
    hipError_t e;
    TMP(dr, float, n * 6) TMP(ds, uint4, n) TMP(dout, float, n * 3)
    RCHK(c, hipMemcpyAsync(dr, rays6, n * 24, hipMemcpyHostToDevice, c->stream), fn);
    RCHK(c, hipMemcpyAsync(ds, seeds, n * 16, hipMemcpyHostToDevice, c->stream), fn);
    uint32_t* cnt = pb.counters.as<uint32_t>();
    unsigned long long* work = pb.work.as<unsigned long long>();
    RCHK(c, pt_reset(cnt, work, c->stream), fn);
    RCHK(c, pt_generate_rays(dr, ds, (int)n, lfh::view(pb), pb.rq0.as<float4>(), cnt, c->stream), fn);
    lfhd::ShadeScene sc = lfh::shade_view(c->ps, c->ds, c->opt);
    GuideHook gh{};
    for (int depth = 0; depth < c->opt.max_depth; ++depth) {
        const int cur = depth & 1;
        float4* q[2] = {pb.rq0.as<float4>(), pb.rq1.as<float4>()};
        RCHK(c,
             pt_bounce(sc, depth, pb.cap, q[cur], cnt + cur, q[cur ^ 1], cnt + (cur ^ 1), pb.hq.as<float4>(),
                       pb.sq.as<float4>(), cnt + 2, lfh::view(pb), c->spill.as<uint32_t>(), work, c->rflag.as<int>(),
                       c->cull, c->pt_grid, gh, nullptr, c->stream),
             fn);
    }
    RCHK(c, pt_collect(lfh::view(pb), (int)n, dout, c->stream), fn);
    RCHK(c, hipMemcpyAsync(radiance, dout, n * 12, hipMemcpyDeviceToHost, c->stream), fn);
    RCHK(c, hipStreamSynchronize(c->stream), fn);
    return lfh::check_render_flag(c);
}

}  // extern "C"
