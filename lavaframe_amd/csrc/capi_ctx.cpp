// C ABI: device context lifetime, scene upload, direct traversal.
#include <hip/hip_runtime.h>

#include <cstring>
#include <new>
#include <string>

#include "capi_common.h"
#include "ctx.h"
#include "kernels.h"
#include "render.h"

using lfh::DevBuf;

static int ctx_fail(lfh_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    lfh_internal::set_thread_error(msg);
    return code;
}

static int hip_fail(lfh_ctx* c, hipError_t e, const char* where) {
    std::string m = std::string(where) + ": " + hipGetErrorString(e);
    return ctx_fail(c, e == hipErrorOutOfMemory ? LFH_E_OOM : LFH_E_HIP, m);
}

#define HIPCHK(c, expr, where)                          \
    do {                                                \
        hipError_t _e = (expr);                         \
        if (_e != hipSuccess) return hip_fail(c, _e, where); \
    } while (0)

extern "C" {

const char* lfh_last_error(const lfh_ctx* ctx) {
    if (ctx) return ctx->err.c_str();
    return lfh_internal::thread_error();
}

int lfh_device_memory(int64_t* free_bytes, int64_t* total_bytes) {
    size_t f = 0, t = 0;
    hipError_t e = hipMemGetInfo(&f, &t);
    if (e != hipSuccess) return hip_fail(nullptr, e, "lfh_device_memory");
    if (free_bytes) *free_bytes = (int64_t)f;
    if (total_bytes) *total_bytes = (int64_t)t;
    return LFH_OK;
}

int lfh_create(const lfh_scene_desc* desc, const lfh_options* opt, lfh_ctx** out) {
    if (!desc || !opt || !out) return ctx_fail(nullptr, LFH_E_INVALID, "lfh_create: NULL argument");
    *out = nullptr;
    if (opt->width <= 0 || opt->height <= 0 || opt->max_depth < 0)
        return ctx_fail(nullptr, LFH_E_INVALID, "lfh_create: bad resolution or depth");
    if (opt->estimator != 0 && opt->estimator != 1)
        return ctx_fail(nullptr, LFH_E_INVALID, "lfh_create: estimator must be 0 or 1");
    if (opt->frame_offset < 0) return ctx_fail(nullptr, LFH_E_INVALID, "lfh_create: negative frame offset");
    lfh_ctx* c = new (std::nothrow) lfh_ctx;
    if (!c) return ctx_fail(nullptr, LFH_E_OOM, "lfh_create: out of host memory");
    c->opt = *opt;
    c->frameOffset = opt->frame_offset;
    std::string err = lfh::pack_scene(*desc, &c->ps);
    if (!err.empty()) {
        delete c;
        return ctx_fail(nullptr, LFH_E_INVALID, "lfh_create: " + err);
    }
    hipError_t e = hipGetDevice(&c->device);
    if (e == hipSuccess) e = lfh::noalloc_setup(c->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        int r = hip_fail(nullptr, e, "lfh_create: stream");
        delete c;
        return r;
    }
    c->stream = c->own_stream;
    std::string berr;
    e = lfh::upload_scene(c->ps, &c->ds, c->stream, &berr);
    if (!berr.empty()) {
        lfh_destroy(c);
        return ctx_fail(nullptr, LFH_E_INVALID, "lfh_create: device BVH build: " + berr);
    }
    if (e == hipSuccess) {
        c->grid_blocks = lfh_trace_grid_blocks();
        size_t lanes = (size_t)c->grid_blocks * lfh_trace_threads_per_block();
        e = c->spill.alloc(lanes * lfh_trace_spill_stride() * sizeof(uint32_t));
        if (e == hipSuccess) e = c->work.alloc(lfhd::kFetchWords * 8);
        if (e == hipSuccess) e = c->errflag.alloc(64);
        if (e == hipSuccess) e = hipMemsetAsync(c->errflag.p, 0, 64, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    }
    if (e != hipSuccess) {
        int r = hip_fail(nullptr, e, "lfh_create: upload");
        lfh_destroy(c);
        return r;
    }
    int r = lfh::render_init(c);
    if (r != LFH_OK) {
        std::string m = c->err;
        lfh_destroy(c);
        return ctx_fail(nullptr, r, m);
    }
    *out = c;
    return LFH_OK;
}

void lfh_destroy(lfh_ctx* c) {
    if (!c) return;
    if (c->stream) hipStreamSynchronize(c->stream);
    lfh::render_release(c);
    if (c->own_stream) hipStreamDestroy(c->own_stream);
    delete c;
}

int lfh_set_stream(lfh_ctx* c, void* s) {
    if (!c) return ctx_fail(nullptr, LFH_E_INVALID, "lfh_set_stream: NULL ctx");
    HIPCHK(c, hipStreamSynchronize(c->stream), "lfh_set_stream");
    c->stream = s ? static_cast<hipStream_t>(s) : c->own_stream;
    return LFH_OK;
}

int lfh_synchronize(lfh_ctx* c) {
    if (!c) return ctx_fail(nullptr, LFH_E_INVALID, "lfh_synchronize: NULL ctx");
    HIPCHK(c, hipStreamSynchronize(c->stream), "lfh_synchronize");
    return lfh::check_render_flag(c);
}

int lfh_update_instances(lfh_ctx* c, const float* transforms, const float* materials, const float* tlas_nodes,
                         int64_t first_node, int64_t count) {
    if (!c) return ctx_fail(nullptr, LFH_E_INVALID, "lfh_update_instances: NULL ctx");
    if (c->ps.device_build && tlas_nodes)
        return ctx_fail(c, LFH_E_INVALID, "lfh_update_instances: a device-built scene takes no TLAS nodes");
    // A rejected update leaves the scene as it was: the host fields are put
    // back from this snapshot and the device arrays are not touched (the
    // device TLAS build restores its own node range on failure).
    lfh::InstanceSnapshot snap;
    snap.save(c->ps);
    auto reject = [&](int code, const std::string& msg) {
        snap.restore(&c->ps);
        return ctx_fail(c, code, "lfh_update_instances: " + msg);
    };
    HIPCHK(c, hipStreamSynchronize(c->stream), "lfh_update_instances");
    auto cp = [&](DevBuf& b, const std::vector<float>& v) {
        return hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(float), hipMemcpyHostToDevice, c->stream);
    };
    if (c->ps.device_build) {
        // no reference TLAS: the new transforms' TLAS is rebuilt on the device
        std::string err = lfh::repack_instances(&c->ps, transforms, nullptr, 0, 0, materials);
        if (!err.empty()) return reject(LFH_E_INVALID, err);
        err = lfh::device_build_tlas(&c->ps, &c->ds, c->stream);
        if (!err.empty()) return reject(LFH_E_INVALID, err);
    } else {
        std::string err = lfh::repack_instances(&c->ps, transforms, tlas_nodes, first_node, count, materials);
        if (!err.empty()) return reject(LFH_E_INVALID, err);
        HIPCHK(c, cp(c->ds.nodes, c->ps.nodes4), "lfh_update_instances: nodes");
        HIPCHK(c, cp(c->ds.insts, c->ps.insts4), "lfh_update_instances: instances");
    }
    HIPCHK(c, cp(c->ds.xforms, c->ps.xforms), "lfh_update_instances: transforms");
    HIPCHK(c, cp(c->ds.nmats, c->ps.nmats), "lfh_update_instances: normal matrices");
    HIPCHK(c, cp(c->ds.materials, c->ps.materials), "lfh_update_instances: materials");
    HIPCHK(c, hipStreamSynchronize(c->stream), "lfh_update_instances");
    c->scene_version++;
    return LFH_OK;
}

static int check_err_flag(lfh_ctx* c) {
    int flag = 0;
    HIPCHK(c, hipMemcpyAsync(&flag, c->errflag.p, sizeof(int), hipMemcpyDeviceToHost, c->stream), "trace: flag");
    HIPCHK(c, hipStreamSynchronize(c->stream), "trace: flag");
    if (flag & 1) {
        hipMemsetAsync(c->errflag.p, 0, 64, c->stream);
        return ctx_fail(c, LFH_E_UNSUPPORTED, "traversal stack overflow");
    }
    return LFH_OK;
}

static int trace_dev(lfh_ctx* c, const lfh_ray* d_rays, int64_t n, void* d_out, int32_t* d_counts, bool any) {
    if (!c || (n > 0 && (!d_rays || !d_out))) return ctx_fail(c, LFH_E_INVALID, "trace: NULL argument");
    if (n < 0) return ctx_fail(c, LFH_E_INVALID, "trace: negative ray count");
    if (n == 0) return LFH_OK;
    TraceLaunch L;
    L.scene = lfh::trace_view(c->ps, c->ds);
    L.rays = reinterpret_cast<const float4*>(d_rays);
    L.n = n;
    L.hits = any ? nullptr : static_cast<int32_t*>(d_out);
    L.occluded = any ? static_cast<uint8_t*>(d_out) : nullptr;
    L.counts = d_counts;
    L.spill = c->spill.as<uint32_t>();
    L.work = c->work.as<unsigned long long>();
    L.err = c->errflag.as<int>();
    L.cull = c->cull;
    L.any = any;
    L.grid_blocks = c->grid_blocks;
    HIPCHK(c, lfh_launch_trace(L, c->stream), "trace: launch");
    return LFH_OK;
}

int lfh_trace_closest_device(lfh_ctx* c, const lfh_ray* d_rays, int64_t n, lfh_hit* d_hits) {
    return trace_dev(c, d_rays, n, d_hits, nullptr, false);
}

int lfh_trace_any_device(lfh_ctx* c, const lfh_ray* d_rays, int64_t n, uint8_t* d_occ) {
    return trace_dev(c, d_rays, n, d_occ, nullptr, true);
}

int lfh_trace_closest_stats_device(lfh_ctx* c, const lfh_ray* d_rays, int64_t n, lfh_hit* d_hits,
                                   int32_t* d_counts) {
    if (!d_counts) return ctx_fail(c, LFH_E_INVALID, "trace stats: NULL counts");
    return trace_dev(c, d_rays, n, d_hits, d_counts, false);
}

static int trace_host(lfh_ctx* c, const lfh_ray* rays, int64_t n, void* out, size_t out_elem, bool any) {
    if (!c || (n > 0 && (!rays || !out))) return ctx_fail(c, LFH_E_INVALID, "trace: NULL argument");
    if (n <= 0) return n == 0 ? LFH_OK : ctx_fail(c, LFH_E_INVALID, "trace: negative ray count");
    size_t in_b = (size_t)n * sizeof(lfh_ray), out_b = (size_t)n * out_elem;
    if (c->stage_in.bytes < in_b) HIPCHK(c, c->stage_in.alloc(in_b), "trace: staging");
    if (c->stage_out.bytes < out_b) HIPCHK(c, c->stage_out.alloc(out_b), "trace: staging");
    HIPCHK(c, hipMemcpyAsync(c->stage_in.p, rays, in_b, hipMemcpyHostToDevice, c->stream), "trace: copy in");
    int r = trace_dev(c, c->stage_in.as<lfh_ray>(), n, c->stage_out.p, nullptr, any);
    if (r != LFH_OK) return r;
    HIPCHK(c, hipMemcpyAsync(out, c->stage_out.p, out_b, hipMemcpyDeviceToHost, c->stream), "trace: copy out");
    HIPCHK(c, hipStreamSynchronize(c->stream), "trace: sync");
    return check_err_flag(c);
}

int lfh_trace_closest(lfh_ctx* c, const lfh_ray* rays, int64_t n, lfh_hit* hits) {
    return trace_host(c, rays, n, hits, sizeof(lfh_hit), false);
}

int lfh_trace_any(lfh_ctx* c, const lfh_ray* rays, int64_t n, uint8_t* occ) {
    return trace_host(c, rays, n, occ, 1, true);
}

}  // extern "C"
