// Render service: continuous batching of path-tracing jobs on one GPU.
//
// No reference counterpart (the reference renders one interactive scene).
// Jobs arrive with their scene arrays, options, camera and a target sample
// count.  Each lfh_server_step():
//   1. admits queued jobs (priority, then arrival order; a job that does not
//      fit is skipped so smaller ones can backfill) while the device budget
//      and the job slots allow;
//   2. runs ONE wavefront for the running jobs together, each job taking
//      some of its passes (earliest deadline first, within a path budget):
//      the rays of every job go through the same extend / shade / shadow
//      queues and the same persistent kernels (the per-path job id selects
//      scene, camera and accumulation buffers);
//   3. retires jobs that reached their target.
// Device scenes are shared by content hash and stay resident after their
// jobs finish, until the budget needs the space (least recently used first).
// A job's result is bit-identical whether it runs alone or in a mix: its
// paths depend only on its own scene, camera, pixel and frame numbers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <map>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "capi_common.h"
#include "ctx.h"
#include "kernels.h"
#include "pt_kernels.h"
#include "render.h"

namespace lfh {

namespace {

// Device bytes of one pass's path state + queues per path (PassBufs).
constexpr int64_t kPathBytes = 16 * (1 + 1 + 1 + 1 + 1 + 3 + 2 + 2 + 2 + 4) + 4;
constexpr int64_t kAccumBytes = 32;  // sum + compensation, float4 each
#ifndef LFH_STEP_PATHS_LOG2
#define LFH_STEP_PATHS_LOG2 24   // 60-job stream: 2^23 0.027 / 0.133 s p50 / p99, 2^24 0.026 / 0.116 s
#endif
constexpr int64_t kStepPaths = (int64_t)1 << LFH_STEP_PATHS_LOG2;   // paths per server step when the budget allows
constexpr int64_t kStepMaxPasses = 64;    // passes of one job per step, at most
// A job's deadline: submit time + min(samples / nominal rate, max offset).
// Small jobs are ordered by size (short ones overtake long ones); a long job is
// overtaken only by work submitted within the max offset after it, which keeps
// its latency near that of first-come-first-served.  On a 60-job Poisson
// stream: offset cap 0.1 s gave p50 0.045 s / p99 0.22 s; no cap 0.043 / 0.25;
// FIFO order 0.060 / 0.21.
#ifndef LFH_SERVE_RATE
#define LFH_SERVE_RATE 3.0e8
#endif
#ifndef LFH_SERVE_MAX_OFFSET
#define LFH_SERVE_MAX_OFFSET 0.1
#endif
constexpr double kDeadlineSecondsPerSample = 1.0 / LFH_SERVE_RATE;
constexpr double kDeadlineMaxOffset = LFH_SERVE_MAX_OFFSET;

uint64_t fnv(uint64_t h, const void* p, size_t n) {
    const unsigned char* c = static_cast<const unsigned char*>(p);
    for (size_t i = 0; i < n; ++i) {
        h ^= c[i];
        h *= 1099511628211ull;
    }
    return h;
}

uint64_t hash_desc(const lfh_scene_desc& d) {
    uint64_t h = 1469598103934665603ull;
    h = fnv(h, d.nodes, (size_t)d.num_nodes * 36);
    h = fnv(h, &d.top_level_index, 4);
    h = fnv(h, d.vert_indices, (size_t)d.num_tris * 12);
    h = fnv(h, d.vertices_uvx, (size_t)d.num_vertices * 16);
    if (d.normals_uvy) h = fnv(h, d.normals_uvy, (size_t)d.num_vertices * 16);
    h = fnv(h, d.materials, (size_t)d.num_materials * 112);
    h = fnv(h, d.transforms, (size_t)d.num_instances * 64);
    if (d.lights) h = fnv(h, d.lights, (size_t)d.num_lights * 60);
    if (d.hdr_width > 0) {
        size_t n = (size_t)d.hdr_width * d.hdr_height;
        h = fnv(h, d.hdr_cols, n * 12);
        if (d.hdr_marginal) h = fnv(h, d.hdr_marginal, (size_t)d.hdr_height * 8);
        if (d.hdr_conditional) h = fnv(h, d.hdr_conditional, n * 8);
    }
    if (d.textures && d.num_textures > 0)
        h = fnv(h, d.textures, (size_t)d.num_textures * d.tex_width * d.tex_height * 4);
    if (!d.nodes && d.num_meshes > 0) {  // device BVH build: the mesh and instance tables
        h = fnv(h, d.mesh_tri_offsets, (size_t)(d.num_meshes + 1) * 8);
        h = fnv(h, d.instance_meshes, (size_t)d.num_instances * 4);
        h = fnv(h, d.instance_materials, (size_t)d.num_instances * 4);
    }
    return h;
}

int64_t packed_bytes(const PackedScene& ps) {
    int64_t b = (int64_t)(ps.nodes.size() + ps.tris.size() + ps.insts.size() + ps.xforms.size() + ps.nmats.size() +
                          ps.normals.size() + ps.materials.size() + ps.lights.size() + ps.hdr_cols.size() +
                          ps.hdr_marg.size() + ps.hdr_cond.size() + ps.uvs.size()) * 4 +
                (int64_t)ps.vidx.size() * 4 + (int64_t)ps.tex.size();
    if (ps.hdr_tables_on_device) b += (int64_t)ps.hdr_w * ps.hdr_h * 8 + (int64_t)ps.hdr_h * 8;
    if (ps.device_build)  // resident after the build: triangle records, BVH4 nodes (at most one per 3 triangles + TLAS)
        b += ps.num_tris * 48 + (ps.num_tris / 3 + 2 * (int64_t)ps.num_instances + 2) * 128;
    return b;
}

struct SceneEntry {
    uint64_t hash = 0;
    PackedScene ps;
    DeviceScene ds;
    bool resident = false;
    int refs = 0;
    int64_t bytes = 0;
    double last_use = 0;
};

struct Job {
    int64_t id = 0;
    int state = 0;  // 0 queued, 1 running, 2 done, -1 rejected
    int spp_done = 0, target = 0, priority = 0;
    int64_t tag = 0;
    double t_submit = 0, t_admit = 0, t_done = 0;
    lfh_options opt{};
    lfh_camera cam{};
    std::shared_ptr<SceneEntry> scene;
    DevBuf accum, comp;
    int frame = 2;  // first pass uses frame 2, like TiledRenderer after Init + Update
    int64_t npix = 0;
    int64_t charged = 0;  // accum + path share held while running
    std::string error;    // why the job was rejected
};

}  // namespace
}  // namespace lfh

struct lfh_server {
    int64_t budget = 0;
    int max_jobs = 0;
    hipStream_t stream = nullptr;
    std::chrono::steady_clock::time_point t0;
    int64_t next_id = 1;
    std::map<int64_t, std::unique_ptr<lfh::Job>> jobs;
    std::map<uint64_t, std::shared_ptr<lfh::SceneEntry>> scenes;
    lfh::PassBufs pass;
    lfh::DevBuf table, spill, errflag;
    int grid = 0;
    int64_t used = 0;  // resident scenes + accum of admitted jobs + path shares of running jobs
    std::string err;
    double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
};

using lfh::DevBuf;

static int sfail(lfh_server* s, int code, const std::string& m) {
    if (s) s->err = m;
    lfh_internal::set_thread_error(m);
    return code;
}
static int shipfail(lfh_server* s, hipError_t e, const char* where) {
    return sfail(s, e == hipErrorOutOfMemory ? LFH_E_OOM : LFH_E_HIP, std::string(where) + ": " + hipGetErrorString(e));
}
#define SCHK(s, expr, where)                                  \
    do {                                                      \
        hipError_t _e = (expr);                               \
        if (_e != hipSuccess) return shipfail(s, _e, where);  \
    } while (0)

// Evict unreferenced resident scenes (oldest first) until `need` more bytes fit.
static void evict_for(lfh_server* s, int64_t need) {
    while (s->used + need > s->budget) {
        std::shared_ptr<lfh::SceneEntry> victim;
        for (auto& kv : s->scenes) {
            auto& e = kv.second;
            if (e->resident && e->refs == 0 && (!victim || e->last_use < victim->last_use)) victim = e;
        }
        if (!victim) return;
        victim->ds.release();
        victim->resident = false;
        s->used -= victim->bytes;
    }
}

extern "C" {

int lfh_server_create(int64_t budget_bytes, int32_t max_jobs, lfh_server** out) {
    if (!out || max_jobs <= 0 || max_jobs > 64) return sfail(nullptr, LFH_E_INVALID, "lfh_server_create: bad arguments");
    *out = nullptr;
    lfh_server* s = new (std::nothrow) lfh_server;
    if (!s) return sfail(nullptr, LFH_E_OOM, "lfh_server_create: out of host memory");
    s->t0 = std::chrono::steady_clock::now();
    s->max_jobs = max_jobs;
    if (budget_bytes <= 0) {
        size_t f = 0, t = 0;
        if (hipMemGetInfo(&f, &t) != hipSuccess) {
            delete s;
            return sfail(nullptr, LFH_E_HIP, "lfh_server_create: hipMemGetInfo failed");
        }
        budget_bytes = (int64_t)(0.9 * (double)f);
    }
    s->budget = budget_bytes;
    hipError_t e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
    if (e == hipSuccess) {
        s->grid = pt_trace_grid_blocks();
        e = s->spill.alloc((size_t)s->grid * lfh_trace_threads_per_block() * lfh_trace_spill_stride() * 4);
    }
    if (e == hipSuccess) e = s->errflag.alloc(64);
    if (e == hipSuccess) e = hipMemsetAsync(s->errflag.p, 0, 64, s->stream);
    if (e == hipSuccess) e = s->table.alloc(sizeof(lfhd::JobDev) * 64);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) {
        int r = shipfail(nullptr, e, "lfh_server_create");
        lfh_server_destroy(s);
        return r;
    }
    *out = s;
    return LFH_OK;
}

void lfh_server_destroy(lfh_server* s) {
    if (!s) return;
    if (s->stream) hipStreamSynchronize(s->stream);
    s->jobs.clear();
    s->scenes.clear();
    s->pass.release();
    s->table.release();
    s->spill.release();
    s->errflag.release();
    if (s->stream) hipStreamDestroy(s->stream);
    delete s;
}

double lfh_server_clock(const lfh_server* s) { return s ? s->now() : 0.0; }

const char* lfh_server_job_error(lfh_server* s, int64_t id) {
    if (!s) return "";
    auto it = s->jobs.find(id);
    return it == s->jobs.end() ? "unknown job" : it->second->error.c_str();
}

int lfh_server_submit(lfh_server* s, const lfh_job_desc* jd, int64_t* job_id) {
    if (!s || !jd || !jd->scene) return sfail(s, LFH_E_INVALID, "lfh_server_submit: NULL argument");
    const lfh_options& o = jd->options;
    if (o.width <= 0 || o.height <= 0 || o.max_depth < 0 || jd->target_spp <= 0)
        return sfail(s, LFH_E_INVALID, "lfh_server_submit: bad resolution, depth or target spp");
    if (o.guiding) return sfail(s, LFH_E_UNSUPPORTED, "lfh_server_submit: guided jobs are not served");
    auto job = std::make_unique<lfh::Job>();
    job->id = s->next_id++;
    job->opt = o;
    job->cam = jd->camera;
    job->target = jd->target_spp;
    job->priority = jd->priority;
    job->tag = jd->user_tag;
    job->t_submit = s->now();
    job->npix = (int64_t)o.width * o.height;
    uint64_t h = lfh::hash_desc(*jd->scene);
    auto it = s->scenes.find(h);
    std::shared_ptr<lfh::SceneEntry> e;
    if (it != s->scenes.end()) {
        e = it->second;
    } else {
        e = std::make_shared<lfh::SceneEntry>();
        e->hash = h;
        std::string perr = lfh::pack_scene(*jd->scene, &e->ps);
        if (!perr.empty()) {
            job->state = -1;
            job->error = perr;
            if (job_id) *job_id = job->id;
            s->jobs[job->id] = std::move(job);
            return sfail(s, LFH_E_INVALID, "lfh_server_submit: " + perr);
        }
        e->bytes = lfh::packed_bytes(e->ps);
    }
    int64_t need = job->npix * (lfh::kAccumBytes + lfh::kPathBytes) + e->bytes;
    if (job_id) *job_id = job->id;
    if (need > s->budget) {
        job->state = -1;
        job->error = "job exceeds the device budget";
        s->jobs[job->id] = std::move(job);
        return sfail(s, LFH_E_BUDGET, "lfh_server_submit: job needs " + std::to_string(need) +
                                          " device bytes, budget is " + std::to_string(s->budget));
    }
    s->scenes[h] = e;
    job->scene = e;
    e->refs++;
    s->jobs[job->id] = std::move(job);
    return LFH_OK;
}

// Rejects a queued job at admission: it leaves the queue with its reason and
// gives up its scene reference.
static void reject(lfh_server* s, lfh::Job* j, const std::string& why) {
    j->state = -1;
    j->error = why;
    j->t_done = s->now();
    if (j->scene) {
        j->scene->refs--;
        j->scene.reset();
    }
}

int lfh_server_step(lfh_server* s, int32_t* active) {
    if (!s) return sfail(nullptr, LFH_E_INVALID, "lfh_server_step: NULL server");
    // -- admission: priority, then arrival; jobs that do not fit are skipped
    std::vector<lfh::Job*> queued, running;
    for (auto& kv : s->jobs) {
        if (kv.second->state == 0) queued.push_back(kv.second.get());
        if (kv.second->state == 1) running.push_back(kv.second.get());
    }
    std::stable_sort(queued.begin(), queued.end(),
                     [](const lfh::Job* a, const lfh::Job* b) { return a->priority > b->priority; });
    for (lfh::Job* j : queued) {
        if ((int)running.size() >= s->max_jobs) break;
        lfh::SceneEntry& e = *j->scene;
        int64_t need = j->npix * (lfh::kAccumBytes + lfh::kPathBytes) + (e.resident ? 0 : e.bytes);
        if (s->used + need > s->budget) evict_for(s, need);
        if (s->used + need > s->budget) continue;
        // A job whose scene cannot be made resident, or whose buffers cannot
        // be allocated, is rejected with its reason; the other jobs go on.
        if (!e.resident) {
            std::string berr;
            hipError_t ue = lfh::upload_scene(e.ps, &e.ds, s->stream, &berr);
            if (ue == hipSuccess && berr.empty() && lfh_internal::inject_fault(lfh_internal::kFaultSceneUpload))
                ue = hipErrorOutOfMemory;
            if (!berr.empty() || ue != hipSuccess) {
                (void)hipStreamSynchronize(s->stream);
                e.ds.release();
                reject(s, j, !berr.empty() ? "device BVH build: " + berr
                                           : std::string("scene upload: ") + hipGetErrorString(ue));
                continue;
            }
            e.resident = true;
            s->used += e.bytes;
        }
        hipError_t ae = j->accum.alloc((size_t)j->npix * 16);
        if (ae == hipSuccess) ae = j->comp.alloc((size_t)j->npix * 16);
        if (ae == hipSuccess) ae = pt_clear4(j->accum.as<float4>(), j->npix, s->stream);
        if (ae == hipSuccess) ae = pt_clear4(j->comp.as<float4>(), j->npix, s->stream);
        if (ae != hipSuccess) {
            j->accum.release();
            j->comp.release();
            reject(s, j, std::string("accumulation buffers: ") + hipGetErrorString(ae));
            continue;
        }
        j->charged = j->npix * (lfh::kAccumBytes + lfh::kPathBytes);
        s->used += j->charged;
        j->state = 1;
        j->t_admit = s->now();
        running.push_back(j);
    }
    // -- passes for this step, rays of all chosen jobs packed into shared
    //    queues.  Jobs are served earliest deadline first, a job's deadline
    //    being its submit time plus its samples at a nominal rate, at most
    //    kDeadlineMaxOffset: a small job finishes soon after it arrives, and a
    //    large job is not overtaken by work that arrives after its deadline.  In that order each job takes as
    //    many of its remaining passes as the step's path budget still holds (at
    //    most kStepMaxPasses); the first always gets one.  Passes of a job are
    //    accumulated per pixel in frame order, so its image is that of single
    //    passes whatever the grouping.
    if (!running.empty()) {
        int64_t one_pass = 0;
        for (lfh::Job* j : running) one_pass += j->npix;
        const int64_t spare = std::max<int64_t>(0, (s->budget - s->used) / lfh::kPathBytes);
        auto deadline = [](const lfh::Job* j) {
            return j->t_submit + std::min(lfh::kDeadlineMaxOffset,
                                          (double)j->npix * (double)j->target * lfh::kDeadlineSecondsPerSample);
        };
        std::stable_sort(running.begin(), running.end(), [&](const lfh::Job* a, const lfh::Job* b) {
            const double da = deadline(a), db = deadline(b);
            return da < db || (da == db && a->id < b->id);
        });
        int64_t cap = std::min(lfh::kStepPaths, one_pass + spare);
        cap = std::max(cap, running[0]->npix);
        std::vector<lfh::Job*> chosen;
        std::vector<int> passes;
        int64_t total = 0;
        for (lfh::Job* j : running) {
            const int64_t left = (int64_t)j->target - j->spp_done;
            const int64_t give = std::min<int64_t>(std::min<int64_t>(left, lfh::kStepMaxPasses), (cap - total) / j->npix);
            if (give <= 0) continue;
            chosen.push_back(j);
            passes.push_back((int)give);
            total += give * j->npix;
        }
        int max_depth = 0;
        std::vector<lfhd::JobDev> tab(chosen.size());
        int64_t off = 0;
        for (size_t k = 0; k < chosen.size(); ++k) {
            lfh::Job* j = chosen[k];
            lfhd::JobDev& d = tab[k];
            std::memset(&d, 0, sizeof d);
            d.sc = lfh::shade_view(j->scene->ps, j->scene->ds, j->opt);
            lfhd::CamParams& c = d.cam;
            std::memcpy(c.pos, j->cam.position, 12);
            std::memcpy(c.right, j->cam.right, 12);
            std::memcpy(c.up, j->cam.up, 12);
            std::memcpy(c.fwd, j->cam.forward, 12);
            c.fov = j->cam.fov;
            c.focal_dist = j->cam.focal_dist;
            c.aperture = j->cam.aperture;
            c.width = j->opt.width;
            c.height = j->opt.height;
            c.tile_w = j->opt.width;
            c.tile_h = j->opt.height;
            c.tileX = 0;
            c.tileY = 0;
            c.frame = j->frame;
            c.batch = passes[k];
            d.accum = j->accum.as<float4>();
            d.comp = j->comp.as<float4>();
            d.path_offset = (int)off;
            d.kahan = j->opt.kahan;
            off += j->npix * passes[k];
            max_depth = std::max(max_depth, j->opt.max_depth);
            j->scene->last_use = s->now();
        }
        if (total > (1LL << 30)) return sfail(s, LFH_E_UNSUPPORTED, "lfh_server_step: too many pixels in flight");
        if (s->pass.cap < total || s->pass.cap > 4 * total + (1 << 23)) {
            // grow with headroom, shrink only when far too large (a reallocation waits for the GPU)
            const int64_t cap = std::min<int64_t>(total + total / 4, std::max<int64_t>(total, lfh::kStepPaths));
            SCHK(s, hipStreamSynchronize(s->stream), "lfh_server_step: sync");
            SCHK(s, s->pass.alloc((int)cap), "lfh_server_step: path buffers");
        }
        SCHK(s, hipMemcpyAsync(s->table.p, tab.data(), tab.size() * sizeof(lfhd::JobDev), hipMemcpyHostToDevice,
                               s->stream),
             "lfh_server_step: job table");
        const lfhd::JobDev* jt = s->table.as<lfhd::JobDev>();
        lfh::PassBufs& pb = s->pass;
        PathBufs v = lfh::view(pb);
        uint32_t* cnt = pb.counters.as<uint32_t>();
        unsigned long long* work = pb.work.as<unsigned long long>();
        SCHK(s, pt_reset(cnt, work, s->stream), "lfh_server_step: reset");
        SCHK(s, pt_generate_multi(jt, (int)tab.size(), (int)total, v, pb.rq0.as<float4>(), cnt, s->stream),
             "lfh_server_step: generate");
        for (int depth = 0; depth < max_depth; ++depth) {
            const int cur = depth & 1;
            float4* q[2] = {pb.rq0.as<float4>(), pb.rq1.as<float4>()};
            SCHK(s,
                 pt_bounce_multi(jt, depth, pb.cap, q[cur], cnt + cur, q[cur ^ 1], cnt + (cur ^ 1), pb.hq.as<float4>(),
                                 pb.sq.as<float4>(), cnt + 2, v, s->spill.as<uint32_t>(), work, s->errflag.as<int>(),
                                 1.0f + 1.0f / 1024.0f, s->grid, s->stream),
                 "lfh_server_step: bounce");
        }
        SCHK(s, pt_accumulate_multi(jt, (int)total, v, s->stream), "lfh_server_step: accumulate");
        SCHK(s, hipStreamSynchronize(s->stream), "lfh_server_step: sync");
        int flag = 0;
        SCHK(s, hipMemcpy(&flag, s->errflag.p, sizeof(int), hipMemcpyDeviceToHost), "lfh_server_step: flag");
        if (flag) {
            SCHK(s, pt_clear_flag(s->errflag.as<int>(), s->stream), "lfh_server_step: flag");
            SCHK(s, hipStreamSynchronize(s->stream), "lfh_server_step: flag");
            return sfail(s, LFH_E_UNSUPPORTED, "lfh_server_step: traversal stack overflow (step not counted)");
        }
        double t = s->now();
        for (size_t k = 0; k < chosen.size(); ++k) {
            lfh::Job* j = chosen[k];
            j->spp_done += passes[k];
            j->frame += passes[k];
            if (j->spp_done >= j->target) {
                j->state = 2;
                j->t_done = t;
                s->used -= j->npix * lfh::kPathBytes;  // path share back; accum held until release
                j->charged = j->npix * lfh::kAccumBytes;
            }
        }
    }
    int32_t n = 0;
    for (auto& kv : s->jobs)
        if (kv.second->state == 0 || kv.second->state == 1) ++n;
    if (active) *active = n;
    return LFH_OK;
}

int lfh_server_status(lfh_server* s, int64_t id, lfh_job_status* st) {
    if (!s || !st) return sfail(s, LFH_E_INVALID, "lfh_server_status: NULL argument");
    auto it = s->jobs.find(id);
    if (it == s->jobs.end()) return sfail(s, LFH_E_INVALID, "lfh_server_status: unknown job");
    const lfh::Job& j = *it->second;
    st->job_id = j.id;
    st->state = j.state;
    st->spp_done = j.spp_done;
    st->t_submit = j.t_submit;
    st->t_admit = j.t_admit;
    st->t_done = j.t_done;
    st->device_bytes = j.charged;
    return LFH_OK;
}

int lfh_server_result(lfh_server* s, int64_t id, float* dst, int32_t w, int32_t h) {
    if (!s || !dst) return sfail(s, LFH_E_INVALID, "lfh_server_result: NULL argument");
    auto it = s->jobs.find(id);
    if (it == s->jobs.end()) return sfail(s, LFH_E_INVALID, "lfh_server_result: unknown job");
    lfh::Job& j = *it->second;
    if (j.state != 2 && j.state != 1) return sfail(s, LFH_E_STATE, "lfh_server_result: job has no samples");
    if (w != j.opt.width || h != j.opt.height) return sfail(s, LFH_E_INVALID, "lfh_server_result: size mismatch");
    DevBuf out;
    SCHK(s, out.alloc((size_t)j.npix * 12), "lfh_server_result: alloc");
    SCHK(s, pt_output(j.accum.as<float4>(), j.comp.as<float4>(), j.npix, 1.0f / (float)std::max(1, j.spp_done),
                      out.as<float>(), s->stream),
         "lfh_server_result: kernel");
    SCHK(s, hipMemcpyAsync(dst, out.p, (size_t)j.npix * 12, hipMemcpyDeviceToHost, s->stream), "lfh_server_result: copy");
    SCHK(s, hipStreamSynchronize(s->stream), "lfh_server_result: sync");
    return LFH_OK;
}

int lfh_server_release(lfh_server* s, int64_t id) {
    if (!s) return sfail(nullptr, LFH_E_INVALID, "lfh_server_release: NULL server");
    auto it = s->jobs.find(id);
    if (it == s->jobs.end()) return sfail(s, LFH_E_INVALID, "lfh_server_release: unknown job");
    lfh::Job& j = *it->second;
    if (j.state == 1) return sfail(s, LFH_E_STATE, "lfh_server_release: job still running");
    SCHK(s, hipStreamSynchronize(s->stream), "lfh_server_release: sync");
    if (j.state == 2) s->used -= j.charged;
    if (j.scene) j.scene->refs--;
    s->jobs.erase(it);
    return LFH_OK;
}

}  // extern "C"
