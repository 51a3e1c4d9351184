// Render service: continuous batching of path-tracing jobs on one GPU.
//
// No reference counterpart (the reference renders one interactive scene).
// Jobs arrive with their scene arrays, options, camera and a target sample
// count.  Each lfh_server_step():
//   1. admits queued jobs (priority, then arrival order; a job that does not
//      fit is skipped so smaller ones can backfill) while the device budget
//      and the job slots allow.  A scene that is not resident is uploaded
//      (and its BVHs built on the device when its desc has none) by an
//      admission thread on its own stream, so uploads overlap the wavefronts
//      the step loop keeps rendering; its jobs start once it is resident;
//   2. runs the running jobs' passes (earliest deadline first, within a path
//      budget) as wavefronts: one for the plain jobs together, and one per
//      scene with guided jobs, whose rays all go through the same extend /
//      shade / shadow queues and persistent kernels (the per-path job id
//      selects scene, camera and accumulation buffers);
//   3. retires jobs that reached their target.
// Device scenes are shared by content hash and stay resident after their
// jobs finish, until the budget needs the space (least recently used first).
// A guided job (options.guiding, consistent estimator) samples its scene's
// learned guide, which the guided wavefronts of that scene train online, one
// guide per resident scene whichever jobs render it.  A plain job's result is
// bit-identical whether it runs alone or in a mix: its paths depend only on
// its own scene, camera, pixel and frame numbers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <thread>
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
#ifndef LFH_SERVE_GUIDE_LR
// Adam learning rate of a served scene's guide.  Guiding room on one warm server, median
// guided time to RMSE 1e-2 over 32 random sequences (a second set of 32 in brackets):
// 1e-3 0.168 (0.162) s, 1.5e-3 0.156 (0.155), 2e-3 0.161 (0.159), 2.5e-3 0.163, 3e-3 0.164.
#define LFH_SERVE_GUIDE_LR 1.5e-3f
#endif
#ifndef LFH_STEP_PATHS_LOG2
#define LFH_STEP_PATHS_LOG2 24   // 60-job stream: 2^23 0.027 / 0.133 s p50 / p99, 2^24 0.026 / 0.116 s
#endif
constexpr int64_t kStepPaths = (int64_t)1 << LFH_STEP_PATHS_LOG2;   // paths per server step when the budget allows
constexpr int64_t kStepMaxPasses = 64;    // passes of one job per step, at most
// Step policy defaults (lfh_server_set_policy changes them per server).
// A job's deadline: submit time + min(samples / nominal rate, max offset).
// Earliest deadline first orders small jobs by size (short ones overtake long
// ones) while a long job is overtaken only by work whose deadline falls before
// its own.  Fair share offers every running job an equal part of the step.
#ifndef LFH_SERVE_POLICY
#define LFH_SERVE_POLICY 0   // 0: earliest deadline first, 1: fair share
#endif
#ifndef LFH_SERVE_RATE
#define LFH_SERVE_RATE 4.5e8
#endif
#ifndef LFH_SERVE_MAX_OFFSET
#define LFH_SERVE_MAX_OFFSET 1.0e9
#endif

// Content hash of the scene arrays (identical scenes share one resident
// copy).  Each array: 32-byte blocks into four independent multiply-rotate
// lanes (the round shape of xxHash64, here with splitmix64's odd constants),
// the tail bytes zero-padded into a last block, then the lanes, the length
// and the running hash through splitmix64's finaliser.  Several GB/s on one
// core, so a submit of a 20 MB scene spends a few milliseconds hashing.
uint64_t fmix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
inline uint64_t rotl64(uint64_t x, int r) { return (x << r) | (x >> (64 - r)); }
inline uint64_t lane_round(uint64_t acc, uint64_t w) {
    return rotl64(acc + w * 0xBF58476D1CE4E5B9ull, 31) * 0x94D049BB133111EBull;
}
uint64_t content_hash(uint64_t h, const void* p, size_t n) {
    const unsigned char* c = static_cast<const unsigned char*>(p);
    uint64_t v[4] = {0x9E3779B97F4A7C15ull, 0x3C6EF372FE94F82Aull, 0xDAA66D2C7DDF743Full, 0x78DDE6E5FD29F054ull};
    size_t i = 0;
    for (; i + 32 <= n; i += 32) {
        uint64_t w[4];
        std::memcpy(w, c + i, 32);
        v[0] = lane_round(v[0], w[0]);
        v[1] = lane_round(v[1], w[1]);
        v[2] = lane_round(v[2], w[2]);
        v[3] = lane_round(v[3], w[3]);
    }
    if (i < n) {
        uint64_t w[4] = {0, 0, 0, 0};
        std::memcpy(w, c + i, n - i);
        for (int k = 0; k < 4; ++k) v[k] = lane_round(v[k], w[k]);
    }
    const uint64_t x = rotl64(v[0], 1) + rotl64(v[1], 7) + rotl64(v[2], 12) + rotl64(v[3], 18);
    return fmix64(h ^ fmix64(x ^ (uint64_t)n));
}

uint64_t hash_desc(const lfh_scene_desc& d) {
    uint64_t h = 1469598103934665603ull;
    h = content_hash(h, d.nodes, (size_t)d.num_nodes * 36);
    h = content_hash(h, &d.top_level_index, 4);
    h = content_hash(h, d.vert_indices, (size_t)d.num_tris * 12);
    h = content_hash(h, d.vertices_uvx, (size_t)d.num_vertices * 16);
    if (d.normals_uvy) h = content_hash(h, d.normals_uvy, (size_t)d.num_vertices * 16);
    h = content_hash(h, d.materials, (size_t)d.num_materials * 112);
    h = content_hash(h, d.transforms, (size_t)d.num_instances * 64);
    if (d.lights) h = content_hash(h, d.lights, (size_t)d.num_lights * 60);
    if (d.hdr_width > 0) {
        size_t n = (size_t)d.hdr_width * d.hdr_height;
        h = content_hash(h, d.hdr_cols, n * 12);
        if (d.hdr_marginal) h = content_hash(h, d.hdr_marginal, (size_t)d.hdr_height * 8);
        if (d.hdr_conditional) h = content_hash(h, d.hdr_conditional, n * 8);
    }
    if (d.textures && d.num_textures > 0)
        h = content_hash(h, d.textures, (size_t)d.num_textures * d.tex_width * d.tex_height * 4);
    if (!d.nodes && d.num_meshes > 0) {  // device BVH build: the mesh and instance tables
        h = content_hash(h, d.mesh_tri_offsets, (size_t)(d.num_meshes + 1) * 8);
        h = content_hash(h, d.instance_meshes, (size_t)d.num_instances * 4);
        h = content_hash(h, d.instance_materials, (size_t)d.num_instances * 4);
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

// World-space bounds of a scene desc: each instance's mesh box (or, without
// mesh tables, the box of all vertices) through its transform; the guide's
// position encoding is normalised to it.
void scene_bounds(const lfh_scene_desc& d, float lo[3], float hi[3]) {
    for (int a = 0; a < 3; ++a) {
        lo[a] = 3.0e38f;
        hi[a] = -3.0e38f;
    }
    auto box_of = [&](int64_t t0, int64_t t1, float* b) {
        for (int a = 0; a < 3; ++a) {
            b[a] = 3.0e38f;
            b[3 + a] = -3.0e38f;
        }
        for (int64_t t = t0; t < t1; ++t)
            for (int k = 0; k < 3; ++k) {
                const float* v = d.vertices_uvx + (size_t)d.vert_indices[t * 3 + k] * 4;
                for (int a = 0; a < 3; ++a) {
                    b[a] = std::min(b[a], v[a]);
                    b[3 + a] = std::max(b[3 + a], v[a]);
                }
            }
    };
    std::vector<float> mbox;
    const bool meshes = d.num_meshes > 0 && d.mesh_tri_offsets && d.instance_meshes;
    if (meshes) {
        mbox.resize((size_t)d.num_meshes * 6);
        for (int m = 0; m < d.num_meshes; ++m) box_of(d.mesh_tri_offsets[m], d.mesh_tri_offsets[m + 1], &mbox[(size_t)m * 6]);
    } else {
        mbox.resize(6);
        box_of(0, d.num_tris, mbox.data());
    }
    for (int i = 0; i < d.num_instances; ++i) {
        const float* b = &mbox[meshes ? (size_t)d.instance_meshes[i] * 6 : 0];
        if (b[0] > b[3]) continue;
        const float* t = d.transforms + (size_t)i * 16;   // column-major
        for (int c = 0; c < 8; ++c) {
            const float p[3] = {b[(c & 1) ? 3 : 0], b[(c & 2) ? 4 : 1], b[(c & 4) ? 5 : 2]};
            for (int a = 0; a < 3; ++a) {
                const float w = t[a] * p[0] + t[4 + a] * p[1] + t[8 + a] * p[2] + t[12 + a];
                lo[a] = std::min(lo[a], w);
                hi[a] = std::max(hi[a], w);
            }
        }
    }
    for (int a = 0; a < 3; ++a)
        if (lo[a] > hi[a]) lo[a] = hi[a] = 0.0f;
}

enum SceneState { kNotResident = 0, kUploading = 1, kResident = 2, kUploadFailed = 3 };

struct GuideDeleter {
    void operator()(lfh_guide* g) const { lfh_guide_destroy(g); }
};

struct SceneEntry {
    uint64_t hash = 0;
    PackedScene ps;
    DeviceScene ds;
    int state = kNotResident;     // written by the admission thread under the server mutex
    std::string upload_error;
    int refs = 0;
    int64_t bytes = 0;
    double last_use = 0;
    float lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};   // world bounds (guide encoding)
    bool bounds_set = false;
    std::unique_ptr<lfh_guide, GuideDeleter> guide;  // of the scene's guided jobs, created at their first admission
    GuideRun grun;                // that guide's training cadence and pass count (buffers: the server's)
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
    bool guided = false;
};

}  // namespace
}  // namespace lfh

struct lfh_server {
    int64_t budget = 0;
    int max_jobs = 0;
    hipStream_t stream = nullptr;          // wavefronts
    hipStream_t upload_stream = nullptr;   // the admission thread's uploads and device BVH builds
    std::chrono::steady_clock::time_point t0;
    int64_t next_id = 1;
    std::map<int64_t, std::unique_ptr<lfh::Job>> jobs;
    std::map<uint64_t, std::shared_ptr<lfh::SceneEntry>> scenes;
    lfh::PassBufs pass;
    lfh::GuideBufs gbuf;                   // guided wavefronts' buffers (one wavefront at a time)
    lfh::BounceGraph bgraph;               // the plain group's bounce chain
    int64_t gbuf_charged = 0;
    lfh::DevBuf table, spill, errflag;
    lfh::LaneRes lane;                     // the bounce chain's shadow lane
    int grid = 0;
    int64_t used = 0;  // resident scenes + accum of admitted jobs + path shares of running jobs
    std::string err;
    // admission thread: uploads scenes queued by the step loop
    std::mutex mu;
    std::condition_variable cv;            // work for the admission thread
    std::condition_variable done_cv;       // an upload finished
    std::deque<std::shared_ptr<lfh::SceneEntry>> uploads;
    int uploading = 0;                     // scenes queued or being uploaded
    bool stop = false;
    std::thread worker;
    // step policy (lfh_server_set_policy)
    int policy = LFH_SERVE_POLICY;
    double sec_per_sample = 1.0 / LFH_SERVE_RATE;
    double max_offset = LFH_SERVE_MAX_OFFSET;
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

// Evict unreferenced resident scenes (oldest first) until `need` more bytes
// fit.  Called with the server mutex held.
static void evict_for(lfh_server* s, int64_t need) {
    while (s->used + need > s->budget) {
        std::shared_ptr<lfh::SceneEntry> victim;
        for (auto& kv : s->scenes) {
            auto& e = kv.second;
            if (e->state == lfh::kResident && e->refs == 0 && (!victim || e->last_use < victim->last_use)) victim = e;
        }
        if (!victim) return;
        victim->ds.release();
        victim->state = lfh::kNotResident;
        s->used -= victim->bytes;
    }
}

// The admission thread: makes queued scenes resident on its own stream (the
// step loop meanwhile renders on the server stream); a failed upload leaves
// the scene's device arrays released and its error for the jobs waiting on it.
static void admission_loop(lfh_server* s) {
    for (;;) {
        std::shared_ptr<lfh::SceneEntry> e;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv.wait(lk, [&] { return s->stop || !s->uploads.empty(); });
            if (s->stop) return;
            e = s->uploads.front();
            s->uploads.pop_front();
        }
        std::string berr;
        hipError_t ue = lfh::upload_scene(e->ps, &e->ds, s->upload_stream, &berr);
        if (ue == hipSuccess && berr.empty() && lfh_internal::inject_fault(lfh_internal::kFaultSceneUpload))
            ue = hipErrorOutOfMemory;
        const bool ok = ue == hipSuccess && berr.empty();
        if (!ok) {
            (void)hipStreamSynchronize(s->upload_stream);
            e->ds.release();
        }
        {
            std::lock_guard<std::mutex> lk(s->mu);
            if (ok) {
                e->state = lfh::kResident;
            } else {
                e->state = lfh::kUploadFailed;
                e->upload_error = !berr.empty() ? "device BVH build: " + berr
                                                : std::string("scene upload: ") + hipGetErrorString(ue);
                s->used -= e->bytes;
            }
            s->uploading--;
        }
        s->done_cv.notify_all();
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
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&s->upload_stream, hipStreamNonBlocking);
    if (e == hipSuccess) {
        s->grid = pt_trace_grid_blocks();
        const size_t spill_bytes = (size_t)s->grid * lfh_trace_threads_per_block() * lfh_trace_spill_stride() * 4;
        e = s->spill.alloc(spill_bytes);
        if (e == hipSuccess && lfh::shadow_lane_enabled()) e = s->lane.init(spill_bytes);
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
    try {
        s->worker = std::thread(admission_loop, s);
    } catch (...) {
        lfh_server_destroy(s);
        return sfail(nullptr, LFH_E_OOM, "lfh_server_create: cannot start the admission thread");
    }
    *out = s;
    return LFH_OK;
}

void lfh_server_destroy(lfh_server* s) {
    if (!s) return;
    if (s->worker.joinable()) {
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->stop = true;
        }
        s->cv.notify_all();
        s->worker.join();
    }
    if (s->upload_stream) hipStreamSynchronize(s->upload_stream);
    if (s->stream) hipStreamSynchronize(s->stream);
    s->jobs.clear();
    s->scenes.clear();
    s->bgraph.reset();
    s->pass.release();
    s->table.release();
    s->spill.release();
    s->errflag.release();
    if (s->lane.ready()) hipStreamSynchronize(s->lane.s);
    s->lane.release();
    if (s->upload_stream) hipStreamDestroy(s->upload_stream);
    if (s->stream) hipStreamDestroy(s->stream);
    delete s;
}

double lfh_server_clock(const lfh_server* s) { return s ? s->now() : 0.0; }

const char* lfh_server_job_error(lfh_server* s, int64_t id) {
    if (!s) return "";
    auto it = s->jobs.find(id);
    return it == s->jobs.end() ? "unknown job" : it->second->error.c_str();
}

int lfh_server_set_policy(lfh_server* s, int32_t policy, double nominal_rate, double max_offset) {
    if (!s || (policy != 0 && policy != 1) || !(nominal_rate > 0.0) || !(max_offset > 0.0))
        return sfail(s, LFH_E_INVALID, "lfh_server_set_policy: bad arguments");
    s->policy = policy;
    s->sec_per_sample = 1.0 / nominal_rate;
    s->max_offset = max_offset;
    return LFH_OK;
}

int lfh_server_submit(lfh_server* s, const lfh_job_desc* jd, int64_t* job_id) {
    if (!s || !jd || !jd->scene) return sfail(s, LFH_E_INVALID, "lfh_server_submit: NULL argument");
    const lfh_options& o = jd->options;
    if (o.width <= 0 || o.height <= 0 || o.max_depth < 0 || jd->target_spp <= 0)
        return sfail(s, LFH_E_INVALID, "lfh_server_submit: bad resolution, depth or target spp");
    if (o.guiding && o.estimator != 1)
        return sfail(s, LFH_E_INVALID, "lfh_server_submit: guided jobs need options.estimator = 1");
    if (o.estimator != 0 && o.estimator != 1) return sfail(s, LFH_E_INVALID, "lfh_server_submit: bad estimator");
    if (o.frame_offset < 0) return sfail(s, LFH_E_INVALID, "lfh_server_submit: negative frame offset");
    auto job = std::make_unique<lfh::Job>();
    job->guided = o.guiding != 0;
    job->id = s->next_id++;
    job->opt = o;
    job->cam = jd->camera;
    job->target = jd->target_spp;
    job->priority = jd->priority;
    job->tag = jd->user_tag;
    job->t_submit = s->now();
    job->npix = (int64_t)o.width * o.height;
    job->frame += o.frame_offset;
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
    if (job->guided && !e->bounds_set) {
        lfh::scene_bounds(*jd->scene, e->lo, e->hi);
        e->bounds_set = true;
    }
    std::lock_guard<std::mutex> lk(s->mu);
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

namespace {

// One wavefront of a step: jobs (all plain, or all guided on one scene) with
// their passes.
struct Group {
    std::vector<lfh::Job*> jobs;
    std::vector<int> passes;
    lfh::SceneEntry* guided_scene = nullptr;
};

lfhd::JobDev job_dev(const lfh::Job* j, int passes, int64_t off) {
    lfhd::JobDev d;
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
    c.batch = passes;
    d.accum = j->accum.as<float4>();
    d.comp = j->comp.as<float4>();
    d.path_offset = (int)off;
    d.kahan = j->opt.kahan;
    return d;
}

}  // namespace

// Runs one group's wavefront on the server stream: generate, the bounces
// (with the scene guide's inference and training records for a guided group),
// accumulation, and the guide's optimisation steps.
static int run_group(lfh_server* s, const Group& g, const lfhd::JobDev* table_dev, int64_t total, int max_depth) {
    lfh::PassBufs& pb = s->pass;
    PathBufs v = lfh::view(pb);
    uint32_t* cnt = pb.counters.as<uint32_t>();
    unsigned long long* work = pb.work.as<unsigned long long>();
    SCHK(s, pt_reset(cnt, work, s->stream), "lfh_server_step: reset");
    SCHK(s, pt_generate_multi(table_dev, (int)g.jobs.size(), (int)total, v, pb.rq0.as<float4>(), cnt, s->stream),
         "lfh_server_step: generate");
    GuideHook gh{};
    lfhd::GuideDev gd{};
    lfhd::ShadeScene gsc{};
    lfh::GuidePassArgs ga{};
    int batch = 0;
    if (g.guided_scene) {
        lfh::SceneEntry& e = *g.guided_scene;
        for (int p : g.passes) batch = std::max(batch, p);
        ga.run = &e.grun;
        ga.paths = total;
        ga.cap = pb.cap;
        ga.max_depth = max_depth;
        ga.rad = pb.rad.as<float4>();
        ga.stream = s->stream;
        ga.buf = &s->gbuf;
        if (lfh::guide_begin(ga, batch, &gh, &gd) < 0) return sfail(s, LFH_E_HIP, "lfh_server_step: guide buffers");
        gsc = lfh::shade_view(e.ps, e.ds, g.jobs[0]->opt);
    }
    const ShadowLane lane = s->lane.lane();
    const ShadowLane* lp = s->lane.ready() ? &lane : nullptr;
    auto chain = [&]() -> hipError_t {
        for (int depth = 0; depth < max_depth; ++depth) {
            const int cur = depth & 1;
            float4* q[2] = {pb.rq0.as<float4>(), pb.rq1.as<float4>()};
            hipError_t e = pt_bounce_multi(table_dev, depth, pb.cap, q[cur], cnt + cur, q[cur ^ 1], cnt + (cur ^ 1),
                                           pb.hq.as<float4>(), pb.sq.as<float4>(), cnt + 2, v, s->spill.as<uint32_t>(),
                                           work, s->errflag.as<int>(), 1.0f + 1.0f / 1024.0f, s->grid, gh, &gd, &gsc,
                                           s->stream, lp);
            if (e != hipSuccess) return e;
        }
        return lp && max_depth > 0 ? pt_bounce_join(*lp, s->stream) : hipSuccess;
    };
    if (!g.guided_scene && lfh::bounce_graphs_enabled()) {
        // the plain group's chain reads its job table on the device: one
        // graph per table position, path capacity and depth
        const std::vector<uintptr_t> key = {(uintptr_t)table_dev, (uintptr_t)pb.rad.p, (uintptr_t)pb.cap,
                                            (uintptr_t)max_depth, (uintptr_t)s->spill.p, (uintptr_t)s->lane.s};
        SCHK(s, lfh::run_graph(s->bgraph, key, s->stream, chain), "lfh_server_step: bounce graph");
    } else {
        SCHK(s, chain(), "lfh_server_step: bounce");
    }
    SCHK(s, pt_accumulate_multi(table_dev, (int)total, v, s->stream), "lfh_server_step: accumulate");
    if (g.guided_scene) {
        const int r = lfh::guide_end(ga, gh, batch);
        if (r != LFH_OK) return sfail(s, r, std::string("lfh_server_step: guide: ") + lfh_internal::thread_error());
    }
    return LFH_OK;
}

int lfh_server_step(lfh_server* s, int32_t* active) {
    if (!s) return sfail(nullptr, LFH_E_INVALID, "lfh_server_step: NULL server");
    // -- admission: priority, then arrival; jobs that do not fit are skipped.
    //    Scenes that are not resident go to the admission thread; their jobs
    //    start in a later step, once the upload is done.
    std::vector<lfh::Job*> queued, running;
    for (auto& kv : s->jobs) {
        if (kv.second->state == 0) queued.push_back(kv.second.get());
        if (kv.second->state == 1) running.push_back(kv.second.get());
    }
    std::stable_sort(queued.begin(), queued.end(),
                     [](const lfh::Job* a, const lfh::Job* b) { return a->priority > b->priority; });
    bool uploads_pending = false;
    for (lfh::Job* j : queued) {
        if ((int)running.size() >= s->max_jobs) break;
        lfh::SceneEntry& e = *j->scene;
        int state;
        {
            std::lock_guard<std::mutex> lk(s->mu);
            state = e.state;
            if (state == lfh::kUploadFailed) {
                // the jobs waiting on a failed upload are rejected; the scene
                // may be uploaded again for later jobs
                e.state = lfh::kNotResident;
            } else if (state == lfh::kNotResident) {
                const int64_t need = e.bytes + j->npix * (lfh::kAccumBytes + lfh::kPathBytes);
                if (s->used + need > s->budget) evict_for(s, need);
                if (s->used + need <= s->budget) {
                    e.state = lfh::kUploading;
                    s->used += e.bytes;
                    s->uploading++;
                    s->uploads.push_back(j->scene);
                    s->cv.notify_one();
                }
            }
        }
        if (state == lfh::kUploadFailed) {
            reject(s, j, e.upload_error);
            continue;
        }
        if (state != lfh::kResident) {
            uploads_pending = true;
            continue;
        }
        const int64_t need = j->npix * (lfh::kAccumBytes + lfh::kPathBytes);
        {
            std::lock_guard<std::mutex> lk(s->mu);
            if (s->used + need > s->budget) evict_for(s, need);
            if (s->used + need > s->budget) continue;
        }
        hipError_t ae = j->accum.alloc((size_t)j->npix * 16);
        if (ae == hipSuccess) ae = j->comp.alloc((size_t)j->npix * 16);
        if (ae == hipSuccess) ae = pt_clear4(j->accum.as<float4>(), j->npix, s->stream);
        if (ae == hipSuccess) ae = pt_clear4(j->comp.as<float4>(), j->npix, s->stream);
        if (ae == hipSuccess && j->guided && !e.guide) {
            // the scene's guide: created by its first guided job
            lfh_guide* gp = nullptr;
            if (lfh_guide_create(e.lo, e.hi, 0, LFH_SERVE_GUIDE_LR, &gp) != LFH_OK) ae = hipErrorOutOfMemory;
            e.guide.reset(gp);
            e.grun.guide = gp;
            e.grun.train_every = 1;
        }
        if (ae != hipSuccess) {
            j->accum.release();
            j->comp.release();
            reject(s, j, std::string("job buffers: ") + hipGetErrorString(ae));
            continue;
        }
        j->charged = need;
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->used += j->charged;
        }
        j->state = 1;
        j->t_admit = s->now();
        running.push_back(j);
    }
    // -- passes for this step.  Jobs are ordered by deadline, a job's deadline
    //    being its submit time plus its samples at the policy's nominal rate
    //    (at most its max offset): a small job finishes soon after it arrives,
    //    and a large job is not overtaken by work whose deadline is later than
    //    its own.  Earliest deadline first: in that order each job takes as
    //    many of its remaining passes as the step's path budget still holds;
    //    fair share: each is offered an equal part of the budget.  A job takes
    //    at most kStepMaxPasses (one while its scene's guide is warming up);
    //    the earliest deadline always gets one.  Passes of a job are
    //    accumulated per pixel in frame order, so its image is that of single
    //    passes whatever the grouping.
    if (!running.empty()) {
        int64_t one_pass = 0;
        for (lfh::Job* j : running) one_pass += j->npix;
        const int64_t spare = std::max<int64_t>(0, (s->budget - s->used) / lfh::kPathBytes);
        auto deadline = [s](const lfh::Job* j) {
            return j->t_submit + std::min(s->max_offset, (double)j->npix * (double)j->target * s->sec_per_sample);
        };
        std::stable_sort(running.begin(), running.end(), [&](const lfh::Job* a, const lfh::Job* b) {
            const double da = deadline(a), db = deadline(b);
            return da < db || (da == db && a->id < b->id);
        });
        int64_t cap = std::min(lfh::kStepPaths, one_pass + spare);
        cap = std::max(cap, running[0]->npix);
        // passes per job: what each job could take this step, then the path
        // budget split between them
        std::vector<int64_t> want(running.size()), give(running.size(), 0);
        for (size_t k = 0; k < running.size(); ++k) {
            lfh::Job* j = running[k];
            want[k] = std::min<int64_t>((int64_t)j->target - j->spp_done, lfh::kStepMaxPasses);
            if (j->guided) want[k] = std::min<int64_t>(want[k], lfh::guide_pass_limit(j->scene->grun));
        }
        if (s->policy == 1) {
            // fair shares (water filling): every job is offered an equal part of
            // the budget; jobs that need less return the rest, which the others
            // share in the next round; the earliest deadline always gets a pass
            int64_t left_budget = cap;
            std::vector<size_t> open;
            for (size_t k = 0; k < running.size(); ++k) open.push_back(k);
            while (!open.empty() && left_budget > 0) {
                const int64_t share = left_budget / (int64_t)open.size();
                std::vector<size_t> still;
                int64_t used_round = 0;
                for (size_t k : open) {
                    const int64_t fit = share / running[k]->npix;
                    if (fit >= want[k]) {
                        give[k] = want[k];
                        used_round += want[k] * running[k]->npix;
                    } else {
                        still.push_back(k);
                    }
                }
                left_budget -= used_round;
                if (still.size() == open.size()) {
                    // nobody is satisfied by an equal share: each takes what its share holds
                    for (size_t k : still) {
                        give[k] = std::min<int64_t>(want[k], share / running[k]->npix);
                        left_budget -= give[k] * running[k]->npix;
                    }
                    break;
                }
                open.swap(still);
            }
            if (give[0] == 0 && want[0] > 0) give[0] = 1;   // running[0]: the earliest deadline
        } else {
            // earliest deadline first: in that order each job takes what the budget still holds
            int64_t total0 = 0;
            for (size_t k = 0; k < running.size(); ++k) {
                give[k] = std::min<int64_t>(want[k], (cap - total0) / running[k]->npix);
                if (k == 0) give[k] = std::max<int64_t>(give[k], std::min<int64_t>(want[k], 1));
                total0 += std::max<int64_t>(give[k], 0) * running[k]->npix;
            }
        }
        // groups: the plain jobs together, the guided jobs of each scene together
        std::vector<Group> groups(1);
        std::map<lfh::SceneEntry*, size_t> gi;
        int64_t total = 0;
        for (size_t rk = 0; rk < running.size(); ++rk) {
            lfh::Job* j = running[rk];
            if (give[rk] <= 0) continue;
            const int64_t gv = give[rk];
            size_t k = 0;
            if (j->guided) {
                auto it = gi.find(j->scene.get());
                if (it == gi.end()) {
                    groups.push_back(Group{});
                    groups.back().guided_scene = j->scene.get();
                    it = gi.emplace(j->scene.get(), groups.size() - 1).first;
                }
                k = it->second;
            }
            groups[k].jobs.push_back(j);
            groups[k].passes.push_back((int)gv);
            total += gv * j->npix;
        }
        // job tables of all groups, uploaded together
        std::vector<lfhd::JobDev> tab;
        std::vector<int64_t> totals(groups.size(), 0);
        std::vector<int> depths(groups.size(), 0);
        std::vector<size_t> first(groups.size(), 0);
        int64_t largest = 0;
        for (size_t gk = 0; gk < groups.size(); ++gk) {
            first[gk] = tab.size();
            int64_t off = 0;
            for (size_t k = 0; k < groups[gk].jobs.size(); ++k) {
                lfh::Job* j = groups[gk].jobs[k];
                tab.push_back(job_dev(j, groups[gk].passes[k], off));
                off += j->npix * groups[gk].passes[k];
                depths[gk] = std::max(depths[gk], j->opt.max_depth);
                j->scene->last_use = s->now();
            }
            totals[gk] = off;
            largest = std::max(largest, off);
        }
        if (largest > (1LL << 30)) return sfail(s, LFH_E_UNSUPPORTED, "lfh_server_step: too many pixels in flight");
        if (largest > 0) {
            if (s->pass.cap < largest || s->pass.cap > 4 * largest + (1 << 23)) {
                // grow with headroom, shrink only when far too large (a reallocation waits for the GPU)
                const int64_t pcap = std::min<int64_t>(largest + largest / 4, std::max<int64_t>(largest, lfh::kStepPaths));
                SCHK(s, hipStreamSynchronize(s->stream), "lfh_server_step: sync");
                SCHK(s, s->pass.alloc((int)pcap), "lfh_server_step: path buffers");
            }
            const bool any_guided = groups.size() > 1;
            if (any_guided) {
                int gdepth = 0;
                for (size_t gk = 1; gk < groups.size(); ++gk) gdepth = std::max(gdepth, depths[gk]);
                SCHK(s, lfh::guide_buffers(s->gbuf, s->pass.cap, gdepth, s->stream), "lfh_server_step: guide buffers");
                const int64_t gb = (int64_t)lfh::guide_buffer_bytes(s->pass.cap, gdepth);
                if (gb > s->gbuf_charged) {
                    std::lock_guard<std::mutex> lk(s->mu);
                    s->used += gb - s->gbuf_charged;
                    s->gbuf_charged = gb;
                }
            }
            SCHK(s, hipMemcpyAsync(s->table.p, tab.data(), tab.size() * sizeof(lfhd::JobDev), hipMemcpyHostToDevice,
                                   s->stream),
                 "lfh_server_step: job table");
            const lfhd::JobDev* jt = s->table.as<lfhd::JobDev>();
            for (size_t gk = 0; gk < groups.size(); ++gk) {
                if (groups[gk].jobs.empty()) continue;
                int r = run_group(s, groups[gk], jt + first[gk], totals[gk], depths[gk]);
                if (r != LFH_OK) return r;
            }
            SCHK(s, hipStreamSynchronize(s->stream), "lfh_server_step: sync");
            int flag = 0;
            SCHK(s, hipMemcpy(&flag, s->errflag.p, sizeof(int), hipMemcpyDeviceToHost), "lfh_server_step: flag");
            if (flag) {
                SCHK(s, pt_clear_flag(s->errflag.as<int>(), s->stream), "lfh_server_step: flag");
                SCHK(s, hipStreamSynchronize(s->stream), "lfh_server_step: flag");
                return sfail(s, LFH_E_UNSUPPORTED, "lfh_server_step: traversal stack overflow (step not counted)");
            }
            const double t = s->now();
            for (const Group& g : groups)
                for (size_t k = 0; k < g.jobs.size(); ++k) {
                    lfh::Job* j = g.jobs[k];
                    j->spp_done += g.passes[k];
                    j->frame += g.passes[k];
                    if (j->spp_done >= j->target) {
                        j->state = 2;
                        j->t_done = t;
                        std::lock_guard<std::mutex> lk(s->mu);
                        s->used -= j->npix * lfh::kPathBytes;  // path share back; accum held until release
                        j->charged = j->npix * lfh::kAccumBytes;
                    }
                }
        }
    } else if (uploads_pending) {
        // nothing to render until an upload finishes: wait for one (bounded)
        std::unique_lock<std::mutex> lk(s->mu);
        s->done_cv.wait_for(lk, std::chrono::milliseconds(2), [&] { return s->uploading == 0; });
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
    // a running job is cancelled: the wavefronts that used its buffers have finished
    SCHK(s, hipStreamSynchronize(s->stream), "lfh_server_release: sync");
    if (j.state == 1 || j.state == 2) {
        std::lock_guard<std::mutex> lk(s->mu);
        s->used -= j.charged;
    }
    if (j.scene) j.scene->refs--;
    s->jobs.erase(it);
    return LFH_OK;
}

int lfh_server_reset_guides(lfh_server* s) {
    if (!s) return sfail(nullptr, LFH_E_INVALID, "lfh_server_reset_guides: NULL server");
    for (auto& kv : s->jobs)
        if (kv.second->guided && (kv.second->state == 0 || kv.second->state == 1))
            return sfail(s, LFH_E_STATE, "lfh_server_reset_guides: a guided job is queued or running");
    SCHK(s, hipStreamSynchronize(s->stream), "lfh_server_reset_guides: sync");
    std::lock_guard<std::mutex> lk(s->mu);
    for (auto& kv : s->scenes) {
        lfh::SceneEntry& e = *kv.second;
        e.guide.reset();
        e.grun.guide = nullptr;
        e.grun.pass = 0;
        e.grun.train_every = 0;
    }
    return LFH_OK;
}

}  // extern "C"
