// Learned path guide: lfh_guide_* entry points, online training from render
// passes, and the per-pass hooks of the renderer.  The guide has no
// counterpart in the reference renderer; it adds a second sampling strategy
// (BF16 MFMA MLP -> 8 vMF lobes) that the shading stage mixes 50/50 with
// DisneySample under the consistent one-sample MIS estimator.
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <limits>
#include <vector>

#include "../../include/lfh.h"
#include "capi_common.h"
#include "ctx.h"
#include "guide_kernels.h"
#include "pt_kernels.h"
#include "render.h"

using lfhd::kGuideLobeF4;
using lfhd::kGuideOut;
using lfhd::kGuideParams;

// Recorded paths per wavefront (log2) and training samples kept from it (in
// 65536-sample batches).  Served guiding room, median guided time to RMSE
// 1e-2 over 10 runs: 2^18 / 4 batches 0.182 s, 2^19 / 8 0.186 s (128 warm-up
// steps), 2^20 / 16 0.162 s (64 steps; 0.167 s with 128, 0.189 s with 256).
#ifndef LFH_GUIDE_RECORD_LOG2
#define LFH_GUIDE_RECORD_LOG2 20
#endif
#ifndef LFH_GUIDE_SAMPLE_BATCHES
#define LFH_GUIDE_SAMPLE_BATCHES 16
#endif
#ifndef LFH_GUIDE_SERIAL_WARMUP
#define LFH_GUIDE_SERIAL_WARMUP 0
#endif

namespace {
constexpr int kBatch = 65536;             // samples per optimisation step
constexpr int kMaxSamples = LFH_GUIDE_SAMPLE_BATCHES * kBatch;  // samples kept from one wavefront
constexpr int kRecordedPaths = 1 << LFH_GUIDE_RECORD_LOG2;     // recorded paths per wavefront, at most
#ifndef LFH_GUIDE_WARM_STEPS
// Served guiding room, median over 8-10 independent runs of the guided time
// to RMSE 1e-2 (plain: 0.276 s): warm-up passes / Adam steps / weight cap
// 4/128/32 0.195 s, 4/64/32 0.174, 8/128/32 0.180, 8/64/32 0.167, 4/128/128
// 0.161-0.177, 4/64/128 0.173, 8/64/128 0.175, 8/128/128 0.188, 4/128/64
// 0.191, 8/64/64 0.173, 4/128/1024 0.192; serial warm-up 0.308.  The spread
// of one median over different run sets is about +-0.015 s.  Re-measured on one
// warm server over 32 sequences (learning rate 1e-3): 64 steps 0.168 s, 128 steps
// 0.162; with rate 2e-3: 32 steps 0.160, 64 steps 0.161, 96 steps 0.161, 128 steps 0.168.
#define LFH_GUIDE_WARM_STEPS 64
#endif
#ifndef LFH_GUIDE_WARM_REPEATS
#define LFH_GUIDE_WARM_REPEATS 4
#endif
#ifndef LFH_GUIDE_PLAIN_WARMUP
// > 0: the first passes (one wavefront) sample the BSDF alone and only train
// the guide; 0: serial warm-up (one pass per wavefront until kWarmSteps).
// Served guiding room, median guided time to RMSE 1e-2 over 8 runs (64 Adam
// steps, cap 32): serial 0.308 s, 4 passes 0.174 s, 8 passes 0.167 s; with cap
// 128: 4 passes 0.173 s, 8 passes 0.175 s.  On one warm server over 32 sequences
// (learning rate 1e-3): 2 passes 0.166 s, 4 passes 0.168, 8 passes 0.169; with
// rate 2e-3: 2 passes 0.152 (0.157 on a second set), 3 passes 0.165, 8 passes 0.170.
// With rate 1.5e-3, two sets of 32 sequences: 2 passes 0.152 / 0.156 s, 4 passes
// 0.148 / 0.148 s.
#define LFH_GUIDE_PLAIN_WARMUP 4
#endif
#ifndef LFH_GUIDE_LATE_REPEATS
#define LFH_GUIDE_LATE_REPEATS 1   // median guided time to 1e-2 over 8 runs: 1 step 0.195 s, 2 steps 0.193 s
#endif
constexpr int kWarmSteps = LFH_GUIDE_WARM_STEPS;       // Adam steps taken at the warm-up rate
constexpr int kLateRepeats = LFH_GUIDE_LATE_REPEATS;   // Adam steps per training pass after the warm-up
constexpr int kWarmRepeats = LFH_GUIDE_WARM_REPEATS;   // optimisation passes over a pass's samples while warming up
#ifndef LFH_GUIDE_WEIGHT_CAP
#define LFH_GUIDE_WEIGHT_CAP 128.0f   // the door's rare bright paths need weights far above the mean
#endif
constexpr float kWeightCap = LFH_GUIDE_WEIGHT_CAP;   // online training: weights / their mean, capped
#ifndef LFH_GUIDE_LATE_EVERY
// Served guiding room, median over 6 runs of the guided time to RMSE 1e-2 / 5e-3:
// every pass 0.156 / 0.606 s; every 2nd pass after 128 steps 0.156 / 0.586 s
// (2.106 ms per pass against 2.12); every 2nd or 4th pass after 64 steps 0.167 /
// 0.657 s and 0.187 / 0.624 s.  On one warm server over 32 sequences to 1e-2:
// every pass 0.168 s, every 2nd after 128 steps 0.168 s.
#define LFH_GUIDE_LATE_EVERY 2
#endif
#ifndef LFH_GUIDE_LATE_AFTER
#define LFH_GUIDE_LATE_AFTER 128
#endif
constexpr int kLateEvery = LFH_GUIDE_LATE_EVERY;   // once kLateAfter steps are taken, one training pass in kLateEvery
constexpr int kLateAfter = LFH_GUIDE_LATE_AFTER;
// The guide learns fastest early at a higher rate, and refines better later at a lower
// one.  Served guiding room (rate 1.5e-3), median guided time to RMSE 1e-2 / 5e-3 /
// 2.5e-3 over 8 sequences on one warm server: constant 0.143 / 0.571-0.573 / 2.41-2.42 s;
// halved after 160 steps 0.143 / 0.494 / 2.32; halved after 128 0.144 / 0.552 / 2.40;
// halved after 256 0.143 / 0.568 / 2.22; a third after 160 0.143 / 0.513 / 2.26; constant
// 1e-3: 0.154 / 0.604 / 2.16.
#ifndef LFH_GUIDE_LR_LATE
#define LFH_GUIDE_LR_LATE 0.5f
#endif
#ifndef LFH_GUIDE_LR_DECAY_AFTER
#define LFH_GUIDE_LR_DECAY_AFTER 160
#endif
constexpr float kLrLate = LFH_GUIDE_LR_LATE;            // learning-rate factor once kLrDecayAfter steps are taken
constexpr int kLrDecayAfter = LFH_GUIDE_LR_DECAY_AFTER;
}  // namespace

struct lfh_guide {
    std::string err;
    hipStream_t stream = nullptr;
    lfh::DevBuf master, m, v, wf, wb, bias;
    lfh::DevBuf partial, samples, count, grad, loss, in, aux, out;
    float lo[3], hi[3];
    float lr = 1e-3f;
    int64_t steps = 0;   // Adam steps taken

    lfhd::GuideDev dev() const {
        lfhd::GuideDev g;
        g.wf = wf.as<const uint16_t>();
        g.wb = wb.as<const uint16_t>();
        g.bias = bias.as<const float>();
        for (int a = 0; a < 3; ++a) {
            g.lo[a] = lo[a];
            float e = hi[a] - lo[a];
            g.inv_ext[a] = 1.0f / (e > 1e-6f ? e : 1e-6f);
        }
        return g;
    }
    lfhd::GuideAdam adam(bool step) {
        lfhd::GuideAdam a;
        a.master = master.as<float>();
        a.m = m.as<float>();
        a.v = v.as<float>();
        a.wf = wf.as<uint16_t>();
        a.wb = wb.as<uint16_t>();
        a.bias = bias.as<float>();
        a.lr = steps >= kLrDecayAfter ? lr * kLrLate : lr;
        a.b1 = 0.9f;
        a.b2 = 0.999f;
        a.eps = 1e-8f;
        const double t = (double)(step ? steps + 1 : 1);
        a.c1 = (float)(1.0 / (1.0 - std::pow(0.9, t)));
        a.c2 = (float)(1.0 / (1.0 - std::pow(0.999, t)));
        return a;
    }
};

namespace {

int gfail(lfh_guide* g, int code, const std::string& msg) {
    if (g) g->err = msg;
    lfh_internal::set_thread_error(msg);
    return code;
}
int ghip(lfh_guide* g, hipError_t e, const char* where) {
    return gfail(g, e == hipErrorOutOfMemory ? LFH_E_OOM : LFH_E_HIP,
                 std::string(where) + ": " + hipGetErrorString(e));
}
#define GCHK(g, expr, where)                          \
    do {                                              \
        hipError_t _e = (expr);                       \
        if (_e != hipSuccess) return ghip(g, _e, where); \
    } while (0)

hipError_t ensure(lfh::DevBuf& b, size_t bytes) { return b.bytes >= bytes ? hipSuccess : b.alloc(bytes); }

uint64_t splitmix(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// He-uniform weights (bound sqrt(6 / fan_in)), output layer scaled by 0.1,
// zero biases.
std::vector<float> init_params(uint64_t seed) {
    std::vector<float> p(kGuideParams, 0.0f);
    uint64_t s = seed;
    for (int l = 0; l < lfhd::kGuideLayers; ++l) {
        const int nout = lfhd::guide_layer_out(l), off = lfhd::guide_layer_offset(l);
        const float bound = std::sqrt(6.0f / 64.0f) * (l == 3 ? 0.1f : 1.0f);
        for (int i = 0; i < 64 * nout; ++i) {
            const double u = (double)(splitmix(s) >> 11) * (1.0 / 9007199254740992.0);
            p[off + i] = (float)((2.0 * u - 1.0) * bound);
        }
    }
    return p;
}

// Uploads params as the master copy, resets Adam, rebuilds the BF16 copies.
int load_params(lfh_guide* g, const float* p) {
    GCHK(g, hipMemcpyAsync(g->master.p, p, kGuideParams * sizeof(float), hipMemcpyHostToDevice, g->stream),
         "guide: upload");
    GCHK(g, hipMemsetAsync(g->m.p, 0, kGuideParams * sizeof(float), g->stream), "guide: reset");
    GCHK(g, hipMemsetAsync(g->v.p, 0, kGuideParams * sizeof(float), g->stream), "guide: reset");
    GCHK(g, lfhd::guide_refresh(g->master.as<float>(), g->wf.as<uint16_t>(), g->wb.as<uint16_t>(),
                                g->bias.as<float>(), g->stream),
         "guide: refresh");
    GCHK(g, hipStreamSynchronize(g->stream), "guide: sync");
    g->steps = 0;
    return LFH_OK;
}

// Host samples (pos, n, wo) + dirs + weights -> device sample records.
int upload_batch(lfh_guide* g, const float* in9, const float* dirs3, const float* w, int64_t n) {
    std::vector<float> h((size_t)n * 16);
    for (int64_t i = 0; i < n; ++i) {
        float* r = &h[(size_t)i * 16];
        const float* x = in9 + i * 9;
        r[0] = x[0], r[1] = x[1], r[2] = x[2], r[3] = w[i];
        r[4] = x[3], r[5] = x[4], r[6] = x[5], r[7] = 0.0f;
        r[8] = x[6], r[9] = x[7], r[10] = x[8], r[11] = 0.0f;
        r[12] = dirs3[i * 3], r[13] = dirs3[i * 3 + 1], r[14] = dirs3[i * 3 + 2], r[15] = 0.0f;
    }
    GCHK(g, ensure(g->samples, h.size() * sizeof(float)), "guide: alloc");
    GCHK(g, hipMemcpyAsync(g->samples.p, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice, g->stream),
         "guide: upload");
    const int cnt = (int)n;
    GCHK(g, hipMemcpyAsync(g->count.p, &cnt, sizeof(int), hipMemcpyHostToDevice, g->stream), "guide: upload");
    GCHK(g, hipStreamSynchronize(g->stream), "guide: sync");
    return LFH_OK;
}

int check_batch(lfh_guide* g, const char* fn, const void* a, const void* b, int64_t n) {
    if (!g) return gfail(nullptr, LFH_E_INVALID, std::string(fn) + ": NULL guide");
    if (n < 0 || n > (int64_t)1 << 24) return gfail(g, LFH_E_INVALID, std::string(fn) + ": bad batch size");
    if (n > 0 && (!a || !b)) return gfail(g, LFH_E_INVALID, std::string(fn) + ": NULL argument");
    return LFH_OK;
}

// Forward of host queries into g->aux (lobes) and/or g->out (raw outputs).
int forward_host(lfh_guide* g, const float* in9, int64_t n, bool raw, bool lobes) {
    GCHK(g, ensure(g->in, (size_t)n * 9 * sizeof(float)), "guide: alloc");
    if (raw) GCHK(g, ensure(g->out, (size_t)n * kGuideOut * sizeof(float)), "guide: alloc");
    if (lobes) GCHK(g, ensure(g->aux, (size_t)n * kGuideLobeF4 * sizeof(float4)), "guide: alloc");
    GCHK(g, hipMemcpyAsync(g->in.p, in9, (size_t)n * 9 * sizeof(float), hipMemcpyHostToDevice, g->stream),
         "guide: upload");
    GCHK(g, lfhd::guide_forward_host(g->dev(), g->in.as<float>(), (int)n, raw ? g->out.as<float>() : nullptr,
                                     lobes ? g->aux.as<float4>() : nullptr, g->stream),
         "guide: forward");
    return LFH_OK;
}

int run_batch(lfh_guide* g, int64_t n, bool step, float* loss, float* grads) {
    const int batch = (int)(n > 0 ? n : 1);
    GCHK(g, ensure(g->partial, lfhd::guide_partial_floats(batch) * sizeof(float)), "guide: alloc");
    lfhd::GuideAdam ad = g->adam(step);
    GCHK(g, lfhd::guide_train(g->dev(), g->samples.as<float4>(), g->count.as<int>(), 0, batch,
                              g->partial.as<float>(), ad, step ? nullptr : g->grad.as<float>(), g->loss.as<float>(),
                              g->stream),
         "guide: train");
    if (step && n > 0) g->steps++;
    if (grads)
        GCHK(g, hipMemcpyAsync(grads, g->grad.p, kGuideParams * sizeof(float), hipMemcpyDeviceToHost, g->stream),
             "guide: download");
    if (loss) GCHK(g, hipMemcpyAsync(loss, g->loss.p, sizeof(float), hipMemcpyDeviceToHost, g->stream), "guide: download");
    GCHK(g, hipStreamSynchronize(g->stream), "guide: sync");
    return LFH_OK;
}

}  // namespace

namespace lfh {

// Per-pass hooks of the renderer ------------------------------------------------------
// Passes that may share one wavefront: the BSDF-only warm-up passes in one,
// then any number (with the serial warm-up: one at a time until the warm-up
// steps are taken, each pass sampling the guide the previous pass trained).
// Wavefronts of at most 4, 8 or 6 passes for the first 16-24 steps after the
// warm-up measured no better (median guided time to 1e-2 over 12 runs: 0.191,
// 0.187, 0.175 s against 0.171 s).
int guide_pass_limit(const GuideRun& run) {
    const lfh_guide* g = run.guide;
    if (LFH_GUIDE_PLAIN_WARMUP > 0 && g && run.train_every > 0 && run.pass < LFH_GUIDE_PLAIN_WARMUP)
        return LFH_GUIDE_PLAIN_WARMUP - run.pass;   // the warm-up passes share one wavefront
    if (LFH_GUIDE_SERIAL_WARMUP && g && run.train_every > 0 && g->steps < kWarmSteps) return 1;
    return std::numeric_limits<int>::max();
}

namespace {
// Whether a wavefront starting at run.pass belongs to the BSDF-only warm-up.
bool plain_warmup(const GuideRun& run) {
    return LFH_GUIDE_PLAIN_WARMUP > 0 && run.train_every > 0 && run.pass < LFH_GUIDE_PLAIN_WARMUP;
}
}  // namespace

namespace {
// Training passes among the wavefront's passes run.pass .. run.pass + batch - 1.
int training_passes(const GuideRun& run, int max_depth, int batch) {
    if (run.train_every <= 0 || max_depth <= 0) return 0;
    const bool late = run.guide && run.guide->steps >= kLateAfter;
    const int every = run.train_every * (late ? kLateEvery : 1);
    int n = 0;
    for (int i = 0; i < batch; ++i) n += ((run.pass + i) % every) == 0;
    return n;
}

// Recorded paths: every stride-th of the wavefront's paths, at most kRecordedPaths.
struct RecordPlan {
    int64_t paths;
    int stride, recorded, nslots;
};
RecordPlan record_plan(const GuidePassArgs& a) {
    RecordPlan r;
    r.paths = a.paths;
    r.stride = (int)std::max<int64_t>(1, (r.paths + kRecordedPaths - 1) / kRecordedPaths);
    r.recorded = (int)((r.paths + r.stride - 1) / r.stride);
    r.nslots = r.recorded * std::max(1, a.max_depth);
    return r;
}
}  // namespace

int guide_begin(const GuidePassArgs& a, int batch, GuideHook* gh, lfhd::GuideDev* gd) {
    *gh = GuideHook{};
    GuideRun& run = *a.run;
    lfh_guide* g = run.guide;
    if (!g) return 0;
    GuideBufs& b = a.buf ? *a.buf : run.buf;
    // lobes per hit-queue entry: the queues grow with multi-pass wavefronts
    const size_t lobe_bytes = (size_t)a.cap * kGuideLobeF4 * sizeof(float4);
    if (b.lobes.bytes < lobe_bytes) {
        if (hipStreamSynchronize(a.stream) != hipSuccess || b.lobes.alloc(lobe_bytes) != hipSuccess) return -1;
    }
    gh->enabled = 1;
    gh->sample = plain_warmup(run) ? 0 : 1;
    gh->lobes = b.lobes.as<const float4>();
    gh->lobe_stride = a.cap;
    gh->max_depth = a.max_depth;
    *gd = g->dev();
    if (training_passes(run, a.max_depth, batch) > 0) {
        const RecordPlan rp = record_plan(a);
        if (rp.nslots > b.nslots) return -1;
        gh->train = 1;
        gh->stride = rp.stride;
        gh->phase = (run.pass / run.train_every) % rp.stride;
        gh->rec = b.rec.as<float4>();
        if (pt_clear4(b.rec.as<float4>(), (int64_t)rp.nslots * 4, a.stream) != hipSuccess) return -1;
    }
    return 1;
}

int guide_end(const GuidePassArgs& a, const GuideHook& gh, int batch) {
    GuideRun& run = *a.run;
    lfh_guide* g = run.guide;
    int ntrain = training_passes(run, a.max_depth, batch);
    const bool warm_wavefront = plain_warmup(run);
    run.pass += batch;
    if (!g || !gh.train) return LFH_OK;
    GuideBufs& b = a.buf ? *a.buf : run.buf;
    const RecordPlan rp = record_plan(a);
    // collect, then train on the device-side sample count: no host round trip
    hipError_t e = lfhd::guide_collect(b.rec.as<float4>(), rp.nslots, a.max_depth, gh.stride, gh.phase, a.rad,
                                       (int)rp.paths, b.block_counts.as<int>(), b.block_offsets.as<int>(),
                                       b.total.as<int>(), b.samples.as<float4>(), kMaxSamples,
                                       b.block_wsum.as<float>(), b.total.as<float>() + 1,
                                       reinterpret_cast<unsigned long long*>(b.total.as<char>() + 8), a.stream);
    if (e != hipSuccess) return ghip(g, e, "guide: collect");
    // one Adam step per training pass on a strided batch of the wavefront's
    // samples; several while the guide is young
    if (warm_wavefront) ntrain = 1;   // one round of the warm-up steps on the BSDF-sampled passes' samples
    for (int t = 0; t < ntrain; ++t) {
        const int reps = warm_wavefront ? kWarmSteps : (g->steps < kWarmSteps ? kWarmRepeats : kLateRepeats);
        for (int r = 0; r < reps; ++r) {
            lfhd::GuideAdam ad = g->adam(true);
            e = lfhd::guide_train(g->dev(), b.samples.as<float4>(), b.total.as<int>(), (int)(g->steps & 0x7fffffff),
                                  kBatch, b.partial.as<float>(), ad, nullptr, nullptr, a.stream,
                                  b.total.as<float>() + 1, kWeightCap);
            if (e != hipSuccess) return ghip(g, e, "guide: train");
            g->steps++;
        }
    }
    return LFH_OK;
}

// Buffers guided wavefronts of up to `cap` paths and paths of up to
// max_depth vertices need; the counters are cleared when first allocated.
hipError_t guide_buffers(GuideBufs& b, int cap, int max_depth, hipStream_t s) {
    // records for the largest wavefront: kRecordedPaths paths x max_depth
    b.nslots = std::max(b.nslots, kRecordedPaths * (max_depth > 0 ? max_depth : 1));
    const int nblocks = lfhd::guide_collect_blocks(b.nslots);
    hipError_t e = ensure(b.lobes, (size_t)cap * kGuideLobeF4 * sizeof(float4));
    if (e == hipSuccess) e = ensure(b.rec, (size_t)b.nslots * 4 * sizeof(float4));
    if (e == hipSuccess) e = ensure(b.samples, (size_t)kMaxSamples * 4 * sizeof(float4));
    if (e == hipSuccess) e = ensure(b.block_counts, (size_t)(nblocks + 1) * sizeof(int));
    if (e == hipSuccess) e = ensure(b.block_offsets, (size_t)(nblocks + 1) * sizeof(int));
    if (e == hipSuccess) e = ensure(b.block_wsum, (size_t)(nblocks + 1) * sizeof(float));
    if (e == hipSuccess && !b.total.p) {
        e = b.total.alloc(64);   // int count, float weight sum, u64 running count at byte 8
        if (e == hipSuccess) e = hipMemsetAsync(b.total.p, 0, 64, s);
    }
    if (e == hipSuccess) e = ensure(b.partial, lfhd::guide_partial_floats(kBatch) * sizeof(float));
    return e;
}

size_t guide_buffer_bytes(int cap, int max_depth) {
    const size_t nslots = (size_t)kRecordedPaths * (max_depth > 0 ? max_depth : 1);
    const size_t nblocks = (size_t)lfhd::guide_collect_blocks((int)nslots) + 1;
    return (size_t)cap * kGuideLobeF4 * sizeof(float4) + nslots * 4 * sizeof(float4) +
           (size_t)kMaxSamples * 4 * sizeof(float4) + nblocks * 12 + 64 +
           lfhd::guide_partial_floats(kBatch) * sizeof(float);
}

hipError_t guide_attach(GuideRun& run, lfh_guide* g, int train_every, int cap, int max_depth, hipStream_t s) {
    run.guide = g;
    run.train_every = train_every;
    if (!g) return hipSuccess;
    hipError_t e = guide_buffers(run.buf, cap, max_depth, s);
    if (e != hipSuccess) run.guide = nullptr;
    return e;
}

int guide_samples_seen(const GuideRun& run, hipStream_t s, int64_t* n) {
    unsigned long long v = 0;
    if (run.buf.total.p) {
        if (hipMemcpyAsync(&v, run.buf.total.as<char>() + 8, sizeof v, hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            return LFH_E_HIP;
    }
    *n = (int64_t)v;
    return LFH_OK;
}


// The context's hooks: its path buffers and its one-tile wavefronts.
namespace {
GuidePassArgs ctx_args(lfh_ctx* c, int batch) {
    GuidePassArgs a;
    a.run = &c->grun;
    a.paths = (int64_t)c->rs.npaths * batch;
    a.cap = c->pass.cap;
    a.max_depth = c->opt.max_depth;
    a.rad = c->pass.rad.as<float4>();
    a.stream = c->stream;
    return a;
}
}  // namespace

int guide_pass_begin(lfh_ctx* c, int batch, GuideHook* gh, lfhd::GuideDev* gd) {
    return guide_begin(ctx_args(c, batch), batch, gh, gd);
}
int guide_pass_end(lfh_ctx* c, const GuideHook& gh, int batch) { return guide_end(ctx_args(c, batch), gh, batch); }
int guide_batch_limit(const lfh_ctx* c) { return guide_pass_limit(c->grun); }

}  // namespace lfh

extern "C" {

int lfh_guide_create(const float* scene_min3, const float* scene_max3, uint64_t seed, float learning_rate,
                     lfh_guide** out) {
    if (!scene_min3 || !scene_max3 || !out) return gfail(nullptr, LFH_E_INVALID, "lfh_guide_create: NULL argument");
    *out = nullptr;
    if (!(learning_rate > 0.0f)) return gfail(nullptr, LFH_E_INVALID, "lfh_guide_create: learning rate must be > 0");
    lfh_guide* g = new (std::nothrow) lfh_guide;
    if (!g) return gfail(nullptr, LFH_E_OOM, "lfh_guide_create: out of host memory");
    for (int a = 0; a < 3; ++a) {
        g->lo[a] = scene_min3[a];
        g->hi[a] = scene_max3[a];
    }
    g->lr = learning_rate;
    hipError_t e = hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = g->master.alloc(kGuideParams * sizeof(float));
    if (e == hipSuccess) e = g->m.alloc(kGuideParams * sizeof(float));
    if (e == hipSuccess) e = g->v.alloc(kGuideParams * sizeof(float));
    if (e == hipSuccess) e = g->wf.alloc(2 * lfhd::kGuideWImage * sizeof(uint16_t));
    if (e == hipSuccess) e = g->wb.alloc(4 * 4096 * sizeof(uint16_t));
    if (e == hipSuccess) e = g->bias.alloc(2 * lfhd::kGuideBImage * sizeof(float));
    if (e == hipSuccess) e = g->count.alloc(64);
    if (e == hipSuccess) e = g->grad.alloc((kGuideParams + 1) * sizeof(float));
    if (e == hipSuccess) e = g->loss.alloc(64);
    if (e == hipSuccess) e = hipMemsetAsync(g->wf.p, 0, g->wf.bytes, g->stream);
    if (e == hipSuccess) e = hipMemsetAsync(g->wb.p, 0, g->wb.bytes, g->stream);
    if (e == hipSuccess) e = hipMemsetAsync(g->bias.p, 0, g->bias.bytes, g->stream);
    if (e != hipSuccess) {
        int r = ghip(nullptr, e, "lfh_guide_create");
        lfh_guide_destroy(g);
        return r;
    }
    std::vector<float> p = init_params(seed);
    int r = load_params(g, p.data());
    if (r != LFH_OK) {
        lfh_guide_destroy(g);
        return r;
    }
    *out = g;
    return LFH_OK;
}

void lfh_guide_destroy(lfh_guide* g) {
    if (!g) return;
    if (g->stream) {
        (void)hipStreamSynchronize(g->stream);
        (void)hipStreamDestroy(g->stream);
    }
    delete g;
}

int lfh_guide_get_params(lfh_guide* g, float* params, int64_t* n_params) {
    if (!g || !n_params) return gfail(g, LFH_E_INVALID, "lfh_guide_get_params: NULL argument");
    if (!params) {
        *n_params = kGuideParams;
        return LFH_OK;
    }
    if (*n_params < kGuideParams) return gfail(g, LFH_E_INVALID, "lfh_guide_get_params: buffer too small");
    *n_params = kGuideParams;
    GCHK(g, hipMemcpyAsync(params, g->master.p, kGuideParams * sizeof(float), hipMemcpyDeviceToHost, g->stream),
         "lfh_guide_get_params");
    GCHK(g, hipStreamSynchronize(g->stream), "lfh_guide_get_params");
    return LFH_OK;
}

int lfh_guide_set_params(lfh_guide* g, const float* params, int64_t n_params) {
    if (!g || !params) return gfail(g, LFH_E_INVALID, "lfh_guide_set_params: NULL argument");
    if (n_params != kGuideParams) return gfail(g, LFH_E_INVALID, "lfh_guide_set_params: wrong parameter count");
    return load_params(g, params);
}

int lfh_guide_forward(lfh_guide* g, const float* inputs9, int64_t n, float* out40) {
    int r = check_batch(g, "lfh_guide_forward", inputs9, out40, n);
    if (r != LFH_OK || n == 0) return r;
    if ((r = forward_host(g, inputs9, n, true, false)) != LFH_OK) return r;
    GCHK(g, hipMemcpyAsync(out40, g->out.p, (size_t)n * kGuideOut * sizeof(float), hipMemcpyDeviceToHost, g->stream),
         "lfh_guide_forward");
    GCHK(g, hipStreamSynchronize(g->stream), "lfh_guide_forward");
    return LFH_OK;
}

int lfh_guide_forward_device(lfh_guide* g, const float* inputs9, int64_t n, float* out40, void* stream) {
    int r = check_batch(g, "lfh_guide_forward_device", inputs9, out40, n);
    if (r != LFH_OK || n == 0) return r;
    GCHK(g, hipStreamSynchronize(g->stream), "lfh_guide_forward_device");  // parameter uploads
    GCHK(g, lfhd::guide_forward_host(g->dev(), inputs9, (int)n, out40, nullptr, static_cast<hipStream_t>(stream)),
         "lfh_guide_forward_device");
    return LFH_OK;
}

int lfh_guide_sample(lfh_guide* g, const float* inputs9, const uint32_t* seeds, int64_t n, float* dirs3,
                     float* pdf) {
    int r = check_batch(g, "lfh_guide_sample", inputs9, seeds, n);
    if (r != LFH_OK || n == 0) return r;
    if (!dirs3 || !pdf) return gfail(g, LFH_E_INVALID, "lfh_guide_sample: NULL argument");
    if ((r = forward_host(g, inputs9, n, false, true)) != LFH_OK) return r;
    GCHK(g, ensure(g->out, (size_t)n * 16 * sizeof(float)), "lfh_guide_sample");
    float* d_seeds = g->out.as<float>();
    float* d_dirs = d_seeds + (size_t)n * 4;
    float* d_pdf = d_seeds + (size_t)n * 7;
    GCHK(g, hipMemcpyAsync(d_seeds, seeds, (size_t)n * 16, hipMemcpyHostToDevice, g->stream), "lfh_guide_sample");
    GCHK(g, lfhd::guide_sample(g->aux.as<float4>(), reinterpret_cast<const uint4*>(d_seeds), (int)n, d_dirs, d_pdf,
                               g->stream),
         "lfh_guide_sample");
    GCHK(g, hipMemcpyAsync(dirs3, d_dirs, (size_t)n * 12, hipMemcpyDeviceToHost, g->stream), "lfh_guide_sample");
    GCHK(g, hipMemcpyAsync(pdf, d_pdf, (size_t)n * 4, hipMemcpyDeviceToHost, g->stream), "lfh_guide_sample");
    GCHK(g, hipStreamSynchronize(g->stream), "lfh_guide_sample");
    return LFH_OK;
}

int lfh_guide_pdf(lfh_guide* g, const float* inputs9, const float* dirs3, int64_t n, float* pdf) {
    int r = check_batch(g, "lfh_guide_pdf", inputs9, dirs3, n);
    if (r != LFH_OK || n == 0) return r;
    if (!pdf) return gfail(g, LFH_E_INVALID, "lfh_guide_pdf: NULL argument");
    if ((r = forward_host(g, inputs9, n, false, true)) != LFH_OK) return r;
    GCHK(g, ensure(g->out, (size_t)n * 4 * sizeof(float)), "lfh_guide_pdf");
    float* d_dirs = g->out.as<float>();
    float* d_pdf = d_dirs + (size_t)n * 3;
    GCHK(g, hipMemcpyAsync(d_dirs, dirs3, (size_t)n * 12, hipMemcpyHostToDevice, g->stream), "lfh_guide_pdf");
    GCHK(g, lfhd::guide_pdf(g->aux.as<float4>(), d_dirs, (int)n, d_pdf, g->stream), "lfh_guide_pdf");
    GCHK(g, hipMemcpyAsync(pdf, d_pdf, (size_t)n * 4, hipMemcpyDeviceToHost, g->stream), "lfh_guide_pdf");
    GCHK(g, hipStreamSynchronize(g->stream), "lfh_guide_pdf");
    return LFH_OK;
}

int lfh_guide_loss_grad(lfh_guide* g, const float* inputs9, const float* dirs3, const float* weights, int64_t n,
                        float* loss, float* grads) {
    int r = check_batch(g, "lfh_guide_loss_grad", inputs9, dirs3, n);
    if (r != LFH_OK) return r;
    if (n > 0 && !weights) return gfail(g, LFH_E_INVALID, "lfh_guide_loss_grad: NULL argument");
    if (n > kBatch) return gfail(g, LFH_E_INVALID, "lfh_guide_loss_grad: at most 65536 samples per batch");
    if ((r = upload_batch(g, inputs9, dirs3, weights, n)) != LFH_OK) return r;
    return run_batch(g, n, false, loss, grads);
}

int lfh_guide_train_step(lfh_guide* g, const float* inputs9, const float* dirs3, const float* weights, int64_t n,
                         float* loss) {
    int r = check_batch(g, "lfh_guide_train_step", inputs9, dirs3, n);
    if (r != LFH_OK) return r;
    if (n > 0 && !weights) return gfail(g, LFH_E_INVALID, "lfh_guide_train_step: NULL argument");
    if (n > kBatch) return gfail(g, LFH_E_INVALID, "lfh_guide_train_step: at most 65536 samples per batch");
    if (n == 0) return LFH_OK;
    if ((r = upload_batch(g, inputs9, dirs3, weights, n)) != LFH_OK) return r;
    return run_batch(g, n, true, loss, nullptr);
}

int lfh_guide_steps(const lfh_guide* g, int64_t* steps) {
    if (!g || !steps) return gfail(nullptr, LFH_E_INVALID, "lfh_guide_steps: NULL argument");
    *steps = g->steps;
    return LFH_OK;
}

int lfh_set_guide(lfh_ctx* c, lfh_guide* g, int32_t train_every) {
    if (!c) return gfail(nullptr, LFH_E_INVALID, "lfh_set_guide: NULL ctx");
    if (train_every < 0) return gfail(g, LFH_E_INVALID, "lfh_set_guide: train_every must be >= 0");
    if (g && !c->opt.guiding)
        return gfail(g, LFH_E_STATE, "lfh_set_guide: the context was created without options.guiding");
    if (g && c->opt.estimator != 1)
        return gfail(g, LFH_E_STATE, "lfh_set_guide: guided sampling needs options.estimator = 1");
    if (hipStreamSynchronize(c->stream) != hipSuccess) return gfail(g, LFH_E_HIP, "lfh_set_guide: sync");
    hipError_t e = lfh::guide_attach(c->grun, g, train_every, c->pass.cap, c->opt.max_depth, c->stream);
    if (e != hipSuccess) return ghip(g, e, "lfh_set_guide");
    return LFH_OK;
}

int lfh_guide_samples_seen(const lfh_ctx* c, int64_t* n) {
    if (!c || !n) return gfail(nullptr, LFH_E_INVALID, "lfh_guide_samples_seen: NULL argument");
    if (lfh::guide_samples_seen(c->grun, c->stream, n) != LFH_OK)
        return gfail(nullptr, LFH_E_HIP, "lfh_guide_samples_seen: copy");
    return LFH_OK;
}

}  // extern "C"
