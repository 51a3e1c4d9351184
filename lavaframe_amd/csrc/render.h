// Path-tracing renderer behind lfh_update / lfh_render (TiledRenderer
// semantics) and the per-stage parity entry points.
#pragma once
#include "ctx.h"
#include "dev_guide.h"
#include "pt_kernels.h"

#include <functional>

namespace lfh {
int render_init(lfh_ctx* c);
// Replays g's captured launch chain on s, capturing `launch` first when the key differs.
hipError_t run_graph(BounceGraph& g, const std::vector<uintptr_t>& key, hipStream_t s,
                     const std::function<hipError_t()>& launch);
void render_release(lfh_ctx* c);
// Device views of a pass's path buffers.
PathBufs view(const PassBufs& b);
// Learned-guide hooks of one pass: begin fills the hook and the weights view
// (returns 1 when a guide is attached, 0 if not, -1 on a HIP error); end
// collects the pass's training samples and runs the optimisation steps.
int guide_pass_begin(lfh_ctx* c, int batch, GuideHook* gh, lfhd::GuideDev* gd);
int guide_pass_end(lfh_ctx* c, const GuideHook& gh, int batch);
int guide_batch_limit(const lfh_ctx* c);
// The same hooks for any owner of a guided wavefront (the server's scenes).
struct GuidePassArgs {
    GuideRun* run;
    int64_t paths;        // paths of the wavefront, all passes
    int cap;              // path capacity of its queues (lobe stride)
    int max_depth;
    const float4* rad;    // final path radiance, after accumulation
    hipStream_t stream;
    GuideBufs* buf = nullptr;   // the wavefront buffers (NULL: run->buf)
};
int guide_begin(const GuidePassArgs& a, int batch, GuideHook* gh, lfhd::GuideDev* gd);
int guide_end(const GuidePassArgs& a, const GuideHook& gh, int batch);
int guide_pass_limit(const GuideRun& run);
hipError_t guide_attach(GuideRun& run, struct lfh_guide* g, int train_every, int cap, int max_depth, hipStream_t s);
hipError_t guide_buffers(GuideBufs& b, int cap, int max_depth, hipStream_t s);
size_t guide_buffer_bytes(int cap, int max_depth);
int guide_samples_seen(const GuideRun& run, hipStream_t s, int64_t* n);
// LFH_E_UNSUPPORTED (and the flag cleared) when a pass overflowed a traversal stack.
int check_render_flag(lfh_ctx* c);
// Device address of the no-alloc entry point's scratch array on `device`,
// taken once per device (called by lfh_create with the device current).
hipError_t noalloc_setup(int device);
}  // namespace lfh
