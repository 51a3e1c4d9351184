// Path-tracing renderer behind lfh_update / lfh_render (TiledRenderer
// semantics) and the per-stage parity entry points.
#pragma once
#include "ctx.h"
#include "dev_guide.h"
#include "pt_kernels.h"

namespace lfh {
int render_init(lfh_ctx* c);
void render_release(lfh_ctx* c);
// Device views of a pass's path buffers.
PathBufs view(const PassBufs& b);
// Learned-guide hooks of one pass: begin fills the hook and the weights view
// (returns 1 when a guide is attached, 0 if not, -1 on a HIP error); end
// collects the pass's training samples and runs the optimisation steps.
int guide_pass_begin(lfh_ctx* c, int batch, GuideHook* gh, lfhd::GuideDev* gd);
int guide_pass_end(lfh_ctx* c, const GuideHook& gh, int batch);
int guide_batch_limit(const lfh_ctx* c);
// LFH_E_UNSUPPORTED (and the flag cleared) when a pass overflowed a traversal stack.
int check_render_flag(lfh_ctx* c);
}  // namespace lfh
