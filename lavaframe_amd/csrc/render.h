// Path-tracing renderer behind lfh_update / lfh_render (TiledRenderer
// semantics) and the per-stage parity entry points.
#pragma once
#include "ctx.h"
#include "pt_kernels.h"

namespace lfh {
int render_init(lfh_ctx* c);
void render_release(lfh_ctx* c);
// Device views of a pass's path buffers.
PathBufs view(const PassBufs& b);
}  // namespace lfh
