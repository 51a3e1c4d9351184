// Entry points of include/lfh.h not implemented in this translation unit set yet.
#include "../../include/lfh.h"
#include "capi_common.h"
#include "ctx.h"
#include "render.h"

static int unsupported(const char* fn) {
    lfh_internal::set_thread_error(std::string(fn) + ": not available in this build");
    return LFH_E_UNSUPPORTED;
}

extern "C" {
int lfh_guide_create(const float*, const float*, uint64_t, float, lfh_guide**) { return unsupported("lfh_guide_create"); }
void lfh_guide_destroy(lfh_guide*) {}
int lfh_guide_get_params(lfh_guide*, float*, int64_t*) { return unsupported("lfh_guide_get_params"); }
int lfh_guide_set_params(lfh_guide*, const float*, int64_t) { return unsupported("lfh_guide_set_params"); }
int lfh_guide_forward(lfh_guide*, const float*, int64_t, float*) { return unsupported("lfh_guide_forward"); }
int lfh_guide_loss_grad(lfh_guide*, const float*, const float*, const float*, int64_t, float*, float*) {
    return unsupported("lfh_guide_loss_grad");
}
int lfh_guide_train_step(lfh_guide*, const float*, const float*, const float*, int64_t, float*) {
    return unsupported("lfh_guide_train_step");
}
int lfh_set_guide(lfh_ctx*, lfh_guide*, int32_t) { return unsupported("lfh_set_guide"); }
}
