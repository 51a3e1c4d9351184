// Thread-local error message for ctx-less calls; failure injection for tests.
#include <atomic>

#include "../../include/lfh.h"
#include "capi_common.h"

namespace lfh_internal {
static thread_local std::string g_thread_error;
void set_thread_error(const std::string& msg) { g_thread_error = msg; }
const char* thread_error() { return g_thread_error.c_str(); }

static std::atomic<int> g_armed[kFaultSites];
bool inject_fault(int site) {
    if (site <= 0 || site >= kFaultSites) return false;
    int n = g_armed[site].load();
    while (n > 0)
        if (g_armed[site].compare_exchange_weak(n, n - 1)) return true;
    return false;
}
}  // namespace lfh_internal

extern "C" int lfh_debug_fail_next(int32_t site, int32_t count) {
    if (site <= 0 || site >= lfh_internal::kFaultSites || count < 0) {
        lfh_internal::set_thread_error("lfh_debug_fail_next: bad site or count");
        return LFH_E_INVALID;
    }
    lfh_internal::g_armed[site].store(count);
    return LFH_OK;
}
