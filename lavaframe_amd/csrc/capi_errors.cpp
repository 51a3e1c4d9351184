// Thread-local error message for ctx-less calls.
#include "capi_common.h"

namespace lfh_internal {
static thread_local std::string g_thread_error;
void set_thread_error(const std::string& msg) { g_thread_error = msg; }
const char* thread_error() { return g_thread_error.c_str(); }
}  // namespace lfh_internal
