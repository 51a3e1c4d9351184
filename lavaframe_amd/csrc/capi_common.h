// Error plumbing shared by the C ABI translation units.
#pragma once
#include <string>

namespace lfh_internal {
void set_thread_error(const std::string& msg);
const char* thread_error();
}  // namespace lfh_internal
