// Error plumbing shared by the C ABI translation units.
#pragma once
#include <string>

namespace lfh_internal {
void set_thread_error(const std::string& msg);
const char* thread_error();

// Failure injection for tests of the error paths (lfh_debug_fail_next):
// returns true, and uses up one armed failure, when a failure is armed for
// `site`.
enum FaultSite { kFaultPathAlloc = 1, kFaultSceneUpload = 2, kFaultTlasBuild = 3, kFaultSites = 4 };
bool inject_fault(int site);
}  // namespace lfh_internal
