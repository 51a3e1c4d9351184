// lfh_path_trace_rays_device_noalloc: lfh_path_trace_rays on device arrays
// with no device allocation in the call, for callers that cannot allocate
// (code under HIP graph capture, tight serving loops).
//
// The path state and queues of the wavefront (what lfh_path_trace_rays
// allocates per call as PassBufs) live in `scratch`, a file-scope device
// array of kScratchBytes that every GPU the library runs on holds until the
// process ends.  Its per-device address is taken once, with the device
// current, when the first context on that device is created
// (noalloc_setup, called by lfh_create), and only that address is handed to
// kernels.  Every call on one GPU shares that GPU's array: calls that overlap
// on one GPU (two streams, or two host threads) race.  The kernels are the
// ones lfh_path_trace_rays launches, so the radiance is bitwise the same.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "capi_common.h"
#include "ctx.h"
#include "pt_kernels.h"
#include "render.h"

namespace {

constexpr int64_t kNoAllocPaths = (int64_t)1 << 16;   // paths per call, at most
// per path: rad, thr, absorb, rng, misc (16 B each), nee (48), two ray queues
// (32 each), hit queue (32), shadow queue (64), job id (4)
constexpr size_t kAlign = 256;
constexpr size_t align_up(size_t x) { return (x + kAlign - 1) / kAlign * kAlign; }
constexpr size_t kF4 = (size_t)kNoAllocPaths * 16;
constexpr size_t kCountersBytes = 64;
constexpr size_t kWorkBytes = align_up((size_t)lfhd::kWorkWords * 8);
constexpr size_t kScratchBytes = kF4 * (5 + 3 + 2 + 2 + 2 + 4) + align_up((size_t)kNoAllocPaths * 4) +
                                 kCountersBytes * 4 + kWorkBytes;   // 18.9 MiB

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};

int fail(lfh_ctx* c, int code, const std::string& m) {
    if (c) c->err = m;
    lfh_internal::set_thread_error(m);
    return code;
}

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace lfh {

// Once per device, with that device current: the device address of this
// device's copy of `scratch`.  The array's name is a host-side stand-in and is
// never passed to a kernel or a copy.
hipError_t noalloc_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_ptr[device]) return hipSuccess;
    int cur = -1;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess) return e;
    if (cur != device && (e = hipSetDevice(device)) != hipSuccess) return e;
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != device) (void)hipSetDevice(cur);
    if (e != hipSuccess) return e;
    g_ptr[device] = p;
    return hipSuccess;
}

}  // namespace lfh

extern "C" int lfh_path_trace_rays_device_noalloc(lfh_ctx* c, const float* d_rays6, const uint32_t* d_seeds, int64_t n,
                                                  float* d_radiance) {
    const char* fn = "lfh_path_trace_rays_device_noalloc";
    if (!c) return fail(c, LFH_E_INVALID, std::string(fn) + ": NULL ctx");
    if (n < 0) return fail(c, LFH_E_INVALID, std::string(fn) + ": negative query count");
    if (n > kNoAllocPaths)
        return fail(c, LFH_E_INVALID, std::string(fn) + ": " + std::to_string(n) + " queries exceed the maximum of " +
                                          std::to_string(kNoAllocPaths) + " per call");
    if (n == 0) return LFH_OK;
    if (!d_rays6 || !d_seeds || !d_radiance) return fail(c, LFH_E_INVALID, std::string(fn) + ": NULL array");
    unsigned char* base = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        if (c->device >= 0 && c->device < kMaxDevices) base = static_cast<unsigned char*>(g_ptr[c->device]);
    }
    if (!base) return fail(c, LFH_E_STATE, std::string(fn) + ": scratch address of this device was not set up");
    // carve the path state out of the device's scratch (capacity kNoAllocPaths)
    size_t off = 0;
    auto take = [&](size_t bytes) {
        unsigned char* p = base + off;
        off += align_up(bytes);
        return p;
    };
    PathBufs pb;
    pb.rad = reinterpret_cast<float4*>(take(kF4));
    pb.thr = reinterpret_cast<float4*>(take(kF4));
    pb.absorb = reinterpret_cast<float4*>(take(kF4));
    pb.rng = reinterpret_cast<uint4*>(take(kF4));
    pb.misc = reinterpret_cast<float4*>(take(kF4));
    pb.nee = reinterpret_cast<float4*>(take(3 * kF4));
    float4* rq0 = reinterpret_cast<float4*>(take(2 * kF4));
    float4* rq1 = reinterpret_cast<float4*>(take(2 * kF4));
    float4* hq = reinterpret_cast<float4*>(take(2 * kF4));
    float4* sq = reinterpret_cast<float4*>(take(4 * kF4));
    pb.job = reinterpret_cast<int*>(take((size_t)kNoAllocPaths * 4));
    uint32_t* cnt = reinterpret_cast<uint32_t*>(take(kCountersBytes));
    unsigned long long* work = reinterpret_cast<unsigned long long*>(take(kWorkBytes));
    if (off > kScratchBytes) return fail(c, LFH_E_INVALID, std::string(fn) + ": scratch layout exceeds its size");
    const int cap = (int)kNoAllocPaths;
    hipError_t e = pt_reset(cnt, work, c->stream);   // counters start at zero (a kernel, no memset)
    if (e == hipSuccess)
        e = pt_generate_rays(d_rays6, reinterpret_cast<const uint4*>(d_seeds), (int)n, pb, rq0, cnt, c->stream);
    lfhd::ShadeScene sc = lfh::shade_view(c->ps, c->ds, c->opt);
    GuideHook gh{};
    for (int depth = 0; e == hipSuccess && depth < c->opt.max_depth; ++depth) {
        const int cur = depth & 1;
        float4* q[2] = {rq0, rq1};
        e = pt_bounce(sc, depth, cap, q[cur], cnt + cur, q[cur ^ 1], cnt + (cur ^ 1), hq, sq, cnt + 2, pb,
                      c->spill.as<uint32_t>(), work, c->rflag.as<int>(), c->cull, c->pt_grid, gh, nullptr, c->stream);
    }
    if (e == hipSuccess) e = pt_collect(pb, (int)n, d_radiance, c->stream);
    if (e != hipSuccess)
        return fail(c, e == hipErrorOutOfMemory ? LFH_E_OOM : LFH_E_HIP, std::string(fn) + ": " + hipGetErrorString(e));
    return LFH_OK;
}
