// Device context behind lfh_ctx: resident scene, render state, stream.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/lfh.h"
#include "capi_common.h"
#include "dev_common.h"
#include "pt_kernels.h"
#include "pt_types.h"

namespace lfh {

// RAII device buffer.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    hipError_t alloc(size_t n) {
        release();
        if (n == 0) n = 16;
        hipError_t e = hipMalloc(&p, n);
        if (e == hipSuccess) bytes = n;
        else p = nullptr;
        return e;
    }
    template <class T> T* as() const { return static_cast<T*>(p); }
};

// Host-side packed copy of a scene (built from an lfh_scene_desc).
struct PackedScene {
    std::vector<float> nodes;     // 16 floats / internal BVH2 node (host intermediate)
    std::vector<float> nodes4;    // 32 floats / BVH4 node (device layout)
    std::vector<float> insts4;    // insts with BLAS refs into nodes4 (device layout)
    uint32_t root4 = lfhd::kRefSentinel;
    int stack_need = 0;           // worst-case traversal stack entries of nodes4
    std::vector<float> tris;      // 12 floats / triangle slot
    std::vector<float> tri_box;   // 6 floats / triangle slot: its vertices' box (host)
    std::vector<float> insts;     // 16 floats / instance (inverse rows + meta)
    std::vector<float> xforms;    // 16 floats / instance (forward, column-major)
    std::vector<float> nmats;     // 12 floats / instance (normal matrix rows)
    std::vector<int32_t> vidx;    // 4 int / triangle slot
    std::vector<float> normals;   // 4 floats / vertex
    std::vector<float> materials; // 28 floats / material
    std::vector<float> lights;    // 28 floats / light (raw 16 + derived 12)
    std::vector<float> hdr_cols;  // 4 floats / texel
    std::vector<float> hdr_marg;  // 2 floats / row
    std::vector<float> hdr_cond;  // 2 floats / texel
    int hdr_w = 0, hdr_h = 0;
    bool hdr_tables_on_device = false;   // desc without HDR tables: built at upload
    std::vector<uint8_t> tex;     // RGBA8 texture layers
    int tex_w = 0, tex_h = 0, tex_layers = 0;
    std::vector<float> uvs;       // 2 floats / vertex, only with textures
    uint32_t root = lfhd::kRefSentinel;
    int64_t num_internal = 0, num_tris = 0, num_vertices = 0;
    int32_t num_instances = 0, num_materials = 0, num_lights = 0;
    int max_depth_levels = 0;
    // Node index maps for partial TLAS updates: flat index -> packed index.
    std::vector<int32_t> flat_to_packed;
    std::vector<int32_t> flat_lrleaf;   // cached LRLeaf (3 ints / flat node)
    int32_t top_level_index = 0;
    // Device BVH build (a desc without BVH nodes): inputs, then what the build
    // leaves for instance updates.  nodes4 / tris / insts4 stay empty on the
    // host; the build writes them on the device.
    bool device_build = false;
    std::vector<float> verts;           // 4 floats / vertex (x, y, z, u)
    std::vector<int64_t> mesh_offsets;  // num_meshes + 1 triangle offsets
    std::vector<int32_t> inst_mesh;     // mesh of each instance
    std::vector<uint32_t> mesh_ref4;    // BVH4 ref of each mesh's BLAS
    std::vector<float> mesh_box;        // 6 floats / mesh: BLAS root box
    int blas_need = 0;                  // deepest BLAS stack need
    int64_t blas_nodes4 = 0;            // BVH4 nodes of the BLAS part; the TLAS region follows
    int64_t tlas_cap4 = 0;              // BVH4 nodes reserved for the TLAS
};

// The host fields an instance update rewrites, saved before the update and
// put back when it is rejected, so a failed update leaves the scene (host and
// device) as it was.
struct InstanceSnapshot {
    std::vector<float> nodes, nodes4, insts, insts4, xforms, nmats, materials;
    std::vector<int32_t> flat_to_packed, flat_lrleaf;
    uint32_t root = 0, root4 = 0;
    int stack_need = 0;
    void save(const PackedScene& ps) {
        nodes = ps.nodes; nodes4 = ps.nodes4; insts = ps.insts; insts4 = ps.insts4;
        xforms = ps.xforms; nmats = ps.nmats; materials = ps.materials;
        flat_to_packed = ps.flat_to_packed; flat_lrleaf = ps.flat_lrleaf;
        root = ps.root; root4 = ps.root4; stack_need = ps.stack_need;
    }
    void restore(PackedScene* ps) {
        ps->nodes.swap(nodes); ps->nodes4.swap(nodes4); ps->insts.swap(insts); ps->insts4.swap(insts4);
        ps->xforms.swap(xforms); ps->nmats.swap(nmats); ps->materials.swap(materials);
        ps->flat_to_packed.swap(flat_to_packed); ps->flat_lrleaf.swap(flat_lrleaf);
        ps->root = root; ps->root4 = root4; ps->stack_need = stack_need;
    }
};

// Build the device layout from the reference flat arrays; "" on success.
std::string pack_scene(const lfh_scene_desc& d, PackedScene* out);
// BVH4 device nodes from the packed BVH2 (called by pack_scene and after
// instance updates); "" on success.
std::string widen_bvh(PackedScene* ps);
// Re-pack the instance-dependent parts after Renderer::Update-style edits.
std::string repack_instances(PackedScene* ps, const float* transforms, const float* tlas_nodes,
                             int64_t first_node, int64_t count, const float* materials);
// Instance records of a device-built scene: ps->insts with each instance's BLAS ref.
void device_insts4(PackedScene* ps);

struct DeviceScene {
    DevBuf nodes, tris, insts, xforms, nmats, vidx, normals, materials, lights, hdr_cols, hdr_marg, hdr_cond, tex, uvs;
    void release() {
        DevBuf* b[14] = {&nodes, &tris, &insts, &xforms, &nmats, &vidx, &normals, &materials, &lights,
                         &hdr_cols, &hdr_marg, &hdr_cond, &tex, &uvs};
        for (DevBuf* x : b) x->release();
    }
    size_t bytes() const {
        return nodes.bytes + tris.bytes + insts.bytes + xforms.bytes + nmats.bytes + vidx.bytes + normals.bytes +
               materials.bytes + lights.bytes + hdr_cols.bytes + hdr_marg.bytes + hdr_cond.bytes + tex.bytes +
               uvs.bytes;
    }
};

// Accumulation buffers for the whole frame.
struct RenderState {
    DevBuf accum, comp;   // float4 per pixel: running sum, Kahan compensation
    int npaths = 0;       // tile pixels
    int64_t npix = 0;
};

// Path state + queues of one wavefront pass for up to `cap` paths.
struct PassBufs {
    DevBuf rad, thr, absorb, rng, misc, nee;  // per path (nee: 3 float4)
    DevBuf rq0, rq1, hq, sq;                  // ray queues, hit queue, shadow queue
    DevBuf counters, work;
    DevBuf job;                               // int per path (multi-job passes)
    int cap = 0;
    // All or nothing: on any failure every buffer is released and cap is 0,
    // so the next caller that needs paths allocates again instead of
    // launching on a half-allocated set.
    hipError_t alloc(int n) {
        release();
        hipError_t e = alloc_all(n);
        if (e != hipSuccess) {
            release();
            return e;
        }
        cap = n;
        return hipSuccess;
    }
    void release() {
        DevBuf* b[13] = {&rad, &thr, &absorb, &rng, &misc, &nee, &rq0, &rq1, &hq, &sq, &counters, &work, &job};
        for (DevBuf* x : b) x->release();
        cap = 0;
    }
    size_t bytes() const {
        return rad.bytes + thr.bytes + absorb.bytes + rng.bytes + misc.bytes + nee.bytes + rq0.bytes + rq1.bytes +
               hq.bytes + sq.bytes + counters.bytes + work.bytes + job.bytes;
    }

private:
    hipError_t alloc_all(int n) {
        if (lfh_internal::inject_fault(lfh_internal::kFaultPathAlloc)) return hipErrorOutOfMemory;
        hipError_t e;
        const size_t f4 = 16;
        DevBuf* b[10] = {&rad, &thr, &absorb, &rng, &misc, &nee, &rq0, &rq1, &hq, &sq};
        const size_t m[10] = {1, 1, 1, 1, 1, 3, 2, 2, 2, 4};
        for (int i = 0; i < 10; ++i)
            if ((e = b[i]->alloc(f4 * m[i] * (size_t)n)) != hipSuccess) return e;
        if ((e = counters.alloc(64)) != hipSuccess) return e;
        if ((e = work.alloc(lfhd::kWorkWords * 8)) != hipSuccess) return e;
        if ((e = hipMemset(work.p, 0, lfhd::kWorkWords * 8)) != hipSuccess) return e;
        if ((e = job.alloc(4 * (size_t)n)) != hipSuccess) return e;
        return hipSuccess;
    }
};
// Per-context buffers of a learned guide attached with lfh_set_guide.
struct GuideBufs {
    DevBuf lobes;                  // kGuideLobeF4 float4 per hit-queue entry
    DevBuf rec;                    // training records, 4 float4 per slot
    DevBuf samples;                // compacted training samples, 4 float4 each
    DevBuf block_counts, block_offsets, block_wsum, partial;
    DevBuf total;                  // int sample count, then float weight sum
    int nslots = 0;                // record slots: recorded paths x max_depth, largest wavefront
};

// The per-depth launch chain of a wavefront (trace, shade, shadow, advance for
// every depth) captured once as a HIP graph and replayed: its kernels read the
// queue counters on the device, so the chain is the same for every wavefront
// of one path capacity, scene, job table and depth.  `key` lists what the
// chain's arguments depend on; a different key recaptures.
struct BounceGraph {
    hipGraphExec_t exec = nullptr;
    std::vector<uintptr_t> key;
    BounceGraph() = default;
    BounceGraph(const BounceGraph&) = delete;
    BounceGraph& operator=(const BounceGraph&) = delete;
    ~BounceGraph() { reset(); }
    void reset() {
        if (exec) (void)hipGraphExecDestroy(exec);
        exec = nullptr;
        key.clear();
    }
};
// Bounce graphs are on unless LFH_BOUNCE_GRAPHS=0 (read once).
bool bounce_graphs_enabled();

// The stream, events and spill stack of a bounce chain's shadow lane
// (ShadowLane).  Created with the first render; the stream has the default
// priority.
struct LaneRes {
    hipStream_t s = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    DevBuf spill;
    LaneRes() = default;
    LaneRes(const LaneRes&) = delete;
    LaneRes& operator=(const LaneRes&) = delete;
    ~LaneRes() { release(); }
    void release() {
        for (hipEvent_t& e : ev) {
            if (e) (void)hipEventDestroy(e);
            e = nullptr;
        }
        if (s) (void)hipStreamDestroy(s);
        s = nullptr;
        spill.release();
    }
    // spill_bytes: the size of the main stream's spill stack
    hipError_t init(size_t spill_bytes) {
        hipError_t e = hipSuccess;
        if (!s) e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
        for (hipEvent_t& v : ev)
            if (e == hipSuccess && !v) e = hipEventCreateWithFlags(&v, hipEventDisableTiming);
        if (e == hipSuccess && spill.bytes < spill_bytes) e = spill.alloc(spill_bytes);
        if (e != hipSuccess) release();
        return e;
    }
    bool ready() const { return s != nullptr; }
    ShadowLane lane() const { return ShadowLane{s, spill.as<uint32_t>(), ev[0], ev[1]}; }
};
// The shadow lane is used unless LFH_SHADOW_LANE=0 (read once).
bool shadow_lane_enabled();

// One owner's guided rendering (a context, or a resident scene of the render
// server): the guide, its training cadence, the passes rendered with it and
// the buffers its wavefronts use.
struct GuideRun {
    struct lfh_guide* guide = nullptr;
    int train_every = 0;   // every train_every-th pass records training samples; 0: frozen
    int pass = 0;          // passes rendered with the guide attached
    GuideBufs buf;
};

// Uploads the packed scene; for a desc without BVH nodes it also runs the
// device BVH build (on failure *err, if given, says why).
hipError_t upload_scene(PackedScene& ps, DeviceScene* ds, hipStream_t s, std::string* err = nullptr);
// Device BVH build: BLAS of every mesh, then the TLAS; "" on success.
std::string device_build(PackedScene* ps, DeviceScene* ds, hipStream_t s);
// TLAS rebuild after instance edits of a device-built scene; "" on success.
std::string device_build_tlas(PackedScene* ps, DeviceScene* ds, hipStream_t s);
lfhd::TraceScene trace_view(const PackedScene& ps, const DeviceScene& ds);
lfhd::ShadeScene shade_view(const PackedScene& ps, const DeviceScene& ds, const lfh_options& o);

}  // namespace lfh

struct lfh_ctx {
    std::string err;
    lfh_options opt{};
    lfh::PackedScene ps;
    lfh::DeviceScene ds;
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    int device = 0;
    // traversal scratch
    lfh::DevBuf spill, work, errflag;
    lfh::DevBuf rflag;            // traversal-overflow flag of the path-tracing passes
    lfh::DevBuf preview;          // float4 per preview pixel (lfh_render_preview)
    int grid_blocks = 0;
    float cull = 1.0f + 1.0f / 1024.0f;
    // host staging for the host-array entry points
    lfh::DevBuf stage_in, stage_out;
    // render state
    lfh::RenderState rs;
    lfh::PassBufs pass;
    lfh::DevBuf out;   // HDR output staging
    int pt_grid = 0;
    lfh_camera cam{};
    bool has_cam = false;
    // TiledRenderer counters (TiledRenderer.cpp:48-66, 471-501)
    int tileX = -1, tileY = 0, numTilesX = 1, numTilesY = 1, tileW = 0, tileH = 0;
    int sampleCounter = 1, frameCounter = 1;
    int frameOffset = 0;           // added to the frame number that seeds the RNG
    // guiding
    lfh::GuideRun grun;
    lfh::BounceGraph bgraph;       // the unguided bounce chain of render_tile
    lfh::LaneRes lane;             // render_tile's shadow lane
    uint64_t scene_version = 0;    // bumped by instance updates (the chain's scene view changes)
};
