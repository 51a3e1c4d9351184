// lfh_renderer.hpp — C++ drop-in for LavaFrame's Renderer / TiledRenderer.
//
// HipRendererT<SceneT> has the method set and semantics of the abstract
// Renderer (LavaFrame/Renderer.h:71-117) as implemented by TiledRenderer
// (LavaFrame/TiledRenderer.h:57-71), and reaches HIP only through the C ABI
// of lfh.h.  SceneT is any type with the reference Scene's public members
// (LavaFrame/Scene.h:67-101): the reference's LavaFrame::Scene itself, or
// this build's lfh::Scene (lavaframe_amd/csrc/host_scene.h).
//
//   Init()                 Renderer::Init + TiledRenderer::Init: uploads the
//                          flattened scene (Renderer.cpp:76-188)
//   Update(secondsElapsed) TiledRenderer::Update: re-uploads instances when
//                          scene->instancesModified, resets on camera motion,
//                          else advances tile / sample / frame counters
//   Render()               TiledRenderer::Render: one tile, one sample/pixel
//   GetSampleCount(), GetProgress()
//   GetOutputBufferHDR()   the post-processed image (postprocess.glsl: accum /
//                          samples, tonemap, chromatic aberration, vignette),
//                          RGB float, bottom-up rows; allocates with new[] like
//                          the reference (the caller frees with delete[])
//   GetOutputBuffer()      its RGB8 read-back, round(255 clamp(c))
//   Present(), SetViewport(), Denoise()  headless no-ops
#pragma once

#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "lfh.h"

namespace lfh {

template <class SceneT>
class HipRendererT {
public:
    HipRendererT(SceneT* scene, const std::string& shadersDirectory = std::string())
        : scene_(scene), shadersDirectory_(shadersDirectory) {}
    virtual ~HipRendererT() { Finish(); }
    HipRendererT(const HipRendererT&) = delete;
    HipRendererT& operator=(const HipRendererT&) = delete;

    // Build the BVHs on the device at Init (from the scene's meshes and
    // instances) instead of uploading the scene's own BVH arrays.
    void SetDeviceBvh(bool on) { device_bvh_ = on; }
    // GlobalState.previewScale and useDofInPreview (GlobalState.h, Main.cpp:525-527).
    void SetPreview(float scale, bool use_dof) {
        preview_scale_ = scale;
        preview_dof_ = use_dof;
    }
    // The last preview frame drawn by Render() while the camera moved:
    // linear RGB, rows bottom to top (what Present shows after the post pass).
    void GetPreviewBufferHDR(float** data, int& w, int& h) {
        *data = new float[preview_.size()];
        std::memcpy(*data, preview_.data(), preview_.size() * sizeof(float));
        w = preview_w_;
        h = preview_h_;
    }

    virtual void Init() {
        if (ctx_) return;
        if (!scene_) throw std::runtime_error("HipRenderer: no scene");
        lfh_scene_desc d = describe();
        lfh_options o = options();
        int r = lfh_create(&d, &o, &ctx_);
        if (r != LFH_OK) throw std::runtime_error(std::string("HipRenderer::Init: ") + lfh_last_error(nullptr));
        width_ = o.width;
        height_ = o.height;
    }

    virtual void Finish() {
        if (ctx_) lfh_destroy(ctx_);
        ctx_ = nullptr;
    }

    virtual void Update(float /*secondsElapsed*/) {
        if (!ctx_) return;
        bool reset = scene_->camera->isMoving || scene_->instancesModified;
        if (scene_->instancesModified && device_bvh_) {
            // transforms and materials; the TLAS is rebuilt on the device
            check(lfh_update_instances(ctx_, reinterpret_cast<const float*>(scene_->transforms.data()),
                                       reinterpret_cast<const float*>(scene_->materials.data()), nullptr, 0, 0),
                  "Update");
            scene_->instancesModified = false;
        } else if (scene_->instancesModified) {
            // Renderer::Update: transforms, materials and the TLAS node range
            const int64_t top = scene_->bvhTranslator.topLevelIndex;
            const int64_t count = (int64_t)scene_->bvhTranslator.nodes.size() - top;
            check(lfh_update_instances(ctx_, reinterpret_cast<const float*>(scene_->transforms.data()),
                                       reinterpret_cast<const float*>(scene_->materials.data()),
                                       reinterpret_cast<const float*>(&scene_->bvhTranslator.nodes[top]), top, count),
                  "Update");
            scene_->instancesModified = false;  // the reference clears it in Render
        }
        lfh_camera c = camera();
        check(lfh_update(ctx_, &c, reset ? 1 : 0), "Update");
        moving_ = reset;
    }

    virtual void Render() {
        if (!ctx_) return;
        if (moving_) render_preview();
        else check(lfh_render(ctx_), "Render");
    }

    virtual void Present() const {}

    virtual float GetProgress() const {
        float p = 0.f;
        if (ctx_) lfh_get_progress(ctx_, &p);
        return p;
    }

    virtual int GetSampleCount() const {
        int32_t n = 0;
        if (ctx_) lfh_get_sample_count(ctx_, &n);
        return n;
    }

    // The post-processed image (postprocess.glsl with the scene's tonemap,
    // chromatic aberration and vignette options) as RGB float.
    virtual void GetOutputBufferHDR(float** data, int& w, int& h) {
        w = width_;
        h = height_;
        *data = new float[(size_t)w * h * 3];
        lfh_postprocess pp = postprocess();
        check(lfh_get_output_post(ctx_, &pp, 0.0f, *data, w, h), "GetOutputBufferHDR");
    }

    // Its RGB8 read-back.
    virtual void GetOutputBuffer(unsigned char** data, int& w, int& h) {
        w = width_;
        h = height_;
        *data = new unsigned char[(size_t)w * h * 3];
        lfh_postprocess pp = postprocess();
        check(lfh_get_output_ldr(ctx_, &pp, 0.0f, *data, w, h), "GetOutputBuffer");
    }

    lfh_postprocess postprocess() const {
        const auto& r = scene_->renderOptions;
        lfh_postprocess p;
        p.tonemap_index = r.tonemapIndex;
        p.use_ca = r.useCA ? 1 : 0;
        p.use_ca_distortion = r.useCADistortion ? 1 : 0;
        p.ca_distance = r.caDistance;
        p.ca_p1 = r.caP1;
        p.ca_p2 = r.caP2;
        p.ca_p3 = r.caP3;
        p.use_vignette = r.useVignette ? 1 : 0;
        p.vignette_intensity = r.vignetteIntensity;
        p.vignette_power = r.vignettePower;
        return p;
    }

    virtual uint32_t SetViewport(int, int) { return 0; }
    virtual uint32_t Denoise() { return 0; }

    lfh_ctx* context() const { return ctx_; }

protected:
    lfh_scene_desc describe() const {
        const SceneT& s = *scene_;
        lfh_scene_desc d;
        std::memset(&d, 0, sizeof d);
        d.nodes = reinterpret_cast<const float*>(s.bvhTranslator.nodes.data());
        d.num_nodes = (int64_t)s.bvhTranslator.nodes.size();
        d.top_level_index = s.bvhTranslator.topLevelIndex;
        d.vert_indices = reinterpret_cast<const int32_t*>(s.vertIndices.data());
        d.num_tris = (int64_t)s.vertIndices.size();
        d.vertices_uvx = reinterpret_cast<const float*>(s.verticesUVX.data());
        d.normals_uvy = reinterpret_cast<const float*>(s.normalsUVY.data());
        d.num_vertices = (int64_t)s.verticesUVX.size();
        d.materials = reinterpret_cast<const float*>(s.materials.data());
        d.num_materials = (int32_t)s.materials.size();
        d.transforms = reinterpret_cast<const float*>(s.transforms.data());
        d.num_instances = (int32_t)s.transforms.size();
        d.lights = reinterpret_cast<const float*>(s.lights.data());
        d.num_lights = (int32_t)s.lights.size();
        if (s.hdrData && s.renderOptions.useEnvMap) {
            d.hdr_width = s.hdrData->width;
            d.hdr_height = s.hdrData->height;
            d.hdr_cols = hdr_cols(*s.hdrData);
            d.hdr_marginal = reinterpret_cast<const float*>(hdr_ptr(s.hdrData->marginalDistData));
            d.hdr_conditional = reinterpret_cast<const float*>(hdr_ptr(s.hdrData->conditionalDistData));
        }
        if (!s.textures.empty()) {   // textureMapsArray (Scene.cpp:223-230)
            d.textures = reinterpret_cast<const uint8_t*>(s.textureMapsArray.data());
            d.tex_width = s.texWidth;
            d.tex_height = s.texHeight;
            d.num_textures = (int32_t)s.textures.size();
        }
        if (device_bvh_) {
            // mesh triangle ranges follow vertIndices' mesh order (Scene.cpp:198-214)
            mesh_offsets_.assign(1, 0);
            for (const auto* m : s.meshes) mesh_offsets_.push_back(mesh_offsets_.back() + (int64_t)(m->verticesUVX.size() / 3));
            inst_mesh_.clear();
            inst_mat_.clear();
            for (const auto& mi : s.meshInstances) {
                inst_mesh_.push_back(mi.meshID);
                inst_mat_.push_back(mi.materialID);
            }
            d.nodes = nullptr;
            d.num_nodes = 0;
            d.mesh_tri_offsets = mesh_offsets_.data();
            d.num_meshes = (int32_t)s.meshes.size();
            d.instance_meshes = inst_mesh_.data();
            d.instance_materials = inst_mat_.data();
        }
        return d;
    }

    lfh_options options() const {
        const auto& r = scene_->renderOptions;
        lfh_options o;
        std::memset(&o, 0, sizeof o);
        o.width = r.resolution.x;
        o.height = r.resolution.y;
        o.tile_width = r.tileWidth;
        o.tile_height = r.tileHeight;
        o.max_depth = r.maxDepth;
        o.enable_rr = r.enableRR ? 1 : 0;
        o.rr_depth = r.RRDepth;
        o.use_env_map = (r.useEnvMap && scene_->hdrData) ? 1 : 0;
        o.use_constant_bg = r.useConstantBg ? 1 : 0;
        o.bg_color[0] = r.bgColor.x;
        o.bg_color[1] = r.bgColor.y;
        o.bg_color[2] = r.bgColor.z;
        o.hdr_multiplier = r.hdrMultiplier;
        o.kahan = 1;
        return o;
    }

    lfh_camera camera() const {
        const auto& k = *scene_->camera;
        lfh_camera c;
        c.position[0] = k.position.x; c.position[1] = k.position.y; c.position[2] = k.position.z;
        c.right[0] = k.right.x; c.right[1] = k.right.y; c.right[2] = k.right.z;
        c.up[0] = k.up.x; c.up[1] = k.up.y; c.up[2] = k.up.z;
        c.forward[0] = k.forward.x; c.forward[1] = k.forward.y; c.forward[2] = k.forward.z;
        c.fov = k.fov;
        c.focal_dist = k.focalDist;
        c.aperture = k.aperture;
        return c;
    }

    void check(int r, const char* what) const {
        if (r != LFH_OK) throw std::runtime_error(std::string("HipRenderer::") + what + ": " + lfh_last_error(ctx_));
    }

    // HDR data is a raw array in the reference (float* / Vec2*) and a vector here.
    template <class T> static const T* hdr_ptr(const T* p) { return p; }
    template <class V> static const typename V::value_type* hdr_ptr(const V& v) { return v.data(); }
    template <class H> static const float* hdr_cols(const H& h) { return hdr_ptr(h.cols); }

    // Render() while the camera moves: the preview engine's frame
    // (TiledRenderer.cpp:327-333) instead of a path-traced tile.
    void render_preview() {
        int32_t w = (int32_t)((float)width_ * preview_scale_), h = (int32_t)((float)height_ * preview_scale_);
        preview_.resize((size_t)(w > 0 ? w : 0) * (h > 0 ? h : 0) * 3);
        check(lfh_render_preview(ctx_, preview_scale_, preview_dof_ ? 1 : 0, preview_.data(), &w, &h), "Render");
        preview_w_ = w;
        preview_h_ = h;
    }

    SceneT* scene_;
    std::string shadersDirectory_;
    lfh_ctx* ctx_ = nullptr;
    int width_ = 0, height_ = 0;
    bool device_bvh_ = false;
    bool moving_ = false;
    float preview_scale_ = 1.0f;
    bool preview_dof_ = false;
    std::vector<float> preview_;
    int preview_w_ = 0, preview_h_ = 0;
    mutable std::vector<int64_t> mesh_offsets_;
    mutable std::vector<int32_t> inst_mesh_, inst_mat_;
};

}  // namespace lfh
