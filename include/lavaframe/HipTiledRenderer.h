// HipTiledRenderer.h — the MI355X path tracer behind LavaFrame's Renderer
// interface.  Added next to TiledRenderer.h in the reference tree (with this
// repository's include/ on the include path and liblfh.so linked); the one
// construction site, InitRenderer (LavaFrame/Main.cpp:88-97), then chooses
// it instead of TiledRenderer.
//
// Renderer's constructor only stores members and every GL upload happens in
// its virtual Init() (Renderer.cpp:24-43, 76-188), so overriding every
// virtual of Renderer (Renderer.h:71-117) keeps GL out of this class.  The
// methods forward to lfh::HipRendererT<Scene>, which reaches the GPU only
// through the C ABI of lfh.h and follows TiledRenderer's semantics
// (TiledRenderer.h:57-71).
#pragma once

#include <cstdint>
#include <string>

#include "Renderer.h"
#include "Scene.h"
#include "lfh_renderer.hpp"

namespace LavaFrame {

class HipTiledRenderer : public Renderer {
    lfh::HipRendererT<Scene> impl;

public:
    HipTiledRenderer(Scene* s, const std::string& shadersDirectory)
        : Renderer(s, shadersDirectory), impl(s, shadersDirectory) {}
    // Finish here, before ~Renderer (which would call its GL Finish when
    // initialized is still set).
    ~HipTiledRenderer() override { Finish(); }

    void Init() override {
        impl.Init();
        initialized = true;
    }
    void Finish() override {
        impl.Finish();
        initialized = false;
    }
    void Render() override { impl.Render(); }
    void Present() const override {}
    void Update(float secondsElapsed) override { impl.Update(secondsElapsed); }
    float GetProgress() const override { return impl.GetProgress(); }
    int GetSampleCount() const override { return impl.GetSampleCount(); }
    void GetOutputBuffer(unsigned char** data, int& w, int& h) override { impl.GetOutputBuffer(data, w, h); }
    void GetOutputBufferHDR(float** data, int& w, int& h) override { impl.GetOutputBufferHDR(data, w, h); }
    uint32_t SetViewport(int, int) override { return 0; }
    uint32_t Denoise() override { return 0; }

    // The preview engine's options (GlobalState.previewScale, useDofInPreview).
    void SetPreview(float scale, bool useDof) { impl.SetPreview(scale, useDof); }
};

}  // namespace LavaFrame
