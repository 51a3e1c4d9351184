// Compile check of include/lavaframe/HipTiledRenderer.h against the
// reference's own headers (test infrastructure: built by oracle/Makefile
// into _ref/, never linked into the product).  It instantiates
// lfh::HipRendererT over LavaFrame::Scene and derives HipTiledRenderer from
// LavaFrame::Renderer, so every Scene member, Renderer virtual and
// RenderOptions field the adapter uses must exist with the type it assumes.
#include "Scene.h"
#include "lavaframe/HipTiledRenderer.h"

template class lfh::HipRendererT<LavaFrame::Scene>;

LavaFrame::Renderer* lfh_adapter_make(LavaFrame::Scene* scene, const std::string& shadersDirectory) {
    return new LavaFrame::HipTiledRenderer(scene, shadersDirectory);
}
