// Test-infrastructure harness around the reference's own C++ scene pipeline.
//
// ORACLE ONLY: this file is compiled together with the reference's unmodified
// sources (LavaFrame/Scene.cpp, Loader.cpp, Mesh.cpp, Camera.cpp,
// hdrloader.cpp, Texture.cpp and thirdparty/RadeonRays/*.cpp, read in place
// from the reference checkout by oracle/Makefile) into oracle/_ref/libref.so.
// Only the tests, smoke() and the CPU baseline of bench.py load it, and only
// to check the product's host-side scene surface against the reference
// (loader, OBJ ingest, BVH builders + flattening, HDR importance tables,
// camera basis).
//
// It defines the application-state global that the reference's Main.cpp
// normally defines (GlobalState, declared in LavaFrame/GlobalState.h) and
// compiles the vendored stb_image implementation that Main.cpp pulls in via
// Export.h; nothing else of the reference is replaced.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "GlobalState.h"
#include "Loader.h"
#include "Scene.h"
#include "hdrloader.h"

#define STB_IMAGE_IMPLEMENTATION
#include "stb_image.h"

LavaFrameState GlobalState;

using namespace LavaFrame;

namespace {
int quiet_log(const char*, ...) { return 0; }
struct RefScene {
    Scene scene;
    RenderOptions options;
    bool built = false;
};
}  // namespace

extern "C" {

void* ref_scene_new(void) {
    GlobalState.useDebug = false;
    RefScene* r = new RefScene;
    r->scene.AddMaterial(Material());
    return r;
}

void ref_scene_free(void* h) { delete static_cast<RefScene*>(h); }

// LoadSceneFromFile + CreateAccelerationStructures (Loader.cpp:45-386).
void* ref_scene_load(const char* path) {
    GlobalState.useDebug = false;
    RefScene* r = new RefScene;
    LavaFrame::Log = quiet_log;
    if (!LoadSceneFromFile(path, &r->scene, r->options)) {
        delete r;
        return nullptr;
    }
    r->built = true;
    return r;
}

// A mesh from de-indexed positions (3 vertices per triangle) and normals.
int ref_scene_add_mesh(void* h, const float* pos, const float* nrm, int ntris) {
    RefScene* r = static_cast<RefScene*>(h);
    Mesh* m = new Mesh;
    m->name = "mesh" + std::to_string(r->scene.meshes.size());
    for (int i = 0; i < 3 * ntris; ++i) {
        m->verticesUVX.push_back(Vec4(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], 0.f));
        m->normalsUVY.push_back(Vec4(nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2], 0.f));
    }
    r->scene.meshes.push_back(m);
    return (int)r->scene.meshes.size() - 1;
}

int ref_scene_add_material(void* h, const float* m28) {
    Material m;
    std::memcpy(&m, m28, sizeof(Material));
    return static_cast<RefScene*>(h)->scene.AddMaterial(m);
}

int ref_scene_add_instance(void* h, int mesh, const float* t16, int mat) {
    Mat4 x;
    std::memcpy(&x.data[0][0], t16, 16 * sizeof(float));
    RefScene* r = static_cast<RefScene*>(h);
    return r->scene.AddMeshInstance(MeshInstance("i", mesh, x, mat));
}

void ref_scene_build(void* h) {
    RefScene* r = static_cast<RefScene*>(h);
    r->scene.CreateAccelerationStructures();
    r->built = true;
}

// Pointers into the flattened arrays + counts.
// counts: [num_nodes, top_level_index, num_tris, num_vertices, num_materials,
//          num_instances, num_lights, res_x, res_y, max_depth, tile_w, tile_h,
//          enable_rr, rr_depth, num_meshes]
void ref_scene_arrays(void* h, const float** nodes, const int** vidx, const float** verts,
                      const float** nrms, const float** mats, const float** xforms,
                      const float** lights, int64_t* counts) {
    RefScene* r = static_cast<RefScene*>(h);
    Scene& s = r->scene;
    *nodes = reinterpret_cast<const float*>(s.bvhTranslator.nodes.data());
    *vidx = reinterpret_cast<const int*>(s.vertIndices.data());
    *verts = reinterpret_cast<const float*>(s.verticesUVX.data());
    *nrms = reinterpret_cast<const float*>(s.normalsUVY.data());
    *mats = reinterpret_cast<const float*>(s.materials.data());
    *xforms = reinterpret_cast<const float*>(s.transforms.data());
    *lights = reinterpret_cast<const float*>(s.lights.data());
    counts[0] = (int64_t)s.bvhTranslator.nodes.size();
    counts[1] = s.bvhTranslator.topLevelIndex;
    counts[2] = (int64_t)s.vertIndices.size();
    counts[3] = (int64_t)s.verticesUVX.size();
    counts[4] = (int64_t)s.materials.size();
    counts[5] = (int64_t)s.transforms.size();
    counts[6] = (int64_t)s.lights.size();
    counts[7] = r->options.resolution.x;
    counts[8] = r->options.resolution.y;
    counts[9] = r->options.maxDepth;
    counts[10] = r->options.tileWidth;
    counts[11] = r->options.tileHeight;
    counts[12] = r->options.enableRR;
    counts[13] = r->options.RRDepth;
    counts[14] = (int64_t)s.meshes.size();
}

// Camera basis after AddCamera (Camera.cpp:101-169):
// out = position, right, up, forward, (fov, focalDist, aperture).
int ref_scene_camera(void* h, float* out15) {
    Scene& s = static_cast<RefScene*>(h)->scene;
    if (!s.camera) return -1;
    const Vec3* v[4] = {&s.camera->position, &s.camera->right, &s.camera->up, &s.camera->forward};
    for (int i = 0; i < 4; ++i) {
        out15[3 * i] = v[i]->x;
        out15[3 * i + 1] = v[i]->y;
        out15[3 * i + 2] = v[i]->z;
    }
    out15[12] = s.camera->fov;
    out15[13] = s.camera->focalDist;
    out15[14] = s.camera->aperture;
    return 0;
}

int ref_make_camera(const float* eye, const float* lookat, float fov, float* out15) {
    Camera c(Vec3(eye[0], eye[1], eye[2]), Vec3(lookat[0], lookat[1], lookat[2]), fov);
    const Vec3* v[4] = {&c.position, &c.right, &c.up, &c.forward};
    for (int i = 0; i < 4; ++i) {
        out15[3 * i] = v[i]->x;
        out15[3 * i + 1] = v[i]->y;
        out15[3 * i + 2] = v[i]->z;
    }
    out15[12] = c.fov;
    out15[13] = c.focalDist;
    out15[14] = c.aperture;
    return 0;
}

// HDRLoader::load on a Radiance file; copies the decoded image and the two
// importance tables out.  Returns 0 on success.
int ref_hdr_load(const char* path, int* w, int* h, float* cols, float* marginal, float* conditional,
                 int capacity_texels) {
    HDRData* d = HDRLoader::load(path);
    if (!d) return -1;
    *w = d->width;
    *h = d->height;
    int n = d->width * d->height;
    if (n > capacity_texels) { delete d; return -2; }
    std::memcpy(cols, d->cols, sizeof(float) * 3 * n);
    std::memcpy(marginal, d->marginalDistData, sizeof(float) * 2 * d->height);
    std::memcpy(conditional, d->conditionalDistData, sizeof(float) * 2 * n);
    delete d;
    return 0;
}

}  // extern "C"
