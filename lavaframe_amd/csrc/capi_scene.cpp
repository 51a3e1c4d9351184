// C ABI over the host scene surface (lfh_scene_*).
#include <cstring>
#include <string>

#include "../../include/lfh.h"
#include "capi_common.h"
#include "host_scene.h"

struct lfh_scene {
    lfh::Scene scene;
    bool built = false;
};

using lfh::Scene;

static int fail_scene(lfh_scene* s, int code, const std::string& msg) {
    lfh_internal::set_thread_error(msg);
    if (s) s->scene.lastError = msg;
    return code;
}

extern "C" {

int lfh_abi_version(void) { return LFH_ABI_VERSION; }

int lfh_scene_create(lfh_scene** out) {
    if (!out) return fail_scene(nullptr, LFH_E_INVALID, "lfh_scene_create: out is NULL");
    *out = new (std::nothrow) lfh_scene;
    if (!*out) return fail_scene(nullptr, LFH_E_OOM, "lfh_scene_create: out of host memory");
    return LFH_OK;
}

void lfh_scene_destroy(lfh_scene* s) { delete s; }

int lfh_scene_load(lfh_scene* s, const char* path) {
    if (!s || !path) return fail_scene(s, LFH_E_INVALID, "lfh_scene_load: NULL argument");
    std::string err;
    if (!lfh::LoadSceneFromFile(path, &s->scene, s->scene.renderOptions, &err))
        return fail_scene(s, LFH_E_IO, "lfh_scene_load: " + err);
    if (!s->scene.PackTextures()) return fail_scene(s, LFH_E_INVALID, "lfh_scene_load: " + s->scene.lastError);
    s->built = true;
    return LFH_OK;
}

int lfh_scene_add_texture(lfh_scene* s, const uint8_t* rgba8, int32_t width, int32_t height, const char* name,
                          int32_t* id) {
    if (!s || !rgba8) return fail_scene(s, LFH_E_INVALID, "lfh_scene_add_texture: NULL argument");
    if (width <= 0 || height <= 0 || width > 32768 || height > 32768)
        return fail_scene(s, LFH_E_INVALID, "lfh_scene_add_texture: bad size");
    int r = s->scene.AddTexture(rgba8, width, height, name ? std::string(name) : std::string());
    if (id) *id = r;
    return LFH_OK;
}

int lfh_scene_add_texture_file(lfh_scene* s, const char* path, int32_t* id) {
    if (!s || !path) return fail_scene(s, LFH_E_INVALID, "lfh_scene_add_texture_file: NULL argument");
    int r = s->scene.AddTexture(std::string(path));
    if (id) *id = r;
    if (r < 0) return fail_scene(s, LFH_E_IO, "lfh_scene_add_texture_file: " + s->scene.lastError);
    return LFH_OK;
}

int lfh_scene_add_mesh_file(lfh_scene* s, const char* path, int32_t* id) {
    if (!s || !path) return fail_scene(s, LFH_E_INVALID, "lfh_scene_add_mesh_file: NULL argument");
    int r = s->scene.AddMesh(std::string(path));
    if (id) *id = r;
    if (r < 0) return fail_scene(s, LFH_E_IO, "lfh_scene_add_mesh_file: " + s->scene.lastError);
    return LFH_OK;
}

int lfh_scene_add_mesh(lfh_scene* s, const float* positions, const float* normals, const float* uvs,
                       int64_t num_tris, int32_t* id) {
    if (!s || !positions || num_tris <= 0)
        return fail_scene(s, LFH_E_INVALID, "lfh_scene_add_mesh: need positions and num_tris > 0");
    if (num_tris > (1LL << 24) / 3)
        return fail_scene(s, LFH_E_INVALID,
                          "lfh_scene_add_mesh: more than 2^24 vertices cannot be indexed by float-encoded nodes");
    int r = s->scene.AddMesh(positions, normals, uvs, num_tris,
                             "mesh" + std::to_string(s->scene.meshes.size()));
    if (id) *id = r;
    return LFH_OK;
}

int lfh_scene_add_material(lfh_scene* s, const float* m28, int32_t* id) {
    if (!s || !m28) return fail_scene(s, LFH_E_INVALID, "lfh_scene_add_material: NULL argument");
    lfh::Material m;
    std::memcpy(&m, m28, sizeof m);
    int r = s->scene.AddMaterial(m);
    if (id) *id = r;
    return LFH_OK;
}

int lfh_scene_add_light(lfh_scene* s, const float* l15, int32_t* id) {
    if (!s || !l15) return fail_scene(s, LFH_E_INVALID, "lfh_scene_add_light: NULL argument");
    lfh::Light l;
    std::memcpy(&l, l15, sizeof l);
    int r = s->scene.AddLight(l);
    if (id) *id = r;
    return LFH_OK;
}

int lfh_scene_add_instance(lfh_scene* s, int32_t mesh_id, const float* t16, int32_t material_id, int32_t* id) {
    if (!s || !t16) return fail_scene(s, LFH_E_INVALID, "lfh_scene_add_instance: NULL argument");
    if (mesh_id < 0 || mesh_id >= (int)s->scene.meshes.size())
        return fail_scene(s, LFH_E_INVALID, "lfh_scene_add_instance: mesh id out of range");
    lfh::Mat4 m;
    std::memcpy(m.data, t16, sizeof m.data);
    int r = s->scene.AddMeshInstance(lfh::MeshInstance("inst" + std::to_string(s->scene.meshInstances.size()),
                                                       mesh_id, m, material_id));
    if (id) *id = r;
    return LFH_OK;
}

int lfh_scene_set_camera(lfh_scene* s, const float* eye, const float* lookat, float fov_deg, float aperture,
                         float focal_dist) {
    if (!s || !eye || !lookat) return fail_scene(s, LFH_E_INVALID, "lfh_scene_set_camera: NULL argument");
    s->scene.AddCamera(lfh::Vec3(eye[0], eye[1], eye[2]), lfh::Vec3(lookat[0], lookat[1], lookat[2]), fov_deg);
    s->scene.camera->aperture = aperture;
    s->scene.camera->focalDist = focal_dist;
    return LFH_OK;
}

int lfh_scene_set_hdr(lfh_scene* s, const float* rgb, int32_t w, int32_t h) {
    if (!s || !rgb || w <= 0 || h <= 0) return fail_scene(s, LFH_E_INVALID, "lfh_scene_set_hdr: bad image");
    s->scene.AddHDR(rgb, w, h);
    return LFH_OK;
}

int lfh_scene_set_options(lfh_scene* s, const lfh_options* o) {
    if (!s || !o) return fail_scene(s, LFH_E_INVALID, "lfh_scene_set_options: NULL argument");
    auto& r = s->scene.renderOptions;
    r.resolution = {o->width, o->height};
    r.tileWidth = o->tile_width > 0 ? o->tile_width : o->width;
    r.tileHeight = o->tile_height > 0 ? o->tile_height : o->height;
    r.maxDepth = o->max_depth;
    r.enableRR = o->enable_rr != 0;
    r.RRDepth = o->rr_depth;
    r.useEnvMap = o->use_env_map != 0;
    r.useConstantBg = o->use_constant_bg != 0;
    r.bgColor = lfh::Vec3(o->bg_color[0], o->bg_color[1], o->bg_color[2]);
    r.hdrMultiplier = o->hdr_multiplier;
    return LFH_OK;
}

int lfh_scene_get_options(const lfh_scene* s, lfh_options* o) {
    if (!s || !o) return fail_scene(nullptr, LFH_E_INVALID, "lfh_scene_get_options: NULL argument");
    const auto& r = s->scene.renderOptions;
    std::memset(o, 0, sizeof *o);
    o->width = r.resolution.x;
    o->height = r.resolution.y;
    o->tile_width = r.tileWidth;
    o->tile_height = r.tileHeight;
    o->max_depth = r.maxDepth;
    o->enable_rr = r.enableRR;
    o->rr_depth = r.RRDepth;
    o->use_env_map = r.useEnvMap && s->scene.hdrData != nullptr;
    o->use_constant_bg = r.useConstantBg;
    o->bg_color[0] = r.bgColor.x;
    o->bg_color[1] = r.bgColor.y;
    o->bg_color[2] = r.bgColor.z;
    o->hdr_multiplier = r.hdrMultiplier;
    o->kahan = 1;
    return LFH_OK;
}

int lfh_scene_get_postprocess(const lfh_scene* s, lfh_postprocess* pp) {
    if (!s || !pp) return fail_scene(nullptr, LFH_E_INVALID, "lfh_scene_get_postprocess: NULL argument");
    const auto& r = s->scene.renderOptions;
    pp->tonemap_index = r.tonemapIndex;
    pp->use_ca = r.useCA;
    pp->use_ca_distortion = r.useCADistortion;
    pp->ca_distance = r.caDistance;
    pp->ca_p1 = r.caP1;
    pp->ca_p2 = r.caP2;
    pp->ca_p3 = r.caP3;
    pp->use_vignette = r.useVignette;
    pp->vignette_intensity = r.vignetteIntensity;
    pp->vignette_power = r.vignettePower;
    return LFH_OK;
}

int lfh_scene_get_camera(const lfh_scene* s, lfh_camera* c) {
    if (!s || !c) return fail_scene(nullptr, LFH_E_INVALID, "lfh_scene_get_camera: NULL argument");
    if (!s->scene.camera) return fail_scene(nullptr, LFH_E_STATE, "lfh_scene_get_camera: scene has no camera");
    const lfh::Camera& k = *s->scene.camera;
    const lfh::Vec3* v[4] = {&k.position, &k.right, &k.up, &k.forward};
    float* d[4] = {c->position, c->right, c->up, c->forward};
    for (int i = 0; i < 4; ++i) { d[i][0] = v[i]->x; d[i][1] = v[i]->y; d[i][2] = v[i]->z; }
    c->fov = k.fov;
    c->focal_dist = k.focalDist;
    c->aperture = k.aperture;
    return LFH_OK;
}

int lfh_scene_build(lfh_scene* s) {
    if (!s) return fail_scene(nullptr, LFH_E_INVALID, "lfh_scene_build: NULL scene");
    if (s->scene.meshInstances.empty()) return fail_scene(s, LFH_E_STATE, "lfh_scene_build: no instances");
    for (const auto& mi : s->scene.meshInstances)
        if (mi.materialID < 0 || mi.materialID >= (int)s->scene.materials.size())
            return fail_scene(s, LFH_E_INVALID, "lfh_scene_build: instance material id out of range");
    s->scene.CreateAccelerationStructures();
    if (!s->scene.PackTextures()) return fail_scene(s, LFH_E_INVALID, "lfh_scene_build: " + s->scene.lastError);
    s->built = true;
    return LFH_OK;
}

int lfh_scene_flatten(lfh_scene* s) {
    if (!s) return fail_scene(nullptr, LFH_E_INVALID, "lfh_scene_flatten: NULL scene");
    if (s->scene.meshInstances.empty()) return fail_scene(s, LFH_E_STATE, "lfh_scene_flatten: no instances");
    for (const auto& mi : s->scene.meshInstances)
        if (mi.materialID < 0 || mi.materialID >= (int)s->scene.materials.size())
            return fail_scene(s, LFH_E_INVALID, "lfh_scene_flatten: instance material id out of range");
    s->scene.Flatten();
    if (!s->scene.PackTextures()) return fail_scene(s, LFH_E_INVALID, "lfh_scene_flatten: " + s->scene.lastError);
    s->built = true;
    return LFH_OK;
}

int lfh_scene_set_instance_transform(lfh_scene* s, int32_t inst, const float* t16) {
    if (!s || !t16 || inst < 0 || inst >= (int)s->scene.meshInstances.size())
        return fail_scene(s, LFH_E_INVALID, "lfh_scene_set_instance_transform: bad instance");
    std::memcpy(s->scene.meshInstances[inst].transform.data, t16, 16 * sizeof(float));
    return LFH_OK;
}

int lfh_scene_rebuild_instances(lfh_scene* s) {
    if (!s || !s->built) return fail_scene(s, LFH_E_STATE, "lfh_scene_rebuild_instances: scene not built");
    s->scene.RebuildInstances();
    return LFH_OK;
}

int lfh_scene_get_desc(const lfh_scene* s, lfh_scene_desc* d) {
    if (!s || !d) return fail_scene(nullptr, LFH_E_INVALID, "lfh_scene_get_desc: NULL argument");
    if (!s->built) return fail_scene(nullptr, LFH_E_STATE, "lfh_scene_get_desc: scene not built");
    const Scene& sc = s->scene;
    std::memset(d, 0, sizeof *d);
    if (!sc.flatOnly) {
        d->nodes = reinterpret_cast<const float*>(sc.bvhTranslator.nodes.data());
        d->num_nodes = (int64_t)sc.bvhTranslator.nodes.size();
        d->top_level_index = sc.bvhTranslator.topLevelIndex;
    }
    d->mesh_tri_offsets = sc.meshTriOffsets.data();
    d->num_meshes = (int32_t)sc.meshTriOffsets.size() - 1;
    d->instance_meshes = sc.instanceMeshes.data();
    d->instance_materials = sc.instanceMaterials.data();
    if (!sc.textures.empty()) {
        d->textures = sc.textureMapsArray.data();
        d->tex_width = sc.texWidth;
        d->tex_height = sc.texHeight;
        d->num_textures = (int32_t)sc.textures.size();
    }
    d->vert_indices = reinterpret_cast<const int32_t*>(sc.vertIndices.data());
    d->num_tris = (int64_t)sc.vertIndices.size();
    d->vertices_uvx = reinterpret_cast<const float*>(sc.verticesUVX.data());
    d->normals_uvy = reinterpret_cast<const float*>(sc.normalsUVY.data());
    d->num_vertices = (int64_t)sc.verticesUVX.size();
    d->materials = reinterpret_cast<const float*>(sc.materials.data());
    d->num_materials = (int32_t)sc.materials.size();
    d->transforms = reinterpret_cast<const float*>(sc.transforms.data());
    d->num_instances = (int32_t)sc.transforms.size();
    d->lights = reinterpret_cast<const float*>(sc.lights.data());
    d->num_lights = (int32_t)sc.lights.size();
    if (sc.hdrData && sc.renderOptions.useEnvMap) {
        d->hdr_width = sc.hdrData->width;
        d->hdr_height = sc.hdrData->height;
        d->hdr_cols = sc.hdrData->cols.data();
        d->hdr_marginal = reinterpret_cast<const float*>(sc.hdrData->marginalDistData.data());
        d->hdr_conditional = reinterpret_cast<const float*>(sc.hdrData->conditionalDistData.data());
    }
    return LFH_OK;
}

int lfh_scene_counts(const lfh_scene* s, int32_t* nm, int32_t* ni) {
    if (!s) return fail_scene(nullptr, LFH_E_INVALID, "lfh_scene_counts: NULL scene");
    if (nm) *nm = (int32_t)s->scene.meshes.size();
    if (ni) *ni = (int32_t)s->scene.meshInstances.size();
    return LFH_OK;
}

}  // extern "C"
