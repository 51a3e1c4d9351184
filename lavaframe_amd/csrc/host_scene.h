// Host-side scene surface of the MI355X path-tracing core.
//
// Mirrors the reference's Scene / Mesh / MeshInstance / Material / Light /
// Camera / RenderOptions types (LavaFrame/Scene.h:24-107, Mesh.h:13-51,
// Material.h:13-76, Camera.h:14-49, Renderer.h:18-67) with the same public
// member names, so code written against the reference Scene (the Renderer's
// upload path, Renderer.cpp:76-188) compiles against this one unchanged.
//
// The BVH builders reproduce the reference's tree exactly (RadeonRays
// SplitBvh(2, 64, 0, 0.001, 0) for meshes, Mesh.h:18; Bvh(10, 64, false) for
// instances, Scene.h:51) because the flattened node array is part of the data
// contract and "identical BVH" is what bit-exact traversal is defined over.
#pragma once

#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

namespace lfh {

struct Vec2 {
    float x = 0, y = 0;
    Vec2() = default;
    Vec2(float a, float b) : x(a), y(b) {}
};

struct Vec4 {
    float x = 0, y = 0, z = 0, w = 0;
    Vec4() = default;
    Vec4(float a, float b, float c, float d) : x(a), y(b), z(c), w(d) {}
};

struct Vec3 {
    float x = 0, y = 0, z = 0;
    Vec3() = default;
    Vec3(float a, float b, float c) : x(a), y(b), z(c) {}
    explicit Vec3(const Vec4& v) : x(v.x), y(v.y), z(v.z) {}
    Vec3 operator+(const Vec3& b) const { return Vec3(x + b.x, y + b.y, z + b.z); }
    Vec3 operator-(const Vec3& b) const { return Vec3(x - b.x, y - b.y, z - b.z); }
    Vec3 operator*(const Vec3& b) const { return Vec3(x * b.x, y * b.y, z * b.z); }
    Vec3 operator*(float s) const { return Vec3(x * s, y * s, z * s); }
    float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    float& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
    // std::min / std::max semantics: min(a, b) = (b < a) ? b : a.
    static Vec3 Min(const Vec3& a, const Vec3& b) {
        return Vec3(b.x < a.x ? b.x : a.x, b.y < a.y ? b.y : a.y, b.z < a.z ? b.z : a.z);
    }
    static Vec3 Max(const Vec3& a, const Vec3& b) {
        return Vec3(a.x < b.x ? b.x : a.x, a.y < b.y ? b.y : a.y, a.z < b.z ? b.z : a.z);
    }
    static Vec3 Cross(const Vec3& a, const Vec3& b) {
        return Vec3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
    }
    static float Dot(const Vec3& a, const Vec3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
    static float Length(const Vec3& a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
    static Vec3 Normalize(const Vec3& a) {
        float l = Length(a);
        return Vec3(a.x / l, a.y / l, a.z / l);
    }
};

// Column-major 4x4: data[col][row]; translation in data[3][0..2] (Mat4.h:24,35-42).
struct Mat4 {
    float data[4][4];
    Mat4() {
        for (int c = 0; c < 4; ++c)
            for (int r = 0; r < 4; ++r) data[c][r] = (c == r) ? 1.f : 0.f;
    }
    float (&operator[](int c))[4] { return data[c]; }
    const float (&operator[](int c) const)[4] { return data[c]; }
};

// 28-float Disney parameter block, same order and defaults as Material.h:16-46.
struct Material {
    Vec3 albedo{1.f, 1.f, 1.f};
    float specular = 0.5f;
    Vec3 emission{0.f, 0.f, 0.f};
    float anisotropic = 0.f;
    float metallic = 0.f;
    float roughness = 0.5f;
    float subsurface = 0.f;
    float specularTint = 0.f;
    float sheen = 0.f;
    float sheenTint = 0.f;
    float clearcoat = 0.f;
    float clearcoatRoughness = 0.f;
    float transmission = 0.f;
    float ior = 1.45f;
    float atDistance = 1.f;
    float padding1 = 0.f;
    Vec3 extinction{1.f, 1.f, 1.f};
    float padding2 = 0.f;
    float albedoTexID = -1.f;
    float metallicRoughnessTexID = -1.f;
    float normalmapTexID = -1.f;
    float emissionmapTexID = -1.f;
};
static_assert(sizeof(Material) == 28 * sizeof(float), "Material must be 28 floats");

enum LightType { RectLight = 0, SphereLight = 1, DistantLight = 2 };

// 15 floats (Scene.h:31-40).  radius/area/type are zero-initialised here (the
// reference leaves them uninitialised for some light kinds, Loader.cpp:162).
struct Light {
    Vec3 position, emission, u, v;
    float radius = 0.f, area = 0.f, type = 0.f;
};
static_assert(sizeof(Light) == 15 * sizeof(float), "Light must be 15 floats");

struct Indices { int x, y, z; };

struct iVec2 { int x, y; };

struct BBox {
    Vec3 pmin{3.402823466e+38f, 3.402823466e+38f, 3.402823466e+38f};
    Vec3 pmax{-3.402823466e+38f, -3.402823466e+38f, -3.402823466e+38f};
    Vec3 center() const { return (pmax + pmin) * 0.5f; }
    Vec3 extents() const { return pmax - pmin; }
    float surface_area() const {
        Vec3 e = extents();
        return 2.f * (e.x * e.y + e.x * e.z + e.y * e.z);
    }
    int maxdim() const {
        Vec3 e = extents();
        if (e.x >= e.y && e.x >= e.z) return 0;
        if (e.y >= e.x && e.y >= e.z) return 1;
        if (e.z >= e.x && e.z >= e.y) return 2;
        return 0;
    }
    void grow(const Vec3& p) { pmin = Vec3::Min(pmin, p); pmax = Vec3::Max(pmax, p); }
    void grow(const BBox& b) { pmin = Vec3::Min(pmin, b.pmin); pmax = Vec3::Max(pmax, b.pmax); }
};

// Binary BVH over primitive boxes, index-linked.  Two build policies:
//   kObjectSah   — binned object SAH, leaves of < 4 prims, right subtree built
//                  (and its prims packed) first  (split_bvh.cpp:8-194, 299-402)
//   kMedianSplit — centroid-bounds midpoint split, 1-prim leaves, left first
//                  (bvh.cpp:41-243, usesah = false)
class Bvh {
public:
    enum Policy { kObjectSah, kMedianSplit };
    struct Node {
        BBox bounds;
        bool leaf = false;
        int lc = -1, rc = -1;        // internal: children
        int startidx = 0, numprims = 0;  // leaf: range of packed indices
    };

    Bvh(Policy p, float traversal_cost, int num_bins)
        : policy_(p), traversal_cost_(traversal_cost), num_bins_(num_bins) {}

    void Build(const BBox* bounds, int numbounds);
    const BBox& Bounds() const { return bounds_; }
    const std::vector<Node>& nodes() const { return nodes_; }
    int root() const { return root_; }
    int node_count() const { return (int)nodes_.size(); }
    const std::vector<int>& packed_indices() const { return packed_; }
    int height() const { return height_; }

private:
    struct Request {
        int startidx, numprims;
        BBox bounds, centroid_bounds;
        int level;
    };
    struct PrimRef { BBox bounds; Vec3 center; int idx; };
    struct SahSplit { int dim; float split; float sah; };

    int BuildSah(const Request& req);
    int BuildMedian(const Request& req);
    SahSplit FindObjectSahSplit(const Request& req) const;

    Policy policy_;
    float traversal_cost_;
    int num_bins_;
    BBox bounds_;
    std::vector<Node> nodes_;
    std::vector<int> packed_;
    int root_ = -1;
    int height_ = 0;
    // SAH state
    std::vector<PrimRef> refs_;
    // median state
    const BBox* in_bounds_ = nullptr;
    std::vector<Vec3> centroids_;
    std::vector<int> primindices_;
};

class Mesh {
public:
    Mesh() : bvh(Bvh::kObjectSah, 2.0f, 64) {}
    bool LoadFromFile(const std::string& filename, std::string* err = nullptr);
    void BuildBVH();
    std::vector<Vec4> verticesUVX;  // xyz + u
    std::vector<Vec4> normalsUVY;   // normal + v
    Bvh bvh;
    std::string name;
};

class MeshInstance {
public:
    MeshInstance(std::string n, int meshId, Mat4 xform, int matId)
        : transform(xform), name(std::move(n)), materialID(matId), meshID(meshId) {}
    Mat4 transform;
    std::string name;
    int materialID;
    int meshID;
};

// Orbit camera (Camera.cpp:101-169); fov stored in radians.
class Camera {
public:
    Camera(Vec3 eye, Vec3 lookat, float fov_deg);
    Vec3 position, up, right, forward;
    float focalDist = 0.1f;
    float aperture = 0.f;
    float fov = 0.f;
    bool isMoving = false;
private:
    void UpdateCamera();
    Vec3 worldUp{0.f, 1.f, 0.f};
    Vec3 pivot;
    float pitch = 0.f, radius = 0.f, yaw = 0.f;
};

// Environment map with its importance tables (hdrloader.h:19-28).
struct HDRData {
    int width = 0, height = 0;
    std::vector<float> cols;             // RGB, width*height*3
    std::vector<Vec2> marginalDistData;  // (row/H, pdf), H
    std::vector<Vec2> conditionalDistData;  // (col/W, pdf), W*H
};
// HDRLoader::buildDistributions (hdrloader.cpp:57-131).
void BuildHdrDistributions(HDRData* hdr);

struct RenderOptions {
    iVec2 resolution{1280, 720};
    int maxDepth = 2;
    int tileWidth = 512;
    int tileHeight = 512;
    bool useEnvMap = false;
    bool enableRR = true;
    bool useConstantBg = false;
    int RRDepth = 2;
    float hdrMultiplier = 1.f;
    Vec3 bgColor{0.5f, 0.5f, 0.5f};
    int tonemapIndex = 2;
    // post-process (Renderer.h:34-43); the reference leaves the three vignette
    // fields uninitialised (its constructor declares locals of the same
    // names), here they take the values it evidently intended
    bool useCA = false;
    bool useCADistortion = false;
    bool useVignette = false;
    float caDistance = 0.05f;
    float caP1 = 5.0f;
    float caP2 = -0.5f;
    float caP3 = 0.5f;
    float vignetteIntensity = 0.0f;
    float vignettePower = 1.0f;
};

// Flat node of the reference layout (bvh_translator.h:48-53).
struct FlatNode { Vec3 bboxmin, bboxmax, LRLeaf; };
static_assert(sizeof(FlatNode) == 36, "FlatNode must be 36 bytes");

// Pointer-tree -> flat array (bvh_translator.cpp:35-140).
class BvhTranslator {
public:
    void Process(const Bvh* tlas, const std::vector<Mesh*>& meshes,
                 const std::vector<MeshInstance>& instances);
    void UpdateTLAS(const Bvh* tlas, const std::vector<MeshInstance>& instances);
    int topLevelIndex = 0;
    std::vector<FlatNode> nodes;
private:
    int ProcessBLASNodes(const Bvh& bvh, int node, int triOffset);
    int ProcessTLASNodes(const Bvh& tlas, int node, const std::vector<MeshInstance>& inst);
    int curNode_ = 0;
    std::vector<int> bvhRootStartIndices_;
};

class Scene {
public:
    Scene() : sceneBvh(Bvh::kMedianSplit, 10.0f, 64) {}
    ~Scene();
    Scene(const Scene&) = delete;
    Scene& operator=(const Scene&) = delete;

    int AddMesh(const std::string& filename);
    // De-indexed triangles (3 vertices each) from arrays; normals/uvs optional.
    int AddMesh(const float* positions, const float* normals, const float* uvs,
                int64_t num_tris, const std::string& name);
    // Scene::AddTexture (Scene.cpp:45-66): a file (PNG / binary PPM) or RGBA8
    // pixels; the same name twice gives the first id; -1 if the file cannot be read.
    int AddTexture(const std::string& filename);
    int AddTexture(const uint8_t* rgba, int width, int height, const std::string& name);
    bool PackTextures();
    int AddMaterial(const Material& material);
    int AddMeshInstance(const MeshInstance& meshInstance);
    int AddLight(const Light& light);
    void AddCamera(Vec3 eye, Vec3 lookat, float fov);
    // AddHDR from an in-memory RGB float image.
    void AddHDR(const float* rgb, int width, int height);

    void CreateAccelerationStructures();
    // The flattened arrays without BVHs: every mesh's triangles in their own order.
    void Flatten();
    void RebuildInstances();

    RenderOptions renderOptions;
    std::vector<Mesh*> meshes;
    std::vector<Indices> vertIndices;
    std::vector<Vec4> verticesUVX;
    std::vector<Vec4> normalsUVY;
    std::vector<Mat4> transforms;
    std::vector<int64_t> meshTriOffsets;     // triangle range of each mesh in vertIndices
    std::vector<int32_t> instanceMeshes, instanceMaterials;
    bool flatOnly = false;                   // Flatten(): no BVH arrays
    std::vector<Material> materials;
    struct Texture {
        std::string name;
        int width = 0, height = 0;
        std::vector<uint8_t> rgba;
    };
    std::vector<Texture> textures;
    std::vector<uint8_t> textureMapsArray;   // RGBA8 layers, first row first
    int texWidth = 0, texHeight = 0;
    std::vector<MeshInstance> meshInstances;
    bool instancesModified = false;
    std::vector<Light> lights;
    HDRData* hdrData = nullptr;
    Camera* camera = nullptr;
    BvhTranslator bvhTranslator;
    BBox sceneBounds;
    std::string lastError;

private:
    Bvh sceneBvh;
    void createBLAS();
    void createTLAS();
};

// LoadSceneFromFile (Loader.cpp:45-386): same block grammar and defaults.
bool LoadSceneFromFile(const std::string& filename, Scene* scene, RenderOptions& options,
                       std::string* err);

// Parses one OBJ number with the arithmetic of the OBJ reader the reference
// links (tinyobjloader's digit accumulation), so vertices are bit-identical.
double ParseObjReal(const char** s);

}  // namespace lfh
