// Reference flat layout -> device layout (Renderer::Init's upload step,
// Renderer.cpp:76-188, re-designed for gfx950 loads).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <unordered_map>
#include <string>
#include <vector>

#include "ctx.h"
#include "hdr_kernels.h"
#include "xform.h"

namespace lfh {

using lfhd::kRefSentinel;

static inline float u2f(uint32_t u) {
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

static void put_light(const float* l, float* out) {
    // raw: pos(3) emission(3) u(3) v(3) radius area type
    const float* p = l;
    const float* e = l + 3;
    const float* u = l + 6;
    const float* v = l + 9;
    float radius = l[12], area = l[13], type = l[14];
    float o[28] = {p[0], p[1], p[2], type, e[0], e[1], e[2], area, u[0], u[1], u[2], radius, v[0], v[1], v[2], 0.f};
    // derived quad-light quantities (closest_hit.glsl:30-35)
    float cx = u[1] * v[2] - u[2] * v[1];
    float cy = u[2] * v[0] - u[0] * v[2];
    float cz = u[0] * v[1] - u[1] * v[0];
    float len = std::sqrt(cx * cx + cy * cy + cz * cz);
    float nx = cx / len, ny = cy / len, nz = cz / len;
    float pw = nx * p[0] + ny * p[1] + nz * p[2];
    float iu = 1.0f / (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    float iv = 1.0f / (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    float d[12] = {nx, ny, nz, pw, u[0] * iu, u[1] * iu, u[2] * iu, 0.f, v[0] * iv, v[1] * iv, v[2] * iv, 0.f};
    std::memcpy(out, o, 16 * sizeof(float));
    std::memcpy(out + 16, d, 12 * sizeof(float));
}

static void put_instance(PackedScene* ps, int i, const float* t16, uint32_t blasRef, int mat) {
    float m[4][4];
    std::memcpy(m, t16, sizeof m);
    float inv[4][4];
    inverse4(m, inv);
    float* r = &ps->insts[(size_t)i * 16];
    for (int row = 0; row < 3; ++row) {
        r[row * 4 + 0] = inv[0][row];
        r[row * 4 + 1] = inv[1][row];
        r[row * 4 + 2] = inv[2][row];
        r[row * 4 + 3] = inv[3][row];
    }
    r[12] = u2f(blasRef);
    r[13] = u2f((uint32_t)mat);
    r[14] = u2f((uint32_t)i);
    r[15] = 0.f;
    std::memcpy(&ps->xforms[(size_t)i * 16], t16, 16 * sizeof(float));
    float nm[3][3];
    normal_matrix(m, nm);
    float* n = &ps->nmats[(size_t)i * 12];
    for (int row = 0; row < 3; ++row) {
        n[row * 4 + 0] = nm[row][0];
        n[row * 4 + 1] = nm[row][1];
        n[row * 4 + 2] = nm[row][2];
        n[row * 4 + 3] = 0.f;
    }
}

namespace {
struct Builder {
    const lfh_scene_desc& d;
    PackedScene* ps;
    std::vector<int32_t> lr;  // 3 per flat node
    std::vector<uint32_t> inst_blas;
    std::vector<int32_t> inst_mat;
    std::string err;

    explicit Builder(const lfh_scene_desc& dd, PackedScene* p) : d(dd), ps(p) {}

    bool ref_of(int64_t n, uint32_t* out) {
        if (n < 0 || n >= d.num_nodes) { err = "node index out of range"; return false; }
        int32_t leaf = lr[n * 3 + 2];
        if (leaf > 0) {
            int32_t first = lr[n * 3], count = lr[n * 3 + 1];
            if (count < 1 || count > 7) { err = "BLAS leaf with " + std::to_string(count) + " triangles (1..7 supported)"; return false; }
            if (first < 0 || (int64_t)first + count > d.num_tris || first >= (1 << 28)) { err = "BLAS leaf triangle range out of bounds"; return false; }
            *out = lfhd::make_leaf((uint32_t)first, (uint32_t)count);
        } else if (leaf < 0) {
            int64_t inst = -(int64_t)leaf - 1;
            if (inst >= d.num_instances) { err = "TLAS leaf instance out of range"; return false; }
            *out = lfhd::make_inst((uint32_t)inst);
        } else {
            int32_t p = ps->flat_to_packed[n];
            if (p < 0) { err = "internal node not reachable"; return false; }
            *out = (uint32_t)p;
        }
        return true;
    }

    bool run() {
        const int64_t N = d.num_nodes;
        if (!d.nodes || N <= 0) { err = "no BVH nodes"; return false; }
        if (d.top_level_index < 0 || d.top_level_index >= N) { err = "top_level_index out of range"; return false; }
        if (d.num_tris > 0 && (!d.vert_indices || !d.vertices_uvx)) { err = "missing triangle arrays"; return false; }
        if (d.num_instances > 0 && !d.transforms) { err = "missing transforms"; return false; }
        if (d.num_materials <= 0 || !d.materials) { err = "at least one material is required"; return false; }
        lr.resize(N * 3);
        for (int64_t i = 0; i < N; ++i)
            for (int k = 0; k < 3; ++k) lr[i * 3 + k] = (int32_t)d.nodes[i * 9 + 6 + k];
        // Reachable internal nodes (DFS from the top-level node, entering
        // BLAS roots through TLAS leaves).
        std::vector<uint8_t> seen(N, 0);
        std::vector<std::pair<int64_t, int>> stack{{d.top_level_index, 0}};
        inst_blas.assign(d.num_instances, kRefSentinel);
        inst_mat.assign(d.num_instances, 0);
        int maxlev = 0;
        while (!stack.empty()) {
            auto [n, lev] = stack.back();
            stack.pop_back();
            if (n < 0 || n >= N) { err = "child index out of range"; return false; }
            if (seen[n]) continue;
            seen[n] = 1;
            maxlev = std::max(maxlev, lev);
            int32_t leaf = lr[n * 3 + 2];
            if (leaf == 0) {
                stack.push_back({lr[n * 3 + 0], lev + 1});
                stack.push_back({lr[n * 3 + 1], lev + 1});
            } else if (leaf < 0) {
                int64_t inst = -(int64_t)leaf - 1;
                if (inst >= d.num_instances) { err = "TLAS leaf instance out of range"; return false; }
                stack.push_back({lr[n * 3 + 0], lev + 1});
            }
        }
        // Acyclic by construction: every internal node's children come after
        // it (preorder flattening, bvh_translator.cpp:35-89), and no TLAS leaf
        // may be reached again below a BLAS root.  A desc violating either
        // could make the traversal loop, so it is rejected here.
        std::vector<uint8_t> inBlas(N, 0);
        for (int64_t i = 0; i < N; ++i) {
            if (!seen[i]) continue;
            int32_t leaf = lr[i * 3 + 2];
            if (leaf == 0) {
                if (lr[i * 3] <= i || lr[i * 3 + 1] <= i) { err = "BVH child index not after its parent"; return false; }
            } else if (leaf < 0) {
                int64_t b = lr[i * 3];
                if (b < 0 || b >= N) { err = "BLAS root out of range"; return false; }
                inBlas[b] = 1;
            }
        }
        for (int64_t i = 0; i < N; ++i) {  // propagate "below a BLAS root" in index order
            if (!seen[i] || !inBlas[i]) continue;
            int32_t leaf = lr[i * 3 + 2];
            if (leaf < 0) { err = "nested instancing (TLAS leaf inside a BLAS) is not supported"; return false; }
            if (leaf == 0) {
                inBlas[lr[i * 3]] = 1;
                inBlas[lr[i * 3 + 1]] = 1;
            }
        }
        ps->max_depth_levels = maxlev;
        if (maxlev + 4 > (int)96) { err = "BVH deeper than the traversal stack (" + std::to_string(maxlev) + " levels)"; return false; }
        ps->flat_to_packed.assign(N, -1);
        int64_t np = 0;
        for (int64_t i = 0; i < N; ++i)
            if (seen[i] && lr[i * 3 + 2] == 0) ps->flat_to_packed[i] = (int32_t)np++;
        if (np >= (1 << 30)) { err = "too many BVH nodes"; return false; }
        ps->num_internal = np;
        ps->nodes.assign((size_t)np * 16, 0.f);
        for (int64_t i = 0; i < N; ++i) {
            if (!(seen[i] && lr[i * 3 + 2] == 0)) continue;
            int64_t p = ps->flat_to_packed[i];
            int64_t L = lr[i * 3], R = lr[i * 3 + 1];
            uint32_t rl, rr;
            if (!ref_of(L, &rl) || !ref_of(R, &rr)) return false;
            const float* a = &d.nodes[L * 9];
            const float* b = &d.nodes[R * 9];
            float* o = &ps->nodes[p * 16];
            o[0] = a[0]; o[1] = a[1]; o[2] = a[2]; o[3] = a[3];
            o[4] = a[4]; o[5] = a[5]; o[6] = b[0]; o[7] = b[1];
            o[8] = b[2]; o[9] = b[3]; o[10] = b[4]; o[11] = b[5];
            o[12] = u2f(rl); o[13] = u2f(rr); o[14] = 0.f; o[15] = 0.f;
        }
        for (int64_t i = 0; i < N; ++i) {
            if (!seen[i] || lr[i * 3 + 2] >= 0) continue;
            int64_t inst = -(int64_t)lr[i * 3 + 2] - 1;
            uint32_t br;
            if (!ref_of(lr[i * 3], &br)) return false;
            if (lfhd::ref_is_inst(br)) { err = "TLAS leaf points at another TLAS leaf"; return false; }
            inst_blas[inst] = br;
            inst_mat[inst] = lr[i * 3 + 1];
        }
        if (!ref_of(d.top_level_index, &ps->root)) return false;
        ps->top_level_index = d.top_level_index;
        ps->flat_lrleaf = lr;
        return fill_rest(true);
    }

    // Scene arrays without a BVH: validated here, BVHs built on the device.
    bool run_device() {
        if (d.num_tris <= 0 || !d.vert_indices || !d.vertices_uvx) { err = "missing triangle arrays"; return false; }
        if (d.num_tris >= (1 << 28)) { err = "too many triangles for a device build (2^28)"; return false; }
        if (d.num_meshes <= 0 || d.num_meshes > 65535 || !d.mesh_tri_offsets) {
            err = "a device build needs 1..65535 meshes with triangle offsets";
            return false;
        }
        if (d.mesh_tri_offsets[0] != 0 || d.mesh_tri_offsets[d.num_meshes] != d.num_tris) {
            err = "mesh triangle offsets must start at 0 and end at num_tris";
            return false;
        }
        for (int m = 0; m < d.num_meshes; ++m)
            if (d.mesh_tri_offsets[m + 1] <= d.mesh_tri_offsets[m]) { err = "empty mesh or decreasing offsets"; return false; }
        if (d.num_instances <= 0 || !d.transforms || !d.instance_meshes || !d.instance_materials) {
            err = "a device build needs instances with mesh and material ids";
            return false;
        }
        if (d.num_materials <= 0 || !d.materials) { err = "at least one material is required"; return false; }
        for (int64_t v = 0; v < d.num_vertices; ++v)
            for (int a = 0; a < 3; ++a)
                if (!std::isfinite(d.vertices_uvx[v * 4 + a])) { err = "non-finite vertex position"; return false; }
        inst_blas.assign(d.num_instances, kRefSentinel);
        inst_mat.assign(d.num_instances, 0);
        ps->inst_mesh.resize(d.num_instances);
        for (int i = 0; i < d.num_instances; ++i) {
            const int m = d.instance_meshes[i];
            if (m < 0 || m >= d.num_meshes) { err = "instance mesh out of range"; return false; }
            ps->inst_mesh[i] = m;
            inst_mat[i] = d.instance_materials[i];
        }
        ps->device_build = true;
        ps->mesh_offsets.assign(d.mesh_tri_offsets, d.mesh_tri_offsets + d.num_meshes + 1);
        ps->verts.assign(d.vertices_uvx, d.vertices_uvx + (size_t)d.num_vertices * 4);
        return fill_rest(false);
    }

    // Triangles (records only for a host-built BVH), vertices, instances,
    // materials, lights, environment.
    bool fill_rest(bool tris) {
        // triangles
        ps->num_tris = d.num_tris;
        ps->num_vertices = d.num_vertices;
        ps->tris.assign(tris ? (size_t)d.num_tris * 12 : 0, 0.f);
        ps->tri_box.assign(tris ? (size_t)d.num_tris * 6 : 0, 0.f);
        ps->vidx.assign((size_t)d.num_tris * 4, 0);
        for (int64_t k = 0; k < d.num_tris; ++k) {
            int32_t vi[3] = {d.vert_indices[k * 3], d.vert_indices[k * 3 + 1], d.vert_indices[k * 3 + 2]};
            for (int j = 0; j < 3; ++j)
                if (vi[j] < 0 || vi[j] >= d.num_vertices) { err = "vertex index out of range"; return false; }
            ps->vidx[k * 4 + 0] = vi[0];
            ps->vidx[k * 4 + 1] = vi[1];
            ps->vidx[k * 4 + 2] = vi[2];
            if (!tris) continue;
            {
                float* b = &ps->tri_box[(size_t)k * 6];
                for (int a = 0; a < 3; ++a) {
                    const float p0 = d.vertices_uvx[(size_t)vi[0] * 4 + a], p1 = d.vertices_uvx[(size_t)vi[1] * 4 + a],
                                p2 = d.vertices_uvx[(size_t)vi[2] * 4 + a];
                    b[a] = std::min(std::min(p0, p1), p2);
                    b[3 + a] = std::max(std::max(p0, p1), p2);
                }
            }
            const float* v0 = &d.vertices_uvx[(size_t)vi[0] * 4];
            const float* v1 = &d.vertices_uvx[(size_t)vi[1] * 4];
            const float* v2 = &d.vertices_uvx[(size_t)vi[2] * 4];
            float* t = &ps->tris[k * 12];
            t[0] = v0[0]; t[1] = v0[1]; t[2] = v0[2]; t[3] = u2f((uint32_t)(vi[0] / 3));
            t[4] = v1[0] - v0[0]; t[5] = v1[1] - v0[1]; t[6] = v1[2] - v0[2]; t[7] = u2f((uint32_t)k);
            t[8] = v2[0] - v0[0]; t[9] = v2[1] - v0[1]; t[10] = v2[2] - v0[2]; t[11] = 0.f;
        }
        ps->normals.assign((size_t)d.num_vertices * 4, 0.f);
        if (d.normals_uvy && d.num_vertices > 0)
            std::memcpy(ps->normals.data(), d.normals_uvy, (size_t)d.num_vertices * 16);

        // instances
        ps->num_instances = d.num_instances;
        ps->insts.assign((size_t)d.num_instances * 16, 0.f);
        ps->xforms.assign((size_t)d.num_instances * 16, 0.f);
        ps->nmats.assign((size_t)d.num_instances * 12, 0.f);
        for (int i = 0; i < d.num_instances; ++i) {
            if (inst_mat[i] < 0 || inst_mat[i] >= d.num_materials) { err = "instance material out of range"; return false; }
            put_instance(ps, i, &d.transforms[(size_t)i * 16], inst_blas[i], inst_mat[i]);
        }
        // materials
        ps->num_materials = d.num_materials;
        ps->materials.assign(d.materials, d.materials + (size_t)d.num_materials * 28);
        // texture maps: every material's ids must name a layer (or be -1)
        const int ntex = d.textures ? d.num_textures : 0;
        if (ntex > 0 && (d.tex_width <= 0 || d.tex_height <= 0)) { err = "bad texture size"; return false; }
        for (int m = 0; m < d.num_materials; ++m)
            for (int k = 24; k < 28; ++k) {
                const float id = d.materials[(size_t)m * 28 + k];
                if (!(id == -1.0f || (id >= 0.0f && id < (float)ntex && id == std::floor(id)))) {
                    err = "material " + std::to_string(m) + " names texture " + std::to_string(id) + " of " +
                          std::to_string(ntex);
                    return false;
                }
            }
        ps->tex_layers = ntex;
        ps->tex_w = ntex ? d.tex_width : 0;
        ps->tex_h = ntex ? d.tex_height : 0;
        ps->tex.assign(d.textures, d.textures + (size_t)ntex * ps->tex_w * ps->tex_h * 4);
        ps->uvs.clear();
        if (ntex > 0) {
            ps->uvs.resize((size_t)d.num_vertices * 2);
            for (int64_t v = 0; v < d.num_vertices; ++v) {
                ps->uvs[v * 2] = d.vertices_uvx[v * 4 + 3];
                ps->uvs[v * 2 + 1] = d.normals_uvy ? d.normals_uvy[v * 4 + 3] : 0.f;
            }
        }
        // lights
        ps->num_lights = d.num_lights;
        ps->lights.assign((size_t)d.num_lights * 28, 0.f);
        for (int i = 0; i < d.num_lights; ++i) put_light(&d.lights[(size_t)i * 15], &ps->lights[(size_t)i * 28]);
        // environment
        ps->hdr_w = d.hdr_width;
        ps->hdr_h = d.hdr_height;
        if (d.hdr_width > 0 && d.hdr_height > 0) {
            if (!d.hdr_cols) { err = "HDR size without colours"; return false; }
            if (!d.hdr_marginal != !d.hdr_conditional) { err = "give both HDR tables or neither"; return false; }
            size_t n = (size_t)d.hdr_width * d.hdr_height;
            ps->hdr_cols.assign(n * 4, 0.f);
            for (size_t i = 0; i < n; ++i)
                for (int c = 0; c < 3; ++c) ps->hdr_cols[i * 4 + c] = d.hdr_cols[i * 3 + c];
            ps->hdr_tables_on_device = !d.hdr_marginal;   // built at upload (kernels_hdr.hip)
            if (d.hdr_marginal) {
                ps->hdr_marg.assign(d.hdr_marginal, d.hdr_marginal + (size_t)d.hdr_height * 2);
                ps->hdr_cond.assign(d.hdr_conditional, d.hdr_conditional + n * 2);
            }
        }
        return true;
    }
};
}  // namespace

namespace {
// Collapses the packed BVH2 into a BVH4: each BVH4 node takes the two children
// of a BVH2 node and keeps replacing its internal child of largest surface
// area by that child's two children until it has four.  Every such child keeps
// the exact float box of its BVH2 node, and a BVH2 box contains its children's
// boxes, so a ray passes a collapsed child's slab test only if it passes the
// tests of the BVH2 nodes skipped above it.  A node left with room after that
// (all its children leaves) splits its multi-triangle leaves into one child
// per triangle, boxed by its vertices as the reference builder boxes
// primitives; a multi-triangle leaf that finds no such room becomes a node of
// one-triangle children.  Fewer triangle tests, the same closest hits: a box
// of exact vertex bounds is entered exactly when the reference's leaf box
// around it would lead to the triangle test (flat boxes included, which a ray
// leaving their plane at distance 0 does not enter).
// Nodes are allocated in preorder, so a node's first child follows it in memory.
#ifndef LFH_WIDEN_SAH
#define LFH_WIDEN_SAH 1   // children chosen by surface-area cost (1) or by largest area first (0)
#endif
#ifndef LFH_WIDEN_TRI_COST
#define LFH_WIDEN_TRI_COST 0.4f   // triangle test cost relative to a BVH4 node visit
#endif
struct Widen {
    const std::vector<float>& n2;
    std::vector<float>& out;
    std::unordered_map<uint32_t, std::pair<uint32_t, int>> memo;  // BLAS root -> (ref4, need)
    const std::vector<float>& insts;
    const std::vector<float>& tbox;   // 6 floats per triangle slot: its vertices' box

    struct Child {
        float box[6];
        uint32_t ref;
    };
    static float area(const float* b) {
        float dx = b[3] - b[0], dy = b[4] - b[1], dz = b[5] - b[2];
        return dx * dy + dy * dz + dz * dx;
    }
    void children_of(uint32_t r, Child* c) const {
        const float* o = &n2[(size_t)r * 16];
        std::memcpy(c[0].box, o, 6 * sizeof(float));
        std::memcpy(c[1].box, o + 6, 6 * sizeof(float));
        std::memcpy(&c[0].ref, o + 12, 4);
        std::memcpy(&c[1].ref, o + 13, 4);
    }
    static bool internal(uint32_t r) { return r != lfhd::kRefSentinel && !lfhd::ref_is_leaf(r) && !lfhd::ref_is_inst(r); }
    void tri_box(uint32_t slot, float* b) const { std::memcpy(b, &tbox[(size_t)slot * 6], 6 * sizeof(float)); }

    // Returns the BVH4 ref of BVH2 ref r and, in *need, the largest number of
    // stack entries a traversal below it can hold.
    uint32_t build(uint32_t r, int* need) {
        if (lfhd::ref_is_inst(r)) {
            uint32_t inst = r & 0x3FFFFFFFu, blas;
            std::memcpy(&blas, &insts[(size_t)inst * 16 + 12], 4);
            *need = 1 + blas_ref(blas).second;  // the BLAS-return marker, then the BLAS
            return r;
        }
#ifndef LFH_LEAF_NODE_MIN
#define LFH_LEAF_NODE_MIN 2   // on the 1M-triangle soup: 2 -> 1528, 3 -> 1421, off -> 1287 Mrays/s
#endif
        if (LFH_LEAF_NODE_MIN > 0 && r != lfhd::kRefSentinel && lfhd::ref_is_leaf(r) &&
            (int)lfhd::leaf_count(r) >= LFH_LEAF_NODE_MIN && lfhd::leaf_count(r) <= 4) {
            // a multi-triangle leaf the parent had no room to split: a node of
            // one-triangle children
            const uint32_t first = lfhd::leaf_first(r), cnt = lfhd::leaf_count(r);
            const size_t idx = out.size() / 32;
            out.resize(out.size() + 32, 0.0f);
            float* o = &out[idx * 32];
            for (uint32_t k = 0; k < 4; ++k) {
                float b[6] = {1.f, 1.f, 1.f, -1.f, -1.f, -1.f};
                uint32_t ref = lfhd::kRefSentinel;
                if (k < cnt) {
                    tri_box(first + k, b);
                    ref = lfhd::make_leaf(first + k, 1);
                }
                for (int a = 0; a < 6; ++a) o[a * 4 + k] = b[a];
                std::memcpy(&o[24 + k], &ref, 4);
            }
            *need = (int)cnt - 1;
            return (uint32_t)idx;
        }
        if (!internal(r)) {
            *need = 0;
            return r;
        }
        Child ch[4];
        int n = 2;
        if (LFH_WIDEN_SAH) {
            n = sah_children(r, ch);
        } else {
            children_of(r, ch);
        }
        while (!LFH_WIDEN_SAH && n < 4) {
            int best = -1;
            float ba = -1.0f;
            for (int k = 0; k < n; ++k)
                if (internal(ch[k].ref) && area(ch[k].box) > ba) {
                    ba = area(ch[k].box);
                    best = k;
                }
            if (best < 0) break;
            Child sub[2];
            children_of(ch[best].ref, sub);
            for (int k = n; k > best + 1; --k) ch[k] = ch[k - 1];
            ch[best] = sub[0];
            ch[best + 1] = sub[1];
            ++n;
        }
        while (!LFH_WIDEN_SAH && n < 4) {
            int best = -1;
            float ba = -1.0f;
            for (int k = 0; k < n; ++k) {
                const uint32_t r = ch[k].ref;
                if (r != lfhd::kRefSentinel && lfhd::ref_is_leaf(r) && lfhd::leaf_count(r) >= 2 &&
                    n - 1 + (int)lfhd::leaf_count(r) <= 4 && area(ch[k].box) > ba) {
                    ba = area(ch[k].box);
                    best = k;
                }
            }
            if (best < 0) break;
            const uint32_t first = lfhd::leaf_first(ch[best].ref), cnt = lfhd::leaf_count(ch[best].ref);
            Child sub[4];
            for (uint32_t t = 0; t < cnt; ++t) {
                sub[t].ref = lfhd::make_leaf(first + t, 1);
                tri_box(first + t, sub[t].box);
            }
            for (int k = n - 1; k > best; --k) ch[k + (int)cnt - 1] = ch[k];
            for (uint32_t t = 0; t < cnt; ++t) ch[best + (int)t] = sub[t];
            n += (int)cnt - 1;
        }
        const size_t idx = out.size() / 32;
        out.resize(out.size() + 32, 0.0f);
        uint32_t refs[4] = {lfhd::kRefSentinel, lfhd::kRefSentinel, lfhd::kRefSentinel, lfhd::kRefSentinel};
        int worst = 0;
        for (int k = 0; k < n; ++k) {
            int nk = 0;
            refs[k] = build(ch[k].ref, &nk);
            worst = std::max(worst, nk);
        }
        float* o = &out[idx * 32];
        for (int k = 0; k < 4; ++k) {
            const float* b = k < n ? ch[k].box : nullptr;
            for (int a = 0; a < 6; ++a) o[a * 4 + k] = b ? b[a] : (a < 3 ? 1.0f : -1.0f);
            std::memcpy(&o[24 + k], &refs[k], 4);
        }
        *need = (n - 1) + worst;
        return (uint32_t)idx;
    }
    // ---- surface-area-cost collapse ----------------------------------------
    // slots(c, j): the least expected cost of representing child c (a BVH2
    // ref seen through its box in the parent) with at most j of a BVH4 node's
    // child slots.  Costs are box areas times the work a ray does once it
    // enters the box: kCostNode for a BVH4 node visit, kCostTri for a triangle
    // test (the surface-area heuristic's conditional probabilities, all
    // relative to the same root).  A BVH2 leaf of c triangles takes c slots
    // (one child per triangle, boxed by its vertices) or one slot as a node of
    // one-triangle children; an internal BVH2 node takes one slot as a BVH4
    // node or lends its children to the parent's slots.  Any such choice of
    // children keeps each child's exact box from the BVH2 tree (or the
    // triangle's vertex box inside its leaf's box), so the triangles a ray
    // tests are among those the BVH2 traversal tests.
    static constexpr float kCostNode = 1.0f;
    static constexpr float kCostTri = LFH_WIDEN_TRI_COST;
    static constexpr float kCostInst = 0.5f;
    static constexpr float kInf = 3.0e38f;
    std::vector<float> dp;   // per internal BVH2 node: [node cost, split(2), split(3), split(4)]; < 0: not yet computed
    float box_area(const float* b) const { return std::max(area(b), 0.0f); }
    float leaf_tris_cost(uint32_t r) const {
        float c = 0.0f;
        for (uint32_t k = 0; k < lfhd::leaf_count(r); ++k) {
            float b[6];
            tri_box(lfhd::leaf_first(r) + k, b);
            c += box_area(b) * kCostTri;
        }
        return c;
    }
    // cost of child ref r with box area a using at most j slots (j = 0..4)
    float slots(uint32_t r, float a, int j) {
        if (r == lfhd::kRefSentinel) return 0.0f;
        if (j <= 0) return kInf;
        if (lfhd::ref_is_inst(r)) return a * kCostInst;
        if (lfhd::ref_is_leaf(r)) {
            const int c = (int)lfhd::leaf_count(r);
            const float tris = leaf_tris_cost(r);
            return (j >= c) ? tris : a * kCostNode + tris;
        }
        const float* d = sah(r, a);
        float best = d[0];
        for (int k = 2; k <= j; ++k) best = std::min(best, d[k - 1]);
        return best;
    }
    const float* sah(uint32_t r, float a) {
        if (dp.empty()) dp.assign(n2.size() / 16 * 4, -1.0f);
        float* d = &dp[(size_t)r * 4];
        if (d[0] >= 0.0f) return d;
        Child c[2];
        children_of(r, c);
        const float al = box_area(c[0].box), ar = box_area(c[1].box);
        float sl[5], sr[5];
        for (int k = 0; k <= 4; ++k) {
            sl[k] = slots(c[0].ref, al, k);
            sr[k] = slots(c[1].ref, ar, k);
        }
        for (int j = 2; j <= 4; ++j) {
            float b = kInf;
            for (int k = 0; k <= j; ++k) b = std::min(b, sl[k] + sr[j - k]);
            d[j - 1] = b;
        }
        d[0] = a * kCostNode + d[3];
        return d;
    }
    // Appends the children that child ref c (box b) contributes with at most j slots.
    void expand(const Child& c, int j, Child* out, int* n) {
        if (c.ref == lfhd::kRefSentinel) return;
        const float a = box_area(c.box);
        if (lfhd::ref_is_leaf(c.ref)) {
            const int cnt = (int)lfhd::leaf_count(c.ref);
            if (cnt > 1 && cnt <= j) {
                for (int t = 0; t < cnt; ++t) {
                    out[*n].ref = lfhd::make_leaf(lfhd::leaf_first(c.ref) + (uint32_t)t, 1);
                    tri_box(lfhd::leaf_first(c.ref) + (uint32_t)t, out[*n].box);
                    ++*n;
                }
                return;
            }
            out[(*n)++] = c;   // one slot (a multi-triangle leaf becomes a node of one-triangle children)
            return;
        }
        if (!internal(c.ref) || j <= 1) {
            out[(*n)++] = c;
            return;
        }
        const float* d = sah(c.ref, a);
        int bj = 1;
        float best = d[0];
        for (int k = 2; k <= j; ++k)
            if (d[k - 1] < best) {
                best = d[k - 1];
                bj = k;
            }
        if (bj == 1) {
            out[(*n)++] = c;
            return;
        }
        split_into(c.ref, bj, out, n);
    }
    // Splits internal node r's j slots between its two children as the cost table chose.
    void split_into(uint32_t r, int j, Child* out, int* n) {
        Child c[2];
        children_of(r, c);
        const float al = box_area(c[0].box), ar = box_area(c[1].box);
        int bk = -1;
        float best = kInf;
        for (int k = 0; k <= j; ++k) {
            const float v = slots(c[0].ref, al, k) + slots(c[1].ref, ar, j - k);
            if (v < best) {
                best = v;
                bk = k;
            }
        }
        expand(c[0], bk, out, n);
        expand(c[1], j - bk, out, n);
    }
    // The children of internal BVH2 node r as a BVH4 node (2..4 of them).
    int sah_children(uint32_t r, Child* ch) {
        int n = 0;
        split_into(r, 4, ch, &n);
        return n;
    }

    std::pair<uint32_t, int> blas_ref(uint32_t r) {
        auto it = memo.find(r);
        if (it != memo.end()) return it->second;
        int need = 0;
        uint32_t r4 = build(r, &need);
        memo[r] = {r4, need};
        return {r4, need};
    }
};
}  // namespace

std::string widen_bvh(PackedScene* ps) {
    ps->nodes4.clear();
    Widen w{ps->nodes, ps->nodes4, {}, ps->insts, ps->tri_box};
    // BLAS trees first, in instance order, then the TLAS
    for (int i = 0; i < ps->num_instances; ++i) {
        uint32_t blas;
        std::memcpy(&blas, &ps->insts[(size_t)i * 16 + 12], 4);
        w.blas_ref(blas);
    }
    const size_t blas4 = ps->nodes4.size() / 32;
    int need = 0;
    ps->root4 = w.build(ps->root, &need);
    // The TLAS part gets room for one BVH4 node per TLAS BVH2 node (an upper
    // bound for any boxes), so instance updates never change the array size.
    size_t tlas2 = 0;
    std::vector<uint32_t> st;
    if (Widen::internal(ps->root)) st.push_back(ps->root);
    while (!st.empty()) {
        uint32_t r = st.back();
        st.pop_back();
        ++tlas2;
        Widen::Child c[2];
        w.children_of(r, c);
        for (auto& k : c)
            if (Widen::internal(k.ref)) st.push_back(k.ref);
    }
    const size_t cap = blas4 + std::max<size_t>(tlas2, 1);
    if (ps->nodes4.size() / 32 < cap) ps->nodes4.resize(cap * 32, 0.0f);
    if ((uint64_t)ps->nodes4.size() / 32 >= (1ull << 30)) return "too many BVH4 nodes";
    ps->stack_need = need;
    if (need > (int)lfhd::kSpillStride)
        return "BVH too deep for the traversal stack (" + std::to_string(need) + " entries needed, " +
               std::to_string(lfhd::kSpillStride) + " available)";
    ps->insts4 = ps->insts;
    for (int i = 0; i < ps->num_instances; ++i) {
        uint32_t blas;
        std::memcpy(&blas, &ps->insts[(size_t)i * 16 + 12], 4);
        uint32_t b4 = w.memo.count(blas) ? w.memo[blas].first : blas;
        std::memcpy(&ps->insts4[(size_t)i * 16 + 12], &b4, 4);
    }
    return std::string();
}

std::string pack_scene(const lfh_scene_desc& d, PackedScene* out) {
    Builder b(d, out);
    if (!d.nodes) {
        if (!b.run_device()) return b.err.empty() ? std::string("invalid scene") : b.err;
        return std::string();
    }
    if (!b.run()) return b.err.empty() ? std::string("invalid scene") : b.err;
    return widen_bvh(out);
}

void device_insts4(PackedScene* ps) {
    ps->insts4 = ps->insts;
    for (int i = 0; i < ps->num_instances; ++i) {
        const uint32_t r = ps->mesh_ref4[ps->inst_mesh[i]];
        std::memcpy(&ps->insts4[(size_t)i * 16 + 12], &r, 4);
    }
}

// Renderer::Update re-uploads transforms, materials and the TLAS node range
// (Renderer.cpp:192-203); here the TLAS nodes are re-packed from the new flat
// nodes, which must keep the BLAS part and the packed node count unchanged.
std::string repack_instances(PackedScene* ps, const float* transforms, const float* tlas_nodes, int64_t first_node,
                             int64_t count, const float* materials) {
    if (materials) ps->materials.assign(materials, materials + (size_t)ps->num_materials * 28);
    std::vector<uint32_t> blas(ps->num_instances);
    std::vector<int32_t> mat(ps->num_instances);
    for (int i = 0; i < ps->num_instances; ++i) {
        uint32_t b;
        std::memcpy(&b, &ps->insts[(size_t)i * 16 + 12], 4);
        blas[i] = b;
        uint32_t m;
        std::memcpy(&m, &ps->insts[(size_t)i * 16 + 13], 4);
        mat[i] = (int32_t)m;
    }
    if (tlas_nodes && count > 0) {
        const int64_t N = (int64_t)ps->flat_to_packed.size();
        if (first_node < 0 || first_node + count > N) return "TLAS node range out of bounds";
        // RebuildInstances may change the TLAS's shape (which flat nodes are
        // internal) but not its size: the packed slots of its internal nodes
        // are handed to the new internal nodes in flat order.
        std::vector<int32_t> slots;
        for (int64_t i = 0; i < count; ++i) {
            const int64_t n = first_node + i;
            if (ps->flat_lrleaf[n * 3 + 2] == 0 && ps->flat_to_packed[n] >= 0) slots.push_back(ps->flat_to_packed[n]);
            ps->flat_to_packed[n] = -1;
        }
        std::sort(slots.begin(), slots.end());
        for (int64_t i = 0; i < count; ++i)
            for (int k = 0; k < 3; ++k) ps->flat_lrleaf[(first_node + i) * 3 + k] = (int32_t)tlas_nodes[i * 9 + 6 + k];
        // only nodes reachable from the top level are packed (the spare
        // nodes of the TLAS area are zero and look internal)
        std::vector<uint8_t> reach((size_t)count, 0);
        std::vector<int64_t> st{ps->top_level_index};
        while (!st.empty()) {
            const int64_t n = st.back();
            st.pop_back();
            if (n < first_node || n >= first_node + count || reach[n - first_node]) continue;
            reach[n - first_node] = 1;
            if (ps->flat_lrleaf[n * 3 + 2] == 0) {
                st.push_back(ps->flat_lrleaf[n * 3]);
                st.push_back(ps->flat_lrleaf[n * 3 + 1]);
            }
        }
        size_t used = 0;
        for (int64_t i = 0; i < count; ++i) {
            const int64_t n = first_node + i;
            if (!reach[i] || ps->flat_lrleaf[n * 3 + 2] != 0) continue;
            if (used >= slots.size()) return "TLAS update adds internal nodes";
            ps->flat_to_packed[n] = slots[used++];
        }
        if (used != slots.size()) return "TLAS update removes internal nodes";
        if (ps->top_level_index >= first_node && ps->top_level_index < first_node + count) {
            const int32_t lf = ps->flat_lrleaf[(int64_t)ps->top_level_index * 3 + 2];
            ps->root = lf < 0 ? lfhd::make_inst((uint32_t)(-(int64_t)lf - 1))
                              : (uint32_t)ps->flat_to_packed[ps->top_level_index];
        }
        for (int64_t i = 0; i < count; ++i) {
            int64_t n = first_node + i;
            int32_t leaf = ps->flat_lrleaf[n * 3 + 2];
            if (reach[i] && leaf < 0) {
                int64_t inst = -(int64_t)leaf - 1;
                if (inst >= ps->num_instances) return "TLAS leaf instance out of range";
                mat[inst] = ps->flat_lrleaf[n * 3 + 1];
            }
        }
        auto box_of = [&](int64_t c, float* dst) -> bool {
            if (c >= first_node && c < first_node + count) {
                std::memcpy(dst, &tlas_nodes[(c - first_node) * 9], 6 * sizeof(float));
                return true;
            }
            return false;
        };
        for (int64_t i = 0; i < count; ++i) {
            int64_t n = first_node + i;
            if (!reach[i] || ps->flat_lrleaf[n * 3 + 2] != 0) continue;
            int32_t p = ps->flat_to_packed[n];
            if (p < 0) return "TLAS update changes the node structure";
            int64_t L = ps->flat_lrleaf[n * 3], R = ps->flat_lrleaf[n * 3 + 1];
            float a[6], b[6];
            if (!box_of(L, a) || !box_of(R, b)) return "TLAS update references nodes outside the range";
            auto ref = [&](int64_t c, uint32_t* r) -> bool {
                int32_t lf = ps->flat_lrleaf[c * 3 + 2];
                if (lf < 0) { *r = lfhd::make_inst((uint32_t)(-(int64_t)lf - 1)); return true; }
                if (lf > 0) return false;
                if (ps->flat_to_packed[c] < 0) return false;
                *r = (uint32_t)ps->flat_to_packed[c];
                return true;
            };
            uint32_t rl, rr;
            if (!ref(L, &rl) || !ref(R, &rr)) return "TLAS update changes the node structure";
            float* o = &ps->nodes[(size_t)p * 16];
            o[0] = a[0]; o[1] = a[1]; o[2] = a[2]; o[3] = a[3];
            o[4] = a[4]; o[5] = a[5]; o[6] = b[0]; o[7] = b[1];
            o[8] = b[2]; o[9] = b[3]; o[10] = b[4]; o[11] = b[5];
            o[12] = u2f(rl); o[13] = u2f(rr);
        }
    }
    if (transforms)
        for (int i = 0; i < ps->num_instances; ++i) put_instance(ps, i, &transforms[(size_t)i * 16], blas[i], mat[i]);
    else
        for (int i = 0; i < ps->num_instances; ++i) {
            uint32_t m = (uint32_t)mat[i];
            std::memcpy(&ps->insts[(size_t)i * 16 + 13], &m, 4);
        }
    if (ps->device_build) return std::string();  // the caller rebuilds the TLAS on the device
    const size_t n4 = ps->nodes4.size();
    std::string err = widen_bvh(ps);
    if (!err.empty()) return err;
    if (ps->nodes4.size() != n4) return "TLAS update changes the node structure";
    return std::string();
}

template <class T>
static hipError_t up(DevBuf& b, const std::vector<T>& v, hipStream_t s) {
    hipError_t e = b.alloc(v.size() * sizeof(T));
    if (e != hipSuccess) return e;
    if (!v.empty()) return hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s);
    return hipSuccess;
}

hipError_t upload_scene(PackedScene& ps, DeviceScene* ds, hipStream_t s, std::string* err) {
    hipError_t e;
    if (!ps.device_build) {
        if ((e = up(ds->nodes, ps.nodes4, s)) != hipSuccess) return e;
        if ((e = up(ds->tris, ps.tris, s)) != hipSuccess) return e;
        if ((e = up(ds->insts, ps.insts4, s)) != hipSuccess) return e;
    }
    if ((e = up(ds->xforms, ps.xforms, s)) != hipSuccess) return e;
    if ((e = up(ds->nmats, ps.nmats, s)) != hipSuccess) return e;
    if ((e = up(ds->vidx, ps.vidx, s)) != hipSuccess) return e;
    if ((e = up(ds->normals, ps.normals, s)) != hipSuccess) return e;
    if ((e = up(ds->materials, ps.materials, s)) != hipSuccess) return e;
    if ((e = up(ds->lights, ps.lights, s)) != hipSuccess) return e;
    if ((e = up(ds->hdr_cols, ps.hdr_cols, s)) != hipSuccess) return e;
    if ((e = up(ds->hdr_marg, ps.hdr_marg, s)) != hipSuccess) return e;
    if ((e = up(ds->hdr_cond, ps.hdr_cond, s)) != hipSuccess) return e;
    if (ps.hdr_tables_on_device && ps.hdr_w > 0) {
        const size_t n = (size_t)ps.hdr_w * ps.hdr_h;
        DevBuf scratch;
        if ((e = ds->hdr_marg.alloc((size_t)ps.hdr_h * 8)) != hipSuccess) return e;
        if ((e = ds->hdr_cond.alloc(n * 8)) != hipSuccess) return e;
        if ((e = scratch.alloc(lfhd::hdr_scratch_floats(ps.hdr_w, ps.hdr_h) * 4)) != hipSuccess) return e;
        if ((e = lfhd::hdr_tables(ds->hdr_cols.as<const float4>(), ps.hdr_w, ps.hdr_h, scratch.as<float>(),
                                  ds->hdr_marg.as<float2>(), ds->hdr_cond.as<float2>(), s)) != hipSuccess)
            return e;
        if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;   // before the scratch is freed
    }
    if ((e = up(ds->tex, ps.tex, s)) != hipSuccess) return e;
    if ((e = up(ds->uvs, ps.uvs, s)) != hipSuccess) return e;
    if (ps.device_build) {
        std::string m = device_build(&ps, ds, s);
        if (!m.empty()) {
            if (err) *err = m;
            return hipErrorInvalidValue;
        }
    }
    return hipStreamSynchronize(s);
}

lfhd::TraceScene trace_view(const PackedScene& ps, const DeviceScene& ds) {
    lfhd::TraceScene t;
    t.nodes = ds.nodes.as<const float4>();
    t.tris = ds.tris.as<const float4>();
    t.insts = ds.insts.as<const float4>();
    t.root = ps.root4;
    t.spill_stride = lfhd::kSpillStride;
    return t;
}

lfhd::ShadeScene shade_view(const PackedScene& ps, const DeviceScene& ds, const lfh_options& o) {
    lfhd::ShadeScene s;
    s.tr = trace_view(ps, ds);
    s.xforms = ds.xforms.as<const float4>();
    s.nmats = ds.nmats.as<const float4>();
    s.vidx = ds.vidx.as<const int4>();
    s.normals = ds.normals.as<const float4>();
    s.materials = ds.materials.as<const float4>();
    s.lights = ds.lights.as<const float4>();
    s.hdr_cols = ds.hdr_cols.as<const float4>();
    s.hdr_marg = ds.hdr_marg.as<const float2>();
    s.hdr_cond = ds.hdr_cond.as<const float2>();
    s.hdr_w = ps.hdr_w;
    s.hdr_h = ps.hdr_h;
    s.hdr_res = (float)(ps.hdr_w * ps.hdr_h);
    s.hdr_mult = o.hdr_multiplier;
    s.num_lights = ps.num_lights;
    s.max_depth = o.max_depth;
    s.rr = o.enable_rr;
    s.rr_depth = o.rr_depth;
    bool env = o.use_env_map && ps.hdr_w > 0;
    s.use_env = (env && !o.use_constant_bg) ? 1 : 0;
    s.const_bg = o.use_constant_bg ? 1 : 0;
    s.bg[0] = o.bg_color[0];
    s.bg[1] = o.bg_color[1];
    s.bg[2] = o.bg_color[2];
    s.estimator = o.estimator;
    s.tex_layers = ps.tex_layers;
    s.tex_w = ps.tex_w;
    s.tex_h = ps.tex_h;
    s.tex = ps.tex_layers ? ds.tex.as<const uint32_t>() : nullptr;
    s.uvs = ps.tex_layers ? ds.uvs.as<const float2>() : nullptr;
    return s;
}

}  // namespace lfh
