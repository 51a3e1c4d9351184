// Host-side scene surface: builders, flattening, OBJ ingest, HDR tables,
// camera.  See host_scene.h for the reference types mirrored here.
#include "host_scene.h"
#include "host_image.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <utility>

namespace lfh {

static const float kPi = 3.14159265358979323846f;

// ---------------------------------------------------------------------------
// BVH builders
// ---------------------------------------------------------------------------

void Bvh::Build(const BBox* bounds, int numbounds) {
    nodes_.clear();
    packed_.clear();
    root_ = -1;
    height_ = 0;
    bounds_ = BBox();
    for (int i = 0; i < numbounds; ++i) bounds_.grow(bounds[i]);
    if (numbounds <= 0) return;

    if (policy_ == kObjectSah) {
        // Primitive references with their centroids (split_bvh.cpp:8-31).
        refs_.resize(numbounds);
        BBox cb;
        for (int i = 0; i < numbounds; ++i) {
            refs_[i] = PrimRef{bounds[i], bounds[i].center(), i};
            cb.grow(bounds[i].center());
        }
        nodes_.reserve(2 * (size_t)numbounds);
        packed_.reserve(numbounds);
        Request init{0, numbounds, bounds_, cb, 0};
        root_ = BuildSah(init);
        refs_.clear();
        refs_.shrink_to_fit();
    } else {
        in_bounds_ = bounds;
        centroids_.resize(numbounds);
        primindices_.resize(numbounds);
        BBox cb;
        for (int i = 0; i < numbounds; ++i) {
            primindices_[i] = i;
            Vec3 c = bounds[i].center();
            cb.grow(c);
            centroids_[i] = c;
        }
        nodes_.reserve(2 * (size_t)numbounds);
        Request init{0, numbounds, bounds_, cb, 0};
        root_ = BuildMedian(init);
        in_bounds_ = nullptr;
        centroids_.clear();
        primindices_.clear();
    }
}

// Binned object-SAH split over centroid bins (split_bvh.cpp:299-402): cost =
// traversal_cost + (nL * area(L) + nR * area(R)) / area(parent); the split
// plane is the upper border of the best bin.  NaN split = no candidate.
Bvh::SahSplit Bvh::FindObjectSahSplit(const Request& req) const {
    SahSplit split;
    split.dim = 0;
    split.split = std::numeric_limits<float>::quiet_NaN();
    split.sah = std::numeric_limits<float>::max();
    float sah = std::numeric_limits<float>::max();
    int splitidx = -1;

    Vec3 cext = req.centroid_bounds.extents();
    if (Vec3::Dot(cext, cext) == 0.f) return split;

    struct Bin { BBox bounds; int count; };
    std::vector<Bin> bins(num_bins_);
    std::vector<BBox> rightbounds(num_bins_ - 1);
    const float invarea = 1.f / req.bounds.surface_area();
    const Vec3 rootmin = req.centroid_bounds.pmin;

    for (int axis = 0; axis < 3; ++axis) {
        const float rootminc = rootmin[axis];
        const float rng = cext[axis];
        const float invrng = 1.f / rng;
        if (rng == 0.f) continue;
        for (int i = 0; i < num_bins_; ++i) { bins[i].count = 0; bins[i].bounds = BBox(); }
        const float nb = static_cast<float>(num_bins_);
        const float nbm1 = static_cast<float>(num_bins_ - 1);
        for (int i = req.startidx; i < req.startidx + req.numprims; ++i) {
            float f = nb * ((refs_[i].center[axis] - rootminc) * invrng);
            int b = (int)std::min<float>(f, nbm1);
            ++bins[b].count;
            bins[b].bounds.grow(refs_[i].bounds);
        }
        BBox rightbox;
        for (int i = num_bins_ - 1; i > 0; --i) {
            rightbox.grow(bins[i].bounds);
            rightbounds[i - 1] = rightbox;
        }
        BBox leftbox;
        int leftcount = 0, rightcount = req.numprims;
        for (int i = 0; i < num_bins_ - 1; ++i) {
            leftbox.grow(bins[i].bounds);
            leftcount += bins[i].count;
            rightcount -= bins[i].count;
            float s = traversal_cost_ +
                      (leftcount * leftbox.surface_area() + rightcount * rightbounds[i].surface_area()) *
                          invarea;
            if (s < sah) {
                split.dim = axis;
                splitidx = i;
                sah = s;
            }
        }
    }
    if (splitidx != -1) {
        split.split = rootmin[split.dim] + (splitidx + 1) * (cext[split.dim] / num_bins_);
        split.sah = sah;
    }
    return split;
}

// Two-sided in-place partition around `border` on `axis`, growing the child
// boxes as it goes.  The side that keeps the "< border" centroids alternates
// with the parity of (start + count) (split_bvh.cpp:112-151, bvh.cpp:132-204);
// the order of swaps decides the final primitive order, so it is kept as is.
template <class CenterOf, class BoundsOf, class Swap>
static int PartitionRange(int start, int num, int axis, float border, bool near2far,
                          CenterOf center, BoundsOf bnd, Swap swp,
                          BBox& lb, BBox& rb, BBox& lcb, BBox& rcb) {
    auto lt = [](float a, float b) { return a < b; };
    auto ge = [](float a, float b) { return a >= b; };
    int first = start, last = start + num;
    for (;;) {
        while (first != last && (near2far ? lt(center(first)[axis], border) : ge(center(first)[axis], border))) {
            lb.grow(bnd(first));
            lcb.grow(center(first));
            ++first;
        }
        if (first == last--) break;
        rb.grow(bnd(first));
        rcb.grow(center(first));
        while (first != last && (near2far ? ge(center(last)[axis], border) : lt(center(last)[axis], border))) {
            rb.grow(bnd(last));
            rcb.grow(center(last));
            --last;
        }
        if (first == last) break;
        lb.grow(bnd(last));
        lcb.grow(center(last));
        swp(first++, last);
    }
    return first;
}

int Bvh::BuildSah(const Request& req) {
    height_ = std::max(height_, req.level);
    int self = (int)nodes_.size();
    nodes_.emplace_back();
    nodes_[self].bounds = req.bounds;

    if (req.numprims < 4) {
        nodes_[self].leaf = true;
        nodes_[self].startidx = (int)packed_.size();
        nodes_[self].numprims = req.numprims;
        for (int i = req.startidx; i < req.startidx + req.numprims; ++i) packed_.push_back(refs_[i].idx);
        return self;
    }

    int axis = req.centroid_bounds.maxdim();
    float border = req.centroid_bounds.center()[axis];
    SahSplit os = FindObjectSahSplit(req);
    if (!std::isnan(os.split)) {
        border = os.split;
        axis = os.dim;
    }

    BBox lb, rb, lcb, rcb;
    int splitidx = req.startidx;
    bool near2far = ((req.numprims + req.startidx) & 0x1) != 0;
    if (req.centroid_bounds.extents()[axis] > 0.f) {
        splitidx = PartitionRange(
            req.startidx, req.numprims, axis, border, near2far,
            [&](int i) -> const Vec3& { return refs_[i].center; },
            [&](int i) -> const BBox& { return refs_[i].bounds; },
            [&](int a, int b) { std::swap(refs_[a], refs_[b]); }, lb, rb, lcb, rcb);
    }
    if (splitidx == req.startidx || splitidx == req.startidx + req.numprims) {
        splitidx = req.startidx + (req.numprims >> 1);
        for (int i = req.startidx; i < splitidx; ++i) { lb.grow(refs_[i].bounds); lcb.grow(refs_[i].center); }
        for (int i = splitidx; i < req.startidx + req.numprims; ++i) { rb.grow(refs_[i].bounds); rcb.grow(refs_[i].center); }
    }
    Request left{req.startidx, splitidx - req.startidx, lb, lcb, req.level + 1};
    Request right{splitidx, req.numprims - (splitidx - req.startidx), rb, rcb, req.level + 1};
    // Right subtree first: its leaves come first in the packed index order.
    int r = BuildSah(right);
    int l = BuildSah(left);
    nodes_[self].lc = l;
    nodes_[self].rc = r;
    return self;
}

int Bvh::BuildMedian(const Request& req) {
    height_ = std::max(height_, req.level);
    int self = (int)nodes_.size();
    nodes_.emplace_back();
    nodes_[self].bounds = req.bounds;

    if (req.numprims < 2) {
        nodes_[self].leaf = true;
        nodes_[self].startidx = (int)packed_.size();
        nodes_[self].numprims = req.numprims;
        for (int i = 0; i < req.numprims; ++i) packed_.push_back(primindices_[req.startidx + i]);
        return self;
    }
    int axis = req.centroid_bounds.maxdim();
    float border = req.centroid_bounds.center()[axis];

    BBox lb, rb, lcb, rcb;
    int splitidx = req.startidx;
    bool near2far = ((req.numprims + req.startidx) & 0x1) != 0;
    if (req.centroid_bounds.extents()[axis] > 0.f) {
        splitidx = PartitionRange(
            req.startidx, req.numprims, axis, border, near2far,
            [&](int i) -> const Vec3& { return centroids_[primindices_[i]]; },
            [&](int i) -> const BBox& { return in_bounds_[primindices_[i]]; },
            [&](int a, int b) { std::swap(primindices_[a], primindices_[b]); }, lb, rb, lcb, rcb);
    }
    if (splitidx == req.startidx || splitidx == req.startidx + req.numprims) {
        splitidx = req.startidx + (req.numprims >> 1);
        for (int i = req.startidx; i < splitidx; ++i) {
            lb.grow(in_bounds_[primindices_[i]]);
            lcb.grow(centroids_[primindices_[i]]);
        }
        for (int i = splitidx; i < req.startidx + req.numprims; ++i) {
            rb.grow(in_bounds_[primindices_[i]]);
            rcb.grow(centroids_[primindices_[i]]);
        }
    }
    Request left{req.startidx, splitidx - req.startidx, lb, lcb, req.level + 1};
    Request right{splitidx, req.numprims - (splitidx - req.startidx), rb, rcb, req.level + 1};
    int l = BuildMedian(left);
    int r = BuildMedian(right);
    nodes_[self].lc = l;
    nodes_[self].rc = r;
    return self;
}

// ---------------------------------------------------------------------------
// Flattening: preorder, left child first; all BLAS node blocks mesh by mesh,
// then the TLAS at topLevelIndex.  Leaf LRLeaf: BLAS (firstTri, count, 1);
// TLAS (blasRoot, materialID, -instance-1)  (bvh_translator.cpp:35-140).
// ---------------------------------------------------------------------------

int BvhTranslator::ProcessBLASNodes(const Bvh& bvh, int n, int triOffset) {
    const Bvh::Node& node = bvh.nodes()[n];
    int index = curNode_;
    nodes[index].bboxmin = node.bounds.pmin;
    nodes[index].bboxmax = node.bounds.pmax;
    nodes[index].LRLeaf.z = 0;
    if (node.leaf) {
        nodes[index].LRLeaf.x = (float)(triOffset + node.startidx);
        nodes[index].LRLeaf.y = (float)node.numprims;
        nodes[index].LRLeaf.z = 1;
    } else {
        curNode_++;
        nodes[index].LRLeaf.x = (float)ProcessBLASNodes(bvh, node.lc, triOffset);
        curNode_++;
        nodes[index].LRLeaf.y = (float)ProcessBLASNodes(bvh, node.rc, triOffset);
    }
    return index;
}

int BvhTranslator::ProcessTLASNodes(const Bvh& tlas, int n, const std::vector<MeshInstance>& inst) {
    const Bvh::Node& node = tlas.nodes()[n];
    int index = curNode_;
    nodes[index].bboxmin = node.bounds.pmin;
    nodes[index].bboxmax = node.bounds.pmax;
    nodes[index].LRLeaf.z = 0;
    if (node.leaf) {
        int instanceIndex = tlas.packed_indices()[node.startidx];
        int meshIndex = inst[instanceIndex].meshID;
        nodes[index].LRLeaf.x = (float)bvhRootStartIndices_[meshIndex];
        nodes[index].LRLeaf.y = (float)inst[instanceIndex].materialID;
        nodes[index].LRLeaf.z = (float)(-instanceIndex - 1);
    } else {
        curNode_++;
        nodes[index].LRLeaf.x = (float)ProcessTLASNodes(tlas, node.lc, inst);
        curNode_++;
        nodes[index].LRLeaf.y = (float)ProcessTLASNodes(tlas, node.rc, inst);
    }
    return index;
}

void BvhTranslator::Process(const Bvh* tlas, const std::vector<Mesh*>& meshes,
                            const std::vector<MeshInstance>& instances) {
    int nodeCnt = 0;
    for (const Mesh* m : meshes) nodeCnt += m->bvh.node_count();
    topLevelIndex = nodeCnt;
    // Room for 2 * instances TLAS nodes (one more than a full binary tree
    // needs; the spare stays zero like the reference's resize).
    nodeCnt += 2 * (int)instances.size();
    nodes.assign(nodeCnt, FlatNode());
    bvhRootStartIndices_.clear();
    int bvhRootIndex = 0, curTriIndex = 0;
    for (const Mesh* m : meshes) {
        curNode_ = bvhRootIndex;
        bvhRootStartIndices_.push_back(bvhRootIndex);
        bvhRootIndex += m->bvh.node_count();
        if (m->bvh.root() >= 0) ProcessBLASNodes(m->bvh, m->bvh.root(), curTriIndex);
        curTriIndex += (int)m->bvh.packed_indices().size();
    }
    curNode_ = topLevelIndex;
    if (tlas->root() >= 0) ProcessTLASNodes(*tlas, tlas->root(), instances);
}

void BvhTranslator::UpdateTLAS(const Bvh* tlas, const std::vector<MeshInstance>& instances) {
    curNode_ = topLevelIndex;
    if (tlas->root() >= 0) ProcessTLASNodes(*tlas, tlas->root(), instances);
}

// ---------------------------------------------------------------------------
// OBJ ingest
// ---------------------------------------------------------------------------

// Digit-accumulation number parser with the arithmetic of tinyobjloader 2.0.0-rc
// (the OBJ reader the reference links, tiny_obj_loader.h:866-997): integer
// digits m = 10 m + d in double, fraction digits m += d * 10^-k from a table of
// double constants, exponent via ldexp(m * 5^e, e).  Kept alike so vertex
// positions are bit-identical to the reference mesh arrays.  It follows
// tinyobjloader's tryParseDouble step for step and is distributed under that
// library's licence:
//
//   The MIT License (MIT)
//   Copyright (c) 2012-Present, Syoyo Fujita and many contributors.
//
//   Permission is hereby granted, free of charge, to any person obtaining a
//   copy of this software and associated documentation files (the
//   "Software"), to deal in the Software without restriction, including
//   without limitation the rights to use, copy, modify, merge, publish,
//   distribute, sublicense, and/or sell copies of the Software, and to permit
//   persons to whom the Software is furnished to do so, subject to the
//   following conditions:
//
//   The above copyright notice and this permission notice shall be included
//   in all copies or substantial portions of the Software.
//
//   THE SOFTWARE IS PROVIDED "AS IS", WITHOUT WARRANTY OF ANY KIND, EXPRESS
//   OR IMPLIED, INCLUDING BUT NOT LIMITED TO THE WARRANTIES OF
//   MERCHANTABILITY, FITNESS FOR A PARTICULAR PURPOSE AND NONINFRINGEMENT. IN
//   NO EVENT SHALL THE AUTHORS OR COPYRIGHT HOLDERS BE LIABLE FOR ANY CLAIM,
//   DAMAGES OR OTHER LIABILITY, WHETHER IN AN ACTION OF CONTRACT, TORT OR
//   OTHERWISE, ARISING FROM, OUT OF OR IN CONNECTION WITH THE SOFTWARE OR THE
//   USE OR OTHER DEALINGS IN THE SOFTWARE.
double ParseObjReal(const char** token) {
    const char* s = *token;
    s += std::strspn(s, " \t");
    const char* end = s + std::strcspn(s, " \t\r\n");
    *token = end;
    const char* curr = s;
    if (curr >= end) return 0.0;
    double mantissa = 0.0;
    int exponent = 0;
    char sign = '+', exp_sign = '+';
    int read = 0;
    bool leading_dot = false;
    auto isdig = [](char c) { return c >= '0' && c <= '9'; };
    if (*curr == '+' || *curr == '-') {
        sign = *curr;
        curr++;
        if (curr != end && *curr == '.') leading_dot = true;
    } else if (isdig(*curr)) {
    } else if (*curr == '.') {
        leading_dot = true;
    } else {
        return 0.0;
    }
    bool more = (curr != end);
    if (!leading_dot) {
        while (more && isdig(*curr)) {
            mantissa *= 10;
            mantissa += static_cast<int>(*curr - '0');
            curr++;
            read++;
            more = (curr != end);
        }
        if (read == 0) return 0.0;
    }
    if (!more) goto assemble;
    if (*curr == '.') {
        curr++;
        read = 1;
        more = (curr != end);
        static const double kPow[] = {1.0, 0.1, 0.01, 0.001, 0.0001, 0.00001, 0.000001, 0.0000001};
        while (more && isdig(*curr)) {
            mantissa += static_cast<int>(*curr - '0') * (read < 8 ? kPow[read] : std::pow(10.0, -read));
            read++;
            curr++;
            more = (curr != end);
        }
    } else if (*curr == 'e' || *curr == 'E') {
    } else {
        goto assemble;
    }
    if (!more) goto assemble;
    if (*curr == 'e' || *curr == 'E') {
        curr++;
        more = (curr != end);
        if (more && (*curr == '+' || *curr == '-')) {
            exp_sign = *curr;
            curr++;
        } else if (more && isdig(*curr)) {
        } else {
            return 0.0;
        }
        read = 0;
        more = (curr != end);
        while (more && isdig(*curr)) {
            if (exponent > (2147483647 / 10)) return 0.0;
            exponent *= 10;
            exponent += static_cast<int>(*curr - '0');
            curr++;
            read++;
            more = (curr != end);
        }
        exponent *= (exp_sign == '+' ? 1 : -1);
        if (read == 0) return 0.0;
    }
assemble:
    return (sign == '+' ? 1 : -1) *
           (exponent ? std::ldexp(mantissa * std::pow(5.0, exponent), exponent) : mantissa);
}

namespace {
struct ObjIndex { int v = -1, vt = -1, vn = -1; };

// OBJ index: 1-based, negative = relative to the current count, 0 = invalid.
bool FixIndex(int idx, int n, int* ret) {
    if (idx > 0) { *ret = idx - 1; return true; }
    if (idx == 0) return false;
    *ret = n + idx;
    return *ret >= 0;
}

bool ParseTriple(const char** token, int nv, int nvn, int nvt, ObjIndex* out) {
    const char* s = *token;
    ObjIndex r;
    char* e;
    long a = std::strtol(s, &e, 10);
    if (e == s) return false;
    if (!FixIndex((int)a, nv, &r.v)) return false;
    s = e;
    if (*s == '/') {
        s++;
        if (*s == '/') {
            s++;
            long c = std::strtol(s, &e, 10);
            if (e != s) { FixIndex((int)c, nvn, &r.vn); s = e; }
        } else {
            long b = std::strtol(s, &e, 10);
            if (e != s) { FixIndex((int)b, nvt, &r.vt); s = e; }
            if (*s == '/') {
                s++;
                long c = std::strtol(s, &e, 10);
                if (e != s) { FixIndex((int)c, nvn, &r.vn); s = e; }
            }
        }
    }
    *out = r;
    *token = s;
    return true;
}
}  // namespace

// Reads positions / normals / texcoords and triangulated faces, then
// de-indexes them (3 vertices per triangle, uv packed in .w) in file order
// (Mesh.cpp:28-91).  Quads are split along the shorter diagonal like the
// reference's OBJ reader (tiny_obj_loader.h:1457-1548); larger polygons are
// fanned.  A face without normals gets its geometric normal (the reference
// reads attrib.normals[-3] there, Mesh.cpp:65-67).
bool Mesh::LoadFromFile(const std::string& filename, std::string* err) {
    name = filename;
    FILE* f = std::fopen(filename.c_str(), "r");
    if (!f) {
        if (err) *err = "cannot open mesh file " + filename;
        return false;
    }
    std::vector<float> v, vn, vt;
    std::vector<ObjIndex> tri;  // 3 per triangle
    char line[8192];
    while (std::fgets(line, sizeof line, f)) {
        const char* s = line;
        s += std::strspn(s, " \t");
        if (s[0] == 'v' && (s[1] == ' ' || s[1] == '\t')) {
            s += 2;
            for (int k = 0; k < 3; ++k) v.push_back((float)ParseObjReal(&s));
        } else if (s[0] == 'v' && s[1] == 'n' && (s[2] == ' ' || s[2] == '\t')) {
            s += 3;
            for (int k = 0; k < 3; ++k) vn.push_back((float)ParseObjReal(&s));
        } else if (s[0] == 'v' && s[1] == 't' && (s[2] == ' ' || s[2] == '\t')) {
            s += 3;
            for (int k = 0; k < 2; ++k) vt.push_back((float)ParseObjReal(&s));
        } else if (s[0] == 'f' && (s[1] == ' ' || s[1] == '\t')) {
            s += 2;
            std::vector<ObjIndex> face;
            int nv = (int)(v.size() / 3), nvn = (int)(vn.size() / 3), nvt = (int)(vt.size() / 2);
            for (;;) {
                s += std::strspn(s, " \t");
                if (*s == '\0' || *s == '\r' || *s == '\n') break;
                ObjIndex idx;
                if (!ParseTriple(&s, nv, nvn, nvt, &idx)) break;
                face.push_back(idx);
                s += std::strcspn(s, " \t\r\n");
            }
            if (face.size() < 3) continue;
            if (face.size() == 3) {
                // Triangles go through the reader's ear clipper, which emits
                // them rotated: (1,2,0) when the projected ring sums
                // positive, else (1,0,2); a ring with zero projected area
                // yields no triangle (earcut.hpp:152-345 as called from
                // tiny_obj_loader.h:1567-1660).
                auto P = [&](int i) { return Vec3(v[3 * face[i].v], v[3 * face[i].v + 1], v[3 * face[i].v + 2]); };
                int ax0 = 1, ax1 = 2;
                for (int k = 0; k < 3; ++k) {
                    Vec3 a = P(k), b = P((k + 1) % 3), c = P((k + 2) % 3);
                    float e0x = b.x - a.x, e0y = b.y - a.y, e0z = b.z - a.z;
                    float e1x = c.x - b.x, e1y = c.y - b.y, e1z = c.z - b.z;
                    float cx = std::fabs(e0y * e1z - e0z * e1y);
                    float cy = std::fabs(e0z * e1x - e0x * e1z);
                    float cz = std::fabs(e0x * e1y - e0y * e1x);
                    const float eps = std::numeric_limits<float>::epsilon();
                    if (cx > eps || cy > eps || cz > eps) {
                        if (!(cx > cy && cx > cz)) {
                            ax0 = 0;
                            if (cz > cx && cz > cy) ax1 = 1;
                        }
                        break;
                    }
                }
                double px[3], py[3];
                for (int k = 0; k < 3; ++k) {
                    Vec3 q = P(k);
                    px[k] = (double)q[ax0];
                    py[k] = (double)q[ax1];
                }
                double sum = 0;
                for (int i = 0, j = 2; i < 3; j = i++) sum += (px[j] - px[i]) * (py[i] + py[j]);
                int o[3];
                if (sum > 0) { o[0] = 1; o[1] = 2; o[2] = 0; } else { o[0] = 1; o[1] = 0; o[2] = 2; }
                double area = (py[o[1]] - py[o[0]]) * (px[o[2]] - px[o[1]]) -
                              (px[o[1]] - px[o[0]]) * (py[o[2]] - py[o[1]]);
                if (area >= 0) continue;
                tri.insert(tri.end(), {face[o[0]], face[o[1]], face[o[2]]});
            } else if (face.size() == 4) {
                auto P = [&](int i) { return Vec3(v[3 * face[i].v], v[3 * face[i].v + 1], v[3 * face[i].v + 2]); };
                Vec3 p0 = P(0), p1 = P(1), p2 = P(2), p3 = P(3);
                float e02x = p2.x - p0.x, e02y = p2.y - p0.y, e02z = p2.z - p0.z;
                float e13x = p3.x - p1.x, e13y = p3.y - p1.y, e13z = p3.z - p1.z;
                float sqr02 = e02x * e02x + e02y * e02y + e02z * e02z;
                float sqr13 = e13x * e13x + e13y * e13y + e13z * e13z;
                if (sqr02 < sqr13) {
                    tri.insert(tri.end(), {face[0], face[1], face[2], face[0], face[2], face[3]});
                } else {
                    tri.insert(tri.end(), {face[0], face[1], face[3], face[1], face[2], face[3]});
                }
            } else {
                for (size_t k = 1; k + 1 < face.size(); ++k)
                    tri.insert(tri.end(), {face[0], face[k], face[k + 1]});
            }
        }
    }
    std::fclose(f);
    verticesUVX.clear();
    normalsUVY.clear();
    verticesUVX.reserve(tri.size());
    normalsUVY.reserve(tri.size());
    for (size_t t = 0; t < tri.size(); t += 3) {
        Vec3 fn;
        bool needFn = tri[t].vn < 0 || tri[t + 1].vn < 0 || tri[t + 2].vn < 0;
        if (needFn) {
            Vec3 a(v[3 * tri[t].v], v[3 * tri[t].v + 1], v[3 * tri[t].v + 2]);
            Vec3 b(v[3 * tri[t + 1].v], v[3 * tri[t + 1].v + 1], v[3 * tri[t + 1].v + 2]);
            Vec3 c(v[3 * tri[t + 2].v], v[3 * tri[t + 2].v + 1], v[3 * tri[t + 2].v + 2]);
            fn = Vec3::Normalize(Vec3::Cross(b - a, c - a));
        }
        for (int k = 0; k < 3; ++k) {
            const ObjIndex& id = tri[t + k];
            float tx = 0, ty = 0;
            if (!vt.empty() && id.vt >= 0) { tx = vt[2 * id.vt]; ty = vt[2 * id.vt + 1]; }
            verticesUVX.emplace_back(v[3 * id.v], v[3 * id.v + 1], v[3 * id.v + 2], tx);
            if (id.vn >= 0)
                normalsUVY.emplace_back(vn[3 * id.vn], vn[3 * id.vn + 1], vn[3 * id.vn + 2], ty);
            else
                normalsUVY.emplace_back(fn.x, fn.y, fn.z, ty);
        }
    }
    return true;
}

// One box per triangle from its three vertices (Mesh.cpp:93-111).
void Mesh::BuildBVH() {
    const int numTris = (int)(verticesUVX.size() / 3);
    std::vector<BBox> bounds(numTris);
    for (int i = 0; i < numTris; ++i) {
        bounds[i].grow(Vec3(verticesUVX[i * 3 + 0]));
        bounds[i].grow(Vec3(verticesUVX[i * 3 + 1]));
        bounds[i].grow(Vec3(verticesUVX[i * 3 + 2]));
    }
    bvh.Build(bounds.data(), numTris);
}

// ---------------------------------------------------------------------------
// Camera (Camera.cpp:101-169)
// ---------------------------------------------------------------------------

Camera::Camera(Vec3 eye, Vec3 lookat, float fov_deg) {
    position = eye;
    pivot = lookat;
    Vec3 dir = Vec3::Normalize(pivot - position);
    pitch = std::asin(dir.y) * (180.f / kPi);
    yaw = std::atan2(dir.z, dir.x) * (180.f / kPi);
    radius = Vec3::Length(eye - lookat);
    fov = fov_deg * (kPi / 180.f);
    focalDist = 0.1f;
    aperture = 0.f;
    UpdateCamera();
}

void Camera::UpdateCamera() {
    Vec3 f;
    float yr = yaw * (kPi / 180.f), pr = pitch * (kPi / 180.f);
    f.x = std::cos(yr) * std::cos(pr);
    f.y = std::sin(pr);
    f.z = std::sin(yr) * std::cos(pr);
    forward = Vec3::Normalize(f);
    position = pivot + (forward * -1.0f) * radius;
    right = Vec3::Normalize(Vec3::Cross(forward, worldUp));
    up = Vec3::Normalize(Vec3::Cross(right, forward));
}

// ---------------------------------------------------------------------------
// HDR importance tables (hdrloader.cpp:33-131): luminance 0.3R + 0.6G + 0.1B,
// per-row CDF, marginal CDF over rows, inverse CDFs tabulated at (i+1)/N with
// a lower-bound search.
// ---------------------------------------------------------------------------

static int LowerBound(const float* a, int lo, int hi, float value) {
    while (lo < hi) {
        int mid = lo + (hi - lo) / 2;
        if (a[mid] < value) lo = mid + 1; else hi = mid;
    }
    return lo;
}

void BuildHdrDistributions(HDRData* res) {
    const int W = res->width, H = res->height;
    std::vector<float> pdf2D((size_t)W * H), cdf2D((size_t)W * H), pdf1D(H), cdf1D(H);
    res->marginalDistData.assign(H, Vec2());
    res->conditionalDistData.assign((size_t)W * H, Vec2());
    float colWeightSum = 0.f;
    for (int j = 0; j < H; ++j) {
        float rowWeightSum = 0.f;
        for (int i = 0; i < W; ++i) {
            const float* c = &res->cols[((size_t)j * W + i) * 3];
            float weight = c[0] * 0.3f + c[1] * 0.6f + c[2] * 0.1f;
            rowWeightSum += weight;
            pdf2D[(size_t)j * W + i] = weight;
            cdf2D[(size_t)j * W + i] = rowWeightSum;
        }
        for (int i = 0; i < W; ++i) {
            pdf2D[(size_t)j * W + i] /= rowWeightSum;
            cdf2D[(size_t)j * W + i] /= rowWeightSum;
        }
        colWeightSum += rowWeightSum;
        pdf1D[j] = rowWeightSum;
        cdf1D[j] = colWeightSum;
    }
    for (int j = 0; j < H; ++j) {
        cdf1D[j] /= colWeightSum;
        pdf1D[j] /= colWeightSum;
    }
    for (int i = 0; i < H; ++i) {
        float inv = (float)(i + 1) / H;
        int row = LowerBound(cdf1D.data(), 0, H, inv);
        res->marginalDistData[i].x = row / (float)H;
        res->marginalDistData[i].y = pdf1D[i];
    }
    for (int j = 0; j < H; ++j) {
        for (int i = 0; i < W; ++i) {
            float inv = (float)(i + 1) / W;
            int col = LowerBound(cdf2D.data(), j * W, (j + 1) * W, inv) - j * W;
            res->conditionalDistData[(size_t)j * W + i].x = col / (float)W;
            res->conditionalDistData[(size_t)j * W + i].y = pdf2D[(size_t)j * W + i];
        }
    }
}

// ---------------------------------------------------------------------------
// Scene (Scene.cpp:15-231)
// ---------------------------------------------------------------------------

Scene::~Scene() {
    for (Mesh* m : meshes) delete m;
    delete camera;
    delete hdrData;
}

int Scene::AddMesh(const std::string& filename) {
    for (size_t i = 0; i < meshes.size(); ++i)
        if (meshes[i]->name == filename) return (int)i;
    Mesh* mesh = new Mesh;
    std::string err;
    if (!mesh->LoadFromFile(filename, &err)) {
        lastError = err;
        delete mesh;
        return -1;
    }
    meshes.push_back(mesh);
    return (int)meshes.size() - 1;
}

int Scene::AddMesh(const float* positions, const float* normals, const float* uvs,
                   int64_t num_tris, const std::string& name) {
    Mesh* mesh = new Mesh;
    mesh->name = name;
    mesh->verticesUVX.resize((size_t)num_tris * 3);
    mesh->normalsUVY.resize((size_t)num_tris * 3);
    for (int64_t t = 0; t < num_tris; ++t) {
        Vec3 p[3];
        for (int k = 0; k < 3; ++k)
            p[k] = Vec3(positions[(t * 3 + k) * 3], positions[(t * 3 + k) * 3 + 1], positions[(t * 3 + k) * 3 + 2]);
        Vec3 fn;
        if (!normals) fn = Vec3::Normalize(Vec3::Cross(p[1] - p[0], p[2] - p[0]));
        for (int k = 0; k < 3; ++k) {
            size_t vi = (size_t)t * 3 + k;
            float u = uvs ? uvs[vi * 2] : 0.f, v = uvs ? uvs[vi * 2 + 1] : 0.f;
            mesh->verticesUVX[vi] = Vec4(p[k].x, p[k].y, p[k].z, u);
            if (normals)
                mesh->normalsUVY[vi] = Vec4(normals[vi * 3], normals[vi * 3 + 1], normals[vi * 3 + 2], v);
            else
                mesh->normalsUVY[vi] = Vec4(fn.x, fn.y, fn.z, v);
        }
    }
    meshes.push_back(mesh);
    return (int)meshes.size() - 1;
}

int Scene::AddTexture(const std::string& filename) {
    for (size_t i = 0; i < textures.size(); ++i)
        if (textures[i].name == filename) return (int)i;
    Texture t;
    Image img;
    std::string e;
    if (!load_image(filename, &img, &e)) {
        lastError = e;
        return -1;
    }
    t.name = filename;
    t.width = img.width;
    t.height = img.height;
    t.rgba = std::move(img.rgba);
    textures.push_back(std::move(t));
    return (int)textures.size() - 1;
}

int Scene::AddTexture(const uint8_t* rgba, int width, int height, const std::string& name) {
    if (!name.empty())
        for (size_t i = 0; i < textures.size(); ++i)
            if (textures[i].name == name) return (int)i;
    Texture t;
    t.name = name;
    t.width = width;
    t.height = height;
    t.rgba.assign(rgba, rgba + (size_t)width * height * 4);
    textures.push_back(std::move(t));
    return (int)textures.size() - 1;
}

// Textures copied into one RGBA8 array of equally sized layers (Scene.cpp:223-230;
// the GL array texture takes one width and height for every layer).
bool Scene::PackTextures() {
    textureMapsArray.clear();
    texWidth = texHeight = 0;
    for (const Texture& t : textures) {
        if (texWidth && (t.width != texWidth || t.height != texHeight)) {
            lastError = "all textures must have one size (one RGBA8 array texture holds them): " + t.name;
            return false;
        }
        texWidth = t.width;
        texHeight = t.height;
        textureMapsArray.insert(textureMapsArray.end(), t.rgba.begin(), t.rgba.end());
    }
    return true;
}

int Scene::AddMaterial(const Material& m) {
    materials.push_back(m);
    return (int)materials.size() - 1;
}

int Scene::AddMeshInstance(const MeshInstance& mi) {
    meshInstances.push_back(mi);
    return (int)meshInstances.size() - 1;
}

int Scene::AddLight(const Light& l) {
    lights.push_back(l);
    return (int)lights.size() - 1;
}

void Scene::AddCamera(Vec3 eye, Vec3 lookat, float fov) {
    delete camera;
    camera = new Camera(eye, lookat, fov);
}

void Scene::AddHDR(const float* rgb, int width, int height) {
    delete hdrData;
    hdrData = new HDRData;
    hdrData->width = width;
    hdrData->height = height;
    hdrData->cols.assign(rgb, rgb + (size_t)width * height * 3);
    BuildHdrDistributions(hdrData);
    renderOptions.useEnvMap = true;
}

// Instance world box: the mesh box's corners pushed through the affine part
// axis by axis (Scene.cpp:106-142), then a midpoint-split BVH over instances.
void Scene::createTLAS() {
    std::vector<BBox> bounds(meshInstances.size());
    for (size_t i = 0; i < meshInstances.size(); ++i) {
        const BBox& bb = meshes[meshInstances[i].meshID]->bvh.Bounds();
        const Mat4& m = meshInstances[i].transform;
        Vec3 minB = bb.pmin, maxB = bb.pmax;
        Vec3 right(m[0][0], m[0][1], m[0][2]);
        Vec3 up(m[1][0], m[1][1], m[1][2]);
        Vec3 fwd(m[2][0], m[2][1], m[2][2]);
        Vec3 tr(m[3][0], m[3][1], m[3][2]);
        Vec3 xa = right * minB.x, xb = right * maxB.x;
        Vec3 ya = up * minB.y, yb = up * maxB.y;
        Vec3 za = fwd * minB.z, zb = fwd * maxB.z;
        bounds[i].pmin = Vec3::Min(xa, xb) + Vec3::Min(ya, yb) + Vec3::Min(za, zb) + tr;
        bounds[i].pmax = Vec3::Max(xa, xb) + Vec3::Max(ya, yb) + Vec3::Max(za, zb) + tr;
    }
    sceneBvh.Build(bounds.data(), (int)bounds.size());
    sceneBounds = sceneBvh.Bounds();
}

void Scene::createBLAS() {
    for (Mesh* m : meshes) m->BuildBVH();
}

void Scene::RebuildInstances() {
    if (flatOnly) {  // no host TLAS: a device-built scene rebuilds it from the transforms
        for (size_t i = 0; i < meshInstances.size(); ++i) transforms[i] = meshInstances[i].transform;
        instancesModified = true;
        return;
    }
    sceneBvh = Bvh(Bvh::kMedianSplit, 10.0f, 64);
    createTLAS();
    bvhTranslator.UpdateTLAS(&sceneBvh, meshInstances);
    for (size_t i = 0; i < meshInstances.size(); ++i) transforms[i] = meshInstances[i].transform;
    instancesModified = true;
}

void Scene::CreateAccelerationStructures() {
    createBLAS();
    createTLAS();
    bvhTranslator.Process(&sceneBvh, meshes, meshInstances);
    vertIndices.clear();
    verticesUVX.clear();
    normalsUVY.clear();
    meshTriOffsets.assign(1, 0);
    int verticesCnt = 0;
    for (Mesh* m : meshes) {
        const std::vector<int>& tri = m->bvh.packed_indices();
        for (int index : tri)
            vertIndices.push_back(Indices{index * 3 + 0 + verticesCnt, index * 3 + 1 + verticesCnt,
                                          index * 3 + 2 + verticesCnt});
        verticesUVX.insert(verticesUVX.end(), m->verticesUVX.begin(), m->verticesUVX.end());
        normalsUVY.insert(normalsUVY.end(), m->normalsUVY.begin(), m->normalsUVY.end());
        verticesCnt += (int)m->verticesUVX.size();
        meshTriOffsets.push_back((int64_t)vertIndices.size());
    }
    transforms.resize(meshInstances.size());
    instanceMeshes.resize(meshInstances.size());
    instanceMaterials.resize(meshInstances.size());
    for (size_t i = 0; i < meshInstances.size(); ++i) {
        transforms[i] = meshInstances[i].transform;
        instanceMeshes[i] = meshInstances[i].meshID;
        instanceMaterials[i] = meshInstances[i].materialID;
    }
    flatOnly = false;
}

void Scene::Flatten() {
    vertIndices.clear();
    verticesUVX.clear();
    normalsUVY.clear();
    meshTriOffsets.assign(1, 0);
    int verticesCnt = 0;
    for (Mesh* m : meshes) {
        const int ntri = (int)(m->verticesUVX.size() / 3);
        for (int index = 0; index < ntri; ++index)
            vertIndices.push_back(Indices{index * 3 + 0 + verticesCnt, index * 3 + 1 + verticesCnt,
                                          index * 3 + 2 + verticesCnt});
        verticesUVX.insert(verticesUVX.end(), m->verticesUVX.begin(), m->verticesUVX.end());
        normalsUVY.insert(normalsUVY.end(), m->normalsUVY.begin(), m->normalsUVY.end());
        verticesCnt += (int)m->verticesUVX.size();
        meshTriOffsets.push_back((int64_t)vertIndices.size());
    }
    transforms.resize(meshInstances.size());
    instanceMeshes.resize(meshInstances.size());
    instanceMaterials.resize(meshInstances.size());
    for (size_t i = 0; i < meshInstances.size(); ++i) {
        transforms[i] = meshInstances[i].transform;
        instanceMeshes[i] = meshInstances[i].meshID;
        instanceMaterials[i] = meshInstances[i].materialID;
    }
    flatOnly = true;
}

// ---------------------------------------------------------------------------
// Scene file loader (Loader.cpp:45-386): blocks `material <name> {}`,
// `light {}`, `Camera {}`, `Renderer {}`, `mesh {}`; keys parsed per line with
// scanf patterns; material 0 is the default material; unknown material names
// fall back to 0; meshes that fail to load are skipped.
// ---------------------------------------------------------------------------

bool LoadSceneFromFile(const std::string& filename, Scene* scene, RenderOptions& ro, std::string* err) {
    FILE* file = std::fopen(filename.c_str(), "r");
    if (!file) {
        if (err) *err = "could not open scene file " + filename;
        return false;
    }
    struct MaterialData { Material mat; int id; };
    std::vector<std::pair<std::string, MaterialData>> materialMap;
    auto findMat = [&](const std::string& n) -> MaterialData* {
        for (auto& kv : materialMap) if (kv.first == n) return &kv.second;
        return nullptr;
    };
    size_t cut = filename.find_last_of("/\\");
    std::string path = (cut == std::string::npos ? filename : filename.substr(0, cut)) + "/";
    bool overrideTile = false;
    int legacyAces = 2;
    char line[2048];

    scene->AddMaterial(Material());
    bool cameraAdded = false;

    while (std::fgets(line, sizeof line, file)) {
        if (line[0] == '#') continue;
        char name[2048] = {0};

        if (std::sscanf(line, " material %s", name) == 1) {
            Material material;
            char albedoTex[100] = "None", mrTex[100] = "None", nTex[100] = "None", eTex[100] = "None";
            while (std::fgets(line, sizeof line, file)) {
                if (std::strchr(line, '}')) break;
                std::sscanf(line, " name %s", name);
                std::sscanf(line, " color %f %f %f", &material.albedo.x, &material.albedo.y, &material.albedo.z);
                std::sscanf(line, " albedo %f %f %f", &material.albedo.x, &material.albedo.y, &material.albedo.z);
                std::sscanf(line, " emission %f %f %f", &material.emission.x, &material.emission.y, &material.emission.z);
                std::sscanf(line, " metallic %f", &material.metallic);
                std::sscanf(line, " roughness %f", &material.roughness);
                std::sscanf(line, " subsurface %f", &material.subsurface);
                std::sscanf(line, " specular %f", &material.specular);
                std::sscanf(line, " specularTint %f", &material.specularTint);
                std::sscanf(line, " anisotropic %f", &material.anisotropic);
                std::sscanf(line, " sheen %f", &material.sheen);
                std::sscanf(line, " sheenTint %f", &material.sheenTint);
                std::sscanf(line, " clearcoat %f", &material.clearcoat);
                std::sscanf(line, " clearcoatRoughness %f", &material.clearcoatRoughness);
                std::sscanf(line, " transmission %f", &material.transmission);
                std::sscanf(line, " ior %f", &material.ior);
                std::sscanf(line, " extinction %f %f %f", &material.extinction.x, &material.extinction.y, &material.extinction.z);
                std::sscanf(line, " albedoTexture %99s", albedoTex);
                std::sscanf(line, " metallicRoughnessTexture %99s", mrTex);
                std::sscanf(line, " normalTexture %99s", nTex);
                std::sscanf(line, " emissionTexture %99s", eTex);
            }
            // Texture maps (Loader.cpp:133-147): a file that cannot be read
            // leaves the id at -1, as Scene::AddTexture does (Scene.cpp:45-66).
            if (std::strcmp(albedoTex, "None") != 0) material.albedoTexID = (float)scene->AddTexture(path + albedoTex);
            if (std::strcmp(mrTex, "None") != 0) material.metallicRoughnessTexID = (float)scene->AddTexture(path + mrTex);
            if (std::strcmp(nTex, "None") != 0) material.normalmapTexID = (float)scene->AddTexture(path + nTex);
            if (std::strcmp(eTex, "None") != 0) material.emissionmapTexID = (float)scene->AddTexture(path + eTex);
            if (!findMat(name)) {
                int id = scene->AddMaterial(material);
                materialMap.emplace_back(name, MaterialData{material, id});
            }
        }

        if (std::strstr(line, "light")) {
            Light light;
            Vec3 v1, v2;
            char type[20] = "None";
            while (std::fgets(line, sizeof line, file)) {
                if (std::strchr(line, '}')) break;
                std::sscanf(line, " position %f %f %f", &light.position.x, &light.position.y, &light.position.z);
                std::sscanf(line, " emission %f %f %f", &light.emission.x, &light.emission.y, &light.emission.z);
                std::sscanf(line, " radius %f", &light.radius);
                std::sscanf(line, " v1 %f %f %f", &v1.x, &v1.y, &v1.z);
                std::sscanf(line, " v2 %f %f %f", &v2.x, &v2.y, &v2.z);
                std::sscanf(line, " type %19s", type);
            }
            if (std::strcmp(type, "Quad") == 0) {
                light.type = RectLight;
                light.u = v1 - light.position;
                light.v = v2 - light.position;
                light.area = Vec3::Length(Vec3::Cross(light.u, light.v));
            } else if (std::strcmp(type, "Sphere") == 0) {
                light.type = SphereLight;
                light.area = 4.0f * kPi * light.radius * light.radius;
            }
            scene->AddLight(light);
        }

        if (std::strstr(line, "Camera")) {
            Vec3 position, lookAt;
            float fov = 0.f, aperture = 0.f, focalDist = 1.f;
            while (std::fgets(line, sizeof line, file)) {
                if (std::strchr(line, '}')) break;
                std::sscanf(line, " position %f %f %f", &position.x, &position.y, &position.z);
                std::sscanf(line, " lookAt %f %f %f", &lookAt.x, &lookAt.y, &lookAt.z);
                std::sscanf(line, " aperture %f ", &aperture);
                std::sscanf(line, " focalDistance %f", &focalDist);
                std::sscanf(line, " focaldist %f", &focalDist);
                std::sscanf(line, " fov %f", &fov);
            }
            scene->AddCamera(position, lookAt, fov);
            scene->camera->aperture = aperture;
            scene->camera->focalDist = focalDist;
            cameraAdded = true;
        }

        if (std::strstr(line, "Renderer")) {
            char envMap[200] = "None";
            char enableRR[10] = "None";
            int useCA = ro.useCA ? 1 : 0, useCAD = ro.useCADistortion ? 1 : 0, useVig = ro.useVignette ? 1 : 0;
            while (std::fgets(line, sizeof line, file)) {
                if (std::strchr(line, '}')) break;
                std::sscanf(line, " envMap %199s", envMap);
                std::sscanf(line, " hdriMap %199s", envMap);
                std::sscanf(line, " resolution %d %d", &ro.resolution.x, &ro.resolution.y);
                std::sscanf(line, " hdrMultiplier %f", &ro.hdrMultiplier);
                std::sscanf(line, " hdriMultiplier %f", &ro.hdrMultiplier);
                std::sscanf(line, " maxDepth %i", &ro.maxDepth);
                if (std::sscanf(line, " tileWidth %i", &ro.tileWidth) == 1) overrideTile = true;
                if (std::sscanf(line, " tileHeight %i", &ro.tileHeight) == 1) overrideTile = true;
                std::sscanf(line, " enableRR %9s", enableRR);
                std::sscanf(line, " RRDepth %i", &ro.RRDepth);
                std::sscanf(line, " tonemapIndex %i", &ro.tonemapIndex);
                std::sscanf(line, " useAces %i", &legacyAces);
                if (legacyAces == 1) ro.tonemapIndex = 1;
                if (legacyAces == 0) ro.tonemapIndex = 0;
                // Loader.cpp:272-282 reads these flags with %i into bool fields;
                // here each is read as an int and kept as (value != 0)
                std::sscanf(line, " useChromaticAberration %i", &useCA);
                std::sscanf(line, " CADistance %f", &ro.caDistance);
                std::sscanf(line, " useCAdistortion %i", &useCAD);
                std::sscanf(line, " CAAngularity %f", &ro.caP1);
                std::sscanf(line, " CADirectionality %f", &ro.caP2);
                std::sscanf(line, " CACenter %f", &ro.caP3);
                std::sscanf(line, " useVignette %i", &useVig);
                std::sscanf(line, " vignetteIntensity %f", &ro.vignetteIntensity);
                std::sscanf(line, " vignettePower %f", &ro.vignettePower);
            }
            ro.useCA = useCA != 0;
            ro.useCADistortion = useCAD != 0;
            ro.useVignette = useVig != 0;
            if (std::strcmp(envMap, "None") != 0) {
                // Radiance .hdr parsing is outside this build; the env map
                // stays disabled, as when the reference fails to load it
                // (Scene.cpp:80-82).
            }
            if (std::strcmp(enableRR, "False") == 0) ro.enableRR = false;
            else if (std::strcmp(enableRR, "True") == 0) ro.enableRR = true;
        }

        if (std::strstr(line, "mesh")) {
            std::string meshFile;
            Mat4 xform;
            int material_id = 0;
            char meshName[200] = "None";
            while (std::fgets(line, sizeof line, file)) {
                if (std::strchr(line, '}')) break;
                char fileTok[2048];
                char matName[100];
                std::sscanf(line, " name %199[^\t\n]s", meshName);
                if (std::sscanf(line, " file %2047s", fileTok) == 1) meshFile = path + fileTok;
                if (std::sscanf(line, " material %99s", matName) == 1) {
                    if (MaterialData* md = findMat(matName)) material_id = md->id;
                }
                std::sscanf(line, " position %f %f %f", &xform[3][0], &xform[3][1], &xform[3][2]);
                std::sscanf(line, " scale %f %f %f", &xform[0][0], &xform[1][1], &xform[2][2]);
            }
            if (!meshFile.empty()) {
                int mesh_id = scene->AddMesh(meshFile);
                if (mesh_id != -1) {
                    std::string instName;
                    if (std::strcmp(meshName, "None") != 0) instName = meshName;
                    else {
                        size_t p = meshFile.find_last_of("/\\");
                        instName = meshFile.substr(p + 1);
                    }
                    scene->AddMeshInstance(MeshInstance(instName, mesh_id, xform, material_id));
                }
            }
        }
    }
    std::fclose(file);
    if (!overrideTile) {
        ro.tileHeight = ro.resolution.y;
        ro.tileWidth = ro.resolution.x;
    }
    if (!cameraAdded) scene->AddCamera(Vec3(0.f, 0.f, 10.f), Vec3(0.f, 0.f, -10.f), 35.f);
    if (scene->meshInstances.empty()) {
        if (err) *err = "scene has no mesh instances";
        return false;
    }
    scene->CreateAccelerationStructures();
    if (!scene->PackTextures()) {
        if (err) *err = scene->lastError;
        return false;
    }
    return true;
}

}  // namespace lfh
