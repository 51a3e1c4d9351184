// ORACLE — test infrastructure only (see oracle.h).
//
// CPU restatement of the reference GLSL path tracer, templated on the
// arithmetic type T (float: parity checker and CPU baseline; double: the
// converged reference).  Function-by-function citations are given inline.
#include "oracle.h"

#include <cmath>
#include <cstring>
#include <vector>
#include <omp.h>

namespace {

// ---------------------------------------------------------------------------
// vector helpers (GLSL vec3 semantics with the conventions of oracle.h)
// ---------------------------------------------------------------------------
template <class T> struct V3 {
    T x, y, z;
};
template <class T> inline V3<T> mk(T x, T y, T z) { return V3<T>{x, y, z}; }
template <class T> inline V3<T> operator+(V3<T> a, V3<T> b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
template <class T> inline V3<T> operator-(V3<T> a, V3<T> b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
template <class T> inline V3<T> operator-(V3<T> a) { return mk(-a.x, -a.y, -a.z); }
template <class T> inline V3<T> operator*(V3<T> a, V3<T> b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
template <class T> inline V3<T> operator*(V3<T> a, T s) { return mk(a.x * s, a.y * s, a.z * s); }
template <class T> inline V3<T> operator*(T s, V3<T> a) { return mk(s * a.x, s * a.y, s * a.z); }
template <class T> inline V3<T> operator/(V3<T> a, T s) { return mk(a.x / s, a.y / s, a.z / s); }
template <class T> inline T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> inline V3<T> cross(V3<T> a, V3<T> b) {
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
template <class T> inline T sq(T x) { return std::sqrt(x); }
template <class T> inline V3<T> normalize(V3<T> a) {
    T l = std::sqrt(dot(a, a));
    return mk(a.x / l, a.y / l, a.z / l);
}
template <class T> inline T length(V3<T> a) { return std::sqrt(dot(a, a)); }
template <class T> inline T mixs(T x, T y, T a) { return x * (T(1) - a) + y * a; }
template <class T> inline V3<T> mixv(V3<T> x, V3<T> y, T a) { return x * (T(1) - a) + y * a; }
template <class T> inline T clampT(T x, T lo, T hi) { return std::fmin(std::fmax(x, lo), hi); }
// reflect(I, N) = I - 2 dot(N, I) N
template <class T> inline V3<T> reflect(V3<T> I, V3<T> N) { return I - (T(2) * dot(N, I)) * N; }
// refract(I, N, eta): k = 1 - eta^2 (1 - dot(N,I)^2); eta I - (eta dot(N,I) + sqrt(k)) N
template <class T> inline V3<T> refract(V3<T> I, V3<T> N, T eta) {
    T d = dot(N, I);
    T k = T(1) - eta * eta * (T(1) - d * d);
    if (k < T(0)) return mk(T(0), T(0), T(0));
    return eta * I - (eta * d + std::sqrt(k)) * N;
}
template <class T> inline T fminT(T a, T b) { return std::fmin(a, b); }
template <class T> inline T fmaxT(T a, T b) { return std::fmax(a, b); }

template <class T> struct C {
    static constexpr T PI = T(3.14159265358979323);
    static constexpr T TWO_PI = T(6.28318530717958648);
    static constexpr T INF = T(1000000.0);
    static constexpr T EPS = T(0.001);
};

// ---------------------------------------------------------------------------
// RNG (globals.glsl:116-133)
// ---------------------------------------------------------------------------
inline void pcg4d(uint32_t* v) {
    v[0] = v[0] * 1664525u + 1013904223u;
    v[1] = v[1] * 1664525u + 1013904223u;
    v[2] = v[2] * 1664525u + 1013904223u;
    v[3] = v[3] * 1664525u + 1013904223u;
    v[0] += v[1] * v[3];
    v[1] += v[2] * v[0];
    v[2] += v[0] * v[1];
    v[3] += v[1] * v[2];
    v[0] ^= v[0] >> 16u;
    v[1] ^= v[1] >> 16u;
    v[2] ^= v[2] >> 16u;
    v[3] ^= v[3] >> 16u;
    v[0] += v[1] * v[3];
    v[1] += v[2] * v[0];
    v[2] += v[0] * v[1];
    v[3] += v[1] * v[2];
}
template <class T> inline T rnd(uint32_t* s);
template <> inline float rnd<float>(uint32_t* s) {
    pcg4d(s);
    return (float)s[0] / (float)0xffffffffu;
}
template <> inline double rnd<double>(uint32_t* s) {
    pcg4d(s);
    return (double)s[0] / 4294967295.0;
}

}  // namespace

struct orc_scene {
    std::vector<float> nodes;
    int32_t top = 0;
    std::vector<int32_t> vidx;
    std::vector<float> verts, nrms, mats, xforms, lights;
    int32_t nmat = 0, ninst = 0, nlights = 0;
    int32_t hw = 0, hh = 0;
    std::vector<float> hcols, hmarg, hcond;
    std::vector<uint8_t> tex;   // RGBA8 texture layers, first row first
    int32_t tw = 0, th = 0, tl = 0;
};

namespace {

// ---------------------------------------------------------------------------
// matrices
// ---------------------------------------------------------------------------
template <class T> struct M4 {
    T m[4][4];  // [col][row]
};
template <class T> inline T det3(T a00, T a01, T a02, T a10, T a11, T a12, T a20, T a21, T a22) {
    return a00 * (a11 * a22 - a12 * a21) - a01 * (a10 * a22 - a12 * a20) + a02 * (a10 * a21 - a11 * a20);
}
// inverse(mat4) = adj * (1/det), cofactor expansion (closest_hit.glsl:159-160)
template <class T> M4<T> inverse4(const M4<T>& a) {
    auto A = [&](int r, int c) { return a.m[c][r]; };
    T Cf[4][4];
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            int rr[3], cc[3];
            for (int k = 0, q = 0; k < 4; ++k)
                if (k != i) rr[q++] = k;
            for (int k = 0, q = 0; k < 4; ++k)
                if (k != j) cc[q++] = k;
            T d = det3(A(rr[0], cc[0]), A(rr[0], cc[1]), A(rr[0], cc[2]), A(rr[1], cc[0]), A(rr[1], cc[1]),
                       A(rr[1], cc[2]), A(rr[2], cc[0]), A(rr[2], cc[1]), A(rr[2], cc[2]));
            Cf[i][j] = ((i + j) & 1) ? -d : d;
        }
    T det = A(0, 0) * Cf[0][0] + A(0, 1) * Cf[0][1] + A(0, 2) * Cf[0][2] + A(0, 3) * Cf[0][3];
    T s = T(1) / det;
    M4<T> r;
    for (int row = 0; row < 4; ++row)
        for (int col = 0; col < 4; ++col) r.m[col][row] = Cf[col][row] * s;
    return r;
}
// M * vec4(p, w): ((c0 x + c1 y) + c2 z) + c3 w, xyz only
template <class T> inline V3<T> mulp(const M4<T>& a, V3<T> p, T w) {
    return mk(a.m[0][0] * p.x + a.m[1][0] * p.y + a.m[2][0] * p.z + a.m[3][0] * w,
              a.m[0][1] * p.x + a.m[1][1] * p.y + a.m[2][1] * p.z + a.m[3][1] * w,
              a.m[0][2] * p.x + a.m[1][2] * p.y + a.m[2][2] * p.z + a.m[3][2] * w);
}
// transpose(inverse(mat3(M))) * n (pathtrace.glsl:31-32)
template <class T> V3<T> normal_xform(const M4<T>& a, V3<T> n) {
    auto A = [&](int r, int c) { return a.m[c][r]; };
    T Cf[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            int rr[2], cc[2];
            for (int k = 0, q = 0; k < 3; ++k)
                if (k != i) rr[q++] = k;
            for (int k = 0, q = 0; k < 3; ++k)
                if (k != j) cc[q++] = k;
            T d = A(rr[0], cc[0]) * A(rr[1], cc[1]) - A(rr[0], cc[1]) * A(rr[1], cc[0]);
            Cf[i][j] = ((i + j) & 1) ? -d : d;
        }
    T det = A(0, 0) * Cf[0][0] + A(0, 1) * Cf[0][1] + A(0, 2) * Cf[0][2];
    T s = T(1) / det;
    T nm[3][3];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) nm[r][c] = Cf[r][c] * s;
    // mat3 * vec3 sums columns: component r = nm(r,0) n.x + nm(r,1) n.y + nm(r,2) n.z
    return mk(nm[0][0] * n.x + nm[0][1] * n.y + nm[0][2] * n.z, nm[1][0] * n.x + nm[1][1] * n.y + nm[1][2] * n.z,
              nm[2][0] * n.x + nm[2][1] * n.y + nm[2][2] * n.z);
}

template <class T> M4<T> xform_of(const orc_scene* s, int inst) {
    M4<T> r;
    for (int c = 0; c < 4; ++c)
        for (int k = 0; k < 4; ++k) r.m[c][k] = (T)s->xforms[(size_t)inst * 16 + c * 4 + k];
    return r;
}

// ---------------------------------------------------------------------------
// intersection.glsl
// ---------------------------------------------------------------------------
template <class T> struct Ray {
    V3<T> origin, direction;
};

// AABBIntersect (intersection.glsl:56-72)
template <class T> inline T aabb(V3<T> mn, V3<T> mx, const Ray<T>& r) {
    V3<T> invdir = mk(T(1) / r.direction.x, T(1) / r.direction.y, T(1) / r.direction.z);
    V3<T> f = (mx - r.origin) * invdir;
    V3<T> n = (mn - r.origin) * invdir;
    V3<T> tmax = mk(fmaxT(f.x, n.x), fmaxT(f.y, n.y), fmaxT(f.z, n.z));
    V3<T> tmin = mk(fminT(f.x, n.x), fminT(f.y, n.y), fminT(f.z, n.z));
    T t1 = fminT(tmax.x, fminT(tmax.y, tmax.z));
    T t0 = fmaxT(tmin.x, fmaxT(tmin.y, tmin.z));
    return (t1 >= t0) ? (t0 > T(0) ? t0 : t1) : T(-1);
}

// RectIntersect (intersection.glsl:29-50)
template <class T> inline T rect(V3<T> pos, V3<T> u, V3<T> v, V3<T> n, T pw, const Ray<T>& r) {
    T dt = dot(r.direction, n);
    T t = (pw - dot(n, r.origin)) / dt;
    if (t > C<T>::EPS) {
        V3<T> p = r.origin + r.direction * t;
        V3<T> vi = p - pos;
        T a1 = dot(u, vi);
        if (a1 >= T(0) && a1 <= T(1)) {
            T a2 = dot(v, vi);
            if (a2 >= T(0) && a2 <= T(1)) return t;
        }
    }
    return C<T>::INF;
}

// SphereIntersect (intersection.glsl:7-26)
template <class T> inline T sphere(T rad, V3<T> pos, const Ray<T>& r) {
    V3<T> op = pos - r.origin;
    T eps = T(0.001);
    T b = dot(op, r.direction);
    T det = b * b - dot(op, op) + rad * rad;
    if (det < T(0)) return C<T>::INF;
    det = std::sqrt(det);
    T t1 = b - det;
    if (t1 > eps) return t1;
    T t2 = b + det;
    if (t2 > eps) return t2;
    return C<T>::INF;
}

template <class T> struct Light {
    V3<T> position, emission, u, v;
    T radius, area, type;
};
template <class T> Light<T> light_of(const orc_scene* s, int i) {
    const float* l = &s->lights[(size_t)i * 15];
    Light<T> L;
    L.position = mk<T>(l[0], l[1], l[2]);
    L.emission = mk<T>(l[3], l[4], l[5]);
    L.u = mk<T>(l[6], l[7], l[8]);
    L.v = mk<T>(l[9], l[10], l[11]);
    L.radius = l[12];
    L.area = l[13];
    L.type = l[14];
    return L;
}

template <class T> inline V3<T> node_v(const orc_scene* s, int64_t n, int k) {
    const float* p = &s->nodes[(size_t)n * 9 + k * 3];
    return mk<T>(p[0], p[1], p[2]);
}

template <class T> struct Material {
    V3<T> albedo;
    T specular;
    V3<T> emission;
    T anisotropic, metallic, roughness, subsurface, specularTint, sheen, sheenTint, clearcoat, clearcoatRoughness,
        specTrans, ior, atDistance;
    V3<T> extinction;
};

template <class T> struct State {
    int depth;
    T eta, hitDist;
    V3<T> fhp, normal, ffnormal, tangent, bitangent;
    bool isEmitter;
    V3<T> bary;
    T texCoord[2];
    int triID[3];
    int matID;
    int inst;
    Material<T> mat;
};

template <class T> struct LightSampleRec {
    V3<T> normal, emission, direction;
    T dist, pdf;
};
template <class T> struct BsdfSampleRec {
    V3<T> L, f;
    T pdf;
};

struct Counts {
    int nint = 0, ntri = 0, ninst = 0;
    int64_t ties = 0;
};

// ---------------------------------------------------------------------------
// ClosestHit (closest_hit.glsl:7-205).  `lights`: the LIGHTS loop; tie_mode as
// in orc_trace_closest.  Besides the reference outputs it returns the hit's
// instance and (u, v).
// ---------------------------------------------------------------------------
template <class T>
T closest_hit(const orc_scene* s, const Ray<T>& r, State<T>& state, LightSampleRec<T>& lrec, bool lights,
              int tie_mode, T tmin, T tinit, Counts* cnt, int* hitPrim, T* hu, T* hv, M4<T>* hitXform) {
    T t = tinit;
    if (lights) {
        for (int i = 0; i < s->nlights; ++i) {
            Light<T> L = light_of<T>(s, i);
            if (L.type == T(0)) {
                V3<T> normal = normalize(cross(L.u, L.v));
                if (dot(normal, r.direction) > T(0)) continue;
                T pw = dot(normal, L.position);
                V3<T> u = L.u * (T(1) / dot(L.u, L.u));
                V3<T> v = L.v * (T(1) / dot(L.v, L.v));
                T d = rect(L.position, u, v, normal, pw, r);
                if (d < T(0)) d = C<T>::INF;
                if (d < t) {
                    t = d;
                    T cosTheta = dot(-r.direction, normal);
                    T pdf = (t * t) / (L.area * cosTheta);
                    lrec.emission = L.emission;
                    lrec.pdf = pdf;
                    state.isEmitter = true;
                }
            }
            if (L.type == T(1)) {
                T d = sphere(L.radius, L.position, r);
                if (d < T(0)) d = C<T>::INF;
                if (d < t) {
                    t = d;
                    T pdf = (t * t) / L.area;
                    lrec.emission = L.emission;
                    lrec.pdf = pdf;
                    state.isEmitter = true;
                }
            }
        }
    }
    int stack[64];
    int ptr = 0;
    stack[ptr++] = -1;
    int idx = s->top;
    int currMatID = 0, currInst = -1, bestInst = -1, bestPrim = -1;
    bool meshBVH = false;
    Ray<T> rt = r;
    M4<T> temp;
    for (int c = 0; c < 4; ++c)
        for (int k = 0; k < 4; ++k) temp.m[c][k] = (c == k) ? T(1) : T(0);
    while (idx > -1 || meshBVH) {
        int n = idx;
        if (meshBVH && idx < 0) {
            meshBVH = false;
            idx = stack[--ptr];
            rt = r;
            continue;
        }
        const float* lr = &s->nodes[(size_t)n * 9 + 6];
        int leftIndex = (int)lr[0], rightIndex = (int)lr[1], leaf = (int)lr[2];
        if (leaf > 0) {
            for (int i = 0; i < rightIndex; i++) {
                int index = leftIndex + i;
                const int32_t* vi = &s->vidx[(size_t)index * 3];
                const float* a = &s->verts[(size_t)vi[0] * 4];
                const float* b = &s->verts[(size_t)vi[1] * 4];
                const float* c = &s->verts[(size_t)vi[2] * 4];
                V3<T> v0 = mk<T>(a[0], a[1], a[2]), v1 = mk<T>(b[0], b[1], b[2]), v2 = mk<T>(c[0], c[1], c[2]);
                if (cnt) cnt->ntri++;
                V3<T> e0 = v1 - v0, e1 = v2 - v0;
                V3<T> pv = cross(rt.direction, e1);
                T det = dot(e0, pv);
                V3<T> tv = rt.origin - v0;
                V3<T> qv = cross(tv, e0);
                T ux = dot(tv, pv), uy = dot(rt.direction, qv), uz = dot(e1, qv);
                ux = ux / det;
                uy = uy / det;
                uz = uz / det;
                T uw = T(1) - ux - uy;
                bool ok = ux >= T(0) && uy >= T(0) && uz >= T(0) && uw >= T(0) && uz >= tmin;
                if (!ok) continue;
                int prim = vi[0] / 3;
                bool better = uz < t;
                if (!better && uz == t && bestPrim >= 0) {
                    if (cnt) cnt->ties++;
                    if (tie_mode == 1)
                        better = currInst < bestInst || (currInst == bestInst && prim < bestPrim);
                }
                if (better) {
                    t = uz;
                    state.isEmitter = false;
                    state.triID[0] = vi[0];
                    state.triID[1] = vi[1];
                    state.triID[2] = vi[2];
                    state.matID = currMatID;
                    state.inst = currInst;
                    V3<T> p = rt.origin + rt.direction * t;
                    state.bary = mk(uw, ux, uy);
                    state.fhp = mulp(temp, p, T(1));
                    bestInst = currInst;
                    bestPrim = prim;
                    if (hitPrim) *hitPrim = prim;
                    if (hu) *hu = ux;
                    if (hv) *hv = uy;
                    if (hitXform) *hitXform = temp;
                }
            }
        } else if (leaf < 0) {
            idx = leftIndex;
            int inst = -leaf - 1;
            if (cnt) cnt->ninst++;
            temp = xform_of<T>(s, inst);
            M4<T> inv = inverse4(temp);
            rt.origin = mulp(inv, r.origin, T(1));
            rt.direction = mulp(inv, r.direction, T(0));
            stack[ptr++] = -1;
            meshBVH = true;
            currMatID = rightIndex;
            currInst = inst;
            continue;
        } else {
            if (cnt) cnt->nint++;
            T leftHit = aabb(node_v<T>(s, leftIndex, 0), node_v<T>(s, leftIndex, 1), rt);
            T rightHit = aabb(node_v<T>(s, rightIndex, 0), node_v<T>(s, rightIndex, 1), rt);
            if (leftHit > T(0) && rightHit > T(0)) {
                int deferred;
                if (leftHit > rightHit) {
                    idx = rightIndex;
                    deferred = leftIndex;
                } else {
                    idx = leftIndex;
                    deferred = rightIndex;
                }
                stack[ptr++] = deferred;
                continue;
            } else if (leftHit > T(0)) {
                idx = leftIndex;
                continue;
            } else if (rightHit > T(0)) {
                idx = rightIndex;
                continue;
            }
        }
        idx = stack[--ptr];
    }
    state.hitDist = t;
    if (hitPrim && bestPrim < 0 && !state.isEmitter) *hitPrim = -1;
    return t;
}

// AnyHit (anyhit.glsl:7-172)
template <class T> bool any_hit(const orc_scene* s, const Ray<T>& r, T maxDist, bool lights, T tmin) {
    if (lights) {
        for (int i = 0; i < s->nlights; ++i) {
            Light<T> L = light_of<T>(s, i);
            if (L.type == T(0)) {
                V3<T> normal = normalize(cross(L.u, L.v));
                T pw = dot(normal, L.position);
                V3<T> u = L.u * (T(1) / dot(L.u, L.u));
                V3<T> v = L.v * (T(1) / dot(L.v, L.v));
                T d = rect(L.position, u, v, normal, pw, r);
                if (d > T(0) && d < maxDist) return true;
            }
            if (L.type == T(1)) {
                T d = sphere(L.radius, L.position, r);
                if (d > T(0) && d < maxDist) return true;
            }
        }
    }
    int stack[64];
    int ptr = 0;
    stack[ptr++] = -1;
    int idx = s->top;
    bool meshBVH = false;
    Ray<T> rt = r;
    while (idx > -1 || meshBVH) {
        int n = idx;
        if (meshBVH && idx < 0) {
            meshBVH = false;
            idx = stack[--ptr];
            rt = r;
            continue;
        }
        const float* lr = &s->nodes[(size_t)n * 9 + 6];
        int leftIndex = (int)lr[0], rightIndex = (int)lr[1], leaf = (int)lr[2];
        if (leaf > 0) {
            for (int i = 0; i < rightIndex; i++) {
                int index = leftIndex + i;
                const int32_t* vi = &s->vidx[(size_t)index * 3];
                const float* a = &s->verts[(size_t)vi[0] * 4];
                const float* b = &s->verts[(size_t)vi[1] * 4];
                const float* c = &s->verts[(size_t)vi[2] * 4];
                V3<T> v0 = mk<T>(a[0], a[1], a[2]), v1 = mk<T>(b[0], b[1], b[2]), v2 = mk<T>(c[0], c[1], c[2]);
                V3<T> e0 = v1 - v0, e1 = v2 - v0;
                V3<T> pv = cross(rt.direction, e1);
                T det = dot(e0, pv);
                V3<T> tv = rt.origin - v0;
                V3<T> qv = cross(tv, e0);
                T ux = dot(tv, pv) / det, uy = dot(rt.direction, qv) / det, uz = dot(e1, qv) / det;
                T uw = T(1) - ux - uy;
                if (ux >= T(0) && uy >= T(0) && uz >= T(0) && uw >= T(0) && uz >= tmin && uz < maxDist) return true;
            }
        } else if (leaf < 0) {
            idx = leftIndex;
            M4<T> inv = inverse4(xform_of<T>(s, -leaf - 1));
            rt.origin = mulp(inv, r.origin, T(1));
            rt.direction = mulp(inv, r.direction, T(0));
            stack[ptr++] = -1;
            meshBVH = true;
            continue;
        } else {
            T leftHit = aabb(node_v<T>(s, leftIndex, 0), node_v<T>(s, leftIndex, 1), rt);
            T rightHit = aabb(node_v<T>(s, rightIndex, 0), node_v<T>(s, rightIndex, 1), rt);
            if (leftHit > T(0) && rightHit > T(0)) {
                int deferred;
                if (leftHit > rightHit) {
                    idx = rightIndex;
                    deferred = leftIndex;
                } else {
                    idx = leftIndex;
                    deferred = rightIndex;
                }
                stack[ptr++] = deferred;
                continue;
            } else if (leftHit > T(0)) {
                idx = leftIndex;
                continue;
            } else if (rightHit > T(0)) {
                idx = rightIndex;
                continue;
            }
        }
        idx = stack[--ptr];
    }
    return false;
}

// ---------------------------------------------------------------------------
// sampling.glsl helpers
// ---------------------------------------------------------------------------
template <class T> V3<T> ImportanceSampleGTR1(T rgh, T r1, T r2) {
    T a = std::fmax(T(0.001), rgh);
    T a2 = a * a;
    T phi = r1 * C<T>::TWO_PI;
    T cosTheta = std::sqrt((T(1) - std::pow(a2, T(1) - r1)) / (T(1) - a2));
    T sinTheta = clampT(std::sqrt(T(1) - (cosTheta * cosTheta)), T(0), T(1));
    T sinPhi = std::sin(phi), cosPhi = std::cos(phi);
    return mk(sinTheta * cosPhi, sinTheta * sinPhi, cosTheta);
}
template <class T> V3<T> ImportanceSampleGTR2(T rgh, T r1, T r2) {
    T a = std::fmax(T(0.001), rgh);
    T phi = r1 * C<T>::TWO_PI;
    T cosTheta = std::sqrt((T(1) - r2) / (T(1) + (a * a - T(1)) * r2));
    T sinTheta = clampT(std::sqrt(T(1) - (cosTheta * cosTheta)), T(0), T(1));
    T sinPhi = std::sin(phi), cosPhi = std::cos(phi);
    return mk(sinTheta * cosPhi, sinTheta * sinPhi, cosTheta);
}
template <class T> T SchlickFresnel(T u) {
    T m = clampT(T(1) - u, T(0), T(1));
    T m2 = m * m;
    return m2 * m2 * m;
}
template <class T> T DielectricFresnel(T cos_theta_i, T eta) {
    T sinThetaTSq = eta * eta * (T(1) - cos_theta_i * cos_theta_i);
    if (sinThetaTSq > T(1)) return T(1);
    T cos_theta_t = std::sqrt(std::fmax(T(1) - sinThetaTSq, T(0)));
    T rs = (eta * cos_theta_t - cos_theta_i) / (eta * cos_theta_t + cos_theta_i);
    T rp = (eta * cos_theta_i - cos_theta_t) / (eta * cos_theta_i + cos_theta_t);
    return T(0.5) * (rs * rs + rp * rp);
}
template <class T> T GTR1(T NDotH, T a) {
    if (a >= T(1)) return (T(1) / C<T>::PI);
    T a2 = a * a;
    T t = T(1) + (a2 - T(1)) * NDotH * NDotH;
    return (a2 - T(1)) / (C<T>::PI * std::log(a2) * t);
}
template <class T> T GTR2(T NDotH, T a) {
    T a2 = a * a;
    T t = T(1) + (a2 - T(1)) * NDotH * NDotH;
    return a2 / (C<T>::PI * t * t);
}
template <class T> T SmithG_GGX(T NDotV, T alphaG) {
    T a = alphaG * alphaG;
    T b = NDotV * NDotV;
    return T(1) / (NDotV + std::sqrt(a + b - a * b));
}
template <class T> V3<T> CosineSampleHemisphere(T r1, T r2) {
    V3<T> dir;
    T r = std::sqrt(r1);
    T phi = C<T>::TWO_PI * r2;
    dir.x = r * std::cos(phi);
    dir.y = r * std::sin(phi);
    dir.z = std::sqrt(std::fmax(T(0), T(1) - dir.x * dir.x - dir.y * dir.y));
    return dir;
}
template <class T> V3<T> UniformSampleSphere(T r1, T r2) {
    T z = T(1) - T(2) * r1;
    T r = std::sqrt(std::fmax(T(0), T(1) - z * z));
    T phi = C<T>::TWO_PI * r2;
    return mk(r * std::cos(phi), r * std::sin(phi), z);
}
template <class T> T powerHeuristic(T a, T b) {
    T t = a * a;
    return t / (b * b + t);
}

// sampleSphereLight / sampleRectLight / sampleDistantLight (sampling.glsl:172-216)
template <class T>
void sampleOneLight(const orc_scene* s, const Light<T>& light, V3<T> surfacePos, LightSampleRec<T>& rec, uint32_t* seed) {
    int type = (int)light.type;
    T nl = (T)s->nlights;
    if (type == 0) {
        T r1 = rnd<T>(seed), r2 = rnd<T>(seed);
        V3<T> lsp = light.position + light.u * r1 + light.v * r2;
        rec.direction = lsp - surfacePos;
        rec.dist = length(rec.direction);
        T distSq = rec.dist * rec.dist;
        rec.direction = rec.direction / rec.dist;
        rec.normal = normalize(cross(light.u, light.v));
        rec.emission = light.emission * nl;
        rec.pdf = distSq / (light.area * std::fabs(dot(rec.normal, rec.direction)));
    } else if (type == 1) {
        T r1 = rnd<T>(seed), r2 = rnd<T>(seed);
        V3<T> lsp = light.position + UniformSampleSphere(r1, r2) * light.radius;
        rec.direction = lsp - surfacePos;
        rec.dist = length(rec.direction);
        T distSq = rec.dist * rec.dist;
        rec.direction = rec.direction / rec.dist;
        rec.normal = normalize(lsp - light.position);
        rec.emission = light.emission * nl;
        rec.pdf = distSq / (light.area * std::fabs(dot(rec.normal, rec.direction)));
    } else {
        rec.direction = normalize(light.position - mk(T(0), T(0), T(0)));
        rec.normal = normalize(surfacePos - light.position);
        rec.emission = light.emission * nl;
        rec.dist = C<T>::INF;
        rec.pdf = T(1);
    }
}

// GL texture lookups on the HDR tables (GL_REPEAT wrap).
inline int wrapi(int i, int n) {
    int r = i % n;
    return r < 0 ? r + n : r;
}
template <class T> inline void tex_nearest2(const std::vector<float>& tab, int W, int H, T s, T t, T* out) {
    T u = s * (T)W, v = t * (T)H;
    int i = wrapi((int)std::floor(u), W), j = wrapi((int)std::floor(v), H);
    out[0] = tab[((size_t)j * W + i) * 2];
    out[1] = tab[((size_t)j * W + i) * 2 + 1];
}
template <class T> inline V3<T> tex_linear3(const std::vector<float>& cols, int W, int H, T s, T t) {
    T u = s * (T)W - T(0.5), v = t * (T)H - T(0.5);
    T fu = std::floor(u), fv = std::floor(v);
    T a = u - fu, b = v - fv;
    int i0 = wrapi((int)fu, W), i1 = wrapi((int)fu + 1, W);
    int j0 = wrapi((int)fv, H), j1 = wrapi((int)fv + 1, H);
    auto px = [&](int i, int j) {
        const float* c = &cols[((size_t)j * W + i) * 3];
        return mk<T>(c[0], c[1], c[2]);
    };
    V3<T> top = px(i0, j0) * (T(1) - a) + px(i1, j0) * a;
    V3<T> bot = px(i0, j1) * (T(1) - a) + px(i1, j1) * a;
    return top * (T(1) - b) + bot * b;
}

// EnvPdf (sampling.glsl:236-243)
template <class T> T EnvPdf(const orc_scene* s, const Ray<T>& r) {
    T theta = std::acos(clampT(r.direction.y, T(-1), T(1)));
    T u = (C<T>::PI + std::atan2(r.direction.z, r.direction.x)) * (T(1) / C<T>::TWO_PI);
    T v = theta * (T(1) / C<T>::PI);
    T c[2], m[2];
    tex_nearest2<T>(s->hcond, s->hw, s->hh, u, v, c);
    tex_nearest2<T>(s->hmarg, s->hh, 1, v, T(0), m);
    T pdf = c[1] * m[1];
    T hres = (T)(s->hw * s->hh);
    return (pdf * hres) / (T(2) * C<T>::PI * C<T>::PI * std::sin(theta));
}

// EnvSample (sampling.glsl:246-265); returns dir + pdf, colour out
template <class T> void EnvSample(const orc_scene* s, uint32_t* seed, T mult, V3<T>& color, V3<T>& dir, T& pdfOut) {
    T r1 = rnd<T>(seed), r2 = rnd<T>(seed);
    T m[2], c[2];
    tex_nearest2<T>(s->hmarg, s->hh, 1, r1, T(0), m);
    T v = m[0];
    tex_nearest2<T>(s->hcond, s->hw, s->hh, r2, v, c);
    T u = c[0];
    color = tex_linear3<T>(s->hcols, s->hw, s->hh, u, v) * mult;
    T cp[2], mp[2];
    tex_nearest2<T>(s->hcond, s->hw, s->hh, u, v, cp);
    tex_nearest2<T>(s->hmarg, s->hh, 1, v, T(0), mp);
    T pdf = cp[1] * mp[1];
    T phi = u * C<T>::TWO_PI;
    T theta = v * C<T>::PI;
    if (std::sin(theta) == T(0)) pdf = T(0);
    T hres = (T)(s->hw * s->hh);
    dir = mk(-std::sin(theta) * std::cos(phi), std::cos(theta), -std::sin(theta) * std::sin(phi));
    pdfOut = (pdf * hres) / (T(2) * C<T>::PI * C<T>::PI * std::sin(theta));
}

// ---------------------------------------------------------------------------
// disney.glsl
// ---------------------------------------------------------------------------
template <class T> V3<T> EvalDielectricReflection(const State<T>& st, V3<T> V, V3<T> N, V3<T> L, V3<T> H, T& pdf) {
    pdf = T(0);
    if (dot(N, L) <= T(0)) return mk(T(0), T(0), T(0));
    T F = DielectricFresnel(dot(V, H), st.eta);
    T D = GTR2(dot(N, H), st.mat.roughness);
    pdf = D * dot(N, H) * F / (T(4) * std::fabs(dot(V, H)));
    T G = SmithG_GGX(std::fabs(dot(N, L)), st.mat.roughness) * SmithG_GGX(std::fabs(dot(N, V)), st.mat.roughness);
    return st.mat.albedo * F * D * G;
}
template <class T> V3<T> EvalDielectricRefraction(const State<T>& st, V3<T> V, V3<T> N, V3<T> L, V3<T> H, T& pdf) {
    pdf = T(0);
    if (dot(N, L) >= T(0)) return mk(T(0), T(0), T(0));
    T F = DielectricFresnel(std::fabs(dot(V, H)), st.eta);
    T D = GTR2(dot(N, H), st.mat.roughness);
    T denomSqrt = dot(L, H) + dot(V, H) * st.eta;
    pdf = D * dot(N, H) * (T(1) - F) * std::fabs(dot(L, H)) / (denomSqrt * denomSqrt);
    T G = SmithG_GGX(std::fabs(dot(N, L)), st.mat.roughness) * SmithG_GGX(std::fabs(dot(N, V)), st.mat.roughness);
    return st.mat.albedo * (T(1) - F) * D * G * std::fabs(dot(V, H)) * std::fabs(dot(L, H)) * T(4) * st.eta * st.eta /
           (denomSqrt * denomSqrt);
}
template <class T> V3<T> EvalSpecular(const State<T>& st, V3<T> Cspec0, V3<T> V, V3<T> N, V3<T> L, V3<T> H, T& pdf) {
    pdf = T(0);
    if (dot(N, L) <= T(0)) return mk(T(0), T(0), T(0));
    T D = GTR2(dot(N, H), st.mat.roughness);
    pdf = D * dot(N, H) / (T(4) * dot(V, H));
    T FH = SchlickFresnel(dot(L, H));
    V3<T> F = mixv(Cspec0, mk(T(1), T(1), T(1)), FH);
    T G = SmithG_GGX(std::fabs(dot(N, L)), st.mat.roughness) * SmithG_GGX(std::fabs(dot(N, V)), st.mat.roughness);
    return F * D * G;
}
template <class T> V3<T> EvalClearcoat(const State<T>& st, V3<T> V, V3<T> N, V3<T> L, V3<T> H, T& pdf) {
    pdf = T(0);
    if (dot(N, L) <= T(0)) return mk(T(0), T(0), T(0));
    T D = GTR1(dot(N, H), mixs(T(0.1), T(0.001), st.mat.clearcoatRoughness));
    pdf = D * dot(N, H) / (T(4) * dot(V, H));
    T FH = SchlickFresnel(dot(L, H));
    T F = mixs(T(0.04), T(1), FH);
    T G = SmithG_GGX(dot(N, L), T(0.25)) * SmithG_GGX(dot(N, V), T(0.25));
    T c = T(0.25) * st.mat.clearcoat * F * D * G;
    return mk(c, c, c);
}
template <class T> V3<T> EvalDiffuse(const State<T>& st, V3<T> Csheen, V3<T> V, V3<T> N, V3<T> L, V3<T> H, T& pdf) {
    pdf = T(0);
    if (dot(N, L) <= T(0)) return mk(T(0), T(0), T(0));
    pdf = dot(N, L) * (T(1) / C<T>::PI);
    T FL = SchlickFresnel(dot(N, L));
    T FV = SchlickFresnel(dot(N, V));
    T FH = SchlickFresnel(dot(L, H));
    T Fss90 = dot(L, H) * dot(L, H) * st.mat.roughness;
    T Fss = mixs(T(1), Fss90, FL) * mixs(T(1), Fss90, FV);
    T ss = T(1) * (Fss * (T(1) / (dot(N, L) + dot(N, V)) - T(0.5)) + T(0.5));
    V3<T> Fsheen = (FH * st.mat.sheen) * Csheen;
    return ((T(1) / C<T>::PI) * (ss + st.mat.subsurface) * st.mat.albedo + Fsheen) * (T(1) - st.mat.metallic);
}

template <class T> void tints(const State<T>& st, V3<T>& Cspec0, V3<T>& Csheen) {
    V3<T> Cdlin = st.mat.albedo;
    T Cdlum = T(0.3) * Cdlin.x + T(0.6) * Cdlin.y + T(0.1) * Cdlin.z;
    V3<T> Ctint = Cdlum > T(0) ? Cdlin / Cdlum : mk(T(1), T(1), T(1));
    V3<T> one = mk(T(1), T(1), T(1));
    Cspec0 = mixv((st.mat.specular * T(0.08)) * mixv(one, Ctint, st.mat.specularTint), Cdlin, st.mat.metallic);
    Csheen = mixv(one, Ctint, st.mat.sheenTint);
}

// DisneySample (disney.glsl:128-225)
template <class T> V3<T> DisneySample(State<T>& st, V3<T> V, V3<T> N, V3<T>& L, T& pdf, uint32_t* seed) {
    pdf = T(0);
    V3<T> f = mk(T(0), T(0), T(0));
    T r1 = rnd<T>(seed), r2 = rnd<T>(seed);
    T diffuseRatio = T(0.5) * (T(1) - st.mat.metallic);
    T transWeight = (T(1) - st.mat.metallic) * st.mat.specTrans;
    V3<T> Cspec0, Csheen;
    tints(st, Cspec0, Csheen);
    if (rnd<T>(seed) < transWeight) {
        V3<T> H = ImportanceSampleGTR2(st.mat.roughness, r1, r2);
        H = st.tangent * H.x + st.bitangent * H.y + N * H.z;
        if (dot(V, H) < T(0)) H = -H;
        V3<T> R = reflect(-V, H);
        T F = DielectricFresnel(std::fabs(dot(R, H)), st.eta);
        if (rnd<T>(seed) < F) {
            L = normalize(R);
            f = EvalDielectricReflection(st, V, N, L, H, pdf);
        } else {
            L = normalize(refract(-V, H, st.eta));
            f = EvalDielectricRefraction(st, V, N, L, H, pdf);
        }
        f = f * transWeight;
        pdf *= transWeight;
    } else {
        if (rnd<T>(seed) < diffuseRatio) {
            L = CosineSampleHemisphere(r1, r2);
            L = st.tangent * L.x + st.bitangent * L.y + N * L.z;
            V3<T> H = normalize(L + V);
            f = EvalDiffuse(st, Csheen, V, N, L, H, pdf);
            pdf *= diffuseRatio;
        } else {
            T primarySpecRatio = T(1) / (T(1) + st.mat.clearcoat);
            if (rnd<T>(seed) < primarySpecRatio) {
                V3<T> H = ImportanceSampleGTR2(st.mat.roughness, r1, r2);
                H = st.tangent * H.x + st.bitangent * H.y + N * H.z;
                if (dot(V, H) < T(0)) H = -H;
                L = normalize(reflect(-V, H));
                f = EvalSpecular(st, Cspec0, V, N, L, H, pdf);
                pdf *= primarySpecRatio * (T(1) - diffuseRatio);
            } else {
                V3<T> H = ImportanceSampleGTR1(mixs(T(0.1), T(0.001), st.mat.clearcoatRoughness), r1, r2);
                H = st.tangent * H.x + st.bitangent * H.y + N * H.z;
                if (dot(V, H) < T(0)) H = -H;
                L = normalize(reflect(-V, H));
                f = EvalClearcoat(st, V, N, L, H, pdf);
                pdf *= (T(1) - primarySpecRatio) * (T(1) - diffuseRatio);
            }
        }
        f = f * (T(1) - transWeight);
        pdf *= (T(1) - transWeight);
    }
    return f;
}

// DisneyEval (disney.glsl:228-291)
template <class T> V3<T> DisneyEval(const State<T>& st, V3<T> V, V3<T> N, V3<T> L, T& pdf) {
    V3<T> H;
    bool refl = dot(N, L) > T(0);
    if (refl) H = normalize(L + V);
    else H = normalize(L + V * st.eta);
    if (dot(V, H) < T(0)) H = -H;
    T diffuseRatio = T(0.5) * (T(1) - st.mat.metallic);
    T primarySpecRatio = T(1) / (T(1) + st.mat.clearcoat);
    T transWeight = (T(1) - st.mat.metallic) * st.mat.specTrans;
    V3<T> brdf = mk(T(0), T(0), T(0)), bsdf = mk(T(0), T(0), T(0));
    T brdfPdf = T(0), bsdfPdf = T(0);
    if (transWeight > T(0)) {
        if (refl) bsdf = EvalDielectricReflection(st, V, N, L, H, bsdfPdf);
        else bsdf = EvalDielectricRefraction(st, V, N, L, H, bsdfPdf);
    }
    T m_pdf;
    if (transWeight < T(1)) {
        V3<T> Cspec0, Csheen;
        tints(st, Cspec0, Csheen);
        brdf = brdf + EvalDiffuse(st, Csheen, V, N, L, H, m_pdf);
        brdfPdf += m_pdf * diffuseRatio;
        brdf = brdf + EvalSpecular(st, Cspec0, V, N, L, H, m_pdf);
        brdfPdf += m_pdf * primarySpecRatio * (T(1) - diffuseRatio);
        brdf = brdf + EvalClearcoat(st, V, N, L, H, m_pdf);
        brdfPdf += m_pdf * (T(1) - primarySpecRatio) * (T(1) - diffuseRatio);
    }
    pdf = mixs(brdfPdf, bsdfPdf, transWeight);
    return mixv(brdf, bsdf, transWeight);
}

// ---------------------------------------------------------------------------
// pathtrace.glsl
// ---------------------------------------------------------------------------
template <class T> void Onb(V3<T> N, V3<T>& Tg, V3<T>& B) {
    V3<T> up = std::fabs(N.z) < T(0.999) ? mk(T(0), T(0), T(1)) : mk(T(1), T(0), T(0));
    Tg = normalize(cross(up, N));
    B = cross(N, Tg);
}

template <class T> Material<T> material_of(const orc_scene* s, int id) {
    const float* p = &s->mats[(size_t)id * 28];
    Material<T> m;
    m.albedo = mk<T>(p[0], p[1], p[2]);
    m.specular = p[3];
    m.emission = mk<T>(p[4], p[5], p[6]);
    m.anisotropic = p[7];
    m.metallic = p[8];
    m.roughness = std::fmax((T)p[9], T(0.001));
    m.subsurface = p[10];
    m.specularTint = p[11];
    m.sheen = p[12];
    m.sheenTint = p[13];
    m.clearcoat = p[14];
    m.clearcoatRoughness = p[15];
    m.specTrans = p[16];
    m.ior = p[17];
    m.atDistance = p[18];
    m.extinction = mk<T>(p[20], p[21], p[22]);
    return m;
}

// GetNormalsAndTexCoord (pathtrace.glsl:16-37)
template <class T> void GetNormals(const orc_scene* s, State<T>& st, const Ray<T>& r, const M4<T>& xf) {
    const float* n1 = &s->nrms[(size_t)st.triID[0] * 4];
    const float* n2 = &s->nrms[(size_t)st.triID[1] * 4];
    const float* n3 = &s->nrms[(size_t)st.triID[2] * 4];
    const float u1 = s->verts[(size_t)st.triID[0] * 4 + 3], u2 = s->verts[(size_t)st.triID[1] * 4 + 3],
                u3 = s->verts[(size_t)st.triID[2] * 4 + 3];
    st.texCoord[0] = (T)u1 * st.bary.x + (T)u2 * st.bary.y + (T)u3 * st.bary.z;
    st.texCoord[1] = (T)n1[3] * st.bary.x + (T)n2[3] * st.bary.y + (T)n3[3] * st.bary.z;
    V3<T> a = mk<T>(n1[0], n1[1], n1[2]), b = mk<T>(n2[0], n2[1], n2[2]), c = mk<T>(n3[0], n3[1], n3[2]);
    V3<T> normal = normalize(a * st.bary.x + b * st.bary.y + c * st.bary.z);
    normal = normalize(normal_xform(xf, normal));
    st.normal = normal;
    st.ffnormal = dot(normal, r.direction) <= T(0) ? normal : normal * T(-1);
    Onb(st.normal, st.tangent, st.bitangent);
}

// texture() on the RGBA8 array texture (Renderer.cpp:151-160): GL_LINEAR,
// GL_REPEAT (the default wrap), layer clamped to the array (OpenGL 4.6,
// 8.14.2); unsigned normalized channels c / 255.
template <class T> void tex_rgba8(const orc_scene* s, T sx, T ty, int layer, T out[4]) {
    const int W = s->tw, H = s->th;
    layer = layer < 0 ? 0 : (layer >= s->tl ? s->tl - 1 : layer);
    T u = sx * (T)W - T(0.5), v = ty * (T)H - T(0.5);
    T fu = std::floor(u), fv = std::floor(v);
    T a = u - fu, b = v - fv;
    fu -= (T)W * std::floor(fu / (T)W);
    fv -= (T)H * std::floor(fv / (T)H);
    int i0 = wrapi((int)fu, W), i1 = wrapi((int)fu + 1, W);
    int j0 = wrapi((int)fv, H), j1 = wrapi((int)fv + 1, H);
    const uint8_t* L = &s->tex[(size_t)layer * W * H * 4];
    auto px = [&](int i, int j, int c) { return (T)L[((size_t)j * W + i) * 4 + c] / T(255); };
    for (int c = 0; c < 4; ++c) {
        T top = px(i0, j0, c) * (T(1) - a) + px(i1, j0, c) * a;
        T bot = px(i0, j1, c) * (T(1) - a) + px(i1, j1, c) * a;
        out[c] = top * (T(1) - b) + bot * b;
    }
}

// GetMaterialsAndTextures (pathtrace.glsl:40-123)
template <class T> void GetMaterial(const orc_scene* s, State<T>& st, const Ray<T>& r) {
    st.mat = material_of<T>(s, st.matID);
    if (s->tl > 0) {
        const float* ids = &s->mats[(size_t)st.matID * 28 + 24];
        const T sx = st.texCoord[0], ty = T(1) - st.texCoord[1];
        T c[4];
        auto pw = [](T x) { return std::pow(x, T(2.2)); };
        if ((int)ids[0] >= 0) {
            tex_rgba8<T>(s, sx, ty, (int)ids[0], c);
            st.mat.albedo = st.mat.albedo * mk(pw(c[0]), pw(c[1]), pw(c[2]));
        }
        if ((int)ids[1] >= 0) {
            tex_rgba8<T>(s, sx, ty, (int)ids[1], c);
            st.mat.metallic = c[0];
            st.mat.roughness = std::fmax(c[1] * c[1], T(0.001));
        }
        if ((int)ids[2] >= 0) {
            tex_rgba8<T>(s, sx, ty, (int)ids[2], c);
            V3<T> n = normalize(mk(c[0] * T(2) - T(1), c[1] * T(2) - T(1), c[2] * T(2) - T(1)));
            V3<T> Tg, B;
            Onb(st.normal, Tg, B);
            st.normal = normalize(Tg * n.x + B * n.y + st.normal * n.z);
            st.ffnormal = dot(st.normal, r.direction) <= T(0) ? st.normal : st.normal * T(-1);
            Onb(st.normal, st.tangent, st.bitangent);
        }
        if (ids[3] >= 0.0f) {
            tex_rgba8<T>(s, sx, ty, (int)ids[3], c);
            st.mat.emission = mk(pw(c[0]), pw(c[1]), pw(c[2]));
        }
    }
    st.eta = dot(r.direction, st.normal) < T(0) ? (T(1) / st.mat.ior) : st.mat.ior;
}

// DirectLight (pathtrace.glsl:126-204)
template <class T>
V3<T> DirectLight(const orc_scene* s, const orc_options* o, const Ray<T>& r, const State<T>& st, uint32_t* seed) {
    V3<T> Li = mk(T(0), T(0), T(0));
    // estimator 1: offset to the side the ray came from (the reference uses the
    // unflipped normal, pathtrace.glsl:130)
    V3<T> surfacePos = st.fhp + (o->estimator == 1 ? st.ffnormal : st.normal) * C<T>::EPS;
    BsdfSampleRec<T> b;
    if (o->envmap && !o->constant_bg) {
        V3<T> color, lightDir;
        T lightPdf;
        EnvSample<T>(s, seed, (T)o->hdr_multiplier, color, lightDir, lightPdf);
        Ray<T> shadow{surfacePos, lightDir};
        bool inShadow = any_hit<T>(s, shadow, C<T>::INF - C<T>::EPS, o->lights != 0, T(0));
        if (!inShadow) {
            b.f = DisneyEval(st, -r.direction, st.ffnormal, lightDir, b.pdf);
            if (b.pdf > T(0)) {
                T misWeight = powerHeuristic(lightPdf, b.pdf);
                if (misWeight > T(0)) Li = Li + misWeight * b.f * std::fabs(dot(lightDir, st.ffnormal)) * color / lightPdf;
            }
        }
    }
    if (o->lights) {
        LightSampleRec<T> lr;
        int index = (int)(rnd<T>(seed) * (T)s->nlights);
        // rand() == 1.0 picks index numOfLights (trap in the reference, which
        // then reads past the light texture); clamped to the last light here.
        if (index >= s->nlights) index = s->nlights - 1;
        Light<T> light = light_of<T>(s, index);
        sampleOneLight(s, light, surfacePos, lr, seed);
        if (dot(lr.direction, lr.normal) < T(0)) {
            Ray<T> shadow{surfacePos, lr.direction};
            bool inShadow = any_hit<T>(s, shadow, lr.dist - C<T>::EPS, true, T(0));
            if (!inShadow) {
                b.f = DisneyEval(st, -r.direction, st.ffnormal, lr.direction, b.pdf);
                T weight = T(1);
                if (light.area > T(0)) weight = powerHeuristic(lr.pdf, b.pdf);
                if (b.pdf > T(0)) Li = Li + weight * b.f * std::fabs(dot(st.ffnormal, lr.direction)) * lr.emission / lr.pdf;
            }
        }
    }
    return Li;
}

// PathTrace (pathtrace.glsl:208-295)
template <class T> V3<T> PathTrace(const orc_scene* s, const orc_options* o, Ray<T> r, uint32_t* seed) {
    V3<T> radiance = mk(T(0), T(0), T(0));
    V3<T> throughput = mk(T(1), T(1), T(1));
    State<T> st{};
    st.triID[0] = 0;
    st.triID[1] = 1;
    st.triID[2] = 2;
    st.matID = 0;
    st.inst = -1;
    LightSampleRec<T> lrec{};
    BsdfSampleRec<T> brec{};
    V3<T> absorption = mk(T(0), T(0), T(0));
    T misPdf = T(0);   // density of the BSDF strategy in the MIS weights of the next hit
    M4<T> xf;
    for (int c = 0; c < 4; ++c)
        for (int k = 0; k < 4; ++k) xf.m[c][k] = (c == k) ? T(1) : T(0);
    for (int depth = 0; depth < o->max_depth; depth++) {
        st.depth = depth;
        M4<T> hx = xf;
        T t = closest_hit<T>(s, r, st, lrec, o->lights != 0, 1, T(0), C<T>::INF, nullptr, nullptr, nullptr, nullptr, &hx);
        if (!st.isEmitter) xf = hx;
        if (t == C<T>::INF) {
            if (o->constant_bg) {
                radiance = radiance + mk<T>(o->bg[0], o->bg[1], o->bg[2]) * throughput;
            } else if (o->envmap) {
                T misWeight = T(1);
                T u = (C<T>::PI + std::atan2(r.direction.z, r.direction.x)) * (T(1) / C<T>::TWO_PI);
                T v = std::acos(r.direction.y) * (T(1) / C<T>::PI);
                if (depth > 0) {
                    T lightPdf = EnvPdf<T>(s, r);
                    misWeight = powerHeuristic(misPdf, lightPdf);
                }
                radiance = radiance + misWeight * tex_linear3<T>(s->hcols, s->hw, s->hh, u, v) * throughput *
                                          (T)o->hdr_multiplier;
            }
            return radiance;
        }
        GetNormals(s, st, r, xf);
        GetMaterial(s, st, r);
        if (dot(st.normal, st.ffnormal) > T(0)) absorption = mk(T(0), T(0), T(0));
        radiance = radiance + st.mat.emission * throughput;
        if (o->lights && st.isEmitter) {
            V3<T> Le = (st.depth == 0) ? lrec.emission : powerHeuristic(misPdf, lrec.pdf) * lrec.emission;
            radiance = radiance + Le * throughput;
            break;
        }
        throughput = throughput * mk(std::exp(-absorption.x * t), std::exp(-absorption.y * t), std::exp(-absorption.z * t));
        radiance = radiance + DirectLight(s, o, r, st, seed) * throughput;
        brec.f = DisneySample(st, -r.direction, st.ffnormal, brec.L, brec.pdf, seed);
        misPdf = brec.pdf;
        if (o->estimator == 1 && brec.pdf > T(0)) {
            // consistent MIS: the lobe-mixture pdf of DisneyEval (disney.glsl:289-290)
            T mp;
            DisneyEval(st, -r.direction, st.ffnormal, brec.L, mp);
            misPdf = mp;
        }
        if (dot(st.ffnormal, brec.L) < T(0))
            absorption = mk(-std::log(st.mat.extinction.x) / st.mat.atDistance, -std::log(st.mat.extinction.y) / st.mat.atDistance,
                            -std::log(st.mat.extinction.z) / st.mat.atDistance);
        if (brec.pdf > T(0))
            throughput = throughput * ((brec.f * std::fabs(dot(st.ffnormal, brec.L))) / brec.pdf);
        else
            break;
        if (o->enable_rr && depth >= o->rr_depth) {
            T q = std::fmin(std::fmax(throughput.x, std::fmax(throughput.y, throughput.z)) + T(0.001), T(0.95));
            if (rnd<T>(seed) > q) break;
            throughput = throughput / q;
        }
        r.direction = brec.L;
        r.origin = st.fhp + r.direction * C<T>::EPS;
    }
    return radiance;
}

// renderer.glsl:main for one tile pixel
inline float mapf(float value, float low1, float high1, float low2, float high2) {
    return low2 + ((value - low1) * (high2 - low2)) / (high1 - low1);
}

template <class T>
void camera_ray(const orc_options* o, int tileX, int tileY, int lx, int ly, int frame, Ray<T>& ray, uint32_t* seed) {
    float W = (float)o->width, H = (float)o->height;
    float invNumTilesX = 1.0f / (W / (float)o->tile_width);
    float invNumTilesY = 1.0f / (H / (float)o->tile_height);
    float tcx = ((float)lx + 0.5f) / (float)o->tile_width;
    float tcy = ((float)ly + 0.5f) / (float)o->tile_height;
    float xoffset = -1.0f + 2.0f * invNumTilesX * (float)tileX;
    float yoffset = -1.0f + 2.0f * invNumTilesY * (float)tileY;
    float ctx = mapf(tcx, 0.0f, 1.0f, xoffset, xoffset + 2.0f * invNumTilesX);
    float cty = mapf(tcy, 0.0f, 1.0f, yoffset, yoffset + 2.0f * invNumTilesY);
    float fsx = mapf(tcx, 0.0f, 1.0f, invNumTilesX * (float)tileX, invNumTilesX * (float)tileX + invNumTilesX);
    float fsy = mapf(tcy, 0.0f, 1.0f, invNumTilesY * (float)tileY, invNumTilesY * (float)tileY + invNumTilesY);
    float px = fsx * W, py = fsy * H;
    seed[0] = (uint32_t)px;
    seed[1] = (uint32_t)py;
    seed[2] = (uint32_t)frame;
    seed[3] = (uint32_t)px + (uint32_t)py;
    T r1 = T(2) * rnd<T>(seed);
    T r2 = T(2) * rnd<T>(seed);
    T jx = r1 < T(1) ? std::sqrt(r1) - T(1) : T(1) - std::sqrt(T(2) - r1);
    T jy = r2 < T(1) ? std::sqrt(r2) - T(1) : T(1) - std::sqrt(T(2) - r2);
    jx = jx / ((T)W * T(0.5));
    jy = jy / ((T)H * T(0.5));
    T dx = (T)ctx + jx, dy = (T)cty + jy;
    T scale = std::tan((T)o->fov * T(0.5));
    dy = dy * ((T)H / (T)W * scale);
    dx = dx * scale;
    V3<T> right = mk<T>(o->cam_right[0], o->cam_right[1], o->cam_right[2]);
    V3<T> up = mk<T>(o->cam_up[0], o->cam_up[1], o->cam_up[2]);
    V3<T> fwd = mk<T>(o->cam_fwd[0], o->cam_fwd[1], o->cam_fwd[2]);
    V3<T> pos = mk<T>(o->cam_pos[0], o->cam_pos[1], o->cam_pos[2]);
    V3<T> rayDir = normalize(dx * right + dy * up + fwd);
    V3<T> focalPoint = (T)o->focal_dist * rayDir;
    T cam_r1 = rnd<T>(seed) * C<T>::TWO_PI;
    T cam_r2 = rnd<T>(seed) * (T)o->aperture;
    V3<T> ap = (std::cos(cam_r1) * right + std::sin(cam_r1) * up) * std::sqrt(cam_r2);
    V3<T> fin = normalize(focalPoint - ap);
    ray.origin = pos + ap;
    ray.direction = fin;
}

// preview_flareon.glsl:20-60 for pixel (x, y) of a pw x ph preview of the
// o->width x o->height screen
template <class T>
void camera_ray_preview(const orc_options* o, int pw, int ph, int x, int y, int dof, Ray<T>& ray, uint32_t* seed) {
    float W = (float)o->width, H = (float)o->height;
    float tcx = ((float)x + 0.5f) / (float)pw;
    float tcy = ((float)y + 0.5f) / (float)ph;
    const uint32_t ux = (uint32_t)((float)x + 0.5f), uy = (uint32_t)((float)y + 0.5f);
    seed[0] = ux;
    seed[1] = uy;
    seed[2] = 1u;
    seed[3] = ux + uy;
    T r1 = T(2) * rnd<T>(seed);
    T r2 = T(2) * rnd<T>(seed);
    T jx = r1 < T(1) ? std::sqrt(r1) - T(1) : T(1) - std::sqrt(T(2) - r1);
    T jy = r2 < T(1) ? std::sqrt(r2) - T(1) : T(1) - std::sqrt(T(2) - r2);
    jx = jx / (T)W;
    jy = jy / (T)H;
    T dx = (T(2) * (T)tcx - T(1)) + jx, dy = (T(2) * (T)tcy - T(1)) + jy;
    T scale = std::tan((T)o->fov * T(0.5));
    dy = dy * ((T)H / (T)W * scale);
    dx = dx * scale;
    V3<T> right = mk<T>(o->cam_right[0], o->cam_right[1], o->cam_right[2]);
    V3<T> up = mk<T>(o->cam_up[0], o->cam_up[1], o->cam_up[2]);
    V3<T> fwd = mk<T>(o->cam_fwd[0], o->cam_fwd[1], o->cam_fwd[2]);
    V3<T> pos = mk<T>(o->cam_pos[0], o->cam_pos[1], o->cam_pos[2]);
    V3<T> rayDir = normalize(dx * right + dy * up + fwd);
    V3<T> focalPoint = (T)o->focal_dist * rayDir;
    T cam_r1 = rnd<T>(seed) * C<T>::TWO_PI;
    T cam_r2 = rnd<T>(seed) * (T)o->aperture;
    if (dof) {
        V3<T> ap = (std::cos(cam_r1) * right + std::sin(cam_r1) * up) * std::sqrt(cam_r2);
        ray.origin = pos + ap;
        ray.direction = normalize(focalPoint - ap);
    } else {
        ray.origin = pos;
        ray.direction = normalize(focalPoint);
    }
}

template <class T> State<T> frame_state(const orc_scene* s, const float* fr) {
    State<T> st{};
    st.normal = mk<T>(fr[0], fr[1], fr[2]);
    st.ffnormal = st.normal;
    st.tangent = mk<T>(fr[3], fr[4], fr[5]);
    st.bitangent = mk<T>(fr[6], fr[7], fr[8]);
    st.eta = fr[12];
    st.matID = (int)fr[13];
    st.mat = material_of<T>(s, st.matID);
    return st;
}

template <class T> void put3(float* d, V3<T> v) {
    d[0] = (float)v.x;
    d[1] = (float)v.y;
    d[2] = (float)v.z;
}

}  // namespace

// ===========================================================================
// C interface
// ===========================================================================
extern "C" {

void orc_pcg4d(uint32_t* v4, int n) {
    for (int i = 0; i < n; ++i) pcg4d(v4);
}

float orc_rand(uint32_t* v4) { return rnd<float>(v4); }

orc_scene* orc_scene_create(const float* nodes, int64_t num_nodes, int32_t top, const int32_t* vert_indices,
                            int64_t num_tris, const float* vertices_uvx, const float* normals_uvy, int64_t num_vertices,
                            const float* materials, int32_t num_materials, const float* transforms,
                            int32_t num_instances, const float* lights, int32_t num_lights, int32_t hdr_w, int32_t hdr_h,
                            const float* hdr_cols, const float* hdr_marginal, const float* hdr_conditional) {
    orc_scene* s = new orc_scene;
    s->nodes.assign(nodes, nodes + num_nodes * 9);
    s->top = top;
    s->vidx.assign(vert_indices, vert_indices + num_tris * 3);
    s->verts.assign(vertices_uvx, vertices_uvx + num_vertices * 4);
    if (normals_uvy) s->nrms.assign(normals_uvy, normals_uvy + num_vertices * 4);
    else s->nrms.assign(num_vertices * 4, 0.f);
    s->mats.assign(materials, materials + (size_t)num_materials * 28);
    s->nmat = num_materials;
    s->xforms.assign(transforms, transforms + (size_t)num_instances * 16);
    s->ninst = num_instances;
    if (lights && num_lights > 0) s->lights.assign(lights, lights + (size_t)num_lights * 15);
    s->nlights = num_lights;
    s->hw = hdr_w;
    s->hh = hdr_h;
    if (hdr_w > 0 && hdr_h > 0) {
        s->hcols.assign(hdr_cols, hdr_cols + (size_t)hdr_w * hdr_h * 3);
        s->hmarg.assign(hdr_marginal, hdr_marginal + (size_t)hdr_h * 2);
        s->hcond.assign(hdr_conditional, hdr_conditional + (size_t)hdr_w * hdr_h * 2);
    }
    return s;
}

void orc_texture_sample(const orc_scene* s, const double* st, const int32_t* layer, int64_t n, double* out4) {
    for (int64_t i = 0; i < n; ++i) tex_rgba8<double>(s, st[i * 2], st[i * 2 + 1], layer[i], out4 + i * 4);
}

void orc_scene_set_textures(orc_scene* s, const uint8_t* rgba8, int32_t width, int32_t height, int32_t layers) {
    s->tw = width;
    s->th = height;
    s->tl = layers;
    s->tex.assign(rgba8, rgba8 + (size_t)width * height * layers * 4);
}

void orc_scene_destroy(orc_scene* s) { delete s; }

int64_t orc_trace_closest(const orc_scene* s, const float* rays, int64_t n, void* hits, int tie_mode, int32_t* counts,
                          int threads) {
    int64_t ties = 0;
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : ties)
    for (int64_t i = 0; i < n; ++i) {
        const float* q = rays + i * 8;
        Ray<float> r{mk(q[0], q[1], q[2]), mk(q[4], q[5], q[6])};
        State<float> st{};
        st.isEmitter = false;
        st.matID = -1;
        st.inst = -1;
        LightSampleRec<float> lr{};
        Counts c;
        int prim = -1;
        float u = 0.f, v = 0.f;
        float t = closest_hit<float>(s, r, st, lr, false, tie_mode, q[3], q[7], &c, &prim, &u, &v, nullptr);
        int32_t* h = static_cast<int32_t*>(hits) + i * 6;
        if (prim >= 0) {
            h[0] = prim;
            h[1] = st.inst;
            h[2] = st.matID;
        } else {
            h[0] = h[1] = h[2] = -1;
            u = v = 0.f;
        }
        std::memcpy(&h[3], &t, 4);
        std::memcpy(&h[4], &u, 4);
        std::memcpy(&h[5], &v, 4);
        if (counts) {
            counts[i * 3] = c.nint;
            counts[i * 3 + 1] = c.ntri;
            counts[i * 3 + 2] = c.ninst;
        }
        ties += c.ties;
    }
    return ties;
}

void orc_trace_any(const orc_scene* s, const float* rays, int64_t n, uint8_t* occ, int threads) {
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < n; ++i) {
        const float* q = rays + i * 8;
        Ray<float> r{mk(q[0], q[1], q[2]), mk(q[4], q[5], q[6])};
        occ[i] = any_hit<float>(s, r, q[7], false, q[3]) ? 1 : 0;
    }
}

void orc_eval_bsdf(const orc_scene* s, const float* frames, const float* L, int64_t n, float* f, float* pdf,
                   int precision) {
    for (int64_t i = 0; i < n; ++i) {
        const float* fr = frames + i * 16;
        if (precision == 64) {
            State<double> st = frame_state<double>(s, fr);
            double p;
            V3<double> r = DisneyEval(st, mk<double>(fr[9], fr[10], fr[11]), st.ffnormal,
                                      mk<double>(L[i * 3], L[i * 3 + 1], L[i * 3 + 2]), p);
            put3(f + i * 3, r);
            pdf[i] = (float)p;
        } else {
            State<float> st = frame_state<float>(s, fr);
            float p;
            V3<float> r = DisneyEval(st, mk(fr[9], fr[10], fr[11]), st.ffnormal, mk(L[i * 3], L[i * 3 + 1], L[i * 3 + 2]), p);
            put3(f + i * 3, r);
            pdf[i] = p;
        }
    }
}

void orc_sample_bsdf(const orc_scene* s, const float* frames, uint32_t* seeds, int64_t n, float* L, float* f,
                     float* pdf, int precision) {
    for (int64_t i = 0; i < n; ++i) {
        const float* fr = frames + i * 16;
        if (precision == 64) {
            State<double> st = frame_state<double>(s, fr);
            V3<double> l{0, 0, 0};
            double p;
            V3<double> r = DisneySample(st, mk<double>(fr[9], fr[10], fr[11]), st.ffnormal, l, p, seeds + i * 4);
            put3(L + i * 3, l);
            put3(f + i * 3, r);
            pdf[i] = (float)p;
        } else {
            State<float> st = frame_state<float>(s, fr);
            V3<float> l{0, 0, 0};
            float p;
            V3<float> r = DisneySample(st, mk(fr[9], fr[10], fr[11]), st.ffnormal, l, p, seeds + i * 4);
            put3(L + i * 3, l);
            put3(f + i * 3, r);
            pdf[i] = p;
        }
    }
}

void orc_sample_light(const orc_scene* s, const float* pos, uint32_t* seeds, int64_t n, float* dir, float* normal,
                      float* emission, float* dist, float* pdf, int precision) {
    for (int64_t i = 0; i < n; ++i) {
        uint32_t* seed = seeds + i * 4;
        if (precision == 64) {
            int index = (int)(rnd<double>(seed) * (double)s->nlights);
            if (index >= s->nlights) index = s->nlights - 1;
            LightSampleRec<double> r{};
            sampleOneLight(s, light_of<double>(s, index), mk<double>(pos[i * 3], pos[i * 3 + 1], pos[i * 3 + 2]), r, seed);
            put3(dir + i * 3, r.direction);
            put3(normal + i * 3, r.normal);
            put3(emission + i * 3, r.emission);
            dist[i] = (float)r.dist;
            pdf[i] = (float)r.pdf;
        } else {
            int index = (int)(rnd<float>(seed) * (float)s->nlights);
            if (index >= s->nlights) index = s->nlights - 1;
            LightSampleRec<float> r{};
            sampleOneLight(s, light_of<float>(s, index), mk(pos[i * 3], pos[i * 3 + 1], pos[i * 3 + 2]), r, seed);
            put3(dir + i * 3, r.direction);
            put3(normal + i * 3, r.normal);
            put3(emission + i * 3, r.emission);
            dist[i] = r.dist;
            pdf[i] = r.pdf;
        }
    }
}

void orc_sample_env(const orc_scene* s, uint32_t* seeds, int64_t n, float* dir, float* pdf, float* color, float mult,
                    int precision) {
    for (int64_t i = 0; i < n; ++i) {
        if (precision == 64) {
            V3<double> c, d;
            double p;
            EnvSample<double>(s, seeds + i * 4, mult, c, d, p);
            put3(dir + i * 3, d);
            put3(color + i * 3, c);
            pdf[i] = (float)p;
        } else {
            V3<float> c, d;
            float p;
            EnvSample<float>(s, seeds + i * 4, mult, c, d, p);
            put3(dir + i * 3, d);
            put3(color + i * 3, c);
            pdf[i] = p;
        }
    }
}

void orc_env_pdf(const orc_scene* s, const float* dir, int64_t n, float* pdf, int precision) {
    for (int64_t i = 0; i < n; ++i) {
        if (precision == 64) {
            Ray<double> r{mk(0.0, 0.0, 0.0), mk<double>(dir[i * 3], dir[i * 3 + 1], dir[i * 3 + 2])};
            pdf[i] = (float)EnvPdf<double>(s, r);
        } else {
            Ray<float> r{mk(0.f, 0.f, 0.f), mk(dir[i * 3], dir[i * 3 + 1], dir[i * 3 + 2])};
            pdf[i] = EnvPdf<float>(s, r);
        }
    }
}

void orc_path_trace_rays(const orc_scene* s, const orc_options* o, const float* rays, uint32_t* seeds, int64_t n,
                         float* radiance, int precision, int threads) {
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < n; ++i) {
        const float* q = rays + i * 6;
        if (precision == 64) {
            Ray<double> r{mk<double>(q[0], q[1], q[2]), mk<double>(q[3], q[4], q[5])};
            put3(radiance + i * 3, PathTrace<double>(s, o, r, seeds + i * 4));
        } else {
            Ray<float> r{mk(q[0], q[1], q[2]), mk(q[3], q[4], q[5])};
            put3(radiance + i * 3, PathTrace<float>(s, o, r, seeds + i * 4));
        }
    }
}

void orc_camera_rays(const orc_options* o, int32_t tileX, int32_t tileY, int32_t frame, double* rays6, int precision) {
    const int tw = o->tile_width, th = o->tile_height;
    for (int ly = 0; ly < th; ++ly)
        for (int lx = 0; lx < tw; ++lx) {
            uint32_t seed[4];
            double* r = rays6 + ((size_t)ly * tw + lx) * 6;
            if (precision == 64) {
                Ray<double> ray;
                camera_ray<double>(o, tileX, tileY, lx, ly, frame, ray, seed);
                r[0] = ray.origin.x, r[1] = ray.origin.y, r[2] = ray.origin.z;
                r[3] = ray.direction.x, r[4] = ray.direction.y, r[5] = ray.direction.z;
            } else {
                Ray<float> ray;
                camera_ray<float>(o, tileX, tileY, lx, ly, frame, ray, seed);
                r[0] = ray.origin.x, r[1] = ray.origin.y, r[2] = ray.origin.z;
                r[3] = ray.direction.x, r[4] = ray.direction.y, r[5] = ray.direction.z;
            }
        }
}

void orc_render(const orc_scene* s, const orc_options* o, int32_t frame0, int32_t passes, double* accum, int precision,
                int threads) {
    if (threads > 0) omp_set_num_threads(threads);
    const int W = o->width, H = o->height, tw = o->tile_width, th = o->tile_height;
    const int ntx = (W + tw - 1) / tw, nty = (H + th - 1) / th;
    int frame = frame0;
    for (int p = 0; p < passes; ++p) {
        for (int ty = nty - 1; ty >= 0; --ty) {
            for (int tx = 0; tx < ntx; ++tx, ++frame) {
#pragma omp parallel for schedule(dynamic, 16) collapse(2)
                for (int ly = 0; ly < th; ++ly) {
                    for (int lx = 0; lx < tw; ++lx) {
                        int gx = tx * tw + lx, gy = ty * th + ly;
                        if (gx >= W || gy >= H) continue;
                        uint32_t seed[4];
                        double* a = accum + ((size_t)gy * W + gx) * 3;
                        if (precision == 64) {
                            Ray<double> r;
                            camera_ray<double>(o, tx, ty, lx, ly, frame, r, seed);
                            V3<double> c = PathTrace<double>(s, o, r, seed);
                            a[0] += c.x;
                            a[1] += c.y;
                            a[2] += c.z;
                        } else {
                            Ray<float> r;
                            camera_ray<float>(o, tx, ty, lx, ly, frame, r, seed);
                            V3<float> c = PathTrace<float>(s, o, r, seed);
                            a[0] += c.x;
                            a[1] += c.y;
                            a[2] += c.z;
                        }
                    }
                }
            }
        }
    }
}

void orc_preview(const orc_scene* s, const orc_options* o, int32_t pw, int32_t ph, int32_t dof, double* out3,
                 int precision, int threads) {
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 16) collapse(2)
    for (int y = 0; y < ph; ++y)
        for (int x = 0; x < pw; ++x) {
            uint32_t seed[4];
            double* a = out3 + ((size_t)y * pw + x) * 3;
            if (precision == 64) {
                Ray<double> r;
                camera_ray_preview<double>(o, pw, ph, x, y, dof, r, seed);
                V3<double> c = PathTrace<double>(s, o, r, seed);
                a[0] = c.x, a[1] = c.y, a[2] = c.z;
            } else {
                Ray<float> r;
                camera_ray_preview<float>(o, pw, ph, x, y, dof, r, seed);
                V3<float> c = PathTrace<float>(s, o, r, seed);
                a[0] = c.x, a[1] = c.y, a[2] = c.z;
            }
        }
}

void orc_render_pixels(const orc_scene* s, const orc_options* o, const int32_t* pixels, int64_t npix, int32_t frame0,
                       int32_t spp, double* sum, double* sumsq, int precision, int threads) {
    if (threads > 0) omp_set_num_threads(threads);
    const int W = o->width;
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t i = 0; i < npix; ++i) {
        int gx = pixels[i] % W, gy = pixels[i] / W;
        double acc[3] = {0, 0, 0}, acc2[3] = {0, 0, 0};
        for (int k = 0; k < spp; ++k) {
            uint32_t seed[4];
            double c[3];
            if (precision == 64) {
                Ray<double> r;
                camera_ray<double>(o, 0, 0, gx, gy, frame0 + k, r, seed);
                V3<double> v = PathTrace<double>(s, o, r, seed);
                c[0] = v.x; c[1] = v.y; c[2] = v.z;
            } else {
                Ray<float> r;
                camera_ray<float>(o, 0, 0, gx, gy, frame0 + k, r, seed);
                V3<float> v = PathTrace<float>(s, o, r, seed);
                c[0] = v.x; c[1] = v.y; c[2] = v.z;
            }
            for (int j = 0; j < 3; ++j) {
                acc[j] += c[j];
                acc2[j] += c[j] * c[j];
            }
        }
        for (int j = 0; j < 3; ++j) {
            sum[i * 3 + j] = acc[j];
            sumsq[i * 3 + j] = acc2[j];
        }
    }
}

}  // extern "C"
