// Shading on gfx950: RNG, camera rays, Disney BSDF, light and environment
// sampling — the arithmetic of the reference GLSL evaluated in its order
// (each function cites the GLSL it implements).
#pragma once
#include "dev_trace.h"
#include "pt_types.h"

namespace lfhd {

// ---- RNG (globals.glsl:116-133) -----------------------------------------------
LFH_DEV void pcg4d(uint4& v) {
    v.x = v.x * 1664525u + 1013904223u;
    v.y = v.y * 1664525u + 1013904223u;
    v.z = v.z * 1664525u + 1013904223u;
    v.w = v.w * 1664525u + 1013904223u;
    v.x += v.y * v.w;
    v.y += v.z * v.x;
    v.z += v.x * v.y;
    v.w += v.y * v.z;
    v.x ^= v.x >> 16u;
    v.y ^= v.y >> 16u;
    v.z ^= v.z >> 16u;
    v.w ^= v.w >> 16u;
    v.x += v.y * v.w;
    v.y += v.z * v.x;
    v.z += v.x * v.y;
    v.w += v.y * v.z;
}
// float(seed.x) / float(0xffffffffu): u32 -> f32 round-to-nearest, then an
// IEEE division by 2^32 (0xffffffff rounds to 2^32); 1.0 is reachable.
LFH_DEV float rnd(uint4& s) {
    pcg4d(s);
    return __uint2float_rn(s.x) / 4294967296.0f;
}

// ---- small helpers ---------------------------------------------------------------
LFH_DEV float mixs(float x, float y, float a) { return x * (1.0f - a) + y * a; }
LFH_DEV V3 mixv(V3 x, V3 y, float a) { return x * (1.0f - a) + y * a; }
LFH_DEV float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
LFH_DEV V3 reflect(V3 I, V3 N) { return I - (2.0f * dot(N, I)) * N; }
LFH_DEV V3 refract(V3 I, V3 N, float eta) {
    float d = dot(N, I);
    float k = 1.0f - eta * eta * (1.0f - d * d);
    if (k < 0.0f) return v3(0.f, 0.f, 0.f);
    return eta * I - (eta * d + sqrtf(k)) * N;
}
LFH_DEV V3 ld3(const float4& f) { return v3(f.x, f.y, f.z); }

// ---- material (pathtrace.glsl:43-77) ------------------------------------------------
struct Mat {
    V3 albedo;
    float specular;
    V3 emission;
    float metallic, roughness, subsurface, specularTint, sheen, sheenTint, clearcoat, clearcoatRoughness,
        specTrans, ior, atDistance;
    V3 extinction;
};

LFH_DEV Mat load_material(const float4* mats, int id) {
    const float4* p = mats + (size_t)id * kMaterialF4;
    float4 a = p[0], b = p[1], c = p[2], d = p[3], e = p[4], f = p[5];
    Mat m;
    m.albedo = v3(a.x, a.y, a.z);
    m.specular = a.w;
    m.emission = v3(b.x, b.y, b.z);
    m.metallic = c.x;
    m.roughness = fmaxf(c.y, 0.001f);
    m.subsurface = c.z;
    m.specularTint = c.w;
    m.sheen = d.x;
    m.sheenTint = d.y;
    m.clearcoat = d.z;
    m.clearcoatRoughness = d.w;
    m.specTrans = e.x;
    m.ior = e.y;
    m.atDistance = e.z;
    m.extinction = v3(f.x, f.y, f.z);
    return m;
}

// Onb (pathtrace.glsl:7-13): tangent frame around N.
LFH_DEV void onb(V3 N, V3& T, V3& B) {
    V3 up = fabsf(N.z) < 0.999f ? v3(0.f, 0.f, 1.f) : v3(1.f, 0.f, 0.f);
    T = normalize(cross(up, N));
    B = cross(N, T);
}

struct Frame {
    V3 N;        // ffnormal (the BSDF's N)
    V3 T, B;     // tangent frame built from state.normal
    float eta;
};

// ---- sampling.glsl helpers ------------------------------------------------------
LFH_DEV V3 ImportanceSampleGTR1(float rgh, float r1, float r2) {
    float a = fmaxf(0.001f, rgh);
    float a2 = a * a;
    float phi = r1 * kTwoPi;
    float cosTheta = sqrtf((1.0f - powf(a2, 1.0f - r1)) / (1.0f - a2));
    float sinTheta = clampf(sqrtf(1.0f - (cosTheta * cosTheta)), 0.0f, 1.0f);
    float sinPhi = sinf(phi), cosPhi = cosf(phi);
    return v3(sinTheta * cosPhi, sinTheta * sinPhi, cosTheta);
}
LFH_DEV V3 ImportanceSampleGTR2(float rgh, float r1, float r2) {
    float a = fmaxf(0.001f, rgh);
    float phi = r1 * kTwoPi;
    float cosTheta = sqrtf((1.0f - r2) / (1.0f + (a * a - 1.0f) * r2));
    float sinTheta = clampf(sqrtf(1.0f - (cosTheta * cosTheta)), 0.0f, 1.0f);
    float sinPhi = sinf(phi), cosPhi = cosf(phi);
    return v3(sinTheta * cosPhi, sinTheta * sinPhi, cosTheta);
}
LFH_DEV float SchlickFresnel(float u) {
    float m = clampf(1.0f - u, 0.0f, 1.0f);
    float m2 = m * m;
    return m2 * m2 * m;
}
LFH_DEV float DielectricFresnel(float cos_theta_i, float eta) {
    float sinThetaTSq = eta * eta * (1.0f - cos_theta_i * cos_theta_i);
    if (sinThetaTSq > 1.0f) return 1.0f;
    float cos_theta_t = sqrtf(fmaxf(1.0f - sinThetaTSq, 0.0f));
    float rs = (eta * cos_theta_t - cos_theta_i) / (eta * cos_theta_t + cos_theta_i);
    float rp = (eta * cos_theta_i - cos_theta_t) / (eta * cos_theta_i + cos_theta_t);
    return 0.5f * (rs * rs + rp * rp);
}
LFH_DEV float GTR1(float NDotH, float a) {
    if (a >= 1.0f) return (1.0f / kPi);
    float a2 = a * a;
    float t = 1.0f + (a2 - 1.0f) * NDotH * NDotH;
    return (a2 - 1.0f) / (kPi * logf(a2) * t);
}
LFH_DEV float GTR2(float NDotH, float a) {
    float a2 = a * a;
    float t = 1.0f + (a2 - 1.0f) * NDotH * NDotH;
    return a2 / (kPi * t * t);
}
LFH_DEV float SmithG_GGX(float NDotV, float alphaG) {
    float a = alphaG * alphaG;
    float b = NDotV * NDotV;
    return 1.0f / (NDotV + sqrtf(a + b - a * b));
}
LFH_DEV V3 CosineSampleHemisphere(float r1, float r2) {
    float r = sqrtf(r1);
    float phi = kTwoPi * r2;
    float x = r * cosf(phi), y = r * sinf(phi);
    return v3(x, y, sqrtf(fmaxf(0.0f, 1.0f - x * x - y * y)));
}
LFH_DEV V3 UniformSampleSphere(float r1, float r2) {
    float z = 1.0f - 2.0f * r1;
    float r = sqrtf(fmaxf(0.0f, 1.0f - z * z));
    float phi = kTwoPi * r2;
    return v3(r * cosf(phi), r * sinf(phi), z);
}
LFH_DEV float powerHeuristic(float a, float b) {
    float t = a * a;
    return t / (b * b + t);
}

// ---- Disney BSDF (disney.glsl:18-291) ----------------------------------------------
LFH_DEV V3 EvalDielectricReflection(const Mat& m, float eta, V3 V, V3 N, V3 L, V3 H, float& pdf) {
    pdf = 0.0f;
    if (dot(N, L) <= 0.0f) return v3(0.f, 0.f, 0.f);
    float F = DielectricFresnel(dot(V, H), eta);
    float D = GTR2(dot(N, H), m.roughness);
    pdf = D * dot(N, H) * F / (4.0f * fabsf(dot(V, H)));
    float G = SmithG_GGX(fabsf(dot(N, L)), m.roughness) * SmithG_GGX(fabsf(dot(N, V)), m.roughness);
    return m.albedo * F * D * G;
}
LFH_DEV V3 EvalDielectricRefraction(const Mat& m, float eta, V3 V, V3 N, V3 L, V3 H, float& pdf) {
    pdf = 0.0f;
    if (dot(N, L) >= 0.0f) return v3(0.f, 0.f, 0.f);
    float F = DielectricFresnel(fabsf(dot(V, H)), eta);
    float D = GTR2(dot(N, H), m.roughness);
    float denomSqrt = dot(L, H) + dot(V, H) * eta;
    pdf = D * dot(N, H) * (1.0f - F) * fabsf(dot(L, H)) / (denomSqrt * denomSqrt);
    float G = SmithG_GGX(fabsf(dot(N, L)), m.roughness) * SmithG_GGX(fabsf(dot(N, V)), m.roughness);
    return m.albedo * (1.0f - F) * D * G * fabsf(dot(V, H)) * fabsf(dot(L, H)) * 4.0f * eta * eta /
           (denomSqrt * denomSqrt);
}
LFH_DEV V3 EvalSpecular(const Mat& m, V3 Cspec0, V3 V, V3 N, V3 L, V3 H, float& pdf) {
    pdf = 0.0f;
    if (dot(N, L) <= 0.0f) return v3(0.f, 0.f, 0.f);
    float D = GTR2(dot(N, H), m.roughness);
    pdf = D * dot(N, H) / (4.0f * dot(V, H));
    float FH = SchlickFresnel(dot(L, H));
    V3 F = mixv(Cspec0, v3(1.f, 1.f, 1.f), FH);
    float G = SmithG_GGX(fabsf(dot(N, L)), m.roughness) * SmithG_GGX(fabsf(dot(N, V)), m.roughness);
    return F * D * G;
}
LFH_DEV V3 EvalClearcoat(const Mat& m, V3 V, V3 N, V3 L, V3 H, float& pdf) {
    pdf = 0.0f;
    if (dot(N, L) <= 0.0f) return v3(0.f, 0.f, 0.f);
    float D = GTR1(dot(N, H), mixs(0.1f, 0.001f, m.clearcoatRoughness));
    pdf = D * dot(N, H) / (4.0f * dot(V, H));
    float FH = SchlickFresnel(dot(L, H));
    float F = mixs(0.04f, 1.0f, FH);
    float G = SmithG_GGX(dot(N, L), 0.25f) * SmithG_GGX(dot(N, V), 0.25f);
    float c = 0.25f * m.clearcoat * F * D * G;
    return v3(c, c, c);
}
// LavaFrame's diffuse: (1/pi) (ss + subsurface) albedo + sheen, times (1 - metallic)
LFH_DEV V3 EvalDiffuse(const Mat& m, V3 Csheen, V3 V, V3 N, V3 L, V3 H, float& pdf) {
    pdf = 0.0f;
    if (dot(N, L) <= 0.0f) return v3(0.f, 0.f, 0.f);
    pdf = dot(N, L) * (1.0f / kPi);
    float FL = SchlickFresnel(dot(N, L));
    float FV = SchlickFresnel(dot(N, V));
    float FH = SchlickFresnel(dot(L, H));
    float Fss90 = dot(L, H) * dot(L, H) * m.roughness;
    float Fss = mixs(1.0f, Fss90, FL) * mixs(1.0f, Fss90, FV);
    float ss = 1.0f * (Fss * (1.0f / (dot(N, L) + dot(N, V)) - 0.5f) + 0.5f);
    V3 Fsheen = (FH * m.sheen) * Csheen;
    return ((1.0f / kPi) * (ss + m.subsurface) * m.albedo + Fsheen) * (1.0f - m.metallic);
}
LFH_DEV void tints(const Mat& m, V3& Cspec0, V3& Csheen) {
    V3 Cdlin = m.albedo;
    float Cdlum = 0.3f * Cdlin.x + 0.6f * Cdlin.y + 0.1f * Cdlin.z;
    V3 Ctint = Cdlum > 0.0f ? Cdlin / Cdlum : v3(1.f, 1.f, 1.f);
    V3 one = v3(1.f, 1.f, 1.f);
    Cspec0 = mixv((m.specular * 0.08f) * mixv(one, Ctint, m.specularTint), Cdlin, m.metallic);
    Csheen = mixv(one, Ctint, m.sheenTint);
}

// DisneySample (disney.glsl:128-225)
LFH_DEV V3 DisneySample(const Mat& m, const Frame& fr, V3 V, V3& L, float& pdf, uint4& rng) {
    const V3 N = fr.N;
    pdf = 0.0f;
    V3 f = v3(0.f, 0.f, 0.f);
    float r1 = rnd(rng), r2 = rnd(rng);
    float diffuseRatio = 0.5f * (1.0f - m.metallic);
    float transWeight = (1.0f - m.metallic) * m.specTrans;
    V3 Cspec0, Csheen;
    tints(m, Cspec0, Csheen);
    if (rnd(rng) < transWeight) {
        V3 H = ImportanceSampleGTR2(m.roughness, r1, r2);
        H = fr.T * H.x + fr.B * H.y + N * H.z;
        if (dot(V, H) < 0.0f) H = -H;
        V3 R = reflect(-V, H);
        float F = DielectricFresnel(fabsf(dot(R, H)), fr.eta);
        if (rnd(rng) < F) {
            L = normalize(R);
            f = EvalDielectricReflection(m, fr.eta, V, N, L, H, pdf);
        } else {
            L = normalize(refract(-V, H, fr.eta));
            f = EvalDielectricRefraction(m, fr.eta, V, N, L, H, pdf);
        }
        f = f * transWeight;
        pdf *= transWeight;
    } else {
        if (rnd(rng) < diffuseRatio) {
            L = CosineSampleHemisphere(r1, r2);
            L = fr.T * L.x + fr.B * L.y + N * L.z;
            V3 H = normalize(L + V);
            f = EvalDiffuse(m, Csheen, V, N, L, H, pdf);
            pdf *= diffuseRatio;
        } else {
            float primarySpecRatio = 1.0f / (1.0f + m.clearcoat);
            if (rnd(rng) < primarySpecRatio) {
                V3 H = ImportanceSampleGTR2(m.roughness, r1, r2);
                H = fr.T * H.x + fr.B * H.y + N * H.z;
                if (dot(V, H) < 0.0f) H = -H;
                L = normalize(reflect(-V, H));
                f = EvalSpecular(m, Cspec0, V, N, L, H, pdf);
                pdf *= primarySpecRatio * (1.0f - diffuseRatio);
            } else {
                V3 H = ImportanceSampleGTR1(mixs(0.1f, 0.001f, m.clearcoatRoughness), r1, r2);
                H = fr.T * H.x + fr.B * H.y + N * H.z;
                if (dot(V, H) < 0.0f) H = -H;
                L = normalize(reflect(-V, H));
                f = EvalClearcoat(m, V, N, L, H, pdf);
                pdf *= (1.0f - primarySpecRatio) * (1.0f - diffuseRatio);
            }
        }
        f = f * (1.0f - transWeight);
        pdf *= (1.0f - transWeight);
    }
    return f;
}

// DisneyEval (disney.glsl:228-291): mixture f and pdf
LFH_DEV V3 DisneyEval(const Mat& m, float eta, V3 V, V3 N, V3 L, float& pdf) {
    V3 H;
    bool refl = dot(N, L) > 0.0f;
    if (refl) H = normalize(L + V);
    else H = normalize(L + V * eta);
    if (dot(V, H) < 0.0f) H = -H;
    float diffuseRatio = 0.5f * (1.0f - m.metallic);
    float primarySpecRatio = 1.0f / (1.0f + m.clearcoat);
    float transWeight = (1.0f - m.metallic) * m.specTrans;
    V3 brdf = v3(0.f, 0.f, 0.f), bsdf = v3(0.f, 0.f, 0.f);
    float brdfPdf = 0.0f, bsdfPdf = 0.0f;
    if (transWeight > 0.0f) {
        if (refl) bsdf = EvalDielectricReflection(m, eta, V, N, L, H, bsdfPdf);
        else bsdf = EvalDielectricRefraction(m, eta, V, N, L, H, bsdfPdf);
    }
    float m_pdf;
    if (transWeight < 1.0f) {
        V3 Cspec0, Csheen;
        tints(m, Cspec0, Csheen);
        brdf = brdf + EvalDiffuse(m, Csheen, V, N, L, H, m_pdf);
        brdfPdf += m_pdf * diffuseRatio;
        brdf = brdf + EvalSpecular(m, Cspec0, V, N, L, H, m_pdf);
        brdfPdf += m_pdf * primarySpecRatio * (1.0f - diffuseRatio);
        brdf = brdf + EvalClearcoat(m, V, N, L, H, m_pdf);
        brdfPdf += m_pdf * (1.0f - primarySpecRatio) * (1.0f - diffuseRatio);
    }
    pdf = mixs(brdfPdf, bsdfPdf, transWeight);
    return mixv(brdf, bsdf, transWeight);
}

// ---- lights (sampling.glsl:172-230, intersection.glsl:7-50) -----------------------------
struct LightRec {
    V3 normal, emission, direction;
    float dist, pdf;
};

LFH_DEV void sampleOneLight(const ShadeScene& sc, int index, V3 surfacePos, LightRec& rec, uint4& rng) {
    const float4* lp = sc.lights + (size_t)index * kLightF4;
    float4 a = lp[0], b = lp[1], c = lp[2], d = lp[3];
    V3 pos = ld3(a), em = ld3(b), u = ld3(c), v = ld3(d);
    float type = a.w, area = b.w, radius = c.w;
    int t = (int)type;
    float nl = (float)sc.num_lights;
    if (t == 0) {
        float r1 = rnd(rng), r2 = rnd(rng);
        V3 lsp = pos + u * r1 + v * r2;
        rec.direction = lsp - surfacePos;
        rec.dist = length(rec.direction);
        float distSq = rec.dist * rec.dist;
        rec.direction = rec.direction / rec.dist;
        rec.normal = normalize(cross(u, v));
        rec.emission = em * nl;
        rec.pdf = distSq / (area * fabsf(dot(rec.normal, rec.direction)));
    } else if (t == 1) {
        float r1 = rnd(rng), r2 = rnd(rng);
        V3 lsp = pos + UniformSampleSphere(r1, r2) * radius;
        rec.direction = lsp - surfacePos;
        rec.dist = length(rec.direction);
        float distSq = rec.dist * rec.dist;
        rec.direction = rec.direction / rec.dist;
        rec.normal = normalize(lsp - pos);
        rec.emission = em * nl;
        rec.pdf = distSq / (area * fabsf(dot(rec.normal, rec.direction)));
    } else {
        rec.direction = normalize(pos - v3(0.f, 0.f, 0.f));
        rec.normal = normalize(surfacePos - pos);
        rec.emission = em * nl;
        rec.dist = kInfinity;
        rec.pdf = 1.0f;
    }
}

// RectIntersect with the per-light derived plane / scaled edges.
LFH_DEV float rect_hit(const float4* lp, V3 o, V3 d) {
    float4 n4 = lp[4], u4 = lp[5], v4 = lp[6], p4 = lp[0];
    V3 n = ld3(n4);
    float dt = dot(d, n);
    float t = (n4.w - dot(n, o)) / dt;
    if (t > kEps) {
        V3 p = o + d * t;
        V3 vi = p - ld3(p4);
        float a1 = dot(ld3(u4), vi);
        if (a1 >= 0.0f && a1 <= 1.0f) {
            float a2 = dot(ld3(v4), vi);
            if (a2 >= 0.0f && a2 <= 1.0f) return t;
        }
    }
    return kInfinity;
}
LFH_DEV float sphere_hit(float rad, V3 pos, V3 o, V3 d) {
    V3 op = pos - o;
    float eps = 0.001f;
    float b = dot(op, d);
    float det = b * b - dot(op, op) + rad * rad;
    if (det < 0.0f) return kInfinity;
    det = sqrtf(det);
    float t1 = b - det;
    if (t1 > eps) return t1;
    float t2 = b + det;
    if (t2 > eps) return t2;
    return kInfinity;
}

// ---- environment (sampling.glsl:236-265) ------------------------------------------
LFH_DEV int wrapi(int i, int n) {
    int r = i % n;
    return r < 0 ? r + n : r;
}
// NEAREST lookup of an RG32F table with GL_REPEAT.
LFH_DEV float2 tex_nearest(const float2* tab, int W, int H, float s, float t) {
    float u = s * (float)W, v = t * (float)H;
    int i = wrapi((int)floorf(u), W), j = wrapi((int)floorf(v), H);
    return tab[(size_t)j * W + i];
}
// LINEAR lookup of the RGB32F radiance with GL_REPEAT.
LFH_DEV V3 tex_linear(const float4* cols, int W, int H, float s, float t) {
    float u = s * (float)W - 0.5f, v = t * (float)H - 0.5f;
    float fu = floorf(u), fv = floorf(v);
    float a = u - fu, b = v - fv;
    int i0 = wrapi((int)fu, W), i1 = wrapi((int)fu + 1, W);
    int j0 = wrapi((int)fv, H), j1 = wrapi((int)fv + 1, H);
    V3 c00 = ld3(cols[(size_t)j0 * W + i0]), c10 = ld3(cols[(size_t)j0 * W + i1]);
    V3 c01 = ld3(cols[(size_t)j1 * W + i0]), c11 = ld3(cols[(size_t)j1 * W + i1]);
    V3 top = c00 * (1.0f - a) + c10 * a;
    V3 bot = c01 * (1.0f - a) + c11 * a;
    return top * (1.0f - b) + bot * b;
}
// ---- texture maps (GetMaterialsAndTextures, pathtrace.glsl:78-113) -------------------
// GL_LINEAR read of RGBA8 layer `layer` (clamped to the array, as GL clamps the
// layer coordinate) with GL_REPEAT wrapping, channels normalized to [0, 1].
LFH_DEV float4 tex_rgba8(const ShadeScene& sc, float s, float t, int layer) {
    const int W = sc.tex_w, H = sc.tex_h;
    layer = layer < 0 ? 0 : (layer >= sc.tex_layers ? sc.tex_layers - 1 : layer);
    float u = s * (float)W - 0.5f, v = t * (float)H - 0.5f;
    float fu = floorf(u), fv = floorf(v);
    float a = u - fu, b = v - fv;
    // the integer texel index modulo the size, in float first (texture
    // coordinates far outside [0, 1] repeat)
    fu -= (float)W * floorf(fu / (float)W);
    fv -= (float)H * floorf(fv / (float)H);
    int i0 = wrapi((int)fu, W), i1 = wrapi((int)fu + 1, W);
    int j0 = wrapi((int)fv, H), j1 = wrapi((int)fv + 1, H);
    const uint32_t* L = sc.tex + (size_t)layer * W * H;
    auto px = [&](int i, int j) {
        const uint32_t c = L[(size_t)j * W + i];
        const float k = 1.0f / 255.0f;
        return make_float4((float)(c & 0xFFu) * k, (float)((c >> 8) & 0xFFu) * k, (float)((c >> 16) & 0xFFu) * k,
                           (float)(c >> 24) * k);
    };
    const float4 c00 = px(i0, j0), c10 = px(i1, j0), c01 = px(i0, j1), c11 = px(i1, j1);
    const float ia = 1.0f - a, ib = 1.0f - b;
    const float4 top = make_float4(c00.x * ia + c10.x * a, c00.y * ia + c10.y * a, c00.z * ia + c10.z * a,
                                   c00.w * ia + c10.w * a);
    const float4 bot = make_float4(c01.x * ia + c11.x * a, c01.y * ia + c11.y * a, c01.z * ia + c11.z * a,
                                   c01.w * ia + c11.w * a);
    return make_float4(top.x * ib + bot.x * b, top.y * ib + bot.y * b, top.z * ib + bot.z * b,
                       top.w * ib + bot.w * b);
}

LFH_DEV V3 pow22(float4 c) { return v3(powf(c.x, 2.2f), powf(c.y, 2.2f), powf(c.z, 2.2f)); }

// The maps of material matId at texture coordinate tc (texUV.y = 1 - tc.y):
// albedo *= map^2.2, metallic / roughness from the R / G channels (roughness
// squared), the tangent-space normal map in the shading normal's frame, and
// emission = map^2.2.
LFH_DEV void apply_textures(const ShadeScene& sc, int matId, float2 tc, Mat& m, V3& nrm) {
    if (!sc.tex) return;
    const float4 ids = sc.materials[(size_t)matId * kMaterialF4 + 6];
    const float s = tc.x, t = 1.0f - tc.y;
    if ((int)ids.x >= 0) m.albedo = m.albedo * pow22(tex_rgba8(sc, s, t, (int)ids.x));
    if ((int)ids.y >= 0) {
        const float4 c = tex_rgba8(sc, s, t, (int)ids.y);
        m.metallic = c.x;
        m.roughness = fmaxf(c.y * c.y, 0.001f);
    }
    if ((int)ids.z >= 0) {
        const float4 c = tex_rgba8(sc, s, t, (int)ids.z);
        const V3 n = normalize(v3(c.x * 2.0f - 1.0f, c.y * 2.0f - 1.0f, c.z * 2.0f - 1.0f));
        V3 T, B;
        onb(nrm, T, B);
        nrm = normalize(T * n.x + B * n.y + nrm * n.z);
    }
    if (ids.w >= 0.0f) m.emission = pow22(tex_rgba8(sc, s, t, (int)ids.w));
}

// Emission of material matId at texture coordinate tc (the emitter-hit case,
// where only the emission of the stale material is used).
LFH_DEV V3 textured_emission(const ShadeScene& sc, int matId, float2 tc, V3 emission) {
    if (!sc.tex) return emission;
    const float4 ids = sc.materials[(size_t)matId * kMaterialF4 + 6];
    if (ids.w >= 0.0f) return pow22(tex_rgba8(sc, tc.x, 1.0f - tc.y, (int)ids.w));
    return emission;
}

LFH_DEV float EnvPdf(const ShadeScene& sc, V3 d) {
    float theta = acosf(clampf(d.y, -1.0f, 1.0f));
    float u = (kPi + atan2f(d.z, d.x)) * (1.0f / kTwoPi);
    float v = theta * (1.0f / kPi);
    float pdf = tex_nearest(sc.hdr_cond, sc.hdr_w, sc.hdr_h, u, v).y *
                tex_nearest(sc.hdr_marg, sc.hdr_h, 1, v, 0.0f).y;
    return (pdf * sc.hdr_res) / (2.0f * kPi * kPi * sinf(theta));
}
LFH_DEV void EnvSample(const ShadeScene& sc, uint4& rng, V3& color, V3& dir, float& pdfOut) {
    float r1 = rnd(rng), r2 = rnd(rng);
    float v = tex_nearest(sc.hdr_marg, sc.hdr_h, 1, r1, 0.0f).x;
    float u = tex_nearest(sc.hdr_cond, sc.hdr_w, sc.hdr_h, r2, v).x;
    color = tex_linear(sc.hdr_cols, sc.hdr_w, sc.hdr_h, u, v) * sc.hdr_mult;
    float pdf = tex_nearest(sc.hdr_cond, sc.hdr_w, sc.hdr_h, u, v).y *
                tex_nearest(sc.hdr_marg, sc.hdr_h, 1, v, 0.0f).y;
    float phi = u * kTwoPi;
    float theta = v * kPi;
    if (sinf(theta) == 0.0f) pdf = 0.0f;
    dir = v3(-sinf(theta) * cosf(phi), cosf(theta), -sinf(theta) * sinf(phi));
    pdfOut = (pdf * sc.hdr_res) / (2.0f * kPi * kPi * sinf(theta));
}

// ---- camera ray for one tile pixel (renderer.glsl:25-62) ----------------------------


LFH_DEV float mapf(float value, float low1, float high1, float low2, float high2) {
    return low2 + ((value - low1) * (high2 - low2)) / (high1 - low1);
}

LFH_DEV void camera_ray(const CamParams& c, int lx, int ly, V3& org, V3& dir, uint4& seed) {
    float W = (float)c.width, H = (float)c.height;
    float invNumTilesX = 1.0f / (W / (float)c.tile_w);
    float invNumTilesY = 1.0f / (H / (float)c.tile_h);
    float tcx = ((float)lx + 0.5f) / (float)c.tile_w;
    float tcy = ((float)ly + 0.5f) / (float)c.tile_h;
    float xoffset = -1.0f + 2.0f * invNumTilesX * (float)c.tileX;
    float yoffset = -1.0f + 2.0f * invNumTilesY * (float)c.tileY;
    float ctx = mapf(tcx, 0.0f, 1.0f, xoffset, xoffset + 2.0f * invNumTilesX);
    float cty = mapf(tcy, 0.0f, 1.0f, yoffset, yoffset + 2.0f * invNumTilesY);
    float fsx = mapf(tcx, 0.0f, 1.0f, invNumTilesX * (float)c.tileX, invNumTilesX * (float)c.tileX + invNumTilesX);
    float fsy = mapf(tcy, 0.0f, 1.0f, invNumTilesY * (float)c.tileY, invNumTilesY * (float)c.tileY + invNumTilesY);
    float px = fsx * W, py = fsy * H;
    uint32_t ux = (uint32_t)px, uy = (uint32_t)py;
    seed = make_uint4(ux, uy, (uint32_t)c.frame, ux + uy);
    float r1 = 2.0f * rnd(seed);
    float r2 = 2.0f * rnd(seed);
    float jx = r1 < 1.0f ? sqrtf(r1) - 1.0f : 1.0f - sqrtf(2.0f - r1);
    float jy = r2 < 1.0f ? sqrtf(r2) - 1.0f : 1.0f - sqrtf(2.0f - r2);
    jx = jx / (W * 0.5f);
    jy = jy / (H * 0.5f);
    float dx = ctx + jx, dy = cty + jy;
    float scale = tanf(c.fov * 0.5f);
    dy = dy * (H / W * scale);
    dx = dx * scale;
    V3 right = v3(c.right[0], c.right[1], c.right[2]);
    V3 up = v3(c.up[0], c.up[1], c.up[2]);
    V3 fwd = v3(c.fwd[0], c.fwd[1], c.fwd[2]);
    V3 pos = v3(c.pos[0], c.pos[1], c.pos[2]);
    V3 rayDir = normalize(dx * right + dy * up + fwd);
    V3 focalPoint = c.focal_dist * rayDir;
    float cam_r1 = rnd(seed) * kTwoPi;
    float cam_r2 = rnd(seed) * c.aperture;
    V3 ap = (cosf(cam_r1) * right + sinf(cam_r1) * up) * sqrtf(cam_r2);
    dir = normalize(focalPoint - ap);
    org = pos + ap;
}

// Camera ray of the preview engine (preview_flareon.glsl:20-60) for pixel
// (x, y) of a tile_w x tile_h preview of a width x height screen: the seed is
// the preview pixel with frame 1, the jitter is divided by the screen size
// (not half of it), the lens samples are drawn but used only with USE_DOF.
LFH_DEV void camera_ray_preview(const CamParams& c, int x, int y, V3& org, V3& dir, uint4& seed) {
    float W = (float)c.width, H = (float)c.height;
    float tcx = ((float)x + 0.5f) / (float)c.tile_w;
    float tcy = ((float)y + 0.5f) / (float)c.tile_h;
    uint32_t ux = (uint32_t)((float)x + 0.5f), uy = (uint32_t)((float)y + 0.5f);
    seed = make_uint4(ux, uy, 1u, ux + uy);
    float r1 = 2.0f * rnd(seed);
    float r2 = 2.0f * rnd(seed);
    float jx = r1 < 1.0f ? sqrtf(r1) - 1.0f : 1.0f - sqrtf(2.0f - r1);
    float jy = r2 < 1.0f ? sqrtf(r2) - 1.0f : 1.0f - sqrtf(2.0f - r2);
    jx = jx / W;
    jy = jy / H;
    float dx = (2.0f * tcx - 1.0f) + jx, dy = (2.0f * tcy - 1.0f) + jy;
    float scale = tanf(c.fov * 0.5f);
    dy = dy * (H / W * scale);
    dx = dx * scale;
    V3 right = v3(c.right[0], c.right[1], c.right[2]);
    V3 up = v3(c.up[0], c.up[1], c.up[2]);
    V3 fwd = v3(c.fwd[0], c.fwd[1], c.fwd[2]);
    V3 pos = v3(c.pos[0], c.pos[1], c.pos[2]);
    V3 rayDir = normalize(dx * right + dy * up + fwd);
    V3 focalPoint = c.focal_dist * rayDir;
    float cam_r1 = rnd(seed) * kTwoPi;
    float cam_r2 = rnd(seed) * c.aperture;
    if (c.dof) {
        V3 ap = (cosf(cam_r1) * right + sinf(cam_r1) * up) * sqrtf(cam_r2);
        dir = normalize(focalPoint - ap);
        org = pos + ap;
    } else {
        dir = normalize(focalPoint);
        org = pos;
    }
}

}  // namespace lfhd
