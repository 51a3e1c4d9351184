// Learned path guide: device model shared by the MLP kernels and the guided
// shading stage.  The guide has no counterpart in the reference renderer
// (LavaFrame samples the BSDF only, pathtrace.glsl:268-291); it adds a second
// sampling strategy, one-sample-MIS mixed with the BSDF.
//
//   encoding  64 features of (x = scene-normalised position in [0,1]^3, n, wo):
//             [0,36)  sin/cos(2^f pi x_a), f = 0..5     index (a*6+f)*2 + {sin,cos}
//             [36,48) sin/cos(2^f pi n_a), f = 0..1     36 + (a*2+f)*2 + {sin,cos}
//             [48,60) sin/cos(2^f pi wo_a), f = 0..1    48 + (a*2+f)*2 + {sin,cos}
//             [60,63) wo, 63: 1
//   MLP       64 -> 64 -> 64 -> 64 -> 40, ReLU after the hidden layers,
//             y = x W + b (W [in][out]); BF16 operands, FP32 accumulation
//   lobes     k = 0..7 from outputs 5k..5k+4: mu = normalize(r), kappa =
//             exp(clamp(s, -3, 7)), pi = softmax(a)
//   pdf       sum_k pi_k kappa_k / (2 pi (1 - e^{-2 kappa_k})) e^{kappa_k (mu_k.d - 1)}
#pragma once
#include "dev_common.h"

namespace lfhd {

constexpr int kGuideIn = 64;
constexpr int kGuideHid = 64;
constexpr int kGuideLobes = 8;
constexpr int kGuideOut = 40;
constexpr int kGuideLayers = 4;
// flat FP32 parameter vector: per layer W [in][out] row-major, then b [out]
constexpr int kGuideParams = 3 * (64 * 64 + 64) + 64 * 40 + 40;  // 15080
constexpr float kLogKMin = -3.0f;
constexpr float kLogKMax = 7.0f;
constexpr float kGuideFrac = 0.5f;   // probability of drawing from the guide
constexpr int kGuideLobeF4 = 10;     // processed lobes per query: 8 x (mu, kappa), 2 x (pi * vmf_norm)[4]

LFH_HD int guide_layer_out(int l) { return l == 3 ? kGuideOut : kGuideHid; }
LFH_HD int guide_layer_offset(int l) { return l * (64 * 64 + 64); }

// Device view of the guide's weights.  wf: per layer [64 out][64 in] BF16 (the
// forward product's B operand, zero rows past the layer's outputs); wb: per
// layer [64 in][64 out] BF16 (the backward product's B operand); bias: per
// layer 64 FP32 (zero-padded).
struct GuideDev {
    const uint16_t* wf;
    const uint16_t* wb;
    const float* bias;
    float lo[3];
    float inv_ext[3];
};

LFH_HD float guide_norm_coord(float p, float lo, float inv_ext) {
    float x = (p - lo) * inv_ext;
    return x < 0.0f ? 0.0f : (x > 1.0f ? 1.0f : x);
}

#if defined(__HIP__)
// The 64 features of one query (x already normalised).  Each coordinate's
// base angle pi*c gets one sincos; the higher octaves follow by the double-
// angle identities sin 2a = 2 sin a cos a, cos 2a = (cos a - sin a)(cos a + sin a),
// which keeps the absolute error below 1e-5 at the top octave (2^5).
LFH_DEV void guide_encode(V3 x, V3 n, V3 wo, float* f) {
    const float xs[3] = {x.x, x.y, x.z}, ns[3] = {n.x, n.y, n.z}, ws[3] = {wo.x, wo.y, wo.z};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float sn, cs;
        sincosf(kPi * xs[a], &sn, &cs);
#pragma unroll
        for (int fr = 0; fr < 6; ++fr) {
            f[(a * 6 + fr) * 2] = sn;
            f[(a * 6 + fr) * 2 + 1] = cs;
            const float s2 = 2.0f * sn * cs, c2 = (cs - sn) * (cs + sn);
            sn = s2;
            cs = c2;
        }
        sincosf(kPi * ns[a], &sn, &cs);
        f[36 + a * 4 + 0] = sn;
        f[36 + a * 4 + 1] = cs;
        f[36 + a * 4 + 2] = 2.0f * sn * cs;
        f[36 + a * 4 + 3] = (cs - sn) * (cs + sn);
        sincosf(kPi * ws[a], &sn, &cs);
        f[48 + a * 4 + 0] = sn;
        f[48 + a * 4 + 1] = cs;
        f[48 + a * 4 + 2] = 2.0f * sn * cs;
        f[48 + a * 4 + 3] = (cs - sn) * (cs + sn);
    }
    f[60] = wo.x;
    f[61] = wo.y;
    f[62] = wo.z;
    f[63] = 1.0f;
}

LFH_DEV uint16_t to_bf16(float v) {
    uint32_t u = __float_as_uint(v);
    u += 0x7FFFu + ((u >> 16) & 1u);   // round to nearest even
    return (uint16_t)(u >> 16);
}
LFH_DEV float from_bf16(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// Normalisation of a vMF lobe on the sphere: kappa / (2 pi (1 - e^{-2 kappa})).
LFH_DEV float vmf_norm(float kappa) { return kappa / (kTwoPi * -expm1f(-2.0f * kappa)); }

// 40 raw outputs -> 10 float4: lobes (mu, kappa), then per lobe its mixture
// weight times its normalisation, w_k = pi_k * vmf_norm(kappa_k) (the
// factor every pdf evaluation needs; the sampler divides it out again).
LFH_DEV void guide_lobes(const float* o, float4* out) {
    float a[kGuideLobes];
    float amax = -3.0e38f;
#pragma unroll
    for (int k = 0; k < kGuideLobes; ++k) {
        a[k] = o[5 * k + 4];
        amax = fmaxf(amax, a[k]);
    }
    float sum = 0.0f;
#pragma unroll
    for (int k = 0; k < kGuideLobes; ++k) {
        a[k] = expf(a[k] - amax);
        sum += a[k];
    }
    const float inv = 1.0f / sum;
#pragma unroll
    for (int k = 0; k < kGuideLobes; ++k) {
        float rx = o[5 * k], ry = o[5 * k + 1], rz = o[5 * k + 2];
        float rn = sqrtf(rx * rx + ry * ry + rz * rz);
        float ir = rn > 1e-12f ? 1.0f / rn : 0.0f;
        if (rn <= 1e-12f) rz = 1.0f, ir = 1.0f;
        float s = fminf(fmaxf(o[5 * k + 3], kLogKMin), kLogKMax);
        out[k] = make_float4(rx * ir, ry * ir, rz * ir, expf(s));
    }
    float w[kGuideLobes];
#pragma unroll
    for (int k = 0; k < kGuideLobes; ++k) w[k] = a[k] * inv * vmf_norm(out[k].w);
    out[8] = make_float4(w[0], w[1], w[2], w[3]);
    out[9] = make_float4(w[4], w[5], w[6], w[7]);
    // a query whose outputs are not all finite gets no guide: weights -1 mark it
    bool ok = isfinite(inv);
#pragma unroll
    for (int k = 0; k < kGuideLobes; ++k) ok = ok && isfinite(w[k]);
#pragma unroll
    for (int k = 0; k < kGuideLobes; ++k)
        ok = ok && isfinite(out[k].x) && isfinite(out[k].y) && isfinite(out[k].z) && isfinite(out[k].w);
    if (!ok) {
        out[8] = make_float4(-1.f, -1.f, -1.f, -1.f);
        out[9] = make_float4(-1.f, -1.f, -1.f, -1.f);
    }
}

// The processed lobes of one query are lob[k * st], k = 0..9 (st = 1: ten
// consecutive float4; larger st: one query of a structure-of-arrays buffer).
// True when they are usable (see guide_lobes).
LFH_DEV bool guide_valid(const float4* __restrict__ lob, size_t st = 1) { return lob[8 * st].x >= 0.0f; }

// Stores master parameter j (flat layout) into the BF16 operand copies:
// wf[layer][out][in], wb[layer][in][out], or the FP32 bias.
LFH_DEV void guide_store_param(uint16_t* wf, uint16_t* wb, float* bias, int j, float w) {
    const int layer = j < 3 * (64 * 64 + 64) ? j / (64 * 64 + 64) : 3;
    const int r = j - guide_layer_offset(layer);
    const int nout = guide_layer_out(layer);
    if (r < 64 * nout) {
        const int in = r / nout, out = r - in * nout;
        const uint16_t h = to_bf16(w);
        wf[layer * 4096 + out * 64 + in] = h;
        wb[layer * 4096 + in * 64 + out] = h;
    } else {
        bias[layer * 64 + (r - 64 * nout)] = w;
    }
}

// sum_k pi_k vmf_norm(kappa_k) exp(kappa_k (mu_k . d - 1)), the first product
// stored per lobe by guide_lobes.
LFH_DEV float guide_pdf(const float4* __restrict__ lob, V3 d, size_t st = 1) {
    const float4 pa = lob[8 * st], pb = lob[9 * st];
    const float w[8] = {pa.x, pa.y, pa.z, pa.w, pb.x, pb.y, pb.z, pb.w};
    float p = 0.0f;
#pragma unroll
    for (int k = 0; k < kGuideLobes; ++k) {
        const float4 l = lob[k * st];
        const float c = l.x * d.x + l.y * d.y + l.z * d.z;
        p += w[k] * expf(l.w * (c - 1.0f));
    }
    return p;
}

// Orthonormal basis around a unit vector (Duff et al. 2017, "Building an
// orthonormal basis, revisited").
LFH_DEV void guide_frame(V3 n, V3& t, V3& b) {
    const float sgn = n.z >= 0.0f ? 1.0f : -1.0f;
    const float a = -1.0f / (sgn + n.z);
    const float bb = n.x * n.y * a;
    t = v3(1.0f + sgn * n.x * n.x * a, sgn * bb, -sgn * n.x);
    b = v3(bb, sgn + n.y * n.y * a, -n.y);
}

// Draws a lobe by its weight (u0), then a direction from that vMF lobe by the
// inverse CDF of cos(theta) (Jakob 2012, "Numerically stable sampling of the
// von Mises Fisher distribution on S2"): w = 1 + log(u1 + (1-u1) e^{-2 kappa}) / kappa,
// phi = 2 pi u2.
LFH_DEV V3 guide_sample(const float4* __restrict__ lob, float u0, float u1, float u2, size_t st = 1) {
    const float4 pa = lob[8 * st], pb = lob[9 * st];
    const float wn[8] = {pa.x, pa.y, pa.z, pa.w, pb.x, pb.y, pb.z, pb.w};
    float pis[8];
#pragma unroll
    for (int j = 0; j < kGuideLobes; ++j) pis[j] = wn[j] / vmf_norm(lob[j * st].w);
    int k = kGuideLobes - 1;
    float cum = 0.0f;
#pragma unroll
    for (int j = 0; j < kGuideLobes - 1; ++j) {
        cum += pis[j];
        if (u0 < cum && k == kGuideLobes - 1) k = j;
    }
    float4 l = lob[0];
#pragma unroll
    for (int j = 1; j < kGuideLobes; ++j)
        if (j == k) l = lob[j * st];
    const float kappa = l.w;
    float w = 1.0f + logf(u1 + (1.0f - u1) * expf(-2.0f * kappa)) / kappa;
    w = fminf(fmaxf(w, -1.0f), 1.0f);
    const float s = sqrtf(fmaxf(0.0f, 1.0f - w * w));
    const V3 mu = v3(l.x, l.y, l.z);
    V3 t, b;
    guide_frame(mu, t, b);
    const float ph = kTwoPi * u2;
    const float cp = cosf(ph), sp = sinf(ph);
    return v3(w * mu.x + s * (cp * t.x + sp * b.x), w * mu.y + s * (cp * t.y + sp * b.y),
              w * mu.z + s * (cp * t.z + sp * b.z));
}
#endif  // __HIP__

}  // namespace lfhd
