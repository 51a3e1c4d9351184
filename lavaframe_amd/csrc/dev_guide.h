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
#ifndef LFH_GUIDE_FRAC
// Served guiding room, median guided time to RMSE 1e-2 over two sets of 32 sequences:
// 0.35: 0.165 / 0.161 s, 0.5: 0.156 / 0.155, 0.65: 0.164 / 0.166, 0.75: 0.158 / 0.160.
#define LFH_GUIDE_FRAC 0.5f
#endif
constexpr float kGuideFrac = LFH_GUIDE_FRAC;   // probability of drawing from the guide
constexpr int kGuideLobeF4 = 6;      // processed lobes per query: 4 x two packed (mu, kappa), 2 x (pi * vmf_norm)[4]

LFH_HD int guide_layer_out(int l) { return l == 3 ? kGuideOut : kGuideHid; }
LFH_HD int guide_layer_offset(int l) { return l * (64 * 64 + 64); }

// Device view of the guide's weights.  wf holds two BF16 images of the
// weights: [0, kGuideWImage) per layer [64 out][64 in] (the training forward's
// B operand, zero rows past the layer's outputs) and [kGuideWImage,
// 2 kGuideWImage) the inference image (guide_infer_frag); wb: per layer
// [64 in][64 out] BF16 (the backward product's B operand); bias: per layer 64
// FP32 in output order (zero-padded), then the same in inference row order.
constexpr int kGuideWImage = 4 * 4096;
constexpr int kGuideBImage = 4 * 64;
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
// sin and cos of pi v for v in [-1, 1]: the hardware's v_sin_f32 / v_cos_f32
// take the angle in revolutions (sin 2 pi r), here r = v / 2 (exact).
LFH_DEV void sincos_pi(float v, float* s, float* c) {
    const float r = 0.5f * v;
    *s = __builtin_amdgcn_sinf(r);
    *c = __builtin_amdgcn_cosf(r);
}

// The 64 features of one query (x already normalised).  Each coordinate's
// base angle pi*c gets one hardware sin/cos pair; the higher octaves follow
// by the double-angle identities sin 2a = 2 sin a cos a, cos 2a = (cos a -
// sin a)(cos a + sin a), which multiply the base error by at most 2^5 at the
// top octave: a few 1e-5 absolute, far below the BF16 step (2^-8 near 1) the
// features are rounded to.
LFH_DEV void guide_encode(V3 x, V3 n, V3 wo, float* f) {
    const float xs[3] = {x.x, x.y, x.z}, ns[3] = {n.x, n.y, n.z}, ws[3] = {wo.x, wo.y, wo.z};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float sn, cs;
        sincos_pi(xs[a], &sn, &cs);
#pragma unroll
        for (int fr = 0; fr < 6; ++fr) {
            f[(a * 6 + fr) * 2] = sn;
            f[(a * 6 + fr) * 2 + 1] = cs;
            const float s2 = 2.0f * sn * cs, c2 = (cs - sn) * (cs + sn);
            sn = s2;
            cs = c2;
        }
        sincos_pi(ns[a], &sn, &cs);
        f[36 + a * 4 + 0] = sn;
        f[36 + a * 4 + 1] = cs;
        f[36 + a * 4 + 2] = 2.0f * sn * cs;
        f[36 + a * 4 + 3] = (cs - sn) * (cs + sn);
        sincos_pi(ws[a], &sn, &cs);
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
// Every kappa here is exp(s) with s >= -3, so 1 - e^{-2 kappa} >= 0.095 and
// the subtraction loses under 1e-6 relative: hardware exp and reciprocal.
LFH_DEV float vmf_norm(float kappa) {
    return kappa * __builtin_amdgcn_rcpf(kTwoPi * (1.0f - __expf(-2.0f * kappa)));
}
// 1 / vmf_norm, for the sampler's choice of lobe.
LFH_DEV float vmf_inv_norm(float kappa) { return kTwoPi * (1.0f - __expf(-2.0f * kappa)) * __builtin_amdgcn_rcpf(kappa); }

// Processed lobes, kGuideLobeF4 float4 per query: float4 r < 4 holds lobes 2r
// and 2r+1 packed (guide_pack_lobe), each the (mu, kappa) of outputs 5k..5k+4
// = (r, s, a), mu = normalize(r), kappa = exp(clamp(s, kLogKMin, kLogKMax));
// float4 4 and 5 hold w_k = pi_k * vmf_norm(kappa_k) of the packed kappa, pi =
// softmax(a) (the factor every pdf evaluation needs; the sampler divides it
// out again), or -1 for a query whose outputs are not all finite (no guide).
// guide_lobe makes one (mu, kappa).
LFH_DEV float4 guide_lobe(const float* o5) {
    float rx = o5[0], ry = o5[1], rz = o5[2];
    const float r2 = rx * rx + ry * ry + rz * rz;
    float ir = r2 > 1e-24f ? __builtin_amdgcn_rsqf(r2) : 0.0f;
    if (r2 <= 1e-24f) rz = 1.0f, ir = 1.0f;
    const float sk = fminf(fmaxf(o5[3], kLogKMin), kLogKMax);
    return make_float4(rx * ir, ry * ir, rz * ir, __expf(sk));
}

// A lobe in 8 bytes: mu.x, mu.y, mu.z and kappa as binary16, rounded to
// nearest even.  The sampler and every pdf evaluation read the lobe back
// through guide_unpack_lobe, so the mixture they use is one and the same.
LFH_DEV uint2 guide_pack_lobe(float4 l) {
    const uint32_t x = __builtin_bit_cast(uint16_t, (_Float16)l.x), y = __builtin_bit_cast(uint16_t, (_Float16)l.y);
    const uint32_t z = __builtin_bit_cast(uint16_t, (_Float16)l.z), k = __builtin_bit_cast(uint16_t, (_Float16)l.w);
    return make_uint2(x | (y << 16), z | (k << 16));
}
LFH_DEV float half_lo(uint32_t v) { return (float)__builtin_bit_cast(_Float16, (uint16_t)(v & 0xFFFFu)); }
LFH_DEV float half_hi(uint32_t v) { return (float)__builtin_bit_cast(_Float16, (uint16_t)(v >> 16)); }
LFH_DEV float guide_packed_kappa(uint2 p) { return half_hi(p.y); }
// (mu renormalised after the rounding, kappa)
LFH_DEV float4 guide_unpack_lobe(uint2 p) {
    const float x = half_lo(p.x), y = half_hi(p.x), z = half_lo(p.y);
    const float ir = __builtin_amdgcn_rsqf(x * x + y * y + z * z);
    return make_float4(x * ir, y * ir, z * ir, half_hi(p.y));
}
// Packed lobe j of one query (lob as in guide_valid).
LFH_DEV uint2 guide_packed(const float4* __restrict__ lob, int j, size_t st) {
    const uint4 u = __builtin_bit_cast(uint4, lob[(j >> 1) * st]);
    return (j & 1) ? make_uint2(u.z, u.w) : make_uint2(u.x, u.y);
}

// The processed lobes of one query are lob[r * st], r = 0..5 (st = 1: six
// consecutive float4; larger st: one query of a structure-of-arrays buffer).
// True when they are usable.
LFH_DEV bool guide_valid(const float4* __restrict__ lob, size_t st = 1) { return lob[4 * st].x >= 0.0f; }

// Inference image.  Inference runs the MLP transposed, Y^T = W^T X^T, on
// 16-query tiles with v_mfma_f32_16x16x32_bf16 (weights = A operand, 16 rows
// x 32 K per fragment; activations = B operand, 32 K x 16 queries): a layer's
// accumulators then hold, in lane group g = lane / 16, rows 4g .. 4g+3 of each
// 16-row tile for query lane % 16.  The image orders rows and columns so that
// no activation leaves its lane between layers:
//   row p of layers 0-2 is hidden unit p; row p of layer 3 is output
//   10 (q / 4) + 4 t + q % 4 (q = p % 16, t = p / 16; none where 4 t + q % 4
//   >= 10), so lane group g ends with lobes 2g and 2g+1 of its query;
//   column k = 32 ks + 8 g + j of layer 0 is input feature k; of layers 1-3,
//   hidden unit 16 (2 ks + j / 4) + 4 g + j % 4: the B fragment a lane packs
//   from its own accumulators of tiles 2 ks and 2 ks + 1.
LFH_HD int guide_infer_row(int layer, int out) {
    if (layer < 3) return out;
    const int g = out / 10, s = out - 10 * g;
    return (s >> 2) * 16 + g * 4 + (s & 3);
}
LFH_HD int guide_infer_col(int layer, int in) {
    if (layer == 0) return in;
    const int t = in >> 4, g = (in >> 2) & 3;
    return (t >> 1) * 32 + g * 8 + (t & 1) * 4 + (in & 3);
}
// Element (row p, column k) of a layer's image: A fragments stored
// [layer][tile p/16][K step k/32][lane 16 ((k % 32) / 8) + p % 16][k % 8].
LFH_HD int guide_infer_frag(int layer, int p, int k) {
    return (((layer * 4 + (p >> 4)) * 2 + (k >> 5)) * 64 + ((k & 31) >> 3) * 16 + (p & 15)) * 8 + (k & 7);
}

// Stores master parameter j (flat layout) into the BF16 operand copies (both
// images of wf, and wb) or the FP32 biases (both orders).
LFH_DEV void guide_store_param(uint16_t* wf, uint16_t* wb, float* bias, int j, float w) {
    const int layer = j < 3 * (64 * 64 + 64) ? j / (64 * 64 + 64) : 3;
    const int r = j - guide_layer_offset(layer);
    const int nout = guide_layer_out(layer);
    if (r < 64 * nout) {
        const int in = r / nout, out = r - in * nout;
        const uint16_t h = to_bf16(w);
        wf[layer * 4096 + out * 64 + in] = h;
        wf[kGuideWImage + guide_infer_frag(layer, guide_infer_row(layer, out), guide_infer_col(layer, in))] = h;
        wb[layer * 4096 + in * 64 + out] = h;
    } else {
        const int out = r - 64 * nout;
        bias[layer * 64 + out] = w;
        bias[kGuideBImage + layer * 64 + guide_infer_row(layer, out)] = w;
    }
}

// sum_k pi_k vmf_norm(kappa_k) exp(kappa_k (mu_k . d - 1)), the first product
// stored per lobe in the processed lobes.
LFH_DEV float guide_pdf(const float4* __restrict__ lob, V3 d, size_t st = 1) {
    const float4 pa = lob[4 * st], pb = lob[5 * st];
    const float w[8] = {pa.x, pa.y, pa.z, pa.w, pb.x, pb.y, pb.z, pb.w};
    float p = 0.0f;
#pragma unroll
    for (int k = 0; k < kGuideLobes; ++k) {
        const float4 l = guide_unpack_lobe(guide_packed(lob, k, st));
        const float c = l.x * d.x + l.y * d.y + l.z * d.z;
        p += w[k] * __expf(l.w * (c - 1.0f));
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
    const float4 pa = lob[4 * st], pb = lob[5 * st];
    const float wn[8] = {pa.x, pa.y, pa.z, pa.w, pb.x, pb.y, pb.z, pb.w};
    float pis[8];
#pragma unroll
    for (int j = 0; j < kGuideLobes; ++j) pis[j] = wn[j] * vmf_inv_norm(guide_packed_kappa(guide_packed(lob, j, st)));
    int k = kGuideLobes - 1;
    float cum = 0.0f;
#pragma unroll
    for (int j = 0; j < kGuideLobes - 1; ++j) {
        cum += pis[j];
        if (u0 < cum && k == kGuideLobes - 1) k = j;
    }
    // the chosen lobe read again (the same cache lines): no eight packed lobes held in registers
    const float4 l = guide_unpack_lobe(guide_packed(lob, k, st));
    const float kappa = l.w;
    float w = 1.0f + __logf(u1 + (1.0f - u1) * __expf(-2.0f * kappa)) * __builtin_amdgcn_rcpf(kappa);
    w = fminf(fmaxf(w, -1.0f), 1.0f);
    const float s = sqrtf(fmaxf(0.0f, 1.0f - w * w));
    const V3 mu = v3(l.x, l.y, l.z);
    V3 t, b;
    guide_frame(mu, t, b);
    // phi = 2 pi u2: the hardware sin / cos take revolutions
    const float cp = __builtin_amdgcn_cosf(u2), sp = __builtin_amdgcn_sinf(u2);
    return v3(w * mu.x + s * (cp * t.x + sp * b.x), w * mu.y + s * (cp * t.y + sp * b.y),
              w * mu.z + s * (cp * t.z + sp * b.z));
}
#endif  // __HIP__

}  // namespace lfhd
