// Learned path guide on gfx950: BF16 MFMA (v_mfma_f32_16x16x32_bf16) MLP
// inference and training, vMF-mixture sampling entry points, deterministic
// training-sample compaction and Adam.
//
// MFMA operand layout (16x16x32, one wave of 64 lanes; lane l):
//   A (16 x 32):  row l%16, k = (l/16)*8 .. +7     (8 BF16 = 16 B per lane)
//   B (32 x 16):  col l%16, k = (l/16)*8 .. +7
//   D (16 x 16):  col l%16, rows (l/16)*4 .. +3    (4 FP32)
// Training: activations live in LDS as BF16 rows of kPitch elements (64 + 8
// padding so that the 16 B fragment reads of 16 consecutive rows spread over
// the banks); weights are read as B fragments straight from global memory
// (32 KiB per direction, resident in L1/L2).  Inference: the encoded inputs
// pass through LDS once, then the MLP runs transposed with the weights held in
// registers and the activations never leaving their lanes (k_guide_infer).
#include <algorithm>

#include "dev_guide.h"
#include "dev_shade.h"
#include "guide_kernels.h"
#include "pt_kernels.h"

namespace lfhd {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef short i16x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

constexpr int kPitch = 72;            // BF16 elements per activation row
constexpr int kOutPitch = 41;         // FP32 elements per output row
#ifndef LFH_INFER_LDS_W
// 1: layers 0 and 3 of the inference weights in LDS (52 KB per block), budgeted for 3 waves per
// SIMD (162 VGPRs).  Measured 575 against 600 TFLOP/s on 1M host queries and a 1 % slower guided
// C3 pass: the per-MFMA LDS reads cost more than the third wave hides.
#define LFH_INFER_LDS_W 0
#endif
#ifndef LFH_INFER_TILES
// 16-query tiles a wave runs through the MLP together (1 or 2).  Two (218 VGPRs, still 2
// waves per SIMD) measured 576 against 588 TFLOP/s on 1M host queries and the same guided
// C3 time: the four independent accumulator chains of one tile already fill the MFMA pipe.
#define LFH_INFER_TILES 1
#endif
constexpr int kInferBlock = 256;      // 4 persistent waves
constexpr int kInferWaves = kInferBlock / 64;
constexpr int kTrainRows = 64;        // samples per training block
constexpr int kTrainBlock = 256;      // 4 waves: one 16-row M tile each
constexpr int kPartStride = kGuideParams + 1;  // per-block partial: grads + loss

LFH_DEV bf16x8 ld_frag(const uint16_t* p) { return *reinterpret_cast<const bf16x8*>(p); }

LFH_DEV f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0); }

// Two FP32 as a BF16 pair, round to nearest even (v_cvt_pk_bf16_f32).
LFH_DEV uint32_t pack_bf16x2(float a, float b) {
    return __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2){a, b}, bf16x2));
}

// Writes the 64 encoded features of one query as BF16 into an LDS row.
LFH_DEV void encode_row(uint16_t* row, const GuideDev& g, V3 p, V3 n, V3 wo) {
    const V3 x = v3(guide_norm_coord(p.x, g.lo[0], g.inv_ext[0]), guide_norm_coord(p.y, g.lo[1], g.inv_ext[1]),
                    guide_norm_coord(p.z, g.lo[2], g.inv_ext[2]));
    float f[kGuideIn];
    guide_encode(x, n, wo, f);
#pragma unroll
    for (int k = 0; k < kGuideIn; k += 8) {
        const u32x4 v = {pack_bf16x2(f[k], f[k + 1]), pack_bf16x2(f[k + 2], f[k + 3]), pack_bf16x2(f[k + 4], f[k + 5]),
                         pack_bf16x2(f[k + 6], f[k + 7])};
        *reinterpret_cast<u32x4*>(row + k) = v;
    }
}

// ---- inference ------------------------------------------------------------------
// Source of the queries: host-provided (position, normal, wo) triples, or the
// current hit queue of the path tracer.
struct InferQueue {
    ShadeScene sc;
    const float4* rayq;
    const float4* hitq;
    const uint32_t* count;
    GuideHook gh;
    int depth;
};

// Query inputs of hit-queue entry i: the hit point, the face-forwarded
// geometric normal in world space, and wo = -dir.  Returns false for misses
// and light hits (no lobes needed).
// The ray and hit records of entry i (the first loads of a query's chain).
struct QueueRec {
    float4 r0, r1, h0, h1;
};
LFH_DEV QueueRec load_queue_rec(const InferQueue& q, uint32_t i) {
    return QueueRec{q.rayq[(size_t)i * 2], q.rayq[(size_t)i * 2 + 1], q.hitq[(size_t)i * 2], q.hitq[(size_t)i * 2 + 1]};
}
LFH_DEV bool queue_query(const InferQueue& q, const QueueRec& rec, V3& pos, V3& nrm, V3& wo, int& path) {
    const float4 r0 = rec.r0, r1 = rec.r1, h0 = rec.h0, h1 = rec.h1;
    path = __float_as_int(r0.w);
    const V3 ro = ld3(r0), rd = ld3(r1);
    wo = -rd;
    pos = ro + rd * h0.x;
    const int slot = __float_as_int(h0.w);
    if (h0.x == kInfinity || slot < 0) {
        nrm = v3(0.f, 0.f, 1.f);
        return false;
    }
    const int inst = __float_as_int(h1.x);
    const float4* tri = q.sc.tr.tris + (size_t)slot * 3;
    const V3 e0 = ld3(tri[1]), e1 = ld3(tri[2]);
    const V3 no = cross(e0, e1);
    const float4* np = q.sc.nmats + (size_t)inst * 3;
    const float4 m0 = np[0], m1 = np[1], m2 = np[2];
    V3 n = v3(m0.x * no.x + m0.y * no.y + m0.z * no.z, m1.x * no.x + m1.y * no.y + m1.z * no.z,
              m2.x * no.x + m2.y * no.y + m2.z * no.z);
    const float l2 = dot(n, n);
    n = l2 > 0.0f ? n * (1.0f / sqrtf(l2)) : v3(0.f, 0.f, 1.f);
    nrm = dot(n, rd) > 0.0f ? -n : n;
    return true;
}

// Training record of a recorded path's vertex: the query inputs; k_shade adds
// the sampled direction and the radiance terms.
LFH_DEV void record_query(const GuideHook& gh, int path, int depth, V3 pos, V3 nrm, V3 wo) {
    float4* r = gh.rec + ((size_t)(path / gh.stride) * gh.max_depth + depth) * 4;
    r[0] = make_float4(pos.x, pos.y, pos.z, 0.0f);
    r[1] = make_float4(nrm.x, nrm.y, nrm.z, 0.0f);
    r[2] = make_float4(wo.x, wo.y, wo.z, -1.0f);
}

// Accumulators of tiles 2ks and 2ks+1 after bias -> BF16 (v_cvt_pk_bf16_f32,
// round to nearest even) -> ReLU, as the next layer's B fragment of K step ks.
// The ReLU is a max with 0 of the BF16 pairs read as signed 16-bit integers:
// a BF16 with its sign bit set is a negative integer.

LFH_DEV uint32_t pack_relu2(float a, float b) {
    const i16x2 h = __builtin_bit_cast(i16x2, pack_bf16x2(a, b));
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(h, (i16x2){0, 0}));
}
LFH_DEV bf16x8 next_operand(f32x4 lo, f32x4 hi) {
    const u32x4 u = {pack_relu2(lo[0], lo[1]), pack_relu2(lo[2], lo[3]), pack_relu2(hi[0], hi[1]),
                     pack_relu2(hi[2], hi[3])};
    return __builtin_bit_cast(bf16x8, u);
}

// Persistent waves, 64 queries a round: lane l fetches and encodes query
// base + l into the wave's LDS rows; then each tile of 16 queries runs the
// four layers on the inference image (weights as A operand, loaded into
// registers once per wave), the accumulators of one layer packed in place
// into the next layer's B operand.  Lane group g = lane / 16 ends with lobes
// 2g, 2g+1 of query lane % 16 of the tile; the softmax over the 8 lobes and
// the validity flag are reduced over the 4 lane groups by shuffles.  Output:
// raw40 (40 outputs per query) and/or the processed lobes (dev_guide.h;
// QUEUE: structure of arrays with stride gh.lobe_stride).
template <bool QUEUE>
__global__ __launch_bounds__(kInferBlock, LFH_INFER_LDS_W ? 3 : 1) void k_guide_infer(GuideDev g, const float* __restrict__ in9, int n_host,
                                                             InferQueue q, float* __restrict__ raw40,
                                                             float4* __restrict__ lobes) {
    __shared__ __attribute__((aligned(16))) uint16_t sX[kInferWaves][64 * kPitch];
    __shared__ float4 sBias[kGuideBImage / 4];
#if LFH_INFER_LDS_W
    // layers 0 and 3 (14 fragments) from LDS, layers 1 and 2 in registers
    __shared__ bf16x8 sW[14][64];
    for (int i = threadIdx.x; i < 14 * 64; i += kInferBlock) {
        const int f = i >> 6, ln = i & 63;
        const int l = f < 8 ? 0 : 3, nt = (f & 7) >> 1, ks = f & 1;
        sW[f][ln] = ld_frag(g.wf + kGuideWImage + ((l * 4 + nt) * 2 + ks) * 512 + ln * 8);
    }
#endif
    const uint32_t n = QUEUE ? *q.count : (uint32_t)n_host;
    if ((uint32_t)blockIdx.x * kInferBlock >= n) return;  // block-uniform
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int lr = lane & 15, grp = lane >> 4;
    if (threadIdx.x < kGuideBImage / 4)
        sBias[threadIdx.x] = reinterpret_cast<const float4*>(g.bias + kGuideBImage)[threadIdx.x];
    bf16x8 w[kGuideLayers][4][2];
    {
        const uint16_t* wi = g.wf + kGuideWImage + lane * 8;
#pragma unroll
        for (int l = 0; l < kGuideLayers; ++l)
#pragma unroll
            for (int nt = 0; nt < 4; ++nt)
#pragma unroll
                for (int ks = 0; ks < 2; ++ks)
                    w[l][nt][ks] = (l < 3 || nt < 3) && !(LFH_INFER_LDS_W && (l == 0 || l == 3))
                                       ? ld_frag(wi + ((l * 4 + nt) * 2 + ks) * 512)
                                       : bf16x8{};
    }
    __syncthreads();
    uint16_t* X = sX[wave];
    const uint32_t step = gridDim.x * kInferBlock;
    const uint32_t first = (blockIdx.x * kInferWaves + wave) * 64u;
    // hit-queue records one round ahead: their loads are in flight while the
    // current round runs its tiles.  Lanes past the end load entry n - 1
    // (unused), so the loads need no branch and the wait before their use
    // counts only them.
    QueueRec rec{};
    if (QUEUE) rec = load_queue_rec(q, min(first + lane, n - 1));
    for (uint32_t base = first; base < n; base += step) {
        const uint32_t i = base + lane;
        QueueRec next{};
        if (QUEUE) next = load_queue_rec(q, min(base + step + lane, n - 1));
        V3 pos = v3(0.f, 0.f, 0.f), nrm = v3(0.f, 0.f, 1.f), wo = v3(0.f, 0.f, 1.f);
        if (i < n) {
            if (QUEUE) {
                int path;
                const bool surf = queue_query(q, rec, pos, nrm, wo, path);
                const GuideHook& gh = q.gh;
                if (surf && gh.train && (path % gh.stride) == gh.phase) record_query(gh, path, q.depth, pos, nrm, wo);
            } else {
                const float* s = in9 + (size_t)i * 9;
                pos = v3(s[0], s[1], s[2]);
                nrm = v3(s[3], s[4], s[5]);
                wo = v3(s[6], s[7], s[8]);
            }
        }
        // the rows are read back by other lanes of this wave only: LDS
        // operations of one wave complete in order
        encode_row(X + lane * kPitch, g, pos, nrm, wo);
        // Outputs of one 16-query tile: raw40 and / or the processed lobes.
        auto emit = [&](const f32x4 (&acc)[4], int t) {
            // outputs 10 grp + s, s = 4 nt + i, in acc[nt][i]
            float o[10];
#pragma unroll
            for (int s2 = 0; s2 < 10; ++s2) o[s2] = acc[s2 >> 2][s2 & 3];
            const uint32_t qi = base + t * 16 + lr;
            if (raw40 && qi < n) {
                float* dst = raw40 + (size_t)qi * kGuideOut + grp * 10;
#pragma unroll
                for (int s2 = 0; s2 < 10; ++s2) dst[s2] = o[s2];
            }
            if (!lobes) return;
            // the processed lobes (dev_guide.h) over the query's four lane
            // groups: lobes 2 grp, 2 grp + 1 here
            float m = fmaxf(o[4], o[9]);
            m = fmaxf(m, __shfl_xor(m, 16));
            m = fmaxf(m, __shfl_xor(m, 32));
            const float e0 = __expf(o[4] - m), e1 = __expf(o[9] - m);
            float sum = e0 + e1;
            sum += __shfl_xor(sum, 16);
            sum += __shfl_xor(sum, 32);
            const float inv = __builtin_amdgcn_rcpf(sum);
            const float4 lb[2] = {guide_lobe(o), guide_lobe(o + 5)};
            const uint2 pk0 = guide_pack_lobe(lb[0]), pk1 = guide_pack_lobe(lb[1]);
            // the weights carry the normalisation of the stored (rounded) kappa
            const float wk0 = e0 * inv * vmf_norm(guide_packed_kappa(pk0));
            const float wk1 = e1 * inv * vmf_norm(guide_packed_kappa(pk1));
            float wk[2] = {wk0, wk1};
            bool ok = isfinite(inv) && isfinite(wk0) && isfinite(wk1);
#pragma unroll
            for (int j = 0; j < 2; ++j)
                ok = ok && isfinite(lb[j].x) && isfinite(lb[j].y) && isfinite(lb[j].z) && isfinite(lb[j].w);
            int okw = ok ? 1 : 0;
            okw &= __shfl_xor(okw, 16);
            okw &= __shfl_xor(okw, 32);
            if (!okw) wk[0] = wk[1] = -1.0f;   // a query with a non-finite output gets no guide
            if (qi < n) {
                const size_t st = QUEUE ? (size_t)q.gh.lobe_stride : 1;
                float4* dst = lobes + (QUEUE ? (size_t)qi : (size_t)qi * kGuideLobeF4);
                *reinterpret_cast<uint4*>(dst + grp * st) = make_uint4(pk0.x, pk0.y, pk1.x, pk1.y);
                reinterpret_cast<float2*>(dst + (4 + (grp >> 1)) * st)[grp & 1] = make_float2(wk[0], wk[1]);
            }
        };
        // The four layers of one tile from its two B fragments.
        auto layer = [&](int l, bf16x8& b0, bf16x8& b1, f32x4 (&acc)[4]) {
            const int nN = l < 3 ? 4 : 3;
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) {
                if (nt >= nN) continue;
                const float4 bb = sBias[(l * 64 + nt * 16 + grp * 4) >> 2];
                acc[nt] = f32x4{bb.x, bb.y, bb.z, bb.w};
#if LFH_INFER_LDS_W
                if (l == 0 || l == 3) {
                    const int f = (l == 0 ? 0 : 8) + nt * 2;
                    acc[nt] = mfma(sW[f][lane], b0, acc[nt]);
                    acc[nt] = mfma(sW[f + 1][lane], b1, acc[nt]);
                    continue;
                }
#endif
                acc[nt] = mfma(w[l][nt][0], b0, acc[nt]);
                acc[nt] = mfma(w[l][nt][1], b1, acc[nt]);
            }
        };
#if LFH_INFER_TILES == 2
        // two tiles at a time: each one's MFMA chain overlaps the other's packing
#pragma unroll 1
        for (int t = 0; t < 4; t += 2) {
            if (base + t * 16 >= n) break;   // wave-uniform
            const uint16_t* xa = X + (t * 16 + lr) * kPitch + grp * 8;
            const uint16_t* xb = xa + 16 * kPitch;
            bf16x8 a0 = ld_frag(xa), a1 = ld_frag(xa + 32), c0 = ld_frag(xb), c1 = ld_frag(xb + 32);
            f32x4 accA[4], accB[4];
#pragma unroll
            for (int l = 0; l < kGuideLayers; ++l) {
                layer(l, a0, a1, accA);
                layer(l, c0, c1, accB);
                if (l < 3) {
                    a0 = next_operand(accA[0], accA[1]);
                    a1 = next_operand(accA[2], accA[3]);
                    c0 = next_operand(accB[0], accB[1]);
                    c1 = next_operand(accB[2], accB[3]);
                }
            }
            emit(accA, t);
            emit(accB, t + 1);
        }
#else
#pragma unroll 1
        for (int t = 0; t < 4; ++t) {
            if (base + t * 16 >= n) break;   // wave-uniform
            const uint16_t* xr = X + (t * 16 + lr) * kPitch + grp * 8;
            bf16x8 b0 = ld_frag(xr), b1 = ld_frag(xr + 32);
            f32x4 acc[4];
#pragma unroll
            for (int l = 0; l < kGuideLayers; ++l) {
                layer(l, b0, b1, acc);
                if (l < 3) {
                    b0 = next_operand(acc[0], acc[1]);
                    b1 = next_operand(acc[2], acc[3]);
                }
            }
            emit(acc, t);
        }
#endif
        rec = next;
    }
}

// Passes that draw no direction from the guide (the BSDF-only warm-up) need no
// lobes, only the training records of the hit queue's recorded paths.
__global__ __launch_bounds__(256) void k_guide_record(InferQueue q) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= *q.count) return;
    V3 pos, nrm, wo;
    int path;
    if (queue_query(q, load_queue_rec(q, i), pos, nrm, wo, path) && (path % q.gh.stride) == q.gh.phase)
        record_query(q.gh, path, q.depth, pos, nrm, wo);
}

// ---- mixture sampling / pdf on processed lobes (checks of the device sampler) ----
__global__ void k_guide_sample(const float4* __restrict__ lobes, const uint4* __restrict__ seeds, int n,
                               float* __restrict__ dirs, float* __restrict__ pdf) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint4 s = seeds[i];
    const float u0 = rnd(s), u1 = rnd(s), u2 = rnd(s);
    const float4* lb = lobes + (size_t)i * kGuideLobeF4;
    const V3 d = guide_sample(lb, u0, u1, u2);
    dirs[(size_t)i * 3] = d.x;
    dirs[(size_t)i * 3 + 1] = d.y;
    dirs[(size_t)i * 3 + 2] = d.z;
    pdf[i] = guide_pdf(lb, d);
}

__global__ void k_guide_pdf(const float4* __restrict__ lobes, const float* __restrict__ dirs, int n,
                            float* __restrict__ pdf) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    pdf[i] = guide_pdf(lobes + (size_t)i * kGuideLobeF4, v3(dirs[(size_t)i * 3], dirs[(size_t)i * 3 + 1],
                                                              dirs[(size_t)i * 3 + 2]));
}

// ---- training -----------------------------------------------------------------------
// One block = 64 samples (4 waves x one 16-row tile).  Forward with the
// activations kept in LDS, analytic d loss / d outputs per sample, then
// backward: weight gradients dW = X^T dY (K = the block's samples) and
// dX = (dY W^T) masked by ReLU, all as MFMA tiles.  Each block writes its
// partial gradient; k_guide_reduce_adam sums the partials in block order, so
// the result does not depend on scheduling.
//
// Samples: 4 float4 each: (pos, w), (n, 0), (wo, 0), (d, 0).
struct TrainArgs {
    const float4* samples;
    const int* count;      // device: number of valid samples
    int start;             // first sample of this batch
    int batch;             // batch capacity (blocks x 64)
    float* partial;        // [blocks][kPartStride]
    const float* wsum;     // device: sum of the sample weights (NULL: use the weights as given)
    float wcap;            // with wsum: weights are w / mean(w), capped at wcap
};

// The batch takes every stride-th sample of the list, stride = ceil(count /
// batch), starting at start % stride: a batch spreads over the whole list (the
// list is in path order), and successive starts visit all of it.
LFH_DEV int batch_stride(const TrainArgs& t) {
    const int c = *t.count;
    return c <= t.batch ? 1 : (c + t.batch - 1) / t.batch;
}
LFH_DEV int batch_phase(const TrainArgs& t) { return t.start % batch_stride(t); }
LFH_DEV int batch_size(const TrainArgs& t) {
    const int c = *t.count, st = batch_stride(t), ph = batch_phase(t);
    const int n = c > ph ? (c - ph + st - 1) / st : 0;
    return n > t.batch ? t.batch : n;
}

__global__ __launch_bounds__(kTrainBlock) void k_guide_train(GuideDev g, TrainArgs ta) {
    __shared__ __attribute__((aligned(16))) uint16_t sA[4][kTrainRows * kPitch];  // E, H1, H2, H3
    __shared__ __attribute__((aligned(16))) uint16_t sG0[kTrainRows * kPitch];
    // the FP32 outputs (forward, loss) and the second gradient buffer
    // (backward) are never live together
    __shared__ __attribute__((aligned(16))) float sOG[kTrainRows * kOutPitch];
    static_assert(sizeof(float) * kTrainRows * kOutPitch >= sizeof(uint16_t) * kTrainRows * kPitch, "alias size");
    float* sO = sOG;
    uint16_t* sG1 = reinterpret_cast<uint16_t*>(sOG);
    __shared__ float sD[kTrainRows][4];
    __shared__ float sLoss[kTrainRows];
    __shared__ float sDb[4][64];   // quarter sums of the bias gradient
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int lr = lane & 15, lk = (lane >> 4) * 8, lq = (lane >> 4) * 4;
    const int N = batch_size(ta);
    const int base = blockIdx.x * kTrainRows;
    float* part = ta.partial + (size_t)blockIdx.x * kPartStride;
    if (base >= N) return;  // block-uniform; the reduction reads only active blocks

    // 1. encode: one thread per sample (rows past the batch are zero with w = 0)
    if (tid < kTrainRows) {
        const int s = tid;
        const bool ok = base + s < N;
        uint16_t* row = sA[0] + s * kPitch;
        if (ok) {
            const float4* sp = ta.samples + ((size_t)(base + s) * batch_stride(ta) + batch_phase(ta)) * 4;
            const float4 a = sp[0], b = sp[1], c = sp[2], d = sp[3];
            sD[s][0] = d.x;
            sD[s][1] = d.y;
            sD[s][2] = d.z;
            float w = a.w;
            if (ta.wsum) {
                const float mean = *ta.wsum / (float)*ta.count;
                w = mean > 0.0f ? fminf(w / mean, ta.wcap) : 0.0f;
            }
            sD[s][3] = w;
            encode_row(row, g, ld3(a), ld3(b), ld3(c));
        } else {
            sD[s][0] = sD[s][1] = sD[s][2] = sD[s][3] = 0.0f;
            for (int f = 0; f < kGuideIn; ++f) row[f] = 0;
        }
    }
    __syncthreads();

    // 2. forward, wave w on rows 16w..16w+15; layer l reads sA[l], writes sA[l+1] (or sO)
#pragma unroll
    for (int layer = 0; layer < kGuideLayers; ++layer) {
        const uint16_t* W = g.wf + layer * 4096;
        const float* bias = g.bias + layer * 64;
        const int nN = layer == 3 ? 3 : 4;
        const uint16_t* X = sA[layer];
        f32x4 acc[4];
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) acc[nt] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) {
            const bf16x8 a = ld_frag(X + (wave * 16 + lr) * kPitch + ks * 32 + lk);
#pragma unroll
            for (int nt = 0; nt < 4; ++nt)
                if (nt < nN) acc[nt] = mfma(a, ld_frag(W + (nt * 16 + lr) * 64 + ks * 32 + lk), acc[nt]);
        }
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
            if (nt >= nN) continue;
            const int col = nt * 16 + lr;
            const float b = bias[col];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int row = wave * 16 + lq + j;
                const float v = acc[nt][j] + b;
                if (layer < 3) sA[layer + 1][row * kPitch + col] = to_bf16(fmaxf(v, 0.0f));
                else if (col < kGuideOut) sO[row * kOutPitch + col] = v;
            }
        }
    }
    __syncthreads();

    // 3. d loss / d outputs, BF16 into sG0 (cols 40..63 zero).  Four threads per
    // sample, in one wave: sample 16 wave + lane % 16, lane group g = lane / 16
    // takes lobes 2g and 2g+1; the softmax and log-sum-exp over the 8 lobes
    // reduce over the 4 groups with shuffles (every group ends with the same
    // sums: the two-step xor tree adds the same pairs in every lane).
    const float invN = 1.0f / (float)N;
    {
        const int s = wave * 16 + lr, grp = lane >> 4;
        const float* o = sO + s * kOutPitch;
        uint16_t* grow = sG0 + s * kPitch;
        const float w = sD[s][3];
        for (int c = kGuideOut + grp * 6; c < kGuideOut + grp * 6 + 6; ++c) grow[c] = 0;
        const V3 d = v3(sD[s][0], sD[s][1], sD[s][2]);
        const int k0 = 2 * grp;
        float amax = fmaxf(o[5 * k0 + 4], o[5 * k0 + 9]);
        amax = fmaxf(amax, __shfl_xor(amax, 16));
        amax = fmaxf(amax, __shfl_xor(amax, 32));
        float pi[2], lw[2], kap[2], cs[2], rn[2];
        V3 mu[2];
#pragma unroll
        for (int j = 0; j < 2; ++j) pi[j] = expf(o[5 * (k0 + j) + 4] - amax);
        float psum = pi[0] + pi[1];
        psum += __shfl_xor(psum, 16);
        psum += __shfl_xor(psum, 32);
        float lmax = -3.0e38f;
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int k = k0 + j;
            pi[j] /= psum;
            const V3 r = v3(o[5 * k], o[5 * k + 1], o[5 * k + 2]);
            rn[j] = fmaxf(sqrtf(dot(r, r)), 1e-12f);
            mu[j] = r * (1.0f / rn[j]);
            kap[j] = expf(fminf(fmaxf(o[5 * k + 3], kLogKMin), kLogKMax));
            cs[j] = dot(mu[j], d);
            lw[j] = logf(pi[j]) + logf(vmf_norm(kap[j])) + kap[j] * (cs[j] - 1.0f);
            lmax = fmaxf(lmax, lw[j]);
        }
        lmax = fmaxf(lmax, __shfl_xor(lmax, 16));
        lmax = fmaxf(lmax, __shfl_xor(lmax, 32));
        float gs = expf(lw[0] - lmax) + expf(lw[1] - lmax);
        gs += __shfl_xor(gs, 16);
        gs += __shfl_xor(gs, 32);
        const float sc = w * invN;
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int k = k0 + j;
            if (w != 0.0f) {
                const float gam = expf(lw[j] - lmax) / gs;
                const float e2 = expf(-2.0f * kap[j]);
                const float dlogv = 1.0f / kap[j] - 2.0f * e2 / (-expm1f(-2.0f * kap[j])) + (cs[j] - 1.0f);
                const float sraw = o[5 * k + 3];
                const bool inside = sraw > kLogKMin && sraw < kLogKMax;
                const float gr = -sc * gam * kap[j] / rn[j];
                grow[5 * k + 0] = to_bf16(gr * (d.x - cs[j] * mu[j].x));
                grow[5 * k + 1] = to_bf16(gr * (d.y - cs[j] * mu[j].y));
                grow[5 * k + 2] = to_bf16(gr * (d.z - cs[j] * mu[j].z));
                grow[5 * k + 3] = to_bf16(inside ? -sc * gam * kap[j] * dlogv : 0.0f);
                grow[5 * k + 4] = to_bf16(sc * (pi[j] - gam));
            } else {
#pragma unroll
                for (int c = 0; c < 5; ++c) grow[5 * k + c] = 0;
            }
        }
        if (grp == 0) sLoss[s] = w != 0.0f ? -w * (lmax + logf(gs)) * invN : 0.0f;
    }
    __syncthreads();
    if (tid < 64) {
        float l = sLoss[tid];
        for (int off = 32; off > 0; off >>= 1) l += __shfl_xor(l, off);
        if (tid == 0) part[kGuideParams] = l;
    }

    // 4. backward, layers 3..0.  dY of layer l is in (l odd ? sG0 : sG1) — l = 3: sG0.
#pragma unroll
    for (int layer = 3; layer >= 0; --layer) {
        uint16_t* dY = (layer & 1) ? sG0 : sG1;
        uint16_t* dX = (layer & 1) ? sG1 : sG0;
        const uint16_t* X = sA[layer];
        const int nout = guide_layer_out(layer);
        const int off = guide_layer_offset(layer);
        // dW[i][o] = sum_s X[s][i] dY[s][o]: wave w computes in-tile it = w for every out tile
        {
            const int it = wave;
            const int nOt = layer == 3 ? 3 : 4;
#pragma unroll
            for (int ot = 0; ot < 4; ++ot) {
                if (ot >= nOt) continue;
                f32x4 acc = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int ks = 0; ks < 2; ++ks) {
                    bf16x8 a, b;
#pragma unroll
                    for (int j = 0; j < 8; ++j) {
                        const int s = ks * 32 + lk + j;
                        uint16_t xa = X[s * kPitch + it * 16 + lr];
                        uint16_t yb = dY[s * kPitch + ot * 16 + lr];
                        a[j] = __builtin_bit_cast(__bf16, xa);
                        b[j] = __builtin_bit_cast(__bf16, yb);
                    }
                    acc = mfma(a, b, acc);
                }
                const int col = ot * 16 + lr;
                if (col < nout) {
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const int in = it * 16 + lq + j;
                        part[off + in * nout + col] = acc[j];
                    }
                }
            }
        }
        // db[o] = sum_s dY[s][o]: four quarter sums of 16 rows per output,
        // then added in quarter order
        {
            const int o = tid & 63, q = tid >> 6;
            float sum = 0.0f;
#pragma unroll
            for (int s = 0; s < kTrainRows / 4; ++s) sum += from_bf16(dY[(q * (kTrainRows / 4) + s) * kPitch + o]);
            sDb[q][o] = sum;
        }
        __syncthreads();
        if (tid < nout) part[off + 64 * nout + tid] = ((sDb[0][tid] + sDb[1][tid]) + sDb[2][tid]) + sDb[3][tid];
        if (layer > 0) {
            // dX[s][i] = (sum_o dY[s][o] W[i][o]) * (X[s][i] > 0), rows of this wave
            const uint16_t* WB = g.wb + layer * 4096;
            f32x4 acc[4];
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) acc[nt] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int ks = 0; ks < 2; ++ks) {
                const bf16x8 a = ld_frag(dY + (wave * 16 + lr) * kPitch + ks * 32 + lk);
#pragma unroll
                for (int nt = 0; nt < 4; ++nt) acc[nt] = mfma(a, ld_frag(WB + (nt * 16 + lr) * 64 + ks * 32 + lk), acc[nt]);
            }
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) {
                const int col = nt * 16 + lr;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int row = wave * 16 + lq + j;
                    const bool on = from_bf16(X[row * kPitch + col]) > 0.0f;
                    dX[row * kPitch + col] = to_bf16(on ? acc[nt][j] : 0.0f);
                }
            }
        }
        __syncthreads();
    }
}

// Sums the partials of the active blocks in block order, then either stores
// the gradient (grad_out) or applies one Adam step (Kingma & Ba 2015,
// Algorithm 1) to the FP32 master weights and refreshes the BF16 copies.
struct AdamArgs {
    float* master;
    float* m;
    float* v;
    uint16_t* wf;
    uint16_t* wb;
    float* bias;
    float lr, b1, b2, eps;
    float c1, c2;   // 1 / (1 - b1^t), 1 / (1 - b2^t)
};

// Stage 1: chunk r of the active blocks' partials, summed in block order.
constexpr int kReduceChunks = 32;
__global__ void k_guide_reduce1(TrainArgs ta, float* __restrict__ chunks) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j > kGuideParams) return;
    const int N = batch_size(ta);
    const int nb = (N + kTrainRows - 1) / kTrainRows;
    const int per = (nb + kReduceChunks - 1) / kReduceChunks;
    const int b0 = blockIdx.y * per, b1 = min(nb, b0 + per);
    float sum = 0.0f;
    for (int b = b0; b < b1; ++b) sum += ta.partial[(size_t)b * kPartStride + j];
    chunks[(size_t)blockIdx.y * kPartStride + j] = sum;
}

// Stage 2: the chunks in order; then either store the gradient (grad_out) or
// apply one Adam step (Kingma & Ba 2015, Algorithm 1) to the FP32 master
// weights and refresh the BF16 copies.
__global__ void k_guide_reduce_adam(TrainArgs ta, AdamArgs ad, const float* __restrict__ chunks,
                                    float* __restrict__ grad_out, float* __restrict__ loss_out) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j > kGuideParams) return;
    const int N = batch_size(ta);
    const int nb = (N + kTrainRows - 1) / kTrainRows;
    float sum = 0.0f;
    for (int r = 0; r < kReduceChunks; ++r) sum += chunks[(size_t)r * kPartStride + j];
    if (j == kGuideParams) {
        if (loss_out) *loss_out = sum;
        return;
    }
    if (grad_out) {
        grad_out[j] = sum;
        return;
    }
    if (nb == 0) return;
    const float gm = ad.b1 * ad.m[j] + (1.0f - ad.b1) * sum;
    const float gv = ad.b2 * ad.v[j] + (1.0f - ad.b2) * sum * sum;
    ad.m[j] = gm;
    ad.v[j] = gv;
    const float w = ad.master[j] - ad.lr * (gm * ad.c1) / (sqrtf(gv * ad.c2) + ad.eps);
    ad.master[j] = w;
    guide_store_param(ad.wf, ad.wb, ad.bias, j, w);
}

__global__ void k_guide_refresh(const float* __restrict__ master, uint16_t* wf, uint16_t* wb, float* bias) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < kGuideParams) guide_store_param(wf, wb, bias, j, master[j]);
}

// ---- training-sample collection ----------------------------------------------------
// Records (4 float4 per recorded path vertex): (pos, valid), (n, lum T_before),
// (wo, lum R_base), (wi, 0).  A sample is emitted when the vertex continued
// the path (valid) and the next vertex stored the radiance reached before it
// (R_base >= 0); its weight is the luminance of the radiance added through wi,
// relative to the throughput before the vertex: (lum R_final - R_base) / lum T.
struct CollectArgs {
    const float4* rec;
    int nslots;        // recorded paths x max_depth
    int max_depth;
    int stride, phase;
    const float4* rad; // final path radiance
    int npaths;
    int* block_counts;
    int* block_offsets;
    int* total;
    float4* samples;
    int max_samples;
    float* block_wsum; // per block: sum of the weights of its samples
    float* wsum;       // sum of the weights of the kept samples
    unsigned long long* seen;  // running count of kept samples (may be NULL)
};

LFH_DEV bool collect_one(const CollectArgs& c, int slot, float4* out, float* wout = nullptr) {
    if (slot >= c.nslots) return false;
    const float4* r = c.rec + (size_t)slot * 4;
    const float4 r0 = r[0];
    if (r0.w != 1.0f) return false;
    const float4 r1 = r[1], r2 = r[2];
    if (r2.w < 0.0f || !(r1.w > 0.0f)) return false;
    const int path = (slot / c.max_depth) * c.stride + c.phase;
    if (path >= c.npaths) return false;
    const float4 f = c.rad[path];
    const float lf = 0.2126f * f.x + 0.7152f * f.y + 0.0722f * f.z;
    const float w = (lf - r2.w) / r1.w;
    if (!(w > 0.0f) || !isfinite(w)) return false;
    if (wout) *wout = w;
    if (out) {
        const float4 r3 = r[3];
        out[0] = make_float4(r0.x, r0.y, r0.z, w);
        out[1] = make_float4(r1.x, r1.y, r1.z, 0.0f);
        out[2] = make_float4(r2.x, r2.y, r2.z, 0.0f);
        out[3] = make_float4(r3.x, r3.y, r3.z, 0.0f);
    }
    return true;
}

__global__ __launch_bounds__(256) void k_guide_count(CollectArgs c) {
    __shared__ float sw[256];
    const int slot = blockIdx.x * 256 + threadIdx.x;
    float w = 0.0f;
    const int ok = collect_one(c, slot, nullptr, &w) ? 1 : 0;
    sw[threadIdx.x] = ok ? w : 0.0f;
    const int cnt = __syncthreads_count(ok);
    for (int off = 128; off > 0; off >>= 1) {   // fixed-order tree sum
        if (threadIdx.x < off) sw[threadIdx.x] += sw[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        c.block_counts[blockIdx.x] = cnt;
        c.block_wsum[blockIdx.x] = sw[0];
    }
}

// Exclusive scan of the block counts (one block of 1024 threads) and the
// fixed-order sum of the block weight sums over the blocks whose samples are
// all kept.
__global__ __launch_bounds__(1024) void k_guide_scan(CollectArgs c, int nblocks) {
    __shared__ int s[1024];
    __shared__ float sws[1024];
    const int t = threadIdx.x;
    const int per = (nblocks + 1023) / 1024;
    const int b0 = t * per;
    int local = 0;
    for (int b = b0; b < b0 + per && b < nblocks; ++b) local += c.block_counts[b];
    s[t] = local;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const int v = t >= off ? s[t - off] : 0;
        __syncthreads();
        s[t] += v;
        __syncthreads();
    }
    int run = s[t] - local;
    float wl = 0.0f;
    for (int b = b0; b < b0 + per && b < nblocks; ++b) {
        c.block_offsets[b] = run;
        if (run + c.block_counts[b] <= c.max_samples) wl += c.block_wsum[b];
        run += c.block_counts[b];
    }
    sws[t] = wl;
    __syncthreads();
    for (int off = 512; off > 0; off >>= 1) {
        if (t < off) sws[t] += sws[t + off];
        __syncthreads();
    }
    if (t == 1023) {
        const int kept = s[1023] < c.max_samples ? s[1023] : c.max_samples;
        *c.total = kept;
        if (c.seen) *c.seen += (unsigned long long)kept;
    }
    if (t == 0) *c.wsum = sws[0];
}

__global__ __launch_bounds__(256) void k_guide_compact(CollectArgs c) {
    __shared__ int sw[4];
    const int slot = blockIdx.x * 256 + threadIdx.x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const bool ok = collect_one(c, slot, nullptr);
    const uint64_t m = __ballot(ok);
    if (lane == 0) sw[wave] = __popcll(m);
    __syncthreads();
    int off = c.block_offsets[blockIdx.x];
    for (int w = 0; w < wave; ++w) off += sw[w];
    off += __popcll(m & ((1ull << lane) - 1ull));
    if (ok && off < c.max_samples) collect_one(c, slot, c.samples + (size_t)off * 4);
}

// ---- launchers ------------------------------------------------------------------------
static inline int nblk(int64_t n, int b) { return (int)((n + b - 1) / b); }

// Blocks of a persistent inference launch for n queries: what the chip holds
// at once, at most one block per 256 queries.
static int infer_blocks(int64_t n) {
    static int resident = 0;
    if (resident == 0) {
        int dev = 0, per_cu = 0, cus = 256;
        hipDeviceProp_t prop;
        if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
            cus = prop.multiProcessorCount;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k_guide_infer<true>, kInferBlock, 0) !=
                hipSuccess || per_cu <= 0)
            per_cu = 2;
        resident = per_cu * cus;
    }
    return (int)std::min<int64_t>(resident, nblk(n, kInferBlock));
}

hipError_t guide_forward_host(const GuideDev& g, const float* in9, int n, float* raw40, float4* lobes,
                              hipStream_t s) {
    if (n <= 0) return hipSuccess;
    InferQueue q{};
    hipLaunchKernelGGL(k_guide_infer<false>, dim3(infer_blocks(n)), dim3(kInferBlock), 0, s, g, in9, n, q, raw40,
                       lobes);
    return hipGetLastError();
}

hipError_t guide_infer_queue(const GuideDev& g, const ShadeScene& sc, const float4* rayq, const float4* hitq,
                             const uint32_t* count, int maxq, const GuideHook& gh, int depth, float4* lobes,
                             hipStream_t s) {
    if (maxq <= 0) return hipSuccess;
    InferQueue q{sc, rayq, hitq, count, gh, depth};
    if (!gh.sample) {
        if (gh.train) hipLaunchKernelGGL(k_guide_record, dim3(nblk(maxq, 256)), dim3(256), 0, s, q);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(k_guide_infer<true>, dim3(infer_blocks(maxq)), dim3(kInferBlock), 0, s, g,
                       (const float*)nullptr, 0, q, (float*)nullptr, lobes);
    return hipGetLastError();
}

hipError_t guide_sample(const float4* lobes, const uint4* seeds, int n, float* dirs, float* pdf, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_guide_sample, dim3(nblk(n, 256)), dim3(256), 0, s, lobes, seeds, n, dirs, pdf);
    return hipGetLastError();
}

hipError_t guide_pdf(const float4* lobes, const float* dirs, int n, float* pdf, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_guide_pdf, dim3(nblk(n, 256)), dim3(256), 0, s, lobes, dirs, n, pdf);
    return hipGetLastError();
}

int guide_train_blocks(int batch) { return nblk(batch, kTrainRows); }
// per-block partials, then the stage-1 chunks
size_t guide_partial_floats(int batch) { return (size_t)(guide_train_blocks(batch) + kReduceChunks) * kPartStride; }

hipError_t guide_train(const GuideDev& g, const float4* samples, const int* count, int start, int batch,
                       float* partial, const GuideAdam& adam, float* grad_out, float* loss_out, hipStream_t s,
                       const float* wsum, float wcap) {
    TrainArgs ta{samples, count, start, batch, partial, wsum, wcap};
    hipLaunchKernelGGL(k_guide_train, dim3(guide_train_blocks(batch)), dim3(kTrainBlock), 0, s, g, ta);
    AdamArgs ad{adam.master, adam.m, adam.v, adam.wf, adam.wb, adam.bias, adam.lr, adam.b1, adam.b2, adam.eps,
                adam.c1, adam.c2};
    float* chunks = partial + (size_t)guide_train_blocks(batch) * kPartStride;
    hipLaunchKernelGGL(k_guide_reduce1, dim3(nblk(kGuideParams + 1, 256), kReduceChunks), dim3(256), 0, s, ta,
                       chunks);
    hipLaunchKernelGGL(k_guide_reduce_adam, dim3(nblk(kGuideParams + 1, 256)), dim3(256), 0, s, ta, ad,
                       (const float*)chunks, grad_out, loss_out);
    return hipGetLastError();
}

hipError_t guide_refresh(const float* master, uint16_t* wf, uint16_t* wb, float* bias, hipStream_t s) {
    hipLaunchKernelGGL(k_guide_refresh, dim3(nblk(kGuideParams, 256)), dim3(256), 0, s, master, wf, wb, bias);
    return hipGetLastError();
}

hipError_t guide_collect(const float4* rec, int nslots, int max_depth, int stride, int phase, const float4* rad,
                         int npaths, int* block_counts, int* block_offsets, int* total, float4* samples,
                         int max_samples, float* block_wsum, float* wsum, unsigned long long* seen,
                         hipStream_t s) {
    CollectArgs c{rec, nslots, max_depth, stride, phase, rad, npaths, block_counts, block_offsets, total, samples,
                  max_samples, block_wsum, wsum, seen};
    const int nb = nblk(nslots, 256);
    if (nb <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_guide_count, dim3(nb), dim3(256), 0, s, c);
    hipLaunchKernelGGL(k_guide_scan, dim3(1), dim3(1024), 0, s, c, nb);
    hipLaunchKernelGGL(k_guide_compact, dim3(nb), dim3(256), 0, s, c);
    return hipGetLastError();
}

int guide_collect_blocks(int nslots) { return nblk(nslots, 256); }

}  // namespace lfhd
