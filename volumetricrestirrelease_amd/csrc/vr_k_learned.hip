// vr_k_learned.hip -- the learned free-flight proposal's network on gfx950:
// BF16 MFMA inference for every camera ray (its 16-bit CDF over the analytic
// tracker's CDF coordinate), the cross-entropy gradient of the training rays,
// Adam on the FP32 master weights, and the packing of those weights into MFMA
// operand fragments. No counterpart in the reference (see vr_learned.hpp).
//
// Matrix layout. Every product runs on v_mfma_f32_32x32x16_bf16 with the rays
// as the N dimension: a layer computes H^T (features x rays) = W^T (features x
// inputs) . X^T (inputs x rays), so a wave holds 32 rays, lane l owning ray
// l % 32, and the accumulator of a 32-feature tile leaves in lane l register r
// the feature phi(r, l / 32) = 8 (r / 4) + 4 (l / 32) + r % 4. The next layer
// reads those registers directly as its B operand: B element j of k-step s in
// lane half h is the accumulator's register 8 (s % 2) + j, i.e. feature
// perm(s, h, j) = 16 s + 8 (j / 4) + 4 h + j % 4. Packing W with that same
// permutation along K (kLearnedPack) makes the layers chain with no data
// movement. Backward products dH = W . dZ chain the same way; the weight
// gradients, which sum over rays (K = rays), read activations and dZ from LDS
// stored feature-major.
#include <hip/hip_runtime.h>

#include "vr_launch.h"
#include "vr_learned.hpp"

namespace vr {
namespace learned {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// Fragment table: forward A operands (W^T, K permuted) of layers 1-4, backward
// A operands (W, K permuted) of layers 4, 3, 2; 64 lanes x 8 bf16 each.
constexpr int kFragL1 = 0, kFragL2 = 6, kFragL3 = 14, kFragL4 = 22, kFragB4 = 26, kFragB3 = 30, kFragB2 = 38;
static_assert(kFragCount == 46, "fragment count");
constexpr int kStatLoss = kParams, kStatPlain = kParams + 1, kStatNet = kParams + 2;

VR_DEV int phi(int r, int h) { return 8 * (r >> 2) + 4 * h + (r & 3); }
VR_DEV int perm(int s, int h, int j) { return 16 * s + 8 * (j >> 2) + 4 * h + (j & 3); }

VR_DEV f32x16 mfma(bf16x8 a, bf16x8 b, f32x16 c) { return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0); }
VR_DEV f32x16 zero16() {
    f32x16 z;
#pragma unroll
    for (int i = 0; i < 16; i++) z[i] = 0.f;
    return z;
}
VR_DEV bf16x8 loadFrag(const uint4* frags, int f, int lane) { return __builtin_bit_cast(bf16x8, frags[f * 64 + lane]); }
// B operand of k-step q (0, 1) of one 32-feature accumulator tile.
VR_DEV bf16x8 operandOf(const f32x16& acc, int q) {
    bf16x8 b;
#pragma unroll
    for (int j = 0; j < 8; j++) b[j] = (__bf16)acc[8 * q + j];
    return b;
}

// Frequency encoding (sin / cos of pi 2^k x): position 3 x 3 octaves, direction
// 3 x 2, length 3; feature f as listed in order.
VR_DEV void encode(const float* raw, float* enc) {
    int n = 0;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const float a = raw[i] * (float)(1 << k);
            enc[n++] = sinpif(a);
            enc[n++] = cospif(a);
        }
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int k = 0; k < 2; k++) {
            const float a = raw[3 + i] * (float)(1 << k);
            enc[n++] = sinpif(a);
            enc[n++] = cospif(a);
        }
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const float a = raw[6] * (float)(1 << k);
        enc[n++] = sinpif(a);
        enc[n++] = cospif(a);
    }
}
// Layer-1 B operand of k-step s: inputs 16 s + 8 h + j (zero past 36).
VR_DEV bf16x8 inputOperand(const float* enc, int s, int h) {
    bf16x8 b;
#pragma unroll
    for (int j = 0; j < 8; j++) {
        const int f0 = 16 * s + j, f1 = 16 * s + 8 + j;
        const float v0 = f0 < kIn ? enc[f0] : 0.f, v1 = f1 < kIn ? enc[f1] : 0.f;
        b[j] = (__bf16)(h ? v1 : v0);
    }
    return b;
}

// acc + bias, then ReLU (hidden layers), for tile mt of a layer whose biases start at bias.
VR_DEV void biasRelu(f32x16& acc, const float* bias, int mt, int h, bool relu) {
#pragma unroll
    for (int r = 0; r < 16; r++) {
        float v = acc[r] + bias[32 * mt + phi(r, h)];
        acc[r] = relu ? fmaxf(v, 0.f) : v;
    }
}

// The forward pass of the 32 rays of a wave from their B operands of layer 1;
// returns the logits tile. Optionally stores each hidden layer's bf16
// activations feature-major in LDS (act[layer][feature * kLds + ray]).
constexpr int kLds = 40;  // row stride (bf16) of the feature-major LDS tiles: 32 rays + padding
template <bool STORE>
VR_DEV f32x16 forwardWave(const bf16x8 x[3], const uint4* frags, const float* prm, int lane, __bf16* act) {
    const int h = lane >> 5, n = lane & 31;
    f32x16 a0 = zero16(), a1 = zero16();
#pragma unroll
    for (int s = 0; s < 3; s++) {
        a0 = mfma(loadFrag(frags, kFragL1 + s, lane), x[s], a0);
        a1 = mfma(loadFrag(frags, kFragL1 + 3 + s, lane), x[s], a1);
    }
    biasRelu(a0, prm + kB1, 0, h, true);
    biasRelu(a1, prm + kB1, 1, h, true);
    const int fragL[2] = {kFragL2, kFragL3};
    const int biasL[2] = {kB2, kB3};
#pragma unroll
    for (int L = 0; L < 2; L++) {
        if (STORE) {
#pragma unroll
            for (int r = 0; r < 16; r++) {
                act[(L * 64 + phi(r, h)) * kLds + n] = (__bf16)a0[r];
                act[(L * 64 + 32 + phi(r, h)) * kLds + n] = (__bf16)a1[r];
            }
        }
        bf16x8 b[4] = {operandOf(a0, 0), operandOf(a0, 1), operandOf(a1, 0), operandOf(a1, 1)};
        f32x16 c0 = zero16(), c1 = zero16();
#pragma unroll
        for (int s = 0; s < 4; s++) {
            c0 = mfma(loadFrag(frags, fragL[L] + s, lane), b[s], c0);
            c1 = mfma(loadFrag(frags, fragL[L] + 4 + s, lane), b[s], c1);
        }
        biasRelu(c0, prm + biasL[L], 0, h, true);
        biasRelu(c1, prm + biasL[L], 1, h, true);
        a0 = c0;
        a1 = c1;
    }
    if (STORE) {
#pragma unroll
        for (int r = 0; r < 16; r++) {
            act[(128 + phi(r, h)) * kLds + n] = (__bf16)a0[r];
            act[(128 + 32 + phi(r, h)) * kLds + n] = (__bf16)a1[r];
        }
    }
    bf16x8 b[4] = {operandOf(a0, 0), operandOf(a0, 1), operandOf(a1, 0), operandOf(a1, 1)};
    f32x16 z = zero16();
#pragma unroll
    for (int s = 0; s < 4; s++) z = mfma(loadFrag(frags, kFragL4 + s, lane), b[s], z);
    biasRelu(z, prm + kB4, 0, h, false);
    return z;
}

// Softmax over the 32 logits of a ray (16 in each lane of the pair l, l ^ 32)
// and the mixture p = alpha / 32 + (1 - alpha) q.
VR_DEV void mixture(const f32x16& z, float alpha, float q[16], float p[16]) {
    float m = z[0];
#pragma unroll
    for (int r = 1; r < 16; r++) m = fmaxf(m, z[r]);
    m = fmaxf(m, __shfl_xor(m, 32));
    float s = 0.f;
#pragma unroll
    for (int r = 0; r < 16; r++) { q[r] = expf(z[r] - m); s += q[r]; }
    s += __shfl_xor(s, 32);
#pragma unroll
    for (int r = 0; r < 16; r++) {
        q[r] = q[r] / s;
        p[r] = alpha / (float)kOut + (1.f - alpha) * q[r];
    }
}

// ------------------------------------------------------------------ inference
// One wave per 32 rays, looping over chunks of 32 rays; SRC 0: the frame's
// camera rays (pixel = ray index), 1: raw inputs from rawIn. Writes each ray's
// 16-bit CDF (half h of the lane pair writes cdf[16 h .. 16 h + 15]) and
// optionally its logits. inUse == 0: the uniform CDF (the analytic proposal).
template <int SRC>
__global__ __launch_bounds__(64) void kLearnedInfer(const Scene* __restrict__ sc, const Frame fr,
                                                    const float* __restrict__ rawIn, int n,
                                                    const uint4* __restrict__ frags, const float* __restrict__ prm,
                                                    const int* __restrict__ inUse, float alpha,
                                                    uint16_t* __restrict__ cdfOut, float* __restrict__ logitsOut) {
    const int lane = threadIdx.x, h = lane >> 5, nl = lane & 31;
    const bool use = inUse == nullptr || *inUse != 0;
    for (int base = blockIdx.x * 32; base < n; base += gridDim.x * 32) {
        const int ray = base + nl;
        const bool valid = ray < n;
        uint32_t packed[8];
        if (!use) {
#pragma unroll
            for (int i = 0; i < 8; i++) {
                const uint32_t j = 16 * h + 2 * i;
                packed[i] = (2048u * j) | ((2048u * (j + 1)) << 16);
            }
        } else {
            float raw[kRaw] = {0.5f, 0.5f, 0.5f, 0.f, 0.f, 1.f, 0.f, 0.f};
            if (valid) {
                if (SRC == 0) {
                    const Scene& S = *sc;
                    const int px = ray % fr.width, py = ray / fr.width;
                    Ray r{ld3(fr.cam.pos), rayDir(fr.cam, px, py, fr.width, fr.height), 0.f, kTMax};
                    rawInputs(S, r, raw);
                } else {
#pragma unroll
                    for (int i = 0; i < kRaw; i++) raw[i] = rawIn[(size_t)ray * kRaw + i];
                }
            }
            float enc[kIn];
            encode(raw, enc);
            const bf16x8 x[3] = {inputOperand(enc, 0, h), inputOperand(enc, 1, h), inputOperand(enc, 2, h)};
            f32x16 z = forwardWave<false>(x, frags, prm, lane, nullptr);
            if (logitsOut && valid) {
#pragma unroll
                for (int r = 0; r < 16; r++) logitsOut[(size_t)ray * kOut + phi(r, h)] = z[r];
            }
            float q[16], p[16];
            mixture(z, alpha, q, p);
            // n_j = round(p_j 2^16) (at least 1) for j < 31, the last bin the rest;
            // the pair exchanges its counts so each lane has all 32 in order
            int cnt[16], other[16];
#pragma unroll
            for (int r = 0; r < 16; r++) cnt[r] = max(1, (int)floorf(p[r] * 65536.f + 0.5f));
#pragma unroll
            for (int r = 0; r < 16; r++) other[r] = __shfl_xor(cnt[r], 32);
            int all[kOut];
#pragma unroll
            for (int r = 0; r < 16; r++) {
                all[phi(r, h)] = cnt[r];
                all[phi(r, 1 - h)] = other[r];
            }
            int acc = 0, cdf[kOut];
#pragma unroll
            for (int j = 0; j < kOut; j++) {
                cdf[j] = acc;
                int c = j < kOut - 1 ? all[j] : 65536 - acc;
                acc += max(c, 1);
            }
#pragma unroll
            for (int i = 0; i < 8; i++) {
                const int j = 16 * h + 2 * i;
                const int lo = h ? cdf[16 + 2 * i] : cdf[2 * i], hi = h ? cdf[17 + 2 * i] : cdf[2 * i + 1];
                (void)j;
                packed[i] = (uint32_t)lo | ((uint32_t)hi << 16);
            }
        }
        if (valid) {
            uint4* o = reinterpret_cast<uint4*>(cdfOut + (size_t)ray * kOut + 16 * h);
            o[0] = make_uint4(packed[0], packed[1], packed[2], packed[3]);
            o[1] = make_uint4(packed[4], packed[5], packed[6], packed[7]);
        }
    }
}

// ------------------------------------------------------------------ gradient
// One wave per 32 training records: forward (activations to LDS), the
// cross-entropy's logit gradient, backward through the layers, and the
// weight gradients as MFMA products over the rays. Each wave writes its own
// partial (kPartialStride floats: gradient, scale x loss, the two second
// moments), summed in a fixed order by the reduction, so results do not depend
// on scheduling.
VR_DEV void storeTile(__bf16* dst, const f32x16& v, int row0, int h, int n) {
#pragma unroll
    for (int r = 0; r < 16; r++) dst[(row0 + phi(r, h)) * kLds + n] = (__bf16)v[r];
}
// A operand (rows row0 + l % 32, rays 16 s + 8 h ..) or B operand (columns
// row0 + l % 32, same rays) of a product over rays, from a feature-major tile.
VR_DEV bf16x8 rayOperand(const __bf16* src, int row0, int s, int lane) {
    const int h = lane >> 5, i = lane & 31;
    return *reinterpret_cast<const bf16x8*>(src + (row0 + i) * kLds + 16 * s + 8 * h);
}
// dW[in][out] (+ rows of tile mt, columns of tile nt) over the wave's 32 rays.
VR_DEV void weightGrad(const __bf16* actIn, const __bf16* dz, int mtiles, int ntiles, int nout, int nin, float* out,
                       int lane) {
    const int h = lane >> 5, col = lane & 31;
    for (int mt = 0; mt < mtiles; mt++)
        for (int nt = 0; nt < ntiles; nt++) {
            f32x16 c = zero16();
#pragma unroll
            for (int s = 0; s < 2; s++) c = mfma(rayOperand(actIn, 32 * mt, s, lane), rayOperand(dz, 32 * nt, s, lane), c);
#pragma unroll
            for (int r = 0; r < 16; r++) {
                const int row = 32 * mt + phi(r, h);
                if (row < nin) out[row * nout + 32 * nt + col] = c[r];
            }
        }
}
// db[o] = sum over rays of dz[o] (the bf16 values the products used).
VR_DEV void biasGrad(const __bf16* dz, int nout, float* out, int lane) {
    if (lane < nout) {
        float s = 0.f;
        for (int k = 0; k < 32; k++) s += (float)dz[lane * kLds + k];
        out[lane] = s;
    }
}

__global__ __launch_bounds__(64) void kLearnedGrad(const float* __restrict__ recs, int n, float alpha, float scale,
                                                   const uint4* __restrict__ frags, const float* __restrict__ prm,
                                                   float* __restrict__ partial) {
    __shared__ __align__(16) __bf16 act[4 * 64 * kLds];  // inputs, h1, h2, h3 (feature-major)
    __shared__ __align__(16) __bf16 dz[64 * kLds];
    const int lane = threadIdx.x, h = lane >> 5, nl = lane & 31;
    const int ray = blockIdx.x * 32 + nl;
    const float* rec = recs + (size_t)min(ray, n - 1) * kRecFloats;
    const float W = rec[7];
    const bool valid = ray < n && W > 0.f && isfinite(W);
    float* out = partial + (size_t)blockIdx.x * kPartialStride;
    // inputs
    float raw[kRaw];
#pragma unroll
    for (int i = 0; i < kRaw; i++) raw[i] = rec[i];
    float enc[kIn];
    encode(raw, enc);
    bf16x8 x[3] = {inputOperand(enc, 0, h), inputOperand(enc, 1, h), inputOperand(enc, 2, h)};
    {
        // feature-major copy of the inputs (rows 36..63 zero)
        __bf16* xin = act;
#pragma unroll
        for (int s = 0; s < 3; s++)
#pragma unroll
            for (int j = 0; j < 8; j++) xin[(16 * s + 8 * h + j) * kLds + nl] = x[s][j];
        for (int f = 48 + h; f < 64; f += 2)
            for (int k = 0; k < 1; k++) xin[f * kLds + nl] = (__bf16)0.f;
    }
    f32x16 z = forwardWave<true>(x, frags, prm, lane, act + 64 * kLds);
    // logit gradient: dz_j = scale (1 - alpha) (q_j R - r_j), r_j = h_j q_j / p_j
    float q[16], p[16];
    mixture(z, alpha, q, p);
    float hist[16], rr[16], loss = 0.f, sPlain = 0.f, sNet = 0.f, R = 0.f;
    const float invW = valid ? 1.f / W : 0.f;
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const int j = phi(r, h);
        hist[r] = rec[kRecHist + j] * invW;
        const float sq = rec[kRecSq + j] * invW * invW;
        loss -= hist[r] * logf((float)kOut * p[r]);
        sPlain += sq;
        sNet += sq / ((float)kOut * p[r]);
        rr[r] = hist[r] * q[r] / p[r];
        R += rr[r];
    }
    R += __shfl_xor(R, 32);
    f32x16 d4;
#pragma unroll
    for (int r = 0; r < 16; r++) d4[r] = valid ? scale * (1.f - alpha) * (q[r] * R - rr[r]) : 0.f;
    // per-wave statistics (ray sums; lane pairs hold the two halves of a ray)
    float st[3] = {valid ? scale * loss : 0.f, valid ? sPlain : 0.f, valid ? sNet : 0.f};
#pragma unroll
    for (int k = 0; k < 3; k++) {
        float v = st[k];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
        st[k] = v;
    }
    if (lane == 0) {
        out[kStatLoss] = st[0];
        out[kStatPlain] = st[1];
        out[kStatNet] = st[2];
    }
    // layer 4
    storeTile(dz, d4, 0, h, nl);
    __syncthreads();
    const __bf16* h1 = act + 64 * kLds;
    const __bf16* h2 = act + 128 * kLds;
    const __bf16* h3 = act + 192 * kLds;
    weightGrad(h3, dz, 2, 1, kOut, kHid, out + kW4, lane);
    biasGrad(dz, kOut, out + kB4, lane);
    // dH3 = W4 dz4, masked by h3 > 0
    bf16x8 bd[4] = {operandOf(d4, 0), operandOf(d4, 1), {}, {}};
    f32x16 e0 = zero16(), e1 = zero16();
#pragma unroll
    for (int s = 0; s < 2; s++) {
        e0 = mfma(loadFrag(frags, kFragB4 + s, lane), bd[s], e0);
        e1 = mfma(loadFrag(frags, kFragB4 + 2 + s, lane), bd[s], e1);
    }
    const __bf16* acts[3] = {h1, h2, h3};
    const int fragB[2] = {kFragB3, kFragB2};
    const int gW[3] = {kW1, kW2, kW3}, gB[3] = {kB1, kB2, kB3};
    for (int L = 2; L >= 0; L--) {
        // dz of layer L + 1 (hidden L): mask by the layer's activation
        const __bf16* a = acts[L];
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const int f0 = phi(r, h), f1 = 32 + phi(r, h);
            if (!((float)a[f0 * kLds + nl] > 0.f)) e0[r] = 0.f;
            if (!((float)a[f1 * kLds + nl] > 0.f)) e1[r] = 0.f;
        }
        __syncthreads();  // every lane has read dz of the layer above
        storeTile(dz, e0, 0, h, nl);
        storeTile(dz, e1, 32, h, nl);
        __syncthreads();
        const __bf16* ain = L == 0 ? act : acts[L - 1];
        weightGrad(ain, dz, 2, 2, kHid, L == 0 ? kIn : kHid, out + gW[L], lane);
        biasGrad(dz, kHid, out + gB[L], lane);
        if (L > 0) {
            bf16x8 b[4] = {operandOf(e0, 0), operandOf(e0, 1), operandOf(e1, 0), operandOf(e1, 1)};
            f32x16 c0 = zero16(), c1 = zero16();
#pragma unroll
            for (int s = 0; s < 4; s++) {
                c0 = mfma(loadFrag(frags, fragB[2 - L] + s, lane), b[s], c0);
                c1 = mfma(loadFrag(frags, fragB[2 - L] + 4 + s, lane), b[s], c1);
            }
            e0 = c0;
            e1 = c1;
        }
    }
}

// Sum of the waves' partials in wave order; with adam != 0 also the Adam step on
// the master weights (bias corrections c1 = 1 - b1^t, c2 = 1 - b2^t from the
// host), and, in thread 0, the statistics and the judge's decision.
__global__ __launch_bounds__(256) void kLearnedReduce(const float* __restrict__ partial, int waves,
                                                      float* __restrict__ out, int adam, float* __restrict__ prm,
                                                      float* __restrict__ m, float* __restrict__ v, float lr, float c1,
                                                      float c2, float* __restrict__ stats, int* __restrict__ inUse,
                                                      int judge, float step) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < kParams + 3) {
        float g = 0.f;
        for (int w = 0; w < waves; w++) g += partial[(size_t)w * kPartialStride + i];
        if (out) out[i] = g;
        if (adam && i < kParams) {
            const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
            const float mi = b1 * m[i] + (1.f - b1) * g;
            const float vi = b2 * v[i] + (1.f - b2) * g * g;
            m[i] = mi;
            v[i] = vi;
            const float mh = mi / c1, vh = vi / c2;
            prm[i] -= lr * mh / (sqrtf(vh) + eps);
        }
        if (adam && i >= kParams) stats[i - kParams] = g;
    }
    if (adam && i == 0) {
        // the judge reads the statistics this kernel's threads kParams.. write;
        // thread 0 recomputes them instead of waiting on another workgroup
        float s[3] = {0.f, 0.f, 0.f};
        for (int w = 0; w < waves; w++)
            for (int k = 0; k < 3; k++) s[k] += partial[(size_t)w * kPartialStride + kParams + k];
        const int use = !judge || s[2] < s[1];
        *inUse = use;
        stats[3] = use ? 1.f : 0.f;
        stats[4] = step;
    }
}

// W (FP32 master) -> bf16 MFMA fragments: 46 x 64 lanes x 8 values.
__global__ __launch_bounds__(256) void kLearnedPack(const float* __restrict__ prm, uint4* __restrict__ frags) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= kFragCount * 64) return;
    const int f = t / 64, lane = t % 64, i = lane & 31, h = lane >> 5;
    bf16x8 b;
#pragma unroll
    for (int j = 0; j < 8; j++) {
        float val;
        if (f < kFragL2) {  // W1^T: rows = outputs, K = inputs 16 s + 8 h + j
            const int mt = f / 3, s = f % 3, in = 16 * s + 8 * h + j, o = 32 * mt + i;
            val = in < kIn ? prm[kW1 + in * kHid + o] : 0.f;
        } else if (f < kFragL4) {  // W2^T, W3^T with K permuted
            const int g = (f - kFragL2) % 8, mt = g / 4, s = g % 4;
            const int base = f < kFragL3 ? kW2 : kW3;
            val = prm[base + perm(s, h, j) * kHid + 32 * mt + i];
        } else if (f < kFragB4) {  // W4^T
            const int s = f - kFragL4;
            val = prm[kW4 + perm(s, h, j) * kOut + i];
        } else if (f < kFragB3) {  // W4: rows = hidden, K = logits permuted
            const int g = f - kFragB4, mt = g / 2, s = g % 2;
            val = prm[kW4 + (32 * mt + i) * kOut + perm(s, h, j)];
        } else {  // W3, W2: rows = inputs of the layer, K = outputs permuted
            const int g = (f - kFragB3) % 8, mt = g / 4, s = g % 4;
            const int base = f < kFragB2 ? kW3 : kW2;
            val = prm[base + (32 * mt + i) * kHid + perm(s, h, j)];
        }
        b[j] = (__bf16)val;
    }
    frags[t] = __builtin_bit_cast(uint4, b);
}

}  // namespace learned

// ------------------------------------------------------------------ launchers
hipError_t launchLearnedInfer(hipStream_t st, const Scene* scene, const Frame& frame, const float* rawIn, int n,
                              const uint4* frags, const float* prm, const int* inUse, float alpha, uint16_t* cdf,
                              float* logits) {
    if (n <= 0) return hipSuccess;
    const int chunks = (n + 31) / 32;
    const int grid = chunks < 8192 ? chunks : 8192;
    if (rawIn)
        hipLaunchKernelGGL(learned::kLearnedInfer<1>, dim3(grid), dim3(64), 0, st, scene, frame, rawIn, n, frags, prm,
                           inUse, alpha, cdf, logits);
    else
        hipLaunchKernelGGL(learned::kLearnedInfer<0>, dim3(grid), dim3(64), 0, st, scene, frame, rawIn, n, frags, prm,
                           inUse, alpha, cdf, logits);
    return hipGetLastError();
}

hipError_t launchLearnedGrad(hipStream_t st, const float* recs, int n, float alpha, float scale, const uint4* frags,
                             const float* prm, float* partial) {
    if (n <= 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(learned::kLearnedGrad, dim3((n + 31) / 32), dim3(64), 0, st, recs, n, alpha, scale, frags, prm,
                       partial);
    return hipGetLastError();
}

hipError_t launchLearnedReduce(hipStream_t st, const float* partial, int waves, float* out, bool adam, float* prm,
                               float* m, float* v, float lr, float c1, float c2, float* stats, int* inUse, bool judge,
                               int step) {
    const int total = learned::kParams + 3;
    hipLaunchKernelGGL(learned::kLearnedReduce, dim3((total + 255) / 256), dim3(256), 0, st, partial, waves, out,
                       adam ? 1 : 0, prm, m, v, lr, c1, c2, stats, inUse, judge ? 1 : 0, (float)step);
    return hipGetLastError();
}

hipError_t launchLearnedPack(hipStream_t st, const float* prm, uint4* frags) {
    hipLaunchKernelGGL(learned::kLearnedPack, dim3((learned::kFragCount * 64 + 255) / 256), dim3(256), 0, st, prm,
                       frags);
    return hipGetLastError();
}

}  // namespace vr
