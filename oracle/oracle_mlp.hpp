// oracle_mlp.hpp -- TEST INFRASTRUCTURE ONLY (the CPU oracle).
//
// Scalar restatement of the learned free-flight proposal: the frequency
// encoding, the 36 -> 64 -> 64 -> 64 -> 32 ReLU network, the 32-bin mixture
// pdf over the analytic tracker's CDF coordinate c = 1 - exp(-tau), its 16-bit
// CDF, the cross-entropy gradient of one training ray and the Adam step. The
// learned proposal has no counterpart in the reference: it replaces the
// distance pdf of the initial candidates (pdfDist, VR/ComputeInitialSample.slang:
// 86,107), drawn by SampleMediumAnalyticAdapterGVDB
// (VR/VolumeTrackingAdapterGVDB.slang:212-436).
//
// Two precisions: T = double without rounding is the mathematical network
// (pinned by finite differences and closed forms in tests/); T = float with
// bf16 rounding of every matrix operand reproduces the rounding points of the
// HIP kernels (bf16 operands, FP32 accumulation). Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg load the oracle.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>

namespace orc {
namespace mlp {

constexpr int kIn = 36, kHid = 64, kOut = 32;
// Parameter layout (FP32 master copy and exchange format): W1[36][64], b1[64],
// W2[64][64], b2[64], W3[64][64], b3[64], W4[64][32], b4[32]; W[i][o] multiplies
// input i into output o.
constexpr int kW1 = 0, kB1 = kW1 + kIn * kHid, kW2 = kB1 + kHid, kB2 = kW2 + kHid * kHid, kW3 = kB2 + kHid,
              kB3 = kW3 + kHid * kHid, kW4 = kB3 + kHid, kB4 = kW4 + kHid * kOut, kParams = kB4 + kOut;
static_assert(kParams == 12768, "parameter count");
// Raw inputs of one ray (7 used): entry point in the volume's normalised box
// [0,1]^3, unit direction, segment length / box diagonal.
constexpr int kRaw = 8;
// Training record of one ray: raw inputs (7) + the ray's summed RIS weight W,
// then per CDF bin the summed weights w and the summed w^2 * g_used.
constexpr int kRecHist = 8, kRecSq = kRecHist + kOut, kRecFloats = kRecSq + kOut;

// bf16 round to nearest even (as v_cvt_pk_bf16_f32 on gfx950).
static inline float bf16(float v) {
    if (std::isnan(v)) return v;
    uint32_t u;
    std::memcpy(&u, &v, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    u &= 0xFFFF0000u;
    float r;
    std::memcpy(&r, &u, 4);
    return r;
}

// Frequency encoding: position e_i (3 axes x 3 octaves x sin/cos), direction
// d_i (3 x 2 octaves x sin/cos), length L (3 octaves x sin/cos), each as
// sin/cos(pi 2^k x); computed in double and rounded once to float.
static inline void encode(const float* raw, float* out) {
    const double pi = 3.14159265358979323846;
    int n = 0;
    for (int i = 0; i < 3; i++)
        for (int k = 0; k < 3; k++) {
            double a = pi * std::ldexp((double)raw[i], k);
            out[n++] = (float)std::sin(a);
            out[n++] = (float)std::cos(a);
        }
    for (int i = 0; i < 3; i++)
        for (int k = 0; k < 2; k++) {
            double a = pi * std::ldexp((double)raw[3 + i], k);
            out[n++] = (float)std::sin(a);
            out[n++] = (float)std::cos(a);
        }
    for (int k = 0; k < 3; k++) {
        double a = pi * std::ldexp((double)raw[6], k);
        out[n++] = (float)std::sin(a);
        out[n++] = (float)std::cos(a);
    }
}

template <class T>
struct Acts {
    T x[kIn], h1[kHid], h2[kHid], h3[kHid], z[kOut], q[kOut], p[kOut];
};

// Forward pass. emulate: bf16 operands of every product (inputs, weights,
// hidden activations), products summed exactly and rounded once to T.
template <class T>
static inline void forward(const float* prm, const float* raw, float alpha, bool emulate, Acts<T>& a) {
    auto rnd = [&](double v) -> T { return emulate ? (T)bf16((float)v) : (T)v; };
    auto w = [&](int idx) -> double { return emulate ? (double)bf16(prm[idx]) : (double)prm[idx]; };
    float enc[kIn];
    encode(raw, enc);
    for (int i = 0; i < kIn; i++) a.x[i] = rnd(enc[i]);
    auto layer = [&](const T* in, int nin, int W, int B, T* out, int nout, bool relu) {
        for (int o = 0; o < nout; o++) {
            double s = 0.0;
            for (int i = 0; i < nin; i++) s += (double)in[i] * w(W + i * nout + o);
            T acc = (T)s;
            T v = acc + (T)prm[B + o];
            if (relu) v = v > (T)0 ? v : (T)0;
            out[o] = relu ? rnd((double)v) : v;
        }
    };
    layer(a.x, kIn, kW1, kB1, a.h1, kHid, true);
    layer(a.h1, kHid, kW2, kB2, a.h2, kHid, true);
    layer(a.h2, kHid, kW3, kB3, a.h3, kHid, true);
    layer(a.h3, kHid, kW4, kB4, a.z, kOut, false);
    T m = a.z[0];
    for (int j = 1; j < kOut; j++) m = a.z[j] > m ? a.z[j] : m;
    T s = 0;
    for (int j = 0; j < kOut; j++) { a.q[j] = std::exp(a.z[j] - m); s += a.q[j]; }
    for (int j = 0; j < kOut; j++) {
        a.q[j] = a.q[j] / s;
        a.p[j] = (T)alpha / (T)kOut + ((T)1 - (T)alpha) * a.q[j];
    }
}

// 16-bit CDF of the bin probabilities: n_j = round(p_j 2^16) (at least 1) for
// j < 31, the last bin takes the rest; cdf[j] = sum_{k<j} n_k, cdf[32] = 65536
// is implicit. Sampling and the pdf both read these integers, so the pdf the
// pass divides by is exactly the one it samples from.
static inline void quantize(const float* p, uint16_t* cdf) {
    int acc = 0;
    for (int j = 0; j < kOut; j++) {
        cdf[j] = (uint16_t)acc;
        int n = j < kOut - 1 ? (int)std::floor(p[j] * 65536.f + 0.5f) : 65536 - acc;
        if (n < 1) n = 1;
        acc += n;
    }
}
static inline int cdfAt(const uint16_t* cdf, int j) { return j >= kOut ? 65536 : (int)cdf[j]; }

// Inverse CDF: u in [0,1) -> c in [0,1), the bin and the density g(c) = 32 n_j / 2^16.
static inline float sampleC(const uint16_t* cdf, float u, int& bin, float& g) {
    const float U = u * 65536.f;
    int j = 0;
    for (int k = 1; k < kOut; k++) j += (float)cdf[k] <= U ? 1 : 0;
    const int lo = cdfAt(cdf, j), n = cdfAt(cdf, j + 1) - lo;
    const float f = (U - (float)lo) / (float)n;
    bin = j;
    g = (float)n * (32.f / 65536.f);
    return ((float)j + f) * (1.f / 32.f);
}
// Probability mass of c >= x (the escape mass when x = 1 - Tr_total).
static inline float massAbove(const uint16_t* cdf, float x) {
    if (!(x > 0.f)) return 1.f;
    if (x >= 1.f) return 0.f;
    const float s = x * 32.f;
    int j = (int)s;
    if (j > kOut - 1) j = kOut - 1;
    const int lo = cdfAt(cdf, j), n = cdfAt(cdf, j + 1) - lo;
    const float G = ((float)lo + (float)n * (s - (float)j)) * (1.f / 65536.f);
    return 1.f - G;
}

// Cross-entropy of one training record against the mixture density g(c) = 32 p:
// loss = -sum_j h_j log(32 p_j) with h = hist / W. Returns 0 (and adds nothing)
// for a record without weight.
template <class T>
static inline bool recordTargets(const float* rec, T* h, T* sq) {
    const double W = rec[7];
    if (!(W > 0.0) || !std::isfinite(W)) return false;
    for (int j = 0; j < kOut; j++) {
        h[j] = (T)(rec[kRecHist + j] / W);
        sq[j] = (T)(rec[kRecSq + j] / (W * W));
    }
    return true;
}
template <class T>
static inline T recordLoss(const float* prm, const float* rec, float alpha, bool emulate) {
    T h[kOut], sq[kOut];
    if (!recordTargets(rec, h, sq)) return 0;
    Acts<T> a;
    forward<T>(prm, rec, alpha, emulate, a);
    T l = 0;
    for (int j = 0; j < kOut; j++) l -= h[j] * std::log((T)kOut * a.p[j]);
    return l;
}

// Gradient of scale * loss of one record, added to grad[kParams]; stats[0] +=
// scale * loss, stats[1] += sum_j sq_j (the analytic proposal's normalised
// second moment of the RIS weights), stats[2] += sum_j sq_j / (32 p_j) (the
// network's, by importance reweighting of the same samples).
template <class T>
static inline void backward(const float* prm, const float* rec, float alpha, float scale, bool emulate, double* grad,
                            double* stats) {
    T h[kOut], sq[kOut];
    if (!recordTargets(rec, h, sq)) return;
    auto rnd = [&](T v) -> T { return emulate ? (T)bf16((float)v) : v; };
    auto w = [&](int idx) -> double { return emulate ? (double)bf16(prm[idx]) : (double)prm[idx]; };
    Acts<T> a;
    forward<T>(prm, rec, alpha, emulate, a);
    T loss = 0, sPlain = 0, sNet = 0, R = 0, r[kOut];
    for (int j = 0; j < kOut; j++) {
        loss -= h[j] * std::log((T)kOut * a.p[j]);
        sPlain += sq[j];
        sNet += sq[j] / ((T)kOut * a.p[j]);
        r[j] = h[j] * a.q[j] / a.p[j];
        R += r[j];
    }
    stats[0] += (double)scale * (double)loss;
    stats[1] += (double)sPlain;
    stats[2] += (double)sNet;
    T dz4[kOut], dz3[kHid], dz2[kHid], dz1[kHid];
    for (int j = 0; j < kOut; j++) dz4[j] = rnd((T)scale * ((T)1 - (T)alpha) * (a.q[j] * R - r[j]));
    // dL/dh_in = W dz (products of bf16 operands summed exactly, rounded once), masked by ReLU
    auto back = [&](const T* dz, int nout, int W, const T* hin, int nin, T* dzIn) {
        for (int i = 0; i < nin; i++) {
            double s = 0.0;
            for (int o = 0; o < nout; o++) s += w(W + i * nout + o) * (double)dz[o];
            dzIn[i] = hin[i] > (T)0 ? rnd((T)s) : (T)0;
        }
    };
    auto outer = [&](const T* in, int nin, const T* dz, int nout, int W, int B) {
        for (int i = 0; i < nin; i++)
            for (int o = 0; o < nout; o++) grad[W + i * nout + o] += (double)in[i] * (double)dz[o];
        for (int o = 0; o < nout; o++) grad[B + o] += (double)dz[o];
    };
    outer(a.h3, kHid, dz4, kOut, kW4, kB4);
    back(dz4, kOut, kW4, a.h3, kHid, dz3);
    outer(a.h2, kHid, dz3, kHid, kW3, kB3);
    back(dz3, kHid, kW3, a.h2, kHid, dz2);
    outer(a.h1, kHid, dz2, kHid, kW2, kB2);
    back(dz2, kHid, kW2, a.h1, kHid, dz1);
    outer(a.x, kIn, dz1, kHid, kW1, kB1);
}

// Which pixel of its block trains the network this frame: one per BX x BY block
// (clipped at the image edge), chosen by a hash of the block and the frame
// (lowbias32, C. Wellons' integer hash). Returns the record slot or -1.
static inline uint32_t trainHash(uint32_t a) {
    a ^= a >> 16; a *= 0x7feb352du; a ^= a >> 15; a *= 0x846ca68bu; a ^= a >> 16;
    return a;
}
static inline int trainSlot(int x, int y, int W, int H, int BX, int BY, int frame) {
    const int bx = x / BX, by = y / BY, nbx = (W + BX - 1) / BX;
    const int slot = by * nbx + bx;
    const int wx = W - bx * BX < BX ? W - bx * BX : BX, wy = H - by * BY < BY ? H - by * BY : BY;
    const uint32_t hsh = trainHash((uint32_t)slot ^ ((uint32_t)frame * 0x9E3779B9u));
    const int ox = (int)(hsh % (uint32_t)wx), oy = (int)((hsh >> 16) % (uint32_t)wy);
    return (x - bx * BX == ox && y - by * BY == oy) ? slot : -1;
}
static inline int trainSlots(int W, int H, int BX, int BY) { return ((W + BX - 1) / BX) * ((H + BY - 1) / BY); }

// Adam (Kingma & Ba 2015, Algorithm 1) on the FP32 master copy; step t >= 1.
// The bias corrections 1 - b^t are computed in double and rounded once.
constexpr float kAdamB1 = 0.9f, kAdamB2 = 0.999f, kAdamEps = 1e-8f;
static inline void adam(float* prm, float* m, float* v, const float* g, int n, int t, float lr, float b1, float b2,
                        float eps) {
    const float c1 = (float)(1.0 - std::pow((double)b1, (double)t)), c2 = (float)(1.0 - std::pow((double)b2, (double)t));
    for (int i = 0; i < n; i++) {
        m[i] = b1 * m[i] + (1.f - b1) * g[i];
        v[i] = b2 * v[i] + (1.f - b2) * g[i] * g[i];
        const float mh = m[i] / c1, vh = v[i] / c2;
        prm[i] -= lr * mh / (std::sqrt(vh) + eps);
    }
}

}  // namespace mlp
}  // namespace orc
