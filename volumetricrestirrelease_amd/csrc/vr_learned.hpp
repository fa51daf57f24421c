// vr_learned.hpp -- device-side pieces of the learned free-flight proposal that
// TraceRays uses: which pixel of its block records training data, the
// network's raw inputs for a camera ray, and sampling / evaluating the 16-bit
// CDF the proposal kernel writes per pixel.
//
// The proposal has no counterpart in the reference. Per camera ray it gives a
// 32-bin piecewise-constant pdf g over the analytic tracker's CDF coordinate
// c = 1 - exp(-tau); the tracker then runs to optical depth -log(1 - c), so a
// distance t has source pdf g(c(t)) sigma_t(t) Tr(t) -- the pdfDist of
// VR/ComputeInitialSample.slang:86,107 times g -- and an escape has the mass
// of g above 1 - Tr_total instead of Tr_total.
#pragma once
#include "vr_pass.hpp"

namespace vr {
namespace learned {

constexpr int kIn = 36, kHid = 64, kOut = 32;
// FP32 master parameters: W1[36][64], b1[64], W2[64][64], b2[64], W3[64][64],
// b3[64], W4[64][32], b4[32]; W[i][o] multiplies input i into output o.
constexpr int kW1 = 0, kB1 = kW1 + kIn * kHid, kW2 = kB1 + kHid, kB2 = kW2 + kHid * kHid, kW3 = kB2 + kHid,
              kB3 = kW3 + kHid * kHid, kW4 = kB3 + kHid, kB4 = kW4 + kHid * kOut, kParams = kB4 + kOut;
static_assert(kParams == 12768, "parameter count");
// Raw inputs per ray: entry point in mip 0's normalised index box, world
// direction, segment length over the box diagonal, (unused).
constexpr int kRaw = 8;
// Training record per ray: raw inputs (7), the summed weight W, then per bin
// the summed weights and the summed weight^2 * g_used.
constexpr int kRecHist = 8, kRecSq = kRecHist + kOut, kRecFloats = kRecSq + kOut;

// One training ray per bx x by block (clipped at the image edge), chosen by a
// hash of the block and the frame (lowbias32, C. Wellons' integer hash).
VR_DEV uint32_t trainHash(uint32_t a) {
    a ^= a >> 16; a *= 0x7feb352du; a ^= a >> 15; a *= 0x846ca68bu; a ^= a >> 16;
    return a;
}
VR_DEV int trainSlot(int x, int y, int W, int H, int bx, int by, int frame) {
    const int cx = x / bx, cy = y / by, nbx = (W + bx - 1) / bx;
    const int slot = cy * nbx + cx;
    const int wx = min(bx, W - cx * bx), wy = min(by, H - cy * by);
    const uint32_t hsh = trainHash((uint32_t)slot ^ ((uint32_t)frame * 0x9E3779B9u));
    const int ox = (int)(hsh % (uint32_t)wx), oy = (int)((hsh >> 16) % (uint32_t)wy);
    return (x - cx * bx == ox && y - cy * by == oy) ? slot : -1;
}

// The network's inputs for a world-space ray.
VR_DEV void rawInputs(const Scene& S, const Ray& rw, float raw[kRaw]) {
    const Grid& G = S.grid[0];
    Ray ri{xformP(G.w2i, rw.o), xformV(G.w2i, rw.d), 0.f, kTMax};
    raw[3] = rw.d.x; raw[4] = rw.d.y; raw[5] = rw.d.z; raw[7] = 0.f;
    float t0 = 0.f, t1 = 0.f;
    if (!intersectBounds(G, ri, t0, t1, false)) {
        raw[0] = raw[1] = raw[2] = 0.5f;
        raw[6] = 0.f;
        return;
    }
    const v3 e = ri.o + t0 * ri.d;
    raw[0] = e.x / G.bmax[0]; raw[1] = e.y / G.bmax[1]; raw[2] = e.z / G.bmax[2];
    const float diag = sqrtf(G.bmax[0] * G.bmax[0] + G.bmax[1] * G.bmax[1] + G.bmax[2] * G.bmax[2]);
    raw[6] = (t1 - t0) * len3(ri.d) / diag;
}

// A pixel's CDF: cdf[j] = sum_{k<j} n_k as 16-bit integers, cdf[32] = 65536.
struct Cdf {
    uint32_t w[16];  // pairs: low half cdf[2i], high half cdf[2i+1]
    VR_DEV void load(const uint16_t* p) {
        const uint4* q = reinterpret_cast<const uint4*>(p);
#pragma unroll
        for (int i = 0; i < 4; i++) {
            uint4 v = q[i];
            w[4 * i] = v.x; w[4 * i + 1] = v.y; w[4 * i + 2] = v.z; w[4 * i + 3] = v.w;
        }
    }
    VR_DEV int at(int k) const { return (int)((w[k >> 1] >> ((k & 1) * 16)) & 0xFFFFu); }
    // Inverse CDF: u in [0,1) -> c in [0,1), its bin and g(c) = 32 n_bin / 2^16.
    VR_DEV float sample(float u, int& bin, float& g) const {
        const float U = u * 65536.f;
        int j = 0, lo = 0, hi = 65536;
#pragma unroll
        for (int k = 1; k < kOut; k++) {
            const int c = at(k);
            if ((float)c <= U) { j++; lo = max(lo, c); }
            else hi = min(hi, c);
        }
        const int n = hi - lo;
        bin = j;
        g = (float)n * (32.f / 65536.f);
        return ((float)j + (U - (float)lo) / (float)n) * (1.f / 32.f);
    }
    // Mass of c >= x (an escape's probability when x = 1 - Tr_total).
    VR_DEV float massAbove(float x) const {
        if (!(x > 0.f)) return 1.f;
        if (x >= 1.f) return 0.f;
        const float s = x * 32.f;
        const int j = min((int)s, kOut - 1);
        int lo = 0, hi = 65536;
#pragma unroll
        for (int k = 1; k < kOut; k++) {
            const int c = at(k);
            if (k <= j) lo = c;
            if (k == j + 1) hi = c;
        }
        const float G = ((float)lo + (float)(hi - lo) * (s - (float)j)) * (1.f / 65536.f);
        return 1.f - G;
    }
};

using TraceArgs = LearnedTraceArgs;
constexpr int kFragCount = kLearnedFrags, kPartialStride = kLearnedPartialStride;
static_assert(kRecFloats == kLearnedRecFloats && kParams == kLearnedParams && kRaw == kLearnedRaw, "sizes");

}  // namespace learned
}  // namespace vr
