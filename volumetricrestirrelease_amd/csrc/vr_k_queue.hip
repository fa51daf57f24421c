// vr_k_queue.hip -- balanced trace of recorded transmittance rays.
//
// The record pass of a reuse kernel leaves, per pixel, the rays its target
// function evaluations need (TraceRec, in evaluation order). Here each wave
// takes 64 consecutive pixels and treats their rays as one queue: a lane whose
// ray is finished takes the next ray of the queue, so the wave runs until the
// queue is empty instead of paying, for every evaluation, the longest ray among
// its 64 pixels. Idle lanes are refilled together once a quarter of the wave is
// idle, so the set-up code runs at a useful width. The transmittance of each
// ray is written at its record's index for the replay pass.
#include "vr_stages.hpp"
#include "vr_walker.hpp"

namespace vr {

static constexpr int kRefillLanes = 16;

template <int TK>
__global__ __launch_bounds__(256) void kTraceQueue(const Scene* __restrict__ sc, const Frame fr, Options opt,
                                                   int isFinal, const TraceRec* __restrict__ recs,
                                                   const int* __restrict__ counts, float* __restrict__ trs, int cap,
                                                   int numPixels) {
    const Scene& S = *sc;
    occCacheLoad(S, fr.occSlot[0], fr.occSlot[1]);
    __shared__ int pre[4][65];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int p0 = (blockIdx.x * 4 + wv) * 64;
    const int cnt = p0 + lane < numPixels ? counts[p0 + lane] : 0;
    int incl = cnt;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(incl, o);
        if (lane >= o) incl += y;
    }
    pre[wv][lane + 1] = incl;
    if (lane == 0) pre[wv][0] = 0;
    __syncthreads();
    const int total = pre[wv][64];

    const int vs = opt.visibilityMipLevel, ls = isFinal ? 0 : opt.lightingMipLevel;
    const bool oneSlot = vs == ls;
    Walker<TK> w;
    w.active = false;
    bool busy = false;
    int rec = 0, head = 0;
    for (;;) {
        if (busy) {
            walkAdvance(w, S, oneSlot, vs);
            if (!w.active) { trs[rec] = w.Tr; busy = false; }
        }
        const uint64_t idle = __ballot(!busy);
        const int nIdle = __popcll(idle);
        if (head >= total) {
            if (nIdle == 64) break;
            continue;
        }
        if (nIdle < kRefillLanes && nIdle < 64) continue;
        if (!busy) {
            const int r = head + __popcll(idle & ((1ull << lane) - 1ull));
            if (r < total) {
                int lo = 0, hi = 64;  // the pixel q with pre[q] <= r < pre[q + 1]
                while (hi - lo > 1) {
                    const int mid = (lo + hi) >> 1;
                    if (pre[wv][mid] <= r) lo = mid; else hi = mid;
                }
                rec = (p0 + lo) * cap + (r - pre[wv][lo]);
                const TraceRec R = recs[rec];
                const int kind = R.kind;
                const int slot = kind == 0 ? vs : ls;
                const bool linear = kind == 0 ? opt.visibilityUseLinearSampler != 0 : opt.lightingUseLinearSampler != 0;
                const float tss = kind == 0 ? opt.visibilityTStepScale : opt.lightingTStepScale;
                walkBegin(w, S, oneSlot, vs, slot, Ray{V3(R.ox, R.oy, R.oz), V3(R.dx, R.dy, R.dz), 0.f, R.tMax},
                          linear, tss);
                if (w.active) busy = true;
                else trs[rec] = w.Tr;
            }
        }
        head += nIdle;
    }
}

hipError_t launchTraceQueue(const LaunchArgs& a, const Options& opt, bool isFinal, int tk) {
    const int n = a.width * a.height;
    const dim3 g((n + 255) / 256), blk(256);
    if (tk == 2)
        hipLaunchKernelGGL(kTraceQueue<2>, g, blk, a.occBytes, a.stream, a.scene, a.frame, opt, isFinal ? 1 : 0,
                           a.split.recs, a.split.counts, a.split.trs, a.split.cap, n);
    else if (tk == 1)
        hipLaunchKernelGGL(kTraceQueue<1>, g, blk, a.occBytes, a.stream, a.scene, a.frame, opt, isFinal ? 1 : 0,
                           a.split.recs, a.split.counts, a.split.trs, a.split.cap, n);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

}  // namespace vr
