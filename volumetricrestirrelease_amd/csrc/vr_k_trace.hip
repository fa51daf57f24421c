// vr_k_trace.hip -- launcher of the TraceRays stage (and of the reference path
// tracer, which replaces it when the pass runs in reference mode).
#include "vr_stages.hpp"

namespace vr {

// The volume path tracer on a pixel lattice of the frame: lattice point (i, j)
// is pixel (s i + s/2, s j + s/2) of the W x H frame (same camera ray as the
// full-resolution pass). Chunk z (blockIdx.z) averages `spp` samples from the
// stream of sample number sn + z into image z of `out`, so one launch fills the
// chip even for a sparse lattice. Used to build converged reference images on a
// uniform pixel subsample.
template <int B>
__global__ __launch_bounds__(256) void kReferenceLattice(const Scene* __restrict__ sc, const Frame fr, int stride,
                                                         int lw, int lh, int spp, uint32_t sn,
                                                         float4* __restrict__ out) {
    const Scene& S = *sc;
    const Frame& F = fr;
    occCacheLoad(S, F.occSlot[0], F.occSlot[1]);
    const int i = blockIdx.x * 16 + (threadIdx.x & 15), j = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (i >= lw || j >= lh) return;
    const int x = stride * i + stride / 2, y = stride * j + stride / 2;
    Rng rng = rngCreate(x, y, sn + blockIdx.z);
    Ray ray{ld3(F.cam.pos), rayDir(F.cam, x, y, F.width, F.height), 0.f, kTMax};
    v3 sum = V3(0.f);
    for (int s = 0; s < spp; s++) sum = sum + pathTrace(S, F, ray, rng, max(1, F.initialOpt.lightSamples), B);
    sum = sum / (float)spp;
    out[((size_t)blockIdx.z * lh + j) * lw + i] = make_float4(sum.x, sum.y, sum.z, 1.f);
}

void latticeSize(int width, int height, int stride, int& lw, int& lh) {
    lw = (width - stride / 2 + stride - 1) / stride;
    lh = (height - stride / 2 + stride - 1) / stride;
}

hipError_t launchReferenceLattice(const LaunchArgs& a, int stride, int spp, uint32_t sampleNumber, int chunks,
                                  float4* out) {
    int lw, lh;
    latticeSize(a.width, a.height, stride, lw, lh);
    const dim3 g((lw + 15) / 16, (lh + 15) / 16, chunks), blk(256, 1, 1);
    VR_DISPATCH_B(a, hipLaunchKernelGGL(kReferenceLattice<B>, g, blk, a.occBytes, a.stream, a.scene, a.frame, stride,
                                        lw, lh, spp, sampleNumber, out))
    return hipGetLastError();
}

// Split TraceRays: record the candidates' shadow rays (and keep their free-flight
// samples), trace the rays as brick segments with the initial lighting options,
// replay the candidates with their transmittances, then run the pixels the replay
// could not finish (a transmittance of exactly 0) in one kernel.
static hipError_t launchTraceRaysSplit(const LaunchArgs& a, bool spMarch) {
    const dim3 g = tiles(a.width, a.height), blk(256, 1, 1);
    hipError_t e = launchZeroCounters(a.stream, a.split.fix, 1);
    if (e != hipSuccess) return e;
    VR_DISPATCH_B(a, hipLaunchKernelGGL((kTraceSplit<B, 2, 2, kRecord>), g, blk, a.occBytes, a.stream, a.scene,
                                        a.frame, a.res0, a.extra0, a.split))
    if ((e = hipGetLastError()) != hipSuccess) return e;
    // every record is a shadow ray (kind 1): one grid, filter and step for all segments
    Options lo = a.frame.initialOpt;
    lo.visibilityMipLevel = lo.lightingMipLevel;
    lo.visibilityUseLinearSampler = lo.lightingUseLinearSampler;
    lo.visibilityTStepScale = lo.lightingTStepScale;
    lo.visibilityTrackingMethod = lo.lightingTrackingMethod;
    lo.visibilitySamples = lo.lightSamples;
    if ((e = launchTraceSegments(a, lo, false)) != hipSuccess) return e;
    VR_DISPATCH_B(a, {
        if (spMarch)
            hipLaunchKernelGGL((kTraceSplit<B, 2, 2, kReplay>), g, blk, a.occBytes, a.stream, a.scene, a.frame, a.res0,
                               a.extra0, a.split);
        else
            hipLaunchKernelGGL((kTraceSplit<B, -1, -1, kReplay>), g, blk, a.occBytes, a.stream, a.scene, a.frame,
                               a.res0, a.extra0, a.split);
    })
    if ((e = hipGetLastError()) != hipSuccess) return e;
    const dim3 gf(64), bf(256);
    VR_DISPATCH_B(a, {
        if (spMarch)
            hipLaunchKernelGGL((kTraceFix<B, 2, 2>), gf, bf, a.occBytes, a.stream, a.scene, a.frame, a.res0, a.extra0,
                               a.split);
        else
            hipLaunchKernelGGL((kTraceFix<B, -1, -1>), gf, bf, a.occBytes, a.stream, a.scene, a.frame, a.res0,
                               a.extra0, a.split);
    })
    return hipGetLastError();
}

hipError_t launchTraceRays(const LaunchArgs& a) {
    const dim3 g = tiles(a.width, a.height), blk(256, 1, 1);
    const Options& so = a.frame.spatialOpt;
    const bool spMarch = so.visibilityTrackingMethod == 2u && so.lightingTrackingMethod == 2u;
    const bool inMarch = a.frame.initialOpt.lightingTrackingMethod == 2u;
    if (a.splitTK == 2 && inMarch && !a.learnedOn && !a.frame.useReference && !a.frame.noReuse)
        return launchTraceRaysSplit(a, spMarch);
    VR_DISPATCH_B(a, {
        if (a.frame.useReference)
            hipLaunchKernelGGL(kReference<B>, g, blk, a.occBytes, a.stream, a.scene, a.frame, a.color0);
        else if (a.learnedOn && spMarch && inMarch)
            hipLaunchKernelGGL((kTrace<B, 2, 2, 2, true>), g, blk, a.occBytes, a.stream, a.scene, a.frame, a.res0,
                               a.extra0, a.learned);
        else if (a.learnedOn)
            hipLaunchKernelGGL((kTrace<B, -1, -1, -1, true>), g, blk, a.occBytes, a.stream, a.scene, a.frame, a.res0,
                               a.extra0, a.learned);
        else if (spMarch && inMarch)
            hipLaunchKernelGGL((kTrace<B, 2, 2, 2>), g, blk, a.occBytes, a.stream, a.scene, a.frame, a.res0, a.extra0,
                               a.learned);
        else
            hipLaunchKernelGGL((kTrace<B, -1, -1, -1>), g, blk, a.occBytes, a.stream, a.scene, a.frame, a.res0,
                               a.extra0, a.learned);
    })
    return hipGetLastError();
}

}  // namespace vr
