// vr_k_temporal.hip -- launcher of the TemporalReuse stage.
#include "vr_stages.hpp"

namespace vr {

hipError_t launchTemporalReuse(const LaunchArgs& a) {
    const dim3 g = tiles(a.width, a.height), blk(256, 1, 1);
    const Options& so = a.frame.spatialOpt;
    const bool spMarch = so.visibilityTrackingMethod == 2u && so.lightingTrackingMethod == 2u;
    if (a.splitTK == 2 && spMarch) {
        // record the rays, trace them as brick segments, replay with their transmittances
        VR_DISPATCH_B(a, {
            hipLaunchKernelGGL((kTemporal<B, 2, 2, kRecord>), g, blk, a.occBytes, a.stream, a.scene, a.frame, a.res0,
                               a.extra0, a.resT, a.extraT, a.features, a.featuresT, a.mvec, a.split);
        })
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        e = launchTraceSegments(a, so, false);
        if (e != hipSuccess) return e;
        VR_DISPATCH_B(a, {
            hipLaunchKernelGGL((kTemporal<B, 2, 2, kReplay>), g, blk, a.occBytes, a.stream, a.scene, a.frame, a.res0,
                               a.extra0, a.resT, a.extraT, a.features, a.featuresT, a.mvec, a.split);
        })
        return hipGetLastError();
    }
    VR_DISPATCH_B(a, {
        if (spMarch)
            hipLaunchKernelGGL((kTemporal<B, 2, 2>), g, blk, a.occBytes, a.stream, a.scene, a.frame, a.res0, a.extra0,
                               a.resT, a.extraT, a.features, a.featuresT, a.mvec, a.split);
        else
            hipLaunchKernelGGL((kTemporal<B, -1, -1>), g, blk, a.occBytes, a.stream, a.scene, a.frame, a.res0, a.extra0,
                               a.resT, a.extraT, a.features, a.featuresT, a.mvec, a.split);
    })
    return hipGetLastError();
}

}  // namespace vr
