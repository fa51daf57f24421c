// vr_k_spatial.hip -- launcher of one SpatialReuse round.
#include "vr_stages.hpp"

namespace vr {

hipError_t launchSpatialReuse(const LaunchArgs& a) {
    const dim3 g = tiles(a.width, a.height), blk(256, 1, 1);
    const Options& so = a.frame.spatialOpt;
    const bool spMarch = so.visibilityTrackingMethod == 2u && so.lightingTrackingMethod == 2u;
    if (a.splitTK == 2 && spMarch) {
        // record the rays, trace them from balanced queues, replay with their transmittances
        VR_DISPATCH_B(a, {
            hipLaunchKernelGGL((kSpatial<B, 2, 2, kRecord>), g, blk, a.occBytes, a.stream, a.scene, a.frame, a.round,
                               a.offsets, a.res0, a.res1, a.extra0, a.extra1, a.features, a.split);
        })
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        e = a.splitMode == 2 ? launchTraceSegments(a, so, false) : launchTraceQueue(a, so, false, 2);
        if (e != hipSuccess) return e;
        VR_DISPATCH_B(a, {
            hipLaunchKernelGGL((kSpatial<B, 2, 2, kReplay>), g, blk, a.occBytes, a.stream, a.scene, a.frame, a.round,
                               a.offsets, a.res0, a.res1, a.extra0, a.extra1, a.features, a.split);
        })
        return hipGetLastError();
    }
    VR_DISPATCH_B(a, {
        if (spMarch)
            hipLaunchKernelGGL((kSpatial<B, 2, 2, kDirect>), g, blk, a.occBytes, a.stream, a.scene, a.frame, a.round,
                               a.offsets, a.res0, a.res1, a.extra0, a.extra1, a.features, a.split);
        else
            hipLaunchKernelGGL((kSpatial<B, -1, -1, kDirect>), g, blk, a.occBytes, a.stream, a.scene, a.frame,
                               a.round, a.offsets, a.res0, a.res1, a.extra0, a.extra1, a.features, a.split);
    })
    return hipGetLastError();
}

}  // namespace vr
