// vr_k_spatial.hip -- launcher of one SpatialReuse round.
#include "vr_stages.hpp"

namespace vr {

hipError_t launchSpatialReuse(const LaunchArgs& a) {
    const dim3 g = tiles(a.width, a.height), blk(256, 1, 1);
    const Options& so = a.frame.spatialOpt;
    const bool spMarch = so.visibilityTrackingMethod == 2u && so.lightingTrackingMethod == 2u;
    VR_DISPATCH_B(a, {
        if (spMarch)
            hipLaunchKernelGGL((kSpatial<B, 2, 2>), g, blk, a.occBytes, a.stream, a.scene, a.frame, a.round, a.offsets, a.res0,
                               a.res1, a.extra0, a.extra1, a.features);
        else
            hipLaunchKernelGGL((kSpatial<B, -1, -1>), g, blk, a.occBytes, a.stream, a.scene, a.frame, a.round, a.offsets,
                               a.res0, a.res1, a.extra0, a.extra1, a.features);
    })
    return hipGetLastError();
}

}  // namespace vr
