// vr_k_trace.hip -- launcher of the TraceRays stage (and of the reference path
// tracer, which replaces it when the pass runs in reference mode).
#include "vr_stages.hpp"

namespace vr {

hipError_t launchTraceRays(const LaunchArgs& a) {
    const dim3 g = tiles(a.width, a.height), blk(256, 1, 1);
    const Options& so = a.frame.spatialOpt;
    const bool spMarch = so.visibilityTrackingMethod == 2u && so.lightingTrackingMethod == 2u;
    const bool inMarch = a.frame.initialOpt.lightingTrackingMethod == 2u;
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
