// vr_k_final.hip -- launcher of the FinalShading stage.
#include "vr_stages.hpp"

namespace vr {

hipError_t launchFinalShading(const LaunchArgs& a) {
    const dim3 g = tiles(a.width, a.height), blk(256, 1, 1);
    const Options& fo = a.frame.finalOpt;
    const bool fiAnalytic = fo.visibilityTrackingMethod == 1u && fo.lightingTrackingMethod == 1u;
    if (a.splitTK == 1 && fiAnalytic) {
        // record the transmittance rays, trace them as analytic brick segments, replay
        // with their transmittances
        VR_DISPATCH_B(a, {
            hipLaunchKernelGGL((kFinal<B, 1, 1, kRecord>), g, blk, a.occBytes, a.stream, a.scene, a.frame, a.res0,
                               a.extra0, a.features, a.color0, a.out, a.split);
        })
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        e = launchTraceCells(a, fo, true);
        if (e != hipSuccess) return e;
        VR_DISPATCH_B(a, {
            hipLaunchKernelGGL((kFinal<B, 1, 1, kReplay>), g, blk, a.occBytes, a.stream, a.scene, a.frame, a.res0,
                               a.extra0, a.features, a.color0, a.out, a.split);
        })
        return hipGetLastError();
    }
    VR_DISPATCH_B(a, {
        if (fiAnalytic)
            hipLaunchKernelGGL((kFinal<B, 1, 1>), g, blk, a.occBytes, a.stream, a.scene, a.frame, a.res0, a.extra0,
                               a.features, a.color0, a.out, a.split);
        else
            hipLaunchKernelGGL((kFinal<B, -1, -1>), g, blk, a.occBytes, a.stream, a.scene, a.frame, a.res0, a.extra0,
                               a.features, a.color0, a.out, a.split);
    })
    return hipGetLastError();
}

}  // namespace vr
