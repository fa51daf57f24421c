// vr_k_batch.hip -- batched persistent launches of the per-tile stage bodies:
// one launch runs a stage (or one of its record / replay passes) for every pass
// instance of a batch. Its workgroups take work items -- groups of 16x16-pixel
// tiles of one instance -- from a queue head in device memory until the list is
// drained, so instances of uneven size share the chip without a launch each.
// No counterpart in the reference, which renders one view per pass.
#include "vr_stages.hpp"

namespace vr {

// Copies a tile's features into the temporal history (the features half of
// the copy vrestir_run_stage makes after TemporalReuse).
VR_DEV void copyFeaturesBody(const Frame& F, const float2* __restrict__ features, float2* __restrict__ featuresT,
                             VBlock vb) {
    PixelIdx px = pixelOf(F, vb);
    if (!px.ok) return;
    featuresT[px.id] = features[px.id];
}

template <int OP, int B>
VR_DEV void runOp(const LaunchArgs& a, VBlock vb) {
    if constexpr (OP == kOpFeatures) featuresBody(a.scene, a.frame, a.features, a.vbuf, vb);
    else if constexpr (OP == kOpTrace) traceBody<B, 2, 2, 2, false>(a.scene, a.frame, a.res0, a.extra0, a.learned, vb);
    else if constexpr (OP == kOpTraceLearned)
        traceBody<B, 2, 2, 2, true>(a.scene, a.frame, a.res0, a.extra0, a.learned, vb);
    else if constexpr (OP == kOpTemporalDirect)
        temporalBody<B, 2, 2, kDirect>(a.scene, a.frame, a.res0, a.extra0, a.resT, a.extraT, a.features, a.featuresT,
                                       a.mvec, a.split, vb);
    else if constexpr (OP == kOpTemporalRecord)
        temporalBody<B, 2, 2, kRecord>(a.scene, a.frame, a.res0, a.extra0, a.resT, a.extraT, a.features, a.featuresT,
                                       a.mvec, a.split, vb);
    else if constexpr (OP == kOpTemporalReplay)
        temporalBody<B, 2, 2, kReplay>(a.scene, a.frame, a.res0, a.extra0, a.resT, a.extraT, a.features, a.featuresT,
                                       a.mvec, a.split, vb);
    else if constexpr (OP == kOpCopyFeatures) copyFeaturesBody(a.frame, a.features, a.featuresT, vb);
    else if constexpr (OP == kOpSpatialRecord)
        spatialBody<B, 2, 2, kRecord>(a.scene, a.frame, a.round, a.offsets, a.res0, a.res1, a.extra0, a.extra1,
                                      a.features, a.split, vb);
    else if constexpr (OP == kOpSpatialReplay)
        spatialBody<B, 2, 2, kReplay>(a.scene, a.frame, a.round, a.offsets, a.res0, a.res1, a.extra0, a.extra1,
                                      a.features, a.split, vb);
    else if constexpr (OP == kOpCopyReservoirs) copyBody<B>(a.frame, a.res0, a.extra0, a.resT, a.extraT, vb);
    else if constexpr (OP == kOpFinalRecord)
        finalBody<B, 1, 1, kRecord>(a.scene, a.frame, a.res0, a.extra0, a.features, a.color0, a.out, a.split, vb);
    else if constexpr (OP == kOpFinalReplay)
        finalBody<B, 1, 1, kReplay>(a.scene, a.frame, a.res0, a.extra0, a.features, a.color0, a.out, a.split, vb);
}

// The persistent loop: a barrier before each item (the previous item's bodies
// are done with the workgroup's LDS), one queue fetch per item by thread 0, the
// item's tiles in order. Every thread runs every iteration, so the bodies'
// barriers are reached by the whole workgroup.
template <int OP, int B>
VR_TRACING_KERNEL void kBatched(const LaunchArgs* __restrict__ args, const BatchList* __restrict__ list,
                                int* __restrict__ counter) {
    __shared__ int sItem;
    const int n = list->n, total = list->start[n], per = list->per;
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) sItem = atomicAdd(counter, 1);
        __syncthreads();
        // the same value in every lane: made wave-uniform so that the pass's
        // arguments below are read with scalar loads into scalar registers
        const int item = __builtin_amdgcn_readfirstlane(sItem);
        if (item >= total) break;
        int i = 0;
        while (i + 1 < n && item >= list->start[i + 1]) i++;
        const LaunchArgs& a = args[i];
        const int t0 = (item - list->start[i]) * per, t1 = min(t0 + per, list->tiles[i]), gx = list->gx[i];
        for (int t = t0; t < t1; t++) {
            if (t > t0) __syncthreads();
            runOp<OP, B>(a, VBlock{t % gx, t / gx});
        }
    }
}

template <int B>
static hipError_t launchOp(hipStream_t st, int op, const LaunchArgs* args, const BatchList* list, int* counter,
                           int grid, int occBytes) {
    const dim3 g(grid), blk(256);
#define VR_BATCH_CASE(OP) \
    case OP: hipLaunchKernelGGL((kBatched<OP, B>), g, blk, occBytes, st, args, list, counter); break;
    switch (op) {
        VR_BATCH_CASE(kOpFeatures)
        VR_BATCH_CASE(kOpTrace)
        VR_BATCH_CASE(kOpTemporalRecord)
        VR_BATCH_CASE(kOpTemporalReplay)
        VR_BATCH_CASE(kOpCopyFeatures)
        VR_BATCH_CASE(kOpSpatialRecord)
        VR_BATCH_CASE(kOpSpatialReplay)
        VR_BATCH_CASE(kOpCopyReservoirs)
        VR_BATCH_CASE(kOpFinalRecord)
        VR_BATCH_CASE(kOpFinalReplay)
        VR_BATCH_CASE(kOpTraceLearned)
        VR_BATCH_CASE(kOpTemporalDirect)
        default: return hipErrorInvalidValue;
    }
#undef VR_BATCH_CASE
    return hipGetLastError();
}

hipError_t launchBatched(hipStream_t st, int op, int maxBounces, const LaunchArgs* args, const BatchList* list,
                         int* counter, int grid, int occBytes) {
    if (grid <= 0) return hipErrorInvalidValue;
    switch (maxBounces) {
        case 1: return launchOp<1>(st, op, args, list, counter, grid, occBytes);
        case 2: return launchOp<2>(st, op, args, list, counter, grid, occBytes);
        case 3: return launchOp<3>(st, op, args, list, counter, grid, occBytes);
        case 4: return launchOp<4>(st, op, args, list, counter, grid, occBytes);
        default: return hipErrorInvalidValue;
    }
}

// Queue heads of a batched frame to zero: a kernel, not a memset, so the frame
// can be captured into a graph and replayed.
__global__ void kZeroCounters(int* __restrict__ c, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) c[i] = 0;
}
hipError_t launchZeroCounters(hipStream_t st, int* counters, int n) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(kZeroCounters, dim3((n + 255) / 256), dim3(256), 0, st, counters, n);
    return hipGetLastError();
}

}  // namespace vr
