// vr_launch.h -- host-side launch interface of the HIP kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <functional>
#include <vector>

#include "vr_types.h"

namespace vr {

struct ResStore;

// Buffers of one stage launch. res0/extra0 are the stage's "current" (or input)
// reservoirs, res1/extra1 the spatial-reuse output, resT/extraT the temporal history.
struct LaunchArgs {
    hipStream_t stream;
    int stage, round, maxBounces, width, height;
    int occBytes;  // dynamic LDS of the occupancy tables named in frame.occSlot
    int splitTK;   // > 0: run the reuse stage as record / trace / replay with tracker kind splitTK
    int splitMode; // the trace of that split: 1 = walker queue (kTraceQueue), 2 = brick segments (kTraceSegments)
    SplitBufs split;
    int64_t traceGridBlocks;  // 128^3 blocks of the grid the analytic cell trace walks (host-side count)
    const Scene* scene;
    Frame frame;
    const int* offsets;
    ResStore* res0;
    ResStore* res1;
    ResStore* resT;
    float* extra0;
    float* extra1;
    float* extraT;
    float2* features;
    float2* featuresT;
    VItem* vbuf;             // the V-buffer (surface scenes)
    float4* color0;
    float4* out;
    float2* mvec;
    int learnedOn;           // TraceRays samples its initial distances from the learned proposal
    LearnedTraceArgs learned;
};

hipError_t launchStage(const LaunchArgs& a);
// Per-stage launchers, one translation unit each (vr_k_*.hip).
hipError_t launchTraceRays(const LaunchArgs& a);
void latticeSize(int width, int height, int stride, int& lw, int& lh);
hipError_t launchReferenceLattice(const LaunchArgs& a, int stride, int spp, uint32_t sampleNumber, int chunks,
                                  float4* out);
hipError_t launchTemporalReuse(const LaunchArgs& a);
hipError_t launchSpatialReuse(const LaunchArgs& a);
hipError_t launchFinalShading(const LaunchArgs& a);
// Balanced trace of the rays recorded in a.split (vr_k_queue.hip); tk: 2 ray marching, 1 analytic.
hipError_t launchTraceQueue(const LaunchArgs& a, const Options& opt, bool isFinal, int tk);
hipError_t launchTraceSegments(const LaunchArgs& a, const Options& opt, bool isFinal);
// The same for analytic tracking (kTraceSegments<1>): one grid whose block bits fit LDS, one filter.
hipError_t launchTraceCells(const LaunchArgs& a, const Options& opt, bool isFinal);
// ---- batched launches: many passes' tiles in one persistent launch
// The work list of one batched step: instance i owns work items [start[i],
// start[i+1]); an item is `per` consecutive tiles of that instance, tiles
// row-major over its tile grid of gx columns (16x16-pixel tiles for the stage
// bodies, 256-pixel ranges for the segment traces).
constexpr int kMaxBatch = 64;
struct BatchList {
    int n, per;
    int start[kMaxBatch + 1];
    int tiles[kMaxBatch];
    int gx[kMaxBatch];
};
// The bodies a batched step runs; every instance of a batch shares the path
// length and the tracker kinds that select them.
enum BatchOp {
    kOpFeatures = 0, kOpTrace, kOpTemporalRecord, kOpTemporalReplay, kOpCopyFeatures, kOpSpatialRecord,
    kOpSpatialReplay, kOpCopyReservoirs, kOpFinalRecord, kOpFinalReplay, kOpSegmentsMarch, kOpSegmentsCells,
    kOpTraceLearned, kOpTemporalDirect, kOpCount
};
// One persistent launch of `grid` workgroups draining the step's work list
// (counter: the step's queue head, zero before the launch). args and list are
// device memory; occBytes is the largest dynamic LDS any instance needs.
hipError_t launchBatched(hipStream_t st, int op, int maxBounces, const LaunchArgs* args, const BatchList* list,
                         int* counter, int grid, int occBytes);
hipError_t launchBatchedSegments(hipStream_t st, int op, const LaunchArgs* args, const BatchList* list, int* counter,
                                 int grid, int occBytes);
hipError_t launchZeroCounters(hipStream_t st, int* counters, int n);

// Ground-truth accumulation and relMSE (vr_k_accum.hip).
hipError_t launchAccumulate(hipStream_t st, const float4* frame, double4* sum, int n, bool reset);
hipError_t launchResolve(hipStream_t st, const double4* sum, double inv, float4* out, int n);
int relMsePartialCount(int n);
hipError_t launchRelMse(hipStream_t st, const float4* img, const float4* ref, int n, double* partial, double* out);
// Allocation-free relMSE: partials and result in a static device array of
// kRelMseScratchDoubles doubles (images of up to 32767 * 256 = 8,388,352 pixels).
constexpr int kRelMseScratchDoubles = 32768, kMaxDevices = 64;
constexpr int kRelMseNoAllocMaxPixels = (kRelMseScratchDoubles - 1) * 256;
hipError_t relMseScratchInit(double** ptrOut);
hipError_t relMseNoAlloc(hipStream_t st, const float4* img, const float4* ref, int n, double* hostOut);
hipError_t launchTrack(hipStream_t stream, const Scene* scene, const EnvXf& env, int mode, int mip, int linear,
                       float tStepScale, int numSamples, int n, const float* rays, const uint32_t* seeds, float* out,
                       int occSlot, int occBytes);
// Learned proposal (vr_k_learned.hip): the network's CDF for n rays (rawIn NULL:
// the frame's camera rays, ray = pixel), the per-wave gradient partials of n
// training records, their fixed-order sum (with adam: the Adam step, statistics
// and the judge's flag), and the bf16 operand fragments of the master weights.
hipError_t launchLearnedInfer(hipStream_t st, const Scene* scene, const Frame& frame, const float* rawIn, int n,
                              const uint4* frags, const float* prm, const int* inUse, float alpha, uint16_t* cdf,
                              float* logits);
hipError_t launchLearnedGrad(hipStream_t st, const float* recs, int n, float alpha, float scale, const uint4* frags,
                             const float* prm, float* partial);
hipError_t launchLearnedReduce(hipStream_t st, const float* partial, int waves, float* out, bool adam, float* prm,
                               float* m, float* v, float lr, float c1, float c2, float* stats, int* inUse, bool judge,
                               int step);
hipError_t launchLearnedPack(hipStream_t st, const float* prm, uint4* frags);
// Volume grids built on the GPU from dense device grids (vr_k_build.hip): the
// host build's clamp, conservative dilation and 2x reduction, and one slot
// (tables, occupancy words, pool, bounds) from a dense grid. buildSlotDevice
// appends its allocations to allocs and synchronises the stream.
// Each also leaves the largest value it writes (as float bits, via atomicMax) at maxBits.
hipError_t buildClamp(hipStream_t st, const float* in, float* out, size_t n, int* bad, unsigned* maxBits);
hipError_t buildDilate(hipStream_t st, const float* s, float* d, int nx, int ny, int nz, unsigned* maxBits);
hipError_t buildReduce2(hipStream_t st, const float* s, float* d, int nx, int ny, int nz, unsigned* maxBits);
size_t buildScratchBytes(int nx, int ny, int nz);
// A slot's device buffer of at least `bytes` (kind: one of kBuf*), possibly one
// kept from an earlier build of the same slot.
enum { kBufTable = 0, kBufNodes, kBufWords, kBufBounds, kBufPool, kBufKinds };
using BuildAlloc = std::function<hipError_t(int kind, size_t bytes, void** p)>;
hipError_t buildSlotDevice(hipStream_t st, const float* dense, int nx, int ny, int nz, const unsigned* gmaxBits,
                           bool fp16, void* scratch, Grid& G, uint32_t& numBricks, float& maxValue,
                           const BuildAlloc& allocBuf);
// Importance mip chain of a lat-long map: level k of dim >> k texels at imp + off[k].
hipError_t launchEnvImportance(hipStream_t st, const float* tex, int W, int H, float* imp, const int* off, int dim,
                               int spp);

}  // namespace vr
