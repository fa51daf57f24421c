// vr_types.h -- plain-old-data descriptors shared by the host library and the
// HIP kernels. Everything the kernels read about the scene and the frame is in
// these structs, uploaded once per change into device memory and passed to the
// kernels as `const T* __restrict__` (uniform loads go through the scalar cache).
#pragma once
#include <stdint.h>

namespace vr {

static const uint32_t kEmpty = 0xFFFFFFFFu;
// Grid slots (VR/HostDeviceSharedConstants.slang:19-28): mips 0..7, conservative
// mips 8..15, temperature 16, velocity 17 (18, the super-voxel grid, is not used),
// and the previous frame of an animated sequence: mips at 19 + k, temperature at
// 16 + 11 = 27, velocity at 28.
static const int kMaxSlots = 29;
static const int kTemperatureSlot = 16;
static const int kVelocitySlot = 17;
static const int kFrameSlots = 18;           // slots one frame of a sequence owns
static const int kPrevDensityOffset = 19;    // kPrevFrameDensityGridOffset
static const int kPrevExtraOffset = 11;      // kPrevFrameExtraGridOffset
static const int kBrick = 8;        // voxels per brick side
static const int kApron = 10;       // stored brick side (one-voxel apron on each face)
static const int kBrickVals = 1000; // kApron^3 values per stored brick
static const int kNodeBricks = 16;  // bricks per level-1 node side (128 voxels)
static const int kMaxLights = 256;
static const int kMaxLutRows = 256;  // blackbody table rows (the reference's table has 128)
static const int kMaxImpMips = 12;   // importance map levels: 2048^2 .. 1^2

// One grid slot in HBM: a flat brick table over the grid's brick extent, an
// occupancy table over 128^3-voxel blocks (the hierarchical-DDA level above
// bricks), per-brick density bounds and the brick pool.
struct Grid {
    const uint32_t* brickTable;   // nbx*nby*nbz brick ids (x fastest), kEmpty = no brick
    const uint32_t* nodeTable;    // nnx*nny*nnz, non-zero = block holds a brick
    const float* bounds;          // 4 floats per brick: min, max, avg, 0 (raw values)
    const void* pool;             // 1000 values per brick, fp32 or fp16, x fastest
    int32_t nbx, nby, nbz;
    int32_t nnx, nny, nnz;
    float bmax[3];                // [0, bmax) index-space bounds (nx, ny, nz)
    float w2i[16];                // world -> index, row vector
    float maxValue;
    int32_t fp16;
    int32_t valid;
    int32_t occWords;             // entries of occ
    uint32_t numBricks;
    int32_t channels;             // values per voxel: 1, or 3 for the velocity grid (pool: [brick][channel][1000])
    int32_t pad_;
    // Brick occupancy as 32-bit words over the brick-table index, each with the
    // number of occupied bricks in all earlier words: brick ids are assigned in
    // table order, so id = prefix + popcount(word bits below the brick's bit).
    const uint32_t* occ;          // occWords pairs {word, prefix}
};

struct Scene {
    Grid grid[kMaxSlots];
    float sigma_t;
    float sigma_s[3];
    float sigma_a[3];
    float g;
    float densityScale;
    float tStep;
    float worldScaling;
    float superVoxelDiag;
    int32_t numMips;
    int32_t numLights;
    float env[3];                 // constant environment radiance (used when envTextured == 0)
    float pad_;
    float lights[kMaxLights * 6];  // pos.xyz, intensity.xyz
    // Textured lat-long environment map (LT/EnvMap.slang:34-80) and its hierarchical
    // importance map (LT/EnvMapSampler.slang:94-167): level k (envImpDim >> k texels
    // square, row-major) starts at envImp + envImpOff[k]; level envImpBaseMip is 1x1.
    const float* envTex;          // envW * envH RGBA texels, row 0 at v = 0
    const float* envImp;
    int32_t envTextured, envW, envH, envImpDim, envImpBaseMip;
    float envImpInvDim;
    float envScale[3];            // EnvMapData intensity * tint
    int32_t envImpOff[kMaxImpMips];
    // Volume emission (VR/VolumeBase.slang:196-232): the temperature grid is slot
    // kTemperatureSlot; lut holds lutRows RGB rows of the blackbody table.
    int32_t hasEmission, lutRows;
    float LeScale, tempCutoff, tempScale, pad2_;
    // animated sequences: the previous frame's emission flag, a velocity grid (world units per frame)
    int32_t prevHasEmission, hasVelocity;
    float velocityScale, pad3_;
    float lut[kMaxLutRows * 3];
    // Emissive triangles (LT/LightCollectionShared.slang:41-60, LT/EmissivePowerSampler.slang):
    // emTri holds kEmTriFloats floats per collection triangle (positions p0 p1 p2, face
    // normal as stored, area, emitted radiance); emAlias the power sampler's alias table
    // (4 words per entry: threshold bits, index taken at or above it, index below it);
    // emWeight the flux per triangle and emWeightSum their sum.
    const float* emTri;
    const uint32_t* emAlias;
    const float* emWeight;
    int32_t numEmTris, numEmActive;
    float emWeightSum, emMult;
    // Triangle geometry for ray queries (the acceleration structure of
    // VR/InlineRayTracingHelpers.slang:8): a BVH over the triangles, the triangles in
    // leaf order (kBvhTriFloats floats: v0, scene triangle index bits, v1, material
    // index bits, v2, 0), per scene triangle kTriDataFloats floats (vertex
    // positions, vertex shading normals, material index bits, 0), and per material
    // kMaterialFloats floats (base colour, roughness, metallic, IoR, emissive,
    // double-sided, 0, 0).
    const struct BvhNode* bvh;
    const float* bvhTri;
    const float* triData;
    const float* materials;
    int32_t numTris, numMaterials;
};
constexpr int kEmTriFloats = 16;
constexpr int kBvhTriFloats = 12;
constexpr int kTriDataFloats = 20;
constexpr int kMaterialFloats = 12;
constexpr int kBvhMaxDepth = 30;  // the builder's limit; traversal stacks hold kBvhMaxDepth + 2 entries
// A BVH node: count > 0 is a leaf of count triangles from bvhTri entry first;
// count == 0 is an interior node whose children are nodes first and first + 1.
struct BvhNode {
    float bmin[3];
    int32_t first;
    float bmax[3];
    int32_t count;
};
// A V-buffer entry (VBufferItem, VR/InlineRayTracingHelpers.slang:17-39): scene
// triangle index (-1: no hit), the hit's barycentric weights of vertices 1 and 2,
// and its ray distance (FLT_MAX: no hit).
struct VItem {
    int32_t tri;
    float bx, by, t;
};

// Orientation of the environment map for one frame (EnvMapData::transform and
// invTransform, LT/EnvMapData.slang:35-41): world = toWorld * local, row-major.
struct EnvXf {
    float toWorld[9], toLocal[9];
};

// SamplingOptions (VR/HostDeviceSharedDefinitions.h:100-138).
struct Options {
    uint32_t visibilityTrackingMethod, lightingTrackingMethod;
    int32_t lightSamples, lightingMipLevel, visibilitySamples, visibilityMipLevel;
    int32_t visibilityUseLinearSampler, lightingUseLinearSampler;
    float visibilityTStepScale, lightingTStepScale;
    int32_t useEnv, useAnalytic, useEmissive;
    int32_t vertexReuseStartBounce;
};

struct CameraFrame {
    float pos[3], U[3], V[3], W[3];
};

// Per-frame constants (VR/VolumetricReSTIR.cpp:453-496 and the kernels' cbuffers),
// passed by value as a kernel argument.
struct Frame {
    int32_t width, height;
    int32_t frameCount;
    int32_t numTotalRounds;
    int32_t maxBounces;
    int32_t initialM;
    int32_t useRussianRoulette;
    int32_t noReuse;
    int32_t coarserIndirect;
    int32_t useReference;
    int32_t baselineSpp;
    int32_t visualizeTransmittance;
    int32_t isFirstFrame;
    int32_t temporalMIS;
    int32_t reprojectionMode;
    int32_t reprojectionMip;
    float temporalThreshold;
    int32_t spatialMIS;
    int32_t spatialSampleCount;
    int32_t enableTemporal;
    int32_t enableSpatial;
    int32_t spatialRounds;
    int32_t outputMotionVec;
    int32_t useSurfaceScene;      // GenerateFeatures intersects the triangles and writes the V-buffer
    int32_t occSlot[2];           // grid slots whose occupancy tables the kernel keeps in LDS (-1: none)
    CameraFrame cam, prevCam;
    float prevView[16], prevProj[16];
    EnvXf env[2];                 // environment orientation: [0] this frame, [1] the previous frame
    int32_t hasAnimation;         // the volume advanced a frame of its sequence since the last frame
    int32_t prevGridOffset;       // kPrevDensityOffset when previous-frame grids serve isLastFrame lookups, else 0
    Options initialOpt, spatialOpt, finalOpt;
};
// A transmittance ray recorded by the record pass of a reuse kernel
// (world-space origin and direction, tMax, and which options trace it:
// 0 = the visibility options, 1 = the lighting options).
struct alignas(16) TraceRec {
    float ox, oy, oz, tMax;
    float dx, dy, dz;
    int32_t kind;
};
// Buffers of the record / trace / replay form of a reuse kernel.
struct SplitBufs {
    TraceRec* recs;   // cap records per pixel, in evaluation order
    float* trs;       // their transmittances: ray i of pixel p at i * pixels + p
    float* dens;      // the density lookups of the same evaluations: slot i of pixel p at i * pixels + p
    int32_t* counts;  // records per pixel
    uint32_t* plan;   // per pixel: bit s = candidate s's Talbot evaluations were recorded
                      // (TraceRays: the list of pixels its replay leaves to the one-kernel form)
    int32_t* fix;     // TraceRays: the length of that list
    int32_t cap;
    int32_t pad_;
};
// Density slots one pixel of split TraceRays fills (M candidates, path length B):
// 3 per free-flight sample and 4 (the generator's state) per traversal, with the
// camera ray's samples drawn in rounds of up to 4 and one sample per extra bounce;
// and one 32-bit word per 32 resampling steps (at most B + 1 per candidate).
constexpr int candStashSlots(int M, int B) {
    return 3 * M + 4 * ((M + 3) / 4) + 7 * M * (B - 1) + (M * (B + 1) + 31) / 32;
}
// Neighbour offsets of spatial reuse: [16 frame seeds][kMaxSpatialSamples][2] ints.
static const int kMaxSpatialSamples = 32;

// Learned free-flight proposal: FP32 parameters, bf16 MFMA operand fragments
// (64 lanes x 8 values each), the per-wave gradient partial, a training record.
static const int kLearnedParams = 12768;
static const int kLearnedFrags = 46;
static const int kLearnedPartialStride = 12800;  // gradient, loss, two second moments, padding
static const int kLearnedRecFloats = 72;
static const int kLearnedRaw = 8;
// What TraceRays needs of the learned proposal: each pixel's 16-bit CDF (32
// values) and the training records, one ray per bx x by block.
struct LearnedTraceArgs {
    const uint16_t* cdf;
    float* recs;
    int32_t bx, by;
};

}  // namespace vr
