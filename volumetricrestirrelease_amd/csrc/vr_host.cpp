// vr_host.cpp -- the C ABI of the pass (include/vrestir.h): handle lifecycle,
// scene upload, option surface, and the per-frame dispatch sequence of
// VolumetricReSTIR::execute (VR/VolumetricReSTIR.cpp:303-906) on a HIP stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/vrestir.h"
#include "vr_launch.h"
#include "vr_surface.h"
#include "vr_volume.h"

namespace vr {
struct ResStore { float4 a, b; };
}

using namespace vr;

static thread_local std::string g_createError;

// One frame of a volume sequence: the grid descriptors of slots 0..17 and their allocations.
// A device buffer kept across builds: reallocated only when a build needs more.
struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
};
static hipError_t growBuf(DevBuf& b, size_t bytes, void** out) {
    bytes = std::max<size_t>(bytes, 16);
    if (b.cap < bytes) {
        if (b.p) (void)hipFree(b.p);
        b.p = nullptr;
        b.cap = 0;
        hipError_t e = hipMalloc(&b.p, bytes);
        if (e != hipSuccess) return e;
        b.cap = bytes;
    }
    *out = b.p;
    return hipSuccess;
}
static void freeBuf(DevBuf& b) {
    if (b.p) (void)hipFree(b.p);
    b = DevBuf();
}
struct VolumeFrame {
    Grid grids[kFrameSlots];
    HostGrid meta[kFrameSlots];  // metadata only (pools are released after upload)
    std::vector<void*> allocs;
    DevBuf built[kFrameSlots][kBufKinds];  // the slots of a frame built on the GPU
    int numMips = 0;
    float tStep = 1.f, superVoxelDiag = 1.f;
};

struct vrestir_handle {
    int device = 0;
    std::string err;
    vrestir_params P;
    bool optionsChanged = true;
    bool offsetsDirty = true;
    // scene
    Scene sceneHost;
    Scene* sceneDev = nullptr;
    HostVolume meta;  // slot metadata only (pools are released after upload)
    bool volumeSet = false, camSet = false;
    vrestir_camera camDesc;
    CameraFrame cam{}, prevCam{};
    float view[16], proj[16], prevView[16], prevProj[16];
    bool havePrev = false;
    // frame state
    int W = 0, H = 0, allocB = 0;
    int frameCount = 0;
    bool temporalAccumulated = false;
    int totalRoundId = 0;
    Frame frame{};
    bool frameOpen = false;
    // buffers (VR/VolumetricReSTIR.cpp:419-450)
    ResStore* res[2] = {nullptr, nullptr};
    ResStore* resT = nullptr;
    float* extra[2] = {nullptr, nullptr};
    float* extraT = nullptr;
    float2* features = nullptr;
    float2* featuresT = nullptr;
    float4* color[2] = {nullptr, nullptr};
    float4* outInternal = nullptr;
    int* offsetsDev = nullptr;
    // ground-truth accumulation and relMSE scratch
    double4* accum = nullptr;
    int accumPixels = 0;
    int64_t accumFrames = 0;
    double* relPartial = nullptr;
    int relPartialN = 0;
    double* relOut = nullptr;
    // record / trace / replay buffers of the reuse stages
    int splitReuse = 2;  // 0: one kernel; 1: balanced walker queue; 2: brick-segment trace
    int splitTrace = 1;  // TraceRays' shadow rays as brick segments (with splitReuse == 2)
    bool streamed = false;  // the volume is a ring of two frames built on the GPU
    const vrestir_handle* sharedFrom = nullptr;  // the owner of the bound volume (vrestir_share_volume)
    DevBuf buildTmp[20];    // the GPU build's temporaries (grids per mip level, temperature, flags, scratch)
    bool inBatch = false;  // staged by a batch, whose TraceRays step is one kernel
    SplitBufs split{};
    size_t splitPixels = 0;
    // textured environment map: texels, importance mip chain, orientation now and at the last frame
    float* envTexDev = nullptr;
    float* envImpDev = nullptr;
    // the emissive triangles of vrestir_set_surfaces: collection, alias table, flux
    float* emTriDev = nullptr;
    uint32_t* emAliasDev = nullptr;
    float* emWeightDev = nullptr;
    // their ray-query structures
    void* bvhDev = nullptr;
    float* bvhTriDev = nullptr;
    float* triDataDev = nullptr;
    float* materialsDev = nullptr;
    VItem* vbuf = nullptr;  // the V-buffer of a surface scene
    EnvXf envCur{}, envLast{};
    bool haveEnvLast = false;
    // the volume: frames of a sequence (one for a static volume), each resident in HBM
    std::vector<VolumeFrame> frames;
    int curFrame = 0;
    bool hasAnimation = false;
    bool freezeFrame = false;
    // learned free-flight proposal: FP32 master weights and Adam moments, their
    // bf16 MFMA fragments, gradient partials, statistics and the judge's flag,
    // the per-pixel CDF and the training records
    vrestir_learned_desc L{};
    float* lPrm = nullptr;
    float* lM = nullptr;
    float* lV = nullptr;
    uint4* lFrags = nullptr;
    float* lPartial = nullptr;
    int lPartialWaves = 0;
    float* lStats = nullptr;
    int* lInUse = nullptr;
    uint16_t* lCdf = nullptr;
    size_t lCdfPixels = 0;
    float* lRecs = nullptr;
    int lRecSlots = 0;
    int lStep = 0;
};

static int fail(vrestir_handle* h, int code, const std::string& msg) {
    if (h) h->err = msg;
    else g_createError = msg;
    return code;
}
#define HIPCHK(h, call)                                                                        \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return fail((h), VRESTIR_EHIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <class T>
static void dfree(T*& p) {
    if (p) (void)hipFree((void*)p);
    p = nullptr;
}

static void freeLearned(vrestir_handle* h) {
    dfree(h->lPrm); dfree(h->lM); dfree(h->lV); dfree(h->lFrags); dfree(h->lPartial); dfree(h->lStats);
    dfree(h->lInUse); dfree(h->lCdf); dfree(h->lRecs);
    h->lPartialWaves = 0; h->lCdfPixels = 0; h->lRecSlots = 0;
}

static void freeSplit(vrestir_handle* h) {
    dfree(h->split.recs); dfree(h->split.trs); dfree(h->split.dens); dfree(h->split.counts); dfree(h->split.plan);
    dfree(h->split.fix);
    h->split.cap = 0;
    h->splitPixels = 0;
}
static void freeFrameBuffers(vrestir_handle* h) {
    freeSplit(h);
    for (int i = 0; i < 2; i++) { dfree(h->res[i]); dfree(h->extra[i]); dfree(h->color[i]); }
    dfree(h->resT); dfree(h->extraT); dfree(h->features); dfree(h->featuresT); dfree(h->vbuf); dfree(h->outInternal);
    h->W = h->H = 0;
}
static void freeVolume(vrestir_handle* h) {
    for (VolumeFrame& f : h->frames) {
        for (void* p : f.allocs) (void)hipFree(p);
        for (auto& slot : f.built)
            for (DevBuf& b : slot) freeBuf(b);
    }
    h->frames.clear();
    for (DevBuf& b : h->buildTmp) freeBuf(b);
    h->curFrame = 0;
    h->hasAnimation = false;
}
static void freeEnv(vrestir_handle* h) {
    dfree(h->envTexDev);
    dfree(h->envImpDev);
}
static void freeSurfaces(vrestir_handle* h) {
    dfree(h->emTriDev);
    dfree(h->emAliasDev);
    dfree(h->emWeightDev);
    dfree(h->bvhDev);
    dfree(h->bvhTriDev);
    dfree(h->triDataDev);
    dfree(h->materialsDev);
    Scene& S = h->sceneHost;
    S.bvh = nullptr;
    S.bvhTri = S.triData = S.materials = nullptr;
    S.numTris = S.numMaterials = 0;
    S.emTri = nullptr;
    S.emAlias = nullptr;
    S.emWeight = nullptr;
    S.numEmTris = S.numEmActive = 0;
    S.emWeightSum = 0.f;
    S.emMult = 1.f;
}
// EnvMap::setRotation (LT/EnvMap.cpp:57-72): local -> world = Rz * Ry * Rx of
// glm::eulerAngleX/Y/Z (right-handed rotations by the XYZ angles in degrees),
// computed in double and rounded once; world -> local is its transpose.
static void envRotation(const float deg[3], EnvXf& X) {
    const double k = 3.14159265358979323846 / 180.0;
    const double cx = std::cos(deg[0] * k), sx = std::sin(deg[0] * k);
    const double cy = std::cos(deg[1] * k), sy = std::sin(deg[1] * k);
    const double cz = std::cos(deg[2] * k), sz = std::sin(deg[2] * k);
    const double Rx[9] = {1, 0, 0, 0, cx, -sx, 0, sx, cx};
    const double Ry[9] = {cy, 0, sy, 0, 1, 0, -sy, 0, cy};
    const double Rz[9] = {cz, -sz, 0, sz, cz, 0, 0, 0, 1};
    double T[9], R[9];
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) {
            T[r * 3 + c] = Ry[r * 3] * Rx[c] + Ry[r * 3 + 1] * Rx[3 + c] + Ry[r * 3 + 2] * Rx[6 + c];
        }
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) R[r * 3 + c] = Rz[r * 3] * T[c] + Rz[r * 3 + 1] * T[3 + c] + Rz[r * 3 + 2] * T[6 + c];
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) {
            X.toWorld[r * 3 + c] = (float)R[r * 3 + c];
            X.toLocal[r * 3 + c] = (float)R[c * 3 + r];
        }
}

// ------------------------------------------------------------------ params
extern "C" int32_t vrestir_version(void) { return VRESTIR_ABI_VERSION; }

// Defaults of VolumetricReSTIRParams (VR/VolumetricReSTIR.h:132-206).
extern "C" void vrestir_default_params(vrestir_params* p) {
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->maxBounces = 1;
    p->enableTemporalReuse = 1;
    p->enableSpatialReuse = 1;
    p->vertexReuseStartBounce = 1;
    p->useEnvironmentLights = 1;
    p->baselineSamplePerPixel = 1;
    p->usePrevVolumeForReproj = 1;
    p->initialBaseMipLevel = 1;
    p->initialM = 4;
    p->initialLightSamples = 1;
    p->initialLightingMipLevel = 2;
    p->initialLightingUseLinearSampler = 1;
    p->initialLightingTrackingMethod = VRESTIR_RAY_MARCHING;
    p->initialVisibilityTStepScale = 1.f;
    p->initialLightingTStepScale = 2.f;
    p->initialUseRussianRoulette = 1;
    p->initialUseCoarserGridForIndirectBounce = 1;
    p->temporalReuseMThreshold = 4.f;
    p->temporalReprojectionMode = VRESTIR_REPROJECTION_LINEAR;
    p->temporalMISMethod = VRESTIR_MIS_TALBOT;
    p->temporalReprojectionMipLevel = 1;
    p->spatialReuseRounds = 1;
    p->spatialVisibilityMipLevel = 1;
    p->spatialLightingMipLevel = 1;
    p->spatialVisibilityUseLinearSampler = 1;
    p->spatialLightingUseLinearSampler = 1;
    p->spatialVisibilityTStepScale = 1.f;
    p->spatialLightingTStepScale = 1.f;
    p->spatialVisibilityTrackingMethod = VRESTIR_RAY_MARCHING;
    p->spatialLightingTrackingMethod = VRESTIR_RAY_MARCHING;
    p->randomSamplerType = VRESTIR_SAMPLER_R2;
    p->sampleRadius = 10.f;
    p->spatialSampleCount = 4;
    p->spatialMISMethod = VRESTIR_MIS_TALBOT;
    p->finalLightSamples = 1;
    p->finalVisibilitySamples = 1;
    p->finalVisibilityTrackingMethod = VRESTIR_ANALYTIC_TRACKING;
    p->finalLightTrackingMethod = VRESTIR_ANALYTIC_TRACKING;
    p->finalRandomSamplerType = VRESTIR_SAMPLER_R2;
    p->finalTStepScale = 0.2f;
}

static int checkParams(vrestir_handle* h, const vrestir_params* p) {
    if (!p) return fail(h, VRESTIR_EINVAL, "params is NULL");
    if (p->vertexReuse) return fail(h, VRESTIR_EUNSUPPORTED, "vertex reuse is not implemented");
    if (p->maxBounces < 1 || p->maxBounces > 4) return fail(h, VRESTIR_EINVAL, "maxBounces must be in [1, 4]");
    if (p->initialM < 1) return fail(h, VRESTIR_EINVAL, "initialM must be >= 1");
    if (p->spatialSampleCount < 1 || p->spatialSampleCount > kMaxSpatialSamples)
        return fail(h, VRESTIR_EINVAL, "spatialSampleCount must be in [1, 32]");
    if (p->spatialReuseRounds < 0) return fail(h, VRESTIR_EINVAL, "spatialReuseRounds must be >= 0");
    if (p->baselineSamplePerPixel < 1) return fail(h, VRESTIR_EINVAL, "baselineSamplePerPixel must be >= 1");
    auto method = [](uint32_t m) { return m <= 4u; };
    if (!method(p->initialLightingTrackingMethod) || !method(p->spatialVisibilityTrackingMethod) ||
        !method(p->spatialLightingTrackingMethod) || !method(p->finalVisibilityTrackingMethod) ||
        !method(p->finalLightTrackingMethod))
        return fail(h, VRESTIR_EINVAL, "tracking method out of range");
    if (p->initialBaseMipLevel < 0 || p->initialBaseMipLevel > 7 || p->initialLightingMipLevel < 0 ||
        p->initialLightingMipLevel > 7 || p->spatialVisibilityMipLevel < 0 || p->spatialVisibilityMipLevel > 7 ||
        p->spatialLightingMipLevel < 0 || p->spatialLightingMipLevel > 7 || p->temporalReprojectionMipLevel < 0 ||
        p->temporalReprojectionMipLevel > 7)
        return fail(h, VRESTIR_EINVAL, "mip level out of range");
    return VRESTIR_OK;
}

extern "C" int32_t vrestir_create(const vrestir_params* params, vrestir_handle** out) {
    if (!out) return fail(nullptr, VRESTIR_EINVAL, "out is NULL");
    *out = nullptr;
    vrestir_params p;
    if (params) p = *params;
    else vrestir_default_params(&p);
    int rc = checkParams(nullptr, &p);
    if (rc) return rc;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return fail(nullptr, VRESTIR_EHIP, std::string("hipGetDevice: ") + hipGetErrorString(e));
    vrestir_handle* h = new vrestir_handle();
    h->device = dev;
    h->P = p;
    std::memset(&h->sceneHost, 0, sizeof(Scene));
    h->sceneHost.emMult = 1.f;
    const float zero[3] = {0.f, 0.f, 0.f};
    envRotation(zero, h->envCur);
    h->envLast = h->envCur;
    if ((e = relMseScratchInit(nullptr)) != hipSuccess) {
        delete h;
        return fail(nullptr, VRESTIR_EHIP, std::string("scratch address: ") + hipGetErrorString(e));
    }
    if (hipMalloc((void**)&h->sceneDev, sizeof(Scene)) != hipSuccess ||
        hipMalloc((void**)&h->offsetsDev, sizeof(int) * 16 * kMaxSpatialSamples * 2) != hipSuccess) {
        dfree(h->sceneDev);
        delete h;
        return fail(nullptr, VRESTIR_ENOMEM, "device allocation failed");
    }
    *out = h;
    return VRESTIR_OK;
}

extern "C" int32_t vrestir_destroy(vrestir_handle* h) {
    if (!h) return VRESTIR_OK;
    (void)hipSetDevice(h->device);
    (void)hipDeviceSynchronize();
    freeFrameBuffers(h);
    freeVolume(h);
    freeEnv(h);
    freeSurfaces(h);
    dfree(h->sceneDev);
    dfree(h->offsetsDev);
    dfree(h->accum);
    dfree(h->relPartial);
    dfree(h->relOut);
    freeSplit(h);
    freeLearned(h);
    delete h;
    return VRESTIR_OK;
}

extern "C" const char* vrestir_last_error(const vrestir_handle* h) {
    return h ? h->err.c_str() : g_createError.c_str();
}

extern "C" int32_t vrestir_set_params(vrestir_handle* h, const vrestir_params* p) {
    if (!h) return VRESTIR_EINVAL;
    int rc = checkParams(h, p);
    if (rc) return rc;
    h->P = *p;
    h->optionsChanged = true;
    h->offsetsDirty = true;
    return VRESTIR_OK;
}
extern "C" int32_t vrestir_get_params(const vrestir_handle* h, vrestir_params* out) {
    if (!h || !out) return VRESTIR_EINVAL;
    *out = h->P;
    return VRESTIR_OK;
}
extern "C" int32_t vrestir_reset_history(vrestir_handle* h) {
    if (!h) return VRESTIR_EINVAL;
    h->optionsChanged = true;
    return VRESTIR_OK;
}
extern "C" int32_t vrestir_frame_count(const vrestir_handle* h, int32_t* out) {
    if (!h || !out) return VRESTIR_EINVAL;
    *out = h->frameCount;
    return VRESTIR_OK;
}

// ------------------------------------------------------------------ scene
static int uploadScene(vrestir_handle* h) {
    HIPCHK(h, hipSetDevice(h->device));
    HIPCHK(h, hipDeviceSynchronize());
    HIPCHK(h, hipMemcpy(h->sceneDev, &h->sceneHost, sizeof(Scene), hipMemcpyHostToDevice));
    return VRESTIR_OK;
}

// Uploads one built slot into G (metadata into m); device allocations are recorded in allocs.
static int uploadGrid(vrestir_handle* h, Grid& G, HostGrid& m, HostGrid& hg, std::vector<void*>& allocs) {
    std::memset(&G, 0, sizeof(Grid));
    m = HostGrid();
    if (!hg.valid) return VRESTIR_OK;
    auto up = [&](const void* src, size_t bytes, const void** dst) -> int {
        void* p = nullptr;
        HIPCHK(h, hipMalloc(&p, std::max<size_t>(bytes, 16)));
        allocs.push_back(p);
        if (bytes) HIPCHK(h, hipMemcpy(p, src, bytes, hipMemcpyHostToDevice));
        *dst = p;
        return VRESTIR_OK;
    };
    int rc;
    const void* p;
    if ((rc = up(hg.brickTable.data(), hg.brickTable.size() * 4, &p))) return rc;
    G.brickTable = (const uint32_t*)p;
    if ((rc = up(hg.nodeTable.data(), hg.nodeTable.size() * 4, &p))) return rc;
    G.nodeTable = (const uint32_t*)p;
    if ((rc = up(hg.occ.data(), hg.occ.size() * 4, &p))) return rc;
    G.occ = (const uint32_t*)p;
    G.occWords = (int32_t)(hg.occ.size() / 2);
    G.numBricks = hg.numBricks;
    if ((rc = up(hg.bounds.data(), hg.bounds.size() * 4, &p))) return rc;
    G.bounds = (const float*)p;
    if (hg.fp16) { if ((rc = up(hg.pool16.data(), hg.pool16.size() * 2, &p))) return rc; }
    else { if ((rc = up(hg.pool32.data(), hg.pool32.size() * 4, &p))) return rc; }
    G.pool = p;
    G.nbx = hg.nbx; G.nby = hg.nby; G.nbz = hg.nbz;
    G.nnx = hg.nnx; G.nny = hg.nny; G.nnz = hg.nnz;
    G.bmax[0] = (float)hg.nx; G.bmax[1] = (float)hg.ny; G.bmax[2] = (float)hg.nz;
    std::memcpy(G.w2i, hg.w2i, sizeof(G.w2i));
    G.maxValue = hg.maxValue;
    G.fp16 = hg.fp16 ? 1 : 0;
    G.channels = hg.channels;
    G.valid = 1;
    // keep metadata, drop the host copies of the pools
    m.valid = true; m.nx = hg.nx; m.ny = hg.ny; m.nz = hg.nz; m.numBricks = hg.numBricks;
    m.maxValue = hg.maxValue; m.fp16 = hg.fp16; m.channels = hg.channels;
    m.nbx = hg.nbx; m.nby = hg.nby; m.nbz = hg.nbz;
    m.nnx = hg.nnx; m.nny = hg.nny; m.nnz = hg.nnz;
    return VRESTIR_OK;
}

static int uploadFrame(vrestir_handle* h, HostVolume& hv, VolumeFrame& f) {
    f.numMips = hv.numMips;
    f.tStep = hv.tStep;
    f.superVoxelDiag = hv.superVoxelDiag;
    for (int s = 0; s < kFrameSlots; s++) {
        int rc = uploadGrid(h, f.grids[s], f.meta[s], hv.slots[s], f.allocs);
        if (rc) return rc;
    }
    return VRESTIR_OK;
}

// Binds frame `cur` of the sequence and, in the previous-frame slots, the frame
// before it in sequence order (the first frame's previous is the last,
// SC/Scene.cpp:3342-3350). animating: the sequence advanced since the last frame
// (VolumeDesc::hasAnimation, SC/Scene.cpp:825-863).
static int bindFrame(vrestir_handle* h, int cur, bool animating) {
    const int n = (int)h->frames.size();
    const VolumeFrame& c = h->frames[cur];
    const VolumeFrame& pv = h->frames[(cur + n - 1) % n];
    Scene& S = h->sceneHost;
    for (int s = 0; s < kMaxSlots; s++) { std::memset(&S.grid[s], 0, sizeof(Grid)); h->meta.slots[s] = HostGrid(); }
    for (int s = 0; s < kFrameSlots; s++) { S.grid[s] = c.grids[s]; h->meta.slots[s] = c.meta[s]; }
    for (int k = 0; k < 8; k++) {
        S.grid[kPrevDensityOffset + k] = pv.grids[k];
        h->meta.slots[kPrevDensityOffset + k] = pv.meta[k];
    }
    for (int s : {kTemperatureSlot, kVelocitySlot}) {
        S.grid[s + kPrevExtraOffset] = pv.grids[s];
        h->meta.slots[s + kPrevExtraOffset] = pv.meta[s];
    }
    h->meta.numMips = c.numMips;
    S.numMips = c.numMips;
    S.tStep = c.tStep;
    S.superVoxelDiag = c.superVoxelDiag;
    S.hasEmission = c.grids[kTemperatureSlot].valid ? 1 : 0;
    S.prevHasEmission = pv.grids[kTemperatureSlot].valid ? 1 : 0;
    S.hasVelocity = c.grids[kVelocitySlot].valid ? 1 : 0;
    h->curFrame = cur;
    h->hasAnimation = animating;
    return uploadScene(h);
}

// The material of a volume (a sequence keeps its first frame's).
static void setMaterial(vrestir_handle* h, const vrestir_volume_desc* d) {
    Scene& S = h->sceneHost;
    S.sigma_t = d->sigmaS[0] + d->sigmaA[0];
    for (int i = 0; i < 3; i++) { S.sigma_s[i] = d->sigmaS[i]; S.sigma_a[i] = d->sigmaA[i]; }
    S.g = d->g;
    S.densityScale = d->densityScale;
    S.worldScaling = 1.f;
    // addGVDBVolume(hasEmission, LeScale, temperatureCutOff, temperatureScale), SC/Scene.cpp:2754-2790
    S.LeScale = d->LeScale;
    S.tempCutoff = d->temperatureCutOff;
    S.tempScale = d->temperatureScale;
    S.velocityScale = d->velocityScale;
}

// Uploads a built volume as frame 0 of a new sequence with the material of d.
static int installVolume(vrestir_handle* h, HostVolume& hv, const vrestir_volume_desc* d) {
    int rc;
    HIPCHK(h, hipSetDevice(h->device));
    HIPCHK(h, hipDeviceSynchronize());
    freeVolume(h);
    h->streamed = false;
    h->sharedFrom = nullptr;
    h->frames.emplace_back();
    if ((rc = uploadFrame(h, hv, h->frames.back()))) return rc;
    setMaterial(h, d);
    h->volumeSet = true;
    h->optionsChanged = true;
    return bindFrame(h, 0, false);
}
extern "C" int32_t vrestir_set_volume(vrestir_handle* h, const vrestir_volume_desc* d) {
    if (!h || !d) return VRESTIR_EINVAL;
    HostVolume hv;
    int rc = buildVolume(*d, hv, h->err);
    if (rc) return rc;
    return installVolume(h, hv, d);
}

// ------------------------------------------------------------------ volumes built on the GPU
// (no counterpart in the reference, which loads every frame from prebuilt files)
// Frame f from dense device grids: the slots buildVolume makes (mips, conservative
// grids, temperature), built by the kernels of vr_k_build.hip value for value.
static void freeFrame(VolumeFrame& f) {
    for (void* p : f.allocs) (void)hipFree(p);
    for (auto& slot : f.built)
        for (DevBuf& b : slot) freeBuf(b);
    f = VolumeFrame();
}
static int buildFrameDevice(vrestir_handle* h, hipStream_t st, const vrestir_volume_desc& d, VolumeFrame& f) {
    int rc = checkVolumeDesc(d, h->err);
    if (rc) return rc;
    if (d.velocity) return fail(h, VRESTIR_EUNSUPPORTED, "a volume built on the GPU takes no velocity grid");
    // the frame keeps its slots' buffers and the handle its temporaries from one
    // build to the next: a streamed frame of the same volume reallocates nothing
    // unless a slot needs more bricks than before
    for (void* p : f.allocs) (void)hipFree(p);
    f.allocs.clear();
    for (int s = 0; s < kFrameSlots; s++) {
        std::memset(&f.grids[s], 0, sizeof(Grid));
        f.meta[s] = HostGrid();
    }
    const bool fp16 = d.brickFormat == VRESTIR_BRICK_FP16;
    const size_t n = (size_t)d.nx * d.ny * d.nz;
    auto tmp = [&](int key, size_t bytes, void** p) -> int {
        HIPCHK(h, growBuf(h->buildTmp[key], bytes, p));
        return VRESTIR_OK;
    };
    auto done = [&](int r) {
        (void)hipStreamSynchronize(st);
        if (r) freeFrame(f);
        return r;
    };
    float *cur = nullptr, *cons = nullptr;
    // per-build counters: [0] and [19] non-finite density / temperature values; the
    // largest value of each grid as float bits: mip k at [1 + k], conservative mip k
    // at [9 + k], temperature at [17] (each left by the kernel that writes the grid)
    unsigned* stats = nullptr;
    void* scratch = nullptr;
    if ((rc = tmp(0, n * 4, (void**)&cur)) || (rc = tmp(1, n * 4, (void**)&cons)) ||
        (rc = tmp(17, 20 * sizeof(unsigned), (void**)&stats)) ||
        (rc = tmp(18, buildScratchBytes(d.nx, d.ny, d.nz), &scratch)))
        return done(rc);
    HIPCHK(h, hipMemsetAsync(stats, 0, 20 * sizeof(unsigned), st));
    // values are clamped to >= 0; a non-finite one is refused, as the host build refuses it
    auto clampChecked = [&](const float* src, float* dst, int badAt, int maxAt, const char* what) -> int {
        int nbad = 0;
        HIPCHK(h, buildClamp(st, src, dst, n, (int*)stats + badAt, stats + maxAt));
        HIPCHK(h, hipMemcpyAsync(&nbad, stats + badAt, 4, hipMemcpyDeviceToHost, st));
        HIPCHK(h, hipStreamSynchronize(st));
        if (nbad) return fail(h, VRESTIR_EINVAL, std::string("the ") + what + " grid holds a non-finite value");
        return VRESTIR_OK;
    };
    auto slot = [&](int s, const float* g, int nx, int ny, int nz, int scale, int maxAt) -> int {
        uint32_t nb = 0;
        float mv = 0.f;
        const BuildAlloc alloc = [&f, s](int kind, size_t bytes, void** p) { return growBuf(f.built[s][kind], bytes, p); };
        HIPCHK(h, buildSlotDevice(st, g, nx, ny, nz, stats + maxAt, fp16, scratch, f.grids[s], nb, mv, alloc));
        mipTransform(d.indexToWorld, scale, f.grids[s].w2i);
        HostGrid& m = f.meta[s];
        m = HostGrid();
        m.valid = true; m.nx = nx; m.ny = ny; m.nz = nz; m.numBricks = nb; m.maxValue = mv; m.fp16 = fp16;
        m.channels = 1; m.nbx = nx / 8; m.nby = ny / 8; m.nbz = nz / 8;
        m.nnx = f.grids[s].nnx; m.nny = f.grids[s].nny; m.nnz = f.grids[s].nnz;
        return VRESTIR_OK;
    };
    // a failed launch leaves through done(), which releases the temporaries
    auto ck = [&](hipError_t e, const char* what) -> int {
        return e == hipSuccess ? VRESTIR_OK : fail(h, VRESTIR_EHIP, std::string(what) + ": " + hipGetErrorString(e));
    };
    if ((rc = clampChecked(d.density, cur, 0, 1, "density"))) return done(rc);
    if ((rc = ck(buildDilate(st, cur, cons, d.nx, d.ny, d.nz, stats + 9), "buildDilate"))) return done(rc);
    int nx = d.nx, ny = d.ny, nz = d.nz;
    for (int k = 0; k < d.numMips; k++) {
        if (k > 0) {
            float *n2 = nullptr, *c2 = nullptr;
            const size_t m = (size_t)(nx / 2) * (ny / 2) * (nz / 2);
            if ((rc = tmp(2 * k, m * 4, (void**)&n2)) || (rc = tmp(2 * k + 1, m * 4, (void**)&c2))) return done(rc);
            if ((rc = ck(buildReduce2(st, cur, n2, nx, ny, nz, stats + 1 + k), "buildReduce2")) ||
                (rc = ck(buildReduce2(st, cons, c2, nx, ny, nz, stats + 9 + k), "buildReduce2")))
                return done(rc);
            cur = n2; cons = c2;
            nx /= 2; ny /= 2; nz /= 2;
        }
        if ((rc = slot(k, cur, nx, ny, nz, 1 << k, 1 + k)) || (rc = slot(8 + k, cons, nx, ny, nz, 1 << k, 9 + k)))
            return done(rc);
    }
    if (d.temperature) {
        // one full-resolution slot with mip 0's transform
        float* t = nullptr;
        if ((rc = tmp(16, n * 4, (void**)&t))) return done(rc);
        if ((rc = clampChecked(d.temperature, t, 19, 17, "temperature"))) return done(rc);
        if ((rc = slot(kTemperatureSlot, t, d.nx, d.ny, d.nz, 1, 17))) return done(rc);
    }
    HostVolume hv;
    volumeSpacing(d.indexToWorld, hv);
    f.numMips = d.numMips;
    f.tStep = hv.tStep;
    f.superVoxelDiag = hv.superVoxelDiag;
    return done(VRESTIR_OK);
}

// Binds another handle's volume: its current and previous frame's grids, mips,
// material and animation state, read-only (the owner keeps the memory). Lights,
// environment, camera and options stay h's own. Copied at the call, so a view
// of a streamed volume calls it again after each vrestir_stream_volume_frame of
// the owner.
extern "C" int32_t vrestir_share_volume(vrestir_handle* h, const vrestir_handle* owner) {
    if (!h || !owner || h == owner) return VRESTIR_EINVAL;
    if (!owner->volumeSet) return fail(h, VRESTIR_ESTATE, "the owner has no volume");
    if (owner->device != h->device) return fail(h, VRESTIR_EINVAL, "the owner is on another device");
    if (!h->frames.empty()) {
        HIPCHK(h, hipSetDevice(h->device));
        HIPCHK(h, hipDeviceSynchronize());
        freeVolume(h);
    }
    h->streamed = false;
    Scene& S = h->sceneHost;
    const Scene& O = owner->sceneHost;
    std::memcpy(S.grid, O.grid, sizeof(S.grid));
    S.sigma_t = O.sigma_t;
    std::memcpy(S.sigma_s, O.sigma_s, sizeof(S.sigma_s));
    std::memcpy(S.sigma_a, O.sigma_a, sizeof(S.sigma_a));
    S.g = O.g;
    S.densityScale = O.densityScale;
    S.tStep = O.tStep;
    S.worldScaling = O.worldScaling;
    S.superVoxelDiag = O.superVoxelDiag;
    S.numMips = O.numMips;
    S.hasEmission = O.hasEmission;
    S.LeScale = O.LeScale;
    S.tempCutoff = O.tempCutoff;
    S.tempScale = O.tempScale;
    S.prevHasEmission = O.prevHasEmission;
    S.hasVelocity = O.hasVelocity;
    S.velocityScale = O.velocityScale;
    h->meta = owner->meta;
    h->hasAnimation = owner->hasAnimation;
    h->volumeSet = true;
    // a new volume resets the history; the owner's next frame is an animation step
    if (h->sharedFrom != owner) h->optionsChanged = true;
    h->sharedFrom = owner;
    return uploadScene(h);
}

extern "C" int32_t vrestir_stream_volume_frame(vrestir_handle* h, void* stream, const vrestir_volume_desc* d) {
    if (!h || !d) return VRESTIR_EINVAL;
    hipStream_t st = (hipStream_t)stream;
    HIPCHK(h, hipSetDevice(h->device));
    // the frame rebuilt below may still be read by work queued on any stream
    HIPCHK(h, hipDeviceSynchronize());
    int rc;
    if (!h->volumeSet || !h->streamed) {
        freeVolume(h);
        h->frames.emplace_back();
        if ((rc = buildFrameDevice(h, st, *d, h->frames[0]))) {
            freeVolume(h);
            h->volumeSet = false;
            return rc;
        }
        setMaterial(h, d);
        h->streamed = true;
        h->sharedFrom = nullptr;
        h->volumeSet = true;
        h->optionsChanged = true;
        return bindFrame(h, 0, false);
    }
    const HostGrid& m0 = h->frames[h->curFrame].meta[0];
    if (d->nx != m0.nx || d->ny != m0.ny || d->nz != m0.nz || d->numMips != h->frames[h->curFrame].numMips ||
        (d->brickFormat == VRESTIR_BRICK_FP16) != m0.fp16 ||
        (d->temperature != nullptr) != (h->frames[h->curFrame].grids[kTemperatureSlot].valid != 0))
        return fail(h, VRESTIR_EINVAL, "every streamed frame needs the first frame's dimensions, mips, format and grids");
    int target = 1 - h->curFrame;
    if (h->frames.size() == 1) {
        h->frames.emplace_back();
        target = 1;
    }
    if ((rc = buildFrameDevice(h, st, *d, h->frames[target]))) {
        // the bound scene's previous-frame slots named the frame just released:
        // keep the current frame alone, as a one-frame sequence
        const std::string err = h->err;
        h->frames.erase(h->frames.begin() + target);
        (void)bindFrame(h, 0, false);
        h->err = err;
        return rc;
    }
    return bindFrame(h, target, true);
}

// addGVDBVolumeSequence (SC/Scene.cpp:3329-3355): one more frame of the sequence,
// kept resident in HBM. Its material (sigma, g, density and emission scales)
// is the first frame's, as the reference keeps them across frames (SC/Scene.cpp:829-855).
extern "C" int32_t vrestir_add_volume_frame(vrestir_handle* h, const vrestir_volume_desc* d, int32_t* index) {
    if (!h || !d) return VRESTIR_EINVAL;
    if (!h->volumeSet) return fail(h, VRESTIR_ESTATE, "bind the first frame with vrestir_set_volume");
    if (h->streamed) return fail(h, VRESTIR_ESTATE, "a streamed volume takes frames from vrestir_stream_volume_frame");
    if (h->frames.empty()) return fail(h, VRESTIR_ESTATE, "the bound volume is another handle's (vrestir_share_volume)");
    if (d->numMips != h->frames[0].numMips) return fail(h, VRESTIR_EINVAL, "every frame needs the first frame's numMips");
    HostVolume hv;
    int rc = buildVolume(*d, hv, h->err);
    if (rc) return rc;
    HIPCHK(h, hipSetDevice(h->device));
    h->frames.emplace_back();
    if ((rc = uploadFrame(h, hv, h->frames.back()))) return rc;
    if (index) *index = (int32_t)h->frames.size() - 1;
    return bindFrame(h, h->curFrame, h->hasAnimation);
}

// The scene's per-frame VDB update (SC/Scene.cpp:825-863): frame `index` becomes
// current; animating = 0 is the paused state (mPauseVDBAnimation), in which the
// previous-frame grids and the velocity grid are not used.
extern "C" int32_t vrestir_set_volume_frame(vrestir_handle* h, int32_t index, int32_t animating) {
    if (!h) return VRESTIR_EINVAL;
    if (!h->volumeSet) return fail(h, VRESTIR_ESTATE, "no volume bound");
    if (index < 0 || index >= (int)h->frames.size()) return fail(h, VRESTIR_EINVAL, "frame index out of range");
    HIPCHK(h, hipSetDevice(h->device));
    HIPCHK(h, hipDeviceSynchronize());
    return bindFrame(h, index, animating != 0);
}

extern "C" int32_t vrestir_volume_frame_count(const vrestir_handle* h, int32_t* count) {
    if (!h || !count) return VRESTIR_EINVAL;
    *count = (int32_t)h->frames.size();
    return VRESTIR_OK;
}

// <prefix>_temperature.vbx next to <prefix>_mip0.vbx (SC/Scene.cpp:2789-2791, 2808-2811),
// attached to the current frame.
extern "C" int32_t vrestir_load_vbx_emission(vrestir_handle* h, const char* pathPrefix, float LeScale,
                                             float temperatureCutOff, float temperatureScale) {
    if (!h || !pathPrefix) return VRESTIR_EINVAL;
    if (!h->volumeSet) return fail(h, VRESTIR_ESTATE, "bind the density volume first");
    if (h->frames.empty()) return fail(h, VRESTIR_ESTATE, "the bound volume is another handle's (vrestir_share_volume)");
    std::string path = std::string(pathPrefix) + "_temperature.vbx";
    int nx = 0, ny = 0, nz = 0;
    float i2w[16];
    int rc = vrestir_read_vbx(path.c_str(), nullptr, &nx, &ny, &nz, i2w);
    if (rc) return fail(h, rc, "cannot read " + path);
    std::vector<float> t((size_t)nx * ny * nz);
    rc = vrestir_read_vbx(path.c_str(), t.data(), &nx, &ny, &nz, i2w);
    if (rc) return fail(h, rc, "cannot read " + path);
    HostGrid hg;
    VolumeFrame& f = h->frames[h->curFrame];
    rc = buildTemperatureSlot(t.data(), nx, ny, nz, i2w, f.meta[0].fp16, hg, h->err);
    if (rc) return rc;
    HIPCHK(h, hipSetDevice(h->device));
    HIPCHK(h, hipDeviceSynchronize());
    if ((rc = uploadGrid(h, f.grids[kTemperatureSlot], f.meta[kTemperatureSlot], hg, f.allocs))) return rc;
    Scene& S = h->sceneHost;
    S.LeScale = LeScale;
    S.tempCutoff = temperatureCutOff;
    S.tempScale = temperatureScale;
    return bindFrame(h, h->curFrame, h->hasAnimation);
}

// The blackbody table of SC/Scene.cpp:2877-2893 (LUT/BlackBodyRadiationRGB_50K-6400K.txt,
// 128 rows of "r g b"), handed over by the caller.
extern "C" int32_t vrestir_set_blackbody_lut(vrestir_handle* h, const float* rgb, int32_t rows) {
    if (!h) return VRESTIR_EINVAL;
    if (rows < 1 || rows > kMaxLutRows || !rgb) return fail(h, VRESTIR_EINVAL, "blackbody table needs 1..256 rows");
    for (int i = 0; i < 3 * rows; i++)
        if (!std::isfinite(rgb[i])) return fail(h, VRESTIR_EINVAL, "blackbody table holds a non-finite value");
    Scene& S = h->sceneHost;
    std::memset(S.lut, 0, sizeof(S.lut));
    std::memcpy(S.lut, rgb, sizeof(float) * 3 * rows);
    S.lutRows = rows;
    HIPCHK(h, hipSetDevice(h->device));
    HIPCHK(h, hipDeviceSynchronize());
    return uploadScene(h);
}

// ------------------------------------------------------------------ environment map
// EnvMap::create + EnvMapSampler::create (LT/EnvMap.cpp:107-122, LT/EnvMapSampler.cpp:61-110):
// texels to HBM as RGBA32F, the importance mip chain built on the device.
extern "C" int32_t vrestir_set_env_map(vrestir_handle* h, const vrestir_env_map_desc* d) {
    if (!h) return VRESTIR_EINVAL;
    Scene& S = h->sceneHost;
    if (!d) {
        HIPCHK(h, hipSetDevice(h->device));
        HIPCHK(h, hipDeviceSynchronize());
        freeEnv(h);
        S.envTextured = 0;
        S.envTex = S.envImp = nullptr;
        return uploadScene(h);
    }
    const int W = d->width, H = d->height;
    if (W < 1 || H < 1 || (int64_t)W * H > (int64_t)1 << 26 || !d->rgb)
        return fail(h, VRESTIR_EINVAL, "environment map needs 1..2^26 texels and rgb data");
    const int dim = d->importanceDim ? d->importanceDim : 512;
    const int spp = d->importanceSpp ? d->importanceSpp : 64;
    if (dim < 2 || dim > 2048 || (dim & (dim - 1)))
        return fail(h, VRESTIR_EINVAL, "importanceDim must be a power of two in [2, 2048]");
    int sx = 1;
    while ((sx * 2) * (sx * 2) <= spp) sx *= 2;
    if (spp < 1 || spp > 1024 || (spp & (spp - 1)) || sx * sx != spp)
        return fail(h, VRESTIR_EINVAL, "importanceSpp must be 1, 4, 16, 64, 256 or 1024");
    std::vector<float> rgba((size_t)W * H * 4);
    float lumMax = 0.f;
    for (size_t i = 0; i < (size_t)W * H; i++) {
        const float r = d->rgb[3 * i], g = d->rgb[3 * i + 1], b = d->rgb[3 * i + 2];
        if (!(std::isfinite(r) && std::isfinite(g) && std::isfinite(b)) || r < 0.f || g < 0.f || b < 0.f)
            return fail(h, VRESTIR_EINVAL, "environment texels must be finite and non-negative");
        rgba[4 * i] = r; rgba[4 * i + 1] = g; rgba[4 * i + 2] = b; rgba[4 * i + 3] = 1.f;
        lumMax = std::max(lumMax, r * 0.2126f + g * 0.7152f + b * 0.0722f);
    }
    if (!(lumMax > 0.f)) return fail(h, VRESTIR_EINVAL, "environment map is black (bind a zero constant environment instead)");
    int mips = 0, off[kMaxImpMips];
    size_t total = 0;
    for (int dd = dim; dd >= 1; dd /= 2) { off[mips++] = (int)total; total += (size_t)dd * dd; }
    HIPCHK(h, hipSetDevice(h->device));
    HIPCHK(h, hipDeviceSynchronize());
    freeEnv(h);
    HIPCHK(h, hipMalloc((void**)&h->envTexDev, rgba.size() * sizeof(float)));
    HIPCHK(h, hipMalloc((void**)&h->envImpDev, total * sizeof(float)));
    HIPCHK(h, hipMemcpy(h->envTexDev, rgba.data(), rgba.size() * sizeof(float), hipMemcpyHostToDevice));
    HIPCHK(h, launchEnvImportance(nullptr, h->envTexDev, W, H, h->envImpDev, off, dim, spp));
    HIPCHK(h, hipDeviceSynchronize());
    float avg = 0.f;
    HIPCHK(h, hipMemcpy(&avg, h->envImpDev + off[mips - 1], sizeof(float), hipMemcpyDeviceToHost));
    if (!(avg > 0.f)) {
        freeEnv(h);
        S.envTextured = 0;
        (void)uploadScene(h);
        return fail(h, VRESTIR_EINVAL, "the importance map of the environment is zero everywhere");
    }
    S.envTex = h->envTexDev;
    S.envImp = h->envImpDev;
    S.envTextured = 1;
    S.envW = W; S.envH = H;
    S.envImpDim = dim;
    S.envImpBaseMip = mips - 1;
    S.envImpInvDim = 1.f / (float)dim;
    for (int k = 0; k < kMaxImpMips; k++) S.envImpOff[k] = k < mips ? off[k] : 0;
    for (int k = 0; k < 3; k++) S.envScale[k] = d->intensity * d->tint[k];
    envRotation(d->rotationDegrees, h->envCur);
    h->envLast = h->envCur;
    return uploadScene(h);
}

// EnvMap::setRotation: takes effect at the next frame; the frame after sees this
// orientation as the previous frame's (Scene::updateEnvMap, SC/Scene.cpp:761-775).
extern "C" int32_t vrestir_set_env_rotation(vrestir_handle* h, const float degreesXYZ[3]) {
    if (!h || !degreesXYZ) return VRESTIR_EINVAL;
    envRotation(degreesXYZ, h->envCur);
    return VRESTIR_OK;
}

// EnvMap::setIntensity / setTint.
extern "C" int32_t vrestir_set_env_intensity(vrestir_handle* h, float intensity, const float tint[3]) {
    if (!h || !tint) return VRESTIR_EINVAL;
    Scene& S = h->sceneHost;
    for (int k = 0; k < 3; k++) S.envScale[k] = intensity * tint[k];
    HIPCHK(h, hipSetDevice(h->device));
    HIPCHK(h, hipDeviceSynchronize());
    return uploadScene(h);
}

// addGVDBVolume's file set (SC/Scene.cpp:2794-2850): <prefix>_mip{k}.vbx and
// <prefix>_mip{k}c.vbx per mip, built straight from the files' bricks; levels
// whose files are missing are derived from the level above (vr_volume.cpp:
// buildVolumeFromFiles).
extern "C" int32_t vrestir_load_vbx(vrestir_handle* h, const char* pathPrefix, int32_t numMips, int32_t brickFormat,
                                    const float sigmaA[3], const float sigmaS[3], float g, float densityScale) {
    if (!h || !pathPrefix || !sigmaA || !sigmaS) return VRESTIR_EINVAL;
    if (brickFormat != VRESTIR_BRICK_FP32 && brickFormat != VRESTIR_BRICK_FP16)
        return fail(h, VRESTIR_EINVAL, "unknown brickFormat");
    if (std::fabs((sigmaS[0] + sigmaA[0]) - (sigmaS[1] + sigmaA[1])) > 1e-6f * (sigmaS[0] + sigmaA[0]) ||
        std::fabs((sigmaS[1] + sigmaA[1]) - (sigmaS[2] + sigmaA[2])) > 1e-6f * (sigmaS[0] + sigmaA[0]))
        return fail(h, VRESTIR_EINVAL, "sigma_a + sigma_s must be grey");
    HostVolume hv;
    vrestir_volume_desc d;
    std::memset(&d, 0, sizeof(d));
    int rc = buildVolumeFromFiles(pathPrefix, numMips, brickFormat == VRESTIR_BRICK_FP16, hv, d.indexToWorld, h->err);
    if (rc) return rc;
    d.nx = hv.slots[0].nx; d.ny = hv.slots[0].ny; d.nz = hv.slots[0].nz;
    d.numMips = hv.numMips;
    d.brickFormat = brickFormat;
    for (int i = 0; i < 3; i++) { d.sigmaA[i] = sigmaA[i]; d.sigmaS[i] = sigmaS[i]; }
    d.g = g;
    d.densityScale = densityScale;
    return installVolume(h, hv, &d);
}

extern "C" int32_t vrestir_set_lights(vrestir_handle* h, const vrestir_point_light* lights, int32_t count,
                                      const float env[3]) {
    if (!h) return VRESTIR_EINVAL;
    if (count < 0 || count > kMaxLights || (count > 0 && !lights))
        return fail(h, VRESTIR_EINVAL, "light count must be in [0, 256]");
    Scene& S = h->sceneHost;
    S.numLights = count;
    for (int i = 0; i < count; i++)
        for (int k = 0; k < 3; k++) {
            S.lights[6 * i + k] = lights[i].posW[k];
            S.lights[6 * i + 3 + k] = lights[i].intensity[k];
        }
    for (int k = 0; k < 3; k++) S.env[k] = env ? env[k] : 0.f;
    return uploadScene(h);
}

extern "C" int32_t vrestir_set_surfaces(vrestir_handle* h, const vrestir_surfaces_desc* d) {
    if (!h) return VRESTIR_EINVAL;
    HIPCHK(h, hipSetDevice(h->device));
    HIPCHK(h, hipDeviceSynchronize());
    if (d) {
        const std::string err = checkSurfacesDesc(*d);
        if (!err.empty()) return fail(h, VRESTIR_EINVAL, err);
    }
    freeSurfaces(h);
    if (!d) return uploadScene(h);
    LightCollectionHost lc;
    buildLightCollection(*d, lc);
    Scene& S = h->sceneHost;
    S.emMult = d->emissiveIntensityMultiplier;
    const int n = (int)lc.weight.size();
    if (n > 0) {
        auto up = [&](const void* src, size_t bytes, void** dst) -> hipError_t {
            hipError_t e = hipMalloc(dst, bytes);
            if (e != hipSuccess) return e;
            return hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice);
        };
        HIPCHK(h, up(lc.tri.data(), lc.tri.size() * sizeof(float), (void**)&h->emTriDev));
        HIPCHK(h, up(lc.alias.data(), lc.alias.size() * sizeof(uint32_t), (void**)&h->emAliasDev));
        HIPCHK(h, up(lc.weight.data(), lc.weight.size() * sizeof(float), (void**)&h->emWeightDev));
        S.emTri = h->emTriDev;
        S.emAlias = h->emAliasDev;
        S.emWeight = h->emWeightDev;
        S.numEmTris = n;
        S.numEmActive = lc.numActive;
        S.emWeightSum = (float)lc.weightSum;
    }
    SurfaceGeometryHost sg;
    buildSurfaceGeometry(*d, sg);
    if (d->numTriangles > 0) {
        auto up = [&](const void* src, size_t bytes, void** dst) -> hipError_t {
            hipError_t e = hipMalloc(dst, bytes);
            if (e != hipSuccess) return e;
            return hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice);
        };
        HIPCHK(h, up(sg.nodes.data(), sg.nodes.size() * sizeof(float), &h->bvhDev));
        HIPCHK(h, up(sg.tris.data(), sg.tris.size() * sizeof(float), (void**)&h->bvhTriDev));
        HIPCHK(h, up(sg.triData.data(), sg.triData.size() * sizeof(float), (void**)&h->triDataDev));
        HIPCHK(h, up(sg.materials.data(), sg.materials.size() * sizeof(float), (void**)&h->materialsDev));
        S.bvh = (const BvhNode*)h->bvhDev;
        S.bvhTri = h->bvhTriDev;
        S.triData = h->triDataDev;
        S.materials = h->materialsDev;
        S.numTris = d->numTriangles;
        S.numMaterials = d->numMaterials;
    }
    h->optionsChanged = true;
    return uploadScene(h);
}

extern "C" int32_t vrestir_set_camera(vrestir_handle* h, const vrestir_camera* c) {
    if (!h || !c) return VRESTIR_EINVAL;
    if (!(c->fovY > 0.f && c->fovY < 3.14159f)) return fail(h, VRESTIR_EINVAL, "fovY must be in (0, pi)");
    h->camDesc = *c;
    h->camSet = true;
    return VRESTIR_OK;
}

// UVW frame (SC/Camera/Camera.cpp:183-189) and lookAt / perspective matrices
// in double precision, stored as float (row-vector convention).
static void buildCamera(const vrestir_camera& c, float aspect, CameraFrame& cf, float* view, float* proj) {
    double P[3] = {c.posW[0], c.posW[1], c.posW[2]};
    double w[3] = {(double)c.target[0] - P[0], (double)c.target[1] - P[1], (double)c.target[2] - P[2]};
    double up[3] = {c.up[0], c.up[1], c.up[2]};
    auto nrm = [](double* v) { double l = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); v[0] /= l; v[1] /= l; v[2] /= l; };
    auto crs = [](const double* a, const double* b, double* o) {
        o[0] = a[1] * b[2] - a[2] * b[1]; o[1] = a[2] * b[0] - a[0] * b[2]; o[2] = a[0] * b[1] - a[1] * b[0];
    };
    nrm(w);
    double u[3], v[3];
    crs(w, up, u); nrm(u);
    crs(u, w, v); nrm(v);
    double th = std::tan((double)c.fovY * 0.5);
    double ul = th * aspect, vl = th;
    for (int i = 0; i < 3; i++) {
        cf.pos[i] = c.posW[i];
        cf.W[i] = (float)w[i];
        cf.U[i] = (float)(u[i] * ul);
        cf.V[i] = (float)(v[i] * vl);
    }
    double vw[16] = {u[0], v[0], -w[0], 0, u[1], v[1], -w[1], 0, u[2], v[2], -w[2], 0,
                     -(u[0] * P[0] + u[1] * P[1] + u[2] * P[2]), -(v[0] * P[0] + v[1] * P[1] + v[2] * P[2]),
                     (w[0] * P[0] + w[1] * P[1] + w[2] * P[2]), 1};
    double pr[16] = {0};
    pr[0] = 1.0 / (aspect * th);
    pr[5] = 1.0 / th;
    pr[10] = c.farZ / (c.nearZ - c.farZ);
    pr[11] = -1.0;
    pr[14] = -((double)c.farZ * c.nearZ) / ((double)c.farZ - c.nearZ);
    for (int i = 0; i < 16; i++) { view[i] = (float)vw[i]; proj[i] = (float)pr[i]; }
}

// Neighbour offsets (VR/SpatialReuse.cs.slang:64-81) for the 16 R2 frame seeds:
// u from the R2 sequence (single-precision constants, double arithmetic) or
// Hammersley, then int(radius * sample_disk(u)).
static void neighbourOffsets(const vrestir_params& p, std::vector<int>& out) {
    out.assign(16 * kMaxSpatialSamples * 2, 0);
    int K = p.spatialSampleCount;
    for (int seed = 0; seed < 16; seed++)
        for (int s = 0; s < K; s++) {
            float ux, uy;
            if (p.randomSamplerType == VRESTIR_SAMPLER_HAMMERSLEY) {
                uint32_t i = (uint32_t)s;
                i = (i & 0x55555555u) << 1 | (i & 0xAAAAAAAAu) >> 1;
                i = (i & 0x33333333u) << 2 | (i & 0xCCCCCCCCu) >> 2;
                i = (i & 0x0F0F0F0Fu) << 4 | (i & 0xF0F0F0F0u) >> 4;
                i = (i & 0x00FF00FFu) << 8 | (i & 0xFF00FF00u) >> 8;
                i = (i << 16) | (i >> 16);
                ux = (float)s / (float)K;
                uy = (float)i * 2.3283064365386963e-10f;
            } else if (s == 0) {
                ux = 0.f; uy = 0.f;
            } else {
                double m = (double)seed * p.spatialSampleCount + s;
                double a = (double)0.754877669f * m, b = (double)0.569840296f * m;
                ux = (float)(a - std::floor(a));
                uy = (float)(b - std::floor(b));
            }
            float r = std::sqrt(ux);
            float phi = 2.f * 3.14159265358979323846f * uy;
            out[(seed * kMaxSpatialSamples + s) * 2 + 0] = (int)(p.sampleRadius * (r * std::cos(phi)));
            out[(seed * kMaxSpatialSamples + s) * 2 + 1] = (int)(p.sampleRadius * (r * std::sin(phi)));
        }
}

static void fillOptions(Options& o, uint32_t vm, uint32_t lm, int ls, int lmip, int vs, int vmip, bool vlin,
                        bool llin, float vts, float lts, const vrestir_params& p) {
    o.visibilityTrackingMethod = vm; o.lightingTrackingMethod = lm;
    o.lightSamples = ls; o.lightingMipLevel = lmip; o.visibilitySamples = vs; o.visibilityMipLevel = vmip;
    o.visibilityUseLinearSampler = vlin; o.lightingUseLinearSampler = llin;
    o.visibilityTStepScale = vts; o.lightingTStepScale = lts;
    o.useEnv = p.useEnvironmentLights; o.useAnalytic = p.useAnalyticLights; o.useEmissive = p.useEmissiveLights;
    o.vertexReuseStartBounce = p.vertexReuseStartBounce;
}

static int allocFrameBuffers(vrestir_handle* h, int W, int H, int B) {
    freeFrameBuffers(h);
    size_t n = (size_t)W * H, ne = n * (size_t)std::max(B - 1, 1);
    auto al = [&](void** p, size_t bytes) -> int {
        HIPCHK(h, hipMalloc(p, bytes));
        HIPCHK(h, hipMemset(*p, 0, bytes));
        return VRESTIR_OK;
    };
    int rc;
    for (int i = 0; i < 2; i++) {
        if ((rc = al((void**)&h->res[i], n * sizeof(ResStore)))) return rc;
        if ((rc = al((void**)&h->extra[i], ne * 12))) return rc;
        if ((rc = al((void**)&h->color[i], n * 16))) return rc;
    }
    if ((rc = al((void**)&h->resT, n * sizeof(ResStore)))) return rc;
    if ((rc = al((void**)&h->extraT, ne * 12))) return rc;
    if ((rc = al((void**)&h->features, n * 8))) return rc;
    if ((rc = al((void**)&h->featuresT, n * 8))) return rc;
    if ((rc = al((void**)&h->vbuf, n * sizeof(VItem)))) return rc;
    if ((rc = al((void**)&h->outInternal, n * 16))) return rc;
    h->W = W; h->H = H; h->allocB = B;
    return VRESTIR_OK;
}

// ------------------------------------------------------------------ learned free-flight proposal
// (an addition of this project: no counterpart in the reference)
static bool learnedActive(const vrestir_handle* h) {
    return h->L.enable && h->lPrm && !h->P.useReference && (h->P.enableSpatialReuse || h->P.enableTemporalReuse);
}
static int learnedSlots(int W, int H, const vrestir_learned_desc& L) {
    return ((W + L.trainBlockX - 1) / L.trainBlockX) * ((H + L.trainBlockY - 1) / L.trainBlockY);
}
static int learnedFrameBuffers(vrestir_handle* h) {
    const size_t n = (size_t)h->W * h->H;
    if (h->lCdfPixels != n) {
        dfree(h->lCdf);
        HIPCHK(h, hipMalloc((void**)&h->lCdf, n * 64));
        HIPCHK(h, hipMemset(h->lCdf, 0, n * 64));
        h->lCdfPixels = n;
    }
    const int slots = learnedSlots(h->W, h->H, h->L);
    if (h->lRecSlots != slots) {
        dfree(h->lRecs);
        HIPCHK(h, hipMalloc((void**)&h->lRecs, (size_t)slots * kLearnedRecFloats * 4));
        HIPCHK(h, hipMemset(h->lRecs, 0, (size_t)slots * kLearnedRecFloats * 4));
        h->lRecSlots = slots;
    }
    return VRESTIR_OK;
}
static int learnedPartials(vrestir_handle* h, int records) {
    const int waves = (records + 31) / 32;
    if (waves > h->lPartialWaves) {
        dfree(h->lPartial);
        HIPCHK(h, hipMalloc((void**)&h->lPartial, (size_t)waves * kLearnedPartialStride * 4));
        h->lPartialWaves = waves;
    }
    return VRESTIR_OK;
}
// One Adam step on the mean cross-entropy of n records (device memory).
static int learnedStep(vrestir_handle* h, hipStream_t st, const float* recs, int n) {
    if (n <= 0) return VRESTIR_OK;
    int rc = learnedPartials(h, n);
    if (rc) return rc;
    const vrestir_learned_desc& L = h->L;
    HIPCHK(h, launchLearnedGrad(st, recs, n, L.defensiveFraction, 1.f / (float)n, h->lFrags, h->lPrm, h->lPartial));
    h->lStep++;
    const float c1 = (float)(1.0 - std::pow(0.9, (double)h->lStep));
    const float c2 = (float)(1.0 - std::pow(0.999, (double)h->lStep));
    HIPCHK(h, launchLearnedReduce(st, h->lPartial, (n + 31) / 32, nullptr, true, h->lPrm, h->lM, h->lV, L.learningRate,
                                  c1, c2, h->lStats, h->lInUse, L.judge != 0, h->lStep));
    HIPCHK(h, launchLearnedPack(st, h->lPrm, h->lFrags));
    return VRESTIR_OK;
}
// Stage: the network's CDF for every camera ray of the frame.
static int learnedProposal(vrestir_handle* h, hipStream_t st) {
    if (!learnedActive(h)) return fail(h, VRESTIR_ESTATE, "the learned proposal is not enabled for these options");
    int rc = learnedFrameBuffers(h);
    if (rc) return rc;
    const int n = h->W * h->H;
    HIPCHK(h, launchLearnedInfer(st, h->sceneDev, h->frame, nullptr, n, h->lFrags, h->lPrm, h->lInUse,
                                 h->L.defensiveFraction, h->lCdf, nullptr));
    return VRESTIR_OK;
}
// Stage: Adam steps on this frame's training rays, split into trainStepsPerFrame minibatches.
static int learnedTrainFrame(vrestir_handle* h, hipStream_t st) {
    if (!learnedActive(h)) return fail(h, VRESTIR_ESTATE, "the learned proposal is not enabled for these options");
    if (!h->L.train) return VRESTIR_OK;
    int rc0 = learnedFrameBuffers(h);
    if (rc0) return rc0;
    const int n = h->lRecSlots, K = std::max(1, h->L.trainStepsPerFrame);
    for (int k = 0; k < K; k++) {
        const int b0 = (int)((int64_t)n * k / K), b1 = (int)((int64_t)n * (k + 1) / K);
        int rc = learnedStep(h, st, h->lRecs + (size_t)b0 * kLearnedRecFloats, b1 - b0);
        if (rc) return rc;
    }
    return VRESTIR_OK;
}
// He-uniform hidden layers from SplitMix64(seed), zero output layer and biases:
// the untrained network is the uniform pdf, i.e. the analytic proposal.
static void learnedInitWeights(uint64_t seed, std::vector<float>& w) {
    w.assign(kLearnedParams, 0.f);
    uint64_t st = seed;
    auto next = [&]() {
        uint64_t z = (st += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return (double)(z >> 11) * (1.0 / 9007199254740992.0);
    };
    const int sizes[3][3] = {{0, 36, 64}, {2368, 64, 64}, {6528, 64, 64}};
    for (const auto& L : sizes) {
        const double lim = std::sqrt(6.0 / (double)L[1]);
        for (int i = 0; i < L[1] * L[2]; i++) w[L[0] + i] = (float)((2.0 * next() - 1.0) * lim);
    }
}
static int learnedUpload(vrestir_handle* h, const float* w) {
    HIPCHK(h, hipMemcpy(h->lPrm, w, kLearnedParams * 4, hipMemcpyHostToDevice));
    HIPCHK(h, hipMemset(h->lM, 0, kLearnedParams * 4));
    HIPCHK(h, hipMemset(h->lV, 0, kLearnedParams * 4));
    const float st[5] = {0.f, 0.f, 0.f, 1.f, 0.f};
    const int one = 1;
    HIPCHK(h, hipMemcpy(h->lStats, st, sizeof(st), hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(h->lInUse, &one, sizeof(int), hipMemcpyHostToDevice));
    h->lStep = 0;
    HIPCHK(h, launchLearnedPack(nullptr, h->lPrm, h->lFrags));
    HIPCHK(h, hipDeviceSynchronize());
    return VRESTIR_OK;
}

extern "C" void vrestir_default_learned(vrestir_learned_desc* d) {
    if (!d) return;
    std::memset(d, 0, sizeof(*d));
    d->enable = 0;
    d->train = 1;
    d->trainBlockX = 16;
    d->trainBlockY = 8;
    d->trainStepsPerFrame = 1;
    d->learningRate = 2e-3f;
    d->defensiveFraction = 0.25f;
    d->judge = 1;
    d->seed = 1;
}

extern "C" int32_t vrestir_set_learned(vrestir_handle* h, const vrestir_learned_desc* d) {
    if (!h || !d) return VRESTIR_EINVAL;
    if (d->trainBlockX < 1 || d->trainBlockY < 1 || d->trainBlockX > 4096 || d->trainBlockY > 4096)
        return fail(h, VRESTIR_EINVAL, "training block sizes must be in [1, 4096]");
    if (d->trainStepsPerFrame < 1 || d->trainStepsPerFrame > 64)
        return fail(h, VRESTIR_EINVAL, "trainStepsPerFrame must be in [1, 64]");
    if (!(d->learningRate > 0.f && d->learningRate < 1.f)) return fail(h, VRESTIR_EINVAL, "learningRate out of range");
    if (!(d->defensiveFraction >= 1.f / 64.f && d->defensiveFraction <= 1.f))
        return fail(h, VRESTIR_EINVAL, "defensiveFraction must be in [1/64, 1]");
    HIPCHK(h, hipSetDevice(h->device));
    if (!h->lPrm) {
        HIPCHK(h, hipMalloc((void**)&h->lPrm, kLearnedParams * 4));
        HIPCHK(h, hipMalloc((void**)&h->lM, kLearnedParams * 4));
        HIPCHK(h, hipMalloc((void**)&h->lV, kLearnedParams * 4));
        HIPCHK(h, hipMalloc((void**)&h->lFrags, (size_t)kLearnedFrags * 64 * sizeof(uint4)));
        HIPCHK(h, hipMalloc((void**)&h->lStats, 8 * sizeof(float)));
        HIPCHK(h, hipMalloc((void**)&h->lInUse, sizeof(int)));
    }
    const bool sizeChanged = d->trainBlockX != h->L.trainBlockX || d->trainBlockY != h->L.trainBlockY;
    h->L = *d;
    if (sizeChanged) { dfree(h->lRecs); h->lRecSlots = 0; }
    std::vector<float> w;
    learnedInitWeights(d->seed, w);
    return learnedUpload(h, w.data());
}

extern "C" int32_t vrestir_get_learned(const vrestir_handle* h, vrestir_learned_desc* out) {
    if (!h || !out) return VRESTIR_EINVAL;
    *out = h->L;
    return VRESTIR_OK;
}

extern "C" int32_t vrestir_learned_stats(vrestir_handle* h, float stats[5]) {
    if (!h || !stats) return VRESTIR_EINVAL;
    if (!h->lPrm) return fail(h, VRESTIR_ESTATE, "the learned proposal is not set up (vrestir_set_learned)");
    HIPCHK(h, hipSetDevice(h->device));
    HIPCHK(h, hipDeviceSynchronize());
    HIPCHK(h, hipMemcpy(stats, h->lStats, 5 * sizeof(float), hipMemcpyDeviceToHost));
    return VRESTIR_OK;
}

extern "C" int32_t vrestir_learned_get_weights(vrestir_handle* h, float* out, int32_t n) {
    if (!h || !out) return VRESTIR_EINVAL;
    if (n != kLearnedParams) return fail(h, VRESTIR_EINVAL, "the network has VRESTIR_LEARNED_PARAMS parameters");
    if (!h->lPrm) return fail(h, VRESTIR_ESTATE, "the learned proposal is not set up (vrestir_set_learned)");
    HIPCHK(h, hipSetDevice(h->device));
    HIPCHK(h, hipDeviceSynchronize());
    HIPCHK(h, hipMemcpy(out, h->lPrm, kLearnedParams * 4, hipMemcpyDeviceToHost));
    return VRESTIR_OK;
}

extern "C" int32_t vrestir_learned_set_weights(vrestir_handle* h, const float* in, int32_t n) {
    if (!h || !in) return VRESTIR_EINVAL;
    if (n != kLearnedParams) return fail(h, VRESTIR_EINVAL, "the network has VRESTIR_LEARNED_PARAMS parameters");
    if (!h->lPrm) return fail(h, VRESTIR_ESTATE, "the learned proposal is not set up (vrestir_set_learned)");
    for (int i = 0; i < n; i++)
        if (!std::isfinite(in[i])) return fail(h, VRESTIR_EINVAL, "non-finite weight");
    HIPCHK(h, hipSetDevice(h->device));
    HIPCHK(h, hipDeviceSynchronize());
    return learnedUpload(h, in);
}

extern "C" int32_t vrestir_learned_forward(vrestir_handle* h, void* stream, const float* rawDev, int32_t n,
                                          uint16_t* cdfDev, float* logitsDev) {
    if (!h || n < 0 || (n > 0 && (!rawDev || !cdfDev))) return VRESTIR_EINVAL;
    if (!h->lPrm) return fail(h, VRESTIR_ESTATE, "the learned proposal is not set up (vrestir_set_learned)");
    HIPCHK(h, hipSetDevice(h->device));
    HIPCHK(h, launchLearnedInfer((hipStream_t)stream, h->sceneDev, h->frame, rawDev, n, h->lFrags, h->lPrm, nullptr,
                                 h->L.defensiveFraction, cdfDev, logitsDev));
    return VRESTIR_OK;
}

extern "C" int32_t vrestir_learned_gradient(vrestir_handle* h, void* stream, const float* recsDev, int32_t n,
                                           float scale, float* outDev) {
    if (!h || n <= 0 || !recsDev || !outDev) return VRESTIR_EINVAL;
    if (!h->lPrm) return fail(h, VRESTIR_ESTATE, "the learned proposal is not set up (vrestir_set_learned)");
    HIPCHK(h, hipSetDevice(h->device));
    int rc = learnedPartials(h, n);
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;
    HIPCHK(h, launchLearnedGrad(st, recsDev, n, h->L.defensiveFraction, scale, h->lFrags, h->lPrm, h->lPartial));
    HIPCHK(h, launchLearnedReduce(st, h->lPartial, (n + 31) / 32, outDev, false, nullptr, nullptr, nullptr, 0.f, 1.f,
                                  1.f, nullptr, nullptr, false, 0));
    return VRESTIR_OK;
}

extern "C" int32_t vrestir_learned_train(vrestir_handle* h, void* stream, const float* recsDev, int32_t n) {
    if (!h || n <= 0 || !recsDev) return VRESTIR_EINVAL;
    if (!h->lPrm) return fail(h, VRESTIR_ESTATE, "the learned proposal is not set up (vrestir_set_learned)");
    HIPCHK(h, hipSetDevice(h->device));
    return learnedStep(h, (hipStream_t)stream, recsDev, n);
}

// Frame set-up: reallocation, history reset on option change, round
// bookkeeping and the three SamplingOptions (VR/VolumetricReSTIR.cpp:339-496).
extern "C" int32_t vrestir_prepare_frame(vrestir_handle* h, int32_t W, int32_t H) {
    if (!h) return VRESTIR_EINVAL;
    if (!h->volumeSet) return fail(h, VRESTIR_ESTATE, "no volume bound (vrestir_set_volume)");
    if (!h->camSet) return fail(h, VRESTIR_ESTATE, "no camera bound (vrestir_set_camera)");
    if (W <= 0 || H <= 0 || W > 16384 || H > 16384) return fail(h, VRESTIR_EINVAL, "resolution out of range");
    const vrestir_params& p = h->P;
    HIPCHK(h, hipSetDevice(h->device));
    if (h->optionsChanged) {
        h->frameCount = 0;
        h->temporalAccumulated = false;
        h->optionsChanged = false;
    }
    if (W != h->W || H != h->H || p.maxBounces != h->allocB) {
        int rc = allocFrameBuffers(h, W, H, p.maxBounces);
        if (rc) return rc;
        h->temporalAccumulated = false;
    }
    if (h->offsetsDirty) {
        std::vector<int> offs;
        neighbourOffsets(p, offs);
        HIPCHK(h, hipMemcpy(h->offsetsDev, offs.data(), offs.size() * sizeof(int), hipMemcpyHostToDevice));
        h->offsetsDirty = false;
    }
    float aspect = (float)W / (float)H;
    buildCamera(h->camDesc, aspect, h->cam, h->view, h->proj);
    if (!h->havePrev) {
        h->prevCam = h->cam;
        std::memcpy(h->prevView, h->view, sizeof(h->view));
        std::memcpy(h->prevProj, h->proj, sizeof(h->proj));
        h->havePrev = true;
    }
    Frame& F = h->frame;
    std::memset(&F, 0, sizeof(F));
    F.width = W; F.height = H;
    F.frameCount = h->frameCount;
    F.numTotalRounds = (p.enableSpatialReuse ? p.spatialReuseRounds : 0) + (p.enableTemporalReuse ? 1 : 0) + 2;
    F.maxBounces = p.maxBounces;
    F.initialM = p.initialM;
    F.useRussianRoulette = p.initialUseRussianRoulette;
    F.noReuse = !p.enableSpatialReuse && !p.enableTemporalReuse;
    F.coarserIndirect = p.initialUseCoarserGridForIndirectBounce;
    F.useReference = p.useReference;
    F.useSurfaceScene = p.useSurfaceScene;
    F.baselineSpp = p.baselineSamplePerPixel;
    F.visualizeTransmittance = p.visualizeTotalTransmittance;
    F.isFirstFrame = h->temporalAccumulated ? 0 : 1;
    F.temporalMIS = (int)p.temporalMISMethod;
    F.reprojectionMode = (int)p.temporalReprojectionMode;
    F.reprojectionMip = p.temporalReprojectionMipLevel;
    F.temporalThreshold = p.temporalReuseMThreshold;
    F.spatialMIS = (int)p.spatialMISMethod;
    F.spatialSampleCount = p.spatialSampleCount;
    F.enableTemporal = p.enableTemporalReuse;
    F.enableSpatial = p.enableSpatialReuse;
    F.spatialRounds = p.spatialReuseRounds;
    F.cam = h->cam;
    F.prevCam = h->prevCam;
    F.env[0] = h->envCur;
    F.env[1] = h->haveEnvLast ? h->envLast : h->envCur;
    F.hasAnimation = h->hasAnimation ? 1 : 0;
    F.prevGridOffset = (p.usePrevVolumeForReproj && h->hasAnimation) ? kPrevDensityOffset : 0;
    std::memcpy(F.prevView, h->prevView, sizeof(F.prevView));
    std::memcpy(F.prevProj, h->prevProj, sizeof(F.prevProj));
    fillOptions(F.initialOpt, VRESTIR_ANALYTIC_TRACKING, p.initialLightingTrackingMethod, p.initialLightSamples,
                p.initialLightingMipLevel, 1,
                p.initialVisibilityUseLinearSampler ? p.initialBaseMipLevel : p.initialBaseMipLevel + 8,
                p.initialVisibilityUseLinearSampler, p.initialLightingUseLinearSampler, p.initialVisibilityTStepScale,
                p.initialLightingTStepScale, p);
    fillOptions(F.spatialOpt, p.spatialVisibilityTrackingMethod, p.spatialLightingTrackingMethod, 1,
                p.spatialLightingMipLevel, 1, p.spatialVisibilityMipLevel, p.spatialVisibilityUseLinearSampler,
                p.spatialLightingUseLinearSampler, p.spatialVisibilityTStepScale, p.spatialLightingTStepScale, p);
    bool fa = p.finalLightTrackingMethod == 1 || p.finalLightTrackingMethod == 2;
    bool fv = p.finalVisibilityTrackingMethod == 1 || p.finalVisibilityTrackingMethod == 2;
    fillOptions(F.finalOpt, p.finalVisibilityTrackingMethod, p.finalLightTrackingMethod, fa ? 1 : p.finalLightSamples,
                0, fv ? 1 : p.finalVisibilitySamples, 0, true, true, p.finalTStepScale, p.finalTStepScale, p);
    // every mip level the options reference must exist
    int maxMip = std::max({p.initialBaseMipLevel, p.initialLightingMipLevel, p.spatialVisibilityMipLevel,
                           p.spatialLightingMipLevel, p.temporalReprojectionMipLevel});
    if (maxMip >= h->meta.numMips)
        return fail(h, VRESTIR_EINVAL, "an option references a mip level the bound volume does not have");
    if (p.useAnalyticLights && h->sceneHost.numLights == 0)
        return fail(h, VRESTIR_EINVAL, "useAnalyticLights is set but no point light is bound");
    if (h->sceneHost.hasEmission && h->sceneHost.lutRows == 0)
        return fail(h, VRESTIR_ESTATE, "an emission grid is bound but no blackbody table (vrestir_set_blackbody_lut)");
    if (learnedActive(h)) {
        int rc = learnedFrameBuffers(h);
        if (rc) return rc;
    }
    h->totalRoundId = 0;
    h->frameOpen = true;
    return VRESTIR_OK;
}

// Grid slots whose occupancy tables a launch keeps in LDS (see occCacheLoad):
// at most kOccLdsBytes per workgroup, so four 256-lane workgroups still fit a CU.
static const int kOccLdsBytes = 24 * 1024;
static int occPlan(const vrestir_handle* h, int a, int b, int slots[2]) {
    auto words = [&](int s) {
        return (s >= 0 && s < kMaxSlots && h->sceneHost.grid[s].valid) ? h->sceneHost.grid[s].occWords : -1;
    };
    slots[0] = slots[1] = -1;
    int bytes = 0;
    const int wa = words(a), wb = b != a ? words(b) : -1;
    if (wa >= 0 && wa * 8 <= kOccLdsBytes) { slots[0] = a; bytes += wa * 8; }
    if (wb >= 0 && bytes + wb * 8 <= kOccLdsBytes) { slots[1] = b; bytes += wb * 8; }
    return bytes;
}

// The spatial stage runs as record / trace / replay (balanced tracing of its
// transmittance rays) when both of its ray kinds use ray marching, temporal reuse
// likewise (segment form only, and not on the previous frame's grids), and final
// shading when both of its ray kinds use analytic tracking (segment form only,
// when grid 0's blocks fit the trace's LDS bit mask), and the ray records of a
// frame fit kSplitBudget; otherwise each runs in one kernel. The record buffers
// hold the larger per-pixel capacity; each stage strides them by its own.
static const size_t kSplitBudget = (size_t)48 << 30;
static int splitEnsure(vrestir_handle* h, int cap, size_t n) {
    if (h->split.cap >= cap && h->splitPixels == n) return VRESTIR_OK;
    freeSplit(h);
    HIPCHK(h, hipMalloc((void**)&h->split.recs, n * (size_t)cap * sizeof(TraceRec)));
    HIPCHK(h, hipMalloc((void**)&h->split.trs, n * (size_t)cap * sizeof(float)));
    HIPCHK(h, hipMalloc((void**)&h->split.dens, n * (size_t)cap * sizeof(float)));
    HIPCHK(h, hipMalloc((void**)&h->split.counts, n * sizeof(int32_t)));
    HIPCHK(h, hipMalloc((void**)&h->split.plan, n * sizeof(uint32_t)));
    HIPCHK(h, hipMalloc((void**)&h->split.fix, sizeof(int32_t)));
    h->split.cap = cap;
    h->splitPixels = n;
    return VRESTIR_OK;
}
static int splitPlan(vrestir_handle* h, LaunchArgs& a, int stage) {
    a.splitTK = 0;
    a.splitMode = 0;
    if (!h->splitReuse) return VRESTIR_OK;
    const vrestir_params& p = h->P;
    const size_t n = (size_t)h->W * h->H;
    int cap = 0, tk = 0;
    if (stage == VRESTIR_STAGE_TRACE_RAYS) {
        // the shadow rays of the candidates (ray marching only; not with the learned
        // proposal, the reference path tracer, or without reuse)
        const Frame& F = h->frame;
        if (h->splitReuse != 2 || !h->splitTrace || h->inBatch || learnedActive(h) || p.useReference || F.noReuse ||
            F.initialOpt.lightingTrackingMethod != 2u)
            return VRESTIR_OK;
        const int M = std::max(1, p.initialM), B = std::max(1, p.maxBounces);
        cap = std::max(M * B, candStashSlots(M, B));  // one shadow ray per vertex; the free-flight samples
        tk = 2;
    } else if (stage == VRESTIR_STAGE_SPATIAL_REUSE) {
        const Options& so = h->frame.spatialOpt;
        if (so.visibilityTrackingMethod != 2u || so.lightingTrackingMethod != 2u) return VRESTIR_OK;
        const int K = p.spatialSampleCount;
        const int evals = p.spatialMISMethod == VRESTIR_MIS_TALBOT ? K * (K - 1) : K - 1;
        cap = std::max(1, evals * (p.maxBounces + 1));
        tk = 2;
    } else if (stage == VRESTIR_STAGE_TEMPORAL_REUSE) {
        // the segment trace only; rays on the previous frame's grids (animated volumes) run in one kernel
        const Options& so = h->frame.spatialOpt;
        if (h->splitReuse != 2 || so.visibilityTrackingMethod != 2u || so.lightingTrackingMethod != 2u ||
            h->frame.prevGridOffset != 0)
            return VRESTIR_OK;
        cap = 2 * (p.maxBounces + 1);  // two target evaluations of b + 1 rays each
        tk = 2;
    } else if (stage == VRESTIR_STAGE_FINAL_SHADING) {
        const Options& fo = h->frame.finalOpt;
        if (fo.visibilityTrackingMethod != 1u || fo.lightingTrackingMethod != 1u || p.useReference ||
            h->frame.visualizeTransmittance)
            return VRESTIR_OK;
        // the analytic segment trace only (mode 2): one grid whose blocks fit its LDS bit mask, one filter
        const Grid& g0 = h->sceneHost.grid[fo.visibilityMipLevel];
        const bool uniformFilter = fo.visibilityUseLinearSampler == fo.lightingUseLinearSampler;
        if (h->splitReuse != 2 || fo.visibilityMipLevel != 0 || !g0.valid ||
            (size_t)g0.nnx * g0.nny * g0.nnz > 2048 || !uniformFilter)
            return VRESTIR_OK;
        cap = p.maxBounces + 1;  // the primary ray, one per extra bounce, the light ray
        tk = 1;
        a.traceGridBlocks = (int64_t)g0.nnx * g0.nny * g0.nnz;
    } else {
        return VRESTIR_OK;
    }
    // the record buffers must fit the budget and the device's free memory (keeping
    // a quarter of it); a failed allocation falls back to the single kernel
    const size_t need = n * (size_t)cap * (sizeof(TraceRec) + 2 * sizeof(float)) + n * 8;
    const size_t held = h->splitPixels * (size_t)h->split.cap * (sizeof(TraceRec) + 2 * sizeof(float)) +
                        h->splitPixels * 8;
    size_t freeB = 0, totalB = 0;
    if (need > kSplitBudget) return VRESTIR_OK;
    if (need > held && hipMemGetInfo(&freeB, &totalB) == hipSuccess && need > (freeB + held) / 4 * 3)
        return VRESTIR_OK;
    if (splitEnsure(h, cap, n) != VRESTIR_OK) {
        freeSplit(h);
        (void)hipGetLastError();
        h->err.clear();
        return VRESTIR_OK;
    }
    a.split = h->split;
    a.split.cap = cap;
    a.splitTK = tk;
    a.splitMode = h->splitReuse;
    return VRESTIR_OK;
}

// Implementation choices with no counterpart in the reference.
extern "C" int32_t vrestir_set_tuning(vrestir_handle* h, int32_t key, int32_t value) {
    if (!h) return VRESTIR_EINVAL;
    switch (key) {
        case VRESTIR_TUNE_SPLIT_REUSE:
            if (value < 0 || value > 2) return fail(h, VRESTIR_EINVAL, "split reuse mode must be 0, 1 or 2");
            h->splitReuse = value;
            if (value == 0) {
                HIPCHK(h, hipSetDevice(h->device));
                HIPCHK(h, hipDeviceSynchronize());
                freeSplit(h);
            }
            return VRESTIR_OK;
        case VRESTIR_TUNE_SPLIT_TRACE:
            if (value < 0 || value > 1) return fail(h, VRESTIR_EINVAL, "split TraceRays must be 0 or 1");
            h->splitTrace = value;
            return VRESTIR_OK;
        default: return fail(h, VRESTIR_EINVAL, "unknown tuning key");
    }
}

// The volume path tracer (useReference's kernel) on a uniform pixel lattice of
// the width x height frame, for converged references on a pixel subsample. The
// camera is the bound one; the other frame state is the last prepared frame's.
extern "C" int32_t vrestir_reference_lattice(vrestir_handle* h, void* stream, int32_t width, int32_t height,
                                             int32_t stride, int32_t spp, uint32_t sampleNumber, int32_t chunks,
                                             float* outRGBA) {
    if (!h) return VRESTIR_EINVAL;
    if (!h->volumeSet || !h->camSet) return fail(h, VRESTIR_ESTATE, "no volume or camera bound");
    if (width != h->W || height != h->H)
        return fail(h, VRESTIR_ESTATE, "prepare or execute a frame of this size first");
    if (stride < 1 || stride > width || stride > height || spp < 1 || spp > 65536 || chunks < 1 ||
        chunks > 65535 || !outRGBA)
        return fail(h, VRESTIR_EINVAL, "bad lattice arguments");
    if (h->frameOpen) return fail(h, VRESTIR_ESTATE, "a frame is open (vrestir_end_frame first)");
    HIPCHK(h, hipSetDevice(h->device));
    LaunchArgs a;
    std::memset(&a, 0, sizeof(a));
    a.stream = (hipStream_t)stream;
    a.maxBounces = h->P.maxBounces;
    a.width = width;
    a.height = height;
    a.scene = h->sceneDev;
    a.frame = h->frame;
    float view[16], proj[16];
    buildCamera(h->camDesc, (float)width / (float)height, a.frame.cam, view, proj);
    a.occBytes = occPlan(h, 0, -1, a.frame.occSlot);
    HIPCHK(h, launchReferenceLattice(a, stride, spp, sampleNumber, chunks, (float4*)outRGBA));
    return VRESTIR_OK;
}

// The launch arguments of one stage of the open frame: its buffers, the
// occupancy tables it keeps in LDS and its split plan. Launches nothing.
static int stageArgs(vrestir_handle* h, hipStream_t st, int stage, int round, float* outRGBA, LaunchArgs& a) {
    const vrestir_params& p = h->P;
    std::memset(&a, 0, sizeof(a));
    a.stream = st;
    a.stage = stage;
    a.round = round;
    a.maxBounces = p.maxBounces;
    a.width = h->W;
    a.height = h->H;
    a.scene = h->sceneDev;
    a.frame = h->frame;
    a.offsets = h->offsetsDev;
    a.features = h->features;
    a.featuresT = h->featuresT;
    a.vbuf = h->vbuf;
    a.resT = h->resT;
    a.extraT = h->extraT;
    const int t = h->totalRoundId;
    switch (stage) {
        case VRESTIR_STAGE_GENERATE_FEATURES: break;
        case VRESTIR_STAGE_TRACE_RAYS:
            a.res0 = h->res[0]; a.extra0 = h->extra[0]; a.color0 = h->color[0];
            break;
        case VRESTIR_STAGE_TEMPORAL_REUSE:
            a.res0 = h->res[t % 2]; a.extra0 = h->extra[t % 2];
            a.mvec = (float2*)outRGBA;  // motion vectors ride in the output slot for this stage
            a.frame.outputMotionVec = outRGBA != nullptr;
            break;
        case VRESTIR_STAGE_SPATIAL_REUSE:
            a.res0 = h->res[t % 2]; a.res1 = h->res[(t + 1) % 2];
            a.extra0 = h->extra[t % 2]; a.extra1 = h->extra[(t + 1) % 2];
            break;
        case VRESTIR_STAGE_COPY_RESERVOIRS:
            a.res0 = h->res[t % 2]; a.extra0 = h->extra[t % 2];
            break;
        case VRESTIR_STAGE_FINAL_SHADING:
            a.res0 = h->res[t % 2]; a.extra0 = h->extra[t % 2]; a.color0 = h->color[t % 2];
            a.out = outRGBA ? (float4*)outRGBA : h->outInternal;
            break;
        default: return fail(h, VRESTIR_EINVAL, "unknown stage");
    }
    if (stage == VRESTIR_STAGE_TRACE_RAYS && learnedActive(h)) {
        int rc = learnedFrameBuffers(h);  // options may have changed since prepare_frame
        if (rc) return rc;
        a.learnedOn = 1;
        a.learned = {h->lCdf, h->lRecs, h->L.trainBlockX, h->L.trainBlockY};
    }
    // the grids each stage traverses most: its options' visibility and lighting mips
    const Frame& F = h->frame;
    int oa = -1, ob = -1;
    switch (stage) {
        case VRESTIR_STAGE_GENERATE_FEATURES: oa = 0; break;
        case VRESTIR_STAGE_TRACE_RAYS:
            if (p.useReference) oa = 0;
            else { oa = F.initialOpt.visibilityMipLevel; ob = F.initialOpt.lightingMipLevel; }
            break;
        case VRESTIR_STAGE_TEMPORAL_REUSE:
        case VRESTIR_STAGE_SPATIAL_REUSE:
            oa = F.spatialOpt.visibilityMipLevel; ob = F.spatialOpt.lightingMipLevel;
            break;
        case VRESTIR_STAGE_FINAL_SHADING: oa = F.finalOpt.visibilityMipLevel; ob = 0; break;
        default: break;
    }
    a.occBytes = occPlan(h, oa, ob, a.frame.occSlot);
    if (stage == VRESTIR_STAGE_TRACE_RAYS || stage == VRESTIR_STAGE_TEMPORAL_REUSE ||
        stage == VRESTIR_STAGE_SPATIAL_REUSE || stage == VRESTIR_STAGE_FINAL_SHADING) {
        int rc = splitPlan(h, a, stage);
        if (rc) return rc;
    }
    return VRESTIR_OK;
}

extern "C" int32_t vrestir_run_stage(vrestir_handle* h, void* stream, int32_t stage, int32_t round, float* outRGBA) {
    if (!h) return VRESTIR_EINVAL;
    if (!h->frameOpen) return fail(h, VRESTIR_ESTATE, "vrestir_prepare_frame must precede stages");
    const vrestir_params& p = h->P;
    hipStream_t st = (hipStream_t)stream;
    if (p.useSurfaceScene && stage != VRESTIR_STAGE_GENERATE_FEATURES)
        return fail(h, VRESTIR_EUNSUPPORTED, "surface scenes: only GenerateFeatures (V-buffer) is built");
    if (stage == VRESTIR_STAGE_LEARNED_PROPOSAL) return learnedProposal(h, st);
    if (stage == VRESTIR_STAGE_LEARNED_TRAIN) return learnedTrainFrame(h, st);
    LaunchArgs a;
    int rc = stageArgs(h, st, stage, round, outRGBA, a);
    if (rc) return rc;
    const size_t n = (size_t)h->W * h->H, ne = n * (size_t)std::max(p.maxBounces - 1, 1);
    const int t = h->totalRoundId;
    HIPCHK(h, launchStage(a));
    if (stage == VRESTIR_STAGE_TEMPORAL_REUSE) {
        if (!p.enableSpatialReuse) {
            HIPCHK(h, hipMemcpyAsync(h->resT, h->res[t % 2], n * sizeof(ResStore), hipMemcpyDeviceToDevice, st));
            if (p.maxBounces > 1)
                HIPCHK(h, hipMemcpyAsync(h->extraT, h->extra[t % 2], ne * 12, hipMemcpyDeviceToDevice, st));
        }
        HIPCHK(h, hipMemcpyAsync(h->featuresT, h->features, n * 8, hipMemcpyDeviceToDevice, st));
    }
    if (stage == VRESTIR_STAGE_SPATIAL_REUSE) h->totalRoundId++;
    return VRESTIR_OK;
}

extern "C" int32_t vrestir_end_frame(vrestir_handle* h) {
    if (!h) return VRESTIR_EINVAL;
    h->temporalAccumulated = true;
    h->prevCam = h->cam;
    h->envLast = h->envCur;
    h->haveEnvLast = true;
    std::memcpy(h->prevView, h->view, sizeof(h->view));
    std::memcpy(h->prevProj, h->proj, sizeof(h->proj));
    h->frameCount++;
    h->frameOpen = false;
    return VRESTIR_OK;
}

extern "C" int32_t vrestir_execute(vrestir_handle* h, void* stream, int32_t W, int32_t H, float* outRGBA,
                                   float* outMotionVec) {
    if (!h) return VRESTIR_EINVAL;
    if (!outRGBA) return fail(h, VRESTIR_EINVAL, "outRGBA is NULL");
    if (h->P.useSurfaceScene)
        return fail(h, VRESTIR_EUNSUPPORTED, "surface scenes: only GenerateFeatures (V-buffer) is built");
    int rc = vrestir_prepare_frame(h, W, H);
    if (rc) return rc;
    const vrestir_params& p = h->P;
    if (h->freezeFrame) {
        // mFreezeFrame (VR/VolumetricReSTIR.cpp:499-780): only final shading runs, on the
        // reservoirs of totalRoundId = [spatial], with the previous frame count
        h->frame.frameCount = h->frameCount - 1;
        h->totalRoundId = p.enableSpatialReuse ? 1 : 0;
        if ((rc = vrestir_run_stage(h, stream, VRESTIR_STAGE_FINAL_SHADING, 0, outRGBA))) return rc;
        h->frameCount--;  // end_frame advances it; a frozen frame does not
        return vrestir_end_frame(h);
    }
    if ((rc = vrestir_run_stage(h, stream, VRESTIR_STAGE_GENERATE_FEATURES, 0, nullptr))) return rc;
    const bool learned = learnedActive(h);
    if (learned && (rc = vrestir_run_stage(h, stream, VRESTIR_STAGE_LEARNED_PROPOSAL, 0, nullptr))) return rc;
    if ((rc = vrestir_run_stage(h, stream, VRESTIR_STAGE_TRACE_RAYS, 0, nullptr))) return rc;
    if (learned && (rc = vrestir_run_stage(h, stream, VRESTIR_STAGE_LEARNED_TRAIN, 0, nullptr))) return rc;
    if (!p.useReference && p.enableTemporalReuse)
        if ((rc = vrestir_run_stage(h, stream, VRESTIR_STAGE_TEMPORAL_REUSE, 0, outMotionVec))) return rc;
    if (!p.useReference && p.enableSpatialReuse)
        for (int r = 0; r < p.spatialReuseRounds; r++)
            if ((rc = vrestir_run_stage(h, stream, VRESTIR_STAGE_SPATIAL_REUSE, r, nullptr))) return rc;
    if (!p.useReference && p.enableTemporalReuse)
        if ((rc = vrestir_run_stage(h, stream, VRESTIR_STAGE_COPY_RESERVOIRS, 0, nullptr))) return rc;
    if ((rc = vrestir_run_stage(h, stream, VRESTIR_STAGE_FINAL_SHADING, 0, outRGBA))) return rc;
    return vrestir_end_frame(h);
}

// mFreezeFrame: later executes only re-shade the last reservoirs (the camera
// animation modes pause the pass this way).
extern "C" int32_t vrestir_set_freeze_frame(vrestir_handle* h, int32_t freeze) {
    if (!h) return VRESTIR_EINVAL;
    h->freezeFrame = freeze != 0;
    return VRESTIR_OK;
}

// ------------------------------------------------------------------ buffers (exchange format)
static int bufInfo(vrestir_handle* h, int id, void** dev, size_t& elems, int& kind) {
    size_t n = (size_t)h->W * h->H;
    int B = h->allocB;
    size_t ne = n * (size_t)(B > 1 ? B - 1 : 0);
    kind = 0;
    switch (id) {
        case VRESTIR_BUF_FEATURES: *dev = h->features; elems = n * 8; return 0;
        case VRESTIR_BUF_TEMPORAL_FEATURES: *dev = h->featuresT; elems = n * 8; return 0;
        case VRESTIR_BUF_RESERVOIRS0: *dev = h->res[0]; elems = n; kind = 1; return 0;
        case VRESTIR_BUF_RESERVOIRS1: *dev = h->res[1]; elems = n; kind = 1; return 0;
        case VRESTIR_BUF_TEMPORAL_RESERVOIRS: *dev = h->resT; elems = n; kind = 1; return 0;
        case VRESTIR_BUF_EXTRA0: *dev = h->extra[0]; elems = ne * 12; return 0;
        case VRESTIR_BUF_EXTRA1: *dev = h->extra[1]; elems = ne * 12; return 0;
        case VRESTIR_BUF_TEMPORAL_EXTRA: *dev = h->extraT; elems = ne * 12; return 0;
        case VRESTIR_BUF_COLOR: *dev = h->outInternal; elems = n * 16; return 0;
        case VRESTIR_BUF_LEARNED_CDF: *dev = h->lCdf; elems = h->lCdfPixels * 64; return 0;
        case VRESTIR_BUF_TRAIN_RECORDS: *dev = h->lRecs; elems = (size_t)h->lRecSlots * kLearnedRecFloats * 4; return 0;
        case VRESTIR_BUF_TRACE_REDONE: *dev = h->split.fix; elems = 4; return 0;
        case VRESTIR_BUF_VBUFFER: *dev = h->vbuf; elems = n * sizeof(VItem); return 0;
    }
    return -1;
}

extern "C" int32_t vrestir_read_buffer(vrestir_handle* h, int32_t id, void* dst, size_t bytes) {
    if (!h || !dst) return VRESTIR_EINVAL;
    void* dev;
    size_t elems;
    int kind;
    if (bufInfo(h, id, &dev, elems, kind) || !dev) return fail(h, VRESTIR_EINVAL, "unknown or unallocated buffer");
    HIPCHK(h, hipSetDevice(h->device));
    HIPCHK(h, hipDeviceSynchronize());
    if (kind == 0) {
        if (bytes != elems) return fail(h, VRESTIR_EINVAL, "byte count does not match the buffer");
        HIPCHK(h, hipMemcpy(dst, dev, bytes, hipMemcpyDeviceToHost));
        return VRESTIR_OK;
    }
    if (bytes != elems * 36) return fail(h, VRESTIR_EINVAL, "byte count does not match the buffer");
    std::vector<ResStore> tmp(elems);
    HIPCHK(h, hipMemcpy(tmp.data(), dev, elems * sizeof(ResStore), hipMemcpyDeviceToHost));
    int B = h->allocB;
    char* o = (char*)dst;
    for (size_t i = 0; i < elems; i++) {
        float f[8] = {tmp[i].a.x, tmp[i].a.y, tmp[i].a.z, tmp[i].a.w, tmp[i].b.x, tmp[i].b.y, tmp[i].b.z, tmp[i].b.w};
        std::memcpy(o + 36 * i, f, 32);
        int32_t ex = B > 1 ? (int32_t)(i * (size_t)(B - 1)) : 0;
        std::memcpy(o + 36 * i + 32, &ex, 4);
    }
    return VRESTIR_OK;
}

extern "C" int32_t vrestir_write_buffer(vrestir_handle* h, int32_t id, const void* src, size_t bytes) {
    if (!h || !src) return VRESTIR_EINVAL;
    void* dev;
    size_t elems;
    int kind;
    if (bufInfo(h, id, &dev, elems, kind) || !dev) return fail(h, VRESTIR_EINVAL, "unknown or unallocated buffer");
    HIPCHK(h, hipSetDevice(h->device));
    HIPCHK(h, hipDeviceSynchronize());
    if (kind == 0) {
        if (bytes != elems) return fail(h, VRESTIR_EINVAL, "byte count does not match the buffer");
        HIPCHK(h, hipMemcpy(dev, src, bytes, hipMemcpyHostToDevice));
        return VRESTIR_OK;
    }
    if (bytes != elems * 36) return fail(h, VRESTIR_EINVAL, "byte count does not match the buffer");
    std::vector<ResStore> tmp(elems);
    const char* s = (const char*)src;
    for (size_t i = 0; i < elems; i++) {
        float f[8];
        std::memcpy(f, s + 36 * i, 32);
        tmp[i].a = make_float4(f[0], f[1], f[2], f[3]);
        tmp[i].b = make_float4(f[4], f[5], f[6], f[7]);
    }
    HIPCHK(h, hipMemcpy(dev, tmp.data(), elems * sizeof(ResStore), hipMemcpyHostToDevice));
    return VRESTIR_OK;
}

// ------------------------------------------------------------------ ground truth and error
// FP64 accumulation of frames (AccumulatePass's double mode,
// Source/RenderPasses/AccumulatePass/AccumulatePass.cpp:42,93).
extern "C" int32_t vrestir_accumulate(vrestir_handle* h, void* stream, const float* frameRGBA, int32_t width,
                                      int32_t height, int32_t reset) {
    if (!h || !frameRGBA || width <= 0 || height <= 0) return VRESTIR_EINVAL;
    HIPCHK(h, hipSetDevice(h->device));
    const int n = width * height;
    if (n != h->accumPixels) {
        dfree(h->accum);
        HIPCHK(h, hipMalloc((void**)&h->accum, (size_t)n * sizeof(double4)));
        h->accumPixels = n;
        reset = 1;
    }
    if (reset) h->accumFrames = 0;
    HIPCHK(h, launchAccumulate((hipStream_t)stream, (const float4*)frameRGBA, h->accum, n, reset != 0));
    h->accumFrames++;
    return VRESTIR_OK;
}

// Mean of the accumulated frames, rounded once to float.
extern "C" int32_t vrestir_accumulated(vrestir_handle* h, void* stream, float* outRGBA, int64_t* frames) {
    if (!h || !outRGBA) return VRESTIR_EINVAL;
    if (h->accumFrames == 0) return fail(h, VRESTIR_ESTATE, "nothing accumulated");
    HIPCHK(h, hipSetDevice(h->device));
    HIPCHK(h, launchResolve((hipStream_t)stream, h->accum, 1.0 / (double)h->accumFrames, (float4*)outRGBA,
                            h->accumPixels));
    if (frames) *frames = h->accumFrames;
    return VRESTIR_OK;
}

// relMSE of an RGBA image against a reference (ImageCompare's "rmse",
// Source/Tools/ImageCompare/ImageCompare.cpp:172-180). Waits for the stream.
extern "C" int32_t vrestir_relmse(vrestir_handle* h, void* stream, const float* imageRGBA, const float* referenceRGBA,
                                  int32_t width, int32_t height, double* out) {
    if (!h || !imageRGBA || !referenceRGBA || !out || width <= 0 || height <= 0) return VRESTIR_EINVAL;
    HIPCHK(h, hipSetDevice(h->device));
    const int n = width * height;
    const int m = relMsePartialCount(n);
    if (m > h->relPartialN) {
        dfree(h->relPartial);
        HIPCHK(h, hipMalloc((void**)&h->relPartial, (size_t)m * sizeof(double)));
        h->relPartialN = m;
    }
    if (!h->relOut) HIPCHK(h, hipMalloc((void**)&h->relOut, sizeof(double)));
    hipStream_t st = (hipStream_t)stream;
    HIPCHK(h, launchRelMse(st, (const float4*)imageRGBA, (const float4*)referenceRGBA, n, h->relPartial, h->relOut));
    HIPCHK(h, hipMemcpyAsync(out, h->relOut, sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(h, hipStreamSynchronize(st));
    return VRESTIR_OK;
}

static_assert(VRESTIR_RELMSE_NOALLOC_MAX_PIXELS == kRelMseNoAllocMaxPixels, "no-alloc relMSE limit");

// vrestir_relmse without allocating: the partial sums and the result live in the
// library's static device scratch (vr_k_accum.hip), whose address was looked up
// when the handle was created on this device.
extern "C" int32_t vrestir_relmse_noalloc(vrestir_handle* h, void* stream, const float* imageRGBA,
                                          const float* referenceRGBA, int32_t width, int32_t height, double* out) {
    if (!h) return VRESTIR_EINVAL;
    if (!imageRGBA || !referenceRGBA || !out || width <= 0 || height <= 0)
        return fail(h, VRESTIR_EINVAL, "bad image arguments");
    if ((int64_t)width * height > kRelMseNoAllocMaxPixels)
        return fail(h, VRESTIR_EINVAL, "image larger than VRESTIR_RELMSE_NOALLOC_MAX_PIXELS");
    HIPCHK(h, hipSetDevice(h->device));
    HIPCHK(h, relMseNoAlloc((hipStream_t)stream, (const float4*)imageRGBA, (const float4*)referenceRGBA,
                            width * height, out));
    return VRESTIR_OK;
}

extern "C" int32_t vrestir_track_batch(vrestir_handle* h, void* stream, int32_t mode, int32_t mip,
                                       int32_t useLinearSampler, float tStepScale, int32_t numSamples, int32_t n,
                                       const float* rays, const uint32_t* seeds, float* out) {
    if (!h) return VRESTIR_EINVAL;
    if (!h->volumeSet) return fail(h, VRESTIR_ESTATE, "no volume bound");
    if (mode < 0 || mode > VRESTIR_TRACK_SURFACE_SHADE) return fail(h, VRESTIR_EINVAL, "unknown tracking mode");
    if (mode == VRESTIR_TRACK_EMISSION && h->sceneHost.hasEmission && h->sceneHost.lutRows == 0)
        return fail(h, VRESTIR_ESTATE, "no blackbody table bound");
    if (mip < 0 || mip >= kMaxSlots || !h->meta.slots[mip].valid) return fail(h, VRESTIR_EINVAL, "grid slot not bound");
    if (n < 0 || (n > 0 && (!rays || !seeds || !out))) return fail(h, VRESTIR_EINVAL, "bad ray batch");
    HIPCHK(h, hipSetDevice(h->device));
    int occSlots[2];
    const int occBytes = occPlan(h, mip, -1, occSlots);
    HIPCHK(h, launchTrack((hipStream_t)stream, h->sceneDev, h->envCur, mode, mip, useLinearSampler, tStepScale, numSamples, n,
                          rays, seeds, out, occSlots[0], occBytes));
    return VRESTIR_OK;
}

// Size of a bound grid slot: voxel dimensions, occupied bricks, brick cells, and
// the device bytes of its brick pool (with aprons), tables and per-brick bounds.
extern "C" int32_t vrestir_grid_info(const vrestir_handle* h, int32_t slot, int32_t dims[3], int64_t* numBricks,
                                     int64_t* brickCells, int64_t* deviceBytes) {
    if (!h || slot < 0 || slot >= kMaxSlots) return VRESTIR_EINVAL;
    const HostGrid& g = h->meta.slots[slot];
    if (!h->volumeSet || !g.valid) return VRESTIR_EINVAL;
    if (dims) { dims[0] = g.nx; dims[1] = g.ny; dims[2] = g.nz; }
    const int64_t cells = (int64_t)g.nbx * g.nby * g.nbz;
    if (numBricks) *numBricks = g.numBricks;
    if (brickCells) *brickCells = cells;
    if (deviceBytes)
        *deviceBytes = (int64_t)g.numBricks * kBrickVals * (g.fp16 ? 2 : 4) * g.channels + cells * 4 +
                       (int64_t)g.nnx * g.nny * g.nnz * 4 + ((cells + 31) / 32) * 8 + (int64_t)g.numBricks * 16;
    return VRESTIR_OK;
}

// ------------------------------------------------------------------ batches
// Many passes (views and volumes of uneven size) rendered together: each step of
// the frame -- a stage, or one record / trace / replay pass of a split stage --
// is one persistent launch over every pass's tiles (vr_k_batch.hip), and the
// frame's launches are captured into one HIP graph and replayed. The passes keep
// their own scenes, buffers and frame state; they must share the options that
// select kernels: path length, spatial rounds, temporal and spatial reuse on, the
// ray-marching spatial and initial-lighting trackers, analytic final shading, the
// split stage forms, no learned proposal, no reference mode.
struct vrestir_batch {
    std::vector<vrestir_handle*> passes;
    int device = 0;
    int grid = 1024, per = 1;  // persistent workgroups per launch, tiles per queue item
    int useGraph = 1;
    hipStream_t stream = nullptr;  // the batch's own stream (graph capture needs a non-null one)
    hipEvent_t evIn = nullptr, evOut = nullptr;
    int stepCap = 0;
    LaunchArgs* argsDev = nullptr;
    BatchList* listsDev = nullptr;
    int* countersDev = nullptr;
    LaunchArgs* argsHost = nullptr;  // pinned staging
    BatchList* listsHost = nullptr;
    std::vector<int> ops, occ;
    hipGraphExec_t exec = nullptr;
    std::vector<int> capturedKey;
    std::string err;
};

static int bfail(vrestir_batch* b, int code, const std::string& msg) {
    if (b) b->err = msg;
    return code;
}
#define BHIP(b, call)                                                                             \
    do {                                                                                          \
        hipError_t e_ = (call);                                                                   \
        if (e_ != hipSuccess) return bfail((b), VRESTIR_EHIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

static void batchFreeGraph(vrestir_batch* b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    b->exec = nullptr;
    b->capturedKey.clear();
}

extern "C" int32_t vrestir_batch_create(vrestir_handle* const* passes, int32_t count, vrestir_batch** out) {
    if (!out) return VRESTIR_EINVAL;
    *out = nullptr;
    if (!passes || count < 1 || count > kMaxBatch) return fail(nullptr, VRESTIR_EINVAL, "batch of 1..64 passes");
    for (int i = 0; i < count; i++) {
        if (!passes[i]) return fail(nullptr, VRESTIR_EINVAL, "a pass handle is NULL");
        if (passes[i]->device != passes[0]->device) return fail(nullptr, VRESTIR_EINVAL, "passes on different GPUs");
        for (int j = 0; j < i; j++)
            if (passes[j] == passes[i]) return fail(nullptr, VRESTIR_EINVAL, "a pass appears twice");
    }
    vrestir_batch* b = new vrestir_batch();
    b->passes.assign(passes, passes + count);
    b->device = passes[0]->device;
    if (hipSetDevice(b->device) != hipSuccess || hipStreamCreate(&b->stream) != hipSuccess ||
        hipEventCreateWithFlags(&b->evIn, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&b->evOut, hipEventDisableTiming) != hipSuccess) {
        delete b;
        return fail(nullptr, VRESTIR_EHIP, "batch stream / events");
    }
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b->device) == hipSuccess && cus > 0)
        b->grid = 4 * cus;  // four 256-lane workgroups per CU: 16 waves, the tracing kernels' occupancy
    *out = b;
    return VRESTIR_OK;
}

extern "C" int32_t vrestir_batch_destroy(vrestir_batch* b) {
    if (!b) return VRESTIR_OK;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    batchFreeGraph(b);
    dfree(b->argsDev); dfree(b->listsDev); dfree(b->countersDev);
    if (b->argsHost) (void)hipHostFree(b->argsHost);
    if (b->listsHost) (void)hipHostFree(b->listsHost);
    if (b->evIn) (void)hipEventDestroy(b->evIn);
    if (b->evOut) (void)hipEventDestroy(b->evOut);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return VRESTIR_OK;
}

extern "C" const char* vrestir_batch_last_error(const vrestir_batch* b) { return b ? b->err.c_str() : ""; }

extern "C" int32_t vrestir_batch_set_tuning(vrestir_batch* b, int32_t grid, int32_t tilesPerItem, int32_t useGraph) {
    if (!b) return VRESTIR_EINVAL;
    if (grid < 1 || grid > 65536 || tilesPerItem < 1 || tilesPerItem > 64)
        return bfail(b, VRESTIR_EINVAL, "grid in [1, 65536], tiles per item in [1, 64]");
    if (grid != b->grid || tilesPerItem != b->per || (useGraph != 0) != (b->useGraph != 0)) batchFreeGraph(b);
    b->grid = grid; b->per = tilesPerItem; b->useGraph = useGraph != 0;
    return VRESTIR_OK;
}

extern "C" int32_t vrestir_batch_get_tuning(const vrestir_batch* b, int32_t* grid, int32_t* tilesPerItem) {
    if (!b || !grid || !tilesPerItem) return VRESTIR_EINVAL;
    *grid = b->grid; *tilesPerItem = b->per;
    return VRESTIR_OK;
}

static bool sameKernelOptions(const vrestir_params& a, const vrestir_params& c) {
    return a.maxBounces == c.maxBounces && a.spatialReuseRounds == c.spatialReuseRounds &&
           a.enableTemporalReuse == c.enableTemporalReuse && a.enableSpatialReuse == c.enableSpatialReuse &&
           a.useReference == c.useReference;
}

// One batched frame: every pass's frame prepared, its steps' launch arguments
// staged, the launches (or their graph) issued on the batch stream, the frame
// closed. The caller's stream waits for the batch and the batch for the caller.
static int batchFrame(vrestir_batch* b, hipStream_t user, const int32_t* widths, const int32_t* heights,
                      float* const* outs, bool allowGraph) {
    const int n = (int)b->passes.size();
    const vrestir_params& p0 = b->passes[0]->P;
    const bool learned = learnedActive(b->passes[0]);
    for (int i = 0; i < n; i++) {
        vrestir_handle* h = b->passes[i];
        if (!widths || !heights || !outs || !outs[i]) return bfail(b, VRESTIR_EINVAL, "sizes and outputs for every pass");
        if (!sameKernelOptions(h->P, p0)) return bfail(b, VRESTIR_EUNSUPPORTED, "passes differ in kernel options");
        if (learnedActive(h) != learned)
            return bfail(b, VRESTIR_EUNSUPPORTED, "passes differ in whether the learned proposal is on");
        if (!h->P.enableTemporalReuse || !h->P.enableSpatialReuse || h->P.useReference || h->freezeFrame ||
            h->splitReuse != 2 || h->P.useSurfaceScene)
            return bfail(b, VRESTIR_EUNSUPPORTED, "batches need temporal and spatial reuse, split stages, no reference");
    }
    // the learned proposal's Adam step takes host-computed bias corrections, so
    // frames with it run their launches directly instead of replaying a graph
    if (learned) allowGraph = false;
    // a batch runs TraceRays as one kernel: its passes plan no TraceRays split while staged
    struct InBatch {
        const std::vector<vrestir_handle*>& ps;
        explicit InBatch(const std::vector<vrestir_handle*>& v) : ps(v) {
            for (vrestir_handle* h : ps) h->inBatch = true;
        }
        ~InBatch() {
            for (vrestir_handle* h : ps) h->inBatch = false;
        }
    } inBatch(b->passes);
    BHIP(b, hipSetDevice(b->device));
    // the steps of the frame
    // steps with op < 0 are per-pass launches of the learned stages (-1 proposal, -2 training)
    struct Step { int op, stage, round; };
    std::vector<Step> steps = {{kOpFeatures, VRESTIR_STAGE_GENERATE_FEATURES, 0}};
    if (learned) steps.push_back({-1, VRESTIR_STAGE_LEARNED_PROPOSAL, 0});
    steps.push_back({learned ? kOpTraceLearned : kOpTrace, VRESTIR_STAGE_TRACE_RAYS, 0});
    if (learned) steps.push_back({-2, VRESTIR_STAGE_LEARNED_TRAIN, 0});
    steps.push_back({kOpTemporalRecord, VRESTIR_STAGE_TEMPORAL_REUSE, 0});
    steps.push_back({kOpSegmentsMarch, VRESTIR_STAGE_TEMPORAL_REUSE, 0});
    steps.push_back({kOpTemporalReplay, VRESTIR_STAGE_TEMPORAL_REUSE, 0});
    steps.push_back({kOpCopyFeatures, VRESTIR_STAGE_TEMPORAL_REUSE, 0});
    for (int r = 0; r < p0.spatialReuseRounds; r++) {
        steps.push_back({kOpSpatialRecord, VRESTIR_STAGE_SPATIAL_REUSE, r});
        steps.push_back({kOpSegmentsMarch, VRESTIR_STAGE_SPATIAL_REUSE, r});
        steps.push_back({kOpSpatialReplay, VRESTIR_STAGE_SPATIAL_REUSE, r});
    }
    steps.push_back({kOpCopyReservoirs, VRESTIR_STAGE_COPY_RESERVOIRS, 0});
    steps.push_back({kOpFinalRecord, VRESTIR_STAGE_FINAL_SHADING, 0});
    steps.push_back({kOpSegmentsCells, VRESTIR_STAGE_FINAL_SHADING, 0});
    steps.push_back({kOpFinalReplay, VRESTIR_STAGE_FINAL_SHADING, 0});
    int S = (int)steps.size();
    const int stepCap = S;  // the most steps a frame has (the temporal one-kernel form has fewer)
    if (stepCap > b->stepCap) {
        batchFreeGraph(b);
        dfree(b->argsDev); dfree(b->listsDev); dfree(b->countersDev);
        if (b->argsHost) (void)hipHostFree(b->argsHost);
        if (b->listsHost) (void)hipHostFree(b->listsHost);
        b->argsHost = nullptr; b->listsHost = nullptr; b->stepCap = 0;
        BHIP(b, hipMalloc((void**)&b->argsDev, sizeof(LaunchArgs) * (size_t)stepCap * kMaxBatch));
        BHIP(b, hipMalloc((void**)&b->listsDev, sizeof(BatchList) * (size_t)stepCap));
        BHIP(b, hipMalloc((void**)&b->countersDev, sizeof(int) * (size_t)stepCap));
        BHIP(b, hipHostMalloc((void**)&b->argsHost, sizeof(LaunchArgs) * (size_t)stepCap * kMaxBatch));
        BHIP(b, hipHostMalloc((void**)&b->listsHost, sizeof(BatchList) * (size_t)stepCap));
        b->stepCap = stepCap;
    }
    // the previous frame's staging may still be in flight
    BHIP(b, hipStreamSynchronize(b->stream));
    for (int i = 0; i < n; i++) {
        int rc = vrestir_prepare_frame(b->passes[i], widths[i], heights[i]);
        if (rc) return bfail(b, rc, std::string("pass ") + std::to_string(i) + ": " + b->passes[i]->err);
    }
    // Reservation pass: a stage's split plan (re)allocates its pass's record buffers
    // when it needs more rays per pixel than they hold, which would leave the
    // arguments of an earlier step of this frame pointing at freed buffers. Run
    // every step's plan once, so the buffers hold the frame's largest capacity,
    // then build the arguments that are launched.
    int temporalTK = -1;
    {
        std::vector<int> rounds(n);
        for (int i = 0; i < n; i++) rounds[i] = b->passes[i]->totalRoundId;
        for (const Step& sp : steps) {
            if (sp.op < 0) continue;
            for (int i = 0; i < n; i++) {
                LaunchArgs tmp;
                int rc = stageArgs(b->passes[i], b->stream, sp.stage, sp.round, nullptr, tmp);
                if (rc) return bfail(b, rc, std::string("pass ") + std::to_string(i) + ": " + b->passes[i]->err);
                if (sp.op == kOpSpatialReplay) b->passes[i]->totalRoundId++;
                if (sp.op == kOpTemporalRecord) {
                    if (temporalTK >= 0 && temporalTK != tmp.splitTK)
                        return bfail(b, VRESTIR_EUNSUPPORTED, "passes differ in the form of their temporal reuse");
                    temporalTK = tmp.splitTK;
                }
            }
        }
        for (int i = 0; i < n; i++) b->passes[i]->totalRoundId = rounds[i];
    }
    // temporal reuse on the previous frame's grids (an animated volume) runs as one kernel
    if (temporalTK == 0) {
        std::vector<Step> s2;
        for (const Step& sp : steps) {
            if (sp.op == kOpTemporalRecord) s2.push_back({kOpTemporalDirect, VRESTIR_STAGE_TEMPORAL_REUSE, 0});
            else if (!(sp.stage == VRESTIR_STAGE_TEMPORAL_REUSE &&
                       (sp.op == kOpSegmentsMarch || sp.op == kOpTemporalReplay)))
                s2.push_back(sp);
        }
        steps.swap(s2);
        S = (int)steps.size();
    }
    b->ops.assign(S, 0);
    b->occ.assign(S, 0);
    std::vector<int> key = {n, S, b->grid, b->per};
    for (int k = 0; k < S; k++) {
        const Step& sp = steps[k];
        BatchList& L = b->listsHost[k];
        std::memset(&L, 0, sizeof(L));
        L.n = n;
        L.per = b->per;
        b->ops[k] = sp.op;
        key.push_back(sp.op);
        if (sp.op < 0) continue;
        for (int i = 0; i < n; i++) {
            vrestir_handle* h = b->passes[i];
            LaunchArgs& a = b->argsHost[(size_t)k * kMaxBatch + i];
            int rc = stageArgs(h, b->stream, sp.stage, sp.round,
                               sp.stage == VRESTIR_STAGE_FINAL_SHADING ? outs[i] : nullptr, a);
            if (rc) return bfail(b, rc, std::string("pass ") + std::to_string(i) + ": " + h->err);
            const bool needSplit = (sp.stage == VRESTIR_STAGE_TEMPORAL_REUSE && temporalTK != 0) ||
                                   sp.stage == VRESTIR_STAGE_SPATIAL_REUSE || sp.stage == VRESTIR_STAGE_FINAL_SHADING;
            const int wantTK = sp.stage == VRESTIR_STAGE_FINAL_SHADING ? 1 : 2;
            const Options& so = a.frame.spatialOpt;
            if ((sp.op == kOpTrace || sp.op == kOpTraceLearned) && (a.frame.initialOpt.lightingTrackingMethod != 2u ||
                                      so.visibilityTrackingMethod != 2u || so.lightingTrackingMethod != 2u))
                return bfail(b, VRESTIR_EUNSUPPORTED, "batches need the ray-marching spatial and lighting trackers");
            if (needSplit && (a.splitTK != wantTK || a.splitMode != 2))
                return bfail(b, VRESTIR_EUNSUPPORTED, std::string("pass ") + std::to_string(i) +
                                                         ": a stage of this pass does not run in its split form");
            const bool seg = sp.op == kOpSegmentsMarch || sp.op == kOpSegmentsCells;
            const int W = a.width, H = a.height;
            const int gx = seg ? (W * H + 255) / 256 : (W + 15) / 16;
            const int tiles = seg ? gx : gx * ((H + 15) / 16);
            L.gx[i] = gx;
            L.tiles[i] = tiles;
            L.start[i + 1] = L.start[i] + (tiles + L.per - 1) / L.per;
            b->occ[k] = std::max(b->occ[k], a.occBytes);
            key.push_back(a.occBytes);
            key.push_back(tiles);
        }
        if (sp.stage == VRESTIR_STAGE_SPATIAL_REUSE && sp.op == kOpSpatialReplay)
            for (int i = 0; i < n; i++) b->passes[i]->totalRoundId++;
    }
    // every staged record buffer must be the one its pass holds now
    for (int k = 0; k < S; k++)
        for (int i = 0; i < n && b->ops[k] >= 0; i++) {
            const LaunchArgs& a = b->argsHost[(size_t)k * kMaxBatch + i];
            const vrestir_handle* h = b->passes[i];
            if (a.splitTK && (a.split.recs != h->split.recs || a.split.trs != h->split.trs ||
                              a.split.dens != h->split.dens || a.split.counts != h->split.counts))
                return bfail(b, VRESTIR_ESTATE, "record buffers changed while the frame was built");
        }
    // stage the arguments (the graph's kernels read them from device memory)
    BHIP(b, hipEventRecord(b->evIn, user));
    BHIP(b, hipStreamWaitEvent(b->stream, b->evIn, 0));
    for (int k = 0; k < S; k++)
        BHIP(b, hipMemcpyAsync(b->argsDev + (size_t)k * kMaxBatch, b->argsHost + (size_t)k * kMaxBatch,
                               sizeof(LaunchArgs) * n, hipMemcpyHostToDevice, b->stream));
    BHIP(b, hipMemcpyAsync(b->listsDev, b->listsHost, sizeof(BatchList) * S, hipMemcpyHostToDevice, b->stream));
    auto issue = [&]() -> int {
        BHIP(b, launchZeroCounters(b->stream, b->countersDev, S));
        for (int k = 0; k < S; k++) {
            const LaunchArgs* args = b->argsDev + (size_t)k * kMaxBatch;
            const int op = b->ops[k];
            if (op < 0) {
                for (int i = 0; i < n; i++) {
                    vrestir_handle* h = b->passes[i];
                    const int rc = op == -1 ? learnedProposal(h, b->stream) : learnedTrainFrame(h, b->stream);
                    if (rc) return bfail(b, rc, std::string("pass ") + std::to_string(i) + ": " + h->err);
                }
                continue;
            }
            if (op == kOpSegmentsMarch || op == kOpSegmentsCells)
                BHIP(b, launchBatchedSegments(b->stream, op, args, b->listsDev + k, b->countersDev + k, b->grid,
                                              b->occ[k]));
            else
                BHIP(b, launchBatched(b->stream, op, p0.maxBounces, args, b->listsDev + k, b->countersDev + k,
                                      b->grid, b->occ[k]));
        }
        return VRESTIR_OK;
    };
    int rc = VRESTIR_OK;
    if (allowGraph && b->useGraph) {
        if (!b->exec || key != b->capturedKey) {
            batchFreeGraph(b);
            hipGraph_t g = nullptr;
            BHIP(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            rc = issue();
            hipError_t ce = hipStreamEndCapture(b->stream, &g);
            if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
            BHIP(b, ce);
            hipError_t ie = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            BHIP(b, ie);
            b->capturedKey = key;
        }
        BHIP(b, hipGraphLaunch(b->exec, b->stream));
    } else {
        rc = issue();
        if (rc) return rc;
    }
    BHIP(b, hipEventRecord(b->evOut, b->stream));
    BHIP(b, hipStreamWaitEvent(user, b->evOut, 0));
    for (int i = 0; i < n; i++) vrestir_end_frame(b->passes[i]);
    return VRESTIR_OK;
}

extern "C" int32_t vrestir_batch_execute(vrestir_batch* b, void* stream, const int32_t* widths, const int32_t* heights,
                                         float* const* outRGBA) {
    if (!b) return VRESTIR_EINVAL;
    return batchFrame(b, (hipStream_t)stream, widths, heights, outRGBA, true);
}

// The project's autotuner: renders `frames` batched frames with each candidate
// (persistent workgroups x tiles per queue item), keeps the fastest by HIP-event
// time, then resets every pass's history so the caller's sequence starts afresh.
extern "C" int32_t vrestir_batch_autotune(vrestir_batch* b, void* stream, const int32_t* widths,
                                          const int32_t* heights, float* const* outRGBA, int32_t frames,
                                          float* bestMs) {
    if (!b || frames < 1 || frames > 16) return VRESTIR_EINVAL;
    hipStream_t user = (hipStream_t)stream;
    BHIP(b, hipSetDevice(b->device));
    int cus = 256;
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b->device);
    const int grids[3] = {2 * cus, 4 * cus, 8 * cus};
    const int pers[3] = {1, 2, 4};
    hipEvent_t e0, e1;
    BHIP(b, hipEventCreate(&e0));
    BHIP(b, hipEventCreate(&e1));
    float best = 1e30f;
    int bg = b->grid, bp = b->per;
    int rc = VRESTIR_OK;
    // one untimed frame allocates every pass's buffers
    b->grid = grids[1]; b->per = 1;
    rc = batchFrame(b, user, widths, heights, outRGBA, false);
    for (int gi = 0; gi < 3 && rc == VRESTIR_OK; gi++)
        for (int pi = 0; pi < 3 && rc == VRESTIR_OK; pi++) {
            b->grid = grids[gi]; b->per = pers[pi];
            if (hipEventRecord(e0, b->stream) != hipSuccess) { rc = VRESTIR_EHIP; break; }
            for (int f = 0; f < frames && rc == VRESTIR_OK; f++) rc = batchFrame(b, user, widths, heights, outRGBA, false);
            if (rc) break;
            float ms = 0.f;
            if (hipEventRecord(e1, b->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                hipEventElapsedTime(&ms, e0, e1) != hipSuccess) { rc = VRESTIR_EHIP; break; }
            if (ms < best) { best = ms; bg = grids[gi]; bp = pers[pi]; }
        }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    batchFreeGraph(b);
    b->grid = bg; b->per = bp;
    for (vrestir_handle* h : b->passes) vrestir_reset_history(h);
    if (bestMs) *bestMs = best / (float)frames;
    return VRESTIR_OK;
}

// ------------------------------------------------------------------ host helpers
extern "C" int32_t vrestir_synth_fbm(uint64_t seed, int32_t nx, int32_t ny, int32_t nz, int32_t octaves,
                                     float baseFrequency, float lacunarity, float gain, int32_t sphereMask,
                                     float threshold, float amplitude, float* out) {
    return synthFbm(seed, nx, ny, nz, octaves, baseFrequency, lacunarity, gain, sphereMask, threshold, amplitude, out);
}

extern "C" int32_t vrestir_count_bricks(const float* d, int32_t nx, int32_t ny, int32_t nz, int64_t* out) {
    if (!d || !out || nx % 8 || ny % 8 || nz % 8) return VRESTIR_EINVAL;
    int64_t c = 0;
    for (int bz = 0; bz < nz / 8; bz++)
        for (int by = 0; by < ny / 8; by++)
            for (int bx = 0; bx < nx / 8; bx++) {
                bool any = false;
                for (int z = 0; z < 8 && !any; z++)
                    for (int y = 0; y < 8 && !any; y++)
                        for (int x = 0; x < 8 && !any; x++)
                            if (d[(size_t)(bx * 8 + x) + (size_t)nx * ((size_t)(by * 8 + y) + (size_t)ny * (bz * 8 + z))] > 0.f)
                                any = true;
                c += any;
            }
    *out = c;
    return VRESTIR_OK;
}

// Orbit about `center` (camera animation mode 2, VR/VolumetricReSTIR.cpp:879-904):
// rotate the start position about the vertical axis through the point of the
// initial view ray closest to the centre, looking at that point.
// The camera animation modes of VR/VolumetricReSTIR.cpp:801-905, run as the
// reference's per-frame state machine from the toggle (framesMoved = 1, camera
// = base, not frozen) through `updates` end-of-frame updates.
namespace {
struct V3d { double x, y, z; };
V3d v3d(const float* a) { return {a[0], a[1], a[2]}; }
V3d operator+(V3d a, V3d b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
V3d operator-(V3d a, V3d b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
V3d operator*(double s, V3d a) { return {s * a.x, s * a.y, s * a.z}; }
double dotd(V3d a, V3d b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
V3d crossd(V3d a, V3d b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
V3d nrm(V3d a) { double l = std::sqrt(dotd(a, a)); return l > 0 ? (1.0 / l) * a : a; }
}  // namespace

extern "C" int32_t vrestir_camera_animation(const vrestir_camera* base, const vrestir_camera_animation_desc* a,
                                            int32_t updates, vrestir_camera* out, int32_t* freeze) {
    if (!base || !a || !out || updates < 0 || a->frameInterval <= 0) return VRESTIR_EINVAL;
    const double kPi = 3.14159265358979323846;
    V3d pos = v3d(base->posW), target = v3d(base->target), up = v3d(base->up);
    const V3d bpos = pos, btarget = target;
    V3d anchor2 = pos;
    bool frozen = false;
    int moved = 1;
    const int I = a->frameInterval;
    auto moveRight = [&](double dist, V3d anchor) {  // moveCameraRight (:256-263)
        V3d view = nrm(target - pos);
        V3d right = nrm(crossd(view, up));
        pos = anchor + dist * right;
        target = pos + view;
    };
    auto forward = [&](double dist, V3d anchor) {  // _forwardCameraInterval (:265-272)
        V3d view = nrm(target - anchor);
        V3d center = anchor + (double)a->forwardScale * view;
        pos = center - dist * view;
        target = pos + view;
    };
    for (int u = 0; u < updates; u++) {
        if (a->mode == 0) {
            const int roundFrames = a->shakeRoundsBeforePause * I + a->pauseInterval;
            const int total = a->shakeTotalRounds * roundFrames;
            const double speed = 2 * kPi / I;
            if (moved++ < total) {
                if (moved % roundFrames < a->shakeRoundsBeforePause * I) {
                    frozen = false;
                    moveRight(std::sin(speed * moved) * a->moveScale, bpos);
                } else {
                    frozen = true;
                }
            }
        } else if (a->mode == 1) {
            const double speed = 2 * kPi / I;
            if (moved++ < 6 * I) {
                if (moved < I) {
                    moveRight(std::sin(speed * moved) * a->moveScale, bpos);
                } else if (moved < 2 * I) {
                    frozen = true;
                } else if (moved < 3 * I) {
                    frozen = false;
                    moveRight(std::sin(speed * (moved - I)) * a->moveScale, bpos);
                    anchor2 = pos;
                } else if (moved < 3 * I + I / 4) {
                    frozen = false;
                    forward(a->forwardScale * std::cos(speed * (moved - 3 * I)), anchor2);
                } else if (moved < 4 * I + I / 4) {
                    frozen = true;
                } else if (moved < 4 * I + 3 * I / 4) {
                    frozen = false;
                    forward(a->forwardScale * std::cos(speed * (moved - 4 * I)), anchor2);
                } else if (moved < 5 * I + 3 * I / 4) {
                    frozen = true;
                } else {
                    frozen = false;
                    forward(a->forwardScale * std::cos(speed * (moved - 5 * I)), anchor2);
                }
            }
        } else {
            // mode 2: rotation about the point of the initial view ray closest to the volume centre
            V3d vdir = nrm(btarget - bpos);
            V3d rotCenter = bpos + dotd(v3d(a->volumeCenter) - bpos, vdir) * vdir;
            if (moved++ < a->rotationFrames) {
                frozen = false;
                const double ang = moved * (double)a->rotationSpeed * kPi / 180.0;
                V3d X = nrm(bpos - rotCenter);
                V3d Y = nrm(crossd({0, 1, 0}, X));
                X = nrm(crossd(Y, {0, 1, 0}));
                V3d actual = rotCenter + (bpos - rotCenter) - dotd(bpos - rotCenter, X) * X;
                V3d d = bpos - actual;
                double dist = a->rotationDistance <= 0 ? std::sqrt(dotd(d, d)) : (double)a->rotationDistance;
                pos = actual + dist * (std::cos(ang) * X + std::sin(ang) * Y);
                target = rotCenter;
            }
        }
    }
    *out = *base;
    out->posW[0] = (float)pos.x; out->posW[1] = (float)pos.y; out->posW[2] = (float)pos.z;
    out->target[0] = (float)target.x; out->target[1] = (float)target.y; out->target[2] = (float)target.z;
    if (freeze) *freeze = frozen ? 1 : 0;
    return VRESTIR_OK;
}

extern "C" int32_t vrestir_orbit_camera(const vrestir_camera* base, const float center[3], int32_t frame,
                                        float degreesPerFrame, vrestir_camera* out) {
    if (!base || !center || !out) return VRESTIR_EINVAL;
    double P[3] = {base->posW[0], base->posW[1], base->posW[2]}, T[3] = {base->target[0], base->target[1], base->target[2]};
    double f[3] = {T[0] - P[0], T[1] - P[1], T[2] - P[2]};
    double fl = std::sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
    for (double& x : f) x /= fl;
    double cp[3] = {center[0] - P[0], center[1] - P[1], center[2] - P[2]};
    double proj = cp[0] * f[0] + cp[1] * f[1] + cp[2] * f[2];
    double rc[3] = {P[0] + proj * f[0], P[1] + proj * f[1], P[2] + proj * f[2]};
    double X[3] = {P[0] - rc[0], 0.0, P[2] - rc[2]};
    double xl = std::sqrt(X[0] * X[0] + X[2] * X[2]);
    *out = *base;
    if (xl <= 0.0) return VRESTIR_OK;
    X[0] /= xl; X[2] /= xl;
    double Y[3] = {X[2], 0.0, -X[0]};  // cross((0,1,0), X)
    double ang = (double)frame * degreesPerFrame * 3.14159265358979323846 / 180.0;
    double arc[3] = {rc[0], P[1], rc[2]};
    for (int i = 0; i < 3; i++) {
        out->posW[i] = (float)(arc[i] + xl * (X[i] * std::cos(ang) + Y[i] * std::sin(ang)));
        out->target[i] = (float)rc[i];
    }
    return VRESTIR_OK;
}
