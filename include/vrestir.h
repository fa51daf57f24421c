/*
 * vrestir.h -- C ABI of the MI355X-native VolumetricReSTIR render pass.
 *
 * This is the drop-in boundary for the reference's render-pass plugin
 * (Source/RenderPasses/VolumetricReSTIR). Every entry point below names the
 * reference interface it replaces. Plain C types only: no HIP, no torch, no
 * C++ in the signatures. All functions return a vrestir_status; the text of
 * the last failure on a handle is available from vrestir_last_error().
 *
 * Threading: one handle per host thread / HIP stream (the reference executes
 * its pass serially from one render thread, RenderGraphExe.cpp:33-42).
 * Ownership: the handle owns its device buffers (volume brick pools,
 * reservoirs, features); the caller owns the output image and the stream.
 */
#ifndef VRESTIR_H
#define VRESTIR_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VRESTIR_ABI_VERSION 1

typedef enum vrestir_status {
    VRESTIR_OK = 0,
    VRESTIR_EINVAL = 1,       /* bad argument (replaces C++ exceptions / logError) */
    VRESTIR_EIO = 2,          /* file could not be read or parsed (.vbx) */
    VRESTIR_EHIP = 3,         /* a HIP runtime call failed */
    VRESTIR_ENOMEM = 4,       /* device or host allocation failed */
    VRESTIR_ESTATE = 5,       /* call out of order (e.g. execute before set_volume) */
    VRESTIR_EUNSUPPORTED = 6  /* option the pass does not implement (vertex reuse; surface vertices in the reservoir stages) */
} vrestir_status;

/* Tracking methods, MIS, neighbour sampler, reprojection modes:
 * same numeric values as VR/HostDeviceSharedConstants.slang:5-15,30-32. */
enum {
    VRESTIR_RATIO_TRACKING = 0,
    VRESTIR_ANALYTIC_TRACKING = 1,
    VRESTIR_RAY_MARCHING = 2,
    VRESTIR_RESIDUAL_RATIO_TRACKING = 3,
    VRESTIR_ANALOG_RESIDUAL_RATIO_TRACKING = 4
};
enum { VRESTIR_MIS_NONE = 0, VRESTIR_MIS_TALBOT = 1 };
enum { VRESTIR_SAMPLER_HAMMERSLEY = 0, VRESTIR_SAMPLER_R2 = 1 };
enum { VRESTIR_REPROJECTION_LINEAR = 0, VRESTIR_REPROJECTION_NONE = 1, VRESTIR_REPROJECTION_NO_BACKGROUND = 2 };

/* Option surface: one field per VolumetricReSTIRParams member
 * (VR/VolumetricReSTIR.h:130-206), same meaning, same defaults
 * (vrestir_default_params). Booleans are int32 0/1. */
typedef struct vrestir_params {
    int32_t maxBounces;
    int32_t enableTemporalReuse;
    int32_t enableSpatialReuse;
    int32_t vertexReuse;                 /* only 0 is implemented (VRESTIR_EUNSUPPORTED) */
    int32_t vertexReuseStartBounce;
    int32_t useReference;                /* volume path tracer instead of ReSTIR */
    int32_t useEnvironmentLights;
    int32_t useAnalyticLights;
    int32_t useEmissiveLights;           /* emissive triangles of vrestir_set_surfaces as lights */
    int32_t baselineSamplePerPixel;
    int32_t visualizeTotalTransmittance;
    int32_t useSurfaceScene;             /* GenerateFeatures only: V-buffer and surface-limited features;
                                            the other stages refuse (VRESTIR_EUNSUPPORTED) */
    int32_t usePrevVolumeForReproj;
    /* initial sampling */
    int32_t initialBaseMipLevel;
    int32_t initialM;
    int32_t initialLightSamples;
    int32_t initialLightingMipLevel;
    int32_t initialVisibilityUseLinearSampler;
    int32_t initialLightingUseLinearSampler;
    uint32_t initialLightingTrackingMethod;
    float initialVisibilityTStepScale;
    float initialLightingTStepScale;
    int32_t initialUseRussianRoulette;
    int32_t initialUseCoarserGridForIndirectBounce;
    /* temporal reuse */
    float temporalReuseMThreshold;       /* multiplicative: cap = threshold * M_current */
    uint32_t temporalReprojectionMode;
    uint32_t temporalMISMethod;
    int32_t temporalReprojectionMipLevel;
    /* spatial reuse */
    int32_t spatialReuseRounds;
    int32_t spatialVisibilityMipLevel;
    int32_t spatialLightingMipLevel;
    int32_t spatialVisibilityUseLinearSampler;
    int32_t spatialLightingUseLinearSampler;
    float spatialVisibilityTStepScale;
    float spatialLightingTStepScale;
    uint32_t spatialVisibilityTrackingMethod;
    uint32_t spatialLightingTrackingMethod;
    uint32_t randomSamplerType;
    float sampleRadius;
    int32_t spatialSampleCount;          /* neighbours + 1 (the pixel itself) */
    int32_t enableVisibilitySimilarityRejection; /* kept for ABI parity; never read (as in the reference) */
    uint32_t spatialMISMethod;
    /* final shading */
    int32_t finalLightSamples;
    int32_t finalVisibilitySamples;
    uint32_t finalVisibilityTrackingMethod;
    uint32_t finalLightTrackingMethod;
    uint32_t finalRandomSamplerType;     /* kept for ABI parity; never read (as in the reference) */
    float finalTStepScale;
} vrestir_params;

enum { VRESTIR_BRICK_FP32 = 0, VRESTIR_BRICK_FP16 = 1 };

/* A density volume given as a dense mip-0 grid (the form every synthetic
 * workload takes); replaces Scene::addGVDBVolume (SC/Scene.h:481,
 * SC/Scene.cpp:2754-3326) for in-memory grids. Mips 1..numMips-1 are 2x box
 * reductions, conservative grids (slots 8..15) follow the reference's
 * 27-neighbour dilation of mip 0 (GL/gvdb_volume_gvdb.cpp:2753-2800). */
typedef struct vrestir_volume_desc {
    int32_t nx, ny, nz;          /* mip-0 voxels per axis, each a multiple of 8 << (numMips-1) */
    const float* density;        /* host memory, nx*ny*nz values, x fastest; values < 0 are clamped to 0 */
    int32_t numMips;             /* 1..8 */
    int32_t brickFormat;         /* VRESTIR_BRICK_FP32 / VRESTIR_BRICK_FP16 */
    float indexToWorld[16];      /* mip-0 index space -> world, row-vector affine (translation in [12..14]) */
    float sigmaA[3];
    float sigmaS[3];             /* sigma_a + sigma_s must be grey (SC/Scene.cpp:3247) */
    float g;                     /* Henyey-Greenstein anisotropy */
    float densityScale;          /* densityScaleFactorByScaling */
    const float* temperature;    /* optional temperature grid (same dims and transform) or NULL: the volume
                                    emits LeScale * blackbody(min(6400, (T - temperatureCutOff) *
                                    temperatureScale)) (VR/VolumeBase.slang:211-232; needs
                                    vrestir_set_blackbody_lut) */
    float LeScale, temperatureCutOff, temperatureScale;
    const float* velocity;       /* optional velocity grid (same dims and transform) or NULL: 3 floats per voxel,
                                    world-space displacement per frame, used by temporal reprojection while a
                                    sequence animates (VR/TemporalReuse.cs.slang:161-166) */
    float velocityScale;         /* VolumeDesc::velocityScale */
} vrestir_volume_desc;

/* A lat-long environment map (Scene::setEnvMap / setEnvMapIntensity / setEnvMapRotation
 * of the Mogwai scripts; LT/EnvMap.cpp:57-122, LT/EnvMapSampler.cpp:61-110). u runs
 * with atan2(x, -z), v from +y (v = 0) to -y (world_to_latlong_map,
 * UT/Math/MathHelpers.slang:92-99); bilinear filter, wrap in u, clamp in v. */
typedef struct vrestir_env_map_desc {
    int32_t width, height;       /* texels */
    const float* rgb;            /* host memory, width*height RGB triples, row 0 at v = 0; finite, >= 0 */
    float intensity;
    float tint[3];
    float rotationDegrees[3];    /* XYZ Euler angles: local -> world = Rz * Ry * Rx */
    int32_t importanceDim;       /* importance map side, power of two in [2, 2048]; 0 = 512 (the reference's) */
    int32_t importanceSpp;       /* samples per importance texel: 1, 4, 16, 64, 256, 1024; 0 = 64 */
} vrestir_env_map_desc;

/* Point light (LT/LightHelpers.slang:223-244): Li = intensity / dist^2. */
typedef struct vrestir_point_light {
    float posW[3];
    float intensity[3];
} vrestir_point_light;

/* Untextured standard material, metal-rough shading model (SC/Material/Material.h
 * and the MetalRough branch of _prepareShadingData, SC/Shading.slang:155-230). */
typedef struct vrestir_material {
    float baseColor[3];
    float roughness;             /* linear roughness (specular.g) */
    float metallic;              /* specular.b */
    float IoR;                   /* dielectric F0 = ((IoR - 1) / (IoR + 1))^2; 1.5 gives 0.04 */
    float emissive[3];           /* emissive colour x emissive factor; triangles emit on their front side */
    int32_t doubleSided;
} vrestir_material;

/* Triangle geometry in world space (the scene's meshes after their instance
 * transforms). Triangles whose material emits form the light collection
 * (LT/LightCollection.cpp: face normal and area per triangle, flux =
 * luminance(emissive) * area * pi, active when the flux is positive), sampled by
 * the power sampler (LT/EmissivePowerSampler.slang:56-92: an alias table over the
 * flux, then a uniform point on the triangle). */
typedef struct vrestir_surfaces_desc {
    const float* positions;      /* 3 floats per vertex */
    const float* normals;        /* 3 floats per vertex (shading normals), or NULL for face normals */
    int32_t numVertices;
    const int32_t* indices;      /* 3 vertex indices per triangle */
    const int32_t* materialIds;  /* one material index per triangle */
    int32_t numTriangles;
    const vrestir_material* materials;
    int32_t numMaterials;
    float emissiveIntensityMultiplier;  /* Scene::emissiveIntensityMultiplier */
} vrestir_surfaces_desc;

/* Pinhole camera (SC/Camera/Camera.cpp:113-189), no jitter. */
typedef struct vrestir_camera {
    float posW[3];
    float target[3];
    float up[3];
    float fovY;                  /* radians */
    float nearZ, farZ;
} vrestir_camera;

typedef struct vrestir_handle vrestir_handle;

/* ---- lifecycle: getPasses()/VolumetricReSTIR::create (VR/VolumetricReSTIR.cpp:56-70) ---- */
int32_t vrestir_version(void);
void vrestir_default_params(vrestir_params* out);                     /* VR/VolumetricReSTIR.h:132-206 */
int32_t vrestir_create(const vrestir_params* params, vrestir_handle** out); /* binds the current HIP device */
int32_t vrestir_destroy(vrestir_handle* h);
const char* vrestir_last_error(const vrestir_handle* h);               /* NULL handle: last create() error */

/* ---- scene binding: setScene (VR/VolumetricReSTIR.cpp:1218-1256) ---- */
int32_t vrestir_set_volume(vrestir_handle* h, const vrestir_volume_desc* desc);
/* A volume built on the GPU (no counterpart in the reference, which loads prebuilt
 * frames): desc->density and desc->temperature are DEVICE pointers (same layout as
 * above; velocity must be NULL), read on `stream`. The grids equal what
 * vrestir_set_volume builds from the same values, byte for byte. The first call
 * (or the first after vrestir_set_volume / vrestir_load_vbx) starts a new volume
 * with desc's material; each later call builds the next frame of an animation into
 * a ring of two frames and binds it with the frame before it as the previous frame
 * (as vrestir_set_volume_frame(h, i, 1) does for a resident sequence). Later frames
 * keep the first frame's material and need its dimensions, mips, format and grids.
 * Synchronises the device; temporary device memory is about three dense grids. */
int32_t vrestir_stream_volume_frame(vrestir_handle* h, void* stream, const vrestir_volume_desc* desc);
/* Binds owner's volume into h read-only (no counterpart in the reference, where the
 * passes of a frame share the Scene): the current and previous frame's grids, the
 * mips, the material and the animation state; h keeps its own lights, environment,
 * camera and options. A copy of the binding at the call: call it again after the
 * owner's volume changes (each vrestir_stream_volume_frame / vrestir_set_volume_frame).
 * The owner must outlive h's use of the volume. */
int32_t vrestir_share_volume(vrestir_handle* h, const vrestir_handle* owner);
int32_t vrestir_load_vbx(vrestir_handle* h, const char* pathPrefix, int32_t numMips, int32_t brickFormat,
                         const float sigmaA[3], const float sigmaS[3], float g, float densityScale);
int32_t vrestir_set_lights(vrestir_handle* h, const vrestir_point_light* lights, int32_t count,
                           const float envRadiance[3]);
int32_t vrestir_set_camera(vrestir_handle* h, const vrestir_camera* cam);
/* The scene's triangles and materials (Scene::getLightCollection and the emissive
 * sampler's update, VR/VolumetricReSTIR.cpp:157-207); NULL removes them. With
 * useEmissiveLights their emissive triangles are lights of every stage, as in a
 * volume scene of the reference (no surface intersections). */
int32_t vrestir_set_surfaces(vrestir_handle* h, const vrestir_surfaces_desc* desc);
/* Environment map in place of the constant envRadiance of vrestir_set_lights; NULL
 * returns to the constant. The importance map is built on the device. */
int32_t vrestir_set_env_map(vrestir_handle* h, const vrestir_env_map_desc* desc);
/* EnvMap::setRotation: the orientation from the next frame on; temporal reuse
 * evaluates the previous frame's samples under the previous frame's orientation
 * (EnvMap prevData, SC/Scene.cpp:761-775). Also the per-frame env animation of
 * VR/VolumetricReSTIR.cpp:788-799. */
int32_t vrestir_set_env_rotation(vrestir_handle* h, const float degreesXYZ[3]);
int32_t vrestir_set_env_intensity(vrestir_handle* h, float intensity, const float tint[3]); /* setIntensity/setTint */
/* Blackbody table for volume emission (SC/Scene.cpp:2877-2893 loads
 * LUT/BlackBodyRadiationRGB_50K-6400K.txt, 128 RGB rows); rows in [1, 256]. */
int32_t vrestir_set_blackbody_lut(vrestir_handle* h, const float* rgb, int32_t rows);
/* Temperature grid <pathPrefix>_temperature.vbx (own transform) for the bound
 * volume (addGVDBVolume hasEmission=True, SC/Scene.cpp:2789-2811). */
int32_t vrestir_load_vbx_emission(vrestir_handle* h, const char* pathPrefix, float LeScale, float temperatureCutOff,
                                  float temperatureScale);
/* Animated volumes (addGVDBVolumeSequence, SC/Scene.cpp:3329-3355 and the per-frame
 * update of SC/Scene.cpp:825-863). vrestir_set_volume binds frame 0 of a new
 * sequence; each vrestir_add_volume_frame keeps one more frame resident in HBM
 * (material parameters stay the first frame's). vrestir_set_volume_frame makes
 * frame `index` current with the frame before it (in sequence order, the first
 * frame's previous is the last) in the previous-frame slots; animating != 0 is
 * the reference's hasAnimation: temporal reuse then reprojects with the velocity
 * grid and, with usePrevVolumeForReproj, evaluates the previous frame's samples
 * on the previous frame's grids. */
int32_t vrestir_add_volume_frame(vrestir_handle* h, const vrestir_volume_desc* desc, int32_t* index);
int32_t vrestir_set_volume_frame(vrestir_handle* h, int32_t index, int32_t animating);
int32_t vrestir_volume_frame_count(const vrestir_handle* h, int32_t* count);
/* mFreezeFrame (VR/VolumetricReSTIR.cpp:499-780): while set, execute only runs final
 * shading on the last reservoirs and does not advance the frame count. */
int32_t vrestir_set_freeze_frame(vrestir_handle* h, int32_t freeze);

/* ---- options: updateDict / getScriptingDictionary (VR/VolumetricReSTIR.cpp:1280-1347).
 *      Any change resets the temporal history and the frame counter, like mOptionsChanged. ---- */
int32_t vrestir_set_params(vrestir_handle* h, const vrestir_params* params);
int32_t vrestir_get_params(const vrestir_handle* h, vrestir_params* out);
int32_t vrestir_reset_history(vrestir_handle* h);

/* ---- per frame: VolumetricReSTIR::execute (VR/VolumetricReSTIR.cpp:303-906).
 * Enqueues the whole frame on `stream` (a hipStream_t, NULL = default stream).
 * outRGBA: device float4[width*height] ("accumulated_color", VR/VolumetricReSTIR.cpp:39-43);
 * outMotionVec: device float2[width*height] or NULL ("mvec"). ---- */
int32_t vrestir_execute(vrestir_handle* h, void* stream, int32_t width, int32_t height,
                        float* outRGBA, float* outMotionVec);
int32_t vrestir_frame_count(const vrestir_handle* h, int32_t* out);

/* ---- stage-level access, the six dispatches of execute (VR/VolumetricReSTIR.cpp:499-762).
 * Used by parity tests to run one kernel on injected inputs. ---- */
enum {
    VRESTIR_STAGE_GENERATE_FEATURES = 0,  /* VR/GenerateFeatures.cs.slang:57-102 */
    VRESTIR_STAGE_TRACE_RAYS = 1,         /* VR/TraceRays.cs.slang:64-201 */
    VRESTIR_STAGE_TEMPORAL_REUSE = 2,     /* VR/TemporalReuse.cs.slang:80-377 */
    VRESTIR_STAGE_SPATIAL_REUSE = 3,      /* VR/SpatialReuse.cs.slang:94-265 (round = argument) */
    VRESTIR_STAGE_COPY_RESERVOIRS = 4,    /* VR/CopyReservoirs.cs.slang:84-95 */
    VRESTIR_STAGE_FINAL_SHADING = 5,      /* VR/FinalShading.cs.slang:71-141 */
    /* the learned proposal's two dispatches (no counterpart in the reference; see
     * vrestir_learned_desc): the network's CDF for every camera ray, run before
     * TraceRays, and the Adam steps on TraceRays' training rays, run after it */
    VRESTIR_STAGE_LEARNED_PROPOSAL = 6,
    VRESTIR_STAGE_LEARNED_TRAIN = 7
};
/* Buffers in the reference's own identities (VR/VolumetricReSTIR.cpp:419-450).
 * Host exchange format per pixel: features {int32 noReflectiveSurface; float transmittance};
 * reservoir {float runningSum, M, depth, p_y, lightUV[2]; int32 lightID, sampledPixel, extraBounceStartId};
 * extra bounce {float wi_dist[3]} x (maxBounces-1); color float4. */
enum {
    VRESTIR_BUF_FEATURES = 0,
    VRESTIR_BUF_TEMPORAL_FEATURES = 1,
    VRESTIR_BUF_RESERVOIRS0 = 2,
    VRESTIR_BUF_RESERVOIRS1 = 3,
    VRESTIR_BUF_TEMPORAL_RESERVOIRS = 4,
    VRESTIR_BUF_EXTRA0 = 5,
    VRESTIR_BUF_EXTRA1 = 6,
    VRESTIR_BUF_TEMPORAL_EXTRA = 7,
    VRESTIR_BUF_COLOR = 8,
    /* learned proposal: uint16 cdf[32] per pixel (cdf[0] = 0, cdf[32] = 65536 implicit) */
    VRESTIR_BUF_LEARNED_CDF = 9,
    /* learned proposal: float[72] per training ray (VRESTIR_LEARNED_RECORD_FLOATS): raw inputs
     * (entry point in mip 0's normalised box, direction, segment length / box diagonal),
     * the summed weight W, then per bin the summed weights and summed weight^2 * g_used */
    VRESTIR_BUF_TRAIN_RECORDS = 10,
    /* split TraceRays (VRESTIR_TUNE_SPLIT_TRACE): int32 count of the last frame's pixels
     * that were traced again in one kernel (resampling draws that differed from the record's) */
    VRESTIR_BUF_TRACE_REDONE = 11,
    /* surface scenes: the V-buffer of GenerateFeatures, per pixel {int32 triangle (-1: no
     * hit), float barycentric weights of vertices 1 and 2, float ray distance (FLT_MAX)} */
    VRESTIR_BUF_VBUFFER = 12
};
int32_t vrestir_prepare_frame(vrestir_handle* h, int32_t width, int32_t height);
int32_t vrestir_run_stage(vrestir_handle* h, void* stream, int32_t stage, int32_t round, float* outRGBA);
int32_t vrestir_end_frame(vrestir_handle* h);
int32_t vrestir_read_buffer(vrestir_handle* h, int32_t buffer, void* hostDst, size_t bytes);
int32_t vrestir_write_buffer(vrestir_handle* h, int32_t buffer, const void* hostSrc, size_t bytes);

/* ---- batched tracking primitive (the adapters of VR/VolumeTrackingAdapterGVDB.slang run
 * over a ray batch; used by parity tests and the tracking micro-benchmark).
 * rays: device float[8*n] = {origin.xyz, dir.xyz, tMax, unused}; seeds: device uint32[3*n] =
 * {pixel.x, pixel.y, sampleNumber}; out: device float[8*n] (layout per mode, see DESIGN). ---- */
enum {
    VRESTIR_TRACK_TR_ANALYTIC = 0,
    VRESTIR_TRACK_TR_RAY_MARCHING = 1,
    VRESTIR_TRACK_TR_RATIO = 2,
    VRESTIR_TRACK_TR_RESIDUAL_RATIO = 3,
    VRESTIR_TRACK_TR_ANALOG_RESIDUAL_RATIO = 4,
    VRESTIR_TRACK_SAMPLE_ANALYTIC = 5,     /* up to 4 free-flight samples in one traversal */
    VRESTIR_TRACK_SAMPLE_DECOMPOSITION = 6,/* delta tracking against per-brick bounds */
    VRESTIR_TRACK_SAMPLE_CELL_BY_DENSITY = 7,
    VRESTIR_TRACK_FEATURE_MARCH = 8,
    VRESTIR_TRACK_DENSITY_AT_POINT = 9,
    VRESTIR_TRACK_RNG_STREAM = 10,         /* first 8 uint32 of the sample generator (bit-cast to float) */
    VRESTIR_TRACK_ENV_EVAL = 11,           /* environment radiance towards dir: out rgb */
    VRESTIR_TRACK_ENV_SAMPLE = 12,         /* importance sample at u = origin.xy: out dir, pdf, Le, lightID */
    VRESTIR_TRACK_EMISSION = 13,           /* emission at the point origin: out rgb, temperature */
    VRESTIR_TRACK_VELOCITY = 14,           /* velocity grid at the point origin: out xyz */
    /* Brick and voxel indexing at the point origin (GVS/gvdbNodes.slang:261-279,
     * GVS/gvdb.slang:6-29): out = {point-sampled density, trilinear density, the
     * leaf brick's index-space minimum xyz (-1 when no brick holds the point),
     * the voxel floor(index-space point) xyz}. */
    VRESTIR_TRACK_DENSITY_INDEX = 15,
    /* The hierarchical DDA's visit sequence (VR/VolumeUtils.slang:171-282) over the
     * grid slot mipLevel, vertex-centred when useLinearSampler: out = {bricks
     * visited, FNV-1a hash of their index-space minima in visit order (uint32
     * bits), sum of entry distances, first entry, last entry, tNear, tFar, 0}. */
    VRESTIR_TRACK_VISITED_BRICKS = 16,
    /* sampleSceneLights (VR/VolumeUtils.slang:12-149) at the point origin, the light
     * kinds from numSamples' bits (1 environment, 2 point lights, 4 emissive
     * triangles): out = {Li rgb, pdfArea, light index, light uv, ray distance}; all
     * zero with index -9 when no light was sampled. */
    VRESTIR_TRACK_LIGHT_SAMPLE = 17,
    /* The emissive-triangle branch of evaluate_L_in_volume (VR/ReSTIRHelper.slang:
     * 467-476) at the point origin, collection triangle dir.x, light uv dir.yz:
     * out = {Le * cos / dist^2 rgb (no phase, no multiplier), dist, dir xyz, valid}. */
    VRESTIR_TRACK_EMISSIVE_EVAL = 18,
    /* FindSurfaceHit and FindIfOccluded (VR/InlineRayTracingHelpers.slang:42-155) over
     * [0, tMax]: out = {closest t (FLT_MAX: none), its triangle, barycentric weights of
     * vertices 1 and 2, occluded (any hit), 0, 0, 0}. */
    VRESTIR_TRACK_SURFACE_HIT = 19,
    /* Shading at the closest hit over [0, tMax] (computeSurfaceShadingInfo and the
     * legacy standard BSDF, MaterialShading.slang:167-631): numSamples 0 gives
     * sampleBSDF's {wi xyz, pdf, weight rgb, lobe (1 diffuse, 2 specular; -1 no
     * sample)}, numSamples 1 the evaluation at that wi {evalBSDFCosine rgb,
     * evalPdfBSDF, emissive rgb, NdotV}; all zero without a hit. */
    VRESTIR_TRACK_SURFACE_SHADE = 20
};
int32_t vrestir_track_batch(vrestir_handle* h, void* stream, int32_t mode, int32_t mipLevel,
                            int32_t useLinearSampler, float tStepScale, int32_t numSamples, int32_t n,
                            const float* raysDev, const uint32_t* seedsDev, float* outDev);

/* ---- learned free-flight proposal (an addition of this project: no counterpart
 * in the reference). Per camera ray a 36 -> 64 -> 64 -> 64 -> 32 ReLU network on a
 * frequency encoding of (entry point, direction, segment length) predicts a 32-bin
 * pdf g over the analytic tracker's CDF coordinate c = 1 - exp(-tau). The initial
 * candidates' distances are the tracker's, driven by optical-depth targets
 * -log(1 - c) with c drawn from g, so their source pdf -- pdfDist of
 * VR/ComputeInitialSample.slang:86,107 -- becomes g(c) sigma_t Tr (an escape: the
 * mass of g above 1 - Tr), evaluated exactly from the same 16-bit CDF it samples.
 * g mixes the network's softmax with the uniform pdf (defensiveFraction), so the
 * analytic proposal stays in every frame; with judge on, a frame samples with the
 * network only while the network's estimated second moment of the RIS weights
 * (importance-reweighted from the training rays) is below the analytic one's.
 * Training is online: one ray per trainBlockX x trainBlockY block records its
 * candidates' weights, and hand-written BF16 MFMA kernels (FP32 accumulation)
 * take Adam steps on the cross-entropy of g against those weights. ---- */
enum { VRESTIR_LEARNED_PARAMS = 12768, VRESTIR_LEARNED_RECORD_FLOATS = 72, VRESTIR_LEARNED_RAW_FLOATS = 8 };
typedef struct vrestir_learned_desc {
    int32_t enable;              /* 0: the reference's analytic proposal only */
    int32_t train;               /* online training on the pass's own candidates */
    int32_t trainBlockX, trainBlockY; /* one training ray per block of pixels (16 x 8) */
    int32_t trainStepsPerFrame;  /* Adam steps per frame, the frame's training rays split evenly (1) */
    float learningRate;          /* Adam step size (2e-3) */
    float defensiveFraction;     /* weight of the uniform pdf in g, in [1/64, 1] (0.25) */
    int32_t judge;               /* 1: the analytic proposal decides whether the network samples (1) */
    uint64_t seed;               /* initial weights; the output layer starts at zero (g = 1) */
} vrestir_learned_desc;
void vrestir_default_learned(vrestir_learned_desc* out);
/* Sets the options and (re)initialises the network from seed, the Adam state and the judge. */
int32_t vrestir_set_learned(vrestir_handle* h, const vrestir_learned_desc* desc);
int32_t vrestir_get_learned(const vrestir_handle* h, vrestir_learned_desc* out);
/* After the stream drains: [0] loss of the last Adam step, [1] the analytic proposal's and
 * [2] the network's second moment of the per-ray normalised RIS weights over the last
 * step's rays, [3] 1 if the network samples the next frame, [4] Adam steps taken. */
int32_t vrestir_learned_stats(vrestir_handle* h, float stats[5]);
/* FP32 master weights (VRESTIR_LEARNED_PARAMS floats: W1[36][64], b1, W2[64][64], b2,
 * W3[64][64], b3, W4[64][32], b4). Setting them resets the Adam state. */
int32_t vrestir_learned_get_weights(vrestir_handle* h, float* out, int32_t n);
int32_t vrestir_learned_set_weights(vrestir_handle* h, const float* in, int32_t n);
/* The network on n raw inputs (device float[8 n]): the 16-bit CDF (device uint16[32 n])
 * and optionally the logits (device float[32 n] or NULL). Ignores the judge. */
int32_t vrestir_learned_forward(vrestir_handle* h, void* stream, const float* rawDev, int32_t n, uint16_t* cdfDev,
                                float* logitsDev);
/* Sum over n training records (device float[72 n]) of scale x the cross-entropy gradient:
 * device float[VRESTIR_LEARNED_PARAMS + 3] = gradient, scale x loss, and the two second moments. */
int32_t vrestir_learned_gradient(vrestir_handle* h, void* stream, const float* recsDev, int32_t n, float scale,
                                 float* outDev);
/* One Adam step on the mean cross-entropy of n training records (device float[72 n]). */
int32_t vrestir_learned_train(vrestir_handle* h, void* stream, const float* recsDev, int32_t n);

/* ---- implementation choices (no counterpart in the reference) ----
 * VRESTIR_TUNE_SPLIT_REUSE (default 2): 1 and 2 run spatial reuse as three
 * launches -- record its transmittance rays, trace them from per-wave balanced
 * queues, replay the reuse with their values -- when its ray kinds use ray
 * marching; 1 traces each ray in one lane, 2 walks rays and marches their
 * brick segments in separate phases. Mode 2 also splits temporal reuse (same
 * ray kinds, not on an animated volume's previous-frame grids) and final
 * shading when its ray kinds use analytic tracking on a mip-0 grid of at most
 * 2048 128^3 blocks (the segments are integrated cell by cell); otherwise they
 * run as one kernel.
 * 0 runs each stage as one kernel. Results agree to float rounding (2 sums each
 * ray's optical depth per brick).
 * VRESTIR_TUNE_SPLIT_TRACE (default 1; with VRESTIR_TUNE_SPLIT_REUSE 2): TraceRays
 * records its candidates' shadow rays (keeping their free-flight samples), traces
 * them as brick segments and replays the candidates with their transmittances,
 * when the initial lighting rays use ray marching and the learned proposal is off.
 * A pixel whose resampling draws differ from the record's (a candidate weight that
 * its real shadow transmittance makes zero) is traced again in one kernel, so the
 * results agree with one kernel's to float rounding. 0 runs TraceRays as one kernel. */
enum { VRESTIR_TUNE_SPLIT_REUSE = 1, VRESTIR_TUNE_SPLIT_TRACE = 2 };
int32_t vrestir_set_tuning(vrestir_handle* h, int32_t key, int32_t value);

/* ---- ground truth and error (AccumulatePass double mode, AccumulatePass.cpp:42,93;
 *      ImageCompare "rmse", Source/Tools/ImageCompare/ImageCompare.cpp:172-180) ---- */
/* Adds one RGBA32F frame (device memory, width*height pixels) to the handle's FP64
 * running sum; reset != 0 (or a new size) starts a new sum. */
int32_t vrestir_accumulate(vrestir_handle* h, void* stream, const float* frameRGBA, int32_t width, int32_t height,
                           int32_t reset);
/* Writes the mean of the accumulated frames (RGBA32F, device memory); frames may be NULL. */
int32_t vrestir_accumulated(vrestir_handle* h, void* stream, float* outRGBA, int64_t* frames);
/* relMSE = mean over pixels of mean over RGB of (ref - img)^2 / (ref^2 + 1e-3), in FP64;
 * both images RGBA32F in device memory; the result is written to *out (host) after the stream drains. */
int32_t vrestir_relmse(vrestir_handle* h, void* stream, const float* imageRGBA, const float* referenceRGBA,
                       int32_t width, int32_t height, double* out);
/* vrestir_relmse that allocates nothing (no hipMalloc / hipFree): no counterpart in
 * the reference. For callers that cannot allocate (code under HIP graph capture,
 * tight serving loops, harnesses that forbid it). Its partial sums and result use
 * a static 256 KiB device array per GPU, held until the process ends, so images are
 * limited to VRESTIR_RELMSE_NOALLOC_MAX_PIXELS pixels (larger: VRESTIR_EINVAL).
 * Calls that overlap on one GPU share that array and race. Same result as
 * vrestir_relmse, bit for bit. */
#define VRESTIR_RELMSE_NOALLOC_MAX_PIXELS 8388352
int32_t vrestir_relmse_noalloc(vrestir_handle* h, void* stream, const float* imageRGBA, const float* referenceRGBA,
                               int32_t width, int32_t height, double* out);

/* The volume path tracer of useReference (VR/VolumePathTracingFunctions.slang:3-132,
 * VR/TraceRays.cs.slang:85-103) on a uniform pixel lattice: lattice point (i, j) is
 * pixel (stride i + stride/2, stride j + stride/2) of the width x height frame, seen
 * through the bound camera, lw = ceil((width - stride/2) / stride) points per row,
 * lh likewise. Image c of outRGBA (device float4[chunks*lh*lw]) averages spp samples
 * from each pixel's stream of sample number sampleNumber + c. A frame of this size
 * must have been prepared (the path tracer's options are that frame's). Summed over
 * calls with distinct sample numbers (vrestir_accumulate), it gives a converged
 * reference on the lattice, where relMSE is an unbiased estimate of the full
 * frame's. No counterpart in the reference beyond its full-frame reference mode. */
int32_t vrestir_reference_lattice(vrestir_handle* h, void* stream, int32_t width, int32_t height, int32_t stride,
                                  int32_t spp, uint32_t sampleNumber, int32_t chunks, float* outRGBA);

/* Size of bound grid slot `slot` (mip k: k, conservative mip k: 8 + k): voxel
 * dimensions, occupied 8^3 bricks, brick cells of its table, and the device bytes
 * of its brick pool (1000 values per brick with the apron), tables and bounds. */
int32_t vrestir_grid_info(const vrestir_handle* h, int32_t slot, int32_t dims[3], int64_t* numBricks,
                          int64_t* brickCells, int64_t* deviceBytes);

/* ---- batches: many passes in one persistent launch per step (no counterpart in the
 * reference, which renders one view per pass). Each step of a frame -- a stage, or
 * one record / trace / replay pass of a split stage -- runs as ONE persistent
 * launch whose workgroups take 16x16-pixel tiles (256-pixel ranges for the
 * traces) of every pass from a work queue in device memory; the frame's launches
 * are captured into a HIP graph and replayed. The passes keep their own volumes,
 * lights, cameras, resolutions and histories, and give the same images as when
 * rendered one by one; they must share path length and spatial rounds and run
 * with temporal and spatial reuse, the split stage forms (VRESTIR_TUNE_SPLIT_REUSE
 * 2), ray-marching spatial and initial-lighting trackers and analytic final
 * shading (the defaults), without the learned proposal or reference mode
 * (otherwise VRESTIR_EUNSUPPORTED). At most 64 passes, all on one GPU. ---- */
typedef struct vrestir_batch vrestir_batch;
int32_t vrestir_batch_create(vrestir_handle* const* passes, int32_t count, vrestir_batch** out);
int32_t vrestir_batch_destroy(vrestir_batch* b);
const char* vrestir_batch_last_error(const vrestir_batch* b);
/* One frame of every pass: widths/heights/outRGBA per pass (outRGBA[i]: device
 * float4[widths[i]*heights[i]]). Work is ordered after what `stream` holds and
 * `stream` waits for the frame; the batch's own stream runs the graph. */
int32_t vrestir_batch_execute(vrestir_batch* b, void* stream, const int32_t* widths, const int32_t* heights,
                              float* const* outRGBA);
/* Queue shape: persistent workgroups per launch and 16x16 tiles per queue item;
 * useGraph 0 issues the launches directly. */
int32_t vrestir_batch_set_tuning(vrestir_batch* b, int32_t grid, int32_t tilesPerItem, int32_t useGraph);
int32_t vrestir_batch_get_tuning(const vrestir_batch* b, int32_t* grid, int32_t* tilesPerItem);
/* The autotuner: renders `frames` frames per candidate queue shape (2, 4 and 8
 * workgroups per CU x 1, 2 and 4 tiles per item), keeps the fastest, then resets
 * every pass's history (vrestir_reset_history); *bestMs (optional) receives its
 * time per frame. */
int32_t vrestir_batch_autotune(vrestir_batch* b, void* stream, const int32_t* widths, const int32_t* heights,
                               float* const* outRGBA, int32_t frames, float* bestMs);

/* ---- the synthetic animated plume (GPU; no counterpart in the reference, whose
 * animated volumes are files) ----
 * Smoke density and temperature on an n^3 grid over the unit cube, rising with a
 * buoyant updraft and stirred by seeded curl noise: one semi-Lagrangian step per
 * vrestir_plume_advance step, a noisy hot source at the base. Deterministic for
 * a seed. The grids (device memory, x fastest) feed vrestir_stream_volume_frame
 * and stay valid until the next advance. n: a multiple of 8 in [16, 2048]; holds
 * four n^3 float grids. */
typedef struct vrestir_plume vrestir_plume;
int32_t vrestir_plume_create(int32_t n, uint64_t seed, vrestir_plume** out);
int32_t vrestir_plume_destroy(vrestir_plume* p);
int32_t vrestir_plume_advance(vrestir_plume* p, void* stream, int32_t steps);
int32_t vrestir_plume_grids(const vrestir_plume* p, const float** density, const float** temperature,
                            int32_t* steps);
/* Copies the current grids to host memory (n^3 floats each; either may be NULL). */
int32_t vrestir_plume_read(const vrestir_plume* p, float* density, float* temperature);

/* ---- host-side helpers (no GPU needed) ---- */
/* Seeded fBm density (4-octave Perlin by default) for the synthetic workloads. */
int32_t vrestir_synth_fbm(uint64_t seed, int32_t nx, int32_t ny, int32_t nz, int32_t octaves,
                          float baseFrequency, float lacunarity, float gain, int32_t sphereMask,
                          float threshold, float amplitude, float* out);
/* Camera orbit about a centre (camera animation mode 2, VR/VolumetricReSTIR.cpp:879-904). */
int32_t vrestir_orbit_camera(const vrestir_camera* base, const float center[3], int32_t frame,
                             float degreesPerFrame, vrestir_camera* out);
/* Camera animation (VR/VolumetricReSTIR.cpp:801-905, fields as the pass's mCamera* members,
 * defaults in VR/VolumetricReSTIR.h:228-240): the camera and freeze state after `updates`
 * end-of-frame updates from the toggle. mode 0: sideways shake with pauses; 1: shake,
 * forward moves and pauses; 2: rotation about the point of the view ray closest to
 * volumeCenter. A frozen state maps to vrestir_set_freeze_frame(1). */
typedef struct vrestir_camera_animation_desc {
    int32_t mode;                  /* 0, 1, 2 (mCameraAnimationMode) */
    int32_t frameInterval;         /* 50 */
    float moveScale;               /* 0.3 */
    float forwardScale;            /* 1 */
    int32_t pauseInterval;         /* 150 */
    int32_t shakeRoundsBeforePause;/* 3 */
    int32_t shakeTotalRounds;      /* 1 */
    int32_t rotationFrames;        /* 360 */
    float rotationSpeed;           /* 1 degree per frame */
    float rotationDistance;        /* <= 0: the initial distance */
    float volumeCenter[3];         /* Scene::getSceneVolumeCenter */
} vrestir_camera_animation_desc;
int32_t vrestir_camera_animation(const vrestir_camera* base, const vrestir_camera_animation_desc* anim, int32_t updates,
                                 vrestir_camera* out, int32_t* freeze);
/* Brick statistics of a dense grid: number of 8^3 bricks with any non-zero voxel. */
int32_t vrestir_count_bricks(const float* density, int32_t nx, int32_t ny, int32_t nz, int64_t* out);
/* Write / read the .vbx container (gvdb-voxel-src/GVDB_FILESPEC.txt, minor version 12). */
int32_t vrestir_write_vbx(const char* path, const float* density, int32_t nx, int32_t ny, int32_t nz,
                          const float indexToWorld[16]);
int32_t vrestir_read_vbx(const char* path, float* densityOrNull, int32_t* nx, int32_t* ny, int32_t* nz,
                         float indexToWorld[16]);

#ifdef __cplusplus
}
#endif
#endif /* VRESTIR_H */
