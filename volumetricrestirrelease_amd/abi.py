"""ctypes mirror of include/vrestir.h and the loader of libvrestir.so.

The struct layouts below are field-for-field copies of the C declarations;
``tests/test_abi.py`` checks their sizes against the compiled library.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_char_p, c_float, c_int32, c_int64, c_size_t, c_uint32, c_uint64, c_void_p

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvrestir.so")

# status codes
VRESTIR_OK = 0
VRESTIR_EINVAL = 1
VRESTIR_EIO = 2
VRESTIR_EHIP = 3
VRESTIR_ENOMEM = 4
VRESTIR_ESTATE = 5
VRESTIR_EUNSUPPORTED = 6
# largest image vrestir_relmse_noalloc accepts (its static scratch: 32767 partial sums of 256 pixels)
VRESTIR_RELMSE_NOALLOC_MAX_PIXELS = 8388352

# tracking methods / MIS / sampler / reprojection (VR/HostDeviceSharedConstants.slang:5-15, 30-32)
RATIO_TRACKING, ANALYTIC_TRACKING, RAY_MARCHING, RESIDUAL_RATIO_TRACKING, ANALOG_RESIDUAL_RATIO_TRACKING = range(5)
MIS_NONE, MIS_TALBOT = 0, 1
SAMPLER_HAMMERSLEY, SAMPLER_R2 = 0, 1
REPROJECTION_LINEAR, REPROJECTION_NONE, REPROJECTION_NO_BACKGROUND = 0, 1, 2
BRICK_FP32, BRICK_FP16 = 0, 1

TUNE_SPLIT_REUSE = 1
TUNE_SPLIT_TRACE = 2

STAGE_GENERATE_FEATURES, STAGE_TRACE_RAYS, STAGE_TEMPORAL_REUSE, STAGE_SPATIAL_REUSE, \
    STAGE_COPY_RESERVOIRS, STAGE_FINAL_SHADING = range(6)
STAGE_LEARNED_PROPOSAL, STAGE_LEARNED_TRAIN = 6, 7
BUF_FEATURES, BUF_TEMPORAL_FEATURES, BUF_RESERVOIRS0, BUF_RESERVOIRS1, BUF_TEMPORAL_RESERVOIRS, \
    BUF_EXTRA0, BUF_EXTRA1, BUF_TEMPORAL_EXTRA, BUF_COLOR, BUF_LEARNED_CDF, BUF_TRAIN_RECORDS, \
    BUF_TRACE_REDONE, BUF_VBUFFER = range(13)

TRACK_TR_ANALYTIC, TRACK_TR_RAY_MARCHING, TRACK_TR_RATIO, TRACK_TR_RESIDUAL_RATIO, \
    TRACK_TR_ANALOG_RESIDUAL_RATIO, TRACK_SAMPLE_ANALYTIC, TRACK_SAMPLE_DECOMPOSITION, \
    TRACK_SAMPLE_CELL_BY_DENSITY, TRACK_FEATURE_MARCH, TRACK_DENSITY_AT_POINT, TRACK_RNG_STREAM, \
    TRACK_ENV_EVAL, TRACK_ENV_SAMPLE, TRACK_EMISSION, TRACK_VELOCITY, TRACK_DENSITY_INDEX, \
    TRACK_VISITED_BRICKS, TRACK_LIGHT_SAMPLE, TRACK_EMISSIVE_EVAL, TRACK_SURFACE_HIT, \
    TRACK_SURFACE_SHADE = range(21)

# (name, ctype) in declaration order of vrestir_params
PARAM_FIELDS = [
    ("maxBounces", c_int32), ("enableTemporalReuse", c_int32), ("enableSpatialReuse", c_int32),
    ("vertexReuse", c_int32), ("vertexReuseStartBounce", c_int32), ("useReference", c_int32),
    ("useEnvironmentLights", c_int32), ("useAnalyticLights", c_int32), ("useEmissiveLights", c_int32),
    ("baselineSamplePerPixel", c_int32), ("visualizeTotalTransmittance", c_int32), ("useSurfaceScene", c_int32),
    ("usePrevVolumeForReproj", c_int32),
    ("initialBaseMipLevel", c_int32), ("initialM", c_int32), ("initialLightSamples", c_int32),
    ("initialLightingMipLevel", c_int32), ("initialVisibilityUseLinearSampler", c_int32),
    ("initialLightingUseLinearSampler", c_int32), ("initialLightingTrackingMethod", c_uint32),
    ("initialVisibilityTStepScale", c_float), ("initialLightingTStepScale", c_float),
    ("initialUseRussianRoulette", c_int32), ("initialUseCoarserGridForIndirectBounce", c_int32),
    ("temporalReuseMThreshold", c_float), ("temporalReprojectionMode", c_uint32),
    ("temporalMISMethod", c_uint32), ("temporalReprojectionMipLevel", c_int32),
    ("spatialReuseRounds", c_int32), ("spatialVisibilityMipLevel", c_int32), ("spatialLightingMipLevel", c_int32),
    ("spatialVisibilityUseLinearSampler", c_int32), ("spatialLightingUseLinearSampler", c_int32),
    ("spatialVisibilityTStepScale", c_float), ("spatialLightingTStepScale", c_float),
    ("spatialVisibilityTrackingMethod", c_uint32), ("spatialLightingTrackingMethod", c_uint32),
    ("randomSamplerType", c_uint32), ("sampleRadius", c_float), ("spatialSampleCount", c_int32),
    ("enableVisibilitySimilarityRejection", c_int32), ("spatialMISMethod", c_uint32),
    ("finalLightSamples", c_int32), ("finalVisibilitySamples", c_int32),
    ("finalVisibilityTrackingMethod", c_uint32), ("finalLightTrackingMethod", c_uint32),
    ("finalRandomSamplerType", c_uint32), ("finalTStepScale", c_float),
]


class Params(ctypes.Structure):
    _fields_ = PARAM_FIELDS


class VolumeDesc(ctypes.Structure):
    _fields_ = [
        ("nx", c_int32), ("ny", c_int32), ("nz", c_int32),
        ("density", POINTER(c_float)),
        ("numMips", c_int32), ("brickFormat", c_int32),
        ("indexToWorld", c_float * 16),
        ("sigmaA", c_float * 3), ("sigmaS", c_float * 3),
        ("g", c_float), ("densityScale", c_float),
        ("temperature", POINTER(c_float)),
        ("LeScale", c_float), ("temperatureCutOff", c_float), ("temperatureScale", c_float),
        ("velocity", POINTER(c_float)), ("velocityScale", c_float),
    ]


class EnvMapDesc(ctypes.Structure):
    _fields_ = [
        ("width", c_int32), ("height", c_int32),
        ("rgb", POINTER(c_float)),
        ("intensity", c_float), ("tint", c_float * 3), ("rotationDegrees", c_float * 3),
        ("importanceDim", c_int32), ("importanceSpp", c_int32),
    ]


class CameraAnimationDesc(ctypes.Structure):
    _fields_ = [
        ("mode", c_int32), ("frameInterval", c_int32), ("moveScale", c_float), ("forwardScale", c_float),
        ("pauseInterval", c_int32), ("shakeRoundsBeforePause", c_int32), ("shakeTotalRounds", c_int32),
        ("rotationFrames", c_int32), ("rotationSpeed", c_float), ("rotationDistance", c_float),
        ("volumeCenter", c_float * 3),
    ]


def camera_animation_desc(mode=0, **kw) -> CameraAnimationDesc:
    """The pass's camera-animation members with their defaults (VR/VolumetricReSTIR.h:228-240)."""
    d = CameraAnimationDesc(mode=mode, frameInterval=50, moveScale=0.3, forwardScale=1.0, pauseInterval=150,
                            shakeRoundsBeforePause=3, shakeTotalRounds=1, rotationFrames=360, rotationSpeed=1.0,
                            rotationDistance=0.0)
    for k, v in kw.items():
        if k == "volumeCenter":
            for i in range(3):
                d.volumeCenter[i] = v[i]
        else:
            setattr(d, k, v)
    return d


class PointLight(ctypes.Structure):
    _fields_ = [("posW", c_float * 3), ("intensity", c_float * 3)]


class Material(ctypes.Structure):
    """vrestir_material: untextured standard material (metal-rough)."""
    _fields_ = [("baseColor", c_float * 3), ("roughness", c_float), ("metallic", c_float), ("IoR", c_float),
                ("emissive", c_float * 3), ("doubleSided", ctypes.c_int32)]


class SurfacesDesc(ctypes.Structure):
    """vrestir_surfaces_desc: world-space triangles and their materials."""
    _fields_ = [("positions", POINTER(c_float)), ("normals", POINTER(c_float)), ("numVertices", ctypes.c_int32),
                ("indices", POINTER(ctypes.c_int32)), ("materialIds", POINTER(ctypes.c_int32)),
                ("numTriangles", ctypes.c_int32), ("materials", POINTER(Material)),
                ("numMaterials", ctypes.c_int32), ("emissiveIntensityMultiplier", c_float)]


class Camera(ctypes.Structure):
    _fields_ = [("posW", c_float * 3), ("target", c_float * 3), ("up", c_float * 3),
                ("fovY", c_float), ("nearZ", c_float), ("farZ", c_float)]


class LearnedDesc(ctypes.Structure):
    """vrestir_learned_desc: the learned free-flight proposal's options."""
    _fields_ = [("enable", ctypes.c_int32), ("train", ctypes.c_int32), ("trainBlockX", ctypes.c_int32),
                ("trainBlockY", ctypes.c_int32), ("trainStepsPerFrame", ctypes.c_int32),
                ("learningRate", ctypes.c_float), ("defensiveFraction", ctypes.c_float), ("judge", ctypes.c_int32),
                ("seed", ctypes.c_uint64)]


LEARNED_PARAMS, LEARNED_RECORD_FLOATS, LEARNED_RAW_FLOATS = 12768, 72, 8


def default_learned() -> "LearnedDesc":
    d = LearnedDesc()
    load_library().vrestir_default_learned(ctypes.byref(d))
    return d


class LibraryMissing(RuntimeError):
    pass


_lib = None


def load_library(path: str = LIB_PATH) -> ctypes.CDLL:
    """Load libvrestir.so (built by __graft_entry__.build()); raise if it is absent.

    There is no fallback implementation: without the library the pass cannot run.
    """
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise LibraryMissing(f"{path} not built; run __graft_entry__.build() (make -C volumetricrestirrelease_amd/csrc)")
    # libvrestir.so binds to whichever libamdhip64.so.7 the process has loaded. PyTorch
    # ships its own copy; load torch first so the library and torch share one HIP runtime
    # (loading the system runtime first leaves torch unable to initialise its devices).
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.CDLL(path)
    vp, i32, f32, u32 = c_void_p, c_int32, c_float, c_uint32
    sig = {
        "vrestir_version": (i32, []),
        "vrestir_default_params": (None, [POINTER(Params)]),
        "vrestir_create": (i32, [POINTER(Params), POINTER(vp)]),
        "vrestir_destroy": (i32, [vp]),
        "vrestir_last_error": (c_char_p, [vp]),
        "vrestir_set_volume": (i32, [vp, POINTER(VolumeDesc)]),
        "vrestir_stream_volume_frame": (i32, [vp, vp, POINTER(VolumeDesc)]),
        "vrestir_share_volume": (i32, [vp, vp]),
        "vrestir_plume_create": (i32, [i32, c_uint64, POINTER(vp)]),
        "vrestir_plume_destroy": (i32, [vp]),
        "vrestir_plume_advance": (i32, [vp, vp, i32]),
        "vrestir_plume_grids": (i32, [vp, POINTER(vp), POINTER(vp), POINTER(i32)]),
        "vrestir_plume_read": (i32, [vp, vp, vp]),
        "vrestir_load_vbx": (i32, [vp, c_char_p, i32, i32, POINTER(f32), POINTER(f32), f32, f32]),
        "vrestir_set_lights": (i32, [vp, POINTER(PointLight), i32, POINTER(f32)]),
        "vrestir_set_surfaces": (i32, [vp, POINTER(SurfacesDesc)]),
        "vrestir_set_camera": (i32, [vp, POINTER(Camera)]),
        "vrestir_set_env_map": (i32, [vp, POINTER(EnvMapDesc)]),
        "vrestir_set_env_rotation": (i32, [vp, POINTER(f32)]),
        "vrestir_set_env_intensity": (i32, [vp, f32, POINTER(f32)]),
        "vrestir_set_blackbody_lut": (i32, [vp, POINTER(f32), i32]),
        "vrestir_load_vbx_emission": (i32, [vp, c_char_p, f32, f32, f32]),
        "vrestir_add_volume_frame": (i32, [vp, POINTER(VolumeDesc), POINTER(i32)]),
        "vrestir_set_volume_frame": (i32, [vp, i32, i32]),
        "vrestir_volume_frame_count": (i32, [vp, POINTER(i32)]),
        "vrestir_set_freeze_frame": (i32, [vp, i32]),
        "vrestir_camera_animation": (i32, [POINTER(Camera), POINTER(CameraAnimationDesc), i32, POINTER(Camera),
                                           POINTER(i32)]),
        "vrestir_set_params": (i32, [vp, POINTER(Params)]),
        "vrestir_get_params": (i32, [vp, POINTER(Params)]),
        "vrestir_reset_history": (i32, [vp]),
        "vrestir_execute": (i32, [vp, vp, i32, i32, vp, vp]),
        "vrestir_frame_count": (i32, [vp, POINTER(i32)]),
        "vrestir_prepare_frame": (i32, [vp, i32, i32]),
        "vrestir_run_stage": (i32, [vp, vp, i32, i32, vp]),
        "vrestir_end_frame": (i32, [vp]),
        "vrestir_read_buffer": (i32, [vp, i32, vp, c_size_t]),
        "vrestir_write_buffer": (i32, [vp, i32, vp, c_size_t]),
        "vrestir_track_batch": (i32, [vp, vp, i32, i32, i32, f32, i32, i32, vp, vp, vp]),
        "vrestir_set_tuning": (i32, [vp, i32, i32]),
        "vrestir_accumulate": (i32, [vp, vp, vp, i32, i32, i32]),
        "vrestir_accumulated": (i32, [vp, vp, vp, POINTER(c_int64)]),
        "vrestir_relmse": (i32, [vp, vp, vp, vp, i32, i32, POINTER(ctypes.c_double)]),
        "vrestir_relmse_noalloc": (i32, [vp, vp, vp, vp, i32, i32, POINTER(ctypes.c_double)]),
        "vrestir_reference_lattice": (i32, [vp, vp, i32, i32, i32, i32, ctypes.c_uint32, i32, vp]),
        "vrestir_grid_info": (i32, [vp, i32, POINTER(i32), POINTER(c_int64), POINTER(c_int64), POINTER(c_int64)]),
        "vrestir_batch_create": (i32, [vp, i32, POINTER(vp)]),
        "vrestir_batch_destroy": (i32, [vp]),
        "vrestir_batch_last_error": (c_char_p, [vp]),
        "vrestir_batch_execute": (i32, [vp, vp, POINTER(i32), POINTER(i32), vp]),
        "vrestir_batch_set_tuning": (i32, [vp, i32, i32, i32]),
        "vrestir_batch_get_tuning": (i32, [vp, POINTER(i32), POINTER(i32)]),
        "vrestir_batch_autotune": (i32, [vp, vp, POINTER(i32), POINTER(i32), vp, i32, POINTER(f32)]),
        "vrestir_default_learned": (None, [POINTER(LearnedDesc)]),
        "vrestir_set_learned": (i32, [vp, POINTER(LearnedDesc)]),
        "vrestir_get_learned": (i32, [vp, POINTER(LearnedDesc)]),
        "vrestir_learned_stats": (i32, [vp, POINTER(f32)]),
        "vrestir_learned_get_weights": (i32, [vp, POINTER(f32), i32]),
        "vrestir_learned_set_weights": (i32, [vp, POINTER(f32), i32]),
        "vrestir_learned_forward": (i32, [vp, vp, vp, i32, vp, vp]),
        "vrestir_learned_gradient": (i32, [vp, vp, vp, i32, f32, vp]),
        "vrestir_learned_train": (i32, [vp, vp, vp, i32]),
        "vrestir_synth_fbm": (i32, [c_uint64, i32, i32, i32, i32, f32, f32, f32, i32, f32, f32, POINTER(f32)]),
        "vrestir_orbit_camera": (i32, [POINTER(Camera), POINTER(f32), i32, f32, POINTER(Camera)]),
        "vrestir_count_bricks": (i32, [POINTER(f32), i32, i32, i32, POINTER(c_int64)]),
        "vrestir_write_vbx": (i32, [c_char_p, POINTER(f32), i32, i32, i32, POINTER(f32)]),
        "vrestir_read_vbx": (i32, [c_char_p, POINTER(f32), POINTER(i32), POINTER(i32), POINTER(i32), POINTER(f32)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


EXPORTED_SYMBOLS = [
    "vrestir_version", "vrestir_default_params", "vrestir_create", "vrestir_destroy", "vrestir_last_error",
    "vrestir_set_volume", "vrestir_stream_volume_frame", "vrestir_share_volume", "vrestir_load_vbx", "vrestir_set_lights", "vrestir_set_surfaces", "vrestir_set_camera", "vrestir_set_env_map",
    "vrestir_set_env_rotation", "vrestir_set_env_intensity", "vrestir_set_blackbody_lut", "vrestir_load_vbx_emission",
    "vrestir_add_volume_frame", "vrestir_set_volume_frame", "vrestir_volume_frame_count", "vrestir_set_freeze_frame",
    "vrestir_camera_animation",
    "vrestir_set_params",
    "vrestir_get_params", "vrestir_reset_history", "vrestir_execute", "vrestir_frame_count",
    "vrestir_prepare_frame", "vrestir_run_stage", "vrestir_end_frame", "vrestir_read_buffer",
    "vrestir_write_buffer", "vrestir_track_batch", "vrestir_set_tuning", "vrestir_accumulate", "vrestir_accumulated", "vrestir_relmse",
    "vrestir_relmse_noalloc", "vrestir_reference_lattice", "vrestir_grid_info",
    "vrestir_batch_create", "vrestir_batch_destroy", "vrestir_batch_last_error", "vrestir_batch_execute",
    "vrestir_batch_set_tuning", "vrestir_batch_get_tuning", "vrestir_batch_autotune",
    "vrestir_default_learned", "vrestir_set_learned", "vrestir_get_learned", "vrestir_learned_stats",
    "vrestir_learned_get_weights", "vrestir_learned_set_weights", "vrestir_learned_forward",
    "vrestir_learned_gradient", "vrestir_learned_train",
    "vrestir_plume_create", "vrestir_plume_destroy", "vrestir_plume_advance", "vrestir_plume_grids",
    "vrestir_plume_read",
    "vrestir_synth_fbm", "vrestir_orbit_camera",
    "vrestir_count_bricks", "vrestir_write_vbx", "vrestir_read_vbx",
]


def default_params() -> Params:
    p = Params()
    load_library().vrestir_default_params(ctypes.byref(p))
    return p


def params_dict(p: Params) -> dict:
    return {name: getattr(p, name) for name, _ in PARAM_FIELDS}
