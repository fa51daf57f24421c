"""The C ABI (include/vrestir.h): library loads, exports every declared entry point,
struct layouts agree, defaults equal the reference's, host-side helpers work.
No kernel is launched here."""
import ctypes
import os
import re

import numpy as np

from volumetricrestirrelease_amd import abi, scenes
from volumetricrestirrelease_amd.volumetric_restir import VolumetricReSTIRParams

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_entry_points():
    src = open(os.path.join(ROOT, "include", "vrestir.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int32_t|void|const char\*)\s+(vrestir_\w+)\s*\(", src, re.M)))


def test_library_exports_every_declared_symbol():
    lib = abi.load_library()
    names = declared_entry_points()
    assert len(names) >= 25
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(abi.EXPORTED_SYMBOLS)


def test_version():
    assert abi.load_library().vrestir_version() == 1


def test_params_struct_layout():
    # 48 four-byte fields in declaration order, no padding
    assert ctypes.sizeof(abi.Params) == 4 * len(abi.PARAM_FIELDS) == 192
    assert ctypes.sizeof(abi.PointLight) == 24
    assert ctypes.sizeof(abi.Camera) == 48


# Defaults of VolumetricReSTIRParams, VR/VolumetricReSTIR.h:132-206
REFERENCE_DEFAULTS = dict(
    mMaxBounces=1, mEnableTemporalReuse=1, mEnableSpatialReuse=1, mVertexReuse=0, mVertexReuseStartBounce=1,
    mUseReference=0, mUseEnvironmentLights=1, mUseAnalyticLights=0, mUseEmissiveLights=0, mBaselineSamplePerPixel=1,
    mVisualizeTotalTransmittance=0, mUseSurfaceScene=0, mUsePrevVolumeForReproj=1, mInitialBaseMipLevel=1,
    mInitialM=4, mInitialLightSamples=1, mInitialLightingMipLevel=2, mInitialVisibilityUseLinearSampler=0,
    mInitialLightingUseLinearSampler=1, mInitialLightingTrackingMethod=2, mInitialVisibilityTStepScale=1.0,
    mInitialLightingTStepScale=2.0, mInitialUseRussianRoulette=1, mInitialUseCoarserGridForIndirectBounce=1,
    mTemporalReuseMThreshold=4.0, mTemporalReprojectionMode=0, mTemporalMISMethod=1, mTemporalReprojectionMipLevel=1,
    mSpatialReuseRounds=1, mSpatialVisibilityMipLevel=1, mSpatialLightingMipLevel=1,
    mSpatialVisibilityUseLinearSampler=1, mSpatialLightingUseLinearSampler=1, mSpatialVisibilityTStepScale=1.0,
    mSpatialLightingTStepScale=1.0, mSpatialVisibilityTrackingMethod=2, mSpatialLightingTrackingMethod=2,
    mRandomSamplerType=1, mSampleRadius=10.0, mSpatialSampleCount=4, mEnableVisibilitySimilarityRejection=0,
    mSpatialMISMethod=1, mFinalLightSamples=1, mFinalVisibilitySamples=1, mFinalVisibilityTrackingMethod=1,
    mFinalLightTrackingMethod=1, mFinalRandomSamplerType=1, mFinalTStepScale=0.2)


def test_default_params_match_reference():
    d = VolumetricReSTIRParams().to_dict()
    assert set(d) == set(REFERENCE_DEFAULTS)
    for k, v in REFERENCE_DEFAULTS.items():
        assert abs(d[k] - v) < 1e-7, k


def test_oracle_defaults_equal_product_defaults():
    from oracle import binding as ob
    p = abi.Params()
    ob.lib().vrestir_default_params_oracle(ctypes.byref(p))
    assert abi.params_dict(p) == abi.params_dict(abi.default_params())


def test_params_reference_names_roundtrip():
    p = VolumetricReSTIRParams(mMaxBounces=3, mSampleRadius=7.5)
    assert p.mMaxBounces == 3 and abs(p.mSampleRadius - 7.5) < 1e-7
    try:
        p.mNoSuchField = 1
        raise AssertionError("unknown field accepted")
    except AttributeError:
        pass


def test_synth_fbm_deterministic_and_seeded():
    a = scenes.fbm(7, (32, 32, 32), threshold=-0.1)
    b = scenes.fbm(7, (32, 32, 32), threshold=-0.1)
    c = scenes.fbm(8, (32, 32, 32), threshold=-0.1)
    assert np.array_equal(a, b)
    assert not np.array_equal(a, c)
    assert a.min() >= 0.0 and a.max() > 0.0
    # sphere mask: nothing outside the inscribed sphere
    z, y, x = np.meshgrid(*(np.arange(32) + 0.5,) * 3, indexing="ij")
    r = np.sqrt((x - 16) ** 2 + (y - 16) ** 2 + (z - 16) ** 2)
    assert np.all(a[r > 16.0 + 1e-3] == 0)


def test_count_bricks_matches_numpy():
    d = scenes.fbm(3, (64, 32, 16), threshold=0.1)
    n = ctypes.c_int64()
    assert abi.load_library().vrestir_count_bricks(d.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), 64, 32, 16,
                                                   ctypes.byref(n)) == 0
    occ = (d.reshape(2, 8, 4, 8, 8, 8) > 0).any(axis=(1, 3, 5)).sum()
    assert n.value == occ


def test_orbit_camera_keeps_distance_and_target():
    spec = scenes.config0(res=32, width=16, height=16)
    c0 = spec.camera_struct(0)
    for f in (1, 45, 90, 180):
        out = abi.Camera()
        center = (ctypes.c_float * 3)(0.0, 0.0, 0.0)
        assert abi.load_library().vrestir_orbit_camera(ctypes.byref(c0), center, f, 1.0, ctypes.byref(out)) == 0
        p0 = np.array(c0.posW[:]); p = np.array(out.posW[:]); t = np.array(out.target[:])
        # same height, same horizontal distance to the rotation centre
        assert abs(p[1] - p0[1]) < 1e-5
        assert abs(np.hypot(p[0] - t[0], p[2] - t[2]) - np.hypot(p0[0] - t[0], p0[2] - t[2])) < 1e-4
        a, b = p0[[0, 2]] - t[[0, 2]], p[[0, 2]] - t[[0, 2]]
        ang = np.degrees(np.arctan2(a[0] * b[1] - a[1] * b[0], np.dot(a, b)))
        assert abs(abs(ang) - (f if f <= 180 else 360 - f)) < 1e-2


def test_create_without_device_fails_cleanly_or_succeeds_with_one():
    lib = abi.load_library()
    h = ctypes.c_void_p()
    rc = lib.vrestir_create(None, ctypes.byref(h))
    if rc == 0:
        assert lib.vrestir_destroy(h) == 0
    else:
        assert rc == abi.VRESTIR_EHIP
        assert lib.vrestir_last_error(None)


def test_invalid_params_rejected_without_device():
    lib = abi.load_library()
    h = ctypes.c_void_p()
    for field, val, code in [("vertexReuse", 1, abi.VRESTIR_EUNSUPPORTED),
                             ("maxBounces", 0, abi.VRESTIR_EINVAL),
                             ("maxBounces", 9, abi.VRESTIR_EINVAL),
                             ("spatialSampleCount", 33, abi.VRESTIR_EINVAL),
                             ("initialLightingTrackingMethod", 7, abi.VRESTIR_EINVAL)]:
        p = abi.default_params()
        setattr(p, field, val)
        assert lib.vrestir_create(ctypes.byref(p), ctypes.byref(h)) == code, field
        assert lib.vrestir_last_error(None)
