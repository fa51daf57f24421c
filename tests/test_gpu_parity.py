"""Parity of the HIP pass with the CPU oracle, stage by stage on identical inputs
(the oracle's buffers are injected before each GPU stage), and over whole frames.

Bar: integer reservoir fields (light ids, path tags, indirect-bounce counts, M)
identical and float fields within 2e-3 on >= 99.5% of pixels per stage; final
per-pixel radiance within 1e-4 relative wherever the reservoirs match (control
flow matches); whole uninjected frames within 1e-3 on >= 97% of pixels."""
import numpy as np
import pytest

from gpu_common import RADIANCE_RTOL, Pair, compare_reservoirs, rel, run_injected_frames, say, small_scene
from volumetricrestirrelease_amd import abi
from volumetricrestirrelease_amd.volumetric_restir import FEATURES_DTYPE

pytestmark = pytest.mark.gpu

def test_stage_parity_config0_small():
    say("case: config0 64^3 fp32, 64x48, b=2, M=4, 1 spatial round k=4")
    run_injected_frames(Pair(small_scene()))


def test_stage_parity_fp16_three_bounces_two_rounds():
    say("case: 64^3 fp16 bricks, 64x48, b=3, M=8, 2 spatial rounds k=6 (config1 options)")
    spec = small_scene(fp16=True, maxBounces=3, initialM=8, temporalReuseMThreshold=2.5, spatialReuseRounds=2,
                       spatialSampleCount=6)
    spec.orbit_degrees_per_frame = 2.0
    run_injected_frames(Pair(spec), frames=3)


@pytest.mark.parametrize("variant", [
    dict(maxBounces=1),
    dict(maxBounces=4, initialM=5),
    dict(randomSamplerType=abi.SAMPLER_HAMMERSLEY, spatialMISMethod=abi.MIS_NONE, temporalMISMethod=abi.MIS_NONE),
    dict(temporalReprojectionMode=abi.REPROJECTION_NONE, initialVisibilityUseLinearSampler=1),
    dict(spatialVisibilityTrackingMethod=abi.ANALYTIC_TRACKING, spatialLightingTrackingMethod=abi.RESIDUAL_RATIO_TRACKING,
         initialLightingTrackingMethod=abi.RATIO_TRACKING, finalLightTrackingMethod=abi.RAY_MARCHING),
    dict(enableTemporalReuse=0),
    dict(enableSpatialReuse=0),
    dict(useEnvironmentLights=0),
    dict(useAnalyticLights=0, initialUseCoarserGridForIndirectBounce=0),
    # spatial rays on two grids and filters: the segment trace's hierarchical walk and per-segment march
    dict(spatialVisibilityMipLevel=0, spatialLightingMipLevel=1, spatialVisibilityUseLinearSampler=0),
])
def test_stage_parity_option_surface(variant):
    say("case: option variant", variant)
    spec = small_scene(res=32, width=40, height=32, **variant)
    spec.orbit_degrees_per_frame = 1.5
    run_injected_frames(Pair(spec), frames=2)


def test_no_reuse_and_reference_modes_per_pixel():
    say("case: no reuse (delta tracking candidates) and the reference path tracer")
    for variant in (dict(enableTemporalReuse=0, enableSpatialReuse=0), dict(useReference=1, baselineSamplePerPixel=2)):
        pair = Pair(small_scene(res=32, width=40, height=32, **variant))
        pair.begin(0)
        pair.stage(abi.STAGE_GENERATE_FEATURES)
        pair.stage(abi.STAGE_TRACE_RAYS)
        if not variant.get("useReference"):
            pair.inject(abi.BUF_RESERVOIRS0, pair.nres)
            pair.inject(abi.BUF_EXTRA0, pair.nex)
        pair.stage(abi.STAGE_FINAL_SHADING)
        co, cg = pair.color()
        r = rel(co[:, :3], cg[:, :3]).max(axis=1)
        say("   ", variant, "frac within 1e-4:", float((r <= RADIANCE_RTOL).mean()))
        assert (r <= 1e-3).mean() >= 0.99


def test_whole_frames_uninjected():
    say("case: 4 whole frames, orbiting camera, no injection")
    spec = small_scene(maxBounces=2)
    spec.orbit_degrees_per_frame = 1.0
    pair = Pair(spec)
    import torch
    out = torch.zeros((pair.H, pair.W, 4), device="cuda")
    for f in range(4):
        pair.o.set_camera(spec, f)
        pair.g.setCamera(spec, f)
        co = pair.o.execute(pair.W, pair.H, nthreads=8).reshape(-1, 4)
        pair.g.execute(out)
        torch.cuda.synchronize()
        cg = out.cpu().numpy().reshape(-1, 4)
        r = rel(co[:, :3], cg[:, :3]).max(axis=1)
        frac = float((r <= 1e-3).mean())
        say(f"   frame {f}: frac within 1e-3 {frac:.4f}, mean radiance oracle {co[:, :3].mean():.5f} gpu {cg[:, :3].mean():.5f}")
        assert frac >= 0.97
        assert abs(co[:, :3].mean() - cg[:, :3].mean()) <= 2e-3 * co[:, :3].mean()


@pytest.mark.parametrize("mode,rtol,frac_min,mixed", [(1, 1e-5, 0.999, False), (2, 1e-4, 0.999, False),
                                                     (2, 1e-4, 0.999, True)])
def test_split_spatial_reuse_matches_single_kernel(mode, rtol, frac_min, mixed):
    """Mode 1 traces each recorded ray in one lane: the same arithmetic as the
    single kernel. Mode 2 sums each ray's optical depth per brick segment, so its
    transmittances differ by float summation order (a few 1e-6 relative) and a
    reservoir selection can flip where a random number falls that close to its
    threshold."""
    say(f"case: spatial reuse as record / trace (mode {mode}) / replay vs one kernel, fp16 b=3, 2 rounds, 3 frames")
    import torch
    from volumetricrestirrelease_amd.volumetric_restir import VolumetricReSTIR
    spec = small_scene(fp16=True, maxBounces=3, initialM=8, temporalReuseMThreshold=2.5, spatialReuseRounds=2,
                       spatialSampleCount=6)
    if mixed:  # visibility rays point-sampled at mip 0, lighting rays trilinear at mip 1
        spec.params.update(spatialVisibilityMipLevel=0, spatialLightingMipLevel=1, spatialVisibilityUseLinearSampler=0)
    spec.orbit_degrees_per_frame = 2.0
    outs = []
    for split in (0, mode):
        g = VolumetricReSTIR()
        g.setScene(spec)
        assert abi.load_library().vrestir_set_tuning(g.h, abi.TUNE_SPLIT_REUSE, split) == 0
        out = torch.zeros((spec.height, spec.width, 4), device="cuda")
        frames = []
        for f in range(3):
            g.setCamera(spec, f)
            g.execute(out)
            torch.cuda.synchronize()
            frames.append(out.cpu().numpy().copy())
        outs.append(frames)
    for f in range(3):
        a, b = outs[0][f][..., :3], outs[1][f][..., :3]
        frac = float((rel(a.reshape(-1, 3), b.reshape(-1, 3)).max(axis=1) <= rtol).mean())
        say(f"   frame {f}: pixels equal within {rtol}: {frac:.4f}")
        assert frac >= frac_min


@pytest.mark.parametrize("layout", ["sparse", "dense"])
def test_split_spatial_reuse_block_borders(layout):
    """A 256^3 grid (2 x 2 x 2 blocks of 128^3) traced at mip 0, brick segments
    against the single kernel. sparse: two clouds in opposite blocks, so the walk
    steps over the six empty blocks between them; dense: one cloud through all
    eight blocks, so walks cross block borders inside it, where the hierarchical
    walk re-prepares 0.01 past the border (skipping a corner brick the ray clips
    for less than that; a brick walk ignoring borders differed on 3 % of rays).
    At mip 0 the march step is also shorter than that offset, so a ray's last
    brick can hold samples before tFar although its offset entry lies past tFar
    (a walk that stopped at the offset entry put 6 % of sparse pixels off by more
    than 1e-3)."""
    say(f"case: spatial reuse at mip 0, brick segments vs one kernel on a {layout} 256^3 grid of 2 x 2 x 2 blocks")
    import torch
    from volumetricrestirrelease_amd.volumetric_restir import VolumetricReSTIR
    spec = small_scene(res=256, width=64, height=48, fp16=True, maxBounces=3, initialM=6, spatialReuseRounds=2,
                       spatialSampleCount=5, spatialVisibilityMipLevel=0, spatialLightingMipLevel=0)
    if layout == "sparse":
        d = np.zeros_like(spec.density)
        d[16:112, 16:112, 16:112] = spec.density[16:112, 16:112, 16:112]
        d[144:240, 144:240, 144:240] = spec.density[144:240, 144:240, 144:240]
        spec.density = d
    outs = []
    for split in (0, 2):
        g = VolumetricReSTIR()
        g.setScene(spec)
        assert abi.load_library().vrestir_set_tuning(g.h, abi.TUNE_SPLIT_REUSE, split) == 0
        out = torch.zeros((spec.height, spec.width, 4), device="cuda")
        frames = []
        for f in range(2):
            g.setCamera(spec, f)
            g.execute(out)
            torch.cuda.synchronize()
            frames.append(out.cpu().numpy().copy())
        outs.append(frames)
    for f in range(2):
        a, b = outs[0][f][..., :3], outs[1][f][..., :3]
        assert a.mean() > 0
        frac = float((rel(a.reshape(-1, 3), b.reshape(-1, 3)).max(axis=1) <= 1e-4).mean())
        say(f"   frame {f}: pixels equal within 1e-4: {frac:.4f}")
        assert frac >= 0.999


@pytest.mark.parametrize("variant", ["temporal", "final"])
def test_split_temporal_and_final_match_single_kernel(variant):
    """TemporalReuse (Talbot MIS, b=3, orbiting camera, previous-camera target
    evaluations) and analytic FinalShading as record / trace / replay against their
    single kernels over whole frames. Brick segments sum each ray's optical depth
    per segment, so transmittances differ by float summation order; the images
    agree within 1e-4 on >= 99.9 % of pixels."""
    say(f"case: {variant} as record / brick-segment trace / replay vs one kernel, b=3, 4 orbiting frames")
    import torch
    from volumetricrestirrelease_amd.volumetric_restir import VolumetricReSTIR
    params = dict(maxBounces=3, initialM=8, enableSpatialReuse=0, temporalMISMethod=abi.MIS_TALBOT,
                  temporalReuseMThreshold=2.5)
    if variant == "final":
        params.update(enableTemporalReuse=0)
    spec = small_scene(fp16=True, **params)
    spec.orbit_degrees_per_frame = 2.0
    outs = []
    for split in (0, 2):
        g = VolumetricReSTIR()
        g.setScene(spec)
        assert abi.load_library().vrestir_set_tuning(g.h, abi.TUNE_SPLIT_REUSE, split) == 0
        out = torch.zeros((spec.height, spec.width, 4), device="cuda")
        frames = []
        for f in range(4):
            g.setCamera(spec, f)
            g.execute(out)
            torch.cuda.synchronize()
            frames.append(out.cpu().numpy().copy())
        outs.append(frames)
    for f in range(4):
        a, b = outs[0][f][..., :3], outs[1][f][..., :3]
        assert a.mean() > 0
        frac = float((rel(a.reshape(-1, 3), b.reshape(-1, 3)).max(axis=1) <= 1e-4).mean())
        say(f"   frame {f}: pixels equal within 1e-4: {frac:.4f}")
        assert frac >= 0.999


def test_default_path_repeats_exactly():
    """The same frames rendered twice by two instances on the default path
    (split stages, brick-segment traces) give identical images. The segment
    traces add a ray's per-segment optical depths with LDS atomics, but only the
    lanes of the wave that owns the ray's queue slot add to it, in the order that
    wave's instructions issue, so the sums do not depend on how waves are
    scheduled."""
    say("case: the default path twice, fp16 b=3, 2 spatial rounds, 3 orbiting frames")
    import torch
    from volumetricrestirrelease_amd.volumetric_restir import VolumetricReSTIR
    spec = small_scene(fp16=True, maxBounces=3, initialM=8, temporalReuseMThreshold=2.5, spatialReuseRounds=2,
                       spatialSampleCount=6)
    spec.orbit_degrees_per_frame = 2.0
    runs = []
    for _ in range(2):
        g = VolumetricReSTIR()
        g.setScene(spec)
        out = torch.zeros((spec.height, spec.width, 4), device="cuda")
        frames = []
        for f in range(3):
            g.setCamera(spec, f)
            g.execute(out)
            torch.cuda.synchronize()
            frames.append(out.cpu().numpy().copy())
        runs.append(frames)
    for f in range(3):
        same = float(np.mean(np.all(runs[0][f] == runs[1][f], axis=-1)))
        say(f"   frame {f}: identical pixels {same:.6f}")
        assert np.array_equal(runs[0][f], runs[1][f])
