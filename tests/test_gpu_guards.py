"""The numerical guards of the pass forced on the HIP path and the oracle alike
(the failure cases a CPU-side test can force): a non-finite density is refused
at load by both; bricks of constant density (min == max, a zero residual
majorant for decomposition tracking, VR/VolumeTrackingAdapterGVDB.slang:685-697)
through the path tracer and stage by stage; every tracker on axis-parallel rays
(1/0 in the slab test and the DDA, VR/VolumeBase.slang:140); and a camera that
turns 30 degrees a frame, so that temporal reprojection lands off-screen for many
pixels (VR/TemporalReuse.cs.slang:214-219)."""
import ctypes

import numpy as np
import pytest

from gpu_common import RADIANCE_RTOL, Pair, rel, run_injected_frames, say
from oracle.binding import Oracle, lib
from test_oracle_guards import constant_block_scene
from volumetricrestirrelease_amd import abi, scenes
from volumetricrestirrelease_amd.volumetric_restir import VolumetricReSTIR

pytestmark = pytest.mark.gpu


def test_non_finite_density_refused_by_both():
    say("case: NaN and +inf voxels in the density grid")
    g = VolumetricReSTIR()
    L = abi.load_library()
    for bad in (np.nan, np.inf):
        spec = scenes.config0(res=32, width=8, height=8)
        spec.density = spec.density.copy()
        spec.density[5, 6, 7] = bad
        d = spec.volume_desc()
        assert L.vrestir_set_volume(g.h, ctypes.byref(d)) == abi.VRESTIR_EINVAL
        assert b"non-finite" in L.vrestir_last_error(g.h)
        assert lib().orc_set_volume(Oracle().h, ctypes.byref(d)) == abi.VRESTIR_EINVAL


def test_constant_bricks_path_tracer_and_stages():
    say("case: constant-density bricks (zero-width majorants): path tracer per pixel, then stage parity")
    spec = constant_block_scene()
    spec.params.update(useReference=1, baselineSamplePerPixel=2)
    pair = Pair(spec)
    pair.begin(0)
    pair.stage(abi.STAGE_GENERATE_FEATURES)
    pair.stage(abi.STAGE_TRACE_RAYS)
    pair.stage(abi.STAGE_FINAL_SHADING)
    co, cg = pair.color()
    r = rel(co[:, :3], cg[:, :3]).max(axis=1)
    say(f"   path tracer: within 1e-4 on {(r <= RADIANCE_RTOL).mean():.4f} of pixels")
    assert np.isfinite(cg).all()
    assert (r <= RADIANCE_RTOL).mean() >= 0.99
    spec = constant_block_scene()
    spec.orbit_degrees_per_frame = 2.0
    run_injected_frames(Pair(spec), frames=2)


def test_axis_parallel_rays_every_tracker():
    say("case: every tracker mode on axis-parallel and in-plane rays")
    spec = scenes.config0(res=64, width=8, height=8)
    o = Oracle()
    o.set_scene(spec)
    g = VolumetricReSTIR()
    g.setScene(spec)
    import torch
    rng = np.random.default_rng(21)
    n = 3000
    rays = np.zeros((n, 8), np.float32)
    rays[:, :3] = rng.uniform(-0.7, 0.7, size=(n, 3))
    axis = rng.integers(0, 3, size=n)
    sign = np.where(rng.uniform(size=n) < 0.5, -1.0, 1.0)
    rays[np.arange(n), 3 + axis] = sign
    plane = np.arange(n) % 3 == 0  # a third in a coordinate plane instead
    d2 = rng.normal(size=(n, 3))
    d2[np.arange(n), axis] = 0.0
    d2 /= np.linalg.norm(d2, axis=1, keepdims=True)
    rays[plane, 3:6] = d2[plane]
    rays[:, 6] = np.where(rng.uniform(size=n) < 0.5, 3.4e38, rng.uniform(0.5, 3.0, size=n))
    seeds = np.stack([np.arange(n) % 97, np.arange(n) // 97, np.full(n, 9)], 1).astype(np.uint32)
    R = torch.from_numpy(rays).cuda()
    S = torch.from_numpy(seeds.view(np.int32)).cuda()
    for mode, slot, lin in ((abi.TRACK_TR_ANALYTIC, 0, 1), (abi.TRACK_TR_RAY_MARCHING, 1, 1),
                            (abi.TRACK_TR_RATIO, 0, 1), (abi.TRACK_TR_RESIDUAL_RATIO, 0, 1),
                            (abi.TRACK_SAMPLE_ANALYTIC, 9, 0), (abi.TRACK_SAMPLE_DECOMPOSITION, 0, 1),
                            (abi.TRACK_FEATURE_MARCH, 0, 1), (abi.TRACK_VISITED_BRICKS, 0, 1)):
        O = torch.zeros((n, 8), device="cuda")
        g.track(mode, slot, lin, 1.0, 4 if mode == abi.TRACK_SAMPLE_ANALYTIC else 1, R, S, O)
        torch.cuda.synchronize()
        b = O.cpu().numpy()
        a = o.track(mode, slot, lin, 1.0, 4 if mode == abi.TRACK_SAMPLE_ANALYTIC else 1, rays, seeds)
        assert np.isfinite(b[:, :7]).all() or mode == abi.TRACK_VISITED_BRICKS, mode
        if mode == abi.TRACK_VISITED_BRICKS:
            ok = (a[:, 0] == b[:, 0]) & (a[:, 1].view(np.uint32) == b[:, 1].view(np.uint32))
        else:
            ok = np.all((rel(a, b) <= 1e-4) | (np.abs(a - b) <= 1e-6), axis=1)
        say(f"   mode {mode}: {ok.mean():.4f} of rays match")
        assert ok.mean() >= 0.995, mode


def test_offscreen_reprojection_stage_parity():
    say("case: camera turning 30 degrees a frame (temporal reprojection off-screen), 3 frames, stage parity")
    spec = scenes.config0(res=64, width=48, height=32)
    spec.orbit_degrees_per_frame = 30.0
    report = run_injected_frames(Pair(spec), frames=3)
    assert any(r[0] == "temporal" for r in report)
