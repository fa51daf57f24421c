"""Surface scenes on the GPU against the oracle: ray queries over the BVH equal
the oracle's brute force over every triangle for the same rays (same watertight
test, contraction off: hit triangle, distance and barycentric weights bit for
bit), rays into a closed mesh at its vertices and edges all hit, and
GenerateFeatures' V-buffer and features match the oracle's. The reservoir stages
refuse surface scenes."""
import ctypes

import numpy as np
import pytest

from gpu_common import check_features, rel, say
from oracle import binding as ob
from volumetricrestirrelease_amd import abi, scenes
from volumetricrestirrelease_amd.volumetric_restir import FEATURES_DTYPE, VBUFFER_DTYPE, VolumetricReSTIR

pytestmark = pytest.mark.gpu


def soup(seed=4, nt=2000):
    rng = np.random.default_rng(seed)
    c = rng.uniform(-1.5, 1.5, (nt, 1, 3))
    P = (c + rng.normal(0, 0.08, (nt, 3, 3))).astype(np.float32)
    sp, sn, si = scenes.icosphere(3, 0.7, (0.2, -0.1, 0.3))
    pos = np.concatenate([P.reshape(-1, 3), sp])
    idx = np.concatenate([np.arange(3 * nt).reshape(nt, 3), si + 3 * nt]).astype(np.int32)
    return dict(positions=pos, indices=idx, material_ids=np.zeros(len(idx), np.int32), materials=[dict()]), sp


def pair_for(surfaces, **params):
    spec = scenes.config0(res=32, width=48, height=32)
    spec.surfaces = surfaces
    spec.params.update(params)
    o = ob.Oracle()
    o.set_scene(spec)
    g = VolumetricReSTIR()
    g.setScene(spec)
    return spec, o, g


def gpu_hits(g, rays):
    import torch
    n = rays.shape[0]
    out = torch.zeros((n, 8), device="cuda")
    seeds = torch.zeros((n, 3), dtype=torch.int32, device="cuda")
    g.track(abi.TRACK_SURFACE_HIT, 0, 0, 1.0, 0, torch.from_numpy(rays).cuda(), seeds, out)
    torch.cuda.synchronize()
    return out.cpu().numpy()


def test_bvh_queries_equal_brute_force():
    say("case: 2000-triangle soup + a 1280-triangle icosphere, 20000 random rays, BVH vs brute force")
    surf, _ = soup()
    spec, o, g = pair_for(surf)
    rng = np.random.default_rng(9)
    n = 20000
    rays = np.zeros((n, 8), np.float32)
    rays[:, :3] = rng.uniform(-2.5, 2.5, (n, 3))
    d = rng.normal(size=(n, 3))
    rays[:, 3:6] = d / np.linalg.norm(d, axis=1, keepdims=True)
    rays[:, 6] = rng.uniform(0.2, 8.0, n)
    rays[: n // 10, 6] = 3.402823466e38
    og = gpu_hits(g, rays)
    oo = o.surface_hit(rays)
    same = (og[:, 1] == oo[:, 1]) & (og[:, 0] == oo[:, 0]) & (og[:, 2] == oo[:, 2]) & (og[:, 3] == oo[:, 3])
    say(f"   hits {float((oo[:, 1] >= 0).mean()):.3f}, identical {same.mean():.6f}, any-hit identical "
        f"{float((og[:, 4] == oo[:, 4]).mean()):.6f}")
    assert same.mean() >= 0.9999
    assert np.array_equal(og[:, 4], oo[:, 4])


def test_closed_mesh_watertight_on_gpu():
    say("case: rays from inside an icosphere at its vertices and edge midpoints")
    pos, nrm, idx = scenes.icosphere(3, 1.0)
    surf = dict(positions=pos, normals=nrm, indices=idx, material_ids=np.zeros(len(idx), np.int32),
                materials=[dict()])
    spec, o, g = pair_for(surf)
    T = np.concatenate([pos, 0.5 * (pos[idx[:, 0]] + pos[idx[:, 1]])]).astype(np.float64)
    d = T - np.array([0.1, -0.2, 0.05])
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((len(T), 8), np.float32)
    rays[:, :3] = (0.1, -0.2, 0.05)
    rays[:, 3:6] = d
    rays[:, 6] = 10.0
    og = gpu_hits(g, rays)
    assert (og[:, 1] >= 0).all() and (og[:, 4] == 1).all()


def test_surface_features_equal_oracle():
    say("case: GenerateFeatures with a surface scene (floor, sphere, emissive panels), 48x32")
    spec, o, g = pair_for(scenes.surface_scene_panels(), useSurfaceScene=1)
    W, H = spec.width, spec.height
    o.prepare(W, H)
    g.prepare_frame(W, H)
    o.run_stage(abi.STAGE_GENERATE_FEATURES)
    g.run_stage(abi.STAGE_GENERATE_FEATURES)
    import torch
    torch.cuda.synchronize()
    vo = o.read(abi.BUF_VBUFFER, W * H * 16).view(VBUFFER_DTYPE)
    vg = g.read(abi.BUF_VBUFFER, W * H * 16).view(VBUFFER_DTYPE)
    say(f"   hit pixels {float((vo['tri'] >= 0).mean()):.3f}, identical {float((vo == vg).mean()):.4f}")
    # the camera rays themselves differ in their last bits (ray generation is not
    # contraction-free), so the hits agree to rounding, not bit for bit
    same = vo["tri"] == vg["tri"]
    assert same.mean() >= 0.999
    assert (rel(vo["t"][same], vg["t"][same]) <= 1e-5).all()
    assert np.abs(vo["bary"][same] - vg["bary"][same]).max() <= 1e-4
    fo = o.read(abi.BUF_FEATURES, W * H * 8).view(FEATURES_DTYPE)
    fg = g.read(abi.BUF_FEATURES, W * H * 8).view(FEATURES_DTYPE)
    check_features(fo, fg)
    assert (fo["noReflectiveSurface"] == 0).any()
    # the reservoir stages are not built for surface scenes
    lib = abi.load_library()
    assert lib.vrestir_run_stage(g.h, None, abi.STAGE_TRACE_RAYS, 0, None) == abi.VRESTIR_EUNSUPPORTED
