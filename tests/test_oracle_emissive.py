"""Pins the oracle's emissive-triangle lights (the light collection, the power
sampler, sampleTriangle and the emissive branch of evaluate_L_in_volume) with
closed forms and statistics that a wrong pdf measure, a dropped cosine, a
mis-built alias table or a flipped normal breaks:
  * E[Li] of emissive light samples at a point is the sum over front-facing
    triangles of Le * solid angle (Van Oosterom and Strackee's formula), also
    with every light kind on (then the emissive part is a third of the draws);
  * triangles are picked in proportion to their flux luminance(Le) * area (chi-square
    over the collection), the zero-area and non-emissive triangles never;
  * pdfArea is the flux share over the triangle's area; sampled points lie on the
    triangle (barycentric weights in [0, 1]);
  * the evaluation branch, Le cos / d^2 at uniform uv, times the area, averages to
    Le * solid angle; a triangle facing away contributes nothing;
  * the emissive intensity multiplier scales Li linearly.
"""
import math

import numpy as np
from scipy import stats

from oracle import binding as ob
from volumetricrestirrelease_amd import scenes


def solid_angle(p, tri):
    """Van Oosterom & Strackee (1983): solid angle of triangle tri seen from p."""
    a, b, c = (np.asarray(v, np.float64) - p for v in tri)
    la, lb, lc = (np.linalg.norm(v) for v in (a, b, c))
    num = abs(np.dot(a, np.cross(b, c)))
    den = la * lb * lc + np.dot(a, b) * lc + np.dot(a, c) * lb + np.dot(b, c) * la
    om = 2.0 * math.atan2(num, den)
    return om


def front_facing(p, tri):
    a, b, c = (np.asarray(v, np.float64) for v in tri)
    n = np.cross(b - a, c - a)
    return np.dot(n, p - a) > 0.0


def setup(mult=1.0):
    spec = scenes.config0(res=32, width=8, height=8)
    spec.surfaces = scenes.emissive_panels()
    spec.surfaces["emissive_multiplier"] = mult
    spec.params.update(useEmissiveLights=1)
    o = ob.Oracle()
    o.set_scene(spec)
    return o, spec


def expected_irradiance(spec, p):
    sf = spec.surfaces
    e = np.zeros(3)
    for t, m in zip(sf["indices"], sf["material_ids"]):
        Le = np.array(sf["materials"][m].get("emissive", (0, 0, 0)), np.float64)
        tri = [sf["positions"][i] for i in t]
        if Le.max() > 0 and front_facing(p, tri):
            e += Le * solid_angle(p, tri)
    return e * sf.get("emissive_multiplier", 1.0)


def seeds(n, tag):
    s = np.zeros((n, 3), np.uint32)
    s[:, 0] = np.arange(n) % 4096
    s[:, 1] = np.arange(n) // 4096
    s[:, 2] = tag
    return s


def test_emissive_samples_estimate_le_times_solid_angle():
    o, spec = setup()
    n = 200_000
    for p in ((0.0, 0.0, 0.0), (0.4, -0.3, 0.2), (-0.5, 0.6, -0.4)):
        pts = np.tile(np.array(p, np.float32), (n, 1))
        out = o.sample_lights(4, pts, seeds(n, 1))
        Li = out[:, :3].astype(np.float64)
        Li[out[:, 4] == -9] = 0.0
        est, se = Li.mean(axis=0), Li.std(axis=0) / math.sqrt(n)
        want = expected_irradiance(spec, np.array(p))
        assert np.all(np.abs(est - want) <= 4 * se + 1e-4 * want), (p, est, want, se)
        # every light kind on: a third of the draws go to the triangles, whose Li carries 1/(1/3)
        out = o.sample_lights(7, pts, seeds(n, 2))
        em = out[:, 4] >= len(spec.lights)
        Lie = np.where(em[:, None], out[:, :3], 0.0).astype(np.float64)
        est, se = Lie.mean(axis=0), Lie.std(axis=0) / math.sqrt(n)
        assert np.all(np.abs(est - want) <= 4 * se + 1e-4 * want), (p, est, want, se)


def test_power_sampler_picks_triangles_by_flux():
    o, spec = setup()
    sf = spec.surfaces
    n = 100_000
    # a point the downward quad and the side triangle both face
    pts = np.tile(np.array((0.2, 0.1, 0.0), np.float32), (n, 1))
    out = o.sample_lights(4, pts, seeds(n, 3))
    ok = out[:, 4] != -9
    tri = out[ok, 4].astype(int) - len(spec.lights)
    # collection order: triangles of emissive materials in scene order
    coll = [t for t, m in enumerate(sf["material_ids"]) if max(sf["materials"][m].get("emissive", (0, 0, 0))) > 0]
    lum = lambda c: 0.2126 * c[0] + 0.7152 * c[1] + 0.0722 * c[2]
    flux, area = [], []
    for t in coll:
        P = [np.asarray(sf["positions"][i], np.float64) for i in sf["indices"][t]]
        A = 0.5 * np.linalg.norm(np.cross(P[1] - P[0], P[2] - P[0]))
        area.append(A)
        flux.append(lum(sf["materials"][sf["material_ids"][t]]["emissive"]) * A * math.pi)
    flux, area = np.array(flux), np.array(area)
    counts = np.bincount(tri, minlength=len(coll))
    # the triangle facing away yields invalid samples: count its draws from the invalid ones
    facing = np.array([front_facing(np.array((0.2, 0.1, 0.0)), [sf["positions"][i] for i in sf["indices"][t]])
                       for t in coll])
    assert counts[~facing].sum() == 0 and counts[flux == 0].sum() == 0
    p = flux / flux.sum()
    sel = facing & (flux > 0)
    expected = p[sel] / p[sel].sum() * counts[sel].sum()
    chi = stats.chisquare(counts[sel], expected)
    assert chi.pvalue > 1e-3, (counts, expected)
    assert abs((~ok).mean() - p[~facing].sum()) < 5 * math.sqrt(p[~facing].sum() / n) + 1e-3
    # pdfArea = flux share / area
    pa = out[ok, 3].astype(np.float64)
    want = p[tri] / area[tri]
    assert np.allclose(pa, want, rtol=2e-6)


def test_emissive_eval_integrates_to_le_times_solid_angle():
    o, spec = setup()
    sf = spec.surfaces
    coll = [t for t, m in enumerate(sf["material_ids"]) if max(sf["materials"][m].get("emissive", (0, 0, 0))) > 0]
    rng = np.random.default_rng(5)
    n = 200_000
    p = np.array((0.1, -0.2, 0.3))
    for k, t in enumerate(coll):
        P = [np.asarray(sf["positions"][i], np.float64) for i in sf["indices"][t]]
        A = 0.5 * np.linalg.norm(np.cross(P[1] - P[0], P[2] - P[0]))
        uv = rng.random((n, 2)).astype(np.float32)
        out = o.emissive_eval(np.tile(p.astype(np.float32), (n, 1)), np.full(n, k, np.int32), uv)
        Le = np.array(sf["materials"][sf["material_ids"][t]]["emissive"], np.float64)
        if A == 0.0 or not front_facing(p, P):
            assert (out[:, 7] == 0).all()
            continue
        assert (out[:, 7] == 1).all()
        est = out[:, :3].astype(np.float64).mean(axis=0) * A
        se = out[:, :3].astype(np.float64).std(axis=0) * A / math.sqrt(n)
        want = Le * solid_angle(p, P)
        assert np.all(np.abs(est - want) <= 4 * se + 1e-4 * want), (t, est, want)
        # the returned distance and direction reach the triangle's plane
        hit = p + out[:, 4:7].astype(np.float64) * out[:, 3:4].astype(np.float64)
        nrm = np.cross(P[1] - P[0], P[2] - P[0])
        nrm /= np.linalg.norm(nrm)
        assert np.abs((hit - P[0]) @ nrm).max() < 1e-4


def test_emissive_multiplier_scales_li():
    o1, _ = setup(1.0)
    o3, _ = setup(3.0)
    n = 4096
    pts = np.tile(np.array((0.0, 0.2, 0.1), np.float32), (n, 1))
    a = o1.sample_lights(4, pts, seeds(n, 9))
    b = o3.sample_lights(4, pts, seeds(n, 9))
    assert np.array_equal(a[:, 4], b[:, 4])
    assert np.allclose(b[:, :3], 3.0 * a[:, :3], rtol=1e-6)
