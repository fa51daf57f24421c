"""Pins the oracle's surface shading (ShadingData of an untextured metal-rough
material and Falcor's legacy standard BSDF) at hits on a quad in the z = 0 plane:
  * every sample's pdf is evalPdfBSDF(wi), and the sample weights (the sampled
    lobe's f cos / pdf over its probability) average to the same albedo as
    evalBSDFCosine(wi) / pdf (sampling and evaluation are separate code: a dropped
    lobe probability, a Jacobian left out of the GGX pdf or a mismatched Fresnel
    argument breaks one of them);
  * the samples follow their pdf over the whole hemisphere: E[1 / pdf] = 2 pi;
  * energy: the mean weight (the directional albedo) stays at or below 1 for white
    dielectrics and metals, and a smooth white metal at normal incidence keeps
    almost all of it;
  * a single-sided surface seen from behind scatters nothing (zero weights and
    evaluations), a double-sided one flips its normal; emission shows on the front
    side only.
"""
import math

import numpy as np

from oracle import binding as ob
from volumetricrestirrelease_amd import scenes

QUAD = np.array([(-5, -5, 0), (5, -5, 0), (5, 5, 0), (-5, 5, 0)], np.float32)


def oracle_quad(**mat):
    spec = scenes.config0(res=32, width=8, height=8)
    spec.surfaces = dict(positions=QUAD, indices=np.array([(0, 1, 2), (0, 2, 3)], np.int32),
                         material_ids=np.zeros(2, np.int32), materials=[mat])
    o = ob.Oracle()
    o.set_scene(spec)
    return o


def rays_towards_origin(theta_deg, n, from_below=False):
    th = math.radians(theta_deg)
    d = np.array([math.sin(th), 0.0, -math.cos(th)])
    if from_below:
        d[2] = -d[2]
    r = np.zeros((n, 8), np.float32)
    r[:, :3] = (-2.0 * d)
    r[:, 3:6] = d
    r[:, 6] = 10.0
    s = np.stack([np.arange(n) % 1024, np.arange(n) // 1024, np.full(n, 3)], axis=1).astype(np.uint32)
    return r, s


def test_sample_weight_equals_eval_over_pdf():
    for mat in (dict(baseColor=(0.8, 0.5, 0.3), roughness=0.5, metallic=0.0),
                dict(baseColor=(0.9, 0.9, 0.9), roughness=0.2, metallic=1.0),
                dict(baseColor=(0.3, 0.6, 0.9), roughness=0.8, metallic=0.4, IoR=1.33)):
        o = oracle_quad(**mat)
        for th in (0.0, 35.0, 70.0):
            r, s = rays_towards_origin(th, 20000)
            smp = o.surface_shade(0, r, s)
            ev = o.surface_shade(1, r, s)
            ok = smp[:, 7] > 0
            assert ok.mean() > 0.5, (mat, th)
            pdf, w = smp[ok, 3].astype(np.float64), smp[ok, 4:7].astype(np.float64)
            f, epdf = ev[ok, :3].astype(np.float64), ev[ok, 3].astype(np.float64)
            # the sampler takes NdotH from its draw, the evaluation from normalize(V + L):
            # a sharp lobe (alpha 0.04) magnifies that rounding
            assert np.allclose(pdf, epdf, rtol=2e-3, atol=1e-7), (mat, th)
            # the weight is the sampled lobe's f cos / pdf over the lobe's probability:
            # per sample it is not f / pdf of the mixture, but both estimate the albedo
            n = len(r)
            wa = np.where(ok[:, None], smp[:, 4:7], 0.0).astype(np.float64)
            fa = np.where(ok[:, None], ev[:, :3] / np.maximum(ev[:, 3:4], 1e-30), 0.0).astype(np.float64)
            d = wa - fa
            assert (np.abs(d.mean(axis=0)) <= 4 * d.std(axis=0) / math.sqrt(n) + 1e-5).all(), (mat, th)
            # unit directions above the surface
            wi = smp[ok, :3].astype(np.float64)
            assert np.allclose(np.linalg.norm(wi, axis=1), 1.0, atol=1e-5) and (wi[:, 2] > 0).all()


def test_samples_cover_the_hemisphere_by_their_pdf():
    o = oracle_quad(baseColor=(0.7, 0.7, 0.7), roughness=0.4, metallic=0.0)
    n = 200_000
    r, s = rays_towards_origin(25.0, n)
    smp = o.surface_shade(0, r, s)
    ok = smp[:, 7] > 0
    inv = np.where(ok, 1.0 / np.maximum(smp[:, 3].astype(np.float64), 1e-30), 0.0)
    est, se = inv.mean(), inv.std() / math.sqrt(n)
    assert abs(est - 2 * math.pi) <= 4 * se + 1e-3, (est, se)


def test_directional_albedo_at_most_one():
    n = 100_000
    for mat, lo in ((dict(baseColor=(1.0, 1.0, 1.0), roughness=0.5, metallic=0.0), 0.5),
                    (dict(baseColor=(1.0, 1.0, 1.0), roughness=0.05, metallic=1.0), 0.95)):
        o = oracle_quad(**mat)
        r, s = rays_towards_origin(0.0, n)
        smp = o.surface_shade(0, r, s)
        w = np.where((smp[:, 7] > 0)[:, None], smp[:, 4:7], 0.0).astype(np.float64)
        a, se = w.mean(axis=0), w.std(axis=0) / math.sqrt(n)
        assert (a <= 1.0 + 4 * se + 1e-4).all(), (mat, a)
        assert (a >= lo).all(), (mat, a)


def test_back_faces_and_emission():
    r, s = rays_towards_origin(20.0, 2000, from_below=True)
    o1 = oracle_quad(baseColor=(0.5, 0.5, 0.5), emissive=(2.0, 1.0, 0.5))
    smp, ev = o1.surface_shade(0, r, s), o1.surface_shade(1, r, s)
    # nothing scattered (a specular draw can still report the diffuse pdf of its
    # direction, as evalPdfDiffuse tests only the light side), nothing emitted
    assert (smp[:, 4:7] == 0).all() and (ev[:, :3] == 0).all()
    assert (ev[:, 4:7] == 0).all() and (ev[:, 7] < 0).all()
    o2 = oracle_quad(baseColor=(0.5, 0.5, 0.5), emissive=(2.0, 1.0, 0.5), doubleSided=1)
    smp, ev = o2.surface_shade(0, r, s), o2.surface_shade(1, r, s)
    ok = smp[:, 7] > 0
    assert ok.mean() > 0.9 and (smp[ok, 2] < 0).all()     # scattered back below the quad
    assert (ev[:, 7] > 0).all() and (ev[:, 4:7] == 0).all()  # flipped normal, still no emission
    rf, sf = rays_towards_origin(20.0, 100)
    assert np.allclose(o1.surface_shade(1, rf, sf)[:, 4:7], (2.0, 1.0, 0.5))
