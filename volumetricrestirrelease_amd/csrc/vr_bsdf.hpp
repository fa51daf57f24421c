// vr_bsdf.hpp -- shading at a triangle hit: ShadingData of an untextured
// metal-rough material (SC/Shading.slang:155-245, computeSurfaceShadingInfo of
// VR/InlineRayTracingHelpers.slang:157-173) and Falcor's legacy standard BSDF
// (Experimental/Scene/Material/MaterialShading.slang:167-631, the default
// configuration: Frostbite diffuse, height-correlated Smith-GGX, Schlick Fresnel):
// evalBSDFCosine, evalPdfBSDF and sampleBSDF.
//
// The tangent frame: the reference takes the mesh's tangents; triangles here carry
// none, so the frame around the shading normal is the branchless orthonormal basis
// of Duff et al. ("Building an Orthonormal Basis, Revisited", JCGT 6(1), 2017).
// The BSDF is isotropic, so evaluation does not depend on it; sampling does.
#pragma once
#include "vr_rayquery.hpp"

namespace vr {

struct ShadingData {
    v3 posW, V, N, faceN, T, B;
    bool frontFacing;
    float NdotV;
    v3 diffuse, specular, emissive;
    float linearRoughness, ggxAlpha, metallic;
};
constexpr float kMinGGXAlpha = 0.0064f;
constexpr float kMinCosTheta = 1e-6f;

VR_DEV void onb(v3 n, v3& t, v3& b) {
    const float sign = copysignf(1.f, n.z);
    const float a = -1.f / (sign + n.z);
    const float c = n.x * n.y * a;
    t = V3(1.f + sign * n.x * n.x * a, sign * c, -sign * n.x);
    b = V3(c, sign + n.y * n.y * a, -n.y);
}
VR_DEV v3 fromLocal(v3 v, const ShadingData& sd) { return sd.T * v.x + sd.B * v.y + sd.N * v.z; }

// computeSurfaceShadingInfo for a V-buffer hit (tri, weights of vertices 1 and 2)
// seen along viewDir (towards the viewer).
VR_DEV ShadingData surfaceShading(const Scene& S, int tri, float bx, float by, v3 viewDir) {
    const float* T = S.triData + (size_t)kTriDataFloats * tri;
    const float* M = S.materials + (size_t)kMaterialFloats * __float_as_int(T[18]);
    const float b0 = 1.f - bx - by;
    ShadingData sd;
    const v3 p0 = ld3(T), p1 = ld3(T + 3), p2 = ld3(T + 6);
    sd.posW = p0 * b0 + p1 * bx + p2 * by;
    sd.faceN = norm3(cross3(p1 - p0, p2 - p0));
    sd.N = norm3(ld3(T + 9) * b0 + ld3(T + 12) * bx + ld3(T + 15) * by);
    sd.V = viewDir;
    sd.frontFacing = dot3(sd.V, sd.faceN) >= 0.f;
    const v3 base = ld3(M);
    const float rough = M[3], metal = M[4], ior = M[5];
    // lerp(a, b, t) = a + (b - a) t, as HLSL's
    sd.diffuse = base + (V3(0.f) - base) * metal;
    const float f = (ior - 1.f) / (ior + 1.f), F0 = f * f;
    sd.specular = V3(F0) + (base - V3(F0)) * metal;
    sd.linearRoughness = rough;
    sd.metallic = metal;
    sd.ggxAlpha = fmaxf(kMinGGXAlpha, rough * rough);
    sd.emissive = sd.frontFacing ? ld3(M + 6) : V3(0.f);
    sd.NdotV = dot3(sd.N, sd.V);
    if (!sd.frontFacing && M[9] != 0.f) {
        sd.N = -sd.N;
        sd.NdotV = -sd.NdotV;
    }
    onb(sd.N, sd.T, sd.B);
    sd.posW = rayOriginOffset(sd.posW, sd.frontFacing ? sd.faceN : -sd.faceN);
    return sd;
}

VR_DEV v3 fresnelSchlick(v3 f0, float f90, float cosTheta) {
    const float p = powf(fmaxf(1.f - cosTheta, 0.f), 5.f);
    return f0 + (V3(f90) - f0) * p;
}
VR_DEV float fresnelSchlick1(float f0, float f90, float cosTheta) {
    return f0 + (f90 - f0) * powf(fmaxf(1.f - cosTheta, 0.f), 5.f);
}
VR_DEV float ndfGGX(float alpha, float cosTheta) {
    const float a2 = alpha * alpha;
    const float d = ((cosTheta * a2 - cosTheta) * cosTheta + 1.f);
    return a2 / (d * d * kPiF);
}
VR_DEV float smithGGXCorrelated(float alpha, float NdotL, float NdotV) {
    const float a2 = alpha * alpha;
    const float lambdaV = NdotL * sqrtf((-NdotV * a2 + NdotV) * NdotV + a2);
    const float lambdaL = NdotV * sqrtf((-NdotL * a2 + NdotL) * NdotL + a2);
    return 0.5f / (lambdaV + lambdaL);
}
VR_DEV v3 diffuseFrostbite(const ShadingData& sd, float NdotL, float NdotV, float LdotH) {
    const float energyBias = 0.5f * sd.linearRoughness;
    const float energyFactor = 1.f + (1.f / 1.51f - 1.f) * sd.linearRoughness;
    const float fd90 = energyBias + 2.f * LdotH * LdotH * sd.linearRoughness;
    const float lightScatter = fresnelSchlick1(1.f, fd90, NdotL);
    const float viewScatter = fresnelSchlick1(1.f, fd90, NdotV);
    return sd.diffuse * (viewScatter * lightScatter * energyFactor * (1.f / kPiF));
}
// lobes: 1 diffuse, 2 specular
VR_DEV v3 evalBSDFLobes(const ShadingData& sd, v3 L, int lobes) {
    float NdotL = dot3(sd.N, L);
    if (fminf(sd.NdotV, NdotL) < kMinCosTheta) return V3(0.f);
    const v3 H = norm3(sd.V + L);
    const float NdotH = __saturatef(dot3(sd.N, H));
    const float LdotH = __saturatef(dot3(L, H));
    const float NdotV = __saturatef(sd.NdotV);
    NdotL = __saturatef(NdotL);
    v3 f = V3(0.f);
    if (lobes & 1) f = f + diffuseFrostbite(sd, NdotL, NdotV, LdotH);
    if (lobes & 2) {
        const float alpha = fmaxf(kMinGGXAlpha, sd.ggxAlpha);
        const float D = ndfGGX(alpha, NdotH);
        const float G = smithGGXCorrelated(alpha, NdotL, NdotV);
        f = f + fresnelSchlick(sd.specular, 1.f, LdotH) * (D * G);
    }
    return f;
}
VR_DEV v3 evalBSDFCosine(const ShadingData& sd, v3 L) { return evalBSDFLobes(sd, L, 3) * __saturatef(dot3(sd.N, L)); }
VR_DEV float pdfDiffuse(const ShadingData& sd, v3 L) {
    const float NdotL = dot3(sd.N, L);
    return NdotL < kMinCosTheta ? 0.f : NdotL * (1.f / kPiF);
}
VR_DEV float pdfSpecular(const ShadingData& sd, v3 L) {
    const float NdotL = dot3(sd.N, L);
    if (fminf(sd.NdotV, NdotL) < kMinCosTheta) return 0.f;
    const v3 H = norm3(sd.V + L);
    const float NdotH = __saturatef(dot3(sd.N, H));
    const float VdotH = __saturatef(dot3(sd.V, H));
    const float alpha = fmaxf(kMinGGXAlpha, sd.ggxAlpha);
    return ndfGGX(alpha, NdotH) * NdotH / (4.f * VdotH);
}
VR_DEV float diffuseProbability(const ShadingData& sd) {
    const float dw = (1.f - sd.metallic) * lum(sd.diffuse);
    const float sw = lum(fresnelSchlick(sd.specular, 1.f, dot3(sd.V, sd.N)));
    return (dw + sw) > 0.f ? dw / (dw + sw) : 1.f;
}
VR_DEV float evalPdfBSDF(const ShadingData& sd, v3 L) {
    const float pd = diffuseProbability(sd);
    return pd * pdfDiffuse(sd, L) + (1.f - pd) * pdfSpecular(sd, L);
}
struct BSDFSample { v3 wi, weight; float pdf; int lobe; };
// sample_disk_concentric / sample_cosine_hemisphere_concentric (UT/Math/MathHelpers.slang:293-322)
VR_DEV v3 cosineHemisphereConcentric(float ux, float uy, float& pdf) {
    float x = 2.f * ux - 1.f, y = 2.f * uy - 1.f;
    float dx = 0.f, dy = 0.f;
    if (!(x == 0.f && y == 0.f)) {
        float r, phi;
        if (fabsf(x) > fabsf(y)) { r = x; phi = (y / x) * (kPiF / 4.f); }
        else { r = y; phi = kPiF / 2.f - (x / y) * (kPiF / 4.f); }
        dx = r * cosf(phi);
        dy = r * sinf(phi);
    }
    const float z = sqrtf(fmaxf(0.f, 1.f - (dx * dx + dy * dy)));
    pdf = z * (1.f / kPiF);
    return V3(dx, dy, z);
}
// sampleBSDF (MaterialShading.slang:566-611): u = sampleNext2D, then uSelect.
VR_DEV bool sampleBSDF(const ShadingData& sd, Rng& rng, BSDFSample& r) {
    const float ux = u01(rng);
    const float uy = u01(rng);
    const float uSelect = u01(rng);
    const float pd = diffuseProbability(sd);
    float pmf, pdfOther;
    r.weight = V3(0.f);
    if (uSelect < pd) {
        float pdf;
        r.wi = fromLocal(cosineHemisphereConcentric(ux, uy, pdf), sd);
        r.pdf = pdf;
        r.lobe = 1;
        const float NdotL = dot3(sd.N, r.wi);
        if (fminf(sd.NdotV, NdotL) < kMinCosTheta) { r.pdf = 0.f; }
        else r.weight = evalBSDFLobes(sd, r.wi, 1) * kPiF;
        pmf = pd;
        pdfOther = pdfSpecular(sd, r.wi);
    } else {
        const float alpha = fmaxf(kMinGGXAlpha, sd.ggxAlpha);
        const float a2 = alpha * alpha;
        const float cos2 = fminf((1.f - ux) / ((a2 - 1.f) * ux + 1.f), 1.f);
        const float cosH = sqrtf(cos2), sinH = sqrtf(1.f - cos2);
        const float phi = uy * 2.f * kPiF;
        const v3 H = fromLocal(V3(sinH * cosf(phi), sinH * sinf(phi), cosH), sd);
        float VdotH = dot3(sd.V, H);
        r.wi = norm3(2.f * VdotH * H - sd.V);
        const float d = (a2 - 1.f) * cos2 + 1.f;
        r.pdf = (a2 * cosH) / (d * d * VdotH * 4.f * kPiF);
        r.lobe = 2;
        float NdotL = dot3(sd.N, r.wi);
        if (NdotL < kMinCosTheta) r.pdf = 0.f;
        if (fminf(sd.NdotV, NdotL) < kMinCosTheta) {
            r.pdf = 0.f;
        } else {
            const float NdotV = __saturatef(sd.NdotV);
            NdotL = __saturatef(NdotL);
            const float NdotH = __saturatef(cosH);
            VdotH = __saturatef(VdotH);
            const float G = smithGGXCorrelated(alpha, NdotL, NdotV);
            r.weight = fresnelSchlick(sd.specular, 1.f, VdotH) * (G * NdotL * VdotH * 4.f) / NdotH;
        }
        pmf = 1.f - pd;
        pdfOther = pdfDiffuse(sd, r.wi);
    }
    r.pdf = pmf * r.pdf + (1.f - pmf) * pdfOther;
    r.weight = r.weight / pmf;
    return r.pdf > 0.f;
}

}  // namespace vr
