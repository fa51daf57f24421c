// oracle_pass.hpp -- TEST INFRASTRUCTURE ONLY (the CPU oracle).
//
// Lights, camera, reservoirs, target function and the six kernels of the
// VolumetricReSTIR frame, restated one function per reference function:
//   VR/VolumeUtils.slang:12-149, 454-492       sampleSceneLights, SampleDirectLighting
//   VR/Reservoir.slang:8-57                    reservoir init / WRS update
//   VR/ReSTIRHelper.slang:10-607               packing, evaluate_F_, p-hat helpers
//   VR/ComputeInitialSample.slang:4-395        candidate construction
//   VR/TraceRays.cs.slang, TemporalReuse.cs.slang, SpatialReuse.cs.slang,
//   CopyReservoirs.cs.slang, FinalShading.cs.slang, GenerateFeatures.cs.slang
//   VR/VolumePathTracingFunctions.slang:3-132  reference path tracer
#pragma once
#include <vector>

#include "oracle_tracking.hpp"

namespace orc {

// SamplingOptions (VR/HostDeviceSharedDefinitions.h:100-138).
struct SamplingOptions {
    uint32_t visibilityTrackingMethod, lightingTrackingMethod;
    int lightSamples, lightingMipLevel, visibilitySamples, visibilityMipLevel;
    bool visibilityUseLinearSampler, lightingUseLinearSampler;
    float visibilityTStepScale, lightingTStepScale;
    bool useEnv, useAnalytic, useEmissive;
    int vertexReuseStartBounce;
};

// Reservoir (VR/HostDeviceSharedDefinitions.h:16-45) with the b>1 extraBounceStartId.
struct Reservoir {
    float runningSum, M, depth, p_y;
    f2 lightUV;
    int lightID, sampledPixel, extraBounceStartId;
};
static_assert(sizeof(Reservoir) == 36, "exchange layout");
struct Features { int noReflectiveSurface; float transmittance; };

struct PointLight { f3 pos, intensity; };

struct Camera {
    f3 pos, U, V, W;
    float view[16], proj[16];  // row-vector matrices: clip = [p,1] * view * proj
};

struct SceneState {
    Medium med;
    std::vector<PointLight> lights;
    f3 env{0, 0, 0};
    Camera cam, prevCam;
};

// ---------------------------------------------------------------- camera
// UVW frame (SC/Camera/Camera.cpp:183-189) and glm-style lookAt/perspective,
// both in double and rounded once.
static inline void buildCamera(Camera& c, const float pos[3], const float target[3], const float up[3], float fovY,
                               float aspect, float nearZ, float farZ) {
    double P[3] = {pos[0], pos[1], pos[2]}, T[3] = {target[0], target[1], target[2]}, Up[3] = {up[0], up[1], up[2]};
    auto nrm = [](double* v) { double l = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); v[0] /= l; v[1] /= l; v[2] /= l; };
    auto crs = [](const double* a, const double* b, double* o) {
        o[0] = a[1] * b[2] - a[2] * b[1]; o[1] = a[2] * b[0] - a[0] * b[2]; o[2] = a[0] * b[1] - a[1] * b[0];
    };
    double w[3] = {T[0] - P[0], T[1] - P[1], T[2] - P[2]};
    nrm(w);
    double u[3], v[3];
    crs(w, Up, u); nrm(u);
    crs(u, w, v); nrm(v);
    double th = std::tan((double)fovY * 0.5);
    double ulen = th * aspect, vlen = th;
    c.pos = {pos[0], pos[1], pos[2]};
    c.W = {(float)w[0], (float)w[1], (float)w[2]};
    c.U = {(float)(u[0] * ulen), (float)(u[1] * ulen), (float)(u[2] * ulen)};
    c.V = {(float)(v[0] * vlen), (float)(v[1] * vlen), (float)(v[2] * vlen)};
    // view (row-vector form of glm::lookAt)
    double view[16] = {u[0], v[0], -w[0], 0, u[1], v[1], -w[1], 0, u[2], v[2], -w[2], 0,
                       -(u[0] * P[0] + u[1] * P[1] + u[2] * P[2]), -(v[0] * P[0] + v[1] * P[1] + v[2] * P[2]),
                       (w[0] * P[0] + w[1] * P[1] + w[2] * P[2]), 1};
    double proj[16] = {0};
    proj[0] = 1.0 / (aspect * th);
    proj[5] = 1.0 / th;
    proj[10] = farZ / (nearZ - farZ);
    proj[11] = -1.0;
    proj[14] = -(farZ * nearZ) / (farZ - nearZ);
    for (int i = 0; i < 16; i++) { c.view[i] = (float)view[i]; c.proj[i] = (float)proj[i]; }
}
// computeNonNormalizedRayDirPinhole (SC/Camera/Camera.slang:172-182), no jitter.
static inline f3 rayDirNN(const Camera& c, int px, int py, int W, int H) {
    float p0 = ((float)px + 0.5f) / (float)W, p1 = ((float)py + 0.5f) / (float)H;
    float nx = 2.f * p0 + -1.f, ny = -2.f * p1 + 1.f;
    return nx * c.U + ny * c.V + c.W;
}
static inline f3 rayDir(const Camera& c, int px, int py, int W, int H) { return normalize(rayDirNN(c, px, py, W, H)); }

// ---------------------------------------------------------------- env + lights
// oct_to_ndir_equal_area_unorm (UT/Math/MathHelpers.slang:200-219).
static inline float sgn(float x) { return x > 0.f ? 1.f : (x < 0.f ? -1.f : 0.f); }
static inline f3 octToDirEqualArea(f2 p) {
    p.x = p.x * 2.f - 1.f; p.y = p.y * 2.f - 1.f;
    float d = 1.f - (std::fabs(p.x) + std::fabs(p.y));
    float r = 1.f - std::fabs(d);
    float phi = (r > 0.f) ? ((std::fabs(p.y) - std::fabs(p.x)) / r + 1.f) * (kPi / 4.f) : 0.f;
    float f = r * std::sqrt(2.f - r * r);
    return {f * sgn(p.x) * std::cos(phi), f * sgn(p.y) * std::sin(phi), sgn(d) * (1.f - r * r)};
}
// EnvMapSampler::encodeLightUV / decodeLightUV (LT/EnvMapSampler.slang:75-90); env rotation = identity.
static inline f2 encodeLightUV(f3 dir, int& id) { id = dir.z < 0 ? -2 : -1; return {dir.x, dir.y}; }
static inline f3 decodeLightUV(f2 uv, int id) {
    f3 d{uv.x, uv.y, 0};
    d.z = std::sqrt(1 - uv.x * uv.x - uv.y * uv.y);
    if (std::isnan(d.z)) d.z = 0.f;
    if (id == -2) d.z = -d.z;
    return d;
}

struct SceneLightSample { f3 dir; float distance; f3 Li; float pdf, pdfArea; f3 rayDir; float rayDistance; };

// sampleSceneLights (VR/VolumeUtils.slang:12-149). The environment is constant:
// the hierarchical warp of LT/EnvMapSampler.slang:94-167 through a constant
// importance map is the identity on [0,1)^2 and its pdf is 1/(4 pi).
static inline bool sampleSceneLights(const SceneState& S, f3 origin, bool useEnv, bool useAnalytic, bool useEmissive,
                                     SampleGen& sg, SceneLightSample& ls, int& lightIndex, f2& lightUV) {
    if (!useEnv && !useAnalytic && !useEmissive) return false;
    float p[3] = {useEnv ? 1.f : 0.f, useAnalytic ? 1.f : 0.f, useEmissive ? 1.f : 0.f};
    float sum = p[0] + p[1] + p[2];
    if (sum == 0.f) return false;
    float inv = 1.f / sum;
    p[0] *= inv; p[1] *= inv; p[2] *= inv;
    float u = next1D(sg);
    if (useEnv) {
        if (u < p[0]) {
            float sel = p[0];
            f2 rnd;
            rnd.x = next1D(sg);
            rnd.y = next1D(sg);
            f3 dir = octToDirEqualArea(rnd);
            float lpdf = 1.f * kInv4Pi;
            f3 Le = S.env;
            float pdf = sel * lpdf;
            ls.rayDir = ls.dir = dir;
            ls.rayDistance = ls.distance = kRayTMax;
            ls.pdf = pdf;
            ls.pdfArea = pdf;
            ls.Li = pdf > 0.f ? Le / pdf : mk3(0.f);
            lightIndex = -1;
            lightUV = encodeLightUV(ls.rayDir, lightIndex);
            return !(std::isnan(dir.x) || std::isnan(dir.y) || std::isnan(dir.z));
        }
        u -= p[0];
    }
    if (useAnalytic) {
        if (u < p[1]) {
            u /= p[1];
            uint32_t count = (uint32_t)S.lights.size();
            uint32_t idx = std::min((uint32_t)(u * (float)count), count - 1);
            float sel = p[1] / (float)count;
            const PointLight& L = S.lights[idx];
            f3 toL = L.pos - origin;
            float d2 = std::max(dot(toL, toL), 1e-9f);
            float dist = std::sqrt(d2);
            f3 dir = toL / dist;
            f3 Li = L.intensity / d2;
            float lpdf = 1.f;  // Dirac pdf 0 is replaced by 1 (VR/VolumeUtils.slang:93)
            lightIndex = (int)idx;
            ls.rayDir = ls.dir = dir;
            ls.rayDistance = ls.distance = dist;
            ls.pdf = sel * lpdf;
            ls.pdfArea = ls.pdf;
            ls.Li = Li / ls.pdf;
            lightUV = {0, 0};
            return true;
        }
        u -= p[1];
    }
    return false;
}

// getAnalyticalLightSample (VR/VolumeUtils.slang:151-168) for a point light.
static inline SceneLightSample analyticLightSample(const SceneState& S, int id, f3 origin) {
    const PointLight& L = S.lights[id];
    f3 toL = L.pos - origin;
    float d2 = std::max(dot(toL, toL), 1e-9f);
    SceneLightSample ls;
    ls.distance = std::sqrt(d2);
    ls.dir = toL / ls.distance;
    ls.rayDir = ls.dir;
    ls.rayDistance = ls.distance;
    ls.pdf = 0.f;
    ls.Li = L.intensity / d2;
    return ls;
}

// SampleDirectLighting (VR/VolumeUtils.slang:454-492).
static inline f3 sampleDirectLighting(const SceneState& S, SampleGen& sg, float& pdf, const MediumInteraction& mi,
                                      bool useEnv, bool useAnalytic, bool useEmissive, bool enableShadow, int mip,
                                      uint32_t method, bool linear, int visSamples, float tStepScale, int& lightIndex,
                                      f2& lightUV, float& outVis) {
    pdf = 0.f;
    SceneLightSample ls{};
    bool valid = sampleSceneLights(S, mi.p, useEnv, useAnalytic, useEmissive, sg, ls, lightIndex, lightUV);
    pdf = ls.pdfArea;
    if (!valid) { pdf = 0.f; return mk3(0.f); }
    ls.pdf = 1.f;
    Ray sh{mi.p, ls.rayDir, 0, ls.rayDistance};
    if (enableShadow) {
        float vis = computeVisibility(S.med, sh, sg, visSamples, mip, linear, method, tStepScale);
        outVis = vis;
        ls.Li = ls.Li * vis;
    }
    float p = phaseHG(dot(mi.wo, ls.dir), mi.g);
    return mk3(p) * ls.Li / ls.pdf;
}

// directLighting (VR/VolumeUtils.slang:416-450), used by the path tracer.
static inline f3 directLighting(const SceneState& S, SampleGen& sg, const MediumInteraction& mi, int nSamples,
                                bool useEnv, bool useAnalytic, bool useEmissive, int mip, uint32_t method) {
    f3 Ld = mk3(0.f);
    for (int i = 0; i < nSamples; i++) {
        SceneLightSample ls{};
        int idx; f2 uv;
        bool valid = sampleSceneLights(S, mi.p, useEnv, useAnalytic, useEmissive, sg, ls, idx, uv);
        if (!valid) continue;
        ls.pdf = 1.f;
        Ray sh{mi.p, ls.rayDir, 0, ls.rayDistance};
        float vis = computeVisibility(S.med, sh, sg, 1, mip, true, method, 1.f);
        ls.Li = ls.Li * vis;
        float p = phaseHG(dot(mi.wo, ls.dir), mi.g);
        Ld = Ld + mk3(p) * ls.Li / ls.pdf;
    }
    return Ld / (float)nSamples;
}

// ---------------------------------------------------------------- reservoirs (VR/Reservoir.slang:8-87)
static inline Reservoir newReservoir() { return {0.f, 0.f, kRayTMax, 0.f, {0, 0}, 0, 0, 0}; }
static inline bool resampleStep(const Reservoir& r, Reservoir& st, SampleGen& sg) {
    float w = r.runningSum;
    st.M += r.M;
    if (w <= 0.0f) return false;
    st.runningSum += w;
    bool sel = next1D(sg) * st.runningSum < w;
    if (sel) {
        st.depth = r.depth; st.p_y = r.p_y; st.lightUV = r.lightUV; st.lightID = r.lightID;
        st.extraBounceStartId = r.extraBounceStartId; st.sampledPixel = r.sampledPixel;
    }
    return sel;
}
static inline bool resampleStepMaxM(const Reservoir& r, float cap, Reservoir& st, SampleGen& sg) {
    float cM = std::min(cap, r.M);
    float w = cM == 0.0f ? 0.0f : cM / r.M * r.runningSum;
    st.M += cM;
    if (w <= 0.0f) return false;
    st.runningSum += w;
    bool sel = next1D(sg) * st.runningSum < w;
    if (sel) {
        st.depth = r.depth; st.p_y = r.p_y; st.lightUV = r.lightUV; st.lightID = r.lightID;
        st.extraBounceStartId = r.extraBounceStartId; st.sampledPixel = r.sampledPixel;
    }
    return sel;
}

// bit packing (VR/ReSTIRHelper.slang:10-89)
static inline int decodeMaxIndirect(int s) { return s >> 20; }
static inline int encodeMaxIndirect(int s, int b) { return (b << 20) | (s & 0xFFFFF); }
static inline int decodePathTag(int s) { return (s >> 16) & 0xF; }
static inline int encodePathTag(int s, int tag) { return (tag << 16) | (int)((uint32_t)s & 0xFFF0FFFFu); }
static inline f3 decodeEmissivePosition(int id, f2 uv) { return {uv.x, uv.y, asfloat(id)}; }
static inline void encodeEmissivePosition(f3 p, int& id, f2& uv) { id = asint(p.z); uv = {p.x, p.y}; }
static inline f4 decodeWiDist(f3 in) {
    f3 wi{in.x, in.y, std::sqrt(1 - (in.x * in.x + in.y * in.y))};
    if (std::isnan(wi.z)) wi.z = 0.f;
    if (in.z < 0) { wi.z = -wi.z; in.z = -in.z; }
    return {wi.x, wi.y, wi.z, in.z};
}
static inline f3 encodeWiDist(f4 in) {
    f3 wi{in.x, in.y, in.w};
    if (in.z < 0) wi.z = -wi.z;
    return wi;
}

}  // namespace orc
