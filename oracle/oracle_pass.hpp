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

// Orientation of the environment map for one frame (EnvMapData transform /
// invTransform): world = toWorld * local, 3x3 row-major.
struct EnvXf { float toWorld[9], toLocal[9]; };

// A textured lat-long environment map and its importance mip chain
// (LT/EnvMap.cpp, LT/EnvMapSampler.cpp:61-110).
struct EnvMapState {
    bool textured = false;
    int w = 0, h = 0;
    std::vector<float> rgb;                // w * h RGB, row 0 at v = 0
    f3 scale{1, 1, 1};                     // intensity * tint
    int impDim = 0, impBaseMip = 0;
    std::vector<std::vector<float>> imp;   // level k: (impDim >> k)^2 texels, row-major
};

// Emissive triangle of the light collection (LT/LightCollectionShared.slang:41-60),
// with its face normal as the collection stores it, and an alias-table entry
// (UT/Sampling/AliasTable.slang:34-45).
struct EmissiveTriangle { f3 p[3]; f3 normal; float area; f3 Le; };
struct AliasItem { float threshold; uint32_t indexA, indexB; };
// A scene triangle (positions, vertex shading normals, material) and an untextured
// metal-rough material.
struct SurfTri { f3 p[3]; f3 n[3]; int material; };
struct SurfMaterial { f3 baseColor; float roughness, metallic, IoR; f3 emissive; bool doubleSided; };
struct VItem4 { int tri; float bx, by, t; };  // the V-buffer entry of the exchange format

struct SceneState {
    Medium med;
    std::vector<PointLight> lights;
    std::vector<EmissiveTriangle> emTris;  // the light collection, in scene order
    std::vector<float> emWeight;           // flux per collection triangle
    std::vector<AliasItem> emAlias;        // EmissivePowerSampler's table over emWeight
    int emActive = 0;                      // triangles with positive flux
    float emWeightSum = 0.f;
    float emMult = 1.f;                    // Scene::emissiveIntensityMultiplier
    std::vector<SurfTri> surfTris;         // every scene triangle, for ray queries
    std::vector<SurfMaterial> surfMats;
    f3 env{0, 0, 0};                       // constant environment (envMap.textured == false)
    EnvMapState envMap;
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
// EnvMap::setRotation (LT/EnvMap.cpp:57-72): toWorld = Rz * Ry * Rx with the
// right-handed axis rotations of glm::eulerAngleX/Y/Z, in double, rounded once;
// toLocal is its transpose (the inverse of a rotation).
static inline EnvXf envRotation(const float deg[3]) {
    const double k = 3.14159265358979323846 / 180.0;
    double a = deg[0] * k, b = deg[1] * k, c = deg[2] * k;
    // Rz(c) * Ry(b) * Rx(a), multiplied out
    double ca = std::cos(a), sa = std::sin(a), cb = std::cos(b), sb = std::sin(b), cc = std::cos(c), sc = std::sin(c);
    double R[9] = {cc * cb, cc * sb * sa - sc * ca, cc * sb * ca + sc * sa,
                   sc * cb, sc * sb * sa + cc * ca, sc * sb * ca - cc * sa,
                   -sb,     cb * sa,                cb * ca};
    EnvXf X;
    for (int r = 0; r < 3; r++)
        for (int q = 0; q < 3; q++) {
            X.toWorld[r * 3 + q] = (float)R[r * 3 + q];
            X.toLocal[r * 3 + q] = (float)R[q * 3 + r];
        }
    return X;
}
static inline f3 mat3v(const float* M, f3 d) {
    return {M[0] * d.x + M[1] * d.y + M[2] * d.z, M[3] * d.x + M[4] * d.y + M[5] * d.z,
            M[6] * d.x + M[7] * d.y + M[8] * d.z};
}
// world_to_latlong_map (UT/Math/MathHelpers.slang:92-99).
static inline f2 worldToLatlong(f3 dir) {
    f3 p = normalize(dir);
    float py = std::min(std::max(p.y, -1.f), 1.f);
    return {std::atan2(p.x, -p.z) * (0.5f / kPi) + 0.5f, std::acos(py) * (1.f / kPi)};
}
// ndir_to_oct_equal_area_unorm (UT/Math/MathHelpers.slang:174-193).
static inline f2 dirToOctEqualArea(f3 n) {
    float r = std::sqrt(1.f - std::fabs(n.z));
    float phi = std::atan2(std::fabs(n.y), std::fabs(n.x));
    f2 p;
    p.y = r * phi * (2.f / kPi);
    p.x = r - p.y;
    if (n.z < 0.f) p = {1.f - p.y, 1.f - p.x};
    p.x *= sgn(n.x);
    p.y *= sgn(n.y);
    return {p.x * 0.5f + 0.5f, p.y * 0.5f + 0.5f};
}
// Texture2D SampleLevel(lod 0) with a linear filter, wrap in u and clamp in v
// (LT/EnvMap.cpp:113-121): bilinear over texel centres.
static inline f3 envTexel(const EnvMapState& E, int x, int y) {
    x = ((x % E.w) + E.w) % E.w;
    y = std::min(std::max(y, 0), E.h - 1);
    const float* t = &E.rgb[3 * ((size_t)y * E.w + x)];
    return {t[0], t[1], t[2]};
}
static inline f3 envBilinear(const EnvMapState& E, f2 uv) {
    float x = uv.x * (float)E.w - 0.5f, y = uv.y * (float)E.h - 0.5f;
    if (!(std::fabs(x) < 1e30f) || !(std::fabs(y) < 1e30f)) return mk3(0.f);
    float x0 = std::floor(x), y0 = std::floor(y);
    float fx = x - x0, fy = y - y0;
    long xi = (long)std::fmod(x0, (float)E.w);
    int yi = (int)std::min(std::max(y0, -1.f), (float)E.h);
    f3 a = envTexel(E, (int)xi, yi), b = envTexel(E, (int)xi + 1, yi);
    f3 c = envTexel(E, (int)xi, yi + 1), d = envTexel(E, (int)xi + 1, yi + 1);
    f3 top = a + (b - a) * fx, bot = c + (d - c) * fx;
    return top + (bot - top) * fy;
}
// EnvMap::eval (LT/EnvMap.slang:45-50) for a textured map, the constant otherwise.
static inline f3 envEval(const SceneState& S, const EnvXf& X, f3 dWorld) {
    if (!S.envMap.textured) return S.env;
    return S.envMap.scale * envBilinear(S.envMap, worldToLatlong(mat3v(X.toLocal, dWorld)));
}
// EnvMapSampler::encodeLightUV / decodeLightUV (LT/EnvMapSampler.slang:75-90): the
// local direction's xy, the sign of its z in the light id. A constant environment
// has no orientation (local = world).
static inline f2 encodeLightUV(f3 dir, int& id) { id = dir.z < 0 ? -2 : -1; return {dir.x, dir.y}; }
static inline f3 decodeLightUV(f2 uv, int id) {
    f3 d{uv.x, uv.y, 0};
    d.z = std::sqrt(1 - uv.x * uv.x - uv.y * uv.y);
    if (std::isnan(d.z)) d.z = 0.f;
    if (id == -2) d.z = -d.z;
    return d;
}
static inline f2 encodeLightUV(const SceneState& S, const EnvXf& X, f3 dWorld, int& id) {
    return encodeLightUV(S.envMap.textured ? mat3v(X.toLocal, dWorld) : dWorld, id);
}
static inline f3 decodeLightUV(const SceneState& S, const EnvXf& X, f2 uv, int id) {
    f3 d = decodeLightUV(uv, id);
    return S.envMap.textured ? mat3v(X.toWorld, d) : d;
}
// EnvMapSampler::createImportanceMap + EnvMapSamplerSetup.cs.slang:48-70: the base
// level averages the luminance of spp stratified samples per texel (taken at the
// equal-area octahedral directions), each further level the 2x2 blocks below it.
static inline void buildImportance(EnvMapState& E, int dim, int spp) {
    int sx = std::max(1, (int)std::sqrt((double)spp));
    int sy = spp / sx;
    E.impDim = dim;
    E.imp.clear();
    std::vector<float> base((size_t)dim * dim);
    for (int y = 0; y < dim; y++)
        for (int x = 0; x < dim; x++) {
            float L = 0.f;
            for (int j = 0; j < sy; j++)
                for (int i = 0; i < sx; i++) {
                    f2 p{((float)(x * sx + i) + 0.5f) / (float)(dim * sx), ((float)(y * sy + j) + 0.5f) / (float)(dim * sy)};
                    L += luminance(envBilinear(E, worldToLatlong(octToDirEqualArea(p))));
                }
            base[(size_t)y * dim + x] = L * (1.f / (float)(sx * sy));
        }
    E.imp.push_back(std::move(base));
    for (int d = dim / 2; d >= 1; d /= 2) {
        const std::vector<float>& s = E.imp.back();
        std::vector<float> n((size_t)d * d);
        for (int y = 0; y < d; y++)
            for (int x = 0; x < d; x++) {
                const size_t o = (size_t)(2 * y) * (2 * d) + 2 * x;
                n[(size_t)y * d + x] = ((s[o] + s[o + 1]) + (s[o + 2 * d] + s[o + 2 * d + 1])) * 0.25f;
            }
        E.imp.push_back(std::move(n));
    }
    E.impBaseMip = (int)E.imp.size() - 1;
}
// EnvMapSampler::sample (LT/EnvMapSampler.slang:94-146). Returns the world direction
// and its solid-angle pdf; a constant environment is sampled uniformly.
static inline f3 envSample(const SceneState& S, const EnvXf& X, f2 p, float& pdf) {
    if (!S.envMap.textured) {
        pdf = 1.f * kInv4Pi;
        return octToDirEqualArea(p);
    }
    const EnvMapState& E = S.envMap;
    int px = 0, py = 0;
    for (int mip = E.impBaseMip - 1; mip >= 0; mip--) {
        px *= 2; py *= 2;
        const std::vector<float>& L = E.imp[mip];
        const int d = E.impDim >> mip;
        float w[4] = {L[(size_t)py * d + px], L[(size_t)py * d + px + 1], L[(size_t)(py + 1) * d + px],
                      L[(size_t)(py + 1) * d + px + 1]};
        float q[2] = {w[0] + w[2], w[1] + w[3]};
        int ox, oy;
        float dd = q[0] / (q[0] + q[1]);
        if (p.x < dd) { ox = 0; p.x = p.x / dd; }
        else { ox = 1; p.x = (p.x - dd) / (1.f - dd); }
        float e = w[ox] / q[ox];
        if (p.y < e) { oy = 0; p.y = p.y / e; }
        else { oy = 1; p.y = (p.y - e) / (1.f - e); }
        px += ox; py += oy;
    }
    const float inv = 1.f / (float)E.impDim;
    f2 uv{((float)px + p.x) * inv, ((float)py + p.y) * inv};
    f3 dir = octToDirEqualArea(uv);
    float avg = E.imp[E.impBaseMip][0];
    pdf = E.imp[0][(size_t)py * E.impDim + px] / avg * kInv4Pi;
    return mat3v(X.toWorld, dir);
}
// EnvMapSampler::evalPdf (LT/EnvMapSampler.slang:148-160): point lookup (clamp)
// of the base level at the direction's octahedral coordinate.
static inline float envEvalPdf(const SceneState& S, const EnvXf& X, f3 dWorld) {
    if (!S.envMap.textured) return kInv4Pi;
    const EnvMapState& E = S.envMap;
    f2 uv = dirToOctEqualArea(mat3v(X.toLocal, dWorld));
    int x = std::min(std::max((int)std::floor(uv.x * E.impDim), 0), E.impDim - 1);
    int y = std::min(std::max((int)std::floor(uv.y * E.impDim), 0), E.impDim - 1);
    return E.imp[0][(size_t)y * E.impDim + x] / E.imp[E.impBaseMip][0] * kInv4Pi;
}

// ---------------------------------------------------------------- volume emission
// TemperatureWorldSpace + ConvertTempToColor + EmissionWorldSpace
// (VR/VolumeBase.slang:196-232): trilinear temperature lookup in the temperature
// grid, T' = min(6400, (T - cutoff) * scale), and the blackbody table as a 1-D
// texture over u = (T' - 25) / 6400 with a linear filter and border colour 0.
static inline float temperatureWorld(const Medium& m, f3 pW, int slot = kTemperatureSlot) {
    return valueAtPoint(m.grid[slot], worldToMediumP(m, pW, slot), true);
}
static inline f3 blackbody(const Medium& m, float temp) {
    temp = std::min(6400.f, (temp - m.tempCutoff) * m.tempScale);
    float q = (temp - 25.f) / 6400.f;
    const int N = (int)m.lut.size();
    float x = q * (float)N - 0.5f;
    if (!(x > -1.f && x < (float)N)) return mk3(0.f);
    float x0 = std::floor(x);
    float f = x - x0;
    int i0 = (int)x0, i1 = i0 + 1;
    f3 a = i0 >= 0 ? m.lut[i0] : mk3(0.f), b = i1 < N ? m.lut[i1] : mk3(0.f);
    return m.LeScale * (a + (b - a) * f);
}
// last: the previous frame's temperature grid (useLastFrameGrid).
static inline f3 emissionWorld(const Medium& m, f3 pW, bool last = false) {
    if (last ? !m.prevHasEmission : !m.hasEmission) return mk3(0.f);
    return blackbody(m, temperatureWorld(m, pW, last ? kTemperatureSlot + kPrevExtraOffset : kTemperatureSlot));
}

// VelocityWorld (VR/VolumeBase.slang:183-194) on the dense field: trilinear over
// voxel centres with zero outside the grid, zero in a brick with no velocity.
static inline f3 velocityWorld(const Medium& m, f3 pW) {
    const VelocityField& V = m.velocity;
    if (!V.valid) return mk3(0.f);
    f4 q = mulRow({pW.x, pW.y, pW.z, 1.f}, V.w2i);
    f3 p{q.x / q.w, q.y / q.w, q.z / q.w};
    const int bnx = (V.nx + 7) / 8, bny = (V.ny + 7) / 8, bnz = (V.nz + 7) / 8;
    if (!(p.x >= 0.f && p.y >= 0.f && p.z >= 0.f) || !(p.x < bnx * 8.f && p.y < bny * 8.f && p.z < bnz * 8.f))
        return mk3(0.f);
    int bx = (int)p.x / 8, by = (int)p.y / 8, bz = (int)p.z / 8;
    if (!V.brickOcc[(size_t)bx + (size_t)bnx * ((size_t)by + (size_t)bny * bz)]) return mk3(0.f);
    auto at = [&](int x, int y, int z, int c) -> float {
        if (x < 0 || y < 0 || z < 0 || x >= V.nx || y >= V.ny || z >= V.nz) return 0.f;
        return V.v[3 * ((size_t)x + (size_t)V.nx * ((size_t)y + (size_t)V.ny * z)) + c];
    };
    f3 x = p - 0.5f;
    f3 i0 = f3floor(x);
    f3 f = x - i0;
    int ix = (int)i0.x, iy = (int)i0.y, iz = (int)i0.z;
    float out[3];
    for (int c = 0; c < 3; c++) {
        float v000 = at(ix, iy, iz, c), v100 = at(ix + 1, iy, iz, c);
        float v010 = at(ix, iy + 1, iz, c), v110 = at(ix + 1, iy + 1, iz, c);
        float v001 = at(ix, iy, iz + 1, c), v101 = at(ix + 1, iy, iz + 1, c);
        float v011 = at(ix, iy + 1, iz + 1, c), v111 = at(ix + 1, iy + 1, iz + 1, c);
        float a0 = v000 + (v100 - v000) * f.x, a1 = v010 + (v110 - v010) * f.x;
        float a2 = v001 + (v101 - v001) * f.x, a3 = v011 + (v111 - v011) * f.x;
        float b0 = a0 + (a1 - a0) * f.y, b1 = a2 + (a3 - a2) * f.y;
        out[c] = b0 + (b1 - b0) * f.z;
    }
    return {out[0], out[1], out[2]};
}

struct SceneLightSample { f3 dir; float distance; f3 Li; float pdf, pdfArea; f3 rayDir; float rayDistance; };

// ---------------------------------------------------------------- ray queries
// FindSurfaceHit / FindIfOccluded (VR/InlineRayTracingHelpers.slang:42-155) by brute
// force over every triangle, with the watertight test of Woop, Benthin and Wald
// (JCGT 2(1), 2013). Closest hit: smallest t, ties to the smallest triangle index.
struct HitInfo { int tri = -1; float t = kRayTMax; f2 bary{0.f, 0.f}; };
static inline float axisOf(f3 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }
static inline bool watertightHit(f3 o, f3 d, const SurfTri& T, float tMin, float tMax, float& t, f2& bary) {
    float ad[3] = {std::fabs(d.x), std::fabs(d.y), std::fabs(d.z)};
    // the axis of the largest |d| component (ties go to the later axis)
    int kz = (ad[0] > ad[1]) ? (ad[0] > ad[2] ? 0 : 2) : (ad[1] > ad[2] ? 1 : 2);
    int kx = (kz + 1) % 3, ky = (kx + 1) % 3;
    if (axisOf(d, kz) < 0.f) std::swap(kx, ky);
    const float Sx = axisOf(d, kx) / axisOf(d, kz), Sy = axisOf(d, ky) / axisOf(d, kz), Sz = 1.f / axisOf(d, kz);
    const f3 A = T.p[0] - o, B = T.p[1] - o, C = T.p[2] - o;
    const float Ax = axisOf(A, kx) - Sx * axisOf(A, kz), Ay = axisOf(A, ky) - Sy * axisOf(A, kz);
    const float Bx = axisOf(B, kx) - Sx * axisOf(B, kz), By = axisOf(B, ky) - Sy * axisOf(B, kz);
    const float Cx = axisOf(C, kx) - Sx * axisOf(C, kz), Cy = axisOf(C, ky) - Sy * axisOf(C, kz);
    float U = Cx * By - Cy * Bx, V = Ax * Cy - Ay * Cx, W = Bx * Ay - By * Ax;
    if (U == 0.f || V == 0.f || W == 0.f) {
        U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
        V = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
        W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
    }
    if ((U < 0.f || V < 0.f || W < 0.f) && (U > 0.f || V > 0.f || W > 0.f)) return false;
    const float det = U + V + W;
    if (det == 0.f) return false;
    const float Az = Sz * axisOf(A, kz), Bz = Sz * axisOf(B, kz), Cz = Sz * axisOf(C, kz);
    const float Tn = U * Az + V * Bz + W * Cz;
    const float inv = 1.f / det;
    t = Tn * inv;
    if (!(t >= tMin && t <= tMax)) return false;
    bary = {V * inv, W * inv};
    return true;
}
static inline HitInfo findSurfaceHit(const SceneState& S, f3 o, f3 d, float tMin, float tMax) {
    HitInfo h;
    for (size_t i = 0; i < S.surfTris.size(); i++) {
        float t;
        f2 b;
        if (!watertightHit(o, d, S.surfTris[i], std::max(0.f, tMin), tMax, t, b)) continue;
        if (h.tri < 0 || t < h.t) { h.tri = (int)i; h.t = t; h.bary = b; }
    }
    return h;
}
static inline bool findIfOccluded(const SceneState& S, f3 o, f3 d, float tMin, float tMax) {
    for (size_t i = 0; i < S.surfTris.size(); i++) {
        float t;
        f2 b;
        if (watertightHit(o, d, S.surfTris[i], std::max(0.f, tMin), tMax, t, b)) return true;
    }
    return false;
}
// ---------------------------------------------------------------- surface shading
// computeSurfaceShadingInfo (VR/InlineRayTracingHelpers.slang:157-173) with the
// metal-rough branch of _prepareShadingData (SC/Shading.slang:155-245), and the
// legacy standard BSDF (Experimental/Scene/Material/MaterialShading.slang:167-631:
// Frostbite diffuse, height-correlated Smith-GGX, Schlick Fresnel). The tangent
// frame is Duff et al.'s orthonormal basis around N (the triangles carry no tangents).
static inline f3 computeRayOrigin(f3 pos, f3 normal);
struct ShadingData {
    f3 posW, V, N, faceN, T, B;
    bool frontFacing;
    float NdotV;
    f3 diffuse, specular, emissive;
    float linearRoughness, ggxAlpha, metallic;
};
static const float kMinGGXAlpha = 0.0064f, kMinCosTheta = 1e-6f;
static inline float saturate(float x) { return std::min(std::max(x, 0.f), 1.f); }
static inline f3 lerp3(f3 a, f3 b, float t) { return a + (b - a) * t; }
static inline ShadingData computeSurfaceShadingInfo(const SceneState& S, int tri, f2 bary, f3 viewDir) {
    const SurfTri& T = S.surfTris[tri];
    const SurfMaterial& M = S.surfMats[T.material];
    float b0 = 1.f - bary.x - bary.y;
    ShadingData sd;
    sd.posW = T.p[0] * b0 + T.p[1] * bary.x + T.p[2] * bary.y;
    sd.faceN = normalize(cross(T.p[1] - T.p[0], T.p[2] - T.p[0]));
    sd.N = normalize(T.n[0] * b0 + T.n[1] * bary.x + T.n[2] * bary.y);
    sd.V = viewDir;
    sd.frontFacing = dot(sd.V, sd.faceN) >= 0.f;
    sd.diffuse = lerp3(M.baseColor, mk3(0.f), M.metallic);
    float f = (M.IoR - 1.f) / (M.IoR + 1.f), F0 = f * f;
    sd.specular = lerp3(mk3(F0), M.baseColor, M.metallic);
    sd.linearRoughness = M.roughness;
    sd.metallic = M.metallic;
    sd.ggxAlpha = std::max(0.0064f, M.roughness * M.roughness);
    sd.emissive = sd.frontFacing ? M.emissive : mk3(0.f);
    sd.NdotV = dot(sd.N, sd.V);
    if (!sd.frontFacing && M.doubleSided) {
        sd.N = -sd.N;
        sd.NdotV = -sd.NdotV;
    }
    // Duff, Burgess, Christensen, Hery, Kensler, Liani, Villemin (JCGT 2017)
    float sign = std::copysign(1.f, sd.N.z);
    float a = -1.f / (sign + sd.N.z), c = sd.N.x * sd.N.y * a;
    sd.T = {1.f + sign * sd.N.x * sd.N.x * a, sign * c, -sign * sd.N.x};
    sd.B = {c, sign + sd.N.y * sd.N.y * a, -sd.N.y};
    sd.posW = computeRayOrigin(sd.posW, sd.frontFacing ? sd.faceN : -sd.faceN);
    return sd;
}
static inline f3 fromLocal(f3 v, const ShadingData& sd) { return sd.T * v.x + sd.B * v.y + sd.N * v.z; }
static inline f3 evalFresnelSchlick(f3 f0, f3 f90, float cosTheta) {
    return f0 + (f90 - f0) * std::pow(std::max(1.f - cosTheta, 0.f), 5.f);
}
static inline float evalNdfGGX(float alpha, float cosTheta) {
    float a2 = alpha * alpha;
    float d = ((cosTheta * a2 - cosTheta) * cosTheta + 1.f);
    return a2 / (d * d * (float)M_PI);
}
static inline float evalMaskingSmithGGXCorrelated(float alpha, float NdotL, float NdotV) {
    float a2 = alpha * alpha;
    float lambdaV = NdotL * std::sqrt((-NdotV * a2 + NdotV) * NdotV + a2);
    float lambdaL = NdotV * std::sqrt((-NdotL * a2 + NdotL) * NdotL + a2);
    return 0.5f / (lambdaV + lambdaL);
}
static inline f3 evalDiffuseFrostbite(const ShadingData& sd, float NdotL, float NdotV, float LdotH) {
    float energyBias = 0.f + (0.5f - 0.f) * sd.linearRoughness;
    float energyFactor = 1.f + (1.f / 1.51f - 1.f) * sd.linearRoughness;
    float fd90 = energyBias + 2.f * LdotH * LdotH * sd.linearRoughness;
    float lightScatter = evalFresnelSchlick(mk3(1.f), mk3(fd90), NdotL).x;
    float viewScatter = evalFresnelSchlick(mk3(1.f), mk3(fd90), NdotV).x;
    return sd.diffuse * (viewScatter * lightScatter * energyFactor * (1.f / (float)M_PI));
}
enum { kLobeDiffuse = 1, kLobeSpecular = 2 };
static inline f3 evalBSDFLobes(const ShadingData& sd, f3 L, int lobes) {
    float NdotL = dot(sd.N, L);
    if (std::min(sd.NdotV, NdotL) < kMinCosTheta) return mk3(0.f);
    f3 H = normalize(sd.V + L);
    float NdotH = saturate(dot(sd.N, H)), LdotH = saturate(dot(L, H)), NdotV = saturate(sd.NdotV);
    NdotL = saturate(NdotL);
    f3 thp = mk3(0.f);
    if (lobes & kLobeDiffuse) thp = thp + evalDiffuseFrostbite(sd, NdotL, NdotV, LdotH);
    if (lobes & kLobeSpecular) {
        float alpha = std::max(kMinGGXAlpha, sd.ggxAlpha);
        float D = evalNdfGGX(alpha, NdotH);
        float G = evalMaskingSmithGGXCorrelated(alpha, NdotL, NdotV);
        thp = thp + evalFresnelSchlick(sd.specular, mk3(1.f), LdotH) * (D * G);
    }
    return thp;
}
static inline f3 evalBSDFCosine(const ShadingData& sd, f3 L) {
    return evalBSDFLobes(sd, L, 3) * saturate(dot(sd.N, L));
}
static inline float evalPdfDiffuse(const ShadingData& sd, f3 L) {
    float NdotL = dot(sd.N, L);
    return NdotL < kMinCosTheta ? 0.f : NdotL * (float)M_1_PI;
}
static inline float evalPdfSpecular(const ShadingData& sd, f3 L) {
    float NdotL = dot(sd.N, L);
    if (std::min(sd.NdotV, NdotL) < kMinCosTheta) return 0.f;
    f3 H = normalize(sd.V + L);
    float NdotH = saturate(dot(sd.N, H)), VdotH = saturate(dot(sd.V, H));
    float alpha = std::max(kMinGGXAlpha, sd.ggxAlpha);
    return evalNdfGGX(alpha, NdotH) * NdotH / (4.f * VdotH);
}
static inline float getDiffuseProbability(const ShadingData& sd) {
    float dw = (1.f - sd.metallic) * luminance(sd.diffuse);
    float sw = luminance(evalFresnelSchlick(sd.specular, mk3(1.f), dot(sd.V, sd.N)));
    return (dw + sw) > 0.f ? dw / (dw + sw) : 1.f;
}
static inline float evalPdfBSDF(const ShadingData& sd, f3 L) {
    float pd = getDiffuseProbability(sd);
    return pd * evalPdfDiffuse(sd, L) + (1.f - pd) * evalPdfSpecular(sd, L);
}
struct BSDFSample { f3 wi, weight; float pdf; int lobe; };
static inline bool sampleBSDF(const ShadingData& sd, SampleGen& sg, BSDFSample& r) {
    f2 u;
    u.x = next1D(sg);
    u.y = next1D(sg);
    float uSelect = next1D(sg);
    float pDiffuse = getDiffuseProbability(sd);
    float pmf, pdfOther;
    r.weight = mk3(0.f);
    if (uSelect < pDiffuse) {
        // sampleDiffuse: sample_cosine_hemisphere_concentric (UT/Math/MathHelpers.slang:293-322)
        float x = 2.f * u.x - 1.f, y = 2.f * u.y - 1.f, dx = 0.f, dy = 0.f;
        if (!(x == 0.f && y == 0.f)) {
            float rr, phi;
            if (std::fabs(x) > std::fabs(y)) { rr = x; phi = (y / x) * ((float)M_PI / 4.f); }
            else { rr = y; phi = (float)M_PI / 2.f - (x / y) * ((float)M_PI / 4.f); }
            dx = rr * std::cos(phi);
            dy = rr * std::sin(phi);
        }
        float z = std::sqrt(std::max(0.f, 1.f - (dx * dx + dy * dy)));
        r.pdf = z * (float)M_1_PI;
        r.wi = fromLocal({dx, dy, z}, sd);
        r.lobe = kLobeDiffuse;
        float NdotL = dot(sd.N, r.wi);
        if (std::min(sd.NdotV, NdotL) < kMinCosTheta) r.pdf = 0.f;
        else r.weight = evalBSDFLobes(sd, r.wi, kLobeDiffuse) * (float)M_PI;
        pmf = pDiffuse;
        pdfOther = evalPdfSpecular(sd, r.wi);
    } else {
        // sampleSpecular with sampleNdfGGX_Walter
        float alpha = std::max(kMinGGXAlpha, sd.ggxAlpha);
        float a2 = alpha * alpha;
        float cos2 = std::min((1.f - u.x) / ((a2 - 1.f) * u.x + 1.f), 1.0f);
        float cosH = std::sqrt(cos2), sinH = std::sqrt(1.f - cos2);
        float phi = u.y * 2.f * (float)M_PI;
        f3 H = fromLocal({sinH * std::cos(phi), sinH * std::sin(phi), cosH}, sd);
        float VdotH = dot(sd.V, H);
        r.wi = normalize(H * (2.f * VdotH) - sd.V);
        float d = (a2 - 1.f) * cos2 + 1.f;
        r.pdf = (a2 * cosH) / (d * d * VdotH * 4.f * (float)M_PI);
        r.lobe = kLobeSpecular;
        float NdotL = dot(sd.N, r.wi);
        if (NdotL < kMinCosTheta) r.pdf = 0.f;
        if (std::min(sd.NdotV, NdotL) < kMinCosTheta) {
            r.pdf = 0.f;
        } else {
            float NdotV = saturate(sd.NdotV);
            NdotL = saturate(NdotL);
            float NdotH = saturate(cosH);
            VdotH = saturate(VdotH);
            float G = evalMaskingSmithGGXCorrelated(alpha, NdotL, NdotV);
            r.weight = evalFresnelSchlick(sd.specular, mk3(1.f), VdotH) * ((G * NdotL * VdotH * 4.f) / NdotH);
        }
        pmf = 1.f - pDiffuse;
        pdfOther = evalPdfDiffuse(sd, r.wi);
    }
    r.pdf = pmf * r.pdf + (1.f - pmf) * pdfOther;
    r.weight = r.weight / pmf;
    return r.pdf > 0.f;
}

// _prepareShadingData's frontFacing and front-only emission (SC/Shading.slang:172, 229-232).
static inline bool emitsTowards(const SceneState& S, int tri, f3 V) {
    const SurfTri& T = S.surfTris[tri];
    f3 faceN = cross(T.p[1] - T.p[0], T.p[2] - T.p[0]);
    if (!(dot(V, faceN) >= 0.f)) return false;
    f3 e = S.surfMats[T.material].emissive;
    return e.x > 0.f || e.y > 0.f || e.z > 0.f;
}

// computeRayOrigin (UT/Helpers.slang:92-105).
static inline f3 computeRayOrigin(f3 pos, f3 normal) {
    const float origin = 1.f / 32.f, fScale = 1.f / 65536.f, iScale = 256.f;
    float P[3] = {pos.x, pos.y, pos.z}, N[3] = {normal.x, normal.y, normal.z}, R[3];
    for (int k = 0; k < 3; k++) {
        int iOff = (int)(N[k] * iScale);
        float iPos = asfloat(asint(P[k]) + (P[k] < 0.f ? -iOff : iOff));
        float fOff = N[k] * fScale;
        R[k] = std::fabs(P[k]) < origin ? P[k] + fOff : iPos;
    }
    return {R[0], R[1], R[2]};
}

struct TriangleLightSample { f3 posW, normalW, dir, Le; float distance, pdf, pdfArea, cosTheta; f2 uv; uint32_t tri; };

// sampleTriangle (LT/EmissiveLightSamplerHelpers.slang:56-94) with sample_triangle
// (UT/Math/MathHelpers.slang:356-361).
static inline bool sampleTriangle(const SceneState& S, f3 posW, uint32_t tri, f2 u, TriangleLightSample& ls) {
    ls = TriangleLightSample{};
    ls.tri = tri;
    const EmissiveTriangle& T = S.emTris[tri];
    float su = std::sqrt(u.x);
    float bx = 1.f - su, by = u.y * su;
    f3 bary{1.f - bx - by, bx, by};
    ls.uv = u;
    ls.posW = T.p[0] * bary.x + T.p[1] * bary.y + T.p[2] * bary.z;
    ls.posW = computeRayOrigin(ls.posW, T.normal);
    f3 toLight = ls.posW - posW;
    const float distSqr = std::max(FLT_MIN, dot(toLight, toLight));
    ls.distance = std::sqrt(distSqr);
    ls.dir = toLight / ls.distance;
    ls.normalW = T.normal;
    float cosTheta = dot(ls.normalW, -ls.dir);
    if (!(cosTheta > 0.f)) return false;
    ls.Le = T.Le;
    float denom = std::max(FLT_MIN, cosTheta * T.area);
    ls.pdf = distSqr / denom;
    ls.cosTheta = cosTheta;
    ls.pdfArea = 1.f / T.area;
    return true;
}

// EmissivePowerSampler::sampleLight (LT/EmissivePowerSampler.slang:56-79).
static inline bool powerSamplerSampleLight(const SceneState& S, f3 posW, SampleGen& sg, TriangleLightSample& ls) {
    if (S.emActive == 0) return false;
    f2 r;
    r.x = next1D(sg);
    r.y = next1D(sg);
    const uint32_t count = (uint32_t)S.emAlias.size();
    uint32_t index = std::min(count - 1, (uint32_t)(r.x * (float)count));
    const AliasItem& it = S.emAlias[index];
    uint32_t tri = r.y >= it.threshold ? it.indexA : it.indexB;
    float sel = S.emWeight[tri] / S.emWeightSum;
    f2 u;
    u.x = next1D(sg);
    u.y = next1D(sg);
    if (!sampleTriangle(S, posW, tri, u, ls)) return false;
    ls.pdf *= sel;
    ls.pdfArea *= sel;
    return true;
}

// EmissivePowerSampler::getEmissiveLightSample (LT/EmissivePowerSampler.slang:81-92).
static inline bool powerSamplerGetSample(const SceneState& S, f3 posW, int lightID, f2 uv, TriangleLightSample& ls) {
    if (S.emActive == 0) return false;
    if (lightID < 0 || lightID >= (int)S.emTris.size()) return false;  // a light id of another scene
    return sampleTriangle(S, posW, (uint32_t)lightID, uv, ls);
}

// sampleSceneLights (VR/VolumeUtils.slang:12-149): environment (constant or
// textured), point lights and emissive triangles (the power sampler).
static inline bool sampleSceneLights(const SceneState& S, const EnvXf& X, f3 origin, bool useEnv, bool useAnalytic,
                                     bool useEmissive, SampleGen& sg, SceneLightSample& ls, int& lightIndex,
                                     f2& lightUV) {
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
            float lpdf;
            f3 dir = envSample(S, X, rnd, lpdf);
            f3 Le = envEval(S, X, dir);
            float pdf = sel * lpdf;
            ls.rayDir = ls.dir = dir;
            ls.rayDistance = ls.distance = kRayTMax;
            ls.pdf = pdf;
            ls.pdfArea = pdf;
            ls.Li = pdf > 0.f ? Le / pdf : mk3(0.f);
            lightIndex = -1;
            lightUV = encodeLightUV(S, X, ls.rayDir, lightIndex);
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
    if (useEmissive) {
        if (u < p[2]) {
            float sel = p[2];
            TriangleLightSample t;
            bool valid = powerSamplerSampleLight(S, origin, sg, t);
            if (S.emActive == 0) return false;
            float pdf = sel * t.pdf;
            float pdfArea = sel * t.pdfArea;
            f3 offsetPos = computeRayOrigin(t.posW, t.normalW);
            f3 toLight = offsetPos - origin;
            ls.rayDistance = length(toLight);
            ls.rayDir = normalize(toLight);
            ls.dir = t.dir;
            ls.distance = t.distance;
            ls.pdf = pdf;
            ls.pdfArea = pdfArea;
            ls.Li = pdf > 0.f ? t.Le * S.emMult / pdf : mk3(0.f);
            lightIndex = (int)t.tri + (int)S.lights.size();
            lightUV = t.uv;
            return valid;
        }
        u -= p[2];
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
static inline f3 sampleDirectLighting(const SceneState& S, const EnvXf& X, SampleGen& sg, float& pdf,
                                      const MediumInteraction& mi,
                                      bool useEnv, bool useAnalytic, bool useEmissive, bool enableShadow, int mip,
                                      uint32_t method, bool linear, int visSamples, float tStepScale, int& lightIndex,
                                      f2& lightUV, float& outVis) {
    pdf = 0.f;
    SceneLightSample ls{};
    bool valid = sampleSceneLights(S, X, mi.p, useEnv, useAnalytic, useEmissive, sg, ls, lightIndex, lightUV);
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
static inline f3 directLighting(const SceneState& S, const EnvXf& X, SampleGen& sg, const MediumInteraction& mi,
                                int nSamples,
                                bool useEnv, bool useAnalytic, bool useEmissive, int mip, uint32_t method) {
    f3 Ld = mk3(0.f);
    for (int i = 0; i < nSamples; i++) {
        SceneLightSample ls{};
        int idx; f2 uv;
        bool valid = sampleSceneLights(S, X, mi.p, useEnv, useAnalytic, useEmissive, sg, ls, idx, uv);
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
