// vr_pass.hpp -- device-side ReSTIR machinery: lights, reservoirs, the target
// function and candidate construction, templated on the path length B
// (MAX_BOUNCES of the reference) so per-bounce arrays live in registers.
#pragma once
#include "vr_device.hpp"
#include "vr_rayquery.hpp"
#include "vr_bsdf.hpp"

namespace vr {

// Reservoir (VR/HostDeviceSharedDefinitions.h:16-45). In HBM a reservoir is 32 B
// (two float4); extraBounceStartId is not stored because it always equals
// sourcePixel * (B - 1) -- the kernels carry it in a register instead.
struct Res {
    float runningSum, M, depth, p_y;
    float uvx, uvy;
    int lightID, sampledPixel;
    int exStart;
};
struct ResStore { float4 a, b; };

VR_DEV Res newRes() { return {0.f, 0.f, kTMax, 0.f, 0.f, 0.f, 0, 0, 0}; }
VR_DEV Res loadRes(const ResStore* buf, int pixel, int B) {
    ResStore s = buf[pixel];
    Res r;
    r.runningSum = s.a.x; r.M = s.a.y; r.depth = s.a.z; r.p_y = s.a.w;
    r.uvx = s.b.x; r.uvy = s.b.y;
    r.lightID = __float_as_int(s.b.z);
    r.sampledPixel = __float_as_int(s.b.w);
    r.exStart = pixel * (B - 1);
    return r;
}
VR_DEV void storeRes(ResStore* buf, int pixel, const Res& r) {
    ResStore s;
    s.a = make_float4(r.runningSum, r.M, r.depth, r.p_y);
    s.b = make_float4(r.uvx, r.uvy, __int_as_float(r.lightID), __int_as_float(r.sampledPixel));
    buf[pixel] = s;
}
VR_DEV void copySample(Res& st, const Res& r) {
    st.depth = r.depth; st.p_y = r.p_y; st.uvx = r.uvx; st.uvy = r.uvy; st.lightID = r.lightID;
    st.exStart = r.exStart; st.sampledPixel = r.sampledPixel;
}
// simpleResampleStep / simpleResampleStepWithMaxM (VR/Reservoir.slang:26-87).
VR_DEV bool resample(const Res& r, Res& st, Rng& rng) {
    float w = r.runningSum;
    st.M += r.M;
    if (w <= 0.0f) return false;
    st.runningSum += w;
    bool sel = u01(rng) * st.runningSum < w;
    if (sel) copySample(st, r);
    return sel;
}
VR_DEV bool resampleMaxM(const Res& r, float cap, Res& st, Rng& rng) {
    float cM = fminf(cap, r.M);
    float w = cM == 0.0f ? 0.0f : cM / r.M * r.runningSum;
    st.M += cM;
    if (w <= 0.0f) return false;
    st.runningSum += w;
    bool sel = u01(rng) * st.runningSum < w;
    if (sel) copySample(st, r);
    return sel;
}

// Bit packing (VR/ReSTIRHelper.slang:10-89).
VR_DEV int decodeMaxIndirect(int s) { return s >> 20; }
VR_DEV int encodeMaxIndirect(int s, int b) { return (b << 20) | (s & 0xFFFFF); }
VR_DEV int decodePathTag(int s) { return (s >> 16) & 0xF; }
VR_DEV int encodePathTag(int s, int tag) { return (tag << 16) | (int)((uint32_t)s & 0xFFF0FFFFu); }
VR_DEV v3 decodeWi(v3 in, float& dist) {
    v3 wi{in.x, in.y, sqrtf(1 - (in.x * in.x + in.y * in.y))};
    if (isnan(wi.z)) wi.z = 0.f;
    if (in.z < 0) { wi.z = -wi.z; in.z = -in.z; }
    dist = in.z;
    return wi;
}
VR_DEV v3 encodeWiDist(v3 dir, float dist) {
    v3 r{dir.x, dir.y, dist};
    if (dir.z < 0) r.z = -r.z;
    return r;
}

// ------------------------------------------------------------------ camera (SC/Camera/Camera.slang:172-228), no jitter
VR_DEV v3 rayDirNN(const CameraFrame& c, int px, int py, int W, int H) {
    float p0 = ((float)px + 0.5f) / (float)W, p1 = ((float)py + 0.5f) / (float)H;
    float nx = 2.f * p0 + -1.f, ny = -2.f * p1 + 1.f;
    return nx * ld3(c.U) + ny * ld3(c.V) + ld3(c.W);
}
VR_DEV v3 rayDir(const CameraFrame& c, int px, int py, int W, int H) { return norm3(rayDirNN(c, px, py, W, H)); }

// ------------------------------------------------------------------ lights
VR_DEV float sgnf(float x) { return x > 0.f ? 1.f : (x < 0.f ? -1.f : 0.f); }
// oct_to_ndir_equal_area_unorm (UT/Math/MathHelpers.slang:200-219).
VR_DEV v3 octToDir(float px, float py) {
    px = px * 2.f - 1.f; py = py * 2.f - 1.f;
    float d = 1.f - (fabsf(px) + fabsf(py));
    float r = 1.f - fabsf(d);
    float phi = (r > 0.f) ? ((fabsf(py) - fabsf(px)) / r + 1.f) * (kPiF / 4.f) : 0.f;
    float f = r * sqrtf(2.f - r * r);
    return {f * sgnf(px) * cosf(phi), f * sgnf(py) * sinf(phi), sgnf(d) * (1.f - r * r)};
}
// ------------------------------------------------------------------ environment map
// A constant environment (envTextured == 0) has no orientation: light UVs hold
// the world direction. A textured one follows EnvMap / EnvMapSampler.
VR_DEV v3 mat3v(const float* M, v3 d) {
    return V3(M[0] * d.x + M[1] * d.y + M[2] * d.z, M[3] * d.x + M[4] * d.y + M[5] * d.z,
              M[6] * d.x + M[7] * d.y + M[8] * d.z);
}
// world_to_latlong_map (UT/Math/MathHelpers.slang:92-99); p.y is clamped to
// [-1, 1] so acos stays defined for directions normalised to just above 1.
VR_DEV void latlong(v3 dir, float& u, float& v) {
    v3 p = norm3(dir);
    u = atan2f(p.x, -p.z) * (0.5f / kPiF) + 0.5f;
    v = acosf(fminf(fmaxf(p.y, -1.f), 1.f)) * (1.f / kPiF);
}
// envMap.SampleLevel(envSampler, uv, 0): bilinear over texel centres, wrap in u,
// clamp in v (LT/EnvMap.cpp:113-121). A NaN coordinate reads black.
VR_DEV v3 texBilinear(const float* tex, int W, int H, float u, float v) {
    float x = u * (float)W - 0.5f, y = v * (float)H - 0.5f;
    if (!(fabsf(x) < 1e30f) || !(fabsf(y) < 1e30f)) return V3(0.f);
    float x0f = floorf(x), y0f = floorf(y);
    float fx = x - x0f, fy = y - y0f;
    int x0 = (int)fmodf(x0f, (float)W);
    if (x0 < 0) x0 += W;
    int x1 = x0 + 1 == W ? 0 : x0 + 1;
    const int yi = (int)fminf(fmaxf(y0f, -1.f), (float)H);
    int y0 = min(max(yi, 0), H - 1);
    int y1 = min(max(yi + 1, 0), H - 1);
    gptr<float> t = gbl<float>(tex);
    const size_t r0 = (size_t)y0 * W, r1 = (size_t)y1 * W;
    float a[3], b[3], c[3], d[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        a[k] = t[4 * (r0 + x0) + k]; b[k] = t[4 * (r0 + x1) + k];
        c[k] = t[4 * (r1 + x0) + k]; d[k] = t[4 * (r1 + x1) + k];
    }
    v3 top = V3(a[0] + (b[0] - a[0]) * fx, a[1] + (b[1] - a[1]) * fx, a[2] + (b[2] - a[2]) * fx);
    v3 bot = V3(c[0] + (d[0] - c[0]) * fx, c[1] + (d[1] - c[1]) * fx, c[2] + (d[2] - c[2]) * fx);
    return top + (bot - top) * fy;
}
VR_DEV v3 envBilinear(const Scene& S, float u, float v) { return texBilinear(S.envTex, S.envW, S.envH, u, v); }
// EnvMap::evalLocal / eval (LT/EnvMap.slang:45-57): intensity * tint * texel.
VR_DEV v3 envEvalLocal(const Scene& S, v3 dLocal) {
    float u, v;
    latlong(dLocal, u, v);
    return ld3(S.envScale) * envBilinear(S, u, v);
}
VR_DEV v3 envEval(const Scene& S, const EnvXf& X, v3 dWorld) {
    if (!S.envTextured) return ld3(S.env);
    return envEvalLocal(S, mat3v(X.toLocal, dWorld));
}
// EnvMapSampler::encodeLightUV / decodeLightUV (LT/EnvMapSampler.slang:75-90).
VR_DEV void encodeLightUV(const Scene& S, const EnvXf& X, v3 dWorld, int& id, float& u, float& v) {
    v3 d = S.envTextured ? mat3v(X.toLocal, dWorld) : dWorld;
    id = d.z < 0 ? -2 : -1;
    u = d.x; v = d.y;
}
VR_DEV v3 decodeLightUV(float u, float v, int id) {
    v3 d{u, v, sqrtf(1 - u * u - v * v)};
    if (isnan(d.z)) d.z = 0.f;
    if (id == -2) d.z = -d.z;
    return d;
}
VR_DEV v3 decodeLightUV(const Scene& S, const EnvXf& X, float u, float v, int id) {
    v3 d = decodeLightUV(u, v, id);
    return S.envTextured ? mat3v(X.toWorld, d) : d;
}
// EnvMapSampler::sample (LT/EnvMapSampler.slang:94-146): hierarchical 2x2 warping
// of (r0, r1) down the importance mip chain, then the equal-area octahedral map.
// Returns the world direction; pdf is per solid angle.
VR_DEV v3 envSample(const Scene& S, const EnvXf& X, float r0, float r1, float& pdf) {
    if (!S.envTextured) {
        pdf = kInv4PiF;
        return octToDir(r0, r1);
    }
    gptr<float> imp = gbl<float>(S.envImp);
    int px = 0, py = 0;
    for (int mip = S.envImpBaseMip - 1; mip >= 0; mip--) {
        px *= 2; py *= 2;
        const int dk = S.envImpDim >> mip;
        const size_t o = (size_t)S.envImpOff[mip] + (size_t)py * dk + px;
        float w0 = imp[o], w1 = imp[o + 1], w2 = imp[o + dk], w3 = imp[o + dk + 1];
        float q0 = w0 + w2, q1 = w1 + w3;
        float d = q0 / (q0 + q1);
        int ox;
        if (r0 < d) { ox = 0; r0 = r0 / d; }
        else { ox = 1; r0 = (r0 - d) / (1.f - d); }
        float e = ox ? w1 / q1 : w0 / q0;
        int oy;
        if (r1 < e) { oy = 0; r1 = r1 / e; }
        else { oy = 1; r1 = (r1 - e) / (1.f - e); }
        px += ox; py += oy;
    }
    float u = ((float)px + r0) * S.envImpInvDim, v = ((float)py + r1) * S.envImpInvDim;
    v3 dir = octToDir(u, v);
    float avg = imp[S.envImpOff[S.envImpBaseMip]];
    pdf = imp[(size_t)S.envImpOff[0] + (size_t)py * S.envImpDim + px] / avg * kInv4PiF;
    return mat3v(X.toWorld, dir);
}

struct LightSample { v3 dir, Li, rayDir; float pdf, pdfArea, rayDistance; };

// ------------------------------------------------------------------ emissive triangles
struct TriLightSample { v3 posW, normal, dir, Le; float distance, pdf, pdfArea, cosTheta; };
// sampleTriangle (LT/EmissiveLightSamplerHelpers.slang:56-94): the point of
// sample_triangle(u) (UT/Math/MathHelpers.slang:356-361) on collection triangle
// tri, offset off the triangle; false when the triangle faces away.
VR_DEV bool sampleTriangle(const Scene& S, v3 posW, int tri, float ux, float uy, TriLightSample& ls) {
    const float* T = S.emTri + (size_t)kEmTriFloats * tri;
    const float su = sqrtf(ux);
    const float b1 = 1.f - su, b2 = uy * su, b0 = 1.f - b1 - b2;
    ls.normal = ld3(T + 9);
    ls.posW = ld3(T) * b0 + ld3(T + 3) * b1 + ld3(T + 6) * b2;
    ls.posW = rayOriginOffset(ls.posW, ls.normal);
    const v3 toLight = ls.posW - posW;
    const float distSqr = fmaxf(1.17549435e-38f, dot3(toLight, toLight));
    ls.distance = sqrtf(distSqr);
    ls.dir = toLight / ls.distance;
    const float cosTheta = dot3(ls.normal, -ls.dir);
    if (!(cosTheta > 0.f)) return false;
    ls.Le = ld3(T + 13);
    const float area = T[12];
    ls.pdf = distSqr / fmaxf(1.17549435e-38f, cosTheta * area);
    ls.cosTheta = cosTheta;
    ls.pdfArea = 1.f / area;
    return true;
}
// EmissivePowerSampler::sampleLight (LT/EmissivePowerSampler.slang:56-79): a
// collection triangle from the alias table (UT/Sampling/AliasTable.slang:55-73),
// then a uniform point on it.
VR_DEV bool samplePowerLight(const Scene& S, v3 posW, Rng& rng, int& tri, float& ux, float& uy, TriLightSample& ls) {
    if (S.numEmActive == 0) return false;
    const float a0 = u01(rng);
    const float a1 = u01(rng);
    const uint32_t count = (uint32_t)S.numEmTris;
    const uint32_t slot = min(count - 1u, (uint32_t)(a0 * (float)count));
    const uint32_t* item = S.emAlias + 4 * (size_t)slot;
    tri = (int)(a1 >= __uint_as_float(item[0]) ? item[1] : item[2]);
    const float sel = S.emWeight[tri] / S.emWeightSum;
    ux = u01(rng);
    uy = u01(rng);
    if (!sampleTriangle(S, posW, tri, ux, uy, ls)) return false;
    ls.pdf *= sel;
    ls.pdfArea *= sel;
    return true;
}

// sampleSceneLights (VR/VolumeUtils.slang:12-149) for the environment (constant
// or textured), point lights and emissive triangles.
VR_DEV bool sampleLights(const Scene& S, const EnvXf& X, v3 origin, bool useEnv, bool useAnalytic, bool useEmissive,
                         Rng& rng, LightSample& ls, int& lightIndex, float& uvx, float& uvy) {
    if (!useEnv && !useAnalytic && !useEmissive) return false;
    float p0 = useEnv ? 1.f : 0.f, p1 = useAnalytic ? 1.f : 0.f, p2 = useEmissive ? 1.f : 0.f;
    float sum = p0 + p1 + p2;
    if (sum == 0.f) return false;
    float inv = 1.f / sum;
    p0 *= inv; p1 *= inv; p2 *= inv;
    float u = u01(rng);
    if (useEnv) {
        if (u < p0) {
            float r0 = u01(rng);
            float r1 = u01(rng);
            float lpdf;
            v3 dir = envSample(S, X, r0, r1, lpdf);
            v3 Le = envEval(S, X, dir);
            float pdf = p0 * lpdf;
            ls.rayDir = ls.dir = dir;
            ls.rayDistance = kTMax;
            ls.pdf = pdf;
            ls.pdfArea = pdf;
            ls.Li = pdf > 0.f ? Le / pdf : V3(0.f);
            encodeLightUV(S, X, dir, lightIndex, uvx, uvy);
            return !(isnan(dir.x) || isnan(dir.y) || isnan(dir.z));
        }
        u -= p0;
    }
    if (useAnalytic) {
        if (u < p1) {
            u /= p1;
            uint32_t count = (uint32_t)S.numLights;
            uint32_t idx = min((uint32_t)(u * (float)count), count - 1);
            float sel = p1 / (float)count;
            const float* L = S.lights + 6 * idx;
            v3 toL = ld3(L) - origin;
            float d2 = fmaxf(dot3(toL, toL), 1e-9f);
            float dist = sqrtf(d2);
            v3 dir = toL / dist;
            v3 Li = ld3(L + 3) / d2;
            lightIndex = (int)idx;
            ls.rayDir = ls.dir = dir;
            ls.rayDistance = dist;
            ls.pdf = sel * 1.f;
            ls.pdfArea = ls.pdf;
            ls.Li = Li / ls.pdf;
            uvx = 0.f; uvy = 0.f;
            return true;
        }
        u -= p1;
    }
    if (useEmissive) {
        if (u < p2) {
            // the emissive branch: the ray ends at the sampled point offset once more
            int tri = 0;
            float ux = 0.f, uy = 0.f;
            TriLightSample ts;
            ts.pdf = ts.pdfArea = 0.f;
            ts.posW = ts.normal = ts.dir = ts.Le = V3(0.f);
            ts.distance = 0.f;
            if (S.numEmActive == 0) return false;
            const bool valid = samplePowerLight(S, origin, rng, tri, ux, uy, ts);
            const float pdf = p2 * ts.pdf, pdfArea = p2 * ts.pdfArea;
            const v3 toLight = rayOriginOffset(ts.posW, ts.normal) - origin;
            ls.rayDistance = len3(toLight);
            ls.rayDir = norm3(toLight);
            ls.dir = ts.dir;
            ls.pdf = pdf;
            ls.pdfArea = pdfArea;
            ls.Li = pdf > 0.f ? ts.Le * S.emMult / pdf : V3(0.f);
            lightIndex = tri + S.numLights;
            uvx = ux; uvy = uy;
            return valid;
        }
    }
    return false;
}

// ------------------------------------------------------------------ volume emission
// getValueAtPoint (GVS/gvdb.slang:6-29) in world space: trilinear lookup of a
// grid slot with no bounds test beyond the grid's bricks. Returns the brick id
// (kEmpty outside every brick) and the brick-local point.
VR_DEV uint32_t brickAtWorld(const Scene& S, int slot, v3 pW, v3& q) {
    const Grid& G = S.grid[slot];
    v3 p = xformP(G.w2i, pW);
    if (!(p.x >= 0.f && p.y >= 0.f && p.z >= 0.f)) return kEmpty;
    if (!(p.x < (float)(G.nbx * 8) && p.y < (float)(G.nby * 8) && p.z < (float)(G.nbz * 8))) return kEmpty;
    int bx = (int)p.x / 8, by = (int)p.y / 8, bz = (int)p.z / 8;
    q = p - V3((float)(bx * 8), (float)(by * 8), (float)(bz * 8));
    return brickLookup(G, slot, bx, by, bz);
}
// TemperatureWorldSpace (VR/VolumeBase.slang:196-209); slot 16, or 27 for the
// previous frame of a sequence.
VR_DEV float temperatureWorld(const Scene& S, v3 pW, int slot = kTemperatureSlot) {
    v3 q;
    uint32_t b = brickAtWorld(S, slot, pW, q);
    return b == kEmpty ? 0.f : sampleBrick(S.grid[slot], b, q, true);
}
// VelocityWorld (VR/VolumeBase.slang:183-194): trilinear lookup of the three
// channels of the velocity grid. The grid holds world-space displacement per
// frame (the reference's model-space vector times the volume's model-to-world).
VR_DEV v3 velocityWorld(const Scene& S, v3 pW) {
    v3 q;
    uint32_t b = brickAtWorld(S, kVelocitySlot, pW, q);
    if (b == kEmpty) return V3(0.f);
    const Grid& G = S.grid[kVelocitySlot];
    return V3(sampleBrick(G, 3 * b, q, true), sampleBrick(G, 3 * b + 1, q, true), sampleBrick(G, 3 * b + 2, q, true));
}
// ConvertTempToColor + EmissionWorldSpace (VR/VolumeBase.slang:211-232): the
// blackbody table is a 1-D texture over [25 K, 6425 K) sampled with a linear
// filter and border colour 0.
VR_DEV v3 blackbody(const Scene& S, float temp) {
    temp = fminf(6400.f, (temp - S.tempCutoff) * S.tempScale);
    float q = (temp - 25.f) / 6400.f;
    const int N = S.lutRows;
    float x = q * (float)N - 0.5f;
    if (!(x > -1.f && x < (float)N)) return V3(0.f);  // both taps on the border (or NaN)
    float x0f = floorf(x);
    float f = x - x0f;
    int i0 = (int)x0f, i1 = i0 + 1;
    v3 a = i0 >= 0 ? ld3(S.lut + 3 * i0) : V3(0.f);
    v3 b = i1 < N ? ld3(S.lut + 3 * i1) : V3(0.f);
    return S.LeScale * (a + (b - a) * f);
}
// last: the previous frame's temperature grid (useLastFrameGrid).
VR_DEV v3 emissionWorld(const Scene& S, v3 pW, bool last = false) {
    if (last ? !S.prevHasEmission : !S.hasEmission) return V3(0.f);
    return blackbody(S, temperatureWorld(S, pW, last ? kTemperatureSlot + kPrevExtraOffset : kTemperatureSlot));
}

// directLighting (VR/VolumeUtils.slang:416-450): NEE of the reference path tracer.
VR_DEV v3 directLightingPT(const Scene& S, const EnvXf& X, Rng& rng, const MediumInteraction& mi, int nSamples,
                           const Options& o, int mip, uint32_t method) {
    v3 Ld = V3(0.f);
    for (int i = 0; i < nSamples; i++) {
        LightSample ls;
        int idx; float ux, uy;
        bool valid = sampleLights(S, X, mi.p, o.useEnv, o.useAnalytic, o.useEmissive, rng, ls, idx, ux, uy);
        if (!valid) continue;
        Ray sh{mi.p, ls.rayDir, 0.f, ls.rayDistance};
        float vis = visibilityT<3>(S, sh, rng, 1, mip, true, method, 1.f);
        ls.Li = ls.Li * vis;
        float p = phaseHG(dot3(mi.wo, ls.dir), mi.g);
        Ld = Ld + V3(p) * ls.Li / 1.f;
    }
    return Ld / (float)nSamples;
}

// ------------------------------------------------------------------ target function
// How evaluateF obtains the visibility of a transmittance ray. The direct
// policy traces it on the spot (visibilityT). The record and replay policies
// split one reuse kernel into three launches: the record pass runs the pixel
// program with every visibility taken as 1 and writes each ray it would trace;
// a separate kernel traces all recorded rays from a balanced queue; the replay
// pass runs the program again and reads each ray's transmittance back in the
// same order. Only deterministic trackers (ray marching, analytic) qualify:
// they draw no random numbers, so the split changes no result.
//
// To make record and replay enumerate the same rays, their early exits look
// only at the transmittance-free product Fn (sample, phase, density and light
// factors), never at a visibility: a zero visibility makes F zero and every
// later factor is finite, so continuing changes the result by nothing.
//
// The record pass also keeps every density lookup of its target evaluations,
// and the replay reads them back in the same order instead of fetching the
// volume again. An evaluation makes at most b lookups and may record b + 1
// rays, so a pixel's ray capacity bounds its lookups too.
template <int VM, int LM>
struct DirectTrace {
    static constexpr bool kSplit = false;
    // kind 0: visibility options; 1: lighting options (grid 0 in final shading).
    // off: kPrevDensityOffset for the previous frame's grids, else 0.
    VR_DEV float vis(const Scene& S, const Ray& r, Rng& rng, const Options& o, int kind, bool isFinal, int off) {
        if (kind == 0)
            return visibilityT<VM>(S, r, rng, o.visibilitySamples, o.visibilityMipLevel + off,
                                   o.visibilityUseLinearSampler != 0, o.visibilityTrackingMethod,
                                   o.visibilityTStepScale);
        return visibilityT<LM>(S, r, rng, o.lightSamples, isFinal ? 0 : o.lightingMipLevel + off,
                               o.lightingUseLinearSampler != 0, o.lightingTrackingMethod, o.lightingTStepScale);
    }
    VR_DEV float density(const Scene& S, int off, v3 p) { return densityWorld(S, off, p); }
};
// visibilityT's average of `samples` identical transmittances (0/0 for none).
VR_DEV float sameSamples(float tr, int samples) {
    float v = 0.f;
    for (int i = 0; i < samples; i++) v += tr;
    return v / (float)samples;
}
struct RecordTrace {
    static constexpr bool kSplit = true;
    TraceRec* out;  // this pixel's record slots
    int n;          // records written
    VR_DEV float vis(const Scene&, const Ray& r, Rng&, const Options& o, int kind, bool, int) {
        const int samples = kind == 0 ? o.visibilitySamples : o.lightSamples;
        if (samples <= 0) return sameSamples(1.f, samples);
        TraceRec rec;
        rec.ox = r.o.x; rec.oy = r.o.y; rec.oz = r.o.z; rec.tMax = r.tMax;
        rec.dx = r.d.x; rec.dy = r.d.y; rec.dz = r.d.z; rec.kind = kind;
        out[n++] = rec;
        return sameSamples(1.f, samples);
    }
    float* dens;       // this pixel's first density slot; slot i is i * dstride further
    size_t dstride;
    int nd;
    VR_DEV float density(const Scene& S, int off, v3 p) {
        const float d = densityWorld(S, off, p);
        dens[(size_t)nd++ * dstride] = d;
        return d;
    }
};
struct ReplayTrace {
    static constexpr bool kSplit = true;
    const float* tr;  // this pixel's first transmittance; the i-th recorded ray's is i * trStride further
    size_t trStride;
    int n;
    VR_DEV float vis(const Scene&, const Ray&, Rng&, const Options& o, int kind, bool, int) {
        const int samples = kind == 0 ? o.visibilitySamples : o.lightSamples;
        if (samples <= 0) return sameSamples(1.f, samples);
        return sameSamples(tr[(size_t)n++ * trStride], samples);
    }
    const float* dens;  // the record pass's density lookups, in the same order and layout
    size_t dstride;
    int nd;
    VR_DEV float density(const Scene&, int, v3) { return dens[(size_t)nd++ * dstride]; }
};

// SampleDirectLighting (VR/VolumeUtils.slang:454-492); the shadow ray's
// transmittance comes from the trace policy tp (lighting options, kind 1).
template <class TP>
VR_DEV v3 sampleDirectT(const Scene& S, const EnvXf& X, Rng& rng, float& pdf, const MediumInteraction& mi,
                        const Options& o, bool enableShadow, int& lightIndex, float& uvx, float& uvy, TP& tp) {
    pdf = 0.f;
    LightSample ls;
    ls.pdfArea = 0.f;
    bool valid = sampleLights(S, X, mi.p, o.useEnv, o.useAnalytic, o.useEmissive, rng, ls, lightIndex, uvx, uvy);
    pdf = ls.pdfArea;
    if (!valid) { pdf = 0.f; return V3(0.f); }
    Ray sh{mi.p, ls.rayDir, 0.f, ls.rayDistance};
    if (enableShadow) {
        float vis = tp.vis(S, sh, rng, o, 1, false, 0);
        ls.Li = ls.Li * vis;
    }
    float p = phaseHG(dot3(mi.wo, ls.dir), mi.g);
    return V3(p) * ls.Li / 1.f;
}

// How TraceRays' candidate generation obtains its free-flight samples, its
// shadow rays' transmittances and its resampling draws. CandDirect traverses the
// grid for both kinds of ray. The split form runs the candidate loop twice around
// a balanced trace of the shadow rays:
//   * CandRecord draws every free-flight sample (its traversal), keeps the
//     results and the generator's state after the draw in the pixel's density
//     slots (from slot 0 up), writes each shadow ray as a record with its
//     visibility taken as 1, and keeps one bit per resampling step: whether that
//     step drew a random number (a positive weight; from the last slot down);
//   * CandReplay reads the samples and transmittances back in the same order
//     instead of tracing.
// This needs the shadow rays to be ray-marched (that tracker draws no random
// numbers) and the random draws to depend on no shadow transmittance. They do in
// one way only: a resampling step draws for a positive weight and not for a zero
// one, and a weight that is positive with visibility 1 can be zero with the real
// transmittance (exactly 0, or a product that underflows). The replay compares
// each step with the record's bit and flags a pixel where one differs (zero);
// such pixels are run again in one kernel.
template <int IM>
struct CandDirect : DirectTrace<IM, IM> {
    VR_DEV void analytic(const Scene& S, const Ray& r, Rng& rng, bool linear, int slot, int ns, float h[4],
                         float pd[4], float tr[4]) {
        sampleAnalytic(S, r, rng, linear, slot, ns, h, pd, tr);
    }
    VR_DEV bool resample(const Res& r, Res& st, Rng& rng) { return vr::resample(r, st, rng); }
};
struct CandRecord : RecordTrace {
    int cap;          // the pixel's density slots
    uint32_t mask;    // the current word of draw bits
    int bits, words;  // bits in it, words stored
    VR_DEV void put(float v) { dens[(size_t)nd++ * dstride] = v; }
    VR_DEV void analytic(const Scene& S, const Ray& r, Rng& rng, bool linear, int slot, int ns, float h[4],
                         float pd[4], float tr[4]) {
        sampleAnalytic(S, r, rng, linear, slot, ns, h, pd, tr);
#pragma unroll
        for (int i = 0; i < 4; i++)
            if (i < ns) { put(h[i]); put(pd[i]); put(tr[i]); }
        put(__uint_as_float(rng.s0)); put(__uint_as_float(rng.s1));
        put(__uint_as_float(rng.s2)); put(__uint_as_float(rng.s3));
    }
    VR_DEV void flushMask() {
        dens[(size_t)(cap - 1 - words++) * dstride] = __uint_as_float(mask);
        mask = 0u;
        bits = 0;
    }
    VR_DEV bool resample(const Res& r, Res& st, Rng& rng) {
        mask |= (r.runningSum > 0.f ? 1u : 0u) << bits;
        if (++bits == 32) flushMask();
        return vr::resample(r, st, rng);
    }
    VR_DEV void finish() {
        if (bits) flushMask();
    }
};
struct CandReplay : ReplayTrace {
    int cap;
    uint32_t mask;
    int bits, words;  // bits of mask used (32: load the next word), words loaded
    bool zero;        // a resampling step drew where the record's did not, or the reverse
    VR_DEV float get() { return dens[(size_t)nd++ * dstride]; }
    VR_DEV void analytic(const Scene&, const Ray&, Rng& rng, bool, int, int ns, float h[4], float pd[4],
                         float tr[4]) {
#pragma unroll
        for (int i = 0; i < 4; i++)
            if (i < ns) { h[i] = get(); pd[i] = get(); tr[i] = get(); }
        rng.s0 = __float_as_uint(get()); rng.s1 = __float_as_uint(get());
        rng.s2 = __float_as_uint(get()); rng.s3 = __float_as_uint(get());
    }
    VR_DEV bool resample(const Res& r, Res& st, Rng& rng) {
        if (bits == 32) {
            mask = __float_as_uint(dens[(size_t)(cap - 1 - words++) * dstride]);
            bits = 0;
        }
        const bool drew = ((mask >> bits++) & 1u) != 0u;
        zero = zero || drew != (r.runningSum > 0.f);
        return vr::resample(r, st, rng);
    }
};

// evaluate_F_ (VR/ReSTIRHelper.slang:91-423) + evaluate_L_in_volume (:443-496).
// `extra` points at the extra-bounce records (3 floats each) the tap indexes;
// X is the environment orientation of the frame evaluated (isLastFrame picks the
// previous frame's); gridOff = kPrevDensityOffset evaluates on the previous
// frame's grids (useLastFrameGrid, VR/ReSTIRHelper.slang:95-96), else 0.
template <int B, class TP>
VR_DEV v3 evaluateF(const Scene& S, const EnvXf& X, Res tap, const float* extra, Ray ray, Rng& rng, const Options& opt,
                    bool noReuse, bool isFinal, TP& tp, int gridOff = 0) {
    bool isBg = tap.depth == kTMax;
    ray.tMax = tap.depth;
    v3 F = V3(1.f), Fn = V3(1.f);
    auto D = [&]() -> const v3& { return TP::kSplit ? Fn : F; };  // what the early exits look at
    int maxInd = B > 1 ? decodeMaxIndirect(tap.sampledPixel) : 0;
    bool isSelfEmission = B > 1 ? (maxInd == 0 && tap.lightID == -3) : tap.lightID == -3;
    float vis = 1.f;
    v3 pW = at(ray, ray.tMax);
    MediumInteraction mi{pW, -ray.d, S.g, true};
    {
        float density = (isBg || noReuse) ? 1.f : tp.density(S, gridOff, pW);
        if (density == 0.f) return V3(0.f);
        if (!noReuse) vis = tp.vis(S, ray, rng, opt, 0, isFinal, gridOff);
        v3 ss = isBg ? V3(1.f) : (isSelfEmission ? ld3(S.sigma_a) : ld3(S.sigma_s));
        if (noReuse && !isBg) ss = ss / S.sigma_t;
        F = F * (V3(vis * density) * ss);
        if (TP::kSplit) Fn = Fn * (V3(density) * ss);
    }
    if (anyPos(D())) {
        if (isBg) {
            F = F * envEval(S, X, ray.d);
        } else if (isSelfEmission) {
            F = F * emissionWorld(S, pW, gridOff != 0);
        } else {
            bool scatterEmission = false;
            if (B > 1 && maxInd > 0) {
                Ray sr{V3(0.f), V3(0.f), 0.f, 0.f};
                scatterEmission = decodePathTag(tap.sampledPixel) == 1;
                for (int bi = 0; bi < B - 1; bi++) {
                    if (bi >= maxInd) break;
                    bool curEmissive = scatterEmission && bi == maxInd - 1;
                    const float* e = extra + 3 * (size_t)(tap.exStart + bi);
                    float wdist;
                    v3 wdir = decodeWi(V3(e[0], e[1], e[2]), wdist);
                    if (curEmissive) {
                        wdir = V3(tap.uvx, tap.uvy, __int_as_float(tap.lightID));
                        wdist = -1.f;
                    }
                    if (wdist == kTMax) return V3(0.f);
                    float dist = 1.f;
                    if (wdist == -1.f) {
                        v3 disp = wdir - pW;
                        dist = len3(disp);
                        sr = {pW, norm3(disp), 0.f, dist};
                    } else {
                        sr = {pW, wdir, 0.f, wdist};
                    }
                    const float ph = phaseHG(dot3(mi.wo, sr.d), mi.g);
                    F = F * ph;
                    if (TP::kSplit) Fn = Fn * ph;
                    if (allZero(D())) return V3(0.f);
                    pW = wdist == -1.f ? wdir : at(sr, sr.tMax);
                    v3 ss;
                    float sd;
                    if (!noReuse) {
                        ss = curEmissive ? ld3(S.sigma_a) : ld3(S.sigma_s);
                        sd = fmaxf(0.f, tp.density(S, gridOff, pW));
                    } else {
                        ss = curEmissive ? ld3(S.sigma_a) / S.sigma_t : ld3(S.sigma_s) / S.sigma_t;
                        sd = 1.f;
                    }
                    F = F * (V3(sd) * ss);
                    if (TP::kSplit) Fn = Fn * (V3(sd) * ss);
                    if (curEmissive) {
                        F = F * (1.f / (dist * dist));
                        if (TP::kSplit) Fn = Fn * (1.f / (dist * dist));
                    }
                    if (allZero(D())) return V3(0.f);
                    float sv = 1.f;
                    if (!noReuse) sv = tp.vis(S, sr, rng, opt, 0, isFinal, gridOff);
                    F = F * sv;
                    mi.wo = -sr.d;
                    mi.p = pW;
                    if (allZero(D())) return V3(0.f);
                }
                if (scatterEmission) F = F * emissionWorld(S, pW, gridOff != 0);
                else mi = {at(sr, sr.tMax), -sr.d, S.g, true};
            }
            if (!scatterEmission && anyPos(D())) {
                Ray sh;
                v3 Ld;
                if (tap.lightID < 0) {
                    v3 wi = decodeLightUV(S, X, tap.uvx, tap.uvy, tap.lightID);
                    sh = {mi.p, wi, 0.f, kTMax};
                    Ld = envEval(S, X, wi) * phaseHG(dot3(mi.wo, wi), mi.g);
                } else if (tap.lightID < S.numLights) {
                    const float* L = S.lights + 6 * tap.lightID;
                    v3 toL = ld3(L) - mi.p;
                    float d2 = fmaxf(dot3(toL, toL), 1e-9f);
                    float dist = sqrtf(d2);
                    v3 dir = toL / dist;
                    v3 Li = ld3(L + 3) / d2;
                    sh = {mi.p, dir, 0.f, dist};
                    Ld = Li * phaseHG(dot3(mi.wo, dir), mi.g);
                } else {
                    // EmissivePowerSampler::getEmissiveLightSample (LT/EmissivePowerSampler.slang:81-92)
                    const int tri = tap.lightID - S.numLights;
                    if (S.numEmActive == 0 || tri >= S.numEmTris) return V3(0.f);
                    TriLightSample ts;
                    if (!sampleTriangle(S, mi.p, tri, tap.uvx, tap.uvy, ts)) return V3(0.f);
                    sh = {mi.p, ts.dir, 0.f, ts.distance};
                    Ld = ts.Le * S.emMult * phaseHG(dot3(mi.wo, ts.dir), mi.g) * ts.cosTheta /
                         (ts.distance * ts.distance);
                }
                float Tr = tp.vis(S, sh, rng, opt, 1, isFinal, gridOff);
                F = F * (Ld * Tr);
            }
        }
    }
    return F;
}
// The direct form: traces every transmittance ray where evaluateF needs it.
template <int B, int VM, int LM>
VR_DEV v3 evaluateF(const Scene& S, const EnvXf& X, const Res& tap, const float* extra, const Ray& ray, Rng& rng,
                    const Options& opt, bool noReuse, bool isFinal, int gridOff = 0) {
    DirectTrace<VM, LM> tp;
    return evaluateF<B>(S, X, tap, extra, ray, rng, opt, noReuse, isFinal, tp, gridOff);
}

// ComputeInitialSample (VR/ComputeInitialSample.slang:4-395). ex[] receives the
// candidate's extra-bounce records; IM is the initial lighting tracking method.
// cp: the candidate policy (CandDirect, CandRecord or CandReplay).
template <int B, int IM, class CP>
VR_DEV Res computeInitialSample(const Scene& S, const Frame& Fr, Ray ray, float preHit, float prePdf, float preTr,
                                Rng& rng, const Options& opt, v3 ex[B > 1 ? B - 1 : 1], CP& cp) {
    float pathPdf = 1.f, pathPHat = 1.f;
    Res combined = newRes();
    Res out = newRes();
    float primaryDepth = 0.f;
    const bool noReuse = Fr.noReuse != 0;
    for (int bounce = 0; bounce < B; bounce++) {
        out = newRes();
        out.M = 1;
        float curHit, pdfDist = 0.f, Tr;
        MediumInteraction mi{V3(0.f), V3(0.f), S.g, false};
        if (bounce >= 1 || noReuse) {
            if (noReuse) {
                sampleDecomposition(S, ray, rng, mi, 0);
                pdfDist = 1.f;
                Tr = 1.f;
                curHit = mi.valid ? len3(mi.p - ray.o) : kTMax;
            } else {
                int vm = opt.visibilityMipLevel;
                if (Fr.coarserIndirect) vm = min((opt.visibilityMipLevel >= 8 ? 8 : 0) + S.numMips - 1, vm + 1);
                float h[4], pd[4], tr[4];
                cp.analytic(S, ray, rng, opt.visibilityUseLinearSampler != 0, vm, 1, h, pd, tr);
                curHit = h[0]; pdfDist = pd[0]; Tr = tr[0];
                mi = {at(ray, curHit), -ray.d, S.g, curHit != kTMax};
            }
        } else {
            curHit = preHit; pdfDist = prePdf; Tr = preTr;
            mi = {at(ray, curHit), -ray.d, S.g, curHit != kTMax};
        }
        pathPdf *= pdfDist;
        bool hitEmpty = false;
        float actualDensity = 0.f;
        if (bounce == 0) {
            if (B > 1) out.sampledPixel = encodeMaxIndirect(out.sampledPixel, 0);
            out.depth = mi.valid ? curHit : kTMax;
            primaryDepth = out.depth;
            out.p_y = pathPdf;
        } else {
            out.depth = primaryDepth;
            out.sampledPixel = encodeMaxIndirect(out.sampledPixel, bounce);
            if (B > 1) {
                v3 e = encodeWiDist(ray.d, !mi.valid ? kTMax : curHit);
#pragma unroll
                for (int k = 0; k < (B > 1 ? B - 1 : 1); k++)
                    if (k == bounce - 1) ex[k] = e;
            }
            out.p_y = pathPdf;
        }
        if (mi.valid) actualDensity = noReuse ? 1.f : densityWorld(S, 0, mi.p);
        if ((!mi.valid && bounce > 0) || (mi.valid && actualDensity == 0)) {
            out.p_y = 0.f;
            out.runningSum = 0.f;
            hitEmpty = true;
        }
        float pdfDir = 1.f;
        if (!hitEmpty) {
            if (mi.valid) {
                v3 albedo = ld3(S.sigma_s) / S.sigma_t;
                out.lightID = -1;
                out.uvx = 0.f; out.uvy = 0.f;
                float outLightPdf = 0.f;
                v3 Le = actualDensity > 0.f ? emissionWorld(S, mi.p) : V3(0.f);
                v3 oma = V3(1.f) - albedo;
                bool shadow = opt.lightSamples == 0 ? false : true;
                v3 Ld = sampleDirectT(S, Fr.env[0], rng, outLightPdf, mi, opt, shadow, out.lightID, out.uvx, out.uvy,
                                      cp);
                v3 wi = V3(0.f);
                if (B > 1) {
                    float u0 = u01(rng);
                    float u1 = u01(rng);
                    pdfDir = samplePhase(mi.g, -ray.d, wi, u0, u1);
                }
                float p_src = out.p_y;
                float er = lum(oma * Le) / (lum(oma * Le) + lum(albedo * Ld));
                if (isnan(er)) er = 0.f;
                if (u01(rng) < er) {
                    p_src *= er;
                    out.lightID = -3;
                } else {
                    p_src *= outLightPdf * (1 - er);
                }
                out.runningSum = p_src == 0.f ? 0.f : 1.f;
                out.p_y = p_src;
                {
                    float p_y;
                    pathPHat *= Tr;
                    pathPHat *= actualDensity;
                    if (out.lightID == -3) p_y = pathPHat * lum(ld3(S.sigma_a) * Le);
                    else p_y = pathPHat * lum(ld3(S.sigma_s) * Ld * outLightPdf);
                    pathPHat *= lum(ld3(S.sigma_s)) * pdfDir;
                    if (noReuse) { p_y /= S.sigma_t; pathPHat /= S.sigma_t; }
                    if (out.runningSum > 0.f) {
                        out.runningSum = out.p_y == 0.f ? 0.f : p_y / out.p_y;
                        if (out.lightID == -3 && bounce > 0) {
                            out.lightID = __float_as_int(mi.p.z);
                            out.uvx = mi.p.x; out.uvy = mi.p.y;
                            p_y /= (curHit * curHit);
                            out.sampledPixel = encodePathTag(out.sampledPixel, 1);
                        }
                        out.p_y = p_y;
                    }
                }
                pathPdf *= pdfDir;
                if (B > 1 && bounce < B - 1) {
                    ray = {mi.p, wi, 0.f, kTMax};
                    if (Fr.useRussianRoulette && bounce >= 2) {
                        if (u01(rng) < albedo.x) pathPdf *= albedo.x;
                        else { hitEmpty = true; combined.M++; }
                    }
                }
            } else {
                pathPHat *= Tr;
                float p_y = pathPHat * lum(envEval(S, Fr.env[0], ray.d));
                out.runningSum = out.p_y == 0.f ? 0.f : p_y / out.p_y;
                out.p_y = p_y;
                hitEmpty = true;
            }
        }
        if (B > 1) cp.resample(out, combined, rng);
        if (hitEmpty) break;
    }
    if (B > 1) { combined.M = 1; return combined; }
    return out;
}

}  // namespace vr
