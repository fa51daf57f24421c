// vr_pass.hpp -- device-side ReSTIR machinery: lights, reservoirs, the target
// function and candidate construction, templated on the path length B
// (MAX_BOUNCES of the reference) so per-bounce arrays live in registers.
#pragma once
#include "vr_device.hpp"

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
VR_DEV v3 decodeLightUV(float u, float v, int id) {
    v3 d{u, v, sqrtf(1 - u * u - v * v)};
    if (isnan(d.z)) d.z = 0.f;
    if (id == -2) d.z = -d.z;
    return d;
}
struct LightSample { v3 dir, Li, rayDir; float pdf, pdfArea, rayDistance; };

// sampleSceneLights (VR/VolumeUtils.slang:12-149) for a constant environment
// (uniform octahedral warp, pdf 1/(4 pi)) and point lights.
VR_DEV bool sampleLights(const Scene& S, v3 origin, bool useEnv, bool useAnalytic, bool useEmissive, Rng& rng,
                         LightSample& ls, int& lightIndex, float& uvx, float& uvy) {
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
            v3 dir = octToDir(r0, r1);
            float pdf = p0 * (1.f * kInv4PiF);
            ls.rayDir = ls.dir = dir;
            ls.rayDistance = kTMax;
            ls.pdf = pdf;
            ls.pdfArea = pdf;
            ls.Li = pdf > 0.f ? ld3(S.env) / pdf : V3(0.f);
            lightIndex = dir.z < 0 ? -2 : -1;
            uvx = dir.x; uvy = dir.y;
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
    }
    return false;
}

// SampleDirectLighting (VR/VolumeUtils.slang:454-492).
template <int LM>
VR_DEV v3 sampleDirect(const Scene& S, Rng& rng, float& pdf, const MediumInteraction& mi, const Options& o,
                       bool enableShadow, int& lightIndex, float& uvx, float& uvy) {
    pdf = 0.f;
    LightSample ls;
    ls.pdfArea = 0.f;
    bool valid = sampleLights(S, mi.p, o.useEnv, o.useAnalytic, o.useEmissive, rng, ls, lightIndex, uvx, uvy);
    pdf = ls.pdfArea;
    if (!valid) { pdf = 0.f; return V3(0.f); }
    Ray sh{mi.p, ls.rayDir, 0.f, ls.rayDistance};
    if (enableShadow) {
        float vis = visibilityT<LM>(S, sh, rng, o.lightSamples, o.lightingMipLevel, o.lightingUseLinearSampler != 0,
                                    o.lightingTrackingMethod, o.lightingTStepScale);
        ls.Li = ls.Li * vis;
    }
    float p = phaseHG(dot3(mi.wo, ls.dir), mi.g);
    return V3(p) * ls.Li / 1.f;
}

// directLighting (VR/VolumeUtils.slang:416-450): NEE of the reference path tracer.
VR_DEV v3 directLightingPT(const Scene& S, Rng& rng, const MediumInteraction& mi, int nSamples, const Options& o,
                           int mip, uint32_t method) {
    v3 Ld = V3(0.f);
    for (int i = 0; i < nSamples; i++) {
        LightSample ls;
        int idx; float ux, uy;
        bool valid = sampleLights(S, mi.p, o.useEnv, o.useAnalytic, o.useEmissive, rng, ls, idx, ux, uy);
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
template <int VM, int LM>
struct DirectTrace {
    static constexpr bool kSplit = false;
    // kind 0: visibility options; 1: lighting options (grid 0 in final shading)
    VR_DEV float vis(const Scene& S, const Ray& r, Rng& rng, const Options& o, int kind, bool isFinal) {
        if (kind == 0)
            return visibilityT<VM>(S, r, rng, o.visibilitySamples, o.visibilityMipLevel,
                                   o.visibilityUseLinearSampler != 0, o.visibilityTrackingMethod,
                                   o.visibilityTStepScale);
        return visibilityT<LM>(S, r, rng, o.lightSamples, isFinal ? 0 : o.lightingMipLevel,
                               o.lightingUseLinearSampler != 0, o.lightingTrackingMethod, o.lightingTStepScale);
    }
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
    VR_DEV float vis(const Scene&, const Ray& r, Rng&, const Options& o, int kind, bool) {
        const int samples = kind == 0 ? o.visibilitySamples : o.lightSamples;
        if (samples <= 0) return sameSamples(1.f, samples);
        TraceRec rec;
        rec.ox = r.o.x; rec.oy = r.o.y; rec.oz = r.o.z; rec.tMax = r.tMax;
        rec.dx = r.d.x; rec.dy = r.d.y; rec.dz = r.d.z; rec.kind = kind;
        out[n++] = rec;
        return sameSamples(1.f, samples);
    }
};
struct ReplayTrace {
    static constexpr bool kSplit = true;
    const float* tr;  // this pixel's transmittances, in record order
    int n;
    VR_DEV float vis(const Scene&, const Ray&, Rng&, const Options& o, int kind, bool) {
        const int samples = kind == 0 ? o.visibilitySamples : o.lightSamples;
        if (samples <= 0) return sameSamples(1.f, samples);
        return sameSamples(tr[n++], samples);
    }
};

// evaluate_F_ (VR/ReSTIRHelper.slang:91-423) + evaluate_L_in_volume (:443-496).
// `extra` points at the extra-bounce records (3 floats each) the tap indexes.
template <int B, class TP>
VR_DEV v3 evaluateF(const Scene& S, Res tap, const float* extra, Ray ray, Rng& rng, const Options& opt, bool noReuse,
                    bool isFinal, TP& tp) {
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
        float density = (isBg || noReuse) ? 1.f : densityWorld(S, 0, pW);
        if (density == 0.f) return V3(0.f);
        if (!noReuse) vis = tp.vis(S, ray, rng, opt, 0, isFinal);
        v3 ss = isBg ? V3(1.f) : (isSelfEmission ? ld3(S.sigma_a) : ld3(S.sigma_s));
        if (noReuse && !isBg) ss = ss / S.sigma_t;
        F = F * (V3(vis * density) * ss);
        if (TP::kSplit) Fn = Fn * (V3(density) * ss);
    }
    if (anyPos(D())) {
        if (isBg) {
            F = F * ld3(S.env);
        } else if (isSelfEmission) {
            F = F * V3(0.f);  // emission grids are not bound for static density volumes
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
                        sd = fmaxf(0.f, densityWorld(S, 0, pW));
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
                    if (!noReuse) sv = tp.vis(S, sr, rng, opt, 0, isFinal);
                    F = F * sv;
                    mi.wo = -sr.d;
                    mi.p = pW;
                    if (allZero(D())) return V3(0.f);
                }
                if (scatterEmission) F = F * V3(0.f);
                else mi = {at(sr, sr.tMax), -sr.d, S.g, true};
            }
            if (!scatterEmission && anyPos(D())) {
                Ray sh;
                v3 Ld;
                if (tap.lightID < 0) {
                    v3 wi = decodeLightUV(tap.uvx, tap.uvy, tap.lightID);
                    sh = {mi.p, wi, 0.f, kTMax};
                    Ld = ld3(S.env) * phaseHG(dot3(mi.wo, wi), mi.g);
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
                    return V3(0.f);
                }
                float Tr = tp.vis(S, sh, rng, opt, 1, isFinal);
                F = F * (Ld * Tr);
            }
        }
    }
    return F;
}
// The direct form: traces every transmittance ray where evaluateF needs it.
template <int B, int VM, int LM>
VR_DEV v3 evaluateF(const Scene& S, const Res& tap, const float* extra, const Ray& ray, Rng& rng, const Options& opt,
                    bool noReuse, bool isFinal) {
    DirectTrace<VM, LM> tp;
    return evaluateF<B>(S, tap, extra, ray, rng, opt, noReuse, isFinal, tp);
}

// ComputeInitialSample (VR/ComputeInitialSample.slang:4-395). ex[] receives the
// candidate's extra-bounce records; IM is the initial lighting tracking method.
template <int B, int IM>
VR_DEV Res computeInitialSample(const Scene& S, const Frame& Fr, Ray ray, float preHit, float prePdf, float preTr,
                                Rng& rng, const Options& opt, v3 ex[B > 1 ? B - 1 : 1]) {
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
                sampleAnalytic(S, ray, rng, opt.visibilityUseLinearSampler != 0, vm, 1, h, pd, tr);
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
                v3 Le = V3(0.f);
                v3 oma = V3(1.f) - albedo;
                bool shadow = opt.lightSamples == 0 ? false : true;
                v3 Ld = sampleDirect<IM>(S, rng, outLightPdf, mi, opt, shadow, out.lightID, out.uvx, out.uvy);
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
                float p_y = pathPHat * lum(ld3(S.env));
                out.runningSum = out.p_y == 0.f ? 0.f : p_y / out.p_y;
                out.p_y = p_y;
                hitEmpty = true;
            }
        }
        if (B > 1) resample(out, combined, rng);
        if (hitEmpty) break;
    }
    if (B > 1) { combined.M = 1; return combined; }
    return out;
}

}  // namespace vr
