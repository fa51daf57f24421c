// oracle.cpp -- TEST INFRASTRUCTURE ONLY: the CPU oracle of the VolumetricReSTIR pass.
//
// A scalar FP32 restatement of the reference's shaders (FP64 only where the
// reference uses it: the R2 neighbour sequence). It is the checker for the
// HIP product and the CPU baseline of bench.py; nothing in the product links,
// loads or calls it. Functions below cite the reference file:line they follow.
//
// Pinned by closed forms / golden vectors in tests/: the RNG (the reference's
// own xoshiro / splitmix64 sources), indexing and mips, the trackers, the
// neighbour offsets, codecs, WRS, camera, reprojection and the HDDA walk.
// Parity unpinned: D3D12 trilinear / BC4 and Slang transcendental ulps (not
// reproducible without the reference's GPU path); the composition of the
// estimator (evaluateF's product, Talbot weights, RNG draw order across a
// stage), checked only against the HIP path.
#include "vrestir_oracle.h"

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <thread>

#include "oracle_pass.hpp"
#include "oracle_mlp.hpp"

namespace orc {

// IEEE binary16 round-to-nearest-even of a float (storage precision of fp16 bricks).
float roundToHalf(float v) {
    if (std::isnan(v) || std::isinf(v)) return v;
    float a = std::fabs(v);
    if (a >= 65520.f) return std::copysign(INFINITY, v);
    if (a < 0x1p-24f * 0.5f) return std::copysign(0.f, v);
    int e;
    std::frexp(a, &e);             // a = m * 2^e, m in [0.5,1)
    int exp2 = e - 1;              // a in [2^exp2, 2^(exp2+1))
    if (exp2 < -14) exp2 = -14;    // subnormal spacing
    float q = std::ldexp(1.f, exp2 - 10);
    float r = std::nearbyint(a / q) * q;  // default rounding mode: nearest even
    return std::copysign(r, v);
}

struct Oracle {
    SceneState S;
    vrestir_params P;
    bool volumeSet = false, camSet = false;
    int W = 0, H = 0;
    int frameCount = 0;
    bool temporalAccumulated = false;
    bool optionsChanged = true;
    // buffers (VR/VolumetricReSTIR.cpp:419-450)
    std::vector<Features> features, temporalFeatures;
    std::vector<VItem4> vbuffer;  // surface scenes: {triangle, bary, t} per pixel
    std::vector<Reservoir> res[2], temporalRes;
    std::vector<f3> extra[2], temporalExtra;
    std::vector<f4> color[2], output;
    std::vector<f2> mvec;
    int rowBegin = 0, rowEnd = 0, nthreads = 1;
    // per-frame constants
    int numTotalRounds = 0;
    SamplingOptions initialOpt, spatialOpt, finalOpt;
    int totalRoundId = 0;
    double stageSeconds[6] = {0, 0, 0, 0, 0, 0};
    // environment orientation: set now, used by the last frame, and this frame's [current, previous]
    EnvXf envCur{}, envLast{}, frameEnv[2]{};
    bool haveEnvLast = false;
    // volume sequence: frames of slots 0..17 and the velocity field; the bound frame
    struct VolFrame {
        GridSlot grids[18];
        VelocityField velocity;
        int numMips = 0;
        float tStep = 1.f, superVoxelDiag = 1.f;
    };
    std::vector<VolFrame> frames;
    int curFrame = 0;
    bool hasAnimation = false;
    int prevGridOffset = 0;
    bool freezeFrame = false;
    // learned free-flight proposal: per-pixel 16-bit CDF over c, training records
    bool learned = false;
    float learnedAlpha = 0.25f;
    int trainBX = 16, trainBY = 8;
    std::vector<uint16_t> learnedCdf;
    std::vector<float> trainRecs;
    bool learnedTrain = true, learnedJudge = true, learnedInUse = true;
    int learnedSteps = 1, adamStep = 0;
    float learnedLr = 2e-3f;
    std::vector<float> mlpParams = std::vector<float>(mlp::kParams, 0.f), mlpM = std::vector<float>(mlp::kParams, 0.f),
                       mlpV = std::vector<float>(mlp::kParams, 0.f);
    float learnedStats[5] = {0, 0, 0, 1, 0};
};

static int maxB(const Oracle& O) { return O.P.maxBounces; }

template <class F>
static void forRows(Oracle& O, F&& fn) {
    int y0 = O.rowBegin, y1 = O.rowEnd;
    int nt = std::max(1, O.nthreads);
    if (nt == 1 || y1 - y0 < 2) {
        for (int y = y0; y < y1; y++)
            for (int x = 0; x < O.W; x++) fn(x, y);
        return;
    }
    std::vector<std::thread> th;
    for (int t = 0; t < nt; t++)
        th.emplace_back([&, t]() {
            for (int y = y0 + t; y < y1; y += nt)
                for (int x = 0; x < O.W; x++) fn(x, y);
        });
    for (auto& t : th) t.join();
}

// ---------------------------------------------------------------------------
// evaluate_F_ (VR/ReSTIRHelper.slang:91-423), no surface scene, no vertex reuse.
// mipOff = kPrevDensityOffset evaluates on the previous frame's grids
// (useLastFrameGrid, VR/ReSTIRHelper.slang:95-96), else 0.
static f3 evaluateF(const Oracle& O, const EnvXf& X, Reservoir tap, const f3* extraBuf, Ray ray, SampleGen& sg,
                    const SamplingOptions& opt, bool noReuse, bool isFinalShading, int mipOff = 0) {
    const Medium& m = O.S.med;
    bool isBg = tap.depth == kRayTMax;
    ray.tMax = tap.depth;
    f3 F = mk3(1.f);
    int maxInd = maxB(O) > 1 ? decodeMaxIndirect(tap.sampledPixel) : 0;
    bool isSelfEmission = maxB(O) > 1 ? (maxInd == 0 && tap.lightID == -3) : tap.lightID == -3;
    float visibility = 1.f;
    f3 pW = ray.at(ray.tMax);
    MediumInteraction mi{ray.at(ray.tMax), -ray.dir, m.g, true};
    {
        f3 sigma_s = mk3(1.f);
        float density = (isBg || noReuse) ? 1.f : densityWorld(m, pW, mipOff);
        if (density == 0.f) return mk3(0.f);
        if (!noReuse) {
            visibility = computeVisibility(m, ray, sg, opt.visibilitySamples, opt.visibilityMipLevel + mipOff,
                                           opt.visibilityUseLinearSampler, opt.visibilityTrackingMethod,
                                           opt.visibilityTStepScale);
        }
        sigma_s = isBg ? mk3(1.f) : (isSelfEmission ? m.sigma_a : m.sigma_s);
        if (noReuse && !isBg) sigma_s = sigma_s / m.sigma_t;
        F = F * (mk3(visibility * density) * sigma_s);
    }
    int bounceId = 0;
    if (anyPositive(F)) {
        if (isBg) {
            F = F * envEval(O.S, X, ray.dir);
        } else if (isSelfEmission) {
            F = F * emissionWorld(m, pW, mipOff != 0);
        } else {
            bool isScatterSelfEmission = false;
            if (maxB(O) > 1 && maxInd > 0) {
                Ray scatterRay{};
                isScatterSelfEmission = decodePathTag(tap.sampledPixel) == 1;
                int nInd = maxInd;
                for (; bounceId < nInd; bounceId++) {
                    bool curEmissive = isScatterSelfEmission && bounceId == nInd - 1;
                    f4 wd = decodeWiDist(extraBuf[tap.extraBounceStartId + bounceId]);
                    if (curEmissive) {
                        f3 ep = decodeEmissivePosition(tap.lightID, tap.lightUV);
                        wd = {ep.x, ep.y, ep.z, -1.f};
                    }
                    if (wd.w == kRayTMax) return mk3(0.f);
                    float dist = 1.f;
                    if (wd.w == -1.f) {
                        f3 disp = mk3(wd.x, wd.y, wd.z) - pW;
                        dist = length(disp);
                        f3 dir = normalize(disp);
                        scatterRay = {pW, dir, 0.f, dist};
                    } else {
                        scatterRay = {pW, mk3(wd.x, wd.y, wd.z), 0.f, wd.w};
                    }
                    float bsdf = phaseHG(dot(mi.wo, scatterRay.dir), mi.g);
                    F = F * bsdf;
                    if (allZero(F)) return mk3(0.f);
                    if (wd.w == -1.f) pW = mk3(wd.x, wd.y, wd.z);
                    else pW = scatterRay.at(scatterRay.tMax);
                    f3 ss = mk3(1.f);
                    float scatterDensity = 1.f;
                    if (!noReuse) {
                        ss = curEmissive ? m.sigma_a : m.sigma_s;
                        scatterDensity = std::max(0.f, densityWorld(m, pW, mipOff));
                    } else {
                        ss = curEmissive ? m.sigma_a / m.sigma_t : m.sigma_s / m.sigma_t;
                        scatterDensity = 1.f;
                    }
                    F = F * (mk3(scatterDensity) * ss);
                    if (curEmissive) F = F * (1.f / (dist * dist));
                    if (allZero(F)) return mk3(0.f);
                    float scatterVis = 1.f;
                    if (!noReuse)
                        scatterVis = computeVisibility(m, scatterRay, sg, opt.visibilitySamples,
                                                       opt.visibilityMipLevel + mipOff, opt.visibilityUseLinearSampler,
                                                       opt.visibilityTrackingMethod, opt.visibilityTStepScale);
                    F = F * scatterVis;
                    mi.wo = -scatterRay.dir;
                    mi.p = pW;
                    if (allZero(F)) return mk3(0.f);
                }
                if (isScatterSelfEmission) F = F * emissionWorld(m, pW, mipOff != 0);
                else mi = {scatterRay.at(scatterRay.tMax), -scatterRay.dir, m.g, true};
            }
            if (!isScatterSelfEmission && anyPositive(F)) {
                // evaluate_L_in_volume (VR/ReSTIRHelper.slang:443-496)
                bool cull = !isFinalShading;
                Ray sh{};
                f3 Ld = mk3(0.f);
                if (tap.lightID < 0) {
                    f3 wi = decodeLightUV(O.S, X, tap.lightUV, tap.lightID);
                    sh = {mi.p, wi, 0, kRayTMax};
                    Ld = envEval(O.S, X, wi) * phaseHG(dot(mi.wo, wi), mi.g);
                } else if (tap.lightID < (int)O.S.lights.size()) {
                    SceneLightSample ls = analyticLightSample(O.S, tap.lightID, mi.p);
                    sh = {mi.p, ls.rayDir, 0, ls.rayDistance};
                    Ld = ls.Li * phaseHG(dot(mi.wo, ls.rayDir), mi.g);
                } else {
                    TriangleLightSample ls;
                    if (!powerSamplerGetSample(O.S, mi.p, tap.lightID - (int)O.S.lights.size(), tap.lightUV, ls))
                        return mk3(0.f);
                    sh = {mi.p, ls.dir, 0, ls.distance};
                    Ld = ls.Le * O.S.emMult * phaseHG(dot(mi.wo, ls.dir), mi.g) * ls.cosTheta /
                         (ls.distance * ls.distance);
                }
                float Tr = computeVisibility(m, sh, sg, opt.lightSamples, cull ? opt.lightingMipLevel + mipOff : 0,
                                             opt.lightingUseLinearSampler, opt.lightingTrackingMethod,
                                             opt.lightingTStepScale);
                F = F * (Ld * Tr);
            }
        }
    }
    return F;
}

static float evaluatePHat(const Oracle& O, const EnvXf& X, const Ray& ray, SampleGen& sg, const f3* extraBuf,
                          const SamplingOptions& opt, const Reservoir& tap, int mipOff = 0) {
    f3 F = evaluateF(O, X, tap, extraBuf, ray, sg, opt, false, false, mipOff);
    return luminance(F);
}

// ComputeInitialSample (VR/ComputeInitialSample.slang:4-395), no surface scene, no vertex reuse.
static Reservoir computeInitialSample(const Oracle& O, Ray primaryRay, float preHit, float prePdf, float preTr,
                                      int maxBounces, SampleGen& sg, const SamplingOptions& opt,
                                      bool coarserIndirect, bool useRR, bool noReuse, f3* extra) {
    const Medium& m = O.S.med;
    float pathPdf = 1.f, pathPHat = 1.f;
    Ray ray = primaryRay;
    Reservoir combined = newReservoir();
    Reservoir out{};
    float primaryDepth = 0;
    for (int bounce = 0; bounce < maxBounces; bounce++) {
        out = newReservoir();
        out.M = 1;
        float curHit;
        float pdfDist = 0;
        MediumInteraction mi{};
        float Tr;
        if (bounce >= 1 || noReuse) {
            if (noReuse) {
                sampleDecomposition(m, ray, sg, mi, 0);
                pdfDist = 1.f;
                Tr = 1.f;
                curHit = mi.valid ? length(mi.p - ray.origin) : kRayTMax;
            } else {
                int vm = opt.visibilityMipLevel;
                if (coarserIndirect) vm = std::min((opt.visibilityMipLevel >= 8 ? 8 : 0) + m.numMips - 1, vm + 1);
                float h[4], pd[4], tr[4];
                sampleAnalytic(m, ray, sg, opt.visibilityUseLinearSampler, h, vm, pd, tr, 1);
                curHit = h[0];
                pdfDist = pd[0];
                Tr = tr[0];
                mi = {ray.at(curHit), -ray.dir, m.g, curHit != kRayTMax};
            }
        } else {
            curHit = preHit;
            pdfDist = prePdf;
            Tr = preTr;
            mi = {ray.at(curHit), -ray.dir, m.g, curHit != kRayTMax};
        }
        pathPdf *= pdfDist;
        bool hitEmpty = false;
        float actualDensity = 0.f;
        if (bounce == 0) {
            if (maxBounces > 1) out.sampledPixel = encodeMaxIndirect(out.sampledPixel, 0);
            out.depth = mi.valid ? curHit : kRayTMax;
            primaryDepth = out.depth;
            out.p_y = pathPdf;
        } else {
            out.depth = primaryDepth;
            out.sampledPixel = encodeMaxIndirect(out.sampledPixel, bounce);
            extra[bounce - 1] = encodeWiDist({ray.dir.x, ray.dir.y, ray.dir.z, !mi.valid ? kRayTMax : curHit});
            out.p_y = pathPdf;
        }
        if (mi.valid) actualDensity = noReuse ? 1.f : densityWorld(m, mi.p, 0);
        if ((!mi.valid && bounce > 0) || (mi.valid && actualDensity == 0)) {
            out.p_y = 0.f;
            out.runningSum = 0.f;
            hitEmpty = true;
        }
        float pdfDir = 1.f;
        if (!hitEmpty) {
            if (mi.valid) {
                f3 albedo = m.sigma_s / m.sigma_t;
                out.lightID = -1;
                out.lightUV = {0, 0};
                float outLightPdf = 0.f, outVis = 1.f;
                f3 Ld = mk3(0.f), Le = mk3(0.f);
                if (m.hasEmission && actualDensity > 0.f) Le = emissionWorld(m, mi.p);
                f3 oma = mk3(1.f) - albedo;
                bool shouldComputeLightVisibility = opt.lightSamples == 0 ? false : true;
                Ld = sampleDirectLighting(O.S, O.frameEnv[0], sg, outLightPdf, mi, opt.useEnv, opt.useAnalytic, opt.useEmissive,
                                          shouldComputeLightVisibility, opt.lightingMipLevel,
                                          opt.lightingTrackingMethod, opt.lightingUseLinearSampler, opt.lightSamples,
                                          opt.lightingTStepScale, out.lightID, out.lightUV, outVis);
                f3 wo = -ray.dir, wi = mk3(0.f);
                if (maxBounces > 1) {
                    f2 u;
                    u.x = next1D(sg);
                    u.y = next1D(sg);
                    pdfDir = samplePhase(mi.g, wo, wi, u);
                }
                float p_src = out.p_y;
                float er = luminance(oma * Le) / (luminance(oma * Le) + luminance(albedo * Ld));
                if (std::isnan(er)) er = 0.f;
                if (next1D(sg) < er) {
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
                    if (out.lightID == -3) p_y = pathPHat * luminance(m.sigma_a * Le);
                    else p_y = pathPHat * luminance(m.sigma_s * Ld * outLightPdf);
                    pathPHat *= luminance(m.sigma_s) * pdfDir;
                    if (noReuse) {
                        p_y /= m.sigma_t;
                        pathPHat /= m.sigma_t;
                    }
                    if (out.runningSum > 0.f) {
                        out.runningSum = out.p_y == 0.f ? 0.f : p_y / out.p_y;
                        if (out.lightID == -3 && bounce > 0) {
                            encodeEmissivePosition(mi.p, out.lightID, out.lightUV);
                            p_y /= (curHit * curHit);
                            out.sampledPixel = encodePathTag(out.sampledPixel, 1);
                        }
                        out.p_y = p_y;
                    }
                }
                pathPdf *= pdfDir;
                if (maxBounces > 1 && bounce < maxBounces - 1) {
                    ray = {mi.p, wi, 0, kRayTMax};
                    if (useRR && bounce >= 2) {
                        if (next1D(sg) < albedo.x) pathPdf *= albedo.x;
                        else { hitEmpty = true; combined.M++; }
                    }
                }
            } else {
                f3 Le = envEval(O.S, O.frameEnv[0], ray.dir);
                pathPHat *= Tr;
                float p_y = pathPHat * luminance(Le);
                out.runningSum = out.p_y == 0.f ? 0.f : p_y / out.p_y;
                out.p_y = p_y;
                hitEmpty = true;
            }
        }
        if (maxBounces > 1) resampleStep(out, combined, sg);
        if (hitEmpty) break;
    }
    if (maxBounces > 1) { combined.M = 1; return combined; }
    return out;
}

// IntegrateByVolumePathTracing (VR/VolumePathTracingFunctions.slang:3-132), NEE on.
static f3 pathTrace(const Oracle& O, Ray ray, SampleGen& sg, int lightSamples, bool useEnv, bool useAnalytic,
                    bool useEmissive, int mip, int maxBounces, bool useRR) {
    const Medium& m = O.S.med;
    MediumInteraction mi{};
    f3 beta = mk3(1.f), L = mk3(0.f);
    for (int bounce = 0; bounce < maxBounces; bounce++) {
        mi.valid = false;
        sampleDecomposition(m, ray, sg, mi, mip);
        if (mi.valid) {
            f3 albedo = m.sigma_s / m.sigma_t;
            {
                f3 Le = emissionWorld(m, mi.p);
                L = L + Le * (mk3(1.f) - albedo) * beta;
            }
            beta = beta * albedo;
            f3 Ld = directLighting(O.S, O.frameEnv[0], sg, mi, lightSamples, useEnv, useAnalytic, useEmissive, mip, 3);
            L = L + beta * Ld;
            f3 wo = -ray.dir, wi = mk3(0.f);
            if (maxBounces > 1) {
                f2 u;
                u.x = next1D(sg);
                u.y = next1D(sg);
                samplePhase(mi.g, wo, wi, u);
            }
            if (bounce < maxBounces - 1) {
                ray = {mi.p, wi, 0, kRayTMax};
                if (useRR && bounce >= 2) {
                    if (next1D(sg) < albedo.x) beta = beta / albedo.x;
                    else bounce = maxBounces;
                }
            }
        } else {
            f3 Le = envEval(O.S, O.frameEnv[0], ray.dir);
            if (bounce == 0) L = L + beta * Le;
            bounce = maxBounces;
        }
    }
    return L;
}

static inline bool inRange(int x, int y, int W, int H) { return x >= 0 && x < W && y >= 0 && y < H; }

// ---------------------------------------------------------------------------
// Kernels.

// GenerateFeatures (VR/GenerateFeatures.cs.slang:57-102).
// With a surface scene: the V-buffer, and the march and flag of VR/GenerateFeatures.cs.slang:78-99.
static void stageFeatures(Oracle& O) {
    forRows(O, [&](int x, int y) {
        SampleGen sg{};  // unused by the adapter
        f3 d = rayDir(O.S.cam, x, y, O.W, O.H);
        Ray r{O.S.cam.pos, d, 0, kRayTMax};
        int noSurface = 1;
        if (O.P.useSurfaceScene) {
            HitInfo hi = findSurfaceHit(O.S, r.origin, r.dir, 0.f, kRayTMax);
            O.vbuffer[y * O.W + x] = VItem4{hi.tri, hi.bary.x, hi.bary.y, hi.t};
            r.tMax = hi.t;
            if (hi.tri >= 0) noSurface = emitsTowards(O.S, hi.tri, -r.dir) ? 1 : 0;
        }
        if (O.P.useReference) return;
        float accT = featureMarch(O.S.med, r, 0, true, O.initialOpt.visibilityTStepScale, sg);
        O.features[y * O.W + x] = {noSurface, accT};
    });
}

// Inputs of the learned proposal for a camera ray: entry point into mip 0's
// index box normalised to [0,1]^3, the world direction, the segment length over
// the box diagonal (a ray that misses: centre, length 0).
static void learnedRaw(const Medium& m, const Ray& rw, float* raw) {
    Ray ri = worldToMedium(m, rw, 0);
    const GridSlot& g = m.grid[0];
    float t0 = 0.f, t1 = 0.f;
    raw[3] = rw.dir.x; raw[4] = rw.dir.y; raw[5] = rw.dir.z; raw[7] = 0.f;
    if (!intersectVolumeBound(m, ri, t0, t1, 0, false)) {
        raw[0] = raw[1] = raw[2] = 0.5f;
        raw[6] = 0.f;
        return;
    }
    f3 e = ri.origin + t0 * ri.dir;
    raw[0] = e.x / g.bmax.x; raw[1] = e.y / g.bmax.y; raw[2] = e.z / g.bmax.z;
    raw[6] = (t1 - t0) * length(ri.dir) / length(g.bmax);
}

// TraceRays (VR/TraceRays.cs.slang:64-201).
static void stageTrace(Oracle& O) {
    const int b = maxB(O);
    forRows(O, [&](int x, int y) {
        SampleGen sg = makeSampleGen(x, y, (uint32_t)(O.numTotalRounds * O.frameCount));
        int rid = y * O.W + x;
        f3 d = rayDir(O.S.cam, x, y, O.W, O.H);
        Ray ray{O.S.cam.pos, d, 0, kRayTMax};
        if (O.P.useReference) {
            f3 avg = mk3(0.f);
            for (int s = 0; s < O.P.baselineSamplePerPixel; s++)
                avg = avg + pathTrace(O, ray, sg, std::max(1, O.initialOpt.lightSamples), O.initialOpt.useEnv,
                                      O.initialOpt.useAnalytic, O.initialOpt.useEmissive, 0, b,
                                      O.P.initialUseRussianRoulette != 0);
            avg = avg / (float)O.P.baselineSamplePerPixel;
            O.color[0][rid] = {avg.x, avg.y, avg.z, 1.f};
            return;
        }
        f3 finalExtra[8] = {};
        Reservoir fin = newReservoir();
        int M = O.P.initialM;
        int rounds = (M + 3) / 4;
        bool noReuse = !O.P.enableSpatialReuse && !O.P.enableTemporalReuse;
        // Learned proposal (no counterpart in the reference): the distances are the
        // analytic tracker's, driven by optical-depth targets -log(1 - c) with c
        // drawn from the pixel's 32-bin CDF, so the source pdf is g(c) times the
        // tracker's sigma_t Tr (and an escape has the CDF mass above 1 - Tr).
        const bool learned = O.learned && !noReuse;
        const uint16_t* cdf = learned ? &O.learnedCdf[(size_t)rid * mlp::kOut] : nullptr;
        float* rec = nullptr;
        if (learned) {
            int slot = mlp::trainSlot(x, y, O.W, O.H, O.trainBX, O.trainBY, O.frameCount);
            if (slot >= 0) {
                rec = &O.trainRecs[(size_t)slot * mlp::kRecFloats];
                for (int k = 0; k < mlp::kRecFloats; k++) rec[k] = 0.f;
                learnedRaw(O.S.med, ray, rec);
                if (!(rec[6] > 0.f)) rec = nullptr;  // a ray that misses the volume does not train
            }
        }
        for (int r = 0; r < rounds; r++) {
            float h[4] = {0, 0, 0, 0}, pd[4] = {0, 0, 0, 0}, tr[4] = {0, 0, 0, 0};
            int rs = r == rounds - 1 ? (M - 4 * (rounds - 1)) : 4;
            float gc[4] = {1, 1, 1, 1};
            int bins[4] = {0, 0, 0, 0};
            if (learned) {
                float targets[4] = {0, 0, 0, 0};
                for (int s = 0; s < rs; s++) {
                    float c = mlp::sampleC(cdf, next1D(sg), bins[s], gc[s]);
                    targets[s] = -std::log(1 - c);
                }
                sampleAnalyticTargets(O.S.med, ray, sg, O.initialOpt.visibilityUseLinearSampler,
                                      O.initialOpt.visibilityMipLevel, targets, rs, h, pd, tr);
            } else if (!noReuse)
                sampleAnalytic(O.S.med, ray, sg, O.initialOpt.visibilityUseLinearSampler, h,
                               O.initialOpt.visibilityMipLevel, pd, tr, rs);
            for (int s = 0; s < rs; s++) {
                f3 ex[8] = {};
                // ratio of the learned source pdf to the analytic one
                float ratio = 1.f;
                if (learned) {
                    if (h[s] == kRayTMax) ratio = tr[s] > 0.f ? mlp::massAbove(cdf, 1.f - tr[s]) / tr[s] : gc[s];
                    else ratio = gc[s];
                }
                Reservoir o = computeInitialSample(O, ray, h[s], pd[s] * ratio, tr[s], b, sg, O.initialOpt,
                                                   O.P.initialUseCoarserGridForIndirectBounce != 0,
                                                   O.P.initialUseRussianRoulette != 0, noReuse, ex);
                if (rec) {
                    // the sample's weight against the analytic proposal over g(c): its
                    // contribution to the c-space target density
                    float wt = o.runningSum * (ratio / gc[s]);
                    rec[7] += wt;
                    rec[mlp::kRecHist + bins[s]] += wt;
                    rec[mlp::kRecSq + bins[s]] += wt * wt * gc[s];
                }
                bool sel = resampleStep(o, fin, sg);
                if (sel && b > 1) {
                    int mi = decodeMaxIndirect(fin.sampledPixel);
                    for (int k = 0; k < mi; k++) finalExtra[k] = ex[k];
                }
            }
        }
        // p-hat re-evaluation with the spatial options (VR/TraceRays.cs.slang:169-183)
        fin.extraBounceStartId = 0;
        float ph = evaluatePHat(O, O.frameEnv[0], ray, sg, finalExtra, O.spatialOpt, fin);
        if (fin.runningSum > 0.f) {
            fin.runningSum *= fin.p_y == 0.f ? 0.f : ph / fin.p_y;
            fin.p_y = ph;
        }
        fin.extraBounceStartId = b > 1 ? rid * (b - 1) : 0;
        O.res[0][rid] = fin;
        if (b > 1) {
            int mi = decodeMaxIndirect(fin.sampledPixel);
            for (int k = 0; k < mi; k++) O.extra[0][fin.extraBounceStartId + k] = finalExtra[k];
        }
    });
}

static Features featuresAt(const std::vector<Features>& f, int id) {
    if (id < 0 || id >= (int)f.size()) return {0, 0.f};  // out-of-range structured-buffer loads read 0
    return f[id];
}

// World point -> continuous pixel coordinates through camera c's view and
// projection (VR/TemporalReuse.cs.slang:168-185).
static f2 projectToPixel(const Camera& c, f3 p, int W, int H) {
    f4 vp = mulRow({p.x, p.y, p.z, 1.f}, c.view);
    f4 cp = mulRow(vp, c.proj);
    float sx = cp.x / cp.w, sy = cp.y / cp.w;
    sx = 0.5f * sx + 0.5f;
    sy = -0.5f * sy + 0.5f;
    sx *= (float)W; sy *= (float)H;
    return {sx, sy};
}

// TemporalReuse (VR/TemporalReuse.cs.slang:80-377).
static void stageTemporal(Oracle& O, int curIdx) {
    if (!O.temporalAccumulated) return;  // gIsFirstFrame
    const int b = maxB(O);
    const Camera& pc = O.S.prevCam;
    std::vector<Reservoir>& cur = O.res[curIdx];
    std::vector<f3>& curExtra = O.extra[curIdx];
    // Threads write only their own pixel of `cur` and read other pixels of the
    // temporal buffers only, so the row-parallel loop is race-free.
    forRows(O, [&](int x, int y) {
        SampleGen sg = makeSampleGen(x, y, (uint32_t)(O.numTotalRounds * O.frameCount + 1));
        int selectedId = -1;
        int pixelId = y * O.W + x;
        int numUsed = 1;
        Reservoir taps[2];
        taps[0] = cur[pixelId];
        f3 d = rayDir(O.S.cam, x, y, O.W, O.H);
        Ray ray{O.S.cam.pos, d, 0, kRayTMax};
        bool talbot = O.P.temporalMISMethod == 1;
        Reservoir output = talbot ? newReservoir() : taps[0];
        int centerExtra = taps[0].extraBounceStartId;
        float temporalOriginalDepth = 0.f;
        int rsx = 0, rsy = 0;
        Features cf = O.features[pixelId];
        bool isBgRes = cf.transmittance == 1.f && cf.noReflectiveSurface;
        bool useFallback = true;
        if (O.P.temporalReprojectionMode != 1) {
            float reprojDepth = taps[0].depth;
            if (reprojDepth == kRayTMax && O.P.temporalReprojectionMode != 2 && !isBgRes)
                reprojDepth = rejectionSampleDepth(O.S.med, ray, sg, 8 + O.P.temporalReprojectionMipLevel);
            f3 p = ray.origin + ray.dir * reprojDepth;
            // VR/TemporalReuse.cs.slang:161-166
            if (O.S.med.velocity.valid && O.hasAnimation) p = p - velocityWorld(O.S.med, p) * O.S.med.velocityScale;
            float sx, sy;
            int sxi, syi;
            if (reprojDepth == kRayTMax) {
                sx = (float)x + 0.5f; sy = (float)y + 0.5f;
                sxi = x; syi = y;
            } else {
                f2 sp = projectToPixel(pc, p, O.W, O.H);
                sx = sp.x; sy = sp.y;
                sxi = (int)sx; syi = (int)sy;
            }
            {
                Features tf = featuresAt(O.temporalFeatures, syi * O.W + sxi);
                bool isTapBg = tf.transmittance == 1.f && tf.noReflectiveSurface;
                if (isBgRes && !isTapBg) {
                    if (!O.mvec.empty()) O.mvec[pixelId] = {(float)(rsx - x) / O.W, (float)(rsy - y) / O.H};
                    return;
                }
            }
            sxi = (int)sx; syi = (int)sy;
            rsx = sxi; rsy = syi;
            if (inRange(sxi, syi, O.W, O.H)) {
                numUsed++;
                taps[1] = O.temporalRes[syi * O.W + sxi];
            }
            if (numUsed > 1) useFallback = false;
        }
        if (useFallback) {
            numUsed++;
            rsx = x; rsy = y;
            taps[1] = O.temporalRes[pixelId];
        }
        if (!O.mvec.empty()) O.mvec[pixelId] = {(float)(rsx - x) / O.W, (float)(rsy - y) / O.H};
        float curM = taps[0].M;
        float maxPrevM = O.P.temporalReuseMThreshold * curM;
        if (numUsed == 2) {
            temporalOriginalDepth = taps[1].depth;
            if (taps[1].depth != kRayTMax) {
                f3 pd = normalize(rayDirNN(pc, rsx, rsy, O.W, O.H));
                f3 wp = pc.pos + taps[1].depth * pd;
                taps[1].depth = length(wp - ray.origin);
            }
        }
        float centerPrevDepth = taps[0].depth;
        if (centerPrevDepth != kRayTMax) {
            f3 wp = ray.at(centerPrevDepth);
            centerPrevDepth = length(wp - pc.pos);
        }
        bool hasSel = output.runningSum > 0.f;
        int start = talbot ? 0 : 1;
        if (start == 1) selectedId = 0;
        for (int i = start; i < numUsed; i++) {
            float talbotW = 1.f;
            float nPy = 0.f;
            if (taps[i].p_y > 0.f) {
                nPy = taps[i].p_y;
                if (std::isnan(taps[i].runningSum) || std::isinf(taps[i].runningSum)) taps[i].runningSum = 0.f;
                if (i > 0 && taps[i].runningSum != 0.f) {
                    // resampleNeighbor (VR/ReSTIRHelper.slang:582-598)
                    float ph = evaluatePHat(O, O.frameEnv[0], ray, sg, O.temporalExtra.data(), O.spatialOpt, taps[i]);
                    float w = ph / taps[i].p_y;
                    if (std::isinf(w) || std::isnan(w)) w = 0.f;
                    taps[i].runningSum *= w;
                    taps[i].p_y = ph;
                }
            } else {
                taps[i].p_y = 0.f;
                taps[i].runningSum = 0.f;
            }
            if (talbot && taps[i].runningSum > 0.f) {
                float psum = 0, pqi = 0, k = 0;
                for (int j = 0; j < numUsed; j++) {
                    int tx = j == 0 ? x : rsx, ty = j == 0 ? y : rsy;
                    float cM = std::min(maxPrevM, taps[j].M);
                    k += cM;
                    if (j == 0) {
                        pqi = taps[i].p_y;
                        psum += taps[i].p_y * cM;
                    } else if (i == j) {
                        pqi = nPy;
                        psum += nPy * cM;
                    } else {
                        Ray nr;
                        if (j == 0) nr = {O.S.cam.pos, rayDir(O.S.cam, tx, ty, O.W, O.H), 0, 0};
                        else nr = {pc.pos, normalize(rayDirNN(pc, tx, ty, O.W, O.H)), 0, 0};
                        float usedDepth = j == 0 ? taps[i].depth : (i == 0 ? centerPrevDepth : temporalOriginalDepth);
                        nr.tMax = usedDepth;
                        float backup = taps[i].depth;
                        taps[i].depth = usedDepth;
                        // isLastFrame = j > 0 (VR/TemporalReuse.cs.slang:335-338)
                        float py = evaluatePHat(O, O.frameEnv[j > 0 ? 1 : 0], nr, sg,
                                                i == 0 ? curExtra.data() : O.temporalExtra.data(), O.spatialOpt,
                                                taps[i], j > 0 ? O.prevGridOffset : 0);
                        taps[i].depth = backup;
                        if (std::isinf(py) || std::isnan(py)) py = 0.f;
                        psum += py * cM;
                    }
                }
                if (psum > 0) talbotW = pqi * k / psum;
            }
            taps[i].runningSum *= talbotW;
            bool curSel = resampleStepMaxM(taps[i], maxPrevM, output, sg);
            hasSel |= curSel;
            if (curSel) selectedId = i;
        }
        if (b > 1) {
            if (hasSel && selectedId > 0) {
                int mi = decodeMaxIndirect(output.sampledPixel);
                for (int k = 0; k < mi; k++) curExtra[centerExtra + k] = O.temporalExtra[output.extraBounceStartId + k];
            }
            output.extraBounceStartId = centerExtra;
        }
        cur[pixelId] = output;
    });
}

// R2 / Hammersley neighbour offsets (VR/SpatialReuse.cs.slang:64-81). The R2
// constants are single-precision literals widened to double.
static void neighborOffset(int sampleId, int sampleCount, float radius, int frameSeed, int upperBound, int sampler,
                           int& ox, int& oy) {
    f2 u;
    if (sampler == 0) {
        uint32_t i = (uint32_t)sampleId;
        i = (i & 0x55555555u) << 1 | (i & 0xAAAAAAAAu) >> 1;
        i = (i & 0x33333333u) << 2 | (i & 0xCCCCCCCCu) >> 2;
        i = (i & 0x0F0F0F0Fu) << 4 | (i & 0xF0F0F0F0u) >> 4;
        i = (i & 0x00FF00FFu) << 8 | (i & 0xFF00FF00u) >> 8;
        i = (i << 16) | (i >> 16);
        u = {(float)sampleId / (float)sampleCount, (float)i * 2.3283064365386963e-10f};
    } else {
        double mult = (double)frameSeed * upperBound + sampleId;
        if (sampleId == 0) u = {0, 0};
        else {
            double a = (double)0.754877669f * mult, bb = (double)0.569840296f * mult;
            u = {(float)(a - std::floor(a)), (float)(bb - std::floor(bb))};
        }
    }
    float r = std::sqrt(u.x);
    float phi = 2.f * kPi * u.y;
    ox = (int)(radius * (r * std::cos(phi)));
    oy = (int)(radius * (r * std::sin(phi)));
}

// SpatialReuse (VR/SpatialReuse.cs.slang:94-265).
static void stageSpatial(Oracle& O, int roundId) {
    const int b = maxB(O);
    int inIdx = O.totalRoundId % 2, outIdx = (O.totalRoundId + 1) % 2;
    const std::vector<Reservoir>& in = O.res[inIdx];
    std::vector<Reservoir>& out = O.res[outIdx];
    const std::vector<f3>& inExtra = O.extra[inIdx];
    std::vector<f3>& outExtra = O.extra[outIdx];
    int roundOffset = (O.P.enableTemporalReuse ? 1 : 0) + 1;
    int numRounds = O.P.spatialReuseRounds + roundOffset + 1;
    int rid = roundId + roundOffset;
    int r2Seed = ((O.P.spatialReuseRounds + 1) * O.frameCount + roundId) % 16;
    bool talbot = O.P.spatialMISMethod == 1;
    int K = O.P.spatialSampleCount;
    std::vector<int> offx(K), offy(K);
    for (int s = 0; s < K; s++)
        neighborOffset(s, K, O.P.sampleRadius, r2Seed, O.P.spatialSampleCount, (int)O.P.randomSamplerType, offx[s],
                       offy[s]);
    forRows(O, [&](int x, int y) {
        SampleGen sg = makeSampleGen(x, y, (uint32_t)(numRounds * O.frameCount + rid));
        int pixelId = y * O.W + x;
        Reservoir output = in[pixelId];
        int centerExtra = output.extraBounceStartId;
        f3 d = rayDir(O.S.cam, x, y, O.W, O.H);
        Ray ray{O.S.cam.pos, d, 0, kRayTMax};
        if (talbot) output = newReservoir();
        bool hasSel = output.runningSum > 0.f;
        Features cf = O.features[pixelId];
        bool selfBg = !(cf.transmittance != 1.f);
        if (selfBg) {
            out[pixelId] = talbot ? in[pixelId] : output;
            if (b > 1) {
                int mi = decodeMaxIndirect(output.sampledPixel);
                for (int k = 0; k < mi; k++) outExtra[output.extraBounceStartId + k] = inExtra[output.extraBounceStartId + k];
            }
            return;
        }
        int start = talbot ? 0 : 1;
        for (int s = start; s < K; s++) {
            int tx = x + offx[s], ty = y + offy[s];
            if (!inRange(tx, ty, O.W, O.H)) continue;
            Reservoir tap = in[ty * O.W + tx];
            float misW = 1.f;
            if (s > 0) {
                // resampleNeighborSpatialReuse (VR/ReSTIRHelper.slang:563-580)
                if (tap.runningSum != 0.f) {
                    float ph = evaluatePHat(O, O.frameEnv[0], ray, sg, inExtra.data(), O.spatialOpt, tap);
                    float w = ph / tap.p_y;
                    if (std::isinf(w) || std::isnan(w)) w = 0.f;
                    tap.runningSum *= w;
                    tap.p_y = ph;
                }
            }
            if (talbot && tap.runningSum > 0.f) {
                float psum = 0, pqi = 0, k = 0;
                for (int j = 0; j < K; j++) {
                    int t2x = x + offx[j], t2y = y + offy[j];
                    if (!inRange(t2x, t2y, O.W, O.H)) continue;
                    const Reservoir& tap2 = in[t2y * O.W + t2x];
                    k += tap2.M;
                    if (j == 0) {
                        pqi = tap.p_y;
                        psum += tap.p_y * tap2.M;
                    } else if (s == j) {
                        pqi = tap2.p_y;
                        psum += tap2.p_y * tap2.M;
                    } else {
                        Ray nr{ray.origin, rayDir(O.S.cam, t2x, t2y, O.W, O.H), 0, tap.depth};
                        float py = evaluatePHat(O, O.frameEnv[0], nr, sg, inExtra.data(), O.spatialOpt, tap);
                        if (std::isinf(py) || std::isnan(py)) py = 0.f;
                        psum += py * tap2.M;
                    }
                }
                if (psum > 0) misW = pqi * k / psum;
            }
            tap.runningSum *= misW;
            if (resampleStep(tap, output, sg)) hasSel = true;
        }
        if (b > 1 && hasSel) {
            int mi = decodeMaxIndirect(output.sampledPixel);
            for (int k = 0; k < mi; k++) outExtra[centerExtra + k] = inExtra[output.extraBounceStartId + k];
        }
        output.extraBounceStartId = b > 1 ? centerExtra : output.extraBounceStartId;
        out[pixelId] = output;
    });
}

// CopyReservoirs (VR/CopyReservoirs.cs.slang:84-95).
static void stageCopy(Oracle& O) {
    const int b = maxB(O);
    int idx = O.totalRoundId % 2;
    forRows(O, [&](int x, int y) {
        int pid = y * O.W + x;
        Reservoir r = O.res[idx][pid];
        O.temporalRes[pid] = r;
        if (b > 1) {
            int mi = decodeMaxIndirect(r.sampledPixel);
            for (int k = 0; k < mi; k++) O.temporalExtra[r.extraBounceStartId + k] = O.extra[idx][r.extraBounceStartId + k];
        }
    });
}

// FinalShading (VR/FinalShading.cs.slang:71-141).
static void stageFinal(Oracle& O) {
    int idx = O.totalRoundId % 2;
    bool noReuse = !O.P.enableSpatialReuse && !O.P.enableTemporalReuse;
    forRows(O, [&](int x, int y) {
        int pid = y * O.W + x;
        SampleGen sg = makeSampleGen(x, y, (uint32_t)(O.numTotalRounds * O.frameCount + O.numTotalRounds - 1));
        f3 outC = mk3(0.f);
        if (O.P.useReference) {
            f4 c = O.color[idx][pid];
            outC = {c.x, c.y, c.z};
        } else if (O.P.visualizeTotalTransmittance) {
            outC = mk3(std::pow(O.features[pid].transmittance, 2.2f));
        } else {
            Reservoir r = O.res[idx][pid];
            if (r.runningSum > 0.f) {
                f3 d = rayDir(O.S.cam, x, y, O.W, O.H);
                Ray ray{O.S.cam.pos, d, 0, r.depth};
                f3 c = evaluateF(O, O.frameEnv[0], r, O.extra[idx].data(), ray, sg, O.finalOpt, noReuse, true);
                float Wt = r.p_y == 0.0f ? 1.f : r.runningSum / (r.p_y * r.M);
                outC = outC + c * Wt;
            }
        }
        f4 col{outC.x, outC.y, outC.z, 1.f};
        if (std::isnan(col.x) || std::isnan(col.y) || std::isnan(col.z) || std::isinf(col.x) || std::isinf(col.y) ||
            std::isinf(col.z))
            col = {0, 0, 0, 0};
        O.output[pid] = col;
    });
}

// Frame set-up: option structs and round bookkeeping (VR/VolumetricReSTIR.cpp:339-496).
static void prepareFrame(Oracle& O, int W, int H) {
    const vrestir_params& p = O.P;
    bool sizeChanged = W != O.W || H != O.H;
    if (O.optionsChanged) {
        O.frameCount = 0;
        O.temporalAccumulated = false;
        O.optionsChanged = false;
    }
    if (sizeChanged || O.res[0].empty()) {
        O.W = W; O.H = H;
        size_t n = (size_t)W * H;
        O.features.assign(n, {0, 0.f});
        O.vbuffer.assign(n, VItem4{-1, 0.f, 0.f, kRayTMax});
        O.temporalFeatures.assign(n, {0, 0.f});
        for (int i = 0; i < 2; i++) {
            O.res[i].assign(n, Reservoir{0, 0, 0, 0, {0, 0}, 0, 0, 0});
            O.color[i].assign(n, {0, 0, 0, 0});
        }
        O.temporalRes.assign(n, Reservoir{0, 0, 0, 0, {0, 0}, 0, 0, 0});
        size_t ne = n * (size_t)std::max(0, p.maxBounces - 1);
        for (int i = 0; i < 2; i++) O.extra[i].assign(ne, {0, 0, 0});
        O.temporalExtra.assign(ne, {0, 0, 0});
        O.output.assign(n, {0, 0, 0, 0});
        O.mvec.assign(n, {0, 0});
        O.temporalAccumulated = false;
        O.learnedCdf.assign(n * mlp::kOut, 0);
        for (size_t i = 0; i < n; i++)
            for (int j = 0; j < mlp::kOut; j++) O.learnedCdf[i * mlp::kOut + j] = (uint16_t)(2048 * j);
    }
    {
        size_t nr = (size_t)mlp::trainSlots(W, H, O.trainBX, O.trainBY) * mlp::kRecFloats;
        if (O.trainRecs.size() != nr) O.trainRecs.assign(nr, 0.f);
    }
    O.rowBegin = std::max(0, std::min(O.rowBegin, H));
    O.rowEnd = O.rowEnd <= 0 ? H : std::min(O.rowEnd, H);
    O.numTotalRounds = (p.enableSpatialReuse ? p.spatialReuseRounds : 0) + (p.enableTemporalReuse ? 1 : 0) + 1 + 1;
    O.initialOpt = {VRESTIR_ANALYTIC_TRACKING, p.initialLightingTrackingMethod, p.initialLightSamples,
                    p.initialLightingMipLevel, 1,
                    p.initialVisibilityUseLinearSampler ? p.initialBaseMipLevel : p.initialBaseMipLevel + 8,
                    p.initialVisibilityUseLinearSampler != 0, p.initialLightingUseLinearSampler != 0,
                    p.initialVisibilityTStepScale, p.initialLightingTStepScale, p.useEnvironmentLights != 0,
                    p.useAnalyticLights != 0, p.useEmissiveLights != 0, p.vertexReuseStartBounce};
    O.spatialOpt = {p.spatialVisibilityTrackingMethod, p.spatialLightingTrackingMethod, 1, p.spatialLightingMipLevel, 1,
                    p.spatialVisibilityMipLevel, p.spatialVisibilityUseLinearSampler != 0,
                    p.spatialLightingUseLinearSampler != 0, p.spatialVisibilityTStepScale, p.spatialLightingTStepScale,
                    p.useEnvironmentLights != 0, p.useAnalyticLights != 0, p.useEmissiveLights != 0,
                    p.vertexReuseStartBounce};
    bool fa = p.finalLightTrackingMethod == 1 || p.finalLightTrackingMethod == 2;
    bool fv = p.finalVisibilityTrackingMethod == 1 || p.finalVisibilityTrackingMethod == 2;
    O.finalOpt = {p.finalVisibilityTrackingMethod, p.finalLightTrackingMethod, fa ? 1 : p.finalLightSamples, 0,
                  fv ? 1 : p.finalVisibilitySamples, 0, true, true, p.finalTStepScale, p.finalTStepScale,
                  p.useEnvironmentLights != 0, p.useAnalyticLights != 0, p.useEmissiveLights != 0,
                  p.vertexReuseStartBounce};
    O.totalRoundId = 0;
    O.prevGridOffset = (p.usePrevVolumeForReproj && O.hasAnimation) ? kPrevDensityOffset : 0;
    // EnvMap data / prevData of this frame (SC/Scene.cpp:761-775)
    O.frameEnv[0] = O.envCur;
    O.frameEnv[1] = O.haveEnvLast ? O.envLast : O.envCur;
}

// The learned proposal's CDF for every camera ray (stage 6): the network in its
// bf16-operand form, or the uniform CDF while the judge keeps it out.
static void stageLearnedProposal(Oracle& O) {
    forRows(O, [&](int x, int y) {
        uint16_t* cdf = &O.learnedCdf[((size_t)y * O.W + x) * mlp::kOut];
        if (!O.learnedInUse) {
            for (int j = 0; j < mlp::kOut; j++) cdf[j] = (uint16_t)(2048 * j);
            return;
        }
        Ray ray{O.S.cam.pos, rayDir(O.S.cam, x, y, O.W, O.H), 0, kRayTMax};
        float raw[mlp::kRaw];
        learnedRaw(O.S.med, ray, raw);
        mlp::Acts<float> a;
        mlp::forward<float>(O.mlpParams.data(), raw, O.learnedAlpha, true, a);
        mlp::quantize(a.p, cdf);
    });
}

// Adam steps on this frame's training rays (stage 7): the records split into
// learnedSteps consecutive minibatches, each step on the mean cross-entropy of
// its minibatch; the judge compares the two second moments of the last step.
// The records are consumed (their weight cleared).
static void stageLearnedTrain(Oracle& O) {
    if (!O.learnedTrain) return;
    const int n = mlp::trainSlots(O.W, O.H, O.trainBX, O.trainBY);
    const int K = std::max(1, O.learnedSteps);
    for (int k = 0; k < K; k++) {
        const int b0 = (int)((int64_t)n * k / K), b1 = (int)((int64_t)n * (k + 1) / K);
        if (b1 <= b0) continue;
        std::vector<double> g(mlp::kParams, 0.0);
        double st[3] = {0, 0, 0};
        const float scale = 1.f / (float)(b1 - b0);
        for (int r = b0; r < b1; r++)
            mlp::backward<float>(O.mlpParams.data(), &O.trainRecs[(size_t)r * mlp::kRecFloats], O.learnedAlpha, scale,
                                 true, g.data(), st);
        std::vector<float> gf(g.begin(), g.end());
        O.adamStep++;
        mlp::adam(O.mlpParams.data(), O.mlpM.data(), O.mlpV.data(), gf.data(), mlp::kParams, O.adamStep, O.learnedLr,
                  mlp::kAdamB1, mlp::kAdamB2, mlp::kAdamEps);
        O.learnedStats[0] = (float)st[0];
        O.learnedStats[1] = (float)st[1];
        O.learnedStats[2] = (float)st[2];
    }
    O.learnedInUse = !O.learnedJudge || O.learnedStats[2] < O.learnedStats[1];
    O.learnedStats[3] = O.learnedInUse ? 1.f : 0.f;
    O.learnedStats[4] = (float)O.adamStep;
    for (int r = 0; r < n; r++) O.trainRecs[(size_t)r * mlp::kRecFloats + 7] = 0.f;
}

static void endFrame(Oracle& O) {
    O.temporalAccumulated = true;
    O.S.prevCam = O.S.cam;
    O.envLast = O.envCur;
    O.haveEnvLast = true;
    O.frameCount++;
}

template <class F>
static void timed(Oracle& O, int stage, F&& fn) {
    auto t0 = std::chrono::steady_clock::now();
    fn();
    O.stageSeconds[stage] += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

static void runStage(Oracle& O, int stage, int round) {
    const vrestir_params& p = O.P;
    switch (stage) {
        case VRESTIR_STAGE_GENERATE_FEATURES: timed(O, 0, [&] { stageFeatures(O); }); break;
        case VRESTIR_STAGE_TRACE_RAYS: timed(O, 1, [&] { stageTrace(O); }); break;
        case VRESTIR_STAGE_TEMPORAL_REUSE:
            timed(O, 2, [&] {
                stageTemporal(O, O.totalRoundId);
                if (!p.enableSpatialReuse) {
                    O.temporalRes = O.res[O.totalRoundId];
                    O.temporalExtra = O.extra[O.totalRoundId];
                }
                O.temporalFeatures = O.features;
            });
            break;
        case VRESTIR_STAGE_SPATIAL_REUSE:
            timed(O, 3, [&] { stageSpatial(O, round); });
            O.totalRoundId++;
            break;
        case VRESTIR_STAGE_COPY_RESERVOIRS: timed(O, 4, [&] { stageCopy(O); }); break;
        case VRESTIR_STAGE_FINAL_SHADING: timed(O, 5, [&] { stageFinal(O); }); break;
        case VRESTIR_STAGE_LEARNED_PROPOSAL: timed(O, 1, [&] { stageLearnedProposal(O); }); break;
        case VRESTIR_STAGE_LEARNED_TRAIN: timed(O, 1, [&] { stageLearnedTrain(O); }); break;
    }
}
static bool learnedActive(const Oracle& O) {
    return O.learned && !O.P.useReference && (O.P.enableSpatialReuse || O.P.enableTemporalReuse);
}

// VolumetricReSTIR::execute (VR/VolumetricReSTIR.cpp:303-772), dispatch order.
static void executeFrame(Oracle& O, int W, int H) {
    const vrestir_params& p = O.P;
    prepareFrame(O, W, H);
    if (O.freezeFrame) {
        // mFreezeFrame (VR/VolumetricReSTIR.cpp:499-780): only final shading runs, on the
        // reservoirs of totalRoundId = [spatial], with the previous frame count
        O.totalRoundId = p.enableSpatialReuse ? 1 : 0;
        O.frameCount--;
        runStage(O, VRESTIR_STAGE_FINAL_SHADING, 0);
        O.frameCount++;
        O.temporalAccumulated = true;
        O.S.prevCam = O.S.cam;
        O.envLast = O.envCur;
        O.haveEnvLast = true;
        return;
    }
    runStage(O, VRESTIR_STAGE_GENERATE_FEATURES, 0);
    if (learnedActive(O)) runStage(O, VRESTIR_STAGE_LEARNED_PROPOSAL, 0);
    runStage(O, VRESTIR_STAGE_TRACE_RAYS, 0);
    if (learnedActive(O)) runStage(O, VRESTIR_STAGE_LEARNED_TRAIN, 0);
    if (!p.useReference && p.enableTemporalReuse) runStage(O, VRESTIR_STAGE_TEMPORAL_REUSE, 0);
    if (!p.useReference && p.enableSpatialReuse)
        for (int r = 0; r < p.spatialReuseRounds; r++) runStage(O, VRESTIR_STAGE_SPATIAL_REUSE, r);
    if (!p.useReference && p.enableTemporalReuse) runStage(O, VRESTIR_STAGE_COPY_RESERVOIRS, 0);
    runStage(O, VRESTIR_STAGE_FINAL_SHADING, 0);
    endFrame(O);
}

}  // namespace orc

using namespace orc;

// ---------------------------------------------------------------------------
// C interface for the tests (ctypes) and bench.py's CPU baseline.
extern "C" {

void* orc_create(void) {
    Oracle* O = new Oracle();
    vrestir_default_params_oracle(&O->P);
    const float zero[3] = {0.f, 0.f, 0.f};
    O->envCur = O->envLast = envRotation(zero);
    O->frameEnv[0] = O->frameEnv[1] = O->envCur;
    return O;
}
void orc_destroy(void* h) { delete (Oracle*)h; }

// Defaults of VolumetricReSTIRParams (VR/VolumetricReSTIR.h:132-206).
void vrestir_default_params_oracle(vrestir_params* p) {
    std::memset(p, 0, sizeof(*p));
    p->maxBounces = 1;
    p->enableTemporalReuse = 1;
    p->enableSpatialReuse = 1;
    p->vertexReuse = 0;
    p->vertexReuseStartBounce = 1;
    p->useReference = 0;
    p->useEnvironmentLights = 1;
    p->useAnalyticLights = 0;
    p->useEmissiveLights = 0;
    p->baselineSamplePerPixel = 1;
    p->visualizeTotalTransmittance = 0;
    p->useSurfaceScene = 0;
    p->usePrevVolumeForReproj = 1;
    p->initialBaseMipLevel = 1;
    p->initialM = 4;
    p->initialLightSamples = 1;
    p->initialLightingMipLevel = 2;
    p->initialVisibilityUseLinearSampler = 0;
    p->initialLightingUseLinearSampler = 1;
    p->initialLightingTrackingMethod = VRESTIR_RAY_MARCHING;
    p->initialVisibilityTStepScale = 1.f;
    p->initialLightingTStepScale = 2.f;
    p->initialUseRussianRoulette = 1;
    p->initialUseCoarserGridForIndirectBounce = 1;
    p->temporalReuseMThreshold = 4.f;
    p->temporalReprojectionMode = VRESTIR_REPROJECTION_LINEAR;
    p->temporalMISMethod = VRESTIR_MIS_TALBOT;
    p->temporalReprojectionMipLevel = 1;
    p->spatialReuseRounds = 1;
    p->spatialVisibilityMipLevel = 1;
    p->spatialLightingMipLevel = 1;
    p->spatialVisibilityUseLinearSampler = 1;
    p->spatialLightingUseLinearSampler = 1;
    p->spatialVisibilityTStepScale = 1.f;
    p->spatialLightingTStepScale = 1.f;
    p->spatialVisibilityTrackingMethod = VRESTIR_RAY_MARCHING;
    p->spatialLightingTrackingMethod = VRESTIR_RAY_MARCHING;
    p->randomSamplerType = VRESTIR_SAMPLER_R2;
    p->sampleRadius = 10.f;
    p->spatialSampleCount = 4;
    p->enableVisibilitySimilarityRejection = 0;
    p->spatialMISMethod = VRESTIR_MIS_TALBOT;
    p->finalLightSamples = 1;
    p->finalVisibilitySamples = 1;
    p->finalVisibilityTrackingMethod = VRESTIR_ANALYTIC_TRACKING;
    p->finalLightTrackingMethod = VRESTIR_ANALYTIC_TRACKING;
    p->finalRandomSamplerType = VRESTIR_SAMPLER_R2;
    p->finalTStepScale = 0.2f;
}

int orc_set_params(void* h, const vrestir_params* p) {
    Oracle* O = (Oracle*)h;
    if (p->vertexReuse) return VRESTIR_EUNSUPPORTED;
    if (p->maxBounces < 1 || p->maxBounces > 8) return VRESTIR_EINVAL;
    bool bChanged = p->maxBounces != O->P.maxBounces;
    O->P = *p;
    O->optionsChanged = true;
    if (bChanged) { O->W = 0; O->H = 0; }
    return VRESTIR_OK;
}

// One frame of a volume: mips, conservative mips, temperature and velocity.
static int buildFrame(const vrestir_volume_desc* d, Oracle::VolFrame& f) {
    if (d->numMips < 1 || d->numMips > 8) return VRESTIR_EINVAL;
    int align = 8 << (d->numMips - 1);
    if (d->nx % align || d->ny % align || d->nz % align) return VRESTIR_EINVAL;
    f = Oracle::VolFrame();
    DenseGrid base;
    base.nx = d->nx; base.ny = d->ny; base.nz = d->nz;
    base.v.assign(d->density, d->density + (size_t)d->nx * d->ny * d->nz);
    for (float& v : base.v) {
        if (!std::isfinite(v)) return VRESTIR_EINVAL;
        v = v > 0.f ? v : 0.f;
    }
    bool fp16 = d->brickFormat == VRESTIR_BRICK_FP16;
    DenseGrid cons = dilateConservative(base);
    DenseGrid curN = base, curC = cons;
    for (int k = 0; k < d->numMips; k++) {
        if (k > 0) { curN = downsample2(curN); curC = downsample2(curC); }
        buildSlot(f.grids[k], curN, d->indexToWorld, 1 << k, fp16);
        buildSlot(f.grids[8 + k], curC, d->indexToWorld, 1 << k, fp16);
    }
    f.numMips = d->numMips;
    const float* X = f.grids[0].xform;
    auto rl = [&](int r) { return std::sqrt(X[r * 4] * X[r * 4] + X[r * 4 + 1] * X[r * 4 + 1] + X[r * 4 + 2] * X[r * 4 + 2]); };
    f.tStep = (rl(0) + rl(1) + rl(2)) / 3.f;
    auto rd = [&](int r) { return X[r * 4] * X[r * 4] + X[r * 4 + 1] * X[r * 4 + 1] + X[r * 4 + 2] * X[r * 4 + 2]; };
    f.superVoxelDiag = 8.f * std::sqrt(rd(0) + rd(1) + rd(2));
    if (d->temperature) {
        DenseGrid g;
        g.nx = d->nx; g.ny = d->ny; g.nz = d->nz;
        g.v.assign(d->temperature, d->temperature + (size_t)d->nx * d->ny * d->nz);
        for (float& v : g.v) v = v > 0.f ? v : 0.f;
        buildSlot(f.grids[kTemperatureSlot], g, d->indexToWorld, 1, fp16);
    }
    if (d->velocity) {
        VelocityField& V = f.velocity;
        V.valid = true;
        V.nx = d->nx; V.ny = d->ny; V.nz = d->nz;
        V.v.assign(d->velocity, d->velocity + 3 * (size_t)d->nx * d->ny * d->nz);
        // world -> index of the grid: the inverse of indexToWorld (mip 0)
        std::memcpy(V.w2i, f.grids[0].invxform, sizeof(V.w2i));
        const int bnx = (V.nx + 7) / 8, bny = (V.ny + 7) / 8, bnz = (V.nz + 7) / 8;
        V.brickOcc.assign((size_t)bnx * bny * bnz, 0);
        for (int z = 0; z < V.nz; z++)
            for (int y = 0; y < V.ny; y++)
                for (int x = 0; x < V.nx; x++) {
                    const float* q = &V.v[3 * ((size_t)x + (size_t)V.nx * ((size_t)y + (size_t)V.ny * z))];
                    if (q[0] != 0.f || q[1] != 0.f || q[2] != 0.f)
                        V.brickOcc[(size_t)(x / 8) + (size_t)bnx * ((size_t)(y / 8) + (size_t)bny * (z / 8))] = 1;
                }
    }
    return VRESTIR_OK;
}

// Binds frame `cur` and, in the previous-frame slots, the frame before it in
// sequence order (the first frame's previous is the last, SC/Scene.cpp:3342-3350).
static void bindFrame(Oracle& O, int cur, bool animating) {
    const int n = (int)O.frames.size();
    const Oracle::VolFrame& c = O.frames[cur];
    const Oracle::VolFrame& pv = O.frames[(cur + n - 1) % n];
    Medium& m = O.S.med;
    for (auto& g : m.grid) g = GridSlot();
    for (int s = 0; s < 18; s++) m.grid[s] = c.grids[s];
    for (int k = 0; k < 8; k++) m.grid[kPrevDensityOffset + k] = pv.grids[k];
    m.grid[kTemperatureSlot + kPrevExtraOffset] = pv.grids[kTemperatureSlot];
    m.velocity = c.velocity;
    m.numMips = c.numMips;
    m.tStep = c.tStep;
    m.superVoxelDiag = c.superVoxelDiag;
    m.hasEmission = c.grids[kTemperatureSlot].valid;
    m.prevHasEmission = pv.grids[kTemperatureSlot].valid;
    O.curFrame = cur;
    O.hasAnimation = animating;
}

int orc_set_volume(void* h, const vrestir_volume_desc* d) {
    Oracle* O = (Oracle*)h;
    Oracle::VolFrame f;
    int rc = buildFrame(d, f);
    if (rc) return rc;
    O->frames.clear();
    O->frames.push_back(std::move(f));
    Medium& m = O->S.med;
    m.sigma_s = {d->sigmaS[0], d->sigmaS[1], d->sigmaS[2]};
    m.sigma_a = {d->sigmaA[0], d->sigmaA[1], d->sigmaA[2]};
    m.sigma_t = d->sigmaS[0] + d->sigmaA[0];
    m.g = d->g;
    m.densityScale = d->densityScale;
    m.worldScaling = 1.f;
    m.LeScale = d->LeScale;
    m.tempCutoff = d->temperatureCutOff;
    m.tempScale = d->temperatureScale;
    m.velocityScale = d->velocityScale;
    bindFrame(*O, 0, false);
    O->volumeSet = true;
    O->optionsChanged = true;
    return VRESTIR_OK;
}

// addGVDBVolumeSequence: one more frame (material parameters stay the first frame's).
int orc_add_volume_frame(void* h, const vrestir_volume_desc* d) {
    Oracle* O = (Oracle*)h;
    if (!O->volumeSet || d->numMips != O->frames[0].numMips) return VRESTIR_EINVAL;
    Oracle::VolFrame f;
    int rc = buildFrame(d, f);
    if (rc) return rc;
    O->frames.push_back(std::move(f));
    bindFrame(*O, O->curFrame, O->hasAnimation);
    return (int)O->frames.size() - 1;
}

int orc_set_volume_frame(void* h, int index, int animating) {
    Oracle* O = (Oracle*)h;
    if (index < 0 || index >= (int)O->frames.size()) return VRESTIR_EINVAL;
    bindFrame(*O, index, animating != 0);
    return VRESTIR_OK;
}

int orc_set_freeze_frame(void* h, int freeze) {
    ((Oracle*)h)->freezeFrame = freeze != 0;
    return VRESTIR_OK;
}

// Temperature grid (slot 16) of the current frame; values <= 0 are stored as 0 and
// the dimensions padded with zeros to whole bricks.
int orc_set_temperature(void* h, const float* t, int nx, int ny, int nz, const float* indexToWorld, int fp16,
                        float LeScale, float cutoff, float scale) {
    Oracle* O = (Oracle*)h;
    if (!t || nx <= 0 || ny <= 0 || nz <= 0 || O->frames.empty()) return VRESTIR_EINVAL;
    Medium& m = O->S.med;
    DenseGrid g;
    g.nx = (nx + 7) / 8 * 8; g.ny = (ny + 7) / 8 * 8; g.nz = (nz + 7) / 8 * 8;
    g.v.assign((size_t)g.nx * g.ny * g.nz, 0.f);
    for (int z = 0; z < nz; z++)
        for (int y = 0; y < ny; y++)
            for (int x = 0; x < nx; x++) {
                float v = t[(size_t)x + (size_t)nx * ((size_t)y + (size_t)ny * z)];
                g.v[(size_t)x + (size_t)g.nx * ((size_t)y + (size_t)g.ny * z)] = v > 0.f ? v : 0.f;
            }
    buildSlot(O->frames[O->curFrame].grids[kTemperatureSlot], g, indexToWorld, 1, fp16 != 0);
    m.LeScale = LeScale;
    m.tempCutoff = cutoff;
    m.tempScale = scale;
    bindFrame(*O, O->curFrame, O->hasAnimation);
    return VRESTIR_OK;
}

// Replaces grid slot `slot` (mip k: k, conservative mip k: 8 + k) of the current
// frame with a dense grid of its own and its own index -> world transform, as a
// file set with precomputed mips binds them (SC/Scene.cpp:2794-2850).
int orc_set_slot(void* h, int slot, const float* v, int nx, int ny, int nz, const float* indexToWorld, int fp16) {
    Oracle* O = (Oracle*)h;
    if (!v || slot < 0 || slot >= 16 || nx <= 0 || ny <= 0 || nz <= 0 || nx % 8 || ny % 8 || nz % 8 ||
        O->frames.empty())
        return VRESTIR_EINVAL;
    DenseGrid g;
    g.nx = nx; g.ny = ny; g.nz = nz;
    g.v.assign(v, v + (size_t)nx * ny * nz);
    for (float& x : g.v) x = x > 0.f ? x : 0.f;
    buildSlot(O->frames[O->curFrame].grids[slot], g, indexToWorld, 1, fp16 != 0);
    bindFrame(*O, O->curFrame, O->hasAnimation);
    return VRESTIR_OK;
}

int orc_set_blackbody_lut(void* h, const float* rgb, int rows) {
    Oracle* O = (Oracle*)h;
    if (rows < 1 || rows > 256 || !rgb) return VRESTIR_EINVAL;
    O->S.med.lut.clear();
    for (int i = 0; i < rows; i++) O->S.med.lut.push_back({rgb[3 * i], rgb[3 * i + 1], rgb[3 * i + 2]});
    return VRESTIR_OK;
}

int orc_set_env_map(void* h, int w, int hgt, const float* rgb, float intensity, const float tint[3],
                    const float rotationDegrees[3], int impDim, int impSpp) {
    Oracle* O = (Oracle*)h;
    EnvMapState& E = O->S.envMap;
    if (!rgb) { E = EnvMapState(); return VRESTIR_OK; }
    if (w < 1 || hgt < 1) return VRESTIR_EINVAL;
    E = EnvMapState();
    E.textured = true;
    E.w = w; E.h = hgt;
    E.rgb.assign(rgb, rgb + (size_t)w * hgt * 3);
    E.scale = {intensity * tint[0], intensity * tint[1], intensity * tint[2]};
    buildImportance(E, impDim ? impDim : 512, impSpp ? impSpp : 64);
    O->envCur = O->envLast = envRotation(rotationDegrees);
    return VRESTIR_OK;
}
int orc_set_env_rotation(void* h, const float deg[3]) {
    ((Oracle*)h)->envCur = envRotation(deg);
    return VRESTIR_OK;
}
int orc_set_env_intensity(void* h, float intensity, const float tint[3]) {
    ((Oracle*)h)->S.envMap.scale = {intensity * tint[0], intensity * tint[1], intensity * tint[2]};
    return VRESTIR_OK;
}
// Test hooks on the current orientation (or the previous frame's, prev != 0).
void orc_env_rotation(void* h, int prev, float toWorld[9], float toLocal[9]) {
    Oracle* O = (Oracle*)h;
    const EnvXf& X = prev ? O->envLast : O->envCur;
    std::memcpy(toWorld, X.toWorld, sizeof(X.toWorld));
    std::memcpy(toLocal, X.toLocal, sizeof(X.toLocal));
}
void orc_env_eval(void* h, const float dir[3], float out[3]) {
    Oracle* O = (Oracle*)h;
    f3 L = envEval(O->S, O->envCur, {dir[0], dir[1], dir[2]});
    out[0] = L.x; out[1] = L.y; out[2] = L.z;
}
// out: dir(3), pdf, Le(3), lightID
void orc_env_sample(void* h, float u, float v, float out[8]) {
    Oracle* O = (Oracle*)h;
    float pdf;
    f3 d = envSample(O->S, O->envCur, {u, v}, pdf);
    f3 L = envEval(O->S, O->envCur, d);
    int id;
    (void)encodeLightUV(O->S, O->envCur, d, id);
    out[0] = d.x; out[1] = d.y; out[2] = d.z; out[3] = pdf;
    out[4] = L.x; out[5] = L.y; out[6] = L.z; out[7] = (float)id;
}
float orc_env_pdf(void* h, const float dir[3]) {
    Oracle* O = (Oracle*)h;
    return envEvalPdf(O->S, O->envCur, {dir[0], dir[1], dir[2]});
}
// Importance level `mip` into out ((dim >> mip)^2 floats); returns its side or -1.
int orc_env_importance(void* h, int mip, float* out) {
    const EnvMapState& E = ((Oracle*)h)->S.envMap;
    if (!E.textured || mip < 0 || mip > E.impBaseMip) return -1;
    if (out) std::memcpy(out, E.imp[mip].data(), E.imp[mip].size() * sizeof(float));
    return E.impDim >> mip;
}
void orc_latlong(const float dir[3], float uv[2]) {
    f2 r = worldToLatlong({dir[0], dir[1], dir[2]});
    uv[0] = r.x; uv[1] = r.y;
}
void orc_dir_to_oct(const float dir[3], float uv[2]) {
    f2 r = dirToOctEqualArea({dir[0], dir[1], dir[2]});
    uv[0] = r.x; uv[1] = r.y;
}
// out: emission rgb, temperature
void orc_emission(void* h, const float p[3], float out[4]) {
    const Medium& m = ((Oracle*)h)->S.med;
    f3 pW{p[0], p[1], p[2]};
    f3 E = emissionWorld(m, pW);
    out[0] = E.x; out[1] = E.y; out[2] = E.z;
    out[3] = m.hasEmission ? temperatureWorld(m, pW) : 0.f;
}
void orc_blackbody(void* h, float temperature, float out[3]) {
    f3 c = blackbody(((Oracle*)h)->S.med, temperature);
    out[0] = c.x; out[1] = c.y; out[2] = c.z;
}

int orc_set_lights(void* h, const vrestir_point_light* L, int n, const float env[3]) {
    Oracle* O = (Oracle*)h;
    O->S.lights.clear();
    for (int i = 0; i < n; i++)
        O->S.lights.push_back({{L[i].posW[0], L[i].posW[1], L[i].posW[2]},
                               {L[i].intensity[0], L[i].intensity[1], L[i].intensity[2]}});
    O->S.env = {env[0], env[1], env[2]};
    return VRESTIR_OK;
}

// The light collection (LT/LightCollection.cpp, LT/BuildTriangleList.cs.slang,
// LT/FinalizeIntegration.cs.slang:60-76) and the power sampler's alias table
// (LT/EmissivePowerSampler.cpp:62-79, UT/Sampling/AliasTable.cpp:46-125, without
// the final shuffle of the slots).
int orc_set_surfaces(void* h, const vrestir_surfaces_desc* d) {
    Oracle* O = (Oracle*)h;
    SceneState& S = O->S;
    S.emTris.clear();
    S.emWeight.clear();
    S.emAlias.clear();
    S.emActive = 0;
    S.emWeightSum = 0.f;
    S.emMult = 1.f;
    S.surfTris.clear();
    S.surfMats.clear();
    if (!d) return VRESTIR_OK;
    S.emMult = d->emissiveIntensityMultiplier;
    for (int m = 0; m < d->numMaterials; m++) {
        const vrestir_material& M = d->materials[m];
        S.surfMats.push_back({{M.baseColor[0], M.baseColor[1], M.baseColor[2]}, M.roughness, M.metallic, M.IoR,
                              {M.emissive[0], M.emissive[1], M.emissive[2]}, M.doubleSided != 0});
    }
    for (int t = 0; t < d->numTriangles; t++) {
        SurfTri T;
        for (int k = 0; k < 3; k++) {
            const int v = d->indices[3 * t + k];
            T.p[k] = {d->positions[3 * v], d->positions[3 * v + 1], d->positions[3 * v + 2]};
        }
        f3 fn = cross(T.p[1] - T.p[0], T.p[2] - T.p[0]);
        fn = dot(fn, fn) > 0.f ? normalize(fn) : f3{0.f, 0.f, 1.f};
        for (int k = 0; k < 3; k++) {
            const int v = d->indices[3 * t + k];
            T.n[k] = d->normals ? f3{d->normals[3 * v], d->normals[3 * v + 1], d->normals[3 * v + 2]} : fn;
        }
        T.material = d->materialIds[t];
        S.surfTris.push_back(T);
    }
    auto P = [&](int v) { return f3{d->positions[3 * v], d->positions[3 * v + 1], d->positions[3 * v + 2]}; };
    for (int t = 0; t < d->numTriangles; t++) {
        const vrestir_material& m = d->materials[d->materialIds[t]];
        f3 Le{m.emissive[0], m.emissive[1], m.emissive[2]};
        if (!(Le.x > 0.f || Le.y > 0.f || Le.z > 0.f)) continue;
        EmissiveTriangle T;
        for (int k = 0; k < 3; k++) T.p[k] = P(d->indices[3 * t + k]);
        // computeFaceNormalAndAreaW (SC/Scene.slang:343-358)
        f3 N = cross(T.p[1] - T.p[0], T.p[2] - T.p[0]);
        T.area = 0.5f * length(N);
        T.normal = mk3(0.f);
        if (T.area > 0.f) {
            f3 n = normalize(N);
            // encodeNormal2x16 / decodeNormal2x16 (UT/Math/PackedFormats.slang:34-46)
            auto wrap = [](float x, float y, float& ox, float& oy) {
                ox = (1.f - std::fabs(y)) * (x >= 0.f ? 1.f : -1.f);
                oy = (1.f - std::fabs(x)) * (y >= 0.f ? 1.f : -1.f);
            };
            float inv = 1.f / (std::fabs(n.x) + std::fabs(n.y) + std::fabs(n.z));
            float ox = n.x * inv, oy = n.y * inv;
            if (n.z < 0.f) wrap(n.x * inv, n.y * inv, ox, oy);
            auto toSnorm = [](float v) {
                v = std::isnan(v) ? 0.f : std::min(std::max(v, -1.f), 1.f);
                return (int)std::trunc(v * 32767.f + (v >= 0.f ? 0.5f : -0.5f));
            };
            auto fromSnorm = [](int b) { return std::max((float)(int16_t)b / 32767.f, -1.f); };
            f3 q{fromSnorm(toSnorm(ox) & 0xffff), fromSnorm(toSnorm(oy) & 0xffff), 0.f};
            q.z = 1.f - std::fabs(q.x) - std::fabs(q.y);
            if (q.z < 0.f) wrap(q.x, q.y, q.x, q.y);
            T.normal = normalize(q);
        }
        T.Le = Le;
        S.emTris.push_back(T);
        float flux = luminance(Le) * T.area * (float)M_PI;
        S.emWeight.push_back(flux);
        if (flux > 0.f) S.emActive++;
    }
    const uint32_t count = (uint32_t)S.emWeight.size();
    if (count == 0) return VRESTIR_OK;
    double sum = 0.0;
    for (float f : S.emWeight) sum += f;
    S.emWeightSum = (float)sum;
    S.emAlias.assign(count, AliasItem{1.f, 0u, 0u});
    for (uint32_t i = 0; i < count; i++) S.emAlias[i] = {1.f, i, i};
    if (!(sum > 0.0)) {
        for (uint32_t i = 0; i < count; i++) S.emAlias[i] = {0.f, 0u, 0u};
        return VRESTIR_OK;
    }
    std::vector<float> w(S.emWeight);
    const double factor = count / sum;
    for (float& f : w) f = (float)(f * factor);
    std::vector<uint32_t> order(count);
    for (uint32_t i = 0; i < count; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return w[a] < w[b]; });
    uint32_t lo = 0, hi = count - 1;
    while (lo != hi) {
        uint32_t i = order[lo], j = order[hi];
        S.emAlias[i].threshold = w[i];
        S.emAlias[i].indexA = j;
        w[j] -= 1.f - w[i];
        if (lo == hi - 1) {
            S.emAlias[j].threshold = 1.f;
            S.emAlias[j].indexA = j;
            break;
        } else if (w[j] < 1.f) {
            std::swap(order[lo], order[hi]);
            hi--;
        } else {
            lo++;
        }
    }
    return VRESTIR_OK;
}

// FindSurfaceHit and FindIfOccluded over [0, tMax] per ray (8 floats: origin, dir,
// tMax, 0): out[8] = {t, triangle, bary, occluded, 0, 0, 0}.
void orc_surface_hit(void* h, int n, const float* rays, float* out) {
    Oracle* O = (Oracle*)h;
    for (int i = 0; i < n; i++) {
        const float* r = rays + 8 * i;
        f3 o{r[0], r[1], r[2]}, d{r[3], r[4], r[5]};
        HitInfo hi = findSurfaceHit(O->S, o, d, 0.f, r[6]);
        float* q = out + 8 * i;
        q[0] = hi.t; q[1] = (float)hi.tri; q[2] = hi.bary.x; q[3] = hi.bary.y;
        q[4] = findIfOccluded(O->S, o, d, 0.f, r[6]) ? 1.f : 0.f;
        q[5] = q[6] = q[7] = 0.f;
    }
}

// Shading at each ray's closest hit (rays of 8 floats, seeds of 3 words): variant 0
// out = sampleBSDF {wi, pdf, weight, lobe or -1}; variant 1 = the evaluation at that wi
// {evalBSDFCosine, evalPdfBSDF, emissive, NdotV}; zeros without a hit.
void orc_surface_shade(void* h, int variant, int n, const float* rays, const uint32_t* seeds, float* out) {
    Oracle* O = (Oracle*)h;
    for (int i = 0; i < n; i++) {
        const float* r = rays + 8 * i;
        float* q = out + 8 * i;
        for (int k = 0; k < 8; k++) q[k] = 0.f;
        f3 o{r[0], r[1], r[2]}, d{r[3], r[4], r[5]};
        HitInfo hi = findSurfaceHit(O->S, o, d, 0.f, r[6]);
        if (hi.tri < 0) continue;
        ShadingData sd = computeSurfaceShadingInfo(O->S, hi.tri, hi.bary, -d);
        SampleGen sg = makeSampleGen(seeds[3 * i], seeds[3 * i + 1], seeds[3 * i + 2]);
        BSDFSample bs;
        bool ok = sampleBSDF(sd, sg, bs);
        if (variant == 0) {
            q[0] = bs.wi.x; q[1] = bs.wi.y; q[2] = bs.wi.z; q[3] = bs.pdf;
            q[4] = bs.weight.x; q[5] = bs.weight.y; q[6] = bs.weight.z; q[7] = ok ? (float)bs.lobe : -1.f;
        } else {
            f3 f = evalBSDFCosine(sd, bs.wi);
            q[0] = f.x; q[1] = f.y; q[2] = f.z; q[3] = evalPdfBSDF(sd, bs.wi);
            q[4] = sd.emissive.x; q[5] = sd.emissive.y; q[6] = sd.emissive.z; q[7] = sd.NdotV;
        }
    }
}

// sampleSceneLights at points (light kinds in `kinds` bits: 1 env, 2 point, 4 emissive):
// out[8] = {Li rgb, pdfArea, light index, uv, ray distance}; index -9 when none.
void orc_sample_lights(void* h, int kinds, int n, const float* points, const uint32_t* seeds, float* out) {
    Oracle* O = (Oracle*)h;
    for (int i = 0; i < n; i++) {
        SampleGen sg = makeSampleGen(seeds[3 * i], seeds[3 * i + 1], seeds[3 * i + 2]);
        SceneLightSample ls{};
        int idx = -9;
        f2 uv{0.f, 0.f};
        float* o = out + 8 * i;
        for (int k = 0; k < 8; k++) o[k] = 0.f;
        f3 p{points[3 * i], points[3 * i + 1], points[3 * i + 2]};
        if (sampleSceneLights(O->S, O->frameEnv[0], p, kinds & 1, (kinds & 2) != 0, (kinds & 4) != 0, sg, ls, idx,
                              uv)) {
            o[0] = ls.Li.x; o[1] = ls.Li.y; o[2] = ls.Li.z; o[3] = ls.pdfArea;
            o[4] = (float)idx; o[5] = uv.x; o[6] = uv.y; o[7] = ls.rayDistance;
        } else {
            o[4] = -9.f;
        }
    }
}

// The emissive-triangle branch of evaluate_L_in_volume at points: out[8] =
// {Le cos / dist^2 rgb, dist, dir xyz, valid}.
void orc_emissive_eval(void* h, int n, const float* points, const int* tris, const float* uvs, float* out) {
    Oracle* O = (Oracle*)h;
    for (int i = 0; i < n; i++) {
        float* o = out + 8 * i;
        for (int k = 0; k < 8; k++) o[k] = 0.f;
        TriangleLightSample ls;
        f3 p{points[3 * i], points[3 * i + 1], points[3 * i + 2]};
        if (!powerSamplerGetSample(O->S, p, tris[i], {uvs[2 * i], uvs[2 * i + 1]}, ls)) continue;
        f3 L = ls.Le * ls.cosTheta / (ls.distance * ls.distance);
        o[0] = L.x; o[1] = L.y; o[2] = L.z; o[3] = ls.distance;
        o[4] = ls.dir.x; o[5] = ls.dir.y; o[6] = ls.dir.z; o[7] = 1.f;
    }
}

int orc_set_camera(void* h, const vrestir_camera* c, int W, int H) {
    Oracle* O = (Oracle*)h;
    buildCamera(O->S.cam, c->posW, c->target, c->up, c->fovY, (float)W / (float)H, c->nearZ, c->farZ);
    if (!O->camSet) O->S.prevCam = O->S.cam;
    O->camSet = true;
    return VRESTIR_OK;
}

int orc_execute(void* h, int W, int H, float* outRGBA, int rowBegin, int rowEnd, int nthreads) {
    Oracle* O = (Oracle*)h;
    if (!O->volumeSet || !O->camSet) return VRESTIR_ESTATE;
    O->rowBegin = rowBegin; O->rowEnd = rowEnd; O->nthreads = nthreads;
    executeFrame(*O, W, H);
    if (outRGBA) std::memcpy(outRGBA, O->output.data(), sizeof(f4) * O->output.size());
    return VRESTIR_OK;
}

int orc_prepare_frame(void* h, int W, int H, int nthreads) {
    Oracle* O = (Oracle*)h;
    if (!O->volumeSet || !O->camSet) return VRESTIR_ESTATE;
    O->rowBegin = 0; O->rowEnd = H; O->nthreads = nthreads;
    prepareFrame(*O, W, H);
    return VRESTIR_OK;
}
int orc_set_rows(void* h, int rowBegin, int rowEnd) {
    Oracle* O = (Oracle*)h;
    if (rowBegin < 0 || rowEnd > O->H || rowBegin >= rowEnd) return VRESTIR_EINVAL;
    O->rowBegin = rowBegin; O->rowEnd = rowEnd;
    return VRESTIR_OK;
}
int orc_run_stage(void* h, int stage, int round) {
    if (((Oracle*)h)->P.useSurfaceScene && stage != VRESTIR_STAGE_GENERATE_FEATURES) return VRESTIR_EUNSUPPORTED;
    runStage(*(Oracle*)h, stage, round);
    return VRESTIR_OK;
}
int orc_end_frame(void* h) { endFrame(*(Oracle*)h); return VRESTIR_OK; }

static void* bufPtr(Oracle* O, int id, size_t& bytes) {
    switch (id) {
        case VRESTIR_BUF_LEARNED_CDF: bytes = O->learnedCdf.size() * 2; return O->learnedCdf.data();
        case VRESTIR_BUF_TRAIN_RECORDS: bytes = O->trainRecs.size() * 4; return O->trainRecs.data();
        case VRESTIR_BUF_FEATURES: bytes = O->features.size() * 8; return O->features.data();
        case VRESTIR_BUF_VBUFFER: bytes = O->vbuffer.size() * 16; return O->vbuffer.data();
        case VRESTIR_BUF_TEMPORAL_FEATURES: bytes = O->temporalFeatures.size() * 8; return O->temporalFeatures.data();
        case VRESTIR_BUF_RESERVOIRS0: bytes = O->res[0].size() * 36; return O->res[0].data();
        case VRESTIR_BUF_RESERVOIRS1: bytes = O->res[1].size() * 36; return O->res[1].data();
        case VRESTIR_BUF_TEMPORAL_RESERVOIRS: bytes = O->temporalRes.size() * 36; return O->temporalRes.data();
        case VRESTIR_BUF_EXTRA0: bytes = O->extra[0].size() * 12; return O->extra[0].data();
        case VRESTIR_BUF_EXTRA1: bytes = O->extra[1].size() * 12; return O->extra[1].data();
        case VRESTIR_BUF_TEMPORAL_EXTRA: bytes = O->temporalExtra.size() * 12; return O->temporalExtra.data();
        case VRESTIR_BUF_COLOR: bytes = O->output.size() * 16; return O->output.data();
        case ORC_BUF_MVEC: bytes = O->mvec.size() * 8; return O->mvec.data();
    }
    bytes = 0;
    return nullptr;
}
int orc_read_buffer(void* h, int id, void* dst, size_t bytes) {
    size_t n; void* p = bufPtr((Oracle*)h, id, n);
    if (!p || bytes != n) return VRESTIR_EINVAL;
    std::memcpy(dst, p, n);
    return VRESTIR_OK;
}
int orc_write_buffer(void* h, int id, const void* src, size_t bytes) {
    size_t n; void* p = bufPtr((Oracle*)h, id, n);
    if (!p || bytes != n) return VRESTIR_EINVAL;
    std::memcpy(p, src, n);
    return VRESTIR_OK;
}
int orc_frame_count(void* h) { return ((Oracle*)h)->frameCount; }
int orc_total_round_id(void* h) { return ((Oracle*)h)->totalRoundId; }
void orc_stage_seconds(void* h, double out[6]) { std::memcpy(out, ((Oracle*)h)->stageSeconds, sizeof(double) * 6); }

// Per-pixel traversal work of the final-shading stage on the current buffers
// (diagnostics: out[2*i] = DDA steps, out[2*i+1] = cell steps).
int orc_final_costs(void* h, long long* out) {
    Oracle* O = (Oracle*)h;
    int idx = O->totalRoundId % 2;
    bool noReuse = !O->P.enableSpatialReuse && !O->P.enableTemporalReuse;
    for (int y = 0; y < O->H; y++)
        for (int x = 0; x < O->W; x++) {
            int pid = y * O->W + x;
            g_ddaSteps = 0; g_cellSteps = 0;
            SampleGen sg = makeSampleGen(x, y, (uint32_t)(O->numTotalRounds * O->frameCount + O->numTotalRounds - 1));
            Reservoir r = O->res[idx][pid];
            if (r.runningSum > 0.f) {
                f3 d = rayDir(O->S.cam, x, y, O->W, O->H);
                Ray ray{O->S.cam.pos, d, 0, r.depth};
                (void)evaluateF(*O, O->frameEnv[0], r, O->extra[idx].data(), ray, sg, O->finalOpt, noReuse, true);
            }
            out[2 * pid] = g_ddaSteps;
            out[2 * pid + 1] = g_cellSteps;
        }
    return VRESTIR_OK;
}

// Visit sequence of the hierarchical walk (VR/VolumeUtils.slang:171-282): every
// leaf handed to the adapter, FNV-1a-hashed over its index-space minimum.
struct VisitHashAdapter {
    uint32_t count = 0, hash = 2166136261u;
    float tSum = 0.f, tFirst = 0.f, tLast = 0.f, tNear = 0.f, tFar = 0.f;
    void mix(uint32_t v) {
        for (int k = 0; k < 4; k++) { hash ^= (v >> (8 * k)) & 0xFFu; hash *= 16777619u; }
    }
    void setRayInfo(float a, float b, const Ray&) { tNear = a; tFar = b; }
    void start() {}
    bool mainStep(const HDDA&, f3 vminLeaf, f3, f4, int, float& t, SampleGen&) {
        mix((uint32_t)(int)vminLeaf.x); mix((uint32_t)(int)vminLeaf.y); mix((uint32_t)(int)vminLeaf.z);
        if (count == 0) tFirst = t;
        tLast = t;
        tSum += t;
        count++;
        return false;
    }
    void end() {}
};

// getNodeAtPoint + getValueAtPoint (GVS/gvdbNodes.slang:261-279, GVS/gvdb.slang:6-29)
// at a world point: point and trilinear density, the leaf's minimum, the voxel.
static void indexAtPoint(const Medium& m, f3 pW, int slot, float* o) {
    const GridSlot& g = m.grid[slot];
    f3 p = worldToMediumP(m, pW, slot);
    o[1] = densityIndex(m, p, slot);
    o[2] = o[3] = o[4] = -1.f;
    o[5] = std::floor(p.x); o[6] = std::floor(p.y); o[7] = std::floor(p.z);
    if (p.x < g.bmin.x || p.y < g.bmin.y || p.z < g.bmin.z || p.x >= g.bmax.x || p.y >= g.bmax.y || p.z >= g.bmax.z)
        return;
    const LeafNode* leaf = leafAtPoint(g, p);
    if (!leaf) return;
    o[0] = valueAtPoint(g, p, false) * m.densityScale;
    o[2] = (float)leaf->vmin.x; o[3] = (float)leaf->vmin.y; o[4] = (float)leaf->vmin.z;
}

// Batched tracking primitive: same contract as vrestir_track_batch, on host memory.
int orc_track_batch(void* h, int mode, int mip, int linear, float tStepScale, int numSamples, int n, const float* rays,
                    const uint32_t* seeds, float* out) {
    Oracle* O = (Oracle*)h;
    const Medium& m = O->S.med;
    for (int i = 0; i < n; i++) {
        const float* r = rays + 8 * i;
        Ray ray{{r[0], r[1], r[2]}, {r[3], r[4], r[5]}, 0.f, r[6]};
        SampleGen sg = makeSampleGen(seeds[3 * i], seeds[3 * i + 1], seeds[3 * i + 2]);
        float* o = out + 8 * i;
        for (int k = 0; k < 8; k++) o[k] = 0.f;
        switch (mode) {
            case VRESTIR_TRACK_TR_ANALYTIC: o[0] = trAnalytic(m, ray, mip, sg, linear != 0); break;
            case VRESTIR_TRACK_TR_RAY_MARCHING: o[0] = trRayMarch(m, ray, mip, linear != 0, tStepScale, sg); break;
            case VRESTIR_TRACK_TR_RATIO: o[0] = trResidualRatio(m, ray, mip, sg, false, true); break;
            case VRESTIR_TRACK_TR_RESIDUAL_RATIO: o[0] = trResidualRatio(m, ray, mip, sg, false, false); break;
            case VRESTIR_TRACK_TR_ANALOG_RESIDUAL_RATIO: o[0] = trResidualRatio(m, ray, mip, sg, true, false); break;
            case VRESTIR_TRACK_SAMPLE_ANALYTIC: {
                float hh[4], pd[4], tr[4];
                int ns = std::max(1, std::min(4, numSamples));
                sampleAnalytic(m, ray, sg, linear != 0, hh, mip, pd, tr, ns);
                // packed: hit[0..ns-1] then tr / pdf are reported for the first two samples
                o[0] = hh[0]; o[1] = pd[0]; o[2] = tr[0];
                o[3] = ns > 1 ? hh[1] : 0.f; o[4] = ns > 1 ? pd[1] : 0.f; o[5] = ns > 1 ? tr[1] : 0.f;
                o[6] = ns > 2 ? hh[2] : 0.f; o[7] = ns > 3 ? hh[3] : 0.f;
                break;
            }
            case VRESTIR_TRACK_SAMPLE_DECOMPOSITION: {
                MediumInteraction mi{};
                sampleDecomposition(m, ray, sg, mi, mip);
                o[0] = mi.valid ? 1.f : 0.f;
                o[1] = mi.p.x; o[2] = mi.p.y; o[3] = mi.p.z;
                break;
            }
            case VRESTIR_TRACK_SAMPLE_CELL_BY_DENSITY: {
                float bound = 0.f;
                f2 s = cellByDensity(m, ray, sg, mip, bound);
                o[0] = s.x; o[1] = s.y; o[2] = bound;
                o[3] = rejectionSampleDepth(m, ray, sg, mip);
                break;
            }
            case VRESTIR_TRACK_FEATURE_MARCH: o[0] = featureMarch(m, ray, mip, linear != 0, tStepScale, sg); break;
            case VRESTIR_TRACK_DENSITY_AT_POINT: o[0] = densityWorld(m, ray.origin, mip); break;
            case VRESTIR_TRACK_RNG_STREAM:
                for (int k = 0; k < 8; k++) { uint32_t u = xoshiroNext(sg); std::memcpy(&o[k], &u, 4); }
                break;
            case VRESTIR_TRACK_ENV_EVAL: orc_env_eval(h, r + 3, o); break;
            case VRESTIR_TRACK_ENV_SAMPLE: orc_env_sample(h, r[0], r[1], o); break;
            case VRESTIR_TRACK_EMISSION: orc_emission(h, r, o); break;
            case VRESTIR_TRACK_VELOCITY: {
                f3 v = velocityWorld(m, ray.origin);
                o[0] = v.x; o[1] = v.y; o[2] = v.z;
                break;
            }
            case VRESTIR_TRACK_DENSITY_INDEX: indexAtPoint(m, ray.origin, mip, o); break;
            case VRESTIR_TRACK_VISITED_BRICKS: {
                VisitHashAdapter vh;
                trackGVDB(m, ray, mip, sg, vh, linear != 0);
                uint32_t hb = vh.hash;
                o[0] = (float)vh.count; std::memcpy(&o[1], &hb, 4); o[2] = vh.tSum;
                o[3] = vh.tFirst; o[4] = vh.tLast; o[5] = vh.tNear; o[6] = vh.tFar;
                break;
            }
            default: return VRESTIR_EINVAL;
        }
    }
    return VRESTIR_OK;
}

// RNG stream of the pass's sample generator for one pixel / sample number.
void orc_rng_stream(uint32_t px, uint32_t py, uint32_t sampleNumber, int n, uint32_t* out) {
    SampleGen g = makeSampleGen(px, py, sampleNumber);
    for (int i = 0; i < n; i++) out[i] = xoshiroNext(g);
}
uint32_t orc_morton(uint32_t x, uint32_t y) { return interleave32(x, y); }
void orc_splitmix64(uint64_t seed, int n, uint64_t* out) {
    for (int i = 0; i < n; i++) out[i] = splitmix64Next(seed);
}
void orc_xoshiro128ss(const uint32_t seed[4], int n, uint32_t* out) {
    SampleGen g;
    for (int i = 0; i < 4; i++) g.s[i] = seed[i];
    for (int i = 0; i < n; i++) out[i] = xoshiroNext(g);
}
float orc_round_to_half(float v) { return roundToHalf(v); }

// Henyey-Greenstein sampling (MediumInteraction::Sample_p): wi and the returned pdf.
float orc_sample_phase(float g, const float wo[3], float u0, float u1, float wi[3]) {
    f3 w{0, 0, 0};
    float pdf = samplePhase(g, {wo[0], wo[1], wo[2]}, w, {u0, u1});
    wi[0] = w.x; wi[1] = w.y; wi[2] = w.z;
    return pdf;
}
float orc_phase_hg(float cosTheta, float g) { return phaseHG(cosTheta, g); }

// Neighbour offsets of one spatial round.
void orc_neighbor_offsets(int sampleCount, float radius, int frameSeed, int sampler, int* out) {
    for (int s = 0; s < sampleCount; s++) neighborOffset(s, sampleCount, radius, frameSeed, sampleCount, sampler, out[2 * s], out[2 * s + 1]);
}

// ---- helpers exposed so the tests can check the oracle against closed forms
void orc_oct_to_dir(float u, float v, float out[3]) {
    f3 d = octToDirEqualArea({u, v});
    out[0] = d.x; out[1] = d.y; out[2] = d.z;
}
void orc_light_uv_encode(const float dir[3], float uv[2], int* id) {
    f2 e = encodeLightUV({dir[0], dir[1], dir[2]}, *id);
    uv[0] = e.x; uv[1] = e.y;
}
void orc_light_uv_decode(const float uv[2], int id, float out[3]) {
    f3 d = decodeLightUV({uv[0], uv[1]}, id);
    out[0] = d.x; out[1] = d.y; out[2] = d.z;
}
void orc_wi_dist_encode(const float wiDist[4], float out[3]) {
    f3 e = encodeWiDist({wiDist[0], wiDist[1], wiDist[2], wiDist[3]});
    out[0] = e.x; out[1] = e.y; out[2] = e.z;
}
void orc_wi_dist_decode(const float in[3], float out[4]) {
    f4 d = decodeWiDist({in[0], in[1], in[2]});
    out[0] = d.x; out[1] = d.y; out[2] = d.z; out[3] = d.w;
}
int orc_pack_sampled_pixel(int pixel, int tag, int maxIndirect) {
    return encodePathTag(encodeMaxIndirect(pixel, maxIndirect), tag);
}
void orc_unpack_sampled_pixel(int s, int* tag, int* maxIndirect) {
    *tag = decodePathTag(s);
    *maxIndirect = decodeMaxIndirect(s);
}
// Streams n candidates (weights w, counts M; candidate i carries depth = i)
// through one reservoir with the sample generator of (px, py, sn); cap < 0
// uses the plain step, otherwise the M-capped one. Returns the selected index
// (-1 if none) and the reservoir's running sum and M.
int orc_wrs(int n, const float* w, const float* M, float cap, uint32_t px, uint32_t py, uint32_t sn, float* runningSum,
            float* outM) {
    SampleGen sg = makeSampleGen(px, py, sn);
    Reservoir st = newReservoir();
    st.depth = -1.f;
    for (int i = 0; i < n; i++) {
        Reservoir r = newReservoir();
        r.runningSum = w[i];
        r.M = M[i];
        r.depth = (float)i;
        if (cap < 0.f) resampleStep(r, st, sg);
        else resampleStepMaxM(r, cap, st, sg);
    }
    *runningSum = st.runningSum;
    *outM = st.M;
    return (int)st.depth;
}
// Normalised primary-ray direction of pixel (px, py) for the current camera.
int orc_camera_ray(void* h, int px, int py, int W, int H, float out[3]) {
    Oracle* O = (Oracle*)h;
    if (!O->camSet) return VRESTIR_ESTATE;
    f3 d = rayDir(O->S.cam, px, py, W, H);
    out[0] = d.x; out[1] = d.y; out[2] = d.z;
    return VRESTIR_OK;
}
// Continuous pixel coordinates of world point p through the current camera's
// view and projection, the mapping TemporalReuse applies with the previous camera.
int orc_reproject(void* h, const float p[3], int W, int H, float out[2]) {
    Oracle* O = (Oracle*)h;
    if (!O->camSet) return VRESTIR_ESTATE;
    f2 s = projectToPixel(O->S.cam, {p[0], p[1], p[2]}, W, H);
    out[0] = s.x; out[1] = s.y;
    return VRESTIR_OK;
}
// Occupied bricks the hierarchical DDA hands to a tracker along a world-space
// ray (origin, dir, tMax in ray[0..6]) on grid `slot`: per brick, the entry t,
// exit t and the brick's index-space minimum corner. Returns the count.
struct BrickLog {
    float* out; int maxOut; int n = 0;
    void setRayInfo(float, float, const Ray&) {}
    void start() {}
    bool mainStep(const HDDA& dda, f3 vminLeaf, f3, f4, int, float& t, SampleGen&) {
        if (n < maxOut) {
            float* o = out + 5 * n;
            o[0] = t; o[1] = dda.t.y; o[2] = vminLeaf.x; o[3] = vminLeaf.y; o[4] = vminLeaf.z;
        }
        n++;
        return false;
    }
    void end() {}
};
int orc_trace_bricks(void* h, int slot, const float ray[8], int vertexCenter, int maxOut, float* out) {
    Oracle* O = (Oracle*)h;
    if (!O->volumeSet || slot < 0 || slot >= 16 || !O->S.med.grid[slot].valid) return -1;
    Ray r{{ray[0], ray[1], ray[2]}, {ray[3], ray[4], ray[5]}, 0.f, ray[6]};
    SampleGen sg = makeSampleGen(0, 0, 0);
    BrickLog log{out, maxOut};
    trackGVDB(O->S.med, r, slot, sg, log, vertexCenter != 0);
    return log.n;
}

// Volume introspection for the grid tests.
int orc_grid_info(void* h, int slot, int* dims, float* maxValue, int* numBricks) {
    Oracle* O = (Oracle*)h;
    const GridSlot& g = O->S.med.grid[slot];
    if (!g.valid) return VRESTIR_EINVAL;
    dims[0] = g.nx; dims[1] = g.ny; dims[2] = g.nz;
    *maxValue = g.maxValue;
    *numBricks = (int)g.level0.size();
    return VRESTIR_OK;
}
// Brick bounds of the brick containing mip-local voxel (x,y,z), or -1 if empty.
int orc_brick_bounds(void* h, int slot, int x, int y, int z, float* out4) {
    Oracle* O = (Oracle*)h;
    const GridSlot& g = O->S.med.grid[slot];
    const LeafNode* leaf = leafAtPoint(g, {x + 0.5f, y + 0.5f, z + 0.5f});
    if (!leaf) return -1;
    out4[0] = leaf->bounds.x; out4[1] = leaf->bounds.y; out4[2] = leaf->bounds.z; out4[3] = leaf->bounds.w;
    return 0;
}
// Stored voxel value (point lookup through the tree).
float orc_voxel(void* h, int slot, int x, int y, int z) {
    Oracle* O = (Oracle*)h;
    return valueAtPoint(O->S.med.grid[slot], {x + 0.5f, y + 0.5f, z + 0.5f}, false);
}

}  // extern "C"

// ---------------------------------------------------------------- learned proposal
int orc_set_learned(void* h, int enable, int train, int bx, int by, int steps, float lr, float alpha, int judge) {
    Oracle& O = *(Oracle*)h;
    if (bx < 1 || by < 1 || steps < 1 || !(alpha >= 1.f / 64.f && alpha <= 1.f)) return VRESTIR_EINVAL;
    O.learned = enable != 0;
    O.learnedTrain = train != 0;
    O.trainBX = bx; O.trainBY = by;
    O.learnedSteps = steps;
    O.learnedLr = lr;
    O.learnedAlpha = alpha;
    O.learnedJudge = judge != 0;
    O.learnedInUse = true;
    O.adamStep = 0;
    std::fill(O.mlpM.begin(), O.mlpM.end(), 0.f);
    std::fill(O.mlpV.begin(), O.mlpV.end(), 0.f);
    O.trainRecs.clear();
    return VRESTIR_OK;
}
int orc_learned_set_weights(void* h, const float* w, int n) {
    Oracle& O = *(Oracle*)h;
    if (n != mlp::kParams) return VRESTIR_EINVAL;
    std::memcpy(O.mlpParams.data(), w, sizeof(float) * n);
    std::fill(O.mlpM.begin(), O.mlpM.end(), 0.f);
    std::fill(O.mlpV.begin(), O.mlpV.end(), 0.f);
    O.adamStep = 0;
    return VRESTIR_OK;
}
int orc_learned_get_weights(void* h, float* w, int n) {
    Oracle& O = *(Oracle*)h;
    if (n != mlp::kParams) return VRESTIR_EINVAL;
    std::memcpy(w, O.mlpParams.data(), sizeof(float) * n);
    return VRESTIR_OK;
}
void orc_learned_stats(void* h, float out[5]) { std::memcpy(out, ((Oracle*)h)->learnedStats, sizeof(float) * 5); }
int orc_learned_raw(void* h, int px, int py, int W, int H, float* raw) {
    Oracle& O = *(Oracle*)h;
    Ray ray{O.S.cam.pos, rayDir(O.S.cam, px, py, W, H), 0, kRayTMax};
    learnedRaw(O.S.med, ray, raw);
    return VRESTIR_OK;
}
void orc_mlp_encode(const float* raw, float* out) { mlp::encode(raw, out); }
// emulate != 0: bf16 operands as the HIP kernels; 0: exact (double) network.
void orc_mlp_forward(const float* prm, const float* raw, int n, float alpha, int emulate, float* z, uint16_t* cdf) {
    for (int i = 0; i < n; i++) {
        float p[mlp::kOut];
        if (emulate) {
            mlp::Acts<float> a;
            mlp::forward<float>(prm, raw + (size_t)i * mlp::kRaw, alpha, true, a);
            for (int j = 0; j < mlp::kOut; j++) { p[j] = a.p[j]; if (z) z[(size_t)i * mlp::kOut + j] = a.z[j]; }
        } else {
            mlp::Acts<double> a;
            mlp::forward<double>(prm, raw + (size_t)i * mlp::kRaw, alpha, false, a);
            for (int j = 0; j < mlp::kOut; j++) {
                p[j] = (float)a.p[j];
                if (z) z[(size_t)i * mlp::kOut + j] = (float)a.z[j];
            }
        }
        if (cdf) mlp::quantize(p, cdf + (size_t)i * mlp::kOut);
    }
}
double orc_mlp_loss(const float* prm, const float* recs, int n, float alpha, int emulate) {
    double l = 0.0;
    for (int i = 0; i < n; i++)
        l += emulate ? (double)mlp::recordLoss<float>(prm, recs + (size_t)i * mlp::kRecFloats, alpha, true)
                     : mlp::recordLoss<double>(prm, recs + (size_t)i * mlp::kRecFloats, alpha, false);
    return l;
}
void orc_mlp_backward(const float* prm, const float* recs, int n, float alpha, float scale, int emulate, double* grad,
                      double* stats) {
    for (int i = 0; i < mlp::kParams; i++) grad[i] = 0.0;
    stats[0] = stats[1] = stats[2] = 0.0;
    for (int i = 0; i < n; i++) {
        if (emulate) mlp::backward<float>(prm, recs + (size_t)i * mlp::kRecFloats, alpha, scale, true, grad, stats);
        else mlp::backward<double>(prm, recs + (size_t)i * mlp::kRecFloats, alpha, scale, false, grad, stats);
    }
}
void orc_mlp_adam(float* prm, float* m, float* v, const float* g, int n, int t, float lr) {
    mlp::adam(prm, m, v, g, n, t, lr, mlp::kAdamB1, mlp::kAdamB2, mlp::kAdamEps);
}
float orc_mlp_sample(const uint16_t* cdf, float u, int* bin, float* g) { return mlp::sampleC(cdf, u, *bin, *g); }
float orc_mlp_mass_above(const uint16_t* cdf, float x) { return mlp::massAbove(cdf, x); }
int orc_mlp_train_slot(int x, int y, int W, int H, int bx, int by, int frame) {
    return mlp::trainSlot(x, y, W, H, bx, by, frame);
}
