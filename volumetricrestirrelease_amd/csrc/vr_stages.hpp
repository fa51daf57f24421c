// vr_stages.hpp -- kernel templates of the VolumetricReSTIR pass for gfx950:
// the per-frame stages that trace through the volume, and the reference path
// tracer. Each stage's launcher lives in its own translation unit
// (vr_k_*.hip), which instantiates only that stage's kernels.
//
// One lane per pixel, 256-lane workgroups covering 16x16 pixel tiles (4 wave64
// of 16x4 pixels each, so neighbouring lanes trace neighbouring rays).
// Reservoirs are 32 B per pixel (two float4 loads, coalesced across the wave),
// extra-bounce records 12 B per bounce, features 8 B, colour 16 B.
#pragma once
#include <type_traits>
#include "vr_launch.h"
#include "vr_learned.hpp"
#include "vr_pass.hpp"

namespace vr {

// The tile a workgroup works on: its own block in a plain launch, or a work item
// a batched launch took from its queue. The stage bodies below take it as an
// argument; each kernel is a thin wrapper that passes blockIdx.
struct VBlock { int x, y; };
VR_DEV VBlock launchBlock() { return {(int)blockIdx.x, (int)blockIdx.y}; }

struct PixelIdx { int x, y, id; bool ok; };
VR_DEV PixelIdx pixelOf(const Frame& F, VBlock b) {
    PixelIdx p;
    p.x = b.x * 16 + (threadIdx.x & 15);
    p.y = b.y * 16 + (threadIdx.x >> 4);
    p.ok = p.x < F.width && p.y < F.height;
    p.id = p.y * F.width + p.x;
    return p;
}
VR_DEV bool inRange(int x, int y, int W, int H) { return x >= 0 && x < W && y >= 0 && y < H; }

// The tracing kernels are latency bound (dependent brick-table and voxel
// loads); four resident waves per SIMD (<= 128 VGPRs) hide more of that
// latency than three with fewer spills.
#define VR_TRACING_KERNEL __global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 8)))
// TraceRays' own occupancy target (resident waves per SIMD): its candidate loop
// keeps more state live than the reuse kernels.
#ifndef VR_TRACE_WAVES
#define VR_TRACE_WAVES 4
#endif
#define VR_TRACE_KERNEL __global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(VR_TRACE_WAVES, 8)))

static inline dim3 tiles(int W, int H) { return dim3((W + 15) / 16, (H + 15) / 16, 1); }

// Expands `body` with a compile-time B for the four supported path lengths.
#define VR_DISPATCH_B(a, body)                                   \
    switch ((a).maxBounces) {                                    \
        case 1: { constexpr int B = 1; body; } break;            \
        case 2: { constexpr int B = 2; body; } break;            \
        case 3: { constexpr int B = 3; body; } break;            \
        case 4: { constexpr int B = 4; body; } break;            \
        default: return hipErrorInvalidValue;                    \
    }

// GenerateFeatures (VR/GenerateFeatures.cs.slang:57-102): the camera ray's
// transmittance (feature march) and the "no reflective surface" flag.
// With a surface scene the camera ray's closest triangle hit goes to the V-buffer,
// ends the march, and clears the flag unless the hit is an emitting front face.
VR_DEV void featuresBody(const Scene* __restrict__ sc, const Frame& fr, float2* __restrict__ features,
                         VItem* __restrict__ vbuf, VBlock vb) {
    const Scene& S = *sc;
    const Frame& F = fr;
    occCacheLoad(S, F.occSlot[0], F.occSlot[1]);
    PixelIdx px = pixelOf(F, vb);
    if (!px.ok) return;
    Rng rng{0, 0, 0, 0};
    Ray r{ld3(F.cam.pos), rayDir(F.cam, px.x, px.y, F.width, F.height), 0.f, kTMax};
    int noSurface = 1;
    if (F.useSurfaceScene) {
        SurfHit h;
        surfaceQuery(S, r.o, r.d, 0.f, kTMax, false, h);
        vbuf[px.id] = VItem{h.tri, h.bx, h.by, h.t};
        r.tMax = h.t;
        if (h.tri >= 0) noSurface = surfaceEmitsTowards(S, h.tri, -r.d) ? 1 : 0;
    }
    if (F.useReference) return;
    float accT = featureMarch(S, r, 0, true, F.initialOpt.visibilityTStepScale, rng);
    features[px.id] = make_float2(__int_as_float(noSurface), accT);
}

// IntegrateByVolumePathTracing (VR/VolumePathTracingFunctions.slang:3-132), NEE on.
VR_DEV v3 pathTrace(const Scene& S, const Frame& Fr, Ray ray, Rng& rng, int lightSamples, int maxBounces) {
    MediumInteraction mi{V3(0.f), V3(0.f), S.g, false};
    v3 beta = V3(1.f), L = V3(0.f);
    for (int bounce = 0; bounce < maxBounces; bounce++) {
        mi.valid = false;
        sampleDecomposition(S, ray, rng, mi, 0);
        if (mi.valid) {
            v3 albedo = ld3(S.sigma_s) / S.sigma_t;
            if (S.hasEmission) L = L + emissionWorld(S, mi.p) * (V3(1.f) - albedo) * beta;
            beta = beta * albedo;
            v3 Ld = directLightingPT(S, Fr.env[0], rng, mi, lightSamples, Fr.initialOpt, 0, 3u);
            L = L + beta * Ld;
            v3 wi = V3(0.f);
            if (maxBounces > 1) {
                float u0 = u01(rng);
                float u1 = u01(rng);
                samplePhase(mi.g, -ray.d, wi, u0, u1);
            }
            if (bounce < maxBounces - 1) {
                ray = {mi.p, wi, 0.f, kTMax};
                if (Fr.useRussianRoulette && bounce >= 2) {
                    if (u01(rng) < albedo.x) beta = beta / albedo.x;
                    else bounce = maxBounces;
                }
            }
        } else {
            if (bounce == 0) L = L + beta * envEval(S, Fr.env[0], ray.d);
            bounce = maxBounces;
        }
    }
    return L;
}

// TraceRays in reference mode (VR/TraceRays.cs.slang:85-103): the volume path tracer.
template <int B>
VR_DEV void referenceBody(const Scene* __restrict__ sc, const Frame& fr, float4* __restrict__ outColor, VBlock vb) {
    const Scene& S = *sc;
    const Frame& F = fr;
    occCacheLoad(S, F.occSlot[0], F.occSlot[1]);
    PixelIdx px = pixelOf(F, vb);
    if (!px.ok) return;
    Rng rng = rngCreate(px.x, px.y, (uint32_t)(F.numTotalRounds * F.frameCount));
    Ray ray{ld3(F.cam.pos), rayDir(F.cam, px.x, px.y, F.width, F.height), 0.f, kTMax};
    v3 avg = V3(0.f);
    for (int s = 0; s < F.baselineSpp; s++)
        avg = avg + pathTrace(S, F, ray, rng, max(1, F.initialOpt.lightSamples), B);
    avg = avg / (float)F.baselineSpp;
    outColor[px.id] = make_float4(avg.x, avg.y, avg.z, 1.f);
}
template <int B>
__global__ __launch_bounds__(256) void kReference(const Scene* __restrict__ sc, const Frame fr,
                                                  float4* __restrict__ outColor) {
    referenceBody<B>(sc, fr, outColor, launchBlock());
}

// How a stage kernel obtains transmittances: traced on the spot (kDirect), or
// split around a balanced trace of the recorded rays (kRecord, kReplay).
enum { kDirect = 0, kRecord = 1, kReplay = 2 };

// TraceRays (VR/TraceRays.cs.slang:64-201) for one pixel. IM: initial lighting
// tracking method; VM / LM: the spatial options' methods used by the final p-hat.
// LEARNED: the initial distances come from the learned proposal (the pixel's
// CDF over c drives the analytic tracker's optical-depth targets) and one pixel
// per block writes its candidates' weights as a training record.
// MODE (IM = 2, not LEARNED): kRecord writes the pixel's shadow rays, its
// free-flight samples, its resampling draws and its count of rays (sb) and
// nothing else; kReplay reads them back (CandReplay) and finishes the pixel, or,
// when a resampling step disagrees with the record's, appends the pixel to
// sb.plan / sb.fix instead.
template <int B, int IM, int VM, int LM, bool LEARNED, int MODE>
VR_DEV void tracePixel(const Scene& S, const Frame& F, PixelIdx px, ResStore* __restrict__ outRes,
                       float* __restrict__ outExtra, const LearnedTraceArgs& la, const SplitBufs& sb) {
    static_assert(MODE == kDirect || (IM == 2 && !LEARNED), "split TraceRays needs ray-marched shadow rays");
    using CP = std::conditional_t<MODE == kDirect, CandDirect<IM>,
                                  std::conditional_t<MODE == kRecord, CandRecord, CandReplay>>;
    CP cp{};
    if constexpr (MODE != kDirect) {
        const size_t np = (size_t)F.width * F.height;
        cp.n = 0;
        cp.dens = sb.dens + px.id;
        cp.dstride = np;
        cp.nd = 0;
        cp.cap = sb.cap;
        cp.mask = 0u;
        cp.words = 0;
        if constexpr (MODE == kRecord) {
            cp.out = sb.recs + (size_t)px.id * sb.cap;
            cp.bits = 0;
        } else {
            cp.tr = sb.trs + px.id;
            cp.trStride = np;
            cp.bits = 32;
            cp.zero = false;
        }
    }
    Rng rng = rngCreate(px.x, px.y, (uint32_t)(F.numTotalRounds * F.frameCount));
    Ray ray{ld3(F.cam.pos), rayDir(F.cam, px.x, px.y, F.width, F.height), 0.f, kTMax};
    constexpr int NE = B > 1 ? B - 1 : 1;
    v3 finalEx[NE];
#pragma unroll
    for (int i = 0; i < NE; i++) finalEx[i] = V3(0.f);
    Res fin = newRes();
    const int M = F.initialM;
    const int rounds = (M + 3) / 4;
    float* rec = nullptr;
    if constexpr (LEARNED) {
        const int slot = learned::trainSlot(px.x, px.y, F.width, F.height, la.bx, la.by, F.frameCount);
        if (slot >= 0) {
            rec = la.recs + (size_t)slot * learned::kRecFloats;
            float raw[learned::kRaw];
            learned::rawInputs(S, ray, raw);
            float4* r4 = reinterpret_cast<float4*>(rec);
            r4[0] = make_float4(raw[0], raw[1], raw[2], raw[3]);
            r4[1] = make_float4(raw[4], raw[5], raw[6], 0.f);
#pragma unroll
            for (int i = 2; i < learned::kRecFloats / 4; i++) r4[i] = make_float4(0.f, 0.f, 0.f, 0.f);
            if (!(raw[6] > 0.f)) rec = nullptr;  // a ray that misses the volume does not train
        }
    }
    for (int r = 0; r < rounds; r++) {
        float h[4] = {0, 0, 0, 0}, pd[4] = {0, 0, 0, 0}, tr[4] = {0, 0, 0, 0};
        int rs = r == rounds - 1 ? (M - 4 * (rounds - 1)) : 4;
        float gc[4] = {1.f, 1.f, 1.f, 1.f};
        int bins[4] = {0, 0, 0, 0};
        if constexpr (LEARNED) {
            learned::Cdf cdf;
            cdf.load(la.cdf + (size_t)px.id * learned::kOut);
            float targets[4] = {0, 0, 0, 0};
#pragma unroll
            for (int s = 0; s < 4; s++)
                if (s < rs) targets[s] = -logf(1 - cdf.sample(u01(rng), bins[s], gc[s]));
            sampleAnalyticTargets(S, ray, rng, F.initialOpt.visibilityUseLinearSampler != 0,
                                  F.initialOpt.visibilityMipLevel, rs, targets, h, pd, tr);
        } else if (!F.noReuse)
            cp.analytic(S, ray, rng, F.initialOpt.visibilityUseLinearSampler != 0, F.initialOpt.visibilityMipLevel, rs,
                        h, pd, tr);
        for (int s = 0; s < rs; s++) {
            v3 ex[NE];
#pragma unroll
            for (int i = 0; i < NE; i++) ex[i] = V3(0.f);
            float hs = s == 0 ? h[0] : (s == 1 ? h[1] : (s == 2 ? h[2] : h[3]));
            float ps = s == 0 ? pd[0] : (s == 1 ? pd[1] : (s == 2 ? pd[2] : pd[3]));
            float ts = s == 0 ? tr[0] : (s == 1 ? tr[1] : (s == 2 ? tr[2] : tr[3]));
            float ratio = 1.f, gs = 1.f;
            int bs = 0;
            if constexpr (LEARNED) {
                // the learned source pdf over the analytic one: g(c) inside the
                // volume, the CDF mass above 1 - Tr over Tr for an escape
                gs = s == 0 ? gc[0] : (s == 1 ? gc[1] : (s == 2 ? gc[2] : gc[3]));
                bs = s == 0 ? bins[0] : (s == 1 ? bins[1] : (s == 2 ? bins[2] : bins[3]));
                ratio = gs;
                if (hs == kTMax) {
                    learned::Cdf cdf;
                    cdf.load(la.cdf + (size_t)px.id * learned::kOut);
                    ratio = ts > 0.f ? cdf.massAbove(1.f - ts) / ts : gs;
                }
                ps *= ratio;
            }
            Res o = computeInitialSample<B, IM>(S, F, ray, hs, ps, ts, rng, F.initialOpt, ex, cp);
            if constexpr (LEARNED) {
                if (rec) {
                    // the sample's weight against the analytic proposal over g(c)
                    const float wt = o.runningSum * (ratio / gs);
                    rec[7] += wt;
                    rec[learned::kRecHist + bs] += wt;
                    rec[learned::kRecSq + bs] += wt * wt * gs;
                }
            }
            bool sel = cp.resample(o, fin, rng);
            if (sel && B > 1) {
                int mi = decodeMaxIndirect(fin.sampledPixel);
#pragma unroll
                for (int k = 0; k < NE; k++)
                    if (k < mi) finalEx[k] = ex[k];
            }
        }
    }
    if constexpr (MODE == kRecord) {
        cp.finish();
        sb.counts[px.id] = cp.n;
        return;
    } else if constexpr (MODE == kReplay) {
        if (cp.zero) {
            sb.plan[atomicAdd(sb.fix, 1)] = (uint32_t)px.id;
            return;
        }
    }
    // Write the selected extra-bounce records first: the p-hat re-evaluation
    // below reads them back from the same slots (VR/TraceRays.cs.slang:169-198).
    const int exStart = px.id * (B - 1);
    if (B > 1) {
        int mi = decodeMaxIndirect(fin.sampledPixel);
#pragma unroll
        for (int k = 0; k < NE; k++)
            if (k < mi) {
                float* e = outExtra + 3 * (size_t)(exStart + k);
                e[0] = finalEx[k].x; e[1] = finalEx[k].y; e[2] = finalEx[k].z;
            }
    }
    fin.exStart = exStart;
    float ph = lum(evaluateF<B, VM, LM>(S, F.env[0], fin, outExtra, ray, rng, F.spatialOpt, false, false));
    if (fin.runningSum > 0.f) {
        fin.runningSum *= fin.p_y == 0.f ? 0.f : ph / fin.p_y;
        fin.p_y = ph;
    }
    storeRes(outRes, px.id, fin);
}
template <int B, int IM, int VM, int LM, bool LEARNED = false>
VR_DEV void traceBody(const Scene* __restrict__ sc, const Frame& fr, ResStore* __restrict__ outRes,
                      float* __restrict__ outExtra, const LearnedTraceArgs& la, VBlock vb) {
    const Scene& S = *sc;
    occCacheLoad(S, fr.occSlot[0], fr.occSlot[1]);
    PixelIdx px = pixelOf(fr, vb);
    if (!px.ok) return;
    tracePixel<B, IM, VM, LM, LEARNED, kDirect>(S, fr, px, outRes, outExtra, la, SplitBufs{});
}
template <int B, int IM, int VM, int LM, bool LEARNED = false>
VR_TRACE_KERNEL void kTrace(const Scene* __restrict__ sc, const Frame fr, ResStore* __restrict__ outRes,
                              float* __restrict__ outExtra, const LearnedTraceArgs la) {
    traceBody<B, IM, VM, LM, LEARNED>(sc, fr, outRes, outExtra, la, launchBlock());
}
// Split TraceRays: the record (MODE = kRecord) and replay (kReplay) launches
// around the brick-segment trace of the recorded shadow rays.
template <int B, int VM, int LM, int MODE>
VR_TRACE_KERNEL void kTraceSplit(const Scene* __restrict__ sc, const Frame fr, ResStore* __restrict__ outRes,
                                   float* __restrict__ outExtra, const SplitBufs sb) {
    const Scene& S = *sc;
    occCacheLoad(S, fr.occSlot[0], fr.occSlot[1]);
    PixelIdx px = pixelOf(fr, launchBlock());
    if (!px.ok) return;
    tracePixel<B, 2, VM, LM, false, MODE>(S, fr, px, outRes, outExtra, LearnedTraceArgs{}, sb);
}
// The pixels the replay left (sb.plan[0, *sb.fix)), each traced again in one
// kernel from its first random number; a fixed grid strides over the list.
template <int B, int VM, int LM>
VR_TRACE_KERNEL void kTraceFix(const Scene* __restrict__ sc, const Frame fr, ResStore* __restrict__ outRes,
                                 float* __restrict__ outExtra, const SplitBufs sb) {
    const Scene& S = *sc;
    occCacheLoad(S, fr.occSlot[0], fr.occSlot[1]);
    const int n = *sb.fix;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int id = (int)sb.plan[i];
        const PixelIdx px{id % fr.width, id / fr.width, id, true};
        tracePixel<B, 2, VM, LM, false, kDirect>(S, fr, px, outRes, outExtra, LearnedTraceArgs{}, sb);
    }
}

VR_DEV float2 featuresAt(const float2* f, int id, int n) {
    if (id < 0 || id >= n) return make_float2(0.f, 0.f);  // out-of-range loads read 0
    return f[id];
}
VR_DEV bool isBackground(float2 f) { return f.y == 1.f && __float_as_int(f.x) != 0; }

template <int MODE, int VM, int LM>
using TracePolicy = std::conditional_t<MODE == kDirect, DirectTrace<VM, LM>,
                                       std::conditional_t<MODE == kRecord, RecordTrace, ReplayTrace>>;

// TemporalReuse (VR/TemporalReuse.cs.slang:80-377). MODE: kDirect traces each
// transmittance ray on the spot; kRecord writes them to sb (and nothing else:
// which rays are traced depends on no transmittance, and the ray-marching
// trackers draw no random numbers, so resampling can be left out); kReplay reads
// their transmittances back in the same order.
template <int B, int VM, int LM, int MODE = kDirect>
VR_DEV void temporalBody(const Scene* __restrict__ sc, const Frame& fr, ResStore* __restrict__ cur,
                         float* __restrict__ curExtra, const ResStore* __restrict__ temporal,
                         const float* __restrict__ temporalExtra, const float2* __restrict__ features,
                         const float2* __restrict__ temporalFeatures, float2* __restrict__ mvec, const SplitBufs& sb,
                         VBlock vb) {
    const Scene& S = *sc;
    const Frame& F = fr;
    occCacheLoad(S, F.occSlot[0], F.occSlot[1]);
    PixelIdx px = pixelOf(F, vb);
    if (!px.ok) return;
    TracePolicy<MODE, VM, LM> tp;
    if constexpr (MODE == kRecord) { tp.out = sb.recs + (size_t)px.id * sb.cap; tp.n = 0; tp.dens = sb.dens + px.id; tp.dstride = (size_t)F.width * F.height; tp.nd = 0; sb.counts[px.id] = 0; }
    if constexpr (MODE == kReplay) { tp.tr = sb.trs + px.id; tp.trStride = (size_t)F.width * F.height; tp.n = 0; tp.dens = sb.dens + px.id; tp.dstride = (size_t)F.width * F.height; tp.nd = 0; }
    constexpr bool kWrites = MODE != kRecord;
    if (F.isFirstFrame) return;
    const int W = F.width, H = F.height, N = W * H;
    Rng rng = rngCreate(px.x, px.y, (uint32_t)(F.numTotalRounds * F.frameCount + 1));
    int selectedId = -1;
    int numUsed = 1;
    Res taps[2];
    taps[0] = loadRes(cur, px.id, B);
    taps[1] = newRes();
    Ray ray{ld3(F.cam.pos), rayDir(F.cam, px.x, px.y, W, H), 0.f, kTMax};
    const bool talbot = F.temporalMIS == 1;
    Res output = talbot ? newRes() : taps[0];
    const int centerEx = taps[0].exStart;
    float temporalOriginalDepth = 0.f;
    int rsx = 0, rsy = 0;
    bool isBgRes = isBackground(features[px.id]);
    bool useFallback = true;
    if (F.reprojectionMode != 1) {
        float reprojDepth = taps[0].depth;
        if (reprojDepth == kTMax && F.reprojectionMode != 2 && !isBgRes)
            reprojDepth = rejectionSampleDepth(S, ray, rng, 8 + F.reprojectionMip);
        v3 p = ray.o + ray.d * reprojDepth;
        // last frame's position of a moving medium (VR/TemporalReuse.cs.slang:161-166)
        if (S.hasVelocity && F.hasAnimation) p = p - velocityWorld(S, p) * S.velocityScale;
        const float* V = F.prevView;
        const float* P = F.prevProj;
        float vx = p.x * V[0] + p.y * V[4] + p.z * V[8] + V[12];
        float vy = p.x * V[1] + p.y * V[5] + p.z * V[9] + V[13];
        float vz = p.x * V[2] + p.y * V[6] + p.z * V[10] + V[14];
        float vw = p.x * V[3] + p.y * V[7] + p.z * V[11] + V[15];
        float cx = vx * P[0] + vy * P[4] + vz * P[8] + vw * P[12];
        float cy = vx * P[1] + vy * P[5] + vz * P[9] + vw * P[13];
        float cw = vx * P[3] + vy * P[7] + vz * P[11] + vw * P[15];
        float sx = cx / cw, sy = cy / cw;
        int sxi, syi;
        if (reprojDepth == kTMax) {
            sx = (float)px.x + 0.5f; sy = (float)px.y + 0.5f;
            sxi = px.x; syi = px.y;
        } else {
            sx = 0.5f * sx + 0.5f;
            sy = -0.5f * sy + 0.5f;
            sx *= (float)W; sy *= (float)H;
            sxi = (int)sx; syi = (int)sy;
        }
        {
            float2 tf = featuresAt(temporalFeatures, syi * W + sxi, N);
            if (isBgRes && !isBackground(tf)) {
                if (kWrites && F.outputMotionVec)
                    mvec[px.id] = make_float2((float)(rsx - px.x) / W, (float)(rsy - px.y) / H);
                return;
            }
        }
        sxi = (int)sx; syi = (int)sy;
        rsx = sxi; rsy = syi;
        if (inRange(sxi, syi, W, H)) {
            numUsed++;
            taps[1] = loadRes(temporal, syi * W + sxi, B);
        }
        if (numUsed > 1) useFallback = false;
    }
    if (useFallback) {
        numUsed++;
        rsx = px.x; rsy = px.y;
        taps[1] = loadRes(temporal, px.id, B);
    }
    if (kWrites && F.outputMotionVec) mvec[px.id] = make_float2((float)(rsx - px.x) / W, (float)(rsy - px.y) / H);
    float maxPrevM = F.temporalThreshold * taps[0].M;
    if (numUsed == 2) {
        temporalOriginalDepth = taps[1].depth;
        if (taps[1].depth != kTMax) {
            v3 pd = norm3(rayDirNN(F.prevCam, rsx, rsy, W, H));
            v3 wp = ld3(F.prevCam.pos) + taps[1].depth * pd;
            taps[1].depth = len3(wp - ray.o);
        }
    }
    float centerPrevDepth = taps[0].depth;
    if (centerPrevDepth != kTMax) centerPrevDepth = len3(at(ray, centerPrevDepth) - ld3(F.prevCam.pos));
    bool hasSel = output.runningSum > 0.f;
    int start = talbot ? 0 : 1;
    if (start == 1) selectedId = 0;
    for (int i = start; i < numUsed; i++) {
        Res& ti = i == 0 ? taps[0] : taps[1];
        float talbotW = 1.f;
        float nPy = 0.f;
        if (ti.p_y > 0.f) {
            nPy = ti.p_y;
            if (isnan(ti.runningSum) || isinf(ti.runningSum)) ti.runningSum = 0.f;
            if (i > 0 && ti.runningSum != 0.f) {
                float ph = lum(evaluateF<B>(S, F.env[0], ti, temporalExtra, ray, rng, F.spatialOpt, false, false, tp));
                float w = ph / ti.p_y;
                if (isinf(w) || isnan(w)) w = 0.f;
                ti.runningSum *= w;
                ti.p_y = ph;
            }
        } else {
            ti.p_y = 0.f;
            ti.runningSum = 0.f;
        }
        if (talbot && ti.runningSum > 0.f) {
            float psum = 0.f, pqi = 0.f, k = 0.f;
            for (int j = 0; j < numUsed; j++) {
                const Res& tj = j == 0 ? taps[0] : taps[1];
                float cM = fminf(maxPrevM, tj.M);
                k += cM;
                if (j == 0) {
                    pqi = ti.p_y;
                    psum += ti.p_y * cM;
                } else if (i == j) {
                    pqi = nPy;
                    psum += nPy * cM;
                } else {
                    // j == 1, i == 0: the current sample seen from the previous camera
                    Ray nr{ld3(F.prevCam.pos), norm3(rayDirNN(F.prevCam, rsx, rsy, W, H)), 0.f, centerPrevDepth};
                    Res tmp = ti;
                    tmp.depth = centerPrevDepth;
                    // isLastFrame (VR/TemporalReuse.cs.slang:335-338): previous orientation and grids
                    float py = lum(evaluateF<B>(S, F.env[1], tmp, curExtra, nr, rng, F.spatialOpt, false, false, tp,
                                                F.prevGridOffset));
                    if (isinf(py) || isnan(py)) py = 0.f;
                    psum += py * cM;
                }
            }
            if (psum > 0) talbotW = pqi * k / psum;
        }
        if constexpr (kWrites) {
            ti.runningSum *= talbotW;
            bool curSel = resampleMaxM(ti, maxPrevM, output, rng);
            hasSel |= curSel;
            if (curSel) selectedId = i;
        }
    }
    if constexpr (MODE == kRecord) {
        sb.counts[px.id] = tp.n;
        return;
    }
    if (B > 1) {
        if (hasSel && selectedId > 0) {
            int mi = decodeMaxIndirect(output.sampledPixel);
            for (int k = 0; k < B - 1; k++) {
                if (k >= mi) break;
                const float* s = temporalExtra + 3 * (size_t)(output.exStart + k);
                float* d = curExtra + 3 * (size_t)(centerEx + k);
                d[0] = s[0]; d[1] = s[1]; d[2] = s[2];
            }
        }
        output.exStart = centerEx;
    }
    storeRes(cur, px.id, output);
}
template <int B, int VM, int LM, int MODE = kDirect>
VR_TRACING_KERNEL void kTemporal(const Scene* __restrict__ sc, const Frame fr, ResStore* __restrict__ cur,
                                 float* __restrict__ curExtra, const ResStore* __restrict__ temporal,
                                 const float* __restrict__ temporalExtra, const float2* __restrict__ features,
                                 const float2* __restrict__ temporalFeatures, float2* __restrict__ mvec,
                                 SplitBufs sb) {
    temporalBody<B, VM, LM, MODE>(sc, fr, cur, curExtra, temporal, temporalExtra, features, temporalFeatures, mvec, sb,
                                  launchBlock());
}

// SpatialReuse (VR/SpatialReuse.cs.slang:94-265) for one round. MODE: the
// direct form, or the record / replay passes around a balanced trace of the
// recorded rays (evaluateF's trace policies). The replay pass performs the
// Talbot evaluations the record pass recorded; it uses their weights only
// where the direct form would have made them (the record pass decides with
// every visibility 1, a superset).
template <int B, int VM, int LM, int MODE>
VR_DEV void spatialBody(const Scene* __restrict__ sc, const Frame& fr, int roundId, const int* __restrict__ offsets,
                        const ResStore* __restrict__ in, ResStore* __restrict__ out, const float* __restrict__ inExtra,
                        float* __restrict__ outExtra, const float2* __restrict__ features, const SplitBufs& sb,
                        VBlock vb) {
    const Scene& S = *sc;
    const Frame& F = fr;
    occCacheLoad(S, F.occSlot[0], F.occSlot[1]);
    PixelIdx px = pixelOf(F, vb);
    if (!px.ok) return;
    const int W = F.width, H = F.height;
    const int roundOffset = (F.enableTemporal ? 1 : 0) + 1;
    const int numRounds = F.spatialRounds + roundOffset + 1;
    const int rid = roundId + roundOffset;
    const int r2Seed = ((F.spatialRounds + 1) * F.frameCount + roundId) % 16;
    const int K = F.spatialSampleCount;
    const bool talbot = F.spatialMIS == 1;
    TracePolicy<MODE, VM, LM> tp;
    if constexpr (MODE == kRecord) { tp.out = sb.recs + (size_t)px.id * sb.cap; tp.n = 0; tp.dens = sb.dens + px.id; tp.dstride = (size_t)F.width * F.height; tp.nd = 0; }
    if constexpr (MODE == kReplay) { tp.tr = sb.trs + px.id; tp.trStride = (size_t)F.width * F.height; tp.n = 0; tp.dens = sb.dens + px.id; tp.dstride = (size_t)F.width * F.height; tp.nd = 0; }
    uint32_t plan = 0;
    if constexpr (MODE == kReplay) plan = sb.plan[px.id];
    Rng rng = rngCreate(px.x, px.y, (uint32_t)(numRounds * F.frameCount + rid));
    Res inCenter = loadRes(in, px.id, B);
    Res output = inCenter;
    const int centerEx = output.exStart;
    Ray ray{ld3(F.cam.pos), rayDir(F.cam, px.x, px.y, W, H), 0.f, kTMax};
    if (talbot) output = newRes();
    bool hasSel = output.runningSum > 0.f;
    float2 cf = features[px.id];
    if (!(cf.y != 1.f)) {  // background pixel: pass through
        if constexpr (MODE == kRecord) {
            sb.counts[px.id] = 0;
            sb.plan[px.id] = 0u;
            return;
        }
        storeRes(out, px.id, talbot ? inCenter : output);
        if (B > 1) {
            int mi = decodeMaxIndirect(output.sampledPixel);
            for (int k = 0; k < B - 1; k++) {
                if (k >= mi) break;
                const float* s = inExtra + 3 * (size_t)(output.exStart + k);
                float* d = outExtra + 3 * (size_t)(output.exStart + k);
                d[0] = s[0]; d[1] = s[1]; d[2] = s[2];
            }
        }
        return;
    }
    const int* offs = offsets + r2Seed * kMaxSpatialSamples * 2;
    for (int s = talbot ? 0 : 1; s < K; s++) {
        int tx = px.x + offs[2 * s], ty = px.y + offs[2 * s + 1];
        if (!inRange(tx, ty, W, H)) continue;
        Res tap = loadRes(in, ty * W + tx, B);
        float misW = 1.f;
        if (s > 0 && tap.runningSum != 0.f) {
            float ph = lum(evaluateF<B>(S, F.env[0], tap, inExtra, ray, rng, F.spatialOpt, false, false, tp));
            float w = ph / tap.p_y;
            if (isinf(w) || isnan(w)) w = 0.f;
            tap.runningSum *= w;
            tap.p_y = ph;
        }
        const bool talUse = talbot && tap.runningSum > 0.f;
        bool talRun = talUse;
        if constexpr (MODE == kRecord) { if (talUse) plan |= 1u << s; }
        if constexpr (MODE == kReplay) talRun = (plan >> s) & 1u;
        if (talRun) {
            float psum = 0.f, pqi = 0.f, k = 0.f;
            for (int j = 0; j < K; j++) {
                int t2x = px.x + offs[2 * j], t2y = px.y + offs[2 * j + 1];
                if (!inRange(t2x, t2y, W, H)) continue;
                int t2 = t2y * W + t2x;
                float m2 = in[t2].a.y;
                k += m2;
                if (j == 0) {
                    pqi = tap.p_y;
                    psum += tap.p_y * m2;
                } else if (s == j) {
                    float py2 = in[t2].a.w;
                    pqi = py2;
                    psum += py2 * m2;
                } else {
                    Ray nr{ray.o, rayDir(F.cam, t2x, t2y, W, H), 0.f, tap.depth};
                    float py = lum(evaluateF<B>(S, F.env[0], tap, inExtra, nr, rng, F.spatialOpt, false, false, tp));
                    if (isinf(py) || isnan(py)) py = 0.f;
                    psum += py * m2;
                }
            }
            if (talUse && psum > 0) misW = pqi * k / psum;
        }
        if constexpr (MODE != kRecord) {
            tap.runningSum *= misW;
            if (resample(tap, output, rng)) hasSel = true;
        }
    }
    if constexpr (MODE == kRecord) {
        sb.counts[px.id] = tp.n;
        sb.plan[px.id] = plan;
        return;
    }
    if (B > 1 && hasSel) {
        int mi = decodeMaxIndirect(output.sampledPixel);
        for (int k = 0; k < B - 1; k++) {
            if (k >= mi) break;
            const float* s2 = inExtra + 3 * (size_t)(output.exStart + k);
            float* d = outExtra + 3 * (size_t)(centerEx + k);
            d[0] = s2[0]; d[1] = s2[1]; d[2] = s2[2];
        }
    }
    if (B > 1) output.exStart = centerEx;
    storeRes(out, px.id, output);
}
template <int B, int VM, int LM, int MODE>
VR_TRACING_KERNEL void kSpatial(const Scene* __restrict__ sc, const Frame fr, int roundId,
                                const int* __restrict__ offsets, const ResStore* __restrict__ in,
                                ResStore* __restrict__ out, const float* __restrict__ inExtra,
                                float* __restrict__ outExtra, const float2* __restrict__ features, SplitBufs sb) {
    spatialBody<B, VM, LM, MODE>(sc, fr, roundId, offsets, in, out, inExtra, outExtra, features, sb, launchBlock());
}

// CopyReservoirs (VR/CopyReservoirs.cs.slang:84-95).
template <int B>
VR_DEV void copyBody(const Frame& fr, const ResStore* __restrict__ cur, const float* __restrict__ curExtra,
                     ResStore* __restrict__ temporal, float* __restrict__ temporalExtra, VBlock vb) {
    const Frame& F = fr;
    PixelIdx px = pixelOf(F, vb);
    if (!px.ok) return;
    ResStore r = cur[px.id];
    temporal[px.id] = r;
    if (B > 1) {
        int mi = decodeMaxIndirect(__float_as_int(r.b.w));
        int st = px.id * (B - 1);
        for (int k = 0; k < B - 1; k++) {
            if (k >= mi) break;
            const float* s = curExtra + 3 * (size_t)(st + k);
            float* d = temporalExtra + 3 * (size_t)(st + k);
            d[0] = s[0]; d[1] = s[1]; d[2] = s[2];
        }
    }
}
template <int B>
__global__ __launch_bounds__(256) void kCopy(const Frame fr, const ResStore* __restrict__ cur,
                                             const float* __restrict__ curExtra, ResStore* __restrict__ temporal,
                                             float* __restrict__ temporalExtra) {
    copyBody<B>(fr, cur, curExtra, temporal, temporalExtra, launchBlock());
}

// FinalShading (VR/FinalShading.cs.slang:71-141). MODE: kDirect traces each
// transmittance ray on the spot; kRecord writes them to sb (and nothing else);
// kReplay reads their transmittances back in the same order.
template <int B, int VM, int LM, int MODE = kDirect>
VR_DEV void finalBody(const Scene* __restrict__ sc, const Frame& fr, const ResStore* __restrict__ cur,
                      const float* __restrict__ curExtra, const float2* __restrict__ features,
                      const float4* __restrict__ colorIn, float4* __restrict__ out, const SplitBufs& sb, VBlock vb) {
    const Scene& S = *sc;
    const Frame& F = fr;
    occCacheLoad(S, F.occSlot[0], F.occSlot[1]);
    PixelIdx px = pixelOf(F, vb);
    if (!px.ok) return;
    Rng rng = rngCreate(px.x, px.y, (uint32_t)(F.numTotalRounds * F.frameCount + F.numTotalRounds - 1));
    TracePolicy<MODE, VM, LM> tp;
    if constexpr (MODE == kRecord) { tp.out = sb.recs + (size_t)px.id * sb.cap; tp.n = 0; tp.dens = sb.dens + px.id; tp.dstride = (size_t)F.width * F.height; tp.nd = 0; }
    if constexpr (MODE == kReplay) { tp.tr = sb.trs + px.id; tp.trStride = (size_t)F.width * F.height; tp.n = 0; tp.dens = sb.dens + px.id; tp.dstride = (size_t)F.width * F.height; tp.nd = 0; }
    v3 c = V3(0.f);
    if (F.useReference) {
        float4 v = colorIn[px.id];
        c = V3(v.x, v.y, v.z);
    } else if (F.visualizeTransmittance) {
        c = V3(powf(features[px.id].y, 2.2f));
    } else {
        Res r = loadRes(cur, px.id, B);
        if (r.runningSum > 0.f) {
            Ray ray{ld3(F.cam.pos), rayDir(F.cam, px.x, px.y, F.width, F.height), 0.f, r.depth};
            v3 f = evaluateF<B>(S, F.env[0], r, curExtra, ray, rng, F.finalOpt, F.noReuse != 0, true, tp);
            float Wt = r.p_y == 0.0f ? 1.f : r.runningSum / (r.p_y * r.M);
            c = c + f * Wt;
        }
    }
    if constexpr (MODE == kRecord) {
        sb.counts[px.id] = tp.n;
        return;
    }
    float4 col = make_float4(c.x, c.y, c.z, 1.f);
    if (isnan(c.x) || isnan(c.y) || isnan(c.z) || isinf(c.x) || isinf(c.y) || isinf(c.z))
        col = make_float4(0.f, 0.f, 0.f, 0.f);
    out[px.id] = col;
}
template <int B, int VM, int LM, int MODE = kDirect>
__global__ __launch_bounds__(256) void kFinal(const Scene* __restrict__ sc, const Frame fr,
                                              const ResStore* __restrict__ cur, const float* __restrict__ curExtra,
                                              const float2* __restrict__ features, const float4* __restrict__ colorIn,
                                              float4* __restrict__ out, SplitBufs sb) {
    finalBody<B, VM, LM, MODE>(sc, fr, cur, curExtra, features, colorIn, out, sb, launchBlock());
}


}  // namespace vr
