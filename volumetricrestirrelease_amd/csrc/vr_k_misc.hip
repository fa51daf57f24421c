// vr_k_misc.hip -- GenerateFeatures, the batched tracking primitive, and the
// stage dispatcher (CopyReservoirs is instantiated here).
#include "vr_stages.hpp"

namespace vr {

// GenerateFeatures (VR/GenerateFeatures.cs.slang:57-102); the body is in vr_stages.hpp.
__global__ __launch_bounds__(256) void kFeatures(const Scene* __restrict__ sc, const Frame fr,
                                                 float2* __restrict__ features, VItem* __restrict__ vbuf) {
    featuresBody(sc, fr, features, vbuf, launchBlock());
}


// The hierarchical walk's visit sequence: the occupied bricks handed to a
// tracker, hashed in order (FNV-1a over their index-space minima).
struct VisitHash {
    uint32_t count = 0, hash = 2166136261u;
    float tSum = 0.f, tFirst = 0.f, tLast = 0.f, tNear = 0.f, tFar = 0.f;
    VR_DEV void mix(uint32_t v) {
#pragma unroll
        for (int k = 0; k < 4; k++) { hash ^= (v >> (8 * k)) & 0xFFu; hash *= 16777619u; }
    }
    VR_DEV void setRayInfo(float a, float b, const Ray&) { tNear = a; tFar = b; }
    VR_DEV void start() {}
    VR_DEV bool brickStep(const Dda&, const BrickCtx& bc, int, float& t, Rng&) {
        mix((uint32_t)(int)bc.vmin.x); mix((uint32_t)(int)bc.vmin.y); mix((uint32_t)(int)bc.vmin.z);
        if (count == 0) tFirst = t;
        tLast = t;
        tSum += t;
        count++;
        return false;
    }
    VR_DEV void end() {}
};

// Brick and voxel indexing at an index-space point: point-sampled density, the
// leaf brick's minimum (-1 when none) and the voxel.
VR_DEV void indexAt(const Scene& S, int slot, v3 pW, float o[8]) {
    const Grid& G = S.grid[slot];
    const v3 p = xformP(G.w2i, pW);
    o[1] = densityIndex(S, slot, p);
    o[2] = o[3] = o[4] = -1.f;
    o[5] = floorf(p.x); o[6] = floorf(p.y); o[7] = floorf(p.z);
    if (p.x < 0.f || p.y < 0.f || p.z < 0.f || p.x >= G.bmax[0] || p.y >= G.bmax[1] || p.z >= G.bmax[2]) return;
    const int bx = (int)p.x / 8, by = (int)p.y / 8, bz = (int)p.z / 8;
    const uint32_t b = brickLookup(G, slot, bx, by, bz);
    if (b == kEmpty) return;
    const v3 q = p - V3((float)(bx * 8), (float)(by * 8), (float)(bz * 8));
    o[0] = sampleBrick(G, b, q, false) * S.densityScale;
    o[2] = (float)(bx * 8); o[3] = (float)(by * 8); o[4] = (float)(bz * 8);
}

// Batched tracking primitive: one lane per ray.
__global__ __launch_bounds__(256) void kTrack(const Scene* __restrict__ sc, const EnvXf ex, int mode, int mip,
                                              int linear,
                                              float tStepScale, int numSamples, int n, const float* __restrict__ rays,
                                              const uint32_t* __restrict__ seeds, float* __restrict__ out,
                                              int occSlot) {
    const Scene& S = *sc;
    occCacheLoad(S, occSlot, -1);
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* r = rays + 8 * (size_t)i;
    Ray ray{V3(r[0], r[1], r[2]), V3(r[3], r[4], r[5]), 0.f, r[6]};
    Rng rng = rngCreate(seeds[3 * i], seeds[3 * i + 1], seeds[3 * i + 2]);
    float o[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    switch (mode) {
        case 0: o[0] = trAnalytic(S, ray, mip, rng, linear != 0); break;
        case 1: o[0] = trMarch(S, ray, mip, linear != 0, tStepScale, rng); break;
        case 2: o[0] = trResidual(S, ray, mip, rng, false, true); break;
        case 3: o[0] = trResidual(S, ray, mip, rng, false, false); break;
        case 4: o[0] = trResidual(S, ray, mip, rng, true, false); break;
        case 5: {
            float h[4], pd[4], tr[4];
            int ns = max(1, min(4, numSamples));
            sampleAnalytic(S, ray, rng, linear != 0, mip, ns, h, pd, tr);
            o[0] = h[0]; o[1] = pd[0]; o[2] = tr[0];
            o[3] = ns > 1 ? h[1] : 0.f; o[4] = ns > 1 ? pd[1] : 0.f; o[5] = ns > 1 ? tr[1] : 0.f;
            o[6] = ns > 2 ? h[2] : 0.f; o[7] = ns > 3 ? h[3] : 0.f;
            break;
        }
        case 6: {
            MediumInteraction mi{V3(0.f), V3(0.f), S.g, false};
            sampleDecomposition(S, ray, rng, mi, mip);
            o[0] = mi.valid ? 1.f : 0.f; o[1] = mi.p.x; o[2] = mi.p.y; o[3] = mi.p.z;
            break;
        }
        case 7: {
            float sx, sy, bound;
            cellByDensity(S, ray, rng, mip, sx, sy, bound);
            o[0] = sx; o[1] = sy; o[2] = bound;
            o[3] = rejectionSampleDepth(S, ray, rng, mip);
            break;
        }
        case 8: o[0] = featureMarch(S, ray, mip, linear != 0, tStepScale, rng); break;
        case 9: o[0] = densityWorld(S, mip, ray.o); break;
        case 10:
#pragma unroll
            for (int k = 0; k < 8; k++) o[k] = __uint_as_float(rngNext(rng));
            break;
        case 11: {  // environment radiance towards ray.d
            v3 L = envEval(S, ex, ray.d);
            o[0] = L.x; o[1] = L.y; o[2] = L.z;
            break;
        }
        case 12: {  // environment importance sample at (origin.x, origin.y)
            float pdf;
            v3 d = envSample(S, ex, r[0], r[1], pdf);
            v3 L = envEval(S, ex, d);
            int id; float u, v;
            encodeLightUV(S, ex, d, id, u, v);
            o[0] = d.x; o[1] = d.y; o[2] = d.z; o[3] = pdf;
            o[4] = L.x; o[5] = L.y; o[6] = L.z; o[7] = (float)id;
            break;
        }
        case 14: {  // velocity at the point origin
            v3 v = S.hasVelocity ? velocityWorld(S, ray.o) : V3(0.f);
            o[0] = v.x; o[1] = v.y; o[2] = v.z;
            break;
        }
        case 13: {  // emission and temperature at the point origin
            v3 E = emissionWorld(S, ray.o);
            o[0] = E.x; o[1] = E.y; o[2] = E.z;
            o[3] = S.hasEmission ? temperatureWorld(S, ray.o) : 0.f;
            break;
        }
        case 15: indexAt(S, mip, ray.o, o); break;
        case 16: {
            VisitHash vh;
            traverse(S, ray, mip, rng, vh, linear != 0);
            o[0] = (float)vh.count; o[1] = __uint_as_float(vh.hash); o[2] = vh.tSum;
            o[3] = vh.tFirst; o[4] = vh.tLast; o[5] = vh.tNear; o[6] = vh.tFar;
            break;
        }
        case 17: {  // scene light sample at the point origin
            LightSample ls;
            ls.Li = V3(0.f);
            ls.pdfArea = ls.rayDistance = 0.f;
            int idx = -9;
            float ux = 0.f, uy = 0.f;
            if (sampleLights(S, ex, ray.o, (numSamples & 1) != 0, (numSamples & 2) != 0, (numSamples & 4) != 0, rng,
                             ls, idx, ux, uy)) {
                o[0] = ls.Li.x; o[1] = ls.Li.y; o[2] = ls.Li.z; o[3] = ls.pdfArea;
                o[4] = (float)idx; o[5] = ux; o[6] = uy; o[7] = ls.rayDistance;
            } else {
                o[4] = -9.f;
            }
            break;
        }
        case 18: {  // emissive-triangle light evaluation at the point origin
            const int tri = (int)r[3];
            TriLightSample ts;
            if (S.numEmActive > 0 && tri >= 0 && tri < S.numEmTris && sampleTriangle(S, ray.o, tri, r[4], r[5], ts)) {
                const v3 L = ts.Le * ts.cosTheta / (ts.distance * ts.distance);
                o[0] = L.x; o[1] = L.y; o[2] = L.z; o[3] = ts.distance;
                o[4] = ts.dir.x; o[5] = ts.dir.y; o[6] = ts.dir.z; o[7] = 1.f;
            }
            break;
        }
        case 19: {  // closest and any hit over [0, tMax]
            SurfHit h;
            surfaceQuery(S, ray.o, ray.d, 0.f, ray.tMax, false, h);
            o[0] = h.t; o[1] = (float)h.tri; o[2] = h.bx; o[3] = h.by;
            SurfHit a;
            o[4] = surfaceQuery(S, ray.o, ray.d, 0.f, ray.tMax, true, a) ? 1.f : 0.f;
            break;
        }
        case 20: {  // shading at the closest hit: a BSDF sample (numSamples 0) or its evaluation (1)
            SurfHit h;
            if (!surfaceQuery(S, ray.o, ray.d, 0.f, ray.tMax, false, h)) break;
            const ShadingData sd = surfaceShading(S, h.tri, h.bx, h.by, -ray.d);
            BSDFSample bs;
            const bool ok = sampleBSDF(sd, rng, bs);
            if (numSamples == 0) {
                o[0] = bs.wi.x; o[1] = bs.wi.y; o[2] = bs.wi.z; o[3] = bs.pdf;
                o[4] = bs.weight.x; o[5] = bs.weight.y; o[6] = bs.weight.z; o[7] = ok ? (float)bs.lobe : -1.f;
            } else {
                const v3 f = evalBSDFCosine(sd, bs.wi);
                o[0] = f.x; o[1] = f.y; o[2] = f.z; o[3] = evalPdfBSDF(sd, bs.wi);
                o[4] = sd.emissive.x; o[5] = sd.emissive.y; o[6] = sd.emissive.z; o[7] = sd.NdotV;
            }
            break;
        }
        default: break;
    }
    float* d = out + 8 * (size_t)i;
#pragma unroll
    for (int k = 0; k < 8; k++) d[k] = o[k];
}


hipError_t launchStage(const LaunchArgs& a) {
    const dim3 g = tiles(a.width, a.height), blk(256, 1, 1);
    switch (a.stage) {
        case 0: hipLaunchKernelGGL(kFeatures, g, blk, a.occBytes, a.stream, a.scene, a.frame, a.features, a.vbuf); break;
        case 1: return launchTraceRays(a);
        case 2: return launchTemporalReuse(a);
        case 3: return launchSpatialReuse(a);
        case 4: VR_DISPATCH_B(a, hipLaunchKernelGGL(kCopy<B>, g, blk, 0, a.stream, a.frame, a.res0, a.extra0, a.resT,
                                                     a.extraT)) break;
        case 5: return launchFinalShading(a);
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launchTrack(hipStream_t stream, const Scene* scene, const EnvXf& env, int mode, int mip, int linear,
                       float tStepScale, int numSamples, int n, const float* rays, const uint32_t* seeds, float* out,
                       int occSlot, int occBytes) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(kTrack, dim3((n + 255) / 256), dim3(256), occBytes, stream, scene, env, mode, mip, linear,
                       tStepScale, numSamples, n, rays, seeds, out, occSlot);
    return hipGetLastError();
}

}  // namespace vr
