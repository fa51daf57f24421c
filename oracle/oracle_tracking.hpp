// oracle_tracking.hpp -- TEST INFRASTRUCTURE ONLY (the CPU oracle).
//
// Hierarchical 3D-DDA over the GVDB-shaped tree and the per-brick tracking
// adapters, restated from
//   VR/VolumeUtils.slang:171-378          VolumeTrackingGVDB + *Generic wrappers
//   GVS/gvdbDda.slang:86-157              HDDAState
//   VR/VolumeTrackingAdapterGVDB.slang    the adapters
//   VR/VolumeBase.slang:64-263            phase function, transforms, density
#pragma once
#include "oracle_volume.hpp"

namespace orc {

struct Ray { f3 origin, dir; float tMin, tMax; f3 at(float t) const { return origin + dir * t; } };

// Medium parameters (SC/SceneTypes.slang:85-110) plus the grid slots.
// Grid slots (VR/HostDeviceSharedConstants.slang:19-28).
static const int kTemperatureSlot = 16;
static const int kPrevDensityOffset = 19;  // previous frame's mips at 19 + k
static const int kPrevExtraOffset = 11;    // previous frame's temperature at 27

// A velocity field as dense voxel data (3 floats per voxel, world units per
// frame) with the brick occupancy GVDB would store: a lookup inside a brick that
// holds no non-zero voxel reads 0 (getNodeAtPoint finds no node).
struct VelocityField {
    bool valid = false;
    int nx = 0, ny = 0, nz = 0;
    float w2i[16];
    std::vector<float> v;           // 3 per voxel, x fastest
    std::vector<uint8_t> brickOcc;  // per 8^3 brick
};

struct Medium {
    GridSlot grid[29];                      // mips 0..7, conservative 8..15, temperature 16, previous frame 19..27
    VelocityField velocity;
    float velocityScale = 1.f;
    bool prevHasEmission = false;
    int numMips = 0;
    float sigma_t = 1.f;
    f3 sigma_s{1, 1, 1}, sigma_a{0, 0, 0};
    float g = 0.f;
    float densityScale = 1.f;   // densityScaleFactorByScaling
    float tStep = 1.f;          // mean axis length of xform[0] (SC/Scene.cpp:3253)
    float worldScaling = 1.f;
    float superVoxelDiag = 1.f; // SC/Scene.cpp:3077-3080
    bool hasEmission = false;
    float LeScale = 1.f, tempCutoff = 0.f, tempScale = 0.f;
    std::vector<f3> lut;                    // blackbody table rows
};

// ---------------------------------------------------------------- transforms
static inline Ray worldToMedium(const Medium& m, const Ray& r, int slot) {
    const float* M = m.grid[slot].invxform;
    f4 o = mulRow({r.origin.x, r.origin.y, r.origin.z, 1.f}, M);
    f4 d = mulRow({r.dir.x, r.dir.y, r.dir.z, 0.f}, M);
    return {{o.x / o.w, o.y / o.w, o.z / o.w}, {d.x, d.y, d.z}, r.tMin, r.tMax};
}
static inline f3 worldToMediumP(const Medium& m, f3 p, int slot) {
    f4 o = mulRow({p.x, p.y, p.z, 1.f}, m.grid[slot].invxform);
    return {o.x / o.w, o.y / o.w, o.z / o.w};
}

// Density (VR/VolumeBase.slang:234-243): trilinear, 0 outside [bmin, bmax).
static inline float densityIndex(const Medium& m, f3 p, int slot) {
    const GridSlot& g = m.grid[slot];
    if (p.x < g.bmin.x || p.y < g.bmin.y || p.z < g.bmin.z || p.x >= g.bmax.x || p.y >= g.bmax.y || p.z >= g.bmax.z)
        return 0.f;
    return valueAtPoint(g, p, true) * m.densityScale;
}
static inline float densityWorld(const Medium& m, f3 pW, int slot) { return densityIndex(m, worldToMediumP(m, pW, slot), slot); }

static inline float densityInAtlas(const Medium& m, f3 pAtlas, int slot, bool linear) {
    const GridSlot& g = m.grid[slot];
    return sampleAtlas(g, pAtlas, linear) * g.compressScale * m.densityScale;
}
// FetchEightVoxelsInAtlas (VR/VolumeBase.slang:245-252).
static inline void fetchEight(const Medium& m, f3 pAtlas, int slot, float v[8]) {
    const GridSlot& g = m.grid[slot];
    f3 anchor = f3floor(pAtlas - 0.5f) + 0.5f;
    for (int i = 0; i < 8; i++) {
        f3 c = anchor + mk3((float)(i % 2), (float)((i % 4) / 2), (float)(i / 4));
        v[i] = sampleAtlas(g, c, false) * g.compressScale * m.densityScale;
    }
}

// Bounds3f::IntersectP (VR/VolumeBase.slang:137-158) on [bmin, bmax) (-0.5 for vertex-centred cells).
static inline bool intersectVolumeBound(const Medium& m, const Ray& ray, float& tMin, float& tMax, int slot, bool vc) {
    f3 pMin = m.grid[slot].bmin, pMax = m.grid[slot].bmax;
    if (vc) { pMin = pMin - 0.5f; pMax = pMax - 0.5f; }
    float t0 = 0.f, t1 = ray.tMax;
    for (int i = 0; i < 3; ++i) {
        float inv = 1.f / comp(ray.dir, i);
        float tn = (comp(pMin, i) - comp(ray.origin, i)) * inv;
        float tf = (comp(pMax, i) - comp(ray.origin, i)) * inv;
        if (tn > tf) { float tmp = tn; tn = tf; tf = tmp; }
        t0 = tn > t0 ? tn : t0;
        t1 = tf < t1 ? tf : t1;
        if (t0 > t1) return false;
    }
    tMin = t0; tMax = t1;
    return true;
}

// ---------------------------------------------------------------- HDDA state (GVS/gvdbDda.slang:86-157)
struct HDDA {
    f3 pos, dir;
    i3 pStep;
    f3 tDel;
    f3 t;
    i3 p;
    f3 tSide;
    i3 mask;
    void setFromRay(f3 o, f3 d, f3 t0) {
        pos = o; dir = d;
        pStep = {d.x >= 0 ? 1 : -1, d.y >= 0 ? 1 : -1, d.z >= 0 ? 1 : -1};
        t = t0;
    }
    void prepare(f3 vmin, f3 vdel) {
        tDel = f3abs(vdel / dir);
        f3 pFlt = (pos + dir * t.x - vmin) / vdel;
        f3 fl = f3floor(pFlt);
        tSide = ((fl - pFlt + 0.5f) * f3of(pStep) + 0.5f) * tDel + t.x;
        p = i3floor(pFlt);
    }
    void prepareLeaf(f3 vmin) {
        tDel = f3abs(mk3(1.f) / dir);
        f3 pFlt = pos + dir * t.x - vmin;
        f3 fl = f3floor(pFlt);
        tSide = ((fl - pFlt + 0.5f) * f3of(pStep) + 0.5f) * tDel + t.x;
        p = i3floor(pFlt);
    }
    void next() {
        mask.x = (int)((tSide.x < tSide.y) & (tSide.x <= tSide.z));
        mask.y = (int)((tSide.y < tSide.z) & (tSide.y <= tSide.x));
        mask.z = (int)((tSide.z < tSide.x) & (tSide.z <= tSide.y));
        t.y = mask.x ? tSide.x : (mask.y ? tSide.y : tSide.z);
    }
    // The reference adds float(mask) * tDel; for an axis the ray never crosses
    // (direction component exactly 0, tDel = inf) that is 0 * inf = NaN, which
    // stalls the walk until the 4096-step cap. Adding tDel only on the stepped
    // axis is the same arithmetic everywhere else and keeps such rays walking.
    void step() {
        t.x = t.y;
        tSide = {mask.x ? tSide.x + tDel.x : tSide.x, mask.y ? tSide.y + tDel.y : tSide.y,
                 mask.z ? tSide.z + tDel.z : tSide.z};
        p = {p.x + mask.x * pStep.x, p.y + mask.y * pStep.y, p.z + mask.z * pStep.z};
    }
};

// Work counters (traversal steps, per-brick cell steps) for cost diagnostics.
inline thread_local long long g_ddaSteps = 0, g_cellSteps = 0;

// Tree constants for the 3-level layout: res, log2 res, child extent (voxels).
static const int kRes[3] = {8, 16, 32};
static const int kDim[3] = {3, 4, 5};
static const float kVdel[3] = {1.f, 8.f, 128.f};

// VolumeTrackingGVDB (VR/VolumeUtils.slang:171-282).
template <class A>
void trackGVDB(const Medium& m, const Ray& rWorld, int slot, SampleGen& sg, A& ad, bool vertexCenter) {
    const GridSlot& g = m.grid[slot];
    uint32_t nodeList[4];
    float tMaxL[4] = {0.f, 0.f, 0.f, 0.f};  // [3] is read (never used) when the walk leaves the root
    const float eps = 0.01f;
    int lev = 2, topLev = 2;
    nodeList[lev] = g.root.childList;
    Ray ray = worldToMedium(m, rWorld, slot);
    if (vertexCenter) ray.origin = ray.origin - 0.5f;
    float tNear = 0.f, tFar = 0.f;
    if (!intersectVolumeBound(m, ray, tNear, tFar, slot, vertexCenter)) { ad.end(); return; }
    ad.setRayInfo(tNear, tFar, ray);
    ad.start();
    f3 tStart{tNear, tFar, 0.f};
    f3 vmin = f3of(g.root.vmin);
    tStart.x += eps;
    tMaxL[lev] = tStart.y;
    HDDA dda;
    dda.setFromRay(ray.origin, ray.dir, tStart);
    dda.prepare(vmin, mk3(kVdel[lev]));
    auto outside = [&](int l) {
        return dda.p.x < 0 || dda.p.y < 0 || dda.p.z < 0 || dda.p.x > kRes[l] || dda.p.y > kRes[l] || dda.p.z > kRes[l];
    };
    if (vertexCenter) {
        int it = 0;
        while (it++ < 3 && outside(lev)) { dda.next(); dda.step(); dda.t.x += eps; }
    }
    for (int iter = 0; iter < 4096 && lev > 0 && lev <= topLev && !outside(lev); iter++) {
        g_ddaSteps++;
        dda.next();
        int b = (((dda.p.z << kDim[lev]) + dda.p.y) << kDim[lev]) + dda.p.x;
        uint32_t child = childOf(g, lev, nodeList[lev], b);
        if (child != kUndef) {
            if (lev == 1) {
                float t = dda.t.x - eps;
                const LeafNode& leaf = g.level0[child];
                f4 bd = leaf.bounds;
                f4 bs{bd.x * m.densityScale, bd.y * m.densityScale, bd.z * m.densityScale, bd.w * m.densityScale};
                bool exitNow = ad.mainStep(dda, f3of(leaf.vmin), f3of(leaf.offs), bs, slot, t, sg);
                if (exitNow) return;
                dda.step();
                dda.t.x += eps;
            } else {
                lev--;
                const InnerNode& n1 = g.level1[child];
                nodeList[lev] = n1.childList;
                vmin = f3of(n1.vmin);
                tMaxL[lev] = dda.t.y;
                dda.prepare(vmin, mk3(kVdel[lev]));
            }
        } else {
            dda.step();
            dda.t.x += eps;
        }
        while (dda.t.x > tMaxL[lev] && lev <= topLev) {
            lev++;
            if (lev <= topLev) {
                vmin = f3of(g.root.vmin);  // only the root lives above level 1
                dda.prepare(vmin, mk3(kVdel[lev]));
            }
        }
    }
    ad.end();
}

static inline bool leafInside(const HDDA& d) {
    return d.p.x >= 0 && d.p.y >= 0 && d.p.z >= 0 && d.p.x < 8 && d.p.y < 8 && d.p.z < 8;
}
static inline bool pInside(f3 p) {
    return p.x >= 0 && p.y >= 0 && p.z >= 0 && p.x < 8.f && p.y < 8.f && p.z < 8.f;
}

// Trilinear cell polynomial of sigma_t along a ray segment
// (VR/VolumeTrackingAdapterGVDB.slang:72-100, 286-316).
struct CellPoly { float c3, c2, c1, c0; };
static inline CellPoly cellPoly(const Medium& m, const HDDA& dl, f3 vminLeaf, f3 o, int slot, f3 d) {
    float v[8];
    fetchEight(m, f3of(dl.p) + o + 0.5f, slot, v);
    for (int i = 0; i < 8; i++) v[i] *= m.sigma_t;
    float v000 = v[0], v100 = v[1], v010 = v[2], v110 = v[3], v001 = v[4], v101 = v[5], v011 = v[6], v111 = v[7];
    float mxyz = v111 - v011 - v101 - v110 + v100 + v010 + v001 - v000;
    float mxy = v000 - v100 - v010 + v110;
    float mxz = v000 - v100 - v001 + v101;
    float myz = v000 - v010 - v001 + v011;
    float mx = v100 - v000, my = v010 - v000, mz = v001 - v000;
    f3 p0 = dl.pos + dl.t.x * dl.dir - (f3of(dl.p) + vminLeaf);
    CellPoly c;
    c.c3 = mxyz * d.x * d.y * d.z;
    c.c2 = (p0.z * d.x * d.y + p0.y * d.x * d.z + p0.x * d.y * d.z) * mxyz + mxy * d.x * d.y + mxz * d.x * d.z +
           myz * d.y * d.z;
    c.c1 = (p0.y * p0.z * d.x + p0.x * p0.z * d.y + p0.x * p0.y * d.z) * mxyz + mx * d.x + my * d.y + mz * d.z +
           (p0.y * d.x + p0.x * d.y) * mxy + (p0.z * d.x + p0.x * d.z) * mxz + (p0.z * d.y + p0.y * d.z) * myz;
    c.c0 = p0.x * p0.y * p0.z * mxyz + p0.x * p0.y * mxy + p0.x * p0.z * mxz + p0.y * p0.z * myz + p0.x * mx +
           p0.y * my + p0.z * mz + v000;
    return c;
}
static inline float polyTau(float t, const CellPoly& c) {
    float t2 = t * t, t3 = t2 * t, t4 = t2 * t2;
    return c.c3 * t4 / 4 + c.c2 * t3 / 3 + c.c1 * t2 / 2 + c.c0 * t;
}
static inline float polySigma(float t, const CellPoly& c) {
    float t2 = t * t, t3 = t2 * t;
    return c.c3 * t3 + c.c2 * t2 + c.c1 * t + c.c0;
}

// ---------------------------------------------------------------- adapters
// MediumTrAnalyticAdapterGVDB (VR/VolumeTrackingAdapterGVDB.slang:20-136).
struct TrAnalytic {
    const Medium* m; bool linear; float Tr = 0.f; float tNear = 0, tFar = 0; Ray ray;
    void setRayInfo(float a, float b, const Ray& r) { tNear = a; tFar = b; ray = r; }
    void start() {}
    bool mainStep(const HDDA& dda, f3 vminLeaf, f3 o, f4, int slot, float& t, SampleGen&) {
        HDDA dl = dda;
        dl.prepareLeaf(vminLeaf);
        for (int iter = 0; iter < 128 && leafInside(dl); iter++) {
            g_cellSteps++;
            dl.next();
            float maxDeltaT = dl.t.y - t;
            if (linear) {
                CellPoly c = cellPoly(*m, dl, vminLeaf, o, slot, ray.dir);
                float td = std::min(tFar - t, maxDeltaT);
                float t2 = td * td, t3 = t2 * td, t4 = t2 * t2;
                Tr += -(c.c3 * t4 / 4 + c.c2 * t3 / 3 + c.c1 * t2 / 2 + c.c0 * td);
            } else {
                float dens = densityInAtlas(*m, f3of(dl.p) + o + 0.5f, slot, false);
                float st = dens * m->sigma_t;
                Tr += -std::min(tFar - t, maxDeltaT) * st;
            }
            if (t + maxDeltaT >= tFar) { Tr = std::exp(Tr); return true; }
            t += maxDeltaT;
            dl.step();
        }
        return false;
    }
    void end() { Tr = std::exp(Tr); }
};

// MediumTrRayMarchingAdapterGVDB (VR/VolumeTrackingAdapterGVDB.slang:140-208).
struct TrRayMarch {
    const Medium* m; bool linear; float tStep; float Tr = 0.f; bool init = false; float tNear = 0, tFar = 0; Ray ray;
    void setRayInfo(float a, float b, const Ray& r) { tNear = a; tFar = b; ray = r; }
    void start() { init = true; }
    bool mainStep(const HDDA& dda, f3 vminLeaf, f3 o, f4, int slot, float& t, SampleGen&) {
        t = tNear + (std::floor((t - tNear) / tStep) + 0.5f) * tStep;
        if (t < dda.t.x) t += tStep;
        const float mult = 1.f;
        f3 wp = ray.origin + t * ray.dir;
        f3 p = wp - vminLeaf;
        const f3 wpt = mult * tStep * ray.dir;
        for (int iter = 0; iter < 128 && pInside(p); iter++) {
            if (t >= tFar) { Tr = std::exp(Tr); return true; }
            float dens = densityInAtlas(*m, p + o, slot, linear);
            float st = dens * m->sigma_t;
            Tr += -st * (iter == 0 ? 1.f : mult) * tStep;
            p = p + wpt;
            t += mult * tStep;
        }
        return false;
    }
    void end() { Tr = init ? std::exp(Tr) : 1.f; }
};

// SampleMediumAnalyticAdapterGVDB (VR/VolumeTrackingAdapterGVDB.slang:212-436).
// targets: the point-sampled form also takes pre-drawn optical-depth targets in
// outTr (as the trilinear form does) instead of a fresh uniform per cell; the
// two draw the same distance distribution (the learned proposal's sampler).
struct SampleAnalytic {
    const Medium* m; bool linear; int numSamples; bool targets = false;
    float hit[4], outTr[4], pdf[4];
    float tau = 0.f; bool init = false; float tNear = 0, tFar = 0; Ray ray;
    void setRayInfo(float a, float b, const Ray& r) { tNear = a; tFar = b; ray = r; }
    void start() { for (int i = 0; i < numSamples; i++) hit[i] = -1; init = true; }
    bool mainStep(const HDDA& dda, f3 vminLeaf, f3 o, f4, int slot, float& t, SampleGen& sg) {
        HDDA dl = dda;
        dl.prepareLeaf(vminLeaf);
        for (int iter = 0; iter < 128 && leafInside(dl); iter++) {
            dl.next();
            float maxDeltaT = std::min(tFar - t, dl.t.y - t);
            float curTMax = std::min(tFar, dl.t.y);
            int finished = 0;
            float dTau = 0.f;
            if (linear) {
                CellPoly c = cellPoly(*m, dl, vminLeaf, o, slot, ray.dir);
                dTau = polyTau(maxDeltaT, c);
                for (int s = 0; s < numSamples; s++) {
                    if (hit[s] == -1) {
                        if (tau + dTau >= outTr[s]) {
                            float target = outTr[s] - tau;
                            float tl = 0, th = maxDeltaT, taul = 0, tauh = dTau, tsol = 0;
                            int it = 0;
                            while (it++ < 32 && th - tl > maxDeltaT * 0.001f) {
                                tsol = tl + (th - tl) * (target - taul) / (tauh - taul);
                                float tv = polyTau(tsol, c);
                                if (tv < target) { tl = tsol; taul = tv; } else { th = tsol; tauh = tv; }
                            }
                            hit[s] = t + tsol;
                            outTr[s] = std::exp(-outTr[s]);
                            pdf[s] = polySigma(tsol, c) * outTr[s];
                            finished++;
                        }
                    } else finished++;
                }
            } else if (targets) {
                float dens = densityInAtlas(*m, f3of(dl.p) + o + 0.5f, slot, false);
                float st = dens * m->sigma_t;
                dTau = maxDeltaT * st;
                for (int s = 0; s < numSamples; s++) {
                    if (hit[s] == -1) {
                        if (dTau > 0.f && tau + dTau >= outTr[s]) {
                            hit[s] = t + (outTr[s] - tau) / st;
                            outTr[s] = std::exp(-outTr[s]);
                            pdf[s] = st * outTr[s];
                            finished++;
                        }
                    } else finished++;
                }
            } else {
                float dens = densityInAtlas(*m, f3of(dl.p) + o + 0.5f, slot, false);
                float st = dens * m->sigma_t;
                for (int s = 0; s < numSamples; s++) {
                    if (hit[s] == -1) {
                        float dT = -std::log(1 - next1D(sg)) / st;
                        float curT = t + dT;
                        if (std::isnan(curT) || std::isinf(curT)) curT = kRayTMax;
                        if (curT < curTMax) {
                            hit[s] = curT;
                            outTr[s] = std::exp(-(dT * st + tau));
                            pdf[s] = st * outTr[s];
                            finished++;
                        }
                    } else finished++;
                }
                dTau = maxDeltaT * st;
            }
            if (finished == numSamples) return true;
            t = curTMax;
            tau += dTau;
            if (t >= tFar) { end(); return true; }
            dl.step();
        }
        return false;
    }
    void end() {
        if (init) {
            for (int s = 0; s < numSamples; s++)
                if (hit[s] == -1) { hit[s] = kRayTMax; outTr[s] = std::exp(-tau); pdf[s] = outTr[s]; }
        } else {
            for (int s = 0; s < numSamples; s++) { hit[s] = kRayTMax; outTr[s] = 1.f; pdf[s] = 1.f; }
        }
    }
};

// SampleVolumeCellByDensityAdapterGVDB (VR/VolumeTrackingAdapterGVDB.slang:439-515).
struct CellByDensity {
    const Medium* m; int densityBound = 0; f2 sel{-1, -1}; float tNear = 0, tFar = 0; Ray ray;
    float runningSum = 0.f, Tr = 0.f;
    void setRayInfo(float a, float b, const Ray& r) { tNear = a; tFar = b; ray = r; }
    void start() { sel = {-1, -1}; }
    bool mainStep(const HDDA& dda, f3 vminLeaf, f3 o, f4, int slot, float& t, SampleGen& sg) {
        HDDA dl = dda;
        dl.prepareLeaf(vminLeaf);
        for (int iter = 0; iter < 128 && leafInside(dl); iter++) {
            float dens = densityInAtlas(*m, f3of(dl.p) + o + 0.5f, slot, false);
            dl.next();
            float maxDeltaT = dl.t.y - t;
            float st = dens * m->sigma_t;
            float w = std::exp(Tr) * st;
            runningSum += w;
            if (runningSum > 0.f && next1D(sg) < w / runningSum) {
                densityBound = (int)dens;  // the reference stores it in an int field
                sel = {t, std::min(tFar, t + maxDeltaT)};
            }
            Tr += -maxDeltaT * st;
            if (t + maxDeltaT >= tFar) return true;
            t += maxDeltaT;
            dl.step();
        }
        return false;
    }
    void end() {}
};

// ReservoirFeatureRayMarchingAdapterGVDB (VR/VolumeTrackingAdapterGVDB.slang:518-602).
struct FeatureMarch {
    const Medium* m; bool linear; float tStep; float accuT = 1.f, Tr = 0.f; bool init = false, early = false;
    float tNear = 0, tFar = 0; Ray ray;
    void setRayInfo(float a, float b, const Ray& r) { tNear = a; tFar = b; ray = r; }
    void start() { init = true; }
    bool mainStep(const HDDA& dda, f3 vminLeaf, f3 o, f4, int slot, float& t, SampleGen&) {
        t = tNear + (std::floor((t - tNear) / tStep) + 0.5f) * tStep;
        if (t < dda.t.x) t += tStep;
        f3 wp = ray.origin + t * ray.dir;
        f3 p = wp - vminLeaf;
        const f3 wpt = tStep * ray.dir;
        for (int iter = 0; iter < 128 && pInside(p); iter++) {
            float cur = std::exp(Tr);
            if (cur < 0.01f) { early = true; end(); return true; }
            float dens = densityInAtlas(*m, p + o, slot, linear);
            float st = dens * m->sigma_t;
            float neg = -st * tStep;
            Tr += neg;
            p = p + wpt;
            t += tStep;
        }
        return false;
    }
    void end() { if (init) accuT = std::exp(Tr); }
};

struct MediumInteraction { f3 p, wo; float g; bool valid; };

// DecompositionTrackingGVDBAdapter (VR/VolumeTrackingAdapterGVDB.slang:606-707).
struct Decomposition {
    const Medium* m; Ray rWorld; MediumInteraction mi; float tNear = 0, tFar = 0; Ray ray;
    void setRayInfo(float a, float b, const Ray& r) { tNear = a; tFar = b; ray = r; }
    void start() {}
    bool mainStep(const HDDA& dda, f3 vminLeaf, f3 o, f4 bd, int slot, float& t, SampleGen& sg) {
        float minD = bd.x, maxD = bd.y;
        float curTMax = std::min(tFar, dda.t.y);
        float tControl;
        if (minD == 0.f) tControl = kRayTMax;
        else tControl = t - std::log(1 - next1D(sg)) / (minD * m->sigma_t);
        float invMax = 1.f / (maxD - minD);
        if (maxD - minD > 0.f) {
            while (true) {
                t -= std::log(1 - next1D(sg)) * invMax / m->sigma_t;
                if (t >= tControl || t >= curTMax) break;
                f3 wp = ray.origin + t * ray.dir;
                f3 p = wp - vminLeaf;
                float dens = densityInAtlas(*m, p + o, slot, true);
                if ((dens - minD) * invMax > next1D(sg)) {
                    mi = {rWorld.at(t), -rWorld.dir, m->g, true};
                    return true;
                }
            }
            t = std::min(tControl, t);
            if (t < curTMax) { mi = {rWorld.at(t), -rWorld.dir, m->g, true}; return true; }
            t = curTMax;
            if (t >= tFar) { end(); return true; }
        } else {
            if (tControl < curTMax) { mi = {rWorld.at(tControl), -rWorld.dir, m->g, true}; return true; }
            t = curTMax;
            return false;
        }
        return false;
    }
    void end() { mi.valid = false; }
};

// ResidualRatioTrackingGVDBAdapter (VR/VolumeTrackingAdapterGVDB.slang:710-794).
struct ResidualRatio {
    const Medium* m; float Tr = 1.f; bool analog, global; float tNear = 0, tFar = 0; Ray ray;
    void setRayInfo(float a, float b, const Ray& r) { tNear = a; tFar = b; ray = r; }
    void start() { if (global) analog = true; }
    bool mainStep(const HDDA& dda, f3 vminLeaf, f3 o, f4 bd, int slot, float& t, SampleGen& sg) {
        float muMin = global ? 0.f : bd.x * m->sigma_t;
        float muMax = global ? m->grid[slot].maxValue * m->densityScale * m->sigma_t : bd.y * m->sigma_t;
        float muAvg = bd.z * m->sigma_t;
        float maxDeltaT = std::min(tFar - t, dda.t.y - t);
        float curTMax = std::min(tFar, dda.t.y);
        float muRt = muMax - muMin;
        float D = m->superVoxelDiag;
        float gamma = 2;
        float muC = (muRt == 0.f || analog)
                        ? muMin
                        : std::min(muAvg, std::max(muMin, muMin + muRt * (std::pow(gamma, 1.f / (D * muRt)) - 1)));
        float muR = std::max(muC - muMin, muMax - muC);
        float invMuR = 1.f / muR;
        float Tc = std::exp(-muC * std::min(tFar - t, maxDeltaT));
        float Trr = 1;
        if (muR > 0.f) {
            while (true) {
                t -= std::log(1 - next1D(sg)) * invMuR;
                if (t >= curTMax) break;
                f3 wp = ray.origin + t * ray.dir;
                f3 p = wp - vminLeaf;
                float dens = densityInAtlas(*m, p + o, slot, true);
                float mu = dens * m->sigma_t;
                Trr *= 1 - (mu - muC) * invMuR;
            }
        }
        Tr *= Tc * Trr;
        t = curTMax;
        return t >= tFar;
    }
    void end() {}
};

// ---------------------------------------------------------------- wrappers (VR/VolumeUtils.slang:284-378)
static inline float trAnalytic(const Medium& m, const Ray& r, int slot, SampleGen& sg, bool linear) {
    TrAnalytic a; a.m = &m; a.linear = linear;
    trackGVDB(m, r, slot, sg, a, linear);
    return a.Tr;
}
static inline void sampleAnalytic(const Medium& m, const Ray& r, SampleGen& sg, bool linear, float hit[4], int slot,
                                  float pdf[4], float tr[4], int numSamples) {
    SampleAnalytic a; a.m = &m; a.linear = linear; a.numSamples = numSamples;
    if (linear)
        for (int i = 0; i < numSamples; i++) a.outTr[i] = -std::log(1 - next1D(sg));
    trackGVDB(m, r, slot, sg, a, linear);
    for (int i = 0; i < 4; i++) { hit[i] = a.hit[i]; tr[i] = a.outTr[i]; pdf[i] = a.pdf[i]; }
}
// The same sampler with the optical-depth targets given (point or trilinear form).
static inline void sampleAnalyticTargets(const Medium& m, const Ray& r, SampleGen& sg, bool linear, int slot,
                                         const float targets[4], int numSamples, float hit[4], float pdf[4],
                                         float tr[4]) {
    SampleAnalytic a; a.m = &m; a.linear = linear; a.numSamples = numSamples; a.targets = !linear;
    for (int i = 0; i < numSamples; i++) a.outTr[i] = targets[i];
    trackGVDB(m, r, slot, sg, a, linear);
    for (int i = 0; i < 4; i++) { hit[i] = a.hit[i]; tr[i] = a.outTr[i]; pdf[i] = a.pdf[i]; }
}
static inline void sampleDecomposition(const Medium& m, const Ray& r, SampleGen& sg, MediumInteraction& mi, int slot) {
    Decomposition a; a.m = &m; a.rWorld = r; a.mi = mi;
    trackGVDB(m, r, slot, sg, a, false);
    mi = a.mi;
}
static inline float trResidualRatio(const Medium& m, const Ray& r, int slot, SampleGen& sg, bool analog, bool global) {
    ResidualRatio a; a.m = &m; a.global = global; a.analog = analog || global;
    trackGVDB(m, r, slot, sg, a, false);
    return a.Tr;
}
static inline f2 cellByDensity(const Medium& m, const Ray& r, SampleGen& sg, int slot, float& densityBound) {
    CellByDensity a; a.m = &m;
    trackGVDB(m, r, slot, sg, a, false);
    densityBound = (float)a.densityBound;
    return a.sel;
}
static inline float trRayMarch(const Medium& m, const Ray& r, int slot, bool linear, float tStepScale, SampleGen& sg) {
    int eff = slot >= 19 ? slot - 19 : slot;
    eff = eff >= 8 ? eff - 8 : eff;
    TrRayMarch a; a.m = &m; a.linear = linear;
    a.tStep = m.tStep * m.worldScaling * tStepScale * (float)(eff + 1);
    trackGVDB(m, r, slot, sg, a, false);
    return a.Tr;
}
static inline float featureMarch(const Medium& m, const Ray& r, int slot, bool linear, float tStepScale, SampleGen& sg) {
    FeatureMarch a; a.m = &m; a.linear = linear;
    a.tStep = m.tStep * m.worldScaling * tStepScale;
    trackGVDB(m, r, slot, sg, a, false);
    return a.accuT;
}

// computeVisibility (VR/VolumeUtils.slang:553-571).
static inline float computeVisibility(const Medium& m, const Ray& ray, SampleGen& sg, int samples, int slot, bool linear,
                                      uint32_t method, float tStepScale) {
    float vis = 0;
    for (int i = 0; i < samples; i++) {
        if (method == 0 || method == 3 || method == 4)
            vis += trResidualRatio(m, ray, slot, sg, method == 4, method == 0);
        else if (method == 2)
            vis += trRayMarch(m, ray, slot, linear, tStepScale, sg);
        else if (method == 1)
            vis += trAnalytic(m, ray, slot, sg, linear);
    }
    vis /= samples;
    return vis;
}

// RejectionSampleRandomPointByDensity (VR/VolumeUtils.slang:573-582).
static inline float rejectionSampleDepth(const Medium& m, const Ray& r, SampleGen& sg, int slot) {
    float bound;
    f2 sel = cellByDensity(m, r, sg, slot, bound);
    if (sel.x == -1) return kRayTMax;
    float d = sel.x + (sel.y - sel.x) * next1D(sg);
    (void)(next1D(sg) * bound);
    return d;
}

// ---------------------------------------------------------------- phase function (VR/VolumeBase.slang:50-96)
static inline float phaseHG(float cosTheta, float g) {
    float denom = 1 + g * g + 2 * g * cosTheta;
    const float Inv4Pi = 0.07957747154594766788444188168626f;
    return Inv4Pi * (1 - g * g) / (denom * std::sqrt(denom));
}
static inline void coordinateSystem(f3 v1, f3& v2, f3& v3) {
    if (std::fabs(v1.x) > std::fabs(v1.y)) v2 = mk3(-v1.z, 0, v1.x) / std::sqrt(v1.x * v1.x + v1.z * v1.z);
    else v2 = mk3(0, v1.z, -v1.y) / std::sqrt(v1.y * v1.y + v1.z * v1.z);
    v3 = cross(v1, v2);
}
static inline float samplePhase(float g, f3 wo, f3& wi, f2 u) {
    float cosTheta;
    if (std::fabs(g) < 1e-3f) cosTheta = 1 - 2 * u.x;
    else {
        float sq = (1 - g * g) / (1 + g - 2 * g * u.x);
        cosTheta = -(1 + g * g - sq * sq) / (2 * g);
    }
    float sinTheta = std::sqrt(std::max(0.f, 1 - cosTheta * cosTheta));
    float phi = 2 * kPi * u.y;
    f3 v1, v2;
    coordinateSystem(wo, v1, v2);
    wi = sinTheta * std::cos(phi) * v1 + sinTheta * std::sin(phi) * v2 + cosTheta * wo;
    return phaseHG(cosTheta, g);
}

}  // namespace orc
