// vr_device.hpp -- device-side building blocks of the VolumetricReSTIR pass for gfx950.
//
// Layout in HBM (see vr_types.h): every density grid is a flat brick table plus
// a pool of 10^3-value bricks (one-voxel apron, fp32 or fp16), so a voxel fetch
// is one dependent load of the brick id and one load of the value, and every
// trilinear footprint stays inside one brick. The traversal keeps the
// reference's hierarchical DDA (128-voxel blocks, then 8-voxel bricks, then
// voxels) and its exact stepping arithmetic so the visited cells and the
// t-intervals handed to the per-brick trackers are the reference's.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "vr_types.h"

namespace vr {

#define VR_DEV __device__ __forceinline__

// ------------------------------------------------------------------ vectors
struct v3 { float x, y, z; };
struct v2 { float x, y; };
struct i3v { int x, y, z; };

VR_DEV v3 V3(float a) { return {a, a, a}; }
VR_DEV v3 V3(float a, float b, float c) { return {a, b, c}; }
VR_DEV v3 operator+(v3 a, v3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
VR_DEV v3 operator-(v3 a, v3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
VR_DEV v3 operator*(v3 a, v3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
VR_DEV v3 operator/(v3 a, v3 b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
VR_DEV v3 operator*(v3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
VR_DEV v3 operator*(float s, v3 a) { return {a.x * s, a.y * s, a.z * s}; }
VR_DEV v3 operator/(v3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
VR_DEV v3 operator+(v3 a, float s) { return {a.x + s, a.y + s, a.z + s}; }
VR_DEV v3 operator-(v3 a, float s) { return {a.x - s, a.y - s, a.z - s}; }
VR_DEV v3 operator-(v3 a) { return {-a.x, -a.y, -a.z}; }
VR_DEV float dot3(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
VR_DEV v3 cross3(v3 a, v3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
VR_DEV float len3(v3 a) { return sqrtf(dot3(a, a)); }
VR_DEV v3 norm3(v3 a) { return a / len3(a); }
VR_DEV v3 floor3(v3 a) { return {floorf(a.x), floorf(a.y), floorf(a.z)}; }
VR_DEV v3 abs3(v3 a) { return {fabsf(a.x), fabsf(a.y), fabsf(a.z)}; }
VR_DEV v3 tof(i3v a) { return {(float)a.x, (float)a.y, (float)a.z}; }
VR_DEV bool anyPos(v3 a) { return a.x > 0.f || a.y > 0.f || a.z > 0.f; }
VR_DEV bool allZero(v3 a) { return a.x == 0.f && a.y == 0.f && a.z == 0.f; }
VR_DEV float lum(v3 c) { return c.x * 0.2126f + c.y * 0.7152f + c.z * 0.0722f; }
VR_DEV v3 ld3(const float* p) { return {p[0], p[1], p[2]}; }

static constexpr float kTMax = 3.402823466e+38f;  // FLT_MAX
static constexpr float kPiF = 3.14159265358979323846f;
static constexpr float kInv4PiF = 0.07957747154594766788f;

// ------------------------------------------------------------------ RNG
// SplitMix64-seeded xoshiro128**, seed = (sampleNumber << 32) | Morton(pixel)
// (UT/Sampling/UniformSampleGenerator.slang:49-68).
struct Rng { uint32_t s0, s1, s2, s3; };
VR_DEV uint32_t morton16(uint32_t x, uint32_t y) {
    x &= 0xffffu; y &= 0xffffu;
    x = (x | (x << 8)) & 0x00FF00FFu; x = (x | (x << 4)) & 0x0F0F0F0Fu;
    x = (x | (x << 2)) & 0x33333333u; x = (x | (x << 1)) & 0x55555555u;
    y = (y | (y << 8)) & 0x00FF00FFu; y = (y | (y << 4)) & 0x0F0F0F0Fu;
    y = (y | (y << 2)) & 0x33333333u; y = (y | (y << 1)) & 0x55555555u;
    return x | (y << 1);
}
VR_DEV uint64_t smix(uint64_t& st) {
    uint64_t z = (st += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
VR_DEV Rng rngCreate(uint32_t px, uint32_t py, uint32_t sampleNumber) {
    uint64_t st = (uint64_t)morton16(px, py) | ((uint64_t)sampleNumber << 32);
    uint64_t a = smix(st), b = smix(st);
    return {(uint32_t)a, (uint32_t)(a >> 32), (uint32_t)b, (uint32_t)(b >> 32)};
}
VR_DEV uint32_t rngNext(Rng& g) {
    uint32_t r = g.s0 * 5u;
    r = ((r << 7) | (r >> 25)) * 9u;
    uint32_t t = g.s1 << 9;
    g.s2 ^= g.s0; g.s3 ^= g.s1; g.s1 ^= g.s2; g.s0 ^= g.s3;
    g.s2 ^= t;
    g.s3 = (g.s3 << 11) | (g.s3 >> 21);
    return r;
}
VR_DEV float u01(Rng& g) { return (float)(rngNext(g) >> 8) * 5.9604644775390625e-08f; }

// ------------------------------------------------------------------ volume access
// Pointers into the grid buffers are read from the Scene struct in device
// memory, so the compiler only sees generic pointers; casting them to the
// global address space turns flat loads (which also count against the LDS
// counter and force joint waits) into global loads.
template <class T> using gptr = const __attribute__((address_space(1))) T*;
template <class T> VR_DEV gptr<T> gbl(const void* p) { return (gptr<T>)p; }

// Brick id of brick cell (bx, by, bz), kEmpty outside the table.
VR_DEV uint32_t brickAt(const Grid& G, int bx, int by, int bz) {
    if ((unsigned)bx >= (unsigned)G.nbx || (unsigned)by >= (unsigned)G.nby || (unsigned)bz >= (unsigned)G.nbz)
        return kEmpty;
    return gbl<uint32_t>(G.brickTable)[bx + G.nbx * (by + G.nby * bz)];
}

// ------------------------------------------------------------------ occupancy tables in LDS
// A traversal's empty-brick steps are each a dependent brick-table load. The
// kernels that trace copy the {word, prefix} occupancy tables of up to two grid
// slots into LDS once per workgroup (occCacheLoad, first statement of every
// kernel that traverses), and brickLookup answers from LDS for those slots:
// occupancy is a bit test and the id of an occupied brick its rank
// (prefix + popcount of the lower bits), the id the brick table holds.
extern __shared__ uint2 sOccTab[];        // dynamic LDS: the tables of sOccSlot[0], then sOccSlot[1]
static __shared__ int sOccSlot[2], sOccOff[2];

// slotA / slotB: grid slots to cache (-1: none). The launch must provide
// (occWords(slotA) + occWords(slotB)) * 8 bytes of dynamic LDS.
VR_DEV void occCacheLoad(const Scene& S, int slotA, int slotB) {
    if (slotB == slotA) slotB = -1;
    const int nA = slotA >= 0 ? S.grid[slotA].occWords : 0;
    const int nB = slotB >= 0 ? S.grid[slotB].occWords : 0;
    if (threadIdx.x == 0) {
        sOccSlot[0] = slotA; sOccSlot[1] = slotB;
        sOccOff[0] = 0; sOccOff[1] = nA;
    }
    if (nA > 0) {
        gptr<uint32_t> a = gbl<uint32_t>(S.grid[slotA].occ);
        for (int i = threadIdx.x; i < nA; i += blockDim.x) sOccTab[i] = make_uint2(a[2 * i], a[2 * i + 1]);
    }
    if (nB > 0) {
        gptr<uint32_t> b = gbl<uint32_t>(S.grid[slotB].occ);
        for (int i = threadIdx.x; i < nB; i += blockDim.x) sOccTab[nA + i] = make_uint2(b[2 * i], b[2 * i + 1]);
    }
    __syncthreads();
}

// brickAt for grid `slot`, from LDS when the kernel cached that slot's table.
VR_DEV uint32_t brickLookup(const Grid& G, int slot, int bx, int by, int bz) {
    if ((unsigned)bx >= (unsigned)G.nbx || (unsigned)by >= (unsigned)G.nby || (unsigned)bz >= (unsigned)G.nbz)
        return kEmpty;
    const uint32_t idx = (uint32_t)(bx + G.nbx * (by + G.nby * bz));
    const int k = slot == sOccSlot[0] ? 0 : (slot == sOccSlot[1] ? 1 : -1);
    if (k < 0) return gbl<uint32_t>(G.brickTable)[idx];
    const uint2 e = sOccTab[sOccOff[k] + (idx >> 5)];
    const uint32_t bit = 1u << (idx & 31u);
    if (!(e.x & bit)) return kEmpty;
    const uint32_t id = e.y + (uint32_t)__popc(e.x & (bit - 1u));
    return id < G.numBricks ? id : kEmpty;
}
VR_DEV bool nodeAt(const Grid& G, int nx, int ny, int nz) {
    if ((unsigned)nx >= (unsigned)G.nnx || (unsigned)ny >= (unsigned)G.nny || (unsigned)nz >= (unsigned)G.nnz)
        return false;
    return gbl<uint32_t>(G.nodeTable)[nx + G.nnx * (ny + G.nny * nz)] != 0u;
}
// Stored value at apron index idx in [0,1000) of a brick. The format test is
// uniform (one grid per call), so it costs a scalar branch.
VR_DEV float apronAt(const Grid& G, uint32_t brick, int idx) {
    const size_t o = (size_t)brick * kBrickVals + idx;
    if (G.fp16) return (float)gbl<_Float16>(G.pool)[o];
    return gbl<float>(G.pool)[o];
}
VR_DEV float apronVal(const Grid& G, uint32_t brick, int i, int j, int k) {
    return apronAt(G, brick, i + kApron * (j + kApron * k));
}
// The 2x2x2 footprint whose lowest corner is apron index `base`, in the
// order v000 v100 v010 v110 v001 v101 v011 v111. One format branch for all
// eight loads, so they issue back to back (adjacent x pairs merge into one
// 32-bit load for fp16, 64-bit for fp32).
VR_DEV void footprint(const Grid& G, uint32_t brick, int base, float v[8]) {
    const size_t o = (size_t)brick * kBrickVals + base;
    if (G.fp16) {
        gptr<_Float16> q = gbl<_Float16>(G.pool) + o;
        v[0] = (float)q[0]; v[1] = (float)q[1]; v[2] = (float)q[10]; v[3] = (float)q[11];
        v[4] = (float)q[100]; v[5] = (float)q[101]; v[6] = (float)q[110]; v[7] = (float)q[111];
    } else {
        gptr<float> q = gbl<float>(G.pool) + o;
        v[0] = q[0]; v[1] = q[1]; v[2] = q[10]; v[3] = q[11];
        v[4] = q[100]; v[5] = q[101]; v[6] = q[110]; v[7] = q[111];
    }
}
VR_DEV float trilerp(const float v[8], v3 f) {
    float a0 = v[0] + (v[1] - v[0]) * f.x, a1 = v[2] + (v[3] - v[2]) * f.x;
    float a2 = v[4] + (v[5] - v[4]) * f.x, a3 = v[6] + (v[7] - v[6]) * f.x;
    float b0 = a0 + (a1 - a0) * f.y, b1 = a2 + (a3 - a2) * f.y;
    return b0 + (b1 - b0) * f.z;
}
// Texture-style sampling of a brick at brick-local continuous coordinate q
// (q = p - brickMin, p in index space): point = voxel floor(q); linear =
// trilinear over voxel centres, reading the apron for the +1 neighbours.
VR_DEV float sampleBrick(const Grid& G, uint32_t brick, v3 q, bool linear) {
    if (!linear) {
        int i = (int)floorf(q.x) + 1, j = (int)floorf(q.y) + 1, k = (int)floorf(q.z) + 1;
        if ((unsigned)i >= 10u || (unsigned)j >= 10u || (unsigned)k >= 10u) return 0.f;
        return apronVal(G, brick, i, j, k);
    }
    v3 x = q + 0.5f;
    v3 i0 = floor3(x);
    v3 f = x - i0;
    int ix = (int)i0.x, iy = (int)i0.y, iz = (int)i0.z;
    float v[8];
    if ((unsigned)ix >= 9u || (unsigned)iy >= 9u || (unsigned)iz >= 9u) {
        // only reachable for q outside [0,8]: sample with zero border
#pragma unroll
        for (int c = 0; c < 8; c++) {
            int a = ix + (c & 1), b = iy + ((c >> 1) & 1), d = iz + (c >> 2);
            v[c] = ((unsigned)a >= 10u || (unsigned)b >= 10u || (unsigned)d >= 10u) ? 0.f
                                                                                   : apronVal(G, brick, a, b, d);
        }
        return trilerp(v, f);
    }
    footprint(G, brick, ix + kApron * (iy + kApron * iz), v);
    return trilerp(v, f);
}

struct Ray { v3 o, d; float tMin, tMax; };
VR_DEV v3 at(const Ray& r, float t) { return r.o + r.d * t; }

VR_DEV v3 xformP(const float* M, v3 p) {
    float x = p.x * M[0] + p.y * M[4] + p.z * M[8] + M[12];
    float y = p.x * M[1] + p.y * M[5] + p.z * M[9] + M[13];
    float z = p.x * M[2] + p.y * M[6] + p.z * M[10] + M[14];
    float w = p.x * M[3] + p.y * M[7] + p.z * M[11] + M[15];
    return {x / w, y / w, z / w};
}
VR_DEV v3 xformV(const float* M, v3 v) {
    return {v.x * M[0] + v.y * M[4] + v.z * M[8], v.x * M[1] + v.y * M[5] + v.z * M[9],
            v.x * M[2] + v.y * M[6] + v.z * M[10]};
}

// Density at an index-space point (VR/VolumeBase.slang:234-237 + GVS/gvdb.slang:6-29).
VR_DEV float densityIndex(const Scene& S, int slot, v3 p) {
    const Grid& G = S.grid[slot];
    if (p.x < 0.f || p.y < 0.f || p.z < 0.f || p.x >= G.bmax[0] || p.y >= G.bmax[1] || p.z >= G.bmax[2]) return 0.f;
    int bx = (int)p.x / 8, by = (int)p.y / 8, bz = (int)p.z / 8;
    uint32_t b = brickLookup(G, slot, bx, by, bz);
    if (b == kEmpty) return 0.f;
    v3 q = p - V3((float)(bx * 8), (float)(by * 8), (float)(bz * 8));
    return sampleBrick(G, b, q, true) * S.densityScale;
}
VR_DEV float densityWorld(const Scene& S, int slot, v3 pW) { return densityIndex(S, slot, xformP(S.grid[slot].w2i, pW)); }

VR_DEV bool intersectBounds(const Grid& G, const Ray& r, float& t0o, float& t1o, bool vc) {
    float pmin[3] = {0.f, 0.f, 0.f}, pmax[3] = {G.bmax[0], G.bmax[1], G.bmax[2]};
    if (vc) for (int i = 0; i < 3; i++) { pmin[i] -= 0.5f; pmax[i] -= 0.5f; }
    float o[3] = {r.o.x, r.o.y, r.o.z}, d[3] = {r.d.x, r.d.y, r.d.z};
    float t0 = 0.f, t1 = r.tMax;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        float inv = 1.f / d[i];
        float tn = (pmin[i] - o[i]) * inv, tf = (pmax[i] - o[i]) * inv;
        if (tn > tf) { float tmp = tn; tn = tf; tf = tmp; }
        t0 = tn > t0 ? tn : t0;
        t1 = tf < t1 ? tf : t1;
        if (t0 > t1) return false;
    }
    t0o = t0; t1o = t1;
    return true;
}

// ------------------------------------------------------------------ hierarchical DDA (GVS/gvdbDda.slang:86-157)
struct Dda {
    v3 pos, dir;
    v3 iad;  // |1 / dir|: vdel / dir for a power-of-two cell size vdel is exactly vdel * (1 / dir)
    i3v pStep;
    v3 tDel;
    float tx, ty;
    i3v p;
    v3 tSide;
    i3v mask;
    VR_DEV void setFromRay(v3 o, v3 d, float t0) {
        pos = o; dir = d;
        iad = abs3(V3(1.f) / d);
        pStep = {d.x >= 0 ? 1 : -1, d.y >= 0 ? 1 : -1, d.z >= 0 ? 1 : -1};
        tx = t0;
    }
    // vdel: the level's cell size, a power of two (128 or 8), so tDel = |vdel / dir|
    // and the division by vdel below are exact scalings.
    VR_DEV void prepare(v3 vmin, float vdel) {
        tDel = V3(vdel) * iad;
        v3 pf = (pos + dir * tx - vmin) * V3(1.f / vdel);
        v3 fl = floor3(pf);
        tSide = ((fl - pf + 0.5f) * tof(pStep) + 0.5f) * tDel + tx;
        p = {(int)fl.x, (int)fl.y, (int)fl.z};
    }
    VR_DEV void prepareLeaf(v3 vmin) {
        tDel = iad;
        v3 pf = pos + dir * tx - vmin;
        v3 fl = floor3(pf);
        tSide = ((fl - pf + 0.5f) * tof(pStep) + 0.5f) * tDel + tx;
        p = {(int)fl.x, (int)fl.y, (int)fl.z};
    }
    VR_DEV void next() {
        mask.x = (int)((tSide.x < tSide.y) & (tSide.x <= tSide.z));
        mask.y = (int)((tSide.y < tSide.z) & (tSide.y <= tSide.x));
        mask.z = (int)((tSide.z < tSide.x) & (tSide.z <= tSide.y));
        ty = mask.x ? tSide.x : (mask.y ? tSide.y : tSide.z);
    }
    // Adds tDel on the stepped axis only: float(mask) * tDel would be 0 * inf =
    // NaN on an axis with a zero direction component and stall the walk.
    VR_DEV void step() {
        tx = ty;
        tSide = {mask.x ? tSide.x + tDel.x : tSide.x, mask.y ? tSide.y + tDel.y : tSide.y,
                 mask.z ? tSide.z + tDel.z : tSide.z};
        p = {p.x + mask.x * pStep.x, p.y + mask.y * pStep.y, p.z + mask.z * pStep.z};
    }
};
VR_DEV bool leafInside(const Dda& d) { return (unsigned)d.p.x < 8u && (unsigned)d.p.y < 8u && (unsigned)d.p.z < 8u; }
VR_DEV bool inBrick(v3 p) { return p.x >= 0 && p.y >= 0 && p.z >= 0 && p.x < 8.f && p.y < 8.f && p.z < 8.f; }

// Per-brick context handed to the trackers.
struct BrickCtx {
    uint32_t brick;
    v3 vmin;        // brick minimum in (possibly vertex-shifted) index space
    float bounds[4];  // raw bounds * densityScale
};

// VolumeTrackingGVDB (VR/VolumeUtils.slang:171-282): walks 128^3 blocks, then
// bricks, and hands every occupied brick to the tracker with the DDA state.
// The walk runs at brick level only; it keeps the hierarchical walk's block
// semantics: when tx passes the current block's exit it re-prepares at the
// block level at tx, steps over empty blocks, and prepares the brick level at
// the same tx in the next occupied block (a prepare from the same distance
// gives bit-identical side distances at either level), and it ends where the
// hierarchical walk leaves its top level (tx past tFar at a block border, or
// outside the 32^3-block index range). The trackers use the brick-level DDA
// only through pos, dir, |1/dir|, tx and the brick's minimum.
template <class A>
VR_DEV void traverse(const Scene& S, const Ray& rWorld, int slot, Rng& rng, A& ad, bool vertexCenter) {
    const Grid& G = S.grid[slot];
    const float eps = 0.01f;
    Ray ray{xformP(G.w2i, rWorld.o), xformV(G.w2i, rWorld.d), rWorld.tMin, rWorld.tMax};
    if (vertexCenter) ray.o = ray.o - 0.5f;
    float tNear = 0.f, tFar = 0.f;
    if (!intersectBounds(G, ray, tNear, tFar, vertexCenter)) { ad.end(); return; }
    ad.setRayInfo(tNear, tFar, ray);
    ad.start();
    Dda dda;
    dda.setFromRay(ray.o, ray.d, tNear + eps);
    auto outsideBlocks = [](const Dda& d) {
        return d.p.x < 0 || d.p.y < 0 || d.p.z < 0 || d.p.x > 32 || d.p.y > 32 || d.p.z > 32;
    };
    int iter = 0;
    float tBlock = 0.f;
    // block-level iterations from the prepared block-level DDA nd up to the next
    // occupied block, whose brick level is then prepared at the same tx
    auto enterFrom = [&](Dda nd) -> bool {
        for (;;) {
            if (iter >= 4096 || outsideBlocks(nd)) return false;
            nd.next();
            const int b = (((nd.p.z << 5) + nd.p.y) << 5) + nd.p.x;
            iter++;
            if (b < 32768 && nodeAt(G, b & 31, (b >> 5) & 31, b >> 10)) break;
            nd.step();
            nd.tx += eps;
            if (nd.tx > tFar) return false;
        }
        tBlock = nd.ty;
        dda.tx = nd.tx;
        dda.prepare(V3(0.f), 8.f);
        return true;
    };
    {
        Dda nd = dda;
        nd.prepare(V3(0.f), 128.f);
        if (vertexCenter) {
            int k = 0;
            while (k++ < 3 && outsideBlocks(nd)) { nd.next(); nd.step(); nd.tx += eps; }
        }
        if (!enterFrom(nd)) { ad.end(); return; }
    }
    while (iter < 4096) {
        dda.next();
        const uint32_t brick = brickLookup(G, slot, dda.p.x, dda.p.y, dda.p.z);
        if (brick != kEmpty) {
            float t = dda.tx - eps;
            BrickCtx bc;
            bc.brick = brick;
            bc.vmin = V3((float)(dda.p.x * 8), (float)(dda.p.y * 8), (float)(dda.p.z * 8));
            gptr<float> bd = gbl<float>(G.bounds) + 4 * (size_t)brick;
            bc.bounds[0] = bd[0] * S.densityScale; bc.bounds[1] = bd[1] * S.densityScale;
            bc.bounds[2] = bd[2] * S.densityScale; bc.bounds[3] = bd[3] * S.densityScale;
            if (ad.brickStep(dda, bc, slot, t, rng)) return;
        }
        dda.step();
        dda.tx += eps;
        iter++;
        if (dda.tx > tBlock) {
            if (dda.tx > tFar) break;
            Dda nd = dda;
            nd.prepare(V3(0.f), 128.f);
            if (!enterFrom(nd)) break;
        }
    }
    ad.end();
}

// Trilinear sigma_t polynomial of one cell along the ray
// (VR/VolumeTrackingAdapterGVDB.slang:72-100): 8 corners = voxels p..p+1 of the brick.
struct Poly { float c3, c2, c1, c0; };
VR_DEV Poly cellPoly(const Scene& S, const Grid& G, const Dda& dl, const BrickCtx& bc, v3 d) {
    int base = (dl.p.x + 1) + kApron * ((dl.p.y + 1) + kApron * (dl.p.z + 1));
    float v[8];
    footprint(G, bc.brick, base, v);
    const float ds = S.densityScale, st = S.sigma_t;
    const float v000 = v[0] * ds * st, v100 = v[1] * ds * st, v010 = v[2] * ds * st, v110 = v[3] * ds * st;
    const float v001 = v[4] * ds * st, v101 = v[5] * ds * st, v011 = v[6] * ds * st, v111 = v[7] * ds * st;
    float mxyz = v111 - v011 - v101 - v110 + v100 + v010 + v001 - v000;
    float mxy = v000 - v100 - v010 + v110;
    float mxz = v000 - v100 - v001 + v101;
    float myz = v000 - v010 - v001 + v011;
    float mx = v100 - v000, my = v010 - v000, mz = v001 - v000;
    v3 p0 = dl.pos + dl.tx * dl.dir - (tof(dl.p) + bc.vmin);
    Poly c;
    c.c3 = mxyz * d.x * d.y * d.z;
    c.c2 = (p0.z * d.x * d.y + p0.y * d.x * d.z + p0.x * d.y * d.z) * mxyz + mxy * d.x * d.y + mxz * d.x * d.z +
           myz * d.y * d.z;
    c.c1 = (p0.y * p0.z * d.x + p0.x * p0.z * d.y + p0.x * p0.y * d.z) * mxyz + mx * d.x + my * d.y + mz * d.z +
           (p0.y * d.x + p0.x * d.y) * mxy + (p0.z * d.x + p0.x * d.z) * mxz + (p0.z * d.y + p0.y * d.z) * myz;
    c.c0 = p0.x * p0.y * p0.z * mxyz + p0.x * p0.y * mxy + p0.x * p0.z * mxz + p0.y * p0.z * myz + p0.x * mx +
           p0.y * my + p0.z * mz + v000;
    return c;
}
VR_DEV float polyTau(float t, const Poly& c) {
    float t2 = t * t, t3 = t2 * t, t4 = t2 * t2;
    return c.c3 * t4 / 4 + c.c2 * t3 / 3 + c.c1 * t2 / 2 + c.c0 * t;
}
VR_DEV float polySigma(float t, const Poly& c) {
    float t2 = t * t, t3 = t2 * t;
    return c.c3 * t3 + c.c2 * t2 + c.c1 * t + c.c0;
}
// Point-sampled voxel p of a brick (p in [0,8)^3), value * densityScale.
VR_DEV float voxelDensity(const Scene& S, const Grid& G, const BrickCtx& bc, i3v p) {
    return apronVal(G, bc.brick, p.x + 1, p.y + 1, p.z + 1) * S.densityScale;
}
VR_DEV float brickDensity(const Scene& S, const Grid& G, const BrickCtx& bc, v3 q, bool linear) {
    return sampleBrick(G, bc.brick, q, linear) * S.densityScale;
}

// ------------------------------------------------------------------ trackers
// Analytic transmittance (VR/VolumeTrackingAdapterGVDB.slang:20-136).
struct TrAnalytic {
    const Scene* S; const Grid* G; bool linear; float Tr = 0.f; float tNear = 0, tFar = 0; Ray ray;
    VR_DEV void setRayInfo(float a, float b, const Ray& r) { tNear = a; tFar = b; ray = r; }
    VR_DEV void start() {}
    VR_DEV bool brickStep(const Dda& dda, const BrickCtx& bc, int, float& t, Rng&) {
        Dda dl = dda;
        dl.prepareLeaf(bc.vmin);
        for (int iter = 0; iter < 128 && leafInside(dl); iter++) {
            dl.next();
            float maxDeltaT = dl.ty - t;
            if (linear) {
                Poly c = cellPoly(*S, *G, dl, bc, ray.d);
                float td = fminf(tFar - t, maxDeltaT);
                float t2 = td * td, t3 = t2 * td, t4 = t2 * t2;
                Tr += -(c.c3 * t4 / 4 + c.c2 * t3 / 3 + c.c1 * t2 / 2 + c.c0 * td);
            } else {
                float st = voxelDensity(*S, *G, bc, dl.p) * S->sigma_t;
                Tr += -fminf(tFar - t, maxDeltaT) * st;
            }
            if (t + maxDeltaT >= tFar) { Tr = expf(Tr); return true; }
            t += maxDeltaT;
            dl.step();
        }
        return false;
    }
    VR_DEV void end() { Tr = expf(Tr); }
};

// Ray-marched transmittance (VR/VolumeTrackingAdapterGVDB.slang:140-208).
struct TrMarch {
    const Scene* S; const Grid* G; bool linear; float tStep; float Tr = 0.f; bool init = false;
    float tNear = 0, tFar = 0; Ray ray;
    VR_DEV void setRayInfo(float a, float b, const Ray& r) { tNear = a; tFar = b; ray = r; }
    VR_DEV void start() { init = true; }
    VR_DEV bool brickStep(const Dda& dda, const BrickCtx& bc, int, float& t, Rng&) {
        t = tNear + (floorf((t - tNear) / tStep) + 0.5f) * tStep;
        if (t < dda.tx) t += tStep;
        v3 q = at(ray, t) - bc.vmin;
        const v3 dq = 1.f * tStep * ray.d;
        for (int iter = 0; iter < 128 && inBrick(q); iter++) {
            if (t >= tFar) { Tr = expf(Tr); return true; }
            float st = brickDensity(*S, *G, bc, q, linear) * S->sigma_t;
            Tr += -st * 1.f * tStep;
            q = q + dq;
            t += 1.f * tStep;
        }
        return false;
    }
    VR_DEV void end() { Tr = init ? expf(Tr) : 1.f; }
};

// Up to 4 free-flight samples in one traversal (VR/VolumeTrackingAdapterGVDB.slang:212-436).
// TARGETS: the point-sampled form also takes pre-drawn optical-depth targets in
// tr (as the trilinear form does) instead of a fresh uniform per cell; both draw
// the same distance distribution. The learned proposal drives it this way.
template <bool TARGETS>
struct SampleAnalyticT {
    const Scene* S; const Grid* G; bool linear; int numSamples;
    float hit[4], tr[4], pdf[4];
    float tau = 0.f; bool init = false; float tNear = 0, tFar = 0; Ray ray;
    VR_DEV void setRayInfo(float a, float b, const Ray& r) { tNear = a; tFar = b; ray = r; }
    VR_DEV void start() {
#pragma unroll
        for (int i = 0; i < 4; i++) if (i < numSamples) hit[i] = -1;
        init = true;
    }
    VR_DEV bool brickStep(const Dda& dda, const BrickCtx& bc, int, float& t, Rng& rng) {
        Dda dl = dda;
        dl.prepareLeaf(bc.vmin);
        for (int iter = 0; iter < 128 && leafInside(dl); iter++) {
            dl.next();
            float maxDeltaT = fminf(tFar - t, dl.ty - t);
            float curTMax = fminf(tFar, dl.ty);
            int finished = 0;
            float dTau = 0.f;
            if (linear) {
                Poly c = cellPoly(*S, *G, dl, bc, ray.d);
                dTau = polyTau(maxDeltaT, c);
#pragma unroll
                for (int s = 0; s < 4; s++) {
                    if (s >= numSamples) break;
                    if (hit[s] == -1) {
                        if (tau + dTau >= tr[s]) {
                            float target = tr[s] - tau;
                            float tl = 0, th = maxDeltaT, taul = 0, tauh = dTau, tsol = 0;
                            int it = 0;
                            while (it++ < 32 && th - tl > maxDeltaT * 0.001f) {
                                tsol = tl + (th - tl) * (target - taul) / (tauh - taul);
                                float tv = polyTau(tsol, c);
                                if (tv < target) { tl = tsol; taul = tv; } else { th = tsol; tauh = tv; }
                            }
                            hit[s] = t + tsol;
                            tr[s] = expf(-tr[s]);
                            pdf[s] = polySigma(tsol, c) * tr[s];
                            finished++;
                        }
                    } else finished++;
                }
            } else if (TARGETS) {
                float st = voxelDensity(*S, *G, bc, dl.p) * S->sigma_t;
                dTau = maxDeltaT * st;
#pragma unroll
                for (int s = 0; s < 4; s++) {
                    if (s >= numSamples) break;
                    if (hit[s] == -1) {
                        if (dTau > 0.f && tau + dTau >= tr[s]) {
                            hit[s] = t + (tr[s] - tau) / st;
                            tr[s] = expf(-tr[s]);
                            pdf[s] = st * tr[s];
                            finished++;
                        }
                    } else finished++;
                }
            } else {
                float st = voxelDensity(*S, *G, bc, dl.p) * S->sigma_t;
#pragma unroll
                for (int s = 0; s < 4; s++) {
                    if (s >= numSamples) break;
                    if (hit[s] == -1) {
                        float dT = -logf(1 - u01(rng)) / st;
                        float curT = t + dT;
                        if (isnan(curT) || isinf(curT)) curT = kTMax;
                        if (curT < curTMax) {
                            hit[s] = curT;
                            tr[s] = expf(-(dT * st + tau));
                            pdf[s] = st * tr[s];
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
    VR_DEV void end() {
#pragma unroll
        for (int s = 0; s < 4; s++) {
            if (s >= numSamples) break;
            if (init) {
                if (hit[s] == -1) { hit[s] = kTMax; tr[s] = expf(-tau); pdf[s] = tr[s]; }
            } else { hit[s] = kTMax; tr[s] = 1.f; pdf[s] = 1.f; }
        }
    }
};

// Density-proportional cell selection (VR/VolumeTrackingAdapterGVDB.slang:439-515).
struct CellByDensity {
    const Scene* S; const Grid* G; int densityBound = 0; float selx = -1, sely = -1;
    float runningSum = 0.f, Tr = 0.f; float tNear = 0, tFar = 0; Ray ray;
    VR_DEV void setRayInfo(float a, float b, const Ray& r) { tNear = a; tFar = b; ray = r; }
    VR_DEV void start() { selx = -1; sely = -1; }
    VR_DEV bool brickStep(const Dda& dda, const BrickCtx& bc, int, float& t, Rng& rng) {
        Dda dl = dda;
        dl.prepareLeaf(bc.vmin);
        for (int iter = 0; iter < 128 && leafInside(dl); iter++) {
            float dens = voxelDensity(*S, *G, bc, dl.p);
            dl.next();
            float maxDeltaT = dl.ty - t;
            float st = dens * S->sigma_t;
            float w = expf(Tr) * st;
            runningSum += w;
            if (runningSum > 0.f && u01(rng) < w / runningSum) {
                densityBound = (int)dens;
                selx = t; sely = fminf(tFar, t + maxDeltaT);
            }
            Tr += -maxDeltaT * st;
            if (t + maxDeltaT >= tFar) return true;
            t += maxDeltaT;
            dl.step();
        }
        return false;
    }
    VR_DEV void end() {}
};

// Feature ray march with early termination (VR/VolumeTrackingAdapterGVDB.slang:518-602).
struct FeatureMarch {
    const Scene* S; const Grid* G; bool linear; float tStep; float accuT = 1.f, Tr = 0.f; bool init = false;
    float tNear = 0, tFar = 0; Ray ray;
    VR_DEV void setRayInfo(float a, float b, const Ray& r) { tNear = a; tFar = b; ray = r; }
    VR_DEV void start() { init = true; }
    VR_DEV bool brickStep(const Dda& dda, const BrickCtx& bc, int, float& t, Rng&) {
        t = tNear + (floorf((t - tNear) / tStep) + 0.5f) * tStep;
        if (t < dda.tx) t += tStep;
        v3 q = at(ray, t) - bc.vmin;
        const v3 dq = tStep * ray.d;
        for (int iter = 0; iter < 128 && inBrick(q); iter++) {
            float cur = expf(Tr);
            if (cur < 0.01f) { end(); return true; }
            float st = brickDensity(*S, *G, bc, q, linear) * S->sigma_t;
            Tr += -st * tStep;
            q = q + dq;
            t += tStep;
        }
        return false;
    }
    VR_DEV void end() { if (init) accuT = expf(Tr); }
};

struct MediumInteraction { v3 p, wo; float g; bool valid; };

// Delta tracking against per-brick bounds (VR/VolumeTrackingAdapterGVDB.slang:606-707).
struct Decomposition {
    const Scene* S; const Grid* G; Ray rWorld; MediumInteraction mi; float tNear = 0, tFar = 0; Ray ray;
    VR_DEV void setRayInfo(float a, float b, const Ray& r) { tNear = a; tFar = b; ray = r; }
    VR_DEV void start() {}
    VR_DEV bool brickStep(const Dda& dda, const BrickCtx& bc, int, float& t, Rng& rng) {
        float minD = bc.bounds[0], maxD = bc.bounds[1];
        float curTMax = fminf(tFar, dda.ty);
        float tControl;
        if (minD == 0.f) tControl = kTMax;
        else tControl = t - logf(1 - u01(rng)) / (minD * S->sigma_t);
        float invMax = 1.f / (maxD - minD);
        if (maxD - minD > 0.f) {
            while (true) {
                t -= logf(1 - u01(rng)) * invMax / S->sigma_t;
                if (t >= tControl || t >= curTMax) break;
                v3 q = at(ray, t) - bc.vmin;
                float dens = brickDensity(*S, *G, bc, q, true);
                if ((dens - minD) * invMax > u01(rng)) {
                    mi = {at(rWorld, t), -rWorld.d, S->g, true};
                    return true;
                }
            }
            t = fminf(tControl, t);
            if (t < curTMax) { mi = {at(rWorld, t), -rWorld.d, S->g, true}; return true; }
            t = curTMax;
            if (t >= tFar) { end(); return true; }
        } else {
            if (tControl < curTMax) { mi = {at(rWorld, tControl), -rWorld.d, S->g, true}; return true; }
            t = curTMax;
            return false;
        }
        return false;
    }
    VR_DEV void end() { mi.valid = false; }
};

// (Residual) ratio tracking (VR/VolumeTrackingAdapterGVDB.slang:710-794).
struct ResidualRatio {
    const Scene* S; const Grid* G; float Tr = 1.f; bool analog, global; float tNear = 0, tFar = 0; Ray ray;
    VR_DEV void setRayInfo(float a, float b, const Ray& r) { tNear = a; tFar = b; ray = r; }
    VR_DEV void start() { if (global) analog = true; }
    VR_DEV bool brickStep(const Dda& dda, const BrickCtx& bc, int, float& t, Rng& rng) {
        float muMin = global ? 0.f : bc.bounds[0] * S->sigma_t;
        float muMax = global ? G->maxValue * S->densityScale * S->sigma_t : bc.bounds[1] * S->sigma_t;
        float muAvg = bc.bounds[2] * S->sigma_t;
        float maxDeltaT = fminf(tFar - t, dda.ty - t);
        float curTMax = fminf(tFar, dda.ty);
        float muRt = muMax - muMin;
        float D = S->superVoxelDiag;
        float muC = (muRt == 0.f || analog)
                        ? muMin
                        : fminf(muAvg, fmaxf(muMin, muMin + muRt * (powf(2.f, 1.f / (D * muRt)) - 1)));
        float muR = fmaxf(muC - muMin, muMax - muC);
        float invMuR = 1.f / muR;
        float Tc = expf(-muC * fminf(tFar - t, maxDeltaT));
        float Trr = 1;
        if (muR > 0.f) {
            while (true) {
                t -= logf(1 - u01(rng)) * invMuR;
                if (t >= curTMax) break;
                v3 q = at(ray, t) - bc.vmin;
                float mu = brickDensity(*S, *G, bc, q, true) * S->sigma_t;
                Trr *= 1 - (mu - muC) * invMuR;
            }
        }
        Tr *= Tc * Trr;
        t = curTMax;
        return t >= tFar;
    }
    VR_DEV void end() {}
};

// ------------------------------------------------------------------ tracker entry points (VR/VolumeUtils.slang:284-378)
VR_DEV float trAnalytic(const Scene& S, const Ray& r, int slot, Rng& rng, bool linear) {
    TrAnalytic a; a.S = &S; a.G = &S.grid[slot]; a.linear = linear;
    traverse(S, r, slot, rng, a, linear);
    return a.Tr;
}
VR_DEV float trMarch(const Scene& S, const Ray& r, int slot, bool linear, float tStepScale, Rng& rng) {
    int eff = slot >= 19 ? slot - 19 : slot;
    eff = eff >= 8 ? eff - 8 : eff;
    TrMarch a; a.S = &S; a.G = &S.grid[slot]; a.linear = linear;
    a.tStep = S.tStep * S.worldScaling * tStepScale * (float)(eff + 1);
    traverse(S, r, slot, rng, a, false);
    return a.Tr;
}
VR_DEV float trResidual(const Scene& S, const Ray& r, int slot, Rng& rng, bool analog, bool global) {
    ResidualRatio a; a.S = &S; a.G = &S.grid[slot]; a.global = global; a.analog = analog || global;
    traverse(S, r, slot, rng, a, false);
    return a.Tr;
}
using SampleAnalytic = SampleAnalyticT<false>;
// The same with the optical-depth targets given (point or trilinear form).
VR_DEV void sampleAnalyticTargets(const Scene& S, const Ray& r, Rng& rng, bool linear, int slot, int ns,
                                  const float targets[4], float hit[4], float pdf[4], float tr[4]) {
    SampleAnalyticT<true> a; a.S = &S; a.G = &S.grid[slot]; a.linear = linear; a.numSamples = ns;
#pragma unroll
    for (int i = 0; i < 4; i++)
        if (i < ns) a.tr[i] = targets[i];
    traverse(S, r, slot, rng, a, linear);
#pragma unroll
    for (int i = 0; i < 4; i++) { hit[i] = a.hit[i]; pdf[i] = a.pdf[i]; tr[i] = a.tr[i]; }
}
VR_DEV void sampleAnalytic(const Scene& S, const Ray& r, Rng& rng, bool linear, int slot, int ns, float hit[4],
                           float pdf[4], float tr[4]) {
    SampleAnalytic a; a.S = &S; a.G = &S.grid[slot]; a.linear = linear; a.numSamples = ns;
    if (linear) {
#pragma unroll
        for (int i = 0; i < 4; i++)
            if (i < ns) a.tr[i] = -logf(1 - u01(rng));
    }
    traverse(S, r, slot, rng, a, linear);
#pragma unroll
    for (int i = 0; i < 4; i++) { hit[i] = a.hit[i]; pdf[i] = a.pdf[i]; tr[i] = a.tr[i]; }
}
VR_DEV void sampleDecomposition(const Scene& S, const Ray& r, Rng& rng, MediumInteraction& mi, int slot) {
    Decomposition a; a.S = &S; a.G = &S.grid[slot]; a.rWorld = r; a.mi = mi;
    traverse(S, r, slot, rng, a, false);
    mi = a.mi;
}
VR_DEV void cellByDensity(const Scene& S, const Ray& r, Rng& rng, int slot, float& sx, float& sy, float& bound) {
    CellByDensity a; a.S = &S; a.G = &S.grid[slot];
    traverse(S, r, slot, rng, a, false);
    sx = a.selx; sy = a.sely; bound = (float)a.densityBound;
}
VR_DEV float featureMarch(const Scene& S, const Ray& r, int slot, bool linear, float tStepScale, Rng& rng) {
    FeatureMarch a; a.S = &S; a.G = &S.grid[slot]; a.linear = linear;
    a.tStep = S.tStep * S.worldScaling * tStepScale;
    traverse(S, r, slot, rng, a, false);
    return a.accuT;
}
// computeVisibility (VR/VolumeUtils.slang:553-571). M fixes the tracking method
// at compile time (0..4) so a kernel carries only the tracker it uses; M = -1
// dispatches on `method` at run time.
template <int M>
VR_DEV float visibilityT(const Scene& S, const Ray& r, Rng& rng, int samples, int slot, bool linear, uint32_t method,
                         float tStepScale) {
    float vis = 0.f;
    for (int i = 0; i < samples; i++) {
        if constexpr (M < 0) {
            if (method == 0u || method == 3u || method == 4u) vis += trResidual(S, r, slot, rng, method == 4u, method == 0u);
            else if (method == 2u) vis += trMarch(S, r, slot, linear, tStepScale, rng);
            else if (method == 1u) vis += trAnalytic(S, r, slot, rng, linear);
        } else if constexpr (M == 0 || M == 3 || M == 4) {
            vis += trResidual(S, r, slot, rng, M == 4, M == 0);
        } else if constexpr (M == 2) {
            vis += trMarch(S, r, slot, linear, tStepScale, rng);
        } else {
            vis += trAnalytic(S, r, slot, rng, linear);
        }
    }
    return vis / (float)samples;
}
VR_DEV float visibility(const Scene& S, const Ray& r, Rng& rng, int samples, int slot, bool linear, uint32_t method,
                        float tStepScale) {
    return visibilityT<-1>(S, r, rng, samples, slot, linear, method, tStepScale);
}
// RejectionSampleRandomPointByDensity (VR/VolumeUtils.slang:573-582).
VR_DEV float rejectionSampleDepth(const Scene& S, const Ray& r, Rng& rng, int slot) {
    float sx, sy, bound;
    cellByDensity(S, r, rng, slot, sx, sy, bound);
    if (sx == -1) return kTMax;
    float d = sx + (sy - sx) * u01(rng);
    (void)u01(rng);
    return d;
}

// ------------------------------------------------------------------ phase function (VR/VolumeBase.slang:50-96)
VR_DEV float phaseHG(float cosTheta, float g) {
    float denom = 1 + g * g + 2 * g * cosTheta;
    return 0.07957747154594766788f * (1 - g * g) / (denom * sqrtf(denom));
}
VR_DEV float samplePhase(float g, v3 wo, v3& wi, float u0, float u1) {
    float cosTheta;
    if (fabsf(g) < 1e-3f) cosTheta = 1 - 2 * u0;
    else {
        float sq = (1 - g * g) / (1 + g - 2 * g * u0);
        cosTheta = -(1 + g * g - sq * sq) / (2 * g);
    }
    float sinTheta = sqrtf(fmaxf(0.f, 1 - cosTheta * cosTheta));
    float phi = 2 * kPiF * u1;
    v3 v1, v2;
    if (fabsf(wo.x) > fabsf(wo.y)) v1 = V3(-wo.z, 0, wo.x) / sqrtf(wo.x * wo.x + wo.z * wo.z);
    else v1 = V3(0, wo.z, -wo.y) / sqrtf(wo.y * wo.y + wo.z * wo.z);
    v2 = cross3(wo, v1);
    wi = sinTheta * cosf(phi) * v1 + sinTheta * sinf(phi) * v2 + cosTheta * wo;
    return phaseHG(cosTheta, g);
}

}  // namespace vr
