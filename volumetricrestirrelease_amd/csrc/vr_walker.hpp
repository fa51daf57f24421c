// vr_walker.hpp -- resumable transmittance traversal for the balanced trace
// of recorded rays (vr_k_queue.hip).
//
// A walker holds one ray's hierarchical-DDA and tracker state; advance() does
// one unit of work (one DDA cell, or one ray-march sample / analytic cell
// inside an occupied brick). The arithmetic is traverse<TrMarch> /
// traverse<TrAnalytic> of vr_device.hpp step for step, so a walked ray's
// transmittance is the direct form's.
#pragma once
#include "vr_device.hpp"

namespace vr {

// ------------------------------------------------------------------ resumable traversal
// traverse<TrMarch> (TK = 2) and traverse<TrAnalytic> (TK = 1) of vr_device.hpp
// with the loop state held in the struct: VolumeTrackingGVDB
// (VR/VolumeUtils.slang:171-282) over the adapters of
// VR/VolumeTrackingAdapterGVDB.slang:20-208. The state is kept lean (the DDA's
// axis mask and exit distance are recomputed rather than stored) so many
// waves fit a SIMD.
template <int TK>
struct Walker {
    Dda dda;        // pos / dir: the index-space ray (vertex-shifted for linear analytic tracking)
    float tNear, tFar, tMax1, tMax2, tStep, Tr, t;
    int lev, iter, it, node;  // node: current 128^3 block, x | y << 8 | z << 16
    int slot;
    uint32_t brick;
    bool linear, active, inBk;  // inBk: inside an occupied brick
    v3 q;                       // ray marching: brick-local sample point
    v3 vmin, lDel, lSide;       // analytic: brick minimum and the cell walk inside the brick
    i3v lp;
    float ltx;

    VR_DEV bool outside(int res) const {
        return dda.p.x < 0 || dda.p.y < 0 || dda.p.z < 0 || dda.p.x > res || dda.p.y > res || dda.p.z > res;
    }
    VR_DEV void levelUp() {
        while (lev <= 2 && dda.tx > (lev == 1 ? tMax1 : tMax2)) {
            lev++;
            if (lev <= 2) dda.prepare(V3(0.f), 128.f);
        }
    }
    VR_DEV Dda leaf() const {
        Dda d;
        d.pos = dda.pos; d.dir = dda.dir;
        d.tDel = lDel; d.tSide = lSide; d.p = lp; d.tx = ltx;
        return d;
    }
    VR_DEV void setLeaf(const Dda& d) { lDel = d.tDel; lSide = d.tSide; lp = d.p; ltx = d.tx; }

    // Starts a traversal of world-space ray rWorld through grid G (slot `slot`).
    // When the ray misses the grid the walk is complete on return (active ==
    // false, Tr = 1, what end() gives for an adapter that never started).
    VR_DEV void begin(const Scene& S, const Grid& G, int slot_, const Ray& rWorld, bool linear_, float tStepScale) {
        linear = linear_;
        slot = slot_;
        const bool vc = TK == 1 && linear;
        const float eps = 0.01f;
        Ray ray{xformP(G.w2i, rWorld.o), xformV(G.w2i, rWorld.d), rWorld.tMin, rWorld.tMax};
        if (vc) ray.o = ray.o - 0.5f;
        if constexpr (TK == 2) {
            int eff = slot_ >= 19 ? slot_ - 19 : slot_;
            eff = eff >= 8 ? eff - 8 : eff;
            tStep = S.tStep * S.worldScaling * tStepScale * (float)(eff + 1);
        }
        Tr = 0.f;
        inBk = false;
        if (!intersectBounds(G, ray, tNear, tFar, vc)) {
            Tr = 1.f;
            active = false;
            return;
        }
        lev = 2;
        tMax2 = tFar;
        tMax1 = 0.f;
        node = 0;
        dda.setFromRay(ray.o, ray.d, tNear + eps);
        dda.prepare(V3(0.f), 128.f);
        if (vc) {
            int k = 0;
            while (k++ < 3 && outside(32)) { dda.next(); dda.step(); dda.tx += eps; }
        }
        iter = 0;
        active = true;
    }
    VR_DEV void finish() {
        Tr = expf(Tr);
        active = false;
    }
    // After the brick: the step the DDA would have taken right after entering it
    // (its axis mask is recomputed from the unchanged side distances).
    VR_DEV void leaveBrick() {
        const float eps = 0.01f;
        inBk = false;
        dda.next();
        dda.step();
        dda.tx += eps;
        levelUp();
        iter++;
    }
    // One hierarchical-DDA cell of a ray-marching walk (TK = 2) whose bricks are
    // marched elsewhere (kTraceSegments). When the cell is an occupied brick it
    // returns true with the brick and the march's first sample (t0 and the
    // brick-local point q0, as advance() would start it) and walks on past the
    // brick, as advance() does after the march leaves it. A march started at
    // one brick never depends on another, and a walk's samples at t >= tFar add
    // nothing (the walk finishes at the first such sample inside a brick), so the
    // walk ends once its cells lie beyond tFar: a cell entered at tx - eps puts
    // its first sample after that entry (the snapped start moves up by tStep when
    // it falls before tx), so no sample of it can precede tFar once tx - eps >= tFar.
    // (tx itself may exceed tFar while samples of the cell still precede it: eps
    // is a world-space length, larger than tStep at fine mips.)
    // nodeBits: the grid's block occupancy as a bit mask (in LDS), or null to read the block table.
    VR_DEV bool ddaSegment(const Grid& G, const uint32_t* nodeBits, uint32_t& bk, float& t0, v3& q0) {
        const float eps = 0.01f;
        if (!(iter < 4096 && lev > 0 && lev <= 2 && !outside(lev == 2 ? 32 : 16)) || dda.tx - eps >= tFar) {
            active = false;
            return false;
        }
        dda.next();
        bool occupied;
        int cx, cy, cz;
        if (lev == 2) {
            int b = (((dda.p.z << 5) + dda.p.y) << 5) + dda.p.x;
            if (nodeBits) {
                const int nx = b & 31, ny = (b >> 5) & 31, nz = b >> 10;
                const bool in = b < 32768 && (unsigned)nx < (unsigned)G.nnx && (unsigned)ny < (unsigned)G.nny &&
                                (unsigned)nz < (unsigned)G.nnz;
                const int i = in ? nx + G.nnx * (ny + G.nny * nz) : 0;
                occupied = in && ((nodeBits[i >> 5] >> (i & 31)) & 1u);
            } else {
                occupied = b < 32768 && nodeAt(G, b & 31, (b >> 5) & 31, b >> 10);
            }
            cx = b & 31; cy = (b >> 5) & 31; cz = b >> 10;
        } else {
            int b = (((dda.p.z << 4) + dda.p.y) << 4) + dda.p.x;
            cx = (node & 255) * 16 + (b & 15);
            cy = ((node >> 8) & 255) * 16 + ((b >> 4) & 15);
            cz = (node >> 16) * 16 + (b >> 8);
            occupied = b < 4096;
        }
        bk = kEmpty;
        if (occupied && lev == 1) {
            bk = brickLookup(G, slot, cx, cy, cz);
            occupied = bk != kEmpty;
        }
        const bool seg = occupied && lev == 1;
        if (seg) {
            const v3 bmin = V3((float)(cx * 8), (float)(cy * 8), (float)(cz * 8));
            float t = dda.tx - eps;
            t = tNear + (floorf((t - tNear) / tStep) + 0.5f) * tStep;
            if (t < dda.tx) t += tStep;
            t0 = t;
            q0 = at(Ray{dda.pos, dda.dir, 0.f, 0.f}, t) - bmin;
        }
        if (occupied && lev == 2) {
            lev = 1;
            node = cx | (cy << 8) | (cz << 16);
            tMax1 = dda.ty;
            dda.prepare(V3((float)(cx * 128), (float)(cy * 128), (float)(cz * 128)), 8.f);
        } else {
            // an empty cell, or past the brick: advance()'s leaveBrick() repeats next(), which
            // recomputes the same axis mask from the unchanged side distances
            dda.step();
            dda.tx += eps;
        }
        levelUp();
        iter++;
        return seg;
    }
    // One hierarchical-DDA cell; when it enters an occupied brick, also the
    // brick's first tracker step. G must be the grid begin() was given.
    VR_DEV void advance(const Scene& S, const Grid& G) {
        const float eps = 0.01f;
        if (!inBk) {
            if (!(iter < 4096 && lev > 0 && lev <= 2 && !outside(lev == 2 ? 32 : 16))) {
                finish();
                return;
            }
            dda.next();
            bool occupied;
            int cx, cy, cz;
            if (lev == 2) {
                int b = (((dda.p.z << 5) + dda.p.y) << 5) + dda.p.x;
                occupied = b < 32768 && nodeAt(G, b & 31, (b >> 5) & 31, b >> 10);
                cx = b & 31; cy = (b >> 5) & 31; cz = b >> 10;
            } else {
                int b = (((dda.p.z << 4) + dda.p.y) << 4) + dda.p.x;
                cx = (node & 255) * 16 + (b & 15);
                cy = ((node >> 8) & 255) * 16 + ((b >> 4) & 15);
                cz = (node >> 16) * 16 + (b >> 8);
                occupied = b < 4096;
            }
            uint32_t bk = kEmpty;
            if (occupied && lev == 1) {
                bk = brickLookup(G, slot, cx, cy, cz);
                occupied = bk != kEmpty;
            }
            if (occupied && lev == 1) {
                brick = bk;
                const v3 bmin = V3((float)(cx * 8), (float)(cy * 8), (float)(cz * 8));
                t = dda.tx - eps;
                if constexpr (TK == 2) {
                    t = tNear + (floorf((t - tNear) / tStep) + 0.5f) * tStep;
                    if (t < dda.tx) t += tStep;
                    q = at(Ray{dda.pos, dda.dir, 0.f, 0.f}, t) - bmin;
                } else {
                    vmin = bmin;
                    Dda dl = dda;
                    dl.prepareLeaf(vmin);
                    setLeaf(dl);
                }
                it = 0;
                inBk = true;
            } else {
                if (occupied) {
                    lev = 1;
                    node = cx | (cy << 8) | (cz << 16);
                    tMax1 = dda.ty;
                    dda.prepare(V3((float)(cx * 128), (float)(cy * 128), (float)(cz * 128)), 8.f);
                } else {
                    dda.step();
                    dda.tx += eps;
                }
                levelUp();
                iter++;
                return;
            }
        }
        BrickCtx bc;
        bc.brick = brick;
        if constexpr (TK == 2) {
            // TrMarch::brickStep, one sample
            if (!(it < 128 && inBrick(q))) { leaveBrick(); return; }
            if (t >= tFar) { finish(); return; }
            float st = sampleBrick(G, brick, q, linear) * S.densityScale * S.sigma_t;
            Tr += -st * 1.f * tStep;
            const v3 dq = 1.f * tStep * dda.dir;
            q = q + dq;
            t += 1.f * tStep;
            it++;
        } else {
            // TrAnalytic::brickStep, one cell
            Dda dl = leaf();
            if (!(it < 128 && leafInside(dl))) { leaveBrick(); return; }
            bc.vmin = vmin;
            dl.next();
            float maxDeltaT = dl.ty - t;
            if (linear) {
                Poly c = cellPoly(S, G, dl, bc, dda.dir);
                float td = fminf(tFar - t, maxDeltaT);
                float t2 = td * td, t3 = t2 * td, t4 = t2 * t2;
                Tr += -(c.c3 * t4 / 4 + c.c2 * t3 / 3 + c.c1 * t2 / 2 + c.c0 * td);
            } else {
                float st = voxelDensity(S, G, bc, dl.p) * S.sigma_t;
                Tr += -fminf(tFar - t, maxDeltaT) * st;
            }
            if (t + maxDeltaT >= tFar) { finish(); return; }
            t += maxDeltaT;
            dl.step();
            setLeaf(dl);
            it++;
        }
    }
};

// Starts / advances walker w on the grid of its ray kind. When both kinds use
// the same grid (uniform flag `oneSlot`), the grid descriptor is read once per
// wave through scalar loads instead of per lane.
// ------------------------------------------------------------------ brick-level walk
// Walker<2>::ddaSegment without block-level iterations, for kTraceSegments.
// Inside a 128^3 block the hierarchical walk is a brick-level DDA; at the
// block's exit time it re-prepares at the block level, steps over empty blocks
// and, in the next occupied block, prepares the brick level again at the same
// distance tx (the block's entry + 0.01). FlatWalk does exactly that, but as one
// batched operation per block entry (enterBlock) instead of block-level loop
// iterations that every lane of a wave would run in lock step. The re-prepare
// matters: tx lies 0.01 (a world-space length, up to 1.3 voxels at the fine
// mips) past the block border, so the hierarchical walk starts the block at the
// brick containing tx and skips a corner brick the ray clips for less than
// that, and it leaves a block once tx passes the block's exit time; a brick
// walk that ignored block borders differed from it on 3 % of the rays of a
// dense 256^3 grid. A prepare from the same distance gives bit-identical side
// distances at either level (the block origin shifts the floor and its
// argument by the same exact amount), so the brick-level state equals the
// hierarchical walk's; over empty blocks the block-level side distances
// accumulate as there.

// The walk grid's brick occupancy, with its sizes held in registers.
struct FlatGrid {
    int nbx, nby, nbz, nnx, nny, nnz;
    uint32_t numBricks;
    const uint2* occ;  // the grid's LDS occupancy table (occCacheLoad), or null: read brickTable
    const uint32_t* brickTable;
    VR_DEV void init(const Grid& G, int slot) {
        nbx = G.nbx; nby = G.nby; nbz = G.nbz;
        nnx = G.nnx; nny = G.nny; nnz = G.nnz;
        numBricks = G.numBricks;
        brickTable = G.brickTable;
        const int k = slot == sOccSlot[0] ? 0 : (slot == sOccSlot[1] ? 1 : -1);
        occ = k < 0 ? nullptr : sOccTab + sOccOff[k];
    }
    // brickLookup's answer
    VR_DEV uint32_t lookup(int bx, int by, int bz) const {
        if ((unsigned)bx >= (unsigned)nbx || (unsigned)by >= (unsigned)nby || (unsigned)bz >= (unsigned)nbz)
            return kEmpty;
        const uint32_t idx = (uint32_t)(bx + nbx * (by + nby * bz));
        if (!occ) return gbl<uint32_t>(brickTable)[idx];
        const uint2 e = occ[idx >> 5];
        const uint32_t bit = 1u << (idx & 31u);
        if (!(e.x & bit)) return kEmpty;
        const uint32_t id = e.y + (uint32_t)__popc(e.x & (bit - 1u));
        return id < numBricks ? id : kEmpty;
    }
};

struct FlatWalk {
    Dda dda;
    float tNear, tFar, tStep, invStep;
    float tBlock;      // exit distance of the current 128^3 block
    int iter;
    bool active;
    bool blockEntry;   // tx passed tBlock: enterBlock() runs before the next step
    // Walker<TK>::begin for a walk of grid G (slot `slot`) and the entry of its first
    // block: vc (analytic tracking with the trilinear filter) shifts the ray by half
    // a voxel and takes the hierarchical walk's up to three block-level steps while
    // outside the block range. A ray that misses the grid is complete (Tr = 1).
    VR_DEV void begin(const Scene& S, const Grid& G, int slot, const Ray& rWorld, float tStepScale,
                      const FlatGrid& g, const uint32_t* nodeBits, bool vc = false) {
        const float eps = 0.01f;
        Ray ray{xformP(G.w2i, rWorld.o), xformV(G.w2i, rWorld.d), rWorld.tMin, rWorld.tMax};
        if (vc) ray.o = ray.o - 0.5f;
        int eff = slot >= 19 ? slot - 19 : slot;
        eff = eff >= 8 ? eff - 8 : eff;
        tStep = S.tStep * S.worldScaling * tStepScale * (float)(eff + 1);
        invStep = 1.f / tStep;
        if (!intersectBounds(G, ray, tNear, tFar, vc)) {
            active = false;
            return;
        }
        dda.setFromRay(ray.o, ray.d, tNear + eps);
        iter = 0;
        active = true;
        blockEntry = false;
        Dda nd = dda;
        nd.prepare(V3(0.f), 128.f);
        if (vc) {
            int k = 0;
            while (k++ < 3 && (nd.p.x < 0 || nd.p.y < 0 || nd.p.z < 0 || nd.p.x > 32 || nd.p.y > 32 || nd.p.z > 32)) {
                nd.next();
                nd.step();
                nd.tx += eps;
            }
        }
        enterBlockFrom(nd, g, nodeBits);
    }
    // The hierarchical walk's level-up at tx and its block-level iterations up to
    // the next occupied block, whose brick level it prepares at the same tx (the
    // walk ends when tx passes tFar, leaves the 32^3-block index range or runs out
    // of iterations). nodeBits: the grid's block bits (in LDS).
    VR_DEV void enterBlock(const FlatGrid& g, const uint32_t* nodeBits) {
        blockEntry = false;
        if (dda.tx > tFar) {
            active = false;
            return;
        }
        Dda nd = dda;
        nd.prepare(V3(0.f), 128.f);
        enterBlockFrom(nd, g, nodeBits);
    }
    // enterBlock's block-level iterations from the prepared block-level DDA nd
    VR_DEV void enterBlockFrom(Dda nd, const FlatGrid& g, const uint32_t* nodeBits) {
        const float eps = 0.01f;
        for (;;) {
            if (iter >= 4096 || nd.p.x < 0 || nd.p.y < 0 || nd.p.z < 0 || nd.p.x > 32 || nd.p.y > 32 ||
                nd.p.z > 32) {
                active = false;
                return;
            }
            nd.next();
            const int b = (((nd.p.z << 5) + nd.p.y) << 5) + nd.p.x;
            const int nx = b & 31, ny = (b >> 5) & 31, nz = b >> 10;
            const bool in = b < 32768 && (unsigned)nx < (unsigned)g.nnx && (unsigned)ny < (unsigned)g.nny &&
                            (unsigned)nz < (unsigned)g.nnz;
            const int i = in ? nx + g.nnx * (ny + g.nny * nz) : 0;
            const bool occupied = in && ((nodeBits[i >> 5] >> (i & 31)) & 1u);
            iter++;
            if (occupied) break;
            nd.step();
            nd.tx += eps;
            if (nd.tx > tFar) {
                active = false;
                return;
            }
        }
        tBlock = nd.ty;
        dda.tx = nd.tx;
        dda.prepare(V3(0.f), 8.f);
    }
    // One brick cell of the current block; true with the brick and the march's
    // first sample (t0, q0) when the cell holds an occupied brick. The walk ends
    // once its cells lie beyond tFar (the test of Walker::ddaSegment).
    VR_DEV bool step(const FlatGrid& g, uint32_t& bk, float& t0, v3& q0) {
        const float eps = 0.01f;
        if (!(iter < 4096) || dda.tx - eps >= tFar) {
            active = false;
            return false;
        }
        dda.next();
        const int cx = dda.p.x, cy = dda.p.y, cz = dda.p.z;
        bk = g.lookup(cx, cy, cz);
        const bool seg = bk != kEmpty;
        if (seg) {
            const v3 bmin = V3((float)(cx * 8), (float)(cy * 8), (float)(cz * 8));
            float t = dda.tx - eps;
            // (t - tNear) / tStep through the reciprocal: the product is within 3 ulp
            // of the quotient, so only a product that close to an integer can floor
            // differently, and those take the division
            const float x = t - tNear;
            float qt = x * invStep;
            if (fabsf(qt - rintf(qt)) <= fabsf(qt) * 4.8e-7f) qt = x / tStep;
            t = tNear + (floorf(qt) + 0.5f) * tStep;
            if (t < dda.tx) t += tStep;
            t0 = t;
            q0 = at(Ray{dda.pos, dda.dir, 0.f, 0.f}, t) - bmin;
        }
        dda.step();
        dda.tx += eps;
        iter++;
        blockEntry = dda.tx > tBlock;
        return seg;
    }
    // One brick cell of an analytic-tracking walk: true with the brick, the
    // distance tx its cell walk starts from and its brick coordinates when the
    // cell holds an occupied brick. The tracker integrates each occupied brick
    // cell by cell until a cell reaches tFar, where the traversal ends, and walks
    // on past tFar through empty cells; so the walk ends after the first occupied
    // brick whose exit lies at or past tFar and whose cell walk has a cell (or at
    // a block border past tFar). The cell walk starts at the voxel containing tx
    // = entry + 0.01, which lies past a brick the ray clips for less than that:
    // such a brick adds nothing and the tracker goes on to the next one, where
    // its first cell, already past tFar, adds the clamped term and ends it.
    VR_DEV bool stepCells(const FlatGrid& g, uint32_t& bk, float& txOut, i3v& cell) {
        const float eps = 0.01f;
        if (!(iter < 4096)) {
            active = false;
            return false;
        }
        dda.next();
        cell = dda.p;
        bk = g.lookup(dda.p.x, dda.p.y, dda.p.z);
        const bool seg = bk != kEmpty;
        txOut = dda.tx;
        bool last = false;
        if (seg && dda.ty >= tFar) {
            // Dda::prepareLeaf's first cell, inside the brick?
            const v3 pf = dda.pos + dda.dir * dda.tx - V3((float)(cell.x * 8), (float)(cell.y * 8), (float)(cell.z * 8));
            const v3 fl = floor3(pf);
            const i3v lp = {(int)fl.x, (int)fl.y, (int)fl.z};
            last = (unsigned)lp.x < 8u && (unsigned)lp.y < 8u && (unsigned)lp.z < 8u;
        }
        dda.step();
        dda.tx += eps;
        iter++;
        blockEntry = dda.tx > tBlock;
        if (last) active = false;
        return seg;
    }
};

template <int TK>
VR_DEV void walkBegin(Walker<TK>& w, const Scene& S, bool oneSlot, int uniformSlot, int laneSlot, const Ray& r,
                      bool linear, float tss) {
    if (oneSlot) w.begin(S, S.grid[uniformSlot], uniformSlot, r, linear, tss);
    else w.begin(S, S.grid[laneSlot], laneSlot, r, linear, tss);
}
template <int TK>
VR_DEV void walkAdvance(Walker<TK>& w, const Scene& S, bool oneSlot, int uniformSlot) {
    if (oneSlot) w.advance(S, S.grid[uniformSlot]);
    else w.advance(S, S.grid[w.slot]);
}

}  // namespace vr
