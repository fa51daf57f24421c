// vr_k_queue.hip -- balanced trace of recorded transmittance rays.
//
// The record pass of a reuse kernel leaves, per pixel, the rays its target
// function evaluations need (TraceRec, in evaluation order). Here each wave
// takes 64 consecutive pixels and treats their rays as one queue: a lane whose
// ray is finished takes the next ray of the queue, so the wave runs until the
// queue is empty instead of paying, for every evaluation, the longest ray among
// its 64 pixels. Idle lanes are refilled together once a quarter of the wave is
// idle, so the set-up code runs at a useful width. The transmittance of each
// ray is written at its record's index for the replay pass.
#include "vr_stages.hpp"
#include "vr_walker.hpp"

namespace vr {

static constexpr int kRefillLanes = 16;

template <int TK>
__global__ __launch_bounds__(256) void kTraceQueue(const Scene* __restrict__ sc, const Frame fr, Options opt,
                                                   int isFinal, const TraceRec* __restrict__ recs,
                                                   const int* __restrict__ counts, float* __restrict__ trs, int cap,
                                                   int numPixels) {
    const Scene& S = *sc;
    occCacheLoad(S, fr.occSlot[0], fr.occSlot[1]);
    __shared__ int pre[4][65];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int p0 = (blockIdx.x * 4 + wv) * 64;
    const int cnt = p0 + lane < numPixels ? counts[p0 + lane] : 0;
    int incl = cnt;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(incl, o);
        if (lane >= o) incl += y;
    }
    pre[wv][lane + 1] = incl;
    if (lane == 0) pre[wv][0] = 0;
    __syncthreads();
    const int total = pre[wv][64];

    const int vs = opt.visibilityMipLevel, ls = isFinal ? 0 : opt.lightingMipLevel;
    const bool oneSlot = vs == ls;
    Walker<TK> w;
    w.active = false;
    bool busy = false;
    int tri = 0, head = 0;  // tri: the ray's transmittance slot
    for (;;) {
        if (busy) {
            walkAdvance(w, S, oneSlot, vs);
            if (!w.active) { trs[tri] = w.Tr; busy = false; }
        }
        const uint64_t idle = __ballot(!busy);
        const int nIdle = __popcll(idle);
        if (head >= total) {
            if (nIdle == 64) break;
            continue;
        }
        if (nIdle < kRefillLanes && nIdle < 64) continue;
        if (!busy) {
            const int r = head + __popcll(idle & ((1ull << lane) - 1ull));
            if (r < total) {
                int lo = 0, hi = 64;  // the pixel q with pre[q] <= r < pre[q + 1]
                while (hi - lo > 1) {
                    const int mid = (lo + hi) >> 1;
                    if (pre[wv][mid] <= r) lo = mid; else hi = mid;
                }
                const int rec = (p0 + lo) * cap + (r - pre[wv][lo]);
                tri = (r - pre[wv][lo]) * numPixels + p0 + lo;
                const TraceRec R = recs[rec];
                const int kind = R.kind;
                const int slot = kind == 0 ? vs : ls;
                const bool linear = kind == 0 ? opt.visibilityUseLinearSampler != 0 : opt.lightingUseLinearSampler != 0;
                const float tss = kind == 0 ? opt.visibilityTStepScale : opt.lightingTStepScale;
                walkBegin(w, S, oneSlot, vs, slot, Ray{V3(R.ox, R.oy, R.oz), V3(R.dx, R.dy, R.dz), 0.f, R.tMax},
                          linear, tss);
                if (w.active) busy = true;
                else trs[tri] = w.Tr;
            }
        }
        head += nIdle;
    }
}

// ------------------------------------------------------------------ brick-segment trace
// The same queue of recorded rays, traced in two alternating phases so that the
// lanes of a wave run one code path at a time:
//   * walk: every lane steps the DDA of its own ray, one cell per iteration (at
//     brick level only when the grid's block bits fit LDS, Walker::flatSegment,
//     else the hierarchical walk), and pushes each occupied brick it enters (with
//     the march's first sample) to a per-wave segment queue in LDS; a lane whose
//     walk is over takes the next ray of the queue;
//   * march: lanes take segments from the queue, a new one as soon as theirs is
//     done, and ray-march them (TrMarch's brick loop); each finished segment adds
//     its optical-depth sum to its ray's slot with an LDS float atomic. While
//     walks remain, the march's tail (at most kCarryLanes lanes busy and the
//     queue empty) goes back to the queue for the next march.
// After each march phase the rays whose walk ended and that have no pending
// segment are complete: Tr = exp(sum).
// A ray waits in one of its lane's two slots for the march that completes it;
// a lane with both slots waiting pauses its walks until then.
static constexpr int kRaySlots = 128;           // two per lane
static constexpr int kSegCap = 128;             // segment queue entries per wave
static constexpr int kSegHigh = kSegCap - 64;   // the walk phase stops above this fill (<= 64 pushes per step)
static constexpr int kWalkSteps = 32;           // walk-phase iterations before marching
static constexpr int kNodeWords = 64;           // block-occupancy bits kept in LDS (2048 blocks)
static constexpr int kRayRefill = 16;           // lanes that must want a ray before a refill round
static constexpr int kCarryLanes = 32;          // march lanes left when the rest of the march is carried over
static constexpr int kBlockBatch = 16;          // lanes at a block border before their block entries run
static constexpr int kCellsPerStep = 1;         // empty cells a walking lane may cross per walk iteration

struct SegQueue {
    int rec[kRaySlots];  // the ray's transmittance slot
    float sum[kRaySlots];
    float dq[3][kRaySlots];
    float tStep[kRaySlots], tFar[kRaySlots];
    int gs[kRaySlots];                          // grid slot | linear << 8
    int pend[kRaySlots];                        // 1: the ray has a carried-over segment
    int owner[kSegCap];                         // ray slot | samples already taken << 8
    uint32_t brick[kSegCap];
    float t[kSegCap], q[3][kSegCap];
};

VR_DEV void waveSync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// The footprint of a sample at a point q inside the brick ([0, 8)^3), so every
// cell of it lies in the 10^3 apron brick: sampleBrick's loads without its range
// checks, addressed by a 32-bit byte offset from the pool's base (the launch
// checks that the pool is below 4 GiB). The stored values stay unconverted until
// footprintValue, so the loads can be in flight across other work.
template <int FP16, int LIN>
struct Footprint {
    // fp16 + linear: the four x-pairs as 32-bit words (low half = lower x), left packed
    // so that no unpacking waits for the loads; otherwise the values as stored
    using E = typename std::conditional<FP16 != 0, _Float16, float>::type;
    using W = typename std::conditional<FP16 != 0 && LIN != 0, uint32_t, E>::type;
    W v[LIN ? (FP16 ? 4 : 8) : 1];
    v3 f;  // trilinear weights (LIN)
};
typedef uint32_t __attribute__((aligned(2))) u32pair;  // two fp16 values at a 2-byte-aligned address
// valid == false loads the first footprint of brick 0 instead (the pool holds at
// least one brick), so every lane issues the same loads and the wait for them is
// counted exactly.
template <int FP16, int LIN>
VR_DEV void fetchInside(gptr<char> pool, bool valid, uint32_t brick, v3 q, Footprint<FP16, LIN>& fp) {
    using E = typename Footprint<FP16, LIN>::E;
    constexpr uint32_t es = sizeof(E);
    if constexpr (LIN == 0) {
        const uint32_t i = (uint32_t)((int)floorf(q.x) + 1), j = (uint32_t)((int)floorf(q.y) + 1),
                       k = (uint32_t)((int)floorf(q.z) + 1);
        const uint32_t o = valid ? (brick * (uint32_t)kBrickVals + i + kApron * (j + kApron * k)) * es : 0u;
        fp.v[0] = *(gptr<E>)(pool + o);
    } else {
        v3 x = q + 0.5f;
        v3 i0 = floor3(x);
        fp.f = x - i0;
        const uint32_t ix = (uint32_t)(int)i0.x, iy = (uint32_t)(int)i0.y, iz = (uint32_t)(int)i0.z;
        const uint32_t o = valid ? (brick * (uint32_t)kBrickVals + ix + kApron * (iy + kApron * iz)) * es : 0u;
        const gptr<char> b = pool + o;
        if constexpr (FP16 != 0) {
            fp.v[0] = *(gptr<u32pair>)(b);
            fp.v[1] = *(gptr<u32pair>)(b + 10 * es);
            fp.v[2] = *(gptr<u32pair>)(b + 100 * es);
            fp.v[3] = *(gptr<u32pair>)(b + 110 * es);
        } else {
            gptr<E> p = (gptr<E>)b;
            fp.v[0] = p[0]; fp.v[1] = p[1]; fp.v[2] = p[10]; fp.v[3] = p[11];
            fp.v[4] = p[100]; fp.v[5] = p[101]; fp.v[6] = p[110]; fp.v[7] = p[111];
        }
    }
}
VR_DEV float halfLo(uint32_t w) { return (float)__builtin_bit_cast(_Float16, (uint16_t)(w & 0xffffu)); }
VR_DEV float halfHi(uint32_t w) { return (float)__builtin_bit_cast(_Float16, (uint16_t)(w >> 16)); }
// sampleBrick's value from a fetched footprint
template <int FP16, int LIN>
VR_DEV float footprintValue(const Footprint<FP16, LIN>& fp) {
    if constexpr (LIN == 0) {
        return (float)fp.v[0];
    } else {
        float v[8];
        if constexpr (FP16 != 0) {
#pragma unroll
            for (int c = 0; c < 4; c++) {
                v[2 * c] = halfLo(fp.v[c]);
                v[2 * c + 1] = halfHi(fp.v[c]);
            }
        } else {
#pragma unroll
            for (int c = 0; c < 8; c++) v[c] = fp.v[c];
        }
        return trilerp(v, fp.f);
    }
}

// Drains the wave's segment queue when every segment uses grid Gu with storage
// format FP16 and filter LIN, software-pipelined: each iteration a lane fetches
// the footprint of its segment's next sample (TrMarch's brick loop: up to 128
// samples while inside the brick and before tFar) and steps past it, then adds
// the sample it fetched the iteration before to its running sum, so a sample's
// loads are in flight across a whole iteration. A lane whose segment has no next
// sample takes the next segment of the queue in the same iteration. Samples of
// one ray slot that follow each other in a lane add up in order; the sum goes to
// the slot with an LDS float atomic when the lane moves to another slot.
template <int FP16, int LIN>
VR_DEV void marchPipelined(const Scene& S, const Grid& Gu, SegQueue& Q, int nSeg, uint64_t below) {
    const gptr<char> pool = gbl<char>(Gu.pool);
    const float dScale = S.densityScale, sigT = S.sigma_t;
    int mHead = 0;  // wave-uniform
    // issue side: the segment whose next sample is fetched
    bool iHas = false;
    int iOwner = 0, iIt = 0;
    uint32_t iBrick = 0;
    float iT = 0.f, iTStep = 0.f, iTFar = 0.f;
    v3 iQ = V3(0.f), iDq = V3(0.f);
    // the sample fetched in the previous iteration (its footprint alternates
    // between fpA and fpB, so a fetch never waits to be copied)
    bool fHas = false;
    int fOwner = 0;
    float fStep = 0.f;
    Footprint<FP16, LIN> fpA{}, fpB{};
    // compute side: the running sum of ray slot cOwner
    int cOwner = -1;
    float part = 0.f;
    // one iteration: fetch the next sample into nxt, add the one in cur; false when all is done
    auto iteration = [&](const Footprint<FP16, LIN>& cur, Footprint<FP16, LIN>& nxt) -> bool {
        if (iHas && !(iIt < 128 && inBrick(iQ) && iT < iTFar)) iHas = false;
        const uint64_t nm = __ballot(!iHas);
        if (nm && mHead < nSeg) {
            if (!iHas) {
                const int e = mHead + __popcll(nm & below);
                if (e < nSeg) {
                    const int ow = Q.owner[e];
                    iOwner = ow & 255;
                    iIt = ow >> 8;
                    iBrick = Q.brick[e];
                    iT = Q.t[e];
                    iQ = V3(Q.q[0][e], Q.q[1][e], Q.q[2][e]);
                    iDq = V3(Q.dq[0][iOwner], Q.dq[1][iOwner], Q.dq[2][iOwner]);
                    iTStep = Q.tStep[iOwner];
                    iTFar = Q.tFar[iOwner];
                    // a segment without a sample adds nothing
                    iHas = iIt < 128 && inBrick(iQ) && iT < iTFar;
                }
            }
            mHead += __popcll(nm);
        }
        // fetch the issue side's next sample and step past it
        const bool nHas = iHas;
        const int nOwner = iOwner;
        const float nStep = iTStep;
        fetchInside<FP16, LIN>(pool, iHas, iBrick, iQ, nxt);
        if (iHas) {
            iQ = iQ + iDq;
            iT += 1.f * iTStep;
            iIt++;
        }
        // add the sample fetched last iteration; its value is formed in every lane (from
        // stale registers where nothing was fetched) so that the wait for cur's loads is
        // not inside a branch: after it no path leaves a load pending on cur's registers
        const float st = footprintValue<FP16, LIN>(cur) * dScale * sigT;
        if (fHas && fOwner != cOwner) {
            if (cOwner >= 0) atomicAdd(&Q.sum[cOwner], part);
            cOwner = fOwner;
            part = 0.f;
        }
        part = fHas ? part + -st * 1.f * fStep : part;
        fHas = nHas;
        fOwner = nOwner;
        fStep = nStep;
        return __ballot(fHas || iHas) != 0ull || mHead < nSeg;
    };
    for (;;) {
        if (!iteration(fpA, fpB)) break;
        if (!iteration(fpB, fpA)) break;
    }
    if (cOwner >= 0) atomicAdd(&Q.sum[cOwner], part);
}

// Drains the wave's segment queue: each lane marches a segment (TrMarch's brick
// loop: up to 128 samples while inside the brick and before tFar) and takes the
// next one as soon as it is done; the segment's sum goes to its ray's slot.
// When the queue is empty and at most kCarryLanes lanes are still marching (and
// mayCarry holds), those lanes add their partial sums, write their segments back
// to the front of the queue, flag their rays as pending, and the phase ends; the
// count of such segments is returned. Each segment taken advances at least one
// sample before it can be carried, so every segment ends.
// Each segment names its grid slot and filter (any storage format).
VR_DEV int marchPhase(const Scene& S, SegQueue& Q, int nSeg, uint64_t below, bool mayCarry) {
    const float dScale = S.densityScale, sigT = S.sigma_t;
    int mHead = 0;  // wave-uniform
    bool has = false;
    int owner = 0, it = 0, gs = 0;
    uint32_t brick = 0;
    float t = 0.f, part = 0.f, tStep = 0.f, tFar = 0.f;
    v3 q = V3(0.f), dq = V3(0.f);
    for (;;) {
        const uint64_t nm = __ballot(!has);
        if (nm && mHead < nSeg) {
            if (!has) {
                const int e = mHead + __popcll(nm & below);
                if (e < nSeg) {
                    const int ow = Q.owner[e];
                    owner = ow & 255;
                    it = ow >> 8;
                    brick = Q.brick[e];
                    t = Q.t[e];
                    q = V3(Q.q[0][e], Q.q[1][e], Q.q[2][e]);
                    dq = V3(Q.dq[0][owner], Q.dq[1][owner], Q.dq[2][owner]);
                    tStep = Q.tStep[owner];
                    tFar = Q.tFar[owner];
                    gs = Q.gs[owner];
                    part = 0.f;
                    has = true;
                }
            }
            mHead += __popcll(nm);
        }
        const uint64_t hm = __ballot(has);
        if (!hm) break;
        if (has) {
            const bool done = !(it < 128 && inBrick(q)) || t >= tFar;
            if (!done) {
                const float st = sampleBrick(S.grid[gs & 255], brick, q, (gs & 256) != 0) * dScale * sigT;
                part += -st * 1.f * tStep;
                q = q + dq;
                t += 1.f * tStep;
                it++;
            } else {
                atomicAdd(&Q.sum[owner], part);
                has = false;
            }
        }
        if (mayCarry && mHead >= nSeg && __popcll(hm) <= kCarryLanes) {
            const uint64_t cm = __ballot(has);
            if (has) {
                atomicAdd(&Q.sum[owner], part);
                const int e = __popcll(cm & below);
                Q.owner[e] = owner | (it << 8);
                Q.brick[e] = brick;
                Q.t[e] = t;
                Q.q[0][e] = q.x; Q.q[1][e] = q.y; Q.q[2][e] = q.z;
                Q.pend[owner] = 1;
            }
            return __popcll(cm);
        }
    }
    return 0;
}

// ------------------------------------------------------------------ analytic segments
// The segment trace for analytic tracking (TrAnalytic): a segment is an occupied
// brick and the distance tx its cell walk starts from; the march integrates the
// brick cell by cell (the trilinear cell polynomial or the voxel's constant
// density) up to the cell that reaches tFar.
struct CellQueue {
    int rec[kRaySlots];  // the ray's transmittance slot
    float sum[kRaySlots];
    float pos[3][kRaySlots], dir[3][kRaySlots];  // the index-space ray (vertex-shifted when trilinear)
    float tFar[kRaySlots];
    int pend[kRaySlots];                         // unused: the analytic march carries nothing over
    int owner[kSegCap];                          // ray slot
    uint32_t brick[kSegCap];
    float tx[kSegCap];
    int cell[kSegCap];                           // brick coordinates, x | y << 10 | z << 20
};

// Drains a wave's analytic segment queue: each lane integrates a segment's cells
// (TrAnalytic::brickStep from the segment's tx: the leaf DDA prepared there, t =
// tx - 0.01, at most 128 cells while inside the brick, ending at the cell that
// reaches tFar) and takes the next segment as soon as it is done; the segment's
// sum goes to its ray's slot with an LDS float atomic. LIN: trilinear cells.
template <bool LIN>
VR_DEV void marchCells(const Scene& S, const Grid& G, CellQueue& Q, int nSeg, uint64_t below) {
    const float eps = 0.01f;
    int mHead = 0;  // wave-uniform
    bool has = false;
    int owner = 0, it = 0;
    BrickCtx bc;
    bc.brick = 0;
    bc.vmin = V3(0.f);
    float t = 0.f, tFar = 0.f, part = 0.f;
    Dda dl;
    for (;;) {
        const uint64_t nm = __ballot(!has);
        if (nm && mHead < nSeg) {
            if (!has) {
                const int e = mHead + __popcll(nm & below);
                if (e < nSeg) {
                    owner = Q.owner[e];
                    bc.brick = Q.brick[e];
                    const float tx = Q.tx[e];
                    const int c = Q.cell[e];
                    bc.vmin = V3((float)((c & 1023) * 8), (float)(((c >> 10) & 1023) * 8), (float)((c >> 20) * 8));
                    dl.setFromRay(V3(Q.pos[0][owner], Q.pos[1][owner], Q.pos[2][owner]),
                                  V3(Q.dir[0][owner], Q.dir[1][owner], Q.dir[2][owner]), tx);
                    dl.prepareLeaf(bc.vmin);
                    tFar = Q.tFar[owner];
                    t = tx - eps;
                    it = 0;
                    part = 0.f;
                    has = true;
                }
            }
            mHead += __popcll(nm);
        }
        if (!__ballot(has)) break;
        if (has) {
            bool done = !(it < 128 && leafInside(dl));
            if (!done) {
                dl.next();
                const float maxDeltaT = dl.ty - t;
                if constexpr (LIN) {
                    const Poly c = cellPoly(S, G, dl, bc, dl.dir);
                    const float td = fminf(tFar - t, maxDeltaT);
                    const float t2 = td * td, t3 = t2 * td, t4 = t2 * t2;
                    part += -(c.c3 * t4 / 4 + c.c2 * t3 / 3 + c.c1 * t2 / 2 + c.c0 * td);
                } else {
                    const float st = voxelDensity(S, G, bc, dl.p) * S.sigma_t;
                    part += -fminf(tFar - t, maxDeltaT) * st;
                }
                if (t + maxDeltaT >= tFar) {
                    done = true;
                } else {
                    t += maxDeltaT;
                    dl.step();
                    it++;
                }
            }
            if (done) {
                atomicAdd(&Q.sum[owner], part);
                has = false;
            }
        }
    }
}

// The two walk forms of kTraceSegments behind one interface: the brick-level
// FlatWalk when the walk grid is uniform and its block bits fit LDS, else the
// hierarchical Walker<2>.
struct WalkCtx {
    const Scene* S;
    int vs;                   // the walk grid slot when oneSlot
    bool oneSlot;
    const uint32_t* nb;       // block bits in LDS, or null
    FlatGrid g;               // FlatWalk's grid view
    bool cells;               // analytic tracking (the ray is vertex-shifted when trilinear)
};
VR_DEV void segBegin(Walker<2>& w, const WalkCtx& c, int slot, const Ray& r, bool linear, float tss) {
    walkBegin(w, *c.S, c.oneSlot, c.vs, slot, r, linear, tss);
}
VR_DEV void segBegin(FlatWalk& w, const WalkCtx& c, int, const Ray& r, bool linear, float tss) {
    w.begin(*c.S, c.S->grid[c.vs], c.vs, r, tss, c.g, c.nb, c.cells && linear);
}
VR_DEV float segMissTr(const Walker<2>& w) { return w.Tr; }
VR_DEV float segMissTr(const FlatWalk&) { return 1.f; }
VR_DEV bool segStep(Walker<2>& w, const WalkCtx& c, uint32_t& bk, float& t0, v3& q0) {
    if (c.oneSlot) return w.ddaSegment(c.S->grid[c.vs], c.nb, bk, t0, q0);
    return w.ddaSegment(c.S->grid[w.slot], nullptr, bk, t0, q0);
}
VR_DEV bool segStep(FlatWalk& w, const WalkCtx& c, uint32_t& bk, float& t0, v3& q0) {
    return w.step(c.g, bk, t0, q0);
}
// A walked ray's slot data, by queue kind
template <class W>
VR_DEV void storeRay(SegQueue& Q, int rs, const W& w, int slot, bool linear) {
    const v3 dq = 1.f * w.tStep * w.dda.dir;
    Q.dq[0][rs] = dq.x; Q.dq[1][rs] = dq.y; Q.dq[2][rs] = dq.z;
    Q.tStep[rs] = w.tStep;
    Q.tFar[rs] = w.tFar;
    Q.gs[rs] = slot | (linear ? 256 : 0);
}
template <class W>
VR_DEV void storeRay(CellQueue& Q, int rs, const W& w, int, bool) {
    Q.pos[0][rs] = w.dda.pos.x; Q.pos[1][rs] = w.dda.pos.y; Q.pos[2][rs] = w.dda.pos.z;
    Q.dir[0][rs] = w.dda.dir.x; Q.dir[1][rs] = w.dda.dir.y; Q.dir[2][rs] = w.dda.dir.z;
    Q.tFar[rs] = w.tFar;
}

// FlatWalk's pending block entries (the hierarchical walker has none)
VR_DEV bool segBlockEntry(const Walker<2>&) { return false; }
VR_DEV bool segBlockEntry(const FlatWalk& w) { return w.blockEntry; }
VR_DEV void segEnterBlock(Walker<2>&, const WalkCtx&) {}
VR_DEV void segEnterBlock(FlatWalk& w, const WalkCtx& c) { w.enterBlock(c.g, c.nb); }

// TK = 2: ray marching (SegQueue, brick-sample march); TK = 1: analytic tracking
// (CellQueue, cell integration; launched only on a grid whose block bits fit LDS,
// one grid and one filter for both ray kinds).
template <int TK>
VR_DEV void traceSegmentsBody(const Scene* __restrict__ sc, const Frame& fr, const Options& opt, int isFinal,
                              const TraceRec* __restrict__ recs, const int* __restrict__ counts,
                              float* __restrict__ trs, int cap, int numPixels, int blockX) {
    const Scene& S = *sc;
    occCacheLoad(S, fr.occSlot[0], fr.occSlot[1]);
    __shared__ int pre[4][65];
    using Queue = std::conditional_t<TK == 2, SegQueue, CellQueue>;
    __shared__ Queue queues[4];
    __shared__ uint32_t nodeBits[kNodeWords];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const uint64_t below = (1ull << lane) - 1ull;
    const int p0 = (blockX * 4 + wv) * 64;
    const int cnt = p0 + lane < numPixels ? counts[p0 + lane] : 0;
    int incl = cnt;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(incl, o);
        if (lane >= o) incl += y;
    }
    pre[wv][lane + 1] = incl;
    if (lane == 0) pre[wv][0] = 0;
    __syncthreads();
    const int total = pre[wv][64];
    Queue& Q = queues[wv];

    const int vs = opt.visibilityMipLevel, ls = isFinal ? 0 : opt.lightingMipLevel;
    const bool oneSlot = vs == ls;
    const bool uniformLinear = opt.visibilityUseLinearSampler != 0;
    // one grid, format and filter for every segment; the grid has bricks and byte offsets
    // into its pool fit 32 bits
    const bool uniformMarch = oneSlot && uniformLinear == (opt.lightingUseLinearSampler != 0) &&
                              S.grid[vs].numBricks > 0 &&
                              (uint64_t)S.grid[vs].numBricks * kBrickVals * (S.grid[vs].fp16 ? 2u : 4u) <=
                                  0xffffffffull;
    // the walk grid's block occupancy as bits in LDS, when it fits
    const Grid& Gw = S.grid[vs];
    const int nNodes = Gw.nnx * Gw.nny * Gw.nnz;
    const bool nodeLds = oneSlot && nNodes <= kNodeWords * 32;
    if (nodeLds) {
        for (int i = threadIdx.x; i < kNodeWords; i += blockDim.x) {
            uint32_t bits = 0;
            for (int b = 0; b < 32; b++) {
                const int n = i * 32 + b;
                if (n < nNodes && gbl<uint32_t>(Gw.nodeTable)[n] != 0u) bits |= 1u << b;
            }
            nodeBits[i] = bits;
        }
    }
    __syncthreads();
    const uint32_t* nb = nodeLds ? nodeBits : nullptr;
    // with the block bits in LDS the walk iterates over bricks only (FlatWalk)
    const bool flat = nodeLds;
    WalkCtx wc{&S, vs, oneSlot, nb, {}, TK == 1};
    if (flat) wc.g.init(S.grid[vs], vs);
    auto trace = [&](auto& w) {
        w.active = false;
        int bank = 0;      // the current ray's slot is 2 * lane + bank
        int waitMask = 0;  // bit b: slot 2 * lane + b holds a walked ray awaiting its march
        int head = 0;      // next ray of the wave's queue (wave-uniform)
        int nSeg = 0;      // segments in the queue (wave-uniform)
        Q.pend[2 * lane] = 0;
        Q.pend[2 * lane + 1] = 0;
        for (;;) {
            // ---- walk phase
            for (int step = 0; step < kWalkSteps && nSeg <= kSegHigh; step++) {
                const bool want = !w.active && !((waitMask >> bank) & 1) && head < total;
                const uint64_t wm = __ballot(want);
                // refill in rounds of at least kRayRefill lanes (or when no walk is left), so the
                // wave waits for record loads less often
                if (wm && (__popcll(wm) >= kRayRefill || !__ballot(w.active))) {
                    if (want) {
                        const int r = head + __popcll(wm & below);
                        if (r < total) {
                            int lo = 0, hi = 64;  // the pixel q with pre[q] <= r < pre[q + 1]
                            while (hi - lo > 1) {
                                const int mid = (lo + hi) >> 1;
                                if (pre[wv][mid] <= r) lo = mid; else hi = mid;
                            }
                            const int rec = (p0 + lo) * cap + (r - pre[wv][lo]);
                            const int tri = (r - pre[wv][lo]) * numPixels + p0 + lo;
                            const TraceRec R = recs[rec];
                            const int kind = R.kind;
                            const int slot = kind == 0 ? vs : ls;
                            const bool linear =
                                kind == 0 ? opt.visibilityUseLinearSampler != 0 : opt.lightingUseLinearSampler != 0;
                            const float tss = kind == 0 ? opt.visibilityTStepScale : opt.lightingTStepScale;
                            segBegin(w, wc, slot, Ray{V3(R.ox, R.oy, R.oz), V3(R.dx, R.dy, R.dz), 0.f, R.tMax},
                                     linear, tss);
                            if (w.active) {
                                const int rs = 2 * lane + bank;
                                Q.rec[rs] = tri;
                                Q.sum[rs] = 0.f;
                                storeRay(Q, rs, w, slot, linear);
                            } else {
                                trs[tri] = segMissTr(w);  // the ray misses the grid
                            }
                        }
                    }
                    head += __popcll(wm);
                }
                if (!__ballot(w.active)) {
                    if (head >= total || !__ballot(!((waitMask >> bank) & 1))) break;
                    continue;
                }
                // block entries (FlatWalk), batched: a lane whose walk reached a block
                // border waits until kBlockBatch lanes (or all walking lanes) do
                const bool wasActive = w.active;
                const uint64_t bm = __ballot(w.active && segBlockEntry(w));
                if (bm && (__popcll(bm) >= kBlockBatch || bm == __ballot(w.active))) {
                    if (w.active && segBlockEntry(w)) segEnterBlock(w, wc);
                }
                bool push = false;
                uint32_t bk = kEmpty;
                float t0 = 0.f;
                v3 q0 = V3(0.f);
                i3v cell = {0, 0, 0};
                const int rs = 2 * lane + bank;
                if (w.active && !segBlockEntry(w)) {
                    // up to kCellsPerStep cells while they hold no brick: the iteration's
                    // ballots and queue bookkeeping are paid once for a run of empty cells
                    for (int k = 0; k < kCellsPerStep; k++) {
                        if constexpr (TK == 2) push = segStep(w, wc, bk, t0, q0);
                        else push = w.stepCells(wc.g, bk, t0, cell);
                        if (push || !w.active || segBlockEntry(w)) break;
                    }
                }
                if (wasActive && !w.active) {
                    waitMask |= 1 << bank;
                    bank ^= 1;
                }
                const uint64_t pm = __ballot(push);
                if (push) {
                    const int e = nSeg + __popcll(pm & below);
                    Q.owner[e] = rs;
                    Q.brick[e] = bk;
                    if constexpr (TK == 2) {
                        Q.t[e] = t0;
                        Q.q[0][e] = q0.x; Q.q[1][e] = q0.y; Q.q[2][e] = q0.z;
                    } else {
                        Q.tx[e] = t0;
                        Q.cell[e] = cell.x | (cell.y << 10) | (cell.z << 20);
                    }
                }
                nSeg += __popcll(pm);
            }
            waveSync();
            // ---- march phase: drain the segment queue, or all but a short tail when walks remain
            if constexpr (TK == 1) {
                if (uniformLinear) marchCells<true>(S, S.grid[vs], Q, nSeg, below);
                else marchCells<false>(S, S.grid[vs], Q, nSeg, below);
                nSeg = 0;
            } else if (uniformMarch) {
                const Grid& G = S.grid[vs];
                if (G.fp16) {
                    if (uniformLinear) marchPipelined<1, 1>(S, G, Q, nSeg, below);
                    else marchPipelined<1, 0>(S, G, Q, nSeg, below);
                } else {
                    if (uniformLinear) marchPipelined<0, 1>(S, G, Q, nSeg, below);
                    else marchPipelined<0, 0>(S, G, Q, nSeg, below);
                }
                nSeg = 0;
            } else {
                const bool mayCarry = head < total || __ballot(w.active) != 0ull;
                nSeg = marchPhase(S, Q, nSeg, below, mayCarry);
            }
            waveSync();
            // ---- rays whose walk ended and that have no carried segment are complete
#pragma unroll
            for (int b = 0; b < 2; b++) {
                const int rs = 2 * lane + b;
                if (((waitMask >> b) & 1) && !Q.pend[rs]) {
                    trs[Q.rec[rs]] = expf(Q.sum[rs]);
                    waitMask &= ~(1 << b);
                }
                Q.pend[rs] = 0;
            }
            waveSync();
            if (head >= total && !__ballot(w.active) && !__ballot(waitMask != 0)) break;
        }
    };
    // TK = 1 walks flat only: launchTraceCells refuses a grid it could not walk
    // that way, so no launch leaves a transmittance unwritten
    if (flat) {
        FlatWalk w;
        trace(w);
    } else if constexpr (TK == 2) {
        Walker<2> w;
        trace(w);
    }
}
template <int TK>
__global__ __launch_bounds__(256) void kTraceSegments(const Scene* __restrict__ sc, const Frame fr, Options opt,
                                                      int isFinal, const TraceRec* __restrict__ recs,
                                                      const int* __restrict__ counts, float* __restrict__ trs,
                                                      int cap, int numPixels) {
    traceSegmentsBody<TK>(sc, fr, opt, isFinal, recs, counts, trs, cap, numPixels, (int)blockIdx.x);
}

hipError_t launchTraceSegments(const LaunchArgs& a, const Options& opt, bool isFinal) {
    const int n = a.width * a.height;
    hipLaunchKernelGGL(kTraceSegments<2>, dim3((n + 255) / 256), dim3(256), a.occBytes, a.stream, a.scene, a.frame,
                       opt, isFinal ? 1 : 0, a.split.recs, a.split.counts, a.split.trs, a.split.cap, n);
    return hipGetLastError();
}

// The segment traces of a batched step: work items are 256-pixel ranges of one
// instance (its recorded rays), the same ranges a plain launch gives a workgroup.
// Held to the plain kernel's four waves per SIMD: left to itself the compiler
// gives the persistent loop 175 VGPRs (two waves), and the trace, latency bound,
// runs about a quarter slower.
template <int TK>
VR_TRACING_KERNEL void kBatchedSegments(const LaunchArgs* __restrict__ args,
                                                        const BatchList* __restrict__ list,
                                                        int* __restrict__ counter) {
    __shared__ int sItem;
    const int n = list->n, total = list->start[n], per = list->per;
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) sItem = atomicAdd(counter, 1);
        __syncthreads();
        // the same value in every lane: made wave-uniform so that the pass's
        // arguments below are read with scalar loads into scalar registers
        const int item = __builtin_amdgcn_readfirstlane(sItem);
        if (item >= total) break;
        int i = 0;
        while (i + 1 < n && item >= list->start[i + 1]) i++;
        const LaunchArgs& a = args[i];
        const int t0 = (item - list->start[i]) * per, t1 = min(t0 + per, list->tiles[i]);
        const bool fin = TK == 1;
        for (int t = t0; t < t1; t++) {
            if (t > t0) __syncthreads();
            traceSegmentsBody<TK>(a.scene, a.frame, fin ? a.frame.finalOpt : a.frame.spatialOpt, fin ? 1 : 0,
                                  a.split.recs, a.split.counts, a.split.trs, a.split.cap, a.width * a.height, t);
        }
    }
}

hipError_t launchBatchedSegments(hipStream_t st, int op, const LaunchArgs* args, const BatchList* list, int* counter,
                                 int grid, int occBytes) {
    if (grid <= 0) return hipErrorInvalidValue;
    if (op == kOpSegmentsMarch)
        hipLaunchKernelGGL(kBatchedSegments<2>, dim3(grid), dim3(256), occBytes, st, args, list, counter);
    else if (op == kOpSegmentsCells)
        hipLaunchKernelGGL(kBatchedSegments<1>, dim3(grid), dim3(256), occBytes, st, args, list, counter);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launchTraceCells(const LaunchArgs& a, const Options& opt, bool isFinal) {
    const int n = a.width * a.height;
    // the cell trace walks one grid whose blocks fit its LDS bit mask; it traces
    // nothing otherwise, so refuse what it cannot trace instead of leaving trs unwritten
    if (opt.visibilityMipLevel != opt.lightingMipLevel || a.traceGridBlocks <= 0 ||
        a.traceGridBlocks > kNodeWords * 32)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(kTraceSegments<1>, dim3((n + 255) / 256), dim3(256), a.occBytes, a.stream, a.scene, a.frame,
                       opt, isFinal ? 1 : 0, a.split.recs, a.split.counts, a.split.trs, a.split.cap, n);
    return hipGetLastError();
}

hipError_t launchTraceQueue(const LaunchArgs& a, const Options& opt, bool isFinal, int tk) {
    const int n = a.width * a.height;
    const dim3 g((n + 255) / 256), blk(256);
    // ray marching only: the analytic walker queue has no parity coverage and is not launched
    if (tk != 2) return hipErrorInvalidValue;
    hipLaunchKernelGGL(kTraceQueue<2>, g, blk, a.occBytes, a.stream, a.scene, a.frame, opt, isFinal ? 1 : 0,
                       a.split.recs, a.split.counts, a.split.trs, a.split.cap, n);
    return hipGetLastError();
}

}  // namespace vr
