// vr_volume.cpp -- host-side volume preparation: mip and conservative grids,
// brick pools with apron, brick / block tables, per-brick bounds; the seeded
// fBm generator of the synthetic workloads; camera orbit; .vbx I/O.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/vrestir.h"
#include "vr_volume.h"

namespace vr {

static int hostThreads() {
    unsigned n = std::thread::hardware_concurrency();
    return (int)std::max(1u, std::min(n, 16u));
}

template <class F>
static void parallelFor(int n, F&& fn) {
    int nt = std::min(hostThreads(), std::max(1, n));
    std::vector<std::thread> th;
    for (int t = 0; t < nt; t++)
        th.emplace_back([&, t] {
            for (int i = t; i < n; i += nt) fn(i);
        });
    for (auto& t : th) t.join();
}

// IEEE binary16 bits of a float, round to nearest even.
uint16_t floatToHalfBits(float f) {
    uint32_t x;
    std::memcpy(&x, &f, 4);
    uint32_t sign = (x >> 16) & 0x8000u;
    uint32_t absx = x & 0x7FFFFFFFu;
    if (absx >= 0x7F800000u) return (uint16_t)(sign | (absx > 0x7F800000u ? 0x7E00u : 0x7C00u));
    if (absx >= 0x477FF000u) return (uint16_t)(sign | 0x7C00u);  // rounds to >= 65520: infinity
    if (absx < 0x38800000u) {                                      // subnormal half (|x| < 2^-14)
        if (absx < 0x33000000u) return (uint16_t)sign;             // < 2^-25 rounds to 0
        uint32_t mant = (absx & 0x7FFFFFu) | 0x800000u;
        int e = (int)(absx >> 23);                                 // 102..112
        int shift = 126 - e;                                       // 14..24
        uint32_t hm = mant >> shift;
        uint32_t rem = mant & ((1u << shift) - 1u);
        uint32_t half = 1u << (shift - 1);
        if (rem > half || (rem == half && (hm & 1u))) hm++;
        return (uint16_t)(sign | hm);
    }
    uint32_t e = ((absx >> 23) - 112u) << 10;
    uint32_t m = (absx >> 13) & 0x3FFu;
    uint32_t h = e | m;
    uint32_t rem = absx & 0x1FFFu;
    if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h++;
    return (uint16_t)(sign | h);
}
float halfBitsToFloat(uint16_t h) {
    uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
    uint32_t e = (h >> 10) & 0x1Fu, m = h & 0x3FFu;
    uint32_t x;
    if (e == 0) {
        if (m == 0) x = sign;
        else {
            float v = std::ldexp((float)m, -24);
            std::memcpy(&x, &v, 4);
            x |= sign;
        }
    } else if (e == 31) {
        x = sign | 0x7F800000u | (m << 13);
    } else {
        x = sign | ((e + 112u) << 23) | (m << 13);
    }
    float f;
    std::memcpy(&f, &x, 4);
    return f;
}

// 2x box reduction (x-major sum order, weight 1/8 per tap, negative results to 0),
// the rule of GL/gvdb_volume_gvdb.cpp:2834-2856 for even dimensions.
static DenseF reduce2(const DenseF& s) {
    DenseF d;
    d.nx = s.nx / 2; d.ny = s.ny / 2; d.nz = s.nz / 2;
    d.v.assign((size_t)d.nx * d.ny * d.nz, 0.f);
    parallelFor(d.nz, [&](int k) {
        for (int j = 0; j < d.ny; j++)
            for (int i = 0; i < d.nx; i++) {
                float res = 0.f;
                for (int a = 0; a < 2; a++)
                    for (int b = 0; b < 2; b++)
                        for (int c = 0; c < 2; c++) res += 0.125f * s.at(2 * i + a, 2 * j + b, 2 * k + c);
                d.v[d.idx(i, j, k)] = res > 0.f ? res : 0.f;
            }
    });
    return d;
}

// Conservative mip 0: an empty voxel gets the mean of the positive values of
// its 27-neighbourhood (GL/gvdb_volume_gvdb.cpp:2753-2800).
static DenseF dilate(const DenseF& s) {
    DenseF d = s;
    parallelFor(s.nz, [&](int k) {
        for (int j = 0; j < s.ny; j++)
            for (int i = 0; i < s.nx; i++) {
                if (s.v[s.idx(i, j, k)] != 0.f) continue;
                float avg = 0.f;
                for (int ii = -1; ii <= 1; ii++)
                    for (int jj = -1; jj <= 1; jj++)
                        for (int kk = -1; kk <= 1; kk++) {
                            int x = i + ii, y = j + jj, z = k + kk;
                            if (x < 0 || y < 0 || z < 0 || x >= s.nx || y >= s.ny || z >= s.nz) continue;
                            float tv = s.v[s.idx(x, y, z)];
                            avg += tv > 0.f ? tv : 0.f;
                        }
                avg /= 27.f;
                if (avg > 0.f) d.v[d.idx(i, j, k)] = avg;
            }
    });
    return d;
}

// One slot: stored values (tiny-value clamp of SC/Scene.cpp:3154-3156, then the
// storage precision), brick table, block table, apron pool, bounds
// (min / max / sum/512 over the 10^3 neighbourhood, SC/Scene.cpp:2985-3012).
// Src gives the raw value of any voxel (0 outside the grid) and whether a brick
// cell can hold a non-zero value; the result depends only on the values, so a
// dense grid and the bricks that hold it build the same slot.
template <class Src>
static void buildSlotFrom(HostGrid& g, const Src& src, int nx, int ny, int nz, float gmax, bool fp16) {
    g.nx = nx; g.ny = ny; g.nz = nz;
    g.fp16 = fp16;
    // stored value of a raw value; monotone, so the largest stored value is q(gmax)
    auto q = [&](float x) {
        x = (gmax > 0.f && x / gmax < 1e-9f) ? 0.f : x;
        if (gmax <= 0.f) x = 0.f;
        return fp16 ? halfBitsToFloat(floatToHalfBits(x)) : x;
    };
    g.maxValue = q(gmax);
    g.nbx = nx / 8; g.nby = ny / 8; g.nbz = nz / 8;
    g.nnx = (g.nbx + 15) / 16; g.nny = (g.nby + 15) / 16; g.nnz = (g.nbz + 15) / 16;
    size_t nbTot = (size_t)g.nbx * g.nby * g.nbz;
    g.brickTable.assign(nbTot, kEmpty);
    g.nodeTable.assign((size_t)g.nnx * g.nny * g.nnz, 0u);
    // occupancy in parallel, ids in z,y,x order
    std::vector<uint8_t> occ(nbTot, 0);
    parallelFor(g.nbz, [&](int bz) {
        for (int by = 0; by < g.nby; by++)
            for (int bx = 0; bx < g.nbx; bx++) {
                bool any = false;
                if (src.mayHold(bx, by, bz))
                    for (int z = 0; z < 8 && !any; z++)
                        for (int y = 0; y < 8 && !any; y++)
                            for (int x = 0; x < 8 && !any; x++)
                                if (q(src.at(bx * 8 + x, by * 8 + y, bz * 8 + z)) > 0.f) any = true;
                occ[bx + (size_t)g.nbx * (by + (size_t)g.nby * bz)] = any ? 1 : 0;
            }
    });
    uint32_t nb = 0;
    std::vector<int32_t> brickPos;
    for (int bz = 0; bz < g.nbz; bz++)
        for (int by = 0; by < g.nby; by++)
            for (int bx = 0; bx < g.nbx; bx++) {
                size_t c = bx + (size_t)g.nbx * (by + (size_t)g.nby * bz);
                if (!occ[c]) continue;
                g.brickTable[c] = nb++;
                brickPos.push_back(bx); brickPos.push_back(by); brickPos.push_back(bz);
                g.nodeTable[(bx / 16) + (size_t)g.nnx * ((by / 16) + (size_t)g.nny * (bz / 16))] = 1u;
            }
    g.numBricks = nb;
    const size_t words = (nbTot + 31) / 32;
    g.occ.assign(2 * words, 0u);
    uint32_t before = 0;
    for (size_t w = 0; w < words; w++) {
        uint32_t bits = 0;
        for (int b = 0; b < 32; b++) {
            size_t c = w * 32 + b;
            if (c < nbTot && occ[c]) bits |= 1u << b;
        }
        g.occ[2 * w] = bits;
        g.occ[2 * w + 1] = before;
        before += (uint32_t)__builtin_popcount(bits);
    }
    g.bounds.assign((size_t)nb * 4, 0.f);
    if (fp16) g.pool16.assign((size_t)nb * kBrickVals, 0);
    else g.pool32.assign((size_t)nb * kBrickVals, 0.f);
    parallelFor((int)nb, [&](int id) {
        int bx = brickPos[3 * id], by = brickPos[3 * id + 1], bz = brickPos[3 * id + 2];
        int x0 = bx * 8, y0 = by * 8, z0 = bz * 8;
        float mn = FLT_MAX, mx = 0.f, sum = 0.f;
        for (int i = -1; i <= 8; i++)
            for (int j = -1; j <= 8; j++)
                for (int k = -1; k <= 8; k++) {
                    float v = q(src.at(x0 + i, y0 + j, z0 + k));
                    size_t o = (size_t)id * kBrickVals + (size_t)(i + 1) + kApron * ((size_t)(j + 1) + kApron * (size_t)(k + 1));
                    if (fp16) g.pool16[o] = floatToHalfBits(v);
                    else g.pool32[o] = v;
                    mn = std::min(mn, v);
                    mx = std::max(mx, v);
                    sum += v;
                }
        g.bounds[4 * (size_t)id + 0] = mn;
        g.bounds[4 * (size_t)id + 1] = mx;
        g.bounds[4 * (size_t)id + 2] = sum / 512.f;
        g.bounds[4 * (size_t)id + 3] = 0.f;
    });
}

struct DenseSrc {
    const DenseF& d;
    float at(int x, int y, int z) const { return d.at(x, y, z); }
    bool mayHold(int, int, int) const { return true; }
};

static void buildSlot(HostGrid& g, const DenseF& src, bool fp16) {
    float gmax = 0.f;
    for (float x : src.v) gmax = std::max(gmax, x);
    buildSlotFrom(g, DenseSrc{src}, src.nx, src.ny, src.nz, gmax, fp16);
}

// Voxels of a brick list: a table over brick cells names the brick holding each
// cell (values <= 0 read 0, as a clamped dense grid).
struct BrickSrc {
    const VbxBricks& b;
    int nbx, nby, nbz;
    std::vector<int32_t> cell;
    float at(int x, int y, int z) const {
        if (x < 0 || y < 0 || z < 0) return 0.f;
        const int cx = x >> 3, cy = y >> 3, cz = z >> 3;
        if (cx >= nbx || cy >= nby || cz >= nbz) return 0.f;
        const int32_t id = cell[cx + (size_t)nbx * (cy + (size_t)nby * cz)];
        if (id < 0) return 0.f;
        const float v = b.vals[(size_t)id * 512 + ((z & 7) * 8 + (y & 7)) * 8 + (x & 7)];
        return v > 0.f ? v : 0.f;
    }
    bool mayHold(int bx, int by, int bz) const { return cell[bx + (size_t)nbx * (by + (size_t)nby * bz)] >= 0; }
};

static int brickSource(const VbxBricks& b, BrickSrc& s, int& nx, int& ny, int& nz, float& gmax, std::string& err) {
    if (b.brickDim != 8) { err = "only 8^3 bricks are supported"; return VRESTIR_EUNSUPPORTED; }
    nx = (b.nx + 7) / 8 * 8; ny = (b.ny + 7) / 8 * 8; nz = (b.nz + 7) / 8 * 8;
    if (nx <= 0 || ny <= 0 || nz <= 0 || nx > 4096 || ny > 4096 || nz > 4096) {
        err = "grid dimensions must be in [1, 4096]";
        return VRESTIR_EINVAL;
    }
    s.nbx = nx / 8; s.nby = ny / 8; s.nbz = nz / 8;
    s.cell.assign((size_t)s.nbx * s.nby * s.nbz, -1);
    gmax = 0.f;
    const size_t n = b.pos.size() / 3;
    for (size_t i = 0; i < n; i++) {
        const int x = b.pos[3 * i], y = b.pos[3 * i + 1], z = b.pos[3 * i + 2];
        if ((x | y | z) & 7) { err = "brick positions must be multiples of 8"; return VRESTIR_EUNSUPPORTED; }
        if (x < 0 || y < 0 || z < 0 || x >= nx || y >= ny || z >= nz) continue;  // past the extent: all 0
        s.cell[(x >> 3) + (size_t)s.nbx * ((y >> 3) + (size_t)s.nby * (z >> 3))] = (int32_t)i;
        for (int k = 0; k < 512; k++) {
            const float v = b.vals[i * 512 + k];
            if (!std::isfinite(v)) { err = "the grid holds a non-finite value"; return VRESTIR_EINVAL; }
            gmax = std::max(gmax, v);
        }
    }
    return VRESTIR_OK;
}

int buildSlotFromBricks(HostGrid& g, const VbxBricks& b, bool fp16, std::string& err) {
    BrickSrc s{b, 0, 0, 0, {}};
    int nx, ny, nz;
    float gmax;
    int rc = brickSource(b, s, nx, ny, nz, gmax, err);
    if (rc) return rc;
    g = HostGrid();
    buildSlotFrom(g, s, nx, ny, nz, gmax, fp16);
    g.valid = true;
    return VRESTIR_OK;
}


// world -> index transform of mip k from the mip-0 index -> world transform.
static void mipTransforms(const float* i2w0, int scale, float* w2i) {
    double M[16];
    for (int r = 0; r < 4; r++)
        for (int c = 0; c < 4; c++) M[r * 4 + c] = (r < 3 ? (double)scale : 1.0) * i2w0[r * 4 + c];
    double a = M[0], b = M[1], c = M[2], d = M[4], e = M[5], f = M[6], gg = M[8], h = M[9], k = M[10];
    double det = a * (e * k - f * h) - b * (d * k - f * gg) + c * (d * h - e * gg);
    double I[9] = {(e * k - f * h) / det, -(b * k - c * h) / det, (b * f - c * e) / det,
                   -(d * k - f * gg) / det, (a * k - c * gg) / det, -(a * f - c * d) / det,
                   (d * h - e * gg) / det, -(a * h - b * gg) / det, (a * e - b * d) / det};
    for (int r = 0; r < 3; r++) {
        for (int cc = 0; cc < 3; cc++) w2i[r * 4 + cc] = (float)I[r * 3 + cc];
        w2i[r * 4 + 3] = 0.f;
    }
    for (int cc = 0; cc < 3; cc++)
        w2i[12 + cc] = (float)(-(M[12] * I[0 * 3 + cc] + M[13] * I[1 * 3 + cc] + M[14] * I[2 * 3 + cc]));
    w2i[15] = 1.f;
}

void slotTransform(const float* indexToWorld, float* w2i) { mipTransforms(indexToWorld, 1, w2i); }
void mipTransform(const float* indexToWorld, int scale, float* w2i) { mipTransforms(indexToWorld, scale, w2i); }

// The dense grid a brick list holds (values <= 0 read 0).
static DenseF denseOf(const VbxBricks& b) {
    BrickSrc s{b, 0, 0, 0, {}};
    int nx = 0, ny = 0, nz = 0;
    float gmax;
    std::string err;
    DenseF d;
    if (brickSource(b, s, nx, ny, nz, gmax, err)) return d;
    d.nx = nx; d.ny = ny; d.nz = nz;
    d.v.assign((size_t)nx * ny * nz, 0.f);
    parallelFor(nz, [&](int z) {
        for (int y = 0; y < ny; y++)
            for (int x = 0; x < nx; x++) d.v[d.idx(x, y, z)] = s.at(x, y, z);
    });
    return d;
}

static bool fileExists(const std::string& path) {
    FILE* f = std::fopen(path.c_str(), "rb");
    if (f) std::fclose(f);
    return f != nullptr;
}

// addGVDBVolume's file set (SC/Scene.cpp:2794-2850): <prefix>_mip{k}.vbx and its
// conservative twin <prefix>_mip{k}c.vbx per mip, each with its own index -> world
// transform; slots are built straight from the files' bricks. Where a file is
// missing the level is derived as buildVolume derives it: a conservative twin
// from the dilation of its mip (mip 0) or the 2x reduction of the twin above it,
// a mip from the 2x reduction of the mip above it (through dense grids of those
// levels). Mip 0 must exist.
int buildVolumeFromFiles(const std::string& prefix, int numMips, bool fp16, HostVolume& out, float* i2w0,
                         std::string& err) {
    if (numMips < 1 || numMips > 8) { err = "numMips must be in [1, 8]"; return VRESTIR_EINVAL; }
    out = HostVolume();
    DenseF prevN, prevC;   // the level above, when a later level must be derived from it
    float prevI2W[16];
    for (int k = 0; k < numMips; k++) {
        const std::string base = prefix + "_mip" + std::to_string(k);
        const bool haveN = fileExists(base + ".vbx"), haveC = fileExists(base + "c.vbx");
        if (k == 0 && !haveN) { err = "cannot read " + base + ".vbx"; return VRESTIR_EIO; }
        float i2w[16];
        VbxBricks bn, bc;
        int rc;
        if (haveN) {
            if ((rc = readVbxBricks((base + ".vbx").c_str(), bn))) { err = "cannot read " + base + ".vbx"; return rc; }
            if ((rc = buildSlotFromBricks(out.slots[k], bn, fp16, err))) return rc;
            std::memcpy(i2w, bn.indexToWorld, sizeof(i2w));
        } else {
            for (int r = 0; r < 16; r++) i2w[r] = (r < 12 ? 2.f : 1.f) * prevI2W[r];
            prevN = reduce2(prevN);
            buildSlot(out.slots[k], prevN, fp16);
            out.slots[k].valid = true;
        }
        slotTransform(i2w, out.slots[k].w2i);
        if (k == 0) std::memcpy(i2w0, i2w, sizeof(i2w));
        DenseF cDerived;
        if (haveC) {
            if ((rc = readVbxBricks((base + "c.vbx").c_str(), bc))) { err = "cannot read " + base + "c.vbx"; return rc; }
            if ((rc = buildSlotFromBricks(out.slots[8 + k], bc, fp16, err))) return rc;
            slotTransform(bc.indexToWorld, out.slots[8 + k].w2i);
        } else {
            cDerived = k == 0 ? dilate(denseOf(bn)) : reduce2(prevC);
            buildSlot(out.slots[8 + k], cDerived, fp16);
            out.slots[8 + k].valid = true;
            slotTransform(i2w, out.slots[8 + k].w2i);
        }
        // dense copies of this level, kept only when the next level is derived from them
        const std::string next = prefix + "_mip" + std::to_string(k + 1);
        if (k + 1 < numMips) {
            if (!fileExists(next + ".vbx") && haveN) prevN = denseOf(bn);
            if (!fileExists(next + "c.vbx")) prevC = haveC ? denseOf(bc) : std::move(cDerived);
        }
        std::memcpy(prevI2W, i2w, sizeof(i2w));
        out.numMips = k + 1;
    }
    volumeSpacing(i2w0, out);
    return VRESTIR_OK;
}

// Ray-marching step and supervoxel diagonal from mip 0's index -> world transform
// (SC/Scene.cpp:3074-3080).
void volumeSpacing(const float* X, HostVolume& v) {
    auto rl2 = [&](int r) { return X[r * 4] * X[r * 4] + X[r * 4 + 1] * X[r * 4 + 1] + X[r * 4 + 2] * X[r * 4 + 2]; };
    v.tStep = (std::sqrt(rl2(0)) + std::sqrt(rl2(1)) + std::sqrt(rl2(2))) / 3.f;
    v.superVoxelDiag = 8.f * std::sqrt(rl2(0) + rl2(1) + rl2(2));
}

int checkVolumeDesc(const vrestir_volume_desc& d, std::string& err) {
    if (d.numMips < 1 || d.numMips > 8) { err = "numMips must be in [1, 8]"; return VRESTIR_EINVAL; }
    if (!d.density) { err = "density is NULL"; return VRESTIR_EINVAL; }
    int align = 8 << (d.numMips - 1);
    if (d.nx <= 0 || d.ny <= 0 || d.nz <= 0 || d.nx % align || d.ny % align || d.nz % align) {
        err = "grid dimensions must be positive multiples of 8 << (numMips - 1)";
        return VRESTIR_EINVAL;
    }
    if (d.nx > 4096 || d.ny > 4096 || d.nz > 4096) { err = "grid dimensions above 4096"; return VRESTIR_EINVAL; }
    if (std::fabs((d.sigmaS[0] + d.sigmaA[0]) - (d.sigmaS[1] + d.sigmaA[1])) > 1e-6f * (d.sigmaS[0] + d.sigmaA[0]) ||
        std::fabs((d.sigmaS[1] + d.sigmaA[1]) - (d.sigmaS[2] + d.sigmaA[2])) > 1e-6f * (d.sigmaS[0] + d.sigmaA[0])) {
        err = "sigma_a + sigma_s must be grey";
        return VRESTIR_EINVAL;
    }
    if (d.brickFormat != VRESTIR_BRICK_FP32 && d.brickFormat != VRESTIR_BRICK_FP16) {
        err = "unknown brickFormat";
        return VRESTIR_EINVAL;
    }
    return VRESTIR_OK;
}

int buildVolume(const vrestir_volume_desc& d, HostVolume& out, std::string& err) {
    int rc0 = checkVolumeDesc(d, err);
    if (rc0) return rc0;
    out = HostVolume();
    DenseF base;
    base.nx = d.nx; base.ny = d.ny; base.nz = d.nz;
    base.v.assign(d.density, d.density + (size_t)d.nx * d.ny * d.nz);
    for (float& v : base.v) {
        // the pass's majorants and optical depths assume finite densities
        if (!std::isfinite(v)) { err = "the density grid holds a non-finite value"; return VRESTIR_EINVAL; }
        v = v > 0.f ? v : 0.f;
    }
    bool fp16 = d.brickFormat == VRESTIR_BRICK_FP16;
    DenseF cons = dilate(base);
    DenseF n = std::move(base), c = std::move(cons);
    for (int k = 0; k < d.numMips; k++) {
        if (k > 0) { n = reduce2(n); c = reduce2(c); }
        buildSlot(out.slots[k], n, fp16);
        buildSlot(out.slots[8 + k], c, fp16);
        mipTransforms(d.indexToWorld, 1 << k, out.slots[k].w2i);
        mipTransforms(d.indexToWorld, 1 << k, out.slots[8 + k].w2i);
        out.slots[k].valid = out.slots[8 + k].valid = true;
    }
    out.numMips = d.numMips;
    if (d.temperature) {
        int rc = buildTemperatureSlot(d.temperature, d.nx, d.ny, d.nz, d.indexToWorld, fp16,
                                      out.slots[kTemperatureSlot], err);
        if (rc) return rc;
    }
    if (d.velocity) {
        int rc = buildVelocitySlot(d.velocity, d.nx, d.ny, d.nz, d.indexToWorld, out.slots[kVelocitySlot], err);
        if (rc) return rc;
    }
    volumeSpacing(d.indexToWorld, out);
    return VRESTIR_OK;
}

// The temperature grid of an emissive volume (slot kTemperatureSlot, loaded from
// <name>_temperature.vbx at SC/Scene.cpp:2808-2811): one full-resolution slot with
// its own index -> world transform. Values <= 0 are stored as 0 (they leave no
// brick); dimensions are padded with zeros to whole bricks.
int buildTemperatureSlot(const float* t, int nx, int ny, int nz, const float* indexToWorld, bool fp16,
                         HostGrid& out, std::string& err) {
    if (!t || nx <= 0 || ny <= 0 || nz <= 0 || nx > 4096 || ny > 4096 || nz > 4096) {
        err = "temperature grid dimensions must be in [1, 4096]";
        return VRESTIR_EINVAL;
    }
    DenseF g;
    g.nx = (nx + 7) / 8 * 8; g.ny = (ny + 7) / 8 * 8; g.nz = (nz + 7) / 8 * 8;
    g.v.assign((size_t)g.nx * g.ny * g.nz, 0.f);
    for (int z = 0; z < nz; z++)
        for (int y = 0; y < ny; y++)
            for (int x = 0; x < nx; x++) {
                float v = t[(size_t)x + (size_t)nx * ((size_t)y + (size_t)ny * z)];
                if (!std::isfinite(v)) { err = "temperature grid holds a non-finite value"; return VRESTIR_EINVAL; }
                g.v[g.idx(x, y, z)] = v > 0.f ? v : 0.f;
            }
    out = HostGrid();
    buildSlot(out, g, fp16);
    mipTransforms(indexToWorld, 1, out.w2i);
    out.valid = true;
    return VRESTIR_OK;
}

// The velocity grid of an animated volume (slot kVelocitySlot; the reference loads
// <name>_velocity_{x,y,z}.vbx, SC/Scene.cpp:2812-2820): three fp32 channels per
// voxel sharing one brick topology (a brick exists where any component is
// non-zero), pool laid out [brick][channel][10^3 apron values].
int buildVelocitySlot(const float* v, int nx, int ny, int nz, const float* indexToWorld, HostGrid& g,
                      std::string& err) {
    if (!v || nx <= 0 || ny <= 0 || nz <= 0 || nx > 4096 || ny > 4096 || nz > 4096) {
        err = "velocity grid dimensions must be in [1, 4096]";
        return VRESTIR_EINVAL;
    }
    g = HostGrid();
    const int px = (nx + 7) / 8 * 8, py = (ny + 7) / 8 * 8, pz = (nz + 7) / 8 * 8;
    auto at = [&](int x, int y, int z, int c) -> float {
        if (x < 0 || y < 0 || z < 0 || x >= nx || y >= ny || z >= nz) return 0.f;
        return v[3 * ((size_t)x + (size_t)nx * ((size_t)y + (size_t)ny * z)) + c];
    };
    for (size_t i = 0; i < 3 * (size_t)nx * ny * nz; i++)
        if (!std::isfinite(v[i])) { err = "velocity grid holds a non-finite value"; return VRESTIR_EINVAL; }
    g.nx = px; g.ny = py; g.nz = pz;
    g.fp16 = false;
    g.nbx = px / 8; g.nby = py / 8; g.nbz = pz / 8;
    g.nnx = (g.nbx + 15) / 16; g.nny = (g.nby + 15) / 16; g.nnz = (g.nbz + 15) / 16;
    const size_t nbTot = (size_t)g.nbx * g.nby * g.nbz;
    g.brickTable.assign(nbTot, kEmpty);
    g.nodeTable.assign((size_t)g.nnx * g.nny * g.nnz, 0u);
    std::vector<int32_t> brickPos;
    uint32_t nb = 0;
    for (int bz = 0; bz < g.nbz; bz++)
        for (int by = 0; by < g.nby; by++)
            for (int bx = 0; bx < g.nbx; bx++) {
                bool any = false;
                for (int z = 0; z < 8 && !any; z++)
                    for (int y = 0; y < 8 && !any; y++)
                        for (int x = 0; x < 8 && !any; x++)
                            for (int c = 0; c < 3 && !any; c++)
                                if (at(bx * 8 + x, by * 8 + y, bz * 8 + z, c) != 0.f) any = true;
                if (!any) continue;
                g.brickTable[bx + (size_t)g.nbx * (by + (size_t)g.nby * bz)] = nb++;
                brickPos.push_back(bx); brickPos.push_back(by); brickPos.push_back(bz);
                g.nodeTable[(bx / 16) + (size_t)g.nnx * ((by / 16) + (size_t)g.nny * (bz / 16))] = 1u;
            }
    g.numBricks = nb;
    const size_t words = (nbTot + 31) / 32;
    g.occ.assign(2 * words, 0u);
    uint32_t before = 0;
    for (size_t w = 0; w < words; w++) {
        uint32_t bits = 0;
        for (int b = 0; b < 32; b++) {
            size_t c = w * 32 + b;
            if (c < nbTot && g.brickTable[c] != kEmpty) bits |= 1u << b;
        }
        g.occ[2 * w] = bits;
        g.occ[2 * w + 1] = before;
        before += (uint32_t)__builtin_popcount(bits);
    }
    g.bounds.assign((size_t)nb * 4, 0.f);
    g.pool32.assign((size_t)nb * 3 * kBrickVals, 0.f);
    parallelFor((int)nb, [&](int id) {
        int x0 = brickPos[3 * id] * 8, y0 = brickPos[3 * id + 1] * 8, z0 = brickPos[3 * id + 2] * 8;
        for (int c = 0; c < 3; c++)
            for (int k = -1; k <= 8; k++)
                for (int j = -1; j <= 8; j++)
                    for (int i = -1; i <= 8; i++)
                        g.pool32[((size_t)id * 3 + c) * kBrickVals + (size_t)(i + 1) +
                                 kApron * ((size_t)(j + 1) + kApron * (size_t)(k + 1))] = at(x0 + i, y0 + j, z0 + k, c);
    });
    g.channels = 3;
    mipTransforms(indexToWorld, 1, g.w2i);
    g.valid = true;
    return VRESTIR_OK;
}

// ------------------------------------------------------------------ synthetic fBm
namespace {
struct Perlin {
    int perm[512];
    explicit Perlin(uint64_t seed) {
        int p[256];
        for (int i = 0; i < 256; i++) p[i] = i;
        uint64_t st = seed;
        for (int i = 255; i > 0; i--) {
            uint64_t z = (st += 0x9E3779B97F4A7C15ull);
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            z ^= z >> 31;
            int j = (int)(z % (uint64_t)(i + 1));
            std::swap(p[i], p[j]);
        }
        for (int i = 0; i < 512; i++) perm[i] = p[i & 255];
    }
    static double fade(double t) { return t * t * t * (t * (t * 6 - 15) + 10); }
    static double grad(int h, double x, double y, double z) {
        int hh = h & 15;
        double u = hh < 8 ? x : y, v = hh < 4 ? y : (hh == 12 || hh == 14 ? x : z);
        return ((hh & 1) ? -u : u) + ((hh & 2) ? -v : v);
    }
    static double lerp(double t, double a, double b) { return a + t * (b - a); }
    // Improved gradient noise (K. Perlin, "Improving Noise", SIGGRAPH 2002), period 256.
    double noise(double x, double y, double z) const {
        double fx = std::floor(x), fy = std::floor(y), fz = std::floor(z);
        int X = (int)fx & 255, Y = (int)fy & 255, Z = (int)fz & 255;
        x -= fx; y -= fy; z -= fz;
        double u = fade(x), v = fade(y), w = fade(z);
        int A = perm[X] + Y, AA = perm[A] + Z, AB = perm[A + 1] + Z;
        int B = perm[X + 1] + Y, BA = perm[B] + Z, BB = perm[B + 1] + Z;
        return lerp(w,
                    lerp(v, lerp(u, grad(perm[AA], x, y, z), grad(perm[BA], x - 1, y, z)),
                         lerp(u, grad(perm[AB], x, y - 1, z), grad(perm[BB], x - 1, y - 1, z))),
                    lerp(v, lerp(u, grad(perm[AA + 1], x, y, z - 1), grad(perm[BA + 1], x - 1, y, z - 1)),
                         lerp(u, grad(perm[AB + 1], x, y - 1, z - 1), grad(perm[BB + 1], x - 1, y - 1, z - 1))));
    }
};
}  // namespace

int synthFbm(uint64_t seed, int nx, int ny, int nz, int octaves, float baseFreq, float lacunarity, float gain,
             int sphereMask, float threshold, float amplitude, float* out) {
    if (nx <= 0 || ny <= 0 || nz <= 0 || !out || octaves < 1 || octaves > 16) return VRESTIR_EINVAL;
    Perlin P(seed);
    int nmax = std::max(nx, std::max(ny, nz));
    parallelFor(nz, [&](int k) {
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                // voxel centre in [0,1]^3 of the largest axis
                double x = (i + 0.5) / nmax, y = (j + 0.5) / nmax, z = (k + 0.5) / nmax;
                double f = 0.0, amp = 1.0, fr = baseFreq;
                for (int o = 0; o < octaves; o++) {
                    f += amp * P.noise(x * fr + 17.31 * o, y * fr + 5.13 * o, z * fr + 11.7 * o);
                    amp *= gain;
                    fr *= lacunarity;
                }
                double v = f - threshold;
                if (sphereMask) {
                    double cx = 0.5 * nx / nmax, cy = 0.5 * ny / nmax, cz = 0.5 * nz / nmax;
                    double r = 0.5 * std::min(nx, std::min(ny, nz)) / nmax;
                    double dd = std::sqrt((x - cx) * (x - cx) + (y - cy) * (y - cy) + (z - cz) * (z - cz));
                    if (dd > r) v = 0.0;
                }
                out[(size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * k)] = v > 0.0 ? (float)(amplitude * v) : 0.f;
            }
    });
    return VRESTIR_OK;
}

}  // namespace vr
