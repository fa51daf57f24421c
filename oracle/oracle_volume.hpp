// oracle_volume.hpp -- TEST INFRASTRUCTURE ONLY (the CPU oracle).
//
// A GVDB-shaped volume as the reference binds it to its shaders: a three-level
// tree (root 32^3 children of 128^3 voxels, level-1 nodes 16^3 children of 8^3
// bricks), leaf nodes with packed position / atlas offset and per-brick density
// bounds, and a texture atlas of bricks with a one-voxel apron, sampled with a
// software point or trilinear filter (border colour 0).
//
//   node layout / lookups       GVS/gvdbNodes.slang:38-136, 175-279
//   atlas sampling              GVS/gvdb.slang:6-71, VR/VolumeBase.slang:234-263
//   per-brick density bounds    SC/Scene.cpp:2977-3012
//   mips / conservative grid    GL/gvdb_volume_gvdb.cpp:2753-2863
#pragma once
#include <algorithm>
#include <vector>

#include "oracle_math.hpp"

namespace orc {

static const uint32_t kUndef = 0xFFFFFFFFu;  // GVS/gvdb.slang:4

struct LeafNode {
    i3 vmin;          // index-space minimum of the brick (unpackPos)
    i3 offs;          // atlas-space minimum without apron (unpackValue)
    f4 bounds;        // min, max, avg over the 10^3 neighbourhood, 0
};
struct InnerNode {
    i3 vmin;
    uint32_t childList;  // index into the level's child-list pool
};

// One grid slot (mip k normal: slot k; conservative: slot 8+k).
struct GridSlot {
    bool valid = false;
    int nx = 0, ny = 0, nz = 0;
    f3 bmin{0, 0, 0}, bmax{0, 0, 0};
    float xform[16];      // index -> world (row vector)
    float invxform[16];   // world -> index
    float maxValue = 0.f;
    float compressScale = 1.f;
    // level 2 (root) and level 1: child lists of u32 (level 2: 32^3 per list, level 1: 16^3).
    InnerNode root;
    std::vector<uint32_t> childPool2;  // one list
    std::vector<InnerNode> level1;
    std::vector<uint32_t> childPool1;  // level1.size() lists of 4096
    std::vector<LeafNode> level0;
    // atlas: a near-cubic arrangement of bricks, 10^3 texels each (like the GVDB
    // atlas, so atlas coordinates stay small and keep their float precision)
    int abx = 1, aby = 1, abz = 1;
    int ax = 10, ay = 10, az = 10;
    std::vector<float> atlas;
};

static inline float atlasTexel(const GridSlot& g, int x, int y, int z) {
    if (x < 0 || y < 0 || z < 0 || x >= g.ax || y >= g.ay || z >= g.az) return 0.f;  // border colour 0
    return g.atlas[(size_t)x + (size_t)g.ax * ((size_t)y + (size_t)g.ay * (size_t)z)];
}

// Texture SampleLevel at texel-space coordinate c: point = texel floor(c);
// linear = trilinear over texel centres (c - 0.5).
static inline float sampleAtlas(const GridSlot& g, f3 c, bool linear) {
    if (!linear) return atlasTexel(g, (int)std::floor(c.x), (int)std::floor(c.y), (int)std::floor(c.z));
    f3 x = c - 0.5f;
    f3 i0 = f3floor(x);
    f3 f = x - i0;
    int ix = (int)i0.x, iy = (int)i0.y, iz = (int)i0.z;
    float v000 = atlasTexel(g, ix, iy, iz), v100 = atlasTexel(g, ix + 1, iy, iz);
    float v010 = atlasTexel(g, ix, iy + 1, iz), v110 = atlasTexel(g, ix + 1, iy + 1, iz);
    float v001 = atlasTexel(g, ix, iy, iz + 1), v101 = atlasTexel(g, ix + 1, iy, iz + 1);
    float v011 = atlasTexel(g, ix, iy + 1, iz + 1), v111 = atlasTexel(g, ix + 1, iy + 1, iz + 1);
    float a0 = v000 + (v100 - v000) * f.x, a1 = v010 + (v110 - v010) * f.x;
    float a2 = v001 + (v101 - v001) * f.x, a3 = v011 + (v111 - v011) * f.x;
    float b0 = a0 + (a1 - a0) * f.y, b1 = a2 + (a3 - a2) * f.y;
    return b0 + (b1 - b0) * f.z;
}

// getChild (GVS/gvdbNodes.slang:108-114). An index past the end of the
// node's list reads as empty (the reference's out-of-range buffer load would
// return child 0 of the pool; treated as empty here and in the HIP path).
static inline uint32_t childOf(const GridSlot& g, int lev, uint32_t listId, int b) {
    if (listId == kUndef) return kUndef;
    if (lev == 2) {
        if (b < 0 || b >= 32 * 32 * 32) return kUndef;
        return g.childPool2[b];
    }
    if (b < 0 || b >= 16 * 16 * 16) return kUndef;
    return g.childPool1[(size_t)listId * 4096 + b];
}

// getNodeUnrolledThreeLevel + getNodeAtPoint (GVS/gvdbNodes.slang:175-224, 261-279).
static inline const LeafNode* leafAtPoint(const GridSlot& g, f3 pos) {
    f3 vmin = f3of(g.root.vmin);
    if (pos.x < vmin.x || pos.y < vmin.y || pos.z < vmin.z ||
        pos.x >= vmin.x + 4096.f || pos.y >= vmin.y + 4096.f || pos.z >= vmin.z + 4096.f) return nullptr;
    i3 p{(int)(pos.x - vmin.x) / 128, (int)(pos.y - vmin.y) / 128, (int)(pos.z - vmin.z) / 128};
    int b = (((p.z << 5) + p.y) << 5) + p.x;
    uint32_t n1 = childOf(g, 2, 0, b);
    if (n1 == kUndef) return nullptr;
    const InnerNode& node1 = g.level1[n1];
    vmin = f3of(node1.vmin);
    if (pos.x < vmin.x || pos.y < vmin.y || pos.z < vmin.z ||
        pos.x >= vmin.x + 128.f || pos.y >= vmin.y + 128.f || pos.z >= vmin.z + 128.f) return nullptr;
    p = {(int)(pos.x - vmin.x) / 8, (int)(pos.y - vmin.y) / 8, (int)(pos.z - vmin.z) / 8};
    b = (((p.z << 4) + p.y) << 4) + p.x;
    uint32_t n0 = childOf(g, 1, node1.childList, b);
    if (n0 == kUndef) return nullptr;
    return &g.level0[n0];
}

// getValueAtPoint (GVS/gvdb.slang:6-29).
static inline float valueAtPoint(const GridSlot& g, f3 pos, bool linear) {
    const LeafNode* leaf = leafAtPoint(g, pos);
    if (!leaf) return 0.f;
    f3 tc = f3of(leaf->offs) + (pos - f3of(leaf->vmin));
    return sampleAtlas(g, tc, linear);
}

// ---------------------------------------------------------------------------
// Grid construction from dense data.

struct DenseGrid {
    int nx = 0, ny = 0, nz = 0;
    std::vector<float> v;
    float at(int x, int y, int z) const {
        if (x < 0 || y < 0 || z < 0 || x >= nx || y >= ny || z >= nz) return 0.f;
        return v[(size_t)x + (size_t)nx * ((size_t)y + (size_t)ny * (size_t)z)];
    }
};

// 2x box reduction; loop order and weights follow GL/gvdb_volume_gvdb.cpp:2834-2856
// (x outermost, weight 0.5^3 per tap, values kept only when > 0).
static inline DenseGrid downsample2(const DenseGrid& s) {
    DenseGrid d;
    d.nx = s.nx / 2; d.ny = s.ny / 2; d.nz = s.nz / 2;
    d.v.assign((size_t)d.nx * d.ny * d.nz, 0.f);
    for (int k = 0; k < d.nz; k++)
        for (int j = 0; j < d.ny; j++)
            for (int i = 0; i < d.nx; i++) {
                float res = 0.f;
                for (int a = 0; a < 2; a++)
                    for (int b = 0; b < 2; b++)
                        for (int c = 0; c < 2; c++) res += 0.125f * s.at(2 * i + a, 2 * j + b, 2 * k + c);
                d.v[(size_t)i + (size_t)d.nx * ((size_t)j + (size_t)d.ny * k)] = res > 0.f ? res : 0.f;
            }
    return d;
}

// Conservative mip 0: empty voxels take the mean of the positive values in
// their 27-neighbourhood (GL/gvdb_volume_gvdb.cpp:2753-2800).
static inline DenseGrid dilateConservative(const DenseGrid& s) {
    DenseGrid d = s;
    for (int k = 0; k < s.nz; k++)
        for (int j = 0; j < s.ny; j++)
            for (int i = 0; i < s.nx; i++) {
                if (s.at(i, j, k) != 0.f) continue;
                float avg = 0.f;
                for (int ii = -1; ii <= 1; ii++)
                    for (int jj = -1; jj <= 1; jj++)
                        for (int kk = -1; kk <= 1; kk++) {
                            int x = i + ii, y = j + jj, z = k + kk;
                            if (x < 0 || y < 0 || z < 0 || x >= s.nx || y >= s.ny || z >= s.nz) continue;
                            float tv = s.at(x, y, z);
                            avg += tv > 0.f ? tv : 0.f;
                        }
                avg /= 27.f;
                if (avg > 0.f) d.v[(size_t)i + (size_t)s.nx * ((size_t)j + (size_t)s.ny * k)] = avg;
            }
    return d;
}

float roundToHalf(float v);  // defined in oracle.cpp (IEEE binary16, round to nearest even)

// Build the tree + atlas of one slot (SC/Scene.cpp:2932-3160).
static inline void buildSlot(GridSlot& g, const DenseGrid& src, const float* indexToWorld0, int mipScale,
                             bool fp16) {
    g = GridSlot();
    g.valid = true;
    g.nx = src.nx; g.ny = src.ny; g.nz = src.nz;
    g.bmin = {0, 0, 0};
    g.bmax = {(float)src.nx, (float)src.ny, (float)src.nz};
    // index_k -> world = diag(2^k) * indexToWorld0 (row vectors)
    double M[16], S = (double)mipScale;
    for (int r = 0; r < 4; r++)
        for (int c = 0; c < 4; c++) M[r * 4 + c] = (r < 3 ? S : 1.0) * indexToWorld0[r * 4 + c];
    // affine inverse in double
    double a = M[0], b = M[1], c = M[2], d = M[4], e = M[5], f = M[6], gg = M[8], h = M[9], k = M[10];
    double det = a * (e * k - f * h) - b * (d * k - f * gg) + c * (d * h - e * gg);
    double I[9] = {(e * k - f * h) / det, -(b * k - c * h) / det, (b * f - c * e) / det,
                   -(d * k - f * gg) / det, (a * k - c * gg) / det, -(a * f - c * d) / det,
                   (d * h - e * gg) / det, -(a * h - b * gg) / det, (a * e - b * d) / det};
    double T[3] = {M[12], M[13], M[14]};
    for (int i = 0; i < 16; i++) g.xform[i] = (float)M[i];
    for (int r = 0; r < 3; r++) {
        for (int cc = 0; cc < 3; cc++) g.invxform[r * 4 + cc] = (float)I[r * 3 + cc];
        g.invxform[r * 4 + 3] = 0.f;
    }
    for (int cc = 0; cc < 3; cc++)
        g.invxform[12 + cc] = (float)(-(T[0] * I[0 * 3 + cc] + T[1] * I[1 * 3 + cc] + T[2] * I[2 * 3 + cc]));
    g.invxform[15] = 1.f;

    // stored values: clamp of tiny values (SC/Scene.cpp:3154-3156), then storage precision
    DenseGrid vals = src;
    float gmax = 0.f;
    for (float x : src.v) gmax = std::max(gmax, x);
    for (float& x : vals.v) {
        x = (gmax > 0.f && x / gmax < 1e-9f) ? 0.f : x;
        if (gmax <= 0.f) x = 0.f;
        if (fp16) x = roundToHalf(x);
    }
    float smax = 0.f;
    for (float x : vals.v) smax = std::max(smax, x);
    g.maxValue = smax;

    int nbx = src.nx / 8, nby = src.ny / 8, nbz = src.nz / 8;
    g.root.vmin = {0, 0, 0};
    g.root.childList = 0;
    g.childPool2.assign(32 * 32 * 32, kUndef);
    // allocate bricks in z,y,x order
    std::vector<int> brickIds;
    for (int bz = 0; bz < nbz; bz++)
        for (int by = 0; by < nby; by++)
            for (int bx = 0; bx < nbx; bx++) {
                bool any = false;
                for (int z = 0; z < 8 && !any; z++)
                    for (int y = 0; y < 8 && !any; y++)
                        for (int x = 0; x < 8 && !any; x++)
                            if (vals.at(bx * 8 + x, by * 8 + y, bz * 8 + z) > 0.f) any = true;
                if (!any) continue;
                int n1x = bx / 16, n1y = by / 16, n1z = bz / 16;
                int b2 = (((n1z << 5) + n1y) << 5) + n1x;
                if (g.childPool2[b2] == kUndef) {
                    InnerNode in;
                    in.vmin = {n1x * 128, n1y * 128, n1z * 128};
                    in.childList = (uint32_t)g.level1.size();
                    g.childPool2[b2] = (uint32_t)g.level1.size();
                    g.level1.push_back(in);
                    g.childPool1.resize(g.childPool1.size() + 4096, kUndef);
                }
                uint32_t n1 = g.childPool2[b2];
                int lx = bx % 16, ly = by % 16, lz = bz % 16;
                int b1 = (((lz << 4) + ly) << 4) + lx;
                LeafNode leaf;
                leaf.vmin = {bx * 8, by * 8, bz * 8};
                int id = (int)g.level0.size();
                leaf.offs = {0, 0, 0};  // placed below once the brick count is known
                g.childPool1[(size_t)g.level1[n1].childList * 4096 + b1] = (uint32_t)id;
                g.level0.push_back(leaf);
            }
    int nb = (int)g.level0.size();
    int side = 1;
    while (side * side * side < nb) side++;
    g.abx = side; g.aby = side; g.abz = std::max(1, (nb + side * side - 1) / (side * side));
    g.ax = 10 * g.abx; g.ay = 10 * g.aby; g.az = 10 * g.abz;
    g.atlas.assign((size_t)g.ax * g.ay * g.az, 0.f);
    for (int id = 0; id < nb; id++) {
        LeafNode& leaf = g.level0[id];
        int cx = id % g.abx, cy = (id / g.abx) % g.aby, cz = id / (g.abx * g.aby);
        leaf.offs = {cx * 10 + 1, cy * 10 + 1, cz * 10 + 1};
        float mn = FLT_MAX, mx = 0.f, sum = 0.f;
        // x outermost, as SC/Scene.cpp:2994-3009 (the sum order fixes the avg bits); avg = sum / 512
        for (int i = -1; i <= 8; i++)
            for (int j = -1; j <= 8; j++)
                for (int k = -1; k <= 8; k++) {
                    float v = vals.at(leaf.vmin.x + i, leaf.vmin.y + j, leaf.vmin.z + k);
                    size_t tx = (size_t)(leaf.offs.x + i), ty = (size_t)(leaf.offs.y + j), tz = (size_t)(leaf.offs.z + k);
                    g.atlas[tx + (size_t)g.ax * (ty + (size_t)g.ay * tz)] = v;
                    mn = std::min(mn, v);
                    mx = std::max(mx, v);
                    sum += v;
                }
        leaf.bounds = {mn, mx, sum / 512.f, 0.f};
    }
}

}  // namespace orc
