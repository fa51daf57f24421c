// vr_vbx.cpp -- the .vbx volume container (gvdb-voxel-src/GVDB_FILESPEC.txt),
// version 1.12: the reference's custom minor version, which adds the full
// index-to-world matrix, its inverse, the effective voxel extent and the value
// range after the 1.11 transform fields (GL/gvdb_volume_gvdb.cpp:596-605).
//
// The writer stores one float grid as GVDB topology type 2 with three levels
// (log2 dims 3, 4, 5: 8^3 bricks, 128^3 and 4096^3 nodes), 64-byte nodes, one
// child list of 64-bit pool ids per inner node, and an atlas of bricks with a
// one-voxel apron ("atlas layout"). The reader accepts 1.x and 2.x files of
// that form with any number of levels, walks the tree from the root through
// the child lists, and rebuilds the dense grid from each brick's atlas block.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/vrestir.h"
#include "vr_volume.h"

namespace {

// Pool ids: group (0 = nodes, 1 = child lists) | level << 8 | index << 16.
constexpr uint64_t kUndef64 = 0xFFFFFFFFFFFFFFFFull;
uint64_t poolId(int grp, int lev, uint64_t ndx) { return (uint64_t)grp | ((uint64_t)lev << 8) | (ndx << 16); }
int idLevel(uint64_t id) { return (int)((id >> 8) & 0xFF); }
uint64_t idIndex(uint64_t id) { return id >> 16; }

// One tree node as stored in pool 0: 64 bytes, the 64-bit ids at offsets 40, 48, 56.
#pragma pack(push, 1)
struct NodeRec {
    uint8_t lev, flags, priority, pad;
    int32_t pos[3];      // minimum corner in index space (voxels)
    int32_t value[3];    // bricks: atlas position of the first interior voxel; inner nodes: -1
    float vrange[3];     // min, max, mean of the values below
    uint64_t parent;     // pool id of the parent, kUndef64 at the root
    uint64_t childList;  // pool id of the child list, kUndef64 for bricks
    uint64_t mask;       // bitmask pointer (unused: 0)
};
#pragma pack(pop)
static_assert(sizeof(NodeRec) == 64, "node record");

constexpr int kLevels = 3;
constexpr int kLog2[kLevels] = {3, 4, 5};   // children per side: 8 voxels, 16 bricks, 32 blocks
constexpr int kRange[kLevels] = {8, 128, 4096};
constexpr int kApronW = 1;
constexpr int kT_FLOAT = 3;                 // GVDB channel type code of float

struct File {
    FILE* f = nullptr;
    ~File() { if (f) std::fclose(f); }
};
template <class T> bool put(FILE* f, const T& v) { return std::fwrite(&v, sizeof(T), 1, f) == 1; }
template <class T> bool put(FILE* f, const T* v, size_t n) { return std::fwrite(v, sizeof(T), n, f) == n; }
template <class T> bool get(FILE* f, T& v) { return std::fread(&v, sizeof(T), 1, f) == 1; }
template <class T> bool get(FILE* f, T* v, size_t n) { return std::fread(v, sizeof(T), n, f) == n; }

// 4x4 inverse of a row-vector affine matrix (last column 0,0,0,1).
bool invertAffine(const float* M, float* out) {
    double a[9] = {M[0], M[1], M[2], M[4], M[5], M[6], M[8], M[9], M[10]};
    double det = a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) +
                 a[2] * (a[3] * a[7] - a[4] * a[6]);
    if (det == 0.0 || !std::isfinite(det)) return false;
    double inv[9] = {(a[4] * a[8] - a[5] * a[7]) / det, (a[2] * a[7] - a[1] * a[8]) / det,
                     (a[1] * a[5] - a[2] * a[4]) / det, (a[5] * a[6] - a[3] * a[8]) / det,
                     (a[0] * a[8] - a[2] * a[6]) / det, (a[2] * a[3] - a[0] * a[5]) / det,
                     (a[3] * a[7] - a[4] * a[6]) / det, (a[1] * a[6] - a[0] * a[7]) / det,
                     (a[0] * a[4] - a[1] * a[3]) / det};
    double t[3] = {M[12], M[13], M[14]};
    for (int i = 0; i < 16; i++) out[i] = 0.f;
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) out[r * 4 + c] = (float)inv[r * 3 + c];
    for (int c = 0; c < 3; c++) out[12 + c] = (float)(-(t[0] * inv[c] + t[1] * inv[3 + c] + t[2] * inv[6 + c]));
    out[15] = 1.f;
    return true;
}

}  // namespace

extern "C" int32_t vrestir_write_vbx(const char* path, const float* density, int32_t nx, int32_t ny, int32_t nz,
                                     const float indexToWorld[16]) {
    if (!path || !density || nx <= 0 || ny <= 0 || nz <= 0 || !indexToWorld) return VRESTIR_EINVAL;
    if (nx > kRange[2] || ny > kRange[2] || nz > kRange[2]) return VRESTIR_EINVAL;
    float inv[16];
    if (!invertAffine(indexToWorld, inv)) return VRESTIR_EINVAL;
    auto at = [&](int x, int y, int z) -> float {
        if (x < 0 || y < 0 || z < 0 || x >= nx || y >= ny || z >= nz) return 0.f;
        return density[((size_t)z * ny + y) * nx + x];
    };
    const int bx = (nx + 7) / 8, by = (ny + 7) / 8, bz = (nz + 7) / 8;

    // level 0: bricks with any non-zero voxel, in x-fastest brick order
    std::vector<NodeRec> nodes[kLevels];
    std::vector<std::vector<uint64_t>> lists[kLevels];
    std::map<long long, int> level1;  // 128^3 block key -> level-1 node index
    float gmin = INFINITY, gmax = -INFINITY;
    for (size_t i = 0, n = (size_t)nx * ny * nz; i < n; i++) {
        gmin = std::fmin(gmin, density[i]);
        gmax = std::fmax(gmax, density[i]);
    }
    struct BrickRef { int x, y, z; };
    std::vector<BrickRef> bricks;
    for (int k = 0; k < bz; k++)
        for (int j = 0; j < by; j++)
            for (int i = 0; i < bx; i++) {
                bool any = false;
                for (int z = 0; z < 8 && !any; z++)
                    for (int y = 0; y < 8 && !any; y++)
                        for (int x = 0; x < 8 && !any; x++) any = at(i * 8 + x, j * 8 + y, k * 8 + z) != 0.f;
                if (any) bricks.push_back({i, j, k});
            }
    const int nb = (int)bricks.size();
    // atlas: bricks of 10^3 in an (ax, ay, az) arrangement
    int ax = 1;
    while ((long long)ax * ax * ax < nb) ax++;
    const int ay = ax, az = nb > 0 ? (nb + ax * ay - 1) / (ax * ay) : 1;
    const int bw = 8 + 2 * kApronW;
    const int rx = ax * bw, ry = ay * bw, rz = az * bw;
    std::vector<float> atlas((size_t)rx * ry * rz, 0.f);

    // root
    NodeRec root;
    std::memset(&root, 0, sizeof(root));
    root.lev = 2; root.flags = 1;
    for (int c = 0; c < 3; c++) root.value[c] = -1;
    root.vrange[0] = gmin; root.vrange[1] = gmax; root.vrange[2] = 0.f;
    root.parent = kUndef64;
    root.childList = poolId(1, 2, 0);
    nodes[2].push_back(root);
    lists[2].assign(1, std::vector<uint64_t>((size_t)1 << (3 * kLog2[2]), kUndef64));

    for (int b = 0; b < nb; b++) {
        const BrickRef& r = bricks[b];
        const int px = r.x * 8, py = r.y * 8, pz = r.z * 8;
        // level-1 node of this brick
        const int cx = px / kRange[1], cy = py / kRange[1], cz = pz / kRange[1];
        const long long key = ((long long)cz * 64 + cy) * 64 + cx;
        auto it = level1.find(key);
        int n1;
        if (it == level1.end()) {
            n1 = (int)nodes[1].size();
            level1[key] = n1;
            NodeRec m;
            std::memset(&m, 0, sizeof(m));
            m.lev = 1; m.flags = 1;
            m.pos[0] = cx * kRange[1]; m.pos[1] = cy * kRange[1]; m.pos[2] = cz * kRange[1];
            for (int c = 0; c < 3; c++) m.value[c] = -1;
            m.vrange[0] = INFINITY; m.vrange[1] = -INFINITY; m.vrange[2] = 0.f;
            m.parent = poolId(0, 2, 0);
            m.childList = poolId(1, 1, (uint64_t)n1);
            m.mask = 0;
            nodes[1].push_back(m);
            lists[1].push_back(std::vector<uint64_t>((size_t)1 << (3 * kLog2[1]), kUndef64));
            const int c2 = (cz * 32 + cy) * 32 + cx;  // slot in the root's 32^3 list
            lists[2][0][c2] = poolId(0, 1, (uint64_t)n1);
        } else {
            n1 = it->second;
        }
        // the brick node
        const int sx = b % ax, sy = (b / ax) % ay, sz = b / (ax * ay);
        NodeRec br;
        std::memset(&br, 0, sizeof(br));
        br.lev = 0; br.flags = 1;
        br.pos[0] = px; br.pos[1] = py; br.pos[2] = pz;
        br.value[0] = sx * bw + kApronW; br.value[1] = sy * bw + kApronW; br.value[2] = sz * bw + kApronW;
        float mn = INFINITY, mx = -INFINITY;
        double sum = 0.0;
        for (int z = -kApronW; z < 8 + kApronW; z++)
            for (int y = -kApronW; y < 8 + kApronW; y++)
                for (int x = -kApronW; x < 8 + kApronW; x++) {
                    float v = at(px + x, py + y, pz + z);
                    atlas[((size_t)(br.value[2] + z) * ry + (br.value[1] + y)) * rx + (br.value[0] + x)] = v;
                    if (x >= 0 && y >= 0 && z >= 0 && x < 8 && y < 8 && z < 8) {
                        mn = std::fmin(mn, v); mx = std::fmax(mx, v); sum += v;
                    }
                }
        br.vrange[0] = mn; br.vrange[1] = mx; br.vrange[2] = (float)(sum / 512.0);
        br.parent = poolId(0, 1, (uint64_t)n1);
        br.childList = kUndef64;
        nodes[0].push_back(br);
        NodeRec& m = nodes[1][n1];
        m.vrange[0] = std::fmin(m.vrange[0], mn);
        m.vrange[1] = std::fmax(m.vrange[1], mx);
        const int lx = (px - m.pos[0]) / 8, ly = (py - m.pos[1]) / 8, lz = (pz - m.pos[2]) / 8;
        lists[1][n1][(lz * 16 + ly) * 16 + lx] = poolId(0, 0, (uint64_t)b);
    }

    File fh;
    fh.f = std::fopen(path, "wb");
    if (!fh.f) return VRESTIR_EIO;
    FILE* f = fh.f;
    bool ok = true;
    const uint8_t major = 1, minor = 12;
    ok &= put(f, major) && put(f, minor);
    // 1.11 transform: pre-translation, Euler angles, scale, translation. The
    // matrix fields below carry the exact transform; these describe its
    // axis-aligned part (scale = row lengths, angles 0).
    float pretrans[3] = {0, 0, 0}, angs[3] = {0, 0, 0}, scale[3], trans[3];
    for (int r = 0; r < 3; r++) {
        const float* row = indexToWorld + 4 * r;
        scale[r] = std::sqrt(row[0] * row[0] + row[1] * row[1] + row[2] * row[2]);
        trans[r] = indexToWorld[12 + r];
    }
    ok &= put(f, pretrans, 3) && put(f, angs, 3) && put(f, scale, 3) && put(f, trans, 3);
    ok &= put(f, indexToWorld, 16) && put(f, inv, 16);
    const int32_t vmin[3] = {0, 0, 0}, vmax[3] = {nx, ny, nz};  // effective extent, exclusive max
    ok &= put(f, vmin, 3) && put(f, vmax, 3);
    ok &= put(f, gmin) && put(f, gmax);
    const int32_t numGrids = 1;
    ok &= put(f, numGrids);
    const long tablePos = std::ftell(f);
    uint64_t gridOff = 0;
    ok &= put(f, gridOff);
    gridOff = (uint64_t)std::ftell(f);

    char name[256];
    std::memset(name, 0, sizeof(name));
    std::strncpy(name, "density", sizeof(name) - 1);
    ok &= put(f, name, 256);
    const uint8_t dtype = 'f', comps = 1, compress = 0, topo = 2, layout = 0;
    const float voxelSize[3] = {1.f, 1.f, 1.f};
    const int32_t leafDim[3] = {8, 8, 8}, apron = kApronW, numChan = 1, reuse = 0;
    const int32_t axisCnt[3] = {ax, ay, az}, axisRes[3] = {rx, ry, rz};
    const uint64_t atlasBytes = (uint64_t)atlas.size() * sizeof(float);
    ok &= put(f, dtype) && put(f, comps) && put(f, compress) && put(f, voxelSize, 3);
    ok &= put(f, nb) && put(f, leafDim, 3) && put(f, apron) && put(f, numChan) && put(f, atlasBytes);
    ok &= put(f, topo) && put(f, reuse) && put(f, layout) && put(f, axisCnt, 3) && put(f, axisRes, 3);
    // topology
    const int32_t levels = kLevels;
    const uint64_t rootId = poolId(0, 2, 0);
    ok &= put(f, levels) && put(f, rootId);
    for (int n = 0; n < kLevels; n++) {
        const int32_t ld = kLog2[n], res = 1 << kLog2[n];
        const int32_t range[3] = {kRange[n], kRange[n], kRange[n]};
        const int32_t cnt0 = (int32_t)nodes[n].size(), width0 = (int32_t)sizeof(NodeRec);
        const int32_t cnt1 = (int32_t)lists[n].size(), width1 = (int32_t)(sizeof(uint64_t) << (3 * kLog2[n]));
        ok &= put(f, ld) && put(f, res) && put(f, range, 3) && put(f, cnt0) && put(f, width0) && put(f, cnt1) &&
              put(f, width1);
    }
    for (int n = 0; n < kLevels; n++)
        if (!nodes[n].empty()) ok &= put(f, nodes[n].data(), nodes[n].size());
    for (int n = 0; n < kLevels; n++)
        for (const auto& l : lists[n]) ok &= put(f, l.data(), l.size());
    // atlas: channel type, bytes per voxel, then z slices of x-fastest rows
    const int32_t chanType = kT_FLOAT, chanStride = (int32_t)sizeof(float);
    ok &= put(f, chanType) && put(f, chanStride) && put(f, atlas.data(), atlas.size());
    ok &= std::fseek(f, tablePos, SEEK_SET) == 0 && put(f, gridOff);
    if (!ok) return VRESTIR_EIO;
    return VRESTIR_OK;
}

namespace {

// A parsed file: the dense extent, index -> world, the bricks reachable from the
// root and (when asked for) the atlas.
struct Parsed {
    int ex = 0, ey = 0, ez = 0, bd = 8;
    float M[16];
    int32_t axisRes[3] = {0, 0, 0};
    std::vector<NodeRec> bricks;
    std::vector<float> atlas;
};

int32_t parseVbx(const char* path, bool wantAtlas, Parsed& P) {
    File fh;
    fh.f = std::fopen(path, "rb");
    if (!fh.f) return VRESTIR_EIO;
    FILE* f = fh.f;
    uint8_t major = 0, minor = 0;
    if (!get(f, major) || !get(f, minor)) return VRESTIR_EIO;
    float M[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    int32_t vmin[3] = {0, 0, 0}, vmax[3] = {-1, -1, -1};
    bool haveExtent = false;
    if ((major == 1 && minor >= 11) || major > 1) {
        float pretrans[3], angs[3], scale[3], trans[3];
        if (!get(f, pretrans, 3) || !get(f, angs, 3) || !get(f, scale, 3) || !get(f, trans, 3)) return VRESTIR_EIO;
        // GVDB order: pre-translate, scale, rotate about x then y then z, translate
        const double d2r = 3.14159265358979323846 / 180.0;
        const double cx = std::cos(angs[0] * d2r), sx = std::sin(angs[0] * d2r);
        const double cy = std::cos(angs[1] * d2r), sy = std::sin(angs[1] * d2r);
        const double cz = std::cos(angs[2] * d2r), sz = std::sin(angs[2] * d2r);
        // row-vector rotation R = Rx * Ry * Rz (applied x first)
        const double Rx[9] = {1, 0, 0, 0, cx, sx, 0, -sx, cx};
        const double Ry[9] = {cy, 0, -sy, 0, 1, 0, sy, 0, cy};
        const double Rz[9] = {cz, sz, 0, -sz, cz, 0, 0, 0, 1};
        double T[9], R[9];
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 3; c++) {
                T[r * 3 + c] = Rx[r * 3 + 0] * Ry[0 * 3 + c] + Rx[r * 3 + 1] * Ry[1 * 3 + c] + Rx[r * 3 + 2] * Ry[2 * 3 + c];
            }
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 3; c++)
                R[r * 3 + c] = T[r * 3 + 0] * Rz[0 * 3 + c] + T[r * 3 + 1] * Rz[1 * 3 + c] + T[r * 3 + 2] * Rz[2 * 3 + c];
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 3; c++) M[r * 4 + c] = (float)(scale[r] * R[r * 3 + c]);
        for (int c = 0; c < 3; c++) {
            double t = trans[c];
            for (int r = 0; r < 3; r++) t += pretrans[r] * scale[r] * R[r * 3 + c];
            M[12 + c] = (float)t;
        }
        if (minor == 12) {
            float inv[16];
            float gmin, gmax;
            if (!get(f, M, 16) || !get(f, inv, 16) || !get(f, vmin, 3) || !get(f, vmax, 3) || !get(f, gmin) ||
                !get(f, gmax))
                return VRESTIR_EIO;
            haveExtent = true;
        }
    }
    int32_t numGrids = 0;
    if (!get(f, numGrids) || numGrids < 1) return VRESTIR_EIO;
    if (major >= 2) {
        uint8_t masks;
        if (!get(f, masks)) return VRESTIR_EIO;
    }
    std::vector<uint64_t> offs((size_t)numGrids);
    if (!get(f, offs.data(), offs.size())) return VRESTIR_EIO;
    if (std::fseek(f, (long)offs[0], SEEK_SET) != 0) return VRESTIR_EIO;

    char name[256];
    uint8_t dtype, comps, compress, topo, layout;
    float voxelSize[3];
    int32_t leafCnt, leafDim[3], apron, numChan, reuse, axisCnt[3], axisRes[3];
    uint64_t atlasBytes;
    if (!get(f, name, 256) || !get(f, dtype) || !get(f, comps) || !get(f, compress) || !get(f, voxelSize, 3) ||
        !get(f, leafCnt) || !get(f, leafDim, 3) || !get(f, apron) || !get(f, numChan) || !get(f, atlasBytes) ||
        !get(f, topo) || !get(f, reuse) || !get(f, layout) || !get(f, axisCnt, 3) || !get(f, axisRes, 3))
        return VRESTIR_EIO;
    if (dtype != 'f' || comps != 1 || compress != 0 || topo != 2 || layout != 0 || numChan < 1)
        return VRESTIR_EUNSUPPORTED;
    int32_t levels;
    uint64_t rootId;
    if (!get(f, levels) || !get(f, rootId) || levels < 1 || levels > 8) return VRESTIR_EIO;
    std::vector<int32_t> ld(levels), res(levels), cnt0(levels), width0(levels), cnt1(levels), width1(levels);
    for (int n = 0; n < levels; n++) {
        int32_t range[3];
        if (!get(f, ld[n]) || !get(f, res[n]) || !get(f, range, 3) || !get(f, cnt0[n]) || !get(f, width0[n]) ||
            !get(f, cnt1[n]) || !get(f, width1[n]))
            return VRESTIR_EIO;
        if (width0[n] < (int32_t)sizeof(NodeRec) || cnt0[n] < 0 || cnt1[n] < 0 || width1[n] < 0 || ld[n] < 1 ||
            ld[n] > 10)
            return VRESTIR_EIO;
    }
    if (leafDim[0] != (1 << ld[0]) || leafDim[1] != leafDim[0] || leafDim[2] != leafDim[0]) return VRESTIR_EUNSUPPORTED;
    std::vector<std::vector<uint8_t>> pool0(levels), pool1(levels);
    for (int n = 0; n < levels; n++) {
        pool0[n].resize((size_t)cnt0[n] * width0[n]);
        if (!pool0[n].empty() && !get(f, pool0[n].data(), pool0[n].size())) return VRESTIR_EIO;
    }
    for (int n = 0; n < levels; n++) {
        pool1[n].resize((size_t)cnt1[n] * width1[n]);
        if (!pool1[n].empty() && !get(f, pool1[n].data(), pool1[n].size())) return VRESTIR_EIO;
    }
    int32_t chanType, chanStride;
    if (!get(f, chanType) || !get(f, chanStride)) return VRESTIR_EIO;
    if (chanType != kT_FLOAT || chanStride != 4) return VRESTIR_EUNSUPPORTED;
    const size_t atlasVox = (size_t)axisRes[0] * axisRes[1] * axisRes[2];
    if (axisRes[0] <= 0 || axisRes[1] <= 0 || axisRes[2] <= 0) return VRESTIR_EIO;

    auto node = [&](uint64_t id, NodeRec& out) -> bool {
        const int lev = idLevel(id);
        const uint64_t ndx = idIndex(id);
        if ((id & 0xFF) != 0 || lev >= levels || ndx >= (uint64_t)cnt0[lev]) return false;
        std::memcpy(&out, pool0[lev].data() + ndx * width0[lev], sizeof(NodeRec));
        return true;
    };
    // walk the tree: collect bricks (level-0 nodes reachable from the root)
    std::vector<NodeRec> bricks;
    std::vector<uint64_t> stack{rootId};
    while (!stack.empty()) {
        uint64_t id = stack.back();
        stack.pop_back();
        NodeRec nd;
        if (!node(id, nd)) return VRESTIR_EIO;
        if (idLevel(id) == 0) { bricks.push_back(nd); continue; }
        const uint64_t cl = nd.childList;
        const int lev = idLevel(cl);
        if ((cl & 0xFF) != 1 || lev != idLevel(id) || idIndex(cl) >= (uint64_t)cnt1[lev]) return VRESTIR_EIO;
        const size_t nChild = (size_t)1 << (3 * ld[lev]);
        if ((size_t)width1[lev] < nChild * 8) return VRESTIR_EIO;
        const uint8_t* lst = pool1[lev].data() + idIndex(cl) * width1[lev];
        for (size_t c = 0; c < nChild; c++) {
            uint64_t ch;
            std::memcpy(&ch, lst + 8 * c, 8);
            if (ch == kUndef64) continue;
            if (idLevel(ch) != lev - 1) return VRESTIR_EIO;
            stack.push_back(ch);
        }
        if (stack.size() > ((size_t)1 << 28)) return VRESTIR_EIO;
    }
    // dense extent: the effective voxel box when stored, else the bricks' extent
    const int bd = leafDim[0];
    int ex = 0, ey = 0, ez = 0;
    for (const NodeRec& b : bricks) {
        ex = std::max(ex, b.pos[0] + bd); ey = std::max(ey, b.pos[1] + bd); ez = std::max(ez, b.pos[2] + bd);
    }
    if (haveExtent && vmax[0] > 0 && vmax[1] > 0 && vmax[2] > 0) { ex = vmax[0]; ey = vmax[1]; ez = vmax[2]; }
    P.ex = ex; P.ey = ey; P.ez = ez; P.bd = bd;
    std::memcpy(P.M, M, sizeof(M));
    std::memcpy(P.axisRes, axisRes, sizeof(P.axisRes));
    for (const NodeRec& b : bricks)
        for (int a = 0; a < 3; a++)
            if (b.value[a] < 0 || b.value[a] + bd > axisRes[a]) return VRESTIR_EIO;
    P.bricks = std::move(bricks);
    if (!wantAtlas) return VRESTIR_OK;
    P.atlas.resize(atlasVox);
    if (!get(f, P.atlas.data(), P.atlas.size())) return VRESTIR_EIO;
    return VRESTIR_OK;
}

// Value of voxel (x, y, z) of brick b (0 outside the dense extent).
float brickVoxel(const Parsed& P, const NodeRec& b, int x, int y, int z) {
    const int gx = b.pos[0] + x, gy = b.pos[1] + y, gz = b.pos[2] + z;
    if (gx < 0 || gy < 0 || gz < 0 || gx >= P.ex || gy >= P.ey || gz >= P.ez) return 0.f;
    const int tx = b.value[0] + x, ty = b.value[1] + y, tz = b.value[2] + z;
    return P.atlas[((size_t)tz * P.axisRes[1] + ty) * P.axisRes[0] + tx];
}

}  // namespace

extern "C" int32_t vrestir_read_vbx(const char* path, float* densityOrNull, int32_t* nxOut, int32_t* nyOut,
                                    int32_t* nzOut, float indexToWorld[16]) {
    if (!path || !nxOut || !nyOut || !nzOut) return VRESTIR_EINVAL;
    Parsed P;
    int32_t rc = parseVbx(path, densityOrNull != nullptr, P);
    if (rc) return rc;
    *nxOut = P.ex; *nyOut = P.ey; *nzOut = P.ez;
    if (indexToWorld) std::memcpy(indexToWorld, P.M, sizeof(P.M));
    if (!densityOrNull) return VRESTIR_OK;
    std::memset(densityOrNull, 0, sizeof(float) * (size_t)P.ex * P.ey * P.ez);
    for (const NodeRec& b : P.bricks)
        for (int z = 0; z < P.bd; z++)
            for (int y = 0; y < P.bd; y++)
                for (int x = 0; x < P.bd; x++) {
                    const int gx = b.pos[0] + x, gy = b.pos[1] + y, gz = b.pos[2] + z;
                    if (gx < 0 || gy < 0 || gz < 0 || gx >= P.ex || gy >= P.ey || gz >= P.ez) continue;
                    densityOrNull[((size_t)gz * P.ey + gy) * P.ex + gx] = brickVoxel(P, b, x, y, z);
                }
    return VRESTIR_OK;
}

namespace vr {

// The bricks of a file without a dense grid: each brick's minimum (index space)
// and its bd^3 values (x fastest; 0 past the dense extent).
int32_t readVbxBricks(const char* path, VbxBricks& out) {
    Parsed P;
    int32_t rc = parseVbx(path, true, P);
    if (rc) return rc;
    out.nx = P.ex; out.ny = P.ey; out.nz = P.ez; out.brickDim = P.bd;
    std::memcpy(out.indexToWorld, P.M, sizeof(P.M));
    const size_t nv = (size_t)P.bd * P.bd * P.bd;
    out.pos.resize(P.bricks.size() * 3);
    out.vals.resize(P.bricks.size() * nv);
    for (size_t i = 0; i < P.bricks.size(); i++) {
        const NodeRec& b = P.bricks[i];
        for (int a = 0; a < 3; a++) out.pos[3 * i + a] = b.pos[a];
        float* v = out.vals.data() + i * nv;
        for (int z = 0; z < P.bd; z++)
            for (int y = 0; y < P.bd; y++)
                for (int x = 0; x < P.bd; x++) v[(z * P.bd + y) * P.bd + x] = brickVoxel(P, b, x, y, z);
    }
    return VRESTIR_OK;
}

}  // namespace vr
