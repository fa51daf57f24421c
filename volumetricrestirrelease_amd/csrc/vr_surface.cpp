// vr_surface.cpp -- the scene's triangles on the host: the emissive light
// collection and the power sampler's alias table.
#include "vr_surface.h"

#include <algorithm>
#include <cmath>
#include <cstring>

#include "vr_types.h"

namespace vr {

std::string checkSurfacesDesc(const vrestir_surfaces_desc& d) {
    if (d.numVertices < 0 || d.numTriangles < 0 || d.numMaterials < 0) return "negative count";
    if (d.numTriangles > 0 && (!d.positions || !d.indices || !d.materialIds || !d.materials || d.numMaterials == 0))
        return "triangles need positions, indices, material ids and materials";
    for (int64_t i = 0; i < 3 * (int64_t)d.numTriangles; i++)
        if (d.indices[i] < 0 || d.indices[i] >= d.numVertices) return "vertex index out of range";
    for (int i = 0; i < d.numTriangles; i++)
        if (d.materialIds[i] < 0 || d.materialIds[i] >= d.numMaterials) return "material id out of range";
    for (int64_t i = 0; i < 3 * (int64_t)d.numVertices; i++)
        if (!std::isfinite(d.positions[i])) return "non-finite vertex position";
    if (d.normals)
        for (int64_t i = 0; i < 3 * (int64_t)d.numVertices; i++)
            if (!std::isfinite(d.normals[i])) return "non-finite vertex normal";
    for (int m = 0; m < d.numMaterials; m++)
        for (int k = 0; k < 3; k++)
            if (!(d.materials[m].emissive[k] >= 0.f) || !std::isfinite(d.materials[m].emissive[k]))
                return "emissive colours must be finite and non-negative";
    if (!(d.emissiveIntensityMultiplier >= 0.f) || !std::isfinite(d.emissiveIntensityMultiplier))
        return "emissive intensity multiplier must be finite and non-negative";
    return "";
}

namespace {

struct F3 { float x, y, z; };
F3 sub(F3 a, F3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
float dotf(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
F3 crossf(F3 a, F3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
F3 normalizef(F3 a) {
    const float l = std::sqrt(dotf(a, a));
    return {a.x / l, a.y / l, a.z / l};
}
// floatToSnorm16 (UT/Math/FormatConversion.slang:48-52)
int snorm16(float v) {
    v = std::isnan(v) ? 0.f : std::min(std::max(v, -1.f), 1.f);
    return (int)std::trunc(v * 32767.f + (v >= 0.f ? 0.5f : -0.5f));
}
float unsnorm16(int bits) { return std::max((float)bits / 32767.f, -1.f); }
float signNotNeg(float v) { return v >= 0.f ? 1.f : -1.f; }
// The face normal as the collection stores it: encodeNormal2x16 then
// decodeNormal2x16 (UT/Math/PackedFormats.slang:34-46), octahedral
// (UT/Math/MathHelpers.slang:128-163) in two 16-bit snorms.
F3 storedNormal(F3 n) {
    const float s = 1.f / (std::fabs(n.x) + std::fabs(n.y) + std::fabs(n.z));
    float px = n.x * s, py = n.y * s;
    if (n.z < 0.f) {
        const float wx = (1.f - std::fabs(py)) * signNotNeg(px), wy = (1.f - std::fabs(px)) * signNotNeg(py);
        px = wx; py = wy;
    }
    float qx = unsnorm16((int16_t)snorm16(px)), qy = unsnorm16((int16_t)snorm16(py));
    F3 m{qx, qy, 1.f - std::fabs(qx) - std::fabs(qy)};
    if (m.z < 0.f) {
        const float wx = (1.f - std::fabs(m.y)) * signNotNeg(m.x), wy = (1.f - std::fabs(m.x)) * signNotNeg(m.y);
        m.x = wx; m.y = wy;
    }
    return normalizef(m);
}

}  // namespace

void buildLightCollection(const vrestir_surfaces_desc& d, LightCollectionHost& out) {
    out = LightCollectionHost();
    auto pos = [&](int v) { return F3{d.positions[3 * v], d.positions[3 * v + 1], d.positions[3 * v + 2]}; };
    for (int t = 0; t < d.numTriangles; t++) {
        const vrestir_material& m = d.materials[d.materialIds[t]];
        if (!(m.emissive[0] > 0.f || m.emissive[1] > 0.f || m.emissive[2] > 0.f)) continue;
        const F3 p0 = pos(d.indices[3 * t]), p1 = pos(d.indices[3 * t + 1]), p2 = pos(d.indices[3 * t + 2]);
        // computeFaceNormalAndAreaW (SC/Scene.slang:343-358)
        const F3 N = crossf(sub(p1, p0), sub(p2, p0));
        const float area = 0.5f * std::sqrt(dotf(N, N));
        // a degenerate triangle keeps a zero normal: it has no flux and, read
        // through a stale light id, faces nothing
        const F3 n = area > 0.f ? storedNormal(normalizef(N)) : F3{0.f, 0.f, 0.f};
        const float e[16] = {p0.x, p0.y, p0.z, p1.x, p1.y, p1.z, p2.x, p2.y, p2.z, n.x, n.y, n.z,
                             area, m.emissive[0], m.emissive[1], m.emissive[2]};
        out.tri.insert(out.tri.end(), e, e + kEmTriFloats);
        // FinalizeIntegration.cs.slang:68-76: flux = luminance(radiance) * area * pi
        const float lum = m.emissive[0] * 0.2126f + m.emissive[1] * 0.7152f + m.emissive[2] * 0.0722f;
        const float flux = lum * area * 3.14159265358979323846f;
        out.weight.push_back(flux);
        if (flux > 0.f) out.numActive++;
    }
    const uint32_t count = (uint32_t)out.weight.size();
    if (count == 0) return;
    // AliasTable (UT/Sampling/AliasTable.cpp:46-125): weights scaled to mean 1, sorted
    // ascending, the smallest paired with the largest; each slot keeps its own index
    // (no shuffle of the slots, which changes no probability)
    std::vector<float> w = out.weight;
    for (float f : w) out.weightSum += f;
    if (!(out.weightSum > 0.0)) {
        out.alias.assign(4 * (size_t)count, 0u);
        return;
    }
    const double factor = count / out.weightSum;
    for (float& f : w) f = (float)(f * factor);
    std::vector<uint32_t> perm(count);
    for (uint32_t i = 0; i < count; i++) perm[i] = i;
    std::stable_sort(perm.begin(), perm.end(), [&](uint32_t a, uint32_t b) { return w[a] < w[b]; });
    std::vector<float> threshold(count, 1.f);
    std::vector<uint32_t> redirect(count);
    for (uint32_t i = 0; i < count; i++) redirect[i] = i;
    uint32_t head = 0, tail = count - 1;
    while (head != tail) {
        const uint32_t i = perm[head], j = perm[tail];
        threshold[i] = w[i];
        redirect[i] = j;
        w[j] -= 1.f - w[i];
        if (head == tail - 1) {
            threshold[j] = 1.f;
            redirect[j] = j;
            break;
        } else if (w[j] < 1.f) {
            std::swap(perm[head], perm[tail]);
            tail--;
        } else {
            head++;
        }
    }
    out.alias.resize(4 * (size_t)count);
    for (uint32_t i = 0; i < count; i++) {
        uint32_t bits;
        std::memcpy(&bits, &threshold[i], 4);
        out.alias[4 * i] = bits;
        out.alias[4 * i + 1] = redirect[i];
        out.alias[4 * i + 2] = i;
        out.alias[4 * i + 3] = 0u;
    }
}

namespace {

struct Box {
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    void grow(const float* p) {
        for (int k = 0; k < 3; k++) { lo[k] = std::min(lo[k], p[k]); hi[k] = std::max(hi[k], p[k]); }
    }
    void grow(const Box& b) {
        for (int k = 0; k < 3; k++) { lo[k] = std::min(lo[k], b.lo[k]); hi[k] = std::max(hi[k], b.hi[k]); }
    }
    float area() const {
        if (lo[0] > hi[0]) return 0.f;
        const float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
        return 2.f * (dx * dy + dy * dz + dz * dx);
    }
};
float asFloatBits(int32_t v) {
    float f;
    std::memcpy(&f, &v, 4);
    return f;
}

}  // namespace

void buildSurfaceGeometry(const vrestir_surfaces_desc& d, SurfaceGeometryHost& out) {
    out = SurfaceGeometryHost();
    const int nt = d.numTriangles;
    for (int m = 0; m < d.numMaterials; m++) {
        const vrestir_material& M = d.materials[m];
        const float e[kMaterialFloats] = {M.baseColor[0], M.baseColor[1], M.baseColor[2], M.roughness, M.metallic,
                                          M.IoR, M.emissive[0], M.emissive[1], M.emissive[2],
                                          (float)(M.doubleSided != 0), 0.f, 0.f};
        out.materials.insert(out.materials.end(), e, e + kMaterialFloats);
    }
    if (nt == 0) return;
    auto P = [&](int t, int k) { return d.positions + 3 * (size_t)d.indices[3 * t + k]; };
    // per scene triangle: positions, shading normals (the vertices' own, else the
    // face normal at each vertex), material index
    out.triData.assign((size_t)kTriDataFloats * nt, 0.f);
    for (int t = 0; t < nt; t++) {
        float* o = &out.triData[(size_t)kTriDataFloats * t];
        for (int k = 0; k < 3; k++)
            for (int c = 0; c < 3; c++) o[3 * k + c] = P(t, k)[c];
        if (d.normals) {
            for (int k = 0; k < 3; k++)
                for (int c = 0; c < 3; c++) o[9 + 3 * k + c] = d.normals[3 * (size_t)d.indices[3 * t + k] + c];
        } else {
            const F3 p0{o[0], o[1], o[2]}, p1{o[3], o[4], o[5]}, p2{o[6], o[7], o[8]};
            const F3 N = crossf(sub(p1, p0), sub(p2, p0));
            const F3 n = dotf(N, N) > 0.f ? normalizef(N) : F3{0.f, 0.f, 1.f};
            for (int k = 0; k < 3; k++) { o[9 + 3 * k] = n.x; o[10 + 3 * k] = n.y; o[11 + 3 * k] = n.z; }
        }
        o[18] = asFloatBits(d.materialIds[t]);
    }
    // per-triangle boxes and centroids
    std::vector<Box> tb(nt);
    std::vector<float> cen(3 * (size_t)nt);
    for (int t = 0; t < nt; t++) {
        for (int k = 0; k < 3; k++) tb[t].grow(P(t, k));
        for (int c = 0; c < 3; c++) cen[3 * t + c] = 0.5f * (tb[t].lo[c] + tb[t].hi[c]);
    }
    std::vector<int> order(nt);
    for (int t = 0; t < nt; t++) order[t] = t;
    struct Node { Box b; int first = 0, count = 0; };
    std::vector<Node> nodes(1);
    struct Item { int node, begin, end, depth; };
    std::vector<Item> work{{0, 0, nt, 0}};
    constexpr int kBins = 16, kLeaf = 4;
    while (!work.empty()) {
        const Item it = work.back();
        work.pop_back();
        Box b, cb;
        for (int i = it.begin; i < it.end; i++) {
            b.grow(tb[order[i]]);
            cb.grow(&cen[3 * (size_t)order[i]]);
        }
        nodes[it.node].b = b;
        out.depth = std::max(out.depth, it.depth);
        const int n = it.end - it.begin;
        auto makeLeaf = [&]() { nodes[it.node].first = it.begin; nodes[it.node].count = n; };
        if (n <= kLeaf || it.depth >= kBvhMaxDepth) { makeLeaf(); continue; }
        // binned SAH over the widest centroid axis
        int axis = 0;
        for (int k = 1; k < 3; k++)
            if (cb.hi[k] - cb.lo[k] > cb.hi[axis] - cb.lo[axis]) axis = k;
        const float ext = cb.hi[axis] - cb.lo[axis];
        int mid = it.begin + n / 2;
        if (ext > 0.f) {
            Box bins[kBins];
            int cnt[kBins] = {0};
            auto binOf = [&](int t) {
                const int k = (int)((cen[3 * (size_t)t + axis] - cb.lo[axis]) / ext * kBins);
                return std::min(kBins - 1, std::max(0, k));
            };
            for (int i = it.begin; i < it.end; i++) {
                const int k = binOf(order[i]);
                bins[k].grow(tb[order[i]]);
                cnt[k]++;
            }
            float best = INFINITY;
            int split = -1;
            for (int s = 1; s < kBins; s++) {
                Box l, r;
                int nl = 0, nr = 0;
                for (int k = 0; k < s; k++) { l.grow(bins[k]); nl += cnt[k]; }
                for (int k = s; k < kBins; k++) { r.grow(bins[k]); nr += cnt[k]; }
                if (nl == 0 || nr == 0) continue;
                const float c = l.area() * nl + r.area() * nr;
                if (c < best) { best = c; split = s; }
            }
            if (split > 0) {
                auto pm = std::stable_partition(order.begin() + it.begin, order.begin() + it.end,
                                                [&](int t) { return binOf(t) < split; });
                mid = (int)(pm - order.begin());
            } else {
                std::stable_sort(order.begin() + it.begin, order.begin() + it.end,
                                 [&](int a, int c) { return cen[3 * (size_t)a + axis] < cen[3 * (size_t)c + axis]; });
            }
        } else if (n <= 16) {
            makeLeaf();
            continue;
        }
        const int left = (int)nodes.size();
        nodes.resize(nodes.size() + 2);
        nodes[it.node].first = left;
        nodes[it.node].count = 0;
        work.push_back({left + 1, mid, it.end, it.depth + 1});
        work.push_back({left, it.begin, mid, it.depth + 1});
    }
    out.nodes.resize(8 * nodes.size());
    for (size_t i = 0; i < nodes.size(); i++) {
        float* o = &out.nodes[8 * i];
        for (int k = 0; k < 3; k++) { o[k] = nodes[i].b.lo[k]; o[4 + k] = nodes[i].b.hi[k]; }
        o[3] = asFloatBits(nodes[i].first);
        o[7] = asFloatBits(nodes[i].count);
    }
    out.tris.resize((size_t)kBvhTriFloats * nt);
    for (int i = 0; i < nt; i++) {
        const int t = order[i];
        float* o = &out.tris[(size_t)kBvhTriFloats * i];
        for (int k = 0; k < 3; k++)
            for (int c = 0; c < 3; c++) o[4 * k + c] = P(t, k)[c];
        o[3] = asFloatBits(t);
        o[7] = asFloatBits(d.materialIds[t]);
        o[11] = 0.f;
    }
}

}  // namespace vr
