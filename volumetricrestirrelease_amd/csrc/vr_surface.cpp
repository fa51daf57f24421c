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

}  // namespace vr
