// vr_rayquery.hpp -- ray queries against the scene's triangles: FindSurfaceHit
// (closest hit) and FindIfOccluded (any hit) of VR/InlineRayTracingHelpers.slang:
// 42-155, over a BVH instead of the D3D12 acceleration structure.
//
// The triangle test is the watertight test of Woop, Benthin and Wald ("Watertight
// Ray/Triangle Intersection", JCGT 2(1), 2013): shear the triangle into ray space,
// take the three edge functions (recomputed in double when one is exactly zero),
// accept t in [tMin, tMax]. It runs with contraction off, so a hit's distance and
// barycentric weights do not depend on how the compiler fuses its arithmetic. The
// closest hit is the smallest t, and among equal t the smallest scene triangle
// index, so the result does not depend on the traversal order.
#pragma once
#include "vr_device.hpp"

namespace vr {

// computeRayOrigin (UT/Helpers.slang:92-105): a point moved off its surface along
// the normal, by a few ulps of each coordinate (by a fixed 2^-16 near the origin).
VR_DEV v3 rayOriginOffset(v3 p, v3 n) {
    const float origin = 1.f / 32.f, fScale = 1.f / 65536.f, iScale = 256.f;
    const int ix = (int)(n.x * iScale), iy = (int)(n.y * iScale), iz = (int)(n.z * iScale);
    const float px = __int_as_float(__float_as_int(p.x) + (p.x < 0.f ? -ix : ix));
    const float py = __int_as_float(__float_as_int(p.y) + (p.y < 0.f ? -iy : iy));
    const float pz = __int_as_float(__float_as_int(p.z) + (p.z < 0.f ? -iz : iz));
    return V3(fabsf(p.x) < origin ? p.x + n.x * fScale : px, fabsf(p.y) < origin ? p.y + n.y * fScale : py,
              fabsf(p.z) < origin ? p.z + n.z * fScale : pz);
}
struct RayShear {
    int kx, ky, kz;
    float Sx, Sy, Sz;
};
VR_DEV float comp(v3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }
VR_DEV RayShear rayShear(v3 d) {
    RayShear r;
    const float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    r.kz = (ax > ay) ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
    r.kx = r.kz + 1 == 3 ? 0 : r.kz + 1;
    r.ky = r.kx + 1 == 3 ? 0 : r.kx + 1;
    if (comp(d, r.kz) < 0.f) { const int t = r.kx; r.kx = r.ky; r.ky = t; }
    const float dz = comp(d, r.kz);
    r.Sx = comp(d, r.kx) / dz;
    r.Sy = comp(d, r.ky) / dz;
    r.Sz = 1.f / dz;
    return r;
}
// Watertight test of triangle (a, b, c) against the ray (o, shear); on a hit in
// [tMin, tMax] returns true with t and the weights of b and c.
VR_DEV bool triHit(const RayShear& r, v3 o, v3 a, v3 b, v3 c, float tMin, float tMax, float& t, float& wb, float& wc) {
#pragma clang fp contract(off)
    const v3 A = a - o, B = b - o, C = c - o;
    const float Ax = comp(A, r.kx) - r.Sx * comp(A, r.kz), Ay = comp(A, r.ky) - r.Sy * comp(A, r.kz);
    const float Bx = comp(B, r.kx) - r.Sx * comp(B, r.kz), By = comp(B, r.ky) - r.Sy * comp(B, r.kz);
    const float Cx = comp(C, r.kx) - r.Sx * comp(C, r.kz), Cy = comp(C, r.ky) - r.Sy * comp(C, r.kz);
    float U = Cx * By - Cy * Bx, V = Ax * Cy - Ay * Cx, W = Bx * Ay - By * Ax;
    if (U == 0.f || V == 0.f || W == 0.f) {
        U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
        V = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
        W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
    }
    if ((U < 0.f || V < 0.f || W < 0.f) && (U > 0.f || V > 0.f || W > 0.f)) return false;
    const float det = U + V + W;
    if (det == 0.f) return false;
    const float Az = r.Sz * comp(A, r.kz), Bz = r.Sz * comp(B, r.kz), Cz = r.Sz * comp(C, r.kz);
    const float T = U * Az + V * Bz + W * Cz;
    const float inv = 1.f / det;
    t = T * inv;
    if (!(t >= tMin && t <= tMax)) return false;
    wb = V * inv;
    wc = W * inv;
    return true;
}

struct SurfHit {
    int tri;       // scene triangle index, -1: none
    float t, bx, by;
};

// The closest hit in [tMin, tMax] (anyHit: whether there is one, stopping at the
// first). Stack of kBvhMaxDepth + 2 entries: each interior level pushes at most one.
VR_DEV bool surfaceQuery(const Scene& S, v3 o, v3 d, float tMin, float tMax, bool anyHit, SurfHit& h) {
    h.tri = -1;
    h.t = kTMax;
    h.bx = h.by = 0.f;
    if (S.numTris == 0) return false;
    const RayShear rs = rayShear(d);
    const v3 inv = V3(1.f / d.x, 1.f / d.y, 1.f / d.z);
    // slab distances widened by 1 + 2 gamma(3) (robust traversal, Ize 2013)
    const float widen = 1.f + 2.f * 3.5762786e-07f;
    float best = tMax;
    int stack[kBvhMaxDepth + 2];
    int sp = 0, node = 0;
    auto boxNear = [&](const BvhNode& n, float& tn) -> bool {
        const float tx0 = (n.bmin[0] - o.x) * inv.x, tx1 = (n.bmax[0] - o.x) * inv.x;
        const float ty0 = (n.bmin[1] - o.y) * inv.y, ty1 = (n.bmax[1] - o.y) * inv.y;
        const float tz0 = (n.bmin[2] - o.z) * inv.z, tz1 = (n.bmax[2] - o.z) * inv.z;
        tn = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), tMin));
        const float tf = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fminf(fmaxf(tz0, tz1), best)) * widen;
        return tn <= tf;
    };
    float tRoot;
    if (!boxNear(S.bvh[0], tRoot)) return false;
    for (;;) {
        const BvhNode& n = S.bvh[node];
        if (n.count > 0) {
            for (int i = 0; i < n.count; i++) {
                const float* T = S.bvhTri + (size_t)kBvhTriFloats * (n.first + i);
                float t, wb, wc;
                if (!triHit(rs, o, ld3(T), ld3(T + 4), ld3(T + 8), tMin, best, t, wb, wc)) continue;
                const int id = __float_as_int(T[3]);
                if (t < best || h.tri < 0 || (t == best && id < h.tri)) {
                    best = t;
                    h.tri = id;
                    h.t = t;
                    h.bx = wb;
                    h.by = wc;
                    if (anyHit) return true;
                }
            }
        } else {
            float t0, t1;
            const bool h0 = boxNear(S.bvh[n.first], t0), h1 = boxNear(S.bvh[n.first + 1], t1);
            if (h0 && h1 && sp < kBvhMaxDepth + 2) {  // the depth bound keeps sp below the limit
                const bool firstNear = t0 <= t1;
                stack[sp++] = firstNear ? n.first + 1 : n.first;
                node = firstNear ? n.first : n.first + 1;
                continue;
            }
            if (h0 || h1) {
                node = h0 ? n.first : n.first + 1;
                continue;
            }
        }
        // pop, skipping nodes that no longer come before the closest hit
        bool found = false;
        while (sp > 0) {
            node = stack[--sp];
            float tn;
            if (boxNear(S.bvh[node], tn)) { found = true; break; }
        }
        if (!found) break;
    }
    return h.tri >= 0;
}

// Whether scene triangle tri, seen from direction V (towards the viewer), is an
// emitting front face: _prepareShadingData's frontFacing = dot(V, faceN) >= 0
// (SC/Shading.slang:172) and emissive only there (SC/Shading.slang:229-232).
VR_DEV bool surfaceEmitsTowards(const Scene& S, int tri, v3 V) {
    const float* T = S.triData + (size_t)kTriDataFloats * tri;
    const v3 faceN = cross3(ld3(T + 3) - ld3(T), ld3(T + 6) - ld3(T));
    if (!(dot3(V, faceN) >= 0.f)) return false;
    const float* M = S.materials + (size_t)kMaterialFloats * __float_as_int(T[18]);
    return M[6] > 0.f || M[7] > 0.f || M[8] > 0.f;
}

}  // namespace vr
