// vr_surface.h -- host-side processing of the scene's triangles.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/vrestir.h"

namespace vr {

// The light collection of a triangle scene and its power sampler, in the device
// layout of Scene::emTri / emAlias / emWeight.
struct LightCollectionHost {
    std::vector<float> tri;        // kEmTriFloats per collection triangle
    std::vector<uint32_t> alias;   // 4 words per entry
    std::vector<float> weight;     // flux per triangle
    int numActive = 0;
    double weightSum = 0.0;
};

// The ray-query structures of a triangle scene, in the device layout of
// Scene::bvh / bvhTri / triNormals / materials.
struct SurfaceGeometryHost {
    std::vector<float> nodes;      // 8 words per BvhNode (floats and int bits)
    std::vector<float> tris;       // kBvhTriFloats per triangle, leaf order
    std::vector<float> triData;    // kTriDataFloats per scene triangle
    std::vector<float> materials;  // kMaterialFloats per material
    int depth = 0;                 // the tree's depth (root = 0)
};

// Checks the descriptor; "" when it is valid.
std::string checkSurfacesDesc(const vrestir_surfaces_desc& d);
// LightCollection (LT/LightCollection.cpp, LT/BuildTriangleList.cs.slang,
// LT/FinalizeIntegration.cs.slang) and the power sampler's alias table
// (LT/EmissivePowerSampler.cpp:62-79, UT/Sampling/AliasTable.cpp:46-125).
void buildLightCollection(const vrestir_surfaces_desc& d, LightCollectionHost& out);
// A binned-SAH BVH over the triangles (leaves of at most 4, depth at most
// kBvhMaxDepth), the per-triangle shading normals (face normals where the
// descriptor has none) and the material table.
void buildSurfaceGeometry(const vrestir_surfaces_desc& d, SurfaceGeometryHost& out);

}  // namespace vr
