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

// Checks the descriptor; "" when it is valid.
std::string checkSurfacesDesc(const vrestir_surfaces_desc& d);
// LightCollection (LT/LightCollection.cpp, LT/BuildTriangleList.cs.slang,
// LT/FinalizeIntegration.cs.slang) and the power sampler's alias table
// (LT/EmissivePowerSampler.cpp:62-79, UT/Sampling/AliasTable.cpp:46-125).
void buildLightCollection(const vrestir_surfaces_desc& d, LightCollectionHost& out);

}  // namespace vr
