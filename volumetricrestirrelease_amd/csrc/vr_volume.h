// vr_volume.h -- host-side volume preparation (see vr_volume.cpp).
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/vrestir.h"
#include "vr_types.h"

namespace vr {

struct DenseF {
    int nx = 0, ny = 0, nz = 0;
    std::vector<float> v;
    size_t idx(int x, int y, int z) const { return (size_t)x + (size_t)nx * ((size_t)y + (size_t)ny * (size_t)z); }
    float at(int x, int y, int z) const {
        if (x < 0 || y < 0 || z < 0 || x >= nx || y >= ny || z >= nz) return 0.f;
        return v[idx(x, y, z)];
    }
};

struct HostGrid {
    bool valid = false;
    int nx = 0, ny = 0, nz = 0;
    int nbx = 0, nby = 0, nbz = 0, nnx = 0, nny = 0, nnz = 0;
    bool fp16 = false;
    float maxValue = 0.f;
    float w2i[16];
    uint32_t numBricks = 0;
    int channels = 1;
    std::vector<uint32_t> brickTable, nodeTable;
    std::vector<uint32_t> occ;    // {word, prefix} pairs over the brick-table index
    std::vector<float> bounds;
    std::vector<float> pool32;
    std::vector<uint16_t> pool16;
};

struct HostVolume {
    HostGrid slots[kMaxSlots];
    int numMips = 0;
    float tStep = 1.f;
    float superVoxelDiag = 1.f;
};

// The bricks of a .vbx file (vr_vbx.cpp): dense extent, index -> world, and per
// brick its index-space minimum and brickDim^3 values (x fastest).
struct VbxBricks {
    int nx = 0, ny = 0, nz = 0, brickDim = 8;
    float indexToWorld[16];
    std::vector<int32_t> pos;
    std::vector<float> vals;
};
int32_t readVbxBricks(const char* path, VbxBricks& out);
// One grid slot straight from bricks (no dense grid): the same slot buildVolume
// makes from the dense grid those bricks hold.
int buildSlotFromBricks(HostGrid& g, const VbxBricks& b, bool fp16, std::string& err);
// A slot of a volume loaded from files: its own index -> world transform.
void slotTransform(const float* indexToWorld, float* w2i);
void volumeSpacing(const float* indexToWorld, HostVolume& v);
// A volume from a .vbx file set <prefix>_mip{k}.vbx / _mip{k}c.vbx (vr_volume.cpp);
// i2w0 receives mip 0's index -> world transform.
int buildVolumeFromFiles(const std::string& prefix, int numMips, bool fp16, HostVolume& out, float* i2w0,
                         std::string& err);

uint16_t floatToHalfBits(float f);
float halfBitsToFloat(uint16_t h);
// The checks buildVolume makes of a descriptor (mips, dimensions, grey sigma_t, format).
int checkVolumeDesc(const vrestir_volume_desc& d, std::string& err);
int buildVolume(const vrestir_volume_desc& d, HostVolume& out, std::string& err);
// World -> index of mip level `scale` (1, 2, 4, ...) from mip 0's index -> world.
void mipTransform(const float* indexToWorld, int scale, float* w2i);
int buildTemperatureSlot(const float* t, int nx, int ny, int nz, const float* indexToWorld, bool fp16,
                         HostGrid& out, std::string& err);
int buildVelocitySlot(const float* v, int nx, int ny, int nz, const float* indexToWorld, HostGrid& out,
                      std::string& err);
int synthFbm(uint64_t seed, int nx, int ny, int nz, int octaves, float baseFreq, float lacunarity, float gain,
             int sphereMask, float threshold, float amplitude, float* out);

}  // namespace vr
