// vr_k_env.hip -- set-up of the environment map's hierarchical importance map
// (LT/EnvMapSampler.cpp:83-110, LT/EnvMapSamplerSetup.cs.slang:48-70), run once
// when a map is bound.
#include "vr_pass.hpp"
#include "vr_launch.h"

namespace vr {

// Base level: each texel is the mean luminance of spp = sx * sy stratified
// samples of the lat-long map, taken at the directions of the equal-area
// octahedral map, read with the filter the pass uses (texBilinear).
__global__ __launch_bounds__(256) void kEnvImportanceBase(const float* __restrict__ tex, int W, int H,
                                                          float* __restrict__ out, int dim, int sx, int sy,
                                                          float invSamples) {
    const int x = blockIdx.x * 16 + (threadIdx.x & 15), y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= dim || y >= dim) return;
    const float invX = 1.f / (float)(dim * sx), invY = 1.f / (float)(dim * sy);
    float L = 0.f;
    for (int j = 0; j < sy; j++)
        for (int i = 0; i < sx; i++) {
            float px = ((float)(x * sx + i) + 0.5f) * invX, py = ((float)(y * sy + j) + 0.5f) * invY;
            v3 dir = octToDir(px, py);
            float u, v;
            latlong(dir, u, v);
            L += lum(texBilinear(tex, W, H, u, v));
        }
    out[(size_t)y * dim + x] = L * invSamples;
}

// Next level: the mean of each 2x2 block (the texture's mip generation).
__global__ __launch_bounds__(256) void kEnvImportanceReduce(const float* __restrict__ src, float* __restrict__ dst,
                                                            int dstDim) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= dstDim * dstDim) return;
    const int x = i % dstDim, y = i / dstDim, sd = 2 * dstDim;
    const float* s = src + (size_t)(2 * y) * sd + 2 * x;
    dst[i] = ((s[0] + s[1]) + (s[sd] + s[sd + 1])) * 0.25f;
}

hipError_t launchEnvImportance(hipStream_t st, const float* tex, int W, int H, float* imp, const int* off, int dim,
                               int spp) {
    int sx = 1;
    while ((sx * 2) * (sx * 2) <= spp) sx *= 2;
    const int sy = spp / sx;
    dim3 grid((dim + 15) / 16, (dim + 15) / 16);
    hipLaunchKernelGGL(kEnvImportanceBase, grid, dim3(256), 0, st, tex, W, H, imp + off[0], dim, sx, sy,
                       1.f / (float)(sx * sy));
    for (int k = 1, d = dim / 2; d >= 1; k++, d /= 2)
        hipLaunchKernelGGL(kEnvImportanceReduce, dim3((d * d + 255) / 256), dim3(256), 0, st, imp + off[k - 1],
                           imp + off[k], d);
    return hipGetLastError();
}

}  // namespace vr
