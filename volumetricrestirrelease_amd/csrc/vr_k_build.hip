// vr_k_build.hip -- volume grids built on the GPU from dense device grids: the
// same mip and conservative grids, brick tables, occupancy words, apron pools and
// bounds that the host build (vr_volume.cpp) makes, value for value, so that a
// frame of an animated volume can be rebuilt every frame without a trip through
// host memory.
//
// Every kernel repeats the host build's arithmetic in the host build's order
// (2x box reduction summed x-major, the 27-neighbourhood mean of the conservative
// grid, the tiny-value clamp against the grid maximum, round-to-nearest-even
// fp16, and a brick's bounds summed over its 10^3 apron with x outermost), so a
// device-built slot equals the host-built slot byte for byte.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "vr_launch.h"
#include "vr_volume.h"

namespace vr {

namespace {

constexpr int kBuildThreads = 256;

__device__ __forceinline__ size_t lin(int x, int y, int z, int nx, int ny) {
    return (size_t)x + (size_t)nx * ((size_t)y + (size_t)ny * (size_t)z);
}
__device__ __forceinline__ float at(const float* __restrict__ s, int x, int y, int z, int nx, int ny, int nz) {
    if (x < 0 || y < 0 || z < 0 || x >= nx || y >= ny || z >= nz) return 0.f;
    return s[lin(x, y, z, nx, ny)];
}

// IEEE binary16 bits of a float, round to nearest even (the host build's rule,
// written on the bits so that no device conversion mode can differ from it).
__device__ __forceinline__ uint16_t halfBits(float f) {
    const uint32_t x = __float_as_uint(f);
    const uint32_t sign = (x >> 16) & 0x8000u;
    const uint32_t absx = x & 0x7FFFFFFFu;
    if (absx >= 0x7F800000u) return (uint16_t)(sign | (absx > 0x7F800000u ? 0x7E00u : 0x7C00u));
    if (absx >= 0x477FF000u) return (uint16_t)(sign | 0x7C00u);
    if (absx < 0x38800000u) {
        if (absx < 0x33000000u) return (uint16_t)sign;
        const uint32_t mant = (absx & 0x7FFFFFu) | 0x800000u;
        const int e = (int)(absx >> 23);
        const int shift = 126 - e;
        uint32_t hm = mant >> shift;
        const uint32_t rem = mant & ((1u << shift) - 1u);
        const uint32_t half = 1u << (shift - 1);
        if (rem > half || (rem == half && (hm & 1u))) hm++;
        return (uint16_t)(sign | hm);
    }
    uint32_t h = (((absx >> 23) - 112u) << 10) | ((absx >> 13) & 0x3FFu);
    const uint32_t rem = absx & 0x1FFFu;
    if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h++;
    return (uint16_t)(sign | h);
}
__device__ __forceinline__ float halfValue(uint16_t h) {
    const uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
    const uint32_t e = (h >> 10) & 0x1Fu, m = h & 0x3FFu;
    uint32_t x;
    if (e == 0) {
        if (m == 0) return __uint_as_float(sign);
        return __uint_as_float(__float_as_uint(ldexpf((float)m, -24)) | sign);
    }
    if (e == 31) x = sign | 0x7F800000u | (m << 13);
    else x = sign | ((e + 112u) << 23) | (m << 13);
    return __uint_as_float(x);
}
// The stored value of a raw value (values below 1e-9 of the grid maximum are 0).
__device__ __forceinline__ float stored(float x, float gmax, int fp16) {
    x = (gmax > 0.f && x / gmax < 1e-9f) ? 0.f : x;
    if (gmax <= 0.f) x = 0.f;
    return fp16 ? halfValue(halfBits(x)) : x;
}

// The largest of each wave's values into *maxBits (values are >= 0, so their
// bits order as the values do).
__device__ __forceinline__ void waveMax(unsigned m, unsigned* __restrict__ maxBits) {
    for (int o = 32; o > 0; o >>= 1) m = max(m, (unsigned)__shfl_xor((int)m, o));
    if ((threadIdx.x & 63) == 0 && m) atomicMax(maxBits, m);
}
// out = max(in, 0); bad counts non-finite inputs; maxBits: the largest output
__global__ __launch_bounds__(kBuildThreads) void kClamp(const float* __restrict__ in, float* __restrict__ out,
                                                        size_t n, int* __restrict__ bad, unsigned* __restrict__ maxBits) {
    unsigned m = 0u;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float v = in[i];
        if (!isfinite(v)) atomicAdd(bad, 1);
        const float o = v > 0.f ? v : 0.f;
        out[i] = o;
        if (isfinite(v)) m = max(m, __float_as_uint(o));
    }
    waveMax(m, maxBits);
}
// kClamp on 16-byte-aligned grids whose size is a multiple of 4, four values per load
__global__ __launch_bounds__(kBuildThreads) void kClamp4(const float4* __restrict__ in, float4* __restrict__ out,
                                                         size_t n4, int* __restrict__ bad, unsigned* __restrict__ maxBits) {
    unsigned m = 0u;
    int nonFinite = 0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += (size_t)gridDim.x * blockDim.x) {
        const float4 v = in[i];
        const float e[4] = {v.x, v.y, v.z, v.w};
        float o[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            o[k] = e[k] > 0.f ? e[k] : 0.f;
            if (isfinite(e[k])) m = max(m, __float_as_uint(o[k]));
            else nonFinite++;
        }
        out[i] = make_float4(o[0], o[1], o[2], o[3]);
    }
    if (nonFinite) atomicAdd(bad, nonFinite);
    waveMax(m, maxBits);
}
// Conservative grid: an empty voxel takes the mean of the positive values of its
// 27-neighbourhood (summed x outermost, as the host build does). A workgroup takes
// a 64 x 8 x 8 tile: the tile and its one-voxel halo are read once into LDS, a tile
// with no positive value anywhere in it is copied through, and every other empty
// voxel sums its neighbourhood from LDS in the host build's order.
constexpr int kDilX = 64, kDilY = 8, kDilZ = 8;
__global__ __launch_bounds__(kBuildThreads) void kDilate(const float* __restrict__ s, float* __restrict__ d, int nx,
                                                         int ny, int nz, unsigned* __restrict__ maxBits) {
    constexpr int TX = kDilX + 2, TY = kDilY + 2, TZ = kDilZ + 2;
    __shared__ float t[TZ][TY][TX];
    const int x0 = (int)blockIdx.x * kDilX, y0 = (int)blockIdx.y * kDilY, z0 = (int)blockIdx.z * kDilZ;
    int positive = 0;
    for (int i = threadIdx.x; i < TX * TY * TZ; i += blockDim.x) {
        const int lx = i % TX, ly = (i / TX) % TY, lz = i / (TX * TY);
        const float v = at(s, x0 + lx - 1, y0 + ly - 1, z0 + lz - 1, nx, ny, nz);
        t[lz][ly][lx] = v;
        positive |= v > 0.f ? 1 : 0;
    }
    const int any = __syncthreads_or(positive);
    unsigned m = 0u;
    for (int i = threadIdx.x; i < kDilX * kDilY * kDilZ; i += blockDim.x) {
        const int lx = i % kDilX, ly = (i / kDilX) % kDilY, lz = i / (kDilX * kDilY);
        const int x = x0 + lx, y = y0 + ly, z = z0 + lz;
        if (x >= nx || y >= ny || z >= nz) continue;
        const float v = t[lz + 1][ly + 1][lx + 1];
        float o = v;
        if (v == 0.f && any) {
            float avg = 0.f;
            for (int ii = -1; ii <= 1; ii++)
                for (int jj = -1; jj <= 1; jj++)
                    for (int kk = -1; kk <= 1; kk++) {
                        const float tv = t[lz + 1 + kk][ly + 1 + jj][lx + 1 + ii];
                        avg += tv > 0.f ? tv : 0.f;
                    }
            avg /= 27.f;
            if (avg > 0.f) o = avg;
        }
        d[lin(x, y, z, nx, ny)] = o;
        m = max(m, __float_as_uint(o));
    }
    waveMax(m, maxBits);
}
// 2x box reduction, weight 1/8 per tap summed x-major, negative results to 0; d is nx/2 x ny/2 x nz/2
__global__ __launch_bounds__(kBuildThreads) void kReduce2(const float* __restrict__ s, float* __restrict__ d, int nx,
                                                          int ny, int nz, unsigned* __restrict__ maxBits) {
    const int dx = nx / 2, dy = ny / 2, dz = nz / 2;
    const size_t n = (size_t)dx * dy * dz;
    unsigned m = 0u;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int x = (int)(i % dx), y = (int)((i / dx) % dy), z = (int)(i / ((size_t)dx * dy));
        float res = 0.f;
        for (int a = 0; a < 2; a++)
            for (int b = 0; b < 2; b++)
                for (int c = 0; c < 2; c++)
                    res = __fadd_rn(res, __fmul_rn(0.125f, s[lin(2 * x + a, 2 * y + b, 2 * z + c, nx, ny)]));
        const float o = res > 0.f ? res : 0.f;
        d[i] = o;
        m = max(m, __float_as_uint(o));
    }
    waveMax(m, maxBits);
}
// per brick cell: does any stored value of its 8^3 voxels exceed 0
__global__ __launch_bounds__(kBuildThreads) void kBrickOcc(const float* __restrict__ s, int nx, int ny, int nz,
                                                           int nbx, int nby, int nbz, const unsigned* __restrict__ gmaxBits,
                                                           int fp16, uint8_t* __restrict__ occ) {
    const float gmax = __uint_as_float(*gmaxBits);
    const size_t nb = (size_t)nbx * nby * nbz;
    for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < nb; c += (size_t)gridDim.x * blockDim.x) {
        const int bx = (int)(c % nbx), by = (int)((c / nbx) % nby), bz = (int)(c / ((size_t)nbx * nby));
        bool any = false;
        // a row's 8 loads are issued together, then tested
        for (int z = 0; z < 8 && !any; z++)
            for (int y = 0; y < 8 && !any; y++) {
                float row[8];
                const float* src = s + lin(bx * 8, by * 8 + y, bz * 8 + z, nx, ny);
#pragma unroll
                for (int x = 0; x < 8; x++) row[x] = src[x];
#pragma unroll
                for (int x = 0; x < 8; x++) any |= stored(row[x], gmax, fp16) > 0.f;
            }
        occ[c] = any ? 1 : 0;
    }
}
// occupancy words {bits, prefix}: the bits here, the counts for the scan
__global__ __launch_bounds__(kBuildThreads) void kOccWords(const uint8_t* __restrict__ occ, size_t nb,
                                                           uint32_t* __restrict__ words, uint32_t* __restrict__ cnt,
                                                           size_t nw) {
    for (size_t w = (size_t)blockIdx.x * blockDim.x + threadIdx.x; w < nw; w += (size_t)gridDim.x * blockDim.x) {
        uint32_t bits = 0u;
        for (int b = 0; b < 32; b++) {
            const size_t c = w * 32 + b;
            if (c < nb && occ[c]) bits |= 1u << b;
        }
        words[2 * w] = bits;
        cnt[w] = (uint32_t)__popc(bits);
    }
}
// Exclusive prefix of cnt into the words' second entries, one workgroup: each
// thread sums a contiguous run, the runs' totals are scanned in LDS, then each
// thread writes its run. total receives the number of bricks.
__global__ __launch_bounds__(1024) void kScanWords(uint32_t* __restrict__ words, const uint32_t* __restrict__ cnt,
                                                   size_t nw, uint32_t* __restrict__ total) {
    __shared__ uint32_t part[1024];
    const size_t per = (nw + 1023) / 1024;
    const size_t w0 = threadIdx.x * per, w1 = min(nw, w0 + per);
    uint32_t s = 0u;
    for (size_t w = w0; w < w1; w++) s += cnt[w];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        const uint32_t v = threadIdx.x >= (unsigned)o ? part[threadIdx.x - o] : 0u;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    uint32_t run = part[threadIdx.x] - s;
    for (size_t w = w0; w < w1; w++) {
        words[2 * w + 1] = run;
        run += cnt[w];
    }
    if (threadIdx.x == 1023) *total = part[1023];
}
// brick ids in brick-table order: the table, each id's cell, the block table
__global__ __launch_bounds__(kBuildThreads) void kBrickIds(const uint32_t* __restrict__ words, int nbx, int nby,
                                                           int nbz, int nnx, int nny, uint32_t* __restrict__ table,
                                                           uint32_t* __restrict__ cellOf,
                                                           uint32_t* __restrict__ nodes) {
    const size_t nb = (size_t)nbx * nby * nbz;
    for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < nb; c += (size_t)gridDim.x * blockDim.x) {
        const uint32_t bits = words[2 * (c / 32)], b = (uint32_t)(c % 32);
        if (!((bits >> b) & 1u)) { table[c] = kEmpty; continue; }
        const uint32_t id = words[2 * (c / 32) + 1] + (uint32_t)__popc(bits & ((1u << b) - 1u));
        table[c] = id;
        cellOf[id] = (uint32_t)c;
        const int bx = (int)(c % nbx), by = (int)((c / nbx) % nby), bz = (int)(c / ((size_t)nbx * nby));
        atomicOr(&nodes[(bx / 16) + (size_t)nnx * ((by / 16) + (size_t)nny * (bz / 16))], 1u);
    }
}
// per brick: its 10^3 apron values (stored form) and bounds {min, max, sum / 512, 0}
template <int FP16>
__global__ __launch_bounds__(kBuildThreads) void kBrickFill(const float* __restrict__ s, int nx, int ny, int nz,
                                                            int nbx, int nby, const uint32_t* __restrict__ cellOf,
                                                            const uint32_t* __restrict__ count,
                                                            const unsigned* __restrict__ gmaxBits,
                                                            void* __restrict__ pool, float* __restrict__ bounds) {
    const float gmax = __uint_as_float(*gmaxBits);
    const uint32_t nbricks = *count;
    for (size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x; id < nbricks;
         id += (size_t)gridDim.x * blockDim.x) {
        const uint32_t c = cellOf[id];
        const int x0 = (int)(c % nbx) * 8, y0 = (int)((c / nbx) % nby) * 8, z0 = (int)(c / ((size_t)nbx * nby)) * 8;
        float mn = 3.402823466e+38f, mx = 0.f, sum = 0.f;
        for (int i = -1; i <= 8; i++)
            for (int j = -1; j <= 8; j++) {
                // the 10 loads along z are issued together, then consumed in order
                float raw[kApron];
#pragma unroll
                for (int k = -1; k <= 8; k++) raw[k + 1] = at(s, x0 + i, y0 + j, z0 + k, nx, ny, nz);
#pragma unroll
                for (int k = -1; k <= 8; k++) {
                    const float v = stored(raw[k + 1], gmax, FP16);
                    const size_t o = id * kBrickVals + (size_t)(i + 1) + kApron * ((size_t)(j + 1) + kApron * (size_t)(k + 1));
                    if (FP16) ((uint16_t*)pool)[o] = halfBits(v);
                    else ((float*)pool)[o] = v;
                    mn = fminf(mn, v);
                    mx = fmaxf(mx, v);
                    sum = __fadd_rn(sum, v);
                }
            }
        bounds[4 * id + 0] = mn;
        bounds[4 * id + 1] = mx;
        bounds[4 * id + 2] = sum / 512.f;
        bounds[4 * id + 3] = 0.f;
    }
}

dim3 gridFor(size_t n) { return dim3((unsigned)std::min<size_t>((n + kBuildThreads - 1) / kBuildThreads, 65536)); }

}  // namespace

hipError_t buildClamp(hipStream_t st, const float* in, float* out, size_t n, int* bad, unsigned* maxBits) {
    if (n % 4 == 0 && ((uintptr_t)in % 16) == 0 && ((uintptr_t)out % 16) == 0)
        hipLaunchKernelGGL(kClamp4, gridFor(n / 4), dim3(kBuildThreads), 0, st, (const float4*)in, (float4*)out, n / 4,
                           bad, maxBits);
    else
        hipLaunchKernelGGL(kClamp, gridFor(n), dim3(kBuildThreads), 0, st, in, out, n, bad, maxBits);
    return hipGetLastError();
}
hipError_t buildDilate(hipStream_t st, const float* s, float* d, int nx, int ny, int nz, unsigned* maxBits) {
    const dim3 g((unsigned)((nx + kDilX - 1) / kDilX), (unsigned)((ny + kDilY - 1) / kDilY),
                 (unsigned)((nz + kDilZ - 1) / kDilZ));
    hipLaunchKernelGGL(kDilate, g, dim3(kBuildThreads), 0, st, s, d, nx, ny, nz, maxBits);
    return hipGetLastError();
}
hipError_t buildReduce2(hipStream_t st, const float* s, float* d, int nx, int ny, int nz, unsigned* maxBits) {
    hipLaunchKernelGGL(kReduce2, gridFor((size_t)(nx / 2) * (ny / 2) * (nz / 2)), dim3(kBuildThreads), 0, st, s, d,
                       nx, ny, nz, maxBits);
    return hipGetLastError();
}

// One slot from a dense device grid (nx, ny, nz multiples of 8, values >= 0)
// whose largest value's bits are at gmaxBits (device memory, left there by the
// kernel that made the grid). scratch: a device buffer of at least
// buildScratchBytes(nx, ny, nz) bytes.
// The slot's buffers come from allocBuf (which may hand back a buffer of an
// earlier build); the host-side metadata goes to numBricks / maxValue.
// Synchronises the stream (the brick count sizes the pool).
size_t buildScratchBytes(int nx, int ny, int nz) {
    const size_t nb = (size_t)(nx / 8) * (ny / 8) * (nz / 8), nw = (nb + 31) / 32;
    return 256 + nb + 4 * nw + 4 * nb + 256;
}
hipError_t buildSlotDevice(hipStream_t st, const float* dense, int nx, int ny, int nz, const unsigned* gmaxBits,
                           bool fp16, void* scratch, Grid& G, uint32_t& numBricks, float& maxValue,
                           const BuildAlloc& allocBuf) {
    hipError_t e;
    const int nbx = nx / 8, nby = ny / 8, nbz = nz / 8;
    const int nnx = (nbx + 15) / 16, nny = (nby + 15) / 16, nnz = (nbz + 15) / 16;
    const size_t n = (size_t)nx * ny * nz, nb = (size_t)nbx * nby * nbz, nw = (nb + 31) / 32;
    char* sp = (char*)scratch;
    uint32_t* count = (uint32_t*)sp;                     // the brick count
    uint8_t* occ = (uint8_t*)(sp + 256);
    uint32_t* cnt = (uint32_t*)(sp + 256 + ((nb + 255) / 256) * 256);
    uint32_t* cellOf = cnt + nw;
    uint32_t *table, *nodes, *words;
    if ((e = allocBuf(kBufTable, nb * 4, (void**)&table)) != hipSuccess) return e;
    if ((e = allocBuf(kBufNodes, (size_t)nnx * nny * nnz * 4, (void**)&nodes)) != hipSuccess) return e;
    if ((e = allocBuf(kBufWords, nw * 8, (void**)&words)) != hipSuccess) return e;
    if ((e = hipMemsetAsync(nodes, 0, (size_t)nnx * nny * nnz * 4, st)) != hipSuccess) return e;
    hipLaunchKernelGGL(kBrickOcc, gridFor(nb), dim3(kBuildThreads), 0, st, dense, nx, ny, nz, nbx, nby, nbz, gmaxBits,
                       fp16 ? 1 : 0, occ);
    hipLaunchKernelGGL(kOccWords, gridFor(nw), dim3(kBuildThreads), 0, st, occ, nb, words, cnt, nw);
    hipLaunchKernelGGL(kScanWords, dim3(1), dim3(1024), 0, st, words, cnt, nw, count);
    hipLaunchKernelGGL(kBrickIds, gridFor(nb), dim3(kBuildThreads), 0, st, words, nbx, nby, nbz, nnx, nny, table,
                       cellOf, nodes);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    // one wait per slot: the brick count sizes the pool (the maximum comes along)
    uint32_t nbr = 0;
    unsigned hostMax = 0;
    if ((e = hipMemcpyAsync(&nbr, count, 4, hipMemcpyDeviceToHost, st)) != hipSuccess) return e;
    if ((e = hipMemcpyAsync(&hostMax, gmaxBits, 4, hipMemcpyDeviceToHost, st)) != hipSuccess) return e;
    if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
    float gmax;
    std::memcpy(&gmax, &hostMax, 4);
    float* bounds;
    void* pool;
    if ((e = allocBuf(kBufBounds, (size_t)nbr * 16, (void**)&bounds)) != hipSuccess) return e;
    if ((e = allocBuf(kBufPool, (size_t)nbr * kBrickVals * (fp16 ? 2 : 4), &pool)) != hipSuccess) return e;
    if (nbr > 0) {
        if (fp16)
            hipLaunchKernelGGL(kBrickFill<1>, gridFor(nbr), dim3(kBuildThreads), 0, st, dense, nx, ny, nz, nbx, nby,
                               cellOf, count, gmaxBits, pool, bounds);
        else
            hipLaunchKernelGGL(kBrickFill<0>, gridFor(nbr), dim3(kBuildThreads), 0, st, dense, nx, ny, nz, nbx, nby,
                               cellOf, count, gmaxBits, pool, bounds);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    // the stored form of the maximum (monotone, so the largest stored value)
    float mv = gmax;
    if (!(gmax > 0.f)) mv = 0.f;
    else if (fp16) mv = halfBitsToFloat(floatToHalfBits(gmax));
    std::memset(&G, 0, sizeof(Grid));
    G.brickTable = table;
    G.nodeTable = nodes;
    G.occ = words;
    G.occWords = (int32_t)nw;
    G.numBricks = nbr;
    G.bounds = bounds;
    G.pool = pool;
    G.nbx = nbx; G.nby = nby; G.nbz = nbz;
    G.nnx = nnx; G.nny = nny; G.nnz = nnz;
    G.bmax[0] = (float)nx; G.bmax[1] = (float)ny; G.bmax[2] = (float)nz;
    G.maxValue = mv;
    G.fp16 = fp16 ? 1 : 0;
    G.channels = 1;
    G.valid = 1;
    numBricks = nbr;
    maxValue = mv;
    return hipSuccess;
}

}  // namespace vr
