// vr_k_plume.hip -- the seeded synthetic plume of the animated emissive workload:
// smoke density and temperature on an n^3 grid (the unit cube), rising with a
// buoyant updraft and stirred by curl noise, simulated on the GPU one step per
// frame. Synthetic input only (the reference loads its animated volumes from
// files); the grids feed vrestir_stream_volume_frame.
//
// Each step:
//   * the velocity, on a coarse V^3 lattice: an updraft u(r) = U exp(-r^2 / w^2)
//     around the plume's axis plus A curl(psi), psi a three-component value-noise
//     potential drifting in time (curl by central differences, so the stirring
//     is divergence-free up to the differencing);
//   * semi-Lagrangian advection of both fields on the fine grid: back-trace each
//     voxel centre through the trilinearly interpolated velocity, sample the old
//     field trilinearly there, then decay (density) and cool (temperature), and
//     add the source: a noisy sphere at the base of the axis, hot at its core.
// Empty space is skipped: each step marks the 8^3 blocks where the new density is
// non-zero, the marks are dilated by the farthest a back-trace can reach (from the
// largest lattice speed), and a voxel outside the dilated marks and the source
// writes zeros without tracing, which is what tracing would give it.
// Deterministic: every value a function of the seed and the step (the only atomic
// is a maximum).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/vrestir.h"

namespace {

constexpr int kThreads = 256;
constexpr int kVel = 64;  // coarse velocity lattice

__device__ __forceinline__ uint32_t hash3(int x, int y, int z, uint32_t s) {
    uint32_t h = s ^ 0x9E3779B9u;
    h ^= (uint32_t)x * 0x85EBCA6Bu; h = (h << 13) | (h >> 19); h *= 0xC2B2AE35u;
    h ^= (uint32_t)y * 0x27D4EB2Fu; h = (h << 13) | (h >> 19); h *= 0x165667B1u;
    h ^= (uint32_t)z * 0x9E3779B1u; h ^= h >> 15; h *= 0x85EBCA77u; h ^= h >> 13;
    return h;
}
__device__ __forceinline__ float lattice(int x, int y, int z, uint32_t s) {
    return (float)(hash3(x, y, z, s) >> 8) * (2.f / 16777216.f) - 1.f;  // [-1, 1)
}
__device__ __forceinline__ float quintic(float t) { return t * t * t * (t * (t * 6.f - 15.f) + 10.f); }
// value noise in [-1, 1] with a C2 fade
__device__ float vnoise(float x, float y, float z, uint32_t s) {
    const float fx = floorf(x), fy = floorf(y), fz = floorf(z);
    const int ix = (int)fx, iy = (int)fy, iz = (int)fz;
    const float u = quintic(x - fx), v = quintic(y - fy), w = quintic(z - fz);
    float c[2][2][2];
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++)
            for (int d = 0; d < 2; d++) c[a][b][d] = lattice(ix + a, iy + b, iz + d, s);
    const float x00 = c[0][0][0] + u * (c[1][0][0] - c[0][0][0]), x10 = c[0][1][0] + u * (c[1][1][0] - c[0][1][0]);
    const float x01 = c[0][0][1] + u * (c[1][0][1] - c[0][0][1]), x11 = c[0][1][1] + u * (c[1][1][1] - c[0][1][1]);
    const float y0 = x00 + v * (x10 - x00), y1 = x01 + v * (x11 - x01);
    return y0 + w * (y1 - y0);
}
// two octaves of the potential's component k at p (unit cube) and time t
__device__ float potential(float px, float py, float pz, float t, int k, uint32_t seed) {
    const uint32_t s = seed * 3u + (uint32_t)k * 0x632BE5ABu;
    const float f = 3.f, drift = 0.35f * t;
    return vnoise(px * f, py * f - drift, pz * f + 0.5f * drift, s) +
           0.5f * vnoise(px * 2.f * f + 7.1f, py * 2.f * f - 2.f * drift, pz * 2.f * f, s ^ 0xA511E9B3u);
}

struct PlumeParams {
    int n;            // fine grid side
    uint32_t seed;
    float t;          // simulated time of the step (frames)
    float dt;         // frames per step
    float updraft;    // U, unit-cube lengths per frame
    float width;      // w
    float curl;       // A
    float decay, cool;
    float srcR;       // source radius
    float srcY;       // source centre height (axis at x = z = 0.5)
    float srcDensity, srcTemperature;
};

// velocity on the V^3 lattice (lattice point i at (i + 0.5) / V), unit cube per frame
__global__ __launch_bounds__(kThreads) void kVelocity(PlumeParams P, float4* __restrict__ vel,
                                                      unsigned* __restrict__ vmaxBits) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kVel * kVel * kVel) return;  // kVel^3 is a multiple of the block size
    const int x = i % kVel, y = (i / kVel) % kVel, z = i / (kVel * kVel);
    const float px = (x + 0.5f) / kVel, py = (y + 0.5f) / kVel, pz = (z + 0.5f) / kVel;
    const float h = 0.5f / kVel;
    // curl psi = (dpz/dy - dpy/dz, dpx/dz - dpz/dx, dpy/dx - dpx/dy)
    auto d = [&](int k, int axis) {
        const float ox = axis == 0 ? h : 0.f, oy = axis == 1 ? h : 0.f, oz = axis == 2 ? h : 0.f;
        return (potential(px + ox, py + oy, pz + oz, P.t, k, P.seed) -
                potential(px - ox, py - oy, pz - oz, P.t, k, P.seed)) / (2.f * h);
    };
    const float cx = d(2, 1) - d(1, 2), cy = d(0, 2) - d(2, 0), cz = d(1, 0) - d(0, 1);
    const float rx = px - 0.5f, rz = pz - 0.5f;
    const float up = P.updraft * expf(-(rx * rx + rz * rz) / (P.width * P.width));
    const float4 v = make_float4(P.curl * cx, up + P.curl * cy, P.curl * cz, 0.f);
    vel[i] = v;
    unsigned m = __float_as_uint(sqrtf(v.x * v.x + v.y * v.y + v.z * v.z));
    for (int o = 32; o > 0; o >>= 1) m = max(m, (unsigned)__shfl_xor((int)m, o));
    if ((threadIdx.x & 63) == 0) atomicMax(vmaxBits, m);
}
// One axis of the block-mask dilation: dst[b] = OR of src over blocks b +- k along
// the axis, k up to the reach of a back-trace in blocks: ceil((dt vmax n + 2.5) / 8)
// (a trilinear sample reaches 1.5 voxels past the back-traced point; 1 more covers
// rounding). A reach beyond 8 blocks marks everything.
__global__ __launch_bounds__(kThreads) void kMaskDilate(const uint8_t* __restrict__ src, uint8_t* __restrict__ dst,
                                                        int nb, int axis, float dtn,
                                                        const unsigned* __restrict__ vmaxBits) {
    const size_t total = (size_t)nb * nb * nb;
    const float reach = dtn * __uint_as_float(*vmaxBits) + 2.5f;
    const int r = (int)ceilf(reach / 8.f);
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
        if (!(reach <= 64.f)) { dst[i] = 1; continue; }
        const int c[3] = {(int)(i % nb), (int)((i / nb) % nb), (int)(i / ((size_t)nb * nb))};
        const size_t stride = axis == 0 ? 1 : (axis == 1 ? (size_t)nb : (size_t)nb * nb);
        uint8_t any = 0;
        for (int k = -r; k <= r; k++) {
            const int q = c[axis] + k;
            if (q < 0 || q >= nb) continue;
            any |= src[(size_t)((ptrdiff_t)i + (ptrdiff_t)k * (ptrdiff_t)stride)];
        }
        dst[i] = any ? 1 : 0;
    }
}

__device__ __forceinline__ float4 velAt(const float4* __restrict__ vel, float px, float py, float pz) {
    const float gx = fminf(fmaxf(px * kVel - 0.5f, 0.f), kVel - 1.001f);
    const float gy = fminf(fmaxf(py * kVel - 0.5f, 0.f), kVel - 1.001f);
    const float gz = fminf(fmaxf(pz * kVel - 0.5f, 0.f), kVel - 1.001f);
    const int x = (int)gx, y = (int)gy, z = (int)gz;
    const float fx = gx - x, fy = gy - y, fz = gz - z;
    float4 r = make_float4(0.f, 0.f, 0.f, 0.f);
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++)
            for (int c = 0; c < 2; c++) {
                const float wgt = (a ? fx : 1.f - fx) * (b ? fy : 1.f - fy) * (c ? fz : 1.f - fz);
                const float4 v = vel[(x + a) + kVel * ((y + b) + kVel * (z + c))];
                r.x += wgt * v.x; r.y += wgt * v.y; r.z += wgt * v.z;
            }
    return r;
}
// trilinear sample of a field at fine-grid coordinates g (voxel centres at i + 0.5), 0 outside
__device__ __forceinline__ float sampleField(const float* __restrict__ f, int n, float gx, float gy, float gz) {
    gx -= 0.5f; gy -= 0.5f; gz -= 0.5f;
    const float fx0 = floorf(gx), fy0 = floorf(gy), fz0 = floorf(gz);
    const int x = (int)fx0, y = (int)fy0, z = (int)fz0;
    const float fx = gx - fx0, fy = gy - fy0, fz = gz - fz0;
    float r = 0.f;
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++)
            for (int c = 0; c < 2; c++) {
                const int xi = x + a, yi = y + b, zi = z + c;
                if (xi < 0 || yi < 0 || zi < 0 || xi >= n || yi >= n || zi >= n) continue;
                const float wgt = (a ? fx : 1.f - fx) * (b ? fy : 1.f - fy) * (c ? fz : 1.f - fz);
                r += wgt * f[(size_t)xi + (size_t)n * ((size_t)yi + (size_t)n * zi)];
            }
    return r;
}

__global__ __launch_bounds__(kThreads) void kAdvect(PlumeParams P, const float4* __restrict__ vel,
                                                    const float* __restrict__ d0, const float* __restrict__ t0,
                                                    float* __restrict__ d1, float* __restrict__ t1,
                                                    const uint8_t* __restrict__ reachable,
                                                    uint8_t* __restrict__ marks) {
    const int n = P.n, nb = n / 8;
    const size_t total = (size_t)n * n * n;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
        const int x = (int)(i % n), y = (int)((i / n) % n), z = (int)(i / ((size_t)n * n));
        const float px = (x + 0.5f) / n, py = (y + 0.5f) / n, pz = (z + 0.5f) / n;
        const size_t blk = (size_t)(x / 8) + (size_t)nb * ((size_t)(y / 8) + (size_t)nb * (size_t)(z / 8));
        {
            const float sx = px - 0.5f, sy = py - P.srcY, sz = pz - 0.5f;
            if (!reachable[blk] && !(sqrtf(sx * sx + sy * sy + sz * sz) / P.srcR < 1.f)) {
                d1[i] = 0.f;
                t1[i] = 0.f;
                continue;
            }
        }
        // midpoint back-trace
        const float4 v1 = velAt(vel, px, py, pz);
        const float4 v2 = velAt(vel, px - 0.5f * P.dt * v1.x, py - 0.5f * P.dt * v1.y, pz - 0.5f * P.dt * v1.z);
        const float qx = (px - P.dt * v2.x) * n, qy = (py - P.dt * v2.y) * n, qz = (pz - P.dt * v2.z) * n;
        float dens = sampleField(d0, n, qx, qy, qz) * P.decay;
        float temp = sampleField(t0, n, qx, qy, qz) * P.cool;
        // the source: a noisy sphere at the base, densest and hottest at its core
        const float sx = px - 0.5f, sy = py - P.srcY, sz = pz - 0.5f;
        const float r = sqrtf(sx * sx + sy * sy + sz * sz) / P.srcR;
        if (r < 1.f) {
            const float nz = 0.5f + 0.5f * vnoise(px * 24.f, py * 24.f - 3.f * P.t, pz * 24.f, P.seed ^ 0x5BD1E995u);
            const float fall = (1.f - r * r);
            dens = fmaxf(dens, P.srcDensity * fall * nz);
            temp = fmaxf(temp, P.srcTemperature * fall * fall);
        }
        // the grid's outer voxels stay empty, so the plume never touches the box
        const bool edge = x < 2 || y < 2 || z < 2 || x >= n - 2 || y >= n - 2 || z >= n - 2;
        const bool zero = edge || dens < 1e-3f;
        d1[i] = zero ? 0.f : dens;
        t1[i] = zero ? 0.f : temp;
        if (!zero) marks[blk] = 1;
    }
}

}  // namespace

struct vrestir_plume {
    int n = 0;
    uint64_t seed = 0;
    int steps = 0;
    int device = 0;
    float* d[2] = {nullptr, nullptr};
    float* t[2] = {nullptr, nullptr};
    float4* vel = nullptr;
    uint8_t* marks[2] = {nullptr, nullptr};  // per 8^3 block: non-zero density in d[k]
    uint8_t* reach[2] = {nullptr, nullptr};  // dilation ping-pong; reach[0] ends as the reachable blocks
    unsigned* vmaxBits = nullptr;             // the largest lattice speed of the step
    int cur = 0;
};

extern "C" int32_t vrestir_plume_create(int32_t n, uint64_t seed, vrestir_plume** out) {
    if (!out) return VRESTIR_EINVAL;
    *out = nullptr;
    if (n < 16 || n > 2048 || n % 8) return VRESTIR_EINVAL;
    vrestir_plume* p = new vrestir_plume();
    p->n = n;
    p->seed = seed;
    (void)hipGetDevice(&p->device);
    const size_t bytes = (size_t)n * n * n * sizeof(float);
    bool ok = true;
    for (int k = 0; k < 2 && ok; k++) {
        ok = ok && hipMalloc((void**)&p->d[k], bytes) == hipSuccess;
        ok = ok && hipMalloc((void**)&p->t[k], bytes) == hipSuccess;
    }
    ok = ok && hipMalloc((void**)&p->vel, (size_t)kVel * kVel * kVel * sizeof(float4)) == hipSuccess;
    const size_t nb3 = (size_t)(n / 8) * (n / 8) * (n / 8);
    for (int k = 0; k < 2 && ok; k++) {
        ok = ok && hipMalloc((void**)&p->marks[k], nb3) == hipSuccess && hipMemset(p->marks[k], 0, nb3) == hipSuccess;
        ok = ok && hipMalloc((void**)&p->reach[k], nb3) == hipSuccess;
    }
    ok = ok && hipMalloc((void**)&p->vmaxBits, sizeof(unsigned)) == hipSuccess;
    ok = ok && hipMemset(p->d[0], 0, bytes) == hipSuccess && hipMemset(p->t[0], 0, bytes) == hipSuccess;
    if (!ok) {
        vrestir_plume_destroy(p);
        return VRESTIR_EHIP;
    }
    *out = p;
    return VRESTIR_OK;
}

extern "C" int32_t vrestir_plume_destroy(vrestir_plume* p) {
    if (!p) return VRESTIR_EINVAL;
    (void)hipSetDevice(p->device);
    for (int k = 0; k < 2; k++) {
        if (p->d[k]) (void)hipFree(p->d[k]);
        if (p->t[k]) (void)hipFree(p->t[k]);
    }
    if (p->vel) (void)hipFree(p->vel);
    for (int k = 0; k < 2; k++) {
        if (p->marks[k]) (void)hipFree(p->marks[k]);
        if (p->reach[k]) (void)hipFree(p->reach[k]);
    }
    if (p->vmaxBits) (void)hipFree(p->vmaxBits);
    delete p;
    return VRESTIR_OK;
}

extern "C" int32_t vrestir_plume_advance(vrestir_plume* p, void* stream, int32_t steps) {
    if (!p || steps < 0) return VRESTIR_EINVAL;
    if (hipSetDevice(p->device) != hipSuccess) return VRESTIR_EHIP;
    hipStream_t st = (hipStream_t)stream;
    for (int s = 0; s < steps; s++) {
        PlumeParams P;
        P.n = p->n;
        P.seed = (uint32_t)(p->seed ^ (p->seed >> 32));
        P.t = (float)p->steps;
        P.dt = 1.f;
        P.updraft = 0.015f;
        P.width = 0.16f;
        P.curl = 0.0012f;  // the potential's gradient is ~10 per unit: stirring ~0.012 per frame
        P.decay = 0.992f;
        P.cool = 0.965f;
        P.srcR = 0.11f;
        P.srcY = 0.14f;
        P.srcDensity = 1.f;
        P.srcTemperature = 8.f;
        if (hipMemsetAsync(p->vmaxBits, 0, sizeof(unsigned), st) != hipSuccess) return VRESTIR_EHIP;
        hipLaunchKernelGGL(kVelocity, dim3((kVel * kVel * kVel + kThreads - 1) / kThreads), dim3(kThreads), 0, st, P,
                           p->vel, p->vmaxBits);
        // the blocks a back-trace can reach from marked ones: x, y, z dilations
        const int nb = p->n / 8;
        const size_t nb3 = (size_t)nb * nb * nb;
        const unsigned mblocks = (unsigned)((nb3 + kThreads - 1) / kThreads);
        const float dtn = P.dt * (float)p->n;
        hipLaunchKernelGGL(kMaskDilate, dim3(mblocks), dim3(kThreads), 0, st, p->marks[p->cur], p->reach[1], nb, 0,
                           dtn, p->vmaxBits);
        hipLaunchKernelGGL(kMaskDilate, dim3(mblocks), dim3(kThreads), 0, st, p->reach[1], p->reach[0], nb, 1, dtn,
                           p->vmaxBits);
        hipLaunchKernelGGL(kMaskDilate, dim3(mblocks), dim3(kThreads), 0, st, p->reach[0], p->reach[1], nb, 2, dtn,
                           p->vmaxBits);
        if (hipMemsetAsync(p->marks[1 - p->cur], 0, nb3, st) != hipSuccess) return VRESTIR_EHIP;
        const size_t total = (size_t)p->n * p->n * p->n;
        const unsigned blocks = (unsigned)((total + kThreads - 1) / kThreads < 262144 ? (total + kThreads - 1) / kThreads
                                                                                      : 262144);
        hipLaunchKernelGGL(kAdvect, dim3(blocks), dim3(kThreads), 0, st, P, p->vel, p->d[p->cur], p->t[p->cur],
                           p->d[1 - p->cur], p->t[1 - p->cur], p->reach[1], p->marks[1 - p->cur]);
        if (hipGetLastError() != hipSuccess) return VRESTIR_EHIP;
        p->cur = 1 - p->cur;
        p->steps++;
    }
    return VRESTIR_OK;
}

extern "C" int32_t vrestir_plume_grids(const vrestir_plume* p, const float** density, const float** temperature,
                                       int32_t* steps) {
    if (!p) return VRESTIR_EINVAL;
    if (density) *density = p->d[p->cur];
    if (temperature) *temperature = p->t[p->cur];
    if (steps) *steps = p->steps;
    return VRESTIR_OK;
}

extern "C" int32_t vrestir_plume_read(const vrestir_plume* p, float* density, float* temperature) {
    if (!p) return VRESTIR_EINVAL;
    if (hipSetDevice(p->device) != hipSuccess || hipDeviceSynchronize() != hipSuccess) return VRESTIR_EHIP;
    const size_t bytes = (size_t)p->n * p->n * p->n * sizeof(float);
    if (density && hipMemcpy(density, p->d[p->cur], bytes, hipMemcpyDeviceToHost) != hipSuccess) return VRESTIR_EHIP;
    if (temperature && hipMemcpy(temperature, p->t[p->cur], bytes, hipMemcpyDeviceToHost) != hipSuccess)
        return VRESTIR_EHIP;
    return VRESTIR_OK;
}
