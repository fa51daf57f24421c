// vr_k_accum.hip -- ground-truth accumulation and the relMSE error metric.
//
// Accumulation keeps a per-pixel FP64 running sum of the pass's RGBA frames,
// the double-precision mode of the reference's AccumulatePass
// (Source/RenderPasses/AccumulatePass/AccumulatePass.cpp:42,93), so a 64k-frame
// reference image carries no float round-off from the sum itself.
//
// relMSE is the reference's "rmse" metric (Source/Tools/ImageCompare/
// ImageCompare.cpp:172-180, 202-219): per pixel, the mean over RGB of
// (a - b)^2 / (a^2 + 1e-3) with a the reference, averaged over pixels. Each
// workgroup sums its pixels in FP64 into its own slot; a second single-block
// pass adds the slots in a fixed order, so the result does not depend on
// scheduling.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>

#include "vr_launch.h"

namespace vr {

// Scratch of the allocation-free relMSE: kRelMseScratchDoubles doubles (256 KiB),
// the per-workgroup partial sums from the start and the result in the last slot,
// resident on every device the library runs on until the process ends. Host code never
// passes `scratch` itself to a kernel: its name is a host-side stand-in, and the
// device address comes from hipGetSymbolAddress, once per device.
static __device__ __align__(256) double scratch[kRelMseScratchDoubles];

__global__ __launch_bounds__(256) void kAccumulate(const float4* __restrict__ frame, double4* __restrict__ sum,
                                                   int n, int reset) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 f = frame[i];
    double4 s = reset ? make_double4(0.0, 0.0, 0.0, 0.0) : sum[i];
    s.x += (double)f.x; s.y += (double)f.y; s.z += (double)f.z; s.w += (double)f.w;
    sum[i] = s;
}

__global__ __launch_bounds__(256) void kResolve(const double4* __restrict__ sum, double inv, float4* __restrict__ out,
                                                int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double4 s = sum[i];
    out[i] = make_float4((float)(s.x * inv), (float)(s.y * inv), (float)(s.z * inv), (float)(s.w * inv));
}

__global__ __launch_bounds__(256) void kRelMsePartial(const float4* __restrict__ img, const float4* __restrict__ ref,
                                                      int n, double* __restrict__ partial) {
    __shared__ double red[256];
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    double e = 0.0;
    if (i < n) {
        float4 b = img[i], a = ref[i];
        float d0 = a.x - b.x, d1 = a.y - b.y, d2 = a.z - b.z;
        double px = (double)(d0 * d0) / ((double)(a.x * a.x) + 1e-3) + (double)(d1 * d1) / ((double)(a.y * a.y) + 1e-3) +
                    (double)(d2 * d2) / ((double)(a.z * a.z) + 1e-3);
        e = px / 3.0;
    }
    red[threadIdx.x] = e;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(256) void kSumPartials(const double* __restrict__ partial, int m, double scale,
                                                    double* __restrict__ out) {
    __shared__ double red[256];
    double s = 0.0;
    for (int k = threadIdx.x; k < m; k += 256) s += partial[k];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int t = 128; t > 0; t >>= 1) {
        if ((int)threadIdx.x < t) red[threadIdx.x] += red[threadIdx.x + t];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0] * scale;
}

hipError_t launchAccumulate(hipStream_t st, const float4* frame, double4* sum, int n, bool reset) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(kAccumulate, dim3((n + 255) / 256), dim3(256), 0, st, frame, sum, n, reset ? 1 : 0);
    return hipGetLastError();
}

hipError_t launchResolve(hipStream_t st, const double4* sum, double inv, float4* out, int n) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(kResolve, dim3((n + 255) / 256), dim3(256), 0, st, sum, inv, out, n);
    return hipGetLastError();
}

int relMsePartialCount(int n) { return (n + 255) / 256; }

hipError_t launchRelMse(hipStream_t st, const float4* img, const float4* ref, int n, double* partial, double* out) {
    if (n <= 0) return hipErrorInvalidValue;
    const int m = relMsePartialCount(n);
    hipLaunchKernelGGL(kRelMsePartial, dim3(m), dim3(256), 0, st, img, ref, n, partial);
    hipLaunchKernelGGL(kSumPartials, dim3(1), dim3(256), 0, st, partial, m, 1.0 / (double)n, out);
    return hipGetLastError();
}

// Device address of `scratch` on each device, looked up once per device with
// that device current (each device holds its own copy at its own address).
static std::mutex gScratchLock;
static double* gScratchPtr[kMaxDevices] = {};

hipError_t relMseScratchInit(double** ptrOut) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(gScratchLock);
    if (!gScratchPtr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        gScratchPtr[dev] = (double*)p;
    }
    if (ptrOut) *ptrOut = gScratchPtr[dev];
    return hipSuccess;
}

// relMSE through the same two kernels as launchRelMse, with the partial sums
// and the result in the current device's `scratch`; copies the result to *hostOut
// and waits for the stream. Nothing in `scratch` needs clearing: the first kernel
// writes every partial the second one reads. Calls that overlap on one device
// share the array and race.
hipError_t relMseNoAlloc(hipStream_t st, const float4* img, const float4* ref, int n, double* hostOut) {
    if (n <= 0 || relMsePartialCount(n) + 1 > kRelMseScratchDoubles) return hipErrorInvalidValue;
    double* base = nullptr;
    hipError_t e = relMseScratchInit(&base);
    if (e != hipSuccess) return e;
    double* out = base + (kRelMseScratchDoubles - 1);
    e = launchRelMse(st, img, ref, n, base, out);
    if (e != hipSuccess) return e;
    e = hipMemcpyAsync(hostOut, out, sizeof(double), hipMemcpyDeviceToHost, st);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(st);
}

}  // namespace vr
