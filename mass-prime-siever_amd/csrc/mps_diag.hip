// Issue-rate and latency probes of the VALU operations the worker is built from (gfx950):
// FP64 FMA, 32-bit multiply low/high, 64-bit multiply-add (v_mad_u64_u32) and 32-bit add, each as
// 8 independent chains per lane over a full-chip grid, plus one dependent FP64 FMA chain per lane at
// one wavefront per SIMD (latency).  Used for the roofline peaks reported beside the benchmark.
#include <hip/hip_runtime.h>

#include "mps_gpu.h"

namespace {

constexpr int kChains = 8;

__global__ __launch_bounds__(256) void k_peak_fma64(double* out, int iters, double x, double y) {
    double a[kChains];
    for (int c = 0; c < kChains; ++c) a[c] = (double)(threadIdx.x + c);
    for (int i = 0; i < iters; ++i)
        for (int c = 0; c < kChains; ++c) a[c] = __fma_rn(a[c], x, y);
    double s = 0;
    for (int c = 0; c < kChains; ++c) s += a[c];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ __launch_bounds__(256) void k_peak_mullo(uint32_t* out, int iters, uint32_t k) {
    uint32_t a[kChains];
    for (int c = 0; c < kChains; ++c) a[c] = threadIdx.x * 2654435761u + c;
    for (int i = 0; i < iters; ++i)
        for (int c = 0; c < kChains; ++c) a[c] = a[c] * k;
    uint32_t s = 0;
    for (int c = 0; c < kChains; ++c) s ^= a[c];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ __launch_bounds__(256) void k_peak_mulhi(uint32_t* out, int iters, uint32_t k) {
    uint32_t a[kChains];
    for (int c = 0; c < kChains; ++c) a[c] = threadIdx.x * 2654435761u + c + 1;
    for (int i = 0; i < iters; ++i)
        for (int c = 0; c < kChains; ++c) a[c] = __umulhi(a[c], k) ^ (uint32_t)c;
    uint32_t s = 0;
    for (int c = 0; c < kChains; ++c) s ^= a[c];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ __launch_bounds__(256) void k_peak_mad64(uint64_t* out, int iters, uint32_t k) {
    uint64_t a[kChains];
    for (int c = 0; c < kChains; ++c) a[c] = threadIdx.x * 2654435761u + c;
    for (int i = 0; i < iters; ++i)
        for (int c = 0; c < kChains; ++c) a[c] = (uint64_t)(uint32_t)a[c] * k + (a[c] >> 32);
    uint64_t s = 0;
    for (int c = 0; c < kChains; ++c) s ^= a[c];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ __launch_bounds__(256) void k_peak_add32(uint32_t* out, int iters, uint32_t k) {
    uint32_t a[kChains];
    for (int c = 0; c < kChains; ++c) a[c] = threadIdx.x + c;
    for (int i = 0; i < iters; ++i)
        for (int c = 0; c < kChains; ++c) a[c] = (a[c] + k) ^ (uint32_t)c;
    uint32_t s = 0;
    for (int c = 0; c < kChains; ++c) s ^= a[c];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ __launch_bounds__(64) void k_lat_fma64(double* out, int iters, double x, double y) {
    double a = (double)threadIdx.x;
    for (int i = 0; i < iters; ++i) a = __fma_rn(a, x, y);
    out[blockIdx.x * blockDim.x + threadIdx.x] = a;
}

// Streams n16 16-byte words from src to dst: reads and writes n16 * 16 bytes of HBM once each.
__global__ __launch_bounds__(256) void k_copy(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n16) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

}  // namespace

extern "C" int mps_gpu_copy_probe(int device_index, uint64_t bytes, int reps, double* gbps) {
    if (gbps == nullptr || reps <= 0 || bytes < 16) return MPS_ERROR;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device_index < 0 || device_index >= ndev) return MPS_ERROR;
    if (hipSetDevice(device_index) != hipSuccess) return MPS_ERROR;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device_index) != hipSuccess) return MPS_ERROR;
    const size_t n16 = bytes / 16;
    uint4 *src = nullptr, *dst = nullptr;
    if (hipMalloc(&src, n16 * 16) != hipSuccess) return MPS_ERROR;
    if (hipMalloc(&dst, n16 * 16) != hipSuccess) {
        hipFree(src);
        return MPS_ERROR;
    }
    hipMemset(src, 1, n16 * 16);
    const uint32_t grid = (uint32_t)prop.multiProcessorCount * 16;
    hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, 0, src, dst, n16);  // warm
    hipEvent_t e0, e1;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    hipEventRecord(e0, 0);
    for (int r = 0; r < reps; ++r) hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, 0, src, dst, n16);
    hipEventRecord(e1, 0);
    hipEventSynchronize(e1);
    float ms = 0;
    hipEventElapsedTime(&ms, e0, e1);
    *gbps = 2.0 * (double)n16 * 16 * reps / (ms * 1e-3) / 1e9;
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    const bool ok = hipGetLastError() == hipSuccess;
    hipFree(src);
    hipFree(dst);
    return ok ? MPS_OK : MPS_ERROR;
}

extern "C" int mps_gpu_measure_valu(int device_index, mps_valu_rates* r) {
    if (r == nullptr) return MPS_ERROR;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device_index < 0 || device_index >= ndev) return MPS_ERROR;
    if (hipSetDevice(device_index) != hipSuccess) return MPS_ERROR;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device_index) != hipSuccess) return MPS_ERROR;
    const int cus = prop.multiProcessorCount;
    const int blocks = cus * 8, threads = 256, iters = 2048;
    const size_t n = (size_t)blocks * threads;
    void* buf = nullptr;
    if (hipMalloc(&buf, n * sizeof(uint64_t)) != hipSuccess) return MPS_ERROR;
    hipEvent_t e0, e1;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    auto timed = [&](auto launch) {
        launch();  // warm
        hipEventRecord(e0, 0);
        launch();
        hipEventRecord(e1, 0);
        hipEventSynchronize(e1);
        float ms = 0;
        hipEventElapsedTime(&ms, e0, e1);
        return (double)ms * 1e-3;
    };
    const double ops = (double)n * iters * kChains;
    double t;
    t = timed([&] { hipLaunchKernelGGL(k_peak_fma64, dim3(blocks), dim3(threads), 0, 0, (double*)buf, iters, 0.999999, 1e-3); });
    r->fma64_per_s = ops / t;
    t = timed([&] { hipLaunchKernelGGL(k_peak_mullo, dim3(blocks), dim3(threads), 0, 0, (uint32_t*)buf, iters, 2654435761u); });
    r->mul_lo32_per_s = ops / t;
    t = timed([&] { hipLaunchKernelGGL(k_peak_mulhi, dim3(blocks), dim3(threads), 0, 0, (uint32_t*)buf, iters, 2654435761u); });
    r->mul_hi32_per_s = ops / t;
    t = timed([&] { hipLaunchKernelGGL(k_peak_mad64, dim3(blocks), dim3(threads), 0, 0, (uint64_t*)buf, iters, 2654435761u); });
    r->mad_u64_u32_per_s = ops / t;
    t = timed([&] { hipLaunchKernelGGL(k_peak_add32, dim3(blocks), dim3(threads), 0, 0, (uint32_t*)buf, iters, 2654435761u); });
    r->add32_per_s = ops / t;
    // one wavefront per SIMD, one dependent chain per lane
    const int liters = 1 << 16;
    t = timed([&] { hipLaunchKernelGGL(k_lat_fma64, dim3(cus * 4), dim3(64), 0, 0, (double*)buf, liters, 0.999999, 1e-3); });
    r->fma64_latency_cycles = t * (double)prop.clockRate * 1e3 / liters;
    r->clock_mhz = prop.clockRate / 1000.0;
    r->cus = cus;
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    const bool ok = hipGetLastError() == hipSuccess;
    hipFree(buf);
    return ok ? MPS_OK : MPS_ERROR;
}
