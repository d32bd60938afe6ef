// mc_probe.hip - device throughput probe for the atomic roofline of the
// transport: FP64 (or FP32) atomic adds to uniformly random words of an
// array with the footprint of the path-length estimator.  The transport's
// J update is one such read-modify-write per segment, at addresses that are
// effectively random across the 64 lanes of a wavefront; this kernel issues
// the same requests with a chosen number of active lanes per wavefront, so
// its rate is the ceiling the transport's atomics are measured against.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mocassin_amd.h"

namespace {

// SplitMix64 finaliser (Steele, Lea & Flood 2014): request id -> word
__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

template <bool FP64>
__global__ void __launch_bounds__(256) atomics_kernel(void *buf, uint32_t words, uint64_t requests, int lanes,
                                                      uint64_t seed)
{
    const int lane = threadIdx.x & 63;
    const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    if (lane >= lanes) return;
    const uint64_t stride = nwaves * (uint64_t)lanes;
    for (uint64_t r = wave * (uint64_t)lanes + (uint64_t)lane; r < requests; r += stride) {
        const uint64_t h = mix64(r ^ seed);
        const uint32_t idx = (uint32_t)(((h >> 32) * (uint64_t)words) >> 32);     // < words
        if (FP64) unsafeAtomicAdd((double *)buf + idx, 1.0);
        else unsafeAtomicAdd((float *)buf + idx, 1.0f);
    }
}

// FP64 adds where the lanes of each atomic instruction form 64 / group
// sets of `group` lanes, each set adding to distinct words of one random
// 64-byte line: adjacent lanes (lane / group) or lanes spread across the
// wavefront (lane % (64 / group)).  A request carries the adds of one line.
__global__ void __launch_bounds__(256) line_kernel(double *buf, uint32_t lines, uint64_t requests, int group,
                                                   int spread, uint64_t seed)
{
    const int lane = threadIdx.x & 63;
    const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    const int sets = 64 / group;
    const int g = spread ? lane % sets : lane / group;      // the lane's line in this instruction
    const int w = spread ? lane / sets : lane % group;      // its word in the line, < group <= 8
    for (uint64_t r = wave * 64; r < requests; r += nwaves * 64) {
        const uint64_t h = mix64((r + (uint64_t)g) ^ seed);
        const uint32_t line = (uint32_t)(((h >> 32) * (uint64_t)lines) >> 32);     // < lines
        unsafeAtomicAdd(buf + (size_t)line * 8 + w, 1.0);
    }
}

}  // namespace

extern "C" int mc_probe_atomic_lines(int32_t device, int64_t words, int64_t requests, int32_t group,
                                     int32_t spread, double *kernel_ms)
{
    if (words < 8 || words / 8 > 0xffffffffLL || requests < 64 || !kernel_ms ||
        !(group == 1 || group == 2 || group == 4 || group == 8) || (spread != 0 && spread != 1))
        return MC_EINVAL;
    if (hipSetDevice(device) != hipSuccess) return MC_ENODEV;
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || cus < 1)
        return MC_ENODEV;
    const uint32_t lines = (uint32_t)(words / 8);
    double *buf = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = MC_OK;
    if (hipMalloc(&buf, (size_t)lines * 64) != hipSuccess) return MC_ENOMEM;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
        hipEventCreate(&e1) != hipSuccess || hipMemsetAsync(buf, 0, (size_t)lines * 64, s) != hipSuccess) {
        rc = MC_ENODEV;
    } else {
        const dim3 grid((unsigned)cus * 8), block(256);
        float ms = 0.f;
        for (int k = 0; k < 2 && rc == MC_OK; ++k) {
            if (k == 1 && hipEventRecord(e0, s) != hipSuccess) rc = MC_ENODEV;
            hipLaunchKernelGGL(line_kernel, grid, block, 0, s, buf, lines, (uint64_t)requests, group, spread,
                               (uint64_t)(k + 1));
            if (hipGetLastError() != hipSuccess) rc = MC_ENODEV;
        }
        if (rc == MC_OK && (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                            hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
            rc = MC_ENODEV;
        *kernel_ms = ms;
    }
    if (s) (void)hipStreamSynchronize(s);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (s) (void)hipStreamDestroy(s);
    (void)hipFree(buf);
    return rc;
}

extern "C" int mc_probe_atomics(int32_t device, int64_t words, int64_t requests, int32_t lanes, int32_t fp64,
                                double *kernel_ms)
{
    if (words < 1 || words > 0xffffffffLL || requests < 1 || lanes < 1 || lanes > 64 || !kernel_ms)
        return MC_EINVAL;
    if (hipSetDevice(device) != hipSuccess) return MC_ENODEV;
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || cus < 1)
        return MC_ENODEV;
    const size_t esz = fp64 ? sizeof(double) : sizeof(float);
    void *buf = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = MC_OK;
    if (hipMalloc(&buf, (size_t)words * esz) != hipSuccess) return MC_ENOMEM;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
        hipEventCreate(&e1) != hipSuccess || hipMemsetAsync(buf, 0, (size_t)words * esz, s) != hipSuccess) {
        rc = MC_ENODEV;
    } else {
        // 8 workgroups of 4 wavefronts per CU; one untimed launch first
        const dim3 grid((unsigned)cus * 8), block(256);
        float ms = 0.f;
        for (int k = 0; k < 2 && rc == MC_OK; ++k) {
            if (k == 1 && hipEventRecord(e0, s) != hipSuccess) rc = MC_ENODEV;
            if (fp64) hipLaunchKernelGGL(atomics_kernel<true>, grid, block, 0, s, buf, (uint32_t)words,
                                         (uint64_t)requests, lanes, (uint64_t)(k + 1));
            else hipLaunchKernelGGL(atomics_kernel<false>, grid, block, 0, s, buf, (uint32_t)words,
                                    (uint64_t)requests, lanes, (uint64_t)(k + 1));
            if (hipGetLastError() != hipSuccess) rc = MC_ENODEV;
        }
        if (rc == MC_OK && (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                            hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
            rc = MC_ENODEV;
        *kernel_ms = ms;
    }
    if (s) (void)hipStreamSynchronize(s);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (s) (void)hipStreamDestroy(s);
    (void)hipFree(buf);
    return rc;
}
