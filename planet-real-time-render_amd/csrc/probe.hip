// probe.hip -- measurement probes for the roofline of the atmosphere kernels
// (libatmo_probe.so).  atmo_probe_exp_rate measures the sustained rate of
// v_exp_f32, the transcendental-class instruction that bounds the LUT
// builds and the composite, with independent dependency chains per lane so
// the issue rate rather than the latency is timed.  atmo_probe_copy is a
// plain device copy of known size: the reference point that scales the HBM
// traffic counters (FETCH_SIZE, WRITE_SIZE) of the other kernels.
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace {

constexpr int kChains = 8;

__global__ void __launch_bounds__(256) k_exp_chains(float* __restrict__ out, int iters) {
    float a[kChains];
#pragma unroll
    for (int c = 0; c < kChains; ++c) a[c] = 1.0e-3f * (float)(threadIdx.x + c);
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int c = 0; c < kChains; ++c) a[c] = __builtin_amdgcn_exp2f(-a[c]);  // x -> 2^-x stays in (0, 1]
    }
    float s = 0.0f;
#pragma unroll
    for (int c = 0; c < kChains; ++c) s += a[c];
    out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

// Grid-stride copy of n 16-byte elements.
__global__ void __launch_bounds__(256) k_copy(const float4* __restrict__ src,
                                              float4* __restrict__ dst, long n) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

}  // namespace

extern "C" {

// Copies `bytes` (a multiple of 16) from src to dst on `stream` with
// k_copy: exactly `bytes` read and `bytes` written.  Returns 0 on success,
// otherwise a hipError_t value (hipErrorInvalidValue for a bad size).
int32_t atmo_probe_copy(const void* src, void* dst, int64_t bytes, void* stream) {
    if (!src || !dst || bytes <= 0 || bytes % 16 != 0) return (int32_t)hipErrorInvalidValue;
    const long n = (long)(bytes / 16);
    hipLaunchKernelGGL(k_copy, dim3(4096), dim3(256), 0, (hipStream_t)stream,
                       (const float4*)src, (float4*)dst, n);
    return (int32_t)hipGetLastError();
}

// Runs `blocks` x 256 threads, each `iters` x 8 exp2 evaluations, on the
// null stream; writes the achieved exp2 rate (per second) to *rate.
// Returns 0 on success, otherwise a hipError_t value.
int32_t atmo_probe_exp_rate(int32_t blocks, int32_t iters, double* rate) {
    float* out = nullptr;
    hipEvent_t start, stop;
    hipError_t e = hipMalloc(&out, (size_t)blocks * 256 * sizeof(float));
    if (e != hipSuccess) return (int32_t)e;
    (void)hipEventCreate(&start);
    (void)hipEventCreate(&stop);
    hipLaunchKernelGGL(k_exp_chains, dim3(blocks), dim3(256), 0, 0, out, iters);  // warm-up
    (void)hipEventRecord(start, 0);
    hipLaunchKernelGGL(k_exp_chains, dim3(blocks), dim3(256), 0, 0, out, iters);
    (void)hipEventRecord(stop, 0);
    e = hipEventSynchronize(stop);
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, start, stop);
    (void)hipEventDestroy(start);
    (void)hipEventDestroy(stop);
    (void)hipFree(out);
    if (e != hipSuccess) return (int32_t)e;
    *rate = (double)blocks * 256.0 * (double)iters * kChains / (ms * 1.0e-3);
    return 0;
}

}  // extern "C"
