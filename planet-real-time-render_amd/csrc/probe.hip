// probe.hip -- measurement probes for the roofline of the atmosphere kernels
// (libatmo_probe.so).  atmo_probe_exp_rate measures the sustained rate of
// v_exp_f32, the transcendental-class instruction that bounds the LUT
// builds and the composite, with independent dependency chains per lane so
// the issue rate rather than the latency is timed.  atmo_probe_copy is a
// plain device copy of known size: the reference point that scales the HBM
// traffic counters (FETCH_SIZE, WRITE_SIZE) of the other kernels.
// atmo_probe_arith_check compares the short correctly rounded square root,
// reciprocal and quotient of the atmosphere kernels (atmo::csqrt, crcp,
// cdiv) with the compiler's IEEE expansions over their stated domains, and
// the kernels' float64 acos (atmo::acos_d) rounded to float32 with the
// device library's acos rounded the same way.  atmo_probe_srgb_pow
// evaluates the post-process's sRGB power (post::srgb_pow) on an array, so
// a test can compare it with the CPU oracle on every float of its domain.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "atmo_device.h"
#include "post_device.h"

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

// counts[0]: csqrt != sqrtf over all 2^32 floats within csqrt's domain
// (+-0, NaN, [2^-96, 2^126)); counts[1]: crcp != 1.0f / b over
// every float with |b| in [2^-100, 2^101); counts[2]: cdiv != a / b over
// hashed pairs with both exponents in [-60, 60] (2^32 pairs); counts[3]:
// csqrt != sqrtf on (0, 2^-96), beyond 2^126 and for negative x, for
// information; counts[4]: float(acos_d(x)) != float(acos(x)) over every
// float x in [-1, 1]; counts[5]: cr_sincos(x) != (cr_sin(x), cr_cos(x))
// over every float x in [0, 256).
__device__ __forceinline__ bool same_bits(float a, float b) {
    return __float_as_uint(a) == __float_as_uint(b) || (a != a && b != b);
}

__global__ void __launch_bounds__(256) k_check_sqrt(unsigned long long* counts, uint32_t base) {
    const uint32_t bits = base + blockIdx.x * blockDim.x + threadIdx.x;
    const float x = __uint_as_float(bits);
    if (same_bits(atmo::csqrt(x), sqrtf(x))) return;
    const bool domain = x == 0.0f || x != x || (x >= 0x1p-96f && x < 0x1p126f);
    atomicAdd(&counts[domain ? 0 : 3], 1ull);
}

__global__ void __launch_bounds__(256) k_check_rcp(unsigned long long* counts, int exponent) {
    const uint32_t bits = ((uint32_t)(exponent + 127) << 23) | (blockIdx.x * blockDim.x + threadIdx.x);
    const float b = __uint_as_float(bits);
    if (!same_bits(atmo::crcp(b), 1.0f / b)) atomicAdd(&counts[1], 1ull);
    if (!same_bits(atmo::crcp(-b), 1.0f / -b)) atomicAdd(&counts[1], 1ull);
}

__device__ __forceinline__ uint32_t mix(uint64_t x) {
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdull;
    x ^= x >> 33;
    x *= 0xc4ceb9fe1a85ec53ull;
    x ^= x >> 33;
    return (uint32_t)x;
}

__global__ void __launch_bounds__(256) k_check_div(unsigned long long* counts, uint64_t base) {
    const uint64_t i = base + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t ha = mix(2 * i), hb = mix(2 * i + 1);
    const uint32_t ea = (ha >> 23) % 121u, eb = (hb >> 23) % 121u;
    const float a = __uint_as_float((ha & 0x807fffffu) | ((ea + 67u) << 23));
    const float b = __uint_as_float((hb & 0x807fffffu) | ((eb + 67u) << 23));
    if (!same_bits(atmo::cdiv(a, b), a / b)) atomicAdd(&counts[2], 1ull);
}

// bits in [0, 0x3f800000] (x in [0, 1]) with the given sign bit; one
// atomic per wave with a mismatch
__global__ void __launch_bounds__(256) k_check_acos(unsigned long long* counts, uint32_t base,
                                                    uint32_t sign) {
    const uint32_t i = base + blockIdx.x * blockDim.x + threadIdx.x;
    const float x = __uint_as_float((i <= 0x3f800000u ? i : 0u) | sign);
    const bool miss = i <= 0x3f800000u && !same_bits(atmo::cr_acos(x), (float)acos((double)x));
    const unsigned long long m = __ballot(miss);
    if ((threadIdx.x & 63) == 0 && m) atomicAdd(&counts[4], (unsigned long long)__popcll(m));
}

// every float in [0, 256): bit patterns [0, 0x43800000)
__global__ void __launch_bounds__(256) k_check_sincos(unsigned long long* counts, uint32_t base) {
    const uint32_t i = base + blockIdx.x * blockDim.x + threadIdx.x;
    const float x = __uint_as_float(i < 0x43800000u ? i : 0u);
    float s, c;
    atmo::cr_sincos(x, s, c);
    const bool miss = i < 0x43800000u &&
                      !(same_bits(s, atmo::cr_sin(x)) && same_bits(c, atmo::cr_cos(x)));
    const unsigned long long m = __ballot(miss);
    if ((threadIdx.x & 63) == 0 && m) atomicAdd(&counts[5], (unsigned long long)__popcll(m));
}

}  // namespace

// The display transform's sRGB power, one input per lane.
__global__ void __launch_bounds__(256) k_srgb_pow(const float* __restrict__ in,
                                                  float* __restrict__ out, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) out[i] = post::srgb_pow(in[i]);
}

extern "C" {

// Runs the comparisons above on the null stream and copies the six
// mismatch counts to counts[0..5].  Returns 0 on success, otherwise a
// hipError_t value.
int32_t atmo_probe_arith_check(uint64_t* counts) {
    unsigned long long* d = nullptr;
    hipError_t e = hipMalloc(&d, 6 * sizeof(unsigned long long));
    if (e != hipSuccess) return (int32_t)e;
    e = hipMemset(d, 0, 6 * sizeof(unsigned long long));
    // 2^32 sqrt inputs: 16 launches of 2^20 blocks x 256 threads
    for (uint32_t c = 0; c < 16 && e == hipSuccess; ++c) {
        hipLaunchKernelGGL(k_check_sqrt, dim3(1u << 20), dim3(256), 0, 0, d, c << 28);
        e = hipGetLastError();
    }
    // 2^23 mantissas per binade: 2^15 blocks x 256 threads
    for (int x = -100; x <= 100 && e == hipSuccess; ++x) {
        hipLaunchKernelGGL(k_check_rcp, dim3(1u << 15), dim3(256), 0, 0, d, x);
        e = hipGetLastError();
    }
    for (uint64_t c = 0; c < 16 && e == hipSuccess; ++c) {
        hipLaunchKernelGGL(k_check_div, dim3(1u << 20), dim3(256), 0, 0, d, c << 28);
        e = hipGetLastError();
    }
    // 0x3f800001 patterns per sign: 4 launches of 2^20 blocks x 256 threads cover [0, 2^30)
    for (int sg = 0; sg < 2 && e == hipSuccess; ++sg) {
        for (uint32_t c = 0; c < 4 && e == hipSuccess; ++c) {
            hipLaunchKernelGGL(k_check_acos, dim3(1u << 20), dim3(256), 0, 0, d, c << 28,
                               sg ? 0x80000000u : 0u);
            e = hipGetLastError();
        }
    }
    // 0x43800000 patterns: 5 launches of 2^28 threads cover [0, 5 * 2^28)
    for (uint32_t c = 0; c < 5 && e == hipSuccess; ++c) {
        hipLaunchKernelGGL(k_check_sincos, dim3(1u << 20), dim3(256), 0, 0, d, c << 28);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipDeviceSynchronize();
    unsigned long long h[6] = {0, 0, 0, 0, 0, 0};
    if (e == hipSuccess) e = hipMemcpy(h, d, sizeof h, hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (e != hipSuccess) return (int32_t)e;
    for (int i = 0; i < 6; ++i) counts[i] = (uint64_t)h[i];
    return 0;
}

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

// out[i] = post::srgb_pow(in[i]) for i < n (device pointers, inputs in
// (0.0031308, 1]) on `stream`.  Returns 0 on success, otherwise a
// hipError_t value.
int32_t atmo_probe_srgb_pow(const float* in, float* out, int64_t n, void* stream) {
    if (!in || !out || n <= 0 || n > ((int64_t)1 << 36)) return (int32_t)hipErrorInvalidValue;
    hipLaunchKernelGGL(k_srgb_pow, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       (hipStream_t)stream, in, out, n);
    return (int32_t)hipGetLastError();
}

}  // extern "C"
