// surrogate.hip -- the learned aerial-perspective surrogate of libatmo.so.
//
// A 5 -> 64 -> 64 -> 64 -> 6 SiLU MLP maps (camera altitude, view-zenith
// cosine, sun-zenith cosine, relative azimuth, segment length) to the RGB
// in-scatter per unit irradiance and the RGB transmittance of the
// composite's surface branch (renderer.py:1603-1786, 1863-1880).  This file
// holds:
//  - label generation by the GPU integrator itself (the surface_inscatter /
//    surface_transmittance device functions of the plain composite);
//  - FP32 training kernels: dense forward, backward (data and weights, the
//    weight gradient reduced over the batch in two deterministic stages)
//    and Adam;
//  - the surrogate composite: BF16 MFMA inference (v_mfma_f32_16x16x16bf16,
//    FP32 accumulation) for in-domain surface pixels, the plain integrator
//    for out-of-domain surface pixels, the plain sky branch for sky pixels.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdint.h>
#include <string.h>

#include <new>

#include "atmo_device.h"
#include "atmo_handle.h"

using namespace atmo;

namespace {

constexpr uint32_t kMlpMagic = 0x4d4c5053u;  // "MLPS"
constexpr int kChunk = 128;                  // samples per weight-gradient partial

struct Layer {
    int w, b, in, out;
};
constexpr Layer kLayers[4] = {
    {0, 320, 5, 64},
    {384, 4480, 64, 64},
    {4544, 8640, 64, 64},
    {8704, 9088, 64, 6},
};

}  // namespace

struct AtmoMlp {
    uint32_t magic;
    int max_batch;
    float* params;  // fp32 master weights [ATMO_MLP_PARAMS]
    float* grads;
    float* adam_m;
    float* adam_v;
    float* z[4];     // pre-activations per layer [B, out]
    float* a[3];     // hidden activations silu(z) [B, 64]
    float* dz[4];    // d loss / d z per layer [B, out]
    float* xbuf;     // gathered inputs [B, 5]
    float* wbuf;     // gathered loss weights [B, 6]
    float* partial;  // weight-gradient partial sums [B / kChunk, 64 * 64 + 64]
    float* loss;     // [B / 256 + 1] block partials, loss at [0] after reduce
    uint16_t* frag;  // bf16 MFMA weight fragments of layers 2-4 for inference
    float* frag32;   // fp32 MFMA weight fragments of layer 1
    float* bias;     // fp32 biases in inference order [64 + 64 + 64 + 16]
    AtmoSurrogateTransform transform;
    char error[512];
};

namespace {

bool mlp_ok(const AtmoMlp* m) { return m != nullptr && m->magic == kMlpMagic; }

__device__ __forceinline__ float silu(float z) { return z / (1.0f + expf(-z)); }

__device__ __forceinline__ float dsilu(float z) {
    float s = 1.0f / (1.0f + expf(-z));
    return s * (1.0f + z * (1.0f - s));
}

// ------------------------------------------------------------- features

// Canonical geometry of one sample: the in-scatter and transmittance of a
// ray from altitude h depend on the ray and the sun only through their
// zenith cosines and relative azimuth (spherically symmetric shell, origin
// on the local vertical).
__device__ __forceinline__ void sample_geometry(float mu, float mu_s, float phi, f3& ray, f3& sun) {
    float sr = sqrtf(fmaxf(1.0f - mu * mu, 0.0f));
    float ss = sqrtf(fmaxf(1.0f - mu_s * mu_s, 0.0f));
    ray = mk(sr, mu, 0.0f);
    sun = mk(ss * cr_cos(phi), mu_s, ss * cr_sin(phi));
}

// Cosine of the local horizon seen from altitude h above a planet of
// radius R: -sqrt(h (2R + h)) / (R + h).
__device__ __forceinline__ float horizon_cosine(float R, float h) {
    h = fmaxf(h, 0.0f);
    return -sqrtf(h * (2.0f * R + h)) / (R + h);
}

// Input features in [-1, 1]: log-stretched altitude, view and sun zenith
// cosines (with cosine_stretch s > 0 measured from the local horizon
// cosine c_h and stretched as asinh((c - c_h) / s) / asinh(2 / s), so the
// horizon kink sits at the same feature value at every altitude), relative
// azimuth, log segment length.
// The per-sample constants of the encoding at altitude h.
struct Encoding {
    float x0;      // altitude feature
    float ch;      // local horizon cosine
    float inv;     // 1 / cosine_stretch
    float norm;    // 1 / asinh(2 / cosine_stretch)
    float ld0;     // log(distance_min)
    float dscale;  // 2 / (log(distance_max) - log(distance_min))
};

__device__ __forceinline__ Encoding encoding_at(const AtmoSurrogateTransform& t, float h) {
    Encoding e;
    e.x0 = log1pf(fmaxf(h, 0.0f) / t.altitude_scale_m) *
               (2.0f / log1pf(t.altitude_max_m / t.altitude_scale_m)) -
           1.0f;
    e.ch = horizon_cosine(t.planet_radius_m, h);
    e.inv = t.cosine_stretch > 0.0f ? 1.0f / t.cosine_stretch : 0.0f;
    e.norm = t.cosine_stretch > 0.0f ? 1.0f / asinhf(2.0f * e.inv) : 0.0f;
    e.ld0 = logf(t.distance_min_m);
    e.dscale = 2.0f / (logf(t.distance_max_m) - e.ld0);
    return e;
}

__device__ __forceinline__ float feature_cosine(const AtmoSurrogateTransform& t, const Encoding& e,
                                                float c) {
    return t.cosine_stretch > 0.0f ? asinhf((c - e.ch) * e.inv) * e.norm : c;
}

__device__ __forceinline__ float feature_azimuth(float phi) {
    return phi * (2.0f / 3.14159265358979f) - 1.0f;
}

__device__ __forceinline__ float feature_distance(const Encoding& e, float d) {
    return (logf(d) - e.ld0) * e.dscale - 1.0f;
}

__device__ __forceinline__ void encode(const AtmoSurrogateTransform& t, float h, float mu,
                                       float mu_s, float phi, float d, float x[5]) {
    const Encoding e = encoding_at(t, h);
    x[0] = e.x0;
    x[1] = feature_cosine(t, e, mu);
    x[2] = feature_cosine(t, e, mu_s);
    x[3] = feature_azimuth(phi);
    x[4] = feature_distance(e, d);
}

__device__ __forceinline__ bool in_domain(const AtmoSurrogateTransform& t, float h, float d) {
    return h >= 0.0f && h <= t.altitude_max_m && d >= t.distance_min_m && d <= t.distance_max_m;
}

// ---------------------------------------------------------------- labels

__global__ void __launch_bounds__(256) k_labels(Consts k, long n, const float* __restrict__ raw,
                                                float bottom, float ar,
                                                const float* __restrict__ tlut,
                                                const float* __restrict__ mslut,
                                                float* __restrict__ labels) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* s = raw + i * 5;
    f3 ray, sun;
    sample_geometry(s[1], s[2], s[3], ray, sun);
    const f3 unit = mk(1.0f, 1.0f, 1.0f);
    f3 L = surface_inscatter(k, tlut, mslut, s[0], ray, s[4], sun, unit, bottom, ar);
    f3 T = surface_transmittance(k, tlut, s[0], ray, s[4], bottom);
    float* o = labels + i * 6;
    o[0] = L.x;
    o[1] = L.y;
    o[2] = L.z;
    o[3] = T.x;
    o[4] = T.y;
    o[5] = T.z;
}

__global__ void __launch_bounds__(256) k_encode(long n, const float* __restrict__ raw,
                                                AtmoSurrogateTransform t, float* __restrict__ x) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* s = raw + i * 5;
    float f[5];
    encode(t, s[0], s[1], s[2], s[3], s[4], f);
    for (int q = 0; q < 5; ++q) x[i * 5 + q] = f[q];
}

// ------------------------------------------------------------- training

// Gathers rows idx[s] of src [*, width] into dst [B, width] (idx may be NULL).
__global__ void __launch_bounds__(256) k_gather(int B, int width, const int32_t* __restrict__ idx,
                                                const float* __restrict__ src,
                                                float* __restrict__ dst) {
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long)B * width) return;
    const int s = (int)(e / width), c = (int)(e % width);
    const long row = idx ? (long)idx[s] : (long)s;
    dst[e] = src[row * width + c];
}

// z[s, n] = b[n] + sum_k x[s, k] W[n, k]; a = silu(z) when `a` is non-null.
// Thread per (sample, neuron); W staged transposed in LDS.
__global__ void __launch_bounds__(256) k_dense_forward(int B, int K, int N, const float* __restrict__ x,
                                                       const float* __restrict__ W,
                                                       const float* __restrict__ bias,
                                                       float* __restrict__ z, float* __restrict__ a) {
    __shared__ float wt[64 * 64];
    for (int e = threadIdx.x; e < K * N; e += blockDim.x) {
        int n = e / K, kk = e % K;
        wt[kk * N + n] = W[e];
    }
    __syncthreads();
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long)B * N) return;
    const int s = (int)(e / N), n = (int)(e % N);
    const float* xs = x + (long)s * K;
    float acc = bias[n];
    for (int kk = 0; kk < K; ++kk) acc += xs[kk] * wt[kk * N + n];
    z[e] = acc;
    if (a) a[e] = silu(acc);
}

// Weighted mean-square-error gradient dz = 2 w (y - t) / (B * 6) and
// per-block sums of w (y - t)^2.  `t` and `dz` may alias (the target is
// overwritten by its gradient); `w` NULL means unit weights.
__global__ void __launch_bounds__(256) k_mse(int B, const float* __restrict__ y, const float* t,
                                             const float* __restrict__ w, float* dz,
                                             float* __restrict__ block_loss) {
    __shared__ float red[256];
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long n = (long)B * 6;
    float d = 0.0f, we = 0.0f;
    if (e < n) {
        d = y[e] - t[e];
        we = w ? w[e] : 1.0f;
        dz[e] = 2.0f * we * d / (float)n;
    }
    red[threadIdx.x] = we * d * d;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) block_loss[blockIdx.x + 1] = red[0];
}

__global__ void __launch_bounds__(256) k_loss_reduce(int blocks, long count, float* __restrict__ block_loss) {
    __shared__ float red[256];
    float acc = 0.0f;
    for (int b = threadIdx.x; b < blocks; b += blockDim.x) acc += block_loss[b + 1];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) block_loss[0] = red[0] / (float)count;
}

// dz_prev[s, k] = (sum_n dz[s, n] W[n, k]) * silu'(z_prev[s, k]).
__global__ void __launch_bounds__(256) k_dense_backward_data(int B, int N, int K,
                                                             const float* __restrict__ dz,
                                                             const float* __restrict__ W,
                                                             const float* __restrict__ z_prev,
                                                             float* __restrict__ dz_prev) {
    __shared__ float w[64 * 64];
    for (int e = threadIdx.x; e < K * N; e += blockDim.x) w[e] = W[e];
    __syncthreads();
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long)B * K) return;
    const int s = (int)(e / K), kk = (int)(e % K);
    const float* d = dz + (long)s * N;
    float acc = 0.0f;
    for (int n = 0; n < N; ++n) acc += d[n] * w[n * K + kk];
    dz_prev[e] = acc * dsilu(z_prev[e]);
}

// Partial weight / bias gradients of one chunk of kChunk samples:
// partial[c, n * K + k] = sum_s dz[s, n] x[s, k], partial[c, N * K + n] = sum_s dz[s, n].
__global__ void __launch_bounds__(256) k_dense_backward_weights(int B, int N, int K,
                                                                const float* __restrict__ dz,
                                                                const float* __restrict__ x,
                                                                float* __restrict__ partial) {
    __shared__ float dzs[kChunk * 64];
    __shared__ float xs[kChunk * 64];
    const int c = blockIdx.x;
    const int s0 = c * kChunk;
    const int rows = min(kChunk, B - s0);
    for (int e = threadIdx.x; e < rows * N; e += blockDim.x) dzs[e] = dz[(long)s0 * N + e];
    for (int e = threadIdx.x; e < rows * K; e += blockDim.x) xs[e] = x[(long)s0 * K + e];
    __syncthreads();
    const int count = N * K + N;
    float* out = partial + (long)c * count;
    for (int e = threadIdx.x; e < count; e += blockDim.x) {
        float acc = 0.0f;
        if (e < N * K) {
            const int n = e / K, kk = e % K;
            for (int s = 0; s < rows; ++s) acc += dzs[s * N + n] * xs[s * K + kk];
        } else {
            const int n = e - N * K;
            for (int s = 0; s < rows; ++s) acc += dzs[s * N + n];
        }
        out[e] = acc;
    }
}

// grad[e] = sum over chunks of partial[c, e], in chunk order.
__global__ void __launch_bounds__(256) k_reduce_partials(int chunks, int count,
                                                         const float* __restrict__ partial,
                                                         float* __restrict__ grad) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= count) return;
    float acc = 0.0f;
    for (int c = 0; c < chunks; ++c) acc += partial[(long)c * count + e];
    grad[e] = acc;
}

// Adam (Kingma & Ba 2015, Algorithm 1) with bias corrections bc1 = 1 - b1^t,
// bc2 = 1 - b2^t.
__global__ void __launch_bounds__(256) k_adam(int n, float* __restrict__ p, const float* __restrict__ g,
                                              float* __restrict__ m, float* __restrict__ v, float lr,
                                              float b1, float b2, float eps, float bc1, float bc2) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    float gi = g[e];
    float mi = b1 * m[e] + (1.0f - b1) * gi;
    float vi = b2 * v[e] + (1.0f - b2) * gi * gi;
    m[e] = mi;
    v[e] = vi;
    p[e] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

// -------------------------------------------------------- MFMA inference

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ short to_bf16(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even (finite inputs)
    return (short)(u >> 16);
}

typedef __bf16 bf16v4 __attribute__((ext_vector_type(4)));

// Four floats to packed bf16, round to nearest even: two v_cvt_pk_bf16_f32
// on gfx950.
__device__ __forceinline__ bf16x4 pack_bf16(floatx4 v) {
    return __builtin_bit_cast(bf16x4, __builtin_convertvector(v, bf16v4));
}

__host__ __device__ constexpr int frag_index(int layer, int mt, int kt) {
    // layers 1-2: 4 row tiles x 2 K-steps of 32, layer 3: 1 x 2 bf16
    // fragments (layer 0 runs in fp32)
    return layer == 1 ? mt * 2 + kt : layer == 2 ? 8 + mt * 2 + kt : 16 + kt;
}
constexpr int kFragments = 18;
// fp32 fragments of layer 0: [0, 8) the 5-input layer (4 row tiles x 2
// K-steps of 4); [8, 12) the 3 per-pixel inputs (view cosine, azimuth,
// distance) of the composite, whose altitude and sun cosine are frame
// constants folded into the bias.
constexpr int kFragments32 = 12;
constexpr int kFolded32 = 8;
// 16-pixel chunks the surrogate composite's MLP runs at once (1, 2 or 4).
#ifndef ATMO_SURROGATE_CHUNKS
#define ATMO_SURROGATE_CHUNKS 1
#endif
// At least four waves per SIMD (128 VGPRs): the fused sky branch would
// otherwise take the kernel to 134 VGPRs and three waves (measured 198 ->
// 171 us at 1080p).
#ifndef ATMO_SURROGATE_WAVES
#define ATMO_SURROGATE_WAVES __attribute__((amdgpu_waves_per_eu(4, 8)))
#endif
// Persistent blocks of the surrogate composite (at most; one per 256-pixel
// group on small frames).
#ifndef ATMO_SURROGATE_GRID
#define ATMO_SURROGATE_GRID 2048
#endif

typedef short bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8v __attribute__((ext_vector_type(8)));

// The neuron a K-step position holds.  The output tile mt of a layer leaves
// lane l with neurons 16 mt + 4 (l / 16) + i (i = 0..3) of pixel l % 16;
// the B operand of K-step kt of v_mfma_f32_16x16x32_bf16 takes 8 values per
// lane at K positions 8 (l / 16) + i'.  Lane l's 8 values are the outputs
// of tiles 2 kt and 2 kt + 1, so K position 8 g + i' stands for neuron
// 32 kt + 4 g + i' (i' < 4) or 32 kt + 16 + 4 g + i' - 4 (i' >= 4); the
// weights are packed with the same permutation, which leaves the sum over K
// unchanged and chains the layers with no data movement.
__host__ __device__ constexpr int k_neuron(int kt, int g, int i) {
    return 32 * kt + (i < 4 ? 4 * g + i : 16 + 4 * g + (i - 4));
}

// Packs fp32 parameters into per-lane A-operand fragments.  Layers 1-3,
// v_mfma_f32_16x16x32_bf16: for tile (mt, kt) lane l holds
// W[16 mt + l % 16][k_neuron(kt, l / 16, i)], i = 0..7, in bf16 (zero
// outside W).  Layer 0, v_mfma_f32_16x16x4f32: general form, lane l holds
// W[16 mt + l % 16][4 kt + l / 16] in fp32; folded form, lane l holds
// W[16 mt + l % 16][c] with c = 1, 3, 4 for l / 16 = 0, 1, 2 (0 for 3).
// Also gathers the biases in output-row order.
__global__ void k_pack_fragments(const float* __restrict__ params, uint16_t* __restrict__ frag,
                                 float* __restrict__ frag32, float* __restrict__ bias) {
    const int lane = threadIdx.x;  // 64 threads
    for (int layer = 1; layer < 4; ++layer) {
        const Layer L = kLayers[layer];
        const int mts = (L.out + 15) / 16, kts = L.in / 32;
        for (int mt = 0; mt < mts; ++mt)
            for (int kt = 0; kt < kts; ++kt) {
                uint16_t* f = frag + ((long)frag_index(layer, mt, kt) * 64 + lane) * 8;
                const int row = 16 * mt + lane % 16;
                for (int i = 0; i < 8; ++i) {
                    const int col = k_neuron(kt, lane / 16, i);
                    float w = (row < L.out && col < L.in) ? params[L.w + row * L.in + col] : 0.0f;
                    f[i] = (uint16_t)to_bf16(w);
                }
            }
    }
    const Layer L0 = kLayers[0];
    for (int mt = 0; mt < 4; ++mt) {
        const int row = 16 * mt + lane % 16;
        for (int kt = 0; kt < 2; ++kt) {
            const int col = 4 * kt + lane / 16;
            frag32[(mt * 2 + kt) * 64 + lane] = col < L0.in ? params[L0.w + row * L0.in + col] : 0.0f;
        }
        const int g = lane / 16;
        const int col = g == 0 ? 1 : g == 1 ? 3 : 4;
        frag32[(kFolded32 + mt) * 64 + lane] = g < 3 ? params[L0.w + row * L0.in + col] : 0.0f;
    }
    for (int layer = 0; layer < 4; ++layer) {
        const Layer L = kLayers[layer];
        for (int r = lane; r < (layer < 3 ? 64 : 16); r += 64)
            bias[64 * layer + r] = r < L.out ? params[L.b + r] : 0.0f;
    }
}

__device__ __forceinline__ floatx4 mfma(bf16x8 a, bf16x4 lo, bf16x4 hi, floatx4 c) {
    const bf16x8 b = __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7);
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8v, a),
                                                    __builtin_bit_cast(bf16x8v, b), c, 0, 0, 0);
}

// The MFMA inference runs every layer scaled by kSiluScale = -log2(e): with
// layer 1's weights and every bias multiplied by it, each layer's
// accumulator holds y = -log2(e) z, and SiLU in that scale,
// y / (1 + 2^y) = -log2(e) silu(z), is the next layer's input unchanged:
// the factor carries through the unscaled bf16 weights of layers 2-4, whose
// outputs come out scaled by it too and are multiplied back by
// kSiluUnscale = 1 / kSiluScale.  This saves the multiply of every SiLU's
// exponent.
constexpr float kSiluScale = -1.44269504f;
constexpr float kSiluUnscale = -0.693147182f;

// SiLU of a pre-activation scaled by kSiluScale, y / (1 + 2^y), at
// hardware rate (v_exp_f32, v_rcp_f32), then packed to bf16 as the next
// layer's B operand.
__device__ __forceinline__ bf16x4 silu_bf16(floatx4 c) {
    floatx4 r;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        float y = c[i];
        r[i] = y * __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(y));
    }
    return pack_bf16(r);
}

// Layers 2-4 on the layer-1 activations of 16 pixels: act[mt] holds lane
// l's neurons 16 mt + 4 (l / 16) + i of pixel l % 16.  Returns layer 4's
// output tile (lane l: outputs 4 (l / 16) + i of pixel l % 16).
__device__ __forceinline__ floatx4 mlp_tail(bf16x4 act[4], const bf16x8* __restrict__ frag,
                                            const float* __restrict__ bias, int lane) {
    const int group = lane / 16;
#pragma unroll
    for (int layer = 1; layer <= 2; ++layer) {
        bf16x4 next[4];
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
            floatx4 c;
#pragma unroll
            for (int i = 0; i < 4; ++i) c[i] = bias[64 * layer + 16 * mt + 4 * group + i];
            c = mfma(frag[frag_index(layer, mt, 0) * 64 + lane], act[0], act[1], c);
            c = mfma(frag[frag_index(layer, mt, 1) * 64 + lane], act[2], act[3], c);
            next[mt] = silu_bf16(c);
        }
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) act[mt] = next[mt];
    }
    floatx4 c;
#pragma unroll
    for (int i = 0; i < 4; ++i) c[i] = bias[192 + 4 * group + i];
    c = mfma(frag[frag_index(3, 0, 0) * 64 + lane], act[0], act[1], c);
    return mfma(frag[frag_index(3, 0, 1) * 64 + lane], act[2], act[3], c);
}

// The 6 raw outputs of pixel 16 j + col from layer 4's tile: outputs 0-3 sit
// in lane col, 4-5 in lane col + 16; lane 16 j + col takes them.
__device__ __forceinline__ void gather_outputs(floatx4 c, int j, int lane, float out[6]) {
    const int group = lane / 16, col = lane % 16;
    float o[6];
    o[0] = __shfl(c[0], col, 64);
    o[1] = __shfl(c[1], col, 64);
    o[2] = __shfl(c[2], col, 64);
    o[3] = __shfl(c[3], col, 64);
    o[4] = __shfl(c[0], col + 16, 64);
    o[5] = __shfl(c[1], col + 16, 64);
    if (group == j)
#pragma unroll
        for (int q = 0; q < 6; ++q) out[q] = o[q];
}

// Runs the MLP for the 64 pixels of a wave.  `x` holds the lane's own
// pixel features; returns the 6 raw outputs of the lane's pixel.  The wave
// walks its pixels 16 at a time.  Layer 1 (5 -> 64) runs on
// v_mfma_f32_16x16x4f32 in full fp32 (bf16 inputs would quantise log(d) by
// ~0.4 %, i.e. the segment length by ~3 %): lane l feeds feature
// 4 kt + l / 16 of pixel l % 16 as the B operand.  Layers 2-4 run on
// gfx950's v_mfma_f32_16x16x32_bf16 with the weights as the A operand and
// the activations as the B operand (see k_neuron).  `a32` holds the lane's
// 8 fp32 layer-1 fragments, `bias` the fp32 biases of all layers.
__device__ __forceinline__ void mlp_wave(const float x[5], const bf16x8* __restrict__ frag,
                                         const float a32[8], const float* __restrict__ bias,
                                         float out[6]) {
    const int lane = threadIdx.x % 64;
    const int group = lane / 16, col = lane % 16;
#pragma unroll 1
    for (int j = 0; j < 4; ++j) {
        const int src = 16 * j + col;
        const float f0 = __shfl(x[0], src, 64), f1 = __shfl(x[1], src, 64);
        const float f2 = __shfl(x[2], src, 64), f3 = __shfl(x[3], src, 64);
        const float f4 = __shfl(x[4], src, 64);
        const float b0 = group == 0 ? f0 : group == 1 ? f1 : group == 2 ? f2 : f3;
        const float b1 = group == 0 ? f4 : 0.0f;
        bf16x4 act[4];
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
            floatx4 c;
#pragma unroll
            for (int i = 0; i < 4; ++i) c[i] = bias[16 * mt + 4 * group + i];
            c = __builtin_amdgcn_mfma_f32_16x16x4f32(a32[2 * mt], b0, c, 0, 0, 0);
            c = __builtin_amdgcn_mfma_f32_16x16x4f32(a32[2 * mt + 1], b1, c, 0, 0, 0);
            act[mt] = silu_bf16(c);
        }
        gather_outputs(mlp_tail(act, frag, bias, lane), j, lane, out);
    }
}

// mlp_wave for pixels of one frame, whose altitude and sun-cosine features
// are constants already folded into layer 1's bias (`bias0`): the lane
// feeds its pixel's view-cosine, azimuth and distance features (x1, x3,
// x4), and layer 1 is one K-step per row tile with the folded fragments
// `a32f`.  Layer-major: each layer runs for C 16-pixel chunks at once, so a
// wave has 4 C independent MFMA / SiLU chains in flight and each weight
// fragment is read from LDS once per C chunks; the wave's 64 pixels take
// 4 / C rounds.
template <int C>
__device__ __forceinline__ void mlp_wave_folded(float x1, float x3, float x4,
                                                const bf16x8* __restrict__ frag,
                                                const float a32f[4],
                                                const float* __restrict__ bias0,
                                                const float* __restrict__ bias, float out[6]) {
    const int lane = threadIdx.x % 64;
    const int group = lane / 16, col = lane % 16;
#pragma unroll 1
    for (int round = 0; round < 4 / C; ++round) {
        asm volatile("" ::: "memory");
        bf16x4 act[C][4];
#pragma unroll
        for (int q = 0; q < C; ++q) {
            const int src = 16 * (round * C + q) + col;
            const float f1 = __shfl(x1, src, 64), f3 = __shfl(x3, src, 64);
            const float f4 = __shfl(x4, src, 64);
            const float b = group == 0 ? f1 : group == 1 ? f3 : group == 2 ? f4 : 0.0f;
#pragma unroll
            for (int mt = 0; mt < 4; ++mt) {
                floatx4 c;
#pragma unroll
                for (int i = 0; i < 4; ++i) c[i] = bias0[16 * mt + 4 * group + i];
                c = __builtin_amdgcn_mfma_f32_16x16x4f32(a32f[mt], b, c, 0, 0, 0);
                act[q][mt] = silu_bf16(c);
            }
        }
#pragma unroll
        for (int layer = 1; layer <= 2; ++layer) {
            // The weights and biases are re-read from LDS here rather than
            // held in registers across the pixel loop: an empty asm with a
            // memory clobber keeps the compiler from hoisting the loads.
            asm volatile("" ::: "memory");
            bf16x4 next[C][4];
#pragma unroll
            for (int mt = 0; mt < 4; ++mt) {
                const bf16x8 w0 = frag[frag_index(layer, mt, 0) * 64 + lane];
                const bf16x8 w1 = frag[frag_index(layer, mt, 1) * 64 + lane];
                floatx4 b;
#pragma unroll
                for (int i = 0; i < 4; ++i) b[i] = bias[64 * layer + 16 * mt + 4 * group + i];
#pragma unroll
                for (int q = 0; q < C; ++q) {
                    floatx4 c = mfma(w0, act[q][0], act[q][1], b);
                    c = mfma(w1, act[q][2], act[q][3], c);
                    next[q][mt] = silu_bf16(c);
                }
            }
#pragma unroll
            for (int q = 0; q < C; ++q)
#pragma unroll
                for (int mt = 0; mt < 4; ++mt) act[q][mt] = next[q][mt];
        }
        asm volatile("" ::: "memory");
        const bf16x8 w0 = frag[frag_index(3, 0, 0) * 64 + lane];
        const bf16x8 w1 = frag[frag_index(3, 0, 1) * 64 + lane];
        floatx4 b;
#pragma unroll
        for (int i = 0; i < 4; ++i) b[i] = bias[192 + 4 * group + i];
#pragma unroll
        for (int q = 0; q < C; ++q) {
            floatx4 c = mfma(w0, act[q][0], act[q][1], b);
            c = mfma(w1, act[q][2], act[q][3], c);
            gather_outputs(c, round * C + q, lane, out);
        }
    }
}

// Network outputs -> (in-scatter per unit irradiance L, transmittance T).
// v = y * out_scale + out_mean.  Mode 1 (linear): L = max(v, 0),
// T = clamp(v, 0, 1).  Mode 0 (log): L = exp(v) - eps_inscatter,
// T = exp(v) - eps_transmittance.  Mode 2 (floored log): L = exp(v), zero
// below 1.5 eps_inscatter (the training floor), T as in mode 0.
__device__ __forceinline__ f3 decode_inscatter(const AtmoSurrogateTransform& t,
                                               const float raw[6]) {
    float y[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) y[q] = raw[q] * t.out_scale[q] + t.out_mean[q];
    if (t.output_mode == 1) return maxs(mk(y[0], y[1], y[2]), 0.0f);
    float l[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        const float e = expf(y[q]);
        l[q] = t.output_mode == 2 ? (e < 1.5f * t.eps_inscatter ? 0.0f : e)
                                  : fmaxf(e - t.eps_inscatter, 0.0f);
    }
    return mk(l[0], l[1], l[2]);
}

__device__ __forceinline__ void decode(const AtmoSurrogateTransform& t, const float raw[6], f3& L,
                                       f3& T) {
    L = decode_inscatter(t, raw);
    float y[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) y[q] = raw[3 + q] * t.out_scale[3 + q] + t.out_mean[3 + q];
    if (t.output_mode == 1) {
        T = clamp3(mk(y[0], y[1], y[2]), 0.0f, 1.0f);
        return;
    }
    T = clamp3(mk(expf(y[0]) - t.eps_transmittance, expf(y[1]) - t.eps_transmittance,
                  expf(y[2]) - t.eps_transmittance),
               0.0f, 1.0f);
}

struct SurrogateArgs {
    const float* scene;
    const float* space;
    const int32_t* sid;
    const float* pos;
    float* hdr;
    const float* tlut;
    const float* mslut;
    const float* svlut;
    int width, height;
    float bottom, cam_alt, sky_alt;
    f3 sun, irradiance, disk;
    float ar;
    f3 right, up, forward;
    float thf, c0, c1;
    float aspect;        // width / height
    f3 fwd_t;            // forward / thf
    float pixel_filter;  // the sun disk's pixel filter (composite_frame_constants)
    float sky_horizon;   // sky-view horizon zenith at sky_alt
    const uint16_t* frag;
    const float* frag32;
    const float* bias;
    const float* params;  // fp32 master parameters (layer 1's weights for the bias fold)
    AtmoSurrogateTransform t;
};

// The COMPOSITE view with the surrogate in place of the surface branch's
// in-scatter (and, unless exact_transmittance is set, transmittance); sky
// pixels take the plain sky branch.  Persistent blocks: the 18 KiB of bf16
// weight fragments are staged in LDS once per block, with layer 1's bias
// folded with the frame's constant altitude and sun-cosine features, and
// the block then walks 256-pixel groups.  In a group, surface pixels are
// compacted onto the leading threads and sky pixels onto the threads after
// them, so the network runs in full waves over surface pixels (the MFMA
// chain needs the whole wave; a wave runs it when any lane holds an
// in-domain pixel) and the sky branch in at most two waves.
// Out-of-domain surface pixels, and in-domain ones whose learned in-scatter
// exceeds the transform's route_share of the pixel in some channel, are
// queued for the plain integrator (k_composite_queue_coarse, then
// k_composite_list over the refinement queue for those that need the
// terminator refinement), which
// keeps the integrator's registers out of this kernel.  They are collected
// in a block-local LDS queue and moved to the global queue with one atomic
// per block flush: one global atomic per wave, all on one counter, cost
// 0.14 ms per 1080p frame in serialised atomics.
constexpr int kBlockQueue = 1024;  // entries; a group adds at most 256

__device__ __forceinline__ void flush_block_queue(int* block_queue, int* base_out,
                                                  int* __restrict__ queue, int parity,
                                                  long capacity) {
    const int n = block_queue[0];  // block-uniform after the caller's barrier
    if (n == 0) return;
    if (threadIdx.x == 0) *base_out = atomicAdd(queue + parity, n);
    __syncthreads();
    const long b = *base_out;
    for (int e = threadIdx.x; e < n; e += blockDim.x)
        if (b + e < capacity) queue[2 + b + e] = block_queue[2 + e];
    __syncthreads();  // every thread has read n and the entries
    if (threadIdx.x == 0) block_queue[0] = 0;
    __syncthreads();  // the reset before anyone reads the count again
}

__global__ void __launch_bounds__(256) ATMO_SURROGATE_WAVES k_composite_surrogate(Consts k, SurrogateArgs a,
                                                             int* __restrict__ queue, int parity,
                                                             long capacity) {
    __shared__ bf16x8 frag_s[kFragments * 64];
    __shared__ float bias_s[64 * 3 + 16];
    __shared__ float bias0_s[64];
    __shared__ int list[256];
    __shared__ int wave_counts[8];
    __shared__ int block_queue[2 + kBlockQueue];  // [count, unused, entries]
    __shared__ int block_queue_base;
    if (threadIdx.x == 0) block_queue[0] = 0;
    const bf16x8* fg = reinterpret_cast<const bf16x8*>(a.frag);
    for (int e = threadIdx.x; e < kFragments * 64; e += blockDim.x) frag_s[e] = fg[e];
    for (int e = threadIdx.x; e < 64 * 3 + 16; e += blockDim.x) bias_s[e] = kSiluScale * a.bias[e];
    const Encoding enc = encoding_at(a.t, a.cam_alt);
    if (threadIdx.x < 64) {
        // z = W x + b with x0 (altitude) and x2 (sun cosine) fixed per frame
        const float x2 = feature_cosine(a.t, enc, clampf(a.sun.y, -1.0f, 1.0f));
        const float* w = a.params + kLayers[0].w + threadIdx.x * 5;
        bias0_s[threadIdx.x] = kSiluScale * ((a.bias[threadIdx.x] + w[0] * enc.x0) + w[2] * x2);
    }
    // (the first compaction's barriers order this staging before any read)
    const int lane = threadIdx.x % 64;
    float a32f[4];
#pragma unroll
    for (int f = 0; f < 4; ++f) a32f[f] = kSiluScale * a.frag32[(kFolded32 + f) * 64 + lane];

    // pixel byte offsets in 32 bits: a frame holds at most 2^28 pixels
    const uint32_t npix = (uint32_t)a.width * (uint32_t)a.height;
    const uint32_t height = (uint32_t)a.height;
    const uint32_t groups = (npix + 255u) / 256u;
    for (uint32_t g = blockIdx.x; g < groups; g += gridDim.x) {
        const uint32_t base = g * 256u;
        const uint32_t i = base + threadIdx.x;
        const bool surface = i < npix && a.sid[i] >= 0;
        int nsky;
        const int count = compact_block2(surface, i < npix && !surface, list, wave_counts, nsky);
        if ((int)threadIdx.x >= count && (int)threadIdx.x < count + nsky) {
            const uint32_t p = base + list[threadIdx.x];
            const uint32_t px = p / height;
            const f3 ray = camera_ray((int)px, (int)(p - px * height), a.width, a.height,
                                      a.aspect, a.right, a.up, a.fwd_t);
            const f3 c = composite_sky(k, a.tlut, a.svlut, ray, pix3(a.scene, p), pix3(a.space, p),
                                       a.cam_alt, a.sky_alt, a.sun, a.disk, a.ar, a.pixel_filter,
                                       a.sky_horizon, a.c0, a.c1, a.bottom);
            put3(a.hdr, p, maxs(c, 0.0f));
        }
        const bool has = (int)threadIdx.x < count;
        const uint32_t j = base + (has ? list[threadIdx.x] : 0);
        const uint32_t jx = j / height;
        const f3 ray = camera_ray((int)jx, (int)(j - jx * height), a.width, a.height,
                                  a.aspect, a.right, a.up, a.fwd_t);
        const float sd = has ? norm(pix3(a.pos, j)) : 1.0f;
        const bool learned = has && in_domain(a.t, a.cam_alt, sd);
        bool defer = has && !learned;  // out of the network's domain
        if (__ballot(learned)) {  // wave-uniform branch
            // relative azimuth as in the sky-view sampler (renderer.py:1482-1497)
            float vl = sqrtf(ray.x * ray.x + ray.z * ray.z);
            float sl = sqrtf(a.sun.x * a.sun.x + a.sun.z * a.sun.z);
            float phi = 0.0f;
            if (vl > 1.0e-6f && sl > 1.0e-6f)
                phi = acosf(clampf((ray.x * a.sun.x + ray.z * a.sun.z) / (vl * sl), -1.0f, 1.0f));
            float raw[6];
            mlp_wave_folded<ATMO_SURROGATE_CHUNKS>(feature_cosine(a.t, enc, clampf(ray.y, -1.0f, 1.0f)),
                            feature_azimuth(phi),
                            feature_distance(enc, fmaxf(sd, a.t.distance_min_m)), frag_s, a32f,
                            bias0_s, bias_s, raw);
            if (learned) {
#pragma unroll
                for (int q = 0; q < 6; ++q) raw[q] *= kSiluUnscale;
                f3 L, T;
                if (a.t.exact_transmittance) {
                    L = decode_inscatter(a.t, raw);
                    T = surface_transmittance(k, a.tlut, a.cam_alt, ray, sd, a.bottom);
                } else {
                    decode(a.t, raw, L, T);
                }
                const f3 le = L * a.irradiance;
                const f3 st = pix3(a.scene, j) * T;
                // the learned in-scatter's share of the pixel, per channel,
                // without a division: E L > share (E L + scene T)
                const float rs = a.t.route_share;
                defer = rs > 0.0f && (le.x > rs * (le.x + st.x) || le.y > rs * (le.y + st.y) ||
                                      le.z > rs * (le.z + st.z));
                if (!defer) put3(a.hdr, j, maxs(le + st, 0.0f));
            }
        }
        enqueue_wave(defer, j, block_queue, 0, kBlockQueue);  // LDS atomics
        __syncthreads();  // `list` is rewritten by the next group; the appends are done
        if (block_queue[0] > kBlockQueue - 256)
            flush_block_queue(block_queue, &block_queue_base, queue, parity, capacity);
    }
    flush_block_queue(block_queue, &block_queue_base, queue, parity, capacity);
}

// Batch inference on encoded features (testing and evaluation): x [n, 5]
// -> y [n, 6] raw network outputs, through the same MFMA path.
__global__ void __launch_bounds__(256) k_mlp_infer(long n, const float* __restrict__ x,
                                                   const uint16_t* __restrict__ frag,
                                                   const float* __restrict__ frag32,
                                                   const float* __restrict__ bias,
                                                   float* __restrict__ y) {
    __shared__ bf16x8 frag_s[kFragments * 64];
    __shared__ float bias_s[64 * 3 + 16];
    const bf16x8* fg = reinterpret_cast<const bf16x8*>(frag);
    for (int e = threadIdx.x; e < kFragments * 64; e += blockDim.x) frag_s[e] = fg[e];
    for (int e = threadIdx.x; e < 64 * 3 + 16; e += blockDim.x) bias_s[e] = kSiluScale * bias[e];
    __syncthreads();
    const int lane = threadIdx.x % 64;
    float a32[8];
    for (int f = 0; f < 8; ++f) a32[f] = kSiluScale * frag32[f * 64 + lane];
    const long i_raw = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long i = i_raw < n ? i_raw : n - 1;
    float f[5];
    for (int q = 0; q < 5; ++q) f[q] = x[i * 5 + q];
    float out[6];
    mlp_wave(f, frag_s, a32, bias_s, out);
    if (i_raw < n)
        for (int q = 0; q < 6; ++q) y[i * 6 + q] = kSiluUnscale * out[q];
}

unsigned blocks_for(long threads, int block) { return (unsigned)((threads + block - 1) / block); }

template <typename T>
hipError_t alloc(T** p, size_t count) {
    *p = nullptr;
    return hipMalloc((void**)p, count * sizeof(T));
}

void free_mlp(AtmoMlp* m) {
    void* bufs[] = {m->params, m->grads, m->adam_m, m->adam_v, m->z[0], m->z[1], m->z[2], m->z[3],
                    m->a[0], m->a[1], m->a[2], m->dz[0], m->dz[1], m->dz[2], m->dz[3], m->xbuf,
                    m->wbuf, m->partial, m->loss, m->frag, m->frag32, m->bias};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
}

// Forward pass of the training network on m->xbuf (B rows).
void forward_layers(AtmoMlp* m, int B, hipStream_t s) {
    const float* in = m->xbuf;
    for (int l = 0; l < 4; ++l) {
        const Layer L = kLayers[l];
        hipLaunchKernelGGL(k_dense_forward, dim3(blocks_for((long)B * L.out, 256)), dim3(256), 0, s, B,
                           L.in, L.out, in, m->params + L.w, m->params + L.b, m->z[l],
                           l < 3 ? m->a[l] : nullptr);
        if (l < 3) in = m->a[l];
    }
}

// Backward pass from dz[3] (already holding d loss / d output).
void backward_layers(AtmoMlp* m, int B, hipStream_t s) {
    for (int l = 3; l >= 0; --l) {
        const Layer L = kLayers[l];
        const float* x = l == 0 ? m->xbuf : m->a[l - 1];
        const int chunks = (B + kChunk - 1) / kChunk;
        hipLaunchKernelGGL(k_dense_backward_weights, dim3(chunks), dim3(256), 0, s, B, L.out, L.in,
                           m->dz[l], x, m->partial);
        hipLaunchKernelGGL(k_reduce_partials, dim3(blocks_for(L.out * L.in + L.out, 256)), dim3(256),
                           0, s, chunks, L.out * L.in + L.out, m->partial, m->grads + L.w);
        if (l > 0)
            hipLaunchKernelGGL(k_dense_backward_data, dim3(blocks_for((long)B * L.in, 256)), dim3(256),
                               0, s, B, L.out, L.in, m->dz[l], m->params + L.w, m->z[l - 1],
                               m->dz[l - 1]);
    }
}

int32_t status(AtmoMlp* m, const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        snprintf(m->error, sizeof m->error, "%s: %s", what, hipGetErrorString(e));
        return ATMO_EHIP;
    }
    return ATMO_OK;
}

}  // namespace

extern "C" {

int32_t atmo_mlp_create(int32_t max_batch, AtmoMlp** out) {
    if (!out || max_batch <= 0 || max_batch > (1 << 20)) return ATMO_EINVAL;
    *out = nullptr;
    AtmoMlp* m = new (std::nothrow) AtmoMlp();
    if (!m) return ATMO_ENOMEM;
    memset(m, 0, sizeof *m);
    m->magic = kMlpMagic;
    m->max_batch = max_batch;
    const size_t B = (size_t)max_batch;
    const size_t chunks = (B + kChunk - 1) / kChunk;
    bool ok = alloc(&m->params, ATMO_MLP_PARAMS) == hipSuccess &&
              alloc(&m->grads, ATMO_MLP_PARAMS) == hipSuccess &&
              alloc(&m->adam_m, ATMO_MLP_PARAMS) == hipSuccess &&
              alloc(&m->adam_v, ATMO_MLP_PARAMS) == hipSuccess &&
              alloc(&m->xbuf, B * 5) == hipSuccess &&
              alloc(&m->wbuf, B * 6) == hipSuccess &&
              alloc(&m->partial, chunks * (64 * 64 + 64)) == hipSuccess &&
              alloc(&m->loss, B * 6 / 256 + 2) == hipSuccess &&
              alloc(&m->frag, (size_t)kFragments * 64 * 8) == hipSuccess &&
              alloc(&m->frag32, (size_t)kFragments32 * 64) == hipSuccess &&
              alloc(&m->bias, (size_t)64 * 3 + 16) == hipSuccess;
    for (int l = 0; l < 4 && ok; ++l) {
        ok = alloc(&m->z[l], B * kLayers[l].out) == hipSuccess &&
             alloc(&m->dz[l], B * kLayers[l].out) == hipSuccess;
        if (ok && l < 3) ok = alloc(&m->a[l], B * 64) == hipSuccess;
    }
    if (ok) ok = hipMemset(m->adam_m, 0, ATMO_MLP_PARAMS * 4) == hipSuccess &&
                 hipMemset(m->adam_v, 0, ATMO_MLP_PARAMS * 4) == hipSuccess &&
                 hipMemset(m->params, 0, ATMO_MLP_PARAMS * 4) == hipSuccess &&
                 hipDeviceSynchronize() == hipSuccess;
    if (!ok) {
        free_mlp(m);
        delete m;
        return ATMO_ENOMEM;
    }
    m->transform.altitude_max_m = 100000.0f;
    m->transform.altitude_scale_m = 100.0f;
    m->transform.output_mode = 2;
    m->transform.eps_inscatter = 1.0e-7f;
    m->transform.eps_transmittance = 1.0e-6f;
    m->transform.planet_radius_m = 6360000.0f;
    m->transform.distance_min_m = 1.0f;
    m->transform.distance_max_m = 1.0e6f;
    for (int q = 0; q < 6; ++q) m->transform.out_scale[q] = 1.0f;
    *out = m;
    return ATMO_OK;
}

int32_t atmo_mlp_destroy(AtmoMlp* m) {
    if (!mlp_ok(m)) return ATMO_EHANDLE;
    (void)hipDeviceSynchronize();
    free_mlp(m);
    m->magic = 0;
    delete m;
    return ATMO_OK;
}

const char* atmo_mlp_last_error(const AtmoMlp* m) {
    return mlp_ok(m) ? m->error : "invalid surrogate handle";
}

int32_t atmo_mlp_set_params(AtmoMlp* m, const float* host_params, int64_t count) {
    if (!mlp_ok(m)) return ATMO_EHANDLE;
    if (!host_params || count != ATMO_MLP_PARAMS) return ATMO_ESHAPE;
    if (hipMemcpy(m->params, host_params, count * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(m->adam_m, 0, count * 4) != hipSuccess ||
        hipMemset(m->adam_v, 0, count * 4) != hipSuccess)
        return ATMO_EHIP;
    hipLaunchKernelGGL(k_pack_fragments, dim3(1), dim3(64), 0, 0, m->params, m->frag, m->frag32,
                       m->bias);
    if (status(m, "atmo_mlp_set_params") != ATMO_OK) return ATMO_EHIP;
    return hipDeviceSynchronize() == hipSuccess ? ATMO_OK : ATMO_EHIP;
}

int32_t atmo_mlp_get_params(AtmoMlp* m, float* host_params, int64_t count) {
    if (!mlp_ok(m)) return ATMO_EHANDLE;
    if (!host_params || count != ATMO_MLP_PARAMS) return ATMO_ESHAPE;
    return hipMemcpy(host_params, m->params, count * 4, hipMemcpyDeviceToHost) == hipSuccess
               ? ATMO_OK
               : ATMO_EHIP;
}

int32_t atmo_mlp_set_transform(AtmoMlp* m, const AtmoSurrogateTransform* t) {
    if (!mlp_ok(m)) return ATMO_EHANDLE;
    if (!t || !(t->altitude_max_m > 0.0f) || !(t->altitude_scale_m > 0.0f) ||
        !(t->distance_min_m > 0.0f) ||
        !(t->distance_max_m > t->distance_min_m) || t->output_mode < 0 || t->output_mode > 2 ||
        !(t->planet_radius_m > 0.0f) || !(t->cosine_stretch >= 0.0f) ||
        (t->exact_transmittance != 0 && t->exact_transmittance != 1) ||
        !(t->route_share >= 0.0f && t->route_share < 1.0f))
        return ATMO_EINVAL;
    m->transform = *t;
    return ATMO_OK;
}

int32_t atmo_mlp_forward(AtmoMlp* m, const float* x, int32_t B, float* y, void* stream) {
    if (!mlp_ok(m)) return ATMO_EHANDLE;
    if (!x || !y || B <= 0 || B > m->max_batch) return ATMO_ESHAPE;
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_gather, dim3(blocks_for((long)B * 5, 256)), dim3(256), 0, s, B, 5, nullptr,
                       x, m->xbuf);
    forward_layers(m, B, s);
    if (hipMemcpyAsync(y, m->z[3], (size_t)B * 6 * 4, hipMemcpyDeviceToDevice, s) != hipSuccess)
        return ATMO_EHIP;
    return status(m, "atmo_mlp_forward");
}

int32_t atmo_mlp_gradients(AtmoMlp* m, const float* x, const float* target, const float* weight,
                           const int32_t* index, int32_t B, float* grad_out, float* loss_out,
                           void* stream) {
    if (!mlp_ok(m)) return ATMO_EHANDLE;
    if (!x || !target || B <= 0 || B > m->max_batch) return ATMO_ESHAPE;
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_gather, dim3(blocks_for((long)B * 5, 256)), dim3(256), 0, s, B, 5, index, x,
                       m->xbuf);
    // the targets are gathered into dz[3] first and replaced by the gradient
    hipLaunchKernelGGL(k_gather, dim3(blocks_for((long)B * 6, 256)), dim3(256), 0, s, B, 6, index,
                       target, m->dz[3]);
    if (weight)
        hipLaunchKernelGGL(k_gather, dim3(blocks_for((long)B * 6, 256)), dim3(256), 0, s, B, 6,
                           index, weight, m->wbuf);
    forward_layers(m, B, s);
    const int lblocks = (int)blocks_for((long)B * 6, 256);
    hipLaunchKernelGGL(k_mse, dim3(lblocks), dim3(256), 0, s, B, m->z[3], m->dz[3],
                       weight ? (const float*)m->wbuf : nullptr, m->dz[3], m->loss);
    hipLaunchKernelGGL(k_loss_reduce, dim3(1), dim3(256), 0, s, lblocks, (long)B * 6, m->loss);
    backward_layers(m, B, s);
    if (grad_out &&
        hipMemcpyAsync(grad_out, m->grads, ATMO_MLP_PARAMS * 4, hipMemcpyDeviceToDevice, s) != hipSuccess)
        return ATMO_EHIP;
    if (loss_out && hipMemcpyAsync(loss_out, m->loss, 4, hipMemcpyDeviceToDevice, s) != hipSuccess)
        return ATMO_EHIP;
    return status(m, "atmo_mlp_gradients");
}

int32_t atmo_mlp_train_step(AtmoMlp* m, const float* x, const float* target, const float* weight,
                            const int32_t* index, int32_t B, float lr, int32_t step, float* loss_out,
                            void* stream) {
    if (!mlp_ok(m)) return ATMO_EHANDLE;
    if (step < 1) return ATMO_EINVAL;
    int32_t st = atmo_mlp_gradients(m, x, target, weight, index, B, nullptr, loss_out, stream);
    if (st != ATMO_OK) return st;
    const float b1 = 0.9f, b2 = 0.999f;
    const float bc1 = (float)(1.0 - pow(0.9, step)), bc2 = (float)(1.0 - pow(0.999, step));
    hipLaunchKernelGGL(k_adam, dim3(blocks_for(ATMO_MLP_PARAMS, 256)), dim3(256), 0, (hipStream_t)stream,
                       ATMO_MLP_PARAMS, m->params, m->grads, m->adam_m, m->adam_v, lr, b1, b2, 1.0e-8f,
                       bc1, bc2);
    return status(m, "atmo_mlp_train_step");
}

int32_t atmo_mlp_pack(AtmoMlp* m, void* stream) {
    if (!mlp_ok(m)) return ATMO_EHANDLE;
    hipLaunchKernelGGL(k_pack_fragments, dim3(1), dim3(64), 0, (hipStream_t)stream, m->params, m->frag,
                       m->frag32, m->bias);
    return status(m, "atmo_mlp_pack");
}

int32_t atmo_mlp_infer(AtmoMlp* m, const float* x, int64_t n, float* y, void* stream) {
    if (!mlp_ok(m)) return ATMO_EHANDLE;
    if (!x || !y || n <= 0) return ATMO_ESHAPE;
    hipLaunchKernelGGL(k_mlp_infer, dim3(blocks_for(n, 256)), dim3(256), 0, (hipStream_t)stream, n, x,
                       m->frag, m->frag32, m->bias, y);
    return status(m, "atmo_mlp_infer");
}

int32_t atmo_mlp_encode(AtmoMlp* m, const float* raw, int64_t n, float* x, void* stream) {
    if (!mlp_ok(m)) return ATMO_EHANDLE;
    if (!raw || !x || n <= 0) return ATMO_ESHAPE;
    hipLaunchKernelGGL(k_encode, dim3(blocks_for(n, 256)), dim3(256), 0, (hipStream_t)stream, n, raw,
                       m->transform, x);
    return status(m, "atmo_mlp_encode");
}

int32_t atmo_surrogate_labels(AtmoHandle* h, const float* raw, int64_t n, float bottom,
                              float sun_angular_radius, float* labels, void* stream) {
    if (!atmo_handle_ok(h)) return ATMO_EHANDLE;
    if (!raw || !labels || n <= 0) {
        atmo_set_error(h, "atmo_surrogate_labels: bad arguments");
        return ATMO_EINVAL;
    }
    hipLaunchKernelGGL(k_labels, dim3(blocks_for(n, 256)), dim3(256), 0, (hipStream_t)stream, h->k, (long)n,
                       raw, bottom, sun_angular_radius, h->tlut, h->mslut, labels);
    return atmo_launch_status(h, "atmo_surrogate_labels");
}

int32_t atmo_composite_surrogate(AtmoHandle* h, AtmoMlp* m, const float* scene_hdr,
                                 const float* space_hdr, const int32_t* surface_id,
                                 const float* position_view, float* hdr, float bottom, float cam_alt,
                                 float sky_alt, const float sun[3], const float irradiance[3],
                                 const float disk[3], float ar, const float right[3], const float up[3],
                                 const float forward[3], float thf, float c0, float c1, void* stream) {
    if (!atmo_handle_ok(h)) return ATMO_EHANDLE;
    if (!mlp_ok(m)) {
        atmo_set_error(h, "atmo_composite_surrogate: invalid surrogate handle");
        return ATMO_EINVAL;
    }
    if (!scene_hdr || !space_hdr || !surface_id || !position_view || !hdr || !sun || !irradiance ||
        !disk || !right || !up || !forward) {
        atmo_set_error(h, "atmo_composite_surrogate: NULL argument");
        return ATMO_EINVAL;
    }
    SurrogateArgs a;
    a.scene = scene_hdr;
    a.space = space_hdr;
    a.sid = surface_id;
    a.pos = position_view;
    a.hdr = hdr;
    a.tlut = h->tlut;
    a.mslut = h->mslut;
    a.svlut = h->svlut;
    a.width = h->width;
    a.height = h->height;
    a.bottom = bottom;
    a.cam_alt = cam_alt;
    a.sky_alt = sky_alt;
    a.sun = mk3(sun);
    a.irradiance = mk3(irradiance);
    a.disk = mk3(disk);
    a.ar = ar;
    a.right = mk3(right);
    a.up = mk3(up);
    a.forward = mk3(forward);
    a.thf = thf;
    a.c0 = c0;
    a.c1 = c1;
    a.frag = m->frag;
    a.frag32 = m->frag32;
    a.bias = m->bias;
    a.params = m->params;
    a.t = m->transform;
    const long npix = (long)h->width * h->height;
    const long groups = (npix + 255) / 256;
    hipStream_t s = (hipStream_t)stream;
    CompositeArgs ca;
    ca.scene = scene_hdr;
    ca.space = space_hdr;
    ca.sid = surface_id;
    ca.pos = position_view;
    ca.hdr = hdr;
    ca.tlut = h->tlut;
    ca.mslut = h->mslut;
    ca.svlut = h->svlut;
    ca.aplut = h->aplut;
    ca.width = h->width;
    ca.height = h->height;
    ca.bottom = bottom;
    ca.cam_alt = cam_alt;
    ca.sky_alt = sky_alt;
    ca.sun = a.sun;
    ca.irradiance = a.irradiance;
    ca.disk = a.disk;
    ca.ar = ar;
    ca.right = a.right;
    ca.up = a.up;
    ca.forward = a.forward;
    ca.thf = thf;
    ca.c0 = c0;
    ca.c1 = c1;
    ca.view = ATMO_VIEW_COMPOSITE;
    composite_frame_constants(h->k, ca);
    a.aspect = ca.aspect;
    a.fwd_t = ca.fwd_t;
    a.pixel_filter = ca.pixel_filter;
    a.sky_horizon = ca.sky_horizon;
    if (atmo_begin_queue(h, ca, s) != ATMO_OK) return ATMO_EHIP;
    h->queue_reset = true;  // until every launch of this call has gone out
    hipLaunchKernelGGL(k_composite_surrogate,
                       dim3((unsigned)(groups < ATMO_SURROGATE_GRID ? groups : ATMO_SURROGATE_GRID)),
                       dim3(256), 0, s, h->k, a, h->queue, ca.parity, ca.queue_capacity);
    if (atmo_launch_status(h, "atmo_composite_surrogate") != ATMO_OK) return ATMO_EHIP;
    if (atmo_launch_composite_queue(h, ca, s) != ATMO_OK) return ATMO_EHIP;
    ca.list = ca.refine;  // the queued pixels the coarse pass left to the refinement
    if (atmo_launch_composite_list(h, ca, s) != ATMO_OK) return ATMO_EHIP;
    h->queue_reset = false;
    return ATMO_OK;
}

}  // extern "C"
