// postprocess.hip -- the display post-process of libatmo.so: the kernels
// of the reference's `PostProcessor` (src/planet_renderer/postprocess.py:
// 74-207) on the composite's HDR output.
//
// The reference runs 2 + 2P Taichi kernels per frame: `_extract` (2x2 box
// downsample + bright pass into a half-resolution field), P passes of
// `_blur_horizontal` + `_blur_vertical`, and `_resolve` (exposure,
// bilinear bloom, ACES + sRGB or a [0, 1] clamp).  Here three kernels:
//  - k_post_blur: the bright pass and all P pass pairs, one 48x48 tile of
//    half-resolution texels per 1024-thread block.  The block extracts its
//    tile plus a halo of 2P texels from the HDR frame into LDS (each
//    lane's hdr loads issued before its arithmetic), then runs the passes
//    ping-pong between two LDS copies (100 KB at P = 4, 156 KB at P = 8),
//    each pass over the shrinking region the next one reads.  A lane
//    produces S consecutive outputs of one channel along the pass axis
//    from a window of S + 4, S chosen per pass (8 to 19) so that every
//    lane has at most one window.  Every pass runs over compile-time
//    regions with no clamping: a tile that reaches past the frame (60 of
//    240 at 1080p) fills its cells outside the frame with the values at
//    their clamped positions, after the extraction and after each pass,
//    so its taps read what the reference's clamped taps read.  The tile's
//    interior goes to bloom_a (last vertical pass) and bloom_b (last
//    horizontal pass), the fields the reference leaves;
//  - k_post_resolve_bloom: four consecutive pixels per lane, 16-byte loads
//    and stores, the bloom taps of all four pixels loaded before any
//    arithmetic; without bloom every float of the frame is independent and
//    k_post_resolve_flat streams the frame as float4s.
// Every texel goes through the reference's operations in the reference's
// order, so the result does not depend on the tiling: each pass reads the
// values of the same clamped neighbours the full-frame pass would.
//
// Blur tiles and bloom-resolve chunks are dealt to the 8 XCDs in
// contiguous runs (block b runs on XCD b mod 8), so neighbouring tiles,
// which read each other's halo, and neighbouring pixels, which read the
// same bloom texels, share an L2.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <new>

#include "../../include/atmo.h"
#include "post_device.h"

namespace {

constexpr uint32_t kPostMagic = 0x504f5354u;  // "POST"
// Half-resolution texels per tile side.  At 1920x1080 the bloom frame
// (4 and 8 pass pairs) took 38.6 and 57.9 us with 48 x 48 tiles of 1024
// threads, 38.5 and 67.4 with 32 x 32, 43.7 and 66.5 with 40 x 40, and
// longer with 512 threads: fewer, larger tiles re-read less halo.
constexpr int kTile = 48;
constexpr int kThreads = 256;
constexpr int kBlurThreads = 1024;
constexpr int kXcds = 8;

struct PostArgs {
    const float* hdr;
    float* display;
    float* bloom_a;
    float* bloom_b;
    int32_t width, height, half_width, half_height;
    float width_f, height_f, half_width_f, half_height_f;
    double inv_height;
    float exposure_multiplier, bloom_strength;
    post::BrightPass bright;
    int32_t tone_map;
    int32_t tiles_x, tiles_y, tiles_per_xcd;
};

__device__ __forceinline__ float output_channel(float value, int32_t tone_map) {
    return tone_map ? post::display_channel(value) : fminf(fmaxf(value, 0.0f), 1.0f);
}

// One blur pass, src -> dst, along axis 0 (x, _blur_horizontal) or axis 1
// (y, _blur_vertical) (postprocess.py:121-139), over a region known at
// compile time whose length along the pass axis is a multiple of S.  A
// lane produces S consecutive outputs from a window of S + 4 texels, read
// at fixed offsets from one LDS address; no tap is clamped (cells outside
// the frame already hold their clamped values).
template <int AXIS, int L, int LP, int N, int S, int ACROSS_LO, int ACROSS_N, int SEG_LO, int NSEG>
__device__ __forceinline__ void blur_interior(const float* __restrict__ src,
                                              float* __restrict__ dst) {
    constexpr int STEP = AXIS == 0 ? LP : 1, ASTEP = AXIS == 0 ? 1 : LP;
    constexpr int PER_CHANNEL = NSEG * ACROSS_N;
    for (int item = (int)threadIdx.x; item < 3 * PER_CHANNEL; item += kBlurThreads) {
        const int c = item / PER_CHANNEL;
        const int rest = item - c * PER_CHANNEL;
        const int seg = rest / ACROSS_N;
        const int across = ACROSS_LO + rest - seg * ACROSS_N;
        const int start = SEG_LO + seg * S;
        // element (position p along the pass axis, across) at p * STEP + across * ASTEP
        const int base = c * N + (start - 2) * STEP + across * ASTEP;
        const float* s = src + base;
        float* d = dst + base;
        float w[S + 4];
#pragma unroll
        for (int k = 0; k < S + 4; ++k) w[k] = s[k * STEP];
#pragma unroll
        for (int j = 0; j < S; ++j) {
            float color = w[j + 2] * post::kW0;
            color += w[j + 1] * post::kW1;
            color += w[j + 3] * post::kW1;
            color += w[j] * post::kW2;
            color += w[j + 4] * post::kW2;
            d[(j + 2) * STEP] = color;
        }
    }
}

// Outputs per lane for a pass region `len` texels long and `across` wide:
// the smallest divisor of len (a multiple of 4) that gives every lane at
// most one item, so the pass is one short chain per lane; the largest
// divisor up to 20 if none does.
constexpr int window_outputs(int len, int across) {
    for (int s = 4; s <= len; ++s)
        if (len % s == 0 && 3 * (len / s) * across <= kBlurThreads) return s;
    for (int s = 20; s > 4; --s)
        if (len % s == 0) return s;
    return 4;
}

// The tile's cells inside the frame: local x in [xlo, xhi), y in [ylo, yhi).
struct Frame {
    int xlo, xhi, ylo, yhi;
    bool edge;
};

// Copies onto each cell of [lo0, hi0) x [lo1, hi1) outside the frame the
// value of its clamped position, which lies inside the region whenever a
// cell outside the frame does.  The pass over a padded tile then reads
// what the reference's clamped taps read.
template <int LP, int N>
__device__ __forceinline__ void replicate(float* buf, const Frame& f, int lo0, int hi0, int lo1,
                                          int hi1) {
    const int n1 = hi1 - lo1, n = (hi0 - lo0) * n1;
    for (int item = (int)threadIdx.x; item < n; item += kBlurThreads) {
        const int x = lo0 + item / n1, y = lo1 + item % n1;
        const int cx = min(max(x, f.xlo), f.xhi - 1), cy = min(max(y, f.ylo), f.yhi - 1);
        if (cx == x && cy == y) continue;
        const int d = x * LP + y, s = cx * LP + cy;
#pragma unroll
        for (int c = 0; c < 3; ++c) buf[c * N + d] = buf[c * N + s];
    }
}

// Pass pairs K..P over compile-time regions (pass k: x in [2k, L-2k); the
// horizontal pass keeps y in [2k-2, L-2k+2), the vertical y in
// [2k, L-2k)).  On a tile that reaches past the frame every pass is
// followed by the copy of clamped values onto the cells outside it (the
// last vertical pass's are never read).
template <int L, int LP, int N, int K, int P>
__device__ __forceinline__ void blur_passes(float* A, float* B, const Frame& f) {
    if constexpr (K <= P) {
        constexpr int SH = window_outputs(L - 4 * K, L - 4 * K + 4);
        constexpr int SV = window_outputs(L - 4 * K, L - 4 * K);
        blur_interior<0, L, LP, N, SH, 2 * K - 2, L - 4 * K + 4, 2 * K, (L - 4 * K) / SH>(A, B);
        __syncthreads();
        if (f.edge) {
            replicate<LP, N>(B, f, 2 * K, L - 2 * K, 2 * K - 2, L - 2 * K + 2);
            __syncthreads();
        }
        blur_interior<1, L, LP, N, SV, 2 * K, L - 4 * K, 2 * K, (L - 4 * K) / SV>(B, A);
        __syncthreads();
        if (K < P && f.edge) {
            replicate<LP, N>(A, f, 2 * K, L - 2 * K, 2 * K, L - 2 * K);
            __syncthreads();
        }
        blur_passes<L, LP, N, K + 1, P>(A, B, f);
    }
}

// _extract, then P x (_blur_horizontal, _blur_vertical), on one 48x48 tile
// with its halo of 2P texels, in LDS; writes the tile's bloom_a (last
// vertical pass) and bloom_b (last horizontal pass).
template <int P>
__global__ __launch_bounds__(kBlurThreads) void k_post_blur(PostArgs a) {
    constexpr int R = 2 * P;
    constexpr int L = kTile + 2 * R;
    constexpr int LP = L | 1;             // odd pitch: column walks hit distinct banks
    constexpr int N = L * LP;
    __shared__ float lds[2 * 3 * N];
    float* A = lds;
    float* B = lds + 3 * N;

    const int tile = (int)(blockIdx.x % kXcds) * a.tiles_per_xcd + (int)(blockIdx.x / kXcds);
    if (tile >= a.tiles_x * a.tiles_y) return;  // whole block, before any barrier
    const int X0 = (tile / a.tiles_y) * kTile, Y0 = (tile % a.tiles_y) * kTile;
    const int gx0 = X0 - R, gy0 = Y0 - R;
    const int hw = a.half_width, hh = a.half_height;

    // _extract (postprocess.py:110-119) over the whole local region, in
    // chunks of CH texels per lane whose hdr loads are all issued before
    // the first arithmetic.  A cell outside the frame holds the texel at
    // its clamped position, the value the reference's clamped taps read.
    constexpr int NL = (L * L + kBlurThreads - 1) / kBlurThreads;
    constexpr int CH = 4;
#pragma unroll
    for (int k0 = 0; k0 < NL; k0 += CH) {
        float px[CH][12];
#pragma unroll
        for (int c = 0; c < CH; ++c) {
            const int i = (int)threadIdx.x + (k0 + c) * kBlurThreads;
            const int lx = i / L, ly = i - lx * L;
            const int gx = min(max(gx0 + lx, 0), hw - 1), gy = min(max(gy0 + ly, 0), hh - 1);
            if (k0 + c < NL && i < L * L) {
#pragma unroll
                for (int ox = 0; ox < 2; ++ox) {
                    const int sx = min(gx * 2 + ox, a.width - 1);
#pragma unroll
                    for (int oy = 0; oy < 2; ++oy) {
                        const int sy = min(gy * 2 + oy, a.height - 1);
                        const float* p = a.hdr + ((size_t)sx * a.height + sy) * 3;
#pragma unroll
                        for (int ch = 0; ch < 3; ++ch) px[c][(ox * 2 + oy) * 3 + ch] = p[ch];
                    }
                }
            }
        }
#pragma unroll
        for (int c = 0; c < CH; ++c) {
            const int i = (int)threadIdx.x + (k0 + c) * kBlurThreads;
            if (k0 + c >= NL || i >= L * L) continue;
            float r = 0.0f, g = 0.0f, b = 0.0f;
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                r += px[c][3 * t];
                g += px[c][3 * t + 1];
                b += px[c][3 * t + 2];
            }
            const float s = 0.25f * a.exposure_multiplier;
            r *= s;
            g *= s;
            b *= s;
            post::soft_threshold(a.bright, r, g, b);
            const int lx = i / L, ly = i - lx * L;
            const int li = lx * LP + ly;
            A[li] = r;
            A[N + li] = g;
            A[2 * N + li] = b;
        }
    }
    __syncthreads();
    Frame f;
    f.xlo = max(-gx0, 0);
    f.xhi = min(hw - gx0, L);
    f.ylo = max(-gy0, 0);
    f.yhi = min(hh - gy0, L);
    f.edge = f.xlo > 0 || f.xhi < L || f.ylo > 0 || f.yhi < L;
    blur_passes<L, LP, N, 1, P>(A, B, f);
    for (int i = threadIdx.x; i < kTile * kTile; i += kBlurThreads) {
        const int tx = i / kTile, ty = i - tx * kTile;
        const int gx = X0 + tx, gy = Y0 + ty;
        if (gx >= hw || gy >= hh) continue;
        const int li = (tx + R) * LP + (ty + R);
        const size_t o = ((size_t)gx * hh + gy) * 3;
        a.bloom_a[o] = A[li];
        a.bloom_a[o + 1] = A[N + li];
        a.bloom_a[o + 2] = A[2 * N + li];
        a.bloom_b[o] = B[li];
        a.bloom_b[o + 1] = B[N + li];
        a.bloom_b[o + 2] = B[2 * N + li];
    }
}

// _sample_bloom's texel pair and weight along one axis (postprocess.py:141-150).
struct Lerp {
    int i0, i1;
    float w;
};
__device__ __forceinline__ Lerp bloom_axis(int pixel, float size_f, float half_size_f, int half_size) {
    const float t = ((float)pixel + 0.5f) / size_f;
    const float p = t * half_size_f - 0.5f;
    const float fp = floorf(p);
    Lerp l;
    l.i0 = min(max((int)fp, 0), half_size - 1);
    l.i1 = min(l.i0 + 1, half_size - 1);
    l.w = fminf(fmaxf(p - fp, 0.0f), 1.0f);
    return l;
}

// _resolve with bloom (postprocess.py:155-174): four consecutive pixels per
// lane (16-byte loads and stores when VEC), bloom taps from bloom_a.  Lanes
// past the frame's last pixel compute on clamped taps and store nothing.
template <bool VEC>
__global__ __launch_bounds__(kThreads) void k_post_resolve_bloom(PostArgs a) {
    const size_t n = (size_t)a.width * a.height;
    const size_t chunk = (size_t)(blockIdx.x % kXcds) * a.tiles_per_xcd + blockIdx.x / kXcds;
    const size_t i0 = (chunk * kThreads + threadIdx.x) * 4;
    if (i0 >= n) return;
    const int npx = (int)min((size_t)4, n - i0);
    float v[12];
    if (VEC && npx == 4) {
        const float4* h4 = reinterpret_cast<const float4*>(a.hdr + i0 * 3);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const float4 q = h4[k];
            v[4 * k] = q.x;
            v[4 * k + 1] = q.y;
            v[4 * k + 2] = q.z;
            v[4 * k + 3] = q.w;
        }
    } else {
#pragma unroll
        for (int k = 0; k < 12; ++k) v[k] = k < npx * 3 ? a.hdr[i0 * 3 + k] : 0.0f;
    }
    // pixel i0 is (x, y) with i0 = x * H + y; exact for any frame below
    // 2^51 pixels
    const int x = (int)(((double)i0 + 0.5) * a.inv_height);
    const int y = (int)(i0 - (size_t)x * a.height);
    {
        // bilinear set-up and every tap load of the four pixels first, then
        // the arithmetic (the tone map's rare slow path is a call, which
        // loads do not cross)
        float tap[4][4][3];
        float wx[4], wy[4];
        int px_x = x, px_y = y;
        Lerp lx = bloom_axis(px_x, a.width_f, a.half_width_f, a.half_width);
        const int hh = a.half_height;
#pragma unroll
        for (int p = 0; p < 4; ++p) {
            const Lerp ly = bloom_axis(px_y, a.height_f, a.half_height_f, a.half_height);
            wx[p] = lx.w;
            wy[p] = ly.w;
            const float* b00 = a.bloom_a + ((size_t)lx.i0 * hh + ly.i0) * 3;
            const float* b10 = a.bloom_a + ((size_t)lx.i1 * hh + ly.i0) * 3;
            const float* b01 = a.bloom_a + ((size_t)lx.i0 * hh + ly.i1) * 3;
            const float* b11 = a.bloom_a + ((size_t)lx.i1 * hh + ly.i1) * 3;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                tap[p][0][c] = b00[c];
                tap[p][1][c] = b10[c];
                tap[p][2][c] = b01[c];
                tap[p][3][c] = b11[c];
            }
            if (++px_y == a.height && p < 3) {
                px_y = 0;
                ++px_x;
                lx = bloom_axis(px_x, a.width_f, a.half_width_f, a.half_width);
            }
        }
#pragma unroll
        for (int p = 0; p < 4; ++p) {
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const float low = tap[p][0][c] * (1.0f - wx[p]) + tap[p][1][c] * wx[p];
                const float high = tap[p][2][c] * (1.0f - wx[p]) + tap[p][3][c] * wx[p];
                const float sample = low * (1.0f - wy[p]) + high * wy[p];
                float value = v[3 * p + c] * a.exposure_multiplier;
                value += sample * a.bloom_strength;
                v[3 * p + c] = output_channel(value, a.tone_map);
            }
        }
    }
    if (VEC && npx == 4) {
        float4* d4 = reinterpret_cast<float4*>(a.display + i0 * 3);
#pragma unroll
        for (int k = 0; k < 3; ++k)
            d4[k] = make_float4(v[4 * k], v[4 * k + 1], v[4 * k + 2], v[4 * k + 3]);
    } else {
#pragma unroll
        for (int k = 0; k < 12; ++k)
            if (k < npx * 3) a.display[i0 * 3 + k] = v[k];
    }
}

// _resolve without bloom: every float of the frame independently
// (exposure, then ACES + sRGB or the [0, 1] clamp), four per lane.
template <bool VEC>
__global__ __launch_bounds__(kThreads) void k_post_resolve_flat(PostArgs a) {
    const size_t n = (size_t)a.width * a.height * 3;
    const size_t i0 = ((size_t)blockIdx.x * kThreads + threadIdx.x) * 4;
    if (i0 >= n) return;
    if (VEC && i0 + 4 <= n) {
        const float4 q = *reinterpret_cast<const float4*>(a.hdr + i0);
        float4 o;
        o.x = output_channel(q.x * a.exposure_multiplier, a.tone_map);
        o.y = output_channel(q.y * a.exposure_multiplier, a.tone_map);
        o.z = output_channel(q.z * a.exposure_multiplier, a.tone_map);
        o.w = output_channel(q.w * a.exposure_multiplier, a.tone_map);
        *reinterpret_cast<float4*>(a.display + i0) = o;
    } else {
        for (size_t i = i0; i < i0 + 4 && i < n; ++i)
            a.display[i] = output_channel(a.hdr[i] * a.exposure_multiplier, a.tone_map);
    }
}

template <int P>
hipError_t launch_blur(const PostArgs& a, hipStream_t stream) {
    hipLaunchKernelGGL(k_post_blur<P>, dim3(kXcds * a.tiles_per_xcd), dim3(kBlurThreads), 0, stream, a);
    return hipGetLastError();
}

}  // namespace

struct AtmoPost {
    uint32_t magic;
    int32_t width, height, half_width, half_height;
    AtmoPostParams params;
    post::BrightPass bright;
    float bloom_strength;
    float* bloom[2];
    bool borrowed[2];
    char error[256];
};

namespace {

bool post_ok(const AtmoPost* h) { return h && h->magic == kPostMagic; }

char g_post_create_error[256] = "";

void post_error(AtmoPost* h, const char* message) {
    char* dst = h ? h->error : g_post_create_error;
    snprintf(dst, 256, "%s", message);
}

int32_t validate(const AtmoPostParams* p, char* message, int32_t capacity) {
    auto fail = [&](const char* text) {
        if (message && capacity > 0) snprintf(message, (size_t)capacity, "%s", text);
        return (int32_t)ATMO_EINVAL;
    };
    if (!p) return fail("postprocess parameters are NULL");
    if (!isfinite(p->exposure_ev) || !(p->exposure_ev >= -20.0 && p->exposure_ev <= 20.0))
        return fail("postprocess.exposure_ev must be finite and in [-20, 20]");
    const struct {
        const char* name;
        double value;
    } non_negative[] = {{"bloom_threshold", p->bloom_threshold},
                        {"bloom_knee", p->bloom_knee},
                        {"bloom_clamp", p->bloom_clamp},
                        {"bloom_strength", p->bloom_strength}};
    for (const auto& f : non_negative) {
        if (!isfinite(f.value) || f.value < 0.0) {
            char text[128];
            snprintf(text, sizeof text, "postprocess.%s must be finite and non-negative", f.name);
            return fail(text);
        }
    }
    if (p->bloom_passes < 0 || p->bloom_passes > 8)
        return fail("postprocess.bloom_passes must be in 0..8");
    return ATMO_OK;
}

size_t bloom_floats(const AtmoPost* h) { return (size_t)h->half_width * h->half_height * 3; }

}  // namespace

extern "C" {

int32_t atmo_post_validate_params(const AtmoPostParams* params, char* message,
                                  int32_t message_capacity) {
    return validate(params, message, message_capacity);
}

int32_t atmo_post_create(const AtmoPostParams* params, int32_t width, int32_t height,
                         AtmoPost** out) {
    if (!params || !out) {
        post_error(nullptr, "atmo_post_create: NULL argument");
        return ATMO_EINVAL;
    }
    *out = nullptr;
    if (validate(params, g_post_create_error, sizeof g_post_create_error) != ATMO_OK)
        return ATMO_EINVAL;
    if (width <= 0 || height <= 0) {
        post_error(nullptr, "width and height must be positive");
        return ATMO_EINVAL;
    }
    AtmoPost* h = new (std::nothrow) AtmoPost();
    if (!h) {
        post_error(nullptr, "atmo_post_create: host allocation failed");
        return ATMO_ENOMEM;
    }
    h->magic = kPostMagic;
    h->width = width;
    h->height = height;
    // postprocess.py:81-82
    h->half_width = (width + 1) / 2 > 1 ? (width + 1) / 2 : 1;
    h->half_height = (height + 1) / 2 > 1 ? (height + 1) / 2 : 1;
    h->params = *params;
    // Python double arithmetic, then float32, as `ti.static` folds it
    // (postprocess.py:91-106).
    const double knee = params->bloom_knee;
    h->bright.threshold = (float)params->bloom_threshold;
    h->bright.knee = (float)knee;
    h->bright.two_knee = (float)(2.0 * knee);
    h->bright.knee_denominator = (float)fmax(4.0 * knee, 1.0e-6);
    h->bright.clamp = (float)params->bloom_clamp;
    h->bloom_strength = (float)params->bloom_strength;
    const size_t bytes = bloom_floats(h) * sizeof(float);
    bool ok = true;
    for (float** p : {&h->bloom[0], &h->bloom[1]})
        ok = ok && hipMalloc(p, bytes) == hipSuccess && hipMemset(*p, 0, bytes) == hipSuccess;
    if (!ok) {
        post_error(nullptr, "atmo_post_create: device allocation failed");
        for (float* p : {h->bloom[0], h->bloom[1]})
            if (p) (void)hipFree(p);
        h->magic = 0;
        delete h;
        return ATMO_ENOMEM;
    }
    *out = h;
    return ATMO_OK;
}

int32_t atmo_post_destroy(AtmoPost* h) {
    if (!post_ok(h)) return ATMO_EHANDLE;
    (void)hipDeviceSynchronize();
    for (int i = 0; i < 2; ++i)
        if (h->bloom[i] && !h->borrowed[i]) (void)hipFree(h->bloom[i]);
    h->magic = 0;
    delete h;
    return ATMO_OK;
}

const char* atmo_post_last_error(const AtmoPost* h) {
    if (!h) return g_post_create_error;
    return post_ok(h) ? h->error : "invalid post-process handle";
}

int32_t atmo_post_get_buffer(AtmoPost* h, int32_t which, void** device_ptr, int64_t shape[4]) {
    if (!post_ok(h)) return ATMO_EHANDLE;
    if (which < 0 || which > 1 || !device_ptr || !shape) {
        post_error(h, "atmo_post_get_buffer: unknown buffer id or NULL argument");
        return ATMO_EINVAL;
    }
    *device_ptr = h->bloom[which];
    shape[0] = h->half_width;
    shape[1] = h->half_height;
    shape[2] = 3;
    shape[3] = 0;
    return ATMO_OK;
}

int32_t atmo_post_bind_buffer(AtmoPost* h, int32_t which, void* device_ptr, int64_t bytes) {
    if (!post_ok(h)) return ATMO_EHANDLE;
    if (which < 0 || which > 1 || !device_ptr) {
        post_error(h, "atmo_post_bind_buffer: unknown buffer id or NULL device pointer");
        return ATMO_EINVAL;
    }
    if (bytes != (int64_t)(bloom_floats(h) * sizeof(float))) {
        post_error(h, "atmo_post_bind_buffer: buffer size does not match the field shape");
        return ATMO_ESHAPE;
    }
    if (h->bloom[which] && !h->borrowed[which]) {
        if (hipDeviceSynchronize() != hipSuccess || hipFree(h->bloom[which]) != hipSuccess) {
            post_error(h, "atmo_post_bind_buffer: failed to release the owned buffer");
            return ATMO_EHIP;
        }
    }
    h->bloom[which] = (float*)device_ptr;
    h->borrowed[which] = true;
    return ATMO_OK;
}

int32_t atmo_post_process(AtmoPost* h, const float* hdr, float* display,
                          float exposure_multiplier, int32_t use_bloom, int32_t tone_map,
                          void* stream) {
    if (!post_ok(h)) return ATMO_EHANDLE;
    if (!hdr || !display) {
        post_error(h, "atmo_post_process: NULL argument");
        return ATMO_EINVAL;
    }
    if (use_bloom && h->params.bloom_passes == 0) {
        post_error(h, "atmo_post_process: bloom needs bloom_passes > 0");
        return ATMO_EINVAL;
    }
    PostArgs a{};
    a.hdr = hdr;
    a.display = display;
    a.bloom_a = h->bloom[0];
    a.bloom_b = h->bloom[1];
    a.width = h->width;
    a.height = h->height;
    a.half_width = h->half_width;
    a.half_height = h->half_height;
    a.width_f = (float)h->width;
    a.height_f = (float)h->height;
    a.half_width_f = (float)h->half_width;
    a.half_height_f = (float)h->half_height;
    a.exposure_multiplier = exposure_multiplier;
    a.bloom_strength = h->bloom_strength;
    a.bright = h->bright;
    a.tone_map = tone_map != 0;
    a.inv_height = 1.0 / (double)h->height;
    hipStream_t s = (hipStream_t)stream;
    hipError_t e = hipSuccess;
    if (use_bloom) {
        a.tiles_x = (h->half_width + kTile - 1) / kTile;
        a.tiles_y = (h->half_height + kTile - 1) / kTile;
        a.tiles_per_xcd = (a.tiles_x * a.tiles_y + kXcds - 1) / kXcds;
        switch (h->params.bloom_passes) {
            case 1: e = launch_blur<1>(a, s); break;
            case 2: e = launch_blur<2>(a, s); break;
            case 3: e = launch_blur<3>(a, s); break;
            case 4: e = launch_blur<4>(a, s); break;
            case 5: e = launch_blur<5>(a, s); break;
            case 6: e = launch_blur<6>(a, s); break;
            case 7: e = launch_blur<7>(a, s); break;
            default: e = launch_blur<8>(a, s); break;
        }
    }
    if (e == hipSuccess) {
        const size_t n = (size_t)h->width * h->height;
        const bool vec = (((uintptr_t)hdr | (uintptr_t)display) & 15) == 0;
        if (use_bloom) {
            const size_t chunks = (n + 4 * kThreads - 1) / (4 * kThreads);
            a.tiles_per_xcd = (int32_t)((chunks + kXcds - 1) / kXcds);
            const dim3 grid((unsigned)(kXcds * a.tiles_per_xcd));
            if (vec)
                hipLaunchKernelGGL(k_post_resolve_bloom<true>, grid, dim3(kThreads), 0, s, a);
            else
                hipLaunchKernelGGL(k_post_resolve_bloom<false>, grid, dim3(kThreads), 0, s, a);
        } else {
            const dim3 grid((unsigned)((3 * n + 4 * kThreads - 1) / (4 * kThreads)));
            if (vec)
                hipLaunchKernelGGL((k_post_resolve_flat<true>), grid, dim3(kThreads), 0, s, a);
            else
                hipLaunchKernelGGL((k_post_resolve_flat<false>), grid, dim3(kThreads), 0, s, a);
        }
        e = hipGetLastError();
    }
    if (e != hipSuccess) {
        post_error(h, hipGetErrorString(e));
        return ATMO_EHIP;
    }
    return ATMO_OK;
}

}  // extern "C"
