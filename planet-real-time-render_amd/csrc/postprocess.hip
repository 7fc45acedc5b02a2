// postprocess.hip -- the display post-process of libatmo.so: the kernels
// of the reference's `PostProcessor` (src/planet_renderer/postprocess.py:
// 74-207) on the composite's HDR output.
//
// The reference runs 2 + 2P Taichi kernels per frame: `_extract` (2x2 box
// downsample + bright pass into a half-resolution field), P passes of
// `_blur_horizontal` + `_blur_vertical`, and `_resolve` (exposure,
// bilinear bloom, ACES + sRGB or a [0, 1] clamp).  Here one kernel does
// all of it per tile:
//  - a 32x32 tile of half-resolution texels loads its halo of R = 2P + 1
//    texels, extracts it from the HDR frame into LDS, and runs the P
//    horizontal/vertical pass pairs in LDS (ping-pong, one barrier per
//    pass), each pass over the shrinking region that the next one reads;
//  - the tile's interior is written to bloom_a (last vertical pass) and
//    bloom_b (last horizontal pass), the fields the reference leaves;
//  - the tile's 64x64 full-resolution pixels are resolved from the blurred
//    texels still in LDS (the bilinear footprint of those pixels stays
//    within the tile plus one texel).
// Every texel goes through the reference's operations in the reference's
// order, so the result does not depend on the tiling: each pass reads the
// same clamped neighbours the full-frame pass would.  Without bloom the
// resolve runs alone, one pixel per lane.
//
// Tiles are dealt to the 8 XCDs in contiguous runs (block b runs on XCD
// b mod 8), so neighbouring tiles, which read each other's halo, share an
// L2.
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
constexpr int kTile = 32;                     // half-resolution texels per tile side
constexpr int kThreads = 256;
constexpr int kXcds = 8;

struct PostArgs {
    const float* hdr;
    float* display;
    float* bloom_a;
    float* bloom_b;
    int32_t width, height, half_width, half_height;
    float width_f, height_f, half_width_f, half_height_f;
    float exposure_multiplier, bloom_strength;
    post::BrightPass bright;
    int32_t tone_map;
    int32_t tiles_x, tiles_y, tiles_per_xcd;
};

__device__ __forceinline__ float output_channel(float value, int32_t tone_map) {
    return tone_map ? post::display_channel(value) : fminf(fmaxf(value, 0.0f), 1.0f);
}

// _extract + P x (_blur_horizontal, _blur_vertical) + _resolve, one tile per
// block (postprocess.py:110-174).
template <int P>
__global__ __launch_bounds__(kThreads) void k_post_bloom(PostArgs a) {
    constexpr int R = 2 * P + 1;       // halo: P pass pairs of 2 texels, + 1 for the bilinear
    constexpr int L = kTile + 2 * R;   // local side
    constexpr int N = L * L;
    __shared__ float lds[2 * 3 * N];
    float* A = lds;
    float* B = lds + 3 * N;

    const int tile = (int)(blockIdx.x % kXcds) * a.tiles_per_xcd + (int)(blockIdx.x / kXcds);
    if (tile >= a.tiles_x * a.tiles_y) return;  // whole block, before any barrier
    const int X0 = (tile / a.tiles_y) * kTile, Y0 = (tile % a.tiles_y) * kTile;
    const int gx0 = X0 - R, gy0 = Y0 - R;  // global texel of local (0, 0)
    const int hw = a.half_width, hh = a.half_height;

    // _extract (postprocess.py:110-119) over the whole local region; texels
    // outside the frame are zero and never read.
    for (int i = threadIdx.x; i < N; i += kThreads) {
        const int lx = i / L, ly = i - lx * L;
        const int gx = gx0 + lx, gy = gy0 + ly;
        float r = 0.0f, g = 0.0f, b = 0.0f;
        if (gx >= 0 && gx < hw && gy >= 0 && gy < hh) {
#pragma unroll
            for (int ox = 0; ox < 2; ++ox) {
                const int sx = min(gx * 2 + ox, a.width - 1);
#pragma unroll
                for (int oy = 0; oy < 2; ++oy) {
                    const int sy = min(gy * 2 + oy, a.height - 1);
                    const float* p = a.hdr + ((size_t)sx * a.height + sy) * 3;
                    r += p[0];
                    g += p[1];
                    b += p[2];
                }
            }
            const float s = 0.25f * a.exposure_multiplier;
            r *= s;
            g *= s;
            b *= s;
            post::soft_threshold(a.bright, r, g, b);
        }
        A[i] = r;
        A[N + i] = g;
        A[2 * N + i] = b;
    }
    __syncthreads();

    for (int pass = 1; pass <= P; ++pass) {
        // _blur_horizontal (postprocess.py:121-129): A -> B on the texels
        // the following vertical pass reads.
        {
            const int xlo = max(gx0 + 2 * pass, 0) - gx0;
            const int xhi = min(X0 + kTile + R - 2 * pass, hw) - gx0;
            const int ylo = max(gy0 + 2 * (pass - 1), 0) - gy0;
            const int yhi = min(Y0 + kTile + R - 2 * (pass - 1), hh) - gy0;
            for (int i = threadIdx.x; i < N; i += kThreads) {
                const int lx = i / L, ly = i - lx * L;
                if (lx < xlo || lx >= xhi || ly < ylo || ly >= yhi) continue;
                const int gx = gx0 + lx;
                const int m1 = (max(gx - 1, 0) - gx0) * L + ly;
                const int p1 = (min(gx + 1, hw - 1) - gx0) * L + ly;
                const int m2 = (max(gx - 2, 0) - gx0) * L + ly;
                const int p2 = (min(gx + 2, hw - 1) - gx0) * L + ly;
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const float* s = A + c * N;
                    float color = s[i] * post::kW0;
                    color += s[m1] * post::kW1;
                    color += s[p1] * post::kW1;
                    color += s[m2] * post::kW2;
                    color += s[p2] * post::kW2;
                    B[c * N + i] = color;
                }
            }
        }
        __syncthreads();
        // _blur_vertical (postprocess.py:131-139): B -> A.
        {
            const int xlo = max(gx0 + 2 * pass, 0) - gx0;
            const int xhi = min(X0 + kTile + R - 2 * pass, hw) - gx0;
            const int ylo = max(gy0 + 2 * pass, 0) - gy0;
            const int yhi = min(Y0 + kTile + R - 2 * pass, hh) - gy0;
            for (int i = threadIdx.x; i < N; i += kThreads) {
                const int lx = i / L, ly = i - lx * L;
                if (lx < xlo || lx >= xhi || ly < ylo || ly >= yhi) continue;
                const int gy = gy0 + ly;
                const int row = lx * L;
                const int m1 = row + max(gy - 1, 0) - gy0;
                const int p1 = row + min(gy + 1, hh - 1) - gy0;
                const int m2 = row + max(gy - 2, 0) - gy0;
                const int p2 = row + min(gy + 2, hh - 1) - gy0;
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const float* s = B + c * N;
                    float color = s[i] * post::kW0;
                    color += s[m1] * post::kW1;
                    color += s[p1] * post::kW1;
                    color += s[m2] * post::kW2;
                    color += s[p2] * post::kW2;
                    A[c * N + i] = color;
                }
            }
        }
        __syncthreads();
    }

    // The fields the reference leaves behind: bloom_a after the last
    // vertical pass, bloom_b after the last horizontal pass.
    for (int i = threadIdx.x; i < kTile * kTile; i += kThreads) {
        const int tx = i / kTile, ty = i - tx * kTile;
        const int gx = X0 + tx, gy = Y0 + ty;
        if (gx >= hw || gy >= hh) continue;
        const int li = (tx + R) * L + (ty + R);
        const size_t o = ((size_t)gx * hh + gy) * 3;
        a.bloom_a[o] = A[li];
        a.bloom_a[o + 1] = A[N + li];
        a.bloom_a[o + 2] = A[2 * N + li];
        a.bloom_b[o] = B[li];
        a.bloom_b[o + 1] = B[N + li];
        a.bloom_b[o + 2] = B[2 * N + li];
    }

    // _resolve (postprocess.py:155-174) with _sample_bloom (141-153) for the
    // tile's full-resolution pixels.  Each lane keeps one y (and so one v)
    // and steps x.
    constexpr int F = 2 * kTile;
    const int fy = (int)threadIdx.x % F;
    const int y = 2 * Y0 + fy;
    if (y >= a.height) return;  // no barrier follows
    const float v = ((float)y + 0.5f) / a.height_f;
    const float py = v * a.half_height_f - 0.5f;
    const float fpy = floorf(py);
    const int y0t = min(max((int)fpy, 0), hh - 1);
    const int y1t = min(y0t + 1, hh - 1);
    const float wy = fminf(fmaxf(py - fpy, 0.0f), 1.0f);
    // local rows, clamped into the tile so no access leaves LDS
    const int ly0 = min(max(y0t - gy0, 0), L - 1), ly1 = min(max(y1t - gy0, 0), L - 1);
    for (int fx = (int)threadIdx.x / F; fx < F; fx += kThreads / F) {
        const int x = 2 * X0 + fx;
        if (x >= a.width) break;
        const float u = ((float)x + 0.5f) / a.width_f;
        const float px = u * a.half_width_f - 0.5f;
        const float fpx = floorf(px);
        const int x0t = min(max((int)fpx, 0), hw - 1);
        const int x1t = min(x0t + 1, hw - 1);
        const float wx = fminf(fmaxf(px - fpx, 0.0f), 1.0f);
        const int lx0 = min(max(x0t - gx0, 0), L - 1) * L, lx1 = min(max(x1t - gx0, 0), L - 1) * L;
        const size_t o = ((size_t)x * a.height + y) * 3;
        const float* h = a.hdr + o;
        float* d = a.display + o;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float* s = A + c * N;
            const float low = s[lx0 + ly0] * (1.0f - wx) + s[lx1 + ly0] * wx;
            const float high = s[lx0 + ly1] * (1.0f - wx) + s[lx1 + ly1] * wx;
            const float sample = low * (1.0f - wy) + high * wy;
            float value = h[c] * a.exposure_multiplier;
            value += sample * a.bloom_strength;
            d[c] = output_channel(value, a.tone_map);
        }
    }
}

// _resolve without bloom (postprocess.py:155-174, use_bloom == 0).
__global__ __launch_bounds__(kThreads) void k_post_resolve(PostArgs a) {
    const size_t n = (size_t)a.width * a.height;
    const size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    const float* h = a.hdr + i * 3;
    float* d = a.display + i * 3;
#pragma unroll
    for (int c = 0; c < 3; ++c) d[c] = output_channel(h[c] * a.exposure_multiplier, a.tone_map);
}

template <int P>
hipError_t launch_bloom(const PostArgs& a, hipStream_t stream) {
    hipLaunchKernelGGL(k_post_bloom<P>, dim3(kXcds * a.tiles_per_xcd), dim3(kThreads), 0, stream, a);
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
    for (int i = 0; i < 2; ++i) {
        if (hipMalloc(&h->bloom[i], bloom_floats(h) * sizeof(float)) != hipSuccess ||
            hipMemset(h->bloom[i], 0, bloom_floats(h) * sizeof(float)) != hipSuccess) {
            post_error(nullptr, "atmo_post_create: device allocation failed");
            for (int j = 0; j <= i; ++j)
                if (h->bloom[j]) (void)hipFree(h->bloom[j]);
            h->magic = 0;
            delete h;
            return ATMO_ENOMEM;
        }
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
    hipStream_t s = (hipStream_t)stream;
    hipError_t e;
    if (use_bloom) {
        a.tiles_x = (h->half_width + kTile - 1) / kTile;
        a.tiles_y = (h->half_height + kTile - 1) / kTile;
        a.tiles_per_xcd = (a.tiles_x * a.tiles_y + kXcds - 1) / kXcds;
        switch (h->params.bloom_passes) {
            case 1: e = launch_bloom<1>(a, s); break;
            case 2: e = launch_bloom<2>(a, s); break;
            case 3: e = launch_bloom<3>(a, s); break;
            case 4: e = launch_bloom<4>(a, s); break;
            case 5: e = launch_bloom<5>(a, s); break;
            case 6: e = launch_bloom<6>(a, s); break;
            case 7: e = launch_bloom<7>(a, s); break;
            default: e = launch_bloom<8>(a, s); break;
        }
    } else {
        const size_t n = (size_t)h->width * h->height;
        hipLaunchKernelGGL(k_post_resolve, dim3((unsigned)((n + kThreads - 1) / kThreads)),
                           dim3(kThreads), 0, s, a);
        e = hipGetLastError();
    }
    if (e != hipSuccess) {
        post_error(h, hipGetErrorString(e));
        return ATMO_EHIP;
    }
    return ATMO_OK;
}

}  // extern "C"
