/*
 * post_cpu.c -- CPU restatement of the reference's display post-process,
 * `PostProcessor` in src/planet_renderer/postprocess.py:56-207: the
 * half-resolution bright-pass extraction, the separable 5-tap blur passes,
 * and the resolve (exposure, bilinear bloom, ACES fit, sRGB encoding).
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker of libatmo.so's
 * post-process kernels; the package never loads it.
 *
 * Arithmetic: float32, operation for operation in the reference's order,
 * no FMA contraction (-ffp-contract=off), IEEE division, fmaxf/fminf for
 * ti.max/ti.min.  The one elementary function, the sRGB power
 * (postprocess.py:68), is correctly rounded: evaluated in float64 and
 * rounded once.  post_check_srgb_pow pins that against long double.
 *
 * Layout: the reference's Taichi fields, (X, Y, 3) float32 indexed [x, y],
 * element (x, y) at offset (x * Y + y) * 3.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#define EXPORT __attribute__((visibility("default")))

/* ti.pow(mapped, 1.0 / 2.4): the Python double 1/2.4 becomes a float32
 * constant (postprocess.py:68). */
#define SRGB_EXPONENT 0.416666657f

/* Bright-pass constants as the reference's kernels see them: float32
 * rounding of the Python values that `ti.static` folds
 * (postprocess.py:91-106). */
typedef struct PostConstants {
    float threshold;       /* bloom_threshold */
    float knee;            /* bloom_knee */
    float two_knee;        /* 2.0 * knee, Python double then float32 */
    float knee_denominator;/* max(4.0 * knee, 1e-6), Python double then float32 */
    float clamp;           /* bloom_clamp */
} PostConstants;

static inline float cr_powf(float x, float y) { return (float)pow((double)x, (double)y); }

static inline size_t at(int32_t x, int32_t y, int32_t ny) { return ((size_t)x * (size_t)ny + (size_t)y) * 3; }

/* _display_from_exposed, postprocess.py:56-71: ACES fit, clamp, sRGB. */
EXPORT float post_display_channel(float value) {
    value = fmaxf(value, 0.0f);
    const float a = 2.51f, b = 0.03f, c = 2.43f, d = 0.59f, e = 0.14f;
    float mapped = (value * (a * value + b)) / (value * (c * value + d) + e);
    mapped = fminf(fmaxf(mapped, 0.0f), 1.0f);
    const float lo = mapped * 12.92f;
    const float hi = 1.055f * cr_powf(mapped, SRGB_EXPONENT) - 0.055f;
    return mapped <= 0.0031308f ? lo : hi;
}

/* The sRGB power alone, for every element of an array. */
EXPORT void post_srgb_pow_array(const float* in, float* out, int64_t n) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) out[i] = cr_powf(in[i], SRGB_EXPONENT);
}

/* Number of float32 m in [lo, hi] whose float64 power rounded to float32
 * differs from the long double power rounded to float32: zero means the
 * oracle's power is the correctly rounded one on that range. */
EXPORT int64_t post_check_srgb_pow(float lo, float hi) {
    uint32_t blo, bhi;
    memcpy(&blo, &lo, 4);
    memcpy(&bhi, &hi, 4);
    int64_t bad = 0;
#pragma omp parallel for schedule(static) reduction(+ : bad)
    for (int64_t bits = blo; bits <= (int64_t)bhi; ++bits) {
        const uint32_t b32 = (uint32_t)bits;
        float m;
        memcpy(&m, &b32, 4);
        const float viad = cr_powf(m, SRGB_EXPONENT);
        const float vial = (float)powl((long double)m, (long double)SRGB_EXPONENT);
        bad += viad != vial;
    }
    return bad;
}

/* PostProcessor._soft_threshold, postprocess.py:88-108. */
static void soft_threshold(const PostConstants* k, const float value[3], float out[3]) {
    const float brightness = fmaxf(value[0], fmaxf(value[1], value[2]));
    float soft = fminf(fmaxf(brightness - k->threshold + k->knee, 0.0f), k->two_knee);
    soft = soft * soft / k->knee_denominator;
    float contribution = fmaxf(brightness - k->threshold, soft);
    contribution /= fmaxf(brightness, 1.0e-6f);
    float r[3];
    for (int c = 0; c < 3; ++c) r[c] = value[c] * contribution;
    const float peak = fmaxf(r[0], fmaxf(r[1], r[2]));
    const float scale = fminf(1.0f, k->clamp / fmaxf(peak, 1.0e-6f));
    for (int c = 0; c < 3; ++c) out[c] = r[c] * scale;
}

/* PostProcessor._extract, postprocess.py:110-119. */
EXPORT void post_extract(const float* hdr, int32_t width, int32_t height, float* bloom_a,
                         int32_t half_width, int32_t half_height, float exposure_multiplier,
                         const PostConstants* k) {
#pragma omp parallel for schedule(static)
    for (int32_t x = 0; x < half_width; ++x) {
        for (int32_t y = 0; y < half_height; ++y) {
            float value[3] = {0.0f, 0.0f, 0.0f};
            for (int ox = 0; ox < 2; ++ox) {
                for (int oy = 0; oy < 2; ++oy) {
                    const int32_t sx = x * 2 + ox < width - 1 ? x * 2 + ox : width - 1;
                    const int32_t sy = y * 2 + oy < height - 1 ? y * 2 + oy : height - 1;
                    const float* p = hdr + at(sx, sy, height);
                    for (int c = 0; c < 3; ++c) value[c] += p[c];
                }
            }
            const float s = 0.25f * exposure_multiplier;
            for (int c = 0; c < 3; ++c) value[c] *= s;
            soft_threshold(k, value, bloom_a + at(x, y, half_height));
        }
    }
}

static const float W0 = 0.38774f, W1 = 0.24477f, W2 = 0.06136f;

/* PostProcessor._blur_horizontal, postprocess.py:121-129: src -> dst. */
EXPORT void post_blur_horizontal(const float* src, float* dst, int32_t half_width,
                                 int32_t half_height) {
#pragma omp parallel for schedule(static)
    for (int32_t x = 0; x < half_width; ++x) {
        const int32_t xm1 = x - 1 > 0 ? x - 1 : 0, xm2 = x - 2 > 0 ? x - 2 : 0;
        const int32_t xp1 = x + 1 < half_width - 1 ? x + 1 : half_width - 1;
        const int32_t xp2 = x + 2 < half_width - 1 ? x + 2 : half_width - 1;
        for (int32_t y = 0; y < half_height; ++y) {
            for (int c = 0; c < 3; ++c) {
                float color = src[at(x, y, half_height) + c] * W0;
                color += src[at(xm1, y, half_height) + c] * W1;
                color += src[at(xp1, y, half_height) + c] * W1;
                color += src[at(xm2, y, half_height) + c] * W2;
                color += src[at(xp2, y, half_height) + c] * W2;
                dst[at(x, y, half_height) + c] = color;
            }
        }
    }
}

/* PostProcessor._blur_vertical, postprocess.py:131-139: src -> dst. */
EXPORT void post_blur_vertical(const float* src, float* dst, int32_t half_width,
                               int32_t half_height) {
#pragma omp parallel for schedule(static)
    for (int32_t x = 0; x < half_width; ++x) {
        for (int32_t y = 0; y < half_height; ++y) {
            const int32_t ym1 = y - 1 > 0 ? y - 1 : 0, ym2 = y - 2 > 0 ? y - 2 : 0;
            const int32_t yp1 = y + 1 < half_height - 1 ? y + 1 : half_height - 1;
            const int32_t yp2 = y + 2 < half_height - 1 ? y + 2 : half_height - 1;
            for (int c = 0; c < 3; ++c) {
                float color = src[at(x, y, half_height) + c] * W0;
                color += src[at(x, ym1, half_height) + c] * W1;
                color += src[at(x, yp1, half_height) + c] * W1;
                color += src[at(x, ym2, half_height) + c] * W2;
                color += src[at(x, yp2, half_height) + c] * W2;
                dst[at(x, y, half_height) + c] = color;
            }
        }
    }
}

static inline int32_t clampi(int32_t v, int32_t lo, int32_t hi) {
    return v < lo ? lo : (v > hi ? hi : v);
}

/* PostProcessor._sample_bloom, postprocess.py:141-153. */
static void sample_bloom(const float* bloom_a, int32_t half_width, int32_t half_height,
                         float u, float v, float out[3]) {
    const float px = u * (float)half_width - 0.5f;
    const float py = v * (float)half_height - 0.5f;
    const float fpx = floorf(px), fpy = floorf(py);
    const int32_t x0 = clampi((int32_t)fpx, 0, half_width - 1);
    const int32_t y0 = clampi((int32_t)fpy, 0, half_height - 1);
    const int32_t x1 = x0 + 1 < half_width - 1 ? x0 + 1 : half_width - 1;
    const int32_t y1 = y0 + 1 < half_height - 1 ? y0 + 1 : half_height - 1;
    const float fx = fminf(fmaxf(px - fpx, 0.0f), 1.0f);
    const float fy = fminf(fmaxf(py - fpy, 0.0f), 1.0f);
    for (int c = 0; c < 3; ++c) {
        const float low = bloom_a[at(x0, y0, half_height) + c] * (1.0f - fx) +
                          bloom_a[at(x1, y0, half_height) + c] * fx;
        const float high = bloom_a[at(x0, y1, half_height) + c] * (1.0f - fx) +
                           bloom_a[at(x1, y1, half_height) + c] * fx;
        out[c] = low * (1.0f - fy) + high * fy;
    }
}

/* PostProcessor._resolve, postprocess.py:155-174. */
EXPORT void post_resolve(const float* hdr, float* display, int32_t width, int32_t height,
                         const float* bloom_a, int32_t half_width, int32_t half_height,
                         float exposure_multiplier, float bloom_strength, int32_t use_bloom,
                         int32_t tone_map) {
#pragma omp parallel for schedule(static)
    for (int32_t x = 0; x < width; ++x) {
        for (int32_t y = 0; y < height; ++y) {
            const float* p = hdr + at(x, y, height);
            float value[3];
            for (int c = 0; c < 3; ++c) value[c] = p[c] * exposure_multiplier;
            if (use_bloom != 0) {
                const float u = ((float)x + 0.5f) / (float)width;
                const float v = ((float)y + 0.5f) / (float)height;
                float s[3];
                sample_bloom(bloom_a, half_width, half_height, u, v, s);
                for (int c = 0; c < 3; ++c) value[c] += s[c] * bloom_strength;
            }
            float* o = display + at(x, y, height);
            for (int c = 0; c < 3; ++c)
                o[c] = tone_map != 0 ? post_display_channel(value[c])
                                     : fminf(fmaxf(value[c], 0.0f), 1.0f);
        }
    }
}
