// post_device.h -- device arithmetic of the display post-process
// (src/planet_renderer/postprocess.py:56-174): the bright-pass threshold,
// the 5-tap blur weights, and the ACES + sRGB display transform.
//
// float32 in the reference's operation order (the library is built with
// -ffp-contract=off), IEEE division, fmaxf/fminf for ti.max/ti.min.  The
// sRGB power m^(float32(1/2.4)) is correctly rounded: a float64 refinement
// of a hardware seed, with the float64 library power for the rare inputs
// whose result lies near a float32 rounding boundary.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace post {

// ti.pow(mapped, 1.0 / 2.4): the Python double becomes a float32 constant.
constexpr float kSrgbExponent = 0.416666657f;
// (5/12 - kSrgbExponent) * ln 2.  kSrgbExponent = 5/12 - 2^-25/3 exactly.
constexpr double kDeltaLn2 = 0x1.5555555555555p-27 * 0x1.62e42fefa39efp-1;

// Blur taps of _blur_horizontal / _blur_vertical (postprocess.py:124-138).
constexpr float kW0 = 0.38774f, kW1 = 0.24477f, kW2 = 0.06136f;

// Slow path: the float64 library power, rounded once.
__device__ __noinline__ float srgb_pow_slow(float m) {
    return (float)pow((double)m, (double)kSrgbExponent);
}

// m^kSrgbExponent for m in (0.0031308, 1], correctly rounded to float32.
//
// y0 = exp2(p log2 m) from the hardware log2/exp2 (relative error ~2^-21).
// z = m^(5/12) from y0 by one third-order step on z^12 = m^5:
//   q = m^5 / y0^12 - 1,  z = y0 (1 + q/12 - 11 q^2/288),
// with 1/y0 from the float reciprocal and two float64 Newton steps, so z
// carries ~2^-52.  Then m^p = z * m^-delta, delta = 5/12 - p = 2^-25/3,
// through exp(-delta ln m) = 1 + g + g^2/2 with g = -delta ln m; ln m comes
// from the seed's hardware log2 of m.  |log2 m| <= 8.33 on the domain, so
// a log2 good to 1 ulp (2^-20 there) puts at most delta ln2 2^-20 = 2^-47.1
// into the result, and the other terms ~2^-50.  If the float64 value lies
// within 2^-45 of a float32 rounding boundary, twice that budget even for
// a 2-ulp log2, the library power decides (about one input in 1.4 million).
__device__ __forceinline__ float srgb_pow(float m) {
    const float l2 = __builtin_amdgcn_logf(m);
    const float y0f = __builtin_amdgcn_exp2f(kSrgbExponent * l2);
    const double y0 = y0f;
    double w = __builtin_amdgcn_rcpf(y0f);
    w = fma(w, fma(-y0, w, 1.0), w);
    w = fma(w, fma(-y0, w, 1.0), w);
    const double md = m;
    const double m2 = md * md, m4 = m2 * m2, m5 = m4 * md;
    const double w2 = w * w, w4 = w2 * w2, w8 = w4 * w4, w12 = w8 * w4;
    const double q = fma(m5, w12, -1.0);
    const double t = q * fma(q, -11.0 / 288.0, 1.0 / 12.0);
    const double g = (double)l2 * -kDeltaLn2;
    const double c = fma(g, fma(g, 0.5, 1.0), 1.0);
    const double y = fma(y0, t, y0) * c;
    const float up = (float)(y * (1.0 + 0x1p-45));
    const float down = (float)(y * (1.0 - 0x1p-45));
    return up == down ? up : srgb_pow_slow(m);
}

// _display_from_exposed (postprocess.py:56-71), one channel.
__device__ __forceinline__ float display_channel(float value) {
    value = fmaxf(value, 0.0f);
    float mapped = (value * (2.51f * value + 0.03f)) / (value * (2.43f * value + 0.59f) + 0.14f);
    mapped = fminf(fmaxf(mapped, 0.0f), 1.0f);
    if (mapped <= 0.0031308f) return mapped * 12.92f;
    return 1.055f * srgb_pow(mapped) - 0.055f;
}

// Bright-pass constants as the reference's kernels fold them (float32 of
// the Python values, postprocess.py:91-106).
struct BrightPass {
    float threshold, knee, two_knee, knee_denominator, clamp;
};

// PostProcessor._soft_threshold (postprocess.py:88-108), in place.
__device__ __forceinline__ void soft_threshold(const BrightPass& k, float& r, float& g, float& b) {
    const float brightness = fmaxf(r, fmaxf(g, b));
    float soft = fminf(fmaxf(brightness - k.threshold + k.knee, 0.0f), k.two_knee);
    soft = soft * soft / k.knee_denominator;
    float contribution = fmaxf(brightness - k.threshold, soft);
    contribution /= fmaxf(brightness, 1.0e-6f);
    r *= contribution;
    g *= contribution;
    b *= contribution;
    const float peak = fmaxf(r, fmaxf(g, b));
    const float scale = fminf(1.0f, k.clamp / fmaxf(peak, 1.0e-6f));
    r *= scale;
    g *= scale;
    b *= scale;
}

}  // namespace post
