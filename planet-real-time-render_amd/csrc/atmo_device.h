// atmo_device.h -- device-side atmosphere model for gfx950.
//
// Each __device__ function computes what the @ti.func of the same name in
// the reference's src/planet_renderer/atmosphere/renderer.py computes (line
// ranges cited per function), in float32 with the reference's operation
// order.  The library is compiled with -ffp-contract=off, and every
// division and square root is correctly rounded (the compiler's IEEE
// expansions, or the shorter csqrt/crcp/cdiv/quot forms below, which give
// the same bits on their domains), so the only source of difference from an
// IEEE float32 evaluation of the reference formulas is the ulp-level error
// of the device exp/acos/cos/sin/pow/atan2.
//
// Elementary functions are correctly rounded float32 (cr_*: evaluated in
// float64, rounded once), the definition the CPU restatement uses, wherever
// the reference's formulas amplify last-bit differences: finite-disk
// visibility and the representative sun zenith near the horizon,
// (1 - exp(-x)) / s for small x together with the density exp that feeds
// x, the sky-view horizon coordinate.  Where nothing downstream amplifies a
// last-bit difference (exp of long segments) the hardware-rate function is
// used; each such site says so.  Each correctly rounded site was switched to
// the hardware-rate function in turn and the GPU parity tests run: four
// fail without it (the short-segment exp, the penumbra zeniths, the
// representative cosine above the band, the sky-view sampler's acos), the
// density exp brings the twilight pixel to the 1e-4 bar, and the disk
// visibility quadruples its worst error.  The Mie phase's power and the
// sun-disk distance left every test unchanged but saved no measurable time,
// so they stay correctly rounded like the CPU restatement.
//
// Frame- and LUT-sized arrays use the reference's Taichi field layout:
// (X, Y[, Z], 3) float32, element [x, y] at (x * Y + y) * 3.
//
// Two precision variants of the model functions exist, each in its own
// inline namespace selected by ATMO_ROBUST when a translation unit includes
// this header: atmo::ref (ATMO_ROBUST 0, the default) follows the reference
// operation for operation in float32; atmo::robust (ATMO_ROBUST 1) is the
// same float32 march with float64 islands where float32 cancels at planet
// scale -- the ray-sphere solves near tangency (sphere_roots, shell_roots,
// altitude_on_ray, distance_to_top, the transmittance-LUT cosine), the
// sky-view horizon split and the surface endpoint above 1000 km -- plus
// Kahan-compensated optical depth and an expm1-based segment integral.
// Types (Consts, f3) are shared by both.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace atmo {

// Per-state constants.  Values the reference reads through `ti.static` are
// rounded once from the configuration's float64 value; products of Python
// constants (16*pi, 1+g^2, 2g, 3(1-g^2), 8*pi*(2+g^2)) are folded in
// float64 first, as Python folds them before Taichi sees them
// (renderer.py:792-804).
struct Consts {
    float pi;
    float pi16;
    float large;
    float top;
    float beta_r[3], beta_ms[3], beta_me[3], beta_a[3], albedo[3];
    float h_r, h_m, peak, half_width;
    float mie_a, mie_b, mie_c, mie_d;
    float r_h_r, r_h_m, r_half_width, r_top;  // 1.0f / (h_r, h_m, half_width, top), rounded once
    int tw, th, mw, mh, sw, sh, aw, ah, ad;
    // the bilinear LUTs' sizes as floats, and their largest texel index
    float tw_f, th_f, mw_f, mh_f, sw_f, sh_f;
    float tw1_f, th1_f, mw1_f, mh1_f, sw1_f, sh1_f;
    int t_steps, ms_dirs, ms_steps, sv_steps, ap_spp, hz_steps, term_sub;
};

struct f3 {
    float x, y, z;
};

// a * b + c in one v_fma_f64 that reads c in place.  Left to itself the
// compiler evaluates a polynomial step with the two-address v_fmac_f64,
// whose addend is also its destination: with the coefficient c kept in a
// register across loop iterations, every step then costs a v_mov_b64 copy
// of c as well.  (Arithmetic only: no memory access, no wait.)
__device__ __forceinline__ double fma_c(double a, double b, double c) {
    double d;
    asm("v_fma_f64 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "v"(c));
    return d;
}

// acos in float64 for |x| <= 1, within 1.2 ulp: asin(s) = s + s w P(w),
// w = s^2 <= 1/4, with P the degree-12 Chebyshev fit of
// (asin(sqrt(w)) - sqrt(w)) / w^1.5 on [0, 1/4] (fit error 1.5e-17);
// acos(x) = pi/2 - asin(x) for |x| < 1/2, else 2 asin(sqrt((1 - |x|) / 2))
// reflected for x < 0.  The square root is a float rsqrt seed with two
// float64 Newton corrections.  Half the instructions of the library acos,
// whose |x| >= 1/2 branch carries a double-double reconstruction this
// model does not need: its float64 results are only ever rounded to float32.
__device__ __forceinline__ double asin_poly(double w) {
    double p = 0x1.d72b2bc8155f8p-6;
    p = fma_c(p, w, -0x1.e6aaa8a0a04ccp-7);
    p = fma_c(p, w, 0x1.1d189408314eep-6);
    p = fma_c(p, w, 0x1.65a9c4dfcf8b2p-8);
    p = fma_c(p, w, 0x1.52420b04b37bep-7);
    p = fma_c(p, w, 0x1.782651caa6547p-7);
    p = fma_c(p, w, 0x1.c9cf07674736ap-7);
    p = fma_c(p, w, 0x1.1c4d35cf95421p-6);
    p = fma_c(p, w, 0x1.6e8bb1c8209a2p-6);
    p = fma_c(p, w, 0x1.f1c71c1db0623p-6);
    p = fma_c(p, w, 0x1.6db6db6e31f13p-5);
    p = fma_c(p, w, 0x1.3333333332ecap-4);
    return fma_c(p, w, 0x1.5555555555556p-3);
}

__device__ __forceinline__ double acos_d(double x) {
    const double pio2_hi = 0x1.921fb54442d18p+0, pio2_lo = 0x1.1a62633145c07p-54;
    const double pi_hi = 0x1.921fb54442d18p+1, pi_lo = 0x1.1a62633145c07p-53;
    const double a = fabs(x);
    if (a < 0.5) {
        const double w = x * x;
        return (pio2_hi - x) + (pio2_lo - (x * w) * asin_poly(w));
    }
    const double w = fma(a, -0.5, 0.5);  // (1 - |x|) / 2, exact
    const double y = (double)__builtin_amdgcn_rsqf(fmaxf((float)w, 0x1p-96f));
    const double h = 0.5 * y;
    double s = w * y;
    s = fma(fma(-s, s, w), h, s);
    s = fma(fma(-s, s, w), h, s);
    const double q = fma(s * w, asin_poly(w), s);
    return x > 0.0 ? 2.0 * q : (pi_hi - 2.0 * q) + pi_lo;
}

// float64 square root and quotient for the robust precision's float64
// islands, from float32 seeds: v_rsq_f32 / v_rcp_f32 (2^-23), two Newton
// corrections in float64 (to ~2^-69 relative) and, for the quotient, one
// Markstein correction.  Within an ulp of the float64 result, so their
// float32 roundings are the correctly rounded ones except within ~2^-52 of
// a float32 rounding boundary.  Domain: x = 0 or x in [2^-96, 2^126), and
// |b| in the float32 normal range (squared lengths and radii of the model).
__device__ __forceinline__ double dsqrt(double x) {
    const double y = (double)__builtin_amdgcn_rsqf(fmaxf((float)x, 0x1p-96f));
    const double h = 0.5 * y;
    double g = x * y;
    g = fma(fma(-g, g, x), h, g);
    return fma(fma(-g, g, x), h, g);
}
__device__ __forceinline__ double ddiv(double a, double b) {
    double y = (double)__builtin_amdgcn_rcpf((float)b);
    y = fma(fma(-b, y, 1.0), y, y);
    y = fma(fma(-b, y, 1.0), y, y);
    const double q = a * y;
    return fma(fma(-q, b, a), y, q);
}

__device__ __forceinline__ float cr_exp(float x) { return (float)exp((double)x); }

__device__ __forceinline__ float cr_acos(float x) { return (float)acos_d((double)x); }
__device__ __forceinline__ float cr_cos(float x) { return (float)cos((double)x); }
__device__ __forceinline__ float cr_sin(float x) { return (float)sin((double)x); }
// cr_sin and cr_cos of one argument with one float64 range reduction.
__device__ __forceinline__ void cr_sincos(float x, float& s, float& c) {
    double sd, cd;
    sincos((double)x, &sd, &cd);
    s = (float)sd;
    c = (float)cd;
}
// x^1.5 as x sqrt(x) in float64, rounded once: like a float64 pow rounded
// to float32, it is the correctly rounded value unless x^1.5 lies within a
// few float64 ulps of a float32 rounding boundary.
__device__ __forceinline__ float cr_pow15(float x) {
    const double d = x;
    return (float)(d * sqrt(d));
}
__device__ __forceinline__ float cr_atan2(float y, float x) { return (float)atan2((double)y, (double)x); }

__device__ __forceinline__ f3 mk(float x, float y, float z) { return f3{x, y, z}; }
__device__ __forceinline__ f3 operator+(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ f3 operator-(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ f3 operator*(f3 a, f3 b) { return f3{a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ f3 operator/(f3 a, f3 b) { return f3{a.x / b.x, a.y / b.y, a.z / b.z}; }
__device__ __forceinline__ f3 operator*(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ f3 operator/(f3 a, float s) { return f3{a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ f3 operator-(f3 a) { return f3{-a.x, -a.y, -a.z}; }
__device__ __forceinline__ float dot(f3 a, f3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ __forceinline__ float clampf(float x, float lo, float hi) { return fminf(hi, fmaxf(lo, x)); }

// Short forms of the correctly rounded square root and quotient.  The
// compiler's IEEE float sqrt and division expand to 16 and 11 instructions,
// most of them for arguments this model never produces (subnormals,
// infinities, overflowing quotients); the forms below return the same bits
// on their stated domains in 6 instructions each, without branches.

// sqrt(x): y = v_rsq_f32(max(x, 2^-96)), g = x y, and one residual
// correction g + (x - g^2) y / 2.  Equal to the IEEE sqrtf for x = +-0,
// NaN and every float in [2^-96, 2^126) (all 2^32 inputs compared on
// gfx950).  On (0, 2^-96) it may differ by an ulp of a root below 2^-48;
// negative x is outside its domain (every call site clamps at 0).  The
// arguments here are squared planet-scale lengths and differences of
// cosines, 0 or far above 2^-96.
__device__ __forceinline__ float csqrt(float x) {
    const float y = __builtin_amdgcn_rsqf(fmaxf(x, 0x1p-96f));
    const float g = x * y;
    return fmaf(fmaf(-g, g, x), 0.5f * y, g);
}

// 1 / b correctly rounded: v_rcp_f32 and one Newton step y + y (1 - b y).
// Equal to the IEEE 1.0f / b for every |b| in [2^-100, 2^101) (every
// mantissa of those 201 binades, both signs, compared on gfx950).  The
// caller guarantees the domain.
__device__ __forceinline__ float crcp(float b) {
    const float y = __builtin_amdgcn_rcpf(b);
    return fmaf(fmaf(-b, y, 1.0f), y, y);
}

// a / b, the correctly rounded quotient, for a divisor with a precomputed
// rb = 1.0f / b (itself correctly rounded): q = a rb is within an ulp, the
// remainder a - q b is exact in one FMA, and one Markstein correction
// q + (a - q b) rb rounds to the quotient (Markstein's theorem, with no
// underflow).  Three instructions instead of the eleven of an IEEE
// division; equal to a / b for every float a in [1e-9, 4e6] and the
// constant divisors of this model (checked exhaustively).
__device__ __forceinline__ float quot(float a, float b, float rb) {
    const float q = a * rb;
    return fmaf(fmaf(-q, b, a), rb, q);
}

// a / b correctly rounded, for |b| in [2^-60, 2^60] and |a| <= 2^60 (the
// caller guarantees both: planet-scale lengths and cosines): quot() with
// the reciprocal from crcp().  Equal to the IEEE a / b (2^33 random pairs
// with both exponents in [-60, 60] compared on gfx950, and Markstein's
// theorem) except that -0 / b gives +0 and a subnormal quotient may differ
// by an ulp.
__device__ __forceinline__ float cdiv(float a, float b) { return quot(a, b, crcp(b)); }

// a / b elementwise (Taichi's vector / scalar), correctly rounded, for
// b in [1, 2^60] (radii clamped below at 1 m): one crcp() and three quot().
__device__ __forceinline__ f3 div3(f3 a, float b) {
    const float rb = crcp(b);
    return f3{quot(a.x, b, rb), quot(a.y, b, rb), quot(a.z, b, rb)};
}
// 1 / x: crcp() on its domain, the IEEE quotient elsewhere (zero, tiny,
// huge and non-finite x, on a branch).
__device__ __forceinline__ float rcp_any(float x) {
    if (__builtin_expect(fabsf(x) >= 0x1p-100f && fabsf(x) < 0x1p100f, 1)) return crcp(x);
    return 1.0f / x;
}
__device__ __forceinline__ float norm(f3 a) { return csqrt(dot(a, a)); }
// Taichi: normalized(v) = (1 / norm(v)) * v.
__device__ __forceinline__ f3 normalized(f3 a) {
    float inv = rcp_any(norm(a));
    return f3{inv * a.x, inv * a.y, inv * a.z};
}
__device__ __forceinline__ f3 maxs(f3 a, float s) { return f3{fmaxf(a.x, s), fmaxf(a.y, s), fmaxf(a.z, s)}; }
__device__ __forceinline__ f3 mins(f3 a, float s) { return f3{fminf(a.x, s), fminf(a.y, s), fminf(a.z, s)}; }
__device__ __forceinline__ f3 clamp3(f3 a, float lo, float hi) {
    return f3{clampf(a.x, lo, hi), clampf(a.y, lo, hi), clampf(a.z, lo, hi)};
}
__device__ __forceinline__ f3 exp3(f3 a) { return f3{cr_exp(a.x), cr_exp(a.y), cr_exp(a.z)}; }
__device__ __forceinline__ f3 arr3(const float* a) { return f3{a[0], a[1], a[2]}; }
// The texel at byte offset `off` of a LUT (offsets stay below 2^32).
__device__ __forceinline__ f3 lut3(const float* __restrict__ base, uint32_t off) {
    const float* p = reinterpret_cast<const float*>(reinterpret_cast<const char*>(base) + off);
    return f3{p[0], p[1], p[2]};
}
// Pixel i of a (W, H, 3) float32 frame field by a 32-bit byte offset: a
// frame holds at most 2^28 pixels (atmo_create rejects larger ones), so
// 12 i < 2^32 and the access takes the scalar-base, vector-offset form with
// no 64-bit address arithmetic.
__device__ __forceinline__ f3 pix3(const float* __restrict__ base, uint32_t i) {
    return lut3(base, i * 12u);
}
__device__ __forceinline__ void put3(float* __restrict__ base, uint32_t i, f3 v) {
    float* p = reinterpret_cast<float*>(reinterpret_cast<char*>(base) + i * 12u);
    p[0] = v.x;
    p[1] = v.y;
    p[2] = v.z;
}
__device__ __forceinline__ f3 ld3(const float* __restrict__ base, long i) {
    return f3{base[i * 3 + 0], base[i * 3 + 1], base[i * 3 + 2]};
}
__device__ __forceinline__ void st3(float* __restrict__ base, long i, f3 v) {
    base[i * 3 + 0] = v.x;
    base[i * 3 + 1] = v.y;
    base[i * 3 + 2] = v.z;
}

}  // namespace atmo

#ifndef ATMO_ROBUST
#define ATMO_ROBUST 0
#endif
#if ATMO_ROBUST
#define ATMO_PRECISION robust
#else
#define ATMO_PRECISION ref
#endif

namespace atmo {
inline namespace ATMO_PRECISION {

// ---------------------------------------------------------------- media

// renderer.py:168-192
__device__ __forceinline__ f3 density(const Consts& k, float h) {
    float rayleigh = 0.0f, mie = 0.0f, absorption = 0.0f;
    if (0.0f <= h && h <= k.top) {
        // Correctly rounded exp, as the exponent h / H (quot) is: the
        // extinction must carry the reference's bits.  One ulp of it moves
        // x = sigma * step, and for x < 1/16 that can flip the rounding of
        // the correctly rounded exp(-x) of the segment integral, which
        // (1 - exp(-x)) / sigma amplifies by 1 / x: a hardware-rate exp here
        // put a twilight pixel 9.6e-5 from the float32 restatement; this one
        // 3.8e-6.  h / H must be correctly rounded for the same reason:
        // twilight texels are lit through grazing solar paths of optical
        // depth ~40, where even the 1-2 ulp of h * (1 / H) show as 4e-4.
        rayleigh = cr_exp(-quot(h, k.h_r, k.r_h_r));
        mie = cr_exp(-quot(h, k.h_m, k.r_h_m));
        absorption = fmaxf(1.0f - quot(fabsf(h - k.peak), k.half_width, k.r_half_width), 0.0f);
    }
    return f3{rayleigh, mie, absorption};
}

// renderer.py:194-201, from a density triple.
__device__ __forceinline__ f3 extinction_of(const Consts& k, f3 d) {
    return (arr3(k.beta_r) * d.x + arr3(k.beta_me) * d.y) + arr3(k.beta_a) * d.z;
}

// renderer.py:203-209, from a density triple.
__device__ __forceinline__ f3 scattering_of(const Consts& k, f3 d) {
    return arr3(k.beta_r) * d.x + arr3(k.beta_ms) * d.y;
}

// ------------------------------------------------------------- geometry

// renderer.py:211-226
__device__ __forceinline__ void sphere_roots(const Consts& k, f3 origin, f3 ray, float radius,
                                             float& near, float& far, bool& hit) {
#if ATMO_ROBUST
    {
        // float64 island: |origin - ray (origin . ray)| cancels for
        // near-tangent rays at planet radius
        const double ox = origin.x, oy = origin.y, oz = origin.z;
        const double rx = ray.x, ry = ray.y, rz = ray.z;
        const double pr = (ox * rx + oy * ry) + oz * rz;
        const double px = ox - rx * pr, py = oy - ry * pr, pz = oz - rz * pr;
        const double perp = dsqrt((px * px + py * py) + pz * pz);
        const double clear = (double)radius - perp;
        hit = clear >= 0.0;
        near = k.large;
        far = -k.large;
        if (hit) {
            const double root = dsqrt(fmax(fmax(clear, 0.0) * ((double)radius + perp), 0.0));
            near = (float)(-pr - root);
            far = (float)(-pr + root);
        }
        return;
    }
#endif
    float projected = dot(origin, ray);
    float perpendicular = norm(origin - ray * projected);
    float clearance = radius - perpendicular;
    hit = clearance >= 0.0f;
    near = k.large;
    far = -k.large;
    if (hit) {
        float disc = fmaxf(clearance, 0.0f) * (radius + perpendicular);
        float root = csqrt(fmaxf(disc, 0.0f));
        near = -projected - root;
        far = -projected + root;
    }
}

// renderer.py:228-235
__device__ __forceinline__ float shell_rho(float altitude, float bottom) {
    float s = fmaxf(altitude, 0.0f);
    return csqrt(s * (2.0f * bottom + s));
}

// renderer.py:237-262
__device__ __forceinline__ float altitude_on_ray(float h0, float c, float t, float bottom) {
#if ATMO_ROBUST
    {
        // float64 island: the radial delta cancels near tangency
        const double a = fmax((double)h0, 0.0), b = bottom, cc = c, tt = t;
        const double delta = a * (2.0 * b + a) + 2.0 * (b + a) * cc * tt + tt * tt;
        const double len = dsqrt(fmax(b * b + delta, 0.0));
        return (float)ddiv(delta, fmax(len + b, 1.0));
    }
#endif
    float a = fmaxf(h0, 0.0f);
    float r0 = bottom + a;
    float delta = a * (2.0f * bottom + a) + 2.0f * r0 * c * t + t * t;
    float len = csqrt(fmaxf(bottom * bottom + delta, 0.0f));
    return cdiv(delta, fmaxf(len + bottom, 1.0f));
}

// renderer.py:264-301
__device__ __forceinline__ void shell_roots(const Consts& k, float h0, float c, float shell,
                                            float bottom, float& near, float& far, bool& hit) {
#if ATMO_ROBUST
    {
        // float64 island: the discriminant cancels near tangency
        const double b = bottom, h = h0, sh = shell;
        const double hl = (b + h) * (double)c;
        const double constant = (h - sh) * (2.0 * b + h + sh);
        const double disc = hl * hl - constant;
        hit = disc >= 0.0;
        near = k.large;
        far = -k.large;
        if (hit) {
            const double root = dsqrt(fmax(disc, 0.0));
            const double q = -hl - (hl < 0.0 ? -1.0 : 1.0) * root;
            double r0 = -hl - root, r1 = -hl + root;
            if (fabs(q) > (double)1.0e-12f) {
                r0 = q;
                r1 = ddiv(constant, q);
            }
            near = (float)fmin(r0, r1);
            far = (float)fmax(r0, r1);
        }
        return;
    }
#endif
    float r0 = bottom + h0;
    float hl = r0 * c;
    float constant = (h0 - shell) * (2.0f * bottom + h0 + shell);
    float disc = hl * hl - constant;
    hit = disc >= 0.0f;
    near = k.large;
    far = -k.large;
    if (hit) {
        float root = csqrt(fmaxf(disc, 0.0f));
        float sign = hl < 0.0f ? -1.0f : 1.0f;
        float q = -hl - sign * root;
        float root0 = -hl - root;
        float root1 = -hl + root;
        if (fabsf(q) > 1.0e-12f) {
            root0 = q;
            root1 = cdiv(constant, q);
        }
        near = fminf(root0, root1);
        far = fmaxf(root0, root1);
    }
}

// renderer.py:303-340
__device__ __forceinline__ void segment_from_altitude(const Consts& k, float h, float c,
                                                      float bottom, float& start, float& end,
                                                      bool& valid, bool& ground) {
    float on, of, gn, gf;
    bool oh, gh;
    shell_roots(k, h, c, k.top, bottom, on, of, oh);
    start = fmaxf(on, 0.0f);
    end = of;
    valid = oh && end > start;
    ground = false;
    shell_roots(k, h, c, 0.0f, bottom, gn, gf, gh);
    if (valid && gh) {
        float gd = k.large;
        const float eps = 1.0e-3f;
        if (gn > start + eps)
            gd = gn;
        else if (gf > start + eps)
            gd = gf;
        if (gd < end) {
            end = gd;
            ground = true;
        }
    }
}

// renderer.py:342-374
__device__ __forceinline__ void segment_vec(const Consts& k, f3 origin, f3 ray, float bottom,
                                            float& start, float& end, bool& valid, bool& ground) {
    float top_radius = bottom + k.top;
    float on, of, gn, gf;
    bool oh, gh;
    sphere_roots(k, origin, ray, top_radius, on, of, oh);
    start = fmaxf(on, 0.0f);
    end = of;
    valid = oh && end > start;
    ground = false;
    sphere_roots(k, origin, ray, bottom, gn, gf, gh);
    if (valid && gh) {
        float gd = k.large;
        float eps = fmaxf(0.001f, bottom * 1.0e-8f);
        if (gn > start + eps)
            gd = gn;
        else if (gf > start + eps)
            gd = gf;
        if (gd < end) {
            end = gd;
            ground = true;
        }
    }
}

// renderer.py:376-395
__device__ __forceinline__ void aerial_prefix(const Consts& k, float h, float c, float bottom,
                                              float& start, float& end, bool& valid) {
    bool ground;
    segment_from_altitude(k, h, c, bottom, start, end, valid, ground);
    if (valid && !ground) {
        float closest = -(bottom + h) * c;
        if (closest > start && closest < end) end = closest;
    }
    valid = valid && end > start;
}

// renderer.py:397-408
__device__ __forceinline__ float distance_to_top(const Consts& k, float h, float c, float bottom) {
#if ATMO_ROBUST
    {
        // float64 island: -r c + sqrt(r^2 c^2 + delta) cancels for c > 0
        const double radius = (double)bottom + fmax((double)h, 0.0), cc = c, top = k.top;
        const double delta = (top - (double)h) * (2.0 * (double)bottom + top + (double)h);
        const double disc = radius * radius * cc * cc + delta;
        return (float)fmax(-radius * cc + dsqrt(fmax(disc, 0.0)), 0.0);
    }
#endif
    float radius = bottom + fmaxf(h, 0.0f);
    float shell_delta = (k.top - h) * (2.0f * bottom + k.top + h);
    float disc = radius * radius * c * c + shell_delta;
    return fmaxf(-radius * c + csqrt(fmaxf(disc, 0.0f)), 0.0f);
}

// renderer.py:410-432
__device__ __forceinline__ void transmittance_uv_to_ray(const Consts& k, float u, float v,
                                                        float bottom, float& altitude,
                                                        float& cosine, float& distance) {
    float top = k.top;
    float sh = shell_rho(top, bottom);
    float rho = sh * clampf(v, 0.0f, 1.0f);
    float rl = csqrt(rho * rho + bottom * bottom);
    float alt = cdiv(rho * rho, fmaxf(rl + bottom, 1.0f));
    float radius = bottom + alt;
    float dmin = top - alt;
    float dmax = rho + sh;
    float d = dmin + clampf(u, 0.0f, 1.0f) * (dmax - dmin);
    float cs = 1.0f;
#if ATMO_ROBUST
    // float64 island: the numerator cancels near the horizon
    if (d > 1.0e-6f)
        cs = (float)((((double)top - alt) * (2.0 * (double)bottom + top + alt) - (double)d * d) /
                     (2.0 * (double)radius * d));
#else
    if (d > 1.0e-6f) cs = cdiv((top - alt) * (2.0f * bottom + top + alt) - d * d, 2.0f * radius * d);
#endif
    altitude = alt;
    cosine = clampf(cs, -1.0f, 1.0f);
    distance = d;
}

// renderer.py:434-452, for an altitude h already clamped to [0, top] and
// its rho = shell_rho(h, bottom).
__device__ __forceinline__ void transmittance_uv_of(const Consts& k, float h, float rho, float c,
                                                    float bottom, float& u, float& v) {
    float top = k.top;
    float sh = shell_rho(top, bottom);
    float d = distance_to_top(k, h, c, bottom);
    float dmin = top - h;
    float dmax = rho + sh;
    u = clampf(cdiv(d - dmin, fmaxf(dmax - dmin, 1.0e-6f)), 0.0f, 1.0f);
    const float shm = fmaxf(sh, 1.0e-6f);
    v = clampf(quot(rho, shm, crcp(shm)), 0.0f, 1.0f);
}

// renderer.py:434-452
__device__ __forceinline__ void transmittance_ray_to_uv(const Consts& k, float h, float c,
                                                        float bottom, float& u, float& v) {
    h = clampf(h, 0.0f, k.top);
    transmittance_uv_of(k, h, shell_rho(h, bottom), c, bottom, u, v);
}

// Warped interval boundaries (renderer.py:454-480).  The closest-approach
// classification depends only on the interval, not on the fraction, so it is
// evaluated once per interval; boundary() then reproduces the reference's
// per-fraction arithmetic exactly.
struct Warp {
    float start, end, closest;
    int mode;  // 0: near start, 1: near end, 2: interior
};

__device__ __forceinline__ Warp make_warp(float neg_origin_dot_ray, float start, float end) {
    Warp w;
    w.start = start;
    w.end = end;
    w.closest = clampf(neg_origin_dot_ray, start, end);
    float eps = fmaxf((end - start) * 1.0e-6f, 1.0e-5f);
    if (w.closest <= start + eps)
        w.mode = 0;
    else if (w.closest >= end - eps)
        w.mode = 1;
    else
        w.mode = 2;
    return w;
}

__device__ __forceinline__ float boundary(const Warp& w, float f) {
    if (w.mode == 0) return w.start + (w.end - w.start) * f * f;
    if (w.mode == 1) {
        float inv = 1.0f - f;
        return w.end - (w.end - w.start) * inv * inv;
    }
    if (f < 0.5f) {
        float local = f * 2.0f;
        float inv = 1.0f - local;
        return w.closest - (w.closest - w.start) * inv * inv;
    }
    float local = f * 2.0f - 1.0f;
    return w.closest + (w.end - w.closest) * local * local;
}

// ------------------------------------------------------ LUT coordinates

// renderer.py:482-490
__device__ __forceinline__ float sky_horizon_zenith(const Consts& k, float h, float bottom) {
    float a = fmaxf(h, 0.0f);
    float radius = bottom + a;
    float tc = cdiv(shell_rho(a, bottom), fmaxf(radius, 1.0f));
    return k.pi - cr_acos(clampf(tc, 0.0f, 1.0f));
}

// renderer.py:492-504
__device__ __forceinline__ float sky_v_to_zenith(const Consts& k, float v, float horizon) {
    if (v < 0.5f) {
        float inv = 1.0f - 2.0f * v;
        return horizon * (1.0f - inv * inv);
    }
    float local = 2.0f * v - 1.0f;
    return horizon + (k.pi - horizon) * local * local;
}

// renderer.py:506-520
__device__ __forceinline__ float sky_zenith_to_v(const Consts& k, float zenith, float horizon) {
    if (zenith < horizon) {
        float n = clampf(cdiv(zenith, fmaxf(horizon, 1.0e-6f)), 0.0f, 1.0f);
        return 0.5f * (1.0f - csqrt(fmaxf(1.0f - n, 0.0f)));
    }
    float n = clampf(cdiv(zenith - horizon, fmaxf(k.pi - horizon, 1.0e-6f)), 0.0f, 1.0f);
    return 0.5f + 0.5f * csqrt(n);
}

#if ATMO_ROBUST
// float64 island: the sky-view horizon split.  Near the horizon the view
// zenith and the horizon zenith agree to ~1e-6 rad, and the square-root
// concentration amplifies their float32 difference.
__device__ __forceinline__ double sky_horizon_zenith_w(float h, float bottom) {
    const double a = fmax((double)h, 0.0), b = bottom;
    const double rho = sqrt(a * (2.0 * b + a));
    return 3.14159265358979323846 - acos(fmin(fmax(rho / fmax(b + a, 1.0), 0.0), 1.0));
}

__device__ __forceinline__ float sky_zenith_to_v_w(double zenith, float h, float bottom) {
    const double horizon = sky_horizon_zenith_w(h, bottom);
    double v;
    if (zenith < horizon) {
        const double n = fmin(fmax(zenith / fmax(horizon, 1.0e-6), 0.0), 1.0);
        v = 0.5 * (1.0 - sqrt(fmax(1.0 - n, 0.0)));
    } else {
        const double n = fmin(fmax((zenith - horizon) / fmax(3.14159265358979323846 - horizon, 1.0e-6), 0.0), 1.0);
        v = 0.5 + 0.5 * sqrt(n);
    }
    return (float)v;
}

__device__ __forceinline__ double sky_v_to_zenith_w(float v, float h, float bottom) {
    const double horizon = sky_horizon_zenith_w(h, bottom);
    if (v < 0.5f) {
        const double inv = 1.0 - 2.0 * (double)v;
        return horizon * (1.0 - inv * inv);
    }
    const double local = 2.0 * (double)v - 1.0;
    return horizon + (3.14159265358979323846 - horizon) * local * local;
}
#endif

// renderer.py:522-542
__device__ __forceinline__ float sun_u_to_cosine(float u, float h, float bottom) {
    float radius = bottom + fmaxf(h, 0.0f);
    float hc = cdiv(-shell_rho(h, bottom), fmaxf(radius, 1.0f));
    float c;
    if (u < 0.5f) {
        float inv = 1.0f - 2.0f * u;
        c = hc - (hc + 1.0f) * inv * inv;
    } else {
        float local = 2.0f * u - 1.0f;
        c = hc + (1.0f - hc) * local * local;
    }
    return clampf(c, -1.0f, 1.0f);
}

// renderer.py:544-568
__device__ __forceinline__ float sun_cosine_to_u(float c, float h, float bottom) {
    float radius = bottom + fmaxf(h, 0.0f);
    float hc = cdiv(-shell_rho(h, bottom), fmaxf(radius, 1.0f));
    if (c < hc) {
        float n = clampf(cdiv(hc - c, fmaxf(hc + 1.0f, 1.0e-6f)), 0.0f, 1.0f);
        return 0.5f * (1.0f - csqrt(n));
    }
    float n = clampf(cdiv(c - hc, fmaxf(1.0f - hc, 1.0e-6f)), 0.0f, 1.0f);
    return 0.5f + 0.5f * csqrt(n);
}

// ------------------------------------------------------------- sampling

// Half-texel bilinear fetch (renderer.py:618-640, 655-683, 1503-1525).
// Texels are addressed by 32-bit byte offsets from the LUT's base (every LUT
// is far below 4 GiB), so each corner is one load with a scalar base and a
// vector offset, and the four offsets share one multiply.
__device__ __forceinline__ f3 bilinear(const float* __restrict__ lut, int w, int hgt, float wf,
                                       float hf, float w1f, float h1f, float u, float v) {
    float x = clampf(u * wf - 0.5f, 0.0f, w1f);
    float y = clampf(v * hf - 0.5f, 0.0f, h1f);
    // x, y >= 0: truncation is floor
    int x0 = (int)x;
    int y0 = (int)y;
    int x1 = min(x0 + 1, w - 1);
    int y1 = min(y0 + 1, hgt - 1);
    float fx = x - (float)x0;
    float fy = y - (float)y0;
    const uint32_t row = 12u * (uint32_t)hgt;
    const uint32_t o00 = (uint32_t)x0 * row + 12u * (uint32_t)y0;
    const uint32_t o10 = o00 + row * (uint32_t)(x1 - x0);
    const uint32_t dy = 12u * (uint32_t)(y1 - y0);
    const f3 c00 = lut3(lut, o00), c10 = lut3(lut, o10);
    const f3 c01 = lut3(lut, o00 + dy), c11 = lut3(lut, o10 + dy);
    f3 low = c00 * (1.0f - fx) + c10 * fx;
    f3 high = c01 * (1.0f - fx) + c11 * fx;
    return low * (1.0f - fy) + high * fy;
}

// renderer.py:609-640
__device__ __forceinline__ f3 sample_transmittance(const Consts& k, const float* __restrict__ tlut,
                                                   float h, float c, float bottom) {
    float u, v;
    transmittance_ray_to_uv(k, fmaxf(h, 0.0f), c, bottom, u, v);
    return bilinear(tlut, k.tw, k.th, k.tw_f, k.th_f, k.tw1_f, k.th1_f, u, v);
}

// renderer.py:642-683
__device__ __forceinline__ f3 sample_multi_scattering(const Consts& k,
                                                      const float* __restrict__ mslut, float h,
                                                      float sun_cosine, float bottom) {
    float af = clampf(quot(h, k.top, k.r_top), 0.0f, 1.0f);
    float u = sun_cosine_to_u(sun_cosine, fmaxf(h, 0.0f), bottom);
    float v = csqrt(af);
    return bilinear(mslut, k.mw, k.mh, k.mw_f, k.mh_f, k.mw1_f, k.mh1_f, u, v);
}

// ------------------------------------------------------- sun and phases

// renderer.py:685-726
__device__ __forceinline__ void solar_shadow_roots(f3 origin, f3 ray, f3 sun, float bottom,
                                                   float& r0, float& r1, bool& v0, bool& v1) {
    f3 op = origin - sun * dot(origin, sun);
    f3 rp = ray - sun * dot(ray, sun);
    float quadratic = dot(rp, rp);
    float hl = dot(op, rp);
    float pl = norm(op);
    float constant = (pl - bottom) * (pl + bottom);
    r0 = 0.0f;
    r1 = 0.0f;
    v0 = false;
    v1 = false;
    float disc = hl * hl - quadratic * constant;
    if (quadratic > 1.0e-8f && disc >= 0.0f) {
        float sr = csqrt(fmaxf(disc, 0.0f));
        r0 = cdiv(-hl - sr, quadratic);
        r1 = cdiv(-hl + sr, quadratic);
        f3 p0 = origin + ray * r0;
        f3 p1 = origin + ray * r1;
        v0 = dot(p0, sun) < 0.0f;
        v1 = dot(p1, sun) < 0.0f;
    }
}

// The quantities _solar_disk_visibility and _sample_solar_transmittance
// both derive from (altitude, sun cosine) (renderer.py:728-789), computed
// once and shared.
//
// Outside the penumbra band the visible fraction is exactly 1 or 0, and the
// band is decided from cosines alone: with c_h the horizon cosine, s_h its
// sine and d = 1.05 alpha, the sun is fully up when its cosine exceeds
// cos(acos(c_h) - d) = c_h cos d + s_h sin d and fully down below
// cos(acos(c_h) + d).  The 5 % margin dwarfs every rounding error, so the
// classification agrees with the correctly rounded zeniths.  Fully down, the
// transmittance is multiplied by 0 and nothing else is evaluated.  Fully
// up, the representative zenith is the sun zenith z = float(acos(x)),
// correctly rounded, and its cosine is cos(acos(x) + e) = x - sqrt(1 - x^2) e
// with e = z - acos(x) exact in float64: the correctly rounded cos(z)
// without a float64 cosine.  (A grazing solar ray's optical depth turns the
// ~1e-7 gap between x and cos(z) into ~1e-4 of transmittance, so z's
// rounding must be kept.)  Inside the band both zeniths are correctly
// rounded: there the disk visibility amplifies their last bits.
// float(cos(float(acos(x)))), both roundings correct, with one float64
// acos: cos(acos(x) + e) = x cos e - sqrt(1 - x^2) sin e with
// |e| < 2^-22, so cos e = 1 - O(1e-14) and sin e = e to far below float32
// rounding.  sqrt(1 - x^2) multiplies e, so it needs far less than
// float64 precision: a float rsqrt seed and one float64 Newton correction
// give it to ~2^-45, which moves x - sqrt(1 - x^2) e by < 1e-21.
__device__ __forceinline__ float cos_of_rounded_acos(float x) {
    const double xd = fmin(fmax((double)x, -1.0), 1.0);
    const double t = acos_d(xd);
    const double e = (double)(float)t - t;
    const double v = fmax(1.0 - xd * xd, 0.0);
    const double y = (double)__builtin_amdgcn_rsqf(fmaxf((float)v, 0x1p-96f));
    const double g = v * y;
    const double sine = fma(fma(-g, g, v), 0.5 * y, g);
    return (float)(xd - sine * e);
}

struct SunGeom {
    float visibility;
    float rc;   // cosine of the representative zenith
    float rho;  // shell_rho(h, bottom), shared with the transmittance lookup
    bool fast;  // outside the penumbra band: visibility exactly 0 or 1
};

__device__ __forceinline__ SunGeom sun_geometry(const Consts& k, float h, float sun_cosine,
                                                float bottom, float angular_radius) {
    SunGeom g;
    float radius = bottom + fmaxf(h, 0.0f);
    g.rho = shell_rho(h, bottom);
    const float rr = crcp(fmaxf(radius, 1.0f));
    float hc = quot(-g.rho, fmaxf(radius, 1.0f), rr);  // cdiv(-rho, max(radius, 1))
    {
        const float dl = 1.05f * angular_radius;
        const float cd = 1.0f - 0.5f * dl * dl, sd = dl;  // cos d, sin d to O(d^3)
        // the horizon sine sqrt(1 - hc^2) is bottom / radius, since
        // rho^2 = radius^2 - bottom^2; it only places the band edges
        const float hs = bottom * rr;
        const float up = hc * cd + hs * sd, down = hc * cd - hs * sd;
        if (sun_cosine <= down) {
            g.fast = true;
            g.visibility = 0.0f;
            g.rc = sun_cosine;
            return g;
        }
        if (sun_cosine >= up) {
            g.fast = true;
            g.visibility = 1.0f;
            g.rc = cos_of_rounded_acos(sun_cosine);
            return g;
        }
    }
    g.fast = false;
    // Both zeniths are correctly rounded: near the horizon the representative
    // zenith feeds cos() and the T-LUT distance coordinate, whose slope there
    // is ~planet radius per unit cosine.
    const float horizon_zenith = cr_acos(clampf(hc, -1.0f, 1.0f));
    const float sun_zenith = cr_acos(clampf(sun_cosine, -1.0f, 1.0f));
    float sep = horizon_zenith - sun_zenith;
    float n = clampf(cdiv(sep, fmaxf(angular_radius, 1.0e-6f)), -1.0f, 1.0f);
    if (n >= 1.0f)
        g.visibility = 1.0f;
    else if (n <= -1.0f)
        g.visibility = 0.0f;
    else
        g.visibility = cdiv(cr_acos(-n) + n * csqrt(fmaxf(1.0f - n * n, 0.0f)), k.pi);
    float rep = fminf(sun_zenith, horizon_zenith - angular_radius * 0.25f);
    g.rc = cr_cos(fmaxf(rep, 0.0f));
    return g;
}

// renderer.py:756-789, given the shared geometry.
__device__ __forceinline__ f3 solar_transmittance(const Consts& k, const float* __restrict__ tlut,
                                                  const SunGeom& g, float h, float bottom,
                                                  float angular_radius) {
    if (g.visibility == 0.0f) return mk(0.0f, 0.0f, 0.0f);
    // sample_transmittance(h, rc): shell_rho of clamp(max(h, 0), 0, top) is
    // g.rho below the top and shell_rho(top) above it
    const float hm = fmaxf(h, 0.0f);
    const float rho = hm <= k.top ? g.rho : shell_rho(k.top, bottom);
    float u, v;
    transmittance_uv_of(k, fminf(k.top, hm), rho, g.rc, bottom, u, v);
    f3 t = bilinear(tlut, k.tw, k.th, k.tw_f, k.th_f, k.tw1_f, k.th1_f, u, v);
    return g.fast ? t : t * g.visibility;
}

// renderer.py:791-793
__device__ __forceinline__ float rayleigh_phase(const Consts& k, float c) {
    return cdiv(3.0f * (1.0f + c * c), k.pi16);
}

// renderer.py:795-804
__device__ __forceinline__ float mie_phase(const Consts& k, float c) {
    float den = fmaxf(k.mie_a - k.mie_b * c, 1.0e-8f);
    return cdiv(k.mie_c * (1.0f + c * c), k.mie_d * cr_pow15(den));
}

// exp(-x) for the segment transmittance, x in [0, 80].  For x < 1/16 the
// reference's (1 - exp(-x)) / sigma cancels, so the last bit of exp(-x) is
// amplified by 1/x: there exp(-x) is its Taylor series evaluated in float64
// (truncation < 4e-17) and rounded once, i.e. correctly rounded.  Elsewhere
// the hardware-rate expf is within an ulp, amplified at most 16-fold.
__device__ __forceinline__ float exp_neg_segment(float x) {
    if (x < 0.0625f) {
        const double t = -(double)x;
        double p = 1.0 / 362880.0;
        p = fma_c(p, t, 1.0 / 40320.0);
        p = fma_c(p, t, 1.0 / 5040.0);
        p = fma_c(p, t, 1.0 / 720.0);
        p = fma_c(p, t, 1.0 / 120.0);
        p = fma_c(p, t, 1.0 / 24.0);
        p = fma_c(p, t, 1.0 / 6.0);
        p = fma(p, t, 0.5);
        p = fma(p, t, 1.0);
        p = fma(p, t, 1.0);
        return (float)p;
    }
    return expf(-x);
}

#if ATMO_ROBUST
// One channel of the robust segment: T = e^-x exactly as exp_neg_segment,
// and (1 - e^-x) / sigma without the cancellation of 1 - T for small x.
// For x < 1/16 both come from one polynomial: q = (1 - e^-x) / x is the
// Taylor sum whose last Horner step gives exp_neg_segment's e^-x = 1 - x q,
// so T keeps its bits and 1 - e^-x = x q, formed in float64 and rounded
// once, has no cancellation.  For x >= 1/16, 1 - e^-x >= 0.06 and the
// float32 1 - T carries it to about 1e-6, so no float64 is spent there (a
// float64 exp measured no accuracy gain against the float64 restatement on
// the limb-centred stress frames at any altitude).
__device__ __forceinline__ void robust_segment(float x, float sigma, float step, float& t_out,
                                               float& i_out) {
    if (x < 0.0625f) {
        const double t = -(double)x;
        double q = 1.0 / 362880.0;
        q = fma_c(q, t, 1.0 / 40320.0);
        q = fma_c(q, t, 1.0 / 5040.0);
        q = fma_c(q, t, 1.0 / 720.0);
        q = fma_c(q, t, 1.0 / 120.0);
        q = fma_c(q, t, 1.0 / 24.0);
        q = fma_c(q, t, 1.0 / 6.0);
        q = fma(q, t, 0.5);
        q = fma(q, t, 1.0);
        t_out = (float)fma(q, t, 1.0);
        i_out = sigma > 1.0e-12f ? cdiv((float)((double)x * q), sigma) : step;
        return;
    }
    t_out = expf(-x);
    i_out = sigma > 1.0e-12f ? cdiv(1.0f - t_out, sigma) : step;
}
#endif

// renderer.py:806-816
__device__ __forceinline__ void segment_integral(f3 ext, float step, f3& trans, f3& integ) {
    f3 x = mins(ext * step, 80.0f);
#if ATMO_ROBUST
    // 1 - exp(-x) cancels for small x: it is formed in float64, rounded once
    robust_segment(x.x, ext.x, step, trans.x, integ.x);
    robust_segment(x.y, ext.y, step, trans.y, integ.y);
    robust_segment(x.z, ext.z, step, trans.z, integ.z);
#else
    trans = f3{exp_neg_segment(x.x), exp_neg_segment(x.y), exp_neg_segment(x.z)};
    // (1 - T) * rcp(sigma): within 2 ulp of the quotient, not amplified
    integ.x = ext.x > 1.0e-12f ? (1.0f - trans.x) * __builtin_amdgcn_rcpf(ext.x) : step;
    integ.y = ext.y > 1.0e-12f ? (1.0f - trans.y) * __builtin_amdgcn_rcpf(ext.y) : step;
    integ.z = ext.z > 1.0e-12f ? (1.0f - trans.z) * __builtin_amdgcn_rcpf(ext.z) : step;
#endif
}

// Phase-weighted scattering coefficients of one ray: the view/sun cosine is
// constant along a ray, so P_R and P_M are evaluated once per ray
// (renderer.py:840-848).
struct RayPhase {
    float pr, pm;  // Rayleigh and Cornette-Shanks phase at the ray/sun cosine
};

__device__ __forceinline__ RayPhase ray_phase(const Consts& k, f3 ray, f3 sun) {
    RayPhase p;
    float c = clampf(dot(ray, sun), -1.0f, 1.0f);
    p.pr = rayleigh_phase(k, c);
    p.pm = mie_phase(k, c);
    return p;
}

// renderer.py:818-856.  `point_radius` is |point|; the caller's density
// triple `d` and altitude are those of `point`.
template <bool kMultiple>
__device__ __forceinline__ f3 scattering_source(const Consts& k, const float* __restrict__ tlut,
                                                const float* __restrict__ mslut, f3 point,
                                                float h, f3 d, const RayPhase& ph, f3 sun,
                                                f3 irradiance, float bottom, float ar) {
    float radius = norm(point);
    f3 radial = div3(point, fmaxf(radius, 1.0f));
    float sun_cosine = dot(radial, sun);
    SunGeom g = sun_geometry(k, h, sun_cosine, bottom, ar);
    f3 st = solar_transmittance(k, tlut, g, h, bottom, ar);
    f3 phase_term = (arr3(k.beta_r) * d.x) * ph.pr + (arr3(k.beta_ms) * d.y) * ph.pm;
    f3 source = (irradiance * st) * phase_term;
    if (kMultiple) {
        f3 ms = sample_multi_scattering(k, mslut, h, sun_cosine, bottom);
        source = source + scattering_of(k, d) * ms;
    }
    return source;
}

// One constant-source segment (renderer.py:858-899): returns L and T.
// `oc` is origin_radial . ray for the ray's origin on the local radial axis.
__device__ __forceinline__ void scattering_segment(const Consts& k, const float* __restrict__ tlut,
                                                   const float* __restrict__ mslut, f3 origin,
                                                   float h0, float oc, f3 ray, float start,
                                                   float end, const RayPhase& ph, f3 sun,
                                                   f3 irradiance, float bottom, float ar, f3& L,
                                                   f3& T) {
    float step = fmaxf(end - start, 0.0f);
    float distance = (start + end) * 0.5f;
    f3 point = origin + ray * distance;
    float alt = altitude_on_ray(h0, oc, distance, bottom);
    f3 d = density(k, alt);
    f3 integ;
    segment_integral(extinction_of(k, d), step, T, integ);
    L = scattering_source<true>(k, tlut, mslut, point, alt, d, ph, sun, irradiance, bottom, ar) *
        integ;
}

// One base interval [d0, d1] of the bounded direct limb integral: its
// terminator-refinement weight (_terminator_refinement_weight,
// renderer.py:901-966) and its coarse segment (in-scatter `cl`,
// transmittance `ct`) at the shared midpoint: the same operations as the
// first half of an iteration of surface_inscatter_t's interval loop.
// (surface_inscatter_t keeps its own copy: moved into a function, the
// coarse-only k_composite spills 16 bytes per lane at 128 VGPRs.)
__device__ __forceinline__ void limb_coarse(const Consts& k, const float* __restrict__ tlut,
                                            const float* __restrict__ mslut, float h0, f3 ray,
                                            f3 origin, float oc, float r0, float r1, bool v0,
                                            bool v1, const RayPhase& ph, f3 sun, f3 irradiance,
                                            float bottom, float ar, float d0, float d1,
                                            float& weight, f3& cl, f3& ct) {
    float mid = (d0 + d1) * 0.5f;
    float width = fmaxf(d1 - d0, 1.0e-4f);
    // with neither shadow root the proximity term is exactly 0
    float prox = 0.0f;
    if (v0 || v1) {
        float closest = 1.0e30f;
        if (v0) closest = fminf(closest, fabsf(mid - r0));
        if (v1) closest = fminf(closest, fabsf(mid - r1));
        prox = clampf(2.0f - cdiv(closest, width), 0.0f, 1.0f);
        prox = prox * prox * (3.0f - 2.0f * prox);
    }
    f3 point = origin + ray * mid;
    float radius = fmaxf(norm(point), 1.0f);
    f3 radial = div3(point, radius);
    float alt = altitude_on_ray(h0, oc, mid, bottom);
    float sun_cosine = dot(radial, sun);
    SunGeom g = sun_geometry(k, alt, sun_cosine, bottom, ar);
    weight = fmaxf(prox, 4.0f * g.visibility * (1.0f - g.visibility));
    float step = fmaxf(d1 - d0, 0.0f);
    f3 d = density(k, alt);
    f3 integ;
    segment_integral(extinction_of(k, d), step, ct, integ);
    f3 st = solar_transmittance(k, tlut, g, alt, bottom, ar);
    f3 phase_term = (arr3(k.beta_r) * d.x) * ph.pr + (arr3(k.beta_ms) * d.y) * ph.pm;
    f3 src = (irradiance * st) * phase_term;
    f3 ms = sample_multi_scattering(k, mslut, alt, sun_cosine, bottom);
    src = src + scattering_of(k, d) * ms;
    cl = src * integ;
}

// The bounded direct limb integral of a surface pixel
// (_integrate_surface_scattering, renderer.py:1667-1755, with the origin of
// _surface_aerial_scattering, renderer.py:1757-1786): base intervals with
// continuous terminator refinement (renderer.py:901-1032).  The refinement
// weight and the coarse segment of an interval share one midpoint, so the
// sun geometry at that point is computed once for both.
//
// kRefine = false is the coarse-only variant: it returns false as soon as an
// interval's refinement weight is positive (the pixel then needs the full
// variant), and otherwise produces exactly the full variant's result, since
// an interval with weight 0 contributes its coarse segment unchanged.  It
// carries none of the refinement's registers.
template <bool kRefine>
__device__ __forceinline__ bool surface_inscatter_t(const Consts& k, const float* __restrict__ tlut,
                                                    const float* __restrict__ mslut, float h0,
                                                    f3 ray, float surface_distance, f3 sun,
                                                    f3 irradiance, float bottom, float ar,
                                                    f3& out) {
    f3 radiance = mk(0.0f, 0.0f, 0.0f);
    f3 view_t = mk(1.0f, 1.0f, 1.0f);
    out = radiance;
    float on, of;
    bool oh;
    shell_roots(k, h0, ray.y, k.top, bottom, on, of, oh);
    float start = fmaxf(on, 0.0f);
    float end = fminf(surface_distance, of);
    if (!(oh && end > start)) return true;

    const f3 origin = mk(0.0f, bottom + h0, 0.0f);
    float r0, r1;
    bool v0, v1;
    solar_shadow_roots(origin, ray, sun, bottom, r0, r1, v0, v1);
    const Warp w = make_warp(-dot(origin, ray), start, end);
    const float origin_radius = fmaxf(norm(origin), 1.0f);
    const float oc = dot(div3(origin, origin_radius), ray);
    const RayPhase ph = ray_phase(k, ray, sun);
    const int n = k.hz_steps;
    const float fn = (float)n, rn = 1.0f / fn;
    float d0 = boundary(w, 0.0f / fn);
    for (int i = 0; i < n; ++i) {
        float d1 = boundary(w, quot((float)(i + 1), fn, rn));
        // _terminator_refinement_weight
        float mid = (d0 + d1) * 0.5f;
        float width = fmaxf(d1 - d0, 1.0e-4f);
        // with neither shadow root the proximity term is exactly 0
        float prox = 0.0f;
        if (v0 || v1) {
            float closest = 1.0e30f;
            if (v0) closest = fminf(closest, fabsf(mid - r0));
            if (v1) closest = fminf(closest, fabsf(mid - r1));
            prox = clampf(2.0f - cdiv(closest, width), 0.0f, 1.0f);
            prox = prox * prox * (3.0f - 2.0f * prox);
        }
        f3 point = origin + ray * mid;
        float radius = fmaxf(norm(point), 1.0f);
        f3 radial = div3(point, radius);
        float alt = altitude_on_ray(h0, oc, mid, bottom);
        float sun_cosine = dot(radial, sun);
        SunGeom g = sun_geometry(k, alt, sun_cosine, bottom, ar);
        float weight = fmaxf(prox, 4.0f * g.visibility * (1.0f - g.visibility));
        if (!kRefine && weight > 0.0f) return false;
        // coarse _scattering_segment at the same midpoint
        f3 cl, ct;
        {
            float step = fmaxf(d1 - d0, 0.0f);
            f3 d = density(k, alt);
            f3 integ;
            segment_integral(extinction_of(k, d), step, ct, integ);
            f3 st = solar_transmittance(k, tlut, g, alt, bottom, ar);
            f3 phase_term = (arr3(k.beta_r) * d.x) * ph.pr + (arr3(k.beta_ms) * d.y) * ph.pm;
            f3 src = (irradiance * st) * phase_term;
            f3 ms = sample_multi_scattering(k, mslut, alt, sun_cosine, bottom);
            src = src + scattering_of(k, d) * ms;
            cl = src * integ;
        }
        f3 il = cl, it = ct;
        if (kRefine && weight > 0.0f) {
            f3 fl = mk(0.0f, 0.0f, 0.0f);
            f3 ft = mk(1.0f, 1.0f, 1.0f);
            const int m = k.term_sub;
            for (int j = 0; j < m; ++j) {
                float f0 = (float)j / (float)m;
                float f1 = (float)(j + 1) / (float)m;
                float cs = d0 + (d1 - d0) * f0;
                float ce = d0 + (d1 - d0) * f1;
                f3 sl, stt;
                scattering_segment(k, tlut, mslut, origin, h0, oc, ray, cs, ce, ph, sun,
                                   irradiance, bottom, ar, sl, stt);
                fl = fl + ft * sl;
                ft = ft * stt;
            }
            il = cl * (1.0f - weight) + fl * weight;
            it = ct * (1.0f - weight) + ft * weight;
        }
        radiance = radiance + view_t * il;
        view_t = view_t * it;
        d0 = d1;
    }
    out = maxs(radiance, 0.0f);
    return true;
}

// surface_inscatter_t<true> spread over a group of G lanes (G >= hz_steps,
// a power of two; `sub` is the lane's index in the group, `stage_a` /
// `stage_b` the group's G slots of LDS).  Lane i takes base interval i's
// coarse segment and refinement weight.  The refined intervals' term_sub
// sub-segments each are then dealt out over all G lanes, G per round (an
// interval's sub-segments may span two rounds), and each refined interval's
// lane folds its own sub-segments from LDS in sub-segment order, across the
// rounds, exactly as the sequential loop does.  Last, the
// group folds the intervals' (in-scatter, transmittance) pairs in interval
// order by a butterfly, each step putting the lower half's pair before the
// upper half's.  The value equals surface_inscatter's up to the association
// of that last fold; every lane of the group returns it.  The pixel's
// arguments are group-uniform and the group's lanes are all active.
template <int G>
__device__ __forceinline__ f3 surface_inscatter_lanes(const Consts& k, const float* __restrict__ tlut,
                                                      const float* __restrict__ mslut, float h0,
                                                      f3 ray, float surface_distance, f3 sun,
                                                      f3 irradiance, float bottom, float ar,
                                                      int sub, float4* stage_a, float2* stage_b) {
    float on, of;
    bool oh;
    shell_roots(k, h0, ray.y, k.top, bottom, on, of, oh);
    float start = fmaxf(on, 0.0f);
    float end = fminf(surface_distance, of);
    if (!(oh && end > start)) return mk(0.0f, 0.0f, 0.0f);  // group-uniform

    const f3 origin = mk(0.0f, bottom + h0, 0.0f);
    float r0, r1;
    bool v0, v1;
    solar_shadow_roots(origin, ray, sun, bottom, r0, r1, v0, v1);
    const Warp w = make_warp(-dot(origin, ray), start, end);
    const float origin_radius = fmaxf(norm(origin), 1.0f);
    const float oc = dot(div3(origin, origin_radius), ray);
    const RayPhase ph = ray_phase(k, ray, sun);
    const int n = k.hz_steps;
    float d0 = 0.0f, d1 = 0.0f, weight = 0.0f;
    f3 cl = mk(0.0f, 0.0f, 0.0f), ct = mk(1.0f, 1.0f, 1.0f);
    if (sub < n) {
        const float fn = (float)n, rn = 1.0f / fn;
        // the same boundaries as the sequential loop's
        d0 = sub == 0 ? boundary(w, 0.0f / fn) : boundary(w, quot((float)sub, fn, rn));
        d1 = boundary(w, quot((float)(sub + 1), fn, rn));
        limb_coarse(k, tlut, mslut, h0, ray, origin, oc, r0, r1, v0, v1, ph, sun, irradiance,
                    bottom, ar, d0, d1, weight, cl, ct);
    }
    const bool refined = sub < n && weight > 0.0f;
    const int first_lane = (int)(threadIdx.x % 64) & ~(G - 1);
    const uint64_t ballot = __ballot(refined);
    const uint64_t mask = G == 64 ? ballot : (ballot >> first_lane) & ((1ull << (G & 63)) - 1ull);
    f3 fl = mk(0.0f, 0.0f, 0.0f), ft = mk(1.0f, 1.0f, 1.0f);
    if (mask != 0) {  // group-uniform
        const int m = k.term_sub;
        const int items = __popcll(mask) * m;
        // this lane's sub-segments are items [first, first + m) when it is refined
        const int first = __popcll(mask & ((1ull << sub) - 1ull)) * m;
        for (int round = 0; round < items; round += G) {
            const int t = min(round + sub, items - 1);
            const int kth = t / m, j = t - kth * m;
            uint64_t mm = mask;
            for (int q = 0; q < kth; ++q) mm &= mm - 1ull;
            const int iv = __ffsll((unsigned long long)mm) - 1;
            const float a0 = __shfl(d0, iv, G), a1 = __shfl(d1, iv, G);
            f3 sl = mk(0.0f, 0.0f, 0.0f), stt = mk(1.0f, 1.0f, 1.0f);
            if (round + sub < items) {
                // the sequential loop's sub-segment bounds
                float f0 = (float)j / (float)m;
                float f1 = (float)(j + 1) / (float)m;
                float cs = a0 + (a1 - a0) * f0;
                float ce = a0 + (a1 - a0) * f1;
                scattering_segment(k, tlut, mslut, origin, h0, oc, ray, cs, ce, ph, sun,
                                   irradiance, bottom, ar, sl, stt);
            }
            stage_a[sub] = make_float4(sl.x, sl.y, sl.z, stt.x);
            stage_b[sub] = make_float2(stt.y, stt.z);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            if (refined) {
                const int lo = max(first, round), hi = min(first + m, round + G);
                for (int q = lo; q < hi; ++q) {
                    const float4 va = stage_a[q - round];
                    const float2 vb = stage_b[q - round];
                    fl = fl + ft * mk(va.x, va.y, va.z);
                    ft = ft * mk(va.w, vb.x, vb.y);
                }
            }
            // every lane has read this round's slots before the next round writes them
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
    }
    f3 il = cl, it = ct;
    if (refined) {
        il = cl * (1.0f - weight) + fl * weight;
        it = ct * (1.0f - weight) + ft * weight;
    }
#pragma unroll
    for (int off = 1; off < G; off <<= 1) {
        const f3 pl = mk(__shfl_xor(il.x, off, G), __shfl_xor(il.y, off, G),
                         __shfl_xor(il.z, off, G));
        const f3 pt = mk(__shfl_xor(it.x, off, G), __shfl_xor(it.y, off, G),
                         __shfl_xor(it.z, off, G));
        // one expression for both lanes of a pair, so both hold the same bits
        const bool upper = (sub & off) != 0;
        const f3 lo_l = upper ? pl : il, lo_t = upper ? pt : it;
        const f3 hi_l = upper ? il : pl, hi_t = upper ? it : pt;
        il = lo_l + lo_t * hi_l;
        it = lo_t * hi_t;
    }
    return maxs(il, 0.0f);
}

__device__ __forceinline__ f3 surface_inscatter(const Consts& k, const float* __restrict__ tlut,
                                                const float* __restrict__ mslut, float h0, f3 ray,
                                                float surface_distance, f3 sun, f3 irradiance,
                                                float bottom, float ar) {
    f3 out;
    surface_inscatter_t<true>(k, tlut, mslut, h0, ray, surface_distance, sun, irradiance, bottom,
                              ar, out);
    return out;
}

// Camera-to-surface transmittance from two unoccluded T-LUT samples along
// the reversed ray (_surface_segment_transmittance, renderer.py:1603-1665).
__device__ __forceinline__ f3 surface_transmittance(const Consts& k,
                                                    const float* __restrict__ tlut, float h0,
                                                    f3 ray, float surface_distance,
                                                    float bottom) {
    f3 t = mk(1.0f, 1.0f, 1.0f);
    float on, of;
    bool oh;
    shell_roots(k, h0, ray.y, k.top, bottom, on, of, oh);
    float start = fmaxf(on, 0.0f);
    float end = fminf(surface_distance, of);
    if (oh && end > start) {
        const f3 origin = mk(0.0f, bottom + h0, 0.0f);
        f3 near_point = origin + ray * start;
        f3 surface_point = origin + ray * end;
        f3 reverse = -ray;
        f3 near_radial = div3(near_point, fmaxf(norm(near_point), 1.0f));
        f3 surface_radial = div3(surface_point, fmaxf(norm(surface_point), 1.0f));
        float near_alt = altitude_on_ray(h0, ray.y, start, bottom);
        float surface_alt = altitude_on_ray(h0, ray.y, end, bottom);
        f3 s2t = sample_transmittance(k, tlut, surface_alt, dot(surface_radial, reverse), bottom);
        f3 n2t = sample_transmittance(k, tlut, near_alt, dot(near_radial, reverse), bottom);
        const f3 q = maxs(n2t, 1.0e-6f);
        t = clamp3(mk(cdiv(s2t.x, q.x), cdiv(s2t.y, q.y), cdiv(s2t.z, q.z)), 0.0f, 1.0f);
    }
    return t;
}

// renderer.py:1788-1825
__device__ __forceinline__ f3 camera_transmittance(const Consts& k,
                                                   const float* __restrict__ tlut, float h,
                                                   f3 ray, float bottom) {
    float start, end;
    bool valid, ground;
    segment_from_altitude(k, h, ray.y, bottom, start, end, valid, ground);
    f3 t = mk(1.0f, 1.0f, 1.0f);
    if (valid) {
        if (ground) {
            t = mk(0.0f, 0.0f, 0.0f);
        } else {
            f3 origin = mk(0.0f, bottom + h, 0.0f);
            f3 entry = origin + ray * start;
            f3 radial = div3(entry, fmaxf(norm(entry), 1.0f));
            float ea = altitude_on_ray(h, ray.y, start, bottom);
            t = sample_transmittance(k, tlut, ea, dot(radial, ray), bottom);
        }
    }
    return t;
}

// renderer.py:1474-1525
// `horizon` is sky_horizon_zenith(k, h, bottom), constant per frame and
// computed once by the caller.
__device__ __forceinline__ f3 sample_sky_view(const Consts& k, const float* __restrict__ svlut,
                                              f3 ray, f3 sun, float h, float bottom,
                                              float horizon) {
    float zenith = cr_acos(clampf(ray.y, -1.0f, 1.0f));
    float vl = csqrt(ray.x * ray.x + ray.z * ray.z);
    float sl = csqrt(sun.x * sun.x + sun.z * sun.z);
    float rel = 0.0f;
    if (vl > 1.0e-6f && sl > 1.0e-6f)
        rel = cr_acos(clampf(cdiv(ray.x * sun.x + ray.z * sun.z, vl * sl), -1.0f, 1.0f));
    float u = cdiv(rel, k.pi);
#if ATMO_ROBUST
    (void)zenith;
    (void)horizon;
    float v = sky_zenith_to_v_w(acos(fmin(fmax((double)ray.y, -1.0), 1.0)), h, bottom);
#else
    (void)h;
    float v = sky_zenith_to_v(k, zenith, horizon);
#endif
    return bilinear(svlut, k.sw, k.sh, k.sw_f, k.sh_f, k.sw1_f, k.sh1_f, u, v);
}

// renderer.py:1527-1601
__device__ __forceinline__ f3 sample_aerial(const Consts& k, const float* __restrict__ ap,
                                            float su, float sv, float distance, float h, f3 ray,
                                            float bottom) {
    float start, end;
    bool valid;
    aerial_prefix(k, h, ray.y, bottom, start, end, valid);
    float df = 0.0f;
    if (valid) df = clampf(cdiv(distance - start, fmaxf(end - start, 1.0f)), 0.0f, 1.0f);
    const int aw = k.aw, ah = k.ah, ad = k.ad;
    float x = clampf(su * (float)aw - 0.5f, 0.0f, (float)(aw - 1));
    float y = clampf(sv * (float)ah - 0.5f, 0.0f, (float)(ah - 1));
    float z = clampf(df, 0.0f, 1.0f) * (float)(ad - 1);
    int x0 = (int)floorf(x), y0 = (int)floorf(y), z0 = (int)floorf(z);
    int x1 = min(x0 + 1, aw - 1), y1 = min(y0 + 1, ah - 1), z1 = min(z0 + 1, ad - 1);
    float fx = x - (float)x0, fy = y - (float)y0, fz = z - (float)z0;
    // 32-bit byte offsets of the eight corners (the froxel volume is far below 4 GiB)
    const uint32_t sz = 12u, sy = sz * (uint32_t)ad, sx = sy * (uint32_t)ah;
    const uint32_t o = (uint32_t)x0 * sx + (uint32_t)y0 * sy + (uint32_t)z0 * sz;
    const uint32_t dx = sx * (uint32_t)(x1 - x0), dy = sy * (uint32_t)(y1 - y0),
                   dz = sz * (uint32_t)(z1 - z0);
    f3 a = lut3(ap, o) * (1.0f - fx) + lut3(ap, o + dx) * fx;
    f3 b = lut3(ap, o + dy) * (1.0f - fx) + lut3(ap, o + dx + dy) * fx;
    f3 c = lut3(ap, o + dz) * (1.0f - fx) + lut3(ap, o + dx + dz) * fx;
    f3 e = lut3(ap, o + dy + dz) * (1.0f - fx) + lut3(ap, o + dx + dy + dz) * fx;
    f3 low = a * (1.0f - fy) + b * fy;
    f3 high = c * (1.0f - fy) + e * fy;
    return low * (1.0f - fz) + high * fz;
}

// Sky branch of the COMPOSITE view (renderer.py:1882-1940): sky-view
// radiance, scene and stars through the camera transmittance with the
// star-contrast fade, plus the finite sun disk with a one-pixel filter.
// The frame constants `horizon` (sky_horizon_zenith at the sky-view
// altitude) and `pixel_filter` (atan2(thf sqrt(2) / height, 1)) come from
// the caller.
__device__ __forceinline__ f3 composite_sky(const Consts& k, const float* __restrict__ tlut,
                                            const float* __restrict__ svlut, f3 ray, f3 color,
                                            f3 space, float cam_alt, float sky_alt, f3 sun, f3 disk,
                                            float ar, float pixel_filter, float horizon, float c0,
                                            float c1, float bottom) {
    f3 trans = camera_transmittance(k, tlut, cam_alt, ray, bottom);
    f3 sky = sample_sky_view(k, svlut, ray, sun, sky_alt, bottom, horizon);
    f3 star = space * trans;
    const f3 lum = mk(0.2126f, 0.7152f, 0.0722f);
    float contrast = cdiv(dot(star, lum), fmaxf(dot(sky, lum), 1.0e-5f));
    float vis = clampf(cdiv(contrast - c0, fmaxf(c1 - c0, 1.0e-5f)), 0.0f, 1.0f);
    vis = vis * vis * (3.0f - 2.0f * vis);
    color = (sky + color * trans) + star * vis;
    float angular_distance = cr_acos(clampf(dot(ray, sun), -1.0f, 1.0f));
    float cov = clampf(cdiv(ar + pixel_filter - angular_distance, fmaxf(2.0f * pixel_filter, 1.0e-8f)),
                       0.0f, 1.0f);
    cov = cov * cov * (3.0f - 2.0f * cov);
    return color + (disk * trans) * cov;
}

// Pixel-centre ray (renderer.py:1850-1859), y = 0 the bottom row.  The
// frame constants aspect = width / height and fwd_t = forward / thf come
// from the caller, rounded as the reference rounds them.
__device__ __forceinline__ f3 camera_ray(int px, int py, int width, int height, float aspect,
                                         f3 right, f3 up, f3 fwd_t, float& su, float& sv) {
    su = cdiv((float)px + 0.5f, (float)width);
    sv = cdiv((float)py + 0.5f, (float)height);
    float sx = (su * 2.0f - 1.0f) * aspect;
    float sy = sv * 2.0f - 1.0f;
    return normalized((right * sx + up * sy) + fwd_t);
}
__device__ __forceinline__ f3 camera_ray(int px, int py, int width, int height, float aspect,
                                         f3 right, f3 up, f3 fwd_t) {
    float su, sv;
    return camera_ray(px, py, width, height, aspect, right, up, fwd_t, su, sv);
}

// Compacts the flagged threads of a 256-thread block into list[0..count) in
// thread order (wave ballots and per-wave offsets); returns count.  Every
// thread of the block must call it.
__device__ __forceinline__ int compact_block(bool flag, int* list, int* wave_counts) {
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    const unsigned long long mask = __ballot(flag);
    if (lane == 0) wave_counts[wave] = __popcll(mask);
    __syncthreads();
    int offset = 0, total = 0;
    for (int w = 0; w < 4; ++w) {
        offset += w < wave ? wave_counts[w] : 0;
        total += wave_counts[w];
    }
    if (flag) list[offset + __popcll(mask & ((1ull << lane) - 1ull))] = threadIdx.x;
    __syncthreads();
    return total;
}

// Two-sided compaction of a 256-thread block: the threads flagged `a` go to
// list[0..na) and those flagged `b` (a and b exclusive) to list[na..na+nb),
// each in thread order; returns na and sets nb.  Every thread of the block
// must call it; `wave_counts` holds 8 ints.
__device__ __forceinline__ int compact_block2(bool a, bool b, int* list, int* wave_counts,
                                              int& nb) {
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    const unsigned long long ma = __ballot(a), mb = __ballot(b);
    if (lane == 0) {
        wave_counts[wave] = __popcll(ma);
        wave_counts[4 + wave] = __popcll(mb);
    }
    __syncthreads();
    int offset_a = 0, total_a = 0, offset_b = 0, total_b = 0;
    for (int w = 0; w < 4; ++w) {
        offset_a += w < wave ? wave_counts[w] : 0;
        total_a += wave_counts[w];
        offset_b += w < wave ? wave_counts[4 + w] : 0;
        total_b += wave_counts[4 + w];
    }
    const unsigned long long below = (1ull << lane) - 1ull;
    if (a) list[offset_a + __popcll(ma & below)] = threadIdx.x;
    if (b) list[total_a + offset_b + __popcll(mb & below)] = threadIdx.x;
    __syncthreads();
    nb = total_b;
    return total_a;
}

// Appends the flagged lanes' pixel indices to the deferred-pixel queue (one
// atomic per wave, by the lowest flagged lane).  Every lane of the wave
// must call it.  `capacity` is the number of index slots after the two
// counters; an index past it is dropped (a counter that did not start at
// zero can only come from an earlier call that failed to launch its list
// pass, and the host clears the counters before the next call then).
__device__ __forceinline__ void enqueue_wave(bool flag, long pixel, int* __restrict__ queue,
                                             int parity, long capacity) {
    const unsigned long long mask = __ballot(flag);
    if (!mask) return;
    const int lane = threadIdx.x % 64;
    const int leader = __ffsll((long long)mask) - 1;
    int slot = 0;
    if (lane == leader) slot = atomicAdd(queue + parity, __popcll(mask));
    slot = __shfl(slot, leader, 64);
    // flagged lanes below this one: v_mbcnt over the ballot, no 64-bit lane mask
    const unsigned below = __builtin_amdgcn_mbcnt_hi(
        (unsigned)(mask >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0u));
    const long at = (long)slot + below;
    if (flag && at < capacity) queue[2 + at] = (int)pixel;
}

}  // inline namespace ATMO_PRECISION
}  // namespace atmo
