/*
 * ref_cpu.c -- CPU oracle for the atmosphere hot path.  TEST INFRASTRUCTURE
 * ONLY: nothing in the shipped library calls or links this file.  Tests, the
 * smoke check and the CPU-baseline leg of bench.py load it as the checker.
 *
 * What it is: an operation-for-operation restatement, in C, of the Taichi
 * kernels and @ti.func helpers of the reference's
 * src/planet_renderer/atmosphere/renderer.py (lines 118-2055).  Every
 * function cites the reference lines it follows.  Built twice from this one
 * source by oracle/Makefile:
 *   REAL=float  (libref_cpu_f32.so) -- the float32 parity proxy for the
 *               reference's `ti.init(default_fp=ti.f32)` kernels;
 *   REAL=double (libref_cpu_f64.so) -- the same approximation re-integrated
 *               in float64 (the FP64 oracle; raise step counts and LUT sizes
 *               in AtmoParams for converged quadrature);
 *   REAL=float + ORACLE_ROBUST (libref_cpu_f32r.so) -- the float32 build
 *               with the robust-precision islands of the GPU library's
 *               robust mode: the ray-sphere solves (sphere_roots,
 *               shell_roots_from_altitude, altitude_on_ray, distance_to_top)
 *               and the transmittance-LUT cosine evaluated in float64 and
 *               rounded once, and Kahan-compensated optical-depth sums in
 *               integrate_extinction.  Not the reference's arithmetic: the
 *               checker of the robust GPU path, judged against the f64
 *               build.
 *
 * Semantic assumptions (the Taichi behaviour this restatement fixes):
 *  1. Literal constants are float32 in f32 kernels (default_fp=f32); the
 *     f32 build uses -fsingle-precision-constant, checked by a static
 *     assertion below.  Config values enter as float32(`ti.static(x)`).
 *  2. Products of Python-only constants are folded by Python in float64 and
 *     rounded once (16*pi, 1+g^2, 2g, 3(1-g^2), 8*pi*(2+g^2)); see
 *     make_consts().
 *  3. Arithmetic is IEEE-754 round-to-nearest with no FMA contraction and no
 *     reassociation (-ffp-contract=off, no -ffast-math), and every
 *     elementary function (exp, acos, cos, sin, pow, atan2) is correctly
 *     rounded: evaluated in float64 and rounded once to float32 (the cr_*
 *     helpers).  A libm's own float32 approximations are not used: glibc
 *     2.35's acosf differs from the correctly rounded value on ~8 % of
 *     arguments, and the reference's formulas amplify such last-bit
 *     differences up to ~1e3-fold (the finite-disk visibility near the
 *     horizon, (1 - exp(-x)) / sigma for small x, the sky-view horizon
 *     coordinate).  Taichi's own fast-math code generation is NOT
 *     reproduced: parity with Taichi's exp/acos/pow is unpinned by any
 *     reference test.
 *  4. Vector sums run left to right: dot(a,b) = (ax*bx + ay*by) + az*bz;
 *     norm = sqrt(dot(a,a)); normalized(v) = v * (1 / norm(v)).
 *  5. clamp(x, lo, hi) = min(hi, max(lo, x)); max/min are fmax/fmin.
 *  6. f32 / int promotes the int to f32 (`ti.cast(i, f32) / ti.static(n)`).
 *  7. `x ** 2` with an integer literal is x * x; `x ** 1.5` is pow(x, 1.5).
 *  8. Vector fields are (X, Y[, Z], 3) row-major, element [x, y] at
 *     (x * Y + y) * 3 -- Taichi's default AOS layout seen by to_numpy().
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <tgmath.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/atmo.h"

#ifndef REAL
#define REAL float
#endif
typedef REAL real;

/* Robust-precision islands: float64 evaluation inside the float32 build. */
typedef double wide;
#if defined(ORACLE_ROBUST) && !defined(ORACLE_F32)
#error "ORACLE_ROBUST is a variant of the float32 build"
#endif

#ifdef ORACLE_F32
_Static_assert(sizeof(0.5) == sizeof(float),
               "the f32 oracle must be built with -fsingle-precision-constant");
_Static_assert(sizeof(real) == sizeof(float), "ORACLE_F32 needs REAL=float");
#else
_Static_assert(sizeof(real) == sizeof(double), "f64 oracle needs REAL=double");
#endif

#define EXPORT __attribute__((visibility("default")))

/* Correctly rounded elementary functions (assumption 3). */
#ifdef ORACLE_F32
static inline float cr_exp(float x) { return (float)exp((double)x); }
static inline float cr_acos(float x) { return (float)acos((double)x); }
static inline float cr_cos(float x) { return (float)cos((double)x); }
static inline float cr_sin(float x) { return (float)sin((double)x); }
static inline float cr_pow(float x, float y) { return (float)pow((double)x, (double)y); }
static inline float cr_atan2(float y, float x) { return (float)atan2((double)y, (double)x); }
#else
#define cr_exp exp
#define cr_acos acos
#define cr_cos cos
#define cr_sin sin
#define cr_pow pow
#define cr_atan2 atan2
#endif

/* ------------------------------------------------------------------ */
/* Constants and configuration (renderer.py:12-13, 118-166).           */
/* ------------------------------------------------------------------ */

typedef struct {
    real pi;            /* _PI */
    real pi16;          /* 16.0 * _PI (Python-folded) */
    real large;         /* _LARGE_DISTANCE */
    real top;
    real beta_r[3], beta_ms[3], beta_me[3], beta_a[3], albedo[3];
    real h_r, h_m, peak, half_width;
    real mie_a, mie_b, mie_c, mie_d; /* 1+g^2, 2g, 3(1-g^2), 8 pi (2+g^2) */
    int tw, th, mw, mh, sw, sh, aw, ah, ad;
    int t_steps, ms_dirs, ms_steps, sv_steps, ap_spp, hz_steps, term_sub;
} K;

static K make_consts(const AtmoParams* p) {
    K k;
    /* Python float64 arithmetic, then one rounding to REAL. */
    const double pi_d = acos((double)-1.0);
    const double g = p->mie_phase_g;
    const double two = 2, three = 3, eight = 8, sixteen = 16, one = 1;
    k.pi = (real)pi_d;
    k.pi16 = (real)(sixteen * pi_d);
    k.large = (real)1.0e30;
    k.top = (real)p->top_altitude_m;
    for (int c = 0; c < 3; ++c) {
        k.beta_r[c] = (real)p->rayleigh_scattering_per_m[c];
        k.beta_ms[c] = (real)p->mie_scattering_per_m[c];
        k.beta_me[c] = (real)p->mie_extinction_per_m[c];
        k.beta_a[c] = (real)p->absorption_extinction_per_m[c];
        k.albedo[c] = (real)p->ground_albedo[c];
    }
    k.h_r = (real)p->rayleigh_scale_height_m;
    k.h_m = (real)p->mie_scale_height_m;
    k.peak = (real)p->absorption_peak_altitude_m;
    k.half_width = (real)p->absorption_half_width_m;
    k.mie_a = (real)(one + g * g);
    k.mie_b = (real)(two * g);
    k.mie_c = (real)(three * (one - g * g));
    k.mie_d = (real)(eight * pi_d * (two + g * g));
    k.tw = p->transmittance_lut_width;
    k.th = p->transmittance_lut_height;
    k.mw = p->multi_scattering_lut_width;
    k.mh = p->multi_scattering_lut_height;
    k.sw = p->sky_view_lut_width;
    k.sh = p->sky_view_lut_height;
    k.aw = p->aerial_lut_width;
    k.ah = p->aerial_lut_height;
    k.ad = p->aerial_lut_depth;
    k.t_steps = p->transmittance_steps;
    k.ms_dirs = p->multi_scattering_directions;
    k.ms_steps = p->multi_scattering_steps;
    k.sv_steps = p->sky_view_steps;
    k.ap_spp = p->aerial_steps_per_slice;
    k.hz_steps = p->aerial_horizon_raymarch_steps;
    k.term_sub = p->aerial_terminator_substeps;
    return k;
}

/* ------------------------------------------------------------------ */
/* Small vector type with Taichi's evaluation order.                   */
/* ------------------------------------------------------------------ */

typedef struct { real x, y, z; } v3;

static inline v3 V(real x, real y, real z) { v3 r = {x, y, z}; return r; }
static inline v3 vadd(v3 a, v3 b) { return V(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 vsub(v3 a, v3 b) { return V(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 vmul(v3 a, v3 b) { return V(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline v3 vdiv(v3 a, v3 b) { return V(a.x / b.x, a.y / b.y, a.z / b.z); }
static inline v3 vs(v3 a, real s) { return V(a.x * s, a.y * s, a.z * s); }
static inline v3 vdivs(v3 a, real s) { return V(a.x / s, a.y / s, a.z / s); }
static inline v3 vneg(v3 a) { return V(-a.x, -a.y, -a.z); }
static inline real vdot(v3 a, v3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
static inline real vnorm(v3 a) { return sqrt(vdot(a, a)); }
static inline v3 vnormalized(v3 a) {
    real inv = 1.0 / vnorm(a);
    return V(inv * a.x, inv * a.y, inv * a.z);
}
static inline real clampr(real x, real lo, real hi) { return fmin(hi, fmax(lo, x)); }
static inline v3 vmaxs(v3 a, real s) { return V(fmax(a.x, s), fmax(a.y, s), fmax(a.z, s)); }
static inline v3 vmins(v3 a, real s) { return V(fmin(a.x, s), fmin(a.y, s), fmin(a.z, s)); }
static inline v3 vclamp(v3 a, real lo, real hi) {
    return V(clampr(a.x, lo, hi), clampr(a.y, lo, hi), clampr(a.z, lo, hi));
}
static inline v3 vexp(v3 a) { return V(cr_exp(a.x), cr_exp(a.y), cr_exp(a.z)); }
static inline v3 varr(const real* a) { return V(a[0], a[1], a[2]); }
static inline v3 vload(const real* base, long idx) {
    return V(base[idx * 3 + 0], base[idx * 3 + 1], base[idx * 3 + 2]);
}
static inline void vstore(real* base, long idx, v3 v) {
    base[idx * 3 + 0] = v.x;
    base[idx * 3 + 1] = v.y;
    base[idx * 3 + 2] = v.z;
}

/* ------------------------------------------------------------------ */
/* Media (renderer.py:168-209).                                         */
/* ------------------------------------------------------------------ */

/* renderer.py:168-192 */
static inline v3 density(const K* k, real h) {
    int inside = 0.0 <= h && h <= k->top;
    real rayleigh = 0.0, mie = 0.0, absorption = 0.0;
    if (inside) {
        rayleigh = cr_exp(-h / k->h_r);
        mie = cr_exp(-h / k->h_m);
        absorption = fmax(1.0 - fabs(h - k->peak) / k->half_width, 0.0);
    }
    return V(rayleigh, mie, absorption);
}

/* renderer.py:194-201 */
static inline v3 extinction(const K* k, real h) {
    v3 d = density(k, h);
    return vadd(vadd(vs(varr(k->beta_r), d.x), vs(varr(k->beta_me), d.y)),
                vs(varr(k->beta_a), d.z));
}

/* renderer.py:203-209 */
static inline v3 scattering(const K* k, real h) {
    v3 d = density(k, h);
    return vadd(vs(varr(k->beta_r), d.x), vs(varr(k->beta_ms), d.y));
}

/* ------------------------------------------------------------------ */
/* Geometry (renderer.py:211-480).                                      */
/* ------------------------------------------------------------------ */

/* renderer.py:211-226 */
static inline void sphere_roots(const K* k, v3 origin, v3 ray, real radius,
                                real* near, real* far, int* hit) {
#ifdef ORACLE_ROBUST
    {
        /* float64 island: |origin - ray (origin . ray)| cancels for
         * near-tangent rays at planet radius */
        wide ox = origin.x, oy = origin.y, oz = origin.z;
        wide rx = ray.x, ry = ray.y, rz = ray.z;
        wide pr = (ox * rx + oy * ry) + oz * rz;
        wide px = ox - rx * pr, py = oy - ry * pr, pz = oz - rz * pr;
        wide perp = sqrt((px * px + py * py) + pz * pz);
        wide clear = (wide)radius - perp;
        *hit = clear >= 0.0;
        *near = k->large;
        *far = -k->large;
        if (*hit) {
            wide root = sqrt(fmax(fmax(clear, 0.0) * ((wide)radius + perp), 0.0));
            *near = (real)(-pr - root);
            *far = (real)(-pr + root);
        }
        return;
    }
#endif
    real projected = vdot(origin, ray);
    real perpendicular = vnorm(vsub(origin, vs(ray, projected)));
    real clearance = radius - perpendicular;
    *hit = clearance >= 0.0;
    *near = k->large;
    *far = -k->large;
    if (*hit) {
        real discriminant = fmax(clearance, 0.0) * (radius + perpendicular);
        real root = sqrt(fmax(discriminant, 0.0));
        *near = -projected - root;
        *far = -projected + root;
    }
}

/* renderer.py:228-235 */
static inline real shell_rho(real altitude, real bottom) {
    real safe = fmax(altitude, 0.0);
    return sqrt(safe * (2.0 * bottom + safe));
}

/* renderer.py:237-262 */
static inline real altitude_on_ray(real h0, real c, real t, real bottom) {
#ifdef ORACLE_ROBUST
    {
        /* float64 island: the radial delta cancels near tangency */
        wide a = fmax((wide)h0, 0.0), b = bottom, cc = c, tt = t;
        wide delta = a * (2.0 * b + a) + 2.0 * (b + a) * cc * tt + tt * tt;
        wide len = sqrt(fmax(b * b + delta, 0.0));
        return (real)(delta / fmax(len + b, 1.0));
    }
#endif
    real altitude = fmax(h0, 0.0);
    real origin_radius = bottom + altitude;
    real radial_delta = altitude * (2.0 * bottom + altitude)
                        + 2.0 * origin_radius * c * t + t * t;
    real radial_length = sqrt(fmax(bottom * bottom + radial_delta, 0.0));
    return radial_delta / fmax(radial_length + bottom, 1.0);
}

/* renderer.py:264-301 */
static inline void shell_roots_from_altitude(const K* k, real h0, real c,
                                             real shell, real bottom,
                                             real* near, real* far, int* hit) {
#ifdef ORACLE_ROBUST
    {
        /* float64 island: the discriminant cancels near tangency */
        wide b = bottom, h = h0, sh = shell;
        wide half_linear = (b + h) * (wide)c;
        wide constant = (h - sh) * (2.0 * b + h + sh);
        wide disc = half_linear * half_linear - constant;
        *hit = disc >= 0.0;
        *near = k->large;
        *far = -k->large;
        if (*hit) {
            wide root = sqrt(fmax(disc, 0.0));
            wide q = -half_linear - (half_linear < 0.0 ? -1.0 : 1.0) * root;
            wide r0 = -half_linear - root, r1 = -half_linear + root;
            if (fabs(q) > 1.0e-12) {
                r0 = q;
                r1 = constant / q;
            }
            *near = (real)fmin(r0, r1);
            *far = (real)fmax(r0, r1);
        }
        return;
    }
#endif
    real origin_radius = bottom + h0;
    real half_linear = origin_radius * c;
    real constant = (h0 - shell) * (2.0 * bottom + h0 + shell);
    real discriminant = half_linear * half_linear - constant;
    *hit = discriminant >= 0.0;
    *near = k->large;
    *far = -k->large;
    if (*hit) {
        real root = sqrt(fmax(discriminant, 0.0));
        real sign = 1.0;
        if (half_linear < 0.0) sign = -1.0;
        real q = -half_linear - sign * root;
        real root0 = -half_linear - root;
        real root1 = -half_linear + root;
        if (fabs(q) > 1.0e-12) {
            root0 = q;
            root1 = constant / q;
        }
        *near = fmin(root0, root1);
        *far = fmax(root0, root1);
    }
}

/* renderer.py:303-340 */
static inline void atmosphere_segment_from_altitude(const K* k, real h, real c,
                                                    real bottom, real* start,
                                                    real* end, int* valid,
                                                    int* ground) {
    real on, of, gn, gf;
    int oh, gh;
    shell_roots_from_altitude(k, h, c, k->top, bottom, &on, &of, &oh);
    *start = fmax(on, 0.0);
    *end = of;
    *valid = oh && *end > *start;
    *ground = 0;
    shell_roots_from_altitude(k, h, c, 0.0, bottom, &gn, &gf, &gh);
    if (*valid && gh) {
        real ground_distance = k->large;
        real eps = 1.0e-3;
        if (gn > *start + eps)
            ground_distance = gn;
        else if (gf > *start + eps)
            ground_distance = gf;
        if (ground_distance < *end) {
            *end = ground_distance;
            *ground = 1;
        }
    }
}

/* renderer.py:342-374 */
static inline void atmosphere_segment(const K* k, v3 origin, v3 ray,
                                      real bottom, real* start, real* end,
                                      int* valid, int* ground) {
    real top_radius = bottom + k->top;
    real on, of, gn, gf;
    int oh, gh;
    sphere_roots(k, origin, ray, top_radius, &on, &of, &oh);
    *start = fmax(on, 0.0);
    *end = of;
    *valid = oh && *end > *start;
    *ground = 0;
    sphere_roots(k, origin, ray, bottom, &gn, &gf, &gh);
    if (*valid && gh) {
        real ground_distance = k->large;
        real eps = fmax(0.001, bottom * 1.0e-8);
        if (gn > *start + eps)
            ground_distance = gn;
        else if (gf > *start + eps)
            ground_distance = gf;
        if (ground_distance < *end) {
            *end = ground_distance;
            *ground = 1;
        }
    }
}

/* renderer.py:376-395 */
static inline void aerial_prefix_segment(const K* k, real h, real c,
                                         real bottom, real* start, real* end,
                                         int* valid) {
    int ground;
    atmosphere_segment_from_altitude(k, h, c, bottom, start, end, valid, &ground);
    if (*valid && !ground) {
        real closest = -(bottom + h) * c;
        if (closest > *start && closest < *end) *end = closest;
    }
    *valid = *valid && *end > *start;
}

/* renderer.py:397-408 */
static inline real distance_to_top(const K* k, real h, real c, real bottom) {
#ifdef ORACLE_ROBUST
    {
        /* float64 island: -r c + sqrt(r^2 c^2 + delta) cancels for c > 0 */
        wide radius = (wide)bottom + fmax((wide)h, 0.0), cc = c, top = k->top;
        wide delta = (top - (wide)h) * (2.0 * (wide)bottom + top + (wide)h);
        wide disc = radius * radius * cc * cc + delta;
        return (real)fmax(-radius * cc + sqrt(fmax(disc, 0.0)), 0.0);
    }
#endif
    real radius = bottom + fmax(h, 0.0);
    real shell_delta = (k->top - h) * (2.0 * bottom + k->top + h);
    real discriminant = radius * radius * c * c + shell_delta;
    return fmax(-radius * c + sqrt(fmax(discriminant, 0.0)), 0.0);
}

/* renderer.py:410-432 */
static inline void transmittance_uv_to_ray(const K* k, real u, real v,
                                           real bottom, real* altitude,
                                           real* cosine, real* distance) {
    real top = k->top;
    real shell_horizon = shell_rho(top, bottom);
    real rho = shell_horizon * clampr(v, 0.0, 1.0);
    real radial_length = sqrt(rho * rho + bottom * bottom);
    real alt = rho * rho / fmax(radial_length + bottom, 1.0);
    real radius = bottom + alt;
    real dmin = top - alt;
    real dmax = rho + shell_horizon;
    real d = dmin + clampr(u, 0.0, 1.0) * (dmax - dmin);
    real cs = 1.0;
#ifdef ORACLE_ROBUST
    /* float64 island: the numerator cancels near the horizon */
    if (d > 1.0e-6)
        cs = (real)((((wide)top - alt) * (2.0 * (wide)bottom + top + alt) - (wide)d * d) /
                    (2.0 * (wide)radius * d));
#else
    if (d > 1.0e-6)
        cs = ((top - alt) * (2.0 * bottom + top + alt) - d * d) / (2.0 * radius * d);
#endif
    *altitude = alt;
    *cosine = clampr(cs, -1.0, 1.0);
    *distance = d;
}

/* renderer.py:434-452 */
static inline void transmittance_ray_to_uv(const K* k, real h, real c,
                                           real bottom, real* u, real* v) {
    real top = k->top;
    h = clampr(h, 0.0, top);
    real shell_horizon = shell_rho(top, bottom);
    real rho = shell_rho(h, bottom);
    real d = distance_to_top(k, h, c, bottom);
    real dmin = top - h;
    real dmax = rho + shell_horizon;
    real uu = (d - dmin) / fmax(dmax - dmin, 1.0e-6);
    real vv = rho / fmax(shell_horizon, 1.0e-6);
    *u = clampr(uu, 0.0, 1.0);
    *v = clampr(vv, 0.0, 1.0);
}

/* renderer.py:454-480 */
static inline real warped_boundary(v3 origin, v3 ray, real start, real end,
                                   real fraction) {
    real closest = clampr(-vdot(origin, ray), start, end);
    real distance = start;
    real epsilon = fmax((end - start) * 1.0e-6, 1.0e-5);
    if (closest <= start + epsilon) {
        distance = start + (end - start) * fraction * fraction;
    } else if (closest >= end - epsilon) {
        real inverse = 1.0 - fraction;
        distance = end - (end - start) * inverse * inverse;
    } else if (fraction < 0.5) {
        real local = fraction * 2.0;
        real inverse = 1.0 - local;
        distance = closest - (closest - start) * inverse * inverse;
    } else {
        real local = fraction * 2.0 - 1.0;
        distance = closest + (end - closest) * local * local;
    }
    return distance;
}

/* ------------------------------------------------------------------ */
/* LUT coordinates (renderer.py:482-568).                               */
/* ------------------------------------------------------------------ */

/* renderer.py:482-490 */
static inline real sky_horizon_zenith(const K* k, real h, real bottom) {
    real altitude = fmax(h, 0.0);
    real radius = bottom + altitude;
    real tangent_cosine = shell_rho(altitude, bottom) / fmax(radius, 1.0);
    return k->pi - cr_acos(clampr(tangent_cosine, 0.0, 1.0));
}

/* renderer.py:492-504 */
static inline real sky_v_to_zenith(const K* k, real v, real h, real bottom) {
    real horizon = sky_horizon_zenith(k, h, bottom);
    real zenith;
    if (v < 0.5) {
        real inverse = 1.0 - 2.0 * v;
        zenith = horizon * (1.0 - inverse * inverse);
    } else {
        real local = 2.0 * v - 1.0;
        zenith = horizon + (k->pi - horizon) * local * local;
    }
    return zenith;
}

/* renderer.py:506-520 */
static inline real sky_zenith_to_v(const K* k, real zenith, real h, real bottom) {
    real horizon = sky_horizon_zenith(k, h, bottom);
    real v;
    if (zenith < horizon) {
        real normalized = clampr(zenith / fmax(horizon, 1.0e-6), 0.0, 1.0);
        v = 0.5 * (1.0 - sqrt(fmax(1.0 - normalized, 0.0)));
    } else {
        real normalized =
            clampr((zenith - horizon) / fmax(k->pi - horizon, 1.0e-6), 0.0, 1.0);
        v = 0.5 + 0.5 * sqrt(normalized);
    }
    return v;
}

#ifdef ORACLE_ROBUST
/* float64 island: the sky-view horizon split.  Near the horizon the view
 * zenith and the horizon zenith agree to ~1e-6 rad, and the square-root
 * concentration amplifies their float32 difference. */
static inline wide sky_horizon_zenith_w(const K* k, real h, real bottom) {
    wide a = fmax((wide)h, 0.0), b = bottom;
    wide rho = sqrt(a * (2.0 * b + a));
    return acos((wide)-1.0) - acos(fmin(fmax(rho / fmax(b + a, 1.0), 0.0), 1.0));
}

static inline real sky_zenith_to_v_w(const K* k, wide zenith, real h, real bottom) {
    wide pi = acos((wide)-1.0);
    wide horizon = sky_horizon_zenith_w(k, h, bottom);
    wide v;
    if (zenith < horizon) {
        wide n = fmin(fmax(zenith / fmax(horizon, 1.0e-6), 0.0), 1.0);
        v = 0.5 * (1.0 - sqrt(fmax(1.0 - n, 0.0)));
    } else {
        wide n = fmin(fmax((zenith - horizon) / fmax(pi - horizon, 1.0e-6), 0.0), 1.0);
        v = 0.5 + 0.5 * sqrt(n);
    }
    return (real)v;
}

static inline wide sky_v_to_zenith_w(const K* k, real v, real h, real bottom) {
    wide pi = acos((wide)-1.0);
    wide horizon = sky_horizon_zenith_w(k, h, bottom);
    if (v < 0.5) {
        wide inverse = 1.0 - 2.0 * (wide)v;
        return horizon * (1.0 - inverse * inverse);
    }
    wide local = 2.0 * (wide)v - 1.0;
    return horizon + (pi - horizon) * local * local;
}
#endif

/* renderer.py:522-542 */
static inline real sun_u_to_cosine(real u, real h, real bottom) {
    real radius = bottom + fmax(h, 0.0);
    real horizon_cosine = -shell_rho(h, bottom) / fmax(radius, 1.0);
    real cosine;
    if (u < 0.5) {
        real inverse = 1.0 - 2.0 * u;
        cosine = horizon_cosine - (horizon_cosine + 1.0) * inverse * inverse;
    } else {
        real local = 2.0 * u - 1.0;
        cosine = horizon_cosine + (1.0 - horizon_cosine) * local * local;
    }
    return clampr(cosine, -1.0, 1.0);
}

/* renderer.py:544-568 */
static inline real sun_cosine_to_u(real cosine, real h, real bottom) {
    real radius = bottom + fmax(h, 0.0);
    real horizon_cosine = -shell_rho(h, bottom) / fmax(radius, 1.0);
    real u;
    if (cosine < horizon_cosine) {
        real normalized = clampr((horizon_cosine - cosine)
                                     / fmax(horizon_cosine + 1.0, 1.0e-6),
                                 0.0, 1.0);
        u = 0.5 * (1.0 - sqrt(normalized));
    } else {
        real normalized = clampr((cosine - horizon_cosine)
                                     / fmax(1.0 - horizon_cosine, 1.0e-6),
                                 0.0, 1.0);
        u = 0.5 + 0.5 * sqrt(normalized);
    }
    return u;
}

/* ------------------------------------------------------------------ */
/* Optical depth and LUT sampling (renderer.py:570-683).                */
/* ------------------------------------------------------------------ */

/* renderer.py:570-607 */
/* origin_altitude and origin_cosine (the ray's cosine to the origin's
 * local vertical) are used only by the robust build's float64 altitude. */
static inline v3 integrate_extinction(const K* k, v3 origin, v3 ray, real start,
                                      real end, real bottom, real origin_altitude,
                                      real origin_cosine) {
    v3 optical_depth = V(0.0, 0.0, 0.0);
#ifdef ORACLE_ROBUST
    v3 carry = V(0.0, 0.0, 0.0); /* Kahan compensation */
#endif
    for (int index = 0; index < k->t_steps; ++index) {
        real f0 = (real)index / (real)k->t_steps;
        real f1 = (real)(index + 1) / (real)k->t_steps;
        real d0 = warped_boundary(origin, ray, start, end, f0);
        real d1 = warped_boundary(origin, ray, start, end, f1);
        real step = d1 - d0;
        real distance = (d0 + d1) * 0.5;
#ifdef ORACLE_ROBUST
        /* float64 island: |origin + ray d| - R cancels at planet scale (and
         * the float32 origin radius R + h is itself rounded): the stable
         * altitude along the ray from the origin's altitude, in float64 */
        real altitude = altitude_on_ray(origin_altitude, origin_cosine, distance, bottom);
#else
        real altitude = vnorm(vadd(origin, vs(ray, distance))) - bottom;
#endif
#ifdef ORACLE_ROBUST
        v3 y = vsub(vs(extinction(k, altitude), step), carry);
        v3 t = vadd(optical_depth, y);
        carry = vsub(vsub(t, optical_depth), y);
        optical_depth = t;
#else
        optical_depth = vadd(optical_depth, vs(extinction(k, altitude), step));
#endif
    }
    return vexp(vneg(vmins(optical_depth, 80.0)));
}

/* Half-texel bilinear fetch shared by every 2-D LUT sampler
 * (renderer.py:618-640, 655-683, 1503-1525). */
static inline v3 bilinear(const real* lut, int w, int hgt, real u, real v) {
    real x = clampr(u * (real)w - 0.5, 0.0, (real)(w - 1));
    real y = clampr(v * (real)hgt - 0.5, 0.0, (real)(hgt - 1));
    int x0 = (int)floor(x);
    int y0 = (int)floor(y);
    int x1 = x0 + 1 < w - 1 ? x0 + 1 : w - 1;
    int y1 = y0 + 1 < hgt - 1 ? y0 + 1 : hgt - 1;
    real fx = x - (real)x0;
    real fy = y - (real)y0;
    v3 low = vadd(vs(vload(lut, (long)x0 * hgt + y0), 1.0 - fx),
                  vs(vload(lut, (long)x1 * hgt + y0), fx));
    v3 high = vadd(vs(vload(lut, (long)x0 * hgt + y1), 1.0 - fx),
                   vs(vload(lut, (long)x1 * hgt + y1), fx));
    return vadd(vs(low, 1.0 - fy), vs(high, fy));
}

/* renderer.py:609-640 */
static inline v3 sample_transmittance(const K* k, const real* tlut, real h,
                                      real c, real bottom) {
    real u, v;
    transmittance_ray_to_uv(k, fmax(h, 0.0), c, bottom, &u, &v);
    return bilinear(tlut, k->tw, k->th, u, v);
}

/* renderer.py:642-683 */
static inline v3 sample_multi_scattering(const K* k, const real* mslut, real h,
                                         real sun_cosine, real bottom) {
    real altitude_fraction = clampr(h / k->top, 0.0, 1.0);
    real u = sun_cosine_to_u(sun_cosine, fmax(h, 0.0), bottom);
    real v = sqrt(altitude_fraction);
    return bilinear(mslut, k->mw, k->mh, u, v);
}

/* ------------------------------------------------------------------ */
/* Sun visibility, phases, sources (renderer.py:685-899).              */
/* ------------------------------------------------------------------ */

/* renderer.py:685-726 */
static inline void solar_shadow_roots(v3 origin, v3 ray, v3 sun, real bottom,
                                      real* root0, real* root1, int* valid0,
                                      int* valid1) {
    v3 op = vsub(origin, vs(sun, vdot(origin, sun)));
    v3 rp = vsub(ray, vs(sun, vdot(ray, sun)));
    real quadratic = vdot(rp, rp);
    real half_linear = vdot(op, rp);
    real perpendicular_length = vnorm(op);
    real constant = (perpendicular_length - bottom) * (perpendicular_length + bottom);
    *root0 = 0.0;
    *root1 = 0.0;
    *valid0 = 0;
    *valid1 = 0;
    real discriminant = half_linear * half_linear - quadratic * constant;
    if (quadratic > 1.0e-8 && discriminant >= 0.0) {
        real square_root = sqrt(fmax(discriminant, 0.0));
        *root0 = (-half_linear - square_root) / quadratic;
        *root1 = (-half_linear + square_root) / quadratic;
        v3 p0 = vadd(origin, vs(ray, *root0));
        v3 p1 = vadd(origin, vs(ray, *root1));
        if (vdot(p0, sun) < 0.0) *valid0 = 1;
        if (vdot(p1, sun) < 0.0) *valid1 = 1;
    }
}

/* renderer.py:728-754 */
static inline real solar_disk_visibility(const K* k, real h, real sun_cosine,
                                         real bottom, real angular_radius) {
    real radius = bottom + fmax(h, 0.0);
    real horizon_cosine = -shell_rho(h, bottom) / fmax(radius, 1.0);
    real horizon_zenith = cr_acos(clampr(horizon_cosine, -1.0, 1.0));
    real sun_zenith = cr_acos(clampr(sun_cosine, -1.0, 1.0));
    real separation = horizon_zenith - sun_zenith;
    real n = clampr(separation / fmax(angular_radius, 1.0e-6), -1.0, 1.0);
    return (cr_acos(-n) + n * sqrt(fmax(1.0 - n * n, 0.0))) / k->pi;
}

/* renderer.py:756-789 */
static inline v3 sample_solar_transmittance(const K* k, const real* tlut,
                                            real h, real sun_cosine,
                                            real bottom, real angular_radius) {
    real visibility = solar_disk_visibility(k, h, sun_cosine, bottom, angular_radius);
    real radius = bottom + fmax(h, 0.0);
    real horizon_cosine = -shell_rho(h, bottom) / fmax(radius, 1.0);
    real horizon_zenith = cr_acos(clampr(horizon_cosine, -1.0, 1.0));
    real sun_zenith = cr_acos(clampr(sun_cosine, -1.0, 1.0));
    real representative = fmin(sun_zenith, horizon_zenith - angular_radius * 0.25);
    real representative_cosine = cr_cos(fmax(representative, 0.0));
    return vs(sample_transmittance(k, tlut, h, representative_cosine, bottom),
              visibility);
}

/* renderer.py:791-793 */
static inline real rayleigh_phase(const K* k, real c) {
    return 3.0 * (1.0 + c * c) / k->pi16;
}

/* renderer.py:795-804 */
static inline real mie_phase(const K* k, real c) {
    real denominator = fmax(k->mie_a - k->mie_b * c, 1.0e-8);
    return k->mie_c * (1.0 + c * c) / (k->mie_d * cr_pow(denominator, (real)1.5));
}

/* renderer.py:806-816 */
static inline void segment_integral(v3 ext, real step, v3* transmission,
                                    v3* integral) {
    *transmission = vexp(vneg(vmins(vs(ext, step), 80.0)));
    real e[3] = {ext.x, ext.y, ext.z};
    real t[3] = {transmission->x, transmission->y, transmission->z};
    (void)t;
    real out[3];
    for (int c = 0; c < 3; ++c) {
        out[c] = step;
#ifdef ORACLE_ROBUST
        /* 1 - exp(-x) cancels for small x: there -expm1(-x), rounded once.
           For x >= 1/16 it is at least 0.06, so float32's 1 - T carries it
           to about 1e-6 and needs no float64. */
        if (e[c] > 1.0e-12) {
            const real x = fmin(e[c] * step, 80.0);
            out[c] = x < 0.0625 ? (real)(-expm1(-(wide)x)) / e[c] : (1.0 - t[c]) / e[c];
        }
#else
        if (e[c] > 1.0e-12) out[c] = (1.0 - t[c]) / e[c];
#endif
    }
    *integral = V(out[0], out[1], out[2]);
}

/* renderer.py:818-856 */
static inline v3 scattering_source(const K* k, const real* tlut,
                                   const real* mslut, v3 point, real h, v3 ray,
                                   v3 sun, v3 irradiance, real bottom,
                                   real angular_radius, int include_multiple) {
    real radius = vnorm(point);
    v3 radial = vdivs(point, fmax(radius, 1.0));
    v3 d = density(k, h);
    real sun_cosine = vdot(radial, sun);
    v3 sun_transmission =
        sample_solar_transmittance(k, tlut, h, sun_cosine, bottom, angular_radius);
    real scattering_cosine = clampr(vdot(ray, sun), -1.0, 1.0);
    v3 phase_term = vadd(vs(vs(varr(k->beta_r), d.x), rayleigh_phase(k, scattering_cosine)),
                         vs(vs(varr(k->beta_ms), d.y), mie_phase(k, scattering_cosine)));
    v3 source = vmul(vmul(irradiance, sun_transmission), phase_term);
    if (include_multiple) {
        v3 multiple = sample_multi_scattering(k, mslut, h, sun_cosine, bottom);
        source = vadd(source, vmul(scattering(k, h), multiple));
    }
    return source;
}

/* renderer.py:858-899 */
static inline void scattering_segment(const K* k, const real* tlut,
                                      const real* mslut, v3 origin, real h0,
                                      v3 ray, real start, real end, v3 sun,
                                      v3 irradiance, real bottom,
                                      real angular_radius, v3* radiance,
                                      v3* transmission) {
    real step = fmax(end - start, 0.0);
    real distance = (start + end) * 0.5;
    v3 point = vadd(origin, vs(ray, distance));
    real origin_radius = fmax(vnorm(origin), 1.0);
    v3 origin_radial = vdivs(origin, origin_radius);
    real altitude = altitude_on_ray(h0, vdot(origin_radial, ray), distance, bottom);
    v3 ext = extinction(k, altitude);
    v3 integral;
    segment_integral(ext, step, transmission, &integral);
    v3 source = scattering_source(k, tlut, mslut, point, altitude, ray, sun,
                                  irradiance, bottom, angular_radius, 1);
    *radiance = vmul(source, integral);
}

/* renderer.py:901-966 */
static inline real terminator_refinement_weight(
    const K* k, v3 origin, real h0, v3 ray, real start, real end, v3 sun,
    real bottom, real angular_radius, real root0, real root1, int valid0,
    int valid1) {
    real midpoint = (start + end) * 0.5;
    real interval_width = fmax(end - start, 1.0e-4);
    real closest_root = 1.0e30;
    if (valid0 != 0) closest_root = fmin(closest_root, fabs(midpoint - root0));
    if (valid1 != 0) closest_root = fmin(closest_root, fabs(midpoint - root1));
    real proximity = clampr(2.0 - closest_root / interval_width, 0.0, 1.0);
    proximity = proximity * proximity * (3.0 - 2.0 * proximity);
    v3 point = vadd(origin, vs(ray, midpoint));
    real radius = fmax(vnorm(point), 1.0);
    v3 radial = vdivs(point, radius);
    real origin_radius = fmax(vnorm(origin), 1.0);
    v3 origin_radial = vdivs(origin, origin_radius);
    real altitude = altitude_on_ray(h0, vdot(origin_radial, ray), midpoint, bottom);
    real visibility = solar_disk_visibility(k, altitude, vdot(radial, sun), bottom,
                                            angular_radius);
    real penumbra = 4.0 * visibility * (1.0 - visibility);
    return fmax(proximity, penumbra);
}

/* renderer.py:968-1032 */
static inline void adaptive_interval(const K* k, const real* tlut,
                                     const real* mslut, v3 origin, real h0,
                                     v3 ray, real start, real end, v3 sun,
                                     v3 irradiance, real bottom,
                                     real angular_radius, real weight,
                                     v3* radiance, v3* transmission) {
    v3 coarse_l, coarse_t;
    scattering_segment(k, tlut, mslut, origin, h0, ray, start, end, sun,
                       irradiance, bottom, angular_radius, &coarse_l, &coarse_t);
    *radiance = coarse_l;
    *transmission = coarse_t;
    if (weight > 0.0) {
        v3 fine_l = V(0.0, 0.0, 0.0);
        v3 fine_t = V(1.0, 1.0, 1.0);
        for (int sub = 0; sub < k->term_sub; ++sub) {
            real f0 = (real)sub / (real)k->term_sub;
            real f1 = (real)(sub + 1) / (real)k->term_sub;
            real cs = start + (end - start) * f0;
            real ce = start + (end - start) * f1;
            v3 child_l, child_t;
            scattering_segment(k, tlut, mslut, origin, h0, ray, cs, ce, sun,
                               irradiance, bottom, angular_radius, &child_l,
                               &child_t);
            fine_l = vadd(fine_l, vmul(fine_t, child_l));
            fine_t = vmul(fine_t, child_t);
        }
        *radiance = vadd(vs(coarse_l, 1.0 - weight), vs(fine_l, weight));
        *transmission = vadd(vs(coarse_t, 1.0 - weight), vs(fine_t, weight));
    }
}

/* ------------------------------------------------------------------ */
/* Per-pixel surface path (renderer.py:1603-1825).                      */
/* ------------------------------------------------------------------ */

/* renderer.py:1603-1665 */
static inline v3 surface_segment_transmittance(const K* k, const real* tlut,
                                               v3 origin, real h0, v3 ray,
                                               real surface_distance,
                                               real bottom) {
    v3 transmission = V(1.0, 1.0, 1.0);
    real on, of;
    int oh;
    shell_roots_from_altitude(k, h0, ray.y, k->top, bottom, &on, &of, &oh);
    real start = fmax(on, 0.0);
    real end = fmin(surface_distance, of);
    if (oh && end > start) {
        v3 near_point = vadd(origin, vs(ray, start));
        v3 surface_point = vadd(origin, vs(ray, end));
        v3 reverse_ray = vneg(ray);
        real near_length = vnorm(near_point);
        real surface_length = vnorm(surface_point);
        v3 near_radial = vdivs(near_point, fmax(near_length, 1.0));
        v3 surface_radial = vdivs(surface_point, fmax(surface_length, 1.0));
        real near_altitude = altitude_on_ray(h0, ray.y, start, bottom);
        real surface_altitude = altitude_on_ray(h0, ray.y, end, bottom);
        v3 surface_to_top = sample_transmittance(
            k, tlut, surface_altitude, vdot(surface_radial, reverse_ray), bottom);
        v3 near_to_top = sample_transmittance(
            k, tlut, near_altitude, vdot(near_radial, reverse_ray), bottom);
        transmission = vclamp(vdiv(surface_to_top, vmaxs(near_to_top, 1.0e-6)), 0.0, 1.0);
    }
    return transmission;
}

/* renderer.py:1667-1755; returns the refined-interval count via *refined. */
static inline v3 integrate_surface_scattering(const K* k, const real* tlut,
                                              const real* mslut, v3 origin,
                                              real h0, v3 ray,
                                              real surface_distance, v3 sun,
                                              v3 irradiance, real bottom,
                                              real angular_radius,
                                              long long* refined) {
    v3 radiance = V(0.0, 0.0, 0.0);
    v3 view_t = V(1.0, 1.0, 1.0);
    real on, of;
    int oh;
    shell_roots_from_altitude(k, h0, ray.y, k->top, bottom, &on, &of, &oh);
    real start = fmax(on, 0.0);
    real end = fmin(surface_distance, of);
    if (oh && end > start) {
        real r0, r1;
        int v0, v1;
        solar_shadow_roots(origin, ray, sun, bottom, &r0, &r1, &v0, &v1);
        for (int i = 0; i < k->hz_steps; ++i) {
            real f0 = (real)i / (real)k->hz_steps;
            real f1 = (real)(i + 1) / (real)k->hz_steps;
            real d0 = warped_boundary(origin, ray, start, end, f0);
            real d1 = warped_boundary(origin, ray, start, end, f1);
            real w = terminator_refinement_weight(k, origin, h0, ray, d0, d1, sun,
                                                  bottom, angular_radius, r0, r1,
                                                  v0, v1);
            if (refined && w > 0.0) *refined += 1;
            v3 il, it;
            adaptive_interval(k, tlut, mslut, origin, h0, ray, d0, d1, sun,
                              irradiance, bottom, angular_radius, w, &il, &it);
            radiance = vadd(radiance, vmul(view_t, il));
            view_t = vmul(view_t, it);
        }
    }
    return vmaxs(radiance, 0.0);
}

/* renderer.py:1757-1786 */
static inline v3 surface_aerial_scattering(const K* k, const real* tlut,
                                           const real* mslut, real distance,
                                           real h, v3 ray, v3 sun,
                                           v3 irradiance, real bottom,
                                           real angular_radius,
                                           long long* refined) {
    v3 origin = V(0.0, bottom + h, 0.0);
    return integrate_surface_scattering(k, tlut, mslut, origin, h, ray, distance,
                                        sun, irradiance, bottom, angular_radius,
                                        refined);
}

/* renderer.py:1788-1825 */
static inline v3 camera_transmittance(const K* k, const real* tlut, real h,
                                      v3 ray, real bottom) {
    real start, end;
    int valid, ground;
    atmosphere_segment_from_altitude(k, h, ray.y, bottom, &start, &end, &valid,
                                     &ground);
    v3 transmission = V(1.0, 1.0, 1.0);
    if (valid) {
        if (ground) {
            transmission = V(0.0, 0.0, 0.0);
        } else {
            v3 origin = V(0.0, bottom + h, 0.0);
            v3 entry = vadd(origin, vs(ray, start));
            real radius = vnorm(entry);
            v3 radial = vdivs(entry, fmax(radius, 1.0));
            real entry_altitude = altitude_on_ray(h, ray.y, start, bottom);
            transmission = sample_transmittance(k, tlut, entry_altitude,
                                                vdot(radial, ray), bottom);
        }
    }
    return transmission;
}

/* renderer.py:1474-1525 */
static inline v3 sample_sky_view(const K* k, const real* svlut, v3 ray, v3 sun,
                                 real h, real bottom) {
#ifndef ORACLE_ROBUST
    real zenith = cr_acos(clampr(ray.y, -1.0, 1.0));
#endif
    real vhx = ray.x, vhz = ray.z, shx = sun.x, shz = sun.z;
    real view_length = sqrt(vhx * vhx + vhz * vhz);
    real sun_length = sqrt(shx * shx + shz * shz);
    real relative_azimuth = 0.0;
    if (view_length > 1.0e-6 && sun_length > 1.0e-6)
        relative_azimuth = cr_acos(clampr((vhx * shx + vhz * shz)
                                           / (view_length * sun_length),
                                       -1.0, 1.0));
    real u = relative_azimuth / k->pi;
#ifdef ORACLE_ROBUST
    real v = sky_zenith_to_v_w(k, acos(fmin(fmax((wide)ray.y, -1.0), 1.0)), h, bottom);
#else
    real v = sky_zenith_to_v(k, zenith, h, bottom);
#endif
    return bilinear(svlut, k->sw, k->sh, u, v);
}

/* renderer.py:1527-1571 */
static inline v3 sample_aerial_field(const K* k, const real* ap, real su,
                                     real sv, real depth_fraction) {
    int aw = k->aw, ah = k->ah, ad = k->ad;
    real x = clampr(su * (real)aw - 0.5, 0.0, (real)(aw - 1));
    real y = clampr(sv * (real)ah - 0.5, 0.0, (real)(ah - 1));
    real z = clampr(depth_fraction, 0.0, 1.0) * (real)(ad - 1);
    int x0 = (int)floor(x), y0 = (int)floor(y), z0 = (int)floor(z);
    int x1 = x0 + 1 < aw - 1 ? x0 + 1 : aw - 1;
    int y1 = y0 + 1 < ah - 1 ? y0 + 1 : ah - 1;
    int z1 = z0 + 1 < ad - 1 ? z0 + 1 : ad - 1;
    real fx = x - (real)x0, fy = y - (real)y0, fz = z - (real)z0;
#define APIDX(X, Y, Z) (((long)(X) * ah + (Y)) * ad + (Z))
    v3 zlyl = vadd(vs(vload(ap, APIDX(x0, y0, z0)), 1.0 - fx), vs(vload(ap, APIDX(x1, y0, z0)), fx));
    v3 zlyh = vadd(vs(vload(ap, APIDX(x0, y1, z0)), 1.0 - fx), vs(vload(ap, APIDX(x1, y1, z0)), fx));
    v3 zhyl = vadd(vs(vload(ap, APIDX(x0, y0, z1)), 1.0 - fx), vs(vload(ap, APIDX(x1, y0, z1)), fx));
    v3 zhyh = vadd(vs(vload(ap, APIDX(x0, y1, z1)), 1.0 - fx), vs(vload(ap, APIDX(x1, y1, z1)), fx));
#undef APIDX
    v3 low = vadd(vs(zlyl, 1.0 - fy), vs(zlyh, fy));
    v3 high = vadd(vs(zhyl, 1.0 - fy), vs(zhyh, fy));
    return vadd(vs(low, 1.0 - fz), vs(high, fz));
}

/* renderer.py:1573-1601 */
static inline v3 sample_aerial_scattering(const K* k, const real* ap, real su,
                                          real sv, real distance, real h,
                                          v3 ray, real bottom) {
    real start, end;
    int valid;
    aerial_prefix_segment(k, h, ray.y, bottom, &start, &end, &valid);
    real depth_fraction = 0.0;
    if (valid)
        depth_fraction = clampr((distance - start) / fmax(end - start, 1.0), 0.0, 1.0);
    return sample_aerial_field(k, ap, su, sv, depth_fraction);
}

/* ------------------------------------------------------------------ */
/* Exported kernels.                                                    */
/* ------------------------------------------------------------------ */

EXPORT void oracle_set_threads(int threads) {
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
}

EXPORT int oracle_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

EXPORT int oracle_real_bytes(void) { return (int)sizeof(real); }

/* renderer.py:1034-1063 */
EXPORT int oracle_build_transmittance(const AtmoParams* p, real bottom,
                                      real* tlut) {
    K k = make_consts(p);
    long n = (long)k.tw * k.th;
#pragma omp parallel for schedule(dynamic, 64)
    for (long i = 0; i < n; ++i) {
        int x = (int)(i / k.th), y = (int)(i % k.th);
        real u = ((real)x + 0.5) / (real)k.tw;
        real v = ((real)y + 0.5) / (real)k.th;
        real altitude, cosine, distance;
        transmittance_uv_to_ray(&k, u, v, bottom, &altitude, &cosine, &distance);
        real sine = sqrt(fmax(1.0 - cosine * cosine, 0.0));
        real radius = bottom + altitude;
        v3 origin = V(0.0, radius, 0.0);
        v3 ray = V(sine, cosine, 0.0);
        v3 t = integrate_extinction(&k, origin, ray, 0.0, distance, bottom, altitude, cosine);
        vstore(tlut, i, vclamp(t, 0.0, 1.0));
    }
    return ATMO_OK;
}

/* renderer.py:1065-1206 */
EXPORT int oracle_build_multi_scattering(const AtmoParams* p, real bottom,
                                         const real irradiance_in[3],
                                         real angular_radius, const real* tlut,
                                         real* mslut) {
    K k = make_consts(p);
    v3 irradiance = varr(irradiance_in);
    const int direction_count = k.ms_dirs;
    const int path_steps = k.ms_steps;
    long n = (long)k.mw * k.mh;
#pragma omp parallel for schedule(dynamic, 4)
    for (long i = 0; i < n; ++i) {
        int x = (int)(i / k.mh), y = (int)(i % k.mh);
        real u = ((real)x + 0.5) / (real)k.mw;
        real v = ((real)y + 0.5) / (real)k.mh;
        real origin_altitude = v * v * k.top;
        real radius = bottom + origin_altitude;
        real sun_cosine = sun_u_to_cosine(u, origin_altitude, bottom);
        real sun_sine = sqrt(fmax(1.0 - sun_cosine * sun_cosine, 0.0));
        v3 sun = V(sun_sine, sun_cosine, 0.0);
        v3 origin = V(0.0, radius, 0.0);
        v3 luminance_sum = V(0.0, 0.0, 0.0);
        v3 feedback_sum = V(0.0, 0.0, 0.0);
        for (int di = 0; di < direction_count; ++di) {
            real sample_index = (real)di + 0.5;
            real dc = 1.0 - 2.0 * sample_index / (real)direction_count;
            real ds = sqrt(fmax(1.0 - dc * dc, 0.0));
            real azimuth = (real)di * 2.39996323;
            v3 ray = V(ds * cr_cos(azimuth), dc, ds * cr_sin(azimuth));
            real start, end;
            int valid, ground;
            atmosphere_segment(&k, origin, ray, bottom, &start, &end, &valid, &ground);
            v3 dir_l = V(0.0, 0.0, 0.0);
            v3 dir_f = V(0.0, 0.0, 0.0);
            v3 path_t = V(1.0, 1.0, 1.0);
            if (valid) {
                for (int pi = 0; pi < path_steps; ++pi) {
                    real f0 = (real)pi / (real)path_steps;
                    real f1 = (real)(pi + 1) / (real)path_steps;
                    real d0 = warped_boundary(origin, ray, start, end, f0);
                    real d1 = warped_boundary(origin, ray, start, end, f1);
                    real distance = (d0 + d1) * 0.5;
                    real step = d1 - d0;
                    v3 point = vadd(origin, vs(ray, distance));
                    real altitude = altitude_on_ray(origin_altitude, dc, distance, bottom);
                    v3 ext = extinction(&k, altitude);
                    v3 seg_t, seg_i;
                    segment_integral(ext, step, &seg_t, &seg_i);
                    v3 direct = scattering_source(&k, tlut, NULL, point, altitude, ray,
                                                  sun, irradiance, bottom,
                                                  angular_radius, 0);
                    dir_l = vadd(dir_l, vmul(vmul(path_t, direct), seg_i));
                    dir_f = vadd(dir_f, vmul(vmul(path_t, scattering(&k, altitude)), seg_i));
                    path_t = vmul(path_t, seg_t);
                }
                if (ground) {
                    v3 ground_point = vadd(origin, vs(ray, end));
                    v3 normal = vnormalized(ground_point);
                    real gsc = vdot(normal, sun);
                    v3 gt = sample_solar_transmittance(&k, tlut, 0.0, gsc, bottom,
                                                       angular_radius);
                    v3 ground_l = vdivs(vs(vmul(vmul(varr(k.albedo), irradiance), gt),
                                           fmax(gsc, 0.0)),
                                        k.pi);
                    dir_l = vadd(dir_l, vmul(path_t, ground_l));
                }
            }
            luminance_sum = vadd(luminance_sum, dir_l);
            feedback_sum = vadd(feedback_sum, dir_f);
        }
        v3 avg_l = vdivs(luminance_sum, (real)direction_count);
        v3 avg_f = vclamp(vdivs(feedback_sum, (real)direction_count), 0.0, 0.95);
        vstore(mslut, i, vmaxs(vdiv(avg_l, vsub(V(1.0, 1.0, 1.0), avg_f)), 0.0));
    }
    return ATMO_OK;
}

/* renderer.py:1208-1302 */
EXPORT int oracle_build_sky_view(const AtmoParams* p, real bottom,
                                 real camera_altitude, real sun_zenith_cosine,
                                 const real irradiance_in[3],
                                 real angular_radius, const real* tlut,
                                 const real* mslut, real* svlut) {
    K k = make_consts(p);
    v3 irradiance = varr(irradiance_in);
    real sun_sine = sqrt(fmax(1.0 - sun_zenith_cosine * sun_zenith_cosine, 0.0));
    v3 sun = V(sun_sine, sun_zenith_cosine, 0.0);
    real safe_alt = fmax(camera_altitude, 0.0);
    real safe_radius = bottom + safe_alt;
    v3 origin = V(0.0, safe_radius, 0.0);
    long n = (long)k.sw * k.sh;
#pragma omp parallel for schedule(dynamic, 32)
    for (long i = 0; i < n; ++i) {
        int x = (int)(i / k.sh), y = (int)(i % k.sh);
        real azimuth = ((real)x + 0.5) / (real)k.sw * k.pi;
        real view_v = ((real)y + 0.5) / (real)k.sh;
#ifdef ORACLE_ROBUST
        wide zenith_w = sky_v_to_zenith_w(&k, view_v, safe_alt, bottom);
        real vc = (real)cos(zenith_w);
        real vsn = (real)sin(zenith_w);
#else
        real zenith = sky_v_to_zenith(&k, view_v, safe_alt, bottom);
        real vc = cr_cos(zenith);
        real vsn = cr_sin(zenith);
#endif
        v3 ray = V(vsn * cr_cos(azimuth), vc, vsn * cr_sin(azimuth));
        real start, end;
        int valid, ground;
        atmosphere_segment_from_altitude(&k, safe_alt, vc, bottom, &start, &end,
                                         &valid, &ground);
        v3 radiance = V(0.0, 0.0, 0.0);
        if (valid) {
            v3 view_t = V(1.0, 1.0, 1.0);
            for (int s = 0; s < k.sv_steps; ++s) {
                real f0 = (real)s / (real)k.sv_steps;
                real f1 = (real)(s + 1) / (real)k.sv_steps;
                real d0 = warped_boundary(origin, ray, start, end, f0);
                real d1 = warped_boundary(origin, ray, start, end, f1);
                real distance = (d0 + d1) * 0.5;
                real step = d1 - d0;
                v3 point = vadd(origin, vs(ray, distance));
                real altitude = altitude_on_ray(safe_alt, vc, distance, bottom);
                v3 source = scattering_source(&k, tlut, mslut, point, altitude, ray,
                                              sun, irradiance, bottom,
                                              angular_radius, 1);
                v3 ext = extinction(&k, altitude);
                v3 seg_t, seg_i;
                segment_integral(ext, step, &seg_t, &seg_i);
                radiance = vadd(radiance, vmul(vmul(view_t, source), seg_i));
                view_t = vmul(view_t, seg_t);
            }
        }
        vstore(svlut, i, vmaxs(radiance, 0.0));
    }
    return ATMO_OK;
}

/* renderer.py:1304-1420 */
EXPORT int oracle_build_aerial_perspective(
    const AtmoParams* p, int width, int height, real bottom,
    real camera_altitude, const real sun_in[3], const real irradiance_in[3],
    real angular_radius, const real right_in[3], const real up_in[3],
    const real forward_in[3], real tangent_half_fov, const real* tlut,
    const real* mslut, real* aplut) {
    K k = make_consts(p);
    v3 sun = varr(sun_in), irradiance = varr(irradiance_in);
    v3 right = varr(right_in), up = varr(up_in), forward = varr(forward_in);
    real aspect = (real)width / (real)height;
    v3 origin = V(0.0, bottom + camera_altitude, 0.0);
    long n = (long)k.aw * k.ah;
#pragma omp parallel for schedule(dynamic, 4)
    for (long i = 0; i < n; ++i) {
        int x = (int)(i / k.ah), y = (int)(i % k.ah);
        real u = ((real)x + 0.5) / (real)k.aw;
        real v = ((real)y + 0.5) / (real)k.ah;
        real sx = (u * 2.0 - 1.0) * aspect;
        real sy = v * 2.0 - 1.0;
        v3 ray = vnormalized(vadd(vadd(vs(right, sx), vs(up, sy)),
                                  vdivs(forward, tangent_half_fov)));
        real a_start, a_end;
        int valid;
        aerial_prefix_segment(&k, camera_altitude, ray.y, bottom, &a_start, &a_end, &valid);
        v3 cum_l = V(0.0, 0.0, 0.0);
        v3 cum_t = V(1.0, 1.0, 1.0);
        real previous = a_start;
        for (int di = 0; di < k.ad; ++di) {
            real depth_fraction = (real)di / (real)(k.ad - 1);
            real target = a_start + (a_end - a_start) * depth_fraction;
            if (valid) {
                real is = fmax(previous, a_start);
                real ie = fmin(target, a_end);
                if (ie > is) {
                    for (int s = 0; s < k.ap_spp; ++s) {
                        real f0 = (real)s / (real)k.ap_spp;
                        real f1 = (real)(s + 1) / (real)k.ap_spp;
                        real d0 = warped_boundary(origin, ray, is, ie, f0);
                        real d1 = warped_boundary(origin, ray, is, ie, f1);
                        real distance = (d0 + d1) * 0.5;
                        real step = d1 - d0;
                        v3 point = vadd(origin, vs(ray, distance));
                        real altitude = altitude_on_ray(camera_altitude, ray.y,
                                                        distance, bottom);
                        v3 ext = extinction(&k, altitude);
                        v3 seg_t, seg_i;
                        segment_integral(ext, step, &seg_t, &seg_i);
                        v3 source = scattering_source(&k, tlut, mslut, point, altitude,
                                                      ray, sun, irradiance, bottom,
                                                      angular_radius, 1);
                        cum_l = vadd(cum_l, vmul(vmul(cum_t, source), seg_i));
                        cum_t = vmul(cum_t, seg_t);
                    }
                }
            }
            vstore(aplut, i * k.ad + di, vmaxs(cum_l, 0.0));
            previous = target;
        }
    }
    return ATMO_OK;
}

/* renderer.py:1422-1472 */
EXPORT int oracle_prepare_surface_lighting(
    const AtmoParams* p, int width, int height, const real* position_view,
    const int32_t* surface_id, real bottom, real camera_altitude,
    const real sun_in[3], real angular_radius, const real right_in[3],
    const real up_in[3], const real forward_in[3], const real* tlut,
    const real* mslut, real* sun_t_out, real* sky_l_out) {
    K k = make_consts(p);
    v3 sun = varr(sun_in);
    v3 right = varr(right_in), up = varr(up_in), forward = varr(forward_in);
    v3 camera = V(0.0, bottom + camera_altitude, 0.0);
    long n = (long)width * height;
#pragma omp parallel for schedule(dynamic, 256)
    for (long i = 0; i < n; ++i) {
        v3 solar = V(1.0, 1.0, 1.0);
        v3 sky = V(0.0, 0.0, 0.0);
        if (surface_id[i] >= 0) {
            v3 vp = vload(position_view, i);
            v3 local = vadd(vadd(vs(right, vp.x), vs(up, vp.y)), vs(forward, vp.z));
            v3 planet = vadd(camera, local);
            real radius = vnorm(planet);
            v3 radial = vdivs(planet, fmax(radius, 1.0));
            real surface_distance = vnorm(local);
            v3 local_ray = vdivs(local, fmax(surface_distance, 1.0));
            real sun_cosine = vdot(radial, sun);
#ifdef ORACLE_ROBUST
            if (camera_altitude > 1.0e6) {
                /* float64 island: camera-relative endpoint above 1000 km */
                wide lx = ((wide)right.x * vp.x + (wide)up.x * vp.y) + (wide)forward.x * vp.z;
                wide ly = ((wide)right.y * vp.x + (wide)up.y * vp.y) + (wide)forward.y * vp.z;
                wide lz = ((wide)right.z * vp.x + (wide)up.z * vp.y) + (wide)forward.z * vp.z;
                wide py = ((wide)bottom + camera_altitude) + ly;
                wide pn = fmax(sqrt((lx * lx + py * py) + lz * lz), 1.0);
                sun_cosine = (real)(((lx * sun.x + py * sun.y) + lz * sun.z) / pn);
                wide ld = sqrt((lx * lx + ly * ly) + lz * lz);
                surface_distance = (real)ld;
                local_ray.y = (real)(ly / fmax(ld, 1.0));
            }
#endif
            real altitude = altitude_on_ray(camera_altitude, local_ray.y,
                                            surface_distance, bottom);
            solar = sample_solar_transmittance(&k, tlut, altitude, sun_cosine, bottom,
                                               angular_radius);
            sky = sample_multi_scattering(&k, mslut, altitude, sun_cosine, bottom);
        }
        vstore(sun_t_out, i, solar);
        vstore(sky_l_out, i, vmaxs(sky, 0.0));
    }
    return ATMO_OK;
}

/* renderer.py:1827-2055.  `refined_out` (may be NULL) receives the number
 * of base intervals whose terminator weight was > 0 (the cost-model N_ref). */
EXPORT int oracle_composite(
    const AtmoParams* p, int width, int height, const real* scene_hdr,
    const real* space_hdr, const int32_t* surface_id,
    const real* position_view, real* hdr, real bottom, real camera_altitude,
    real sky_alt, const real sun_in[3], const real irradiance_in[3],
    const real disk_in[3], real angular_radius, const real right_in[3],
    const real up_in[3], const real forward_in[3], real tangent_half_fov,
    real contrast_start, real contrast_end, int view, const real* tlut,
    const real* mslut, const real* svlut, const real* aplut,
    long long* refined_out) {
    K k = make_consts(p);
    v3 sun = varr(sun_in), irradiance = varr(irradiance_in), disk = varr(disk_in);
    v3 right = varr(right_in), up = varr(up_in), forward = varr(forward_in);
    real aspect = (real)width / (real)height;
    long n = (long)width * height;
    long long refined_total = 0;
    v3 lum = V(0.2126, 0.7152, 0.0722);
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : refined_total)
    for (long i = 0; i < n; ++i) {
        int px = (int)(i / height), py = (int)(i % height);
        long long refined = 0;
        v3 color = vload(scene_hdr, i);
        real su = ((real)px + 0.5) / (real)width;
        real sv = ((real)py + 0.5) / (real)height;
        real sx = (su * 2.0 - 1.0) * aspect;
        real sy = sv * 2.0 - 1.0;
        v3 ray = vnormalized(vadd(vadd(vs(right, sx), vs(up, sy)),
                                  vdivs(forward, tangent_half_fov)));
        int surface = surface_id[i] >= 0;
        if (view == ATMO_VIEW_COMPOSITE) {
            if (surface) {
                real sd = vnorm(vload(position_view, i));
                v3 scat = surface_aerial_scattering(&k, tlut, mslut, sd, camera_altitude,
                                                    ray, sun, irradiance, bottom,
                                                    angular_radius, &refined);
                v3 trans = surface_segment_transmittance(
                    &k, tlut, V(0.0, bottom + camera_altitude, 0.0), camera_altitude,
                    ray, sd, bottom);
                color = vadd(scat, vmul(color, trans));
            } else {
                v3 trans = camera_transmittance(&k, tlut, camera_altitude, ray, bottom);
                v3 sky = sample_sky_view(&k, svlut, ray, sun, sky_alt, bottom);
                v3 star = vmul(vload(space_hdr, i), trans);
                real sky_lum = vdot(sky, lum);
                real star_lum = vdot(star, lum);
                real contrast = star_lum / fmax(sky_lum, 1.0e-5);
                real vis = clampr((contrast - contrast_start)
                                      / fmax(contrast_end - contrast_start, 1.0e-5),
                                  0.0, 1.0);
                vis = vis * vis * (3.0 - 2.0 * vis);
                color = vadd(vadd(sky, vmul(color, trans)), vs(star, vis));
                real angular_distance = cr_acos(clampr(vdot(ray, sun), -1.0, 1.0));
                real pixel_filter =
                    cr_atan2(tangent_half_fov * 1.41421356 / (real)height, (real)1.0);
                real coverage = clampr((angular_radius + pixel_filter - angular_distance)
                                           / fmax(2.0 * pixel_filter, 1.0e-8),
                                       0.0, 1.0);
                coverage = coverage * coverage * (3.0 - 2.0 * coverage);
                color = vadd(color, vs(vmul(disk, trans), coverage));
            }
        } else if (view == ATMO_VIEW_SKY_VIEW) {
            color = sample_sky_view(&k, svlut, ray, sun, sky_alt, bottom);
        } else if (view == ATMO_VIEW_CAMERA_TRANSMITTANCE) {
            color = camera_transmittance(&k, tlut, camera_altitude, ray, bottom);
        } else if (view == ATMO_VIEW_TRANSMITTANCE_LUT) {
            int lx = (int)(su * (real)k.tw);
            int ly = (int)(sv * (real)k.th);
            if (lx > k.tw - 1) lx = k.tw - 1;
            if (ly > k.th - 1) ly = k.th - 1;
            color = vload(tlut, (long)lx * k.th + ly);
        } else if (view == ATMO_VIEW_MULTI_SCATTERING_LUT) {
            int lx = (int)(su * (real)k.mw);
            int ly = (int)(sv * (real)k.mh);
            if (lx > k.mw - 1) lx = k.mw - 1;
            if (ly > k.mh - 1) ly = k.mh - 1;
            color = vload(mslut, (long)lx * k.mh + ly);
        } else if (view == ATMO_VIEW_AERIAL_SCATTERING) {
            color = V(0.0, 0.0, 0.0);
            if (surface)
                color = surface_aerial_scattering(&k, tlut, mslut,
                                                  vnorm(vload(position_view, i)),
                                                  camera_altitude, ray, sun, irradiance,
                                                  bottom, angular_radius, NULL);
        } else if (view == ATMO_VIEW_AERIAL_FROXEL_ONLY) {
            color = V(0.0, 0.0, 0.0);
            if (surface)
                color = sample_aerial_scattering(&k, aplut, su, sv,
                                                 vnorm(vload(position_view, i)),
                                                 camera_altitude, ray, bottom);
        } else if (view == ATMO_VIEW_AERIAL_DIRECT_ONLY) {
            color = V(0.0, 0.0, 0.0);
            if (surface) {
                v3 origin = V(0.0, bottom + camera_altitude, 0.0);
                color = integrate_surface_scattering(&k, tlut, mslut, origin,
                                                     camera_altitude, ray,
                                                     vnorm(vload(position_view, i)),
                                                     sun, irradiance, bottom,
                                                     angular_radius, NULL);
            }
        } else if (view == ATMO_VIEW_AERIAL_BLEND_WEIGHT) {
            real w = surface ? 1.0 : 0.0;
            color = V(w, w, w);
        } else if (view == ATMO_VIEW_AERIAL_TRANSMITTANCE) {
            color = V(0.0, 0.0, 0.0);
            if (surface)
                color = surface_segment_transmittance(
                    &k, tlut, V(0.0, bottom + camera_altitude, 0.0), camera_altitude,
                    ray, vnorm(vload(position_view, i)), bottom);
        } else {
            real m = surface ? 1.0 : 0.0;
            color = V(m, m, m);
        }
        vstore(hdr, i, vmaxs(color, 0.0));
        refined_total += refined;
    }
    if (refined_out) *refined_out = refined_total;
    return ATMO_OK;
}

/* Per-sample surface path: the pair (in-scatter A40, transmittance A39) of
 * the COMPOSITE branch for arbitrary (altitude, ray, distance, sun), as used
 * to label and judge the surrogate.  ray and sun are (n,3) arrays in the
 * camera's local frame; in-scatter and transmittance are (n,3) outputs.
 * renderer.py:1603-1786, 1863-1880. */
EXPORT int oracle_surface_path(const AtmoParams* p, long n, const real* altitude,
                               const real* ray_in, const real* distance,
                               const real* sun_in, const real irradiance_in[3],
                               real bottom, real angular_radius,
                               const real* tlut, const real* mslut,
                               real* inscatter_out, real* transmittance_out) {
    K k = make_consts(p);
    v3 irradiance = varr(irradiance_in);
#pragma omp parallel for schedule(dynamic, 64)
    for (long i = 0; i < n; ++i) {
        v3 ray = vload(ray_in, i);
        v3 sun = vload(sun_in, i);
        real h = altitude[i];
        v3 l = surface_aerial_scattering(&k, tlut, mslut, distance[i], h, ray, sun,
                                         irradiance, bottom, angular_radius, NULL);
        v3 t = surface_segment_transmittance(&k, tlut, V(0.0, bottom + h, 0.0), h,
                                             ray, distance[i], bottom);
        vstore(inscatter_out, i, l);
        vstore(transmittance_out, i, t);
    }
    return ATMO_OK;
}

/* ------------------------------------------------------------------ */
/* Scalar helpers exported for the oracle's own closed-form tests.      */
/* ------------------------------------------------------------------ */

EXPORT void oracle_fn_density(const AtmoParams* p, real h, real out[3]) {
    K k = make_consts(p);
    v3 d = density(&k, h);
    out[0] = d.x; out[1] = d.y; out[2] = d.z;
}

EXPORT void oracle_fn_sphere_roots(const real origin[3], const real ray[3],
                                   real radius, real out[3]) {
    AtmoParams dummy;
    memset(&dummy, 0, sizeof dummy);
    K k = make_consts(&dummy);
    int hit;
    sphere_roots(&k, varr(origin), varr(ray), radius, &out[0], &out[1], &hit);
    out[2] = (real)hit;
}

EXPORT void oracle_fn_shell_roots(real h0, real c, real shell, real bottom,
                                  real out[3]) {
    AtmoParams dummy;
    memset(&dummy, 0, sizeof dummy);
    K k = make_consts(&dummy);
    int hit;
    shell_roots_from_altitude(&k, h0, c, shell, bottom, &out[0], &out[1], &hit);
    out[2] = (real)hit;
}

EXPORT void oracle_fn_atmosphere_segment(const AtmoParams* p, real h, real c,
                                         real bottom, real out[4]) {
    K k = make_consts(p);
    int valid, ground;
    atmosphere_segment_from_altitude(&k, h, c, bottom, &out[0], &out[1], &valid, &ground);
    out[2] = (real)valid;
    out[3] = (real)ground;
}

EXPORT void oracle_fn_atmosphere_segment_vec(const AtmoParams* p,
                                             const real origin[3],
                                             const real ray[3], real bottom,
                                             real out[4]) {
    K k = make_consts(p);
    int valid, ground;
    atmosphere_segment(&k, varr(origin), varr(ray), bottom, &out[0], &out[1], &valid, &ground);
    out[2] = (real)valid;
    out[3] = (real)ground;
}

EXPORT real oracle_fn_altitude_on_ray(real h0, real c, real t, real bottom) {
    return altitude_on_ray(h0, c, t, bottom);
}

EXPORT void oracle_fn_transmittance_uv_to_ray(const AtmoParams* p, real u, real v,
                                              real bottom, real out[3]) {
    K k = make_consts(p);
    transmittance_uv_to_ray(&k, u, v, bottom, &out[0], &out[1], &out[2]);
}

EXPORT void oracle_fn_transmittance_ray_to_uv(const AtmoParams* p, real h, real c,
                                              real bottom, real out[2]) {
    K k = make_consts(p);
    transmittance_ray_to_uv(&k, h, c, bottom, &out[0], &out[1]);
}

EXPORT real oracle_fn_sun_u_to_cosine(real u, real h, real bottom) {
    return sun_u_to_cosine(u, h, bottom);
}

EXPORT real oracle_fn_sun_cosine_to_u(real c, real h, real bottom) {
    return sun_cosine_to_u(c, h, bottom);
}

EXPORT real oracle_fn_sky_v_to_zenith(const AtmoParams* p, real v, real h,
                                      real bottom) {
    K k = make_consts(p);
    return sky_v_to_zenith(&k, v, h, bottom);
}

EXPORT real oracle_fn_sky_zenith_to_v(const AtmoParams* p, real z, real h,
                                      real bottom) {
    K k = make_consts(p);
    return sky_zenith_to_v(&k, z, h, bottom);
}

EXPORT real oracle_fn_warped_boundary(const real origin[3], const real ray[3],
                                      real start, real end, real fraction) {
    return warped_boundary(varr(origin), varr(ray), start, end, fraction);
}

EXPORT void oracle_fn_integrate_extinction(const AtmoParams* p,
                                           const real origin[3],
                                           const real ray[3], real start,
                                           real end, real bottom, real out[3]) {
    K k = make_consts(p);
    v3 o = varr(origin), r = varr(ray);
    wide ox = o.x, oy = o.y, oz = o.z;
    wide radius = sqrt((ox * ox + oy * oy) + oz * oz);
    real altitude = (real)(radius - (wide)bottom);
    real cosine = (real)(((ox * r.x + oy * r.y) + oz * r.z) / fmax(radius, 1.0));
    v3 t = integrate_extinction(&k, o, r, start, end, bottom, altitude, cosine);
    out[0] = t.x; out[1] = t.y; out[2] = t.z;
}

EXPORT void oracle_fn_phases(const AtmoParams* p, real c, real out[2]) {
    K k = make_consts(p);
    out[0] = rayleigh_phase(&k, c);
    out[1] = mie_phase(&k, c);
}

EXPORT void oracle_fn_segment_integral(const real ext[3], real step,
                                       real out_t[3], real out_i[3]) {
    v3 t, i;
    segment_integral(varr(ext), step, &t, &i);
    out_t[0] = t.x; out_t[1] = t.y; out_t[2] = t.z;
    out_i[0] = i.x; out_i[1] = i.y; out_i[2] = i.z;
}

EXPORT real oracle_fn_solar_disk_visibility(const AtmoParams* p, real h,
                                            real sun_cosine, real bottom,
                                            real angular_radius) {
    K k = make_consts(p);
    return solar_disk_visibility(&k, h, sun_cosine, bottom, angular_radius);
}

EXPORT void oracle_fn_solar_shadow_roots(const real origin[3], const real ray[3],
                                         const real sun[3], real bottom,
                                         real out[4]) {
    int v0, v1;
    solar_shadow_roots(varr(origin), varr(ray), varr(sun), bottom, &out[0], &out[1],
                       &v0, &v1);
    out[2] = (real)v0;
    out[3] = (real)v1;
}

/* renderer.py:1788-1825 for one ray (camera at altitude h on the local
 * vertical axis). */
EXPORT void oracle_fn_camera_transmittance(const AtmoParams* p, const real* tlut, real h,
                                           const real ray[3], real bottom, real out[3]) {
    K k = make_consts(p);
    v3 t = camera_transmittance(&k, tlut, h, varr(ray), bottom);
    out[0] = t.x; out[1] = t.y; out[2] = t.z;
}

/* renderer.py:901-966 for one base interval [start, end] with the given
 * shadow roots and their validity flags. */
EXPORT real oracle_fn_terminator_weight(const AtmoParams* p, const real origin[3], real h0,
                                        const real ray[3], real start, real end,
                                        const real sun[3], real bottom, real angular_radius,
                                        const real roots[2], const int32_t valid[2]) {
    K k = make_consts(p);
    return terminator_refinement_weight(&k, varr(origin), h0, varr(ray), start, end,
                                        varr(sun), bottom, angular_radius, roots[0],
                                        roots[1], valid[0], valid[1]);
}
