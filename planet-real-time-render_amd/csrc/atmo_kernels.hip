// atmo_kernels.hip -- HIP kernels for gfx950 and the C-ABI of libatmo.so.
//
// One kernel per Taichi kernel of the reference's AtmosphereRenderer
// (src/planet_renderer/atmosphere/renderer.py:1034-2055).  Launch shapes:
//  - transmittance / sky-view: a group of G lanes per texel; each lane
//    marches a contiguous chunk of the texel's steps and the group combines
//    the chunks (a sum of optical depth for T, an ordered (L, T) scan for
//    sky-view).  A single-state LUT has only 16-20 K texels, so one thread
//    per texel would leave most of the 256 CUs idle and expose the full
//    serial march latency.
//  - multi-scattering: one wave64 per texel, lane = sphere direction, the
//    64-direction integral reduced with cross-lane shuffles.
//  - aerial perspective: one wave64 per froxel ray, lane = depth slice; the
//    cumulative in-scatter along the ray is an ordered (L, T) scan.
//  - prepare / composite: one thread per pixel in the (W, H) field order,
//    so consecutive lanes read and write consecutive 12-byte elements.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <new>

#include "../../include/atmo.h"
#include "atmo_device.h"
#include "atmo_handle.h"

using namespace atmo;

namespace {

thread_local char g_create_error[512];

void set_error(AtmoHandle* h, const char* fmt, const char* a = "", const char* b = "") {
    char* dst = h ? h->error : g_create_error;
    snprintf(dst, 512, fmt, a, b);
}


// AtmosphereConfig.__post_init__ (config.py:61-128), same messages.
int32_t validate(const AtmoParams* p, char* msg, size_t cap) {
#define FAIL(...)                          \
    do {                                   \
        if (msg) snprintf(msg, cap, __VA_ARGS__); \
        return ATMO_EINVAL;                \
    } while (0)
    struct {
        const char* name;
        const double* v;
    } groups[] = {
        {"rayleigh_scattering_per_m", p->rayleigh_scattering_per_m},
        {"mie_scattering_per_m", p->mie_scattering_per_m},
        {"mie_extinction_per_m", p->mie_extinction_per_m},
        {"absorption_extinction_per_m", p->absorption_extinction_per_m},
        {"ground_albedo", p->ground_albedo},
    };
    for (auto& g : groups)
        for (int c = 0; c < 3; ++c)
            if (!(isfinite(g.v[c]) && g.v[c] >= 0.0))
                FAIL("atmosphere.%s must contain three non-negative values", g.name);
    for (int c = 0; c < 3; ++c)
        if (p->ground_albedo[c] > 1.0) FAIL("atmosphere.ground_albedo must remain in [0, 1]");
    struct {
        const char* name;
        double v;
    } positive[] = {
        {"top_altitude_m", p->top_altitude_m},
        {"rayleigh_scale_height_m", p->rayleigh_scale_height_m},
        {"mie_scale_height_m", p->mie_scale_height_m},
        {"absorption_half_width_m", p->absorption_half_width_m},
    };
    for (auto& q : positive)
        if (!isfinite(q.v) || q.v <= 0.0) FAIL("atmosphere.%s must be positive and finite", q.name);
    if (!(0.0 <= p->absorption_peak_altitude_m && p->absorption_peak_altitude_m <= p->top_altitude_m))
        FAIL("atmosphere absorption peak must lie inside the shell");
    if (!(-0.99 < p->mie_phase_g && p->mie_phase_g < 0.99))
        FAIL("atmosphere.mie_phase_g must be in (-0.99, 0.99)");
    for (int c = 0; c < 3; ++c)
        if (p->mie_scattering_per_m[c] > p->mie_extinction_per_m[c])
            FAIL("Mie scattering cannot exceed Mie extinction");
    struct {
        int v, lo, hi;
        const char* text;
    } ranges[] = {
        {p->transmittance_lut_width, 16, 2048, "transmittance LUT width must be in 16..2048"},
        {p->transmittance_lut_height, 16, 1024, "transmittance LUT height must be in 16..1024"},
        {p->sky_view_lut_width, 16, 2048, "sky-view LUT width must be in 16..2048"},
        {p->sky_view_lut_height, 16, 2048, "sky-view LUT height must be in 16..2048"},
        {p->multi_scattering_lut_width, 8, 256, "multi-scattering LUT width must be in 8..256"},
        {p->multi_scattering_lut_height, 8, 256, "multi-scattering LUT height must be in 8..256"},
        {p->aerial_lut_width, 4, 256, "aerial LUT width must be in 4..256"},
        {p->aerial_lut_height, 4, 256, "aerial LUT height must be in 4..256"},
        {p->aerial_lut_depth, 8, 256, "aerial LUT depth must be in 8..256"},
        {p->transmittance_steps, 4, 256, "transmittance_steps must be in 4..256"},
        {p->multi_scattering_directions, 4, 128, "multi_scattering_directions must be in 4..128"},
        {p->multi_scattering_steps, 4, 128, "multi_scattering_steps must be in 4..128"},
        {p->sky_view_steps, 4, 256, "sky_view_steps must be in 4..256"},
        {p->aerial_steps_per_slice, 1, 16, "aerial_steps_per_slice must be in 1..16"},
        {p->aerial_horizon_raymarch_steps, 4, 64, "aerial_horizon_raymarch_steps must be in 4..64"},
        {p->aerial_terminator_substeps, 2, 16, "aerial_terminator_substeps must be in 2..16"},
    };
    for (auto& r : ranges)
        if (r.v < r.lo || r.v > r.hi) FAIL("%s", r.text);
#undef FAIL
    return ATMO_OK;
}

Consts make_consts(const AtmoParams& p) {
    Consts k;
    const double pi = acos(-1.0);
    const double g = p.mie_phase_g;
    k.pi = (float)pi;
    k.pi16 = (float)(16.0 * pi);
    k.large = 1.0e30f;
    k.top = (float)p.top_altitude_m;
    for (int c = 0; c < 3; ++c) {
        k.beta_r[c] = (float)p.rayleigh_scattering_per_m[c];
        k.beta_ms[c] = (float)p.mie_scattering_per_m[c];
        k.beta_me[c] = (float)p.mie_extinction_per_m[c];
        k.beta_a[c] = (float)p.absorption_extinction_per_m[c];
        k.albedo[c] = (float)p.ground_albedo[c];
    }
    k.h_r = (float)p.rayleigh_scale_height_m;
    k.h_m = (float)p.mie_scale_height_m;
    k.peak = (float)p.absorption_peak_altitude_m;
    k.half_width = (float)p.absorption_half_width_m;
    k.mie_a = (float)(1.0 + g * g);
    k.mie_b = (float)(2.0 * g);
    k.mie_c = (float)(3.0 * (1.0 - g * g));
    k.mie_d = (float)(8.0 * pi * (2.0 + g * g));
    k.tw = p.transmittance_lut_width;
    k.th = p.transmittance_lut_height;
    k.mw = p.multi_scattering_lut_width;
    k.mh = p.multi_scattering_lut_height;
    k.sw = p.sky_view_lut_width;
    k.sh = p.sky_view_lut_height;
    k.aw = p.aerial_lut_width;
    k.ah = p.aerial_lut_height;
    k.ad = p.aerial_lut_depth;
    k.t_steps = p.transmittance_steps;
    k.ms_dirs = p.multi_scattering_directions;
    k.ms_steps = p.multi_scattering_steps;
    k.sv_steps = p.sky_view_steps;
    k.ap_spp = p.aerial_steps_per_slice;
    k.hz_steps = p.aerial_horizon_raymarch_steps;
    k.term_sub = p.aerial_terminator_substeps;
    return k;
}

f3 host3(const float* a) { return f3{a[0], a[1], a[2]}; }

// Ordered combination of consecutive (L, T) segments within a lane group of
// width G: afterwards lane i holds the composition of segments 0..i
// (earlier segment first: L = L' + T' * L, T = T' * T).
template <int G>
__device__ __forceinline__ void scan_segments(f3& L, f3& T, int lane) {
#pragma unroll
    for (int off = 1; off < G; off <<= 1) {
        f3 pl = mk(__shfl_up(L.x, off, G), __shfl_up(L.y, off, G), __shfl_up(L.z, off, G));
        f3 pt = mk(__shfl_up(T.x, off, G), __shfl_up(T.y, off, G), __shfl_up(T.z, off, G));
        if (lane >= off) {
            L = pl + pt * L;
            T = pt * T;
        }
    }
}

}  // namespace

// ------------------------------------------------------------ kernels

// _build_transmittance (renderer.py:1034-1063) with _integrate_extinction
// (renderer.py:570-607).  G lanes per texel.
template <int G>
__device__ __forceinline__ void transmittance_body(const Consts& k, float bottom,
                                                   float* __restrict__ tlut, long gid) {
    const int lane = threadIdx.x % G;
    const long n = (long)k.tw * k.th;
    const long texel_raw = gid / G;
    const bool active = texel_raw < n;
    const long texel = active ? texel_raw : n - 1;
    const int x = (int)(texel / k.th), y = (int)(texel % k.th);
    float u = ((float)x + 0.5f) / (float)k.tw;
    float v = ((float)y + 0.5f) / (float)k.th;
    float alt, c, distance;
    transmittance_uv_to_ray(k, u, v, bottom, alt, c, distance);
    float sine = sqrtf(fmaxf(1.0f - c * c, 0.0f));
    f3 origin = mk(0.0f, bottom + alt, 0.0f);
    f3 ray = mk(sine, c, 0.0f);
    Warp w = make_warp(-dot(origin, ray), 0.0f, distance);
    const int steps = k.t_steps;
    const int per = (steps + G - 1) / G;
    const int s0 = min(lane * per, steps), s1 = min(s0 + per, steps);
    f3 od = mk(0.0f, 0.0f, 0.0f);
    for (int s = s0; s < s1; ++s) {
        float d0 = boundary(w, (float)s / (float)steps);
        float d1 = boundary(w, (float)(s + 1) / (float)steps);
        float step = d1 - d0;
        float dist = (d0 + d1) * 0.5f;
        float h = norm(origin + ray * dist) - bottom;
        od = od + extinction_of(k, density(k, h)) * step;
    }
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) {
        od.x += __shfl_xor(od.x, off, G);
        od.y += __shfl_xor(od.y, off, G);
        od.z += __shfl_xor(od.z, off, G);
    }
    if (active && lane == 0) st3(tlut, texel, clamp3(exp3(-mins(od, 80.0f)), 0.0f, 1.0f));
}

// _build_multi_scattering (renderer.py:1065-1206).  One wave64 per texel,
// lane = Fibonacci-sphere direction.
__device__ __forceinline__ void multi_scattering_body(const Consts& k, float bottom, f3 irradiance,
                                                      float ar, const float* __restrict__ tlut,
                                                      float* __restrict__ mslut, long gid) {
    const long wave = gid / 64;
    const int lane = threadIdx.x % 64;
    const long n = (long)k.mw * k.mh;
    if (wave >= n) return;  // uniform per wave
    const int x = (int)(wave / k.mh), y = (int)(wave % k.mh);
    float u = ((float)x + 0.5f) / (float)k.mw;
    float v = ((float)y + 0.5f) / (float)k.mh;
    float h0 = v * v * k.top;
    float radius = bottom + h0;
    float sun_c = sun_u_to_cosine(u, h0, bottom);
    float sun_s = sqrtf(fmaxf(1.0f - sun_c * sun_c, 0.0f));
    f3 sun = mk(sun_s, sun_c, 0.0f);
    f3 origin = mk(0.0f, radius, 0.0f);
    const int N = k.ms_dirs;
    const int S = k.ms_steps;
    f3 lum = mk(0.0f, 0.0f, 0.0f), fb = mk(0.0f, 0.0f, 0.0f);
    for (int di = lane; di < N; di += 64) {
        float si = (float)di + 0.5f;
        float dc = 1.0f - 2.0f * si / (float)N;
        float ds = sqrtf(fmaxf(1.0f - dc * dc, 0.0f));
        float az = (float)di * 2.39996323f;
        f3 ray = mk(ds * cr_cos(az), dc, ds * cr_sin(az));
        float start, end;
        bool valid, ground;
        segment_vec(k, origin, ray, bottom, start, end, valid, ground);
        f3 dl = mk(0.0f, 0.0f, 0.0f), df = mk(0.0f, 0.0f, 0.0f), pt = mk(1.0f, 1.0f, 1.0f);
        if (valid) {
            Warp w = make_warp(-dot(origin, ray), start, end);
            RayPhase ph = ray_phase(k, ray, sun);
            float d0 = boundary(w, 0.0f / (float)S);
            for (int s = 0; s < S; ++s) {
                float d1 = boundary(w, (float)(s + 1) / (float)S);
                float dist = (d0 + d1) * 0.5f;
                float step = d1 - d0;
                f3 point = origin + ray * dist;
                float h = altitude_on_ray(h0, dc, dist, bottom);
                f3 d = density(k, h);
                f3 st, si3;
                segment_integral(extinction_of(k, d), step, st, si3);
                f3 direct = scattering_source<false>(k, tlut, nullptr, point, h, d, ph, sun,
                                                     irradiance, bottom, ar);
                dl = dl + (pt * direct) * si3;
                df = df + (pt * scattering_of(k, d)) * si3;
                pt = pt * st;
                d0 = d1;
            }
            if (ground) {
                f3 gp = origin + ray * end;
                f3 gn = normalized(gp);
                float gsc = dot(gn, sun);
                SunGeom g = sun_geometry(k, 0.0f, gsc, bottom, ar);
                f3 gt = solar_transmittance(k, tlut, g, 0.0f, bottom, ar);
                f3 gl = (((arr3(k.albedo) * irradiance) * gt) * fmaxf(gsc, 0.0f)) / k.pi;
                dl = dl + pt * gl;
            }
        }
        lum = lum + dl;
        fb = fb + df;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        lum.x += __shfl_xor(lum.x, off, 64);
        lum.y += __shfl_xor(lum.y, off, 64);
        lum.z += __shfl_xor(lum.z, off, 64);
        fb.x += __shfl_xor(fb.x, off, 64);
        fb.y += __shfl_xor(fb.y, off, 64);
        fb.z += __shfl_xor(fb.z, off, 64);
    }
    if (lane == 0) {
        f3 avg_l = lum / (float)N;
        f3 avg_f = clamp3(fb / (float)N, 0.0f, 0.95f);
        st3(mslut, wave, maxs(avg_l / (mk(1.0f, 1.0f, 1.0f) - avg_f), 0.0f));
    }
}

// _build_sky_view (renderer.py:1208-1302).  G lanes per texel, ordered
// (L, T) scan over the lanes' step chunks.
template <int G>
__device__ __forceinline__ void sky_view_body(const Consts& k, float bottom, float cam_alt,
                                              float sun_zc, f3 irradiance, float ar,
                                              const float* __restrict__ tlut,
                                              const float* __restrict__ mslut,
                                              float* __restrict__ svlut, long gid) {
    const int lane = threadIdx.x % G;
    const long n = (long)k.sw * k.sh;
    const long texel_raw = gid / G;
    const bool active = texel_raw < n;
    const long texel = active ? texel_raw : n - 1;
    const int x = (int)(texel / k.sh), y = (int)(texel % k.sh);
    float sun_s = sqrtf(fmaxf(1.0f - sun_zc * sun_zc, 0.0f));
    f3 sun = mk(sun_s, sun_zc, 0.0f);
    float safe_alt = fmaxf(cam_alt, 0.0f);
    f3 origin = mk(0.0f, bottom + safe_alt, 0.0f);
    float az = ((float)x + 0.5f) / (float)k.sw * k.pi;
    float view_v = ((float)y + 0.5f) / (float)k.sh;
    float zenith = sky_v_to_zenith(k, view_v, sky_horizon_zenith(k, safe_alt, bottom));
    float vc = cr_cos(zenith);
    float vsn = cr_sin(zenith);
    f3 ray = mk(vsn * cr_cos(az), vc, vsn * cr_sin(az));
    float start, end;
    bool valid, ground;
    segment_from_altitude(k, safe_alt, vc, bottom, start, end, valid, ground);
    f3 L = mk(0.0f, 0.0f, 0.0f), T = mk(1.0f, 1.0f, 1.0f);
    if (valid) {
        Warp w = make_warp(-dot(origin, ray), start, end);
        RayPhase ph = ray_phase(k, ray, sun);
        const int steps = k.sv_steps;
        const int per = (steps + G - 1) / G;
        const int s0 = min(lane * per, steps), s1 = min(s0 + per, steps);
        for (int s = s0; s < s1; ++s) {
            float d0 = boundary(w, (float)s / (float)steps);
            float d1 = boundary(w, (float)(s + 1) / (float)steps);
            float dist = (d0 + d1) * 0.5f;
            float step = d1 - d0;
            f3 point = origin + ray * dist;
            float h = altitude_on_ray(safe_alt, vc, dist, bottom);
            f3 d = density(k, h);
            f3 src = scattering_source<true>(k, tlut, mslut, point, h, d, ph, sun, irradiance,
                                             bottom, ar);
            f3 st, si3;
            segment_integral(extinction_of(k, d), step, st, si3);
            L = L + (T * src) * si3;
            T = T * st;
        }
    }
    scan_segments<G>(L, T, lane);
    if (active && lane == G - 1) st3(svlut, texel, maxs(L, 0.0f));
}

// _build_aerial_perspective (renderer.py:1304-1420).  One wave64 per froxel
// ray, lane = depth slice; slices beyond 64 are processed in chunks with the
// previous chunk's composition carried in.
__device__ __forceinline__ void aerial_body(const Consts& k, int width, int height, float bottom,
                                            float cam_alt, f3 sun, f3 irradiance, float ar, f3 right,
                                            f3 up, f3 forward, float thf,
                                            const float* __restrict__ tlut,
                                            const float* __restrict__ mslut,
                                            float* __restrict__ ap, long gid) {
    const long wave = gid / 64;
    const int lane = threadIdx.x % 64;
    const long n = (long)k.aw * k.ah;
    if (wave >= n) return;  // uniform per wave
    const int x = (int)(wave / k.ah), y = (int)(wave % k.ah);
    float aspect = (float)width / (float)height;
    f3 origin = mk(0.0f, bottom + cam_alt, 0.0f);
    float u = ((float)x + 0.5f) / (float)k.aw;
    float v = ((float)y + 0.5f) / (float)k.ah;
    float sx = (u * 2.0f - 1.0f) * aspect;
    float sy = v * 2.0f - 1.0f;
    f3 ray = normalized((right * sx + up * sy) + forward / thf);
    float a_start, a_end;
    bool valid;
    aerial_prefix(k, cam_alt, ray.y, bottom, a_start, a_end, valid);
    RayPhase ph = ray_phase(k, ray, sun);
    const float neg_od = -dot(origin, ray);
    const int D = k.ad;
    const int spp = k.ap_spp;
    f3 carry_l = mk(0.0f, 0.0f, 0.0f), carry_t = mk(1.0f, 1.0f, 1.0f);
    for (int base = 0; base < D; base += 64) {
        const int di = base + lane;
        f3 L = mk(0.0f, 0.0f, 0.0f), T = mk(1.0f, 1.0f, 1.0f);
        if (valid && di < D) {
            float previous = di == 0 ? a_start
                                     : a_start + (a_end - a_start) * ((float)(di - 1) / (float)(D - 1));
            float target = a_start + (a_end - a_start) * ((float)di / (float)(D - 1));
            float is = fmaxf(previous, a_start);
            float ie = fminf(target, a_end);
            if (ie > is) {
                Warp w = make_warp(neg_od, is, ie);
                float d0 = boundary(w, 0.0f / (float)spp);
                for (int s = 0; s < spp; ++s) {
                    float d1 = boundary(w, (float)(s + 1) / (float)spp);
                    float dist = (d0 + d1) * 0.5f;
                    float step = d1 - d0;
                    f3 point = origin + ray * dist;
                    float h = altitude_on_ray(cam_alt, ray.y, dist, bottom);
                    f3 d = density(k, h);
                    f3 st, si3;
                    segment_integral(extinction_of(k, d), step, st, si3);
                    f3 src = scattering_source<true>(k, tlut, mslut, point, h, d, ph, sun,
                                                     irradiance, bottom, ar);
                    L = L + (T * src) * si3;
                    T = T * st;
                    d0 = d1;
                }
            }
        }
        scan_segments<64>(L, T, lane);
        f3 cl = carry_l + carry_t * L;
        f3 ct = carry_t * T;
        if (di < D) st3(ap, wave * D + di, maxs(cl, 0.0f));
        carry_l = mk(__shfl(cl.x, 63, 64), __shfl(cl.y, 63, 64), __shfl(cl.z, 63, 64));
        carry_t = mk(__shfl(ct.x, 63, 64), __shfl(ct.y, 63, 64), __shfl(ct.z, 63, 64));
    }
}


// Single-state kernels: one Taichi kernel call each.
template <int G>
__global__ void __launch_bounds__(256) k_transmittance(Consts k, float bottom, float* __restrict__ tlut) {
    transmittance_body<G>(k, bottom, tlut, (long)blockIdx.x * blockDim.x + threadIdx.x);
}

__global__ void __launch_bounds__(256) k_multi_scattering(Consts k, float bottom, f3 irradiance,
                                                          float ar, const float* __restrict__ tlut,
                                                          float* __restrict__ mslut) {
    multi_scattering_body(k, bottom, irradiance, ar, tlut, mslut,
                          (long)blockIdx.x * blockDim.x + threadIdx.x);
}

template <int G>
__global__ void __launch_bounds__(256) k_sky_view(Consts k, float bottom, float cam_alt, float sun_zc,
                                                  f3 irradiance, float ar,
                                                  const float* __restrict__ tlut,
                                                  const float* __restrict__ mslut,
                                                  float* __restrict__ svlut) {
    sky_view_body<G>(k, bottom, cam_alt, sun_zc, irradiance, ar, tlut, mslut, svlut,
                     (long)blockIdx.x * blockDim.x + threadIdx.x);
}

__global__ void __launch_bounds__(256) k_aerial(Consts k, int width, int height, float bottom,
                                                float cam_alt, f3 sun, f3 irradiance, float ar,
                                                f3 right, f3 up, f3 forward, float thf,
                                                const float* __restrict__ tlut,
                                                const float* __restrict__ mslut,
                                                float* __restrict__ ap) {
    aerial_body(k, width, height, bottom, cam_alt, sun, irradiance, ar, right, up, forward, thf, tlut,
                mslut, ap, (long)blockIdx.x * blockDim.x + threadIdx.x);
}

// Batched kernels: blockIdx.y selects the state, whose float32 kernel
// arguments and LUT slices come from the device state table; the same
// bodies as the single-state kernels, so every state's LUTs are exactly
// what four single-state calls would produce.
template <int G>
__global__ void __launch_bounds__(256) k_transmittance_batched(const BatchState* __restrict__ st,
                                                               float* __restrict__ tluts,
                                                               long stride) {
    const BatchState& b = st[blockIdx.y];
    transmittance_body<G>(b.k, b.bottom, tluts + blockIdx.y * stride,
                          (long)blockIdx.x * blockDim.x + threadIdx.x);
}

__global__ void __launch_bounds__(256) k_multi_scattering_batched(const BatchState* __restrict__ st,
                                                                  const float* __restrict__ tluts,
                                                                  long t_stride,
                                                                  float* __restrict__ msluts,
                                                                  long m_stride) {
    const BatchState& b = st[blockIdx.y];
    multi_scattering_body(b.k, b.bottom, b.irradiance, b.ar_update, tluts + blockIdx.y * t_stride,
                          msluts + blockIdx.y * m_stride, (long)blockIdx.x * blockDim.x + threadIdx.x);
}

template <int G>
__global__ void __launch_bounds__(256) k_sky_view_batched(const BatchState* __restrict__ st,
                                                          const float* __restrict__ tluts, long t_stride,
                                                          const float* __restrict__ msluts,
                                                          long m_stride, float* __restrict__ svluts,
                                                          long s_stride) {
    const BatchState& b = st[blockIdx.y];
    sky_view_body<G>(b.k, b.bottom, b.sky_alt, b.sun_zc, b.irradiance, b.ar_update,
                     tluts + blockIdx.y * t_stride, msluts + blockIdx.y * m_stride,
                     svluts + blockIdx.y * s_stride, (long)blockIdx.x * blockDim.x + threadIdx.x);
}

__global__ void __launch_bounds__(256) k_aerial_batched(const BatchState* __restrict__ st, int width,
                                                        int height, const float* __restrict__ tluts,
                                                        long t_stride, const float* __restrict__ msluts,
                                                        long m_stride, float* __restrict__ aps,
                                                        long a_stride) {
    const BatchState& b = st[blockIdx.y];
    aerial_body(b.k, width, height, b.bottom, b.cam_alt, b.sun, b.irradiance, b.ar, b.right, b.up,
                b.forward, b.thf, tluts + blockIdx.y * t_stride, msluts + blockIdx.y * m_stride,
                aps + blockIdx.y * a_stride, (long)blockIdx.x * blockDim.x + threadIdx.x);
}

// _prepare_surface_lighting (renderer.py:1422-1472).  One thread per pixel.
__global__ void __launch_bounds__(256) k_prepare(Consts k, long npix,
                                                 const float* __restrict__ position_view,
                                                 const int32_t* __restrict__ surface_id,
                                                 float bottom, float cam_alt, f3 sun, float ar,
                                                 f3 right, f3 up, f3 forward,
                                                 const float* __restrict__ tlut,
                                                 const float* __restrict__ mslut,
                                                 float* __restrict__ sun_t,
                                                 float* __restrict__ sky_l) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npix) return;
    f3 solar = mk(1.0f, 1.0f, 1.0f);
    f3 sky = mk(0.0f, 0.0f, 0.0f);
    if (surface_id[i] >= 0) {
        f3 camera = mk(0.0f, bottom + cam_alt, 0.0f);
        f3 vp = ld3(position_view, i);
        f3 local = (right * vp.x + up * vp.y) + forward * vp.z;
        f3 planet = camera + local;
        f3 radial = planet / fmaxf(norm(planet), 1.0f);
        float sd = norm(local);
        f3 local_ray = local / fmaxf(sd, 1.0f);
        float h = altitude_on_ray(cam_alt, local_ray.y, sd, bottom);
        float sc = dot(radial, sun);
        SunGeom g = sun_geometry(k, h, sc, bottom, ar);
        solar = solar_transmittance(k, tlut, g, h, bottom, ar);
        sky = sample_multi_scattering(k, mslut, h, sc, bottom);
    }
    st3(sun_t, i, solar);
    st3(sky_l, i, maxs(sky, 0.0f));
}


__device__ __forceinline__ f3 pixel_ray(const CompositeArgs& a, int px, int py, float& su,
                                        float& sv) {
    float aspect = (float)a.width / (float)a.height;
    su = ((float)px + 0.5f) / (float)a.width;
    sv = ((float)py + 0.5f) / (float)a.height;
    float sx = (su * 2.0f - 1.0f) * aspect;
    float sy = sv * 2.0f - 1.0f;
    return normalized((a.right * sx + a.up * sy) + a.forward / a.thf);
}

// The COMPOSITE view of _composite (renderer.py:1860-1940) as two passes
// over the same pixels: sky pixels (k_composite_sky) and surface pixels
// (k_composite_surface).  Each pass walks 256-pixel groups with persistent
// blocks and compacts its own pixel class onto the leading threads, so
// waves run full and each pass gets the register budget of its own branch
// (the surface branch alone fits three waves per SIMD).  Every pixel is
// written by exactly one thread of one pass.
__global__ void __launch_bounds__(256) k_composite_sky(Consts k, CompositeArgs a) {
    __shared__ int list[256];
    __shared__ int wave_counts[4];
    const long npix = (long)a.width * a.height;
    const long groups = (npix + 255) / 256;
    for (long g = blockIdx.x; g < groups; g += gridDim.x) {
        const long base = g * 256;
        const long i = base + threadIdx.x;
        const bool sky = i < npix && a.sid[i] < 0;
        const int count = compact_block(sky, list, wave_counts);
        if ((int)threadIdx.x < count) {
            const long j = base + list[threadIdx.x];
            f3 ray = camera_ray((int)(j / a.height), (int)(j % a.height), a.width, a.height,
                                a.right, a.up, a.forward, a.thf);
            f3 c = composite_sky(k, a.tlut, a.svlut, ray, ld3(a.scene, j), ld3(a.space, j),
                                 a.cam_alt, a.sky_alt, a.sun, a.disk, a.ar, a.thf, a.height, a.c0,
                                 a.c1, a.bottom);
            st3(a.hdr, j, maxs(c, 0.0f));
        }
        __syncthreads();  // `list` is rewritten by the next group
    }
}

// Three waves per SIMD (at most 168 VGPRs, no spills) measured 15 % faster
// than the compiler's unconstrained two-wave allocation at 1080p.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 8)))
k_composite_surface(Consts k, CompositeArgs a) {
    __shared__ int list[256];
    __shared__ int wave_counts[4];
    const long npix = (long)a.width * a.height;
    const long groups = (npix + 255) / 256;
    for (long g = blockIdx.x; g < groups; g += gridDim.x) {
        const long base = g * 256;
        const long i = base + threadIdx.x;
        const bool surface = i < npix && a.sid[i] >= 0;
        const int count = compact_block(surface, list, wave_counts);
        if ((int)threadIdx.x < count) {
            const long j = base + list[threadIdx.x];
            f3 ray = camera_ray((int)(j / a.height), (int)(j % a.height), a.width, a.height,
                                a.right, a.up, a.forward, a.thf);
            float sd = norm(ld3(a.pos, j));
            f3 scat = surface_inscatter(k, a.tlut, a.mslut, a.cam_alt, ray, sd, a.sun,
                                        a.irradiance, a.bottom, a.ar);
            f3 trans = surface_transmittance(k, a.tlut, a.cam_alt, ray, sd, a.bottom);
            st3(a.hdr, j, maxs(scat + ld3(a.scene, j) * trans, 0.0f));
        }
        __syncthreads();
    }
}

// Diagnostic views 1..10 of _composite (renderer.py:1941-2055).
__global__ void __launch_bounds__(256) k_composite_diagnostic(Consts k, CompositeArgs a) {
    const long npix = (long)a.width * a.height;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npix) return;
    const int px = (int)(i / a.height), py = (int)(i % a.height);
    float su, sv;
    f3 ray = pixel_ray(a, px, py, su, sv);
    const bool surface = a.sid[i] >= 0;
    f3 color = mk(0.0f, 0.0f, 0.0f);
    switch (a.view) {
        case ATMO_VIEW_SKY_VIEW:
            color = sample_sky_view(k, a.svlut, ray, a.sun, a.sky_alt, a.bottom);
            break;
        case ATMO_VIEW_CAMERA_TRANSMITTANCE:
            color = camera_transmittance(k, a.tlut, a.cam_alt, ray, a.bottom);
            break;
        case ATMO_VIEW_TRANSMITTANCE_LUT: {
            int lx = min((int)(su * (float)k.tw), k.tw - 1);
            int ly = min((int)(sv * (float)k.th), k.th - 1);
            color = ld3(a.tlut, (long)lx * k.th + ly);
            break;
        }
        case ATMO_VIEW_MULTI_SCATTERING_LUT: {
            int lx = min((int)(su * (float)k.mw), k.mw - 1);
            int ly = min((int)(sv * (float)k.mh), k.mh - 1);
            color = ld3(a.mslut, (long)lx * k.mh + ly);
            break;
        }
        case ATMO_VIEW_AERIAL_SCATTERING:
        case ATMO_VIEW_AERIAL_DIRECT_ONLY:
            if (surface)
                color = surface_inscatter(k, a.tlut, a.mslut, a.cam_alt, ray, norm(ld3(a.pos, i)),
                                          a.sun, a.irradiance, a.bottom, a.ar);
            break;
        case ATMO_VIEW_AERIAL_FROXEL_ONLY:
            if (surface)
                color = sample_aerial(k, a.aplut, su, sv, norm(ld3(a.pos, i)), a.cam_alt, ray,
                                      a.bottom);
            break;
        case ATMO_VIEW_AERIAL_BLEND_WEIGHT: {
            float w = surface ? 1.0f : 0.0f;
            color = mk(w, w, w);
            break;
        }
        case ATMO_VIEW_AERIAL_TRANSMITTANCE:
            if (surface)
                color = surface_transmittance(k, a.tlut, a.cam_alt, ray, norm(ld3(a.pos, i)),
                                              a.bottom);
            break;
        default: {
            float m = surface ? 1.0f : 0.0f;
            color = mk(m, m, m);
            break;
        }
    }
    st3(a.hdr, i, maxs(color, 0.0f));
}

// ------------------------------------------------------------ C-ABI

namespace {

constexpr int kGroupT = 8;   // lanes per transmittance texel
constexpr int kGroupSV = 8;  // lanes per sky-view texel

unsigned blocks_for(long threads, int block) { return (unsigned)((threads + block - 1) / block); }

// Persistent grid of the composite passes: enough 256-thread blocks for
// several per CU on 256 CUs, never more than there are 256-pixel groups.
unsigned composite_blocks(long npix) {
    const long groups = (npix + 255) / 256;
    return (unsigned)(groups < 2048 ? groups : 2048);
}

float** buffer_slot(AtmoHandle* h, int which) {
    switch (which) {
        case ATMO_BUF_TRANSMITTANCE: return &h->tlut;
        case ATMO_BUF_MULTI_SCATTERING: return &h->mslut;
        case ATMO_BUF_SKY_VIEW: return &h->svlut;
        case ATMO_BUF_AERIAL_SCATTERING: return &h->aplut;
        case ATMO_BUF_SURFACE_SUN_T: return &h->sun_t;
        default: return &h->sky_l;
    }
}

void buffer_shape(const AtmoHandle* h, int which, int64_t s[4]) {
    const Consts& k = h->k;
    s[0] = s[1] = s[2] = s[3] = 0;
    switch (which) {
        case ATMO_BUF_TRANSMITTANCE: s[0] = k.tw; s[1] = k.th; s[2] = 3; break;
        case ATMO_BUF_MULTI_SCATTERING: s[0] = k.mw; s[1] = k.mh; s[2] = 3; break;
        case ATMO_BUF_SKY_VIEW: s[0] = k.sw; s[1] = k.sh; s[2] = 3; break;
        case ATMO_BUF_AERIAL_SCATTERING: s[0] = k.aw; s[1] = k.ah; s[2] = k.ad; s[3] = 3; break;
        default: s[0] = h->width; s[1] = h->height; s[2] = 3; break;
    }
}

size_t buffer_floats(const AtmoHandle* h, int which) {
    int64_t s[4];
    buffer_shape(h, which, s);
    size_t n = 1;
    for (int d = 0; d < 4; ++d)
        if (s[d] > 0) n *= (size_t)s[d];
    return n;
}

void free_buffers(AtmoHandle* h) {
    for (int b = 0; b < 6; ++b) {
        float** slot = buffer_slot(h, b);
        if (*slot && !h->borrowed[b]) (void)hipFree(*slot);
        *slot = nullptr;
    }
    if (h->queue) (void)hipFree(h->queue);
    h->queue = nullptr;
}

}  // namespace

extern "C" {

int32_t atmo_validate_params(const AtmoParams* params, char* message, int32_t capacity) {
    if (!params) {
        if (message && capacity > 0) snprintf(message, (size_t)capacity, "params is NULL");
        return ATMO_EINVAL;
    }
    return validate(params, message, capacity > 0 ? (size_t)capacity : 0);
}

int32_t atmo_create(const AtmoParams* params, int32_t width, int32_t height, AtmoHandle** out) {
    if (!out || !params) {
        set_error(nullptr, "atmo_create: NULL argument");
        return ATMO_EINVAL;
    }
    *out = nullptr;
    if (validate(params, g_create_error, sizeof g_create_error) != ATMO_OK) return ATMO_EINVAL;
    if (width <= 0 || height <= 0) {
        set_error(nullptr, "width and height must be positive");
        return ATMO_EINVAL;
    }
    AtmoHandle* h = new (std::nothrow) AtmoHandle();
    if (!h) {
        set_error(nullptr, "atmo_create: host allocation failed");
        return ATMO_ENOMEM;
    }
    h->magic = kAtmoMagic;
    h->params = *params;
    h->k = make_consts(*params);
    h->width = width;
    h->height = height;
    h->error[0] = 0;
    for (int b = 0; b < 6; ++b) {
        *buffer_slot(h, b) = nullptr;
        h->borrowed[b] = false;
    }
    h->queue = nullptr;
    if (hipGetDevice(&h->device) != hipSuccess) {
        set_error(nullptr, "atmo_create: no HIP device");
        delete h;
        return ATMO_EHIP;
    }
    for (int b = 0; b < 6; ++b) {
        size_t bytes = buffer_floats(h, b) * sizeof(float);
        if (hipMalloc((void**)buffer_slot(h, b), bytes) != hipSuccess ||
            hipMemset(*buffer_slot(h, b), 0, bytes) != hipSuccess) {
            set_error(nullptr, "atmo_create: device allocation failed");
            free_buffers(h);
            delete h;
            return ATMO_ENOMEM;
        }
    }
    h->queue = nullptr;
    if (hipMalloc((void**)&h->queue, ((size_t)width * (size_t)height + 1) * sizeof(int)) != hipSuccess) {
        set_error(nullptr, "atmo_create: device allocation failed");
        free_buffers(h);
        delete h;
        return ATMO_ENOMEM;
    }
    if (hipDeviceSynchronize() != hipSuccess) {
        set_error(nullptr, "atmo_create: device synchronize failed");
        free_buffers(h);
        delete h;
        return ATMO_EHIP;
    }
    *out = h;
    return ATMO_OK;
}

int32_t atmo_destroy(AtmoHandle* h) {
    if (!atmo_handle_ok(h)) return ATMO_EHANDLE;
    free_buffers(h);
    h->magic = 0;
    delete h;
    return ATMO_OK;
}

const char* atmo_last_error(const AtmoHandle* h) {
    if (!h) return g_create_error;
    if (!atmo_handle_ok(h)) return "invalid atmosphere handle";
    return h->error;
}

int32_t atmo_get_buffer(AtmoHandle* h, int32_t which, void** dptr, int64_t shape[4]) {
    if (!atmo_handle_ok(h)) return ATMO_EHANDLE;
    if (!dptr || !shape) {
        set_error(h, "atmo_get_buffer: NULL argument");
        return ATMO_EINVAL;
    }
    if (which < 0 || which > 5) {
        set_error(h, "atmo_get_buffer: unknown buffer id");
        return ATMO_EINVAL;
    }
    buffer_shape(h, which, shape);
    *dptr = *buffer_slot(h, which);
    return ATMO_OK;
}

int32_t atmo_bind_buffer(AtmoHandle* h, int32_t which, void* dptr, int64_t bytes) {
    if (!atmo_handle_ok(h)) return ATMO_EHANDLE;
    if (which < 0 || which > 5) {
        set_error(h, "atmo_bind_buffer: unknown buffer id");
        return ATMO_EINVAL;
    }
    if (!dptr) {
        set_error(h, "atmo_bind_buffer: NULL device pointer");
        return ATMO_EINVAL;
    }
    if (bytes != (int64_t)(buffer_floats(h, which) * sizeof(float))) {
        set_error(h, "atmo_bind_buffer: buffer size does not match the field shape");
        return ATMO_ESHAPE;
    }
    float** slot = buffer_slot(h, which);
    if (*slot && !h->borrowed[which]) {
        if (hipDeviceSynchronize() != hipSuccess || hipFree(*slot) != hipSuccess) {
            set_error(h, "atmo_bind_buffer: failed to release the owned buffer");
            return ATMO_EHIP;
        }
    }
    *slot = (float*)dptr;
    h->borrowed[which] = true;
    return ATMO_OK;
}

int32_t atmo_build_transmittance(AtmoHandle* h, float bottom, void* stream) {
    if (!atmo_handle_ok(h)) return ATMO_EHANDLE;
    const long threads = (long)h->k.tw * h->k.th * kGroupT;
    hipLaunchKernelGGL(k_transmittance<kGroupT>, dim3(blocks_for(threads, 256)), dim3(256), 0,
                       (hipStream_t)stream, h->k, bottom, h->tlut);
    return atmo_launch_status(h, "atmo_build_transmittance");
}

int32_t atmo_build_multi_scattering(AtmoHandle* h, float bottom, const float irradiance[3],
                                    float ar, void* stream) {
    if (!atmo_handle_ok(h)) return ATMO_EHANDLE;
    if (!irradiance) {
        set_error(h, "atmo_build_multi_scattering: NULL irradiance");
        return ATMO_EINVAL;
    }
    const long threads = (long)h->k.mw * h->k.mh * 64;
    hipLaunchKernelGGL(k_multi_scattering, dim3(blocks_for(threads, 256)), dim3(256), 0,
                       (hipStream_t)stream, h->k, bottom, host3(irradiance), ar, h->tlut,
                       h->mslut);
    return atmo_launch_status(h, "atmo_build_multi_scattering");
}

int32_t atmo_build_sky_view(AtmoHandle* h, float bottom, float cam_alt, float sun_zc,
                            const float irradiance[3], float ar, void* stream) {
    if (!atmo_handle_ok(h)) return ATMO_EHANDLE;
    if (!irradiance) {
        set_error(h, "atmo_build_sky_view: NULL irradiance");
        return ATMO_EINVAL;
    }
    const long threads = (long)h->k.sw * h->k.sh * kGroupSV;
    hipLaunchKernelGGL(k_sky_view<kGroupSV>, dim3(blocks_for(threads, 256)), dim3(256), 0,
                       (hipStream_t)stream, h->k, bottom, cam_alt, sun_zc, host3(irradiance), ar,
                       h->tlut, h->mslut, h->svlut);
    return atmo_launch_status(h, "atmo_build_sky_view");
}

int32_t atmo_build_aerial_perspective(AtmoHandle* h, float bottom, float cam_alt,
                                      const float sun[3], const float irradiance[3], float ar,
                                      const float right[3], const float up[3],
                                      const float forward[3], float thf, void* stream) {
    if (!atmo_handle_ok(h)) return ATMO_EHANDLE;
    if (!sun || !irradiance || !right || !up || !forward) {
        set_error(h, "atmo_build_aerial_perspective: NULL vector argument");
        return ATMO_EINVAL;
    }
    const long threads = (long)h->k.aw * h->k.ah * 64;
    hipLaunchKernelGGL(k_aerial, dim3(blocks_for(threads, 256)), dim3(256), 0,
                       (hipStream_t)stream, h->k, h->width, h->height, bottom, cam_alt,
                       host3(sun), host3(irradiance), ar, host3(right), host3(up),
                       host3(forward), thf, h->tlut, h->mslut, h->aplut);
    return atmo_launch_status(h, "atmo_build_aerial_perspective");
}

int32_t atmo_prepare_surface_lighting(AtmoHandle* h, const float* position_view,
                                      const int32_t* surface_id, float bottom, float cam_alt,
                                      const float sun[3], float ar, const float right[3],
                                      const float up[3], const float forward[3], void* stream) {
    if (!atmo_handle_ok(h)) return ATMO_EHANDLE;
    if (!position_view || !surface_id || !sun || !right || !up || !forward) {
        set_error(h, "atmo_prepare_surface_lighting: NULL argument");
        return ATMO_EINVAL;
    }
    const long npix = (long)h->width * h->height;
    hipLaunchKernelGGL(k_prepare, dim3(blocks_for(npix, 256)), dim3(256), 0, (hipStream_t)stream,
                       h->k, npix, position_view, surface_id, bottom, cam_alt, host3(sun), ar,
                       host3(right), host3(up), host3(forward), h->tlut, h->mslut, h->sun_t,
                       h->sky_l);
    return atmo_launch_status(h, "atmo_prepare_surface_lighting");
}

}  // extern "C"

int32_t atmo_launch_composite_sky(AtmoHandle* h, const CompositeArgs& a, hipStream_t stream) {
    const long npix = (long)a.width * a.height;
    hipLaunchKernelGGL(k_composite_sky, dim3(composite_blocks(npix)), dim3(256), 0, stream, h->k, a);
    return atmo_launch_status(h, "k_composite_sky");
}

extern "C" {

int32_t atmo_composite(AtmoHandle* h, const float* scene_hdr, const float* space_hdr,
                       const int32_t* surface_id, const float* position_view, float* hdr,
                       float bottom, float cam_alt, float sky_alt, const float sun[3],
                       const float irradiance[3], const float disk[3], float ar,
                       const float right[3], const float up[3], const float forward[3], float thf,
                       float c0, float c1, int32_t view, void* stream) {
    if (!atmo_handle_ok(h)) return ATMO_EHANDLE;
    if (!scene_hdr || !space_hdr || !surface_id || !position_view || !hdr || !sun ||
        !irradiance || !disk || !right || !up || !forward) {
        set_error(h, "atmo_composite: NULL argument");
        return ATMO_EINVAL;
    }
    if (view < ATMO_VIEW_COMPOSITE || view > ATMO_VIEW_AERIAL_BLEND_WEIGHT) {
        set_error(h, "atmo_composite: diagnostic_view must be in 0..10");
        return ATMO_EINVAL;
    }
    CompositeArgs a;
    a.scene = scene_hdr;
    a.space = space_hdr;
    a.sid = surface_id;
    a.pos = position_view;
    a.hdr = hdr;
    a.tlut = h->tlut;
    a.mslut = h->mslut;
    a.svlut = h->svlut;
    a.aplut = h->aplut;
    a.width = h->width;
    a.height = h->height;
    a.bottom = bottom;
    a.cam_alt = cam_alt;
    a.sky_alt = sky_alt;
    a.sun = host3(sun);
    a.irradiance = host3(irradiance);
    a.disk = host3(disk);
    a.ar = ar;
    a.right = host3(right);
    a.up = host3(up);
    a.forward = host3(forward);
    a.thf = thf;
    a.c0 = c0;
    a.c1 = c1;
    a.view = view;
    const long npix = (long)h->width * h->height;
    if (view == ATMO_VIEW_COMPOSITE) {
        const unsigned blocks = composite_blocks(npix);
        hipLaunchKernelGGL(k_composite_sky, dim3(blocks), dim3(256), 0, (hipStream_t)stream, h->k, a);
        hipLaunchKernelGGL(k_composite_surface, dim3(blocks), dim3(256), 0, (hipStream_t)stream, h->k,
                           a);
    } else
        hipLaunchKernelGGL(k_composite_diagnostic, dim3(blocks_for(npix, 256)), dim3(256), 0,
                           (hipStream_t)stream, h->k, a);
    return atmo_launch_status(h, "atmo_composite");
}

}  // extern "C"

// ------------------------------------------------------------ batches

struct AtmoBatch {
    uint32_t magic;
    int count;
    int width;
    int height;
    Consts k0;           // LUT dimensions shared by every state
    BatchState* states;  // device state table [count]
    float* luts[4];      // T, MS, SV, AP arrays [count * per-state floats]
    bool borrowed[4];    // array supplied by the caller
    long strides[4];     // per-state floats of each LUT
    hipStream_t capture;
    hipGraph_t graph;
    hipGraphExec_t exec;
    bool captured;
    char error[512];
};

namespace {

constexpr uint32_t kBatchMagic = 0x42415443u;  // "BATC"

bool batch_ok(const AtmoBatch* b) { return b != nullptr && b->magic == kBatchMagic; }

void batch_free(AtmoBatch* b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->capture) (void)hipStreamDestroy(b->capture);
    if (b->states) (void)hipFree(b->states);
    for (int l = 0; l < 4; ++l)
        if (b->luts[l] && !b->borrowed[l]) (void)hipFree(b->luts[l]);
}

bool same_dims(const Consts& a, const Consts& b) {
    return a.tw == b.tw && a.th == b.th && a.mw == b.mw && a.mh == b.mh && a.sw == b.sw &&
           a.sh == b.sh && a.aw == b.aw && a.ah == b.ah && a.ad == b.ad;
}

// The reference's host-side rounding of one state's kernel arguments:
// update (renderer.py:2082-2137) and prepare_frame (renderer.py:2153-2184).
BatchState derive_state(const AtmoBatchState& s) {
    BatchState b;
    const double R = s.planet_radius_m, r = s.camera_radius_m;
    const double deg_to_rad = acos(-1.0) / 180.0;
    b.k = make_consts(s.params);
    b.bottom = (float)R;
    b.irradiance = f3{(float)s.solar_irradiance[0], (float)s.solar_irradiance[1],
                      (float)s.solar_irradiance[2]};
    b.ar_update = (float)(s.sun_angular_radius_deg * deg_to_rad);
    const double alt = r - R > 0.0 ? r - R : 0.0;
    b.sky_alt = (float)alt;
    const double* v = s.sun_local;
    double len = sqrt((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]);
    double y = v[1] / (len > 1.0e-12 ? len : 1.0e-12);
    b.sun_zc = (float)(y < -1.0 ? -1.0 : (y > 1.0 ? 1.0 : y));
    b.cam_alt = (float)alt;
    b.sun = f3{(float)v[0], (float)v[1], (float)v[2]};
    b.ar = (float)(s.sun_angular_radius_deg * deg_to_rad);
    b.right = f3{(float)s.right[0], (float)s.right[1], (float)s.right[2]};
    b.up = f3{(float)s.view_up[0], (float)s.view_up[1], (float)s.view_up[2]};
    b.forward = f3{(float)s.forward[0], (float)s.forward[1], (float)s.forward[2]};
    b.thf = (float)s.tangent_half_fov;
    return b;
}

void batch_error(AtmoBatch* b, const char* msg) { snprintf(b->error, sizeof b->error, "%s", msg); }

}  // namespace

extern "C" {

int32_t atmo_batch_create(const AtmoBatchState* states, int32_t count, int32_t width, int32_t height,
                          float* const luts[4], AtmoBatch** out) {
    if (!out || !states || count <= 0 || count > 65535 || width <= 0 || height <= 0) {
        set_error(nullptr, "atmo_batch_create: bad arguments");
        return ATMO_EINVAL;
    }
    *out = nullptr;
    BatchState* host = new (std::nothrow) BatchState[count];
    if (!host) return ATMO_ENOMEM;
    Consts k0 = make_consts(states[0].params);
    for (int i = 0; i < count; ++i) {
        char msg[400];
        if (validate(&states[i].params, msg, sizeof msg) != ATMO_OK) {
            char full[512];
            snprintf(full, sizeof full, "state %d: %s", i, msg);
            set_error(nullptr, "%s", full);
            delete[] host;
            return ATMO_EINVAL;
        }
        const AtmoBatchState& s = states[i];
        if (!(isfinite(s.planet_radius_m) && s.planet_radius_m > 1.0 && isfinite(s.camera_radius_m) &&
              isfinite(s.tangent_half_fov) && s.tangent_half_fov > 0.0)) {
            char full[512];
            snprintf(full, sizeof full, "state %d: planet radius, camera radius and tangent_half_fov "
                     "must be finite and positive", i);
            set_error(nullptr, "%s", full);
            delete[] host;
            return ATMO_EINVAL;
        }
        host[i] = derive_state(s);
        if (!same_dims(host[i].k, k0)) {
            set_error(nullptr, "atmo_batch_create: every state must have the same LUT dimensions");
            delete[] host;
            return ATMO_ESHAPE;
        }
    }
    AtmoBatch* b = new (std::nothrow) AtmoBatch();
    if (!b) {
        delete[] host;
        return ATMO_ENOMEM;
    }
    memset(b, 0, sizeof *b);
    b->magic = kBatchMagic;
    b->count = count;
    b->width = width;
    b->height = height;
    b->k0 = k0;
    b->strides[0] = (long)k0.tw * k0.th * 3;
    b->strides[1] = (long)k0.mw * k0.mh * 3;
    b->strides[2] = (long)k0.sw * k0.sh * 3;
    b->strides[3] = (long)k0.aw * k0.ah * k0.ad * 3;
    bool ok = hipMalloc((void**)&b->states, sizeof(BatchState) * count) == hipSuccess &&
              hipMemcpy(b->states, host, sizeof(BatchState) * count, hipMemcpyHostToDevice) == hipSuccess;
    for (int l = 0; l < 4 && ok; ++l) {
        if (luts && luts[l]) {
            b->luts[l] = luts[l];
            b->borrowed[l] = true;
        } else {
            ok = hipMalloc((void**)&b->luts[l], sizeof(float) * b->strides[l] * count) == hipSuccess;
        }
    }
    delete[] host;
    if (!ok) {
        set_error(nullptr, "atmo_batch_create: device allocation failed");
        batch_free(b);
        delete b;
        return ATMO_ENOMEM;
    }
    *out = b;
    return ATMO_OK;
}

int32_t atmo_batch_destroy(AtmoBatch* b) {
    if (!batch_ok(b)) return ATMO_EHANDLE;
    (void)hipDeviceSynchronize();
    batch_free(b);
    b->magic = 0;
    delete b;
    return ATMO_OK;
}

const char* atmo_batch_last_error(const AtmoBatch* b) {
    if (!b) return g_create_error;
    return batch_ok(b) ? b->error : "invalid batch handle";
}

int32_t atmo_batch_run(AtmoBatch* b, void* stream) {
    if (!batch_ok(b)) return ATMO_EHANDLE;
    if (!b->captured) {
        const Consts& k = b->k0;
        const unsigned n = (unsigned)b->count;
        if (hipStreamCreateWithFlags(&b->capture, hipStreamNonBlocking) != hipSuccess ||
            hipStreamBeginCapture(b->capture, hipStreamCaptureModeThreadLocal) != hipSuccess) {
            batch_error(b, "atmo_batch_run: cannot start graph capture");
            return ATMO_EHIP;
        }
        hipStream_t cs = b->capture;
        hipLaunchKernelGGL(k_transmittance_batched<kGroupT>,
                           dim3(blocks_for((long)k.tw * k.th * kGroupT, 256), n), dim3(256), 0, cs,
                           b->states, b->luts[0], b->strides[0]);
        hipLaunchKernelGGL(k_multi_scattering_batched, dim3(blocks_for((long)k.mw * k.mh * 64, 256), n),
                           dim3(256), 0, cs, b->states, (const float*)b->luts[0], b->strides[0],
                           b->luts[1], b->strides[1]);
        hipLaunchKernelGGL(k_sky_view_batched<kGroupSV>,
                           dim3(blocks_for((long)k.sw * k.sh * kGroupSV, 256), n), dim3(256), 0, cs,
                           b->states, (const float*)b->luts[0], b->strides[0],
                           (const float*)b->luts[1], b->strides[1], b->luts[2], b->strides[2]);
        hipLaunchKernelGGL(k_aerial_batched, dim3(blocks_for((long)k.aw * k.ah * 64, 256), n), dim3(256),
                           0, cs, b->states, b->width, b->height, (const float*)b->luts[0],
                           b->strides[0], (const float*)b->luts[1], b->strides[1], b->luts[3],
                           b->strides[3]);
        hipError_t launch = hipGetLastError();
        hipError_t end = hipStreamEndCapture(cs, &b->graph);
        if (launch != hipSuccess || end != hipSuccess ||
            hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0) != hipSuccess) {
            batch_error(b, "atmo_batch_run: graph capture failed");
            return ATMO_EHIP;
        }
        b->captured = true;
    }
    if (hipGraphLaunch(b->exec, (hipStream_t)stream) != hipSuccess) {
        batch_error(b, "atmo_batch_run: graph launch failed");
        return ATMO_EHIP;
    }
    return ATMO_OK;
}

int32_t atmo_batch_get_buffer(AtmoBatch* b, int32_t which, void** dptr, int64_t shape[5]) {
    if (!batch_ok(b)) return ATMO_EHANDLE;
    if (!dptr || !shape || which < 0 || which > 3) {
        batch_error(b, "atmo_batch_get_buffer: bad arguments");
        return ATMO_EINVAL;
    }
    const Consts& k = b->k0;
    int64_t s[5] = {b->count, 0, 0, 0, 0};
    switch (which) {
        case 0: s[1] = k.tw; s[2] = k.th; s[3] = 3; break;
        case 1: s[1] = k.mw; s[2] = k.mh; s[3] = 3; break;
        case 2: s[1] = k.sw; s[2] = k.sh; s[3] = 3; break;
        default: s[1] = k.aw; s[2] = k.ah; s[3] = k.ad; s[4] = 3; break;
    }
    for (int d = 0; d < 5; ++d) shape[d] = s[d];
    *dptr = b->luts[which];
    return ATMO_OK;
}

}  // extern "C"
