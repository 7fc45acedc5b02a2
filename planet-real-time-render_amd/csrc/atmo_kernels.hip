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
__global__ void __launch_bounds__(256) k_transmittance(Consts k, float bottom,
                                                       float* __restrict__ tlut) {
    const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
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
__global__ void __launch_bounds__(256) k_multi_scattering(Consts k, float bottom, f3 irradiance,
                                                          float ar,
                                                          const float* __restrict__ tlut,
                                                          float* __restrict__ mslut) {
    const long wave = ((long)blockIdx.x * blockDim.x + threadIdx.x) / 64;
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
__global__ void __launch_bounds__(256) k_sky_view(Consts k, float bottom, float cam_alt,
                                                  float sun_zc, f3 irradiance, float ar,
                                                  const float* __restrict__ tlut,
                                                  const float* __restrict__ mslut,
                                                  float* __restrict__ svlut) {
    const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
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
__global__ void __launch_bounds__(256) k_aerial(Consts k, int width, int height, float bottom,
                                                float cam_alt, f3 sun, f3 irradiance, float ar,
                                                f3 right, f3 up, f3 forward, float thf,
                                                const float* __restrict__ tlut,
                                                const float* __restrict__ mslut,
                                                float* __restrict__ ap) {
    const long wave = ((long)blockIdx.x * blockDim.x + threadIdx.x) / 64;
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

struct CompositeArgs {
    const float* scene;
    const float* space;
    const int32_t* sid;
    const float* pos;
    float* hdr;
    const float* tlut;
    const float* mslut;
    const float* svlut;
    const float* aplut;
    int width, height;
    float bottom, cam_alt, sky_alt;
    f3 sun, irradiance, disk;
    float ar;
    f3 right, up, forward;
    float thf, c0, c1;
    int view;
};

__device__ __forceinline__ f3 pixel_ray(const CompositeArgs& a, int px, int py, float& su,
                                        float& sv) {
    float aspect = (float)a.width / (float)a.height;
    su = ((float)px + 0.5f) / (float)a.width;
    sv = ((float)py + 0.5f) / (float)a.height;
    float sx = (su * 2.0f - 1.0f) * aspect;
    float sy = sv * 2.0f - 1.0f;
    return normalized((a.right * sx + a.up * sy) + a.forward / a.thf);
}

// The COMPOSITE view of _composite (renderer.py:1860-1940).  One thread
// per pixel.
__global__ void __launch_bounds__(256) k_composite(Consts k, CompositeArgs a) {
    const long npix = (long)a.width * a.height;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npix) return;
    const int px = (int)(i / a.height), py = (int)(i % a.height);
    float su, sv;
    f3 ray = pixel_ray(a, px, py, su, sv);
    f3 color = ld3(a.scene, i);
    if (a.sid[i] >= 0) {
        float sd = norm(ld3(a.pos, i));
        f3 scat = surface_inscatter(k, a.tlut, a.mslut, a.cam_alt, ray, sd, a.sun, a.irradiance,
                                    a.bottom, a.ar);
        f3 trans = surface_transmittance(k, a.tlut, a.cam_alt, ray, sd, a.bottom);
        color = scat + color * trans;
    } else {
        f3 trans = camera_transmittance(k, a.tlut, a.cam_alt, ray, a.bottom);
        f3 sky = sample_sky_view(k, a.svlut, ray, a.sun, a.sky_alt, a.bottom);
        f3 star = ld3(a.space, i) * trans;
        const f3 lum = mk(0.2126f, 0.7152f, 0.0722f);
        float contrast = dot(star, lum) / fmaxf(dot(sky, lum), 1.0e-5f);
        float vis = clampf((contrast - a.c0) / fmaxf(a.c1 - a.c0, 1.0e-5f), 0.0f, 1.0f);
        vis = vis * vis * (3.0f - 2.0f * vis);
        color = (sky + color * trans) + star * vis;
        float angular_distance = cr_acos(clampf(dot(ray, a.sun), -1.0f, 1.0f));
        float pixel_filter = cr_atan2(a.thf * 1.41421356f / (float)a.height, 1.0f);
        float cov = clampf((a.ar + pixel_filter - angular_distance) / fmaxf(2.0f * pixel_filter, 1.0e-8f),
                           0.0f, 1.0f);
        cov = cov * cov * (3.0f - 2.0f * cov);
        color = color + (a.disk * trans) * cov;
    }
    st3(a.hdr, i, maxs(color, 0.0f));
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
    if (view == ATMO_VIEW_COMPOSITE)
        hipLaunchKernelGGL(k_composite, dim3(blocks_for(npix, 256)), dim3(256), 0,
                           (hipStream_t)stream, h->k, a);
    else
        hipLaunchKernelGGL(k_composite_diagnostic, dim3(blocks_for(npix, 256)), dim3(256), 0,
                           (hipStream_t)stream, h->k, a);
    return atmo_launch_status(h, "atmo_composite");
}

}  // extern "C"
