// atmo_stages.h -- the atmosphere kernels and their launchers, compiled
// once per precision variant (see atmo_device.h): atmo_kernels.hip includes
// it for atmo::ref, atmo_robust.hip (ATMO_ROBUST 1) for atmo::robust.
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
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/atmo.h"

// Waves per SIMD the COMPOSITE kernel is compiled for.
#ifndef ATMO_COMPOSITE_WAVES
#define ATMO_COMPOSITE_WAVES 4
#endif
#include "atmo_device.h"
#include "atmo_handle.h"

namespace atmo {
inline namespace ATMO_PRECISION {

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
    float u = cdiv((float)x + 0.5f, k.tw_f);
    float v = cdiv((float)y + 0.5f, k.th_f);
    float alt, c, distance;
    transmittance_uv_to_ray(k, u, v, bottom, alt, c, distance);
    float sine = csqrt(fmaxf(1.0f - c * c, 0.0f));
    f3 origin = mk(0.0f, bottom + alt, 0.0f);
    f3 ray = mk(sine, c, 0.0f);
    Warp w = make_warp(-dot(origin, ray), 0.0f, distance);
    const int steps = k.t_steps;
    const float fsteps = (float)steps, rsteps = 1.0f / fsteps;
    const int per = (steps + G - 1) / G;
    const int s0 = min(lane * per, steps), s1 = min(s0 + per, steps);
    f3 od = mk(0.0f, 0.0f, 0.0f);
#if ATMO_ROBUST
    f3 carry = mk(0.0f, 0.0f, 0.0f);  // Kahan compensation of the lane's chunk
#endif
    // each step's start is the previous step's end: the same boundary() value
    float d0 = boundary(w, quot((float)s0, fsteps, rsteps));
    for (int s = s0; s < s1; ++s) {
        float d1 = boundary(w, quot((float)(s + 1), fsteps, rsteps));
        float step = d1 - d0;
        float dist = (d0 + d1) * 0.5f;
        d0 = d1;
#if ATMO_ROBUST
        // |origin + ray d| - R cancels at planet scale (and the float32
        // origin radius is itself rounded): the float64 altitude on the ray
        float h = altitude_on_ray(alt, c, dist, bottom);
        f3 y = extinction_of(k, density(k, h)) * step - carry;
        f3 t = od + y;
        carry = (t - od) - y;
        od = t;
#else
        float h = norm(origin + ray * dist) - bottom;
        od = od + extinction_of(k, density(k, h)) * step;
#endif
    }
#if ATMO_ROBUST
    // compensated tree over the lanes: each level adds the partner's sum
    // with its rounding error (TwoSum) folded into the carried compensation
    od = od - carry;
    carry = mk(0.0f, 0.0f, 0.0f);
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) {
        f3 o = mk(__shfl_xor(od.x, off, G), __shfl_xor(od.y, off, G), __shfl_xor(od.z, off, G));
        f3 oc = mk(__shfl_xor(carry.x, off, G), __shfl_xor(carry.y, off, G),
                   __shfl_xor(carry.z, off, G));
        f3 t = od + o;
        f3 bp = t - od;
        f3 err = (od - (t - bp)) + (o - bp);
        carry = (carry + oc) + err;
        od = t;
    }
    od = od + carry;
#else
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) {
        od.x += __shfl_xor(od.x, off, G);
        od.y += __shfl_xor(od.y, off, G);
        od.z += __shfl_xor(od.z, off, G);
    }
#endif
    if (active && lane == 0) st3(tlut, texel, clamp3(exp3(-mins(od, 80.0f)), 0.0f, 1.0f));
}

// Ordered combination of consecutive (L, F, T) segments of one
// multi-scattering direction within a group of G lanes (see scan_segments):
// afterwards lane i holds segments 0..i.
template <int G>
__device__ __forceinline__ void scan_ms_segments(f3& L, f3& F, f3& T, int lane) {
#pragma unroll
    for (int off = 1; off < G; off <<= 1) {
        f3 pl = mk(__shfl_up(L.x, off, G), __shfl_up(L.y, off, G), __shfl_up(L.z, off, G));
        f3 pf = mk(__shfl_up(F.x, off, G), __shfl_up(F.y, off, G), __shfl_up(F.z, off, G));
        f3 pt = mk(__shfl_up(T.x, off, G), __shfl_up(T.y, off, G), __shfl_up(T.z, off, G));
        if (lane >= off) {
            L = pl + pt * L;
            F = pf + pt * F;
            T = pt * T;
        }
    }
}

// _build_multi_scattering (renderer.py:1065-1206).  One 256-thread block per
// texel: 64 Fibonacci-sphere directions at a time, kMsLanes lanes per
// direction, each lane marching a contiguous chunk of the direction's steps;
// the chunks combine by an ordered (L, F, T) scan, the group's last lane adds
// the ground bounce with the direction's full transmittance, and the
// direction sums reduce across the block.
constexpr int kMsLanes = 4;

__device__ __forceinline__ void multi_scattering_body(const Consts& k, float bottom, f3 irradiance,
                                                      float ar, const float* __restrict__ tlut,
                                                      float* __restrict__ mslut, long texel) {
    __shared__ float partial[4][6];
    const int tid = threadIdx.x;
    const int chunk = tid % kMsLanes;
    const int x = (int)(texel / k.mh), y = (int)(texel % k.mh);
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
    const int per = (S + kMsLanes - 1) / kMsLanes;
    const int s0 = min(chunk * per, S), s1 = min(s0 + per, S);
    f3 lum = mk(0.0f, 0.0f, 0.0f), fb = mk(0.0f, 0.0f, 0.0f);
    for (int base = 0; base < N; base += 256 / kMsLanes) {
        const int di = base + tid / kMsLanes;
        f3 dl = mk(0.0f, 0.0f, 0.0f), df = mk(0.0f, 0.0f, 0.0f), pt = mk(1.0f, 1.0f, 1.0f);
        bool valid = false, ground = false;
        float end = 0.0f;
        f3 ray = mk(0.0f, 1.0f, 0.0f);
        if (di < N) {
            float si = (float)di + 0.5f;
            float dc = 1.0f - 2.0f * si / (float)N;
            float ds = sqrtf(fmaxf(1.0f - dc * dc, 0.0f));
            float az = (float)di * 2.39996323f;
            float saz, caz;
            cr_sincos(az, saz, caz);
            ray = mk(ds * caz, dc, ds * saz);
            float start;
            segment_vec(k, origin, ray, bottom, start, end, valid, ground);
            if (valid && s0 < s1) {
                Warp w = make_warp(-dot(origin, ray), start, end);
                RayPhase ph = ray_phase(k, ray, sun);
                const float fS = (float)S, rS = 1.0f / fS;
                float d0 = boundary(w, quot((float)s0, fS, rS));
                for (int s = s0; s < s1; ++s) {
                    float d1 = boundary(w, quot((float)(s + 1), fS, rS));
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
            }
        }
        scan_ms_segments<kMsLanes>(dl, df, pt, chunk);
        if (chunk == kMsLanes - 1 && di < N && valid) {
            if (ground) {
                f3 gp = origin + ray * end;
                f3 gn = normalized(gp);
                float gsc = dot(gn, sun);
                SunGeom g = sun_geometry(k, 0.0f, gsc, bottom, ar);
                f3 gt = solar_transmittance(k, tlut, g, 0.0f, bottom, ar);
                f3 gl = (((arr3(k.albedo) * irradiance) * gt) * fmaxf(gsc, 0.0f)) / k.pi;
                dl = dl + pt * gl;
            }
            lum = lum + dl;
            fb = fb + df;
        }
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
    const int wave = tid / 64;
    if (tid % 64 == 0) {
        partial[wave][0] = lum.x;
        partial[wave][1] = lum.y;
        partial[wave][2] = lum.z;
        partial[wave][3] = fb.x;
        partial[wave][4] = fb.y;
        partial[wave][5] = fb.z;
    }
    __syncthreads();
    if (tid == 0) {
        f3 sl = mk(0.0f, 0.0f, 0.0f), sf = mk(0.0f, 0.0f, 0.0f);
        for (int w = 0; w < 4; ++w) {
            sl = sl + mk(partial[w][0], partial[w][1], partial[w][2]);
            sf = sf + mk(partial[w][3], partial[w][4], partial[w][5]);
        }
        f3 avg_l = sl / (float)N;
        f3 avg_f = clamp3(sf / (float)N, 0.0f, 0.95f);
        st3(mslut, texel, maxs(avg_l / (mk(1.0f, 1.0f, 1.0f) - avg_f), 0.0f));
    }
}

// _build_multi_scattering (renderer.py:1065-1206), one wave per texel: lane
// d marches direction d (then d + 64, ... when there are more than 64) through
// all of its steps in order, as the reference's loop does, adds the ground
// bounce, and the wave sums the directions by a butterfly.  For the batched
// build, whose states and texels give far more waves than the GPU holds, this
// spends no lanes on chunk scans or on repeating each direction's set-up in
// every chunk.  The values equal multi_scattering_body's up to float
// reassociation (the chunk scan and the order of the direction sum).
__device__ __forceinline__ void multi_scattering_wave(const Consts& k, float bottom, f3 irradiance,
                                                      float ar, const float* __restrict__ tlut,
                                                      float* __restrict__ mslut, long texel) {
    const int lane = threadIdx.x % 64;
    const int x = (int)(texel / k.mh), y = (int)(texel % k.mh);
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
    for (int base = 0; base < N; base += 64) {
        const int di = base + lane;
        if (di >= N) continue;
        float si = (float)di + 0.5f;
        float dc = 1.0f - 2.0f * si / (float)N;
        float ds = sqrtf(fmaxf(1.0f - dc * dc, 0.0f));
        float az = (float)di * 2.39996323f;
        float saz, caz;
        cr_sincos(az, saz, caz);
        f3 ray = mk(ds * caz, dc, ds * saz);
        float start, end;
        bool valid, ground;
        segment_vec(k, origin, ray, bottom, start, end, valid, ground);
        if (!valid) continue;
        f3 dl = mk(0.0f, 0.0f, 0.0f), df = mk(0.0f, 0.0f, 0.0f), pt = mk(1.0f, 1.0f, 1.0f);
        Warp w = make_warp(-dot(origin, ray), start, end);
        RayPhase ph = ray_phase(k, ray, sun);
        const float fS = (float)S, rS = 1.0f / fS;
        float d0 = boundary(w, 0.0f);
        for (int s = 0; s < S; ++s) {
            float d1 = boundary(w, quot((float)(s + 1), fS, rS));
            float dist = (d0 + d1) * 0.5f;
            float step = d1 - d0;
            f3 point = origin + ray * dist;
            float h = altitude_on_ray(h0, dc, dist, bottom);
            f3 d = density(k, h);
            f3 st, si3;
            segment_integral(extinction_of(k, d), step, st, si3);
            f3 direct = scattering_source<false>(k, tlut, nullptr, point, h, d, ph, sun, irradiance,
                                                 bottom, ar);
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
        st3(mslut, texel, maxs(avg_l / (mk(1.0f, 1.0f, 1.0f) - avg_f), 0.0f));
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
    float sun_s = csqrt(fmaxf(1.0f - sun_zc * sun_zc, 0.0f));
    f3 sun = mk(sun_s, sun_zc, 0.0f);
    float safe_alt = fmaxf(cam_alt, 0.0f);
    f3 origin = mk(0.0f, bottom + safe_alt, 0.0f);
    float az = cdiv((float)x + 0.5f, k.sw_f) * k.pi;
    float view_v = cdiv((float)y + 0.5f, k.sh_f);
#if ATMO_ROBUST
    const double zenith_w = sky_v_to_zenith_w(view_v, safe_alt, bottom);
    float vc = (float)cos(zenith_w);
    float vsn = (float)sin(zenith_w);
#else
    float zenith = sky_v_to_zenith(k, view_v, sky_horizon_zenith(k, safe_alt, bottom));
    float vc, vsn;
    cr_sincos(zenith, vsn, vc);
#endif
    float saz, caz;
    cr_sincos(az, saz, caz);
    f3 ray = mk(vsn * caz, vc, vsn * saz);
    float start, end;
    bool valid, ground;
    segment_from_altitude(k, safe_alt, vc, bottom, start, end, valid, ground);
    f3 L = mk(0.0f, 0.0f, 0.0f), T = mk(1.0f, 1.0f, 1.0f);
    if (valid) {
        Warp w = make_warp(-dot(origin, ray), start, end);
        RayPhase ph = ray_phase(k, ray, sun);
        const int steps = k.sv_steps;
        const float fsteps = (float)steps, rsteps = 1.0f / fsteps;
        const int per = (steps + G - 1) / G;
        const int s0 = min(lane * per, steps), s1 = min(s0 + per, steps);
        float d0 = boundary(w, quot((float)s0, fsteps, rsteps));
        for (int s = s0; s < s1; ++s) {
            float d1 = boundary(w, quot((float)(s + 1), fsteps, rsteps));
            float dist = (d0 + d1) * 0.5f;
            float step = d1 - d0;
            d0 = d1;
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
// ray.  P lanes per depth slice (P = 4 for up to 16 slices, 2 for up to 32,
// else 1), each marching a contiguous part of the slice's steps, so a
// shallow froxel volume still fills the wave; one ordered (L, T) scan over
// the wave composes the parts and slices, and the last part of each slice
// stores its cumulative value.  Slices beyond a wave are processed in
// chunks with the previous chunk's composition carried in.
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
    const int P = D <= 16 ? 4 : (D <= 32 ? 2 : 1);
    const int part = lane % P;
    const int per = (spp + P - 1) / P;
    const int s0 = min(part * per, spp), s1 = min(s0 + per, spp);
    f3 carry_l = mk(0.0f, 0.0f, 0.0f), carry_t = mk(1.0f, 1.0f, 1.0f);
    for (int base = 0; base < D; base += 64 / P) {
        const int di = base + lane / P;
        f3 L = mk(0.0f, 0.0f, 0.0f), T = mk(1.0f, 1.0f, 1.0f);
        if (valid && di < D) {
            float previous = di == 0 ? a_start
                                     : a_start + (a_end - a_start) * ((float)(di - 1) / (float)(D - 1));
            float target = a_start + (a_end - a_start) * ((float)di / (float)(D - 1));
            float is = fmaxf(previous, a_start);
            float ie = fminf(target, a_end);
            if (ie > is) {
                Warp w = make_warp(neg_od, is, ie);
                float d0 = boundary(w, (float)s0 / (float)spp);
                for (int s = s0; s < s1; ++s) {
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
        if (di < D && part == P - 1) st3(ap, wave * D + di, maxs(cl, 0.0f));
        carry_l = mk(__shfl(cl.x, 63, 64), __shfl(cl.y, 63, 64), __shfl(cl.z, 63, 64));
        carry_t = mk(__shfl(ct.x, 63, 64), __shfl(ct.y, 63, 64), __shfl(ct.z, 63, 64));
    }
}


// Single-state kernels: one Taichi kernel call each.
template <int G>
__global__ void __launch_bounds__(256) k_transmittance(Consts k, float bottom, float* __restrict__ tlut) {
    transmittance_body<G>(k, bottom, tlut, (long)blockIdx.x * blockDim.x + threadIdx.x);
}

// Three waves per SIMD at 1024 texels: 28.1 us, against 29.2 us at two and
// 38.5 us at four (128 VGPRs, a 196-byte spill inside the step loop).  One
// wave per texel, as the batch runs it, takes 42 us here: 1024 waves cannot
// fill 256 CUs, so the single state keeps four lanes per direction.
#ifndef ATMO_MS_WAVES
#define ATMO_MS_WAVES 3
#endif
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(ATMO_MS_WAVES, 8)))
k_multi_scattering(Consts k, float bottom, f3 irradiance,
                                                          float ar, const float* __restrict__ tlut,
                                                          float* __restrict__ mslut) {
    multi_scattering_body(k, bottom, irradiance, ar, tlut, mslut, (long)blockIdx.x);
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

#if !ATMO_ROBUST
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

// One wave per texel, four texels per block.  Four waves per SIMD (128
// VGPRs, a 180-byte spill, 3.57 ms for 256 states) measure 2.5% faster than
// three (168 VGPRs, a 32-byte spill, 3.66 ms); the four-lane layout of the
// single-state build took 5.03 ms.
#ifndef ATMO_MS_BATCH_WAVES
#define ATMO_MS_BATCH_WAVES 4
#endif
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(ATMO_MS_BATCH_WAVES, 8)))
k_multi_scattering_batched(const BatchState* __restrict__ st, const float* __restrict__ tluts,
                           long t_stride, float* __restrict__ msluts, long m_stride) {
    const BatchState& b = st[blockIdx.y];
    const long texel = (long)blockIdx.x * 4 + threadIdx.x / 64;
    if (texel >= (long)b.k.mw * b.k.mh) return;  // wave-uniform
    multi_scattering_wave(b.k, b.bottom, b.irradiance, b.ar_update, tluts + blockIdx.y * t_stride,
                          msluts + blockIdx.y * m_stride, texel);
}
constexpr int kMsBatchTexelsPerBlock = 4;

template <int G>
// Four waves per SIMD (128 VGPRs, a 36-byte spill): 2% faster on the
// 256-state batch than three at 146 VGPRs.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 8)))
k_sky_view_batched(const BatchState* __restrict__ st,
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
#endif  // !ATMO_ROBUST

// A pixel's three floats in a block's LDS tile of 256 pixels.
__device__ __forceinline__ void tile_put(float* tile, uint32_t l, f3 v) {
    tile[3 * l + 0] = v.x;
    tile[3 * l + 1] = v.y;
    tile[3 * l + 2] = v.z;
}
__device__ __forceinline__ f3 tile_get(const float* tile, uint32_t l) {
    return f3{tile[3 * l], tile[3 * l + 1], tile[3 * l + 2]};
}

// _prepare_surface_lighting (renderer.py:1422-1472).  One block per 256
// pixels.  The surface pixels are compacted onto the leading threads, so
// every lane of a lighting wave lights a surface pixel (with a per-pixel sky
// mask, a sky lane would otherwise idle through the whole branch); sky and
// out-of-frame lanes only hold the constants.  The outputs go through two
// LDS tiles in pixel order, so each thread then stores its own pixel.
__global__ void __launch_bounds__(256) k_prepare(Consts k, long npix,
                                                 const float* __restrict__ position_view,
                                                 const int32_t* __restrict__ surface_id,
                                                 float bottom, float cam_alt, f3 sun, float ar,
                                                 f3 right, f3 up, f3 forward,
                                                 const float* __restrict__ tlut,
                                                 const float* __restrict__ mslut,
                                                 float* __restrict__ sun_t,
                                                 float* __restrict__ sky_l) {
    __shared__ int list[256];
    __shared__ int wave_counts[4];
    __shared__ float tile_t[256 * 3], tile_l[256 * 3];
    const uint32_t base = blockIdx.x * 256u;
    const uint32_t i = base + threadIdx.x;  // npix <= 2^28
    const bool live = i < npix;
    tile_put(tile_t, threadIdx.x, mk(1.0f, 1.0f, 1.0f));
    tile_put(tile_l, threadIdx.x, mk(0.0f, 0.0f, 0.0f));
    // (the compaction's barriers order these stores before the surface ones)
    const int count = compact_block(live && surface_id[i] >= 0, list, wave_counts);
    if ((int)threadIdx.x < count) {
        const uint32_t l = list[threadIdx.x];
        const uint32_t p = base + l;
        f3 camera = mk(0.0f, bottom + cam_alt, 0.0f);
        f3 vp = pix3(position_view, p);
        f3 local = (right * vp.x + up * vp.y) + forward * vp.z;
        f3 planet = camera + local;
        f3 radial = div3(planet, fmaxf(norm(planet), 1.0f));
        float sd = norm(local);
        f3 local_ray = div3(local, fmaxf(sd, 1.0f));
        float sc = dot(radial, sun);
#if ATMO_ROBUST
        if (cam_alt > 1.0e6f) {
            // float64 island: camera-relative endpoint above 1000 km, where
            // the float32 camera radius has a metre-sized ulp
            const double lx = ((double)right.x * vp.x + (double)up.x * vp.y) + (double)forward.x * vp.z;
            const double ly = ((double)right.y * vp.x + (double)up.y * vp.y) + (double)forward.y * vp.z;
            const double lz = ((double)right.z * vp.x + (double)up.z * vp.y) + (double)forward.z * vp.z;
            const double py = ((double)bottom + cam_alt) + ly;
            const double pn = fmax(sqrt((lx * lx + py * py) + lz * lz), 1.0);
            sc = (float)(((lx * sun.x + py * sun.y) + lz * sun.z) / pn);
            const double ld = sqrt((lx * lx + ly * ly) + lz * lz);
            sd = (float)ld;
            local_ray.y = (float)(ly / fmax(ld, 1.0));
        }
#endif
        float h = altitude_on_ray(cam_alt, local_ray.y, sd, bottom);
        SunGeom g = sun_geometry(k, h, sc, bottom, ar);
        tile_put(tile_t, l, solar_transmittance(k, tlut, g, h, bottom, ar));
        tile_put(tile_l, l, maxs(sample_multi_scattering(k, mslut, h, sc, bottom), 0.0f));
    }
    __syncthreads();
    if (live) {
        put3(sun_t, i, tile_get(tile_t, threadIdx.x));
        put3(sky_l, i, tile_get(tile_l, threadIdx.x));
    }
}


__device__ __forceinline__ f3 pixel_ray(const CompositeArgs& a, int px, int py, float& su,
                                        float& sv) {
    return camera_ray(px, py, a.width, a.height, a.aspect, a.right, a.up, a.fwd_t, su, sv);
}
__device__ __forceinline__ f3 pixel_ray(const CompositeArgs& a, uint32_t i) {
    float su, sv;
    const uint32_t px = i / (uint32_t)a.height;
    return pixel_ray(a, (int)px, (int)(i - px * (uint32_t)a.height), su, sv);
}

// Sky branch of one COMPOSITE pixel (renderer.py:1882-1940), clamped.
__device__ __forceinline__ f3 composite_sky_pixel(const Consts& k, const CompositeArgs& a,
                                                  uint32_t i) {
    f3 ray = pixel_ray(a, i);
    f3 c = composite_sky(k, a.tlut, a.svlut, ray, pix3(a.scene, i), pix3(a.space, i), a.cam_alt,
                         a.sky_alt, a.sun, a.disk, a.ar, a.pixel_filter, a.sky_horizon, a.c0, a.c1,
                         a.bottom);
    return maxs(c, 0.0f);
}

// The block's 256 output pixels leave through an LDS tile, written to HBM
// by whichever of the block's four waves finishes last (an LDS counter
// decides).  Sky and surface pixels share cache lines, and their threads
// finish tens of microseconds apart: written directly, a line evicted
// between the two writes was written back twice.  Waves that finish early
// still exit at once: there is no block-wide barrier.
__device__ __forceinline__ void store_tile_by_last_wave(const float* tile, int* waves_done,
                                                        float* __restrict__ hdr, uint32_t base,
                                                        uint32_t npix) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");  // this wave's tile writes
    const int lane = threadIdx.x % 64;
    int before = 0;
    if (lane == 0) before = atomicAdd(waves_done, 1);
    before = __shfl(before, 0, 64);
    if (before != 3) return;  // wave-uniform: not the last of the four waves
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");  // the other waves' tile writes
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const uint32_t l = lane + 64 * r;
        if (base + l < npix) put3(hdr, base + l, f3{tile[3 * l], tile[3 * l + 1], tile[3 * l + 2]});
    }
}

// The COMPOSITE view of _composite (renderer.py:1860-1940), one block per
// 256 pixels.  Surface pixels are compacted onto the leading threads and sky
// pixels onto the threads after them, so waves run full (the sky branch in
// at most two waves of a block instead of all four); surface threads run
// the coarse-only limb integral.  A surface pixel with an interval whose
// terminator-refinement weight is positive is deferred to k_composite_list,
// which keeps the refinement's registers out of this kernel; every pixel is
// written by exactly one thread.  Four waves per SIMD (128 VGPRs) is the
// measured optimum at 1080p.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(ATMO_COMPOSITE_WAVES, 8)))
k_composite(Consts k, CompositeArgs a) {
    __shared__ int list[256];
    __shared__ int wave_counts[8];
    __shared__ float4 stash[256];
    __shared__ float tile[256 * 3];
    __shared__ int waves_done;
    const uint32_t npix = (uint32_t)a.width * (uint32_t)a.height;
    const uint32_t base = blockIdx.x * 256u;
    const uint32_t i = base + threadIdx.x;
    const bool live = i < npix;
    const bool surface = live && a.sid[i] >= 0;
    if (threadIdx.x == 0) waves_done = 0;  // ordered by the compaction's barriers
    int nsky;
    const int count = compact_block2(surface, live && !surface, list, wave_counts, nsky);
    if ((int)threadIdx.x >= count && (int)threadIdx.x < count + nsky) {
        const uint32_t l = list[threadIdx.x];
        tile_put(tile, l, composite_sky_pixel(k, a, base + l));
    }
    bool refine = false;
    uint32_t j = 0;
    // the wave's first thread index, wave-uniform (an SGPR)
    const uint32_t wave_base = __builtin_amdgcn_readfirstlane(threadIdx.x & ~63u);
    if ((int)threadIdx.x < count) {
        {
            const uint32_t p = base + list[threadIdx.x];
            const f3 ray = pixel_ray(a, p);
            const float sd = norm(pix3(a.pos, p));
            stash[threadIdx.x] = make_float4(ray.x, ray.y, ray.z, sd);
        }
        f3 scat;
        bool done;
        {
            const float4 r = stash[threadIdx.x];
            done = surface_inscatter_t<false>(k, a.tlut, a.mslut, a.cam_alt, mk(r.x, r.y, r.z), r.w,
                                              a.sun, a.irradiance, a.bottom, a.ar, scat);
        }
        // The pixel, its ray and its distance come back from LDS after the
        // interval loop instead of staying live in registers across it (at
        // four waves per SIMD they would spill to scratch, i.e. to HBM).
        // The empty asm stops the compiler from forwarding the values, and
        // the thread index is rebuilt from the wave base and v_mbcnt so that
        // not even the LDS address stays live.
        asm volatile("" ::: "memory");
        const uint32_t t = wave_base + __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
        const uint32_t l = list[t];
        j = base + l;
        const float4 r = stash[t];
        const f3 ray = mk(r.x, r.y, r.z);
        f3 c = mk(0.0f, 0.0f, 0.0f);  // a deferred pixel: k_composite_list writes it later
        if (done) {
            f3 trans = surface_transmittance(k, a.tlut, a.cam_alt, ray, r.w, a.bottom);
            c = maxs(scat + pix3(a.scene, j) * trans, 0.0f);
        } else {
            refine = true;
        }
        tile_put(tile, l, c);
    }
    enqueue_wave(refine, j, a.queue, a.parity, a.queue_capacity);
    store_tile_by_last_wave(tile, &waves_done, a.hdr, base, npix);
}

// The surrogate composite's queued pixels (out of the network's domain, or
// routed to the plain integrator) through the coarse-only limb integral of
// k_composite, one thread per queue entry at four waves per SIMD.  A
// finished pixel is written; a pixel that needs the terminator refinement
// is appended to the refinement queue `a.refine` (one atomic per wave that
// has any), which k_composite_list works through after this kernel.  No
// block-wide barrier: a wave exits as soon as its pixels are done.  The
// grid covers the largest possible queue; blocks past the queue's count
// exit at once.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(ATMO_COMPOSITE_WAVES, 8)))
k_composite_queue_coarse(Consts k, CompositeArgs a) {
    __shared__ float4 stash[256];
    const long count = min((long)a.queue[a.parity], a.queue_capacity);
    const uint32_t base = blockIdx.x * 256u;
    if ((long)base + threadIdx.x >= count) return;
    // the wave's first thread index, wave-uniform (an SGPR)
    const uint32_t wave_base = __builtin_amdgcn_readfirstlane(threadIdx.x & ~63u);
    {
        const uint32_t p = (uint32_t)a.queue[2 + base + threadIdx.x];
        const f3 ray = pixel_ray(a, p);
        const float sd = norm(pix3(a.pos, p));
        stash[threadIdx.x] = make_float4(ray.x, ray.y, ray.z, sd);
    }
    f3 scat;
    bool done;
    {
        const float4 r = stash[threadIdx.x];
        done = surface_inscatter_t<false>(k, a.tlut, a.mslut, a.cam_alt, mk(r.x, r.y, r.z), r.w,
                                          a.sun, a.irradiance, a.bottom, a.ar, scat);
    }
    // as in k_composite: the entry, ray and distance come back after the
    // interval loop (from the queue and LDS) instead of staying live in
    // registers across it
    asm volatile("" ::: "memory");
    const uint32_t t = wave_base + __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
    const uint32_t j = (uint32_t)a.queue[2 + base + t];
    if (done) {
        const float4 r = stash[t];
        const f3 trans = surface_transmittance(k, a.tlut, a.cam_alt, mk(r.x, r.y, r.z), r.w,
                                               a.bottom);
        put3(a.hdr, j, maxs(scat + pix3(a.scene, j) * trans, 0.0f));
    }
    enqueue_wave(!done, j, a.refine, a.parity, a.queue_capacity);
}

// The surface pixels deferred to a.list (a.queue when null): the full
// surface branch of the COMPOSITE view, refinement included (the pixels
// k_composite or k_composite_queue_coarse could not finish).  G lanes per entry (surface_inscatter_lanes): one base interval
// per lane, then the refined intervals' sub-segments dealt out over the
// group, so a pixel costs one coarse segment and a few rounds of
// sub-segments per lane instead of every segment in one thread, and a queue
// of a few thousand pixels still fills the GPU.
// Block 0 also zeroes the other counter of both queues for the next call.
// Three waves per SIMD (168 VGPRs, 64 bytes of spill): on a twilight
// 960x540 frame with 40 K refined pixels the composite takes 176 us, against
// 191 us at two waves (189 VGPRs) and 219 us at four (a 220-byte spill).
#ifndef ATMO_LIST_WAVES
#define ATMO_LIST_WAVES 3
#endif
template <int G>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(ATMO_LIST_WAVES, 8)))
k_composite_list(Consts k, CompositeArgs a) {
    __shared__ float4 stage_a[256];
    __shared__ float2 stage_b[256];
    const int* const list = a.list ? a.list : a.queue;
    const int count = (int)min((long)list[a.parity], a.queue_capacity);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        a.queue[1 - a.parity] = 0;
        if (a.refine) a.refine[1 - a.parity] = 0;
    }
    const int sub = (int)threadIdx.x % G;
    const int group_first = (int)threadIdx.x - sub;
    const long stride = (long)gridDim.x * (blockDim.x / G);
    for (long e = ((long)blockIdx.x * blockDim.x + threadIdx.x) / G; e < count; e += stride) {
        const uint32_t j = (uint32_t)list[2 + e];
        f3 ray = pixel_ray(a, j);
        float sd = norm(pix3(a.pos, j));
        f3 scat = surface_inscatter_lanes<G>(k, a.tlut, a.mslut, a.cam_alt, ray, sd, a.sun,
                                             a.irradiance, a.bottom, a.ar, sub,
                                             stage_a + group_first, stage_b + group_first);
        if (sub == 0) {
            f3 trans = surface_transmittance(k, a.tlut, a.cam_alt, ray, sd, a.bottom);
            put3(a.hdr, j, maxs(scat + pix3(a.scene, j) * trans, 0.0f));
        }
    }
}

// Diagnostic views 1..10 of _composite (renderer.py:1941-2055).
__global__ void __launch_bounds__(256) k_composite_diagnostic(Consts k, CompositeArgs a) {
    const uint32_t npix = (uint32_t)a.width * (uint32_t)a.height;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npix) return;
    const int px = (int)(i / (uint32_t)a.height), py = (int)(i % (uint32_t)a.height);
    float su, sv;
    f3 ray = pixel_ray(a, px, py, su, sv);
    const bool surface = a.sid[i] >= 0;
    f3 color = mk(0.0f, 0.0f, 0.0f);
    switch (a.view) {
        case ATMO_VIEW_SKY_VIEW:
            color = sample_sky_view(k, a.svlut, ray, a.sun, a.sky_alt, a.bottom, a.sky_horizon);
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
                color = surface_inscatter(k, a.tlut, a.mslut, a.cam_alt, ray, norm(pix3(a.pos, i)),
                                          a.sun, a.irradiance, a.bottom, a.ar);
            break;
        case ATMO_VIEW_AERIAL_FROXEL_ONLY:
            if (surface)
                color = sample_aerial(k, a.aplut, su, sv, norm(pix3(a.pos, i)), a.cam_alt, ray,
                                      a.bottom);
            break;
        case ATMO_VIEW_AERIAL_BLEND_WEIGHT: {
            float w = surface ? 1.0f : 0.0f;
            color = mk(w, w, w);
            break;
        }
        case ATMO_VIEW_AERIAL_TRANSMITTANCE:
            if (surface)
                color = surface_transmittance(k, a.tlut, a.cam_alt, ray, norm(pix3(a.pos, i)),
                                              a.bottom);
            break;
        default: {
            float m = surface ? 1.0f : 0.0f;
            color = mk(m, m, m);
            break;
        }
    }
    put3(a.hdr, i, maxs(color, 0.0f));
}


// ------------------------------------------------------------ launchers

inline unsigned grid_for(long threads) { return (unsigned)((threads + 255) / 256); }

// Lanes per texel of the transmittance and sky-view builds: a power of 2
// up to 16 (1, 2, 4, 8 or 16), chosen per handle or batch by the autotuner
// (default 8).  A single state's few thousand texels need several lanes each
// to fill the GPU; a batch of states fills it with one.
inline bool valid_group(int g) { return g == 1 || g == 2 || g == 4 || g == 8 || g == 16; }

template <int G>
void launch_transmittance_g(const Consts& k, float bottom, float* tlut, hipStream_t s) {
    hipLaunchKernelGGL(k_transmittance<G>, dim3(grid_for((long)k.tw * k.th * G)), dim3(256), 0, s,
                       k, bottom, tlut);
}

void launch_transmittance(const Consts& k, float bottom, float* tlut, hipStream_t s, int g) {
    switch (g) {
        case 1: launch_transmittance_g<1>(k, bottom, tlut, s); break;
        case 2: launch_transmittance_g<2>(k, bottom, tlut, s); break;
        case 4: launch_transmittance_g<4>(k, bottom, tlut, s); break;
        case 16: launch_transmittance_g<16>(k, bottom, tlut, s); break;
        default: launch_transmittance_g<8>(k, bottom, tlut, s); break;
    }
}

void launch_multi_scattering(const Consts& k, float bottom, f3 irradiance, float ar,
                             const float* tlut, float* mslut, hipStream_t s) {
    hipLaunchKernelGGL(k_multi_scattering, dim3((unsigned)(k.mw * k.mh)), dim3(256), 0, s, k, bottom,
                       irradiance, ar, tlut, mslut);
}

template <int G>
void launch_sky_view_g(const Consts& k, float bottom, float cam_alt, float sun_zc, f3 irradiance,
                       float ar, const float* tlut, const float* mslut, float* svlut,
                       hipStream_t s) {
    hipLaunchKernelGGL(k_sky_view<G>, dim3(grid_for((long)k.sw * k.sh * G)), dim3(256), 0, s, k,
                       bottom, cam_alt, sun_zc, irradiance, ar, tlut, mslut, svlut);
}

void launch_sky_view(const Consts& k, float bottom, float cam_alt, float sun_zc, f3 irradiance,
                     float ar, const float* tlut, const float* mslut, float* svlut, hipStream_t s,
                     int g) {
    switch (g) {
        case 1: launch_sky_view_g<1>(k, bottom, cam_alt, sun_zc, irradiance, ar, tlut, mslut, svlut, s); break;
        case 2: launch_sky_view_g<2>(k, bottom, cam_alt, sun_zc, irradiance, ar, tlut, mslut, svlut, s); break;
        case 4: launch_sky_view_g<4>(k, bottom, cam_alt, sun_zc, irradiance, ar, tlut, mslut, svlut, s); break;
        case 16: launch_sky_view_g<16>(k, bottom, cam_alt, sun_zc, irradiance, ar, tlut, mslut, svlut, s); break;
        default: launch_sky_view_g<8>(k, bottom, cam_alt, sun_zc, irradiance, ar, tlut, mslut, svlut, s); break;
    }
}

void launch_aerial(const Consts& k, int width, int height, float bottom, float cam_alt, f3 sun,
                   f3 irradiance, float ar, f3 right, f3 up, f3 forward, float thf,
                   const float* tlut, const float* mslut, float* ap, hipStream_t s) {
    hipLaunchKernelGGL(k_aerial, dim3(grid_for((long)k.aw * k.ah * 64)), dim3(256), 0, s, k, width,
                       height, bottom, cam_alt, sun, irradiance, ar, right, up, forward, thf, tlut,
                       mslut, ap);
}

void launch_prepare(const Consts& k, long npix, const float* position_view,
                    const int32_t* surface_id, float bottom, float cam_alt, f3 sun, float ar,
                    f3 right, f3 up, f3 forward, const float* tlut, const float* mslut,
                    float* sun_t, float* sky_l, hipStream_t s) {
    hipLaunchKernelGGL(k_prepare, dim3(grid_for(npix)), dim3(256), 0, s, k, npix, position_view,
                       surface_id, bottom, cam_alt, sun, ar, right, up, forward, tlut, mslut, sun_t,
                       sky_l);
}

#ifndef ATMO_LIST_BLOCKS
#define ATMO_LIST_BLOCKS 2048
#endif
template <int G>
void launch_composite_list_g(const Consts& k, const CompositeArgs& a, hipStream_t s) {
    // one group per entry up to ATMO_LIST_BLOCKS blocks, then a grid-stride loop
    long blocks = (a.queue_capacity * G + 255) / 256;
    if (blocks > ATMO_LIST_BLOCKS) blocks = ATMO_LIST_BLOCKS;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_composite_list<G>, dim3((unsigned)blocks), dim3(256), 0, s, k, a);
}

void launch_composite_list(const Consts& k, const CompositeArgs& a, hipStream_t s) {
    // the smallest power-of-two group that holds hz_steps (4..64) intervals
    if (k.hz_steps <= 4) launch_composite_list_g<4>(k, a, s);
    else if (k.hz_steps <= 8) launch_composite_list_g<8>(k, a, s);
    else if (k.hz_steps <= 16) launch_composite_list_g<16>(k, a, s);
    else if (k.hz_steps <= 32) launch_composite_list_g<32>(k, a, s);
    else launch_composite_list_g<64>(k, a, s);
}

void launch_composite_queue(const Consts& k, const CompositeArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(k_composite_queue_coarse, dim3(grid_for(a.queue_capacity)), dim3(256), 0, s,
                       k, a);
}

void launch_composite(const Consts& k, const CompositeArgs& a, hipStream_t s) {
    const long npix = (long)a.width * a.height;
    if (a.view == ATMO_VIEW_COMPOSITE) {
        hipLaunchKernelGGL(k_composite, dim3(grid_for(npix)), dim3(256), 0, s, k, a);
        launch_composite_list(k, a, s);
    } else {
        hipLaunchKernelGGL(k_composite_diagnostic, dim3(grid_for(npix)), dim3(256), 0, s, k, a);
    }
}

#if !ATMO_ROBUST
template <int GT, int GS>
void launch_batch_g(const Consts& k, const BatchState* states, unsigned count, int width,
                    int height, float* const luts[4], const long strides[4], hipStream_t s,
                    int stages) {
    if (stages & 1)
        hipLaunchKernelGGL(k_transmittance_batched<GT>,
                           dim3(grid_for((long)k.tw * k.th * GT), count), dim3(256), 0, s, states,
                           luts[0], strides[0]);
    if (stages & 2)
        hipLaunchKernelGGL(k_multi_scattering_batched,
                           dim3((unsigned)((k.mw * k.mh + kMsBatchTexelsPerBlock - 1) /
                                           kMsBatchTexelsPerBlock), count),
                           dim3(256), 0, s, states, (const float*)luts[0], strides[0], luts[1],
                           strides[1]);
    if (stages & 4)
        hipLaunchKernelGGL(k_sky_view_batched<GS>, dim3(grid_for((long)k.sw * k.sh * GS), count),
                           dim3(256), 0, s, states, (const float*)luts[0], strides[0],
                           (const float*)luts[1], strides[1], luts[2], strides[2]);
    if (stages & 8)
        hipLaunchKernelGGL(k_aerial_batched, dim3(grid_for((long)k.aw * k.ah * 64), count),
                           dim3(256), 0, s, states, width, height, (const float*)luts[0],
                           strides[0], (const float*)luts[1], strides[1], luts[3], strides[3]);
}

template <int GT>
void launch_batch_gt(const Consts& k, const BatchState* states, unsigned count, int width,
                     int height, float* const luts[4], const long strides[4], hipStream_t s,
                     int g_sv, int stages) {
    switch (g_sv) {
        case 1: launch_batch_g<GT, 1>(k, states, count, width, height, luts, strides, s, stages); break;
        case 2: launch_batch_g<GT, 2>(k, states, count, width, height, luts, strides, s, stages); break;
        case 4: launch_batch_g<GT, 4>(k, states, count, width, height, luts, strides, s, stages); break;
        case 16: launch_batch_g<GT, 16>(k, states, count, width, height, luts, strides, s, stages); break;
        default: launch_batch_g<GT, 8>(k, states, count, width, height, luts, strides, s, stages); break;
    }
}

void launch_batch(const Consts& k, const BatchState* states, unsigned count, int width, int height,
                  float* const luts[4], const long strides[4], hipStream_t s, int g_t, int g_sv,
                  int stages) {
    switch (g_t) {
        case 1: launch_batch_gt<1>(k, states, count, width, height, luts, strides, s, g_sv, stages); break;
        case 2: launch_batch_gt<2>(k, states, count, width, height, luts, strides, s, g_sv, stages); break;
        case 4: launch_batch_gt<4>(k, states, count, width, height, luts, strides, s, g_sv, stages); break;
        case 16: launch_batch_gt<16>(k, states, count, width, height, luts, strides, s, g_sv, stages); break;
        default: launch_batch_gt<8>(k, states, count, width, height, luts, strides, s, g_sv, stages); break;
    }
}
#endif

}  // inline namespace ATMO_PRECISION
}  // namespace atmo
