// shade.hip -- surface shading of libatmo.so: the reference's
// `PlanetRenderer._shade_surface` (src/planet_renderer/renderer.py:
// 1277-1575), the kernel that consumes prepare_frame's
// surface_sun_transmittance and surface_sky_radiance.
//
// One pixel per lane.  A pixel reads its G-buffer sample and the two
// surface-lighting fields (52 bytes for a lit pixel, 12 bytes out); an
// ocean pixel with the advanced ocean also reads the ocean fields and up
// to four samples of the pre-ocean terrain snapshot around its refracted
// seabed point, up to 48 pixels away.  Blocks are dealt to the XCDs in
// contiguous bands of the frame (block b runs on XCD b mod 8), so those
// lookups land in the L2 that already holds the neighbouring pixels:
// without the banding each XCD fetched the snapshot again and the kernel
// read 3.2x its algorithmic bytes from HBM, with it 1.45x.
//
// Arithmetic follows the reference expression for expression, float32,
// left to right as Taichi evaluates vector expressions, no FMA contraction
// (the library is built with -ffp-contract=off), IEEE division and square
// root; normalized(v) = (1 / |v|) v; exp is float64 rounded once; the
// per-launch tan(sun_angular_radius)^2 is computed once on the host the
// same way.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdint.h>
#include <stdio.h>

#include "../../include/atmo.h"

namespace {

constexpr int kThreads = 256;
constexpr int kXcds = 8;
constexpr int32_t kOceanSurfaceId = 2147000000;  // ocean.py:17

struct v3 {
    float x, y, z;
};
__device__ __forceinline__ v3 V(float x, float y, float z) { return v3{x, y, z}; }
__device__ __forceinline__ v3 add(v3 a, v3 b) { return V(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ v3 sub(v3 a, v3 b) { return V(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ v3 mul(v3 a, v3 b) { return V(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ __forceinline__ v3 div3(v3 a, v3 b) { return V(a.x / b.x, a.y / b.y, a.z / b.z); }
__device__ __forceinline__ v3 scale(v3 a, float s) { return V(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ v3 divs(v3 a, float s) { return V(a.x / s, a.y / s, a.z / s); }
__device__ __forceinline__ v3 neg(v3 a) { return V(-a.x, -a.y, -a.z); }
__device__ __forceinline__ float dot(v3 a, v3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ __forceinline__ float norm(v3 a) { return sqrtf(dot(a, a)); }
__device__ __forceinline__ v3 normalized(v3 a) {
    const float inv = 1.0f / norm(a);
    return V(inv * a.x, inv * a.y, inv * a.z);
}
__device__ __forceinline__ v3 vmax(v3 a, float s) {
    return V(fmaxf(a.x, s), fmaxf(a.y, s), fmaxf(a.z, s));
}
__device__ __forceinline__ v3 load(const float* __restrict__ base, size_t i) {
    return V(base[3 * i], base[3 * i + 1], base[3 * i + 2]);
}
__device__ __forceinline__ v3 arr(const float* a) { return V(a[0], a[1], a[2]); }
__device__ __forceinline__ float exp_cr(float x) { return (float)exp((double)x); }

struct ShadeArgs {
    AtmoShadeFields f;
    AtmoShadeParams p;
    int32_t width, height;
    float width_f, height_f, aspect;
    float solar_slope_variance;  // tan(sun_angular_radius)^2
    double inv_height;
    int32_t blocks_per_xcd;
};

// renderer.py:1301-1575, one pixel per lane.
__global__ __launch_bounds__(kThreads) void k_shade_surface(ShadeArgs a) {
    const AtmoShadeFields& f = a.f;
    const AtmoShadeParams& p = a.p;
    const size_t n = (size_t)a.width * a.height;
    // block b runs on XCD b mod 8: each XCD takes a contiguous band of
    // pixels, so the seabed lookups of neighbouring pixels meet in one L2
    const size_t chunk = (size_t)(blockIdx.x % kXcds) * a.blocks_per_xcd + blockIdx.x / kXcds;
    const size_t q = chunk * kThreads + threadIdx.x;
    if (q >= n) return;
    constexpr float PI = 3.14159265358979323846f;  // math.pi
    constexpr float FOUR_PI = 12.566370614359172f;  // 4.0 * math.pi, folded in float64
    v3 color = V(0.00015f, 0.0002f, 0.00035f);
    const int32_t sid = f.surface_id[q];
    if (sid >= 0) {
        const v3 albedo = load(f.albedo, q);
        color = albedo;
        if (p.mode == 0) {
            const v3 sun = arr(p.sun_global), E = arr(p.solar_irradiance);
            const v3 T = load(f.sun_transmittance, q), sky = load(f.sky_radiance, q);
            const v3 normal = normalized(load(f.normal, q));
            const float ndotl = fmaxf(dot(normal, sun), 0.0f);
            const v3 direct_irradiance = scale(mul(E, T), ndotl);
            color = mul(albedo, add(divs(direct_irradiance, PI), sky));
            if (sid == kOceanSurfaceId && p.advanced_ocean_enabled != 0) {
                // pixel q = (x, y), q = x * H + y; exact below 2^51 pixels
                const int x = (int)(((double)q + 0.5) * a.inv_height);
                const int y = (int)(q - (size_t)x * a.height);
                const v3 right = arr(p.view_right_global), up = arr(p.view_up_global),
                         fwd = arr(p.view_forward_global);
                const v3 position = load(f.position, q);
                const v3 view_direction = neg(normalized(
                    add(add(scale(right, position.x), scale(up, position.y)), scale(fwd, position.z))));
                const float ndotv = fmaxf(dot(normal, view_direction), 1.0e-4f);
                const v3 half_sum = add(view_direction, sun);
                const v3 half_vector = divs(half_sum, fmaxf(norm(half_sum), 1.0e-5f));
                const float ndoth = fmaxf(dot(normal, half_vector), 0.0f);
                const float vdoth = fmaxf(dot(view_direction, half_vector), 0.0f);

                const float f0 = p.ocean_f0;
                const float omv = 1.0f - ndotv;
                const float fresnel_view = f0 + (1.0f - f0) * ((((omv * omv) * omv) * omv) * omv);
                const float omh = 1.0f - vdoth;
                const float fresnel_sun = f0 + (1.0f - f0) * ((((omh * omh) * omh) * omh) * omh);

                const float base_alpha = p.ocean_roughness * p.ocean_roughness;
                float alpha2 = base_alpha * base_alpha + f.ocean_slope_variance[q] +
                               a.solar_slope_variance;
                alpha2 = fminf(fmaxf(alpha2, 1.0e-5f), 1.0f);
                const float denominator = ndoth * ndoth * (alpha2 - 1.0f) + 1.0f;
                const float distribution = alpha2 / fmaxf(PI * denominator * denominator, 1.0e-5f);
                const float effective_roughness = sqrtf(sqrtf(alpha2));
                const float gk1 = effective_roughness + 1.0f;
                const float geometry_k = gk1 * gk1 / 8.0f;
                const float geometry_v =
                    ndotv / fmaxf(ndotv * (1.0f - geometry_k) + geometry_k, 1.0e-5f);
                const float geometry_l =
                    ndotl / fmaxf(ndotl * (1.0f - geometry_k) + geometry_k, 1.0e-5f);
                const v3 sun_specular = divs(
                    scale(scale(scale(scale(scale(mul(E, T), distribution), geometry_v), geometry_l),
                                fresnel_sun),
                          ndotl),
                    fmaxf(4.0f * ndotv * ndotl, 1.0e-4f));
                const v3 reflected_sky = scale(scale(sky, fresnel_view), p.ocean_sky_strength);
                float water_depth = fminf(f.water_depth_m[q], p.ocean_max_depth);
                v3 seabed_albedo = load(f.seabed_albedo, q);
                v3 seabed_normal = normalized(load(f.seabed_normal, q));

                // Refracted seabed lookup in the terrain snapshot
                // (renderer.py:1410-1530).
                const v3 incident_view = normalized(position);
                const v3 normal_view =
                    normalized(V(dot(normal, right), dot(normal, up), dot(normal, fwd)));
                const float cos_incident = fmaxf(-dot(incident_view, normal_view), 0.0f);
                const float eta = 1.0f / p.ocean_refraction_index;
                const float refract_discriminant =
                    1.0f - eta * eta * (1.0f - cos_incident * cos_incident);
                v3 refracted_view = incident_view;
                if (refract_discriminant > 0.0f) {
                    refracted_view = normalized(add(
                        scale(incident_view, eta),
                        scale(normal_view, eta * cos_incident - sqrtf(refract_discriminant))));
                }
                const float strength = p.ocean_refraction_strength;
                refracted_view = normalized(
                    add(scale(incident_view, 1.0f - strength), scale(refracted_view, strength)));
                const float vertical_depth =
                    water_depth * fmaxf(-dot(incident_view, normal_view), 0.05f);
                const float refracted_distance =
                    fminf(vertical_depth / fmaxf(-dot(refracted_view, normal_view), 0.05f),
                          p.ocean_max_depth * 2.0f);
                const v3 endpoint = add(position, scale(refracted_view, refracted_distance));
                if (endpoint.z > 0.1f && strength > 0.0f) {
                    const float thf = p.tangent_half_fov;
                    float spx = (endpoint.x / (endpoint.z * thf * a.aspect) * 0.5f + 0.5f) * a.width_f - 0.5f;
                    float spy = (endpoint.y / (endpoint.z * thf) * 0.5f + 0.5f) * a.height_f - 0.5f;
                    float dx = spx - (float)x, dy = spy - (float)y;
                    const float delta_length = sqrtf(dx * dx + dy * dy);
                    const float max_offset = p.ocean_refraction_max_offset_pixels;
                    if (delta_length > max_offset) {
                        const float s = max_offset / fmaxf(delta_length, 1.0e-5f);
                        dx = dx * s;
                        dy = dy * s;
                        spx = (float)x + dx;
                        spy = (float)y + dy;
                    }
                    // floor of a float far outside the frame is clamped
                    // before the cast; such samples fail the inside test
                    const float fbx = fminf(fmaxf(floorf(spx), -4.0f), a.width_f + 4.0f);
                    const float fby = fminf(fmaxf(floorf(spy), -4.0f), a.height_f + 4.0f);
                    const int bx = (int)fbx, by = (int)fby;
                    const float fx = spx - (float)bx, fy = spy - (float)by;
                    v3 acc_albedo = V(0.0f, 0.0f, 0.0f), acc_normal = V(0.0f, 0.0f, 0.0f),
                       acc_position = V(0.0f, 0.0f, 0.0f);
                    float acc_weight = 0.0f;
#pragma unroll
                    for (int oy = 0; oy < 2; ++oy) {
#pragma unroll
                        for (int ox = 0; ox < 2; ++ox) {
                            const int sx = bx + ox, sy = by + oy;
                            const bool inside = sx >= 0 && sx < a.width && sy >= 0 && sy < a.height;
                            const int cx = min(max(sx, 0), a.width - 1);
                            const int cy = min(max(sy, 0), a.height - 1);
                            const size_t s = (size_t)cx * a.height + cy;
                            if (inside && f.terrain_surface_id[s] >= 0) {
                                const v3 cp = load(f.terrain_position, s);
                                const float cd = norm(sub(cp, position));
                                if (cp.z > position.z + 0.01f && cd <= p.ocean_max_depth * 2.0f) {
                                    const float wx = ox == 0 ? 1.0f - fx : fx;
                                    const float wy = oy == 0 ? 1.0f - fy : fy;
                                    const float weight = wx * wy;
                                    acc_albedo = add(acc_albedo, scale(load(f.terrain_albedo, s), weight));
                                    acc_normal = add(acc_normal, scale(load(f.terrain_normal, s), weight));
                                    acc_position = add(acc_position, scale(cp, weight));
                                    acc_weight += weight;
                                }
                            }
                        }
                    }
                    if (acc_weight > 0.25f) {
                        const float inverse_weight = 1.0f / acc_weight;
                        seabed_albedo = scale(acc_albedo, inverse_weight);
                        seabed_normal = normalized(scale(acc_normal, inverse_weight));
                        const v3 seabed_position = scale(acc_position, inverse_weight);
                        water_depth = fminf(norm(sub(seabed_position, position)), p.ocean_max_depth);
                    }
                }
                const v3 scattering = arr(p.ocean_scattering);
                const v3 water_extinction = add(arr(p.ocean_absorption), scattering);
                const v3 ew = scale(neg(water_extinction), water_depth);
                const v3 water_transmission = V(exp_cr(ew.x), exp_cr(ew.y), exp_cr(ew.z));
                const float seabed_ndotl = fmaxf(dot(seabed_normal, sun), 0.0f);
                const v3 seabed_direct = scale(mul(E, T), seabed_ndotl);
                const v3 lit_seabed = mul(seabed_albedo, add(divs(seabed_direct, PI), sky));
                const v3 seabed = mul(lit_seabed, water_transmission);
                const v3 scattering_albedo = div3(scattering, vmax(water_extinction, 1.0e-6f));
                const v3 incident_water_radiance = add(sky, divs(direct_irradiance, FOUR_PI));
                const v3 water_scattering = mul(mul(incident_water_radiance, scattering_albedo),
                                                sub(V(1.0f, 1.0f, 1.0f), water_transmission));
                const v3 water_body = scale(add(seabed, water_scattering), 1.0f - fresnel_view);
                color = add(add(water_body, reflected_sky), sun_specular);
            }
        }
    }
    color = vmax(color, 0.0f);
    float* o = f.surface_hdr + 3 * q;
    o[0] = color.x;
    o[1] = color.y;
    o[2] = color.z;
}

}  // namespace

extern "C" {

int32_t atmo_shade_surface(const AtmoShadeFields* fields, const AtmoShadeParams* params,
                           int32_t width, int32_t height, void* stream) {
    if (!fields || !params || width <= 0 || height <= 0) return ATMO_EINVAL;
    const AtmoShadeFields& f = *fields;
    if (!f.surface_id || !f.albedo || !f.normal || !f.position || !f.ocean_slope_variance ||
        !f.water_depth_m || !f.seabed_albedo || !f.seabed_normal || !f.terrain_surface_id ||
        !f.terrain_position || !f.terrain_albedo || !f.terrain_normal || !f.sun_transmittance ||
        !f.sky_radiance || !f.surface_hdr)
        return ATMO_EINVAL;
    ShadeArgs a;
    a.f = f;
    a.p = *params;
    a.width = width;
    a.height = height;
    a.width_f = (float)width;
    a.height_f = (float)height;
    a.aspect = (float)width / (float)height;  // ti.cast(W, f32) / H
    const float t = (float)tan((double)params->sun_angular_radius);
    a.solar_slope_variance = t * t;
    a.inv_height = 1.0 / (double)height;
    const size_t n = (size_t)width * height;
    const size_t blocks = (n + kThreads - 1) / kThreads;
    a.blocks_per_xcd = (int32_t)((blocks + kXcds - 1) / kXcds);
    hipLaunchKernelGGL(k_shade_surface, dim3((unsigned)(kXcds * a.blocks_per_xcd)),
                       dim3(kThreads), 0, (hipStream_t)stream, a);
    return hipGetLastError() == hipSuccess ? ATMO_OK : ATMO_EHIP;
}

}  // extern "C"
