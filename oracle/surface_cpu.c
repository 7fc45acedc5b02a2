/*
 * surface_cpu.c -- CPU restatement of the reference's surface shading,
 * `PlanetRenderer._shade_surface` (src/planet_renderer/renderer.py:
 * 1277-1575), float32 operation for operation.
 *
 * TEST INFRASTRUCTURE ONLY: the checker of libatmo.so's atmo_shade_surface;
 * the package never loads it.
 *
 * Conventions (the same as oracle/ref_cpu.c): Taichi's left-to-right
 * evaluation of vector expressions, dot = (x*x' + y*y') + z*z', norm =
 * sqrt(dot), normalized(v) = (1 / norm(v)) * v, vector / scalar divides
 * each component, ti.max/ti.min are fmaxf/fminf, `x ** 2` is x * x, Python
 * constants fold in float64 and round once to float32, no FMA contraction,
 * IEEE division and square root.  exp and tan are correctly rounded
 * (float64, rounded once).
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>

#include "../include/atmo.h"

#define EXPORT __attribute__((visibility("default")))

#define OCEAN_SURFACE_ID 2147000000 /* ocean.py:17 */

typedef struct { float x, y, z; } v3;

static inline v3 V(float x, float y, float z) { v3 r = {x, y, z}; return r; }
static inline v3 add(v3 a, v3 b) { return V(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 sub(v3 a, v3 b) { return V(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 mul(v3 a, v3 b) { return V(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline v3 div3(v3 a, v3 b) { return V(a.x / b.x, a.y / b.y, a.z / b.z); }
static inline v3 scale(v3 a, float s) { return V(a.x * s, a.y * s, a.z * s); }
static inline v3 divs(v3 a, float s) { return V(a.x / s, a.y / s, a.z / s); }
static inline v3 neg(v3 a) { return V(-a.x, -a.y, -a.z); }
static inline float dot(v3 a, v3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
static inline float norm(v3 a) { return sqrtf(dot(a, a)); }
static inline v3 normalized(v3 a) {
    const float inv = 1.0f / norm(a);
    return V(inv * a.x, inv * a.y, inv * a.z);
}
static inline v3 vmax(v3 a, float s) { return V(fmaxf(a.x, s), fmaxf(a.y, s), fmaxf(a.z, s)); }
static inline v3 load(const float* base, size_t i) { return V(base[3 * i], base[3 * i + 1], base[3 * i + 2]); }
static inline v3 arr(const float* a) { return V(a[0], a[1], a[2]); }
static inline float cr_exp(float x) { return (float)exp((double)x); }
static inline float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }

/* renderer.py:1301-1575 for pixel q = (x, y). */
static v3 shade_pixel(const AtmoShadeFields* f, const AtmoShadeParams* p, int32_t width,
                      int32_t height, int32_t x, int32_t y, int32_t* path) {
    const float PI = 3.14159265358979323846f;      /* math.pi */
    const float FOUR_PI = 12.566370614359172f;      /* 4.0 * math.pi, folded in float64 */
    const size_t q = (size_t)x * (size_t)height + (size_t)y;
    v3 color = V(0.00015f, 0.0002f, 0.00035f);
    *path = 0;
    if (f->surface_id[q] < 0) return vmax(color, 0.0f);
    *path = 1;
    const v3 albedo = load(f->albedo, q);
    color = albedo;
    if (p->mode == 0) {
        const v3 sun = arr(p->sun_global), E = arr(p->solar_irradiance);
        const v3 T = load(f->sun_transmittance, q), sky = load(f->sky_radiance, q);
        const v3 normal = normalized(load(f->normal, q));
        const float ndotl = fmaxf(dot(normal, sun), 0.0f);
        const v3 direct_irradiance = scale(mul(E, T), ndotl);
        color = mul(albedo, add(divs(direct_irradiance, PI), sky));
        if (f->surface_id[q] == OCEAN_SURFACE_ID && p->advanced_ocean_enabled != 0) {
            *path = 2;
            const v3 right = arr(p->view_right_global), up = arr(p->view_up_global),
                     fwd = arr(p->view_forward_global);
            const v3 position = load(f->position, q);
            const v3 view_direction = neg(normalized(
                add(add(scale(right, position.x), scale(up, position.y)), scale(fwd, position.z))));
            const float ndotv = fmaxf(dot(normal, view_direction), 1.0e-4f);
            const v3 half_sum = add(view_direction, sun);
            const v3 half_vector = divs(half_sum, fmaxf(norm(half_sum), 1.0e-5f));
            const float ndoth = fmaxf(dot(normal, half_vector), 0.0f);
            const float vdoth = fmaxf(dot(view_direction, half_vector), 0.0f);

            const float f0 = p->ocean_f0;
            const float omv = 1.0f - ndotv;
            const float fresnel_view = f0 + (1.0f - f0) * ((((omv * omv) * omv) * omv) * omv);
            const float omh = 1.0f - vdoth;
            const float fresnel_sun = f0 + (1.0f - f0) * ((((omh * omh) * omh) * omh) * omh);

            const float base_alpha = p->ocean_roughness * p->ocean_roughness;
            const float tan_r = (float)tan((double)p->sun_angular_radius);
            const float solar_slope_variance = tan_r * tan_r;
            float alpha2 = base_alpha * base_alpha + f->ocean_slope_variance[q] + solar_slope_variance;
            alpha2 = clampf(alpha2, 1.0e-5f, 1.0f);
            const float denominator = ndoth * ndoth * (alpha2 - 1.0f) + 1.0f;
            const float distribution = alpha2 / fmaxf(PI * denominator * denominator, 1.0e-5f);
            const float effective_roughness = sqrtf(sqrtf(alpha2));
            const float gk1 = effective_roughness + 1.0f;
            const float geometry_k = gk1 * gk1 / 8.0f;
            const float geometry_v = ndotv / fmaxf(ndotv * (1.0f - geometry_k) + geometry_k, 1.0e-5f);
            const float geometry_l = ndotl / fmaxf(ndotl * (1.0f - geometry_k) + geometry_k, 1.0e-5f);
            const v3 sun_specular = divs(
                scale(scale(scale(scale(scale(mul(E, T), distribution), geometry_v), geometry_l),
                            fresnel_sun), ndotl),
                fmaxf(4.0f * ndotv * ndotl, 1.0e-4f));
            const v3 reflected_sky = scale(scale(sky, fresnel_view), p->ocean_sky_strength);
            float water_depth = fminf(f->water_depth_m[q], p->ocean_max_depth);
            v3 seabed_albedo = load(f->seabed_albedo, q);
            v3 seabed_normal = normalized(load(f->seabed_normal, q));

            const v3 incident_view = normalized(position);
            const v3 normal_view = normalized(V(dot(normal, right), dot(normal, up), dot(normal, fwd)));
            const float cos_incident = fmaxf(-dot(incident_view, normal_view), 0.0f);
            const float eta = 1.0f / p->ocean_refraction_index;
            const float refract_discriminant =
                1.0f - eta * eta * (1.0f - cos_incident * cos_incident);
            v3 refracted_view = incident_view;
            if (refract_discriminant > 0.0f) {
                refracted_view = normalized(add(
                    scale(incident_view, eta),
                    scale(normal_view, eta * cos_incident - sqrtf(refract_discriminant))));
            }
            const float strength = p->ocean_refraction_strength;
            refracted_view = normalized(
                add(scale(incident_view, 1.0f - strength), scale(refracted_view, strength)));
            const float vertical_depth =
                water_depth * fmaxf(-dot(incident_view, normal_view), 0.05f);
            const float refracted_distance =
                fminf(vertical_depth / fmaxf(-dot(refracted_view, normal_view), 0.05f),
                      p->ocean_max_depth * 2.0f);
            const v3 endpoint = add(position, scale(refracted_view, refracted_distance));
            if (endpoint.z > 0.1f && strength > 0.0f) {
                const float aspect = (float)width / (float)height;
                const float thf = p->tangent_half_fov;
                float spx = (endpoint.x / (endpoint.z * thf * aspect) * 0.5f + 0.5f) * (float)width - 0.5f;
                float spy = (endpoint.y / (endpoint.z * thf) * 0.5f + 0.5f) * (float)height - 0.5f;
                float dx = spx - (float)x, dy = spy - (float)y;
                const float delta_length = sqrtf(dx * dx + dy * dy);
                const float max_offset = p->ocean_refraction_max_offset_pixels;
                if (delta_length > max_offset) {
                    const float s = max_offset / fmaxf(delta_length, 1.0e-5f);
                    dx = dx * s;
                    dy = dy * s;
                    spx = (float)x + dx;
                    spy = (float)y + dy;
                }
                const float fbx = floorf(spx), fby = floorf(spy);
                const int32_t bx = (int32_t)fbx, by = (int32_t)fby;
                const float fx = spx - (float)bx, fy = spy - (float)by;
                v3 acc_albedo = V(0.0f, 0.0f, 0.0f), acc_normal = V(0.0f, 0.0f, 0.0f),
                   acc_position = V(0.0f, 0.0f, 0.0f);
                float acc_weight = 0.0f;
                for (int oy = 0; oy < 2; ++oy) {
                    for (int ox = 0; ox < 2; ++ox) {
                        const int32_t sx = bx + ox, sy = by + oy;
                        const int inside = sx >= 0 && sx < width && sy >= 0 && sy < height;
                        const int32_t cx = sx < 0 ? 0 : (sx > width - 1 ? width - 1 : sx);
                        const int32_t cy = sy < 0 ? 0 : (sy > height - 1 ? height - 1 : sy);
                        const size_t s = (size_t)cx * (size_t)height + (size_t)cy;
                        if (inside && f->terrain_surface_id[s] >= 0) {
                            const v3 cp = load(f->terrain_position, s);
                            const float cd = norm(sub(cp, position));
                            const int valid = cp.z > position.z + 0.01f && cd <= p->ocean_max_depth * 2.0f;
                            if (valid) {
                                const float wx = ox == 0 ? 1.0f - fx : fx;
                                const float wy = oy == 0 ? 1.0f - fy : fy;
                                const float weight = wx * wy;
                                acc_albedo = add(acc_albedo, scale(load(f->terrain_albedo, s), weight));
                                acc_normal = add(acc_normal, scale(load(f->terrain_normal, s), weight));
                                acc_position = add(acc_position, scale(cp, weight));
                                acc_weight += weight;
                            }
                        }
                    }
                }
                if (acc_weight > 0.25f) {
                    *path = 3;
                    const float inverse_weight = 1.0f / acc_weight;
                    seabed_albedo = scale(acc_albedo, inverse_weight);
                    seabed_normal = normalized(scale(acc_normal, inverse_weight));
                    const v3 seabed_position = scale(acc_position, inverse_weight);
                    water_depth = fminf(norm(sub(seabed_position, position)), p->ocean_max_depth);
                }
            }
            const v3 absorption = arr(p->ocean_absorption), scattering = arr(p->ocean_scattering);
            const v3 water_extinction = add(absorption, scattering);
            const v3 ew = scale(neg(water_extinction), water_depth);
            const v3 water_transmission = V(cr_exp(ew.x), cr_exp(ew.y), cr_exp(ew.z));
            const float seabed_ndotl = fmaxf(dot(seabed_normal, sun), 0.0f);
            const v3 seabed_direct = scale(mul(E, T), seabed_ndotl);
            const v3 lit_seabed = mul(seabed_albedo, add(divs(seabed_direct, PI), sky));
            const v3 seabed = mul(lit_seabed, water_transmission);
            const v3 scattering_albedo = div3(scattering, vmax(water_extinction, 1.0e-6f));
            const v3 incident_water_radiance = add(sky, divs(direct_irradiance, FOUR_PI));
            const v3 one = V(1.0f, 1.0f, 1.0f);
            const v3 water_scattering =
                mul(mul(incident_water_radiance, scattering_albedo), sub(one, water_transmission));
            const v3 water_body = scale(add(seabed, water_scattering), 1.0f - fresnel_view);
            color = add(add(water_body, reflected_sky), sun_specular);
        }
    }
    return vmax(color, 0.0f);
}

/* _shade_surface over the whole frame.  path (may be NULL) receives per
 * pixel the branch taken: 0 sky, 1 lit surface, 2 ocean with its own
 * seabed, 3 ocean with the refracted seabed lookup. */
EXPORT void surface_shade(const AtmoShadeFields* fields, const AtmoShadeParams* params,
                          int32_t width, int32_t height, int32_t* path) {
#pragma omp parallel for schedule(static)
    for (int32_t x = 0; x < width; ++x) {
        for (int32_t y = 0; y < height; ++y) {
            int32_t taken;
            const v3 c = shade_pixel(fields, params, width, height, x, y, &taken);
            if (path) path[(size_t)x * (size_t)height + (size_t)y] = taken;
            float* o = fields->surface_hdr + ((size_t)x * (size_t)height + (size_t)y) * 3;
            o[0] = c.x;
            o[1] = c.y;
            o[2] = c.z;
        }
    }
}
