// atmo_frame.hip -- the reference's host methods in C: the LUT invalidation
// policy of AtmosphereRenderer.update / prepare_frame / composite /
// set_luts_frozen (src/planet_renderer/atmosphere/renderer.py:2057-2238) as
// pure host functions over an AtmoLutCache, and the handle-level entry points
// that apply it to the handle's LUTs through the kernel entry points.
//
// Host code only.  The float64 -> float32 roundings are the ones the
// reference's Taichi calls make when Python floats meet `ti.f32` kernel
// parameters: round to nearest, once.
#include <hip/hip_runtime.h>

#include <math.h>
#include <string.h>

#include "../../include/atmo.h"
#include "atmo_handle.h"

namespace {

// Python's math.radians: x * (pi / 180) in float64.
inline double radians(double degrees) { return degrees * (3.14159265358979323846 / 180.0); }

inline float f32(double x) { return (float)x; }

// float(np.float32(x)): the float32 rounding, held as a float64 key.
inline double key32(double x) { return (double)(float)x; }

inline bool same(const double* a, const double* b, int n) {
    for (int i = 0; i < n; ++i)
        if (!(a[i] == b[i])) return false;  // NaN never matches
    return true;
}

inline double clip(double x, double lo, double hi) { return x < lo ? lo : (x > hi ? hi : x); }

}  // namespace

extern "C" {

int32_t atmo_lut_cache_init(AtmoLutCache* c) {
    if (!c) return ATMO_EINVAL;
    memset(c, 0, sizeof *c);
    return ATMO_OK;
}

int32_t atmo_lut_cache_set_frozen(AtmoLutCache* c, int32_t frozen) {
    if (!c) return ATMO_EINVAL;
    // renderer.py:2068-2070
    if (c->luts_frozen && !frozen) c->has_sky_key = 0;
    c->luts_frozen = frozen ? 1 : 0;
    return ATMO_OK;
}

int32_t atmo_lut_cache_update(AtmoLutCache* c, double planet_radius_m, double camera_radius_m,
                              const double sun_local[3], const double solar_irradiance[3],
                              double sun_angular_radius_degrees, int32_t force_mask,
                              AtmoUpdatePlan* plan) {
    if (!c || !sun_local || !solar_irradiance || !plan) return ATMO_EINVAL;
    const double R = planet_radius_m;
    const double r = camera_radius_m;
    // renderer.py:2083-2088: the sun normalised in float64; the angular
    // radius rounded to float32 before it enters any key
    double sun[3] = {sun_local[0], sun_local[1], sun_local[2]};
    double norm = sqrt(sun[0] * sun[0] + sun[1] * sun[1] + sun[2] * sun[2]);
    norm = norm > 1.0e-12 ? norm : 1.0e-12;
    for (double& v : sun) v /= norm;
    const double ar = key32(radians(sun_angular_radius_degrees));
    memset(plan, 0, sizeof *plan);
    plan->bottom_radius = f32(R);
    for (int i = 0; i < 3; ++i) plan->solar_irradiance[i] = f32(solar_irradiance[i]);
    plan->sun_angular_radius = f32(ar);

    // renderer.py:2089-2094
    if (!c->has_planet_key || !(c->planet_key == R) || (force_mask & ATMO_LUT_TRANSMITTANCE)) {
        plan->rebuild |= ATMO_LUT_TRANSMITTANCE;
        c->planet_key = R;
        c->has_planet_key = 1;
        c->has_multi_scattering_key = 0;
        c->has_sky_key = 0;
        c->transmittance_rebuilds += 1;
    }
    // renderer.py:2096-2110
    const double ms_key[5] = {R, solar_irradiance[0], solar_irradiance[1], solar_irradiance[2],
                              ar};
    if (!c->has_multi_scattering_key || !same(ms_key, c->multi_scattering_key, 5) ||
        (force_mask & ATMO_LUT_MULTI_SCATTERING)) {
        plan->rebuild |= ATMO_LUT_MULTI_SCATTERING;
        memcpy(c->multi_scattering_key, ms_key, sizeof ms_key);
        c->has_multi_scattering_key = 1;
        c->has_sky_key = 0;
        c->multi_scattering_rebuilds += 1;
    }
    // renderer.py:2112-2137: the altitude is formed in float64, then rounded
    const double alt_key = key32(r - R > 0.0 ? r - R : 0.0);
    const double sun_key = key32(clip(sun[1], -1.0, 1.0));
    const double sky_key[6] = {alt_key, sun_key, solar_irradiance[0], solar_irradiance[1],
                               solar_irradiance[2], ar};
    plan->camera_altitude = f32(alt_key);
    plan->sun_zenith_cosine = f32(sun_key);
    const bool changed = !c->has_sky_key || !same(sky_key, c->sky_key, 6);
    if ((changed && (!c->luts_frozen || !c->has_sky_key)) || (force_mask & ATMO_LUT_SKY_VIEW)) {
        plan->rebuild |= ATMO_LUT_SKY_VIEW;
        memcpy(c->sky_key, sky_key, sizeof sky_key);
        c->has_sky_key = 1;
        c->sky_snapshot_altitude_m = alt_key;
        c->sky_snapshot_sun_cosine = sun_key;
        c->has_sky_snapshot = 1;
        c->sky_view_rebuilds += 1;
    }
    return ATMO_OK;
}

int32_t atmo_lut_cache_prepare(AtmoLutCache* c, double planet_radius_m, double camera_radius_m,
                               int32_t force_mask, int32_t* rebuild_aerial) {
    if (!c || !rebuild_aerial) return ATMO_EINVAL;
    // renderer.py:2173-2187
    *rebuild_aerial = 0;
    if (!c->luts_frozen || !c->aerial_available || (force_mask & ATMO_LUT_AERIAL)) {
        *rebuild_aerial = 1;
        c->aerial_snapshot_altitude_m = camera_radius_m - planet_radius_m;
        c->aerial_available = 1;
        c->aerial_rebuilds += 1;
    }
    return ATMO_OK;
}

double atmo_lut_cache_sky_altitude(const AtmoLutCache* c, double planet_radius_m,
                                   double camera_radius_m) {
    // renderer.py:2210-2217
    const double altitude = camera_radius_m - planet_radius_m;
    const double current = altitude > 0.0 ? altitude : 0.0;
    return (c && c->has_sky_snapshot) ? c->sky_snapshot_altitude_m : current;
}

int32_t atmo_get_lut_cache(AtmoHandle* h, AtmoLutCache* out) {
    if (!atmo_handle_ok(h)) return ATMO_EHANDLE;
    if (!out) return ATMO_EINVAL;
    *out = h->cache;
    return ATMO_OK;
}

int32_t atmo_set_luts_frozen(AtmoHandle* h, int32_t frozen) {
    if (!atmo_handle_ok(h)) return ATMO_EHANDLE;
    return atmo_lut_cache_set_frozen(&h->cache, frozen);
}

int32_t atmo_update(AtmoHandle* h, double planet_radius_m, double camera_radius_m,
                    const double sun_local[3], const double solar_irradiance[3],
                    double sun_angular_radius_degrees, int32_t force_mask, void* stream,
                    int32_t* rebuilt) {
    if (!atmo_handle_ok(h)) return ATMO_EHANDLE;
    if (!sun_local || !solar_irradiance) {
        atmo_set_error(h, "atmo_update: NULL argument");
        return ATMO_EINVAL;
    }
    AtmoLutCache next = h->cache;
    AtmoUpdatePlan p;
    atmo_lut_cache_update(&next, planet_radius_m, camera_radius_m, sun_local, solar_irradiance,
                          sun_angular_radius_degrees, force_mask, &p);
    int32_t status = ATMO_OK;
    if (p.rebuild & ATMO_LUT_TRANSMITTANCE)
        status = atmo_build_transmittance(h, p.bottom_radius, stream);
    if (status == ATMO_OK && (p.rebuild & ATMO_LUT_MULTI_SCATTERING))
        status = atmo_build_multi_scattering(h, p.bottom_radius, p.solar_irradiance,
                                             p.sun_angular_radius, stream);
    if (status == ATMO_OK && (p.rebuild & ATMO_LUT_SKY_VIEW))
        status = atmo_build_sky_view(h, p.bottom_radius, p.camera_altitude, p.sun_zenith_cosine,
                                     p.solar_irradiance, p.sun_angular_radius, stream);
    // keys move only once every launch went out; a failed launch leaves the
    // keys as they were, so the next update tries again
    if (status == ATMO_OK) h->cache = next;
    if (rebuilt) *rebuilt = status == ATMO_OK ? p.rebuild : 0;
    return status;
}

int32_t atmo_prepare_frame(AtmoHandle* h, const float* position_view, const int32_t* surface_id,
                           double planet_radius_m, double camera_radius_m,
                           const double sun_local[3], const double solar_irradiance[3],
                           double sun_angular_radius_degrees, const double view_basis[9],
                           double tangent_half_fov, int32_t force_mask, void* stream,
                           int32_t* rebuilt) {
    if (!atmo_handle_ok(h)) return ATMO_EHANDLE;
    if (!sun_local || !solar_irradiance || !view_basis) {
        atmo_set_error(h, "atmo_prepare_frame: NULL argument");
        return ATMO_EINVAL;
    }
    // renderer.py:2153-2172: altitude max(r - R, 0) and radians(alpha) in
    // float64, every kernel argument rounded once
    const double altitude = camera_radius_m - planet_radius_m;
    const float cam_alt = f32(altitude > 0.0 ? altitude : 0.0);
    const float bottom = f32(planet_radius_m);
    const float ar = f32(radians(sun_angular_radius_degrees));
    float sun[3], e[3], right[3], up[3], forward[3];
    for (int i = 0; i < 3; ++i) {
        sun[i] = f32(sun_local[i]);
        e[i] = f32(solar_irradiance[i]);
        right[i] = f32(view_basis[i]);
        up[i] = f32(view_basis[3 + i]);
        forward[i] = f32(view_basis[6 + i]);
    }
    int32_t status = atmo_prepare_surface_lighting(h, position_view, surface_id, bottom, cam_alt,
                                                   sun, ar, right, up, forward, stream);
    if (rebuilt) *rebuilt = 0;
    if (status != ATMO_OK) return status;
    AtmoLutCache next = h->cache;
    int32_t aerial = 0;
    atmo_lut_cache_prepare(&next, planet_radius_m, camera_radius_m, force_mask, &aerial);
    if (aerial) {
        status = atmo_build_aerial_perspective(h, bottom, cam_alt, sun, e, ar, right, up, forward,
                                               f32(tangent_half_fov), stream);
        if (status != ATMO_OK) return status;
        if (rebuilt) *rebuilt = ATMO_LUT_AERIAL;
    }
    h->cache = next;
    return ATMO_OK;
}

int32_t atmo_composite_frame(AtmoHandle* h, const float* scene_hdr, const float* space_hdr,
                             const int32_t* surface_id, const float* position_view, float* hdr,
                             double planet_radius_m, double camera_radius_m,
                             const double sun_local[3], const double solar_irradiance[3],
                             const double sun_disk_radiance[3],
                             double sun_angular_radius_degrees, const double view_basis[9],
                             double tangent_half_fov, double star_contrast_start,
                             double star_contrast_end, int32_t diagnostic_view, void* stream) {
    if (!atmo_handle_ok(h)) return ATMO_EHANDLE;
    if (!sun_local || !solar_irradiance || !sun_disk_radiance || !view_basis) {
        atmo_set_error(h, "atmo_composite_frame: NULL argument");
        return ATMO_EINVAL;
    }
    // renderer.py:2206-2238
    const double altitude = camera_radius_m - planet_radius_m;
    const double cam_alt = altitude > 0.0 ? altitude : 0.0;
    const double sky_alt = atmo_lut_cache_sky_altitude(&h->cache, planet_radius_m,
                                                       camera_radius_m);
    float sun[3], e[3], disk[3], right[3], up[3], forward[3];
    for (int i = 0; i < 3; ++i) {
        sun[i] = f32(sun_local[i]);
        e[i] = f32(solar_irradiance[i]);
        disk[i] = f32(sun_disk_radiance[i]);
        right[i] = f32(view_basis[i]);
        up[i] = f32(view_basis[3 + i]);
        forward[i] = f32(view_basis[6 + i]);
    }
    return atmo_composite(h, scene_hdr, space_hdr, surface_id, position_view, hdr,
                          f32(planet_radius_m), f32(cam_alt), f32(sky_alt), sun, e, disk,
                          f32(radians(sun_angular_radius_degrees)), right, up, forward,
                          f32(tangent_half_fov), f32(star_contrast_start),
                          f32(star_contrast_end), diagnostic_view, stream);
}

}  // extern "C"
