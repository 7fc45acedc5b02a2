/*
 * atmo.h -- C-ABI of the MI355X atmosphere library (libatmo.so).
 *
 * The library replaces the six Taichi kernels of the reference's
 * `AtmosphereRenderer` (src/planet_renderer/atmosphere/renderer.py) with
 * hand-written HIP kernels for gfx950.  Each entry point below names the
 * reference kernel it replaces.  Host policy (LUT cache keys, freeze state,
 * float64 -> float32 argument rounding) stays in the caller, exactly as the
 * reference keeps it in the Python methods `update`, `prepare_frame` and
 * `composite` (renderer.py:2057-2238).
 *
 * Conventions shared by every entry point:
 *  - All device pointers are plain device addresses (hipMalloc'd memory or
 *    torch tensor data_ptr()).  Frame-sized arrays use the reference's Taichi
 *    field layout: shape (W, H, 3) float32 or (W, H) int32, indexed [x, y],
 *    so element (x, y) lives at offset (x * H + y) * 3.  Screen y = 0 is the
 *    bottom row (renderer.py:1853-1856).
 *  - `stream` is a hipStream_t passed as void* (NULL = the null stream).
 *    Calls are asynchronous on that stream; the caller keeps its buffers
 *    alive until the stream has completed.  A handle is not re-entrant.
 *  - Return value is an ATMO_* status.  atmo_last_error() gives the message
 *    of the last failure on a handle (or of the last failed atmo_create when
 *    the handle argument is NULL).
 *  - Scalar arguments are float32, as the Taichi kernels declare them
 *    (`ti.f32` parameters); the caller rounds its float64 values exactly as
 *    Taichi does when it passes Python floats to an f32 kernel argument.
 */
#ifndef ATMO_H
#define ATMO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    ATMO_OK = 0,
    ATMO_EINVAL = 1,  /* invalid configuration or argument */
    ATMO_ESHAPE = 2,  /* buffer shape / size mismatch */
    ATMO_EHIP = 3,    /* HIP runtime error */
    ATMO_ENOMEM = 4,  /* device allocation failed */
    ATMO_EHANDLE = 5  /* NULL or destroyed handle */
};

/* Mirrors every field of `AtmosphereConfig`
 * (src/planet_renderer/atmosphere/config.py:29-59), in the same order and
 * units (metres, inverse metres).  Optical values are double because the
 * reference holds them as Python floats and rounds them to float32 only
 * where a kernel uses them (`ti.static(...)` constants). */
typedef struct AtmoParams {
    double top_altitude_m;
    double rayleigh_scattering_per_m[3];
    double rayleigh_scale_height_m;
    double mie_scattering_per_m[3];
    double mie_extinction_per_m[3];
    double mie_scale_height_m;
    double mie_phase_g;
    double absorption_extinction_per_m[3];
    double absorption_peak_altitude_m;
    double absorption_half_width_m;
    double ground_albedo[3];
    int32_t transmittance_lut_width;
    int32_t transmittance_lut_height;
    int32_t multi_scattering_lut_width;
    int32_t multi_scattering_lut_height;
    int32_t sky_view_lut_width;
    int32_t sky_view_lut_height;
    int32_t aerial_lut_width;
    int32_t aerial_lut_height;
    int32_t aerial_lut_depth;
    int32_t transmittance_steps;
    int32_t multi_scattering_directions;
    int32_t multi_scattering_steps;
    int32_t sky_view_steps;
    int32_t aerial_steps_per_slice;
    int32_t aerial_horizon_raymarch_steps;
    int32_t aerial_terminator_substeps;
} AtmoParams;

/* Values of `AtmosphereDiagnosticView`
 * (src/planet_renderer/atmosphere/diagnostics.py:12-25). */
enum {
    ATMO_VIEW_COMPOSITE = 0,
    ATMO_VIEW_SKY_VIEW = 1,
    ATMO_VIEW_CAMERA_TRANSMITTANCE = 2,
    ATMO_VIEW_TRANSMITTANCE_LUT = 3,
    ATMO_VIEW_MULTI_SCATTERING_LUT = 4,
    ATMO_VIEW_AERIAL_SCATTERING = 5,
    ATMO_VIEW_AERIAL_TRANSMITTANCE = 6,
    ATMO_VIEW_SURFACE_MASK = 7,
    ATMO_VIEW_AERIAL_FROXEL_ONLY = 8,
    ATMO_VIEW_AERIAL_DIRECT_ONLY = 9,
    ATMO_VIEW_AERIAL_BLEND_WEIGHT = 10
};

/* Handle-owned device buffers, the fields allocated in
 * `AtmosphereRenderer.__init__` (renderer.py:55-96). */
enum {
    ATMO_BUF_TRANSMITTANCE = 0,      /* (TW, TH, 3) f32 */
    ATMO_BUF_MULTI_SCATTERING = 1,   /* (MW, MH, 3) f32 */
    ATMO_BUF_SKY_VIEW = 2,           /* (SW, SH, 3) f32 */
    ATMO_BUF_AERIAL_SCATTERING = 3,  /* (AW, AH, AD, 3) f32 */
    ATMO_BUF_SURFACE_SUN_T = 4,      /* (W, H, 3) f32 */
    ATMO_BUF_SURFACE_SKY_L = 5       /* (W, H, 3) f32 */
};

typedef struct AtmoHandle AtmoHandle;

/* Range checks of `AtmosphereConfig.__post_init__` (config.py:61-128).
 * Writes the reference's ValueError message into `message` on failure. */
int32_t atmo_validate_params(const AtmoParams* params, char* message,
                             int32_t message_capacity);

/* Replaces `AtmosphereRenderer.__init__(width, height, config)`
 * (renderer.py:51-116): validates the configuration and allocates the LUT
 * and surface-lighting buffers on the current HIP device. */
int32_t atmo_create(const AtmoParams* params, int32_t width, int32_t height,
                    AtmoHandle** out);
int32_t atmo_destroy(AtmoHandle* handle);
const char* atmo_last_error(const AtmoHandle* handle);

/* Device address and shape of a handle-owned buffer (ATMO_BUF_*).  Unused
 * trailing shape entries are 0. */
int32_t atmo_get_buffer(AtmoHandle* handle, int32_t which, void** device_ptr,
                        int64_t shape[4]);

/* Replaces a handle-owned buffer with caller-owned device memory of exactly
 * `bytes` bytes (the size atmo_get_buffer's shape implies, float32).  The
 * handle frees its own allocation and never frees the caller's; the caller
 * keeps it alive until atmo_destroy.  This is how the Python facade makes
 * torch tensors the LUT and surface-lighting fields, as the reference's
 * Taichi fields are (renderer.py:55-96). */
int32_t atmo_bind_buffer(AtmoHandle* handle, int32_t which, void* device_ptr,
                         int64_t bytes);

/* Replaces `_build_transmittance(bottom_radius)` (renderer.py:1034-1063). */
int32_t atmo_build_transmittance(AtmoHandle* handle, float bottom_radius,
                                 void* stream);

/* Replaces `_build_multi_scattering(bottom_radius, solar_irradiance,
 * sun_angular_radius)` (renderer.py:1065-1206).  Reads the T-LUT. */
int32_t atmo_build_multi_scattering(AtmoHandle* handle, float bottom_radius,
                                    const float solar_irradiance[3],
                                    float sun_angular_radius, void* stream);

/* Replaces `_build_sky_view(bottom_radius, camera_altitude,
 * sun_zenith_cosine, solar_irradiance, sun_angular_radius)`
 * (renderer.py:1208-1302).  Reads the T- and MS-LUTs. */
int32_t atmo_build_sky_view(AtmoHandle* handle, float bottom_radius,
                            float camera_altitude, float sun_zenith_cosine,
                            const float solar_irradiance[3],
                            float sun_angular_radius, void* stream);

/* Replaces `_build_aerial_perspective(bottom_radius, camera_altitude, sun,
 * solar_irradiance, sun_angular_radius, right, view_up, forward,
 * tangent_half_fov)` (renderer.py:1304-1420). */
int32_t atmo_build_aerial_perspective(
    AtmoHandle* handle, float bottom_radius, float camera_altitude,
    const float sun[3], const float solar_irradiance[3],
    float sun_angular_radius, const float right[3], const float view_up[3],
    const float forward[3], float tangent_half_fov, void* stream);

/* Replaces `_prepare_surface_lighting(position_view, surface_id,
 * bottom_radius, camera_altitude, sun, sun_angular_radius, right, view_up,
 * forward)` (renderer.py:1422-1472).  position_view: (W,H,3) f32,
 * surface_id: (W,H) i32, both device pointers.  Writes the handle's
 * ATMO_BUF_SURFACE_SUN_T and ATMO_BUF_SURFACE_SKY_L buffers. */
int32_t atmo_prepare_surface_lighting(
    AtmoHandle* handle, const float* position_view, const int32_t* surface_id,
    float bottom_radius, float camera_altitude, const float sun[3],
    float sun_angular_radius, const float right[3], const float view_up[3],
    const float forward[3], void* stream);

/* Replaces `_composite(scene_hdr, space_hdr, surface_id, position_view, hdr,
 * bottom_radius, camera_altitude, sky_lut_camera_altitude, sun_local,
 * solar_irradiance, sun_disk_radiance, sun_angular_radius, right, view_up,
 * forward, tangent_half_fov, star_contrast_start, star_contrast_end,
 * diagnostic_view)` (renderer.py:1827-2055), every diagnostic view
 * included.  All frame arrays are (W,H,3) f32 / (W,H) i32 device pointers;
 * hdr is written for every pixel. */
int32_t atmo_composite(
    AtmoHandle* handle, const float* scene_hdr, const float* space_hdr,
    const int32_t* surface_id, const float* position_view, float* hdr,
    float bottom_radius, float camera_altitude, float sky_lut_camera_altitude,
    const float sun_local[3], const float solar_irradiance[3],
    const float sun_disk_radiance[3], float sun_angular_radius,
    const float right[3], const float view_up[3], const float forward[3],
    float tangent_half_fov, float star_contrast_start,
    float star_contrast_end, int32_t diagnostic_view, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* ATMO_H */
