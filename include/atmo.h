/*
 * atmo.h -- C-ABI of the MI355X atmosphere library (libatmo.so).
 *
 * The library replaces the six Taichi kernels of the reference's
 * `AtmosphereRenderer` (src/planet_renderer/atmosphere/renderer.py) with
 * hand-written HIP kernels for gfx950, and with them the two steps either
 * side of the composite: the surface shading that consumes prepare_frame's
 * lighting (`PlanetRenderer._shade_surface`, src/planet_renderer/
 * renderer.py:1277-1575) and the display post-process (`PostProcessor`,
 * src/planet_renderer/postprocess.py).  Each entry point below names the
 * reference kernel it replaces.  Host policy (LUT cache keys, freeze state,
 * float64 -> float32 argument rounding, the exposure and bloom switches)
 * stays in the caller, exactly as the reference keeps it in the Python
 * methods `update`, `prepare_frame`, `composite` (renderer.py:2057-2238)
 * and `PostProcessor.process` (postprocess.py:176-207).
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

/* The defaults of `AtmosphereConfig` (config.py:33-59), for hosts that
 * build an AtmoParams without the Python facade. */
int32_t atmo_default_params(AtmoParams* out);

/* Range checks of `AtmosphereConfig.__post_init__` (config.py:61-128).
 * Writes the reference's ValueError message into `message` on failure. */
int32_t atmo_validate_params(const AtmoParams* params, char* message,
                             int32_t message_capacity);

/* Largest frame a handle accepts, width x height: 2^28 pixels (16384 x
 * 16384), so that a pixel's byte offset in a (W, H, 3) float32 field fits
 * in 32 bits.  The reference sets no limit; its fields at this size would
 * take 3 GiB each. */
#define ATMO_MAX_PIXELS 268435456

/* Replaces `AtmosphereRenderer.__init__(width, height, config)`
 * (renderer.py:51-116): validates the configuration and allocates the LUT
 * and surface-lighting buffers on the current HIP device.  Frames above
 * ATMO_MAX_PIXELS return ATMO_EINVAL. */
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

/* Arithmetic of a handle's kernels.  REFERENCE (the default) follows the
 * reference's float32 kernels operation for operation.  ROBUST has no
 * counterpart in the reference: the same float32 march with float64
 * evaluation where float32 cancels at planet scale -- the ray-sphere solves
 * near tangency (renderer.py:211-301, 397-432), the sky-view horizon split
 * (renderer.py:482-520) and the camera-relative surface endpoint above
 * 1000 km (renderer.py:1422-1472) -- plus Kahan-compensated optical depth
 * (renderer.py:570-607) and the segment integral (1 - e^-x) / sigma
 * (renderer.py:806-816) through expm1.  The batched LUT builds always use
 * REFERENCE. */
enum {
    ATMO_PRECISION_REFERENCE = 0,
    ATMO_PRECISION_ROBUST = 1
};
int32_t atmo_set_precision(AtmoHandle* handle, int32_t precision);

/* Launch-shape autotuner (no counterpart in the reference).  The
 * transmittance and sky-view builds march each texel with a group of G
 * lanes (G = 1, 2, 4, 8 or 16; default 8) that combine their step chunks;
 * atmo_autotune times every G of both builds on `stream` with this
 * handle's LUT sizes and the given arguments (as the handle's next update
 * would pass them), keeps the fastest, and leaves T, MS and SV built with
 * the chosen shapes.  It synchronises the stream.  chosen (may be NULL)
 * receives {G_T, G_SV}.  Different G reassociate the optical-depth sum and
 * the sky-view scan: results move by float32 rounding only. */
int32_t atmo_autotune(AtmoHandle* handle, float bottom_radius, const float irradiance[3],
                      float angular_radius, float camera_altitude, float sun_zenith_cosine,
                      void* stream, int32_t chosen[2]);
int32_t atmo_get_tuning(AtmoHandle* handle, int32_t tuning[2]);
int32_t atmo_set_tuning(AtmoHandle* handle, const int32_t tuning[2]);

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

/* Number of surface pixels the handle's last COMPOSITE-view call queued for
 * the plain integrator's second pass: the pixels that need the terminator
 * refinement in atmo_composite, the out-of-domain and routed pixels in
 * atmo_composite_surrogate.  Waits for `stream`.  No counterpart in the
 * reference (a diagnostic of this library's deferred-pixel queue). */
int32_t atmo_composite_queued(AtmoHandle* handle, int64_t* count, void* stream);

/* ------------------------------------------------------------------------
 * The reference's host methods in C.  `AtmosphereRenderer.update`,
 * `prepare_frame`, `composite` and `set_luts_frozen` (renderer.py:2057-2238)
 * take float64 host values, decide from invalidation keys which LUTs to
 * rebuild, and round every kernel argument to float32.  The entry points
 * below do the same for a C host, with the same keys, freeze rule,
 * snapshots and rebuild counters, so a host without Python reproduces the
 * reference's rebuild semantics exactly.  The policy itself is exported as
 * pure host functions over an AtmoLutCache (no device work), which a handle
 * keeps internally for atmo_update / atmo_prepare_frame / atmo_composite_frame.
 * The planet radius is a per-call argument, as in the reference, not a field
 * of AtmoParams (AtmosphereConfig has none: the radius comes from
 * PlanetModel, src/planet_renderer/planet.py).
 * --------------------------------------------------------------------- */

/* LUT bits: update's T, MS and SV, prepare_frame's AP. */
enum {
    ATMO_LUT_TRANSMITTANCE = 1,
    ATMO_LUT_MULTI_SCATTERING = 2,
    ATMO_LUT_SKY_VIEW = 4,
    ATMO_LUT_AERIAL = 8
};

/* The LUT cache state of `AtmosphereRenderer` (renderer.py:97-116).  Keys
 * compare with float64 ==, so a NaN key never matches. */
typedef struct AtmoLutCache {
    int32_t has_planet_key;            /* _planet_radius_key is not None */
    int32_t has_multi_scattering_key;  /* _multi_scattering_key is not None */
    int32_t has_sky_key;               /* _sky_key is not None */
    int32_t luts_frozen;
    double planet_key;                 /* planet radius (float64) */
    double multi_scattering_key[5];    /* R, E[0..2] (float64), float32(radians(alpha)) */
    double sky_key[6];                 /* float32 altitude, float32 sun cosine, E[0..2],
                                          float32(radians(alpha)) */
    int32_t has_sky_snapshot;          /* sky_snapshot_* are not None */
    int32_t aerial_available;          /* _aerial_available */
    double sky_snapshot_altitude_m;
    double sky_snapshot_sun_cosine;
    double aerial_snapshot_altitude_m;
    int64_t transmittance_rebuilds;
    int64_t multi_scattering_rebuilds;
    int64_t sky_view_rebuilds;
    int64_t aerial_rebuilds;
} AtmoLutCache;

/* What one `update` builds, with the float32 kernel arguments the
 * reference derives (renderer.py:2082-2137). */
typedef struct AtmoUpdatePlan {
    int32_t rebuild;             /* ATMO_LUT_* bits, built in the order T, MS, SV */
    float bottom_radius;         /* float32(R) */
    float solar_irradiance[3];   /* float32(E) */
    float sun_angular_radius;    /* float32(radians(alpha)) */
    float camera_altitude;       /* float32(max(r - R, 0)): the sky-view altitude key */
    float sun_zenith_cosine;     /* float32(clip(y of the normalised sun, -1, 1)) */
} AtmoUpdatePlan;

/* The state of a new AtmosphereRenderer: no keys, not frozen, no
 * snapshots, counters at 0. */
int32_t atmo_lut_cache_init(AtmoLutCache* cache);
/* `set_luts_frozen` (renderer.py:2057-2070): unfreezing drops the sky key. */
int32_t atmo_lut_cache_set_frozen(AtmoLutCache* cache, int32_t frozen);
/* The decisions of `update` (renderer.py:2072-2137): updates the keys,
 * snapshots and counters and fills `plan`.  force_mask (ATMO_LUT_* bits;
 * no counterpart in the reference) treats those keys as changed: forcing T
 * therefore rebuilds MS and SV too, as a new planet radius does, and a
 * forced SV is built even while frozen. */
int32_t atmo_lut_cache_update(AtmoLutCache* cache, double planet_radius_m,
                              double camera_radius_m, const double sun_local[3],
                              const double solar_irradiance[3],
                              double sun_angular_radius_degrees, int32_t force_mask,
                              AtmoUpdatePlan* plan);
/* The AP decision of `prepare_frame` (renderer.py:2173-2187): *rebuild_aerial
 * = 1 unless frozen with an AP already built (or when force_mask has
 * ATMO_LUT_AERIAL); on a rebuild, the snapshot altitude r - R and the
 * counter are updated. */
int32_t atmo_lut_cache_prepare(AtmoLutCache* cache, double planet_radius_m,
                               double camera_radius_m, int32_t force_mask,
                               int32_t* rebuild_aerial);
/* The sky-view altitude `composite` samples at (renderer.py:2210-2217): the
 * snapshot altitude if one exists, else max(r - R, 0). */
double atmo_lut_cache_sky_altitude(const AtmoLutCache* cache, double planet_radius_m,
                                   double camera_radius_m);

/* The handle's copy of the policy.  atmo_autotune and atmo_set_tuning drop
 * its keys (the LUTs were rebuilt with other arguments or shapes). */
int32_t atmo_get_lut_cache(AtmoHandle* handle, AtmoLutCache* out);
int32_t atmo_set_luts_frozen(AtmoHandle* handle, int32_t frozen);

/* `update` (renderer.py:2072-2137) on the handle's LUTs: the plan of
 * atmo_lut_cache_update, then atmo_build_transmittance /
 * _multi_scattering / _sky_view for its bits.  rebuilt (may be NULL)
 * receives the bits. */
int32_t atmo_update(AtmoHandle* handle, double planet_radius_m, double camera_radius_m,
                    const double sun_local[3], const double solar_irradiance[3],
                    double sun_angular_radius_degrees, int32_t force_mask, void* stream,
                    int32_t* rebuilt);
/* `prepare_frame` (renderer.py:2139-2187): surface lighting, then the AP
 * LUT unless frozen.  view_basis is (right, view_up, forward), 9 values. */
int32_t atmo_prepare_frame(AtmoHandle* handle, const float* position_view,
                           const int32_t* surface_id, double planet_radius_m,
                           double camera_radius_m, const double sun_local[3],
                           const double solar_irradiance[3],
                           double sun_angular_radius_degrees, const double view_basis[9],
                           double tangent_half_fov, int32_t force_mask, void* stream,
                           int32_t* rebuilt);
/* `composite` (renderer.py:2189-2238): atmo_composite with the float64
 * host arguments rounded as the reference rounds them and the sky-view LUT
 * sampled at the snapshot altitude. */
int32_t atmo_composite_frame(AtmoHandle* handle, const float* scene_hdr, const float* space_hdr,
                             const int32_t* surface_id, const float* position_view, float* hdr,
                             double planet_radius_m, double camera_radius_m,
                             const double sun_local[3], const double solar_irradiance[3],
                             const double sun_disk_radiance[3],
                             double sun_angular_radius_degrees, const double view_basis[9],
                             double tangent_half_fov, double star_contrast_start,
                             double star_contrast_end, int32_t diagnostic_view, void* stream);

/* No counterpart in the reference.  atmo_composite for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * hosts that forbid hipMalloc): stateless -- no handle; the caller passes
 * the configuration, the four LUTs (the layouts of ATMO_BUF_TRANSMITTANCE ..
 * ATMO_BUF_AERIAL_SCATTERING) and the frame size along with atmo_composite's
 * arguments -- and the deferred-pixel queue of the COMPOSITE view lives in
 * a library-owned static device array of ATMO_NOALLOC_MAX_PIXELS + 2 int32
 * (32 MiB per GPU, held from the library's first use of that GPU until the
 * process ends).  Reference precision.  Frames above ATMO_NOALLOC_MAX_PIXELS
 * pixels return ATMO_ESHAPE; errors are reported by atmo_last_error(NULL).
 * Calls on one GPU share that GPU's queue: calls that overlap in time on
 * one GPU (another stream, another thread) race.  Each COMPOSITE call clears
 * the queue's counters with a one-wave kernel first.  Results are bitwise
 * those of atmo_composite. */
#define ATMO_NOALLOC_MAX_PIXELS 8388608 /* 2^23: covers 3840 x 2160 */
int32_t atmo_composite_noalloc(
    const AtmoParams* params, const float* transmittance_lut, const float* multi_scattering_lut,
    const float* sky_view_lut, const float* aerial_lut, int32_t width, int32_t height,
    const float* scene_hdr, const float* space_hdr, const int32_t* surface_id,
    const float* position_view, float* hdr, float bottom_radius, float camera_altitude,
    float sky_lut_camera_altitude, const float sun_local[3], const float solar_irradiance[3],
    const float sun_disk_radiance[3], float sun_angular_radius, const float right[3],
    const float view_up[3], const float forward[3], float tangent_half_fov,
    float star_contrast_start, float star_contrast_end, int32_t diagnostic_view, void* stream);

/* ------------------------------------------------------------------------
 * Batched LUT builds.  The reference renders one atmosphere state per
 * AtmosphereRenderer; a batch builds the Transmittance, Multi-scattering,
 * Sky-view and Aerial-perspective LUTs of many independent states (optics,
 * planet, camera, sun) at once -- exactly what, per state, a cold `update`
 * (renderer.py:2072-2137) followed by `prepare_frame`'s aerial build
 * (renderer.py:2173-2187) would produce.  The four stage launches are
 * captured once into a HIP graph and replayed by atmo_batch_run.
 * --------------------------------------------------------------------- */

/* One state, in the reference's host units (float64, degrees); the library
 * derives the float32 kernel arguments as the reference's host methods
 * round them.  Every state of a batch must have the same LUT dimensions. */
typedef struct AtmoBatchState {
    AtmoParams params;
    double planet_radius_m;
    double camera_radius_m;
    double sun_local[3];
    double solar_irradiance[3];
    double sun_angular_radius_deg;
    double right[3];
    double view_up[3];
    double forward[3];
    double tangent_half_fov;
} AtmoBatchState;

typedef struct AtmoBatch AtmoBatch;

/* Validates every state (ATMO_EINVAL with the state's message, ATMO_ESHAPE
 * for mismatched LUT sizes).  luts[0..3] are caller-owned device arrays of
 * count copies of the T, MS, SV and AP LUTs (float32, the single-state
 * layouts back to back); a NULL entry (or luts == NULL) makes the library
 * allocate that array.  width / height are the frame size the
 * aerial-perspective rays use. */
int32_t atmo_batch_create(const AtmoBatchState* states, int32_t count, int32_t width,
                          int32_t height, float* const luts[4], AtmoBatch** out);
int32_t atmo_batch_destroy(AtmoBatch* batch);
const char* atmo_batch_last_error(const AtmoBatch* batch);
/* Builds all four LUTs of every state: one graph launch on `stream` (the
 * graph is captured on the first call). */
int32_t atmo_batch_run(AtmoBatch* batch, void* stream);
/* Launches a subset of the batch's four stage kernels directly on `stream`
 * (no graph): stages is an OR of ATMO_LUT_TRANSMITTANCE (1),
 * ATMO_LUT_MULTI_SCATTERING (2), ATMO_LUT_SKY_VIEW (4) and ATMO_LUT_AERIAL
 * (8).  A stage reads the LUTs its predecessors left in the batch arrays,
 * so after one atmo_batch_run any single stage can be timed on its own. */
int32_t atmo_batch_run_stages(AtmoBatch* batch, int32_t stages, void* stream);
/* Times the five lane-group sizes G_T of the batch's transmittance stage
 * alone, then the five G_SV of its sky-view stage alone, on `stream`
 * (synchronising it); keeps the fastest of each for the captured graph
 * (capturing again on the next atmo_batch_run if they changed) and writes
 * them to chosen (may be NULL).  The LUT arrays hold the last trial's
 * results. */
int32_t atmo_batch_autotune(AtmoBatch* batch, void* stream, int32_t chosen[2]);
/* Device address and shape of the batch's LUT array `which`
 * (ATMO_BUF_TRANSMITTANCE .. ATMO_BUF_AERIAL_SCATTERING): shape[0] = count,
 * then the single-state LUT shape. */
int32_t atmo_batch_get_buffer(AtmoBatch* batch, int32_t which, void** device_ptr,
                              int64_t shape[5]);

/* ------------------------------------------------------------------------
 * Learned surrogate of the composite's surface branch.  New in this library:
 * the reference has no learned path.  The surrogate replaces, per surface
 * pixel, the direct limb integral (_integrate_surface_scattering,
 * renderer.py:1667-1786) and the two-sample segment transmittance
 * (_surface_segment_transmittance, renderer.py:1603-1665) with a
 * 5 -> 64 -> 64 -> 64 -> 6 SiLU MLP; out-of-domain pixels keep the plain
 * integrator.
 * --------------------------------------------------------------------- */

/* Parameter count: W1 64x5, b1, W2 64x64, b2, W3 64x64, b3, W4 6x64, b4. */
#define ATMO_MLP_PARAMS 9094

/* Input normalisation, training domain and output decoding.
 * Inputs (camera altitude h, view-zenith cosine, sun-zenith cosine,
 * relative azimuth phi, segment length d), each mapped to [-1, 1]:
 * log(1 + h / altitude_scale_m) over [0, altitude_max_m]; the cosines raw
 * when cosine_stretch is 0, else asinh((c - c_h) / s) / asinh(2 / s) with
 * s = cosine_stretch and c_h the local horizon cosine at h over a planet of
 * planet_radius_m; phi / pi; log d over [distance_min_m, distance_max_m].
 * Outputs y (6) are standardised: v = y * out_scale + out_mean.
 * output_mode 1 (linear): L_c = max(v_c, 0) (in-scatter per unit
 * irradiance, c = 0..2), T_c = clamp(v_{3+c}, 0, 1).  output_mode 0 (log):
 * L_c = exp(v_c) - eps_inscatter, T_c = exp(v_{3+c}) - eps_transmittance.
 * output_mode 2 (floored log, trained on log(max(L, eps_inscatter))):
 * L_c = exp(v_c), zero below 1.5 eps_inscatter; T_c as in mode 0.
 * exact_transmittance = 1 makes the surrogate composite evaluate T with the
 * plain two-sample formula (renderer.py:1604-1665) instead of the network's
 * T outputs.
 * route_share in (0, 1) sends a surface pixel to the plain integrator when,
 * in any channel, the learned in-scatter E L exceeds that share of the
 * pixel E L + scene T: where in-scatter carries the pixel, the network's
 * relative error would show in it.  0 routes no pixel on this criterion. */
typedef struct AtmoSurrogateTransform {
    float altitude_max_m;
    float altitude_scale_m;
    float distance_min_m;
    float distance_max_m;
    float out_mean[6];
    float out_scale[6];
    float eps_inscatter;
    float eps_transmittance;
    int32_t output_mode;
    float planet_radius_m;
    float cosine_stretch;
    int32_t exact_transmittance;
    float route_share;
} AtmoSurrogateTransform;

typedef struct AtmoMlp AtmoMlp;

/* Device-resident network, Adam state and training scratch for batches of
 * up to max_batch samples. */
int32_t atmo_mlp_create(int32_t max_batch, AtmoMlp** out);
int32_t atmo_mlp_destroy(AtmoMlp* mlp);
const char* atmo_mlp_last_error(const AtmoMlp* mlp);
/* Host copies of the fp32 master parameters (count = ATMO_MLP_PARAMS).
 * set_params also resets the Adam moments and repacks the BF16 inference
 * weights. */
int32_t atmo_mlp_set_params(AtmoMlp* mlp, const float* host_params, int64_t count);
int32_t atmo_mlp_get_params(AtmoMlp* mlp, float* host_params, int64_t count);
int32_t atmo_mlp_set_transform(AtmoMlp* mlp, const AtmoSurrogateTransform* transform);
/* Raw samples (n, 5) = (h, view cos, sun cos, phi, d) -> encoded inputs. */
int32_t atmo_mlp_encode(AtmoMlp* mlp, const float* raw, int64_t n, float* x, void* stream);
/* FP32 training forward pass: x (B, 5) encoded -> y (B, 6) raw outputs. */
int32_t atmo_mlp_forward(AtmoMlp* mlp, const float* x, int32_t batch, float* y, void* stream);
/* Gradient of the weighted mean-square error
 * sum(weight * (y - target)^2) / (B * 6) over rows index[0..B) (index may
 * be NULL for rows 0..B) of x (.., 5), target (.., 6) and weight (.., 6;
 * NULL = all ones); grad_out (device, ATMO_MLP_PARAMS) and loss_out (device,
 * 1 float) may be NULL. */
int32_t atmo_mlp_gradients(AtmoMlp* mlp, const float* x, const float* target,
                           const float* weight, const int32_t* index, int32_t batch,
                           float* grad_out, float* loss_out, void* stream);
/* One Adam step (lr, betas 0.9 / 0.999, eps 1e-8; step counts from 1) on
 * the weighted mean-square error of atmo_mlp_gradients. */
int32_t atmo_mlp_train_step(AtmoMlp* mlp, const float* x, const float* target,
                            const float* weight, const int32_t* index, int32_t batch,
                            float lr, int32_t step, float* loss_out, void* stream);
/* Repacks the current parameters into BF16 MFMA fragments for inference. */
int32_t atmo_mlp_pack(AtmoMlp* mlp, void* stream);
/* BF16 MFMA inference (FP32 accumulation) on encoded inputs x (n, 5). */
int32_t atmo_mlp_infer(AtmoMlp* mlp, const float* x, int64_t n, float* y, void* stream);

/* Training labels from the plain integrator: raw samples (n, 5) ->
 * (n, 6) = in-scatter per unit irradiance (RGB) and transmittance (RGB) of
 * the composite's surface branch, using the handle's current T and MS
 * LUTs. */
int32_t atmo_surrogate_labels(AtmoHandle* handle, const float* raw, int64_t n,
                              float bottom_radius, float sun_angular_radius,
                              float* labels, void* stream);

/* The COMPOSITE view of atmo_composite with the surrogate in place of the
 * surface branch's in-scatter and transmittance for in-domain pixels; same
 * arguments as atmo_composite without diagnostic_view.  Out-of-domain
 * pixels, and pixels routed by the transform's route_share, are computed by
 * the plain integrator, bitwise as atmo_composite computes them. */
int32_t atmo_composite_surrogate(
    AtmoHandle* handle, AtmoMlp* mlp, const float* scene_hdr,
    const float* space_hdr, const int32_t* surface_id, const float* position_view,
    float* hdr, float bottom_radius, float camera_altitude,
    float sky_lut_camera_altitude, const float sun_local[3],
    const float solar_irradiance[3], const float sun_disk_radiance[3],
    float sun_angular_radius, const float right[3], const float view_up[3],
    const float forward[3], float tangent_half_fov, float star_contrast_start,
    float star_contrast_end, void* stream);

/* ------------------------------------------------------------------------
 * Display post-process: the reference's `PostProcessor`
 * (src/planet_renderer/postprocess.py:74-207), the step after the composite
 * (renderer.py:1729-1735).  Host policy stays in the caller, as
 * `PostProcessor.process` keeps it (postprocess.py:176-207): the exposure
 * multiplier 2**exposure_ev (1 without tone mapping) and the bloom switch
 * (bloom requested, bloom_passes > 0, bloom_strength > 0).
 * --------------------------------------------------------------------- */

/* Mirrors every field of `PostprocessConfig` (postprocess.py:10-19), in the
 * same order.  exposure_ev is only validated here: the caller derives the
 * exposure multiplier. */
typedef struct AtmoPostParams {
    double exposure_ev;
    double bloom_threshold;
    double bloom_knee;
    double bloom_clamp;
    double bloom_strength;
    int32_t bloom_passes;
} AtmoPostParams;

/* Handle-owned half-resolution fields of `PostProcessor.__init__`
 * (postprocess.py:84-86), (max((W+1)/2, 1), max((H+1)/2, 1), 3) f32. */
enum {
    ATMO_POST_BUF_BLOOM_A = 0,
    ATMO_POST_BUF_BLOOM_B = 1
};

typedef struct AtmoPost AtmoPost;

/* Range checks of `PostprocessConfig.__post_init__` (postprocess.py:21-37),
 * with the reference's ValueError messages. */
int32_t atmo_post_validate_params(const AtmoPostParams* params, char* message,
                                  int32_t message_capacity);
/* Replaces `PostProcessor.__init__(width, height, config)`
 * (postprocess.py:78-86). */
int32_t atmo_post_create(const AtmoPostParams* params, int32_t width, int32_t height,
                         AtmoPost** out);
int32_t atmo_post_destroy(AtmoPost* post);
const char* atmo_post_last_error(const AtmoPost* post);
int32_t atmo_post_get_buffer(AtmoPost* post, int32_t which, void** device_ptr,
                             int64_t shape[4]);
/* Same contract as atmo_bind_buffer, for ATMO_POST_BUF_*. */
int32_t atmo_post_bind_buffer(AtmoPost* post, int32_t which, void* device_ptr,
                              int64_t bytes);
/* Replaces the kernel sequence of `PostProcessor.process`
 * (postprocess.py:195-207): `_extract(hdr, exposure_multiplier)` and
 * bloom_passes x (`_blur_horizontal`, `_blur_vertical`) when use_bloom is
 * non-zero, then `_resolve(hdr, display, exposure_multiplier,
 * bloom_strength, use_bloom, tone_map)`.  hdr and display are (W,H,3) f32
 * device pointers; hdr is only read.  Leaves the bloom fields as the
 * reference does (bloom_a after the last vertical pass, bloom_b after the
 * last horizontal pass). */
int32_t atmo_post_process(AtmoPost* post, const float* hdr, float* display,
                          float exposure_multiplier, int32_t use_bloom, int32_t tone_map,
                          void* stream);

/* ------------------------------------------------------------------------
 * Surface shading: the reference's `PlanetRenderer._shade_surface`
 * (src/planet_renderer/renderer.py:1277-1575), the consumer of
 * prepare_frame's surface_sun_transmittance and surface_sky_radiance
 * (renderer.py:1682-1702).  Lambert terrain lighting, and for ocean pixels
 * (surface_id == 2147000000, ocean.py:17) with the advanced ocean enabled:
 * GGX sun specular, Fresnel sky reflection, a refracted screen-space seabed
 * lookup into the pre-ocean terrain snapshot, and water absorption and
 * scattering.
 * --------------------------------------------------------------------- */

/* The renderer's fields the kernel reads and writes (renderer.py:235-252),
 * all (W,H[,k]) device arrays in the Taichi field layout: f3 = 3 floats. */
typedef struct AtmoShadeFields {
    const int32_t* surface_id;          /* gbuffer_surface_id (W,H) */
    const float* albedo;                /* gbuffer_albedo f3 */
    const float* normal;                /* gbuffer_normal f3 (global) */
    const float* position;              /* gbuffer_position f3 (view) */
    const float* ocean_slope_variance;  /* gbuffer_ocean_slope_variance (W,H) */
    const float* water_depth_m;         /* gbuffer_water_depth_m (W,H) */
    const float* seabed_albedo;         /* gbuffer_seabed_albedo f3 */
    const float* seabed_normal;         /* gbuffer_seabed_normal f3 */
    const int32_t* terrain_surface_id;  /* terrain snapshot (W,H) */
    const float* terrain_position;      /* f3 */
    const float* terrain_albedo;        /* f3 */
    const float* terrain_normal;        /* f3 */
    const float* sun_transmittance;     /* prepare_frame output f3 */
    const float* sky_radiance;          /* prepare_frame output f3 */
    float* surface_hdr;                 /* output f3 */
} AtmoShadeFields;

/* The kernel's scalar arguments (renderer.py:1279-1300), float32 as the
 * `ti.f32` parameters; the caller rounds its float64 values. */
typedef struct AtmoShadeParams {
    float sun_global[3];
    float solar_irradiance[3];
    float view_right_global[3];
    float view_up_global[3];
    float view_forward_global[3];
    int32_t advanced_ocean_enabled;
    float ocean_roughness;
    float ocean_f0;
    float ocean_sky_strength;
    float ocean_absorption[3];
    float ocean_scattering[3];
    float ocean_max_depth;
    float ocean_refraction_index;
    float ocean_refraction_strength;
    float ocean_refraction_max_offset_pixels;
    float sun_angular_radius;   /* radians */
    float tangent_half_fov;
    int32_t mode;               /* debug view: 0 lit, otherwise albedo */
} AtmoShadeParams;

/* Replaces `_shade_surface` (renderer.py:1277-1575) on a W x H frame;
 * writes fields->surface_hdr for every pixel. */
int32_t atmo_shade_surface(const AtmoShadeFields* fields, const AtmoShadeParams* params,
                           int32_t width, int32_t height, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* ATMO_H */
