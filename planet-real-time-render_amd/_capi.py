"""ctypes binding of libatmo.so (include/atmo.h).

The library is the only compute path: when it is missing or cannot be
loaded, every entry point raises instead of falling back to anything else.
"""

from __future__ import annotations

import ctypes
from pathlib import Path

PACKAGE_DIR = Path(__file__).resolve().parent
LIBRARY_PATH = PACKAGE_DIR / "libatmo.so"

ATMO_OK = 0
ATMO_EINVAL = 1
ATMO_ESHAPE = 2
ATMO_EHIP = 3
ATMO_ENOMEM = 4
ATMO_EHANDLE = 5

BUF_TRANSMITTANCE = 0
BUF_MULTI_SCATTERING = 1
BUF_SKY_VIEW = 2
BUF_AERIAL_SCATTERING = 3
BUF_SURFACE_SUN_T = 4
BUF_SURFACE_SKY_L = 5

PRECISION_REFERENCE = 0
PRECISION_ROBUST = 1

MAX_PIXELS = 268435456  # ATMO_MAX_PIXELS
NOALLOC_MAX_PIXELS = 8388608  # ATMO_NOALLOC_MAX_PIXELS

# Every symbol include/atmo.h declares, checked by the CPU test suite.
EXPORTED_SYMBOLS = (
    "atmo_default_params",
    "atmo_validate_params",
    "atmo_create",
    "atmo_destroy",
    "atmo_last_error",
    "atmo_get_buffer",
    "atmo_bind_buffer",
    "atmo_set_precision",
    "atmo_autotune",
    "atmo_get_tuning",
    "atmo_set_tuning",
    "atmo_build_transmittance",
    "atmo_build_multi_scattering",
    "atmo_build_sky_view",
    "atmo_build_aerial_perspective",
    "atmo_prepare_surface_lighting",
    "atmo_composite",
    "atmo_composite_queued",
    "atmo_composite_noalloc",
    "atmo_lut_cache_init",
    "atmo_lut_cache_set_frozen",
    "atmo_lut_cache_update",
    "atmo_lut_cache_prepare",
    "atmo_lut_cache_sky_altitude",
    "atmo_get_lut_cache",
    "atmo_set_luts_frozen",
    "atmo_update",
    "atmo_prepare_frame",
    "atmo_composite_frame",
    "atmo_batch_create",
    "atmo_batch_destroy",
    "atmo_batch_last_error",
    "atmo_batch_run",
    "atmo_batch_run_stages",
    "atmo_batch_autotune",
    "atmo_batch_get_buffer",
    "atmo_mlp_create",
    "atmo_mlp_destroy",
    "atmo_mlp_last_error",
    "atmo_mlp_set_params",
    "atmo_mlp_get_params",
    "atmo_mlp_set_transform",
    "atmo_mlp_encode",
    "atmo_mlp_forward",
    "atmo_mlp_gradients",
    "atmo_mlp_train_step",
    "atmo_mlp_pack",
    "atmo_mlp_infer",
    "atmo_surrogate_labels",
    "atmo_composite_surrogate",
    "atmo_post_validate_params",
    "atmo_post_create",
    "atmo_post_destroy",
    "atmo_post_last_error",
    "atmo_post_get_buffer",
    "atmo_post_bind_buffer",
    "atmo_post_process",
    "atmo_shade_surface",
)

POST_BUF_BLOOM_A = 0
POST_BUF_BLOOM_B = 1

MLP_PARAMS = 9094


class AtmoSurrogateTransform(ctypes.Structure):
    """struct AtmoSurrogateTransform (include/atmo.h)."""

    _fields_ = [
        ("altitude_max_m", ctypes.c_float),
        ("altitude_scale_m", ctypes.c_float),
        ("distance_min_m", ctypes.c_float),
        ("distance_max_m", ctypes.c_float),
        ("out_mean", ctypes.c_float * 6),
        ("out_scale", ctypes.c_float * 6),
        ("eps_inscatter", ctypes.c_float),
        ("eps_transmittance", ctypes.c_float),
        ("output_mode", ctypes.c_int32),
        ("planet_radius_m", ctypes.c_float),
        ("cosine_stretch", ctypes.c_float),
        ("exact_transmittance", ctypes.c_int32),
        ("route_share", ctypes.c_float),
    ]


LUT_TRANSMITTANCE = 1
LUT_MULTI_SCATTERING = 2
LUT_SKY_VIEW = 4
LUT_AERIAL = 8


class AtmoLutCache(ctypes.Structure):
    """Mirror of struct AtmoLutCache (include/atmo.h)."""

    _fields_ = [
        ("has_planet_key", ctypes.c_int32),
        ("has_multi_scattering_key", ctypes.c_int32),
        ("has_sky_key", ctypes.c_int32),
        ("luts_frozen", ctypes.c_int32),
        ("planet_key", ctypes.c_double),
        ("multi_scattering_key", ctypes.c_double * 5),
        ("sky_key", ctypes.c_double * 6),
        ("has_sky_snapshot", ctypes.c_int32),
        ("aerial_available", ctypes.c_int32),
        ("sky_snapshot_altitude_m", ctypes.c_double),
        ("sky_snapshot_sun_cosine", ctypes.c_double),
        ("aerial_snapshot_altitude_m", ctypes.c_double),
        ("transmittance_rebuilds", ctypes.c_int64),
        ("multi_scattering_rebuilds", ctypes.c_int64),
        ("sky_view_rebuilds", ctypes.c_int64),
        ("aerial_rebuilds", ctypes.c_int64),
    ]


class AtmoUpdatePlan(ctypes.Structure):
    """Mirror of struct AtmoUpdatePlan (include/atmo.h)."""

    _fields_ = [
        ("rebuild", ctypes.c_int32),
        ("bottom_radius", ctypes.c_float),
        ("solar_irradiance", ctypes.c_float * 3),
        ("sun_angular_radius", ctypes.c_float),
        ("camera_altitude", ctypes.c_float),
        ("sun_zenith_cosine", ctypes.c_float),
    ]


def d3(values) -> "ctypes.Array[ctypes.c_double]":
    """Three float64 values for the host-method entry points."""
    return (ctypes.c_double * 3)(*[float(v) for v in values])


class AtmoParams(ctypes.Structure):
    """struct AtmoParams (include/atmo.h), field for field."""

    _fields_ = [
        ("top_altitude_m", ctypes.c_double),
        ("rayleigh_scattering_per_m", ctypes.c_double * 3),
        ("rayleigh_scale_height_m", ctypes.c_double),
        ("mie_scattering_per_m", ctypes.c_double * 3),
        ("mie_extinction_per_m", ctypes.c_double * 3),
        ("mie_scale_height_m", ctypes.c_double),
        ("mie_phase_g", ctypes.c_double),
        ("absorption_extinction_per_m", ctypes.c_double * 3),
        ("absorption_peak_altitude_m", ctypes.c_double),
        ("absorption_half_width_m", ctypes.c_double),
        ("ground_albedo", ctypes.c_double * 3),
        ("transmittance_lut_width", ctypes.c_int32),
        ("transmittance_lut_height", ctypes.c_int32),
        ("multi_scattering_lut_width", ctypes.c_int32),
        ("multi_scattering_lut_height", ctypes.c_int32),
        ("sky_view_lut_width", ctypes.c_int32),
        ("sky_view_lut_height", ctypes.c_int32),
        ("aerial_lut_width", ctypes.c_int32),
        ("aerial_lut_height", ctypes.c_int32),
        ("aerial_lut_depth", ctypes.c_int32),
        ("transmittance_steps", ctypes.c_int32),
        ("multi_scattering_directions", ctypes.c_int32),
        ("multi_scattering_steps", ctypes.c_int32),
        ("sky_view_steps", ctypes.c_int32),
        ("aerial_steps_per_slice", ctypes.c_int32),
        ("aerial_horizon_raymarch_steps", ctypes.c_int32),
        ("aerial_terminator_substeps", ctypes.c_int32),
    ]


class AtmoPostParams(ctypes.Structure):
    """struct AtmoPostParams (include/atmo.h), field for field."""

    _fields_ = [
        ("exposure_ev", ctypes.c_double),
        ("bloom_threshold", ctypes.c_double),
        ("bloom_knee", ctypes.c_double),
        ("bloom_clamp", ctypes.c_double),
        ("bloom_strength", ctypes.c_double),
        ("bloom_passes", ctypes.c_int32),
    ]


SHADE_FIELD_NAMES = ("surface_id", "albedo", "normal", "position", "ocean_slope_variance",
                     "water_depth_m", "seabed_albedo", "seabed_normal", "terrain_surface_id",
                     "terrain_position", "terrain_albedo", "terrain_normal", "sun_transmittance",
                     "sky_radiance", "surface_hdr")


class AtmoShadeFields(ctypes.Structure):
    """struct AtmoShadeFields (include/atmo.h): device pointers."""

    _fields_ = [(name, ctypes.c_void_p) for name in SHADE_FIELD_NAMES]


class AtmoShadeParams(ctypes.Structure):
    """struct AtmoShadeParams (include/atmo.h), field for field."""

    _fields_ = [
        ("sun_global", ctypes.c_float * 3),
        ("solar_irradiance", ctypes.c_float * 3),
        ("view_right_global", ctypes.c_float * 3),
        ("view_up_global", ctypes.c_float * 3),
        ("view_forward_global", ctypes.c_float * 3),
        ("advanced_ocean_enabled", ctypes.c_int32),
        ("ocean_roughness", ctypes.c_float),
        ("ocean_f0", ctypes.c_float),
        ("ocean_sky_strength", ctypes.c_float),
        ("ocean_absorption", ctypes.c_float * 3),
        ("ocean_scattering", ctypes.c_float * 3),
        ("ocean_max_depth", ctypes.c_float),
        ("ocean_refraction_index", ctypes.c_float),
        ("ocean_refraction_strength", ctypes.c_float),
        ("ocean_refraction_max_offset_pixels", ctypes.c_float),
        ("sun_angular_radius", ctypes.c_float),
        ("tangent_half_fov", ctypes.c_float),
        ("mode", ctypes.c_int32),
    ]


class AtmoBatchState(ctypes.Structure):
    """struct AtmoBatchState (include/atmo.h): one state in host units
    (float64, degrees)."""

    _fields_ = [
        ("params", AtmoParams),
        ("planet_radius_m", ctypes.c_double),
        ("camera_radius_m", ctypes.c_double),
        ("sun_local", ctypes.c_double * 3),
        ("solar_irradiance", ctypes.c_double * 3),
        ("sun_angular_radius_deg", ctypes.c_double),
        ("right", ctypes.c_double * 3),
        ("view_up", ctypes.c_double * 3),
        ("forward", ctypes.c_double * 3),
        ("tangent_half_fov", ctypes.c_double),
    ]


F3 = ctypes.c_float * 3
_P = ctypes.c_void_p
_F = ctypes.c_float
_I = ctypes.c_int32
_PF3 = ctypes.POINTER(ctypes.c_float)


def _signatures(lib: ctypes.CDLL) -> None:
    lib.atmo_default_params.argtypes = [ctypes.POINTER(AtmoParams)]
    lib.atmo_validate_params.argtypes = [ctypes.POINTER(AtmoParams), ctypes.c_char_p, _I]
    lib.atmo_create.argtypes = [ctypes.POINTER(AtmoParams), _I, _I, ctypes.POINTER(_P)]
    lib.atmo_destroy.argtypes = [_P]
    lib.atmo_last_error.argtypes = [_P]
    lib.atmo_last_error.restype = ctypes.c_char_p
    lib.atmo_get_buffer.argtypes = [_P, _I, ctypes.POINTER(_P), ctypes.POINTER(ctypes.c_int64)]
    lib.atmo_bind_buffer.argtypes = [_P, _I, _P, ctypes.c_int64]
    lib.atmo_set_precision.argtypes = [_P, _I]
    _PI = ctypes.POINTER(ctypes.c_int32)
    lib.atmo_autotune.argtypes = [_P, _F, _PF3, _F, _F, _F, _P, _PI]
    lib.atmo_get_tuning.argtypes = [_P, _PI]
    lib.atmo_set_tuning.argtypes = [_P, _PI]
    lib.atmo_build_transmittance.argtypes = [_P, _F, _P]
    lib.atmo_build_multi_scattering.argtypes = [_P, _F, _PF3, _F, _P]
    lib.atmo_build_sky_view.argtypes = [_P, _F, _F, _F, _PF3, _F, _P]
    lib.atmo_build_aerial_perspective.argtypes = [
        _P, _F, _F, _PF3, _PF3, _F, _PF3, _PF3, _PF3, _F, _P,
    ]
    lib.atmo_prepare_surface_lighting.argtypes = [
        _P, _P, _P, _F, _F, _PF3, _F, _PF3, _PF3, _PF3, _P,
    ]
    lib.atmo_composite.argtypes = [
        _P, _P, _P, _P, _P, _P, _F, _F, _F, _PF3, _PF3, _PF3, _F, _PF3, _PF3, _PF3,
        _F, _F, _F, _I, _P,
    ]
    lib.atmo_composite_queued.argtypes = [_P, ctypes.POINTER(ctypes.c_int64), _P]
    lib.atmo_composite_noalloc.argtypes = [
        ctypes.POINTER(AtmoParams), _P, _P, _P, _P, _I, _I,
        _P, _P, _P, _P, _P, _F, _F, _F, _PF3, _PF3, _PF3, _F, _PF3, _PF3, _PF3,
        _F, _F, _F, _I, _P,
    ]
    _PI = ctypes.POINTER(ctypes.c_int32)
    _PD3 = ctypes.POINTER(ctypes.c_double)
    _D = ctypes.c_double
    _PC = ctypes.POINTER(AtmoLutCache)
    lib.atmo_lut_cache_init.argtypes = [_PC]
    lib.atmo_lut_cache_set_frozen.argtypes = [_PC, _I]
    lib.atmo_lut_cache_update.argtypes = [_PC, _D, _D, _PD3, _PD3, _D, _I,
                                          ctypes.POINTER(AtmoUpdatePlan)]
    lib.atmo_lut_cache_prepare.argtypes = [_PC, _D, _D, _I, _PI]
    lib.atmo_lut_cache_sky_altitude.argtypes = [_PC, _D, _D]
    lib.atmo_lut_cache_sky_altitude.restype = _D
    lib.atmo_get_lut_cache.argtypes = [_P, _PC]
    lib.atmo_set_luts_frozen.argtypes = [_P, _I]
    lib.atmo_update.argtypes = [_P, _D, _D, _PD3, _PD3, _D, _I, _P, _PI]
    lib.atmo_prepare_frame.argtypes = [_P, _P, _P, _D, _D, _PD3, _PD3, _D, _PD3, _D, _I, _P, _PI]
    lib.atmo_composite_frame.argtypes = [_P, _P, _P, _P, _P, _P, _D, _D, _PD3, _PD3, _PD3, _D,
                                         _PD3, _D, _D, _D, _I, _P]
    lib.atmo_batch_create.argtypes = [ctypes.POINTER(AtmoBatchState), _I, _I, _I,
                                      ctypes.POINTER(_P), ctypes.POINTER(_P)]
    lib.atmo_batch_destroy.argtypes = [_P]
    lib.atmo_batch_last_error.argtypes = [_P]
    lib.atmo_batch_last_error.restype = ctypes.c_char_p
    lib.atmo_batch_run.argtypes = [_P, _P]
    lib.atmo_batch_run_stages.argtypes = [_P, _I, _P]
    lib.atmo_batch_autotune.argtypes = [_P, _P, ctypes.POINTER(ctypes.c_int32)]
    lib.atmo_batch_get_buffer.argtypes = [_P, _I, ctypes.POINTER(_P), ctypes.POINTER(ctypes.c_int64)]
    lib.atmo_mlp_create.argtypes = [_I, ctypes.POINTER(_P)]
    lib.atmo_mlp_destroy.argtypes = [_P]
    lib.atmo_mlp_last_error.argtypes = [_P]
    lib.atmo_mlp_last_error.restype = ctypes.c_char_p
    lib.atmo_mlp_set_params.argtypes = [_P, _PF3, ctypes.c_int64]
    lib.atmo_mlp_get_params.argtypes = [_P, _PF3, ctypes.c_int64]
    lib.atmo_mlp_set_transform.argtypes = [_P, ctypes.POINTER(AtmoSurrogateTransform)]
    lib.atmo_mlp_encode.argtypes = [_P, _P, ctypes.c_int64, _P, _P]
    lib.atmo_mlp_forward.argtypes = [_P, _P, _I, _P, _P]
    lib.atmo_mlp_gradients.argtypes = [_P, _P, _P, _P, _P, _I, _P, _P, _P]
    lib.atmo_mlp_train_step.argtypes = [_P, _P, _P, _P, _P, _I, _F, _I, _P, _P]
    lib.atmo_mlp_pack.argtypes = [_P, _P]
    lib.atmo_mlp_infer.argtypes = [_P, _P, ctypes.c_int64, _P, _P]
    lib.atmo_surrogate_labels.argtypes = [_P, _P, ctypes.c_int64, _F, _F, _P, _P]
    lib.atmo_composite_surrogate.argtypes = [
        _P, _P, _P, _P, _P, _P, _P, _F, _F, _F, _PF3, _PF3, _PF3, _F, _PF3, _PF3, _PF3,
        _F, _F, _F, _P,
    ]
    lib.atmo_post_validate_params.argtypes = [ctypes.POINTER(AtmoPostParams), ctypes.c_char_p, _I]
    lib.atmo_post_create.argtypes = [ctypes.POINTER(AtmoPostParams), _I, _I, ctypes.POINTER(_P)]
    lib.atmo_post_destroy.argtypes = [_P]
    lib.atmo_post_last_error.argtypes = [_P]
    lib.atmo_post_last_error.restype = ctypes.c_char_p
    lib.atmo_post_get_buffer.argtypes = [_P, _I, ctypes.POINTER(_P), ctypes.POINTER(ctypes.c_int64)]
    lib.atmo_post_bind_buffer.argtypes = [_P, _I, _P, ctypes.c_int64]
    lib.atmo_post_process.argtypes = [_P, _P, _P, _F, _I, _I, _P]
    lib.atmo_shade_surface.argtypes = [ctypes.POINTER(AtmoShadeFields),
                                       ctypes.POINTER(AtmoShadeParams), _I, _I, _P]
    for name in EXPORTED_SYMBOLS:
        if name not in ("atmo_last_error", "atmo_mlp_last_error", "atmo_batch_last_error",
                        "atmo_post_last_error", "atmo_lut_cache_sky_altitude"):
            getattr(lib, name).restype = ctypes.c_int32


_LIB: ctypes.CDLL | None = None


def library() -> ctypes.CDLL:
    """Load libatmo.so once; raise if it has not been built."""
    global _LIB
    if _LIB is None:
        if not LIBRARY_PATH.exists():
            raise RuntimeError(
                f"{LIBRARY_PATH} is missing: build the HIP library first "
                "(__graft_entry__.build() or python -m planet_real_time_render_amd.build)"
            )
        # torch's HIP runtime initialises first: when libatmo.so's HIP
        # runtime is loaded and used before it, torch finds no device.
        import torch

        torch.cuda.is_available()
        lib = ctypes.CDLL(str(LIBRARY_PATH))
        _signatures(lib)
        _LIB = lib
    return _LIB


def f3(values) -> "ctypes.Array[ctypes.c_float]":
    """Round three Python floats to float32, as Taichi rounds f32 kernel
    arguments."""
    a, b, c = (float(v) for v in values)
    return F3(a, b, c)


class AtmoError(RuntimeError):
    pass


def check(status: int, handle=None) -> None:
    if status == ATMO_OK:
        return
    message = library().atmo_last_error(handle).decode(errors="replace")
    if status in (ATMO_EINVAL, ATMO_ESHAPE):
        raise ValueError(message)
    raise AtmoError(f"atmo status {status}: {message}")


__all__ = [
    "AtmoBatchState",
    "AtmoPostParams",
    "AtmoShadeFields",
    "AtmoShadeParams",
    "AtmoParams",
    "AtmoError",
    "EXPORTED_SYMBOLS",
    "LIBRARY_PATH",
    "check",
    "f3",
    "library",
]
