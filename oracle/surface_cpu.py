"""ctypes wrapper of the surface-shading oracle (oracle/surface_cpu.c).
TEST INFRASTRUCTURE ONLY -- imported by tests/ and the CPU-baseline leg of
bench.py as the checker, never by the shipped package.

`shade(fields, params, width, height)` restates one call of the
reference's `_shade_surface` (src/planet_renderer/renderer.py:1277-1575)
on NumPy arrays in the Taichi field layout; it returns surface_hdr and the
branch each pixel took.
"""

from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

from .ref_cpu import build

ORACLE_DIR = Path(__file__).resolve().parent
LIBRARY = ORACLE_DIR / "libsurface_cpu.so"

FIELD_NAMES = ("surface_id", "albedo", "normal", "position", "ocean_slope_variance",
               "water_depth_m", "seabed_albedo", "seabed_normal", "terrain_surface_id",
               "terrain_position", "terrain_albedo", "terrain_normal", "sun_transmittance",
               "sky_radiance", "surface_hdr")
F3 = ctypes.c_float * 3


class ShadeFields(ctypes.Structure):
    """struct AtmoShadeFields (include/atmo.h)."""

    _fields_ = [(name, ctypes.c_void_p) for name in FIELD_NAMES]


class ShadeParams(ctypes.Structure):
    """struct AtmoShadeParams (include/atmo.h)."""

    _fields_ = [
        ("sun_global", F3), ("solar_irradiance", F3), ("view_right_global", F3),
        ("view_up_global", F3), ("view_forward_global", F3),
        ("advanced_ocean_enabled", ctypes.c_int32), ("ocean_roughness", ctypes.c_float),
        ("ocean_f0", ctypes.c_float), ("ocean_sky_strength", ctypes.c_float),
        ("ocean_absorption", F3), ("ocean_scattering", F3), ("ocean_max_depth", ctypes.c_float),
        ("ocean_refraction_index", ctypes.c_float), ("ocean_refraction_strength", ctypes.c_float),
        ("ocean_refraction_max_offset_pixels", ctypes.c_float),
        ("sun_angular_radius", ctypes.c_float), ("tangent_half_fov", ctypes.c_float),
        ("mode", ctypes.c_int32),
    ]


def make_params(params: dict) -> ShadeParams:
    """Python floats rounded to float32 as Taichi rounds f32 arguments."""
    p = ShadeParams()
    for name, ctype in ShadeParams._fields_:
        value = params[name]
        if ctype is F3:
            setattr(p, name, F3(*(float(v) for v in value)))
        else:
            setattr(p, name, value)
    return p


def _lib(library: Path | None = None) -> ctypes.CDLL:
    if library is None:
        if not LIBRARY.exists():
            build()
        library = LIBRARY
    lib = ctypes.CDLL(str(library))
    lib.surface_shade.argtypes = [ctypes.POINTER(ShadeFields), ctypes.POINTER(ShadeParams),
                                  ctypes.c_int32, ctypes.c_int32, ctypes.c_void_p]
    lib.surface_shade.restype = None
    return lib


def shade(fields: dict, params: dict, width: int, height: int, library: Path | None = None):
    """Returns (surface_hdr (W, H, 3) float32, path (W, H) int32).  `library`
    selects another build of surface_cpu.c (the tests' planted mistakes)."""
    lib = _lib(library)
    arrays = {k: np.ascontiguousarray(v) for k, v in fields.items()}
    out = np.empty((width, height, 3), np.float32)
    path = np.empty((width, height), np.int32)
    arrays["surface_hdr"] = out
    f = ShadeFields(**{name: arrays[name].ctypes.data for name in FIELD_NAMES})
    p = make_params(params)
    lib.surface_shade(ctypes.byref(f), ctypes.byref(p), width, height, path.ctypes.data)
    return out, path
