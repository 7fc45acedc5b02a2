"""Surface shading over libatmo.so: the reference's
`PlanetRenderer._shade_surface` (src/planet_renderer/renderer.py:1277-1575),
the consumer of `AtmosphereRenderer.prepare_frame`'s surface lighting.

`shade_surface(fields, ...)` takes the kernel's own argument list
(renderer.py:1279-1300; called at renderer.py:1682-1702).  `fields` is any
object with the renderer's field attributes -- `gbuffer_surface_id`,
`gbuffer_albedo`, `gbuffer_normal`, `gbuffer_position`,
`gbuffer_ocean_slope_variance`, `gbuffer_water_depth_m`,
`gbuffer_seabed_albedo`, `gbuffer_seabed_normal`, the terrain snapshot
`terrain_surface_id`, `terrain_position`, `terrain_albedo`,
`terrain_normal`, and the output `surface_hdr` -- as torch tensors (or
DeviceFields) on the HIP device in the Taichi field layout; `SurfaceFields`
allocates such a set.  Scalars are rounded to float32 as Taichi rounds
`ti.f32` arguments.  The kernel is libatmo.so's atmo_shade_surface.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass

from . import _capi
from ._capi import library
from .renderer import DeviceField, _device_tensor, _torch

# renderer field name -> AtmoShadeFields member, channels (0 = scalar field)
_FIELDS = (
    ("gbuffer_surface_id", "surface_id", 0),
    ("gbuffer_albedo", "albedo", 3),
    ("gbuffer_normal", "normal", 3),
    ("gbuffer_position", "position", 3),
    ("gbuffer_ocean_slope_variance", "ocean_slope_variance", 0),
    ("gbuffer_water_depth_m", "water_depth_m", 0),
    ("gbuffer_seabed_albedo", "seabed_albedo", 3),
    ("gbuffer_seabed_normal", "seabed_normal", 3),
    ("terrain_surface_id", "terrain_surface_id", 0),
    ("terrain_position", "terrain_position", 3),
    ("terrain_albedo", "terrain_albedo", 3),
    ("terrain_normal", "terrain_normal", 3),
    ("surface_hdr", "surface_hdr", 3),
)
_INT_FIELDS = {"gbuffer_surface_id", "terrain_surface_id"}


@dataclass
class SurfaceFields:
    """The renderer fields `_shade_surface` reads and writes
    (renderer.py:235-252), as DeviceFields."""

    width: int
    height: int
    gbuffer_surface_id: DeviceField
    gbuffer_albedo: DeviceField
    gbuffer_normal: DeviceField
    gbuffer_position: DeviceField
    gbuffer_ocean_slope_variance: DeviceField
    gbuffer_water_depth_m: DeviceField
    gbuffer_seabed_albedo: DeviceField
    gbuffer_seabed_normal: DeviceField
    terrain_surface_id: DeviceField
    terrain_position: DeviceField
    terrain_albedo: DeviceField
    terrain_normal: DeviceField
    surface_hdr: DeviceField

    @classmethod
    def allocate(cls, width: int, height: int, device=None) -> "SurfaceFields":
        torch = _torch()
        device = torch.device(device if device is not None else "cuda")
        made = {}
        for name, _, channels in _FIELDS:
            shape = (width, height, channels) if channels else (width, height)
            dtype = torch.int32 if name in _INT_FIELDS else torch.float32
            made[name] = DeviceField(torch.zeros(shape, dtype=dtype, device=device))
        return cls(width=width, height=height, **made)

    def snapshot_terrain(self) -> None:
        """`_snapshot_terrain_gbuffer` (renderer.py:1270-1275): copy the
        pre-ocean G-buffer into the terrain snapshot."""
        self.terrain_position.tensor.copy_(self.gbuffer_position.tensor)
        self.terrain_normal.tensor.copy_(self.gbuffer_normal.tensor)
        self.terrain_albedo.tensor.copy_(self.gbuffer_albedo.tensor)
        self.terrain_surface_id.tensor.copy_(self.gbuffer_surface_id.tensor)


def _f3(values) -> "ctypes.Array[ctypes.c_float]":
    a, b, c = (float(v) for v in values)
    return (ctypes.c_float * 3)(a, b, c)


def shade_surface(fields, sun_global, solar_irradiance, sun_transmittance, sky_radiance,
                  view_right_global, view_up_global, view_forward_global,
                  advanced_ocean_enabled: int, ocean_roughness: float, ocean_f0: float,
                  ocean_sky_strength: float, ocean_absorption, ocean_scattering,
                  ocean_max_depth: float, ocean_refraction_index: float,
                  ocean_refraction_strength: float, ocean_refraction_max_offset_pixels: float,
                  sun_angular_radius: float, tangent_half_fov: float, mode: int) -> None:
    """renderer.py:1277-1575 on `fields` (see the module docstring);
    sun_transmittance and sky_radiance are prepare_frame's (W, H, 3)
    outputs.  Asynchronous on the current stream."""
    torch = _torch()
    first = getattr(fields, "gbuffer_surface_id")
    first_t = getattr(first, "tensor", first)
    if not isinstance(first_t, torch.Tensor) or first_t.dim() != 2:
        raise ValueError("gbuffer_surface_id must be a (W, H) device tensor or DeviceField")
    width, height = (int(v) for v in first_t.shape)
    device = first_t.device
    if device.type != "cuda":
        raise ValueError("the shading fields must live on the HIP device")
    ptrs = {}
    for name, member, channels in _FIELDS:
        shape = (width, height, channels) if channels else (width, height)
        dtype = torch.int32 if name in _INT_FIELDS else torch.float32
        ptrs[member] = _device_tensor(getattr(fields, name), name, shape, dtype, device).data_ptr()
    for value, member in ((sun_transmittance, "sun_transmittance"), (sky_radiance, "sky_radiance")):
        ptrs[member] = _device_tensor(value, member, (width, height, 3), torch.float32,
                                      device).data_ptr()
    f = _capi.AtmoShadeFields(**ptrs)
    p = _capi.AtmoShadeParams(
        _f3(sun_global), _f3(solar_irradiance), _f3(view_right_global), _f3(view_up_global),
        _f3(view_forward_global), int(advanced_ocean_enabled), float(ocean_roughness),
        float(ocean_f0), float(ocean_sky_strength), _f3(ocean_absorption), _f3(ocean_scattering),
        float(ocean_max_depth), float(ocean_refraction_index), float(ocean_refraction_strength),
        float(ocean_refraction_max_offset_pixels), float(sun_angular_radius),
        float(tangent_half_fov), int(mode))
    stream = torch.cuda.current_stream(device).cuda_stream
    status = library().atmo_shade_surface(ctypes.byref(f), ctypes.byref(p), width, height, stream)
    if status != _capi.ATMO_OK:
        if status == _capi.ATMO_EINVAL:
            raise ValueError("atmo_shade_surface: invalid argument")
        raise _capi.AtmoError(f"atmo_shade_surface: status {status}")


__all__ = ["SurfaceFields", "shade_surface"]
