"""The atmosphere part of one reference frame, end to end on the GPU.

`render_frame` runs the calls `PlanetRenderer.render` makes after the
G-buffer exists (src/planet_renderer/renderer.py:1664-1735), in the same
order and with the same argument derivation:

  AtmosphereRenderer.update -> prepare_frame -> _shade_surface ->
  composite -> PostProcessor.process

The rasteriser, the ocean raster pass and the space background that fill
the G-buffer and `space_hdr` stay with the caller.  `OceanShading` holds
the `OceanConfig` fields `_shade_surface` reads (ocean.py:46-65), with the
reference's defaults; any object with those attribute names can stand in
for it.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

from .diagnostics import AtmosphereDiagnosticView
from .shading import shade_surface


@dataclass(frozen=True)
class OceanShading:
    """The shading-relevant fields of `OceanConfig` (ocean.py:46-65)."""

    enabled: bool = True
    roughness: float = 0.12
    sky_reflection_strength: float = 1.0
    dielectric_f0: float = 0.02
    absorption_per_m: tuple[float, float, float] = (0.14, 0.055, 0.022)
    scattering_per_m: tuple[float, float, float] = (0.0025, 0.005, 0.012)
    max_visible_depth_m: float = 80.0
    refraction_index: float = 1.333
    refraction_strength: float = 1.0
    refraction_max_offset_pixels: float = 48.0


def render_frame(atmosphere, postprocessor, fields, space_hdr, hdr, display, *,
                 planet_radius_m: float, camera_radius_m: float, sun_local, lighting,
                 view_basis, tangent_half_fov: float, star_contrast, exposure_ev: float,
                 sun_global=None, view_basis_global=None, ocean=None,
                 diagnostic_view=AtmosphereDiagnosticView.COMPOSITE, debug_view: int = 0) -> None:
    """renderer.py:1664-1735 from the G-buffer on.

    atmosphere: an AtmosphereRenderer; postprocessor: a PostProcessor;
    fields: the renderer's shading fields (shading.SurfaceFields or any
    object with the same attributes), whose `surface_hdr` becomes the
    composite's scene radiance; space_hdr, hdr, display: (W, H, 3) device
    tensors.  view_basis is (right, up, forward) in the camera's local
    frame; sun_global and view_basis_global default to the local ones (a
    planet frame that coincides with the local frame).  Asynchronous on
    the current stream."""
    ocean = ocean if ocean is not None else OceanShading()
    sun_global = sun_local if sun_global is None else sun_global
    right_g, up_g, forward_g = view_basis if view_basis_global is None else view_basis_global
    irradiance = lighting.solar_irradiance
    alpha_deg = lighting.sun_angular_radius_degrees
    atmosphere.update(planet_radius_m, camera_radius_m, sun_local, irradiance, alpha_deg)
    atmosphere.prepare_frame(fields.gbuffer_position, fields.gbuffer_surface_id, planet_radius_m,
                             camera_radius_m, sun_local, irradiance, alpha_deg, view_basis,
                             tangent_half_fov)
    # renderer.py:1682-1702
    shade_surface(fields, tuple(sun_global), irradiance, atmosphere.surface_sun_transmittance,
                  atmosphere.surface_sky_radiance, tuple(right_g), tuple(up_g), tuple(forward_g),
                  int(ocean.enabled), ocean.roughness, ocean.dielectric_f0,
                  ocean.sky_reflection_strength, ocean.absorption_per_m, ocean.scattering_per_m,
                  ocean.max_visible_depth_m, ocean.refraction_index, ocean.refraction_strength,
                  ocean.refraction_max_offset_pixels, math.radians(alpha_deg), tangent_half_fov,
                  debug_view)
    c0, c1 = star_contrast
    atmosphere.composite(fields.surface_hdr, space_hdr, fields.gbuffer_surface_id,
                         fields.gbuffer_position, hdr, planet_radius_m, camera_radius_m, sun_local,
                         irradiance, lighting.sun_disk_radiance, alpha_deg, view_basis,
                         tangent_half_fov, c0, c1, diagnostic_view=diagnostic_view)
    # renderer.py:1728-1735
    view = AtmosphereDiagnosticView(int(diagnostic_view))
    postprocessor.process(hdr, display, exposure_ev, bloom=view.uses_bloom,
                          tone_map=view.is_radiance)


__all__ = ["OceanShading", "render_frame"]
