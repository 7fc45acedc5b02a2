"""MI355X-native atmosphere hot path: LUT builds and per-pixel composite.

Drop-in for the reference's `planet_renderer.atmosphere` package
(src/planet_renderer/atmosphere/__init__.py): `AtmosphereConfig`,
`AtmosphereDiagnosticView`, `AtmosphereRenderer`, plus `LightingState`
(src/planet_renderer/lighting.py), and the display post-process that
follows the composite, `PostprocessConfig` / `PostProcessor`
(src/planet_renderer/postprocess.py), and the surface shading that consumes
`prepare_frame`'s lighting, `shade_surface` (renderer.py:1277-1575).
Compute runs in libatmo.so (HIP,
gfx950); this package is its host side.
"""

from .config import AtmosphereConfig
from .diagnostics import AtmosphereDiagnosticView
from .frame_io import OCEAN_SURFACE_ID, check_gbuffer, load_hdr, save_hdr
from .lighting import LightingState, StaticLightingProvider
from .pipeline import OceanShading, render_frame
from .postprocess import PostprocessConfig, PostProcessor
from .renderer import AtmosphereRenderer, DeviceField
from .shading import SurfaceFields, shade_surface

__all__ = [
    "AtmosphereConfig",
    "AtmosphereDiagnosticView",
    "AtmosphereRenderer",
    "DeviceField",
    "LightingState",
    "OCEAN_SURFACE_ID",
    "OceanShading",
    "PostProcessor",
    "PostprocessConfig",
    "StaticLightingProvider",
    "SurfaceFields",
    "check_gbuffer",
    "load_hdr",
    "render_frame",
    "save_hdr",
    "shade_surface",
]
