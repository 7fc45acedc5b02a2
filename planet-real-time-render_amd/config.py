"""Configuration contract of the spherical atmosphere.

Same fields, defaults, units (metres, inverse metres) and validation messages
as the reference's `AtmosphereConfig` (src/planet_renderer/atmosphere/
config.py:29-259), so a configuration built for the reference builds here
unchanged.  `to_params()` produces the `AtmoParams` struct of the C-ABI
(include/atmo.h), field for field.
"""

from __future__ import annotations

from dataclasses import dataclass, fields
import math
from typing import Mapping

Rgb = tuple[float, float, float]


def _rgb(value: object, name: str, default: Rgb) -> Rgb:
    """config.py:19-26"""
    source = default if value is None else value
    if not isinstance(source, (list, tuple)) or len(source) != 3:
        raise ValueError(f"atmosphere.{name} must contain three values")
    result = tuple(float(component) for component in source)
    if not all(math.isfinite(component) and component >= 0.0 for component in result):
        raise ValueError(f"atmosphere.{name} must be finite and non-negative")
    return result  # type: ignore[return-value]


# (field, low, high, message) -- config.py:97-128
_INT_RANGES = (
    ("transmittance_lut_width", 16, 2048, "transmittance LUT width must be in 16..2048"),
    ("transmittance_lut_height", 16, 1024, "transmittance LUT height must be in 16..1024"),
    ("sky_view_lut_width", 16, 2048, "sky-view LUT width must be in 16..2048"),
    ("sky_view_lut_height", 16, 2048, "sky-view LUT height must be in 16..2048"),
    ("multi_scattering_lut_width", 8, 256, "multi-scattering LUT width must be in 8..256"),
    ("multi_scattering_lut_height", 8, 256, "multi-scattering LUT height must be in 8..256"),
    ("aerial_lut_width", 4, 256, "aerial LUT width must be in 4..256"),
    ("aerial_lut_height", 4, 256, "aerial LUT height must be in 4..256"),
    ("aerial_lut_depth", 8, 256, "aerial LUT depth must be in 8..256"),
    ("transmittance_steps", 4, 256, "transmittance_steps must be in 4..256"),
    ("multi_scattering_directions", 4, 128, "multi_scattering_directions must be in 4..128"),
    ("multi_scattering_steps", 4, 128, "multi_scattering_steps must be in 4..128"),
    ("sky_view_steps", 4, 256, "sky_view_steps must be in 4..256"),
    ("aerial_steps_per_slice", 1, 16, "aerial_steps_per_slice must be in 1..16"),
    ("aerial_horizon_raymarch_steps", 4, 64, "aerial_horizon_raymarch_steps must be in 4..64"),
    ("aerial_terminator_substeps", 2, 16, "aerial_terminator_substeps must be in 2..16"),
)


@dataclass(frozen=True)
class AtmosphereConfig:
    """Optical parameters and LUT quality for one spherical atmosphere
    (config.py:29-59)."""

    top_altitude_m: float = 100_000.0
    rayleigh_scattering_per_m: Rgb = (5.802e-6, 13.558e-6, 33.100e-6)
    rayleigh_scale_height_m: float = 8_000.0
    mie_scattering_per_m: Rgb = (3.996e-6, 3.996e-6, 3.996e-6)
    mie_extinction_per_m: Rgb = (4.440e-6, 4.440e-6, 4.440e-6)
    mie_scale_height_m: float = 1_200.0
    mie_phase_g: float = 0.80
    absorption_extinction_per_m: Rgb = (0.650e-6, 1.881e-6, 0.085e-6)
    absorption_peak_altitude_m: float = 25_000.0
    absorption_half_width_m: float = 15_000.0
    ground_albedo: Rgb = (0.10, 0.10, 0.10)
    transmittance_lut_width: int = 256
    transmittance_lut_height: int = 64
    multi_scattering_lut_width: int = 32
    multi_scattering_lut_height: int = 32
    sky_view_lut_width: int = 192
    sky_view_lut_height: int = 256
    aerial_lut_width: int = 32
    aerial_lut_height: int = 18
    aerial_lut_depth: int = 32
    transmittance_steps: int = 40
    multi_scattering_directions: int = 16
    multi_scattering_steps: int = 12
    sky_view_steps: int = 24
    aerial_steps_per_slice: int = 2
    aerial_horizon_raymarch_steps: int = 12
    aerial_terminator_substeps: int = 8

    def __post_init__(self) -> None:
        """config.py:61-128"""
        groups = {
            "rayleigh_scattering_per_m": self.rayleigh_scattering_per_m,
            "mie_scattering_per_m": self.mie_scattering_per_m,
            "mie_extinction_per_m": self.mie_extinction_per_m,
            "absorption_extinction_per_m": self.absorption_extinction_per_m,
            "ground_albedo": self.ground_albedo,
        }
        for name, values in groups.items():
            if len(values) != 3 or not all(
                math.isfinite(value) and value >= 0.0 for value in values
            ):
                raise ValueError(f"atmosphere.{name} must contain three non-negative values")
        if any(value > 1.0 for value in self.ground_albedo):
            raise ValueError("atmosphere.ground_albedo must remain in [0, 1]")
        for name in (
            "top_altitude_m",
            "rayleigh_scale_height_m",
            "mie_scale_height_m",
            "absorption_half_width_m",
        ):
            value = getattr(self, name)
            if not math.isfinite(value) or value <= 0.0:
                raise ValueError(f"atmosphere.{name} must be positive and finite")
        if not 0.0 <= self.absorption_peak_altitude_m <= self.top_altitude_m:
            raise ValueError("atmosphere absorption peak must lie inside the shell")
        if not -0.99 < self.mie_phase_g < 0.99:
            raise ValueError("atmosphere.mie_phase_g must be in (-0.99, 0.99)")
        for scattering, extinction in zip(self.mie_scattering_per_m, self.mie_extinction_per_m):
            if scattering > extinction:
                raise ValueError("Mie scattering cannot exceed Mie extinction")
        for name, low, high, message in _INT_RANGES:
            if not low <= getattr(self, name) <= high:
                raise ValueError(message)

    @classmethod
    def from_mapping(cls, values: Mapping[str, object]) -> "AtmosphereConfig":
        """config.py:130-259: absent keys keep their defaults; RGB groups go
        through the same three-value check."""
        defaults = cls()
        kwargs: dict[str, object] = {}
        for f in fields(cls):
            default = getattr(defaults, f.name)
            if isinstance(default, tuple):
                kwargs[f.name] = _rgb(values.get(f.name), f.name, default)
            elif isinstance(default, int):
                kwargs[f.name] = int(values.get(f.name, default))  # type: ignore[arg-type]
            else:
                kwargs[f.name] = float(values.get(f.name, default))  # type: ignore[arg-type]
        return cls(**kwargs)  # type: ignore[arg-type]

    def replace(self, **changes: object) -> "AtmosphereConfig":
        values = {f.name: getattr(self, f.name) for f in fields(self)}
        values.update(changes)
        return AtmosphereConfig(**values)  # type: ignore[arg-type]

    def to_params(self):
        """Fill the C-ABI `AtmoParams` struct (include/atmo.h)."""
        from ._capi import AtmoParams

        params = AtmoParams()
        for f in fields(self):
            value = getattr(self, f.name)
            if isinstance(value, tuple):
                getattr(params, f.name)[:] = [float(v) for v in value]
            else:
                setattr(params, f.name, value)
        return params


__all__ = ["AtmosphereConfig", "Rgb"]
