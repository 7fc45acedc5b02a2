"""Radiometric lighting state consumed by the atmosphere.

Same contract as the reference's `LightingState` / `StaticLightingProvider`
(src/planet_renderer/lighting.py:11-64): a unit sun direction, the solar
angular radius in degrees (0.01..5) and RGB irradiance; the visible disk's
radiance follows from E = L * pi * sin(alpha)^2.
"""

from __future__ import annotations

from dataclasses import dataclass
import math

import numpy as np


def _normalize(value) -> np.ndarray:
    vector = np.asarray(value, dtype=np.float64)
    length = float(np.linalg.norm(vector))
    if vector.shape != (3,) or not np.isfinite(vector).all() or length <= 0.0:
        raise ValueError("direction must be a finite non-zero 3-vector")
    return vector / length


@dataclass(frozen=True)
class LightingState:
    sun_direction_global: np.ndarray
    sun_angular_radius_degrees: float
    solar_irradiance: tuple[float, float, float]

    def __post_init__(self) -> None:
        """lighting.py:17-26"""
        object.__setattr__(self, "sun_direction_global", _normalize(self.sun_direction_global))
        if not 0.01 <= self.sun_angular_radius_degrees <= 5.0:
            raise ValueError("sun angular radius must be in 0.01..5 degrees")
        if len(self.solar_irradiance) != 3 or not all(
            math.isfinite(value) and value >= 0.0 for value in self.solar_irradiance
        ):
            raise ValueError("solar irradiance must contain three finite values")

    @property
    def sun_solid_angle_sr(self) -> float:
        """Solid angle of the solar disk (lighting.py:28-33)."""
        radius = math.radians(self.sun_angular_radius_degrees)
        return 2.0 * math.pi * (1.0 - math.cos(radius))

    @property
    def sun_projected_solid_angle_sr(self) -> float:
        """Cosine-weighted solid angle seen by a sun-facing surface
        (lighting.py:35-40)."""
        radius = math.radians(self.sun_angular_radius_degrees)
        return math.pi * math.sin(radius) ** 2

    @property
    def sun_disk_radiance(self) -> tuple[float, float, float]:
        """Uniform-disk radiance whose integral is the irradiance
        (lighting.py:42-54)."""
        projected = self.sun_projected_solid_angle_sr
        return tuple(value / projected for value in self.solar_irradiance)  # type: ignore[return-value]


@dataclass(frozen=True)
class StaticLightingProvider:
    state: LightingState

    def snapshot(self) -> LightingState:
        return self.state


__all__ = ["LightingState", "StaticLightingProvider"]
