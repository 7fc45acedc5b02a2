"""Float64 NumPy restatement of the reference's validation model
(src/planet_renderer/atmosphere/model.py:19-213).  TEST INFRASTRUCTURE ONLY.

It is the reference's own FP64 oracle for shell intersections, density,
extinction, midpoint transmittance, phase functions and direct single
scattering.  It has no multiple scattering and no finite solar disk: the sun
transmittance is binary at an opaque ground.  The tests pin it with the
reference tests' known answers (tests/test_m3_atmosphere.py:24-99) and with
closed forms (exponential-layer optical depth, phase normalisation).
"""

from __future__ import annotations

from dataclasses import dataclass
import math

import numpy as np


@dataclass(frozen=True)
class RayInterval:
    """model.py:19-27"""

    start_m: float
    end_m: float
    ends_at_ground: bool = False

    @property
    def length_m(self) -> float:
        return max(self.end_m - self.start_m, 0.0)


def unit(value) -> np.ndarray:
    """model.py:30-35"""
    v = np.asarray(value, dtype=np.float64)
    length = float(np.linalg.norm(v))
    if v.shape != (3,) or not np.isfinite(v).all() or length <= 1.0e-15:
        raise ValueError("atmosphere directions must be finite non-zero vectors")
    return v / length


def ray_sphere_roots(origin, direction, radius_m: float):
    """Ordered ray parameters of a planet-centred sphere, or None
    (model.py:38-55)."""
    p = np.asarray(origin, dtype=np.float64)
    d = unit(direction)
    b = float(p @ d)
    disc = b * b - (float(p @ p) - radius_m * radius_m)
    if disc < -1.0e-12 * max(radius_m * radius_m, 1.0):
        return None
    s = math.sqrt(max(disc, 0.0))
    return -b - s, -b + s


class AtmosphereModel:
    """model.py:58-213, planet-centred float64 coordinates."""

    def __init__(self, planet_radius_m: float, config):
        if not math.isfinite(planet_radius_m) or planet_radius_m <= 1.0:
            raise ValueError("planet radius must be positive and finite")
        self.bottom_radius_m = float(planet_radius_m)
        self.top_radius_m = self.bottom_radius_m + config.top_altitude_m
        self.config = config

    def density(self, altitude_m: float):
        """model.py:66-78; zero outside [0, top]."""
        c = self.config
        if altitude_m < 0.0 or altitude_m > c.top_altitude_m:
            return 0.0, 0.0, 0.0
        h = float(altitude_m)
        return (
            math.exp(-h / c.rayleigh_scale_height_m),
            math.exp(-h / c.mie_scale_height_m),
            max(1.0 - abs(h - c.absorption_peak_altitude_m) / c.absorption_half_width_m, 0.0),
        )

    def extinction(self, altitude_m: float) -> np.ndarray:
        """model.py:80-87"""
        r, m, a = self.density(altitude_m)
        c = self.config
        return (np.asarray(c.rayleigh_scattering_per_m) * r
                + np.asarray(c.mie_extinction_per_m) * m
                + np.asarray(c.absorption_extinction_per_m) * a)

    def atmosphere_interval(self, origin, direction, maximum_distance_m: float = math.inf):
        """model.py:89-119"""
        p = np.asarray(origin, dtype=np.float64)
        d = unit(direction)
        outer = ray_sphere_roots(p, d, self.top_radius_m)
        if outer is None or outer[1] <= 0.0:
            return None
        start = max(outer[0], 0.0)
        end = min(outer[1], float(maximum_distance_m))
        ends_at_ground = False
        ground = ray_sphere_roots(p, d, self.bottom_radius_m)
        if ground is not None:
            eps = max(1.0e-6, self.bottom_radius_m * 1.0e-12)
            ahead = [t for t in ground if t > start + eps]
            if ahead and min(ahead) < end:
                end = min(ahead)
                ends_at_ground = True
        if end <= start:
            return None
        return RayInterval(start, end, ends_at_ground)

    def extinction_array(self, altitude_m: np.ndarray) -> np.ndarray:
        """extinction() for an array of altitudes, shape (..., 3)."""
        c = self.config
        h = np.asarray(altitude_m, dtype=np.float64)
        inside = (h >= 0.0) & (h <= c.top_altitude_m)
        hc = np.where(inside, h, 0.0)
        r = np.where(inside, np.exp(-hc / c.rayleigh_scale_height_m), 0.0)
        m = np.where(inside, np.exp(-hc / c.mie_scale_height_m), 0.0)
        a = np.where(inside, np.maximum(
            1.0 - np.abs(hc - c.absorption_peak_altitude_m) / c.absorption_half_width_m, 0.0), 0.0)
        return (r[..., None] * np.asarray(c.rayleigh_scattering_per_m)
                + m[..., None] * np.asarray(c.mie_extinction_per_m)
                + a[..., None] * np.asarray(c.absorption_extinction_per_m))

    def transmittance(self, origin, direction, maximum_distance_m: float = math.inf,
                      steps: int = 128, ground_is_opaque: bool = False) -> np.ndarray:
        """Uniform midpoint optical depth (model.py:121-143); the midpoint sum
        is evaluated as one array sum."""
        interval = self.atmosphere_interval(origin, direction, maximum_distance_m)
        if interval is None:
            return np.ones(3)
        if ground_is_opaque and interval.ends_at_ground:
            return np.zeros(3)
        d = unit(direction)
        n = max(int(steps), 1)
        step = interval.length_m / n
        p = np.asarray(origin, dtype=np.float64)
        t = interval.start_m + (np.arange(n) + 0.5) * step
        h = np.linalg.norm(p[None, :] + t[:, None] * d[None, :], axis=1) - self.bottom_radius_m
        tau = (self.extinction_array(h) * step).sum(axis=0)
        return np.exp(-np.minimum(tau, 80.0))

    @staticmethod
    def rayleigh_phase(cosine: float) -> float:
        """model.py:145-147"""
        return 3.0 * (1.0 + cosine * cosine) / (16.0 * math.pi)

    def mie_phase(self, cosine: float) -> float:
        """Cornette-Shanks (model.py:149-157)."""
        g = self.config.mie_phase_g
        den = max(1.0 + g * g - 2.0 * g * cosine, 1.0e-8)
        return 3.0 * (1.0 - g * g) * (1.0 + cosine * cosine) / (
            8.0 * math.pi * (2.0 + g * g) * den ** 1.5)

    def single_scattering(self, origin, view_direction, sun_direction, solar_irradiance,
                          steps: int = 64, sun_steps: int = 64) -> np.ndarray:
        """Direct single scattering along the visible shell segment
        (model.py:159-213)."""
        interval = self.atmosphere_interval(origin, view_direction)
        if interval is None:
            return np.zeros(3)
        view = unit(view_direction)
        sun = unit(sun_direction)
        p0 = np.asarray(origin, dtype=np.float64)
        n = max(int(steps), 1)
        step = interval.length_m / n
        vt = np.ones(3)
        radiance = np.zeros(3)
        c = float(np.clip(view @ sun, -1.0, 1.0))
        pr, pm = self.rayleigh_phase(c), self.mie_phase(c)
        e = np.asarray(solar_irradiance, dtype=np.float64)
        cfg = self.config
        for i in range(n):
            t = interval.start_m + (i + 0.5) * step
            point = p0 + view * t
            h = float(np.linalg.norm(point) - self.bottom_radius_m)
            dr, dm, _ = self.density(h)
            st = self.transmittance(point, sun, steps=sun_steps, ground_is_opaque=True)
            source = e * st * (np.asarray(cfg.rayleigh_scattering_per_m) * dr * pr
                               + np.asarray(cfg.mie_scattering_per_m) * dm * pm)
            ext = self.extinction(h)
            seg_t = np.exp(-np.minimum(ext * step, 80.0))
            integral = np.where(ext > 1.0e-12, (1.0 - seg_t) / np.where(ext > 0, ext, 1.0), step)
            radiance += vt * source * integral
            vt *= seg_t
        return np.maximum(radiance, 0.0)


__all__ = ["AtmosphereModel", "RayInterval", "ray_sphere_roots", "unit"]
