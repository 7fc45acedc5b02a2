"""Independent float64 formulations of the atmosphere path's integrals, used
to pin the CPU oracle (oracle/ref_cpu.c).  TEST INFRASTRUCTURE ONLY.

They restate what the reference defines -- the medium of `_density`,
`_extinction` and `_scattering` (src/planet_renderer/atmosphere/
renderer.py:169-209), the phases (renderer.py:792-804), the transfer
estimate of `_build_multi_scattering` (renderer.py:1066-1206: the mean over
the Fibonacci sphere of the sun-lit in-scatter L and of the scattering
fraction f, closed as L / (1 - clamp(f, 0, 0.95)), with a Lambertian ground
bounce), and the bounded single scattering that each slice of
`_build_aerial_perspective` accumulates (renderer.py:1305-1420) -- with a
different quadrature from the oracle's:

  * view paths: uniform midpoint sums of T(t) S(t) dt, with T(t) from the
    midpoint optical depth up to t (the oracle uses warped segments and the
    exact constant-extinction integral (1 - e^-x) / sigma per segment);
  * sun transmittance: its own uniform midpoint optical depth from each
    point to the top of the shell, zero behind an opaque ground (the oracle
    samples a transmittance LUT bilinearly, scaled by finite-disk
    visibility);
  * ray intervals: the validation model's ray-sphere roots
    (oracle/model_fp64.py, src/planet_renderer/atmosphere/model.py:38-119).

Where the sun stands well above every point's horizon both definitions
describe the same integral, so the oracle must converge to these values as
its step counts and LUT sizes grow.
"""

from __future__ import annotations

import math

import numpy as np

from oracle.model_fp64 import AtmosphereModel

GOLDEN_ANGLE = 2.39996323  # renderer.py:1104 (the Fibonacci-sphere azimuth step)


class Medium:
    """Density, extinction and scattering of renderer.py:169-209."""

    def __init__(self, cfg):
        self.cfg = cfg
        self.top = float(cfg.top_altitude_m)
        self.beta_r = np.asarray(cfg.rayleigh_scattering_per_m, dtype=np.float64)
        self.beta_ms = np.asarray(cfg.mie_scattering_per_m, dtype=np.float64)
        self.beta_me = np.asarray(cfg.mie_extinction_per_m, dtype=np.float64)
        self.beta_a = np.asarray(cfg.absorption_extinction_per_m, dtype=np.float64)

    def densities(self, h):
        h = np.asarray(h, dtype=np.float64)
        inside = (h >= 0.0) & (h <= self.top)
        hc = np.where(inside, h, 0.0)
        c = self.cfg
        r = np.where(inside, np.exp(-hc / c.rayleigh_scale_height_m), 0.0)
        m = np.where(inside, np.exp(-hc / c.mie_scale_height_m), 0.0)
        a = np.where(inside, np.maximum(
            1.0 - np.abs(hc - c.absorption_peak_altitude_m) / c.absorption_half_width_m, 0.0),
            0.0)
        return r, m, a

    def extinction(self, h):
        r, m, a = self.densities(h)
        return r[..., None] * self.beta_r + m[..., None] * self.beta_me + a[..., None] * self.beta_a

    def scattering(self, h):
        r, m, _ = self.densities(h)
        return r[..., None] * self.beta_r + m[..., None] * self.beta_ms


def rayleigh_phase(c):
    return 3.0 * (1.0 + c * c) / (16.0 * math.pi)


def mie_phase(c, g):
    den = np.maximum(1.0 + g * g - 2.0 * g * c, 1.0e-8)
    return 3.0 * (1.0 - g * g) * (1.0 + c * c) / (8.0 * math.pi * (2.0 + g * g) * den ** 1.5)


def sun_transmittance(points, sun, planet_radius, medium: Medium, steps: int = 400):
    """Transmittance from each point (n, 3) toward the sun to the top of the
    shell: uniform midpoint optical depth, zero where the planet blocks the
    path."""
    p = np.asarray(points, dtype=np.float64).reshape(-1, 3)
    s = np.asarray(sun, dtype=np.float64)
    top = planet_radius + medium.top
    b = p @ s
    pp = np.einsum("ij,ij->i", p, p)
    disc = b * b - (pp - top * top)
    far = -b + np.sqrt(np.maximum(disc, 0.0))
    length = np.maximum(far, 0.0)
    # the planet blocks the sun if the ray meets the ground sphere ahead
    gdisc = b * b - (pp - planet_radius * planet_radius)
    ground_near = -b - np.sqrt(np.maximum(gdisc, 0.0))
    blocked = (gdisc > 0.0) & (ground_near > 1.0e-3)
    frac = (np.arange(steps) + 0.5) / steps
    t = length[:, None] * frac[None, :]
    q = p[:, None, :] + t[..., None] * s[None, None, :]
    h = np.linalg.norm(q, axis=2) - planet_radius
    tau = medium.extinction(h).sum(axis=1) * (length / steps)[:, None]
    out = np.exp(-np.minimum(tau, 80.0))
    out[blocked] = 0.0
    return out


def path_integral(origin, ray, t0, t1, sun, irradiance, planet_radius, medium: Medium,
                  steps: int = 4000, sun_steps: int = 400):
    """Uniform midpoint quadrature along origin + t ray, t in [t0, t1], of
    the sun-lit single-scattering in-scatter L = int T_view E T_sun
    (beta_R rho_R P_R + beta_Ms rho_M P_M) dt and of the scattering fraction
    F = int T_view sigma_s dt.  Returns (L, F, T_view(t1))."""
    o = np.asarray(origin, dtype=np.float64)
    d = np.asarray(ray, dtype=np.float64)
    s = np.asarray(sun, dtype=np.float64)
    zero = np.zeros(3)
    if not t1 > t0:
        return zero, zero, np.ones(3)
    dt = (t1 - t0) / steps
    t = t0 + (np.arange(steps) + 0.5) * dt
    points = o[None, :] + t[:, None] * d[None, :]
    h = np.linalg.norm(points, axis=1) - planet_radius
    ext = medium.extinction(h)
    tau = (np.cumsum(ext, axis=0) - 0.5 * ext) * dt
    view_t = np.exp(-tau)
    c = float(np.clip(d @ s, -1.0, 1.0))
    r, m, _ = medium.densities(h)
    phase = (r[:, None] * medium.beta_r * rayleigh_phase(c)
             + m[:, None] * medium.beta_ms * mie_phase(c, medium.cfg.mie_phase_g))
    st = sun_transmittance(points, s, planet_radius, medium, sun_steps)
    e = np.asarray(irradiance, dtype=np.float64)
    radiance = (view_t * st * phase).sum(axis=0) * e * dt
    feedback = (view_t * medium.scattering(h)).sum(axis=0) * dt
    end_t = np.exp(-ext.sum(axis=0) * dt)
    return radiance, feedback, end_t


def fibonacci_directions(count: int) -> np.ndarray:
    """The direction set of renderer.py:1098-1108: cos = 1 - 2 (i + 1/2) / N,
    azimuth = i * 2.39996323."""
    i = np.arange(count, dtype=np.float64)
    dc = 1.0 - 2.0 * (i + 0.5) / count
    ds = np.sqrt(np.maximum(1.0 - dc * dc, 0.0))
    az = i * GOLDEN_ANGLE
    return np.stack([ds * np.cos(az), dc, ds * np.sin(az)], axis=1)


def multi_scattering_texel(cfg, planet_radius, altitude, sun_cosine, irradiance,
                           steps: int = 2000, sun_steps: int = 300):
    """Psi of one multi-scattering texel at (altitude, sun cosine):
    mean L over the directions, closed by 1 / (1 - clamp(mean F, 0, 0.95)),
    where each direction adds the Lambertian ground bounce
    albedo E T_sun(ground) max(n . s, 0) / pi seen through the path's
    transmittance.  Returns (Psi, mean L, mean F)."""
    medium = Medium(cfg)
    model = AtmosphereModel(planet_radius, cfg)
    origin = np.array([0.0, planet_radius + altitude, 0.0])
    sun = np.array([math.sqrt(max(1.0 - sun_cosine * sun_cosine, 0.0)), sun_cosine, 0.0])
    e = np.asarray(irradiance, dtype=np.float64)
    albedo = np.asarray(cfg.ground_albedo, dtype=np.float64)
    dirs = fibonacci_directions(cfg.multi_scattering_directions)
    sum_l = np.zeros(3)
    sum_f = np.zeros(3)
    for ray in dirs:
        interval = model.atmosphere_interval(origin, ray)
        if interval is None:
            continue
        lum, feedback, end_t = path_integral(origin, ray, interval.start_m, interval.end_m, sun,
                                             e, planet_radius, medium, steps, sun_steps)
        if interval.ends_at_ground:
            ground = origin + ray * interval.end_m
            normal = ground / np.linalg.norm(ground)
            gsc = float(normal @ sun)
            # the sun path starts on the ground sphere: lift it by a millimetre
            gt = sun_transmittance((ground + normal * 1.0e-3)[None], sun, planet_radius, medium,
                                   sun_steps)[0]
            lum = lum + end_t * albedo * e * gt * max(gsc, 0.0) / math.pi
        sum_l += lum
        sum_f += feedback
    n = len(dirs)
    mean_l, mean_f = sum_l / n, sum_f / n
    psi = np.maximum(mean_l / (1.0 - np.clip(mean_f, 0.0, 0.95)), 0.0)
    return psi, mean_l, mean_f


__all__ = ["Medium", "fibonacci_directions", "multi_scattering_texel", "path_integral",
           "sun_transmittance", "rayleigh_phase", "mie_phase"]
