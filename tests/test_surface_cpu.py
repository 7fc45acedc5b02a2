"""Pinning of the surface-shading oracle (oracle/surface_cpu.c), the float32
restatement of the reference's `_shade_surface`
(src/planet_renderer/renderer.py:1277-1575), by cases the mathematics fixes:

  - sky pixels keep the constant sky colour; debug modes show the albedo;
  - Lambert: a white surface facing the sun returns E / pi, and a surface
    facing away returns albedo x sky;
  - the ocean at normal incidence (sun, normal and view aligned): GGX peak
    D = 1 / (pi alpha^2), Smith G = 1, Fresnel = f0, so the sun specular is
    E f0 / (4 pi alpha^2); plus the single-scattering water layer
    E / (4 pi) (s / sigma) (1 - e^(-sigma d)) (1 - f0);
  - a uniformly lit water layer over a seabed of albedo a:
    L (a e^(-sigma d) + (s / sigma)(1 - e^(-sigma d)));
  - a water surface facing the viewer refracts the ray straight on, so
    the screen-space seabed lookup lands on the pixel's own terrain
    sample: every pixel's seabed albedo, recovered from the layer formula,
    is its own terrain albedo (this pins the projection and the field
    indexing of the lookup).
Oblique incidence (GGX, Smith G and Fresnel away from the normal, Snell
bending, the offset clamp) is pinned in test_surface_pins.py.
"""

import math

import numpy as np
import pytest

from oracle.surface_cpu import shade

OCEAN = 2_147_000_000
E = (4.0, 3.9, 3.7)
ABSORPTION = (0.14, 0.055, 0.022)
SCATTERING = (0.0025, 0.005, 0.012)


def base_params(**changes):
    p = {
        "sun_global": (0.0, 1.0, 0.0), "solar_irradiance": E,
        "view_right_global": (1.0, 0.0, 0.0), "view_up_global": (0.0, 1.0, 0.0),
        "view_forward_global": (0.0, 0.0, 1.0), "advanced_ocean_enabled": 1,
        "ocean_roughness": 0.12, "ocean_f0": 0.02, "ocean_sky_strength": 1.0,
        "ocean_absorption": ABSORPTION, "ocean_scattering": SCATTERING,
        "ocean_max_depth": 80.0, "ocean_refraction_index": 1.333,
        "ocean_refraction_strength": 1.0, "ocean_refraction_max_offset_pixels": 48.0,
        "sun_angular_radius": math.radians(0.266), "tangent_half_fov": math.tan(math.radians(30)),
        "mode": 0,
    }
    p.update(changes)
    return p


def fields(w, h, **values):
    f = {
        "surface_id": np.zeros((w, h), np.int32),
        "albedo": np.full((w, h, 3), 0.5, np.float32),
        "normal": np.tile(np.float32([0, 1, 0]), (w, h, 1)),
        "position": np.tile(np.float32([0, 0, 10]), (w, h, 1)),
        "ocean_slope_variance": np.zeros((w, h), np.float32),
        "water_depth_m": np.full((w, h), 5.0, np.float32),
        "seabed_albedo": np.full((w, h, 3), 0.3, np.float32),
        "seabed_normal": np.tile(np.float32([0, 1, 0]), (w, h, 1)),
        "terrain_surface_id": np.zeros((w, h), np.int32),
        "terrain_position": np.tile(np.float32([0, 0, 20]), (w, h, 1)),
        "terrain_albedo": np.full((w, h, 3), 0.2, np.float32),
        "terrain_normal": np.tile(np.float32([0, 1, 0]), (w, h, 1)),
        "sun_transmittance": np.ones((w, h, 3), np.float32),
        "sky_radiance": np.zeros((w, h, 3), np.float32),
    }
    for k, v in values.items():
        f[k] = np.array(np.broadcast_to(np.asarray(v, f[k].dtype), f[k].shape), copy=True)
    return f


def test_sky_pixels_and_debug_modes():
    f = fields(3, 2, surface_id=-1)
    out, path = shade(f, base_params(), 3, 2)
    np.testing.assert_array_equal(out, np.broadcast_to(np.float32([0.00015, 0.0002, 0.00035]), out.shape))
    assert (path == 0).all()
    f = fields(3, 2, albedo=(0.25, -0.5, 2.0))
    out, _ = shade(f, base_params(mode=3), 3, 2)
    np.testing.assert_array_equal(out, np.broadcast_to(np.float32([0.25, 0.0, 2.0]), out.shape))


def test_lambert_closed_forms():
    f = fields(2, 2, albedo=(1.0, 1.0, 1.0), normal=(0.0, 2.0, 0.0))  # unnormalised normal
    out, path = shade(f, base_params(), 2, 2)
    np.testing.assert_allclose(out, np.broadcast_to(np.float32(E) / np.pi, out.shape), rtol=1e-6)
    assert (path == 1).all()
    # facing away from the sun: only the sky term, albedo x sky
    f = fields(2, 2, normal=(0.0, -1.0, 0.0), sky_radiance=(0.1, 0.2, 0.3), albedo=(0.5, 0.25, 1.0))
    out, _ = shade(f, base_params(), 2, 2)
    np.testing.assert_allclose(out[0, 0], (0.05, 0.05, 0.3), rtol=1e-6)


def layer(sky, a_bed, d):
    sigma = np.add(ABSORPTION, SCATTERING)
    t = np.exp(-sigma * d)
    return sky * (a_bed * t + np.divide(SCATTERING, sigma) * (1.0 - t))


def test_ocean_at_normal_incidence():
    """Sun, water normal and view all along -z: GGX peak plus the
    single-scattering water layer lit by E / (4 pi)."""
    n = (0.0, 0.0, -1.0)
    slope = 0.004
    f = fields(1, 1, surface_id=OCEAN, normal=n, position=(0.0, 0.0, 10.0), water_depth_m=7.0,
               seabed_normal=(0.0, 0.0, 1.0), ocean_slope_variance=slope)
    p = base_params(sun_global=n, ocean_refraction_strength=0.0)
    out, path = shade(f, p, 1, 1)
    assert path[0, 0] == 2
    alpha2 = 0.12 ** 4 + slope + math.tan(math.radians(0.266)) ** 2
    f0 = 0.02
    specular = np.float64(E) * f0 / (4.0 * math.pi * alpha2)
    sigma = np.add(ABSORPTION, SCATTERING)
    scatter = np.float64(E) / (4 * math.pi) * np.divide(SCATTERING, sigma) * (1 - np.exp(-sigma * 7.0))
    np.testing.assert_allclose(out[0, 0], specular + scatter * (1.0 - f0), rtol=2e-5)


def test_uniformly_lit_water_layer():
    n = (0.0, 0.0, -1.0)
    f = fields(1, 1, surface_id=OCEAN, normal=n, position=(0.0, 0.0, 10.0), water_depth_m=23.0,
               sun_transmittance=0.0, sky_radiance=(0.2, 0.3, 0.5), seabed_albedo=(0.4, 0.3, 0.2))
    p = base_params(ocean_f0=0.0, ocean_refraction_strength=0.0)
    out, _ = shade(f, p, 1, 1)
    np.testing.assert_allclose(out[0, 0], layer(np.float64([0.2, 0.3, 0.5]), np.float64([0.4, 0.3, 0.2]), 23.0),
                               rtol=2e-6)
    # the depth is capped at ocean_max_depth
    f["water_depth_m"][...] = 500.0
    out, _ = shade(f, p, 1, 1)
    np.testing.assert_allclose(out[0, 0], layer(np.float64([0.2, 0.3, 0.5]), np.float64([0.4, 0.3, 0.2]), 80.0),
                               rtol=2e-6)


@pytest.mark.parametrize("w, h", [(64, 36), (41, 23)])
def test_straight_refraction_reads_the_own_terrain_sample(w, h):
    """Water facing the viewer: every pixel's seabed comes from its own
    terrain-snapshot sample (unique albedo per pixel)."""
    thf = math.tan(math.radians(30))
    aspect = w / h
    sx = ((np.arange(w) + 0.5) / w * 2 - 1) * aspect
    sy = (np.arange(h) + 0.5) / h * 2 - 1
    rays = np.stack(np.broadcast_arrays(sx[:, None], sy[None, :], 1.0 / thf), axis=-1)
    rays /= np.linalg.norm(rays, axis=-1, keepdims=True)
    t, d = 30.0, 12.0
    rng = np.random.default_rng(5)
    bed = rng.uniform(0.05, 0.95, (w, h, 3)).astype(np.float32)
    f = fields(w, h, surface_id=OCEAN, sun_transmittance=0.0, sky_radiance=1.0)
    f["normal"] = (-rays).astype(np.float32)  # view basis = identity: global = view
    f["position"] = (rays * t).astype(np.float32)
    f["terrain_position"] = (rays * (t + d)).astype(np.float32)
    f["terrain_albedo"] = bed
    f["water_depth_m"][...] = d
    out, path = shade(f, base_params(ocean_f0=0.0, tangent_half_fov=thf), w, h)
    assert (path == 3).all()
    sigma = np.add(ABSORPTION, SCATTERING)
    trans = np.exp(-sigma * d)
    recovered = (out - np.divide(SCATTERING, sigma) * (1 - trans)) / trans
    np.testing.assert_allclose(recovered, bed, rtol=0, atol=2e-4)
