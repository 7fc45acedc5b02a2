"""Seeded synthetic workloads for the atmosphere path.

The reference feeds the atmosphere from its rasterizer (G-buffer position in
view coordinates and surface id, src/planet_renderer/renderer.py:1641-1727)
and its camera (view basis, camera.py:24-33).  This module restates only
that call-site contract and generates the inputs of the measured
configurations with NumPy `Generator(PCG64(20261203 + config_index))`:

  config 0  LUT chain: T 256x64/40, MS 32x32 with 64 directions x 20 steps,
            SV 192x108/30, no ozone, Mie extinction 4.40e-6, camera at 2 km,
            sun zenith U[0, 100] deg.
  config 1  full 1920x1080 frame: config 0 plus ozone and a 32x32x32 aerial
            volume; camera altitude logU[10 m, 50 km]; G-buffer with 30 % sky
            pixels and surface ranges logU[10 m, 500 km].
"""

from __future__ import annotations

from dataclasses import dataclass
import math

import numpy as np

from .config import AtmosphereConfig
from .lighting import LightingState

SEED_BASE = 20261203
EARTH_RADIUS_M = 6_360_000.0
SOLAR_IRRADIANCE = (4.0, 3.9, 3.7)  # src/planet_renderer/config.py:29
SUN_ANGULAR_RADIUS_DEG = 0.266  # src/planet_renderer/config.py:28
VERTICAL_FOV_DEG = 60.0  # src/planet_renderer/config.py:25
STAR_CONTRAST = (4.0, 12.0)  # configs/earth.json:29-30
SKY_SCENE_HDR = (0.00015, 0.0002, 0.00035)  # renderer.py:1302


def rng(config_index: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(SEED_BASE + config_index))


def lut_chain_config() -> AtmosphereConfig:
    """Plain LUT chain, single state, sized for a CPU run."""
    return AtmosphereConfig(
        top_altitude_m=100_000.0,
        rayleigh_scattering_per_m=(5.802e-6, 13.558e-6, 33.1e-6),
        rayleigh_scale_height_m=8_000.0,
        mie_scattering_per_m=(3.996e-6,) * 3,
        mie_extinction_per_m=(4.40e-6,) * 3,
        mie_scale_height_m=1_200.0,
        mie_phase_g=0.8,
        absorption_extinction_per_m=(0.0, 0.0, 0.0),
        transmittance_lut_width=256,
        transmittance_lut_height=64,
        transmittance_steps=40,
        multi_scattering_lut_width=32,
        multi_scattering_lut_height=32,
        multi_scattering_directions=64,
        multi_scattering_steps=20,
        sky_view_lut_width=192,
        sky_view_lut_height=108,
        sky_view_steps=30,
    )


def full_frame_config() -> AtmosphereConfig:
    """The LUT chain plus ozone (tent at 25 km, 30 km wide) and a 32^3
    aerial-perspective volume."""
    return lut_chain_config().replace(
        absorption_extinction_per_m=(0.650e-6, 1.881e-6, 0.085e-6),
        absorption_peak_altitude_m=25_000.0,
        absorption_half_width_m=15_000.0,
        aerial_lut_width=32,
        aerial_lut_height=32,
        aerial_lut_depth=32,
        aerial_steps_per_slice=2,
    )


def small_planet_config() -> AtmosphereConfig:
    """The reference tests' small planet shell (tests/test_m3_atmosphere.py:
    11-21; planet radius 50 km) with the tiny LUTs of its CPU render test
    (tests/test_m2_terrain.py:215-230)."""
    return AtmosphereConfig(
        top_altitude_m=15_000.0,
        rayleigh_scale_height_m=3_000.0,
        mie_scale_height_m=800.0,
        absorption_peak_altitude_m=7_500.0,
        absorption_half_width_m=5_000.0,
        transmittance_lut_width=16,
        transmittance_lut_height=16,
        multi_scattering_lut_width=8,
        multi_scattering_lut_height=8,
        sky_view_lut_width=16,
        sky_view_lut_height=16,
        aerial_lut_width=4,
        aerial_lut_height=4,
        aerial_lut_depth=8,
        transmittance_steps=4,
        multi_scattering_directions=4,
        multi_scattering_steps=4,
        sky_view_steps=4,
        aerial_steps_per_slice=1,
    )


def view_basis_local(yaw_degrees: float, pitch_degrees: float):
    """right, view_up, forward in the local East-Up-North frame
    (camera.py:24-33; pitch clamped to +-89 deg as camera.py:22 does)."""
    pitch_degrees = float(np.clip(pitch_degrees, -89.0, 89.0))
    yaw, pitch = math.radians(yaw_degrees), math.radians(pitch_degrees)
    horizontal = np.array([math.sin(yaw), 0.0, math.cos(yaw)])
    forward = horizontal * math.cos(pitch) + np.array([0.0, 1.0, 0.0]) * math.sin(pitch)
    forward /= np.linalg.norm(forward)
    right = np.cross(np.array([0.0, 1.0, 0.0]), forward)
    right /= np.linalg.norm(right)
    view_up = np.cross(forward, right)
    view_up /= np.linalg.norm(view_up)
    return right, view_up, forward


def sun_from_angles(zenith_degrees: float, azimuth_radians: float) -> np.ndarray:
    """Local sun direction (east, up, north) with the azimuth convention of
    cli.py:84-90."""
    z = math.radians(zenith_degrees)
    return np.array([math.sin(azimuth_radians) * math.sin(z), math.cos(z),
                     math.cos(azimuth_radians) * math.sin(z)])


def pixel_rays_view(width: int, height: int, tangent_half_fov: float) -> np.ndarray:
    """Unit pixel-centre rays in view coordinates (x right, y up, z forward),
    shape (W, H, 3), y = 0 the bottom row (renderer.py:1850-1859)."""
    aspect = width / height
    su = (np.arange(width, dtype=np.float64) + 0.5) / width
    sv = (np.arange(height, dtype=np.float64) + 0.5) / height
    sx = (su * 2.0 - 1.0) * aspect
    sy = sv * 2.0 - 1.0
    rays = np.empty((width, height, 3))
    rays[:, :, 0] = sx[:, None]
    rays[:, :, 1] = sy[None, :]
    rays[:, :, 2] = 1.0 / tangent_half_fov
    rays /= np.linalg.norm(rays, axis=2, keepdims=True)
    return rays


@dataclass
class FrameInputs:
    """Everything one atmosphere frame consumes, host-side (NumPy)."""

    config: AtmosphereConfig
    width: int
    height: int
    planet_radius_m: float
    camera_radius_m: float
    sun_local: np.ndarray
    lighting: LightingState
    view_basis: tuple[np.ndarray, np.ndarray, np.ndarray]
    tangent_half_fov: float
    position_view: np.ndarray  # (W, H, 3) float32
    surface_id: np.ndarray  # (W, H) int32
    scene_hdr: np.ndarray  # (W, H, 3) float32
    space_hdr: np.ndarray  # (W, H, 3) float32
    star_contrast: tuple[float, float] = STAR_CONTRAST

    @property
    def camera_altitude_m(self) -> float:
        return self.camera_radius_m - self.planet_radius_m


def make_frame(width: int = 1920, height: int = 1080, config_index: int = 1,
               config: AtmosphereConfig | None = None,
               planet_radius_m: float = EARTH_RADIUS_M,
               camera_altitude_range=(10.0, 50_000.0),
               distance_range=(10.0, 500_000.0), sky_fraction: float = 0.30,
               seed_offset: int = 0) -> FrameInputs:
    """One seeded frame: camera altitude logU(range), sun zenith U[0,100] deg,
    sun azimuth U[0, 2pi), yaw U[0, 360) deg, pitch U[-20, 20] deg, 60 deg
    vertical FOV; a per-pixel sky mask with `sky_fraction` sky pixels;
    surface ranges logU(distance_range) along each pixel ray."""
    g = np.random.Generator(np.random.PCG64(SEED_BASE + config_index + 1000 * seed_offset))
    config = config if config is not None else full_frame_config()
    lo, hi = camera_altitude_range
    altitude = float(math.exp(g.uniform(math.log(lo), math.log(hi))))
    sun = sun_from_angles(g.uniform(0.0, 100.0), g.uniform(0.0, 2.0 * math.pi))
    yaw = g.uniform(0.0, 360.0)
    pitch = g.uniform(-20.0, 20.0)
    basis = view_basis_local(yaw, pitch)
    thf = math.tan(math.radians(VERTICAL_FOV_DEG) / 2.0)
    sky = g.random((width, height)) < sky_fraction
    dlo, dhi = distance_range
    distance = np.exp(g.uniform(math.log(dlo), math.log(dhi), size=(width, height)))
    rays = pixel_rays_view(width, height, thf)
    position_view = (rays * distance[:, :, None]).astype(np.float32)
    surface_id = np.where(sky, -1, 0).astype(np.int32)
    scene = g.random((width, height, 3)).astype(np.float32)
    scene[sky] = np.asarray(SKY_SCENE_HDR, dtype=np.float32)
    space = np.zeros((width, height, 3), dtype=np.float32)
    stars = sky & (g.random((width, height)) < 0.001)
    space[stars] = g.uniform(0.0, 10.0, size=(int(stars.sum()), 3)).astype(np.float32)
    lighting = LightingState(sun, SUN_ANGULAR_RADIUS_DEG, SOLAR_IRRADIANCE)
    return FrameInputs(
        config=config,
        width=width,
        height=height,
        planet_radius_m=planet_radius_m,
        camera_radius_m=planet_radius_m + altitude,
        sun_local=sun,
        lighting=lighting,
        view_basis=basis,
        tangent_half_fov=thf,
        position_view=position_view,
        surface_id=surface_id,
        scene_hdr=scene,
        space_hdr=space,
    )


STRESS_ALTITUDES_M = (0.5, 2_000.0, 100_000.0, 400_000.0, 2_000_000.0)
TANGENT_BAND_DEG = 0.01


def limb_depression_deg(planet_radius_m: float, altitude_m: float) -> float:
    """Angle of the planet's limb below the local horizontal seen from
    altitude h: acos(R / (R + h))."""
    return math.degrees(math.acos(planet_radius_m / (planet_radius_m + altitude_m)))


def make_stress_frame(altitude_m: float, width: int = 3840, height: int = 2160,
                      band_fraction: float = 0.20, max_distance_m: float = 3.0e6,
                      seed_offset: int = 0) -> FrameInputs:
    """Planet-scale precision stress frame: the view is centred on the
    planet's limb with a vertical FOV of 2 x 0.01 deg / band_fraction, so
    `band_fraction` of the pixel rows look within 0.01 deg of tangency;
    surface ranges logU[10 m, max_distance_m]; sun zenith U[0, 100] deg,
    yaw U[0, 360) deg; 30 % sky pixels."""
    g = np.random.Generator(np.random.PCG64(SEED_BASE + 4 + 1000 * seed_offset))
    config = full_frame_config()
    R = EARTH_RADIUS_M
    sun = sun_from_angles(g.uniform(0.0, 100.0), g.uniform(0.0, 2.0 * math.pi))
    yaw = g.uniform(0.0, 360.0)
    pitch = -limb_depression_deg(R, altitude_m)
    basis = view_basis_local(yaw, pitch)
    vfov_deg = 2.0 * TANGENT_BAND_DEG / band_fraction
    thf = math.tan(math.radians(vfov_deg) / 2.0)
    sky = g.random((width, height)) < 0.30
    distance = np.exp(g.uniform(math.log(10.0), math.log(max_distance_m), size=(width, height)))
    rays = pixel_rays_view(width, height, thf)
    position_view = (rays * distance[:, :, None]).astype(np.float32)
    surface_id = np.where(sky, -1, 0).astype(np.int32)
    scene = g.random((width, height, 3)).astype(np.float32)
    scene[sky] = np.asarray(SKY_SCENE_HDR, dtype=np.float32)
    space = np.zeros((width, height, 3), dtype=np.float32)
    stars = sky & (g.random((width, height)) < 0.001)
    space[stars] = g.uniform(0.0, 10.0, size=(int(stars.sum()), 3)).astype(np.float32)
    return FrameInputs(
        config=config, width=width, height=height, planet_radius_m=R,
        camera_radius_m=R + altitude_m, sun_local=sun,
        lighting=LightingState(sun, SUN_ANGULAR_RADIUS_DEG, SOLAR_IRRADIANCE),
        view_basis=basis, tangent_half_fov=thf, position_view=position_view,
        surface_id=surface_id, scene_hdr=scene, space_hdr=space)


def lut_chain_state(seed_offset: int = 0):
    """LUT-chain state: camera at 2 km, sun zenith U[0, 100] deg."""
    g = np.random.Generator(np.random.PCG64(SEED_BASE + 0 + 1000 * seed_offset))
    sun = sun_from_angles(g.uniform(0.0, 100.0), g.uniform(0.0, 2.0 * math.pi))
    return EARTH_RADIUS_M, EARTH_RADIUS_M + 2_000.0, sun


SMALL_PLANET_RADIUS_M = 50_000.0


def batch_states(count: int = 256, seed_offset: int = 0, config: AtmosphereConfig | None = None):
    """`count` independent states for the batched LUT build.  Per state: sun
    zenith U[0, 100] deg, sun azimuth U[0, 2pi), camera altitude
    logU[1 m, 2000 km], yaw U[0, 360) deg, pitch U[-20, 20] deg, and the
    Rayleigh and Mie coefficients each scaled by their own U[0.7, 1.3]
    factor.  The second half of the states sits on a 50 km planet with a
    10 km shell, scale heights and the ozone tent divided by 10."""
    from .batch import AtmosphereState

    g = np.random.Generator(np.random.PCG64(SEED_BASE + 2 + 1000 * seed_offset))
    base = config if config is not None else full_frame_config()
    thf = math.tan(math.radians(VERTICAL_FOV_DEG) / 2.0)
    states = []
    for i in range(count):
        small = i >= count // 2
        sun = sun_from_angles(g.uniform(0.0, 100.0), g.uniform(0.0, 2.0 * math.pi))
        altitude = float(math.exp(g.uniform(math.log(1.0), math.log(2.0e6))))
        yaw, pitch = g.uniform(0.0, 360.0), g.uniform(-20.0, 20.0)
        kr, km = g.uniform(0.7, 1.3), g.uniform(0.7, 1.3)
        cfg = base.replace(
            rayleigh_scattering_per_m=tuple(kr * v for v in base.rayleigh_scattering_per_m),
            mie_scattering_per_m=tuple(km * v for v in base.mie_scattering_per_m),
            mie_extinction_per_m=tuple(km * v for v in base.mie_extinction_per_m),
        )
        radius = EARTH_RADIUS_M
        if small:
            radius = SMALL_PLANET_RADIUS_M
            cfg = cfg.replace(
                top_altitude_m=base.top_altitude_m / 10.0,
                rayleigh_scale_height_m=base.rayleigh_scale_height_m / 10.0,
                mie_scale_height_m=base.mie_scale_height_m / 10.0,
                absorption_peak_altitude_m=base.absorption_peak_altitude_m / 10.0,
                absorption_half_width_m=base.absorption_half_width_m / 10.0,
            )
        states.append(AtmosphereState(
            config=cfg,
            planet_radius_m=radius,
            camera_radius_m=radius + altitude,
            sun_local=sun,
            solar_irradiance=SOLAR_IRRADIANCE,
            sun_angular_radius_degrees=SUN_ANGULAR_RADIUS_DEG,
            view_basis=view_basis_local(yaw, pitch),
            tangent_half_fov=thf,
        ))
    return states


__all__ = [
    "EARTH_RADIUS_M",
    "batch_states",
    "FrameInputs",
    "SOLAR_IRRADIANCE",
    "SUN_ANGULAR_RADIUS_DEG",
    "full_frame_config",
    "lut_chain_config",
    "lut_chain_state",
    "make_frame",
    "make_stress_frame",
    "pixel_rays_view",
    "small_planet_config",
    "sun_from_angles",
    "view_basis_local",
]


# --------------------------------------------------------------- shading

OCEAN_SURFACE_ID = 2_147_000_000  # src/planet_renderer/ocean.py:17
# OceanConfig defaults (src/planet_renderer/ocean.py:46-65)
OCEAN_DEFAULTS = {
    "ocean_roughness": 0.12,
    "ocean_f0": 0.02,
    "ocean_sky_strength": 1.0,
    "ocean_absorption": (0.14, 0.055, 0.022),
    "ocean_scattering": (0.0025, 0.005, 0.012),
    "ocean_max_depth": 80.0,
    "ocean_refraction_index": 1.333,
    "ocean_refraction_strength": 1.0,
    "ocean_refraction_max_offset_pixels": 48.0,
}


@dataclass
class ShadingInputs:
    """The fields and scalars `_shade_surface` consumes
    (src/planet_renderer/renderer.py:1277-1300), host-side (NumPy).
    `fields` holds the G-buffer, the pre-ocean terrain snapshot and the
    surface-lighting outputs under the AtmoShadeFields names; `params` the
    scalar arguments under the AtmoShadeParams names."""

    width: int
    height: int
    fields: dict
    params: dict


def make_shading_inputs(width: int = 1920, height: int = 1080, seed_offset: int = 0,
                        water_plane_m: float = 40.0) -> ShadingInputs:
    """A seeded G-buffer with terrain and ocean pixels.  The ocean is a
    water plane `water_plane_m` below the camera (view coordinates) over a
    smoothly varying seabed 1-70 m deeper, so the refracted seabed lookup
    finds terrain snapshot samples behind the water surface; terrain pixels
    keep the frame generator's positions.  Sky pixels are as make_frame
    draws them.  Local and global frames coincide."""
    frame = make_frame(width=width, height=height, config_index=1, seed_offset=seed_offset)
    g = np.random.Generator(np.random.PCG64(SEED_BASE + 7 + 1000 * seed_offset))
    right, up, forward = frame.view_basis
    rays = pixel_rays_view(width, height, frame.tangent_half_fov)
    sid = frame.surface_id.copy()
    surface = sid >= 0
    sid[surface] = g.integers(0, 1 << 20, size=int(surface.sum()), dtype=np.int32)
    below = rays[:, :, 1] < -0.02
    ocean = surface & below & (g.random((width, height)) < 0.95)
    sid[ocean] = OCEAN_SURFACE_ID
    pos = frame.position_view.astype(np.float64)
    t_water = water_plane_m / np.maximum(-rays[:, :, 1], 1e-6)
    pos[ocean] = rays[ocean] * t_water[ocean, None]
    xs = np.arange(width)[:, None] / width
    ys = np.arange(height)[None, :] / height
    depth = 1.0 + 69.0 * (0.5 + 0.5 * np.sin(9.0 * xs + 4.0 * ys) * np.cos(5.0 * ys - 3.0 * xs))
    t_bed = (water_plane_m + depth) / np.maximum(-rays[:, :, 1], 1e-6)

    def unit(v):
        return v / np.linalg.norm(v, axis=-1, keepdims=True)

    world_up = np.array([0.0, 1.0, 0.0])
    normal = unit(world_up + 0.35 * g.normal(size=(width, height, 3)))
    normal[ocean] = unit(up + 0.06 * g.normal(size=(int(ocean.sum()), 3)))
    albedo = g.uniform(0.02, 0.8, (width, height, 3))
    albedo[ocean] = (0.012, 0.055, 0.16)
    terrain_sid = sid.copy()
    terrain_sid[ocean] = 0
    terrain_sid[ocean & (g.random((width, height)) < 0.05)] = -1  # holes in the snapshot
    terrain_pos = pos.copy()
    terrain_pos[ocean] = rays[ocean] * t_bed[ocean, None]
    terrain_albedo = albedo.copy()
    terrain_albedo[ocean] = g.uniform(0.1, 0.6, (int(ocean.sum()), 3))
    terrain_normal = normal.copy()
    terrain_normal[ocean] = unit(world_up + 0.3 * g.normal(size=(int(ocean.sum()), 3)))
    f32 = lambda a: np.ascontiguousarray(a, dtype=np.float32)  # noqa: E731
    fields = {
        "surface_id": np.ascontiguousarray(sid, dtype=np.int32),
        "albedo": f32(albedo),
        "normal": f32(normal),
        "position": f32(pos),
        "ocean_slope_variance": f32(g.uniform(0.0, 0.02, (width, height))),
        "water_depth_m": f32(np.clip(depth / np.maximum(-rays[:, :, 1], 0.05), 0.0, 500.0)),
        "seabed_albedo": f32(g.uniform(0.1, 0.6, (width, height, 3))),
        "seabed_normal": f32(unit(world_up + 0.3 * g.normal(size=(width, height, 3)))),
        "terrain_surface_id": np.ascontiguousarray(terrain_sid, dtype=np.int32),
        "terrain_position": f32(terrain_pos),
        "terrain_albedo": f32(terrain_albedo),
        "terrain_normal": f32(terrain_normal),
        "sun_transmittance": f32(g.uniform(0.0, 1.0, (width, height, 3))),
        "sky_radiance": f32(g.uniform(0.0, 0.3, (width, height, 3))),
    }
    params = {
        "sun_global": tuple(frame.sun_local),
        "solar_irradiance": SOLAR_IRRADIANCE,
        "view_right_global": tuple(right),
        "view_up_global": tuple(up),
        "view_forward_global": tuple(forward),
        "advanced_ocean_enabled": 1,
        **OCEAN_DEFAULTS,
        "sun_angular_radius": math.radians(SUN_ANGULAR_RADIUS_DEG),
        "tangent_half_fov": frame.tangent_half_fov,
        "mode": 0,
    }
    return ShadingInputs(width=width, height=height, fields=fields, params=params)
