"""GPU parity of the surface shading (libatmo.so's atmo_shade_surface) with
the CPU oracle (oracle/surface_cpu.c), the restatement of the reference's
`_shade_surface` (src/planet_renderer/renderer.py:1277-1575).

Bar: bit-exact.  The kernel is float32 arithmetic in the reference's
operation order with IEEE division and square root; exp and tan are
float64 rounded once on both sides."""

import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

OCEAN = 2_147_000_000
FIELD_MAP = {  # renderer field name -> scene / oracle field name
    "gbuffer_surface_id": "surface_id", "gbuffer_albedo": "albedo", "gbuffer_normal": "normal",
    "gbuffer_position": "position", "gbuffer_ocean_slope_variance": "ocean_slope_variance",
    "gbuffer_water_depth_m": "water_depth_m", "gbuffer_seabed_albedo": "seabed_albedo",
    "gbuffer_seabed_normal": "seabed_normal", "terrain_surface_id": "terrain_surface_id",
    "terrain_position": "terrain_position", "terrain_albedo": "terrain_albedo",
    "terrain_normal": "terrain_normal",
}
ARG_ORDER = ("sun_global", "solar_irradiance", "view_right_global", "view_up_global",
             "view_forward_global", "advanced_ocean_enabled", "ocean_roughness", "ocean_f0",
             "ocean_sky_strength", "ocean_absorption", "ocean_scattering", "ocean_max_depth",
             "ocean_refraction_index", "ocean_refraction_strength",
             "ocean_refraction_max_offset_pixels", "sun_angular_radius", "tangent_half_fov", "mode")


@pytest.fixture(scope="module")
def device():
    import torch

    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch.device("cuda:0")


def gpu_shade(pkg, inputs, params, device):
    import torch

    w, h = inputs.width, inputs.height
    sf = pkg.SurfaceFields.allocate(w, h, device)
    for name, key in FIELD_MAP.items():
        getattr(sf, name).tensor.copy_(torch.from_numpy(inputs.fields[key]))
    sf.surface_hdr.tensor.fill_(float("nan"))
    sun_t = torch.from_numpy(inputs.fields["sun_transmittance"]).to(device)
    sky = torch.from_numpy(inputs.fields["sky_radiance"]).to(device)
    pkg.shade_surface(sf, params["sun_global"], params["solar_irradiance"], sun_t, sky,
                      *(params[k] for k in ARG_ORDER[2:]))
    torch.cuda.synchronize(device)
    return sf.surface_hdr.to_numpy()


def assert_bit_exact(got, want):
    same = got.view(np.uint32) == want.view(np.uint32)
    if not same.all():
        idx = tuple(np.argwhere(~same)[0])
        raise AssertionError(f"{int((~same).sum())} of {same.size} values differ; first at {idx}: "
                             f"{got[idx]!r} vs {want[idx]!r}")


CASES = [
    ((3840, 2160), 9, {}),
    ((1920, 1080), 0, {}),
    ((321, 179), 1, {}),
    ((321, 179), 2, {"ocean_refraction_max_offset_pixels": 3.0}),
    ((200, 120), 3, {"ocean_refraction_strength": 0.0}),
    ((200, 120), 4, {"ocean_refraction_strength": 0.4, "ocean_f0": 0.2, "ocean_roughness": 0.8}),
    ((200, 120), 5, {"advanced_ocean_enabled": 0}),
    ((200, 120), 6, {"mode": 2}),
    ((17, 5), 7, {}),
    ((1, 1), 8, {}),
]


@pytest.mark.parametrize("size, seed, changes", CASES)
def test_shade_surface_bit_exact(pkg, scene, device, size, seed, changes):
    from oracle.surface_cpu import shade

    inputs = scene.make_shading_inputs(*size, seed_offset=seed)
    params = dict(inputs.params, **changes)
    print(f"shade_surface {size} seed {seed} {changes}", flush=True)
    got = gpu_shade(pkg, inputs, params, device)
    want, path = shade(inputs.fields, params, *size)
    assert_bit_exact(got, want)
    if size == (1920, 1080):
        counts = np.bincount(path.ravel(), minlength=4)
        print(f"pixels: sky {counts[0]}, lit {counts[1]}, ocean {counts[2]}, "
              f"ocean with refracted seabed {counts[3]}", flush=True)
        assert counts.min() > 10_000  # every branch is exercised


def test_shade_surface_with_prepare_frame_lighting(pkg, scene, device):
    """The reference's frame order (renderer.py:1664-1702): prepare_frame
    lights the G-buffer, _shade_surface consumes its two fields."""
    import torch

    from oracle.surface_cpu import shade

    inputs = scene.make_shading_inputs(480, 270, seed_offset=11)
    frame = scene.make_frame(width=480, height=270, config_index=1, seed_offset=11)
    r = pkg.AtmosphereRenderer(480, 270, frame.config, device)
    light = frame.lighting
    pos = torch.from_numpy(inputs.fields["position"]).to(device)
    sid = torch.from_numpy(inputs.fields["surface_id"]).to(device)
    r.update(frame.planet_radius_m, frame.camera_radius_m, frame.sun_local,
             light.solar_irradiance, light.sun_angular_radius_degrees)
    r.prepare_frame(pos, sid, frame.planet_radius_m, frame.camera_radius_m, frame.sun_local,
                    light.solar_irradiance, light.sun_angular_radius_degrees, frame.view_basis,
                    frame.tangent_half_fov)
    torch.cuda.synchronize(device)
    fields = dict(inputs.fields)
    fields["sun_transmittance"] = r.surface_sun_transmittance.to_numpy()
    fields["sky_radiance"] = r.surface_sky_radiance.to_numpy()
    r.close()
    lit = type(inputs)(inputs.width, inputs.height, fields, inputs.params)
    got = gpu_shade(pkg, lit, inputs.params, device)
    want, _ = shade(fields, inputs.params, 480, 270)
    assert_bit_exact(got, want)
    assert np.isfinite(got).all() and float(got.max()) > 0.0


def test_shade_surface_argument_errors(pkg, device):
    import torch

    sf = pkg.SurfaceFields.allocate(8, 6, device)
    good = torch.zeros((8, 6, 3), dtype=torch.float32, device=device)
    args = ((0, 1, 0), (4, 3.9, 3.7), good, good, (1, 0, 0), (0, 1, 0), (0, 0, 1), 1, 0.12, 0.02,
            1.0, (0.14, 0.055, 0.022), (0.0025, 0.005, 0.012), 80.0, 1.333, 1.0, 48.0,
            math.radians(0.266), 0.577, 0)
    pkg.shade_surface(sf, *args)
    bad = list(args)
    bad[2] = torch.zeros((6, 8, 3), dtype=torch.float32, device=device)
    with pytest.raises(ValueError):
        pkg.shade_surface(sf, *bad)
    sf.gbuffer_albedo = pkg.DeviceField(torch.zeros((8, 6, 3), dtype=torch.float64, device=device))
    with pytest.raises(ValueError):
        pkg.shade_surface(sf, *args)
