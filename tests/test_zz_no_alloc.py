"""atmo_composite_noalloc (no counterpart in the reference): the composite for
callers that cannot allocate, with its deferred-pixel queue in the library's
static device array.  Called through the C-ABI and compared bit for bit with
atmo_composite on the same LUTs and G-buffer, on two calls in a row, the
second on a smaller frame; both frames put the sun at the horizon so pixels
take the deferred (terminator-refinement) pass through the queue."""

from __future__ import annotations

import ctypes
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def device():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("needs a HIP device")
    return torch.device("cuda:0")


def _dev(a, device):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a)).to(device)


def _frame(pkg, scene, width, height, seed_offset):
    inputs = scene.make_frame(width=width, height=height, config_index=1, seed_offset=seed_offset)
    inputs.sun_local = scene.sun_from_angles(90.5, 0.3)
    inputs.lighting = pkg.LightingState(inputs.sun_local, 0.266, scene.SOLAR_IRRADIANCE)
    inputs.camera_radius_m = inputs.planet_radius_m + 20_000.0
    return inputs


def _both(pkg, capi, inputs, device, view=0):
    """hdr of renderer.composite (atmo_composite) and of
    atmo_composite_noalloc on the renderer's own LUTs."""
    import torch

    w, h = inputs.width, inputs.height
    r = pkg.AtmosphereRenderer(w, h, inputs.config, device)
    pos, sid = _dev(inputs.position_view, device), _dev(inputs.surface_id, device)
    scene_hdr, space_hdr = _dev(inputs.scene_hdr, device), _dev(inputs.space_hdr, device)
    light = inputs.lighting
    R, r_cam = inputs.planet_radius_m, inputs.camera_radius_m
    deg = light.sun_angular_radius_degrees
    r.update(R, r_cam, inputs.sun_local, light.solar_irradiance, deg)
    r.prepare_frame(pos, sid, R, r_cam, inputs.sun_local, light.solar_irradiance, deg,
                    inputs.view_basis, inputs.tangent_half_fov)
    want = torch.zeros_like(scene_hdr)
    r.composite(scene_hdr, space_hdr, sid, pos, want, R, r_cam, inputs.sun_local,
                light.solar_irradiance, light.sun_disk_radiance, deg, inputs.view_basis,
                inputs.tangent_half_fov, *inputs.star_contrast, diagnostic_view=view)
    got = torch.full_like(scene_hdr, float("nan"))
    params = inputs.config.to_params()
    right, up, forward = inputs.view_basis
    cam_alt = max(r_cam - R, 0.0)
    f3 = capi.f3
    lib = capi.library()
    status = lib.atmo_composite_noalloc(
        ctypes.byref(params), r.transmittance_lut.data_ptr(), r.multi_scattering_lut.data_ptr(),
        r.sky_view_lut.data_ptr(), r.aerial_scattering_lut.data_ptr(), w, h,
        scene_hdr.data_ptr(), space_hdr.data_ptr(), sid.data_ptr(), pos.data_ptr(),
        got.data_ptr(), R, cam_alt, r.sky_snapshot_altitude_m, f3(inputs.sun_local),
        f3(light.solar_irradiance), f3(light.sun_disk_radiance), math.radians(deg), f3(right),
        f3(up), f3(forward), inputs.tangent_half_fov, *inputs.star_contrast, view,
        torch.cuda.current_stream(device).cuda_stream)
    assert status == capi.ATMO_OK, lib.atmo_last_error(None)
    torch.cuda.synchronize(device)
    out = want.cpu().numpy(), got.cpu().numpy()
    r.close()
    return out


def test_noalloc_composite_matches_atmo_composite(pkg, scene, ref32, device):
    capi = __import__(pkg.__name__ + "._capi", fromlist=["_capi"])
    for k, (w, h) in enumerate(((320, 180), (200, 112))):
        inputs = _frame(pkg, scene, w, h, seed_offset=40 + k)
        refined = ref32.frame(inputs)["refined_intervals"]
        print(f"call {k}: {w}x{h}, refined intervals (oracle) {refined}", flush=True)
        assert refined > 0  # the deferred pass runs through the queue
        want, got = _both(pkg, capi, inputs, device)
        differ = int((want.view(np.uint32) != got.view(np.uint32)).sum())
        print(f"call {k}: {differ} values differ from atmo_composite", flush=True)
        assert differ == 0


def test_noalloc_diagnostic_view_matches(pkg, scene, device):
    capi = __import__(pkg.__name__ + "._capi", fromlist=["_capi"])
    inputs = _frame(pkg, scene, 96, 64, seed_offset=44)
    want, got = _both(pkg, capi, inputs, device, view=9)
    assert (want.view(np.uint32) == got.view(np.uint32)).all()
