"""The reference's host methods in C (atmo_update, atmo_prepare_frame,
atmo_composite_frame, atmo_set_luts_frozen; renderer.py:2057-2238) on the
GPU: driven with float64 host values, a handle rebuilds exactly the LUTs the
Python facade rebuilds, with the same float32 arguments, so its LUTs, surface
lighting and hdr are bitwise those of the facade, frame after frame, and its
counters and snapshots agree."""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def device():
    import torch

    assert torch.cuda.is_available()
    return torch.device("cuda:0")


def _dev(a, device):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a)).to(device)


def test_c_host_methods_match_the_facade(pkg, scene, device):
    import torch

    capi = __import__(pkg.__name__ + "._capi", fromlist=["_capi"])
    lib = capi.library()
    inputs = scene.make_frame(width=160, height=90, config_index=1, seed_offset=3)
    cfg, light = inputs.config, inputs.lighting
    py = pkg.AtmosphereRenderer(inputs.width, inputs.height, cfg, device)
    cc = pkg.AtmosphereRenderer(inputs.width, inputs.height, cfg, device)
    pos, sid = _dev(inputs.position_view, device), _dev(inputs.surface_id, device)
    scene_hdr, space = _dev(inputs.scene_hdr, device), _dev(inputs.space_hdr, device)
    hdr_py, hdr_c = torch.zeros_like(scene_hdr), torch.zeros_like(scene_hdr)
    R = inputs.planet_radius_m
    deg = light.sun_angular_radius_degrees
    basis9 = (ctypes.c_double * 9)(*np.concatenate([np.asarray(v, dtype=np.float64)
                                                     for v in inputs.view_basis]))
    stream = lambda: torch.cuda.current_stream(device).cuda_stream

    def frame(r_cam, sun, freeze=None):
        if freeze is not None:
            py.set_luts_frozen(freeze)
            assert lib.atmo_set_luts_frozen(cc.handle, int(freeze)) == 0
        py.update(R, r_cam, sun, light.solar_irradiance, deg)
        py.prepare_frame(pos, sid, R, r_cam, sun, light.solar_irradiance, deg, inputs.view_basis,
                         inputs.tangent_half_fov)
        py.composite(scene_hdr, space, sid, pos, hdr_py, R, r_cam, sun, light.solar_irradiance,
                     light.sun_disk_radiance, deg, inputs.view_basis, inputs.tangent_half_fov,
                     *inputs.star_contrast)
        built = ctypes.c_int32()
        assert lib.atmo_update(cc.handle, R, r_cam, capi.d3(sun), capi.d3(light.solar_irradiance),
                               deg, 0, stream(), ctypes.byref(built)) == 0
        assert lib.atmo_prepare_frame(cc.handle, pos.data_ptr(), sid.data_ptr(), R, r_cam,
                                      capi.d3(sun), capi.d3(light.solar_irradiance), deg, basis9,
                                      inputs.tangent_half_fov, 0, stream(), None) == 0
        assert lib.atmo_composite_frame(cc.handle, scene_hdr.data_ptr(), space.data_ptr(),
                                        sid.data_ptr(), pos.data_ptr(), hdr_c.data_ptr(), R, r_cam,
                                        capi.d3(sun), capi.d3(light.solar_irradiance),
                                        capi.d3(light.sun_disk_radiance), deg, basis9,
                                        inputs.tangent_half_fov, inputs.star_contrast[0],
                                        inputs.star_contrast[1], 0, stream()) == 0
        torch.cuda.synchronize(device)
        for name in ("transmittance_lut", "multi_scattering_lut", "sky_view_lut",
                     "aerial_scattering_lut", "surface_sun_transmittance",
                     "surface_sky_radiance"):
            a = getattr(py, name).to_numpy()
            b = getattr(cc, name).to_numpy()
            assert (a.view(np.uint32) == b.view(np.uint32)).all(), name
        assert (hdr_py.cpu().numpy().view(np.uint32) == hdr_c.cpu().numpy().view(np.uint32)).all()
        state = capi.AtmoLutCache()
        assert lib.atmo_get_lut_cache(cc.handle, ctypes.byref(state)) == 0
        assert (state.transmittance_rebuilds, state.multi_scattering_rebuilds,
                state.sky_view_rebuilds, state.aerial_rebuilds) == (
            py.transmittance_rebuilds, py.multi_scattering_rebuilds, py.sky_view_rebuilds,
            py.aerial_rebuilds)
        assert state.sky_snapshot_altitude_m == py.sky_snapshot_altitude_m
        return built.value

    r0, sun0 = inputs.camera_radius_m, np.asarray(inputs.sun_local)
    assert frame(r0, sun0) == 7
    assert frame(r0, sun0) == 0
    assert frame(r0 + 100.0, sun0) == capi.LUT_SKY_VIEW
    assert frame(r0 + 200.0, sun0, freeze=True) == 0
    assert frame(r0 + 200.0, scene.sun_from_angles(70.0, 1.0), freeze=False) == capi.LUT_SKY_VIEW
    # a forced rebuild of everything (the benchmark's cold frame) rebuilds the
    # same LUTs bit for bit
    built = ctypes.c_int32()
    assert lib.atmo_update(cc.handle, R, r0, capi.d3(sun0), capi.d3(light.solar_irradiance), deg,
                           7, stream(), ctypes.byref(built)) == 0
    assert built.value == 7
    py.close()
    cc.close()
