"""Pins the CPU restatement (oracle/ref_cpu.c) before it is trusted as the
parity proxy: the reference tests' known answers, closed forms, inverse
round trips, and convergence to the independent float64 validation model
(oracle/model_fp64.py) where both integrate the same physics with different
quadratures.  A dropped term, a wrong sign or a swapped operand in the
restatement moves one of these by far more than its tolerance."""

import ctypes
import dataclasses
import math

import numpy as np
import pytest

from closed_forms import vertical_optical_depth
from oracle.model_fp64 import AtmosphereModel
from oracle.ref_cpu import params_of


def _arr(ref, values):
    return np.ascontiguousarray(np.asarray(values, dtype=np.float64).astype(ref.dtype))


def _out(ref, n):
    return np.zeros(n, dtype=ref.dtype)


def _params(cfg, **overrides):
    """AtmoParams from a config with overrides beyond the configuration's
    validation ranges (converged step counts for the oracle's own checks)."""
    values = dataclasses.asdict(cfg)
    values.update(overrides)
    return params_of(values)


@pytest.fixture(scope="module")
def small(pkg):
    return pkg.AtmosphereConfig(
        top_altitude_m=15_000.0,
        rayleigh_scale_height_m=3_000.0,
        mie_scale_height_m=800.0,
        absorption_peak_altitude_m=7_500.0,
        absorption_half_width_m=5_000.0,
    )


@pytest.mark.parametrize("prec", ["ref32", "ref64"])
def test_sphere_and_shell_roots_known_answers(request, prec):
    """tests/test_m3_atmosphere.py:24-39 through both root solvers of the
    restatement (vector form renderer.py:211-226, radial-axis form
    renderer.py:264-301)."""
    ref = request.getfixturevalue(prec)
    out = _out(ref, 3)
    ref.lib.oracle_fn_sphere_roots(ref.ptr(_arr(ref, [0, 70_000, 0])), ref.ptr(_arr(ref, [0, -1, 0])),
                                   ref.r(65_000.0), ref.ptr(out))
    np.testing.assert_allclose(out, (5_000.0, 135_000.0, 1.0), atol=1e-8)
    ref.lib.oracle_fn_sphere_roots(ref.ptr(_arr(ref, [0, 55_000, 0])), ref.ptr(_arr(ref, [0, 1, 0])),
                                   ref.r(65_000.0), ref.ptr(out))
    np.testing.assert_allclose(out, (-120_000.0, 10_000.0, 1.0), atol=1e-8)
    # radial-axis form: planet radius 50 km, origins at 20 km / 5 km altitude
    ref.lib.oracle_fn_shell_roots(ref.r(20_000.0), ref.r(-1.0), ref.r(15_000.0), ref.r(50_000.0),
                                  ref.ptr(out))
    np.testing.assert_allclose(out, (5_000.0, 135_000.0, 1.0), atol=1e-8)
    ref.lib.oracle_fn_shell_roots(ref.r(5_000.0), ref.r(1.0), ref.r(15_000.0), ref.r(50_000.0),
                                  ref.ptr(out))
    np.testing.assert_allclose(out, (-120_000.0, 10_000.0, 1.0), atol=1e-8)
    # a ray passing 10 km above a 65 km sphere misses it
    ref.lib.oracle_fn_sphere_roots(ref.ptr(_arr(ref, [0, 75_000, 0])), ref.ptr(_arr(ref, [1, 0, 0])),
                                   ref.r(65_000.0), ref.ptr(out))
    assert out[2] == 0.0


@pytest.mark.parametrize("prec", ["ref32", "ref64"])
def test_segments_stop_at_opaque_ground(request, prec, small):
    """tests/test_m3_atmosphere.py:42-58 through both interval builders
    (renderer.py:303-374)."""
    ref = request.getfixturevalue(prec)
    p = params_of(small)
    out = _out(ref, 4)
    ref.lib.oracle_fn_atmosphere_segment(ctypes.byref(p), ref.r(1_000.0), ref.r(-1.0),
                                         ref.r(50_000.0), ref.ptr(out))
    np.testing.assert_allclose(out, (0.0, 1_000.0, 1.0, 1.0), atol=1e-8)
    ref.lib.oracle_fn_atmosphere_segment(ctypes.byref(p), ref.r(0.0), ref.r(1.0),
                                         ref.r(50_000.0), ref.ptr(out))
    np.testing.assert_allclose(out, (0.0, 15_000.0, 1.0, 0.0), atol=1e-8)
    ref.lib.oracle_fn_atmosphere_segment_vec(ctypes.byref(p), ref.ptr(_arr(ref, [0, 51_000, 0])),
                                             ref.ptr(_arr(ref, [0, -1, 0])), ref.r(50_000.0),
                                             ref.ptr(out))
    np.testing.assert_allclose(out, (0.0, 1_000.0, 1.0, 1.0), atol=1e-8)
    # from outside the shell, looking down through it: entry at 16 km, ground at 31 km
    ref.lib.oracle_fn_atmosphere_segment(ctypes.byref(p), ref.r(31_000.0), ref.r(-1.0),
                                         ref.r(50_000.0), ref.ptr(out))
    np.testing.assert_allclose(out, (16_000.0, 31_000.0, 1.0, 1.0), atol=1e-8)


def test_density_profile(ref64, small):
    p = params_of(small)
    out = _out(ref64, 3)
    expect = {
        0.0: (1.0, 1.0, 0.0),
        3_000.0: (math.exp(-1.0), math.exp(-3_000 / 800), 0.1),
        7_500.0: (math.exp(-2.5), math.exp(-7_500 / 800), 1.0),
        16_000.0: (0.0, 0.0, 0.0),
        -1.0: (0.0, 0.0, 0.0),
    }
    for h, e in expect.items():
        ref64.lib.oracle_fn_density(ctypes.byref(p), ref64.r(h), ref64.ptr(out))
        np.testing.assert_allclose(out, e, rtol=1e-14, atol=1e-15)


@pytest.mark.parametrize("h0", [0.0, 2_000.0, 30_000.0])
def test_vertical_optical_depth_closed_form(ref64, scene, h0):
    """_integrate_extinction (renderer.py:570-607) on a zenith ray equals
    exp(-tau) of the exponential + ozone-tent closed form."""
    cfg = scene.full_frame_config()
    p = _params(cfg, transmittance_steps=20_000)
    R = 6_360_000.0
    out = _out(ref64, 3)
    ref64.lib.oracle_fn_integrate_extinction(
        ctypes.byref(p), ref64.ptr(_arr(ref64, [0, R + h0, 0])), ref64.ptr(_arr(ref64, [0, 1, 0])),
        ref64.r(0.0), ref64.r(cfg.top_altitude_m - h0), ref64.r(R), ref64.ptr(out))
    np.testing.assert_allclose(out, np.exp(-vertical_optical_depth(cfg, h0)), rtol=2e-9)


@pytest.mark.parametrize("cosine", [0.5, 0.05, -0.02])
def test_slant_optical_depth_matches_validation_model(ref64, scene, cosine):
    """Warped-midpoint quadrature of the restatement and uniform midpoint
    quadrature of the validation model converge to the same transmittance
    on slant and grazing rays."""
    cfg = scene.full_frame_config()
    R, h0 = 6_360_000.0, 5_000.0
    ray = np.array([math.sqrt(1 - cosine * cosine), cosine, 0.0])
    origin = np.array([0.0, R + h0, 0.0])
    model = AtmosphereModel(R, cfg)
    interval = model.atmosphere_interval(origin, ray)
    p = _params(cfg, transmittance_steps=20_000)
    out = _out(ref64, 3)
    ref64.lib.oracle_fn_integrate_extinction(ctypes.byref(p), ref64.ptr(origin), ref64.ptr(ray),
                                             ref64.r(interval.start_m), ref64.r(interval.end_m),
                                             ref64.r(R), ref64.ptr(out))
    expect = model.transmittance(origin, ray, steps=200_000)
    np.testing.assert_allclose(out, expect, rtol=2e-6)


def test_phases_match_validation_model(ref64, ref32, scene):
    cfg = scene.full_frame_config()
    model = AtmosphereModel(6_360_000.0, cfg)
    p = params_of(cfg)
    for c in np.linspace(-1.0, 1.0, 21):
        out = _out(ref64, 2)
        ref64.lib.oracle_fn_phases(ctypes.byref(p), ref64.r(c), ref64.ptr(out))
        np.testing.assert_allclose(out, (model.rayleigh_phase(c), model.mie_phase(c)), rtol=1e-13)
        out32 = _out(ref32, 2)
        ref32.lib.oracle_fn_phases(ctypes.byref(p), ref32.r(c), ref32.ptr(out32))
        np.testing.assert_allclose(out32, out, rtol=3e-6)


def test_segment_integral_limits(ref64):
    """_segment_integral (renderer.py:806-816): (1 - e^{-s d}) / s, and d
    itself when s <= 1e-12."""
    t, i = _out(ref64, 3), _out(ref64, 3)
    ext = _arr(ref64, [1e-5, 1e-13, 0.0])
    ref64.lib.oracle_fn_segment_integral(ref64.ptr(ext), ref64.r(2_000.0), ref64.ptr(t), ref64.ptr(i))
    np.testing.assert_allclose(t, np.exp(-np.array([1e-5, 1e-13, 0.0]) * 2_000.0), rtol=1e-15)
    np.testing.assert_allclose(i, [(1 - math.exp(-0.02)) / 1e-5, 2_000.0, 2_000.0], rtol=1e-12)
    ext = _arr(ref64, [1.0, 1.0, 1.0])
    ref64.lib.oracle_fn_segment_integral(ref64.ptr(ext), ref64.r(1_000.0), ref64.ptr(t), ref64.ptr(i))
    np.testing.assert_allclose(t, math.exp(-80.0), rtol=1e-15)  # optical depth clamp


def test_solar_disk_visibility_cases(ref64, scene):
    """renderer.py:728-754: a disk clear of the horizon is fully visible,
    one below it is hidden, a disk centred on the horizon is half visible,
    and the visible fraction is the circular-segment area."""
    p = params_of(scene.full_frame_config())
    R, alpha = 6_360_000.0, math.radians(0.266)
    vis = lambda h, c: ref64.lib.oracle_fn_solar_disk_visibility(
        ctypes.byref(p), ref64.r(h), ref64.r(c), ref64.r(R), ref64.r(alpha))
    ref64.lib.oracle_fn_solar_disk_visibility.restype = ref64.real
    assert vis(0.0, 0.5) == pytest.approx(1.0, abs=1e-15)
    assert vis(0.0, -0.5) == pytest.approx(0.0, abs=1e-15)
    assert vis(0.0, 0.0) == pytest.approx(0.5, abs=1e-12)
    # centre at a quarter radius above the horizon: segment area fraction
    x = 0.25
    c = math.cos(math.pi / 2 - x * alpha)
    expect = (math.acos(-x) + x * math.sqrt(1 - x * x)) / math.pi
    assert vis(0.0, c) == pytest.approx(expect, rel=1e-9)


def test_lut_coordinate_round_trips(ref64, scene):
    cfg = scene.full_frame_config()
    p = params_of(cfg)
    R = 6_360_000.0
    lib = ref64.lib
    for f in ("oracle_fn_sun_u_to_cosine", "oracle_fn_sun_cosine_to_u", "oracle_fn_sky_v_to_zenith",
              "oracle_fn_sky_zenith_to_v", "oracle_fn_altitude_on_ray", "oracle_fn_warped_boundary"):
        getattr(lib, f).restype = ref64.real
    g = np.random.default_rng(7)
    for _ in range(200):
        h = float(g.uniform(0.0, 99_000.0))
        # transmittance (u, v) <-> (altitude, cosine) for rays that reach the top
        c = float(g.uniform(-0.05, 1.0))
        out2, out3 = _out(ref64, 2), _out(ref64, 3)
        horizon = -math.sqrt(h * (2 * R + h)) / (R + h)
        if c < horizon:
            continue
        lib.oracle_fn_transmittance_ray_to_uv(ctypes.byref(p), ref64.r(h), ref64.r(c), ref64.r(R),
                                              ref64.ptr(out2))
        lib.oracle_fn_transmittance_uv_to_ray(ctypes.byref(p), ref64.r(out2[0]), ref64.r(out2[1]),
                                              ref64.r(R), ref64.ptr(out3))
        assert out3[0] == pytest.approx(h, abs=1e-4)
        assert out3[1] == pytest.approx(c, abs=1e-8)
        # multi-scattering sun axis
        sc = float(g.uniform(-1.0, 1.0))
        u = lib.oracle_fn_sun_cosine_to_u(ref64.r(sc), ref64.r(h), ref64.r(R))
        assert lib.oracle_fn_sun_u_to_cosine(ref64.r(u), ref64.r(h), ref64.r(R)) == pytest.approx(sc, abs=1e-9)
        # sky-view zenith axis
        z = float(g.uniform(0.0, math.pi))
        v = lib.oracle_fn_sky_zenith_to_v(ctypes.byref(p), ref64.r(z), ref64.r(h), ref64.r(R))
        assert lib.oracle_fn_sky_v_to_zenith(ctypes.byref(p), ref64.r(v), ref64.r(h), ref64.r(R)) == pytest.approx(z, abs=1e-9)
        # stable altitude along a ray equals |p + t d| - R
        t = float(g.uniform(0.0, 400_000.0))
        exact = math.hypot(R + h + c * t, math.sqrt(max(1 - c * c, 0.0)) * t) - R
        assert lib.oracle_fn_altitude_on_ray(ref64.r(h), ref64.r(c), ref64.r(t), ref64.r(R)) == pytest.approx(exact, abs=1e-6)


def test_warped_boundaries_cover_the_interval(ref64):
    lib = ref64.lib
    lib.oracle_fn_warped_boundary.restype = ref64.real
    R = 6_360_000.0
    for cos in (0.7, -0.01, -0.3):
        ray = _arr(ref64, [math.sqrt(1 - cos * cos), cos, 0.0])
        origin = _arr(ref64, [0.0, R + 10_000.0, 0.0])
        start, end = 100.0, 250_000.0
        f = np.linspace(0.0, 1.0, 101)
        d = np.array([lib.oracle_fn_warped_boundary(ref64.ptr(origin), ref64.ptr(ray), ref64.r(start),
                                                    ref64.r(end), ref64.r(x)) for x in f])
        assert d[0] == pytest.approx(start) and d[-1] == pytest.approx(end)
        assert np.all(np.diff(d) >= 0.0)
        closest = -(R + 10_000.0) * cos
        if start < closest < end:  # boundaries concentrate at the closest approach
            assert d[50] == pytest.approx(closest, rel=1e-12)


def test_multi_scattering_ground_bounce_closed_form(ref64, pkg):
    """With no scattering medium, a texel at ~24 m altitude sees the ground
    over exactly half of the Fibonacci sphere (renderer.py:1096-1195), so
    Psi = 0.5 * albedo * E * cos(sun) / pi (Lambertian ground, unit
    transmittance, no feedback)."""
    cfg = pkg.AtmosphereConfig(
        rayleigh_scattering_per_m=(0.0, 0.0, 0.0), mie_scattering_per_m=(0.0, 0.0, 0.0),
        mie_extinction_per_m=(0.0, 0.0, 0.0), absorption_extinction_per_m=(0.0, 0.0, 0.0),
        ground_albedo=(0.3, 0.2, 0.1), multi_scattering_directions=64, multi_scattering_steps=20)
    R, E = 6_360_000.0, (4.0, 3.9, 3.7)
    ar = math.radians(0.266)
    tlut = ref64.build_transmittance(cfg, R)
    np.testing.assert_array_equal(tlut, 1.0)
    ms = ref64.build_multi_scattering(cfg, R, E, ar, tlut)
    lib = ref64.lib
    lib.oracle_fn_sun_u_to_cosine.restype = ref64.real
    v = 0.5 / cfg.multi_scattering_lut_height
    h = v * v * cfg.top_altitude_m
    for x in (20, 26, 31):
        sun_c = lib.oracle_fn_sun_u_to_cosine(ref64.r((x + 0.5) / cfg.multi_scattering_lut_width),
                                              ref64.r(h), ref64.r(R))
        expect = 0.5 * np.asarray(cfg.ground_albedo) * np.asarray(E) * sun_c / math.pi
        np.testing.assert_allclose(ms[x, 0], expect, rtol=2e-3)


def _sun(zenith_deg):
    z = math.radians(zenith_deg)
    return np.array([math.sin(z), math.cos(z), 0.0])


def test_sky_view_converges_to_single_scattering(ref64, scene):
    """With the multi-scattering LUT zeroed and the sun 60 deg above the
    horizon, a sky-view texel (renderer.py:1208-1302) is the direct single
    scattering integral; the validation model computes it by plain midpoint
    quadrature with its own sun transmittance."""
    cfg = scene.full_frame_config().replace(
        transmittance_lut_width=1024, transmittance_lut_height=512, transmittance_steps=256,
        sky_view_lut_width=16, sky_view_lut_height=32, sky_view_steps=256)
    R, h, E = 6_360_000.0, 1_000.0, (4.0, 3.9, 3.7)
    ar = math.radians(0.266)
    tlut = ref64.build_transmittance(cfg, R)
    zero_ms = np.zeros((cfg.multi_scattering_lut_width, cfg.multi_scattering_lut_height, 3))
    sun_z = 30.0
    sv = ref64.build_sky_view(cfg, R, h, math.cos(math.radians(sun_z)), E, ar, tlut, zero_ms)
    model = AtmosphereModel(R, cfg)
    lib = ref64.lib
    lib.oracle_fn_sky_v_to_zenith.restype = ref64.real
    p = params_of(cfg)
    for x, y in ((3, 4), (8, 10), (12, 12)):
        az = (x + 0.5) / cfg.sky_view_lut_width * math.pi
        zen = lib.oracle_fn_sky_v_to_zenith(ctypes.byref(p), ref64.r((y + 0.5) / cfg.sky_view_lut_height),
                                            ref64.r(h), ref64.r(R))
        assert zen < math.pi / 2  # upward rays: no ground along the view
        ray = np.array([math.sin(zen) * math.cos(az), math.cos(zen), math.sin(zen) * math.sin(az)])
        expect = model.single_scattering(np.array([0.0, R + h, 0.0]), ray, _sun(sun_z), E,
                                         steps=2_000, sun_steps=400)
        np.testing.assert_allclose(sv[x, y], expect, rtol=3e-3)


def test_surface_path_converges_to_validation_model(ref64, scene):
    """The composite's surface branch (renderer.py:1603-1786).  For a camera
    looking down at the surface the reversed ray climbs to the top unoccluded,
    so the ratio of the two T-LUT samples is the segment transmittance: it
    equals the model's bounded midpoint transmittance.  The direct limb
    integral over a path that leaves the atmosphere (MS zeroed, sun high, no
    terminator) equals the model's single scattering."""
    cfg = scene.full_frame_config().replace(
        transmittance_lut_width=1024, transmittance_lut_height=512, transmittance_steps=256,
        aerial_horizon_raymarch_steps=64)
    R, E = 6_360_000.0, (4.0, 3.9, 3.7)
    ar = math.radians(0.266)
    tlut = ref64.build_transmittance(cfg, R)
    zero_ms = np.zeros((cfg.multi_scattering_lut_width, cfg.multi_scattering_lut_height, 3))
    model = AtmosphereModel(R, cfg)
    sun = _sun(40.0)
    cases = [(2_000.0, -0.3, 5_000.0), (8_000.0, -0.1, 40_000.0), (20_000.0, -0.05, 150_000.0)]
    for h, c, d in cases:
        ray = np.array([math.sqrt(1 - c * c), c, 0.0])
        L, T = ref64.surface_path(cfg, np.array([h]), ray[None], np.array([d]), sun[None], E, R, ar,
                                  tlut, zero_ms)
        t_model = model.transmittance(np.array([0.0, R + h, 0.0]), ray, maximum_distance_m=d,
                                      steps=100_000)
        np.testing.assert_allclose(T[0], t_model, rtol=2e-3)
    for h, c in ((300.0, 0.4), (1_500.0, 0.1)):
        ray = np.array([math.sqrt(1 - c * c), c, 0.0])
        L, _ = ref64.surface_path(cfg, np.array([h]), ray[None], np.array([1.0e8]), sun[None], E, R,
                                  ar, tlut, zero_ms)
        expect = model.single_scattering(np.array([0.0, R + h, 0.0]), ray, sun, E, steps=2_000,
                                         sun_steps=400)
        np.testing.assert_allclose(L[0], expect, rtol=3e-3)


def test_small_planet_frame_invariants(ref32, scene, pkg):
    """LUT and HDR invariants of the reference's CPU render test
    (tests/test_m2_terrain.py:252-271) on its tiny LUT configuration."""
    cfg = scene.small_planet_config()
    inputs = scene.make_frame(width=64, height=48, config=cfg, planet_radius_m=50_000.0,
                              camera_altitude_range=(500.0, 2_000.0),
                              distance_range=(10.0, 30_000.0))
    out = ref32.frame(inputs)
    t = out["transmittance"]
    assert np.isfinite(t).all() and np.all((0.0 <= t) & (t <= 1.0))
    for name in ("sky_view", "multi_scattering", "aerial"):
        a = out[name]
        assert np.isfinite(a).all() and np.all(a >= 0.0) and a.max() > 0.0, name
    assert np.isfinite(out["hdr"]).all()


def test_float32_restatement_tracks_float64(ref32, ref64, scene):
    """The float32 build is the same approximation: at the LUT-chain
    configuration its LUTs stay within float32 rounding of the float64
    build (about 1e-3 at worst on grazing transmittance texels)."""
    cfg = scene.lut_chain_config()
    R, r_cam, sun = scene.lut_chain_state()
    E, ar = scene.SOLAR_IRRADIANCE, float(np.float32(math.radians(0.266)))
    out = {}
    for ref in (ref32, ref64):
        t = ref.build_transmittance(cfg, R)
        m = ref.build_multi_scattering(cfg, R, E, ar, t)
        s = ref.build_sky_view(cfg, R, float(np.float32(r_cam - R)), float(np.float32(sun[1])), E,
                               ar, t, m)
        out[ref.precision] = (t, m, s)
    for a, b in zip(out["f32"], out["f64"]):
        rel = np.abs(a - b) / np.maximum(np.abs(b), 1e-6 * np.abs(b).max())
        assert rel.max() < 5e-3 and np.median(rel) < 1e-5
