"""The float64 validation model (oracle/model_fp64.py) against the reference's
own known answers (tests/test_m3_atmosphere.py:24-99) and closed forms."""

import math

import numpy as np
import pytest

from closed_forms import vertical_optical_depth
from oracle.model_fp64 import AtmosphereModel, ray_sphere_roots


@pytest.fixture(scope="module")
def model(pkg):
    """The reference tests' small planet (tests/test_m3_atmosphere.py:11-21)."""
    return AtmosphereModel(
        50_000.0,
        pkg.AtmosphereConfig(
            top_altitude_m=15_000.0,
            rayleigh_scale_height_m=3_000.0,
            mie_scale_height_m=800.0,
            absorption_peak_altitude_m=7_500.0,
            absorption_half_width_m=5_000.0,
        ),
    )


def test_ray_sphere_roots_from_outside_and_inside():
    """tests/test_m3_atmosphere.py:24-39"""
    outside = ray_sphere_roots(np.array([0.0, 70_000.0, 0.0]), np.array([0.0, -1.0, 0.0]), 65_000.0)
    np.testing.assert_allclose(outside, (5_000.0, 135_000.0), atol=1.0e-8)
    inside = ray_sphere_roots(np.array([0.0, 55_000.0, 0.0]), np.array([0.0, 1.0, 0.0]), 65_000.0)
    np.testing.assert_allclose(inside, (-120_000.0, 10_000.0), atol=1.0e-8)
    assert ray_sphere_roots(np.array([0.0, 70_000.0, 0.0]), np.array([1.0, 0.0, 0.0]), 65_000.0) is None


def test_shell_interval_stops_at_opaque_ground(model):
    """tests/test_m3_atmosphere.py:42-58"""
    interval = model.atmosphere_interval(np.array([0.0, 51_000.0, 0.0]), np.array([0.0, -1.0, 0.0]))
    assert interval is not None and interval.ends_at_ground
    assert math.isclose(interval.end_m, 1_000.0, abs_tol=1.0e-8)
    outward = model.atmosphere_interval(np.array([0.0, 50_000.0, 0.0]), np.array([0.0, 1.0, 0.0]))
    assert outward is not None and not outward.ends_at_ground
    assert math.isclose(outward.end_m, 15_000.0, abs_tol=1.0e-8)


def test_density_and_transmittance_ranges(model):
    """tests/test_m3_atmosphere.py:61-84"""
    sea, high = model.density(0.0), model.density(12_000.0)
    assert sea[0] > high[0] >= 0.0 and sea[1] > high[1] >= 0.0
    origin = np.array([0.0, 50_001.0, 0.0])
    t = model.transmittance(origin, np.array([0.0, 1.0, 0.0]), steps=64)
    assert np.isfinite(t).all() and np.all((0.0 <= t) & (t <= 1.0)) and t[2] < t[0]
    blocked = model.transmittance(origin, np.array([0.0, -1.0, 0.0]), steps=16, ground_is_opaque=True)
    np.testing.assert_array_equal(blocked, np.zeros(3))


def test_single_scattering_is_finite_and_non_negative(model):
    """tests/test_m3_atmosphere.py:87-99"""
    radiance = model.single_scattering(np.array([0.0, 50_010.0, 0.0]), np.array([0.0, 0.2, 1.0]),
                                       np.array([0.0, 0.7, 0.7]), (4.0, 3.9, 3.7), steps=12,
                                       sun_steps=12)
    assert np.isfinite(radiance).all() and np.all(radiance >= 0.0) and radiance.max() > 0.0


@pytest.mark.parametrize("h0", [0.0, 1_200.0, 9_000.0])
def test_model_transmittance_matches_closed_form(model, h0):
    tau = vertical_optical_depth(model.config, h0)
    origin = np.array([0.0, model.bottom_radius_m + h0, 0.0])
    t = model.transmittance(origin, np.array([0.0, 1.0, 0.0]), steps=4000)
    np.testing.assert_allclose(t, np.exp(-tau), rtol=1.0e-7)


def test_phase_functions_are_normalised(model):
    """Rayleigh and Cornette-Shanks phases integrate to 1 over the sphere;
    Cornette-Shanks with g = 0 is the Rayleigh phase."""
    c = np.linspace(-1.0, 1.0, 400_001)
    pr = np.array([model.rayleigh_phase(v) for v in c])
    assert abs(2.0 * math.pi * np.trapezoid(pr, c) - 1.0) < 1.0e-9
    pm = np.array([model.mie_phase(v) for v in c])
    assert abs(2.0 * math.pi * np.trapezoid(pm, c) - 1.0) < 1.0e-6
    iso = AtmosphereModel(50_000.0, model.config.replace(mie_phase_g=0.0))
    for v in (-1.0, -0.3, 0.0, 0.5, 1.0):
        assert math.isclose(iso.mie_phase(v), iso.rayleigh_phase(v), rel_tol=1.0e-14)
