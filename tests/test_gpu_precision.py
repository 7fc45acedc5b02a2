"""Planet-scale precision stress on the GPU: the view centred on the limb
(20 % of the rows within 0.01 deg of tangency), camera altitudes 0.5 m to
2000 km, G-buffer endpoints up to 3000 km.

Both precision variants are held to the 1e-4 parity bar against their CPU
restatements (reference -> float32 oracle, robust -> robust oracle), and
their error against the float64 restatement is printed; the robust variant
must track float64 at least as well as the reference arithmetic does."""

import numpy as np
import pytest

from parity import PLAIN_TOLERANCE, rel_error, summary, worst_index
from test_gpu_parity import gpu_frame

pytestmark = pytest.mark.gpu

KEYS = ("transmittance", "multi_scattering", "sky_view", "aerial", "surface_sun_transmittance",
        "surface_sky_radiance", "hdr")


@pytest.fixture(scope="module")
def device():
    import torch

    assert torch.cuda.is_available()
    return torch.device("cuda:0")


@pytest.fixture(scope="module")
def ref32r():
    from oracle.ref_cpu import RefCpu

    return RefCpu("f32r")


@pytest.mark.parametrize("altitude", [0.5, 2_000.0, 100_000.0, 400_000.0, 2_000_000.0])
def test_stress_parity_both_precisions(pkg, scene, ref32, ref32r, ref64, device, altitude):
    inputs = scene.make_stress_frame(altitude, width=480, height=270)
    print(f"stress frame: h {altitude} m, 480x270", flush=True)
    want64 = ref64.frame(inputs)
    p99 = {}
    for precision, oracle in (("reference", ref32), ("robust", ref32r)):
        got = gpu_frame(pkg, inputs, device, precision=precision)
        want = oracle.frame(inputs)
        for key in KEYS:
            err = rel_error(got[key], want[key])
            s = summary(err)
            s64 = summary(rel_error(got[key], want64[key]))
            p99[(precision, key)] = s64["p99"]
            print(f"  {precision:9s} {key:26s} vs own oracle max {s['max']:.2e} | vs f64 "
                  f"p99 {s64['p99']:.2e} max {s64['max']:.2e}", flush=True)
            assert np.isfinite(got[key]).all(), (precision, key)
            assert s["max"] <= PLAIN_TOLERANCE, (precision, key, worst_index(err))
    for key in KEYS:
        assert p99[("robust", key)] <= max(1.5 * p99[("reference", key)], 2e-5), key


def test_precision_argument_is_checked(pkg, scene, device):
    with pytest.raises(ValueError, match="precision must be one of"):
        pkg.AtmosphereRenderer(16, 16, scene.full_frame_config(), device, precision="fp64")


@pytest.mark.parametrize("seed", [2, 3])
def test_robust_matches_reference_oracle_where_float32_is_well_conditioned(pkg, scene, ref32,
                                                                           ref64, device, seed):
    """The robust variant is a restatement of the reference, not another
    model: on every value where the reference's float32 arithmetic is
    well-conditioned -- the float32 oracle within 1e-5 of the float64
    restatement -- the robust GPU path matches the float32 (reference)
    oracle within the 1e-4 parity bar.  Full-frame configuration, 480x270."""
    inputs = scene.make_frame(width=480, height=270, config_index=1, seed_offset=seed)
    print(f"seed {seed}: camera {inputs.camera_altitude_m:.1f} m, sun y "
          f"{inputs.sun_local[1]:.3f}", flush=True)
    got = gpu_frame(pkg, inputs, device, precision="robust")
    want32, want64 = ref32.frame(inputs), ref64.frame(inputs)
    for key in KEYS:
        conditioned = rel_error(want32[key], want64[key]) <= 1.0e-5
        err = rel_error(got[key], want32[key])[conditioned]
        share = float(conditioned.mean())
        worst = float(err.max()) if err.size else 0.0
        print(f"  {key:26s} well-conditioned share {share:.4f}, robust vs float32 oracle there "
              f"max {worst:.2e}", flush=True)
        assert share > 0.4, key  # 45-100 % of each array on these frames
        assert worst <= PLAIN_TOLERANCE, key


# Upper bounds of the shipped surrogate's hdr p99 against float64 on the
# in-domain stress frames (its routing on), about 2x the values measured on
# MI355X at 480x270: 0.187 at 0.5 m, 4.2e-3 at 2 km, 2.9e-3 at 100 km.  The
# limb-centred view from 0.5 m, two thirds near-horizon rays, misses the 1 %
# gate; a regression fails here.
STRESS_SURROGATE_P99 = {0.5: 0.4, 2_000.0: 1e-2, 100_000.0: 1e-2}


@pytest.mark.parametrize("altitude", [0.5, 2_000.0, 100_000.0, 400_000.0, 2_000_000.0])
def test_stress_surrogate_error_vs_fp64(pkg, scene, ref64, device, altitude):
    """The surrogate leg of the precision stress frames: the shipped network
    in the composite at each altitude of the sweep, its hdr error against
    the float64 restatement reported beside the plain path's.  Cameras above
    the network's 100 km training domain fall back to the plain integrator
    for every pixel, so there the surrogate frame is the plain frame bit for
    bit; inside the domain its error (with the shipped routing) is bounded by
    the values measured on these frames."""
    import importlib

    import torch

    sur = importlib.import_module(pkg.__name__ + ".surrogate")
    net = sur.SurrogateMLP.load(sur.default_weights_path(), max_batch=1024, device=device)
    inputs = scene.make_stress_frame(altitude, width=480, height=270)
    want64 = ref64.frame(inputs)["hdr"]
    light = inputs.lighting
    r = pkg.AtmosphereRenderer(inputs.width, inputs.height, inputs.config, device)
    to = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(device)
    pos, sid, sc, sp = (to(inputs.position_view), to(inputs.surface_id), to(inputs.scene_hdr),
                        to(inputs.space_hdr))
    R, rc = inputs.planet_radius_m, inputs.camera_radius_m
    r.update(R, rc, inputs.sun_local, light.solar_irradiance, light.sun_angular_radius_degrees)
    outs = {}
    for name, s in (("plain", None), ("surrogate", net)):
        hdr = torch.zeros_like(sc)
        r.composite(sc, sp, sid, pos, hdr, R, rc, inputs.sun_local, light.solar_irradiance,
                    light.sun_disk_radiance, light.sun_angular_radius_degrees,
                    inputs.view_basis, inputs.tangent_half_fov, *inputs.star_contrast,
                    surrogate=s)
        torch.cuda.synchronize(device)
        outs[name] = hdr.cpu().numpy()
    e_plain = summary(rel_error(outs["plain"], want64))
    e_sur = summary(rel_error(outs["surrogate"], want64))
    print(f"stress h {altitude} m: hdr vs f64 p99 / max  plain {e_plain['p99']:.2e} / "
          f"{e_plain['max']:.2e}  surrogate {e_sur['p99']:.2e} / {e_sur['max']:.2e}", flush=True)
    if altitude > 100_000.0:
        assert (outs["plain"].view(np.uint32) == outs["surrogate"].view(np.uint32)).all()
    else:
        assert np.isfinite(outs["surrogate"]).all()
        assert e_sur["p99"] <= STRESS_SURROGATE_P99[altitude]
    r.close()
    net.close()

