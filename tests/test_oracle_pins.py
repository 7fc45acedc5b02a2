"""Pins of the oracle functions that the reference's own tests leave open
(the multi-scattering closure with a scattering medium, the aerial
perspective slices, the camera transmittance, the sky-view sampler, the
star fade and sun-disk coverage, the terminator refinement weight and its
convergence, the solar shadow roots, the surface lighting, the diagnostic
views), and proof that they pin: each deliberately broken copy of the
float64 oracle below -- one plausible mistake planted in a copy of
oracle/ref_cpu.c, compiled on the CPU -- fails the pin it targets."""

from __future__ import annotations

import concurrent.futures
import subprocess
from pathlib import Path

import pytest

import oracle_pins
from oracle.ref_cpu import RefCpu

REPO = Path(__file__).resolve().parents[1]
SOURCE = REPO / "oracle" / "ref_cpu.c"


@pytest.mark.parametrize("name", sorted(oracle_pins.ALL_PINS))
def test_oracle_meets_pin(ref64, name):
    score = oracle_pins.ALL_PINS[name](ref64)
    print(f"{name}: worst error / tolerance = {score:.3g}")
    assert score <= 1.0


# (mistake, pin that must catch it, text in ref_cpu.c, replacement)
MISTAKES = [
    ("ms_closure_dropped", "multi_scattering",
     "vstore(mslut, i, vmaxs(vdiv(avg_l, vsub(V(1.0, 1.0, 1.0), avg_f)), 0.0));",
     "vstore(mslut, i, vmaxs(avg_l, 0.0));"),
    ("ms_feedback_from_extinction", "multi_scattering",
     "dir_f = vadd(dir_f, vmul(vmul(path_t, scattering(&k, altitude)), seg_i));",
     "dir_f = vadd(dir_f, vmul(vmul(path_t, extinction(&k, altitude)), seg_i));"),
    ("ms_ground_bounce_without_pi", "multi_scattering",
     "fmax(gsc, 0.0)),", "fmax(gsc, 0.0) * k.pi),"),
    ("aerial_depth_one_slice_ahead", "aerial_slices",
     "real depth_fraction = (real)di / (real)(k.ad - 1);",
     "real depth_fraction = (real)(di + 1) / (real)(k.ad - 1);"),
    ("camera_t_ground_lit", "camera_transmittance",
     "transmission = V(0.0, 0.0, 0.0);", "transmission = V(1.0, 1.0, 1.0);"),
    ("sky_view_azimuth_full_turn", "sky_view_sampler",
     "real u = relative_azimuth / k->pi;", "real u = relative_azimuth / (2.0 * k->pi);"),
    ("sun_disk_filter_sign", "sky_composite",
     "(angular_radius + pixel_filter - angular_distance)",
     "(angular_radius - pixel_filter - angular_distance)"),
    ("star_fade_linear", "sky_composite",
     "vis = vis * vis * (3.0 - 2.0 * vis);", "vis = vis * 1.0;"),
    ("star_without_transmittance", "sky_composite",
     "v3 star = vmul(vload(space_hdr, i), trans);", "v3 star = vload(space_hdr, i);"),
    ("terminator_proximity_one_width", "terminator_weight",
     "real proximity = clampr(2.0 - closest_root / interval_width, 0.0, 1.0);",
     "real proximity = clampr(1.0 - closest_root / interval_width, 0.0, 1.0);"),
    ("terminator_penumbra_unscaled", "terminator_weight",
     "real penumbra = 4.0 * visibility * (1.0 - visibility);",
     "real penumbra = visibility * (1.0 - visibility);"),
    ("shadow_root_validity_sign", "shadow_roots",
     "if (vdot(p0, sun) < 0.0) *valid0 = 1;", "if (vdot(p0, sun) > 0.0) *valid0 = 1;"),
    ("refinement_weight_ignored", "terminator_refinement",
     "*radiance = vadd(vs(coarse_l, 1.0 - weight), vs(fine_l, weight));",
     "*radiance = coarse_l;"),
    ("lighting_at_camera_altitude", "prepare_surface_lighting",
     "real altitude = altitude_on_ray(camera_altitude, local_ray.y,",
     "real altitude = camera_altitude + 0.0 * altitude_on_ray(camera_altitude, local_ray.y,"),
    ("ms_sample_linear_altitude", "prepare_surface_lighting",
     "real v = sqrt(altitude_fraction);", "real v = altitude_fraction;"),
    ("froxel_depth_scale", "diagnostic_views",
     "real z = clampr(depth_fraction, 0.0, 1.0) * (real)(ad - 1);",
     "real z = clampr(depth_fraction, 0.0, 1.0) * (real)(ad - 2);"),
    ("transmittance_view_axis", "diagnostic_views",
     "int lx = (int)(su * (real)k.tw);", "int lx = (int)(sv * (real)k.tw);"),
    ("blend_weight_inverted", "diagnostic_views",
     "real w = surface ? 1.0 : 0.0;", "real w = surface ? 0.0 : 1.0;"),
]


def _build(mistake, text, tmp: Path) -> Path:
    name, _, old, new = mistake
    assert text.count(old) == 1, f"{name}: target text must occur once in ref_cpu.c"
    source = text.replace(old, new).replace('#include "../include/atmo.h"',
                                            f'#include "{REPO / "include" / "atmo.h"}"')
    src = tmp / f"{name}.c"
    src.write_text(source)
    out = tmp / f"lib{name}.so"
    cmd = ["gcc", "-O1", "-std=c11", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off",
           "-fno-fast-math", "-fno-math-errno", "-fvisibility=hidden", "-DREAL=double",
           "-w", "-o", str(out), str(src), "-lm"]
    subprocess.run(cmd, check=True, capture_output=True)
    return out


@pytest.fixture(scope="module")
def broken_oracles(tmp_path_factory):
    tmp = tmp_path_factory.mktemp("broken_oracles")
    text = SOURCE.read_text()
    with concurrent.futures.ThreadPoolExecutor(max_workers=8) as pool:
        paths = list(pool.map(lambda m: _build(m, text, tmp), MISTAKES))
    return {m[0]: p for m, p in zip(MISTAKES, paths)}


@pytest.mark.parametrize("mistake", MISTAKES, ids=[m[0] for m in MISTAKES])
def test_planted_mistake_fails_its_pin(broken_oracles, mistake):
    name, pin, _, _ = mistake
    broken = RefCpu("f64", path=broken_oracles[name])
    score = oracle_pins.ALL_PINS[pin](broken)
    print(f"{name}: {pin} worst error / tolerance = {score:.3g}")
    assert score > 1.0, f"the pin {pin} does not catch the planted mistake {name}"
