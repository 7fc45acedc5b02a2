"""Pins of the surface-shading oracle (oracle/surface_cpu.c) away from the
special cases of test_surface_cpu.py: oblique ocean specular, reciprocity
and the refracted seabed lookup at oblique incidence, each checked against
an independent float64 statement of the mathematics.

  - specular: at random oblique sun, view and water-normal directions the
    oracle's sun specular equals Cook-Torrance with the GGX distribution
    (Walter et al. 2007), the Schlick-GGX geometry term with
    k = (r + 1)^2 / 8 (Karis 2013) and Schlick's Fresnel at v.h, with
    alpha^2 the reference's combined slope variance (renderer.py:1327-1376);
  - reciprocity: that BRDF is symmetric in sun and view, so swapping them
    changes the radiance only by the n.l factor;
  - refraction: a tilted water plane over a parallel terrain plane, Snell's
    law by angles (sin t = sin i / n), the reference's depth heuristics
    (renderer.py:1394-1433), the pinhole projection and bilinear sampling of
    a per-pixel terrain albedo give each pixel's seabed
    (renderer.py:1434-1535); and with a small offset limit the lookup point
    is clamped to it (renderer.py:1457-1463).

Each pin returns its worst error over its tolerance; planted mistakes in
copies of surface_cpu.c must raise it above 1."""

from __future__ import annotations

import math
import subprocess
from pathlib import Path

import numpy as np
import pytest

from oracle.surface_cpu import shade

REPO = Path(__file__).resolve().parents[1]
SOURCE = REPO / "oracle" / "surface_cpu.c"
OCEAN = 2_147_000_000
E = np.float64([4.0, 3.9, 3.7])


def _params(**changes):
    p = {
        "sun_global": (0.0, 1.0, 0.0), "solar_irradiance": tuple(E),
        "view_right_global": (1.0, 0.0, 0.0), "view_up_global": (0.0, 1.0, 0.0),
        "view_forward_global": (0.0, 0.0, 1.0), "advanced_ocean_enabled": 1,
        "ocean_roughness": 0.12, "ocean_f0": 0.02, "ocean_sky_strength": 1.0,
        "ocean_absorption": (0.0, 0.0, 0.0), "ocean_scattering": (0.0, 0.0, 0.0),
        "ocean_max_depth": 80.0, "ocean_refraction_index": 1.333,
        "ocean_refraction_strength": 0.0, "ocean_refraction_max_offset_pixels": 48.0,
        "sun_angular_radius": math.radians(0.266), "tangent_half_fov": math.tan(math.radians(30)),
        "mode": 0,
    }
    p.update(changes)
    return p


def _fields(w, h):
    z3 = np.zeros((w, h, 3), np.float32)
    return {
        "surface_id": np.full((w, h), OCEAN, np.int32), "albedo": z3 + 0.5,
        "normal": np.tile(np.float32([0, 1, 0]), (w, h, 1)),
        "position": np.tile(np.float32([0, 0, 10]), (w, h, 1)),
        "ocean_slope_variance": np.zeros((w, h), np.float32),
        "water_depth_m": np.full((w, h), 5.0, np.float32), "seabed_albedo": z3.copy(),
        "seabed_normal": np.tile(np.float32([0, 1, 0]), (w, h, 1)),
        "terrain_surface_id": np.zeros((w, h), np.int32), "terrain_position": z3.copy(),
        "terrain_albedo": z3.copy(), "terrain_normal": np.tile(np.float32([0, 1, 0]), (w, h, 1)),
        "sun_transmittance": z3 + 1.0, "sky_radiance": z3.copy(),
    }


def _unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def textbook_specular(n, v, l, roughness, f0, slope, sun_radius):
    """Cook-Torrance sun radiance per unit irradiance, float64:
    D G F (n.l) / (4 (n.v) (n.l))."""
    h = _unit(v + l)
    nh, nv, nl, vh = (np.sum(a * b, -1) for a, b in ((n, h), (n, v), (n, l), (v, h)))
    a2 = roughness ** 4 + slope + math.tan(sun_radius) ** 2
    d = a2 / (math.pi * (nh * nh * (a2 - 1.0) + 1.0) ** 2)
    k = (a2 ** 0.25 + 1.0) ** 2 / 8.0
    g = (nv / (nv * (1.0 - k) + k)) * (nl / (nl * (1.0 - k) + k))
    f = f0 + (1.0 - f0) * (1.0 - vh) ** 5
    return d * g * f * nl / (4.0 * nv * nl)


def _oblique_configs(count=64, seed=11):
    """Water normals within 50 degrees of up; sun and view on its side with
    n.l, n.v in [0.15, 0.95] and no mirror alignment needed."""
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < count:
        n = _unit(rng.normal(size=3) * [0.5, 0.0, 0.5] + [0.0, 1.0, 0.0])
        v, l = _unit(rng.normal(size=3)), _unit(rng.normal(size=3))
        nv, nl = float(n @ v), float(n @ l)
        if 0.15 <= nv <= 0.95 and 0.15 <= nl <= 0.95:
            out.append((n, v, l, float(rng.uniform(0.0, 0.01))))
    return out


def _specular_run(configs, library, swap=False):
    """Oracle radiance of single ocean pixels whose output is the sun
    specular alone: no sky, no water extinction or scattering, a black
    seabed, no refraction lookup."""
    got = []
    for n, v, l, slope in configs:
        if swap:
            v, l = l, v
        f = _fields(1, 1)
        f["normal"][...] = n
        f["position"][...] = -v * 10.0  # view direction = -normalize(position)
        f["ocean_slope_variance"][...] = slope
        out, path = shade(f, _params(sun_global=tuple(l)), 1, 1, library)
        assert path[0, 0] == 2
        got.append(out[0, 0].astype(np.float64))
    return np.array(got)


def pin_oblique_specular(library=None) -> float:
    configs = _oblique_configs()
    got = _specular_run(configs, library)
    p = _params()
    want = np.array([E * textbook_specular(n, v, l, p["ocean_roughness"], p["ocean_f0"], s,
                                           p["sun_angular_radius"]) for n, v, l, s in configs])
    return float(np.max(np.abs(got - want) / want)) / 2e-5


def pin_specular_reciprocity(library=None) -> float:
    configs = _oblique_configs(seed=12)
    a = _specular_run(configs, library)
    b = _specular_run(configs, library, swap=True)
    nl = np.array([n @ l for n, v, l, _ in configs])[:, None]
    nv = np.array([n @ v for n, v, l, _ in configs])[:, None]
    brdf_a, brdf_b = a / (E * nl), b / (E * nv)  # radiance / (E n.l) with the roles swapped
    # float32 directions: near the GGX peak D's slope amplifies their rounding
    return float(np.max(np.abs(brdf_a - brdf_b) / brdf_a)) / 1e-4


# ------------------------------------------------------------- refraction

W, H = 96, 64
THF = math.tan(math.radians(30))
NORMAL = _unit([0.2, 0.5, -0.8])  # water plane facing the camera, tilted in x and y
DEPTH, INDEX, MAX_DEPTH = 8.0, 1.333, 80.0


def _refraction_scene(seed=7):
    aspect = W / H
    sx = ((np.arange(W) + 0.5) / W * 2 - 1) * THF * aspect
    sy = ((np.arange(H) + 0.5) / H * 2 - 1) * THF
    rays = _unit(np.stack(np.broadcast_arrays(sx[:, None], sy[None, :], 1.0), axis=-1))
    c = NORMAL @ np.float64([0.0, 0.0, 12.0])
    along = rays @ NORMAL  # < 0: every pixel sees the plane's upper side
    assert (along < 0).all()
    water = rays * (c / along)[..., None]
    terrain = rays * ((c - DEPTH) / along)[..., None]
    f = _fields(W, H)
    f["normal"][...] = NORMAL
    f["position"] = water.astype(np.float32)
    f["water_depth_m"][...] = DEPTH
    f["terrain_position"] = terrain.astype(np.float32)
    f["terrain_albedo"] = np.random.default_rng(seed).uniform(0.05, 0.95, (W, H, 3)).astype(np.float32)
    f["sun_transmittance"][...] = 0.0  # no sun: no specular, no direct light
    f["sky_radiance"][...] = 1.0       # the seabed is lit by a unit sky
    return f


def _expected_lookup(f, max_offset):
    """Snell by angles, the reference's depth heuristics, pinhole
    projection, the offset clamp and bilinear sampling, in float64."""
    aspect = W / H
    pos = f["position"].astype(np.float64)
    i = _unit(pos)
    cos_i = np.maximum(-(i @ NORMAL), 0.0)
    sin_i = np.sqrt(1.0 - cos_i ** 2)
    sin_t = sin_i / INDEX
    cos_t = np.sqrt(1.0 - sin_t ** 2)
    tangent = (i + cos_i[..., None] * NORMAL) / sin_i[..., None]
    refracted = sin_t[..., None] * tangent - cos_t[..., None] * NORMAL
    vertical = DEPTH * np.maximum(cos_i, 0.05)
    dist = np.minimum(vertical / np.maximum(cos_t, 0.05), 2.0 * MAX_DEPTH)
    end = pos + refracted * dist[..., None]
    spx = (end[..., 0] / (end[..., 2] * THF * aspect) * 0.5 + 0.5) * W - 0.5
    spy = (end[..., 1] / (end[..., 2] * THF) * 0.5 + 0.5) * H - 0.5
    x, y = np.meshgrid(np.arange(W), np.arange(H), indexing="ij")
    dx, dy = spx - x, spy - y
    length = np.hypot(dx, dy)
    s = np.where(length > max_offset, max_offset / np.maximum(length, 1e-5), 1.0)
    spx, spy = x + dx * s, y + dy * s
    bx, by = np.floor(spx).astype(int), np.floor(spy).astype(int)
    fx, fy = spx - bx, spy - by
    inside = (bx >= 0) & (by >= 0) & (bx + 1 < W) & (by + 1 < H)
    bx, by = np.clip(bx, 0, W - 2), np.clip(by, 0, H - 2)
    alb = f["terrain_albedo"].astype(np.float64)
    want = (alb[bx, by] * ((1 - fx) * (1 - fy))[..., None] + alb[bx + 1, by] * (fx * (1 - fy))[..., None]
            + alb[bx, by + 1] * ((1 - fx) * fy)[..., None] + alb[bx + 1, by + 1] * (fx * fy)[..., None])
    return want, inside, length


def _refraction_score(library, max_offset) -> float:
    f = _refraction_scene()
    p = _params(ocean_f0=0.0, ocean_sky_strength=0.0, ocean_refraction_strength=1.0,
                ocean_refraction_index=INDEX, ocean_refraction_max_offset_pixels=max_offset,
                tangent_half_fov=THF)
    out, path = shade(f, p, W, H, library)
    want, inside, length = _expected_lookup(f, max_offset)
    assert inside.mean() > 0.6 and np.median(length) > 2.0  # a real, bent lookup
    # out = seabed albedo x unit sky x (1 - Schlick Fresnel at n.v), f0 = 0
    nv = np.maximum(-(_unit(f["position"].astype(np.float64)) @ NORMAL), 1e-4)
    got = out.astype(np.float64) / (1.0 - (1.0 - nv) ** 5)[..., None]
    ok = inside & (path == 3)
    if ok.mean() < 0.6:
        return math.inf
    return float(np.max(np.abs(got - want)[ok])) / 2e-4


def pin_oblique_refraction(library=None) -> float:
    return _refraction_score(library, 48.0)


def pin_refraction_offset_clamp(library=None) -> float:
    return _refraction_score(library, 0.75)


PINS = {
    "oblique_specular": pin_oblique_specular,
    "specular_reciprocity": pin_specular_reciprocity,
    "oblique_refraction": pin_oblique_refraction,
    "refraction_offset_clamp": pin_refraction_offset_clamp,
}


@pytest.mark.parametrize("pin", list(PINS))
def test_oracle_meets_pin(pin):
    score = PINS[pin]()
    print(f"{pin}: worst error / tolerance = {score:.3g}")
    assert score <= 1.0


# (name, pin that must catch it, text in surface_cpu.c, replacement)
MISTAKES = [
    ("fresnel_sun_from_view", "oblique_specular",
     "const float omh = 1.0f - vdoth;", "const float omh = 1.0f - ndotv;"),
    ("geometry_k_unsquared", "oblique_specular",
     "const float geometry_k = gk1 * gk1 / 8.0f;", "const float geometry_k = gk1 / 8.0f;"),
    ("ggx_denominator_unsquared", "oblique_specular",
     "fmaxf(PI * denominator * denominator, 1.0e-5f)", "fmaxf(PI * denominator, 1.0e-5f)"),
    ("geometry_view_twice", "specular_reciprocity",
     "const float geometry_l = ndotl / fmaxf(ndotl * (1.0f - geometry_k) + geometry_k, 1.0e-5f);",
     "const float geometry_l = ndotv / fmaxf(ndotv * (1.0f - geometry_k) + geometry_k, 1.0e-5f);"),
    ("refraction_index_not_inverted", "oblique_refraction",
     "const float eta = 1.0f / p->ocean_refraction_index;",
     "const float eta = p->ocean_refraction_index;"),
    ("refraction_root_sign", "oblique_refraction",
     "eta * cos_incident - sqrtf(refract_discriminant)",
     "eta * cos_incident + sqrtf(refract_discriminant)"),
    ("projection_without_aspect", "oblique_refraction",
     "(endpoint.z * thf * aspect)", "(endpoint.z * thf)"),
    ("bilinear_weights_swapped", "oblique_refraction",
     "const float wx = ox == 0 ? 1.0f - fx : fx;", "const float wx = ox == 0 ? fx : 1.0f - fx;"),
    ("offset_clamp_ignored", "refraction_offset_clamp",
     "if (delta_length > max_offset) {", "if (delta_length > 1.0e30f * max_offset) {"),
]


@pytest.fixture(scope="module")
def broken_shading(tmp_path_factory):
    tmp = tmp_path_factory.mktemp("broken_shading")
    text = SOURCE.read_text()
    libs = {}
    for name, _, old, new in MISTAKES:
        assert text.count(old) == 1, f"{name}: target text must occur once in surface_cpu.c"
        src = tmp / f"{name}.c"
        src.write_text(text.replace(old, new).replace(
            '#include "../include/atmo.h"', f'#include "{REPO / "include" / "atmo.h"}"'))
        out = tmp / f"lib{name}.so"
        subprocess.run(["gcc", "-O1", "-std=c11", "-fPIC", "-shared", "-ffp-contract=off",
                        "-fno-fast-math", "-fvisibility=hidden", "-w", "-o", str(out), str(src),
                        "-lm"], check=True, capture_output=True)
        libs[name] = out
    return libs


@pytest.mark.parametrize("mistake", MISTAKES, ids=[m[0] for m in MISTAKES])
def test_planted_mistake_fails_its_pin(broken_shading, mistake):
    name, pin, _, _ = mistake
    score = PINS[pin](broken_shading[name])
    print(f"{name}: {pin} worst error / tolerance = {score:.3g}")
    assert score > 1.0, f"the pin {pin} does not catch the planted mistake {name}"
