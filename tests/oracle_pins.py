"""Pins of the CPU oracle's (oracle/ref_cpu.c) per-pixel samplers, LUT builds
and diagnostic views by what the mathematics fixes.  TEST INFRASTRUCTURE.

Each `pin_*` function evaluates one group of oracle functions on chosen
inputs and returns the worst error divided by its tolerance (<= 1 passes).
The expected values come from:

  * closed forms and constructions whose answer is known (a shadow cylinder
    met at right angles, a root placed exactly at an interval's midpoint, a
    LUT filled with a linear function of its texel indices -- bilinear and
    trilinear interpolation reproduce it exactly, so the sampled value fixes
    the sampler's coordinates);
  * the independent float64 quadratures of tests/fp64_paths.py and the
    validation model (oracle/model_fp64.py), to which the oracle must
    converge where both define the same integral.

tests/test_oracle_pins.py runs every pin on the float64 oracle and on
deliberately broken copies of it, each of which must fail its pin.  The
functions take any RefCpu-like object so both can share them.
"""

from __future__ import annotations

import ctypes
import dataclasses
import functools
import math

import numpy as np

import __graft_entry__ as entry
from fp64_paths import Medium, multi_scattering_texel, path_integral, sun_transmittance
from oracle.model_fp64 import AtmosphereModel
from oracle.ref_cpu import params_of

R_EARTH = 6_360_000.0
E = (4.0, 3.9, 3.7)
ALPHA = math.radians(0.266)
LUM = np.array([0.2126, 0.7152, 0.0722])


def _scene():
    return entry.load_submodule("scene")


class Cfg(dict):
    """A configuration mapping with attribute access: the oracle takes any
    AtmoParams values, including step counts beyond the configuration's
    validation ranges (converged quadrature for these pins)."""

    def __getattr__(self, name):
        try:
            return self[name]
        except KeyError as exc:
            raise AttributeError(name) from exc

    def with_(self, **changes):
        out = Cfg(self)
        out.update(changes)
        return out


def config(**changes) -> Cfg:
    base = dataclasses.asdict(_scene().full_frame_config())
    return Cfg(base).with_(**changes)


def converged_lut_config(**changes) -> Cfg:
    """T-LUT at 1024 x 512 texels and 256 steps: bilinear samples of it agree
    with directly integrated transmittance to about 1e-4."""
    return config(transmittance_lut_width=1024, transmittance_lut_height=512,
                  transmittance_steps=256, **changes)


def rel(got, want, floor_fraction=1.0e-6):
    got = np.asarray(got, dtype=np.float64)
    want = np.asarray(want, dtype=np.float64)
    peak = np.abs(want).reshape(-1, want.shape[-1]).max(axis=0) if want.ndim else abs(want)
    floor = np.maximum(floor_fraction * peak, 1.0e-300)
    return np.abs(got - want) / np.maximum(np.abs(want), floor)


def linear_lut(shape, seed=0):
    """LUT (X, Y[, Z], 3) filled with a + sum_k b_k i_k per channel, plus the
    coefficients (a (3,), b (ndim, 3)); every texel is at least 1, so the
    samplers' clamps at zero never act."""
    g = np.random.default_rng(seed)
    b = g.uniform(0.01, 0.05, (len(shape) - 1, 3)) * g.choice([-1.0, 1.0], (len(shape) - 1, 3))
    a = 1.0 + g.uniform(0.0, 1.0, 3) + (np.abs(b) * (np.asarray(shape[:-1])[:, None] - 1)).sum(0)
    idx = np.indices(shape[:-1]).astype(np.float64)
    lut = a + np.einsum("k...,kc->...c", idx, b)
    return np.ascontiguousarray(lut), a, b


def _restype(ref, name):
    fn = getattr(ref.lib, name)
    fn.restype = ref.real
    return fn


@functools.lru_cache(maxsize=None)
def _real_tlut(key):
    """T-LUT of the unmodified float64 oracle for a configuration key (the
    broken copies never change the T-LUT build, so they share it)."""
    from oracle.ref_cpu import RefCpu

    cfg = Cfg(dict(key))
    return RefCpu("f64").build_transmittance(cfg, R_EARTH)


def tlut_for(cfg):
    key = tuple((k, tuple(v) if isinstance(v, (list, tuple)) else v) for k, v in cfg.items())
    return _real_tlut(key)


# --------------------------------------------------------------------------
# multi-scattering: the closure L / (1 - f) with a scattering medium
# --------------------------------------------------------------------------

MS_TEXELS = ((28, 0), (26, 20), (31, 10))
MS_TOL = 1.0e-3


def ms_config():
    return converged_lut_config(multi_scattering_steps=128)


@functools.lru_cache(maxsize=None)
def _ms_expected():
    cfg = ms_config()
    out = {}
    for x, y in MS_TEXELS:
        v = (y + 0.5) / cfg.multi_scattering_lut_height
        h = v * v * cfg.top_altitude_m
        mu = _sun_u_to_cosine((x + 0.5) / cfg.multi_scattering_lut_width, h, R_EARTH)
        out[(x, y)] = multi_scattering_texel(cfg, R_EARTH, h, mu, E, steps=600, sun_steps=120)
    return out


def _shell_rho(h, bottom):
    return math.sqrt(max(h, 0.0) * (2.0 * bottom + max(h, 0.0)))


def _sun_u_to_cosine(u, h, bottom):
    """renderer.py:522-542 (the multi-scattering LUT's sun axis)."""
    hc = -_shell_rho(h, bottom) / max(bottom + max(h, 0.0), 1.0)
    if u < 0.5:
        c = hc - (hc + 1.0) * (1.0 - 2.0 * u) ** 2
    else:
        c = hc + (1.0 - hc) * (2.0 * u - 1.0) ** 2
    return min(max(c, -1.0), 1.0)


def _sun_cosine_to_u(c, h, bottom):
    """renderer.py:544-568."""
    hc = -_shell_rho(h, bottom) / max(bottom + max(h, 0.0), 1.0)
    if c < hc:
        n = min(max((hc - c) / max(hc + 1.0, 1.0e-6), 0.0), 1.0)
        return 0.5 * (1.0 - math.sqrt(n))
    n = min(max((c - hc) / max(1.0 - hc, 1.0e-6), 0.0), 1.0)
    return 0.5 + 0.5 * math.sqrt(n)


def pin_multi_scattering(ref) -> float:
    """_build_multi_scattering (renderer.py:1066-1206) at three texels with
    the sun 22-70 deg above the horizon, against the independent float64
    transfer estimate; the scattering fraction f is 0.07-0.32 there, so the
    closure moves the texels by 7-47 %."""
    cfg = ms_config()
    ms = ref.build_multi_scattering(cfg, R_EARTH, E, ALPHA, tlut_for(cfg))
    worst = 0.0
    for (x, y), (psi, _, f) in _ms_expected().items():
        assert f.max() > 0.05  # the closure matters at these texels
        worst = max(worst, float(rel(ms[x, y], psi).max()) / MS_TOL)
    return worst


# --------------------------------------------------------------------------
# aerial perspective: each slice is single scattering to its depth
# --------------------------------------------------------------------------

AP_TOL = 2.0e-3
AP_FRAME = (160, 90)
AP_CAMERA_ALT = 1_000.0


def ap_config():
    return converged_lut_config(aerial_lut_width=4, aerial_lut_height=4, aerial_lut_depth=8,
                                aerial_steps_per_slice=16)


def _ap_geometry():
    scene = _scene()
    right, up, forward = scene.view_basis_local(35.0, 0.0)
    sun = scene.sun_from_angles(30.0, 1.1)
    return right, up, forward, sun, math.tan(math.radians(30.0))


def _ap_rays(cfg, right, up, forward, tan_half):
    w, h = AP_FRAME
    aspect = w / h
    rays = {}
    for x in range(cfg.aerial_lut_width):
        for y in range(cfg.aerial_lut_height):
            u = (x + 0.5) / cfg.aerial_lut_width
            v = (y + 0.5) / cfg.aerial_lut_height
            d = right * ((u * 2 - 1) * aspect) + up * (v * 2 - 1) + forward / tan_half
            rays[(x, y)] = d / np.linalg.norm(d)
    return rays


@functools.lru_cache(maxsize=None)
def _ap_expected():
    """Cumulative single scattering at the slice depths i / (D - 1) of each
    froxel ray's prefix segment [0, ground or top]: one uniform-midpoint path
    per ray with a step count divisible by D - 1."""
    cfg = ap_config()
    right, up, forward, sun, tan_half = _ap_geometry()
    medium = Medium(cfg)
    model = AtmosphereModel(R_EARTH, cfg)
    origin = np.array([0.0, R_EARTH + AP_CAMERA_ALT, 0.0])
    depth = cfg.aerial_lut_depth
    per_slice = 300
    steps = per_slice * (depth - 1)
    out = {}
    for key, ray in _ap_rays(cfg, right, up, forward, tan_half).items():
        interval = model.atmosphere_interval(origin, ray)
        start, end = interval.start_m, interval.end_m
        dt = (end - start) / steps
        t = start + (np.arange(steps) + 0.5) * dt
        points = origin + t[:, None] * ray
        hh = np.linalg.norm(points, axis=1) - R_EARTH
        ext = medium.extinction(hh)
        view_t = np.exp(-(np.cumsum(ext, axis=0) - 0.5 * ext) * dt)
        c = float(np.clip(ray @ sun, -1, 1))
        r, m, _ = medium.densities(hh)
        from fp64_paths import mie_phase, rayleigh_phase

        phase = (r[:, None] * medium.beta_r * rayleigh_phase(c)
                 + m[:, None] * medium.beta_ms * mie_phase(c, cfg.mie_phase_g))
        st = sun_transmittance(points, sun, R_EARTH, medium, 200)
        contrib = view_t * st * phase * np.asarray(E) * dt
        cum = np.cumsum(contrib, axis=0)
        slices = np.zeros((depth, 3))
        for i in range(1, depth):
            slices[i] = cum[i * per_slice - 1]
        out[key] = slices
    return out


def pin_aerial_slices(ref) -> float:
    """_build_aerial_perspective (renderer.py:1305-1420) with the
    multi-scattering LUT zeroed: slice 0 is exactly 0, and slice i holds the
    single scattering over the first i / (D - 1) of the ray's prefix segment
    (ground hits for the lower rays, the top of the shell for the upper)."""
    cfg = ap_config()
    right, up, forward, sun, tan_half = _ap_geometry()
    zero_ms = np.zeros((cfg.multi_scattering_lut_width, cfg.multi_scattering_lut_height, 3))
    ap = ref.build_aerial_perspective(cfg, AP_FRAME[0], AP_FRAME[1], R_EARTH, AP_CAMERA_ALT, sun,
                                      E, ALPHA, right, up, forward, tan_half, tlut_for(cfg),
                                      zero_ms)
    worst = 0.0
    for (x, y), want in _ap_expected().items():
        got = ap[x, y]
        if np.any(got[0] != 0.0):
            return math.inf
        worst = max(worst, float(rel(got[1:], want[1:]).max()) / AP_TOL)
    return worst


# --------------------------------------------------------------------------
# camera transmittance
# --------------------------------------------------------------------------

CAMERA_T_TOL = 2.0e-3


def pin_camera_transmittance(ref) -> float:
    """_camera_transmittance (renderer.py:1789-1825): 0 for rays that meet
    the ground, 1 for rays that miss the shell, otherwise the transmittance
    from the atmosphere entry to the top -- from inside the shell the
    camera's own, from above it the whole chord through the shell."""
    cfg = converged_lut_config()
    tlut = tlut_for(cfg)
    p = params_of(cfg)
    model = AtmosphereModel(R_EARTH, cfg)
    worst = 0.0
    out = np.zeros(3, dtype=ref.dtype)
    cases = []
    for h, c in ((1_000.0, 1.0), (1_000.0, 0.5), (1_000.0, 0.1), (1_000.0, 0.03),
                 (20_000.0, -0.05), (150_000.0, -0.3), (150_000.0, 0.4), (1_000.0, -0.5)):
        cases.append((h, c))
    for h, c in cases:
        ray = np.array([math.sqrt(1 - c * c), c, 0.0])
        ref.lib.oracle_fn_camera_transmittance(ctypes.byref(p), ref.ptr(tlut), ref.r(h),
                                               ref.ptr(ray.astype(ref.dtype)), ref.r(R_EARTH),
                                               ref.ptr(out))
        origin = np.array([0.0, R_EARTH + h, 0.0])
        interval = model.atmosphere_interval(origin, ray)
        if interval is None:
            want = np.ones(3)
        elif interval.ends_at_ground:
            want = np.zeros(3)
        else:
            want = model.transmittance(origin, ray, steps=50_000)
        if not want.any() or (want == 1.0).all():
            err = float(np.abs(out - want).max()) / 1e-12
        else:
            err = float(rel(out, want).max()) / CAMERA_T_TOL
        worst = max(worst, err)
    return worst


# --------------------------------------------------------------------------
# sky-view sampler (diagnostic view 1 of the composite)
# --------------------------------------------------------------------------

EXACT_TOL = 1.0e-10


def _frame_rays(w, h, basis, tan_half):
    rays_view = _scene().pixel_rays_view(w, h, tan_half)
    right, up, forward = (np.asarray(v, dtype=np.float64) for v in basis)
    rays = (rays_view[..., 0:1] * right + rays_view[..., 1:2] * up
            + rays_view[..., 2:3] * forward)
    return rays / np.linalg.norm(rays, axis=-1, keepdims=True), rays_view


def _sky_zenith_to_v(zenith, h, bottom):
    """renderer.py:482-490 and 506-520 (the sky-view LUT's zenith axis)."""
    a = max(h, 0.0)
    horizon = math.pi - math.acos(min(max(_shell_rho(a, bottom) / max(bottom + a, 1.0), 0.0), 1.0))
    if zenith < horizon:
        n = min(max(zenith / max(horizon, 1.0e-6), 0.0), 1.0)
        return 0.5 * (1.0 - math.sqrt(max(1.0 - n, 0.0)))
    n = min(max((zenith - horizon) / max(math.pi - horizon, 1.0e-6), 0.0), 1.0)
    return 0.5 + 0.5 * math.sqrt(n)


def _composite(ref, cfg, w, h, sid, pos, scene_hdr, space_hdr, cam_alt, sky_alt, sun, basis,
               tan_half, view, tlut, mslut, svlut, aplut, contrast=(4.0, 12.0), disk=(0, 0, 0)):
    right, up, forward = basis
    hdr, refined = ref.composite(cfg, scene_hdr, space_hdr, sid, pos, R_EARTH, cam_alt, sky_alt,
                                 sun, E, disk, ALPHA, right, up, forward, tan_half, contrast[0],
                                 contrast[1], view, tlut, mslut, svlut, aplut)
    return hdr


def _empty_frame(cfg, w, h):
    sid = -np.ones((w, h), dtype=np.int32)
    z3 = np.zeros((w, h, 3))
    aplut = np.zeros((cfg.aerial_lut_width, cfg.aerial_lut_height, cfg.aerial_lut_depth, 3))
    return sid, z3, aplut


def pin_sky_view_sampler(ref) -> float:
    """_sample_sky_view (renderer.py:1475-1525) through the SKY_VIEW view:
    with a sky-view LUT linear in its texel indices the bilinear fetch
    returns a + b X + c Y exactly, where X = clamp(u W - 1/2, 0, W - 1) with
    u = relative azimuth / pi, and Y likewise with v from the zenith (the
    horizon split at the snapshot altitude, not the camera's)."""
    scene = _scene()
    cfg = config(sky_view_lut_width=24, sky_view_lut_height=20)
    w, h = 48, 32
    basis = scene.view_basis_local(40.0, 10.0)
    tan_half = 1.0
    sun = scene.sun_from_angles(60.0, 2.0)
    rays, _ = _frame_rays(w, h, basis, tan_half)
    svlut, a, b = linear_lut((cfg.sky_view_lut_width, cfg.sky_view_lut_height, 3), seed=1)
    sid, z3, aplut = _empty_frame(cfg, w, h)
    tlut = np.ones((cfg.transmittance_lut_width, cfg.transmittance_lut_height, 3))
    mslut = np.zeros((cfg.multi_scattering_lut_width, cfg.multi_scattering_lut_height, 3))
    sky_alt = 30_000.0
    got = _composite(ref, cfg, w, h, sid, z3, z3, z3, 2_000.0, sky_alt, sun, basis, tan_half, 1,
                     tlut, mslut, svlut, aplut)
    want = np.empty_like(got)
    sw, sh = cfg.sky_view_lut_width, cfg.sky_view_lut_height
    sl = math.hypot(sun[0], sun[2])
    for i in range(w):
        for j in range(h):
            r = rays[i, j]
            vl = math.hypot(r[0], r[2])
            az = 0.0
            if vl > 1e-6 and sl > 1e-6:
                az = math.acos(min(max((r[0] * sun[0] + r[2] * sun[2]) / (vl * sl), -1.0), 1.0))
            u = az / math.pi
            v = _sky_zenith_to_v(math.acos(min(max(r[1], -1.0), 1.0)), sky_alt, R_EARTH)
            x = min(max(u * sw - 0.5, 0.0), sw - 1.0)
            y = min(max(v * sh - 0.5, 0.0), sh - 1.0)
            want[i, j] = a + b[0] * x + b[1] * y
    return float(rel(got, want).max()) / EXACT_TOL


# --------------------------------------------------------------------------
# the composite's sky branch: star fade and sun-disk coverage
# --------------------------------------------------------------------------


def _smoothstep01(x):
    x = np.clip(x, 0.0, 1.0)
    return x * x * (3.0 - 2.0 * x)


def pin_sky_composite(ref) -> float:
    """The COMPOSITE view's sky pixels (renderer.py:1894-1940) in a clear
    medium (every transmittance 1) under a constant sky-view LUT: hdr =
    sky + star * smoothstep((contrast - c0) / (c1 - c0)) + disk *
    smoothstep((alpha + f - theta) / (2 f)), with contrast the luminance
    ratio of star to sky, theta the pixel ray's angle to the sun and f =
    atan(tan_half_fov sqrt 2 / H) the one-pixel filter.  A narrow field of
    view centred on the sun resolves the disk over ~10 pixels."""
    scene = _scene()
    clear = dict(rayleigh_scattering_per_m=(0.0,) * 3, mie_scattering_per_m=(0.0,) * 3,
                 mie_extinction_per_m=(0.0,) * 3, absorption_extinction_per_m=(0.0,) * 3)
    cfg = config(**clear)
    w, h = 40, 32
    basis = scene.view_basis_local(20.0, 55.0)
    tan_half = 0.02
    forward = np.asarray(basis[2])
    # sun a little off the view centre, so the disk spans pixels unevenly
    sun = forward + 0.0011 * np.asarray(basis[0]) + 0.0004 * np.asarray(basis[1])
    sun = sun / np.linalg.norm(sun)
    rays, _ = _frame_rays(w, h, basis, tan_half)
    sky = np.array([0.3, 0.5, 0.9])
    svlut = np.ascontiguousarray(np.broadcast_to(sky, (cfg.sky_view_lut_width,
                                                       cfg.sky_view_lut_height, 3)))
    tlut = np.ones((cfg.transmittance_lut_width, cfg.transmittance_lut_height, 3))
    mslut = np.zeros((cfg.multi_scattering_lut_width, cfg.multi_scattering_lut_height, 3))
    sid, z3, aplut = _empty_frame(cfg, w, h)
    c0, c1 = 4.0, 12.0
    g = np.random.default_rng(5)
    contrast = g.uniform(0.0, 2.0 * c1, (w, h))
    tint = g.uniform(0.5, 1.5, (w, h, 3))
    sky_lum = float(sky @ LUM)
    space = tint * (contrast * sky_lum / (tint @ LUM))[..., None]
    disk = np.array([1.0e4, 0.9e4, 0.8e4])
    got = _composite(ref, cfg, w, h, sid, z3, z3, space, 1_000.0, 1_000.0, sun, basis, tan_half,
                     0, tlut, mslut, svlut, aplut, contrast=(c0, c1), disk=disk)
    star_vis = _smoothstep01(((space @ LUM) / max(sky_lum, 1e-5) - c0) / max(c1 - c0, 1e-5))
    theta = np.arccos(np.clip(rays @ sun, -1.0, 1.0))
    f = math.atan2(tan_half * 1.41421356 / h, 1.0)
    coverage = _smoothstep01((ALPHA + f - theta) / max(2.0 * f, 1e-8))
    assert 5 <= int(((coverage > 0.0) & (coverage < 1.0)).sum())  # the filter band is resolved
    assert int((coverage == 1.0).sum()) >= 3
    want = sky + space * star_vis[..., None] + disk * coverage[..., None]
    worst = float(rel(got, want).max()) / 1.0e-9
    # In a real medium every term but the sky-view sample is seen through
    # the camera transmittance, which the CAMERA_TRANSMITTANCE view (2)
    # shows, and the sky term is the SKY_VIEW view (1).
    cfg = converged_lut_config()
    tlut = tlut_for(cfg)
    basis = scene.view_basis_local(20.0, 3.0)
    sun = scene.sun_from_angles(89.0, 0.35)
    tan_half = 0.05
    forward = np.asarray(basis[2])
    sun = forward + 0.002 * np.asarray(basis[0])
    sun = sun / np.linalg.norm(sun)
    svlut, _, _ = linear_lut((cfg.sky_view_lut_width, cfg.sky_view_lut_height, 3), seed=6)
    svlut = svlut * 1.0e-2
    scene_hdr = np.ascontiguousarray(g.uniform(0.0, 1.0e-3, (w, h, 3)))
    luts = (tlut, mslut, svlut, aplut)
    args = (ref, cfg, w, h, sid, z3, scene_hdr, z3, 1_000.0, 1_000.0, sun, basis, tan_half)
    sky_v = _composite(*args, 1, *luts)
    trans = _composite(*args, 2, *luts)
    assert 0.0 < trans.min() and trans.max() < 0.95  # the medium attenuates every term
    # star radiance whose attenuated luminance spans contrasts 0 .. 2 c1
    space = tint * (contrast * (sky_v @ LUM) / ((tint * trans) @ LUM))[..., None]
    args = (ref, cfg, w, h, sid, z3, scene_hdr, space, 1_000.0, 1_000.0, sun, basis, tan_half)
    got = _composite(*args, 0, *luts, contrast=(c0, c1), disk=disk)
    star = space * trans
    star_vis = _smoothstep01(((star @ LUM) / np.maximum(sky_v @ LUM, 1e-5) - c0)
                             / max(c1 - c0, 1e-5))
    rays, _ = _frame_rays(w, h, basis, tan_half)
    theta = np.arccos(np.clip(rays @ sun, -1.0, 1.0))
    f = math.atan2(tan_half * 1.41421356 / h, 1.0)
    coverage = _smoothstep01((ALPHA + f - theta) / max(2.0 * f, 1e-8))
    assert int((coverage > 0.0).sum()) >= 3
    want = (sky_v + scene_hdr * trans + star * star_vis[..., None]
            + disk * trans * coverage[..., None])
    return max(worst, float(rel(got, want).max()) / 1.0e-9)


# --------------------------------------------------------------------------
# terminator refinement weight and solar shadow roots
# --------------------------------------------------------------------------


def pin_terminator_weight(ref) -> float:
    """_terminator_refinement_weight (renderer.py:901-966): with the sun
    high over the interval the penumbra term is 0 and the weight is the
    smoothstep of clamp(2 - |midpoint - root| / width): 1 with a root
    within one width of the midpoint, 1/2 at 1.5 widths, 0 from 2 widths
    and without valid roots.  With a sun whose disk centre sits on the
    midpoint's horizon the visibility is 1/2 and the weight 4 v (1 - v) = 1;
    with the centre a quarter disk radius above it, 4 v (1 - v) with v the
    circular-segment fraction."""
    fn = _restype(ref, "oracle_fn_terminator_weight")
    cfg = config()
    p = params_of(cfg)
    h0 = 2_000.0
    origin = np.array([0.0, R_EARTH + h0, 0.0])
    ray = np.array([0.6, 0.0, 0.8])
    start, end = 10_000.0, 30_000.0
    mid, width = 20_000.0, 20_000.0

    def weight(sun, roots, valid):
        return fn(ctypes.byref(p), ref.ptr(origin.astype(ref.dtype)), ref.r(h0),
                  ref.ptr(ray.astype(ref.dtype)), ref.r(start), ref.r(end),
                  ref.ptr(np.asarray(sun, dtype=ref.dtype)), ref.r(R_EARTH), ref.r(ALPHA),
                  ref.ptr(np.asarray(roots, dtype=ref.dtype)),
                  np.asarray(valid, dtype=np.int32).ctypes.data_as(ctypes.c_void_p))

    high_sun = np.array([0.0, 1.0, 0.0])
    checks = []
    for k, want in ((0.0, 1.0), (0.5, 1.0), (1.0, 1.0), (1.5, 0.5), (1.75, 0.15625),
                    (2.0, 0.0), (3.0, 0.0)):
        checks.append((weight(high_sun, (mid + k * width, -1e9), (1, 0)), want))
        checks.append((weight(high_sun, (-1e9, mid - k * width), (0, 1)), want))
    # invalid roots never count; the nearer valid root wins
    checks.append((weight(high_sun, (mid, mid), (0, 0)), 0.0))
    checks.append((weight(high_sun, (mid + 3 * width, mid + 1.5 * width), (1, 1)), 0.5))
    # penumbra at the midpoint's horizon
    point = origin + ray * mid
    radial = point / np.linalg.norm(point)
    hm = float(np.linalg.norm(point) - R_EARTH)
    horizon_zenith = math.acos(-_shell_rho(hm, R_EARTH) / (R_EARTH + hm))
    tangent = np.cross(np.cross(radial, ray), radial)
    tangent /= np.linalg.norm(tangent)
    for offset, want_v in ((0.0, 0.5), (0.25, None)):
        zen = horizon_zenith - offset * ALPHA
        sun = radial * math.cos(zen) + tangent * math.sin(zen)
        v = want_v if want_v is not None else (
            (math.acos(-offset) + offset * math.sqrt(1 - offset * offset)) / math.pi)
        checks.append((weight(sun, (0.0, 0.0), (0, 0)), 4.0 * v * (1.0 - v)))
    return max(abs(got - want) for got, want in checks) / 1.0e-9


def pin_shadow_roots(ref) -> float:
    """_solar_shadow_roots (renderer.py:686-726): the view ray against the
    cylinder of the planet's radius along the sun axis, a root valid where
    the point lies behind the terminator plane (p . s < 0).  Constructed
    cases: a ray crossing the axis at right angles meets the cylinder at the
    axis distance -+ R; a ray parallel to the axis has no roots; the roots of
    an oblique ray are those of its projection scaled by the projected
    length."""
    out = np.zeros(4, dtype=ref.dtype)
    worst = 0.0
    R = R_EARTH

    def roots(origin, ray, sun):
        ref.lib.oracle_fn_solar_shadow_roots(
            ref.ptr(np.asarray(origin, dtype=ref.dtype)), ref.ptr(np.asarray(ray, dtype=ref.dtype)),
            ref.ptr(np.asarray(sun, dtype=ref.dtype)), ref.r(R), ref.ptr(out))
        return out.copy()

    s = np.array([0.0, 0.0, 1.0])
    # straight down through the axis, 3000 km behind the planet: roots h, 2R + h
    h = 400_000.0
    got = roots([0.0, R + h, -3.0e6], [0.0, -1.0, 0.0], s)
    worst = max(worst, float(np.abs(got - [h, 2 * R + h, 1, 1]).max()) / 1e-3)
    # the same ray on the sunlit side: same roots, neither valid
    got = roots([0.0, R + h, 3.0e6], [0.0, -1.0, 0.0], s)
    worst = max(worst, float(np.abs(got - [h, 2 * R + h, 0, 0]).max()) / 1e-3)
    # parallel to the axis: no roots
    got = roots([0.0, R / 2, -3.0e6], [0.0, 0.0, 1.0], s)
    worst = max(worst, float(np.abs(got - [0, 0, 0, 0]).max()) / 1e-12)
    # oblique: direction (0, -sin b, cos b) from y = R + h: the projection onto
    # the plane normal to the axis moves at sin b, so roots are h / sin b and
    # (2R + h) / sin b; the second lies in front of the terminator plane
    b = math.radians(25.0)
    z0 = -3.0e6
    got = roots([0.0, R + h, z0], [0.0, -math.sin(b), math.cos(b)], s)
    t0, t1 = h / math.sin(b), (2 * R + h) / math.sin(b)
    valid = [float(z0 + t0 * math.cos(b) < 0.0), float(z0 + t1 * math.cos(b) < 0.0)]
    assert valid == [1.0, 0.0]
    worst = max(worst, float(np.abs(got[:2] - [t0, t1]).max()) / 1e-3,
                float(np.abs(got[2:] - valid).max()) / 1e-12)
    return worst


# --------------------------------------------------------------------------
# terminator refinement: the refined interval converges across the shadow edge
# --------------------------------------------------------------------------


def refinement_case():
    """A camera at 12 km with the sun 3 deg below its horizon, looking away
    from the sun and down at a 220 km surface point: the path enters the
    planet's shadow cylinder."""
    h = 12_000.0
    zen = math.radians(93.0)
    sun = np.array([math.sin(zen), math.cos(zen), 0.0])
    c = -0.04
    ray = np.array([-math.sqrt(1 - c * c), c, 0.0])
    return h, ray, sun, 220_000.0


def pin_terminator_refinement(ref) -> float:
    """_integrate_adaptive_scattering_interval (renderer.py:969-1032) with
    the terminator inside the path: the 12-interval integral with 8-step
    refinement is closer to the converged integral (4000 intervals) than the
    same integral without refinement, and the converged integral matches an
    independent single-scattering quadrature with a binary shadow to 2 %
    (the finite sun disk's penumbra is the difference).  Returns the
    refined error over half the unrefined error."""
    cfg = converged_lut_config()
    tlut = tlut_for(cfg)
    zero_ms = np.zeros((cfg.multi_scattering_lut_width, cfg.multi_scattering_lut_height, 3))
    h, ray, sun, d = refinement_case()

    def inscatter(c):
        L, _ = ref.surface_path(c, np.array([h]), ray[None], np.array([d]), sun[None], E, R_EARTH,
                                ALPHA, tlut, zero_ms)
        return L[0]

    base = inscatter(cfg)
    coarse = inscatter(cfg.with_(aerial_terminator_substeps=1))
    conv = inscatter(cfg.with_(aerial_horizon_raymarch_steps=4000))
    model, _, _ = path_integral(np.array([0.0, R_EARTH + h, 0.0]), ray, 0.0, d, sun, E, R_EARTH,
                                Medium(cfg), steps=20_000, sun_steps=300)
    if float(rel(conv, model).max()) > 2.0e-2:
        return math.inf
    err_refined = float(np.abs(base - conv).max())
    err_coarse = float(np.abs(coarse - conv).max())
    return err_refined / (0.5 * err_coarse)


# --------------------------------------------------------------------------
# surface lighting
# --------------------------------------------------------------------------

PREPARE_T_TOL = 2.0e-3


def pin_prepare_surface_lighting(ref) -> float:
    """_prepare_surface_lighting (renderer.py:1423-1472): sky pixels get
    T = 1 and L = 0; a surface pixel's endpoint is camera + basis .
    position_view, its sun transmittance is the transmittance from that
    endpoint toward a high sun, and its sky radiance the bilinear MS sample
    at (u(sun cosine at the endpoint), sqrt(endpoint altitude / top)) --
    exact for an MS LUT linear in its texel indices."""
    scene = _scene()
    cfg = converged_lut_config()
    tlut = tlut_for(cfg)
    mslut, a, b = linear_lut((cfg.multi_scattering_lut_width, cfg.multi_scattering_lut_height, 3),
                             seed=3)
    w, hh = 12, 8
    cam_alt = 3_000.0
    basis = scene.view_basis_local(70.0, -15.0)
    right, up, forward = (np.asarray(v) for v in basis)
    sun = scene.sun_from_angles(25.0, 0.7)
    g = np.random.default_rng(11)
    rays_view = scene.pixel_rays_view(w, hh, math.tan(math.radians(30.0)))
    dist = np.exp(g.uniform(math.log(50.0), math.log(40_000.0), (w, hh)))
    pos = rays_view * dist[..., None]
    sid = np.where(g.random((w, hh)) < 0.25, -1, 0).astype(np.int32)
    sun_t, sky_l = ref.prepare_surface_lighting(cfg, pos, sid, R_EARTH, cam_alt, sun, ALPHA, right,
                                                up, forward, tlut, mslut)
    medium = Medium(cfg)
    camera = np.array([0.0, R_EARTH + cam_alt, 0.0])
    worst = 0.0
    for i in range(w):
        for j in range(hh):
            if sid[i, j] < 0:
                if (sun_t[i, j] != 1.0).any() or sky_l[i, j].any():
                    return math.inf
                continue
            local = right * pos[i, j, 0] + up * pos[i, j, 1] + forward * pos[i, j, 2]
            end = camera + local
            alt = float(np.linalg.norm(end) - R_EARTH)
            want_t = sun_transmittance(end[None], sun, R_EARTH, medium, 2000)[0]
            worst = max(worst, float(rel(sun_t[i, j], want_t).max()) / PREPARE_T_TOL)
            mu = float(end @ sun / np.linalg.norm(end))
            u = _sun_cosine_to_u(mu, alt, R_EARTH)
            v = math.sqrt(min(max(alt / cfg.top_altitude_m, 0.0), 1.0))
            x = min(max(u * cfg.multi_scattering_lut_width - 0.5, 0.0),
                    cfg.multi_scattering_lut_width - 1.0)
            y = min(max(v * cfg.multi_scattering_lut_height - 0.5, 0.0),
                    cfg.multi_scattering_lut_height - 1.0)
            want_l = a + b[0] * x + b[1] * y
            worst = max(worst, float(rel(sky_l[i, j], want_l).max()) / 1.0e-9)
    return worst


# --------------------------------------------------------------------------
# diagnostic views
# --------------------------------------------------------------------------

VIEW_T_TOL = 2.0e-3


def pin_diagnostic_views(ref) -> float:
    """The diagnostic branches of _composite (renderer.py:1941-2052):
    TRANSMITTANCE_LUT (3) and MULTI_SCATTERING_LUT (4) show texel
    [int(u X), int(v Y)] of their LUT; SURFACE_MASK (7) and
    AERIAL_BLEND_WEIGHT (10) are 1 on surface pixels and 0 on sky;
    AERIAL_SCATTERING (5) and AERIAL_DIRECT_ONLY (9) both equal the
    COMPOSITE view of a black scene on surface pixels and 0 on sky;
    AERIAL_TRANSMITTANCE (6) is the transmittance from the camera to the
    surface point; AERIAL_FROXEL_ONLY (8) is the trilinear froxel sample at
    the pixel's depth fraction of the ray's prefix segment -- exact for an
    aerial LUT linear in its texel indices."""
    scene = _scene()
    cfg = converged_lut_config(aerial_lut_width=6, aerial_lut_height=5, aerial_lut_depth=9,
                               multi_scattering_lut_width=10, multi_scattering_lut_height=9)
    w, hh = 16, 12
    cam_alt = 1_500.0
    tan_half = math.tan(math.radians(30.0))
    basis = scene.view_basis_local(10.0, -30.0)
    sun = scene.sun_from_angles(40.0, 0.3)
    rays, rays_view = _frame_rays(w, hh, basis, tan_half)
    model = AtmosphereModel(R_EARTH, cfg)
    origin = np.array([0.0, R_EARTH + cam_alt, 0.0])
    g = np.random.default_rng(21)
    sid = np.where(g.random((w, hh)) < 0.3, -1, 0).astype(np.int32)
    # surface points in front of the ground hit of each downward ray
    dist = np.zeros((w, hh))
    for i in range(w):
        for j in range(hh):
            itv = model.atmosphere_interval(origin, rays[i, j])
            assert itv is not None and itv.ends_at_ground
            dist[i, j] = itv.end_m * g.uniform(0.2, 0.95)
    pos = rays_view * dist[..., None]
    tlut_r = np.ascontiguousarray(g.uniform(0, 1, (cfg.transmittance_lut_width,
                                                    cfg.transmittance_lut_height, 3)))
    mslut_r = np.ascontiguousarray(g.uniform(0, 1, (cfg.multi_scattering_lut_width,
                                                     cfg.multi_scattering_lut_height, 3)))
    svlut = np.zeros((cfg.sky_view_lut_width, cfg.sky_view_lut_height, 3))
    aplut, a, b = linear_lut((cfg.aerial_lut_width, cfg.aerial_lut_height, cfg.aerial_lut_depth,
                              3), seed=4)
    z3 = np.zeros((w, hh, 3))

    def view(v, tlut, mslut, scene_hdr=z3):
        return _composite(ref, cfg, w, hh, sid, pos, scene_hdr, z3, cam_alt, cam_alt, sun, basis,
                          tan_half, v, tlut, mslut, svlut, aplut)

    worst = 0.0
    su = (np.arange(w) + 0.5) / w
    sv = (np.arange(hh) + 0.5) / hh
    for v, lut in ((3, tlut_r), (4, mslut_r)):
        got = view(v, tlut_r, mslut_r)
        lx = np.minimum((su * lut.shape[0]).astype(int), lut.shape[0] - 1)
        ly = np.minimum((sv * lut.shape[1]).astype(int), lut.shape[1] - 1)
        want = lut[lx[:, None], ly[None, :]]
        if (got != want).any():
            return math.inf
    mask = (sid >= 0).astype(np.float64)[..., None] * np.ones(3)
    for v in (7, 10):
        got = view(v, tlut_r, mslut_r)
        worst = max(worst, 0.0 if (got == mask).all() else math.inf)
    # physical LUTs for the remaining views
    tlut = tlut_for(cfg)
    mslut = np.zeros((cfg.multi_scattering_lut_width, cfg.multi_scattering_lut_height, 3))
    comp = view(0, tlut, mslut)
    for v in (5, 9):
        got = view(v, tlut, mslut)
        want = np.where(sid[..., None] >= 0, comp, 0.0)
        worst = max(worst, float(rel(got, want).max()) / 1e-12)
    got = view(6, tlut, mslut)
    for i in range(w):
        for j in range(hh):
            if sid[i, j] < 0:
                worst = max(worst, 0.0 if not got[i, j].any() else math.inf)
                continue
            want = model.transmittance(origin, rays[i, j], maximum_distance_m=dist[i, j],
                                       steps=20_000)
            worst = max(worst, float(rel(got[i, j], want).max()) / VIEW_T_TOL)
    got = view(8, tlut, mslut)
    aw, ah, ad = cfg.aerial_lut_width, cfg.aerial_lut_height, cfg.aerial_lut_depth
    for i in range(w):
        for j in range(hh):
            if sid[i, j] < 0:
                worst = max(worst, 0.0 if not got[i, j].any() else math.inf)
                continue
            itv = model.atmosphere_interval(origin, rays[i, j])
            frac = min(max((dist[i, j] - itv.start_m) / max(itv.end_m - itv.start_m, 1.0), 0.0),
                       1.0)
            x = min(max(su[i] * aw - 0.5, 0.0), aw - 1.0)
            y = min(max(sv[j] * ah - 0.5, 0.0), ah - 1.0)
            z = frac * (ad - 1)
            want = a + b[0] * x + b[1] * y + b[2] * z
            worst = max(worst, float(rel(got[i, j], want).max()) / 1e-9)
    return worst


ALL_PINS = {
    "multi_scattering": pin_multi_scattering,
    "aerial_slices": pin_aerial_slices,
    "camera_transmittance": pin_camera_transmittance,
    "sky_view_sampler": pin_sky_view_sampler,
    "sky_composite": pin_sky_composite,
    "terminator_weight": pin_terminator_weight,
    "shadow_roots": pin_shadow_roots,
    "terminator_refinement": pin_terminator_refinement,
    "prepare_surface_lighting": pin_prepare_surface_lighting,
    "diagnostic_views": pin_diagnostic_views,
}
