"""The display post-process on the CPU: `PostprocessConfig` and its C-ABI
validation, and the pinning of the post-process oracle (oracle/post_cpu.c)
against what the mathematics fixes.

The reference (src/planet_renderer/postprocess.py) ships no numeric golden
values, so each stage is pinned by closed forms and by an independent
float64 NumPy formulation:
  - the sRGB power, exhaustively against long double;
  - the ACES fit by its slope at 0 (b/e), its saturation point (root of
    (a-c)v^2 + (b-d)v - e) and its inverse, through the standard sRGB
    decode (IEC 61966-2-1);
  - the bright pass by the three regimes of the quadratic soft knee and the
    clamp of the peak;
  - the blur by the P-fold self-convolution of its taps (impulse) and by
    edge-replicated padding (random field);
  - the bilinear bloom lookup by exact reproduction of a linear ramp;
  - the whole process against a float64 NumPy formulation built from
    slicing, padding and gathers instead of the oracle's index loops.
"""

from __future__ import annotations

import ctypes
import math

import numpy as np
import pytest

from oracle.post_cpu import PostOracle, bright_pass_constants

A, B, C, D, E = 2.51, 0.03, 2.43, 0.59, 0.14
TAPS = np.array([0.06136, 0.24477, 0.38774, 0.24477, 0.06136])


@pytest.fixture(scope="module")
def post(pkg):
    return pkg.PostprocessConfig


def oracle(post, width=16, height=12, **changes):
    return PostOracle(width, height, post(**changes))


# --------------------------------------------------------------- config


def test_config_defaults_match_reference(post):
    # postprocess.py:14-19
    c = post()
    assert (c.exposure_ev, c.bloom_threshold, c.bloom_knee, c.bloom_clamp, c.bloom_strength,
            c.bloom_passes) == (0.0, 1.0, 0.5, 8.0, 0.08, 4)


@pytest.mark.parametrize("changes, message", [
    ({"exposure_ev": 21.0}, "postprocess.exposure_ev must be finite and in [-20, 20]"),
    ({"exposure_ev": math.nan}, "postprocess.exposure_ev must be finite and in [-20, 20]"),
    ({"bloom_threshold": -1.0}, "postprocess.bloom_threshold must be finite and non-negative"),
    ({"bloom_knee": math.inf}, "postprocess.bloom_knee must be finite and non-negative"),
    ({"bloom_clamp": -0.5}, "postprocess.bloom_clamp must be finite and non-negative"),
    ({"bloom_strength": math.nan}, "postprocess.bloom_strength must be finite and non-negative"),
    ({"bloom_passes": 9}, "postprocess.bloom_passes must be in 0..8"),
    ({"bloom_passes": -1}, "postprocess.bloom_passes must be in 0..8"),
])
def test_config_validation_messages(pkg, post, changes, message):
    with pytest.raises(ValueError) as err:
        post(**changes)
    assert str(err.value) == message
    # the same check in the C-ABI, on the raw struct
    capi = __import__(pkg.__name__ + "._capi", fromlist=["_capi"])
    fields = dict(exposure_ev=0.0, bloom_threshold=1.0, bloom_knee=0.5, bloom_clamp=8.0,
                  bloom_strength=0.08, bloom_passes=4)
    fields.update(changes)
    params = capi.AtmoPostParams(**fields)
    buf = ctypes.create_string_buffer(256)
    assert capi.library().atmo_post_validate_params(ctypes.byref(params), buf, 256) == \
        capi.ATMO_EINVAL
    assert buf.value.decode() == message
    handle = ctypes.c_void_p()
    assert capi.library().atmo_post_create(ctypes.byref(params), 8, 8, ctypes.byref(handle)) == \
        capi.ATMO_EINVAL
    assert capi.library().atmo_post_last_error(None).decode() == message


def test_config_from_mapping_and_params(pkg, post):
    c = post.from_mapping({"exposure_ev": "1.5", "bloom_passes": 2, "bloom_knee": 0.25})
    assert (c.exposure_ev, c.bloom_passes, c.bloom_knee, c.bloom_threshold) == (1.5, 2, 0.25, 1.0)
    p = c.to_params()
    assert (p.exposure_ev, p.bloom_threshold, p.bloom_knee, p.bloom_clamp, p.bloom_strength,
            p.bloom_passes) == (1.5, 1.0, 0.25, 8.0, 0.08, 2)
    capi = __import__(pkg.__name__ + "._capi", fromlist=["_capi"])
    assert capi.library().atmo_post_validate_params(ctypes.byref(p), None, 0) == capi.ATMO_OK


# -------------------------------------------------- display transform


def test_srgb_power_is_correctly_rounded(post):
    """Every float32 of the encode branch (0.0031308, 1]: the float64 power
    rounded once equals the long double power rounded once."""
    o = oracle(post)
    assert o.check_srgb_pow(np.float32(0.0031308), 1.0) == 0


def srgb_decode(s):
    """IEC 61966-2-1 sRGB decode (the inverse of the encode)."""
    s = np.asarray(s, np.float64)
    return np.where(s <= 0.04045, s / 12.92, ((s + 0.055) / 1.055) ** 2.4)


def test_aces_srgb_closed_forms(post):
    o = oracle(post)
    # zero, and the expansion at the origin inside sRGB's linear segment:
    # (b/e) v (1 + (a/b - d/e) v + O(v^2)), the O(v^2) coefficient below 400
    assert o.display_channel(np.float32(0.0)) == 0.0
    for v in (1e-6, 1e-5, 1e-4):
        got = float(o.display_channel(np.float32(v)))
        want = 12.92 * B / E * v * (1.0 + (A / B - D / E) * v)
        assert got == pytest.approx(want, rel=1e-6 + 400.0 * v * v)
    # saturation: the ratio reaches 1 at the positive root of
    # (a-c) v^2 + (b-d) v - e = 0; saturated pixels encode to the float32
    # 1.055 - 0.055, one unit below 1
    root = (-(B - D) + math.sqrt((B - D) ** 2 + 4 * (A - C) * E)) / (2 * (A - C))
    white = np.float32(1.055) - np.float32(0.055)
    assert white == np.nextafter(np.float32(1.0), np.float32(0.0))
    assert o.display_channel(np.float32(root * 1.001)) == white
    assert o.display_channel(np.float32(1e6)) == white
    # past ~1.2e19 both float32 products overflow: inf / inf is NaN, which
    # the clamp's maxNum maps to 0
    assert o.display_channel(np.float32(1e30)) == 0.0
    assert o.display_channel(np.float32(root * 0.98)) < white
    # negative and NaN radiance map to 0 (ti.max(value, 0), IEEE maxNum)
    assert o.display_channel(np.float32(-3.0)) == 0.0
    assert o.display_channel(np.float32(np.nan)) == 0.0
    # the inverse of the fit (the quadratic in v for a given mapped value m),
    # then the standard decode, recovers m
    m = np.concatenate([np.geomspace(1e-4, 0.999, 300), [0.0031308 * 0.99, 0.0031308 * 1.01]])
    qa, qb, qc = A - m * C, B - m * D, -m * E
    v = (-qb + np.sqrt(qb * qb - 4 * qa * qc)) / (2 * qa)
    got = o.display_channel(v.astype(np.float32))
    np.testing.assert_allclose(srgb_decode(got), m, rtol=2e-6, atol=1e-9)
    # monotone over the whole range
    grid = np.geomspace(1e-7, 50.0, 4000).astype(np.float32)
    assert np.all(np.diff(o.display_channel(grid)) >= 0.0)


# ------------------------------------------------------------ bright pass


def extract_constant(post, value, **changes):
    o = oracle(post, 4, 4, **changes)
    hdr = np.full((4, 4, 3), value, np.float32)
    o.extract(hdr, 1.0)
    return o.bloom_a


def test_bright_pass_soft_knee_regimes(post):
    t, k = 1.0, 0.5
    # below t - k: nothing
    assert np.all(extract_constant(post, 0.49) == 0.0)
    # above t + k: value * (b - t) / b, i.e. b - t per grey channel
    np.testing.assert_allclose(extract_constant(post, 3.0), 2.0, rtol=1e-6)
    # inside the knee: (b - t + k)^2 / (4k) scaled by value / b
    for b in (0.6, 0.9, 1.0, 1.3, 1.49):
        want = (b - t + k) ** 2 / (4 * k)
        np.testing.assert_allclose(extract_constant(post, b), want, rtol=2e-6)
    # knee = 0: a hard threshold (the 1e-6 denominator is never reached)
    np.testing.assert_allclose(extract_constant(post, 1.75, bloom_knee=0.0), 0.75, rtol=1e-6)
    assert np.all(extract_constant(post, 0.999, bloom_knee=0.0) == 0.0)
    # the peak is capped at bloom_clamp, chromaticity kept
    o = oracle(post, 2, 2)
    hdr = np.zeros((2, 2, 3), np.float32)
    hdr[...] = (4.0e4, 2.0e4, 1.0e4)
    o.extract(hdr, 1.0)
    np.testing.assert_allclose(o.bloom_a[0, 0], (8.0, 4.0, 2.0), rtol=1e-6)


def test_bright_pass_constants_follow_python_rounding(post):
    k = bright_pass_constants(post(bloom_knee=0.1, bloom_threshold=0.3))
    assert k.two_knee == np.float32(0.2) and k.knee_denominator == np.float32(0.4)
    assert k.threshold == np.float32(0.3)
    assert bright_pass_constants(post(bloom_knee=0.0)).knee_denominator == np.float32(1e-6)


def test_extract_box_and_edge_clamp(post):
    # odd sizes: the last half-resolution column/row reuses the last pixel
    o = oracle(post, 5, 3, bloom_threshold=0.0, bloom_knee=0.0, bloom_clamp=1e30)
    assert (o.half_width, o.half_height) == (3, 2)
    rng = np.random.default_rng(7)
    hdr = rng.uniform(0.5, 2.0, (5, 3, 3)).astype(np.float32)
    o.extract(hdr, 2.0)
    # threshold 0, knee 0: contribution (b - 0) / b = 1, so the bright pass
    # passes the exposed 2x2 mean through
    pad = np.pad(hdr.astype(np.float64), ((0, 1), (0, 1), (0, 0)), mode="edge")
    mean = 0.25 * (pad[0::2, 0::2] + pad[1::2, 0::2] + pad[0::2, 1::2] + pad[1::2, 1::2])
    np.testing.assert_allclose(o.bloom_a, 2.0 * mean, rtol=1e-6)


# ------------------------------------------------------------------ blur


def test_blur_impulse_is_the_p_fold_tap_convolution(post):
    n = 41
    for passes in (1, 2, 4):
        o = oracle(post, 2 * n, 2 * n)
        o.bloom_a[...] = 0.0
        o.bloom_a[n // 2, n // 2] = (1.0, 2.0, 4.0)
        for _ in range(passes):
            o.blur_pass()
        kernel = np.array([1.0])
        for _ in range(passes):
            kernel = np.convolve(kernel, TAPS)
        want = np.zeros((n, n))
        r = len(kernel) // 2
        want[n // 2 - r:n // 2 + r + 1, n // 2 - r:n // 2 + r + 1] = np.outer(kernel, kernel)
        for c, scale in enumerate((1.0, 2.0, 4.0)):
            np.testing.assert_allclose(o.bloom_a[..., c], scale * want, rtol=0, atol=2e-7 * scale)


def test_blur_clamps_to_edge_each_pass(post):
    rng = np.random.default_rng(3)
    o = oracle(post, 14, 9)  # 7 x 5 half resolution, smaller than two tap widths
    field = rng.uniform(0.0, 1.0, (7, 5, 3))
    o.bloom_a[...] = field
    want = field.copy()
    for _ in range(3):
        o.blur_pass()
        padded = np.pad(want, ((2, 2), (0, 0), (0, 0)), mode="edge")
        want = sum(TAPS[i] * padded[i:i + 7] for i in range(5))
        padded = np.pad(want, ((0, 0), (2, 2), (0, 0)), mode="edge")
        want = sum(TAPS[i] * padded[:, i:i + 5] for i in range(5))
    np.testing.assert_allclose(o.bloom_a, want, rtol=3e-6, atol=1e-7)
    # the taps sum to 1: a constant field stays constant
    o.bloom_a[...] = 0.75
    o.blur_pass()
    np.testing.assert_allclose(o.bloom_a, 0.75, rtol=4e-7)


# --------------------------------------------------------------- resolve


def test_bloom_lookup_reproduces_a_linear_ramp(post):
    w, h = 40, 26
    o = oracle(post, w, h, bloom_strength=1.0)
    hx, hy = np.meshgrid(np.arange(o.half_width), np.arange(o.half_height), indexing="ij")
    ramp = 0.01 * hx + 0.02 * hy + 0.05
    o.bloom_a[...] = ramp[..., None].astype(np.float32)
    hdr = np.zeros((w, h, 3), np.float32)
    out = o.resolve(hdr, 1.0, True, False)[..., 0]
    px = (np.arange(w) + 0.5) / w * o.half_width - 0.5
    py = (np.arange(h) + 0.5) / h * o.half_height - 0.5
    # inside: bilinear interpolation of a linear function is exact
    want = 0.01 * px[:, None] + 0.02 * py[None, :] + 0.05
    np.testing.assert_allclose(out[1:-1, 1:-1], want[1:-1, 1:-1], rtol=2e-6)
    # the first pixel centre lies at -0.25 texels: the lower texel is clamped
    # to 0 but the weight stays 0.75 (postprocess.py:145-149)
    assert out[0, 5] == pytest.approx(0.01 * 0.75 + 0.02 * py[5] + 0.05, rel=2e-6)


def test_raw_views_bypass_exposure_and_clamp(post):
    o = oracle(post, 6, 4)
    rng = np.random.default_rng(1)
    hdr = rng.uniform(-0.5, 1.5, (6, 4, 3)).astype(np.float32)
    before = hdr.copy()
    out = o.process(hdr, 3.0, bloom=False, tone_map=False)
    np.testing.assert_array_equal(out, np.clip(hdr, 0.0, 1.0))
    np.testing.assert_array_equal(hdr, before)
    assert o.exposure_multiplier(3.0, True) == 8.0 and o.exposure_multiplier(3.0, False) == 1.0
    assert not o.use_bloom(True) or o.config.bloom_strength > 0
    assert not oracle(post, bloom_passes=0).use_bloom(True)
    assert not oracle(post, bloom_strength=0.0).use_bloom(True)


# ---------------------------------------------------- whole process, float64


def numpy_process(hdr, config, exposure_ev, bloom, tone_map):
    """float64 formulation of postprocess.py:110-207 by slicing, padding and
    gathers."""
    hdr = hdr.astype(np.float64)
    w, h = hdr.shape[:2]
    hw, hh = max((w + 1) // 2, 1), max((h + 1) // 2, 1)
    em = 2.0 ** exposure_ev if tone_map else 1.0
    value = hdr * em
    if bloom and config.bloom_passes > 0 and config.bloom_strength > 0:
        pad = np.pad(hdr, ((0, 2 * hw - w), (0, 2 * hh - h), (0, 0)), mode="edge")
        mean = 0.25 * (pad[0::2, 0::2] + pad[1::2, 0::2] + pad[0::2, 1::2] + pad[1::2, 1::2])
        x = mean * em
        b = x.max(axis=2, keepdims=True)
        t, k = config.bloom_threshold, config.bloom_knee
        soft = np.clip(b - t + k, 0.0, 2.0 * k) ** 2 / max(4.0 * k, 1e-6)
        x = x * (np.maximum(b - t, soft) / np.maximum(b, 1e-6))
        peak = x.max(axis=2, keepdims=True)
        field = x * np.minimum(1.0, config.bloom_clamp / np.maximum(peak, 1e-6))
        for _ in range(config.bloom_passes):
            p = np.pad(field, ((2, 2), (0, 0), (0, 0)), mode="edge")
            field = sum(TAPS[i] * p[i:i + hw] for i in range(5))
            p = np.pad(field, ((0, 0), (2, 2), (0, 0)), mode="edge")
            field = sum(TAPS[i] * p[:, i:i + hh] for i in range(5))
        px = (np.arange(w) + 0.5) / w * hw - 0.5
        py = (np.arange(h) + 0.5) / h * hh - 0.5
        x0 = np.clip(np.floor(px).astype(int), 0, hw - 1)
        y0 = np.clip(np.floor(py).astype(int), 0, hh - 1)
        x1, y1 = np.minimum(x0 + 1, hw - 1), np.minimum(y0 + 1, hh - 1)
        fx = np.clip(px - np.floor(px), 0, 1)[:, None, None]
        fy = np.clip(py - np.floor(py), 0, 1)[None, :, None]
        g = lambda xs, ys: field[xs[:, None], ys[None, :]]  # noqa: E731
        sample = ((g(x0, y0) * (1 - fx) + g(x1, y0) * fx) * (1 - fy)
                  + (g(x0, y1) * (1 - fx) + g(x1, y1) * fx) * fy)
        value = value + sample * config.bloom_strength
    if not tone_map:
        return np.clip(value, 0.0, 1.0)
    v = np.maximum(value, 0.0)
    m = np.clip(v * (A * v + B) / (v * (C * v + D) + E), 0.0, 1.0)
    return np.where(m <= 0.0031308, 12.92 * m, 1.055 * m ** (1 / 2.4) - 0.055)


@pytest.mark.parametrize("size, changes, ev, bloom, tone_map", [
    ((37, 23), {}, 0.0, True, True),
    ((64, 64), {"bloom_passes": 1}, -2.0, True, True),
    ((31, 8), {"bloom_passes": 8, "bloom_knee": 0.0}, 1.5, True, True),
    ((9, 5), {"bloom_strength": 0.5, "bloom_clamp": 2.0}, 0.0, True, False),
    ((20, 30), {}, 2.0, False, True),
    ((1, 1), {}, 0.0, True, True),
])
def test_process_matches_float64_formulation(post, size, changes, ev, bloom, tone_map):
    rng = np.random.default_rng(11)
    w, h = size
    hdr = np.exp(rng.uniform(np.log(1e-4), np.log(50.0), (w, h, 3))).astype(np.float32)
    o = oracle(post, w, h, **changes)
    got = o.process(hdr, ev, bloom=bloom, tone_map=tone_map)
    want = numpy_process(hdr, o.config, ev, bloom, tone_map)
    np.testing.assert_allclose(got, want, rtol=2e-5, atol=2e-6)


def test_abi_error_paths_without_gpu_work(pkg):
    """Status codes of the new entry points for arguments rejected before
    any device call: a NULL post-process handle, NULL shading fields."""
    capi = __import__(pkg.__name__ + "._capi", fromlist=["_capi"])
    lib = capi.library()
    assert lib.atmo_post_process(None, None, None, 1.0, 1, 1, None) == capi.ATMO_EHANDLE
    assert lib.atmo_post_destroy(None) == capi.ATMO_EHANDLE
    assert lib.atmo_post_bind_buffer(None, 0, None, 0) == capi.ATMO_EHANDLE
    assert lib.atmo_post_last_error(None) is not None
    fields = capi.AtmoShadeFields()  # every pointer NULL
    params = capi.AtmoShadeParams()
    assert lib.atmo_shade_surface(ctypes.byref(fields), ctypes.byref(params), 4, 4, None) == \
        capi.ATMO_EINVAL
    assert lib.atmo_shade_surface(None, ctypes.byref(params), 4, 4, None) == capi.ATMO_EINVAL
    fake = capi.AtmoShadeFields(**{name: 16 for name in capi.SHADE_FIELD_NAMES})
    assert lib.atmo_shade_surface(ctypes.byref(fake), ctypes.byref(params), 0, 4, None) == \
        capi.ATMO_EINVAL


def test_noalloc_composite_rejects_before_device_work(pkg):
    """atmo_composite_noalloc's argument checks run before any device call:
    NULL pointers, an invalid configuration, a bad view, and a frame above
    ATMO_NOALLOC_MAX_PIXELS (the static queue's capacity)."""
    capi = __import__(pkg.__name__ + "._capi", fromlist=["_capi"])
    lib = capi.library()
    params = pkg.AtmosphereConfig().to_params()
    v = capi.f3((0.0, 1.0, 0.0))
    fake = 16

    def call(p, w, h, view=0, lut=fake):
        return lib.atmo_composite_noalloc(
            p, lut, fake, fake, fake, w, h, fake, fake, fake, fake, fake, 6.36e6, 10.0, 10.0, v,
            v, v, 0.0046, v, v, v, 0.5, 4.0, 12.0, view, None)

    assert call(ctypes.byref(params), 64, 64, lut=None) == capi.ATMO_EINVAL
    assert call(None, 64, 64) == capi.ATMO_EINVAL
    assert call(ctypes.byref(params), 64, 64, view=11) == capi.ATMO_EINVAL
    assert call(ctypes.byref(params), 0, 64) == capi.ATMO_EINVAL
    assert call(ctypes.byref(params), 4096, 2049) == capi.ATMO_ESHAPE
    assert b"exceed" in lib.atmo_last_error(None)
    bad = pkg.AtmosphereConfig().to_params()
    bad.transmittance_steps = 1
    assert call(ctypes.byref(bad), 64, 64) == capi.ATMO_EINVAL
    assert b"transmittance_steps" in lib.atmo_last_error(None)
