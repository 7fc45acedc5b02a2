"""GPU parity of the display post-process (libatmo.so's atmo_post_process)
with the CPU oracle (oracle/post_cpu.c), the restatement of the
reference's `PostProcessor` (src/planet_renderer/postprocess.py:74-207).

Bar: bit-exact.  Every stage is float32 arithmetic in the reference's
operation order, and the one elementary function, the sRGB power, is
correctly rounded on both sides (checked on every float of its domain
below), so display, bloom_a and bloom_b must match to the bit."""

import ctypes
import math

import numpy as np
import pytest

from conftest import REPO

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def device():
    import torch

    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch.device("cuda:0")


def hdr_frame(width, height, seed):
    """Linear HDR radiance over eight decades, with a few sun-bright pixels,
    NaN-free and non-negative as the composite writes it."""
    rng = np.random.default_rng(seed)
    hdr = np.exp(rng.uniform(math.log(1e-5), math.log(2e3), (width, height, 3)))
    bright = rng.random((width, height)) < 2e-4
    hdr[bright] = rng.uniform(1e4, 6e4, (int(bright.sum()), 3))
    return hdr.astype(np.float32)


def gpu_process(pkg, hdr, config, ev, bloom, tone_map, device):
    import torch

    w, h = hdr.shape[:2]
    post = pkg.PostProcessor(w, h, config, device)
    src = torch.from_numpy(hdr).to(device)
    display = torch.full_like(src, float("nan"))
    post.process(src, display, ev, bloom=bloom, tone_map=tone_map)
    torch.cuda.synchronize(device)
    out = (display.cpu().numpy(), post.bloom_a.to_numpy(), post.bloom_b.to_numpy())
    assert np.array_equal(src.cpu().numpy().view(np.uint32), hdr.view(np.uint32)), \
        "process modified its HDR source"
    post.close()
    return out


def oracle_process(hdr, config, ev, bloom, tone_map):
    from oracle.post_cpu import PostOracle

    o = PostOracle(hdr.shape[0], hdr.shape[1], config)
    display = o.process(hdr, ev, bloom=bloom, tone_map=tone_map)
    return display, o.bloom_a, o.bloom_b


def assert_bit_exact(name, got, want):
    same = (got.view(np.uint32) == want.view(np.uint32))
    if not same.all():
        idx = np.argwhere(~same)[0]
        raise AssertionError(f"{name}: {int((~same).sum())} of {same.size} values differ; first "
                             f"at {tuple(idx)}: {got[tuple(idx)]!r} vs {want[tuple(idx)]!r}")


CASES = [
    ((3840, 2160), {}, 0.0, True, True),
    ((1920, 1080), {}, 0.0, True, True),
    ((97, 61), {}, 1.0, True, True),
    ((64, 64), {"bloom_passes": 1}, -2.0, True, True),
    ((65, 130), {"bloom_passes": 8, "bloom_knee": 0.0}, 0.5, True, True),
    ((150, 90), {"bloom_passes": 5}, 0.0, True, True),
    ((150, 90), {"bloom_passes": 6, "bloom_threshold": 0.2}, 0.0, True, True),
    ((150, 90), {"bloom_passes": 7}, -1.0, True, True),
    ((151, 91), {"bloom_passes": 2, "bloom_strength": 0.6}, 3.0, True, True),
    ((200, 120), {"bloom_passes": 3, "bloom_strength": 0.5, "bloom_clamp": 2.0}, 0.0, True, False),
    ((33, 1), {}, 0.0, True, True),
    ((1, 1), {}, 0.0, True, True),
    ((2, 3), {"bloom_passes": 2}, 0.0, True, True),
    ((5, 3), {"bloom_passes": 8}, 1.0, True, True),
    ((3, 201), {"bloom_passes": 8}, 0.0, True, True),
    ((300, 170), {}, 2.0, False, True),
    ((300, 170), {}, 0.0, False, False),
    ((300, 170), {"bloom_passes": 0}, 0.0, True, True),
    ((300, 170), {"bloom_strength": 0.0}, 0.0, True, True),
]


@pytest.mark.parametrize("size, changes, ev, bloom, tone_map", CASES)
def test_post_process_bit_exact(pkg, device, size, changes, ev, bloom, tone_map):
    config = pkg.PostprocessConfig(**changes)
    hdr = hdr_frame(*size, seed=size[0] * 7 + size[1])
    print(f"post-process {size} {changes} ev={ev} bloom={bloom} tone_map={tone_map}", flush=True)
    got = gpu_process(pkg, hdr, config, ev, bloom, tone_map, device)
    want = oracle_process(hdr, config, ev, bloom, tone_map)
    assert_bit_exact("display", got[0], want[0])
    assert_bit_exact("bloom_a", got[1], want[1])
    assert_bit_exact("bloom_b", got[2], want[2])


def test_post_process_non_finite_radiance(pkg, device):
    """NaN and infinite HDR pixels go through the same IEEE operations on
    both sides (maxNum drops NaN, inf / inf is NaN, the clamp maps NaN to
    0): the outputs agree value for value, NaN for NaN."""
    hdr = hdr_frame(130, 70, seed=3)
    hdr[10, 10] = np.nan
    hdr[50, 20, 1] = np.inf
    hdr[90, 40] = (np.inf, -np.inf, np.nan)
    hdr[129, 69, 2] = np.nan
    config = pkg.PostprocessConfig()
    for bloom, tone_map in ((True, True), (False, True), (False, False), (True, False)):
        got = gpu_process(pkg, hdr, config, 0.5, bloom, tone_map, device)
        want = oracle_process(hdr, config, 0.5, bloom, tone_map)
        for name, g, w in zip(("display", "bloom_a", "bloom_b"), got, want):
            if not bloom and name != "display":
                continue
            nan_g, nan_w = np.isnan(g), np.isnan(w)
            assert (nan_g == nan_w).all(), f"{name}: NaN positions differ"
            assert_bit_exact(name, np.where(nan_g, 0, g).astype(np.float32),
                             np.where(nan_w, 0, w).astype(np.float32))


def test_post_process_of_a_composite_frame(pkg, scene, device, tmp_path):
    """The frame the composite writes (1920x1080, COMPOSITE view), then the
    post-process the reference runs on it (renderer.py:1728-1735), and the
    linear-HDR file the reference saves from it (cli.py:143-145)."""
    import torch

    from test_gpu_parity import gpu_frame

    inputs = scene.make_frame(width=1920, height=1080, config_index=1)
    print("post-process of the 1080p composite frame", flush=True)
    hdr = gpu_frame(pkg, inputs, device)["hdr"]
    assert np.isfinite(hdr).all()
    view = pkg.AtmosphereDiagnosticView.COMPOSITE
    config = pkg.PostprocessConfig()
    got = gpu_process(pkg, hdr, config, 0.0, view.uses_bloom, view.is_radiance, device)
    want = oracle_process(hdr, config, 0.0, view.uses_bloom, view.is_radiance)
    for name, g, w in zip(("display", "bloom_a", "bloom_b"), got, want):
        assert_bit_exact(name, g, w)
    assert float(want[1].max()) > 0.0  # the frame has a bright pass to blur
    pkg.save_hdr(torch.from_numpy(hdr).to(device), tmp_path / "frame.npy")
    np.save(tmp_path / "reference.npy", hdr)
    assert (tmp_path / "frame.npy").read_bytes() == (tmp_path / "reference.npy").read_bytes()


def test_srgb_power_on_every_float_of_its_domain(device):
    """post::srgb_pow (the device power of the sRGB encode) equals the
    oracle's correctly rounded power for every float in (0.0031308, 1]."""
    import torch

    from oracle.post_cpu import PostOracle

    path = REPO / "planet-real-time-render_amd" / "libatmo_probe.so"
    assert path.exists(), "libatmo_probe.so is missing: build first"
    lib = ctypes.CDLL(str(path))
    lib.atmo_probe_srgb_pow.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                        ctypes.c_void_p]
    lib.atmo_probe_srgb_pow.restype = ctypes.c_int32
    lo = int(np.float32(0.0031308).view(np.int32)) + 1
    hi = int(np.float32(1.0).view(np.int32))
    n = hi - lo + 1
    print(f"srgb power: {n} inputs", flush=True)
    bits = torch.arange(lo, hi + 1, dtype=torch.int32, device=device)
    x = bits.view(torch.float32)
    y = torch.empty_like(x)
    stream = torch.cuda.current_stream(device).cuda_stream
    assert lib.atmo_probe_srgb_pow(x.data_ptr(), y.data_ptr(), n, stream) == 0
    torch.cuda.synchronize(device)
    got = y.cpu().numpy()
    inputs = x.cpu().numpy()
    want = PostOracle(2, 2, None).srgb_pow(inputs)
    bad = np.flatnonzero(got.view(np.uint32) != want.view(np.uint32))
    print(f"srgb power mismatches: {bad.size}", flush=True)
    assert bad.size == 0, f"first mismatch at {inputs[bad[0]]!r}: {got[bad[0]]!r} vs {want[bad[0]]!r}"


def test_post_process_argument_errors(pkg, device):
    import torch

    post = pkg.PostProcessor(8, 6, pkg.PostprocessConfig(), device)
    good = torch.zeros((8, 6, 3), dtype=torch.float32, device=device)
    with pytest.raises(ValueError):
        post.process(torch.zeros((6, 8, 3), dtype=torch.float32, device=device), good, 0.0,
                     bloom=True, tone_map=True)
    with pytest.raises(ValueError):
        post.process(good, torch.zeros((8, 6, 3), dtype=torch.float64, device=device), 0.0,
                     bloom=True, tone_map=True)
    with pytest.raises(ValueError):
        post.process(good.cpu(), good, 0.0, bloom=True, tone_map=True)
    assert post.bloom_a.shape == (4, 3, 3) and post.bloom_b.shape == (4, 3, 3)
    post.close()
