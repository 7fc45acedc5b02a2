"""ctypes wrapper of the post-process oracle (oracle/post_cpu.c).  TEST
INFRASTRUCTURE ONLY -- imported by tests/ and the CPU-baseline leg of
bench.py as the checker, never by the shipped package.

`PostOracle.process` restates `PostProcessor.process`
(src/planet_renderer/postprocess.py:176-207): the host-side exposure
multiplier 2**ev (Python float, then float32 as a `ti.f32` argument), the
bloom switch, the extract / blur / resolve kernel sequence, and the float32
rounding of the `ti.static` bright-pass constants (postprocess.py:88-108).
Arrays are NumPy (X, Y, 3) float32 in the reference's field layout.
"""

from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

from .ref_cpu import build

ORACLE_DIR = Path(__file__).resolve().parent
LIBRARY = ORACLE_DIR / "libpost_cpu.so"


class PostConstants(ctypes.Structure):
    _fields_ = [
        ("threshold", ctypes.c_float),
        ("knee", ctypes.c_float),
        ("two_knee", ctypes.c_float),
        ("knee_denominator", ctypes.c_float),
        ("clamp", ctypes.c_float),
    ]


def bright_pass_constants(config) -> PostConstants:
    """float32 constants of `_soft_threshold`: each Python expression is
    evaluated in float64 and rounded once, as `ti.static` folds it."""
    knee = float(config.bloom_knee)
    return PostConstants(float(config.bloom_threshold), knee, 2.0 * knee,
                         max(4.0 * knee, 1.0e-6), float(config.bloom_clamp))


_F = ctypes.c_float
_I = ctypes.c_int32
_P = ctypes.c_void_p


def _lib() -> ctypes.CDLL:
    if not LIBRARY.exists():
        build()
    lib = ctypes.CDLL(str(LIBRARY))
    lib.post_display_channel.argtypes = [_F]
    lib.post_display_channel.restype = _F
    lib.post_srgb_pow_array.argtypes = [_P, _P, ctypes.c_int64]
    lib.post_check_srgb_pow.argtypes = [_F, _F]
    lib.post_check_srgb_pow.restype = ctypes.c_int64
    lib.post_extract.argtypes = [_P, _I, _I, _P, _I, _I, _F, ctypes.POINTER(PostConstants)]
    lib.post_blur_horizontal.argtypes = [_P, _P, _I, _I]
    lib.post_blur_vertical.argtypes = [_P, _P, _I, _I]
    lib.post_resolve.argtypes = [_P, _P, _I, _I, _P, _I, _I, _F, _F, _I, _I]
    return lib


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


class PostOracle:
    """`PostProcessor(width, height, config)` on the CPU: same fields
    (`bloom_a`, `bloom_b`, half-resolution sizes) and `process`."""

    def __init__(self, width: int, height: int, config):
        self.lib = _lib()
        self.width = int(width)
        self.height = int(height)
        self.half_width = max((self.width + 1) // 2, 1)
        self.half_height = max((self.height + 1) // 2, 1)
        self.config = config
        shape = (self.half_width, self.half_height, 3)
        self.bloom_a = np.zeros(shape, np.float32)
        self.bloom_b = np.zeros(shape, np.float32)

    @staticmethod
    def exposure_multiplier(exposure_ev: float, tone_map: bool) -> float:
        return 2.0 ** float(exposure_ev) if tone_map else 1.0

    def use_bloom(self, bloom: bool) -> bool:
        c = self.config
        return bool(bloom and c.bloom_passes > 0 and c.bloom_strength > 0.0)

    def extract(self, hdr: np.ndarray, exposure_multiplier: float) -> None:
        hdr = _f32(hdr)
        k = bright_pass_constants(self.config)
        self.lib.post_extract(_ptr(hdr), self.width, self.height, _ptr(self.bloom_a),
                              self.half_width, self.half_height, exposure_multiplier,
                              ctypes.byref(k))

    def blur_pass(self) -> None:
        self.lib.post_blur_horizontal(_ptr(self.bloom_a), _ptr(self.bloom_b),
                                      self.half_width, self.half_height)
        self.lib.post_blur_vertical(_ptr(self.bloom_b), _ptr(self.bloom_a),
                                    self.half_width, self.half_height)

    def resolve(self, hdr: np.ndarray, exposure_multiplier: float, use_bloom: bool,
                tone_map: bool) -> np.ndarray:
        hdr = _f32(hdr)
        display = np.empty((self.width, self.height, 3), np.float32)
        self.lib.post_resolve(_ptr(hdr), _ptr(display), self.width, self.height,
                              _ptr(self.bloom_a), self.half_width, self.half_height,
                              exposure_multiplier, float(self.config.bloom_strength),
                              int(use_bloom), int(tone_map))
        return display

    def process(self, hdr: np.ndarray, exposure_ev: float, *, bloom: bool,
                tone_map: bool) -> np.ndarray:
        """postprocess.py:176-207; returns the display array."""
        hdr = _f32(hdr)
        if hdr.shape != (self.width, self.height, 3):
            raise ValueError(f"hdr must have shape {(self.width, self.height, 3)}")
        em = self.exposure_multiplier(exposure_ev, tone_map)
        use_bloom = self.use_bloom(bloom)
        if use_bloom:
            self.extract(hdr, em)
            for _ in range(self.config.bloom_passes):
                self.blur_pass()
        return self.resolve(hdr, em, use_bloom, tone_map)

    # ------------------------------------------------------------ pieces

    def display_channel(self, values) -> np.ndarray:
        """`_display_from_exposed` per element (postprocess.py:56-71)."""
        v = _f32(values).ravel()
        out = np.array([self.lib.post_display_channel(float(x)) for x in v], np.float32)
        return out.reshape(np.shape(values))

    def srgb_pow(self, values: np.ndarray) -> np.ndarray:
        v = _f32(values)
        out = np.empty_like(v)
        self.lib.post_srgb_pow_array(_ptr(v), _ptr(out), v.size)
        return out

    def check_srgb_pow(self, lo: float, hi: float) -> int:
        return int(self.lib.post_check_srgb_pow(lo, hi))
