"""Linear-HDR exposure, bloom, tone mapping and display encoding over
libatmo.so: the reference's `PostprocessConfig` and `PostProcessor`
(src/planet_renderer/postprocess.py).

`PostprocessConfig` has the reference's fields, defaults, validation
messages and `from_mapping`.  `PostProcessor` keeps the reference's host
policy (postprocess.py:176-207): the exposure multiplier 2**exposure_ev as
a Python float rounded to float32 (1 without tone mapping), and bloom only
when requested with bloom_passes > 0 and bloom_strength > 0.  The kernels
are libatmo.so's (atmo_post_process); `bloom_a` and `bloom_b` are the
half-resolution fields, torch tensors bound into the library handle.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass
import math
from typing import Mapping

from . import _capi
from ._capi import library
from .renderer import DeviceField, _device_tensor, _torch


@dataclass(frozen=True)
class PostprocessConfig:
    """Display parameters kept separate from physical scene radiance
    (postprocess.py:10-53)."""

    exposure_ev: float = 0.0
    bloom_threshold: float = 1.0
    bloom_knee: float = 0.5
    bloom_clamp: float = 8.0
    bloom_strength: float = 0.08
    bloom_passes: int = 4

    def __post_init__(self) -> None:
        if not math.isfinite(self.exposure_ev) or not -20.0 <= self.exposure_ev <= 20.0:
            raise ValueError("postprocess.exposure_ev must be finite and in [-20, 20]")
        finite_non_negative = {
            "bloom_threshold": self.bloom_threshold,
            "bloom_knee": self.bloom_knee,
            "bloom_clamp": self.bloom_clamp,
            "bloom_strength": self.bloom_strength,
        }
        for name, value in finite_non_negative.items():
            if not math.isfinite(value) or value < 0.0:
                raise ValueError(f"postprocess.{name} must be finite and non-negative")
        if not 0 <= self.bloom_passes <= 8:
            raise ValueError("postprocess.bloom_passes must be in 0..8")

    @classmethod
    def from_mapping(cls, values: Mapping[str, object]) -> "PostprocessConfig":
        defaults = cls()
        return cls(
            exposure_ev=float(values.get("exposure_ev", defaults.exposure_ev)),
            bloom_threshold=float(values.get("bloom_threshold", defaults.bloom_threshold)),
            bloom_knee=float(values.get("bloom_knee", defaults.bloom_knee)),
            bloom_clamp=float(values.get("bloom_clamp", defaults.bloom_clamp)),
            bloom_strength=float(values.get("bloom_strength", defaults.bloom_strength)),
            bloom_passes=int(values.get("bloom_passes", defaults.bloom_passes)),
        )

    def to_params(self) -> _capi.AtmoPostParams:
        """struct AtmoPostParams of the C-ABI, field for field."""
        return _capi.AtmoPostParams(float(self.exposure_ev), float(self.bloom_threshold),
                                    float(self.bloom_knee), float(self.bloom_clamp),
                                    float(self.bloom_strength), int(self.bloom_passes))


class PostProcessor:
    """Half-resolution bloom followed by the sole display transform
    (postprocess.py:74-207)."""

    def __init__(self, width: int, height: int, config: PostprocessConfig, device=None):
        torch = _torch()
        if not torch.cuda.is_available():
            raise RuntimeError("PostProcessor needs a HIP device (libatmo.so has no CPU path)")
        self.width = int(width)
        self.height = int(height)
        self.half_width = max((self.width + 1) // 2, 1)
        self.half_height = max((self.height + 1) // 2, 1)
        self.config = config
        self.device = torch.device(device if device is not None else "cuda")
        if self.device.index is None:
            self.device = torch.device("cuda", torch.cuda.current_device())
        lib = library()
        self._lib = lib
        self._params = config.to_params()
        handle = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            status = lib.atmo_post_create(ctypes.byref(self._params), self.width, self.height,
                                          ctypes.byref(handle))
            if status != _capi.ATMO_OK:
                message = lib.atmo_post_last_error(None).decode(errors="replace")
                if status == _capi.ATMO_EINVAL:
                    raise ValueError(message)
                raise _capi.AtmoError(f"atmo status {status}: {message}")
            self._handle = handle
            fields = []
            for which in (_capi.POST_BUF_BLOOM_A, _capi.POST_BUF_BLOOM_B):
                tensor = torch.zeros((self.half_width, self.half_height, 3), dtype=torch.float32,
                                     device=self.device)
                torch.cuda.synchronize(self.device)
                self._check(lib.atmo_post_bind_buffer(handle, which, tensor.data_ptr(),
                                                      tensor.numel() * 4))
                fields.append(DeviceField(tensor))
        self.bloom_a, self.bloom_b = fields

    def _check(self, status: int) -> None:
        if status == _capi.ATMO_OK:
            return
        message = self._lib.atmo_post_last_error(self._handle).decode(errors="replace")
        if status in (_capi.ATMO_EINVAL, _capi.ATMO_ESHAPE):
            raise ValueError(message)
        raise _capi.AtmoError(f"atmo status {status}: {message}")

    def close(self) -> None:
        handle = getattr(self, "_handle", None)
        if handle is not None and handle.value:
            _torch().cuda.synchronize(self.device)
            self._lib.atmo_post_destroy(handle)
            self._handle = None

    def __del__(self):  # pragma: no cover - best effort at interpreter exit
        try:
            self.close()
        except Exception:
            pass

    def process(self, hdr, display, exposure_ev: float, *, bloom: bool, tone_map: bool) -> None:
        """Convert one HDR frame without ever modifying the HDR source
        (postprocess.py:176-207).  hdr and display are (W, H, 3) float32
        device tensors (or DeviceFields); the call is asynchronous on the
        current stream."""
        torch = _torch()
        shape = (self.width, self.height, 3)
        hdr_t = _device_tensor(hdr, "hdr", shape, torch.float32, self.device)
        display_t = _device_tensor(display, "display", shape, torch.float32, self.device)
        # Raw transmittance/mask diagnostics bypass exposure (postprocess.py:187-189).
        exposure_multiplier = 2.0 ** float(exposure_ev) if tone_map else 1.0
        use_bloom = bool(bloom and self.config.bloom_passes > 0
                         and self.config.bloom_strength > 0.0)
        stream = torch.cuda.current_stream(self.device).cuda_stream
        self._check(self._lib.atmo_post_process(self._handle, hdr_t.data_ptr(),
                                                display_t.data_ptr(), exposure_multiplier,
                                                int(use_bloom), int(tone_map), stream))


__all__ = ["PostProcessor", "PostprocessConfig"]
