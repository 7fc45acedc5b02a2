"""Many atmosphere states' LUTs in one graph launch.

The reference renders one state per `AtmosphereRenderer`; a state's LUTs
come from a cold `update` (renderer.py:2072-2137, all of T, MS and SV
rebuilt) followed by `prepare_frame`'s aerial-perspective build
(renderer.py:2173-2187).  `LutBatch` builds exactly those four LUTs for
every state of a list at once: the library captures the four batched stage
launches (one grid row of blocks per state) into a HIP graph on the first
`run` and replays it afterwards.  The LUT arrays are torch tensors owned by
this object, shaped (count, *single-state LUT shape).
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _capi
from ._capi import AtmoBatchState, library
from .config import AtmosphereConfig


@dataclass
class AtmosphereState:
    """One state in the reference's host units (float64 metres, degrees):
    the arguments a cold `update` + `prepare_frame` would receive."""

    config: AtmosphereConfig
    planet_radius_m: float
    camera_radius_m: float
    sun_local: np.ndarray
    solar_irradiance: tuple
    sun_angular_radius_degrees: float
    view_basis: tuple  # (right, view_up, forward)
    tangent_half_fov: float

    def to_struct(self) -> AtmoBatchState:
        s = AtmoBatchState()
        s.params = self.config.to_params()
        s.planet_radius_m = float(self.planet_radius_m)
        s.camera_radius_m = float(self.camera_radius_m)
        s.sun_local[:] = [float(v) for v in self.sun_local]
        s.solar_irradiance[:] = [float(v) for v in self.solar_irradiance]
        s.sun_angular_radius_deg = float(self.sun_angular_radius_degrees)
        right, up, forward = self.view_basis
        s.right[:] = [float(v) for v in right]
        s.view_up[:] = [float(v) for v in up]
        s.forward[:] = [float(v) for v in forward]
        s.tangent_half_fov = float(self.tangent_half_fov)
        return s


class LutBatch:
    """T, MS, SV and AP LUTs of `states` (same LUT dimensions for all),
    for a `width` x `height` frame's aerial-perspective rays."""

    NAMES = ("transmittance", "multi_scattering", "sky_view", "aerial_scattering")

    def __init__(self, states, width: int, height: int, device=None):
        import torch

        states = list(states)
        if not states:
            raise ValueError("a batch needs at least one state")
        self.device = torch.device(device if device is not None else "cuda")
        self.count = len(states)
        self.width, self.height = int(width), int(height)
        self._lib = library()
        c = states[0].config
        shapes = ((c.transmittance_lut_width, c.transmittance_lut_height, 3),
                  (c.multi_scattering_lut_width, c.multi_scattering_lut_height, 3),
                  (c.sky_view_lut_width, c.sky_view_lut_height, 3),
                  (c.aerial_lut_width, c.aerial_lut_height, c.aerial_lut_depth, 3))
        self.luts = [torch.zeros((self.count,) + s, dtype=torch.float32, device=self.device)
                     for s in shapes]
        table = (AtmoBatchState * self.count)(*[s.to_struct() for s in states])
        ptrs = (ctypes.c_void_p * 4)(*[t.data_ptr() for t in self.luts])
        handle = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            status = self._lib.atmo_batch_create(table, self.count, self.width, self.height, ptrs,
                                                 ctypes.byref(handle))
        if status != _capi.ATMO_OK:
            message = self._lib.atmo_batch_last_error(None).decode(errors="replace")
            raise (ValueError if status in (_capi.ATMO_EINVAL, _capi.ATMO_ESHAPE)
                   else _capi.AtmoError)(message)
        self._handle = handle

    def run(self, stream=None) -> None:
        """Build every state's four LUTs (asynchronous on `stream`, default
        torch's current stream)."""
        import torch

        s = stream if stream is not None else torch.cuda.current_stream(self.device)
        status = self._lib.atmo_batch_run(self._handle, s.cuda_stream)
        if status != _capi.ATMO_OK:
            raise _capi.AtmoError(self._lib.atmo_batch_last_error(self._handle)
                                  .decode(errors="replace"))

    def run_stages(self, stages: int, stream=None) -> None:
        """Launch only the stage kernels in `stages` (bits 1, 2, 4, 8: T, MS,
        SV, AP) directly, without the graph; each reads what the previous
        stages left in the batch arrays (atmo_batch_run_stages)."""
        import torch

        s = stream if stream is not None else torch.cuda.current_stream(self.device)
        status = self._lib.atmo_batch_run_stages(self._handle, int(stages), s.cuda_stream)
        if status != _capi.ATMO_OK:
            raise _capi.AtmoError(self._lib.atmo_batch_last_error(self._handle)
                                  .decode(errors="replace"))

    def autotune(self, stream=None) -> tuple[int, int]:
        """Time the five lanes-per-texel sizes (1, 2, 4, 8, 16) of the
        batch's transmittance stage, then of its sky-view stage, and keep the
        fastest of each for the captured graph (atmo_batch_autotune).
        Synchronises the stream."""
        import torch

        s = stream if stream is not None else torch.cuda.current_stream(self.device)
        chosen = (ctypes.c_int32 * 2)()
        status = self._lib.atmo_batch_autotune(self._handle, s.cuda_stream, chosen)
        if status != _capi.ATMO_OK:
            raise _capi.AtmoError(self._lib.atmo_batch_last_error(self._handle)
                                  .decode(errors="replace"))
        return int(chosen[0]), int(chosen[1])

    @property
    def transmittance(self):
        return self.luts[0]

    @property
    def multi_scattering(self):
        return self.luts[1]

    @property
    def sky_view(self):
        return self.luts[2]

    @property
    def aerial_scattering(self):
        return self.luts[3]

    def close(self) -> None:
        if getattr(self, "_handle", None) is not None and self._handle.value:
            self._lib.atmo_batch_destroy(self._handle)
            self._handle = None

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass


__all__ = ["AtmosphereState", "LutBatch"]
