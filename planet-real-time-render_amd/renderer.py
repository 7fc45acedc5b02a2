"""`AtmosphereRenderer`: the reference's atmosphere interface over libatmo.so.

The host policy is the reference's own (src/planet_renderer/atmosphere/
renderer.py:2057-2238): which LUTs `update` rebuilds and when, the freeze
rule, the sky-view snapshot altitude, the rebuild counters, and the float64
-> float32 rounding of every kernel argument.  The six kernels behind it are
the HIP kernels of libatmo.so, called through ctypes; the LUT and
surface-lighting fields are torch tensors on the HIP device, bound into the
library handle, and G-buffer / HDR fields are passed as torch tensors in the
reference's Taichi field layout ((W, H, 3) float32, (W, H) int32).
"""

from __future__ import annotations

import ctypes
import math
import warnings

import numpy as np

from . import _capi
from ._capi import check, f3, library
from .config import AtmosphereConfig
from .diagnostics import AtmosphereDiagnosticView


def _torch():
    import torch

    return torch


class DeviceField:
    """A device-resident field with the Taichi field methods the reference's
    callers use (`to_numpy`, `from_numpy`, `shape`)."""

    def __init__(self, tensor):
        self.tensor = tensor

    @property
    def shape(self) -> tuple[int, ...]:
        return tuple(self.tensor.shape)

    def to_numpy(self) -> np.ndarray:
        return self.tensor.detach().cpu().numpy()

    def from_numpy(self, array) -> None:
        torch = _torch()
        self.tensor.copy_(torch.from_numpy(np.ascontiguousarray(array, dtype=np.float32)))

    def data_ptr(self) -> int:
        return self.tensor.data_ptr()


def _device_tensor(value, name: str, shape: tuple[int, ...], dtype, device):
    tensor = value.tensor if isinstance(value, DeviceField) else value
    torch = _torch()
    if not isinstance(tensor, torch.Tensor):
        raise ValueError(f"{name} must be a device tensor or DeviceField")
    if tensor.device.type != "cuda" or (device.index is not None and tensor.device != device):
        raise ValueError(f"{name} must live on {device}")
    if tensor.dtype != dtype:
        raise ValueError(f"{name} must have dtype {dtype}")
    if tuple(tensor.shape) != shape:
        raise ValueError(f"{name} must have shape {shape}, got {tuple(tensor.shape)}")
    if not tensor.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    return tensor


PRECISIONS = {"reference": _capi.PRECISION_REFERENCE, "robust": _capi.PRECISION_ROBUST}


def _warn_if_below_gate(surrogate) -> None:
    """One warning per network whose weight file records that it misses the
    1 % p99 accuracy gate (the surrogate composite then trades accuracy for
    speed; the plain composite is the accurate path)."""
    if getattr(surrogate, "gate_met", None) is False and not getattr(surrogate, "_gate_warned", False):
        validation = surrogate.metadata.get("validation", {})
        warnings.warn(
            f"surrogate weights miss the 1% p99 accuracy gate: hdr p99 relative error "
            f"{validation.get('hdr_p99_vs_fp64')} against the float64 restatement on "
            f"{validation.get('frame')}; the plain composite is the accurate path",
            RuntimeWarning, stacklevel=3)
        surrogate._gate_warned = True


class AtmosphereRenderer:
    """Spherical Rayleigh/Mie atmosphere rendered through compact LUTs
    (renderer.py:42-116).  The camera sits at the floating origin; the planet
    centre is (0, -camera_radius, 0) in the camera's East-Up-North frame.

    `precision` "reference" (default) reproduces the reference's float32
    kernels; "robust" evaluates the planet-scale cancellations (near-tangent
    ray-sphere solves, the sky-view horizon split, endpoints above 1000 km)
    in float64 and compensates the optical-depth sums -- an extension with
    no counterpart in the reference."""

    def __init__(self, width: int, height: int, config: AtmosphereConfig, device=None,
                 precision: str = "reference"):
        torch = _torch()
        if not torch.cuda.is_available():
            raise RuntimeError("AtmosphereRenderer needs a HIP device (libatmo.so has no CPU path)")
        self.width = int(width)
        self.height = int(height)
        self.config = config
        self.device = torch.device(device if device is not None else "cuda", )
        if self.device.index is None:
            self.device = torch.device("cuda", torch.cuda.current_device())
        if precision not in PRECISIONS:
            raise ValueError(f"precision must be one of {sorted(PRECISIONS)}, got {precision!r}")
        self.precision = precision
        lib = library()
        self._lib = lib
        self._params = config.to_params()
        handle = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            check(lib.atmo_create(ctypes.byref(self._params), self.width, self.height,
                                  ctypes.byref(handle)))
            self._handle = handle
            check(lib.atmo_set_precision(handle, PRECISIONS[precision]), handle)
            c = config
            shapes = {
                _capi.BUF_TRANSMITTANCE: (c.transmittance_lut_width, c.transmittance_lut_height, 3),
                _capi.BUF_MULTI_SCATTERING: (c.multi_scattering_lut_width,
                                             c.multi_scattering_lut_height, 3),
                _capi.BUF_SKY_VIEW: (c.sky_view_lut_width, c.sky_view_lut_height, 3),
                _capi.BUF_AERIAL_SCATTERING: (c.aerial_lut_width, c.aerial_lut_height,
                                              c.aerial_lut_depth, 3),
                _capi.BUF_SURFACE_SUN_T: (self.width, self.height, 3),
                _capi.BUF_SURFACE_SKY_L: (self.width, self.height, 3),
            }
            fields = {}
            for which, shape in shapes.items():
                tensor = torch.zeros(shape, dtype=torch.float32, device=self.device)
                torch.cuda.synchronize(self.device)
                check(lib.atmo_bind_buffer(handle, which, tensor.data_ptr(),
                                           tensor.numel() * 4), handle)
                fields[which] = DeviceField(tensor)
        self.transmittance_lut = fields[_capi.BUF_TRANSMITTANCE]
        self.multi_scattering_lut = fields[_capi.BUF_MULTI_SCATTERING]
        self.sky_view_lut = fields[_capi.BUF_SKY_VIEW]
        self.aerial_scattering_lut = fields[_capi.BUF_AERIAL_SCATTERING]
        self.surface_sun_transmittance = fields[_capi.BUF_SURFACE_SUN_T]
        self.surface_sky_radiance = fields[_capi.BUF_SURFACE_SKY_L]
        # LUT invalidation state (renderer.py:97-116).
        self._planet_radius_key: float | None = None
        self._multi_scattering_key = None
        self._sky_key = None
        self.luts_frozen = False
        self.sky_snapshot_altitude_m: float | None = None
        self.sky_snapshot_sun_cosine: float | None = None
        self.aerial_snapshot_altitude_m: float | None = None
        self._aerial_available = False
        self.transmittance_rebuilds = 0
        self.multi_scattering_rebuilds = 0
        self.sky_view_rebuilds = 0
        self.aerial_rebuilds = 0

    # -------------------------------------------------------------- plumbing

    def close(self) -> None:
        handle = getattr(self, "_handle", None)
        if handle is not None and handle.value:
            torch = _torch()
            torch.cuda.synchronize(self.device)
            self._lib.atmo_destroy(handle)
            self._handle = None

    def __del__(self):  # pragma: no cover - best effort at interpreter exit
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self) -> ctypes.c_void_p:
        return self._handle

    def _stream(self) -> int:
        return _torch().cuda.current_stream(self.device).cuda_stream

    def _frame(self, value, name: str, channels: bool = True):
        torch = _torch()
        shape = (self.width, self.height, 3) if channels else (self.width, self.height)
        dtype = torch.float32 if channels else torch.int32
        return _device_tensor(value, name, shape, dtype, self.device)

    # --------------------------------------------------------------- kernels
    # One method per Taichi kernel, same arguments (renderer.py:1034-2055).

    def _build_transmittance(self, bottom_radius: float) -> None:
        check(self._lib.atmo_build_transmittance(self._handle, bottom_radius, self._stream()),
              self._handle)

    def _build_multi_scattering(self, bottom_radius, solar_irradiance, sun_angular_radius) -> None:
        check(self._lib.atmo_build_multi_scattering(
            self._handle, bottom_radius, f3(solar_irradiance), sun_angular_radius,
            self._stream()), self._handle)

    def _build_sky_view(self, bottom_radius, camera_altitude, sun_zenith_cosine,
                        solar_irradiance, sun_angular_radius) -> None:
        check(self._lib.atmo_build_sky_view(
            self._handle, bottom_radius, camera_altitude, sun_zenith_cosine,
            f3(solar_irradiance), sun_angular_radius, self._stream()), self._handle)

    def _build_aerial_perspective(self, bottom_radius, camera_altitude, sun, solar_irradiance,
                                  sun_angular_radius, right, view_up, forward,
                                  tangent_half_fov) -> None:
        check(self._lib.atmo_build_aerial_perspective(
            self._handle, bottom_radius, camera_altitude, f3(sun), f3(solar_irradiance),
            sun_angular_radius, f3(right), f3(view_up), f3(forward), tangent_half_fov,
            self._stream()), self._handle)

    def _prepare_surface_lighting(self, position_view, surface_id, bottom_radius,
                                  camera_altitude, sun, sun_angular_radius, right, view_up,
                                  forward) -> None:
        pos = self._frame(position_view, "position_view")
        sid = self._frame(surface_id, "surface_id", channels=False)
        check(self._lib.atmo_prepare_surface_lighting(
            self._handle, pos.data_ptr(), sid.data_ptr(), bottom_radius, camera_altitude,
            f3(sun), sun_angular_radius, f3(right), f3(view_up), f3(forward),
            self._stream()), self._handle)

    def _composite(self, scene_hdr, space_hdr, surface_id, position_view, hdr, bottom_radius,
                   camera_altitude, sky_lut_camera_altitude, sun_local, solar_irradiance,
                   sun_disk_radiance, sun_angular_radius, right, view_up, forward,
                   tangent_half_fov, star_contrast_start, star_contrast_end,
                   diagnostic_view) -> None:
        scene = self._frame(scene_hdr, "scene_hdr")
        space = self._frame(space_hdr, "space_hdr")
        sid = self._frame(surface_id, "surface_id", channels=False)
        pos = self._frame(position_view, "position_view")
        out = self._frame(hdr, "hdr")
        check(self._lib.atmo_composite(
            self._handle, scene.data_ptr(), space.data_ptr(), sid.data_ptr(), pos.data_ptr(),
            out.data_ptr(), bottom_radius, camera_altitude, sky_lut_camera_altitude,
            f3(sun_local), f3(solar_irradiance), f3(sun_disk_radiance), sun_angular_radius,
            f3(right), f3(view_up), f3(forward), tangent_half_fov, star_contrast_start,
            star_contrast_end, int(diagnostic_view), self._stream()), self._handle)

    def _composite_surrogate(self, surrogate, scene_hdr, space_hdr, surface_id, position_view,
                             hdr, bottom_radius, camera_altitude, sky_lut_camera_altitude,
                             sun_local, solar_irradiance, sun_disk_radiance, sun_angular_radius,
                             right, view_up, forward, tangent_half_fov, star_contrast_start,
                             star_contrast_end) -> None:
        """COMPOSITE view with the learned surface branch (atmo_composite_surrogate)."""
        scene = self._frame(scene_hdr, "scene_hdr")
        space = self._frame(space_hdr, "space_hdr")
        sid = self._frame(surface_id, "surface_id", channels=False)
        pos = self._frame(position_view, "position_view")
        out = self._frame(hdr, "hdr")
        check(self._lib.atmo_composite_surrogate(
            self._handle, surrogate.handle, scene.data_ptr(), space.data_ptr(), sid.data_ptr(),
            pos.data_ptr(), out.data_ptr(), bottom_radius, camera_altitude,
            sky_lut_camera_altitude, f3(sun_local), f3(solar_irradiance), f3(sun_disk_radiance),
            sun_angular_radius, f3(right), f3(view_up), f3(forward), tangent_half_fov,
            star_contrast_start, star_contrast_end, self._stream()), self._handle)

    def composite_queued_pixels(self) -> int:
        """Surface pixels the last COMPOSITE-view call left to the plain
        integrator's second pass (atmo_composite_queued): refinement pixels
        of the plain composite; out-of-domain and routed pixels of the
        surrogate composite.  Waits for the renderer's stream."""
        count = ctypes.c_int64(0)
        check(self._lib.atmo_composite_queued(self._handle, ctypes.byref(count), self._stream()),
              self._handle)
        return int(count.value)

    # ------------------------------------------------------------ autotuner

    def autotune(self, planet_radius_m: float, camera_radius_m: float, sun_local,
                 solar_irradiance, sun_angular_radius_degrees: float) -> tuple[int, int]:
        """Time the transmittance and sky-view launch shapes (lanes per
        texel 1, 2, 4, 8, 16) with this renderer's LUT sizes and keep the fastest
        (atmo_autotune).  Arguments as for `update`; the LUTs are rebuilt,
        so every cache key is reset.  Returns (lanes per T texel, lanes per
        sky-view texel)."""
        planet_radius = float(planet_radius_m)
        sun = np.asarray(sun_local, dtype=np.float64)
        sun = sun / max(float(np.linalg.norm(sun)), 1.0e-12)
        chosen = (ctypes.c_int32 * 2)()
        check(self._lib.atmo_autotune(
            self._handle, planet_radius, f3(solar_irradiance),
            float(np.float32(math.radians(sun_angular_radius_degrees))),
            float(np.float32(max(float(camera_radius_m) - planet_radius, 0.0))),
            float(np.float32(np.clip(sun[1], -1.0, 1.0))), self._stream(), chosen), self._handle)
        self._planet_radius_key = None
        self._multi_scattering_key = None
        self._sky_key = None
        return int(chosen[0]), int(chosen[1])

    @property
    def tuning(self) -> tuple[int, int]:
        t = (ctypes.c_int32 * 2)()
        check(self._lib.atmo_get_tuning(self._handle, t), self._handle)
        return int(t[0]), int(t[1])

    @tuning.setter
    def tuning(self, value) -> None:
        """Set (lanes per T texel, lanes per sky-view texel); the next
        `update` rebuilds every LUT with the new shapes."""
        t = (ctypes.c_int32 * 2)(*[int(v) for v in value])
        check(self._lib.atmo_set_tuning(self._handle, t), self._handle)
        self._planet_radius_key = None
        self._multi_scattering_key = None
        self._sky_key = None

    # ----------------------------------------------------------- host policy

    def set_luts_frozen(self, frozen: bool) -> None:
        """Freeze or resume view-dependent LUT regeneration; unfreezing
        invalidates sky-view (renderer.py:2057-2070)."""
        frozen = bool(frozen)
        if self.luts_frozen and not frozen:
            self._sky_key = None
        self.luts_frozen = frozen

    def update(self, planet_radius_m: float, camera_radius_m: float, sun_local,
               solar_irradiance, sun_angular_radius_degrees: float) -> None:
        """Rebuild only the LUTs invalidated by planet, irradiance, altitude or
        sun state (renderer.py:2072-2137)."""
        planet_radius = float(planet_radius_m)
        camera_radius = float(camera_radius_m)
        sun = np.asarray(sun_local, dtype=np.float64).copy()
        sun /= max(float(np.linalg.norm(sun)), 1.0e-12)
        sun_angular_radius = float(np.float32(math.radians(sun_angular_radius_degrees)))
        if self._planet_radius_key != planet_radius:
            self._build_transmittance(planet_radius)
            self._planet_radius_key = planet_radius
            self._multi_scattering_key = None
            self._sky_key = None
            self.transmittance_rebuilds += 1

        irradiance_key = tuple(float(value) for value in solar_irradiance)
        multi_scattering_key = (planet_radius, irradiance_key, sun_angular_radius)
        if multi_scattering_key != self._multi_scattering_key:
            self._build_multi_scattering(planet_radius, solar_irradiance, sun_angular_radius)
            self._multi_scattering_key = multi_scattering_key
            self._sky_key = None
            self.multi_scattering_rebuilds += 1

        # The altitude is formed in float64 before the float32 rounding, so a
        # sub-metre altitude keeps its precision at Earth radius.
        camera_altitude_key = float(np.float32(max(camera_radius - planet_radius, 0.0)))
        sun_zenith_key = float(np.float32(np.clip(sun[1], -1.0, 1.0)))
        sky_key = (camera_altitude_key, sun_zenith_key, irradiance_key + (sun_angular_radius,))
        if sky_key != self._sky_key and (not self.luts_frozen or self._sky_key is None):
            self._build_sky_view(planet_radius, camera_altitude_key, sun_zenith_key,
                                 solar_irradiance, sun_angular_radius)
            self._sky_key = sky_key
            self.sky_snapshot_altitude_m = camera_altitude_key
            self.sky_snapshot_sun_cosine = sun_zenith_key
            self.sky_view_rebuilds += 1

    def prepare_frame(self, position_view, surface_id, planet_radius_m: float,
                      camera_radius_m: float, sun_local, solar_irradiance,
                      sun_angular_radius_degrees: float, view_basis,
                      tangent_half_fov: float) -> None:
        """Surface lighting inputs, then the aerial-perspective LUT unless
        frozen (renderer.py:2139-2187)."""
        right, view_up, forward = view_basis
        camera_radius = float(camera_radius_m)
        planet_radius = float(planet_radius_m)
        camera_altitude = max(camera_radius - planet_radius, 0.0)
        sun = tuple(float(value) for value in sun_local)
        sun_angular_radius = math.radians(sun_angular_radius_degrees)
        self._prepare_surface_lighting(position_view, surface_id, planet_radius,
                                       camera_altitude, sun, sun_angular_radius, right,
                                       view_up, forward)
        if not self.luts_frozen or not self._aerial_available:
            self._build_aerial_perspective(planet_radius, camera_altitude, sun,
                                           solar_irradiance, sun_angular_radius, right,
                                           view_up, forward, float(tangent_half_fov))
            self.aerial_snapshot_altitude_m = camera_radius - planet_radius
            self._aerial_available = True
            self.aerial_rebuilds += 1

    def composite(self, scene_hdr, space_hdr, surface_id, position_view, hdr,
                  planet_radius_m: float, camera_radius_m: float, sun_local, solar_irradiance,
                  sun_disk_radiance, sun_angular_radius_degrees: float, view_basis,
                  tangent_half_fov: float, star_contrast_start: float,
                  star_contrast_end: float,
                  diagnostic_view: AtmosphereDiagnosticView = AtmosphereDiagnosticView.COMPOSITE,
                  surrogate=None) -> None:
        """Per-pixel composite into `hdr` (renderer.py:2189-2238).  Sky pixels
        sample sky-view at its snapshot altitude.  With a trained
        `surrogate` (surrogate.SurrogateMLP) the COMPOSITE view takes the
        surface in-scatter and transmittance from the network for in-domain
        pixels."""
        right, view_up, forward = view_basis
        camera_altitude = max(float(camera_radius_m) - float(planet_radius_m), 0.0)
        sky_lut_camera_altitude = camera_altitude
        if self.sky_snapshot_altitude_m is not None:
            sky_lut_camera_altitude = self.sky_snapshot_altitude_m
        if surrogate is not None and int(diagnostic_view) == int(AtmosphereDiagnosticView.COMPOSITE):
            _warn_if_below_gate(surrogate)
            self._composite_surrogate(surrogate, scene_hdr, space_hdr, surface_id, position_view,
                                      hdr, float(planet_radius_m), camera_altitude,
                                      sky_lut_camera_altitude, sun_local, solar_irradiance,
                                      sun_disk_radiance, math.radians(sun_angular_radius_degrees),
                                      right, view_up, forward, float(tangent_half_fov),
                                      float(star_contrast_start), float(star_contrast_end))
            return
        self._composite(scene_hdr, space_hdr, surface_id, position_view, hdr,
                        float(planet_radius_m), camera_altitude, sky_lut_camera_altitude,
                        sun_local, solar_irradiance, sun_disk_radiance,
                        math.radians(sun_angular_radius_degrees), right, view_up, forward,
                        float(tangent_half_fov), float(star_contrast_start),
                        float(star_contrast_end), int(diagnostic_view))


__all__ = ["AtmosphereRenderer", "DeviceField", "PRECISIONS"]
