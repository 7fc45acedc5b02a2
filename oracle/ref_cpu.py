"""ctypes wrapper of the CPU oracle (oracle/ref_cpu.c).  TEST INFRASTRUCTURE
ONLY -- imported by tests/, the smoke check and the CPU-baseline leg of
bench.py as the checker, never by the shipped package.

`RefCpu("f32")` is the float32 op-for-op restatement of the reference's
Taichi kernels; `RefCpu("f64")` the same approximation in float64;
`RefCpu("f32r")` the float32 build with the robust-precision islands
(float64 ray-sphere solves, Kahan optical depth), the checker of the GPU
library's robust mode.
`frame()` restates the host side of one reference frame
(src/planet_renderer/atmosphere/renderer.py:2072-2238): the argument
derivation and float32 rounding of `update` (forcing T, MS and SV
rebuilds), `prepare_frame` and `composite`.
"""

from __future__ import annotations

import ctypes
import math
import subprocess
from dataclasses import fields, is_dataclass
from pathlib import Path

import numpy as np

ORACLE_DIR = Path(__file__).resolve().parent
LIBS = {"f32": ORACLE_DIR / "libref_cpu_f32.so", "f64": ORACLE_DIR / "libref_cpu_f64.so",
        "f32r": ORACLE_DIR / "libref_cpu_f32r.so"}


class OracleParams(ctypes.Structure):
    """Mirror of struct AtmoParams (include/atmo.h)."""

    _fields_ = [
        ("top_altitude_m", ctypes.c_double),
        ("rayleigh_scattering_per_m", ctypes.c_double * 3),
        ("rayleigh_scale_height_m", ctypes.c_double),
        ("mie_scattering_per_m", ctypes.c_double * 3),
        ("mie_extinction_per_m", ctypes.c_double * 3),
        ("mie_scale_height_m", ctypes.c_double),
        ("mie_phase_g", ctypes.c_double),
        ("absorption_extinction_per_m", ctypes.c_double * 3),
        ("absorption_peak_altitude_m", ctypes.c_double),
        ("absorption_half_width_m", ctypes.c_double),
        ("ground_albedo", ctypes.c_double * 3),
        ("transmittance_lut_width", ctypes.c_int32),
        ("transmittance_lut_height", ctypes.c_int32),
        ("multi_scattering_lut_width", ctypes.c_int32),
        ("multi_scattering_lut_height", ctypes.c_int32),
        ("sky_view_lut_width", ctypes.c_int32),
        ("sky_view_lut_height", ctypes.c_int32),
        ("aerial_lut_width", ctypes.c_int32),
        ("aerial_lut_height", ctypes.c_int32),
        ("aerial_lut_depth", ctypes.c_int32),
        ("transmittance_steps", ctypes.c_int32),
        ("multi_scattering_directions", ctypes.c_int32),
        ("multi_scattering_steps", ctypes.c_int32),
        ("sky_view_steps", ctypes.c_int32),
        ("aerial_steps_per_slice", ctypes.c_int32),
        ("aerial_horizon_raymarch_steps", ctypes.c_int32),
        ("aerial_terminator_substeps", ctypes.c_int32),
    ]


def params_of(config) -> OracleParams:
    p = OracleParams()
    names = [f.name for f in fields(config)] if is_dataclass(config) else list(config)
    for name in names:
        value = getattr(config, name) if is_dataclass(config) else config[name]
        if isinstance(value, (tuple, list)):
            getattr(p, name)[:] = [float(v) for v in value]
        else:
            setattr(p, name, value)
    return p


def build(force: bool = False) -> None:
    """Compile both oracle libraries with oracle/Makefile."""
    args = ["make", "-s", "-C", str(ORACLE_DIR)]
    if force:
        args.append("-B")
    subprocess.run(args, check=True)


class RefCpu:
    def __init__(self, precision: str = "f32", path: Path | None = None):
        """`path` loads another build of the same source with the given
        precision (the tests' deliberately broken variants)."""
        path = LIBS[precision] if path is None else Path(path)
        if not path.exists():
            raise RuntimeError(f"{path} missing: run `make -C oracle`")
        self.precision = precision
        self.lib = ctypes.CDLL(str(path))
        single = precision in ("f32", "f32r")
        self.dtype = np.float32 if single else np.float64
        self.real = ctypes.c_float if single else ctypes.c_double
        assert self.lib.oracle_real_bytes() == np.dtype(self.dtype).itemsize

    # ---------------------------------------------------------------- utils

    def set_threads(self, n: int) -> None:
        self.lib.oracle_set_threads(int(n))

    def max_threads(self) -> int:
        return int(self.lib.oracle_max_threads())

    def r(self, value) -> "ctypes._SimpleCData":
        """A kernel scalar argument: float32-rounded in the f32 oracle."""
        return self.real(float(self.dtype(value)))

    def v3(self, values) -> np.ndarray:
        return np.ascontiguousarray(np.asarray(values, dtype=np.float64).astype(self.dtype))

    def ptr(self, array: np.ndarray):
        assert array.flags.c_contiguous
        return array.ctypes.data_as(ctypes.c_void_p)

    def zeros(self, shape) -> np.ndarray:
        return np.zeros(shape, dtype=self.dtype)

    # -------------------------------------------------------------- kernels

    def build_transmittance(self, config, bottom_radius) -> np.ndarray:
        out = self.zeros((config.transmittance_lut_width, config.transmittance_lut_height, 3))
        p = params_of(config)
        self.lib.oracle_build_transmittance(ctypes.byref(p), self.r(bottom_radius), self.ptr(out))
        return out

    def build_multi_scattering(self, config, bottom_radius, irradiance, angular_radius,
                               tlut) -> np.ndarray:
        out = self.zeros((config.multi_scattering_lut_width,
                          config.multi_scattering_lut_height, 3))
        p = params_of(config)
        e = self.v3(irradiance)
        self.lib.oracle_build_multi_scattering(ctypes.byref(p), self.r(bottom_radius),
                                               self.ptr(e), self.r(angular_radius),
                                               self.ptr(tlut), self.ptr(out))
        return out

    def build_sky_view(self, config, bottom_radius, camera_altitude, sun_zenith_cosine,
                       irradiance, angular_radius, tlut, mslut) -> np.ndarray:
        out = self.zeros((config.sky_view_lut_width, config.sky_view_lut_height, 3))
        p = params_of(config)
        e = self.v3(irradiance)
        self.lib.oracle_build_sky_view(ctypes.byref(p), self.r(bottom_radius),
                                       self.r(camera_altitude), self.r(sun_zenith_cosine),
                                       self.ptr(e), self.r(angular_radius), self.ptr(tlut),
                                       self.ptr(mslut), self.ptr(out))
        return out

    def build_aerial_perspective(self, config, width, height, bottom_radius, camera_altitude,
                                 sun, irradiance, angular_radius, right, up, forward,
                                 tangent_half_fov, tlut, mslut) -> np.ndarray:
        out = self.zeros((config.aerial_lut_width, config.aerial_lut_height,
                          config.aerial_lut_depth, 3))
        p = params_of(config)
        vecs = [self.v3(v) for v in (sun, irradiance, right, up, forward)]
        self.lib.oracle_build_aerial_perspective(
            ctypes.byref(p), int(width), int(height), self.r(bottom_radius),
            self.r(camera_altitude), self.ptr(vecs[0]), self.ptr(vecs[1]),
            self.r(angular_radius), self.ptr(vecs[2]), self.ptr(vecs[3]), self.ptr(vecs[4]),
            self.r(tangent_half_fov), self.ptr(tlut), self.ptr(mslut), self.ptr(out))
        return out

    def prepare_surface_lighting(self, config, position_view, surface_id, bottom_radius,
                                 camera_altitude, sun, angular_radius, right, up, forward,
                                 tlut, mslut):
        w, h = surface_id.shape
        pos = np.ascontiguousarray(position_view, dtype=self.dtype)
        sid = np.ascontiguousarray(surface_id, dtype=np.int32)
        sun_t = self.zeros((w, h, 3))
        sky_l = self.zeros((w, h, 3))
        p = params_of(config)
        vecs = [self.v3(v) for v in (sun, right, up, forward)]
        self.lib.oracle_prepare_surface_lighting(
            ctypes.byref(p), w, h, self.ptr(pos), self.ptr(sid), self.r(bottom_radius),
            self.r(camera_altitude), self.ptr(vecs[0]), self.r(angular_radius),
            self.ptr(vecs[1]), self.ptr(vecs[2]), self.ptr(vecs[3]), self.ptr(tlut),
            self.ptr(mslut), self.ptr(sun_t), self.ptr(sky_l))
        return sun_t, sky_l

    def composite(self, config, scene_hdr, space_hdr, surface_id, position_view, bottom_radius,
                  camera_altitude, sky_altitude, sun, irradiance, disk, angular_radius, right,
                  up, forward, tangent_half_fov, contrast_start, contrast_end, view, tlut,
                  mslut, svlut, aplut):
        w, h = surface_id.shape
        arrays = [np.ascontiguousarray(a, dtype=self.dtype)
                  for a in (scene_hdr, space_hdr, position_view)]
        sid = np.ascontiguousarray(surface_id, dtype=np.int32)
        hdr = self.zeros((w, h, 3))
        p = params_of(config)
        vecs = [self.v3(v) for v in (sun, irradiance, disk, right, up, forward)]
        refined = ctypes.c_longlong(0)
        self.lib.oracle_composite(
            ctypes.byref(p), w, h, self.ptr(arrays[0]), self.ptr(arrays[1]), self.ptr(sid),
            self.ptr(arrays[2]), self.ptr(hdr), self.r(bottom_radius), self.r(camera_altitude),
            self.r(sky_altitude), self.ptr(vecs[0]), self.ptr(vecs[1]), self.ptr(vecs[2]),
            self.r(angular_radius), self.ptr(vecs[3]), self.ptr(vecs[4]), self.ptr(vecs[5]),
            self.r(tangent_half_fov), self.r(contrast_start), self.r(contrast_end), int(view),
            self.ptr(tlut), self.ptr(mslut), self.ptr(svlut), self.ptr(aplut),
            ctypes.byref(refined))
        return hdr, int(refined.value)

    def surface_path(self, config, altitude, rays, distance, suns, irradiance, bottom_radius,
                     angular_radius, tlut, mslut):
        """(in-scatter, transmittance) of the composite's surface branch for
        arbitrary per-sample inputs."""
        n = int(np.asarray(altitude).shape[0])
        alt = np.ascontiguousarray(altitude, dtype=self.dtype)
        ray = np.ascontiguousarray(rays, dtype=self.dtype)
        dist = np.ascontiguousarray(distance, dtype=self.dtype)
        sun = np.ascontiguousarray(suns, dtype=self.dtype)
        e = self.v3(irradiance)
        lo = self.zeros((n, 3))
        to = self.zeros((n, 3))
        p = params_of(config)
        self.lib.oracle_surface_path(ctypes.byref(p), ctypes.c_long(n), self.ptr(alt),
                                     self.ptr(ray), self.ptr(dist), self.ptr(sun), self.ptr(e),
                                     self.r(bottom_radius), self.r(angular_radius),
                                     self.ptr(tlut), self.ptr(mslut), self.ptr(lo), self.ptr(to))
        return lo, to

    # ------------------------------------------------------ one whole frame

    def frame(self, inputs, view: int = 0, build_aerial: bool = True) -> dict:
        """update (T, MS, SV forced) + prepare_frame + composite of the
        reference host methods on `inputs` (scene.FrameInputs)."""
        cfg = inputs.config
        R = float(inputs.planet_radius_m)
        r_cam = float(inputs.camera_radius_m)
        E = tuple(float(v) for v in inputs.lighting.solar_irradiance)
        deg = float(inputs.lighting.sun_angular_radius_degrees)
        sun_local = np.asarray(inputs.sun_local, dtype=np.float64)
        # update (renderer.py:2082-2137)
        sun_n = sun_local / max(float(np.linalg.norm(sun_local)), 1.0e-12)
        ar_update = float(np.float32(math.radians(deg)))
        alt_key = float(np.float32(max(r_cam - R, 0.0)))
        szc_key = float(np.float32(np.clip(sun_n[1], -1.0, 1.0)))
        tlut = self.build_transmittance(cfg, R)
        mslut = self.build_multi_scattering(cfg, R, E, ar_update, tlut)
        svlut = self.build_sky_view(cfg, R, alt_key, szc_key, E, ar_update, tlut, mslut)
        # prepare_frame (renderer.py:2153-2187)
        right, up, forward = inputs.view_basis
        cam_alt = max(r_cam - R, 0.0)
        ar = math.radians(deg)
        sun_t, sky_l = self.prepare_surface_lighting(cfg, inputs.position_view,
                                                     inputs.surface_id, R, cam_alt, sun_local,
                                                     ar, right, up, forward, tlut, mslut)
        aplut = (self.build_aerial_perspective(cfg, inputs.width, inputs.height, R, cam_alt,
                                               sun_local, E, ar, right, up, forward,
                                               inputs.tangent_half_fov, tlut, mslut)
                 if build_aerial else self.zeros((cfg.aerial_lut_width, cfg.aerial_lut_height,
                                                  cfg.aerial_lut_depth, 3)))
        # composite (renderer.py:2210-2238); sky-view at its snapshot altitude
        disk = inputs.lighting.sun_disk_radiance
        c0, c1 = inputs.star_contrast
        hdr, refined = self.composite(cfg, inputs.scene_hdr, inputs.space_hdr,
                                      inputs.surface_id, inputs.position_view, R, cam_alt,
                                      alt_key, sun_local, E, disk, ar, right, up, forward,
                                      inputs.tangent_half_fov, c0, c1, view, tlut, mslut,
                                      svlut, aplut)
        return {
            "transmittance": tlut,
            "multi_scattering": mslut,
            "sky_view": svlut,
            "aerial": aplut,
            "surface_sun_transmittance": sun_t,
            "surface_sky_radiance": sky_l,
            "hdr": hdr,
            "refined_intervals": refined,
        }


__all__ = ["OracleParams", "RefCpu", "build", "params_of"]
