"""Benchmark of the atmosphere hot path on one MI355X.

One step = one 1920x1080 atmosphere frame of the full-frame configuration:
Transmittance, Multi-scattering and Sky-view LUT rebuilds (forced, as a
cold frame), the surface-lighting pass and the Aerial-perspective LUT
(`prepare_frame`), and the per-pixel composite (COMPOSITE view), all through
libatmo.so on one HIP stream with the G-buffer already resident in HBM.

Prints ONE JSON line on stdout; progress goes to stderr.
"""

from __future__ import annotations

import argparse
import ctypes
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

REPO = Path(__file__).resolve().parent
sys.path.insert(0, str(REPO))

import __graft_entry__ as entry  # noqa: E402

METRIC = "1080p atmosphere frame ms (LUTs+composite): surrogate vs plain @ p99 rel err<=1%"
FP32_PEAK_TFLOPS = 157.3
HBM_PEAK_GBS = 8000.0

# Algorithmic work per unit (FP32 ops F, transcendental-class ops T), the
# hand count of the reference's source per stage; see DESIGN.md.
COST = {
    "surface_pixel": (6746.0, 693.0),
    "sky_pixel": (328.0, 32.0),
    "refined_interval": (2946.0, 336.0),
}
# Composite bytes per pixel: scene, space, position (12 B each) + id (4 B)
# read, hdr (12 B) written -- every pixel reads all inputs in the reference.
COMPOSITE_BYTES_PER_PIXEL = 12 + 12 + 12 + 4 + 12


def log(*args):
    print(*args, file=sys.stderr, flush=True)


def parse_args():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true",
                    help="skip the CPU oracle leg (no cpu_baseline, no refined-interval count)")
    return ap.parse_args()


class PlainFrame:
    """One cold frame through the facade's kernel methods, with the argument
    derivation of the reference's update / prepare_frame / composite."""

    def __init__(self, pkg, inputs, device):
        import torch

        self.torch = torch
        self.inputs = inputs
        self.r = pkg.AtmosphereRenderer(inputs.width, inputs.height, inputs.config, device)
        to_dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(device)
        self.pos = to_dev(inputs.position_view)
        self.sid = to_dev(inputs.surface_id)
        self.scene = to_dev(inputs.scene_hdr)
        self.space = to_dev(inputs.space_hdr)
        self.hdr = torch.zeros_like(self.scene)
        light = inputs.lighting
        R, r_cam = inputs.planet_radius_m, inputs.camera_radius_m
        sun = np.asarray(inputs.sun_local, dtype=np.float64)
        sun_n = sun / max(float(np.linalg.norm(sun)), 1e-12)
        self.R = R
        self.E = light.solar_irradiance
        self.ar_update = float(np.float32(math.radians(light.sun_angular_radius_degrees)))
        self.alt_key = float(np.float32(max(r_cam - R, 0.0)))
        self.szc_key = float(np.float32(np.clip(sun_n[1], -1.0, 1.0)))
        self.cam_alt = max(r_cam - R, 0.0)
        self.sun = tuple(float(v) for v in sun)
        self.ar = math.radians(light.sun_angular_radius_degrees)
        self.basis = inputs.view_basis
        self.disk = light.sun_disk_radiance

    def stages(self):
        r, i = self.r, self.inputs
        right, up, forward = self.basis
        c0, c1 = i.star_contrast
        return [
            ("transmittance", lambda: r._build_transmittance(self.R)),
            ("multi_scattering", lambda: r._build_multi_scattering(self.R, self.E, self.ar_update)),
            ("sky_view", lambda: r._build_sky_view(self.R, self.alt_key, self.szc_key, self.E,
                                                   self.ar_update)),
            ("surface_lighting", lambda: r._prepare_surface_lighting(
                self.pos, self.sid, self.R, self.cam_alt, self.sun, self.ar, right, up, forward)),
            ("aerial_perspective", lambda: r._build_aerial_perspective(
                self.R, self.cam_alt, self.sun, self.E, self.ar, right, up, forward,
                i.tangent_half_fov)),
            ("composite", lambda: r._composite(
                self.scene, self.space, self.sid, self.pos, self.hdr, self.R, self.cam_alt,
                self.alt_key, self.sun, self.E, self.disk, self.ar, right, up, forward,
                i.tangent_half_fov, c0, c1, 0)),
        ]


def probe_exp_rate() -> float | None:
    path = REPO / "planet-real-time-render_amd" / "libatmo_probe.so"
    if not path.exists():
        return None
    lib = ctypes.CDLL(str(path))
    lib.atmo_probe_exp_rate.argtypes = [ctypes.c_int32, ctypes.c_int32,
                                        ctypes.POINTER(ctypes.c_double)]
    lib.atmo_probe_exp_rate.restype = ctypes.c_int32
    rate = ctypes.c_double(0.0)
    status = lib.atmo_probe_exp_rate(256 * 32, 4096, ctypes.byref(rate))
    return rate.value if status == 0 else None


def main() -> int:
    args = parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the atmosphere frame is a single-device path).", file=sys.stderr)
        return 2
    import torch

    if not torch.cuda.is_available():
        print("bench.py needs a HIP device", file=sys.stderr)
        return 2
    pkg = entry.load_package()
    scene = entry.load_submodule("scene")
    device = torch.device("cuda:0")
    inputs = scene.make_frame(width=args.width, height=args.height, config_index=1,
                              seed_offset=args.seed)
    log(f"workload {inputs.width}x{inputs.height}, camera {inputs.camera_altitude_m:.1f} m, "
        f"sun y {inputs.sun_local[1]:.3f}")
    frame = PlainFrame(pkg, inputs, device)
    stages = frame.stages()
    stream = torch.cuda.current_stream(device)

    for _ in range(args.warmup):
        for _, fn in stages:
            fn()
    torch.cuda.synchronize(device)
    log(f"warmup {args.warmup} frames done")

    n_st = len(stages)
    events = [[torch.cuda.Event(enable_timing=True) for _ in range(n_st + 1)]
              for _ in range(args.steps)]
    torch.cuda.synchronize(device)
    t0 = time.perf_counter()
    for k in range(args.steps):
        ev = events[k]
        ev[0].record(stream)
        for j, (_, fn) in enumerate(stages):
            fn()
            ev[j + 1].record(stream)
    torch.cuda.synchronize(device)
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    frame_ms = np.array([events[k][0].elapsed_time(events[k][n_st]) for k in range(args.steps)])
    stage_ms = {name: float(np.mean([events[k][j].elapsed_time(events[k][j + 1])
                                     for k in range(args.steps)]))
                for j, (name, _) in enumerate(stages)}
    log(f"timed {args.steps} frames: {ms_per_step:.4f} ms/frame (wall), "
        f"median {np.median(frame_ms):.4f} ms (events)")
    for name, ms in stage_ms.items():
        log(f"  {name:20s} {ms * 1e3:9.2f} us")

    n_pix = inputs.width * inputs.height
    n_sky = int((inputs.surface_id < 0).sum())
    n_surface = n_pix - n_sky

    # CPU baseline: the float32 oracle on this host's cores, same frame.
    cpu = None
    refined = None
    parity = None
    if not args.no_cpu_baseline:
        from oracle.ref_cpu import RefCpu

        oracle = RefCpu("f32")
        threads = oracle.max_threads()
        log(f"cpu baseline: float32 oracle, {threads} threads, one full frame")
        c0 = time.perf_counter()
        ref = oracle.frame(inputs)
        cpu_ms = (time.perf_counter() - c0) * 1e3
        refined = ref["refined_intervals"]
        got = frame.hdr.cpu().numpy().astype(np.float64)
        want = ref["hdr"].astype(np.float64)
        floor = 1e-6 * np.abs(want).reshape(-1, 3).max(axis=0)
        rel = np.abs(got - want) / np.maximum(np.abs(want), floor)
        parity = {"hdr_max_rel_err_vs_oracle_f32": float(rel.max()),
                  "hdr_p99_rel_err_vs_oracle_f32": float(np.percentile(rel, 99))}
        # one-thread sample: the first 16 pixel columns, scaled to the frame
        oracle.set_threads(1)
        sub = 16
        sample = scene.make_frame(width=args.width, height=args.height, config_index=1,
                                  seed_offset=args.seed)
        for name in ("position_view", "surface_id", "scene_hdr", "space_hdr"):
            setattr(sample, name, np.ascontiguousarray(getattr(sample, name)[:sub]))
        s0 = time.perf_counter()
        oracle.composite(sample.config, sample.scene_hdr, sample.space_hdr, sample.surface_id,
                         sample.position_view, sample.planet_radius_m, frame.cam_alt,
                         frame.alt_key, frame.sun, frame.E, frame.disk, frame.ar,
                         *frame.basis, sample.tangent_half_fov, *sample.star_contrast, 0,
                         ref["transmittance"], ref["multi_scattering"], ref["sky_view"],
                         ref["aerial"])
        one_thread_ms = (time.perf_counter() - s0) * 1e3 * inputs.width / sub
        oracle.set_threads(threads)
        cpu = {
            "value": round(cpu_ms, 3),
            "unit": "ms/frame",
            "cores": threads,
            "kind": "port",
            "sample": (f"one full {inputs.width}x{inputs.height} frame (4 LUT builds + surface "
                       f"lighting + composite) of the float32 C restatement, OpenMP {threads} "
                       f"threads; 1-thread composite {one_thread_ms:.0f} ms/frame extrapolated "
                       f"from {sub} pixel columns"),
        }
        log(f"cpu baseline {cpu_ms:.1f} ms/frame; refined intervals {refined}; "
            f"hdr max rel err vs oracle {parity['hdr_max_rel_err_vs_oracle_f32']:.2e}")

    # Roofline of the dominant kernel (composite).
    comp_s = stage_ms["composite"] * 1e-3
    n_ref = refined if refined is not None else 0
    flops = (COST["surface_pixel"][0] * n_surface + COST["sky_pixel"][0] * n_sky
             + COST["refined_interval"][0] * n_ref)
    tops = (COST["surface_pixel"][1] * n_surface + COST["sky_pixel"][1] * n_sky
            + COST["refined_interval"][1] * n_ref)
    achieved_tflops = flops / comp_s / 1e12
    p_t = probe_exp_rate()
    bytes_per_launch = COMPOSITE_BYTES_PER_PIXEL * n_pix
    roofline = {
        "bound": "valu",
        "achieved": round(achieved_tflops, 3),
        "peak": FP32_PEAK_TFLOPS,
        "unit": "TFLOP/s",
        "frac": round(achieved_tflops / FP32_PEAK_TFLOPS, 4),
        "traffic": None,
        "kernel": "k_composite",
        "kernel_ms": round(stage_ms["composite"], 4),
        "algorithmic_flop": flops,
        "algorithmic_transcendental": tops,
        "refined_intervals": refined,
        "transcendental_rate": round(tops / comp_s / 1e12, 4),
        "transcendental_peak": round(p_t / 1e12, 4) if p_t else None,
        "transcendental_frac": round(tops / comp_s / p_t, 4) if p_t else None,
        "hbm_gbs": round(bytes_per_launch / comp_s / 1e9, 2),
        "hbm_frac": round(bytes_per_launch / comp_s / 1e9 / HBM_PEAK_GBS, 5),
    }
    result = {
        "metric": METRIC,
        "value": round(ms_per_step, 4),
        "unit": "ms/frame",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": (f"{inputs.width}x{inputs.height} atmosphere frame, plain path: T 256x64/40, "
                         "MS 32x32 64 dirs x 20 steps, SV 192x108/30, AP 32x32x32, surface "
                         "lighting, composite; 30% sky pixels"),
            "mode": "plain",
            "frame_ms_median": round(float(np.median(frame_ms)), 4),
            "frame_ms_p99": round(float(np.percentile(frame_ms, 99)), 4),
            "stage_us": {k: round(v * 1e3, 2) for k, v in stage_ms.items()},
            "surface_pixels": n_surface,
            "sky_pixels": n_sky,
            "camera_altitude_m": round(inputs.camera_altitude_m, 3),
            "sun_zenith_cos": round(float(inputs.sun_local[1]), 5),
            "parallelism": "single GPU",
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
        "parity": parity,
    }
    frame.r.close()
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
