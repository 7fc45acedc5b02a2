"""Learned surrogate of the composite's surface branch.

The plain composite integrates, per surface pixel, the direct limb
in-scatter (renderer.py:1667-1786) and the two-sample segment transmittance
(renderer.py:1603-1665).  For a ray from the camera on the local vertical,
both depend only on (camera altitude h, view-zenith cosine mu, sun-zenith
cosine mu_s, relative azimuth phi, segment length d), and the in-scatter is
linear in the solar irradiance (the multi-scattering LUT is too), so a
5 -> 64 -> 64 -> 64 -> 6 SiLU MLP learns (L / E, T) once per atmosphere.

Everything numeric runs in libatmo.so: labels from the GPU integrator
(atmo_surrogate_labels), FP32 forward / backward / Adam training kernels,
BF16 MFMA inference.  This module drives them and handles the target
transform and the weight file (raw float32 parameters + JSON metadata).

Default transform: the view and sun cosines are encoded relative to the
local horizon and asinh-stretched, the in-scatter is learned as
log(max(L, 1e-7)) (relative accuracy wherever L can show in a pixel; the
night side collapses onto the floor), the transmittance as log(T + 1e-6),
and the surrogate composite evaluates the transmittance exactly with the
plain two-sample formula (about 4 % of the plain surface cost) because the
reference's reversed-ray transmittance has kinks a 64-wide network does not
fit to 1 %.
"""

from __future__ import annotations

import ctypes
import json
import math
import time
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from . import _capi
from ._capi import AtmoSurrogateTransform, library

ALTITUDE_MAX_M = 100_000.0
ALTITUDE_SCALE_M = 100.0
DISTANCE_RANGE_M = (1.0, 1.0e6)
# Loss weighting: the composite's output is E * L + scene * T, so the
# relative error of a pixel is about w_L dlog L + w_T dlog T with the shares
# w_L = E L / (E L + s T), w_T = s T / (E L + s T).  E is the solar
# irradiance and s a representative scene radiance.
WEIGHT_IRRADIANCE = (4.0, 3.9, 3.7)
WEIGHT_SCENE = 0.5
WEIGHT_FLOOR = 0.01
# A pixel's radiance never counts as less than this in the relative-error
# weights (a twilight pixel with neither in-scatter nor a lit scene).
WEIGHT_PIXEL_FLOOR = 5.0e-3
WEIGHT_CLIP = 1.0e3
EPS_INSCATTER = 1.0e-8
EPS_TRANSMITTANCE = 1.0e-6
# floor of the floored-log in-scatter target: 1e-7 per unit irradiance is
# below 1e-6 of a daylight pixel
INSCATTER_FLOOR = 1.0e-7
COSINE_STRETCH = 0.02
# Routing: a surface pixel whose learned in-scatter E L exceeds this share of
# the pixel (E L + scene T) in some channel takes the plain integrator, so the
# network's relative in-scatter error shows in no pixel it carries.  With
# 0.02 the shipped network's hdr p99 relative error against the float64
# restatement is 7.1e-3 on the seed-0 1920x1080 full frame (30 % of surface
# pixels routed) and 3.0e-3 to 4.2e-3 on seeds 1-3 at 960x540; 0 routes
# nothing on this criterion.
ROUTE_SHARE = 0.02
OUTPUT_LOG = 0
OUTPUT_LINEAR = 1
OUTPUT_LOG_FLOOR = 2
SEED_BASE = 20261203
MODE_NAMES = {OUTPUT_LOG: "log", OUTPUT_LINEAR: "linear", OUTPUT_LOG_FLOOR: "log_floor"}
WEIGHTS_DIR = Path(__file__).resolve().parent / "weights"


def default_weights_path() -> Path:
    """The shipped network for the full-frame atmosphere (Earth radius
    6360 km, 100 km shell with ozone, 0.266 deg sun): <path>.bin + .json."""
    return WEIGHTS_DIR / "earth_aerial_surrogate"


def _torch():
    import torch

    return torch


def sample_inputs(n: int, seed: int) -> np.ndarray:
    """(n, 5) float32 raw samples over the training domain: h in [0, 100 km]
    (half U[0, 100 km], half logU[1 m, 100 km], so low cameras are as well
    covered as high ones), mu ~ U[-1, 1], mu_s ~ U[-1, 1], phi ~ U[0, pi],
    d ~ logU[1 m, 1000 km]."""
    g = np.random.Generator(np.random.PCG64(seed))
    lo, hi = DISTANCE_RANGE_M
    raw = np.empty((n, 5), dtype=np.float64)
    half = n // 2
    raw[:half, 0] = g.uniform(0.0, ALTITUDE_MAX_M, half)
    raw[half:, 0] = np.exp(g.uniform(0.0, math.log(ALTITUDE_MAX_M), n - half))
    raw[:, 1] = g.uniform(-1.0, 1.0, n)
    raw[:, 2] = g.uniform(-1.0, 1.0, n)
    raw[:, 3] = g.uniform(0.0, math.pi, n)
    raw[:, 4] = np.exp(g.uniform(math.log(lo), math.log(hi), n))
    return raw[g.permutation(n)].astype(np.float32)


def targets_from_labels(labels: np.ndarray, mode: int = OUTPUT_LOG_FLOOR) -> np.ndarray:
    """Training targets before standardisation.  Linear: the labels
    themselves (in-scatter >= 0, transmittance clamped to [0, 1]).  Log:
    log(L + eps) and log(T + eps).  Floored log: log(max(L, floor)) and
    log(T + eps)."""
    labels = np.asarray(labels, dtype=np.float64)
    out = np.empty_like(labels)
    if mode == OUTPUT_LINEAR:
        out[:, :3] = np.maximum(labels[:, :3], 0.0)
        out[:, 3:] = np.clip(labels[:, 3:], 0.0, 1.0)
        return out
    if mode == OUTPUT_LOG_FLOOR:
        out[:, :3] = np.log(np.maximum(labels[:, :3], INSCATTER_FLOOR))
    else:
        out[:, :3] = np.log(np.maximum(labels[:, :3], 0.0) + EPS_INSCATTER)
    out[:, 3:] = np.log(np.clip(labels[:, 3:], 0.0, 1.0) + EPS_TRANSMITTANCE)
    return out


def decode_outputs(raw: np.ndarray, mean: np.ndarray, scale: np.ndarray,
                   mode: int = OUTPUT_LOG_FLOOR):
    """Network outputs -> (L per unit irradiance, T), the inverse of
    targets_from_labels after standardisation (host mirror of the kernel's
    decode)."""
    y = np.asarray(raw, dtype=np.float64) * scale + mean
    if mode == OUTPUT_LINEAR:
        return np.maximum(y[:, :3], 0.0), np.clip(y[:, 3:], 0.0, 1.0)
    e = np.exp(y[:, :3])
    if mode == OUTPUT_LOG_FLOOR:
        L = np.where(e < 1.5 * INSCATTER_FLOOR, 0.0, e)
    else:
        L = np.maximum(e - EPS_INSCATTER, 0.0)
    T = np.clip(np.exp(y[:, 3:]) - EPS_TRANSMITTANCE, 0.0, 1.0)
    return L, T


def loss_weights(labels: np.ndarray, scale: np.ndarray, mode: int = OUTPUT_LINEAR) -> np.ndarray:
    """Per-output weights of the standardised MSE, normalised to mean 1.

    The composite pixel of channel c is P = E_c L_c + s T_c.  In linear mode
    an error of dy standardised units in L_c moves P by E_c sigma dy, so the
    weight (E_c sigma / P)^2 turns the loss into the squared relative pixel
    error (likewise s sigma / P for T_c); P is floored at
    WEIGHT_PIXEL_FLOOR and weights are clipped to WEIGHT_CLIP x their
    median.  In log mode the weight is (share * sigma)^2 with the shares
    E L / P and s T / P, floored at WEIGHT_FLOOR."""
    labels = np.asarray(labels, dtype=np.float64)
    scale = np.asarray(scale, dtype=np.float64)
    el = np.asarray(WEIGHT_IRRADIANCE) * np.maximum(labels[:, :3], 0.0)
    st = WEIGHT_SCENE * np.clip(labels[:, 3:], 0.0, 1.0)
    total = np.maximum(el + st, 1e-30)
    if mode == OUTPUT_LINEAR:
        pixel = np.maximum(total, WEIGHT_PIXEL_FLOOR)
        coef = np.concatenate([np.broadcast_to(np.asarray(WEIGHT_IRRADIANCE), el.shape),
                               np.full(st.shape, WEIGHT_SCENE)], axis=1)
        w = (coef * scale / np.concatenate([pixel, pixel], axis=1)) ** 2
        w = np.minimum(w, WEIGHT_CLIP * np.median(w, axis=0))
    else:
        share = np.concatenate([el / total, st / total], axis=1)
        w = (np.maximum(share, WEIGHT_FLOOR) * scale) ** 2
    return (w / w.mean()).astype(np.float32)


def horizon_cosine(planet_radius_m: float, altitude_m):
    """Cosine of the local horizon's zenith angle at altitude h:
    -sqrt(h (2R + h)) / (R + h)."""
    h = np.maximum(np.asarray(altitude_m, dtype=np.float64), 0.0)
    return -np.sqrt(h * (2.0 * planet_radius_m + h)) / (planet_radius_m + h)


def encode_host(raw: np.ndarray, planet_radius_m: float = 6_360_000.0,
                cosine_stretch: float = COSINE_STRETCH) -> np.ndarray:
    """Host mirror of the kernel's input normalisation (float64)."""
    raw = np.asarray(raw, dtype=np.float64)
    lo, hi = math.log(DISTANCE_RANGE_M[0]), math.log(DISTANCE_RANGE_M[1])
    x = np.empty_like(raw)
    x[:, 0] = np.log1p(np.maximum(raw[:, 0], 0.0) / ALTITUDE_SCALE_M) * (
        2.0 / math.log1p(ALTITUDE_MAX_M / ALTITUDE_SCALE_M)) - 1.0
    if cosine_stretch > 0.0:
        ch = horizon_cosine(planet_radius_m, raw[:, 0])
        norm = 1.0 / math.asinh(2.0 / cosine_stretch)
        x[:, 1] = np.arcsinh((raw[:, 1] - ch) / cosine_stretch) * norm
        x[:, 2] = np.arcsinh((raw[:, 2] - ch) / cosine_stretch) * norm
    else:
        x[:, 1] = raw[:, 1]
        x[:, 2] = raw[:, 2]
    x[:, 3] = raw[:, 3] * (2.0 / math.pi) - 1.0
    x[:, 4] = (np.log(raw[:, 4]) - lo) * (2.0 / (hi - lo)) - 1.0
    return x


def init_params(seed: int = 0) -> np.ndarray:
    """He-uniform weights, zero biases, in the flat layout of include/atmo.h."""
    g = np.random.default_rng(seed)
    parts = []
    for fan_in, fan_out in ((5, 64), (64, 64), (64, 64), (64, 6)):
        bound = math.sqrt(6.0 / fan_in)
        parts.append(g.uniform(-bound, bound, fan_in * fan_out))
        parts.append(np.zeros(fan_out))
    return np.concatenate(parts).astype(np.float32)


@dataclass
class TrainingReport:
    samples: int
    batch: int
    epochs: int
    steps: int
    lr: float
    seconds: float
    label_seconds: float
    losses: list = field(default_factory=list)


class SurrogateMLP:
    """Device-resident network (AtmoMlp handle of libatmo.so)."""

    def __init__(self, max_batch: int = 65536, device=None):
        torch = _torch()
        self.device = torch.device(device if device is not None else "cuda")
        self._lib = library()
        handle = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            status = self._lib.atmo_mlp_create(int(max_batch), ctypes.byref(handle))
        if status != _capi.ATMO_OK:
            raise _capi.AtmoError(f"atmo_mlp_create failed with status {status}")
        self._handle = handle
        self.max_batch = int(max_batch)
        self.out_mean = np.zeros(6)
        self.out_scale = np.ones(6)
        self.output_mode = OUTPUT_LOG_FLOOR
        self.planet_radius_m = 6_360_000.0
        self.cosine_stretch = COSINE_STRETCH
        self.exact_transmittance = True
        self.route_share = 0.0
        self.metadata: dict = {}
        self.set_params(init_params(0))
        self._push_transform()

    # ---------------------------------------------------------------- basics

    @property
    def handle(self) -> ctypes.c_void_p:
        return self._handle

    def close(self) -> None:
        if getattr(self, "_handle", None) is not None and self._handle.value:
            self._lib.atmo_mlp_destroy(self._handle)
            self._handle = None

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass

    def _check(self, status: int) -> None:
        if status != _capi.ATMO_OK:
            message = self._lib.atmo_mlp_last_error(self._handle).decode(errors="replace")
            raise (ValueError if status in (_capi.ATMO_EINVAL, _capi.ATMO_ESHAPE)
                   else _capi.AtmoError)(f"surrogate status {status}: {message}")

    def _stream(self) -> int:
        return _torch().cuda.current_stream(self.device).cuda_stream

    def transform(self) -> AtmoSurrogateTransform:
        t = AtmoSurrogateTransform()
        t.altitude_max_m = ALTITUDE_MAX_M
        t.altitude_scale_m = ALTITUDE_SCALE_M
        t.distance_min_m, t.distance_max_m = DISTANCE_RANGE_M
        t.out_mean[:] = [float(v) for v in self.out_mean]
        t.out_scale[:] = [float(v) for v in self.out_scale]
        t.eps_inscatter = INSCATTER_FLOOR if self.output_mode == OUTPUT_LOG_FLOOR else EPS_INSCATTER
        t.eps_transmittance = EPS_TRANSMITTANCE
        t.output_mode = int(self.output_mode)
        t.planet_radius_m = float(self.planet_radius_m)
        t.cosine_stretch = float(self.cosine_stretch)
        t.exact_transmittance = 1 if self.exact_transmittance else 0
        t.route_share = float(self.route_share)
        return t

    def _push_transform(self) -> None:
        t = self.transform()
        self._check(self._lib.atmo_mlp_set_transform(self._handle, ctypes.byref(t)))

    def set_params(self, params: np.ndarray) -> None:
        p = np.ascontiguousarray(params, dtype=np.float32)
        if p.shape != (_capi.MLP_PARAMS,):
            raise ValueError(f"expected {_capi.MLP_PARAMS} parameters, got {p.shape}")
        self._check(self._lib.atmo_mlp_set_params(
            self._handle, p.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), p.size))

    def get_params(self) -> np.ndarray:
        _torch().cuda.synchronize(self.device)
        p = np.zeros(_capi.MLP_PARAMS, dtype=np.float32)
        self._check(self._lib.atmo_mlp_get_params(
            self._handle, p.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), p.size))
        return p

    def set_output_transform(self, mean, scale, mode: int | None = None) -> None:
        self.out_mean = np.asarray(mean, dtype=np.float64)
        self.out_scale = np.asarray(scale, dtype=np.float64)
        if mode is not None:
            if mode not in (OUTPUT_LOG, OUTPUT_LINEAR, OUTPUT_LOG_FLOOR):
                raise ValueError(f"unknown output mode {mode}")
            self.output_mode = int(mode)
        self._push_transform()

    def set_encoding(self, planet_radius_m: float, cosine_stretch: float) -> None:
        """Planet radius of the horizon-relative cosine encoding and its
        stretch (0 = raw cosines)."""
        self.planet_radius_m = float(planet_radius_m)
        self.cosine_stretch = float(cosine_stretch)
        self._push_transform()

    def set_exact_transmittance(self, exact: bool) -> None:
        """True: the surrogate composite evaluates the transmittance with
        the plain two-sample formula; False: it uses the network's."""
        self.exact_transmittance = bool(exact)
        self._push_transform()

    def set_route_share(self, share: float) -> None:
        """Route to the plain integrator every surface pixel whose learned
        in-scatter exceeds `share` of the pixel in some channel (0 <= share
        < 1; 0 routes none on this criterion)."""
        share = float(share)
        if not 0.0 <= share < 1.0:
            raise ValueError(f"route_share must be in [0, 1), got {share}")
        self.route_share = share
        self._push_transform()

    # ------------------------------------------------------------- compute

    def encode(self, raw):
        torch = _torch()
        x = torch.empty((raw.shape[0], 5), dtype=torch.float32, device=self.device)
        self._check(self._lib.atmo_mlp_encode(self._handle, raw.data_ptr(), raw.shape[0],
                                              x.data_ptr(), self._stream()))
        return x

    def forward(self, x):
        torch = _torch()
        y = torch.empty((x.shape[0], 6), dtype=torch.float32, device=self.device)
        self._check(self._lib.atmo_mlp_forward(self._handle, x.data_ptr(), x.shape[0],
                                               y.data_ptr(), self._stream()))
        return y

    def gradients(self, x, target, index=None, batch=None, weight=None):
        torch = _torch()
        batch = int(batch if batch is not None else x.shape[0])
        grad = torch.empty(_capi.MLP_PARAMS, dtype=torch.float32, device=self.device)
        loss = torch.empty(1, dtype=torch.float32, device=self.device)
        self._check(self._lib.atmo_mlp_gradients(
            self._handle, x.data_ptr(), target.data_ptr(),
            weight.data_ptr() if weight is not None else None,
            index.data_ptr() if index is not None else None, batch, grad.data_ptr(),
            loss.data_ptr(), self._stream()))
        return grad, loss

    def train_step(self, x, target, index, batch: int, lr: float, step: int, loss_out=None,
                   weight=None):
        self._check(self._lib.atmo_mlp_train_step(
            self._handle, x.data_ptr(), target.data_ptr(),
            weight.data_ptr() if weight is not None else None,
            index.data_ptr() if index is not None else None, int(batch), float(lr), int(step),
            loss_out.data_ptr() if loss_out is not None else None, self._stream()))

    def pack(self) -> None:
        self._check(self._lib.atmo_mlp_pack(self._handle, self._stream()))

    def infer(self, x):
        """BF16 MFMA inference of raw outputs for encoded inputs."""
        torch = _torch()
        y = torch.empty((x.shape[0], 6), dtype=torch.float32, device=self.device)
        self._check(self._lib.atmo_mlp_infer(self._handle, x.data_ptr(), x.shape[0],
                                             y.data_ptr(), self._stream()))
        return y

    # ------------------------------------------------------------- file i/o

    def save(self, path) -> None:
        """<path>.bin: float32 parameters (little-endian);
        <path>.json: output transform, domain and training metadata."""
        path = Path(path)
        self.get_params().astype("<f4").tofile(path.with_suffix(".bin"))
        meta = dict(self.metadata)
        meta.update({
            "format": "atmo-surrogate-v1",
            "layers": [[5, 64], [64, 64], [64, 64], [64, 6]],
            "activation": "silu",
            "params": _capi.MLP_PARAMS,
            "output_mode": MODE_NAMES[self.output_mode],
            "planet_radius_m": self.planet_radius_m,
            "cosine_stretch": self.cosine_stretch,
            "exact_transmittance": self.exact_transmittance,
            "route_share": self.route_share,
            "inscatter_floor": INSCATTER_FLOOR,
            "out_mean": [float(v) for v in self.out_mean],
            "out_scale": [float(v) for v in self.out_scale],
            "altitude_max_m": ALTITUDE_MAX_M,
            "altitude_scale_m": ALTITUDE_SCALE_M,
            "distance_range_m": list(DISTANCE_RANGE_M),
            "eps_inscatter": EPS_INSCATTER,
            "eps_transmittance": EPS_TRANSMITTANCE,
        })
        path.with_suffix(".json").write_text(json.dumps(meta, indent=1))

    @property
    def gate_met(self) -> bool | None:
        """Whether the weight file records the network as meeting the 1 % p99
        accuracy gate against the float64 restatement on its validation
        frame (None when the file records no validation)."""
        validation = self.metadata.get("validation")
        return None if validation is None else bool(validation.get("gate_met"))

    @classmethod
    def load(cls, path, max_batch: int = 65536, device=None) -> "SurrogateMLP":
        path = Path(path)
        meta = json.loads(path.with_suffix(".json").read_text())
        if meta.get("format") != "atmo-surrogate-v1":
            raise ValueError(f"{path}: not an atmo surrogate weight file")
        params = np.fromfile(path.with_suffix(".bin"), dtype="<f4")
        net = cls(max_batch=max_batch, device=device)
        net.set_params(params)
        mode = {v: k for k, v in MODE_NAMES.items()}[meta.get("output_mode", "log")]
        net.planet_radius_m = float(meta.get("planet_radius_m", 6_360_000.0))
        net.cosine_stretch = float(meta.get("cosine_stretch", 0.0))
        net.exact_transmittance = bool(meta.get("exact_transmittance", False))
        net.route_share = float(meta.get("route_share", 0.0))
        net.set_output_transform(meta["out_mean"], meta["out_scale"], mode)
        net.metadata = meta
        return net


def generate_labels(renderer, raw_host: np.ndarray, planet_radius_m: float,
                    sun_angular_radius_deg: float):
    """Labels from the GPU integrator: builds the renderer's T- and MS-LUTs
    for unit irradiance, then evaluates the surface branch per sample.
    Returns (raw_dev, labels_dev) torch tensors."""
    torch = _torch()
    ar = float(np.float32(math.radians(sun_angular_radius_deg)))
    renderer._build_transmittance(planet_radius_m)
    renderer._build_multi_scattering(planet_radius_m, (1.0, 1.0, 1.0), ar)
    raw = torch.from_numpy(np.ascontiguousarray(raw_host, dtype=np.float32)).to(renderer.device)
    labels = torch.empty((raw.shape[0], 6), dtype=torch.float32, device=renderer.device)
    stream = torch.cuda.current_stream(renderer.device).cuda_stream
    _capi.check(library().atmo_surrogate_labels(renderer.handle, raw.data_ptr(), raw.shape[0],
                                                float(planet_radius_m), math.radians(
                                                    sun_angular_radius_deg),
                                                labels.data_ptr(), stream), renderer.handle)
    return raw, labels


def train(renderer, planet_radius_m: float, sun_angular_radius_deg: float,
          samples: int = 1 << 22, batch: int = 65536, epochs: int = 100, lr: float = 1.0e-3,
          seed: int = SEED_BASE + 3, schedule: str = "constant", lr_final: float = 1.0e-5,
          weighted: bool = False, mode: int = OUTPUT_LOG_FLOOR,
          cosine_stretch: float = COSINE_STRETCH, exact_transmittance: bool = True,
          route_share: float = ROUTE_SHARE, log=None) -> tuple[SurrogateMLP, TrainingReport]:
    """Label `samples` seeded inputs with the GPU integrator and train the
    MLP with Adam for `epochs` passes in shuffled batches.  `schedule`
    "constant" keeps lr; "cosine" anneals it from lr to lr_final.  With
    `weighted`, each output's squared error is weighted as in loss_weights
    (linear and log modes); `mode` picks the output transform,
    `cosine_stretch` the input encoding, `exact_transmittance` what the
    surrogate composite uses for T and `route_share` its routing to the
    plain integrator."""
    torch = _torch()
    device = renderer.device
    t0 = time.perf_counter()
    raw_host = sample_inputs(samples, seed)
    raw, labels = generate_labels(renderer, raw_host, planet_radius_m, sun_angular_radius_deg)
    torch.cuda.synchronize(device)
    label_s = time.perf_counter() - t0
    labels_host = labels.cpu().numpy()
    targets = targets_from_labels(labels_host, mode)
    mean, scale = targets.mean(axis=0), targets.std(axis=0)
    scale = np.where(scale > 1e-6, scale, 1.0)
    t_std = torch.from_numpy(((targets - mean) / scale).astype(np.float32)).to(device)
    weight = (torch.from_numpy(loss_weights(labels_host, scale, mode)).to(device)
              if weighted and mode != OUTPUT_LOG_FLOOR else None)
    net = SurrogateMLP(max_batch=batch, device=device)
    net.set_encoding(planet_radius_m, cosine_stretch)
    net.set_exact_transmittance(exact_transmittance)
    net.set_route_share(route_share)
    net.set_output_transform(mean, scale, mode)
    net.set_params(init_params(seed))
    x = net.encode(raw)
    g = np.random.Generator(np.random.PCG64(seed + 1))
    steps_per_epoch = samples // batch
    total_steps = steps_per_epoch * epochs
    loss = torch.zeros(1, dtype=torch.float32, device=device)
    report = TrainingReport(samples, batch, epochs, 0, lr, 0.0, label_s)
    step = 0
    t1 = time.perf_counter()
    for epoch in range(epochs):
        perm = torch.from_numpy(g.permutation(samples).astype(np.int32)).to(device)
        for b in range(steps_per_epoch):
            rate = lr
            if schedule == "cosine":
                rate = lr_final + 0.5 * (lr - lr_final) * (1.0 + math.cos(math.pi * step / total_steps))
            step += 1
            net.train_step(x, t_std, perm[b * batch:(b + 1) * batch], batch, rate, step, loss,
                           weight=weight)
        if epoch % 10 == 9 or epoch == epochs - 1:
            value = float(loss.item())
            report.losses.append((epoch + 1, value))
            if log:
                log(f"surrogate epoch {epoch + 1}/{epochs}: loss {value:.3e}")
    net.pack()
    torch.cuda.synchronize(device)
    report.steps = step
    report.seconds = time.perf_counter() - t1
    net.metadata.update({"samples": samples, "batch": batch, "epochs": epochs, "lr": lr,
                         "schedule": schedule, "lr_final": lr_final if schedule == "cosine" else lr,
                         "weighted_loss": weighted, "seed": seed,
                         "output_mode": MODE_NAMES[mode],
                         "planet_radius_m": planet_radius_m,
                         "sun_angular_radius_deg": sun_angular_radius_deg,
                         "final_loss": report.losses[-1][1] if report.losses else None})
    return net, report


__all__ = [
    "SurrogateMLP",
    "TrainingReport",
    "decode_outputs",
    "default_weights_path",
    "encode_host",
    "horizon_cosine",
    "generate_labels",
    "init_params",
    "loss_weights",
    "sample_inputs",
    "targets_from_labels",
    "train",
]
