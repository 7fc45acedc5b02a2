"""Train the aerial-perspective surrogate on the GPU and save its weight file.

    python planet-real-time-render_amd/train_surrogate.py [--samples N] [--epochs E]
        [--batch B] [--lr LR] [--out PATH]

Labels come from the GPU integrator (the plain composite's surface branch)
on the full-frame atmosphere (Earth radius 6360 km, 100 km shell, ozone),
unit irradiance and a 0.266 deg sun; training runs the library's FP32
forward / backward / Adam kernels.  The result is PATH.bin (float32
parameters) + PATH.json (output transform, domain, training record).
"""

from __future__ import annotations

import argparse
import sys
import time
from pathlib import Path

HERE = Path(__file__).resolve().parent


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--samples", type=int, default=1 << 22)
    ap.add_argument("--epochs", type=int, default=200)
    ap.add_argument("--batch", type=int, default=65536)
    ap.add_argument("--lr", type=float, default=1.0e-3)
    ap.add_argument("--seed", type=int, default=None)
    ap.add_argument("--schedule", choices=("constant", "cosine"), default="cosine")
    ap.add_argument("--lr-final", type=float, default=1.0e-5)
    ap.add_argument("--mode", choices=("log_floor", "log", "linear"), default="log_floor",
                    help="output transform (default: floored log)")
    ap.add_argument("--weighted", action="store_true",
                    help="pixel-share weighted loss (linear and log modes)")
    ap.add_argument("--cosine-stretch", type=float, default=None,
                    help="horizon-relative asinh stretch of the cosines (0: raw cosines)")
    ap.add_argument("--network-transmittance", action="store_true",
                    help="composite with the network's transmittance instead of the exact one")
    ap.add_argument("--route-share", type=float, default=None,
                    help="plain integrator for pixels whose learned in-scatter exceeds this "
                         "share of the pixel (default: surrogate.ROUTE_SHARE; 0: no routing)")
    ap.add_argument("--out", type=Path, default=None,
                    help="output path without suffix (default: the shipped weights)")
    args = ap.parse_args(argv)

    sys.path.insert(0, str(HERE.parent))
    import __graft_entry__ as entry

    pkg = entry.load_package()
    scene = entry.load_submodule("scene")
    sur = entry.load_submodule("surrogate")
    import torch

    device = torch.device("cuda:0")
    renderer = pkg.AtmosphereRenderer(16, 16, scene.full_frame_config(), device)
    log = lambda m: print(m, file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    kwargs = {} if args.seed is None else {"seed": args.seed}
    net, report = sur.train(renderer, scene.EARTH_RADIUS_M, scene.SUN_ANGULAR_RADIUS_DEG,
                            samples=args.samples, batch=args.batch, epochs=args.epochs,
                            lr=args.lr, schedule=args.schedule, lr_final=args.lr_final,
                            weighted=args.weighted,
                            mode={v: k for k, v in sur.MODE_NAMES.items()}[args.mode],
                            cosine_stretch=(sur.COSINE_STRETCH if args.cosine_stretch is None
                                            else args.cosine_stretch),
                            exact_transmittance=not args.network_transmittance,
                            route_share=(sur.ROUTE_SHARE if args.route_share is None
                                         else args.route_share),
                            log=log, **kwargs)
    if args.out is None:
        args.out = sur.default_weights_path()
    args.out.parent.mkdir(parents=True, exist_ok=True)
    net.metadata["training_seconds"] = report.seconds
    net.metadata["label_seconds"] = report.label_seconds
    net.metadata["loss_curve"] = report.losses
    net.save(args.out)
    log(f"trained {report.steps} steps in {report.seconds:.1f} s (labels {report.label_seconds:.2f} s, "
        f"total {time.perf_counter() - t0:.1f} s); saved {args.out}.bin/.json")
    net.close()
    renderer.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
