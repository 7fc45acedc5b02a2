"""The data formats on either side of the atmosphere path.

Output: the linear-HDR frame file.  The reference's `save_image`
(src/planet_renderer/cli.py:136-146) writes `np.save(hdr_path,
renderer.hdr_numpy())`: the composite's `hdr` field as a `.npy` array of
shape (W, H, 3), float32, indexed [x, y, channel] with y = 0 the bottom
row.  `save_hdr` writes the same bytes from a device tensor (or a
`DeviceField`, or a NumPy array); `load_hdr` reads such a file back
without unpickling anything.

Input: the G-buffer contract of the composite's caller
(src/planet_renderer/renderer.py:235-251; ocean.py:17, 674-683).
`position_view` is the (W, H, 3) float32 surface position in view
coordinates (x right, y up, z forward, metres from the camera);
`surface_id` is the (W, H) int32 surface id, negative for sky pixels,
0.. for terrain cells and OCEAN_SURFACE_ID for the ocean; the atmosphere
treats every id >= 0 as surface (atmosphere/renderer.py:1441, 1861).
`check_gbuffer` validates a producer's arrays against it.
"""

from __future__ import annotations

from pathlib import Path

import numpy as np

# src/planet_renderer/ocean.py:17
OCEAN_SURFACE_ID = 2_147_000_000


def _host_array(frame) -> np.ndarray:
    tensor = getattr(frame, "tensor", frame)  # DeviceField -> tensor
    if isinstance(tensor, np.ndarray):
        array = tensor
    else:
        import torch

        if not isinstance(tensor, torch.Tensor):
            raise ValueError("frame must be a torch tensor, a DeviceField or a NumPy array")
        array = tensor.detach().to("cpu").numpy()
    if array.dtype != np.float32 or array.ndim != 3 or array.shape[2] != 3:
        raise ValueError(f"an HDR frame is (W, H, 3) float32, got {array.shape} {array.dtype}")
    return np.ascontiguousarray(array)


def save_hdr(frame, path) -> Path:
    """Write a (W, H, 3) float32 HDR frame as the reference's `.npy`
    (cli.py:143-145).  Creates the parent directory, as the reference
    does; like `np.save`, appends `.npy` when the name has no suffix."""
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    array = _host_array(frame)
    np.save(path, array)
    return path if path.suffix == ".npy" else path.with_name(path.name + ".npy")


def load_hdr(path) -> np.ndarray:
    """Read an HDR frame file; pickled payloads are refused."""
    array = np.load(Path(path), allow_pickle=False)
    if array.dtype != np.float32 or array.ndim != 3 or array.shape[2] != 3:
        raise ValueError(f"not an HDR frame file: {array.shape} {array.dtype}")
    return array


def check_gbuffer(position_view, surface_id, width: int, height: int) -> None:
    """Validate a G-buffer producer's arrays (NumPy or torch) against the
    contract above: shapes, dtypes, and finite positions in front of the
    camera on surface pixels.  Raises ValueError."""
    pos = getattr(position_view, "tensor", position_view)
    sid = getattr(surface_id, "tensor", surface_id)
    if not isinstance(pos, np.ndarray):
        pos = pos.detach().to("cpu").numpy()
    if not isinstance(sid, np.ndarray):
        sid = sid.detach().to("cpu").numpy()
    if pos.shape != (width, height, 3) or pos.dtype != np.float32:
        raise ValueError(f"position_view must be ({width}, {height}, 3) float32, "
                         f"got {pos.shape} {pos.dtype}")
    if sid.shape != (width, height) or sid.dtype != np.int32:
        raise ValueError(f"surface_id must be ({width}, {height}) int32, got {sid.shape} {sid.dtype}")
    surface = sid >= 0
    p = pos[surface]
    if not np.isfinite(p).all():
        raise ValueError("position_view must be finite on surface pixels")
    if (p[:, 2] <= 0.0).any():
        raise ValueError("position_view must lie in front of the camera (z > 0) on surface pixels")


__all__ = ["OCEAN_SURFACE_ID", "check_gbuffer", "load_hdr", "save_hdr"]
