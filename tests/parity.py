"""Parity metric shared by the tests: relative error with an absolute floor.

err = |got - ref| / max(|ref|, floor), floor = FLOOR_FRACTION * max|ref| of
the compared array (per channel when `per_channel`).  The floor keeps texels
that are exactly zero or saturated (T = exp(-80), radiance in shadow) from
dividing by zero; 1e-6 of the array's peak is far below any visible value.
"""

from __future__ import annotations

import numpy as np

PLAIN_TOLERANCE = 1.0e-4  # GPU plain path vs the float32 reference restatement
FLOOR_FRACTION = 1.0e-6


def rel_error(got, ref, floor_fraction: float = FLOOR_FRACTION, per_channel: bool = True):
    got = np.asarray(got, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    assert got.shape == ref.shape, (got.shape, ref.shape)
    if per_channel and ref.ndim >= 1 and ref.shape[-1] == 3:
        peak = np.abs(ref).reshape(-1, 3).max(axis=0)
    else:
        peak = np.abs(ref).max()
    floor = np.maximum(floor_fraction * peak, 1.0e-30)
    return np.abs(got - ref) / np.maximum(np.abs(ref), floor)


def summary(err) -> dict:
    err = np.asarray(err).ravel()
    return {
        "max": float(err.max()),
        "p99": float(np.percentile(err, 99)),
        "p999": float(np.percentile(err, 99.9)),
        "mean": float(err.mean()),
    }


def worst_index(err):
    return np.unravel_index(int(np.argmax(err)), np.asarray(err).shape)
