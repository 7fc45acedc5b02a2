"""The short correctly rounded square root, reciprocal and quotient of the
atmosphere kernels (atmo::csqrt, crcp, cdiv) return the bits of the
compiler's IEEE expansions on their stated domains: every float for the
square root and the reciprocal, 2^32 hashed pairs for the quotient.  The
kernels' float64 acos (atmo::acos_d), rounded to float32, equals the device
library's acos rounded the same way for every float in [-1, 1], and the
shared-reduction sincos gives the bits of separate sin and cos for every
float in [0, 256).  This is
what lets the kernels use them while keeping the reference's float32
arithmetic bit for bit."""

import ctypes

import pytest

from conftest import REPO

pytestmark = pytest.mark.gpu


def test_short_sqrt_rcp_div_equal_ieee():
    # torch first, as everywhere else: its HIP runtime must be the one loaded
    # before the probe library's dependencies resolve
    import torch

    assert torch.cuda.is_available()
    path = REPO / "planet-real-time-render_amd" / "libatmo_probe.so"
    assert path.exists(), "libatmo_probe.so is missing: build first"
    lib = ctypes.CDLL(str(path))
    lib.atmo_probe_arith_check.argtypes = [ctypes.POINTER(ctypes.c_uint64)]
    lib.atmo_probe_arith_check.restype = ctypes.c_int32
    counts = (ctypes.c_uint64 * 6)()
    print("arith check: 2^32 sqrt inputs, 201 x 2^24 reciprocals, 2^32 quotients, "
          "2^31 acos inputs, 1.1e9 sincos inputs", flush=True)
    assert lib.atmo_probe_arith_check(counts) == 0
    print(f"mismatches: sqrt {counts[0]}, rcp {counts[1]}, div {counts[2]}, acos {counts[4]}, "
          f"sincos {counts[5]}; "
          f"sqrt outside its domain {counts[3]}", flush=True)
    assert (counts[0], counts[1], counts[2], counts[4], counts[5]) == (0, 0, 0, 0, 0)
