"""The formats on either side of the path: the linear-HDR frame file the
reference's `save_image` writes (src/planet_renderer/cli.py:136-146:
`np.save(hdr_path, renderer.hdr_numpy())`, a (W, H, 3) float32 array) and
the G-buffer contract of the composite's caller (position_view in view
coordinates, surface_id < 0 for sky; renderer.py:235-251, ocean.py:17)."""

import numpy as np
import pytest


def frame(w=7, h=5, seed=0):
    return np.random.default_rng(seed).uniform(0.0, 4.0, (w, h, 3)).astype(np.float32)


def test_hdr_file_is_the_reference_npy(pkg, tmp_path):
    hdr = frame()
    # what the reference writes for the same field
    np.save(tmp_path / "reference.npy", hdr)
    out = pkg.save_hdr(hdr, tmp_path / "sub" / "frame.npy")
    assert out == tmp_path / "sub" / "frame.npy"
    assert out.read_bytes() == (tmp_path / "reference.npy").read_bytes()
    # header: little-endian float32, C order, (W, H, 3)
    with open(out, "rb") as f:
        version = np.lib.format.read_magic(f)
        shape, fortran, dtype = np.lib.format._read_array_header(f, version)
    assert (shape, fortran, dtype) == ((7, 5, 3), False, np.dtype("<f4"))
    np.testing.assert_array_equal(pkg.load_hdr(out), hdr)
    # np.save's suffix rule
    assert pkg.save_hdr(hdr, tmp_path / "bare") == tmp_path / "bare.npy"


def test_hdr_file_from_a_torch_tensor_and_a_field(pkg, tmp_path):
    import torch

    hdr = frame(4, 9, seed=1)
    pkg.save_hdr(torch.from_numpy(hdr), tmp_path / "t.npy")
    pkg.save_hdr(pkg.DeviceField(torch.from_numpy(hdr)), tmp_path / "f.npy")
    np.save(tmp_path / "r.npy", hdr)
    ref = (tmp_path / "r.npy").read_bytes()
    assert (tmp_path / "t.npy").read_bytes() == ref == (tmp_path / "f.npy").read_bytes()


def test_hdr_file_rejections(pkg, tmp_path):
    with pytest.raises(ValueError):
        pkg.save_hdr(frame().astype(np.float64), tmp_path / "x.npy")
    with pytest.raises(ValueError):
        pkg.save_hdr(np.zeros((4, 4), np.float32), tmp_path / "x.npy")
    np.save(tmp_path / "obj.npy", np.array([{"a": 1}], dtype=object), allow_pickle=True)
    with pytest.raises(ValueError):
        pkg.load_hdr(tmp_path / "obj.npy")
    np.save(tmp_path / "i.npy", np.zeros((2, 2, 3), np.int32))
    with pytest.raises(ValueError):
        pkg.load_hdr(tmp_path / "i.npy")


def test_gbuffer_contract(pkg, scene):
    inputs = scene.make_frame(width=48, height=32, config_index=1)
    pkg.check_gbuffer(inputs.position_view, inputs.surface_id, 48, 32)
    # the atmosphere treats every id >= 0 as surface, the ocean's included
    sid = inputs.surface_id.copy()
    sid[sid >= 0] = pkg.OCEAN_SURFACE_ID
    pkg.check_gbuffer(inputs.position_view, sid, 48, 32)
    with pytest.raises(ValueError):
        pkg.check_gbuffer(inputs.position_view.astype(np.float64), inputs.surface_id, 48, 32)
    with pytest.raises(ValueError):
        pkg.check_gbuffer(inputs.position_view, inputs.surface_id.astype(np.int64), 48, 32)
    with pytest.raises(ValueError):
        pkg.check_gbuffer(inputs.position_view, inputs.surface_id, 32, 48)
    behind = inputs.position_view.copy()
    behind[inputs.surface_id >= 0, 2] *= -1.0
    with pytest.raises(ValueError):
        pkg.check_gbuffer(behind, inputs.surface_id, 48, 32)
    nan = inputs.position_view.copy()
    nan[np.argwhere(inputs.surface_id >= 0)[0][0], np.argwhere(inputs.surface_id >= 0)[0][1]] = np.nan
    with pytest.raises(ValueError):
        pkg.check_gbuffer(nan, inputs.surface_id, 48, 32)
