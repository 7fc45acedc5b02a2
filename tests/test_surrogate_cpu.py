"""CPU checks of the surrogate's reference pieces: the float64 MLP
(oracle/mlp_fp64.py) is pinned by central finite differences and by Adam's
first-step property; the host-side input normalisation and output transform
round-trip."""

import math

import numpy as np
import pytest

from oracle import mlp_fp64


def test_param_count_matches_header():
    assert mlp_fp64.PARAM_COUNT == 9094


def test_fp64_gradients_match_finite_differences():
    g = np.random.default_rng(3)
    params = mlp_fp64.init_params(3) + g.normal(0.0, 0.05, mlp_fp64.PARAM_COUNT)
    x = g.uniform(-1, 1, (16, 5))
    t = g.normal(0, 1, (16, 6))
    grad, _ = mlp_fp64.gradients(params, x, t)
    picks = g.choice(mlp_fp64.PARAM_COUNT, 60, replace=False)
    h = 1e-6
    for i in picks:
        p1, p0 = params.copy(), params.copy()
        p1[i] += h
        p0[i] -= h
        fd = (mlp_fp64.loss(mlp_fp64.forward(p1, x)[0], t)
              - mlp_fp64.loss(mlp_fp64.forward(p0, x)[0], t)) / (2 * h)
        assert fd == pytest.approx(grad[i], rel=1e-5, abs=1e-9)


def test_silu_derivative():
    z = np.linspace(-8, 8, 101)
    fd = (mlp_fp64.silu(z + 1e-6) - mlp_fp64.silu(z - 1e-6)) / 2e-6
    np.testing.assert_allclose(mlp_fp64.dsilu(z), fd, rtol=1e-7, atol=1e-10)


def test_adam_first_step_is_signed_lr():
    """With zero moments, Adam's first update is lr * g / (|g| + eps)."""
    g = np.random.default_rng(0)
    p = g.normal(size=100)
    grad = g.normal(size=100)
    new, m, v = mlp_fp64.adam_step(p, grad, np.zeros(100), np.zeros(100), 1, lr=1e-3)
    np.testing.assert_allclose(p - new, 1e-3 * grad / (np.abs(grad) + 1e-8), rtol=1e-9)


def test_input_normalisation_and_target_transform(pkg):
    import importlib

    s = importlib.import_module(pkg.__name__ + ".surrogate")
    raw = s.sample_inputs(10_000, 1)
    assert raw.dtype == np.float32 and raw.shape == (10_000, 5)
    x = s.encode_host(raw)
    assert np.all(np.abs(x) <= 1.0 + 1e-6)
    assert x[:, 4].min() < -0.99 and x[:, 4].max() > 0.99  # log-uniform distance spans the range
    labels = np.column_stack([np.geomspace(1e-7, 0.5, 50)] * 3 + [np.linspace(0.0, 1.0, 50)] * 3)
    for mode in (s.OUTPUT_LINEAR, s.OUTPUT_LOG, s.OUTPUT_LOG_FLOOR):
        t = s.targets_from_labels(labels, mode)
        mean, scale = t.mean(axis=0), t.std(axis=0)
        L, T = s.decode_outputs((t - mean) / scale, mean, scale, mode)
        want = labels[:, :3]
        if mode == s.OUTPUT_LOG_FLOOR:  # values below 1.5 x the floor decode to 0
            want = np.where(want < 1.5 * s.INSCATTER_FLOOR, 0.0, want)
        np.testing.assert_allclose(L, want, rtol=1e-9, atol=1e-15)
        np.testing.assert_allclose(T, labels[:, 3:], rtol=1e-9, atol=1e-12)


def test_horizon_relative_cosine_encoding(pkg):
    """The local horizon cosine equals -sqrt(1 - (R / (R + h))^2) (tangent
    from a sphere), and a cosine on the horizon encodes to 0 at every
    altitude; encoding is monotonic in the cosine and stays in [-1, 1]."""
    import importlib

    s = importlib.import_module(pkg.__name__ + ".surrogate")
    R = 6_360_000.0
    h = np.array([0.0, 0.5, 2_000.0, 100_000.0, 2.0e6])
    np.testing.assert_allclose(s.horizon_cosine(R, h), -np.sqrt(1.0 - (R / (R + h)) ** 2),
                               rtol=1e-9, atol=1e-15)
    raw = np.column_stack([h, s.horizon_cosine(R, h), s.horizon_cosine(R, h),
                           np.full(5, 1.0), np.full(5, 100.0)])
    x = s.encode_host(raw, R)
    np.testing.assert_allclose(x[:, 1:3], 0.0, atol=1e-12)
    mu = np.linspace(-1.0, 1.0, 2001)
    raw = np.column_stack([np.full_like(mu, 300.0), mu, mu, np.zeros_like(mu), np.ones_like(mu)])
    x = s.encode_host(raw, R)
    assert np.all(np.diff(x[:, 1]) > 0.0) and np.all(np.abs(x[:, 1:3]) <= 1.0)


def test_linear_loss_weights_measure_relative_pixel_error(pkg):
    """In linear mode, weight x (standardised error)^2 summed over L_c and
    T_c is the squared relative error of the pixel E L + s T (before the
    mean normalisation), as long as no floor or clip applies."""
    import importlib

    s = importlib.import_module(pkg.__name__ + ".surrogate")
    g = np.random.default_rng(3)
    labels = np.column_stack([g.uniform(0.01, 0.1, (64, 3)), g.uniform(0.2, 1.0, (64, 3))])
    scale = g.uniform(0.05, 0.5, 6)
    w = s.loss_weights(labels, scale, s.OUTPUT_LINEAR).astype(np.float64)
    # undo the normalisation with one pixel computed by hand
    E, sc = np.asarray(s.WEIGHT_IRRADIANCE), s.WEIGHT_SCENE
    pixel = E * labels[:, :3] + sc * labels[:, 3:]
    want = np.concatenate([(E * scale[:3] / pixel) ** 2, (sc * scale[3:] / pixel) ** 2], axis=1)
    np.testing.assert_allclose(w / w[0, 0], want / want[0, 0], rtol=1e-5)
    # a perturbation dy of output L_0 changes channel 0 by E_0 scale_0 dy
    dy = 1e-4
    rel = E[0] * scale[0] * dy / pixel[:, 0]
    np.testing.assert_allclose(np.sqrt(want[:, 0]) * dy, rel, rtol=1e-12)


def test_init_params_layout(pkg):
    import importlib

    s = importlib.import_module(pkg.__name__ + ".surrogate")
    p = s.init_params(0).astype(np.float64)
    layers = mlp_fp64.unpack(p)
    for (w, b), (fan_in, fan_out) in zip(layers, mlp_fp64.LAYERS):
        assert w.shape == (fan_out, fan_in) and np.all(b == 0.0)
        assert np.abs(w).max() <= math.sqrt(6.0 / fan_in)


def test_shipped_weights_record_their_gate_and_the_facade_warns(pkg):
    """The shipped network's weight file records its validation against the
    float64 restatement: with its routing (route_share) it meets the 1 % p99
    gate, while the network alone misses it.  The facade's composite warns
    once per network before using weights that record a missed gate, and
    not for the shipped ones."""
    import copy
    import importlib
    import json
    import warnings

    sur = importlib.import_module(pkg.__name__ + ".surrogate")
    renderer = importlib.import_module(pkg.__name__ + ".renderer")
    meta = json.loads(sur.default_weights_path().with_suffix(".json").read_text())
    validation = meta["validation"]
    assert validation["gate_p99"] == 0.01
    assert validation["gate_met"] is (validation["hdr_p99_vs_fp64"] <= 0.01)
    assert validation["gate_met"] and meta["route_share"] == validation["route_share"] > 0.0
    assert validation["network_alone"]["hdr_p99_vs_fp64"] > 0.01
    missed = copy.deepcopy(meta)
    missed["validation"]["gate_met"] = False

    def net_of(m):  # the attributes the facade reads, without a device
        class Net:
            metadata = m
            gate_met = sur.SurrogateMLP.gate_met.fget(type("M", (), {"metadata": m})())
        return Net()

    shipped, below = net_of(meta), net_of(missed)
    with warnings.catch_warnings(record=True) as caught:
        warnings.simplefilter("always")
        renderer._warn_if_below_gate(shipped)
        renderer._warn_if_below_gate(below)
        renderer._warn_if_below_gate(below)
    assert len(caught) == 1 and "accuracy gate" in str(caught[0].message)
