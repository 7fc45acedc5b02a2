"""Batched LUT builds: every state of a 256-state batch matches what a cold
`update` + `prepare_frame` of its own renderer builds (the transmittance
bitwise; the multi-scattering build marches each direction in one lane
instead of four chunks, so it and the LUTs that read it agree to float
reassociation), a sample of states matches the float32 CPU restatement,
graph replays are stable, and the reference's LUT invariants hold for all
states."""

import math

import numpy as np
import pytest

from parity import PLAIN_TOLERANCE, rel_error, summary

pytestmark = pytest.mark.gpu

WIDTH, HEIGHT = 1920, 1080


@pytest.fixture(scope="module")
def device():
    import torch

    assert torch.cuda.is_available()
    return torch.device("cuda:0")


@pytest.fixture(scope="module")
def batch_mod(pkg):
    import importlib

    return importlib.import_module(pkg.__name__ + ".batch")


@pytest.fixture(scope="module")
def built(scene, batch_mod, device):
    import torch

    states = scene.batch_states(256)
    print(f"batch: {len(states)} states, {WIDTH}x{HEIGHT} aerial rays", flush=True)
    b = batch_mod.LutBatch(states, WIDTH, HEIGHT, device)
    b.run()
    torch.cuda.synchronize(device)
    yield states, b
    b.close()


def _single_state(pkg, s, device):
    """The four LUTs of one state from its own renderer: cold update, then
    the aerial build of prepare_frame (same aspect ratio as the batch)."""
    r = pkg.AtmosphereRenderer(WIDTH // 10, HEIGHT // 10, s.config, device)
    r.update(s.planet_radius_m, s.camera_radius_m, s.sun_local, s.solar_irradiance,
             s.sun_angular_radius_degrees)
    right, up, forward = s.view_basis
    r._build_aerial_perspective(s.planet_radius_m, max(s.camera_radius_m - s.planet_radius_m, 0.0),
                                tuple(float(v) for v in s.sun_local), s.solar_irradiance,
                                math.radians(s.sun_angular_radius_degrees), right, up, forward,
                                s.tangent_half_fov)
    out = [r.transmittance_lut.to_numpy(), r.multi_scattering_lut.to_numpy(),
           r.sky_view_lut.to_numpy(), r.aerial_scattering_lut.to_numpy()]
    r.close()
    return out


# Batch against single-state build: the multi-scattering sums (64 directions
# x 20 steps) differ only in association, about 1e-7 relative.
REASSOCIATION_TOLERANCE = 2e-6


def test_batch_equals_single_state_builds(pkg, built, device):
    """Every 8th state (both planets) against its own renderer: the
    transmittance kernel body is shared, so that LUT is bitwise equal; the
    batched multi-scattering build runs one lane per direction (the
    single-state build four chunks combined by a scan), so MS and the
    sky-view and aerial LUTs built from it agree to float reassociation."""
    states, b = built
    luts = [t.cpu().numpy() for t in b.luts]
    worst = dict.fromkeys(b.NAMES[1:], 0.0)
    for i in range(0, len(states), 8):
        want = _single_state(pkg, states[i], device)
        assert np.array_equal(luts[0][i], want[0]), f"state {i} transmittance"
        for name, got, ref in zip(b.NAMES[1:], luts[1:], want[1:]):
            err = float(rel_error(got[i], ref).max())
            worst[name] = max(worst[name], err)
            assert err <= REASSOCIATION_TOLERANCE, f"state {i} {name}: {err:.2e}"
    print("batch vs single-state max rel err: " +
          ", ".join(f"{k} {v:.2e}" for k, v in worst.items()), flush=True)


def test_batch_matches_cpu_oracle(built, ref32):
    states, b = built
    luts = [t.cpu().numpy() for t in b.luts]
    for i in (0, 37, 128, 201):
        s = states[i]
        cfg, R = s.config, float(s.planet_radius_m)
        E = tuple(float(v) for v in s.solar_irradiance)
        deg = float(s.sun_angular_radius_degrees)
        sun = np.asarray(s.sun_local, dtype=np.float64)
        sun_n = sun / max(float(np.linalg.norm(sun)), 1.0e-12)
        ar_update = float(np.float32(math.radians(deg)))
        alt = max(float(s.camera_radius_m) - R, 0.0)
        alt_key = float(np.float32(alt))
        szc = float(np.float32(np.clip(sun_n[1], -1.0, 1.0)))
        right, up, forward = s.view_basis
        t = ref32.build_transmittance(cfg, R)
        ms = ref32.build_multi_scattering(cfg, R, E, ar_update, t)
        sv = ref32.build_sky_view(cfg, R, alt_key, szc, E, ar_update, t, ms)
        ap = ref32.build_aerial_perspective(cfg, WIDTH, HEIGHT, R, alt, sun, E,
                                            math.radians(deg), right, up, forward,
                                            s.tangent_half_fov, t, ms)
        for name, got, want in zip(b.NAMES, luts, (t, ms, sv, ap)):
            e = summary(rel_error(got[i], want))
            print(f"state {i} (R={R:.0f} m, h={alt:.0f} m) {name}: max {e['max']:.2e}",
                  flush=True)
            assert e["max"] <= PLAIN_TOLERANCE, f"state {i} {name}"


def test_batch_graph_replays_are_stable(built, device):
    """The captured graph replayed with eager work in between gives the same
    LUTs every time."""
    import torch

    states, b = built
    first = [t.clone() for t in b.luts]
    for _ in range(3):
        for t in b.luts:
            t.fill_(-1.0)
        torch.cuda.synchronize(device)
        b.run()
        torch.cuda.synchronize(device)
        for t, f in zip(b.luts, first):
            assert torch.equal(t, f)


def test_batch_stage_launches_equal_the_graph(built, device):
    """atmo_batch_run_stages: the four stage kernels launched one by one, in
    order, give the graph's LUTs bit for bit."""
    import torch

    states, b = built
    b.run()
    torch.cuda.synchronize(device)
    first = [t.clone() for t in b.luts]
    for t in b.luts:
        t.fill_(-1.0)
    for bit in (1, 2, 4, 8):
        b.run_stages(bit)
    torch.cuda.synchronize(device)
    for t, f in zip(b.luts, first):
        assert torch.equal(t, f)


def test_batch_luts_satisfy_reference_invariants(built):
    """tests/test_m2_terrain.py:252-270: T in [0, 1]; MS, SV, AP finite and
    non-negative with a positive maximum, per state."""
    states, b = built
    t, ms, sv, ap = (x.cpu().numpy() for x in b.luts)
    assert np.all(np.isfinite(t)) and t.min() >= 0.0 and t.max() <= 1.0
    for lut in (ms, sv, ap):
        assert np.all(np.isfinite(lut)) and lut.min() >= 0.0 and lut.max() > 0.0
    for i in range(len(states)):
        assert ms[i].max() > 0.0, f"state {i}"


def test_batch_rejects_mismatched_lut_sizes(scene, batch_mod, device):
    states = scene.batch_states(4)
    states[2].config = states[2].config.replace(sky_view_lut_width=96)
    with pytest.raises(ValueError, match="same LUT dimensions"):
        batch_mod.LutBatch(states, WIDTH, HEIGHT, device)

    class OutOfRange:
        """A state whose parameters skip the Python-side validation."""

        def __init__(self, state):
            self.state, self.config = state, state.config

        def to_struct(self):
            s = self.state.to_struct()
            s.params.mie_phase_g = 1.5
            return s

    states = scene.batch_states(4)
    states[1] = OutOfRange(states[1])
    with pytest.raises(ValueError, match="state 1: atmosphere.mie_phase_g"):
        batch_mod.LutBatch(states, WIDTH, HEIGHT, device)


def test_batch_autotune_keeps_parity(scene, batch_mod, ref32, device):
    """The autotuned launch shapes only reassociate sums: a sample of states
    stays within the parity bar of the oracle, and graph replays are
    stable after the re-capture."""
    import torch

    states = scene.batch_states(16)
    b = batch_mod.LutBatch(states, WIDTH, HEIGHT, device)
    chosen = b.autotune()
    print(f"batch autotune (16 states): lanes per T texel {chosen[0]}, per SV texel {chosen[1]}",
          flush=True)
    assert chosen[0] in (1, 2, 4, 8, 16) and chosen[1] in (1, 2, 4, 8, 16)
    b.run()
    torch.cuda.synchronize(device)
    first = [t.clone() for t in b.luts]
    b.run()
    torch.cuda.synchronize(device)
    for t, f in zip(b.luts, first):
        assert torch.equal(t, f)
    s = states[3]
    t = ref32.build_transmittance(s.config, float(s.planet_radius_m))
    e = summary(rel_error(b.transmittance[3].cpu().numpy(), t))
    assert e["max"] <= PLAIN_TOLERANCE
    b.close()
