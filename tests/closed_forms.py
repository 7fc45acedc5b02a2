"""Closed forms used to pin the oracle."""

import math

import numpy as np


def vertical_optical_depth(cfg, h0: float) -> np.ndarray:
    """Closed form of the zenith optical depth from h0 to the top:
    beta H (e^{-h0/H} - e^{-top/H}) per exponential layer plus the exact
    integral of the ozone tent max(1 - |h - p| / w, 0) over [h0, top]."""
    top = cfg.top_altitude_m

    def layer(beta, scale):
        return np.asarray(beta) * scale * (math.exp(-h0 / scale) - math.exp(-top / scale))

    def tent_integral(a, b):
        p, w = cfg.absorption_peak_altitude_m, cfg.absorption_half_width_m

        def antiderivative(h):  # integral of the tent from -inf to h
            if h <= p - w:
                return 0.0
            if h <= p:
                return 0.5 * (h - (p - w)) ** 2 / w
            if h <= p + w:
                return w - 0.5 * ((p + w) - h) ** 2 / w
            return w

        return antiderivative(b) - antiderivative(a)

    return (layer(cfg.rayleigh_scattering_per_m, cfg.rayleigh_scale_height_m)
            + layer(cfg.mie_extinction_per_m, cfg.mie_scale_height_m)
            + np.asarray(cfg.absorption_extinction_per_m) * tent_integral(h0, top))
