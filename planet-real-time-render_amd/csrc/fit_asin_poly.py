"""Fits the coefficients of asin_poly (atmo_device.h): the degree-12
Chebyshev approximation of P(w) = (asin(sqrt(w)) - sqrt(w)) / w^1.5 on
[0, 1/4], computed with mpmath at 50 significant digits and rounded to
float64, highest degree first (the Horner order of asin_poly).  The device
acos evaluates asin(s) = s + s w P(w) with w = s^2 <= 1/4.

    python planet-real-time-render_amd/csrc/fit_asin_poly.py

prints the coefficients as C hex literals and the fit's maximum error.
"""

from __future__ import annotations

import mpmath as mp

DEGREE = 12
DIGITS = 50


def fit():
    """(float64 coefficients, highest degree first; maximum fit error)."""
    with mp.workdps(DIGITS):
        target = lambda w: (mp.asin(mp.sqrt(w)) - mp.sqrt(w)) / w ** mp.mpf(1.5)
        poly, err = mp.chebyfit(target, [0, mp.mpf(1) / 4], DEGREE + 1, error=True)
        return [float(c) for c in poly], float(err)


if __name__ == "__main__":
    coefficients, error = fit()
    for c in coefficients:
        print(c.hex())
    print(f"max fit error {error:.3e}")
