"""Pulse and probe spectra (host, float64).

``pulse_spectrum``: spectrum of an n-cycle windowed sine,
``i * [sinc(T(w - wc)/2pi) - sinc(T(w + wc)/2pi)]`` with ``T = n/fc``
(``spectrum.py:24-50``).

``probe_spectrum``: one-way response, the square root of a generalized-normal
window whose two-way -6 dB points sit at ``wc (1 +- B/2)``; shape exponent
``p = ln 126 / ln(2 wc / wb)``, scale ``sigma = (wb/2) / ln(2)^(1/p)``
(``spectrum.py:53-107``).
"""

from __future__ import annotations

from math import log, pi

import numpy as np


def pulse_spectrum(angular_freq, freq_center: float, tx_n_wavelengths: float = 1.0) -> np.ndarray:
    w = np.asarray(angular_freq, dtype=np.float64)
    duration = tx_n_wavelengths / freq_center
    wc = 2.0 * pi * freq_center
    lo = duration * (w - wc) / 2.0 / pi
    hi = duration * (w + wc) / 2.0 / pi
    return 1j * (np.sinc(lo) - np.sinc(hi))


def probe_spectrum(angular_freq, freq_center: float, bandwidth: float = 0.75) -> np.ndarray:
    if not (0.0 < bandwidth < 2.0):
        raise ValueError(f"bandwidth must be in (0, 2.0), got {bandwidth!r}")
    w = np.asarray(angular_freq, dtype=np.float64)
    wc = 2.0 * pi * freq_center
    wb = bandwidth * wc
    p = log(126) / log(2.0 * wc / wb)
    sigma = wb / 2.0 / (log(2) ** (1.0 / p))
    return np.sqrt(np.exp(-((np.abs(w - wc) / sigma) ** p)))
