"""FP64 CPU checkers for IQ demodulation and delay-and-sum (consumers of the SIMUS RF stack).

TEST INFRASTRUCTURE ONLY: imported by ``tests/`` and by nothing in the product package.

* ``rf2iq``: the recipe of MUST's / PyMUST's ``rf2iq`` (down-mix by exp(-i 2 pi fc t), 5th-order
  Butterworth low-pass at min(2 fc / fs, 0.5) of Nyquist run with ``filtfilt``, times 2), with SciPy's
  ``signal.butter`` and ``signal.filtfilt`` (``scipy`` is present here; PyMUST, which the reference's
  harness calls, ``tests/_picmus_phantom_shared.py:431-440``, is not). Pinned by the library routines
  themselves and by the closed form of a tone at fc (a constant 1 away from the ends).
* ``das``: delay-and-sum as the product defines it (transmit arrival min_e(delay_e + r_e / c), receive
  r_e / c, linear interpolation, exp(i 2 pi fc tau), aperture |x - x_e| <= z / (2 F)), per pixel in
  NumPy float64. Pinned by the closed form of a synthetic point echo (coherent sum = elements in the
  aperture times the echo amplitude at the target pixel). Agreement with ``pymust.dasmtx`` is unpinned.
"""

from __future__ import annotations

import numpy as np
from scipy import signal


def rf2iq(rf, fs, fc, bandwidth=None, order=5):
    """IQ of RF (n_samples, ...) along axis 0, complex128."""
    rf = np.asarray(rf, dtype=np.float64)
    n = rf.shape[0]
    t = np.arange(n) / fs
    mix = np.exp(-1j * 2 * np.pi * fc * t).reshape((n,) + (1,) * (rf.ndim - 1))
    wn = min(2 * fc / fs, 0.5) if bandwidth is None else bandwidth / fs
    b, a = signal.butter(order, wn)
    return signal.filtfilt(b, a, rf * mix, axis=0) * 2


def das(iq, delays, ex, ez, px, pz, fs, fc, c, f_number, t0=0.0):
    """Compound image (sum over firings) of IQ (T, N, E) at pixels px, pz (flat), complex128."""
    iq = np.asarray(iq, dtype=np.complex128)
    T, N, E = iq.shape
    d = np.atleast_2d(np.asarray(delays, dtype=np.float64))
    px = np.asarray(px, dtype=np.float64).reshape(-1)
    pz = np.asarray(pz, dtype=np.float64).reshape(-1)
    r = np.sqrt((px[:, None] - ex[None]) ** 2 + (pz[:, None] - ez[None]) ** 2)  # (P, E)
    amax = np.abs(pz) / (2 * f_number) if f_number > 0 else np.full(px.shape, np.inf)
    inside = np.abs(px[:, None] - ex[None]) <= amax[:, None]
    img = np.zeros(px.size, dtype=np.complex128)
    for t in range(T):
        fin = np.isfinite(d[t])
        ttx = np.min(np.where(fin[None], d[t][None] + r / c, np.inf), axis=1)
        tau = ttx[:, None] + r / c
        si = (tau - t0) * fs
        ok = inside & (si >= 0) & (si <= N - 1) & np.isfinite(ttx)[:, None]
        i0 = np.clip(np.floor(si).astype(np.int64), 0, max(N - 2, 0))
        w = si - i0
        e_idx = np.broadcast_to(np.arange(E)[None], si.shape)
        a = iq[t][i0, e_idx]
        b = iq[t][np.minimum(i0 + 1, N - 1), e_idx]
        v = (a + w * (b - a)) * np.exp(1j * 2 * np.pi * fc * tau)
        img += np.where(ok, v, 0).sum(axis=1)
    return img
