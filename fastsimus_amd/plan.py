"""Frequency plans for ``simus`` and ``pfield`` (host, float64).

All data-dependent sizes are fixed here, before any device work, exactly as
the reference's ``simus_precompute`` (``simus.py:165-260``) and
``pfield_precompute`` (``pfield.py:223-298``) fix them:

* sub-element count ``n_sub = ceil(W / lambda_min)``, ``lambda_min = c / (fc (1 + B/2))``;
* frequency step from the largest scatterer-element distance (element
  centres for simus, clipped sub-element distances for pfield), the two-way
  pulse length and the largest transmit delay;
* the contiguous band whose pulse x probe magnitude exceeds ``db_thresh``;
* simus only: start bin, correction factor and inverse-FFT length.

Multi-transmit (batched) plans take the largest delay over all transmits,
which reproduces the reference's convention of planning once from the
firing with the largest delay span (``tests/_picmus_phantom_shared.py:331-350``).
"""

from __future__ import annotations

from math import ceil, inf, log2, pi
from typing import NamedTuple

import numpy as np

from .geometry import element_positions, subelement_offsets
from .params import MediumParams, TransducerParams
from .spectrum import probe_spectrum, pulse_spectrum

try:
    import torch
except ImportError:  # pragma: no cover
    torch = None


class FrequencyPlan(NamedTuple):
    selected_freqs: np.ndarray
    pulse_spectrum: np.ndarray
    probe_spectrum: np.ndarray
    freq_step: float


class SimusPlan(NamedTuple):
    """Static quantities of one ``simus_compute`` call (field names of ``simus.py:135-162``)."""

    selected_freqs: np.ndarray
    pulse_spectrum: np.ndarray
    probe_spectrum: np.ndarray
    n_sub: int
    seg_length: float
    correction_factor: float
    n_freq_full: int
    freq_idx_start: int
    n_fft: int


class PfieldPlan(NamedTuple):
    """Static quantities of one ``pfield_compute`` call (``pfield.py:58-81``)."""

    selected_freqs: np.ndarray
    pulse_spectrum: np.ndarray
    probe_spectrum: np.ndarray
    n_sub: int
    seg_length: float
    correction_factor: float


def two_way_pulse_duration(freq_center: float, bandwidth: float, tx_n_wavelengths: float) -> float:
    """Duration of pulse * probe^2 above 1/1023 of its peak (``simus.py:49-104``).

    The pulse is built by an inverse real FFT at 1 ns sampling, centred with
    fftshift and normalised; the duration is the length of the reference's
    reversed trim slice times 1 ns, reproduced with the same slice.
    """
    dt = 1e-9
    df = freq_center / tx_n_wavelengths / 32
    n = 2 ** ceil(log2(1.0 / dt / 2.0 / df))
    omega = 2.0 * pi * np.linspace(0, 1.0 / dt / 2.0, n)
    two_way = pulse_spectrum(omega, freq_center, tx_n_wavelengths) * probe_spectrum(omega, freq_center, bandwidth) ** 2
    pulse = np.fft.fftshift(np.fft.irfft(two_way))
    pulse = pulse / np.max(np.abs(pulse))
    above = np.nonzero(pulse > (1.0 / 1023))[0]
    if above.size == 0:
        return tx_n_wavelengths / freq_center
    i1, i2 = int(above[0]), int(above[-1])
    trim = min(i1 + 1, 2 * n - 1 - i2 - 1)
    return float(pulse[-trim: trim - 2: -1].shape[0] * dt)


def select_frequencies(fc: float, bandwidth: float, tx_n_wavelengths: float, db_thresh: float,
                       max_freq_step: float) -> FrequencyPlan:
    """Uniform grid on [0, 2 fc] and the contiguous band above ``db_thresh`` (``_pfield_math.py:121-162``)."""
    n_freq = int(2 * ceil(fc / max_freq_step) + 1)
    freqs = np.linspace(0, 2 * fc, n_freq)
    step = float(freqs[1])
    w = 2.0 * pi * freqs
    mag = np.abs(pulse_spectrum(w, fc, tx_n_wavelengths) * probe_spectrum(w, fc, bandwidth))
    gain = 20.0 * np.log10(1e-200 + mag / np.max(mag))
    mask = gain > db_thresh
    if not mask.any():
        first, last = 0, 0
    else:
        idx = np.nonzero(mask)[0]
        first, last = int(idx[0]), int(idx[-1])
    sel = freqs[first: last + 1]
    ws = 2.0 * pi * sel
    return FrequencyPlan(sel, pulse_spectrum(ws, fc, tx_n_wavelengths), probe_spectrum(ws, fc, bandwidth), step)


def _n_sub(params: TransducerParams, c: float, element_splitting: int | None) -> int:
    if element_splitting is not None:
        return int(element_splitting)
    lambda_min = c / (params.freq_center * (1.0 + params.bandwidth / 2.0))
    return ceil(params.element_width / lambda_min)


def _as_xz_f64(points):
    """(..., 2) NumPy array or torch tensor -> (x, z) float64 on the input's device."""
    if torch is not None and isinstance(points, torch.Tensor):
        p = points.reshape(-1, 2).to(torch.float64)
        return p[:, 0], p[:, 1]
    p = np.asarray(points, dtype=np.float64).reshape(-1, 2)
    return p[:, 0], p[:, 1]


def _max_distance_hip(x, z, ex: np.ndarray, ez: np.ndarray) -> float:
    """Device points: one pass of the library's FP64 reduction (fs_max_distance), one synchronisation."""
    from . import _lib

    lib = _lib.load()
    dev = x.device
    with torch.cuda.device(dev):
        xs = x.contiguous()
        zs = z.contiguous()
        sx = torch.as_tensor(np.ascontiguousarray(ex.reshape(-1)), dtype=torch.float64, device=dev)
        sz = torch.as_tensor(np.ascontiguousarray(ez.reshape(-1)), dtype=torch.float64, device=dev)
        out = torch.empty(1, dtype=torch.float64, device=dev)
        st = lib.fs_max_distance(xs.data_ptr(), zs.data_ptr(), xs.numel(), sx.data_ptr(), sz.data_ptr(), sx.numel(),
                                 out.data_ptr(), torch.cuda.current_stream(dev).cuda_stream)
        _lib.check(st, "fs_max_distance")
        return float(out.item())


def _max_distance(x, z, ex: np.ndarray, ez: np.ndarray, clip: float | None = None) -> float:
    """max over points and sources of the Euclidean distance (optionally clipped below)."""
    best = 0.0
    n = x.shape[0]
    chunk = max(1, (1 << 22) // max(1, ex.size))
    if torch is not None and isinstance(x, torch.Tensor) and x.is_cuda:
        best = _max_distance_hip(x, z, ex, ez)
    elif torch is not None and isinstance(x, torch.Tensor):
        tex = torch.as_tensor(ex.reshape(-1), dtype=torch.float64, device=x.device)
        tez = torch.as_tensor(ez.reshape(-1), dtype=torch.float64, device=x.device)
        for i in range(0, n, chunk):
            d = torch.sqrt((x[i:i + chunk, None] - tex) ** 2 + (z[i:i + chunk, None] - tez) ** 2)
            best = max(best, float(d.max()))
    else:
        fx, fz = ex.reshape(-1), ez.reshape(-1)
        for i in range(0, n, chunk):
            d = np.sqrt((x[i:i + chunk, None] - fx) ** 2 + (z[i:i + chunk, None] - fz) ** 2)
            best = max(best, float(d.max()))
    if clip is not None:
        best = max(best, clip)
    return best


def _max_finite_delay(delays) -> float:
    if torch is not None and isinstance(delays, torch.Tensor):
        d = torch.nan_to_num(delays.to(torch.float64), nan=0.0)
        return float(d.max())
    d = np.asarray(delays, dtype=np.float64)
    return float(np.max(np.where(np.isnan(d), 0.0, d)))


def simus_precompute(scatterers, rc, delays, params: TransducerParams, medium: MediumParams = MediumParams(), *,
                     fs: float | None = None, tx_n_wavelengths: float | int = 1.0, db_thresh: float | int = -60.0,
                     element_splitting: int | None = None, frequency_step: float | int = 1.0) -> SimusPlan:
    """Plan of one simus call; ``delays`` may be (E,) or (N_tx, E)."""
    c = medium.speed_of_sound
    fc = params.freq_center
    if fs is None:
        fs = 4.0 * fc
    n_sub = _n_sub(params, c, element_splitting)
    seg = params.element_width / n_sub
    pos, _, _ = element_positions(params.n_elements, params.pitch, params.radius)
    x, z = _as_xz_f64(scatterers)
    max_d = _max_distance(x, z, pos[:, 0], pos[:, 1])
    if tx_n_wavelengths != float("inf"):
        max_d = max_d + two_way_pulse_duration(fc, params.bandwidth, tx_n_wavelengths) * c
    df = 1.0 / 2.0 / (2.0 * max_d / c + _max_finite_delay(delays))
    df = float(frequency_step) * df
    n_freq_full = int(2 * ceil(fc / df) + 1)
    fp = select_frequencies(fc, params.bandwidth, tx_n_wavelengths, db_thresh, df)
    step = fp.freq_step
    idx0 = round(float(fp.selected_freqs[0]) / step) if step > 0 else 0
    corr = (1.0 if tx_n_wavelengths == float("inf") else step) * params.element_width
    n_fft = ceil(fs / 2.0 / fc * (n_freq_full - 1))
    return SimusPlan(fp.selected_freqs, fp.pulse_spectrum, fp.probe_spectrum, n_sub, seg, corr, n_freq_full, idx0,
                     n_fft)


def pfield_precompute(positions, delays, params: TransducerParams, medium: MediumParams = MediumParams(), *,
                      tx_n_wavelengths: float | int = 1.0, db_thresh: float | int = -60.0,
                      element_splitting: int | None = None, frequency_step: float | int = 1.0) -> PfieldPlan:
    """Plan of one pfield call: df from clipped sub-element distances (``pfield.py:271-286``)."""
    size = positions.numel() if (torch is not None and isinstance(positions, torch.Tensor)) else np.size(positions)
    if size == 0:
        raise ValueError("Grid has no points")
    c = medium.speed_of_sound
    n_sub = _n_sub(params, c, element_splitting)
    seg = params.element_width / n_sub
    pos, theta, _ = element_positions(params.n_elements, params.pitch, params.radius)
    if theta is None:
        theta = np.zeros(params.n_elements)
    off = subelement_offsets(params.element_width, n_sub, theta)
    sx = pos[:, 0][:, None] + off[..., 0]
    sz = pos[:, 1][:, None] + off[..., 1]
    x, z = _as_xz_f64(positions)
    max_r = _max_distance(x, z, sx, sz, clip=c / params.freq_center / 2.0)
    df = 1.0 / (max_r / c + _max_finite_delay(delays))
    df = float(frequency_step) * df
    fp = select_frequencies(params.freq_center, params.bandwidth, tx_n_wavelengths, db_thresh, df)
    step = fp.freq_step
    corr = (1.0 if tx_n_wavelengths == float("inf") else step) * params.element_width
    return PfieldPlan(fp.selected_freqs, fp.pulse_spectrum, fp.probe_spectrum, n_sub, seg, corr)
