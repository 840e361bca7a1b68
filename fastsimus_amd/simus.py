"""Public SIMUS API: ``simus``, ``simus_precompute``, ``simus_compute``.

Signatures follow the reference (``simus.py:165-177, 392-403, 498-513``).
The compute step always runs the HIP kernels of ``libfastsimus.so`` on a
ROCm device; results are torch tensors on that device.

Extensions over the reference signature, all opt-in:

* ``delays`` / ``tx_apodization`` of shape ``(n_tx, n_elements)`` simulate
  ``n_tx`` transmits that share one plan (built from the largest delay, which
  is the reference's multi-firing convention); ``rf`` and ``spectrum`` then
  carry a leading transmit axis.
* ``strategy`` selects the device method: ``SimusStrategy.HIP`` (plain FP32,
  one transmit at a time; also what ``CUDA`` maps to) or
  ``SimusStrategy.HIP_MFMA`` (split-BF16 matrix-core path that shares one
  multistatic matrix per frequency between all transmits). ``None`` (the
  default) takes the matrix-core path from ``AUTO_MFMA_MIN_TX`` transmits on
  (its cost hardly depends on the transmit count, the plain path's grows
  with it; with fewer transmits the plain path is faster) for arrays of up
  to 1024 elements, else the plain path. The reference's host strategies
  ``PYTHON`` / ``SCAN`` / ``METAL`` do not exist on this device and raise
  ``NotImplementedError``.

Intentional divergence: ``scatterers`` of shape ``(*batch, 2)`` with more
than one batch axis are flattened correctly (the reference only handles
``(S, 2)`` in ``simus_compute``, ``simus.py:431-434``).
"""

from __future__ import annotations

from enum import Enum
from typing import NamedTuple

import numpy as np
import torch

from . import kernels as K
from .params import MediumParams, TransducerParams
from .plan import SimusPlan, simus_precompute

_DEFAULT_MEDIUM = MediumParams()


class SimusStrategy(str, Enum):
    """Device method of the frequency sweep (reference names kept for compatibility)."""

    PYTHON = "python"
    SCAN = "scan"
    METAL = "metal"
    CUDA = "cuda"
    HIP = "hip"
    HIP_MFMA = "hip_mfma"


class SimusResult(NamedTuple):
    """``rf`` (n_samples, n_elements) float32 and ``spectrum`` (n_freq_full, n_elements) complex64."""

    rf: torch.Tensor
    spectrum: torch.Tensor


AUTO_MFMA_MIN_TX = 3      # transmits from which strategy=None takes the matrix-core path
AUTO_MFMA_MAX_ELEMENTS = 1024


def _method(strategy, n_tx: int = 1, n_elements: int = 0) -> str:
    if strategy is None:
        return "mfma" if n_tx >= AUTO_MFMA_MIN_TX and n_elements <= AUTO_MFMA_MAX_ELEMENTS else "fp32"
    s = SimusStrategy(strategy)
    if s in (SimusStrategy.HIP, SimusStrategy.CUDA):
        return "fp32"
    if s == SimusStrategy.HIP_MFMA:
        return "mfma"
    raise NotImplementedError(f"strategy {s.value!r} is a host strategy of the reference; this package runs the HIP kernels only")


def _ndim(a) -> int:
    return a.dim() if isinstance(a, torch.Tensor) else np.ndim(a)


def simus_compute(scatterers, rc, delays, plan: SimusPlan, params: TransducerParams,
                  medium: MediumParams = _DEFAULT_MEDIUM, *, tx_apodization=None,
                  full_frequency_directivity: bool = False, strategy: SimusStrategy | str | None = None) -> SimusResult:
    """RF and full spectrum for a precomputed plan (``simus.py:392-495``)."""
    n_tx = int(delays.shape[0]) if _ndim(delays) == 2 else 1
    method = _method(strategy, n_tx, params.n_elements)
    dev = K.hip_device(scatterers, rc, delays, tx_apodization)
    prob = K.Problem(scatterers, rc, delays, tx_apodization, params, medium, n_sub=plan.n_sub,
                     seg_length=plan.seg_length, freqs=plan.selected_freqs, pulse=plan.pulse_spectrum,
                     probe=plan.probe_spectrum, correction=plan.correction_factor,
                     full_frequency_directivity=full_frequency_directivity, device=dev)
    spect = K.run_spectrum(prob, method)
    rf, full, _ = K.run_rf(spect, plan.n_freq_full, plan.freq_idx_start, plan.n_fft)
    if _ndim(delays) == 1:
        return SimusResult(rf=rf[0], spectrum=full[0])
    return SimusResult(rf=rf, spectrum=full)


def simus(scatterers, rc, delays, params: TransducerParams, medium: MediumParams = _DEFAULT_MEDIUM, *,
          fs: float | None = None, tx_apodization=None, tx_n_wavelengths: float | int = 1.0,
          db_thresh: float | int = -60.0, full_frequency_directivity: bool = False,
          element_splitting: int | None = None, frequency_step: float | int = 1.0,
          strategy: SimusStrategy | str | None = None) -> SimusResult:
    """Simulate RF signals of a linear or convex array (``simus.py:498-567``)."""
    plan = simus_precompute(scatterers, rc, delays, params, medium, fs=fs, tx_n_wavelengths=tx_n_wavelengths,
                            db_thresh=db_thresh, element_splitting=element_splitting, frequency_step=frequency_step)
    return simus_compute(scatterers, rc, delays, plan, params, medium, tx_apodization=tx_apodization,
                         full_frequency_directivity=full_frequency_directivity, strategy=strategy)


def rf_stack(firings, *, n_samples: int | None = None) -> torch.Tensor:
    """RF of several firings as one ``(n_samples, n_elements, n_firings)`` array.

    The layout of the reference's multi-firing harness (``tests/_picmus_phantom_shared.py:280-295``:
    samples, channels, firings; every firing zero-padded at the end of the sample axis to the longest
    one, and the stack padded further, never cut, to ``n_samples`` when that is larger).
    ``firings`` is a ``(n_firings, n_samples, n_elements)`` tensor (the ``rf`` of ``simus_compute``
    with 2-D delays, or one phantom of ``simus_batch``) or a sequence of ``(n_samples_i, n_elements)``
    arrays. The result is float32 (or the firings' wider floating type) on the firings' device.
    """
    if isinstance(firings, torch.Tensor) and firings.dim() == 3:
        items = list(firings.unbind(0))
    else:
        items = [f if isinstance(f, torch.Tensor) else torch.as_tensor(np.asarray(f)) for f in firings]
    if not items:
        raise ValueError("rf_stack needs at least one firing")
    n_el = items[0].shape[1]
    if any(f.dim() != 2 or f.shape[1] != n_el for f in items):
        raise ValueError("every firing must be (n_samples, n_elements) with the same n_elements")
    dtype = torch.promote_types(torch.float32, items[0].dtype)
    for f in items[1:]:
        dtype = torch.promote_types(dtype, f.dtype)
    n = max(f.shape[0] for f in items)
    if n_samples is not None:
        n = max(n, int(n_samples))
    out = torch.zeros((n, n_el, len(items)), dtype=dtype, device=items[0].device)
    for i, f in enumerate(items):
        out[: f.shape[0], :, i] = f.to(device=out.device, dtype=dtype)
    return out


__all__ = ["SimusPlan", "SimusResult", "SimusStrategy", "rf_stack", "simus", "simus_compute", "simus_precompute"]
