"""Public PFIELD API: ``pfield``, ``pfield_precompute``, ``pfield_compute``.

RMS pressure field of one transmit on a grid (reference ``pfield.py:223-475``):
``sqrt(corr * sum_f |PP(f) sum_{e,j} G(f, e, x) A(f, e)|^2)``, zero where the
point is out of field. The Green's function and frequency recurrence are the
same device code as the SIMUS transmit kernel (``k_tx`` in mode 1).
"""

from __future__ import annotations

from enum import Enum

import numpy as np
import torch

from . import _lib
from . import kernels as K
from .params import MediumParams, TransducerParams
from .plan import PfieldPlan, pfield_precompute

_DEFAULT_MEDIUM = MediumParams()


class PfieldStrategy(str, Enum):
    """Reference strategy names; every value runs the HIP kernel except the host-only ones."""

    VECTORIZED = "vectorized"
    SCAN = "scan"
    METAL = "metal"
    HIP = "hip"


def pfield_compute(positions, delays, plan: PfieldPlan, params: TransducerParams,
                   medium: MediumParams = _DEFAULT_MEDIUM, *, tx_apodization=None,
                   full_frequency_directivity: bool = False, strategy: PfieldStrategy | str | None = None):
    """RMS pressure for a precomputed plan; output shape ``positions.shape[:-1]``."""
    if strategy is not None and PfieldStrategy(strategy) in (PfieldStrategy.SCAN, PfieldStrategy.METAL):
        raise NotImplementedError(f"strategy {PfieldStrategy(strategy).value!r} is not available on a HIP device")
    shape = tuple(positions.shape[:-1])
    dev = K.hip_device(positions, delays, tx_apodization)
    prob = K.Problem(positions, None, delays, tx_apodization, params, medium, n_sub=plan.n_sub,
                     seg_length=plan.seg_length, freqs=plan.selected_freqs, pulse=plan.pulse_spectrum,
                     probe=plan.probe_spectrum, correction=plan.correction_factor,
                     full_frequency_directivity=full_frequency_directivity, device=dev)
    if prob.n_tx != 1:
        raise ValueError("pfield takes one transmit: delays of shape (n_elements,)")
    lib = _lib.load()
    ws = K._workspace(lib.fs_pfield_workspace_bytes(prob.dref), dev)
    out = torch.empty(prob.x.numel(), dtype=torch.float32, device=dev)
    st = lib.fs_pfield(prob.dref, out.data_ptr(), ws.data_ptr(), ws.numel(), K._stream(dev))
    _lib.check(st, "fs_pfield")
    return out.reshape(shape)


def pfield(positions, delays, params: TransducerParams, medium: MediumParams = _DEFAULT_MEDIUM, *,
           tx_apodization=None, tx_n_wavelengths: float | int = 1.0, db_thresh: float | int = -60.0,
           full_frequency_directivity: bool = False, element_splitting: int | None = None,
           frequency_step: float | int = 1.0, strategy: PfieldStrategy | str | None = None):
    """RMS acoustic pressure field of a transducer array (``pfield.py:381-475``)."""
    if isinstance(positions, np.ndarray) and positions.shape[-1] != 2:
        raise ValueError("positions must have shape (*grid, 2)")
    plan = pfield_precompute(positions, delays, params, medium, tx_n_wavelengths=tx_n_wavelengths,
                             db_thresh=db_thresh, element_splitting=element_splitting, frequency_step=frequency_step)
    return pfield_compute(positions, delays, plan, params, medium, tx_apodization=tx_apodization,
                          full_frequency_directivity=full_frequency_directivity, strategy=strategy)


__all__ = ["PfieldPlan", "PfieldStrategy", "pfield", "pfield_compute", "pfield_precompute"]
