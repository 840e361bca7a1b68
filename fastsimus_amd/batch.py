"""Batches of phantoms: one frequency plan, one persistent launch, optional graph replay.

``simus_batch(phantoms, delays, params, medium)`` simulates every phantom
(scatterers ``(S_q, 2)`` and reflectivities ``(S_q,)``, sizes free to differ)
under every transmit of ``delays`` (``(n_tx, n_elements)`` or ``(n_elements,)``)
and returns ``rf`` of shape ``(n_phantoms, n_tx, n_samples, n_elements)``.
All phantoms share one plan, taken from the largest scatterer distance over
the whole batch and the largest delay over all transmits (the multi-firing
convention of the reference harness, ``tests/_picmus_phantom_shared.py:331-350``,
extended to phantoms), so phantom q's RF equals
``simus_compute(phantom_q, ..., plan)`` with that shared plan.

With ``strategy="hip_mfma"`` the whole batch is one persistent kernel launch
whose work queue runs the largest phantoms first; ``"hip"`` runs the plain
FP32 kernels phantom by phantom. ``SimusBatchGraph`` captures one batched call
(spectrum and RF) in a HIP graph and replays it on new scatterer positions and
reflectivities of the same sizes.

The reference has no batched-phantom API: its harness calls ``simus`` once
per phantom and firing.
"""

from __future__ import annotations

from typing import NamedTuple, Sequence

import numpy as np
import torch

from . import kernels as K
from .params import MediumParams, TransducerParams
from .plan import SimusPlan, simus_precompute
from .simus import _method

MAX_PHANTOMS_PER_CALL = 192  # fs_simus_spectrum_batch limit; larger batches are split


def max_phantoms_per_call(n_elements: int, method: str) -> int:
    """Phantoms per fs_simus_spectrum_batch call: 192 work units per launch, and above 128 elements the
    batched method takes one unit per pair of 64-element groups for every phantom."""
    groups = -(-n_elements // 64)
    pairs = 1 if groups <= 2 else groups * (groups - 1) // 2
    return MAX_PHANTOMS_PER_CALL if method != "mfma" else max(1, MAX_PHANTOMS_PER_CALL // pairs)


class BatchResult(NamedTuple):
    """``rf`` (n_phantoms, n_tx, n_samples, n_elements) float32; ``spectrum`` (n_phantoms, n_tx,
    n_freq_full, n_elements) complex64 or None; ``offsets`` (n_phantoms + 1,) into the concatenation."""

    rf: torch.Tensor
    spectrum: torch.Tensor | None
    offsets: np.ndarray


def _as_points(a):
    if isinstance(a, torch.Tensor):
        return a.detach().reshape(-1, 2)
    return np.asarray(a).reshape(-1, 2)


def _concat(phantoms: Sequence):
    """(points (S, 2), rc (S,), offsets (P + 1,)) of the concatenated batch."""
    if len(phantoms) == 0:
        raise ValueError("empty batch of phantoms")
    pts, rcs, sizes = [], [], []
    torch_in = any(isinstance(p[0], torch.Tensor) or isinstance(p[1], torch.Tensor) for p in phantoms)
    for sc, rc in phantoms:
        sc = _as_points(sc)
        n = sc.shape[0]
        rc_n = rc.numel() if isinstance(rc, torch.Tensor) else np.size(rc)
        if rc_n != n:
            raise ValueError(f"phantom with {n} scatterers has {rc_n} reflectivities")
        if torch_in:
            sc = torch.as_tensor(sc, dtype=torch.float32)
            rc = torch.as_tensor(rc, dtype=torch.float32).reshape(-1)
        else:
            sc = np.asarray(sc, dtype=np.float32)
            rc = np.asarray(rc, dtype=np.float32).reshape(-1)
        pts.append(sc)
        rcs.append(rc)
        sizes.append(n)
    offsets = np.zeros(len(sizes) + 1, dtype=np.int64)
    offsets[1:] = np.cumsum(sizes)
    if torch_in:
        dev = next((t.device for t in pts + rcs if t.is_cuda), pts[0].device)
        return (torch.cat([p.to(dev) for p in pts]), torch.cat([r.to(dev) for r in rcs]), offsets)
    return np.concatenate(pts), np.concatenate(rcs), offsets


def simus_batch_precompute(phantoms: Sequence, delays, params: TransducerParams, medium: MediumParams = MediumParams(),
                           **plan_kw) -> SimusPlan:
    """One plan for the batch (``simus_precompute`` over all phantoms' scatterers)."""
    pts, rc, _ = _concat(phantoms)
    return simus_precompute(pts, rc, delays, params, medium, **plan_kw)


def _problem(pts, rc, delays, apod, plan: SimusPlan, params, medium, fulldir, device):
    return K.Problem(pts, rc, delays, apod, params, medium, n_sub=plan.n_sub, seg_length=plan.seg_length,
                     freqs=plan.selected_freqs, pulse=plan.pulse_spectrum, probe=plan.probe_spectrum,
                     correction=plan.correction_factor, full_frequency_directivity=fulldir, device=device)


def simus_batch_compute(phantoms: Sequence, delays, plan: SimusPlan, params: TransducerParams,
                        medium: MediumParams = MediumParams(), *, tx_apodization=None,
                        full_frequency_directivity: bool = False, strategy="hip_mfma",
                        want_spectrum: bool = False) -> BatchResult:
    """RF of every (phantom, transmit) of the batch for a shared plan."""
    method = _method(strategy)
    pts, rc, offsets = _concat(phantoms)
    dev = K.hip_device(pts, rc, delays, tx_apodization)
    rfs, fulls = [], []
    per_call = max_phantoms_per_call(params.n_elements, method)
    for q0 in range(0, len(offsets) - 1, per_call):
        q1 = min(q0 + per_call, len(offsets) - 1)
        b, e = int(offsets[q0]), int(offsets[q1])
        prob = _problem(pts[b:e], rc[b:e], delays, tx_apodization, plan, params, medium, full_frequency_directivity,
                        dev)
        spect = K.run_spectrum_batch(prob, offsets[q0:q1 + 1] - b, method)
        P, T, NF, E = spect.shape
        rf, full, _ = K.run_rf(spect.reshape(P * T, NF, E), plan.n_freq_full, plan.freq_idx_start, plan.n_fft,
                               want_full=want_spectrum)
        rfs.append(rf.reshape(P, T, *rf.shape[1:]))
        if want_spectrum:
            fulls.append(full.reshape(P, T, *full.shape[1:]))
    rf = torch.cat(rfs) if len(rfs) > 1 else rfs[0]
    full = (torch.cat(fulls) if len(fulls) > 1 else fulls[0]) if want_spectrum else None
    return BatchResult(rf=rf, spectrum=full, offsets=offsets)


def simus_batch(phantoms: Sequence, delays, params: TransducerParams, medium: MediumParams = MediumParams(), *,
                fs: float | None = None, tx_apodization=None, tx_n_wavelengths: float | int = 1.0,
                db_thresh: float | int = -60.0, full_frequency_directivity: bool = False,
                element_splitting: int | None = None, frequency_step: float | int = 1.0,
                strategy="hip_mfma", want_spectrum: bool = False) -> BatchResult:
    """Plan once for the batch, then ``simus_batch_compute``."""
    plan = simus_batch_precompute(phantoms, delays, params, medium, fs=fs, tx_n_wavelengths=tx_n_wavelengths,
                                  db_thresh=db_thresh, element_splitting=element_splitting,
                                  frequency_step=frequency_step)
    return simus_batch_compute(phantoms, delays, plan, params, medium, tx_apodization=tx_apodization,
                               full_frequency_directivity=full_frequency_directivity, strategy=strategy,
                               want_spectrum=want_spectrum)


class SimusBatchGraph:
    """One batched simulation (spectrum + gated RF) captured in a HIP graph.

    The phantom sizes, transmits, plan and options are fixed at construction;
    ``replay(x, z, rc)`` copies new concatenated scatterer coordinates and
    reflectivities (float32, same total length, same phantom boundaries) into
    the graph's static inputs, replays the graph on the current stream and
    returns the static RF tensor ``(n_phantoms, n_tx, n_samples, n_elements)``
    (overwritten by the next replay). At most 192 phantoms (fewer above 128 elements with the batched
    method, ``max_phantoms_per_call``). Kernel timing
    (``fs_set_kernel_timing``) must be off while capturing.
    """

    def __init__(self, phantoms: Sequence, delays, plan: SimusPlan, params: TransducerParams,
                 medium: MediumParams = MediumParams(), *, tx_apodization=None,
                 full_frequency_directivity: bool = False, strategy="hip_mfma"):
        method = _method(strategy)
        pts, rc, offsets = _concat(phantoms)
        per_call = max_phantoms_per_call(params.n_elements, method)
        if len(offsets) - 1 > per_call:
            raise NotImplementedError(f"a captured batch holds at most {per_call} phantoms for this array")
        dev = K.hip_device(pts, rc, delays, tx_apodization)
        self.offsets = offsets
        self.plan = plan
        self.prob = _problem(pts, rc, delays, tx_apodization, plan, params, medium, full_frequency_directivity, dev)
        S = int(offsets[-1])
        self.x = torch.empty(S, dtype=torch.float32, device=dev)
        self.z = torch.empty(S, dtype=torch.float32, device=dev)
        self.rc = torch.empty(S, dtype=torch.float32, device=dev)
        self.x.copy_(self.prob.x)
        self.z.copy_(self.prob.z)
        self.rc.copy_(self.prob.rc)
        self.prob.bind_scatterers(self.x, self.z, self.rc)
        P = len(offsets) - 1
        self._ws = K.batch_workspace(self.prob, offsets, method)
        self._spect = torch.empty((P, self.prob.n_tx, self.prob.n_freq, params.n_elements, 2), dtype=torch.float32,
                                  device=dev)
        self._method = method

        def body():
            spect = K.run_spectrum_batch(self.prob, self.offsets, self._method, out=self._spect, workspace=self._ws)
            Pn, T, NF, E = spect.shape
            rf, _, _ = K.run_rf(spect.reshape(Pn * T, NF, E), plan.n_freq_full, plan.freq_idx_start, plan.n_fft,
                                want_full=False)
            return rf.reshape(Pn, T, *rf.shape[1:])

        side = torch.cuda.Stream(device=dev)
        side.wait_stream(torch.cuda.current_stream(dev))
        with torch.cuda.stream(side):
            body()  # warm-up outside the capture (kernel attributes, code objects)
        torch.cuda.current_stream(dev).wait_stream(side)
        torch.cuda.synchronize(dev)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self.rf = body()

    def replay(self, x=None, z=None, rc=None) -> torch.Tensor:
        """Copy new inputs (any that are given) into the static buffers and replay."""
        for dst, src in ((self.x, x), (self.z, z), (self.rc, rc)):
            if src is not None:
                src = torch.as_tensor(src, dtype=torch.float32)
                if src.numel() != dst.numel():
                    raise ValueError(f"expected {dst.numel()} values, got {src.numel()}")
                dst.copy_(src.reshape(-1))
        self.graph.replay()
        return self.rf


__all__ = ["BatchResult", "SimusBatchGraph", "simus_batch", "simus_batch_compute", "simus_batch_precompute"]
