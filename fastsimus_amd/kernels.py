"""Host side of the HIP operator: torch tensors -> ``fs_simus_desc`` -> C ABI.

``simus_hip`` is the drop-in for the reference's ``simus_cuda``
(``kernels/cuda_simus.py:232-321``): same inputs (already NaN-cleaned delays
and apodization), same output ``(n_freq, n_elements)`` complex64 of the
selected band before the correction factor. ``delays_clean`` may also be
``(n_tx, n_elements)``; the output then gains a leading transmit axis.

Everything here is plumbing: tensors are placed on the HIP device, the
descriptor is filled with their pointers, and the library runs on torch's
current stream. There is no CPU path.
"""

from __future__ import annotations

import ctypes as C
from math import inf

import numpy as np
import torch

from . import _lib
from .geometry import subelement_positions
from .params import MediumParams, TransducerParams, baffle_code

METHODS = {"fp32": _lib.FS_METHOD_FP32, "mfma": _lib.FS_METHOD_MFMA_BF16X3}


def hip_device(*tensors) -> torch.device:
    """The device to run on: the first CUDA (HIP) tensor's device, else cuda:0."""
    for t in tensors:
        if isinstance(t, torch.Tensor) and t.is_cuda:
            return t.device
    if not torch.cuda.is_available():
        raise RuntimeError("fastsimus_amd runs on a ROCm GPU only: no HIP device is visible")
    return torch.device("cuda", torch.cuda.current_device())


def _dev(a, dtype, device) -> torch.Tensor:
    if isinstance(a, torch.Tensor):
        return a.detach().to(device=device, dtype=dtype).contiguous()
    return torch.as_tensor(np.asarray(a), dtype=dtype, device=device).contiguous()


def _ptr(t: torch.Tensor | None) -> int | None:
    return None if t is None else t.data_ptr()


class Problem:
    """Device tensors of one SIMUS (or PFIELD) problem and the C descriptor that points at them."""

    def __init__(self, points, rc, delays, apod, params: TransducerParams, medium: MediumParams, *, n_sub: int,
                 seg_length: float, freqs, pulse, probe, correction: float, full_frequency_directivity: bool,
                 device: torch.device):
        self.device = device
        if isinstance(points, torch.Tensor):
            pts = points.detach().reshape(-1, 2)
        else:
            pts = np.asarray(points).reshape(-1, 2)
        # scatterer coordinates are consumed in float32 (as the reference's GPU path does)
        self.x = _dev(pts[:, 0], torch.float32, device)
        self.z = _dev(pts[:, 1], torch.float32, device)
        self.rc = None if rc is None else _dev(rc, torch.float32, device).reshape(-1)
        if self.rc is not None and self.rc.numel() != self.x.numel():
            raise ValueError(f"rc has {self.rc.numel()} entries for {self.x.numel()} scatterers")
        E = params.n_elements
        d = _dev(delays, torch.float64, device)
        if d.dim() == 1:
            d = d.reshape(1, -1)
        if d.dim() != 2 or d.shape[1] != E:
            raise ValueError(f"delays must have shape (n_elements,) or (n_tx, n_elements) with n_elements={E}")
        a = torch.ones_like(d) if apod is None else _dev(apod, torch.float64, device).expand_as(d).contiguous()
        nan = torch.isnan(d)
        self.delay = torch.where(nan, torch.zeros_like(d), d).contiguous()
        self.apod = torch.where(nan, torch.zeros_like(a), a).contiguous()
        sx, sz, theta, apex = subelement_positions(params, n_sub)
        self.sub_x = torch.as_tensor(sx.reshape(-1), dtype=torch.float64, device=device)
        self.sub_z = torch.as_tensor(sz.reshape(-1), dtype=torch.float64, device=device)
        self.ecos = torch.as_tensor(np.cos(theta), dtype=torch.float64, device=device)
        self.esin = torch.as_tensor(np.sin(theta), dtype=torch.float64, device=device)
        pulse = np.asarray(pulse)
        self.pulse_re = torch.as_tensor(pulse.real.copy(), dtype=torch.float64, device=device)
        self.pulse_im = torch.as_tensor(pulse.imag.copy(), dtype=torch.float64, device=device)
        self.probe = torch.as_tensor(np.asarray(probe, dtype=np.float64), dtype=torch.float64, device=device)
        freqs = np.asarray(freqs, dtype=np.float64)
        self.n_freq = int(freqs.shape[0])
        self.n_tx = int(self.delay.shape[0])
        kind, b = baffle_code(params.baffle)
        desc = _lib.SimusDesc()
        desc.n_scat = self.x.numel()
        desc.n_elem = E
        desc.n_sub = n_sub
        desc.n_freq = self.n_freq
        desc.n_tx = self.n_tx
        desc.full_frequency_directivity = 1 if full_frequency_directivity else 0
        desc.baffle_kind = kind
        desc.scat_x, desc.scat_z, desc.rc = _ptr(self.x), _ptr(self.z), _ptr(self.rc)
        desc.sub_x, desc.sub_z = _ptr(self.sub_x), _ptr(self.sub_z)
        desc.elem_cos, desc.elem_sin = _ptr(self.ecos), _ptr(self.esin)
        desc.tx_delay, desc.tx_apod = _ptr(self.delay), _ptr(self.apod)
        desc.pulse_re, desc.pulse_im, desc.probe = _ptr(self.pulse_re), _ptr(self.pulse_im), _ptr(self.probe)
        desc.f0 = float(freqs[0])
        desc.df = float(freqs[1] - freqs[0]) if freqs.shape[0] > 1 else 0.0
        desc.speed_of_sound = medium.speed_of_sound
        desc.attenuation = medium.attenuation
        desc.freq_center = params.freq_center
        desc.seg_length = seg_length
        desc.radius = params.radius if params.radius != inf else float("inf")
        desc.apex_offset = apex
        desc.baffle_impedance = b
        desc.correction = correction
        self.desc = desc
        self.params = params

    @property
    def dref(self):
        return C.byref(self.desc)

    def bind_scatterers(self, x: torch.Tensor, z: torch.Tensor, rc: torch.Tensor | None) -> None:
        """Point the descriptor at caller-owned float32 device tensors (e.g. the static inputs of a graph)."""
        for t in (x, z) + ((rc,) if rc is not None else ()):
            if t.dtype != torch.float32 or t.device != self.device or not t.is_contiguous() or t.numel() != self.x.numel():
                raise ValueError("scatterer tensors must be contiguous float32 on the problem's device, same length")
        self.x, self.z, self.rc = x, z, rc
        self.desc.scat_x, self.desc.scat_z, self.desc.rc = _ptr(x), _ptr(z), _ptr(rc)


def _stream(device) -> int:
    return torch.cuda.current_stream(device).cuda_stream


def _workspace(nbytes: int, device) -> torch.Tensor:
    return torch.empty(max(int(nbytes), 256), dtype=torch.uint8, device=device)


def run_spectrum(prob: Problem, method: str = "fp32") -> torch.Tensor:
    """Selected-band spectrum times ``desc.correction``: complex64 (n_tx, n_freq, n_elem)."""
    lib = _lib.load()
    m = METHODS[method]
    ws_bytes = lib.fs_simus_workspace_bytes(prob.dref, m)
    if ws_bytes == 0:
        _lib.check(_lib.FS_ERR_INVALID, "fs_simus_workspace_bytes")
    ws = _workspace(ws_bytes, prob.device)
    out = torch.empty((prob.n_tx, prob.n_freq, prob.params.n_elements, 2), dtype=torch.float32, device=prob.device)
    st = lib.fs_simus_spectrum(prob.dref, m, out.data_ptr(), ws.data_ptr(), ws.numel(), _stream(prob.device))
    _lib.check(st, "fs_simus_spectrum")
    return torch.view_as_complex(out)


def _offsets_ptr(offsets: np.ndarray):
    return offsets.ctypes.data_as(C.POINTER(C.c_int64))


def batch_workspace(prob: Problem, offsets: np.ndarray, method: str = "mfma") -> torch.Tensor:
    """Workspace of ``fs_simus_spectrum_batch`` for these phantom offsets."""
    lib = _lib.load()
    n = lib.fs_simus_batch_workspace_bytes(prob.dref, METHODS[method], len(offsets) - 1, _offsets_ptr(offsets))
    if n == 0:
        _lib.check(_lib.FS_ERR_INVALID, "fs_simus_batch_workspace_bytes")
    return _workspace(n, prob.device)


def run_spectrum_batch(prob: Problem, offsets: np.ndarray, method: str = "mfma", *, out: torch.Tensor | None = None,
                       workspace: torch.Tensor | None = None) -> torch.Tensor:
    """Spectra of a batch of phantoms (concatenated in ``prob``; phantom q is
    ``[offsets[q], offsets[q + 1])``): complex64 (n_phantoms, n_tx, n_freq, n_elem), times ``desc.correction``."""
    lib = _lib.load()
    offsets = np.ascontiguousarray(offsets, dtype=np.int64)
    P = len(offsets) - 1
    ws = workspace if workspace is not None else batch_workspace(prob, offsets, method)
    if out is None:
        out = torch.empty((P, prob.n_tx, prob.n_freq, prob.params.n_elements, 2), dtype=torch.float32,
                          device=prob.device)
    st = lib.fs_simus_spectrum_batch(prob.dref, METHODS[method], P, _offsets_ptr(offsets), out.data_ptr(),
                                     ws.data_ptr(), ws.numel(), _stream(prob.device))
    _lib.check(st, "fs_simus_spectrum_batch")
    return torch.view_as_complex(out)


def run_rf(spect: torch.Tensor, n_ff: int, idx0: int, n_fft: int, *, want_full: bool = True,
           want_ungated: bool = False):
    """Place, inverse-DFT and gate: returns (rf (T, n_keep, E), full (T, n_ff, E) | None, ungated | None)."""
    lib = _lib.load()
    T, NF, E = spect.shape
    nkeep = (n_fft + 1) // 2
    dev = spect.device
    s = torch.view_as_real(spect.contiguous())
    rf = torch.empty((T, nkeep, E), dtype=torch.float32, device=dev)
    full = torch.empty((T, n_ff, E, 2), dtype=torch.float32, device=dev) if want_full else None
    ung = torch.empty((T, nkeep, E), dtype=torch.float32, device=dev) if want_ungated else None
    ws = _workspace(lib.fs_rf_workspace_bytes(T, NF, n_fft), dev)
    st = lib.fs_spectrum_to_rf(s.data_ptr(), T, NF, E, n_ff, idx0, n_fft, _ptr(full), rf.data_ptr(), _ptr(ung),
                               ws.data_ptr(), ws.numel(), _stream(dev))
    _lib.check(st, "fs_spectrum_to_rf")
    return rf, (torch.view_as_complex(full) if full is not None else None), ung


def simus_hip(scatterers, rc, params: TransducerParams, plan, medium: MediumParams, delays_clean, tx_apodization,
              *, full_frequency_directivity: bool = False, method: str = "fp32") -> torch.Tensor:
    """Drop-in for ``simus_cuda``: selected spectrum (n_freq, E) complex64, no correction factor applied."""
    dev = hip_device(scatterers, rc, delays_clean)
    prob = Problem(scatterers, rc, delays_clean, tx_apodization, params, medium, n_sub=plan.n_sub,
                   seg_length=plan.seg_length, freqs=plan.selected_freqs, pulse=plan.pulse_spectrum,
                   probe=plan.probe_spectrum, correction=1.0, full_frequency_directivity=full_frequency_directivity,
                   device=dev)
    out = run_spectrum(prob, method)
    nd = delays_clean.dim() if isinstance(delays_clean, torch.Tensor) else np.ndim(delays_clean)
    return out[0] if nd == 1 else out
