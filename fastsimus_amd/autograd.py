"""Differentiable SIMUS: ``simus_rf`` with a hand-written adjoint.

``simus_rf(x, z, rc, speed_of_sound, delays, params, plan, ...)`` returns the
gated RF of every transmit, ``(n_tx, n_samples, n_elements)``, as a
``torch.autograd.Function``. Gradients flow to ``rc``, ``x``, ``z`` (float32,
per scatterer) and ``speed_of_sound`` (0-d tensor). The backward pass runs
``fs_rf_adjoint`` (gate and inverse DFT, reversed) and ``fs_simus_adjoint``
(Green's functions recomputed, never stored). The frequency plan is held
fixed: derivatives are taken at constant frequency grid, band and
correction factor, as for any fixed discretisation.

The reference has no gradient kernel; its only differentiable path is the
JAX scan strategy (``_simus_strategies.py:152-210``), whose algebra this
differentiates, including its per-frequency directivity branch
(``full_frequency_directivity``, ``_simus_strategies.py:177-186``; FP32 vector
adjoint kernels).
"""

from __future__ import annotations

import torch

from . import _lib
from . import kernels as K
from .params import MediumParams, TransducerParams
from .plan import SimusPlan


class _State:
    def __init__(self, params, plan, delays, apod, attenuation, method, adjoint, fulldir=False):
        self.params = params
        self.plan = plan
        self.delays = delays
        self.apod = apod
        self.attenuation = attenuation
        self.method = method
        self.adjoint = adjoint
        self.fulldir = fulldir


# fs_adjoint_method codes, passed with every fs_simus_adjoint call
_ADJOINT = {"auto": _lib.FS_ADJOINT_AUTO, "matrix": _lib.FS_ADJOINT_MATRIX, "vector": _lib.FS_ADJOINT_VECTOR}


# From this many transmits on, the multistatic-matrix (MFMA) forward, whose cost does not grow
# with the transmit count, is faster than the per-transmit FP32 kernels.
MFMA_MIN_TX = 3
MFMA_MAX_ELEMENTS = 1024  # element limit of the batched (multistatic-matrix) method


def _problem(st: _State, x, z, rc, c: float):
    dev = x.device
    medium = MediumParams(speed_of_sound=c, attenuation=st.attenuation)
    pts = torch.stack([x.detach(), z.detach()], -1)
    return K.Problem(pts, rc.detach(), st.delays, st.apod, st.params, medium, n_sub=st.plan.n_sub,
                     seg_length=st.plan.seg_length, freqs=st.plan.selected_freqs, pulse=st.plan.pulse_spectrum,
                     probe=st.plan.probe_spectrum, correction=st.plan.correction_factor,
                     full_frequency_directivity=st.fulldir, device=dev)


class _SimusRF(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, z, rc, c, st: _State):
        prob = _problem(st, x, z, rc, float(c))
        spect = K.run_spectrum(prob, st.method)
        rf, _, rf_u = K.run_rf(spect, st.plan.n_freq_full, st.plan.freq_idx_start, st.plan.n_fft, want_full=False,
                               want_ungated=True)
        ctx.st = st
        ctx.c = float(c)
        ctx.c_meta = (c.dtype, c.device) if isinstance(c, torch.Tensor) else (torch.float64, x.device)
        ctx.save_for_backward(x, z, rc, rf_u)
        return rf

    @staticmethod
    def backward(ctx, d_rf):
        x, z, rc, rf_u = ctx.saved_tensors
        st = ctx.st
        lib = _lib.load()
        prob = _problem(st, x, z, rc, ctx.c)
        T, nkeep, E = rf_u.shape
        nf = prob.n_freq
        dev = x.device
        stream = K._stream(dev)
        d_rf = d_rf.contiguous().to(torch.float32)
        d_spect = torch.empty((T, nf, E, 2), dtype=torch.float32, device=dev)
        ws = K._workspace(lib.fs_rf_adjoint_workspace_bytes(T, nkeep, E), dev)
        s = lib.fs_rf_adjoint(rf_u.data_ptr(), d_rf.data_ptr(), T, nf, E, st.plan.freq_idx_start, st.plan.n_fft,
                              d_spect.data_ptr(), ws.data_ptr(), ws.numel(), stream)
        _lib.check(s, "fs_rf_adjoint")
        S = x.numel()
        d_rc = torch.empty(S, dtype=torch.float32, device=dev)
        d_x = torch.empty(S, dtype=torch.float32, device=dev)
        d_z = torch.empty(S, dtype=torch.float32, device=dev)
        d_c = torch.empty(1, dtype=torch.float64, device=dev)
        method = _ADJOINT[st.adjoint]
        ws2 = K._workspace(lib.fs_simus_adjoint_workspace_bytes(prob.dref, method), dev)
        s = lib.fs_simus_adjoint(prob.dref, method, d_spect.data_ptr(), d_rc.data_ptr(), d_x.data_ptr(),
                                 d_z.data_ptr(), d_c.data_ptr(), ws2.data_ptr(), ws2.numel(), stream)
        _lib.check(s, "fs_simus_adjoint")
        dtype, cdev = ctx.c_meta
        g_c = d_c.reshape(()).to(dtype=dtype, device=cdev)
        return (d_x.to(x.dtype).reshape(x.shape), d_z.to(z.dtype).reshape(z.shape), d_rc.to(rc.dtype).reshape(rc.shape),
                g_c, None)


def simus_rf(x: torch.Tensor, z: torch.Tensor, rc: torch.Tensor, speed_of_sound, delays, params: TransducerParams,
             plan: SimusPlan, *, attenuation: float = 0.0, tx_apodization=None, method: str = "auto",
             adjoint: str = "auto", full_frequency_directivity: bool = False) -> torch.Tensor:
    """Gated RF ``(n_tx, n_samples, n_elements)``, differentiable in x, z, rc and speed_of_sound.

    ``method`` selects the forward: ``"fp32"`` (per-transmit FP32 kernels), ``"mfma"`` (split-BF16
    multistatic matrix, at most 1024 elements) or ``"auto"`` (``"mfma"`` from ``MFMA_MIN_TX``
    transmits on where it applies). ``adjoint`` selects the backward kernels: ``"matrix"`` (element
    contractions as split-BF16 MFMA products, Green's functions generated once for up to 16
    transmits; at most 128 elements and 2 sub-elements), ``"vector"`` (FP32, Green's functions
    regenerated per transmit; any size) or ``"auto"`` (``"matrix"`` where it applies).
    ``full_frequency_directivity`` evaluates the element directivity at every frequency instead of the
    centre frequency (``_simus_strategies.py:177-186``); its gradient runs on the vector kernels
    (``adjoint="matrix"`` raises ``NotImplementedError``).
    """
    if adjoint not in _ADJOINT:
        raise ValueError(f"unknown adjoint {adjoint!r}")
    if not isinstance(speed_of_sound, torch.Tensor):
        speed_of_sound = torch.tensor(float(speed_of_sound), dtype=torch.float64)
    dev = K.hip_device(x, z, rc)
    x = x.to(device=dev, dtype=torch.float32)
    z = z.to(device=dev, dtype=torch.float32)
    rc = rc.to(device=dev, dtype=torch.float32)
    if method == "auto":
        n_tx = 1 if (delays.dim() if isinstance(delays, torch.Tensor) else len(getattr(delays, "shape", ()))) < 2 \
            else int(delays.shape[0])
        method = "mfma" if (n_tx >= MFMA_MIN_TX and params.n_elements <= MFMA_MAX_ELEMENTS) else "fp32"
    if method not in ("fp32", "mfma"):
        raise ValueError(f"unknown method {method!r}")
    return _SimusRF.apply(x, z, rc, speed_of_sound,
                          _State(params, plan, delays, tx_apodization, attenuation, method, adjoint,
                                 bool(full_frequency_directivity)))
