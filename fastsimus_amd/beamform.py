"""IQ demodulation and delay-and-sum beamforming of SIMUS RF on the device.

The reference's multi-firing harness (``tests/_picmus_phantom_shared.py:407-460``) turns every
firing of the RF stack into IQ with ``pymust.rf2iq`` and beamforms it with a ``pymust.dasmtx``
matrix, summing the firings into one complex image. ``beamform`` runs that pipeline here: ``rf2iq``
(``fs_rf2iq``) and ``das`` (``fs_das``) on the ``(n_firings, n_samples, n_elements)`` RF that
``simus_compute`` returns for 2-D delays (or on an ``rf_stack``), with the firing sum in the kernel.

``rf2iq``: down-mixing by exp(-i 2 pi fc t) and a zero-phase 5th-order Butterworth low-pass with
normalised cutoff min(2 fc / fs, 0.5) (or bandwidth / fs), doubled: the recipe of MUST's ``rf2iq``
that PyMUST ports (Butterworth designed here; forward-backward filtering with odd extension and
steady-state initial conditions, the algorithm of ``filtfilt``).

``das``: transmit arrival min_e(delay_e + |p - x_e| / c), receive path |p - x_e| / c, linear
interpolation of IQ, phase rotation exp(i 2 pi fc tau), receive aperture of f-number F
(``f_number_for`` picks F from the element directivity when it is not given).

PyMUST is not installed, so agreement with ``pymust.rf2iq`` / ``pymust.dasmtx`` is not checked; the
kernels are pinned by library routines (SciPy's Butterworth design and ``filtfilt``) and closed forms
(a tone at fc demodulates to a constant, linearity, a point target's image peaks at its position).
"""

from __future__ import annotations

from math import asin, cos, pi, sin, tan

import numpy as np
import torch

from . import _lib
from . import kernels as K
from .geometry import element_positions
from .params import TransducerParams


# ----------------------------------------------------------------------------- filter design (host, float64)
def butterworth_sos(order: int, wn: float) -> np.ndarray:
    """Digital Butterworth low-pass as second-order sections ``(n_sections, 6)`` = b0 b1 b2 a0 a1 a2.

    Analog prototype poles exp(i pi (2k + order + 1) / (2 order)), cutoff prewarped to 4 tan(pi wn / 2)
    (normalised frequency wn in (0, 1), 1 = Nyquist), bilinear transform (zeros at z = -1); conjugate pole
    pairs make the biquads, a real pole (odd order) the last first-order section; every section has unit
    gain at DC, so the cascade does too.
    """
    if not (0.0 < wn < 1.0) or order < 1:
        raise ValueError("order >= 1 and 0 < wn < 1 required")
    k = np.arange(order)
    analog = np.exp(1j * pi * (2 * k + order + 1) / (2 * order))
    warped = 4.0 * tan(pi * wn / 2.0)
    z = (4.0 + warped * analog) / (4.0 - warped * analog)
    upper = sorted([p for p in z if p.imag > 1e-12], key=lambda p: -abs(p))
    real = [p.real for p in z if abs(p.imag) <= 1e-12]
    sos = []
    for p in upper:
        a1, a2 = -2.0 * p.real, abs(p) ** 2
        g = (1.0 + a1 + a2) / 4.0
        sos.append([g, 2 * g, g, 1.0, a1, a2])
    for p in real:
        g = (1.0 - p) / 2.0
        sos.append([g, g, 0.0, 1.0, -p, 0.0])
    return np.asarray(sos, dtype=np.float64)


def section_steady_states(sos: np.ndarray) -> np.ndarray:
    """States (transposed direct form II) of each section after a unit constant input to that section:
    y = G (its DC gain), z1 = G - b0, z2 = b2 - a2 G."""
    out = []
    for b0, b1, b2, a0, a1, a2 in sos:
        b0, b1, b2, a1, a2 = b0 / a0, b1 / a0, b2 / a0, a1 / a0, a2 / a0
        g = (b0 + b1 + b2) / (1.0 + a1 + a2)
        out.append([g - b0, b2 - a2 * g])
    return np.asarray(out, dtype=np.float64)


def rf2iq_filter(fs: float, fc: float, bandwidth: float | None = None, order: int = 5) -> _lib.Iir:
    """The ``fs_iir`` of ``rf2iq``: cutoff min(2 fc / fs, 0.5) of Nyquist (bandwidth / fs when a bandwidth in Hz
    is given), odd-extension length 3 (order + 1) at each end (that of ``filtfilt`` on the transfer function)."""
    wn = min(2.0 * fc / fs, 0.5) if bandwidth is None else float(bandwidth) / fs
    sos = butterworth_sos(order, wn)
    zi = section_steady_states(sos)
    f = _lib.Iir()
    f.n_sections = sos.shape[0]
    f.padlen = 3 * (order + 1)
    for i in range(sos.shape[0]):
        for j in range(6):
            f.sos[i][j] = float(sos[i, j])
        f.zi[i][0], f.zi[i][1] = float(zi[i, 0]), float(zi[i, 1])
    return f


def f_number_for(params: TransducerParams, fc: float | None = None, speed_of_sound: float = 1540.0) -> float:
    """Receive f-number from the element directivity: F = 1 / (2 tan alpha) with alpha the angle where the
    soft-baffle directivity cos(alpha) sinc(W sin(alpha) / lambda) falls to 1 / sqrt(2) at the centre
    frequency (elements see the pixel within their -3 dB cone)."""
    fc = params.freq_center if fc is None else fc
    ratio = params.element_width * fc / speed_of_sound

    def d(a):
        u = pi * ratio * sin(a)
        return cos(a) * (sin(u) / u if u != 0.0 else 1.0)

    lo, hi = 0.0, pi / 2 - 1e-9
    for _ in range(80):
        mid = 0.5 * (lo + hi)
        lo, hi = (mid, hi) if d(mid) > 2 ** -0.5 else (lo, mid)
    return 1.0 / (2.0 * tan(0.5 * (lo + hi)))


# ----------------------------------------------------------------------------- device calls
def _on_device(a, dtype, dev) -> torch.Tensor:
    """``a`` as a contiguous tensor of ``dtype`` on ``dev`` (a copy only where type, device or layout differ)."""
    if isinstance(a, torch.Tensor):
        return a.detach().to(device=dev, dtype=dtype).contiguous()
    return torch.as_tensor(np.asarray(a), dtype=dtype, device=dev)


def rf2iq(rf, fs: float, fc: float, bandwidth: float | None = None, *, channel_major: bool = False) -> torch.Tensor:
    """IQ of RF ``(n_samples, n_elements)`` or ``(n_firings, n_samples, n_elements)``: complex64, same shape, or
    with the last two axes swapped (``channel_major``: ``(n_firings, n_elements, n_samples)``, each sequence
    contiguous, the layout ``das`` reads fastest)."""
    dev = K.hip_device(rf)
    x = rf.detach() if isinstance(rf, torch.Tensor) else torch.as_tensor(np.asarray(rf))
    x = x.to(device=dev, dtype=torch.float32).contiguous()
    squeeze = x.dim() == 2
    if squeeze:
        x = x[None]
    if x.dim() != 3:
        raise ValueError("rf must be (n_samples, n_elements) or (n_firings, n_samples, n_elements)")
    T, N, E = x.shape
    f = rf2iq_filter(fs, fc, bandwidth)
    if N <= f.padlen:
        raise ValueError(f"rf2iq needs more than {f.padlen} samples, got {N}")
    lib = _lib.load()
    import ctypes as C

    out = torch.empty(((T, E, N, 2) if channel_major else (T, N, E, 2)), dtype=torch.float32, device=dev)
    ws = K._workspace(lib.fs_rf2iq_workspace_bytes(T, E, C.byref(f)), dev)
    st = lib.fs_rf2iq(x.data_ptr(), T, N, E, float(fs), float(fc), C.byref(f), 1 if channel_major else 0,
                      out.data_ptr(), ws.data_ptr(), ws.numel(), K._stream(dev))
    _lib.check(st, "fs_rf2iq")
    iq = torch.view_as_complex(out)
    return iq[0] if squeeze else iq


def das(iq, delays, params: TransducerParams, x, z, fs: float, fc: float | None = None, *,
        speed_of_sound: float = 1540.0, f_number: float | None = None, t0: float = 0.0,
        compound: bool = True, channel_major: bool = False) -> torch.Tensor:
    """Delay-and-sum of IQ ``(n_firings, n_samples, n_elements)`` (or one firing ``(n_samples, n_elements)``;
    ``(n_firings, n_elements, n_samples)`` with ``channel_major``, as ``rf2iq(..., channel_major=True)`` writes
    it) fired with ``delays`` ``(n_firings, n_elements)`` (or ``(n_elements,)``) onto pixels ``(x, z)`` (arrays
    of one shape). Returns complex64 of the pixels' shape, summed over firings (``compound``), or with a
    leading firing axis."""
    dev = K.hip_device(iq)
    q = iq.detach() if isinstance(iq, torch.Tensor) else torch.as_tensor(np.asarray(iq))
    q = q.to(device=dev, dtype=torch.complex64)
    if q.dim() == 2:
        q = q[None]
    q = torch.view_as_real(q.contiguous()).contiguous()
    if channel_major:
        T, E, N, _ = q.shape
    else:
        T, N, E, _ = q.shape
    if E != params.n_elements:
        raise ValueError(f"IQ has {E} channels for {params.n_elements} elements")
    # delays and pixels that are already device tensors stay on the device (no host round trip)
    dd = _on_device(delays, torch.float64, dev)
    if dd.dim() == 1:
        dd = dd[None]
    if tuple(dd.shape) != (T, E):
        raise ValueError(f"delays must be ({T}, {E}) for this IQ, got {tuple(dd.shape)}")
    px = _on_device(x, torch.float32, dev)
    pz = _on_device(z, torch.float32, dev)
    if px.shape != pz.shape:
        raise ValueError("x and z must have one shape")
    shape = tuple(px.shape)
    px, pz = px.reshape(-1), pz.reshape(-1)
    pos, _, _ = element_positions(params.n_elements, params.pitch, params.radius)
    fc = params.freq_center if fc is None else float(fc)
    F = f_number_for(params, fc, speed_of_sound) if f_number is None else float(f_number)
    ex = torch.as_tensor(pos[:, 0], dtype=torch.float64, device=dev)
    ez = torch.as_tensor(pos[:, 1], dtype=torch.float64, device=dev)
    P = px.numel()
    out = torch.empty(((1 if compound else T), P, 2), dtype=torch.float32, device=dev)
    lib = _lib.load()
    st = lib.fs_das(q.data_ptr(), T, N, E, 1 if channel_major else 0, ex.data_ptr(), ez.data_ptr(), dd.data_ptr(),
                    px.data_ptr(), pz.data_ptr(), P, float(fs), fc, float(speed_of_sound), float(t0), F,
                    1 if compound else 0, out.data_ptr(), K._stream(dev))
    _lib.check(st, "fs_das")
    img = torch.view_as_complex(out)
    return img[0].reshape(shape) if compound else img.reshape(T, *shape)


def beamform(rf, delays, params: TransducerParams, x, z, fs: float, fc: float | None = None, *,
             speed_of_sound: float = 1540.0, f_number: float | None = None, bandwidth: float | None = None,
             t0: float = 0.0) -> torch.Tensor:
    """The harness's reconstruction: IQ of every firing, beamformed and summed over firings.

    ``rf`` is ``(n_firings, n_samples, n_elements)`` (``simus_compute(...).rf`` for 2-D delays) or an
    ``rf_stack`` ``(n_samples, n_elements, n_firings)`` given with ``delays`` of ``n_firings`` rows."""
    fc = params.freq_center if fc is None else float(fc)
    r = rf if isinstance(rf, torch.Tensor) else torch.as_tensor(np.asarray(rf))
    d = np.atleast_2d(np.asarray(delays.detach().cpu() if isinstance(delays, torch.Tensor) else delays))
    if r.dim() == 3 and r.shape[0] != d.shape[0] and r.shape[2] == d.shape[0]:
        r = r.permute(2, 0, 1)  # rf_stack layout (samples, channels, firings)
    iq = rf2iq(r, fs, fc, bandwidth, channel_major=True)
    return das(iq, d, params, x, z, fs, fc, speed_of_sound=speed_of_sound, f_number=f_number, t0=t0,
               channel_major=True)


__all__ = ["beamform", "butterworth_sos", "das", "f_number_for", "rf2iq", "rf2iq_filter", "section_steady_states"]
