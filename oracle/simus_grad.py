"""FP64 reference gradients of the SIMUS RF (CPU, torch float64 autograd).

TEST INFRASTRUCTURE ONLY (imported by tests/). Restates the forward path of
the reference (simus.py:263-367, _pfield_math.py:71-247,
_simus_strategies.py:23-77) in its direct per-frequency form on torch
float64 tensors so that torch's autograd yields d(loss)/d(rc, x, z, c) at a
fixed frequency plan; the plan dict comes from ``simus_oracle.simus_plan``.
Pinned by central finite differences of this same function
(tests/test_adjoint.py) and by agreement of its forward RF with
``simus_oracle.simus_oracle`` (both with the centre-frequency directivity
and with ``full_frequency_directivity``, the per-frequency sinc of the
reference scan, ``_simus_strategies.py:177-186``).
"""

from __future__ import annotations

from math import pi

import numpy as np
import torch

from .simus_oracle import NEPER_TO_DB, sub_elements


def simus_rf_torch(x, z, rc, c, delays, params, attenuation, plan, apod=None, chunk=64,
                   full_frequency_directivity=False):
    """Gated RF (n_tx, n_keep, E), float64, differentiable in the tensor inputs x, z, rc, c.

    Directivity sinc(k seg / 2 sin(theta) / pi) at the centre frequency (``simus.py:310-313``), or with
    ``full_frequency_directivity`` at every selected frequency (``_simus_strategies.py:110-124, 177-186``)."""
    sx, sz, th, apex = sub_elements(params, plan["n_sub"])
    sx = torch.as_tensor(sx, dtype=torch.float64)
    sz = torch.as_tensor(sz, dtype=torch.float64)
    th = torch.as_tensor(th, dtype=torch.float64)
    fc = params.freq_center
    dx = x[:, None, None] - sx
    dz = z[:, None, None] - sz
    r = torch.sqrt(dx * dx + dz * dz)
    lam2 = c / fc / 2
    rcl = torch.where(r < lam2, lam2 * torch.ones_like(r), r)
    theta = torch.asin((dx + 1e-16) / (r + 1e-16)) - th[None, :, None]
    b = getattr(params.baffle, "value", params.baffle)
    if b == "rigid":
        obl = torch.ones_like(theta)
    elif b == "soft":
        obl = torch.cos(theta)
    else:
        obl = torch.cos(theta) / (torch.cos(theta) + float(b))
    obl = torch.where(theta.abs() >= pi / 2, torch.full_like(theta, 1e-16), obl)
    kc = 2 * pi * fc / c
    sin_t = torch.sin(theta)
    amp = obl / torch.sqrt(rcl) / plan["n_sub"]
    if not full_frequency_directivity:
        amp = amp * torch.sinc(kc * plan["seg"] / 2 * sin_t / pi)
    is_out = z < 0
    if params.radius != float("inf"):
        is_out = is_out | (x**2 + (z + apex) ** 2 <= params.radius**2)
    d = np.atleast_2d(np.asarray(delays, dtype=np.float64))
    w = np.ones_like(d) if apod is None else np.asarray(apod, dtype=np.float64) * np.ones_like(d)
    nan = np.isnan(d)
    d = torch.as_tensor(np.where(nan, 0.0, d))
    w = torch.as_tensor(np.where(nan, 0.0, w))
    f = torch.as_tensor(plan["freqs"])
    pp = torch.as_tensor(plan["pulse"] * plan["probe"])
    pr = torch.as_tensor(plan["probe"])
    alpha = attenuation / NEPER_TO_DB * 1e-4
    rows = []
    for k0 in range(0, f.shape[0], chunk):
        # G of this frequency chunk once, shared by every transmit (same algebra as a loop per transmit)
        fk = f[k0:k0 + chunk]
        ph = (fk[:, None, None, None] * rcl[None] / c)
        a_k = amp[None]
        if full_frequency_directivity:
            kf = 2 * pi * fk[:, None, None, None] / c
            a_k = a_k * torch.sinc(kf * plan["seg"] / 2 * sin_t[None] / pi)
        g = (a_k * torch.exp(-alpha * fk[:, None, None, None] * rcl[None])) * torch.exp(2j * pi * ph)
        G = g.sum(-1)  # (F, S, E)
        a = w[:, None, :] * torch.exp(2j * pi * fk[None, :, None] * d[:, None, :])  # (T, F, E)
        P = pp[None, k0:k0 + chunk, None] * torch.einsum("fse,tfe->tfs", G, a.to(G.dtype))
        P = torch.where(is_out[None, None, :], torch.zeros_like(P), P)
        X = pr[None, k0:k0 + chunk, None] * torch.einsum("tfs,fse->tfe", P * rc[None, None, :], G)
        rows.append(X)
    spec = torch.cat(rows, dim=1) * plan["corr"]  # (T, n_f, E)
    T, nf, E = spec.shape
    full = torch.zeros((T, plan["n_ff"], E), dtype=spec.dtype)
    full = torch.cat([full[:, : plan["idx0"]], spec, full[:, plan["idx0"] + nf:]], dim=1)
    rf = torch.fft.irfft(torch.conj(full), n=plan["n_fft"], dim=1)[:, : (plan["n_fft"] + 1) // 2]
    peak = rf.abs().amax(dim=(1, 2), keepdim=True)
    rel = rf.abs() / (peak + 1e-30)
    return rf * 0.5 * (1 + torch.tanh((rel - 1e-5) / 1e-6))
