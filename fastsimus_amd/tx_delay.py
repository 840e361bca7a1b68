"""Transmit delay laws (host, float64).

Same conventions as ``tx_delay.py:31-276`` of the reference: x lateral, z
axial (positive into tissue), delays shifted so their minimum is zero. Inputs
may be NumPy arrays or torch tensors; the result has the input's type.
"""

from __future__ import annotations

from math import cos, fmod, inf, pi, sin, sqrt, tan

import numpy as np

try:  # torch is optional for these host helpers
    import torch
except ImportError:  # pragma: no cover
    torch = None


def _to_np(a):
    if torch is not None and isinstance(a, torch.Tensor):
        return a.detach().cpu().double().numpy(), a
    return np.asarray(a, dtype=np.float64), None


def _like(out: np.ndarray, ref):
    if ref is not None:
        return torch.as_tensor(out, dtype=ref.dtype if ref.dtype.is_floating_point else torch.float64,
                               device=ref.device)
    return out


def focused(element_positions, focus, *, speed_of_sound: float, radius: float = inf,
            apex_offset: float = 0.0):
    """Delays of a converging (z > 0) or diverging (z < 0) spherical wave."""
    pos, ref = _to_np(element_positions)
    f, _ = _to_np(focus)
    dist = np.sqrt(np.sum((pos - f) ** 2, axis=-1))
    if radius == inf:
        sgn = np.sign(f[-1])
    else:
        if f.shape[0] != 2:
            raise ValueError(f"Convex arrays require 2D coordinates, got dim={f.shape[0]}")
        sgn = np.sign(f[0] ** 2 + (f[-1] + apex_offset) ** 2 - radius**2)
    sgn = -1.0 if sgn == 0 else float(sgn)
    d = -dist * sgn / speed_of_sound
    return _like(d - d.min(), ref)


def plane_wave(element_positions, tilt_rad: float, *, speed_of_sound: float, radius: float = inf,
               apex_offset: float = 0.0):
    """Delays of a plane wave tilted by ``tilt_rad`` from the array normal."""
    if abs(tilt_rad) >= pi / 2:
        raise ValueError("Tilt angles must satisfy |tilt_rad| < π/2")
    pos, ref = _to_np(element_positions)
    x, z = pos[:, 0], pos[:, 1]
    if radius == inf:
        d = x * sin(tilt_rad) / speed_of_sound
    else:
        xn = radius * sin(tilt_rad)
        zn = radius * cos(tilt_rad) - apex_offset
        h = zn + apex_offset
        num = np.abs(z + xn / h * x - xn**2 / h - zn)
        den = sqrt(1 + xn**2 / h**2)
        d = -(num / den) / speed_of_sound
    return _like(d - d.min(), ref)


def _angles_to_virtual_source(aperture_length: float, tilt_rad: float, width_rad: float):
    tilt = fmod(-tilt_rad + pi / 2, 2 * pi) - pi / 2
    sign = 1.0
    if abs(tilt) > pi / 2:
        tilt = pi - tilt
        sign = -1.0
    z0 = sign * aperture_length / (tan(tilt - width_rad / 2) - tan(tilt + width_rad / 2))
    x0 = sign * z0 * tan(width_rad / 2 - tilt) + aperture_length / 2
    return x0, z0


def diverging_wave(element_positions, tilt_rad: float, width_rad: float, *, aperture_length: float,
                   speed_of_sound: float):
    """Delays of a circular wave from a virtual source behind a linear array."""
    if width_rad <= 0 or width_rad >= pi:
        raise ValueError("Width angles must satisfy 0 < width_rad < π")
    x0, z0 = _angles_to_virtual_source(aperture_length, tilt_rad, width_rad)
    return focused(element_positions, np.array([x0, z0]), speed_of_sound=speed_of_sound)
