"""Element layout of linear and convex arrays (host, float64 NumPy).

Follows ``utils/geometry.py:18-88`` of the reference: linear arrays place
element ``i`` at ``x = (i - (N-1)/2) * pitch`` evaluated in **float32** (the
reference builds the index vector as float32 and NumPy keeps that dtype), so
the positions seen by every consumer here are the float32-rounded values.
Convex arrays place elements on an arc whose chord endpoints sit at z = 0.
"""

from __future__ import annotations

from math import asin, atan2, inf, sin, sqrt

import numpy as np


def element_positions(n_elements: int, pitch: float, radius: float = inf, xp=None):
    """Return ``(positions (E,2) float64, theta (E,) or None, apex_offset)``.

    ``xp`` is accepted for signature compatibility and ignored: geometry is a
    host-side table that is uploaded once per call.
    """
    if radius == inf:
        idx = np.arange(n_elements, dtype=np.float32)
        x = (idx - np.float32((n_elements - 1) / 2)) * np.float32(pitch)
        x = x.astype(np.float64)
        z = np.zeros(n_elements)
        return np.stack([x, z], axis=-1), None, 0.0
    half = asin(pitch / 2 / radius)
    total = half * (n_elements - 1)
    chord = 2 * radius * sin(total)
    apex = sqrt(radius**2 - (chord / 2) ** 2)
    t0 = atan2(-chord / 2, apex)
    t1 = atan2(chord / 2, apex)
    theta = np.linspace(t0, t1, n_elements)
    x = radius * np.sin(theta)
    z = radius * np.cos(theta) - apex
    return np.stack([x, z], axis=-1), theta, float(apex)


def subelement_offsets(element_width: float, n_sub: int, theta_e: np.ndarray) -> np.ndarray:
    """Sub-element centroid offsets relative to element centres, shape (E, n_sub, 2).

    Offsets along the element face are ``-W/2 + seg/2 + j*seg``; the face is
    rotated by the element angle: (o cos(theta), -o sin(theta))
    (``_pfield_math.py:41-67``).
    """
    seg = element_width / n_sub
    o = np.array([-element_width / 2.0 + seg / 2.0 + j * seg for j in range(n_sub)])
    dx = o[None, :] * np.cos(theta_e)[:, None]
    dz = o[None, :] * np.sin(-theta_e)[:, None]
    return np.stack([dx, dz], axis=-1)


def subelement_positions(params, n_sub: int):
    """Absolute sub-element positions (E, n_sub) x and z plus per-element angle.

    Returns ``(sub_x, sub_z, theta_e, apex_offset)``; theta_e is zeros for a
    linear array.
    """
    pos, theta, apex = element_positions(params.n_elements, params.pitch, params.radius)
    if theta is None:
        theta = np.zeros(params.n_elements)
    off = subelement_offsets(params.element_width, n_sub, theta)
    sub_x = pos[:, 0][:, None] + off[..., 0]
    sub_z = pos[:, 1][:, None] + off[..., 1]
    return sub_x, sub_z, theta, apex
