"""Verasonics probe presets (values of the reference ``transducer_presets.py:12-102``,
which in turn quotes MUST's ``getparam``)."""

from __future__ import annotations

from math import inf

from .params import TransducerParams


def P4_2v() -> TransducerParams:
    """64-element phased array, 2.72 MHz, pitch 0.3 mm, kerf 50 um, 74 % bandwidth."""
    return TransducerParams(freq_center=2.72e6, pitch=300e-6, n_elements=64, kerf=50e-6,
                            bandwidth=0.74, radius=inf, height=14e-3, elev_focus=60e-3)


def L11_5v() -> TransducerParams:
    """128-element linear array, 7.6 MHz, pitch 0.3 mm, kerf 30 um, 77 % bandwidth."""
    return TransducerParams(freq_center=7.6e6, pitch=300e-6, n_elements=128, kerf=30e-6,
                            bandwidth=0.77, radius=inf, height=5e-3, elev_focus=18e-3)


def L12_3v() -> TransducerParams:
    """192-element linear array, 7.54 MHz, pitch 0.2 mm, kerf 30 um, 93 % bandwidth."""
    return TransducerParams(freq_center=7.54e6, pitch=200e-6, n_elements=192, kerf=30e-6,
                            bandwidth=0.93, radius=inf, height=5e-3, elev_focus=20e-3)


def C5_2v() -> TransducerParams:
    """128-element convex array, 3.57 MHz, pitch 0.508 mm, radius 49.57 mm, 79 % bandwidth."""
    return TransducerParams(freq_center=3.57e6, pitch=508e-6, n_elements=128, kerf=48e-6,
                            bandwidth=0.79, radius=49.57e-3, height=13.5e-3, elev_focus=60e-3)
