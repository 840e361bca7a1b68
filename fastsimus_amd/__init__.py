"""MI355X-native SIMUS / PFIELD hot path (drop-in for charlesbmi/FastSIMUS's ``simus`` / ``pfield``).

Host semantics (parameter models, presets, geometry, spectra, delay laws,
frequency plans) are Python; the per-(frequency, element, scatterer) work, the
transmit / receive contractions and the spectrum-to-RF transform run as
hand-written gfx950 HIP kernels in ``libfastsimus.so`` (C ABI in
``include/fastsimus.h``).
"""

from .geometry import element_positions
from .params import BaffleType, MediumParams, TransducerParams
from .pfield import PfieldPlan, PfieldStrategy, pfield, pfield_compute, pfield_precompute
from .presets import C5_2v, L11_5v, L12_3v, P4_2v
from .simus import SimusPlan, SimusResult, SimusStrategy, rf_stack, simus, simus_compute, simus_precompute
from .tx_delay import diverging_wave, focused, plane_wave

__all__ = [
    "BaffleType", "C5_2v", "L11_5v", "L12_3v", "MediumParams", "P4_2v", "PfieldPlan", "PfieldStrategy",
    "SimusPlan", "SimusResult", "SimusStrategy", "TransducerParams", "diverging_wave", "element_positions",
    "focused", "pfield", "pfield_compute", "pfield_precompute", "plane_wave", "rf_stack", "simus", "simus_compute",
    "simus_precompute",
]
