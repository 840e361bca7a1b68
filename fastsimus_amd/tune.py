"""Autotuner of the multistatic-matrix kernel's tile shapes.

``autotune(prob, offsets=None)`` times ``k_msm`` (HIP events on its stream,
``fs_set_kernel_timing``) for a set of candidate shapes of this problem
signature and stores the fastest in the library's tuning cache
(``fs_tuning_set``), which every later ``fs_simus_spectrum`` /
``fs_simus_spectrum_batch`` call with the same signature consults. Candidates:

* ``msm_fbk``: frequencies per block, 4 (two per wave, the default) or 2 (one
  per wave, half the accumulator registers); 2 is the only shape where the
  Green's functions need more registers (more than two sub-elements, or
  directivity at every frequency);
* ``msm_units`` (one phantom): scatterer chunks, i.e. the granularity of the
  persistent work queue and the number of partial matrices;
* ``msm_item_div`` (batches): the largest work item as a fraction of a queue
  slot's share.

``msm_table`` (not searched: the timing covers ``k_msm`` only) selects
whether the frequency-independent geometry comes from a per-call table
(0: when possible and at most 32 GiB, 1: when possible, 2: never).

The choice changes only the order of floating-point summations over
scatterers (results agree to FP32 rounding). The reference's CUDA tile
constants are fixed by hand (``kernels/cuda_simus.py:39-45``).
"""

from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from . import kernels as K


def get(prob: K.Problem, n_phantoms: int = 1) -> dict:
    """Tuning entry the library would use for this signature (zeros = defaults)."""
    t = _lib.Tuning()
    _lib.check(_lib.load().fs_tuning_get(prob.dref, n_phantoms, C.byref(t)), "fs_tuning_get")
    return {"msm_fbk": t.msm_fbk, "msm_units": t.msm_units, "msm_item_div": t.msm_item_div,
            "msm_table": t.msm_table}


def set_tuning(prob: K.Problem, n_phantoms: int = 1, **fields) -> None:
    """Store an entry (``msm_fbk``, ``msm_units``, ``msm_item_div``, ``msm_table``); no fields forgets the entry."""
    lib = _lib.load()
    if not fields:
        _lib.check(lib.fs_tuning_set(prob.dref, n_phantoms, None), "fs_tuning_set")
        return
    t = _lib.Tuning()
    for k, v in fields.items():
        setattr(t, k, int(v))
    _lib.check(lib.fs_tuning_set(prob.dref, n_phantoms, C.byref(t)), "fs_tuning_set")


def _candidates(prob: K.Problem, n_phantoms: int, n_scat_max: int):
    fbks = [2] if (prob.desc.n_sub > 2 or prob.desc.full_frequency_directivity) else [4, 2]
    if n_phantoms == 1:
        tiles = max(1, -(-n_scat_max // 16))
        units = [0] + [u for u in (1, 2, 4, 6, 8, 12, 16) if u <= max(1, tiles // 8)]
        return [{"msm_fbk": f, "msm_units": u} for f in fbks for u in units]
    return [{"msm_fbk": f, "msm_item_div": d} for f in fbks for d in (0, 2, 4, 16, 32)]


def autotune(prob: K.Problem, offsets=None, *, reps: int = 2, candidates=None) -> dict:
    """Time the candidate shapes for this problem and keep the fastest; returns the choice and the timings."""
    lib = _lib.load()
    P = 1 if offsets is None else len(offsets) - 1
    if offsets is not None:
        offsets = np.ascontiguousarray(offsets, dtype=np.int64)
        n_max = int(np.diff(offsets).max())
    else:
        n_max = int(prob.desc.n_scat)
    cands = candidates if candidates is not None else _candidates(prob, P, n_max)

    def run():
        if offsets is None:
            K.run_spectrum(prob, "mfma")
        else:
            K.run_spectrum_batch(prob, offsets, "mfma")

    timings = []
    for c in cands:
        set_tuning(prob, P, **c)
        run()  # warm-up (workspace size depends on the shape)
        torch.cuda.synchronize(prob.device)
        lib.fs_set_kernel_timing(1)
        _lib.kernel_time_ms()
        for _ in range(reps):
            run()
        ms, n = _lib.kernel_time_ms()
        lib.fs_set_kernel_timing(0)
        timings.append((ms / max(n, 1), c))
    best_ms, best = min(timings, key=lambda t: t[0])
    set_tuning(prob, P, **best)
    return {"chosen": best, "k_msm_ms": best_ms,
            "timings": [{"k_msm_ms": round(ms, 3), **c} for ms, c in timings]}


__all__ = ["autotune", "get", "set_tuning"]
