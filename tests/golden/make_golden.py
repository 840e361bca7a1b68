"""Writes the golden fixtures of this directory from the FP64 oracle.

Each ``.npz`` holds the seed, the float32-rounded inputs, the plan scalars, and
for every transmit the corrected selected spectrum and the gated RF on a fixed
subset of receive channels (``chan``) to keep the files small. They are
regenerated with ``python tests/golden/make_golden.py`` and must stay small
(< 1 MB each).

Cases:
  p4_focused     reference geometry of tests/test_simus.py:68-86 (P4-2v, focus (0, 3 cm), 6 on-axis scatterers)
  l11_plane5     reference geometry of tests/test_simus.py:87-97 (L11-5v, 5 deg plane wave, 6 diagonal scatterers)
  cfg0_sub64     64-scatterer seeded subsample of the 128-element 2 000-scatterer workload (attenuation 0.5)
  brute_tiny     8 elements x 5 scatterers, spectrum from the brute-force (no recurrence) form
  two_tx         2 transmits (plane waves -8 and 12 deg) sharing one plan
  convex_c5      C5-2v convex array, impedance baffle 1.75, 12 scatterers
"""

from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent))

import fastsimus_amd as fs  # noqa: E402
from oracle import simus_oracle as O  # noqa: E402

L11_LIKE = dict(freq_center=7.6e6, pitch=3.0e-4, width=2.7e-4, n_elements=128, bandwidth=0.77)


def _f32(a):
    return np.asarray(a, dtype=np.float32).astype(np.float64)


def _write(name, params, medium, x, z, rc, delays, seed, direct=False):
    delays = np.atleast_2d(delays)
    plan = O.simus_plan(x, z, delays, params, medium)
    rfs, sels = [], []
    E = params.n_elements
    chan = np.unique(np.array([0, 1, E // 3, E // 2 - 1, E // 2, (3 * E) // 4, E - 1]))
    for d in delays:
        rf, _, sel = O.simus_oracle(x, z, rc, d, params, medium, plan, direct=direct)
        rfs.append(rf)
        sels.append(sel)
    np.savez_compressed(
        HERE / f"{name}.npz", seed=seed, x=x, z=z, rc=rc, delays=delays,
        params=np.array(params.model_dump_json()), medium=np.array(medium.model_dump_json()),
        n_f=plan["freqs"].size, n_ff=plan["n_ff"], idx0=plan["idx0"], n_fft=plan["n_fft"], n_sub=plan["n_sub"],
        corr=plan["corr"], chan=chan, spect=np.stack(sels)[:, :, chan].astype(np.complex64),
        rf=np.stack(rfs)[:, :, chan].astype(np.float32),
    )
    print(name, (HERE / f"{name}.npz").stat().st_size, "bytes")


def main():
    p4 = fs.P4_2v()
    pos, _, _ = fs.element_positions(64, p4.pitch, p4.radius)
    d = fs.focused(pos, np.array([0.0, 0.03]), speed_of_sound=1540.0)
    _write("p4_focused", p4, fs.MediumParams(), _f32(np.zeros(6)), _f32(np.linspace(1e-2, 8e-2, 6)), _f32(np.ones(6)), d, 0)

    l11 = fs.L11_5v()
    pos, _, _ = fs.element_positions(128, l11.pitch, l11.radius)
    d = fs.plane_wave(pos, np.deg2rad(5.0), speed_of_sound=1540.0)
    _write("l11_plane5", l11, fs.MediumParams(), _f32(np.linspace(-1e-2, 1e-2, 6)), _f32(np.linspace(0.5e-2, 3e-2, 6)),
           _f32(np.ones(6)), d, 0)

    p = fs.TransducerParams(**L11_LIKE)
    rng = np.random.default_rng(20260)
    x, z, rc = rng.uniform(-0.02, 0.02, 2000), rng.uniform(0.005, 0.045, 2000), rng.standard_normal(2000)
    idx = np.sort(rng.choice(2000, 64, replace=False))
    _write("cfg0_sub64", p, fs.MediumParams(attenuation=0.5), _f32(x[idx]), _f32(z[idx]), _f32(rc[idx]), np.zeros(128),
           20260)

    pt = fs.TransducerParams(freq_center=5e6, pitch=3e-4, n_elements=8, width=2.5e-4, bandwidth=0.6)
    rng = np.random.default_rng(7)
    pos, _, _ = fs.element_positions(8, pt.pitch, pt.radius)
    _write("brute_tiny", pt, fs.MediumParams(attenuation=0.3), _f32(rng.uniform(-4e-3, 4e-3, 5)),
           _f32(rng.uniform(3e-3, 2e-2, 5)), _f32(rng.standard_normal(5)),
           fs.plane_wave(pos, np.deg2rad(7), speed_of_sound=1540.0), 7, direct=True)

    rng = np.random.default_rng(11)
    pos, _, _ = fs.element_positions(128, p.pitch, p.radius)
    D = np.stack([fs.plane_wave(pos, np.deg2rad(a), speed_of_sound=1540.0) for a in (-8, 12)])
    _write("two_tx", p, fs.MediumParams(attenuation=0.5), _f32(rng.uniform(-0.02, 0.02, 24)),
           _f32(rng.uniform(0.005, 0.04, 24)), _f32(rng.standard_normal(24)), D, 11)

    c5 = fs.C5_2v().model_copy(update={"baffle": 1.75})
    rng = np.random.default_rng(13)
    _write("convex_c5", c5, fs.MediumParams(attenuation=0.4), _f32(rng.uniform(-0.03, 0.03, 12)),
           _f32(rng.uniform(0.005, 0.06, 12)), _f32(rng.standard_normal(12)), np.zeros(128), 13)


if __name__ == "__main__":
    main()
