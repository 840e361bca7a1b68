"""Seeded random configurations through both device methods against the FP64 oracle.

Each case draws an array (1 to 192 elements, linear or convex, any baffle kind), a medium, a small
scatterer set (some out of the field), one to three transmits (plane waves or focused beams, some
elements with NaN delays), the sub-element count (default, 1 to 3, or 5 and 6: the grouped kernels)
and the directivity option. The plain FP32 method must match the oracle's RF to 1e-4 relative L2 and
the split-BF16 method to 1e-3, the same bars as the fixed cases.
"""

from __future__ import annotations

from math import inf

import numpy as np
import pytest

import fastsimus_amd as fs
from oracle import simus_oracle as O
from tests.conftest import f32, rel_l2

SEEDS = list(range(1024))


def random_case(seed: int):
    rng = np.random.default_rng(1000 + seed)
    convex = rng.random() < 0.3
    E = int(rng.choice([1, 5, 16, 33, 64, 100, 128, 160, 192]))
    fc = float(rng.uniform(2e6, 9e6))
    c = float(rng.uniform(1450.0, 1600.0))
    pitch = float(rng.uniform(0.8, 1.6) * c / fc / 2)
    width = pitch * float(rng.uniform(0.6, 0.95))
    kind = str(rng.choice(["soft", "rigid", "impedance"]))
    baffle = float(rng.uniform(0.5, 3.0)) if kind == "impedance" else kind
    radius = float(rng.uniform(0.02, 0.08)) if convex else inf
    params = fs.TransducerParams(freq_center=fc, pitch=pitch, width=width, n_elements=E,
                                 bandwidth=float(rng.uniform(0.5, 0.9)), radius=radius, baffle=baffle)
    medium = fs.MediumParams(speed_of_sound=c, attenuation=float(rng.uniform(0.0, 1.0)))
    S = int(rng.integers(1, 25))
    half = max(pitch * E, 0.004)
    x = f32(rng.uniform(-half, half, S))
    z = f32(rng.uniform(-0.002, 0.03, S))          # z < 0: out of the field
    rc = f32(rng.standard_normal(S))
    pos = fs.element_positions(E, pitch, radius)[0]
    T = int(rng.integers(1, 4))
    rows = []
    for _ in range(T):
        if rng.random() < 0.5 and not convex:
            rows.append(fs.plane_wave(pos, float(np.deg2rad(rng.uniform(-15, 15))), speed_of_sound=c))
        else:
            foc = np.array([rng.uniform(-half, half), rng.uniform(0.005, 0.03)])
            rows.append(fs.focused(pos, foc, speed_of_sound=c))
    D = np.stack(rows)
    if rng.random() < 0.25 and E > 2:
        D[:, rng.choice(E, max(1, E // 10), replace=False)] = np.nan
    kw = {}
    nsub = rng.choice([0, 1, 2, 3, 0, 1, 2, 3, 5, 6])
    if nsub:
        kw["element_splitting"] = int(nsub)
    if rng.random() < 0.2:
        kw["full_frequency_directivity"] = True
    return params, medium, x, z, rc, D, kw


def _oracle_rf(x, z, rc, D, params, medium, kw):
    plan = O.simus_plan(x, z, D, params, medium, element_splitting=kw.get("element_splitting"))
    fd = kw.get("full_frequency_directivity", False)
    return np.stack([O.simus_oracle(x, z, rc, d, params, medium, plan, full_frequency_directivity=fd)[0] for d in D])


def test_random_cases_are_valid_on_the_host():
    """The generator only draws inputs the public API accepts (no GPU needed)."""
    for seed in SEEDS:
        params, medium, x, z, rc, D, kw = random_case(seed)
        plan = fs.simus_precompute(np.stack([x, z], -1), rc, D, params, medium,
                                   element_splitting=kw.get("element_splitting"))
        assert plan.selected_freqs.size > 0


@pytest.mark.gpu
@pytest.mark.parametrize("seed", SEEDS)
def test_random_configuration_matches_oracle(seed, cuda):
    params, medium, x, z, rc, D, kw = random_case(seed)
    sc = np.stack([x, z], -1)
    want = _oracle_rf(x, z, rc, D, params, medium, kw)
    print(f"seed {seed}: E={params.n_elements} T={D.shape[0]} S={x.size} convex={params.radius != inf} "
          f"baffle={params.baffle} {kw}", flush=True)
    rf = fs.simus(sc, rc, D, params, medium, strategy="hip", **kw).rf.cpu().numpy().reshape(want.shape)
    e32 = rel_l2(rf, want)
    print(f"  plain FP32 rel L2 {e32:.2e}", flush=True)
    assert e32 < 1e-4
    rf = fs.simus(sc, rc, D, params, medium, strategy="hip_mfma", **kw).rf.cpu().numpy().reshape(want.shape)
    em = rel_l2(rf, want)
    print(f"  split-BF16 rel L2 {em:.2e}", flush=True)
    assert em < 1e-3


@pytest.mark.gpu
@pytest.mark.parametrize("seed", list(range(512)))
def test_random_pfield_matches_oracle(seed, cuda):
    """pfield on random arrays, media and grids (points behind the array included) against the FP64 oracle."""
    params, medium, _, _, _, D, kw = random_case(seed)
    rng = np.random.default_rng(2000 + seed)
    half = max(params.pitch * params.n_elements, 0.004)
    g = f32(np.stack([rng.uniform(-half, half, 60), rng.uniform(-0.002, 0.03, 60)], -1))
    d = D[0]
    print(f"seed {seed}: E={params.n_elements} convex={params.radius != inf} baffle={params.baffle} {kw}", flush=True)
    pf = fs.pfield(g, d, params, medium, **kw).cpu().numpy()
    pl = O.pfield_plan(g[:, 0], g[:, 1], d, params, medium, element_splitting=kw.get("element_splitting"))
    want = O.pfield_oracle(g[:, 0], g[:, 1], d, params, medium, pl,
                           full_frequency_directivity=kw.get("full_frequency_directivity", False))
    e = rel_l2(pf, want)
    print(f"  pfield rel L2 {e:.2e}", flush=True)
    assert e < 1e-4


MEDIUM_SEEDS = list(range(100, 108)) + list(range(200, 240))


@pytest.mark.gpu
@pytest.mark.parametrize("seed", MEDIUM_SEEDS)
def test_random_medium_size_methods_agree(seed, cuda):
    """Thousands of scatterers (several work units and chunks of the persistent launch) on random arrays
    (above 128 elements as pairs of element groups): the split-BF16 method against the plain FP32 one
    (the FP64 oracle is too slow here; both are pinned against it on the small cases)."""
    params, medium, _, _, _, D, kw = random_case(seed)
    rng = np.random.default_rng(3000 + seed)
    S = int(rng.integers(2000, 20000))
    half = max(params.pitch * params.n_elements, 0.004)
    x = f32(rng.uniform(-half, half, S))
    z = f32(rng.uniform(-0.002, 0.03, S))
    rc = f32(rng.standard_normal(S))
    sc = np.stack([x, z], -1)
    print(f"seed {seed}: E={params.n_elements} T={D.shape[0]} S={S} {kw}", flush=True)
    a = fs.simus(sc, rc, D, params, medium, strategy="hip", **kw).rf.cpu().numpy()
    b = fs.simus(sc, rc, D, params, medium, strategy="hip_mfma", **kw).rf.cpu().numpy()
    e = rel_l2(b, a)
    print(f"  split-BF16 vs plain FP32 rel L2 {e:.2e}", flush=True)
    assert e < 1e-4
