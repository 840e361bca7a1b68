"""Batches of phantoms (``fastsimus_amd.batch``): one plan, one persistent launch, graph replay.

CPU: concatenation offsets and the shared plan. GPU (marked ``gpu``): every
phantom of a batch of uneven sizes (empty, below one 16-scatterer tile, ragged
tails, several tiles) against the FP64 oracle under the shared plan, per
phantom; the batched launch equal to single-phantom calls; the plain method's
batch bit-equal to its per-phantom calls; a captured graph replayed several
times with new inputs and eager work in between, each replay bit-equal to the
eager batch on the same inputs.
"""

from __future__ import annotations

import numpy as np
import pytest

import fastsimus_amd as fs
from fastsimus_amd import batch as B
from oracle import simus_oracle as O
from tests.conftest import rel_l2, speckle

MFMA_TOL = 1e-3
FP32_TOL = 1e-4
SIZES = (37, 0, 5, 160, 16, 91)


def _phantoms(seed=40, sizes=SIZES, xlim=(-0.02, 0.02), zlim=(0.005, 0.045)):
    out = []
    for i, n in enumerate(sizes):
        x, z, rc = speckle(seed + i, n, xlim=xlim, zlim=zlim)
        out.append((np.stack([x, z], -1), rc))
    return out


def _pw(params, angles):
    pos = fs.element_positions(params.n_elements, params.pitch, params.radius)[0]
    return np.stack([fs.plane_wave(pos, np.deg2rad(a), speed_of_sound=1540.0) for a in angles])


def test_concat_offsets_and_shared_plan():
    ph = _phantoms()
    pts, rc, off = B._concat(ph)
    assert off.tolist() == [0, 37, 37, 42, 202, 218, 309]
    assert pts.shape == (309, 2) and rc.shape == (309,)
    np.testing.assert_array_equal(pts[42:202], ph[3][0].astype(np.float32))
    p = fs.L11_5v()
    D = _pw(p, (-10, 0, 10))
    plan_b = B.simus_batch_precompute(ph, D, p, fs.MediumParams(attenuation=0.5))
    plan_c = fs.simus_precompute(pts, rc, D, p, fs.MediumParams(attenuation=0.5))
    assert plan_b.n_freq_full == plan_c.n_freq_full and plan_b.n_fft == plan_c.n_fft
    np.testing.assert_array_equal(plan_b.selected_freqs, plan_c.selected_freqs)
    with pytest.raises(ValueError):
        B._concat([(np.zeros((3, 2)), np.zeros(2))])
    with pytest.raises(ValueError):
        B._concat([])


def _oracle_batch(ph, D, params, medium):
    pts, _, _ = B._concat(ph)
    plan_o = O.simus_plan(pts[:, 0].astype(np.float64), pts[:, 1].astype(np.float64), D, params, medium)
    out = []
    for sc, rc in ph:
        out.append(np.stack([O.simus_oracle(sc[:, 0].astype(np.float64), sc[:, 1].astype(np.float64),
                                            rc.astype(np.float64), d, params, medium, plan_o)[0] for d in D]))
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("strategy,tol", [("hip_mfma", MFMA_TOL), ("hip", FP32_TOL)])
def test_batch_matches_oracle_per_phantom(strategy, tol, cuda):
    p = fs.L11_5v()
    m = fs.MediumParams(attenuation=0.5)
    ph = _phantoms()
    D = _pw(p, (-10, 0, 10))
    print(strategy, flush=True)
    res = B.simus_batch(ph, D, p, m, strategy=strategy)
    rf = res.rf.cpu().numpy()
    assert rf.shape[:2] == (len(ph), 3)
    want = _oracle_batch(ph, D, p, m)
    for q, w in enumerate(want):
        if SIZES[q] == 0:
            assert not np.any(rf[q])
        else:
            assert rel_l2(rf[q], w) < tol, q


@pytest.mark.gpu
def test_batch_equals_single_phantom_calls(cuda):
    p = fs.L11_5v()
    m = fs.MediumParams(attenuation=0.3)
    ph = _phantoms(seed=50)
    D = _pw(p, (-5, 5))
    plan = B.simus_batch_precompute(ph, D, p, m)
    for strategy in ("hip_mfma", "hip"):
        print(strategy, flush=True)
        rf = B.simus_batch_compute(ph, D, plan, p, m, strategy=strategy).rf.cpu().numpy()
        for q, (sc, rc) in enumerate(ph):
            one = fs.simus_compute(sc, rc, D, plan, p, m, strategy=strategy).rf.cpu().numpy()
            if strategy == "hip":
                np.testing.assert_array_equal(rf[q], one)
            elif sc.shape[0] == 0:
                assert not np.any(one) and not np.any(rf[q]), q
            else:
                assert np.any(one), q
                assert rel_l2(rf[q], one) < 1e-5, q


@pytest.mark.gpu
def test_batch_above_one_launch_is_split(cuda):
    """More phantoms than one launch holds (192): the batch is split over launches and every phantom,
    including the ones of the second launch and empty ones, equals its own call under the shared plan."""
    p = fs.P4_2v()
    m = fs.MediumParams(attenuation=0.3)
    rng = np.random.default_rng(77)
    sizes = rng.integers(0, 24, 200)
    sizes[[0, 150, 199]] = 0
    ph = []
    for n in sizes:
        sc = np.stack([rng.uniform(-0.02, 0.02, n), rng.uniform(0.01, 0.06, n)], -1).astype(np.float32)
        ph.append((sc, rng.standard_normal(n).astype(np.float32)))
    pos = fs.element_positions(p.n_elements, p.pitch, p.radius)[0]
    D = fs.focused(pos, np.array([0.0, 0.04]), speed_of_sound=1540.0)[None, :]
    plan = B.simus_batch_precompute(ph, D, p, m)
    rf = B.simus_batch_compute(ph, D, plan, p, m, strategy="hip_mfma").rf.cpu().numpy()
    assert rf.shape[0] == 200
    for q in (0, 1, 150, 191, 192, 193, 198, 199):
        sc, rc = ph[q]
        if sc.shape[0] == 0:
            assert not np.any(rf[q])
            continue
        one = fs.simus_compute(sc, rc, D, plan, p, m, strategy="hip_mfma").rf.cpu().numpy()
        assert np.any(one), q
        assert rel_l2(rf[q], one) < 1e-5, q


@pytest.mark.gpu
def test_phased_array_batch_focused_beams(cuda):
    """64-element phased array (one sub-element), focused beams over a sector, uneven phantoms."""
    p = fs.P4_2v()
    m = fs.MediumParams()
    ph = _phantoms(seed=60, sizes=(300, 7, 64, 129), xlim=(-0.04, 0.04), zlim=(0.01, 0.1))
    pos = fs.element_positions(p.n_elements, p.pitch, p.radius)[0]
    ang = np.deg2rad(np.linspace(-40, 40, 5))
    D = np.stack([fs.focused(pos, np.array([0.06 * np.sin(a), 0.06 * np.cos(a)]), speed_of_sound=1540.0) for a in ang])
    rf = B.simus_batch(ph, D, p, m, strategy="hip_mfma").rf.cpu().numpy()
    want = _oracle_batch(ph, D, p, m)
    for q, w in enumerate(want):
        assert rel_l2(rf[q], w) < MFMA_TOL, q


@pytest.mark.gpu
def test_graph_replay_matches_eager(cuda):
    import torch

    p = fs.L11_5v()
    m = fs.MediumParams(attenuation=0.5)
    ph = _phantoms(seed=70, sizes=(50, 3, 120))
    D = _pw(p, (-8, 8))
    plan = B.simus_batch_precompute(ph, D, p, m)
    g = B.SimusBatchGraph(ph, D, plan, p, m, strategy="hip_mfma")
    rng = np.random.default_rng(3)
    for rep in range(3):
        # new scatterers of the same sizes, inside the planned depth range
        new = [(np.stack([rng.uniform(-0.02, 0.02, len(rc)), rng.uniform(0.006, 0.04, len(rc))], -1).astype(np.float32),
                rng.standard_normal(len(rc)).astype(np.float32)) for _, rc in ph]
        pts, rc, _ = B._concat(new)
        print("replay", rep, flush=True)
        got = g.replay(torch.as_tensor(pts[:, 0]), torch.as_tensor(pts[:, 1]), torch.as_tensor(rc)).clone()
        torch.cuda.synchronize()
        eager = B.simus_batch_compute(new, D, plan, p, m, strategy="hip_mfma").rf
        _ = torch.randn(1000, device=cuda).sum().item()  # unrelated eager work between replays
        assert torch.equal(got, eager), rep
