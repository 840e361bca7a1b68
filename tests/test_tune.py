"""Tuning cache and autotuner of the multistatic-matrix kernel (GPU).

Every tile shape the autotuner may pick computes the same spectrum up to the
order of FP32 summations: checked here against the default shape (RF
relative L2 < 1e-5) and the FP64 oracle (< 1e-3), for one phantom and for a
batch.
"""

from __future__ import annotations

import numpy as np
import pytest

import fastsimus_amd as fs
from fastsimus_amd import batch as B
from fastsimus_amd import kernels as K
from fastsimus_amd import tune
from oracle import simus_oracle as O
from tests.conftest import rel_l2, speckle

pytestmark = pytest.mark.gpu


def _setup(n=700, seed=80):
    p = fs.L11_5v()
    m = fs.MediumParams(attenuation=0.5)
    x, z, rc = speckle(seed, n)
    pos = fs.element_positions(p.n_elements, p.pitch, p.radius)[0]
    D = np.stack([fs.plane_wave(pos, np.deg2rad(a), speed_of_sound=1540.0) for a in (-6, 6)])
    sc = np.stack([x, z], -1)
    plan = fs.simus_precompute(sc, rc, D, p, m)
    return p, m, sc, rc, D, plan


def _prob(sc, rc, D, plan, p, m, cuda):
    return K.Problem(sc, rc, D, None, p, m, n_sub=plan.n_sub, seg_length=plan.seg_length, freqs=plan.selected_freqs,
                     pulse=plan.pulse_spectrum, probe=plan.probe_spectrum, correction=plan.correction_factor,
                     full_frequency_directivity=False, device=cuda)


def test_tuning_set_get_and_validation(cuda):
    p, m, sc, rc, D, plan = _setup()
    prob = _prob(sc, rc, D, plan, p, m, cuda)
    tune.set_tuning(prob)  # forget
    assert tune.get(prob) == {"msm_fbk": 0, "msm_units": 0, "msm_item_div": 0, "msm_table": 0}
    tune.set_tuning(prob, msm_fbk=2, msm_units=3)
    assert tune.get(prob) == {"msm_fbk": 2, "msm_units": 3, "msm_item_div": 0, "msm_table": 0}
    assert tune.get(prob, 2)["msm_fbk"] == 0  # other signature untouched
    with pytest.raises(ValueError):
        tune.set_tuning(prob, msm_fbk=3)
    tune.set_tuning(prob)
    assert tune.get(prob)["msm_units"] == 0


def test_every_shape_matches_default_and_oracle(cuda):
    p, m, sc, rc, D, plan = _setup()
    prob = _prob(sc, rc, D, plan, p, m, cuda)
    tune.set_tuning(prob)
    ref = fs.simus_compute(sc, rc, D, plan, p, m, strategy="hip_mfma").rf.cpu().numpy()
    plan_o = O.simus_plan(sc[:, 0].astype(np.float64), sc[:, 1].astype(np.float64), D, p, m)
    want = np.stack([O.simus_oracle(sc[:, 0].astype(np.float64), sc[:, 1].astype(np.float64),
                                    rc.astype(np.float64), d, p, m, plan_o)[0] for d in D])
    assert rel_l2(ref, want) < 1e-3
    for shape in ({"msm_fbk": 2}, {"msm_fbk": 4, "msm_units": 1}, {"msm_fbk": 2, "msm_units": 5},
                  {"msm_units": 40}):
        print(shape, flush=True)
        tune.set_tuning(prob, **shape)
        got = fs.simus_compute(sc, rc, D, plan, p, m, strategy="hip_mfma").rf.cpu().numpy()
        assert rel_l2(got, ref) < 1e-5, shape
    # geometry per frequency block instead of the table: other FP32 rounding of the amplitudes
    for shape in ({"msm_table": 2}, {"msm_table": 2, "msm_fbk": 2}, {"msm_table": 1, "msm_fbk": 2}):
        print(shape, flush=True)
        tune.set_tuning(prob, **shape)
        got = fs.simus_compute(sc, rc, D, plan, p, m, strategy="hip_mfma").rf.cpu().numpy()
        assert rel_l2(got, ref) < 3e-5, shape
        assert rel_l2(got, want) < 1e-3, shape
    tune.set_tuning(prob)


def test_autotune_single_and_batch(cuda):
    p, m, sc, rc, D, plan = _setup(n=900, seed=81)
    prob = _prob(sc, rc, D, plan, p, m, cuda)
    ref = fs.simus_compute(sc, rc, D, plan, p, m, strategy="hip_mfma").rf.cpu().numpy()
    res = tune.autotune(prob, reps=1)
    assert res["chosen"] in [{k: v for k, v in t.items() if k != "k_msm_ms"} for t in res["timings"]]
    assert tune.get(prob)["msm_fbk"] == res["chosen"].get("msm_fbk", 0)
    got = fs.simus_compute(sc, rc, D, plan, p, m, strategy="hip_mfma").rf.cpu().numpy()
    assert rel_l2(got, ref) < 1e-5
    tune.set_tuning(prob)

    ph = [(sc[:300], rc[:300]), (sc[300:310], rc[300:310]), (sc[310:], rc[310:])]
    pts, rcs, off = B._concat(ph)
    bprob = _prob(pts, rcs, D, plan, p, m, cuda)
    before = B.simus_batch_compute(ph, D, plan, p, m).rf.cpu().numpy()
    res = tune.autotune(bprob, off, reps=1)
    after = B.simus_batch_compute(ph, D, plan, p, m).rf.cpu().numpy()
    assert rel_l2(after, before) < 1e-5
    tune.set_tuning(bprob, 3)


@pytest.mark.parametrize("case", ["convex_c5", "impedance_nsub3", "l11_speckle_atten"])
def test_geometry_table_matches_per_block_geometry(case, cuda):
    """The frequency-independent geometry read from the per-call table (default) and evaluated
    per frequency block (msm_table = 2) agree to FP32 rounding of the amplitudes, and both meet
    the FP64 oracle: convex array, impedance baffle with three sub-elements, soft baffle."""
    from tests.test_gpu_simus import CASES, _oracle_rf

    params, medium, x, z, rc, D, kw = CASES[case]()
    D = np.atleast_2d(D)
    sc = np.stack([x, z], -1)
    plan = fs.simus_precompute(sc, rc, D, params, medium, element_splitting=kw.get("element_splitting"))
    prob = K.Problem(sc, rc, D, None, params, medium, n_sub=plan.n_sub, seg_length=plan.seg_length,
                     freqs=plan.selected_freqs, pulse=plan.pulse_spectrum, probe=plan.probe_spectrum,
                     correction=plan.correction_factor, full_frequency_directivity=False, device=cuda)
    print(case, flush=True)
    tune.set_tuning(prob)
    with_table = fs.simus_compute(sc, rc, D, plan, params, medium, strategy="hip_mfma").rf.cpu().numpy()
    tune.set_tuning(prob, msm_table=2)
    per_block = fs.simus_compute(sc, rc, D, plan, params, medium, strategy="hip_mfma").rf.cpu().numpy()
    tune.set_tuning(prob)
    want = _oracle_rf(x, z, rc, D, params, medium, **kw)
    assert rel_l2(with_table, per_block) < 3e-5
    assert rel_l2(with_table, want) < 1e-3
    assert rel_l2(per_block, want) < 1e-3
