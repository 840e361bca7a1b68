"""Arrays beyond the register-resident shapes (marked ``gpu``): more than 4 sub-elements per element, and more
than 128 elements through the batched (multistatic-matrix) method.

The reference takes any sub-element count (``simus.py:205-212``; its CUDA kernel compiles per
``(n_elem, n_sub, n_freq)``, ``kernels/cuda_simus.py:79-116``) and any element count (its own 192-element
``L12_3v`` preset, ``transducer_presets.py:58-72``). Here, sub-elements past 4 run in groups of 4 (plain,
pfield and vector-adjoint kernels) or one at a time (matrix-core kernel), and more than 128 elements run on
the matrix cores as pairs of 64-element groups. Every case is compared with the FP64 oracle under the same
plan: RF relative L2 <= 1e-4 for the plain FP32 method, <= 1e-3 for split-BF16.
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

torch = pytest.importorskip("torch")

TOL = {"hip": 1e-4, "hip_mfma": 1e-3}


def _pw(params, angles):
    pos = fs.element_positions(params.n_elements, params.pitch, params.radius)[0]
    return np.stack([fs.plane_wave(pos, np.deg2rad(a), speed_of_sound=1540.0) for a in angles])


def _check(params, medium, x, z, rc, D, strategies=("hip", "hip_mfma"), **kw):
    sc = np.stack([x, z], -1)
    plan = fs.simus_precompute(sc, rc, D, params, medium, element_splitting=kw.get("element_splitting"))
    plan_o = O.simus_plan(x, z, D, params, medium, element_splitting=kw.get("element_splitting"))
    assert plan.n_sub == plan_o["n_sub"]
    fd = kw.get("full_frequency_directivity", False)
    want = [O.simus_oracle(x, z, rc, d, params, medium, plan_o, full_frequency_directivity=fd)[0] for d in D]
    for strat in strategies:
        rf = fs.simus_compute(sc, rc, D, plan, params, medium, strategy=strat,
                              full_frequency_directivity=fd).rf.cpu().numpy()
        for t in range(D.shape[0]):
            e = rel_l2(rf[t], want[t])
            print(f"  {strat} tx {t}: rel-L2 vs FP64 {e:.2e}", flush=True)
            assert e <= TOL[strat], (strat, t, e)
    return plan


@pytest.mark.gpu
@pytest.mark.parametrize("n_sub,fulldir,baffle", [(5, False, "soft"), (6, False, "rigid"), (7, True, 1.5),
                                                   (9, False, "soft")])
def test_many_subelements_match_oracle(n_sub, fulldir, baffle, cuda):
    p = fs.L11_5v().model_copy(update={"baffle": baffle})
    m = fs.MediumParams(attenuation=0.4)
    x, z, rc = speckle(300 + n_sub, 30, zlim=(0.004, 0.04))
    D = _pw(p, (-6.0, 3.0, 11.0))
    print(f"n_sub {n_sub}, fulldir {fulldir}, baffle {baffle}", flush=True)
    plan = _check(p, m, x, z, rc, D, element_splitting=n_sub, full_frequency_directivity=fulldir)
    assert plan.n_sub == n_sub


@pytest.mark.gpu
def test_many_subelements_convex_match_oracle(cuda):
    p = fs.C5_2v()
    m = fs.MediumParams(attenuation=0.3)
    x, z, rc = speckle(311, 30, xlim=(-0.03, 0.03), zlim=(0.005, 0.07))
    pos = fs.element_positions(p.n_elements, p.pitch, p.radius)[0]
    D = np.stack([fs.focused(pos, np.array([0.0, 0.04]), speed_of_sound=1540.0)])
    print("convex, n_sub 5", flush=True)
    _check(p, m, x, z, rc, D, element_splitting=5)


@pytest.mark.gpu
def test_many_subelements_pfield_matches_oracle(cuda):
    p = fs.L11_5v()
    m = fs.MediumParams(attenuation=0.5)
    pos = fs.element_positions(p.n_elements, p.pitch, p.radius)[0]
    d = fs.focused(pos, np.array([0.002, 0.02]), speed_of_sound=1540.0)
    x, z, _ = speckle(312, 200, zlim=(0.003, 0.04))
    grid = np.stack([x, z], -1)
    plan = fs.pfield_precompute(grid, d, p, m, element_splitting=6)
    plan_o = O.pfield_plan(x, z, d, p, m, element_splitting=6)
    got = fs.pfield(grid, d, p, m, element_splitting=6).cpu().numpy().reshape(-1)
    want = O.pfield_oracle(x, z, d, p, m, plan_o)
    e = rel_l2(got, want)
    print(f"pfield n_sub 6: rel-L2 vs FP64 {e:.2e}", flush=True)
    assert plan.n_sub == 6 and e <= 1e-4


@pytest.mark.gpu
def test_many_subelements_adjoint_matches_fp64_gradient(cuda):
    """Vector adjoint kernels with 5 sub-elements (groups of 4) against the FP64 oracle gradient."""
    from fastsimus_amd.autograd import simus_rf
    from oracle.simus_grad import simus_rf_torch

    p = fs.TransducerParams(freq_center=5e6, pitch=3.0e-4, width=2.7e-4, n_elements=32, bandwidth=0.7)
    m = fs.MediumParams(attenuation=0.5)
    x, z, rc = speckle(313, 8, xlim=(-0.004, 0.004), zlim=(0.006, 0.02))
    rc_t = rc + 0.1 * np.random.default_rng(3).standard_normal(rc.size)
    D = _pw(p, (-4.0, 5.0))
    sc = np.stack([x, z], -1)
    plan = fs.simus_precompute(sc, rc, D, p, m, element_splitting=5)
    plan_o = O.simus_plan(x, z, D, p, m, element_splitting=5)
    with torch.no_grad():
        target = simus_rf_torch(torch.as_tensor(x), torch.as_tensor(z), torch.as_tensor(rc_t), 1560.0, D, p, 0.5,
                                plan_o)
    xt, zt, rt = (torch.tensor(a, requires_grad=True) for a in (x, z, rc))
    ct = torch.tensor(1540.0, dtype=torch.float64, requires_grad=True)
    (0.5 * ((simus_rf_torch(xt, zt, rt, ct, D, p, 0.5, plan_o) - target) ** 2).sum()).backward()
    xg, zg, rg = (torch.tensor(a, dtype=torch.float32, device=cuda, requires_grad=True) for a in (x, z, rc))
    cg = torch.tensor(1540.0, dtype=torch.float64, requires_grad=True)
    rf = simus_rf(xg, zg, rg, cg, D, p, plan, attenuation=0.5, method="fp32")
    (0.5 * ((rf.double() - target.to(cuda)) ** 2).sum()).backward()
    errs = [rel_l2(rg.grad.cpu().numpy(), rt.grad.numpy()), rel_l2(xg.grad.cpu().numpy(), xt.grad.numpy()),
            rel_l2(zg.grad.cpu().numpy(), zt.grad.numpy()), abs(float(cg.grad) / float(ct.grad) - 1)]
    print("n_sub 5 adjoint: rel err rc, x, z, c", [f"{e:.2e}" for e in errs], flush=True)
    assert max(errs) <= 1e-3


@pytest.mark.gpu
@pytest.mark.parametrize("n_elem,table", [(192, 0), (192, 2), (130, 0), (200, 0), (256, 0)])
def test_more_than_128_elements_match_oracle(n_elem, table, cuda):
    """More than 128 elements on the matrix cores (pairs of 64-element groups; 130 and 200 leave a partly
    filled last group), with the geometry table (0) and without it (2), against the FP64 oracle; the plain
    method too for the 192-element L12-3v preset."""
    base = fs.L12_3v()
    p = base if n_elem == 192 else base.model_copy(update={"n_elements": n_elem})
    m = fs.MediumParams(attenuation=0.5)
    half = 0.5 * p.pitch * n_elem
    x, z, rc = speckle(320 + n_elem, 40, xlim=(-half, half), zlim=(0.004, 0.035))
    D = _pw(p, (-9.0, 0.0, 12.0))
    sc = np.stack([x, z], -1)
    plan = fs.simus_precompute(sc, rc, D, p, m)
    prob = K.Problem(sc, rc, D, None, p, m, n_sub=plan.n_sub, seg_length=plan.seg_length,
                     freqs=plan.selected_freqs, pulse=plan.pulse_spectrum, probe=plan.probe_spectrum,
                     correction=plan.correction_factor, full_frequency_directivity=False, device=cuda)
    tune.set_tuning(prob, msm_table=table) if table else tune.set_tuning(prob)
    print(f"{n_elem} elements, table setting {table}, n_sub {plan.n_sub}", flush=True)
    try:
        _check(p, m, x, z, rc, D, strategies=("hip_mfma", "hip") if n_elem == 192 and table == 0 else ("hip_mfma",))
    finally:
        tune.set_tuning(prob)


@pytest.mark.gpu
def test_more_than_128_elements_batch_split_and_equal_to_single_calls(cuda):
    """256 elements (6 pairs of element groups per scatterer chunk): a batch of 40 phantoms is split into calls
    of at most 32 phantoms, and every phantom equals its own call under the shared plan."""
    p = fs.L12_3v().model_copy(update={"n_elements": 256})
    m = fs.MediumParams(attenuation=0.3)
    assert B.max_phantoms_per_call(256, "mfma") == 32
    rng = np.random.default_rng(330)
    sizes = rng.integers(1, 40, 40)
    sizes[[0, 33]] = 0
    ph = []
    for i, n in enumerate(sizes):
        x, z, rc = speckle(400 + i, int(n), xlim=(-0.02, 0.02), zlim=(0.005, 0.03))
        ph.append((np.stack([x, z], -1), rc))
    D = _pw(p, (-5.0, 8.0))
    plan = B.simus_batch_precompute(ph, D, p, m)
    rf = B.simus_batch_compute(ph, D, plan, p, m, strategy="hip_mfma").rf.cpu().numpy()
    for q in (0, 1, 31, 32, 33, 39):
        sc, rc = ph[q]
        if sc.shape[0] == 0:
            assert not np.any(rf[q])
            continue
        one = fs.simus_compute(sc, rc, D, plan, p, m, strategy="hip_mfma").rf.cpu().numpy()
        e = rel_l2(rf[q], one)
        print(f"  phantom {q} ({sc.shape[0]} scatterers): batch vs own call {e:.2e}", flush=True)
        assert np.any(one) and e <= 1e-5, (q, e)
