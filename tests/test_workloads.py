"""The GPU workloads of the benchmark configurations at their full size (marked ``gpu`` and ``slow``).

* Single transmit, 128-element 7.6 MHz linear array: 2 000 speckle scatterers with zero delays, both methods
  against the FP64 oracle on every scatterer; 50 000 speckle scatterers plus 10 bright points under a
  0 degree plane wave, both methods on the full set against each other and against the FP64 oracle on a
  1 010-scatterer sample (1 000 speckle + the 10 points) with the full set's plan.
* Plane-wave compounding: 200 000 speckle scatterers, 75 plane waves in [-16, 16] deg, 128-element 7.6 MHz
  linear array. Split-BF16 against plain FP32 on 3 transmits; both against the FP64 oracle on a
  1 000-scatterer sample of 2 transmits; superposition in rc and the +-angle mirror symmetry at full size.
* Uneven phantom batch: 64 phantoms of 10^U[4,6] scatterers (15.3 M in total), 64-element 2.7 MHz phased
  array, 128 focused beams, one persistent launch. The smallest, the largest and two middle phantoms equal
  their own calls; per-phantom FP64 error on 150-scatterer samples; the whole batch replayed from a captured
  graph is bit-equal to eager runs.
* Differentiable form: 256 x 256 jittered grid (65 536 scatterers), 15 plane waves. Gradients on a 64-point
  sample against the FP64 oracle gradient; at full size the directional central differences along random
  directions, Euler's identity for rc, and the matrix-core against the vector adjoint.

The workloads are the ones ``bench.py`` times (shared definitions). Where the FP64 oracle runs on a sample of
a workload too large for it, it uses the product's plan of the full workload (the plan's own parity with the
oracle's plan is pinned on the CPU, ``tests/test_host.py``).
"""

from __future__ import annotations

import numpy as np
import pytest

import bench
import fastsimus_amd as fs
from fastsimus_amd import batch as B
from fastsimus_amd import kernels as K
from oracle import simus_oracle as O
from tests.conftest import rel_l2

torch = pytest.importorskip("torch")

pytestmark = [pytest.mark.gpu, pytest.mark.slow]


def _oracle_plan(plan, params):
    """The oracle's plan dict holding the product plan's values."""
    return {"freqs": np.asarray(plan.selected_freqs), "df": float(plan.selected_freqs[1] - plan.selected_freqs[0]),
            "n_sub": plan.n_sub, "seg": plan.seg_length, "pulse": np.asarray(plan.pulse_spectrum),
            "probe": np.asarray(plan.probe_spectrum), "n_ff": plan.n_freq_full, "idx0": plan.freq_idx_start,
            "corr": plan.correction_factor, "n_fft": plan.n_fft}


# ----------------------------------------------------------------------------- single transmit
def test_single_2000_scatterers_against_fp64_in_full(cuda):
    params, medium, (x, z, rc, d), _ = bench.single_workloads()
    sc = np.stack([x, z], -1)
    plan = fs.simus_precompute(sc, rc, d, params, medium)
    assert (x.size, plan.n_sub) == (2000, 2)
    print("single 2 000: FP64 oracle on every scatterer", flush=True)
    want, _, _ = O.simus_oracle(x.astype(np.float64), z.astype(np.float64), rc.astype(np.float64), d, params, medium,
                                _oracle_plan(plan, params))
    for strat, tol in (("hip", 1e-4), ("hip_mfma", 1e-3)):
        got = fs.simus_compute(sc, rc, d, plan, params, medium, strategy=strat).rf.cpu().numpy()
        err = rel_l2(got, want)
        print(f"  {strat} vs FP64 rel-L2 {err:.2e}", flush=True)
        assert err <= tol, (strat, err)


def test_single_50010_scatterers_methods_and_fp64_sample(cuda):
    params, medium, _, (x, z, rc, d) = bench.single_workloads()
    sc = np.stack([x, z], -1)
    plan = fs.simus_precompute(sc, rc, d, params, medium)
    assert (x.size, plan.n_sub) == (50_010, 2)
    print("single 50 010: both methods on the full set", flush=True)
    rf_p = fs.simus_compute(sc, rc, d, plan, params, medium, strategy="hip").rf.cpu().numpy()
    rf_m = fs.simus_compute(sc, rc, d, plan, params, medium, strategy="hip_mfma").rf.cpu().numpy()
    e = rel_l2(rf_m, rf_p)
    print(f"  split-BF16 vs plain FP32 rel-L2 {e:.2e}", flush=True)
    assert e <= 1e-3
    assert np.abs(rf_p).max() > 0.0
    idx = np.concatenate([np.sort(np.random.default_rng(3).choice(50_000, 1000, replace=False)),
                          np.arange(50_000, 50_010)])
    want, _, _ = O.simus_oracle(x[idx].astype(np.float64), z[idx].astype(np.float64), rc[idx].astype(np.float64), d,
                                params, medium, _oracle_plan(plan, params))
    for strat, tol in (("hip", 1e-4), ("hip_mfma", 1e-3)):
        got = fs.simus_compute(sc[idx], rc[idx], d, plan, params, medium, strategy=strat).rf.cpu().numpy()
        err = rel_l2(got, want)
        print(f"  {strat} vs FP64 (1 010-scatterer sample) rel-L2 {err:.2e}", flush=True)
        assert err <= tol, (strat, err)


# ----------------------------------------------------------------------------- plane-wave compounding
@pytest.fixture(scope="module")
def planewave75():
    params, medium, x, z, rc, D = bench.workload()
    sc = np.stack([x, z], -1)
    plan = fs.simus_precompute(sc, rc, D, params, medium)
    assert (x.size, D.shape[0], plan.n_sub) == (200_000, 75, 2)
    return params, medium, sc, rc, D, plan


def test_planewave75_mfma_against_plain_and_fp64(planewave75, cuda):
    params, medium, sc, rc, D, plan = planewave75
    print("planewave75: split-BF16, 75 transmits", flush=True)
    rf_m = fs.simus_compute(sc, rc, D, plan, params, medium, strategy="hip_mfma").rf
    sub = [0, 37, 74]
    print("planewave75: plain FP32, transmits", sub, flush=True)
    rf_p = fs.simus_compute(sc, rc, D[sub], plan, params, medium, strategy="hip").rf
    e = rel_l2(rf_m[sub].cpu().numpy(), rf_p.cpu().numpy())
    print(f"  split-BF16 vs plain FP32 (full workload, 3 transmits) rel-L2 {e:.2e}", flush=True)
    assert e <= 1e-3
    del rf_m, rf_p
    # FP64 oracle on a 1 000-scatterer sample, 2 transmits, the workload's plan
    idx = np.sort(np.random.default_rng(8).choice(sc.shape[0], 1000, replace=False))
    plan_o = _oracle_plan(plan, params)
    for t in (5, 61):
        want, _, _ = O.simus_oracle(sc[idx, 0].astype(np.float64), sc[idx, 1].astype(np.float64),
                                    rc[idx].astype(np.float64), D[t], params, medium, plan_o)
        for strat, tol in (("hip", 1e-4), ("hip_mfma", 1e-3)):
            got = fs.simus_compute(sc[idx], rc[idx], D[t], plan, params, medium, strategy=strat).rf.cpu().numpy()
            err = rel_l2(got, want)
            print(f"  tx {t} {strat} vs FP64 (1 000-scatterer sample) rel-L2 {err:.2e}", flush=True)
            assert err <= tol, (t, strat, err)


def test_planewave75_superposition_in_rc(planewave75, cuda):
    """The spectrum is linear in rc: the spectra of two complementary halves of the reflectivities add up to the
    spectrum of the whole, at full size, on the split-BF16 path."""
    params, medium, sc, rc, D, plan = planewave75
    half = np.random.default_rng(9).random(rc.size) < 0.5
    ra = np.where(half, rc, 0).astype(np.float32)
    rb = np.where(half, 0, rc).astype(np.float32)

    def spect(r):
        prob = K.Problem(sc, r, D, None, params, medium, n_sub=plan.n_sub, seg_length=plan.seg_length,
                         freqs=plan.selected_freqs, pulse=plan.pulse_spectrum, probe=plan.probe_spectrum,
                         correction=plan.correction_factor, full_frequency_directivity=False, device=cuda)
        return K.run_spectrum(prob, "mfma")

    print("planewave75: superposition", flush=True)
    s_all = spect(rc)
    s_sum = spect(ra) + spect(rb)
    e = float(torch.linalg.norm(s_all - s_sum) / torch.linalg.norm(s_all))
    print(f"  |S(rc) - S(rc_a) - S(rc_b)| / |S(rc)| = {e:.2e}", flush=True)
    assert e <= 1e-4


def test_planewave75_mirror_symmetry(planewave75, cuda):
    """Mirroring the scene (x -> -x) mirrors the channels. The element positions are antisymmetric and the
    plane-wave delays of -theta are those of theta reversed over the elements, which for the symmetric angle set
    is the transmit order reversed: transmit t of the mirrored scatterers under the reversed transmit list is the
    mirror image of transmit t, so its RF is transmit t's RF with the channels reversed."""
    params, medium, sc, rc, D, plan = planewave75
    mirror = sc * np.array([-1.0, 1.0], dtype=np.float32)
    D_m = D[::-1].copy()
    np.testing.assert_allclose(D_m, D[:, ::-1], atol=1e-15)
    print("planewave75: mirror symmetry", flush=True)
    a = fs.simus_compute(sc, rc, D, plan, params, medium, strategy="hip_mfma").rf
    b = fs.simus_compute(mirror, rc, D_m, plan, params, medium, strategy="hip_mfma").rf
    e = float(torch.linalg.norm(b.flip(2) - a) / torch.linalg.norm(a))
    print(f"  mirrored RF vs RF rel-L2 {e:.2e}", flush=True)
    assert e <= 1e-4


# ----------------------------------------------------------------------------- uneven phantom batch
@pytest.fixture(scope="module")
def phantoms64(cuda):
    params, medium, phantoms, delays = bench.phantom_workload()
    dph = [(torch.as_tensor(p, device=cuda), torch.as_tensor(r, device=cuda)) for p, r in phantoms]
    plan = B.simus_batch_precompute(dph, delays, params, medium)
    sizes = np.array([p.shape[0] for p, _ in phantoms])
    assert len(phantoms) == 64 and sizes.min() >= 10_000 and sizes.max() <= 1_000_000 and plan.n_sub == 1
    return params, medium, phantoms, dph, delays, plan, sizes


def test_phantoms64_batch_equals_single_calls(phantoms64, cuda):
    """Batch and single call split a phantom into different scatterer chunks, so their FP32 sums of up to
    ~60 000 scatterer tiles per chunk differ in order: within 1e-4 (a 938 045-scatterer phantom differs by
    3.9e-5, a 10 376-scatterer one by 1.9e-6), against the 1e-3 bar of the batched form."""
    params, medium, phantoms, dph, delays, plan, sizes = phantoms64
    print(f"phantoms64: one launch, {sizes.sum()} scatterers, {delays.shape[0]} beams", flush=True)
    rf = B.simus_batch_compute(dph, delays, plan, params, medium, strategy="hip_mfma").rf
    order = np.argsort(sizes)
    for q in (order[0], order[31], order[32], order[-1]):
        one = fs.simus_compute(dph[q][0], dph[q][1], delays, plan, params, medium, strategy="hip_mfma").rf
        e = float(torch.linalg.norm(rf[q] - one) / torch.linalg.norm(one))
        print(f"  phantom {q} ({sizes[q]} scatterers) batch vs own call rel-L2 {e:.2e}", flush=True)
        assert torch.count_nonzero(one) > 0
        assert e <= 1e-4, (q, e)


def test_phantoms64_fp64_per_phantom(phantoms64, cuda):
    """Per-phantom FP64 error: the first 150 scatterers of 6 phantoms (the smallest and the largest among them),
    beams 0, 64 and 127, the batch's plan."""
    params, medium, phantoms, dph, delays, plan, sizes = phantoms64
    order = np.argsort(sizes)
    pick = [order[0], order[-1], 3, 17, 40, 63]
    sub = [(phantoms[q][0][:150], phantoms[q][1][:150]) for q in pick]
    beams = delays[[0, 64, 127]]
    rf = B.simus_batch_compute(sub, beams, plan, params, medium, strategy="hip_mfma").rf.cpu().numpy()
    plan_o = _oracle_plan(plan, params)
    for i, (p_, r_) in enumerate(sub):
        want = np.stack([O.simus_oracle(p_[:, 0].astype(np.float64), p_[:, 1].astype(np.float64),
                                        r_.astype(np.float64), d, params, medium, plan_o)[0] for d in beams])
        e = rel_l2(rf[i], want)
        print(f"  phantom {pick[i]} sample vs FP64 rel-L2 {e:.2e}", flush=True)
        assert e <= 1e-3, (pick[i], e)


def test_phantoms64_graph_replay_bit_equal(phantoms64, cuda):
    params, medium, phantoms, dph, delays, plan, sizes = phantoms64
    print("phantoms64: capture", flush=True)
    g = B.SimusBatchGraph(dph, delays, plan, params, medium, strategy="hip_mfma")
    pts, _, _ = B._concat(dph)
    rng = np.random.default_rng(12)
    for rep in range(2):
        rc_new = torch.as_tensor(rng.standard_normal(int(sizes.sum())).astype(np.float32), device=cuda)
        print("  replay", rep, flush=True)
        got = g.replay(pts[:, 0].contiguous(), pts[:, 1].contiguous(), rc_new).clone()
        torch.cuda.synchronize()
        new = [(dph[q][0], rc_new[int(g.offsets[q]):int(g.offsets[q + 1])]) for q in range(len(dph))]
        eager = B.simus_batch_compute(new, delays, plan, params, medium, strategy="hip_mfma").rf
        _ = torch.randn(1000, device=cuda).sum().item()  # unrelated eager work between replays
        assert torch.equal(got, eager), rep
        del got, eager


# ----------------------------------------------------------------------------- differentiable form
@pytest.fixture(scope="module")
def adjoint15():
    params, x, z, rc, rc_t, D, att = bench.adjoint_workload()
    plan = fs.simus_precompute(np.stack([x, z], -1), rc, D, params, fs.MediumParams(attenuation=att))
    return params, x, z, rc, rc_t, D, att, plan


def test_adjoint15_sample_matches_fp64_gradient(adjoint15, cuda):
    """A 64-point sample of the jittered grid with all 15 transmits and the workload's plan: the default path
    (split-BF16 forward, matrix-core adjoint) against the FP64 oracle gradient."""
    from fastsimus_amd.autograd import simus_rf
    from oracle.simus_grad import simus_rf_torch

    params, x, z, rc, rc_t, D, att, plan = adjoint15
    idx = np.sort(np.random.default_rng(13).choice(x.size, 64, replace=False))
    xs, zs, rs, rts = (a[idx].astype(np.float64) for a in (x, z, rc, rc_t))
    plan_o = _oracle_plan(plan, params)
    with torch.no_grad():
        target = simus_rf_torch(torch.as_tensor(xs), torch.as_tensor(zs), torch.as_tensor(rts), 1560.0, D, params,
                                att, plan_o)
    xt, zt, rt = (torch.tensor(a, requires_grad=True) for a in (xs, zs, rs))
    ct = torch.tensor(1540.0, dtype=torch.float64, requires_grad=True)
    loss_o = 0.5 * ((simus_rf_torch(xt, zt, rt, ct, D, params, att, plan_o) - target) ** 2).sum()
    loss_o.backward()
    for adjoint in ("matrix", "vector"):
        xg, zg, rg = (torch.tensor(a, dtype=torch.float32, device=cuda, requires_grad=True) for a in (xs, zs, rs))
        cg = torch.tensor(1540.0, dtype=torch.float64, requires_grad=True)
        print("adjoint15 sample:", adjoint, flush=True)
        rf = simus_rf(xg, zg, rg, cg, D, params, plan, attenuation=att, adjoint=adjoint)
        loss = 0.5 * ((rf.double() - target.to(cuda)) ** 2).sum()
        loss.backward()
        errs = {"rc": rel_l2(rg.grad.cpu().numpy(), rt.grad.numpy()), "x": rel_l2(xg.grad.cpu().numpy(), xt.grad.numpy()),
                "z": rel_l2(zg.grad.cpu().numpy(), zt.grad.numpy()),
                "c": abs(float(cg.grad) - float(ct.grad)) / abs(float(ct.grad))}
        print("  rel err vs FP64 gradient", {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
        assert float(loss.detach()) == pytest.approx(float(loss_o.detach()), rel=1e-3)
        assert max(errs.values()) <= 1e-3, errs


def _full_loss(adjoint15, cuda, rc_v, x_v, z_v, c, target, grad=False, adjoint="auto"):
    from fastsimus_amd.autograd import simus_rf

    params, _, _, _, _, D, att, plan = adjoint15
    xg, zg, rg = (torch.tensor(a, dtype=torch.float32, device=cuda, requires_grad=grad) for a in (x_v, z_v, rc_v))
    cg = torch.tensor(float(c), dtype=torch.float64, requires_grad=grad)
    rf = simus_rf(xg, zg, rg, cg, D, params, plan, attenuation=att, adjoint=adjoint)
    loss = 0.5 * ((rf.double() - target) ** 2).sum()
    if not grad:
        return float(loss.detach()), rf.detach()
    loss.backward()
    return float(loss.detach()), rf.detach(), rg.grad.double(), xg.grad.double(), zg.grad.double(), float(cg.grad)


def test_adjoint15_full_size_gradient_checks(adjoint15, cuda):
    from fastsimus_amd.autograd import simus_rf

    params, x, z, rc, rc_t, D, att, plan = adjoint15
    with torch.no_grad():
        target = simus_rf(torch.tensor(x, device=cuda), torch.tensor(z, device=cuda), torch.tensor(rc_t, device=cuda),
                          1560.0, D, params, plan, attenuation=att).double()
    print("adjoint15 full size: forward + matrix-core adjoint", flush=True)
    L, rf, g_rc, g_x, g_z, g_c = _full_loss(adjoint15, cuda, rc, x, z, 1540.0, target, grad=True)
    # Euler: the gated RF is homogeneous of degree 1 in rc, so sum_s rc_s dL/drc_s = <RF - RF_t, RF>
    euler = float((torch.as_tensor(rc, device=cuda).double() * g_rc).sum())
    want = float(((rf.double() - target) * rf.double()).sum())
    print(f"  Euler: sum rc dL/drc {euler:.6e} vs <RF - RF_t, RF> {want:.6e}", flush=True)
    assert euler == pytest.approx(want, rel=1e-4)
    # matrix-core against vector adjoint on the same forward
    print("adjoint15 full size: vector adjoint", flush=True)
    _, _, v_rc, v_x, v_z, v_c = _full_loss(adjoint15, cuda, rc, x, z, 1540.0, target, grad=True, adjoint="vector")
    for name, a, b in (("rc", g_rc, v_rc), ("x", g_x, v_x), ("z", g_z, v_z)):
        e = float(torch.linalg.norm(a - b) / torch.linalg.norm(b))
        print(f"  matrix vs vector adjoint d{name}: rel-L2 {e:.2e}", flush=True)
        assert e <= 1e-4, (name, e)
    print(f"  dL/dc matrix {g_c:.6e} vector {v_c:.6e}", flush=True)
    assert g_c == pytest.approx(v_c, rel=1e-3)
    # directional central differences along random directions (every scatterer moves); rc: the loss is
    # quadratic in rc (up to the gate), so a large step is exact; positions 1e-6 m and c 2e-5 relative keep the
    # phase changes small (<= 0.1 rad) and the loss change far above the FP32 noise of the forward
    rng = np.random.default_rng(14)
    checks = []
    v = rng.standard_normal(rc.size).astype(np.float32)
    h = 0.5
    lp, _ = _full_loss(adjoint15, cuda, rc + h * v, x, z, 1540.0, target)
    lm, _ = _full_loss(adjoint15, cuda, rc - h * v, x, z, 1540.0, target)
    checks.append(("rc", (lp - lm) / (2 * h), float((g_rc * torch.as_tensor(v, device=cuda).double()).sum())))
    for name, arr, g in (("x", x, g_x), ("z", z, g_z)):
        v = rng.standard_normal(arr.size)
        h = 1e-6
        ap = (arr.astype(np.float64) + h * v).astype(np.float32)
        am = (arr.astype(np.float64) - h * v).astype(np.float32)
        step = torch.as_tensor((ap.astype(np.float64) - am.astype(np.float64)) / 2, device=cuda)  # realised step
        args_p = (ap, z) if name == "x" else (x, ap)
        args_m = (am, z) if name == "x" else (x, am)
        lp, _ = _full_loss(adjoint15, cuda, rc, *args_p, 1540.0, target)
        lm, _ = _full_loss(adjoint15, cuda, rc, *args_m, 1540.0, target)
        checks.append((name, (lp - lm) / 2, float((g * step).sum())))
    h = 2e-5 * 1540.0
    lp, _ = _full_loss(adjoint15, cuda, rc, x, z, 1540.0 + h, target)
    lm, _ = _full_loss(adjoint15, cuda, rc, x, z, 1540.0 - h, target)
    checks.append(("c", (lp - lm) / (2 * h), g_c))
    for name, fd, ad in checks:
        print(f"  directional derivative along a random direction, {name}: central FD {fd:.6e}, adjoint {ad:.6e}",
              flush=True)
    for name, fd, ad in checks:
        assert fd == pytest.approx(ad, rel=5e-3), name
