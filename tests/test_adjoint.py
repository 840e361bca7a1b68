"""Adjoint (d loss / d rc, x, z, c) of the HIP SIMUS path.

CPU: the FP64 torch restatement used as gradient oracle reproduces the
NumPy oracle's RF and its autograd gradient matches central finite
differences of itself (h = 1e-6 relative).
GPU (marked ``gpu``): the hand-written adjoint matches the FP64 oracle
gradient and central finite differences of the FP32 GPU forward.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

import fastsimus_amd as fs
from oracle import simus_oracle as O
from oracle.simus_grad import simus_rf_torch
from tests.conftest import rel_l2, speckle

PARAMS = fs.TransducerParams(freq_center=7.6e6, pitch=3.0e-4, width=2.7e-4, n_elements=128, bandwidth=0.77)
ATT = 0.5


def _scene(n=10, seed=21):
    x, z, rc = speckle(seed, n, zlim=(0.008, 0.03))
    pos, _, _ = fs.element_positions(128, PARAMS.pitch, PARAMS.radius)
    D = np.stack([fs.plane_wave(pos, np.deg2rad(a), speed_of_sound=1540.0) for a in (-5.0, 4.0)])
    rng = np.random.default_rng(seed + 1)
    rc_t = rc + 0.1 * rng.standard_normal(n)
    return x, z, rc, rc_t, D


def _target(x, z, rc_t, D, plan):
    with torch.no_grad():
        return simus_rf_torch(torch.as_tensor(x), torch.as_tensor(z), torch.as_tensor(rc_t), 1560.0, D, PARAMS, ATT,
                              plan)


def _oracle_grads(x, z, rc, D, plan, target):
    xt = torch.tensor(x, requires_grad=True)
    zt = torch.tensor(z, requires_grad=True)
    rt = torch.tensor(rc, requires_grad=True)
    ct = torch.tensor(1540.0, dtype=torch.float64, requires_grad=True)
    rf = simus_rf_torch(xt, zt, rt, ct, D, PARAMS, ATT, plan)
    loss = 0.5 * ((rf - target) ** 2).sum()
    loss.backward()
    return float(loss.detach()), rt.grad.numpy(), xt.grad.numpy(), zt.grad.numpy(), float(ct.grad)


def test_torch_oracle_forward_matches_numpy_oracle():
    x, z, rc, _, D = _scene(4)
    m = fs.MediumParams(attenuation=ATT)
    plan = O.simus_plan(x, z, D, PARAMS, m)
    rf_t = simus_rf_torch(torch.as_tensor(x), torch.as_tensor(z), torch.as_tensor(rc), 1540.0, D, PARAMS, ATT, plan)
    for t in range(2):
        rf_n, _, _ = O.simus_oracle(x, z, rc, D[t], PARAMS, m, plan)
        assert rel_l2(rf_t[t].numpy(), rf_n) < 1e-9


def test_torch_oracle_gradient_matches_finite_differences():
    x, z, rc, rc_t, D = _scene(3)
    plan = O.simus_plan(x, z, D, PARAMS, fs.MediumParams(attenuation=ATT))
    target = _target(x, z, rc_t, D, plan)
    L, g_rc, g_x, g_z, g_c = _oracle_grads(x, z, rc, D, plan, target)

    def loss(xv, zv, rv, cv):
        with torch.no_grad():
            rf = simus_rf_torch(torch.as_tensor(xv), torch.as_tensor(zv), torch.as_tensor(rv), cv, D, PARAMS, ATT, plan)
            return float(0.5 * ((rf - target) ** 2).sum())

    for i in range(3):
        for arr, g in ((x, g_x), (z, g_z), (rc, g_rc)):
            h = 1e-6 * max(abs(arr[i]), 1e-3)
            a, b = arr.copy(), arr.copy()
            a[i] += h
            b[i] -= h
            args_p = [x, z, rc]
            args_m = [x, z, rc]
            k = [id(x), id(z), id(rc)].index(id(arr))
            args_p = [a if j == k else v for j, v in enumerate(args_p)]
            args_m = [b if j == k else v for j, v in enumerate(args_m)]
            fd = (loss(*args_p, 1540.0) - loss(*args_m, 1540.0)) / (2 * h)
            assert fd == pytest.approx(g[i], rel=2e-3, abs=1e-6 * abs(L) / max(abs(arr[i]), 1e-3))
    h = 1e-6 * 1540.0
    fd = (loss(x, z, rc, 1540.0 + h) - loss(x, z, rc, 1540.0 - h)) / (2 * h)
    assert fd == pytest.approx(g_c, rel=2e-3)


@pytest.mark.gpu
@pytest.mark.parametrize("method", ["fp32", "mfma"])
def test_hip_adjoint_matches_fp64_gradient(method, cuda):
    from fastsimus_amd.autograd import simus_rf

    x, z, rc, rc_t, D = _scene(10)
    sc = np.stack([x, z], -1)
    plan = fs.simus_precompute(sc, rc, D, PARAMS, fs.MediumParams(attenuation=ATT))
    plan_o = O.simus_plan(x, z, D, PARAMS, fs.MediumParams(attenuation=ATT))
    target = _target(x, z, rc_t, D, plan_o)
    L, g_rc, g_x, g_z, g_c = _oracle_grads(x, z, rc, D, plan_o, target)

    xt = torch.tensor(x, dtype=torch.float32, device=cuda, requires_grad=True)
    zt = torch.tensor(z, dtype=torch.float32, device=cuda, requires_grad=True)
    rt = torch.tensor(rc, dtype=torch.float32, device=cuda, requires_grad=True)
    ct = torch.tensor(1540.0, dtype=torch.float64, requires_grad=True)
    rf = simus_rf(xt, zt, rt, ct, D, PARAMS, plan, attenuation=ATT, method=method)
    tgt = target.to(torch.float32).to(cuda)
    loss = 0.5 * ((rf - tgt) ** 2).sum()
    loss.backward()
    print("loss", float(loss), L, flush=True)
    assert float(loss) == pytest.approx(L, rel=1e-3)
    assert rel_l2(rt.grad.cpu().numpy(), g_rc) < 1e-3
    assert rel_l2(xt.grad.cpu().numpy(), g_x) < 1e-3
    assert rel_l2(zt.grad.cpu().numpy(), g_z) < 1e-3
    assert float(ct.grad) == pytest.approx(g_c, rel=1e-3)


@pytest.mark.gpu
def test_hip_adjoint_matches_finite_differences(cuda):
    """Directional check on a linear functional L = <w, rf> of the FP32 GPU forward.

    The steps (1e-4 relative for positions and c) keep the loss change far above
    the FP32 rounding noise of the forward. Samples below 1e-3 of the peak get
    zero weight: near the -100 dB gate threshold the RF is far from linear on
    that step (the gate's transition is 1e-6 of the peak wide), which finite
    differences cannot resolve in FP32; the gate's own adjoint is covered by
    the FP64 gradient test above at h = 1e-6.
    """
    from fastsimus_amd.autograd import simus_rf

    x, z, rc, _, D = _scene(6, seed=33)
    sc = np.stack([x, z], -1)
    plan = fs.simus_precompute(sc, rc, D, PARAMS, fs.MediumParams(attenuation=ATT))
    nkeep = (plan.n_fft + 1) // 2
    w = torch.as_tensor(np.random.default_rng(5).standard_normal((2, nkeep, 128)), dtype=torch.float64, device=cuda)
    with torch.no_grad():
        rf0 = simus_rf(torch.tensor(x, dtype=torch.float32, device=cuda), torch.tensor(z, dtype=torch.float32, device=cuda),
                       torch.tensor(rc, dtype=torch.float32, device=cuda), 1540.0, D, PARAMS, plan, attenuation=ATT)
        w = w * (rf0.abs() > 1e-3 * rf0.abs().amax(dim=(1, 2), keepdim=True)).double()

    def run(xv, zv, rv, c, grad=False):
        xt = torch.tensor(xv, dtype=torch.float32, device=cuda, requires_grad=grad)
        zt = torch.tensor(zv, dtype=torch.float32, device=cuda, requires_grad=grad)
        rt = torch.tensor(rv, dtype=torch.float32, device=cuda, requires_grad=grad)
        ct = torch.tensor(c, dtype=torch.float64, requires_grad=grad)
        rf = simus_rf(xt, zt, rt, ct, D, PARAMS, plan, attenuation=ATT)
        loss = (rf.double() * w).sum()
        if grad:
            loss.backward()
            return float(loss), xt.grad.cpu().numpy(), zt.grad.cpu().numpy(), rt.grad.cpu().numpy(), float(ct.grad)
        return float(loss.detach())

    _, gx, gz, grc, gc = run(x, z, rc, 1540.0, grad=True)
    vals = {"x": x, "z": z, "rc": rc}
    for i in (0, 3, 5):
        for which, g, hrel in (("x", gx, 1e-4), ("z", gz, 1e-4), ("rc", grc, 1e-2)):
            base = vals[which].astype(np.float32)
            h = hrel * (1e-2 if which != "rc" else max(abs(float(base[i])), 0.1))
            a, b = base.copy(), base.copy()
            a[i] = np.float32(base[i] + h)
            b[i] = np.float32(base[i] - h)
            step = float(a[i]) - float(b[i])
            args_a = [a if k == which else vals[k] for k in ("x", "z", "rc")]
            args_b = [b if k == which else vals[k] for k in ("x", "z", "rc")]
            fd = (run(*args_a, 1540.0) - run(*args_b, 1540.0)) / step
            print(which, i, fd, g[i], flush=True)
            assert fd == pytest.approx(g[i], rel=2e-2, abs=2e-3 * np.max(np.abs(g)))
    h = 1e-4 * 1540.0
    fd = (run(x, z, rc, 1540.0 + h) - run(x, z, rc, 1540.0 - h)) / (2 * h)
    print("c", fd, gc, flush=True)
    assert fd == pytest.approx(gc, rel=2e-2)
