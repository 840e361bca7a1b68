"""The allocation-free entry point ``fs_simus_spectrum_static`` against ``fs_simus_spectrum`` (FP32 method).

Both are called through the C ABI on the same device inputs. Two calls in a row, the second with a
smaller input (fewer scatterers, fewer transmits, another probe), so that contents the first call
left in the shared static scratch cannot pass for results; each must agree with the workspace entry
point to 1e-5 relative L2 (they run the same kernels with different frequency-block and scatterer
splits, so they differ only in FP32 summation order) and with the FP64 oracle to the plain method's
1e-4. A problem larger than the scratch returns FS_ERR_SCRATCH before any device work.
"""

from __future__ import annotations

import numpy as np
import pytest

import fastsimus_amd as fs
from fastsimus_amd import _lib
from fastsimus_amd import kernels as K
from oracle import simus_oracle as O
from tests.conftest import rel_l2, speckle

torch = pytest.importorskip("torch")


def _problem(params, medium, x, z, rc, D, device, plan=None):
    if plan is None:
        plan = fs.simus_precompute(np.stack([x, z], -1), rc, D, params, medium)
    prob = K.Problem(np.stack([x, z], -1), rc, D, None, params, medium, n_sub=plan.n_sub,
                     seg_length=plan.seg_length, freqs=plan.selected_freqs, pulse=plan.pulse_spectrum,
                     probe=plan.probe_spectrum, correction=plan.correction_factor,
                     full_frequency_directivity=False, device=device)
    return prob, plan


def _static(prob):
    lib = _lib.load()
    out = torch.empty((prob.n_tx, prob.n_freq, prob.params.n_elements, 2), dtype=torch.float32, device=prob.device)
    st = lib.fs_simus_spectrum_static(prob.dref, out.data_ptr(), K._stream(prob.device))
    _lib.check(st, "fs_simus_spectrum_static")
    torch.cuda.synchronize()
    return torch.view_as_complex(out).cpu().numpy()


def _cases():
    p1, m1 = fs.L11_5v(), fs.MediumParams(attenuation=0.5)
    x1, z1, rc1 = speckle(900, 20_000, zlim=(0.005, 0.05))
    pos1 = fs.element_positions(p1.n_elements, p1.pitch, p1.radius)[0]
    D1 = np.stack([fs.plane_wave(pos1, np.deg2rad(a), speed_of_sound=1540.0) for a in (-12.0, 0.0, 7.0)])
    p2, m2 = fs.P4_2v(), fs.MediumParams(attenuation=0.3)
    x2, z2, rc2 = speckle(901, 3_000, xlim=(-0.03, 0.03), zlim=(0.01, 0.08))
    pos2 = fs.element_positions(p2.n_elements, p2.pitch, p2.radius)[0]
    D2 = np.stack([fs.focused(pos2, np.array([0.01 * s, 0.05]), speed_of_sound=1540.0) for s in (-1, 1)])
    return [("L11-5v 20000 x 3 tx", p1, m1, x1, z1, rc1, D1), ("P4-2v 3000 x 2 tx", p2, m2, x2, z2, rc2, D2)]


@pytest.mark.gpu
def test_static_entry_matches_workspace_entry_twice(cuda):
    for name, p, m, x, z, rc, D in _cases():
        print("static scratch case:", name, flush=True)
        prob, plan = _problem(p, m, x, z, rc, D, cuda)
        got = _static(prob)
        want = K.run_spectrum(prob, "fp32").cpu().numpy()
        e = rel_l2(got, want)
        print(f"  static vs workspace entry rel-L2 {e:.2e}", flush=True)
        assert np.all(np.isfinite(got))
        assert e <= 1e-5, (name, e)
        # FP64 oracle on a scatterer sample, spectrum of every transmit under the same plan
        sel = np.random.default_rng(5).choice(x.size, 300, replace=False)
        sub, _ = _problem(p, m, x[sel], z[sel], rc[sel], D, cuda, plan=plan)
        got_s = _static(sub)
        plan_o = O.simus_plan(x, z, D, p, m)
        assert plan_o["freqs"].size == plan.selected_freqs.size
        for t in range(D.shape[0]):
            o = O.spectrum_recurrence(x[sel], z[sel], rc[sel], D[t], p, m, plan_o) * plan_o["corr"]
            eo = rel_l2(got_s[t], o)
            print(f"  tx {t}: static vs FP64 oracle rel-L2 {eo:.2e}", flush=True)
            assert eo <= 1e-4, (name, t, eo)


@pytest.mark.gpu
def test_static_entry_rejects_problems_beyond_its_scratch(cuda):
    p, m = fs.L11_5v(), fs.MediumParams()
    x, z, rc = speckle(902, 300_000)
    prob, _ = _problem(p, m, x, z, rc, np.zeros(p.n_elements), cuda)
    lib = _lib.load()
    out = torch.empty((1, prob.n_freq, p.n_elements, 2), dtype=torch.float32, device=cuda)
    st = lib.fs_simus_spectrum_static(prob.dref, out.data_ptr(), K._stream(cuda))
    assert st == _lib.FS_ERR_SCRATCH
    with pytest.raises(RuntimeError):
        _lib.check(st, "fs_simus_spectrum_static")
