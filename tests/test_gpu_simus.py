"""GPU parity of the HIP SIMUS / PFIELD kernels against the FP64 oracle.

Tolerances (RF relative L2 vs FP64): plain FP32 method 1e-4, split-BF16 MFMA
method 1e-3. Cases follow the reference's tests: its two PyMUST geometries
(tests/test_simus.py:68-97), edge cases (single scatterer, zero rc, attenuation,
NaN delays, out-of-field scatterers), convex arrays, baffles, sub-element
counts, full-frequency directivity, several transmits. At full workload size
the checks are size-independent properties (linearity, mirror symmetry,
MFMA vs FP32 agreement, run-to-run reproducibility).
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

import fastsimus_amd as fs
from oracle import simus_oracle as O
from tests.conftest import f32, rel_l2, speckle

pytestmark = pytest.mark.gpu
FP32_TOL = 1e-4
MFMA_TOL = 1e-3


def _oracle_rf(x, z, rc, D, params, medium, **kw):
    D = np.atleast_2d(D)
    plan = O.simus_plan(x, z, D, params, medium, element_splitting=kw.get("element_splitting"))
    fd = kw.get("full_frequency_directivity", False)
    return np.stack([O.simus_oracle(x, z, rc, d, params, medium, plan, full_frequency_directivity=fd)[0] for d in D])


def _run(strategy, x, z, rc, D, params, medium, **kw):
    res = fs.simus(np.stack([x, z], -1), rc, D, params, medium, strategy=strategy, **kw)
    rf = res.rf.cpu().numpy()
    return rf.reshape(-1, *rf.shape[-2:]), res


def _pos(p):
    return fs.element_positions(p.n_elements, p.pitch, p.radius)[0]


CASES = {
    "p4_focused_ref": lambda: (fs.P4_2v(), fs.MediumParams(), f32(np.zeros(6)), f32(np.linspace(1e-2, 8e-2, 6)),
                               f32(np.ones(6)), fs.focused(_pos(fs.P4_2v()), np.array([0.0, 0.03]), speed_of_sound=1540.0), {}),
    "l11_plane5_ref": lambda: (fs.L11_5v(), fs.MediumParams(), f32(np.linspace(-1e-2, 1e-2, 6)),
                               f32(np.linspace(0.5e-2, 3e-2, 6)), f32(np.ones(6)),
                               fs.plane_wave(_pos(fs.L11_5v()), np.deg2rad(5), speed_of_sound=1540.0), {}),
    "l11_speckle_atten": lambda: (fs.L11_5v(), fs.MediumParams(attenuation=0.5), *speckle(1, 80), np.zeros(128), {}),
    "rigid_nsub1": lambda: (fs.L11_5v().model_copy(update={"baffle": "rigid"}), fs.MediumParams(), *speckle(2, 30),
                            np.zeros(128), {"element_splitting": 1}),
    "impedance_nsub3": lambda: (fs.L11_5v().model_copy(update={"baffle": 1.75}), fs.MediumParams(attenuation=0.3),
                                *speckle(3, 30), np.zeros(128), {"element_splitting": 3}),
    "convex_c5": lambda: (fs.C5_2v(), fs.MediumParams(attenuation=0.4), *speckle(4, 40, xlim=(-0.03, 0.03), zlim=(0.005, 0.06)),
                          np.zeros(128), {}),
    "full_freq_directivity": lambda: (fs.L11_5v(), fs.MediumParams(), *speckle(5, 20), np.zeros(128),
                                      {"full_frequency_directivity": True}),
    "nan_delays": lambda: (fs.L11_5v(), fs.MediumParams(), *speckle(6, 20),
                           np.where(np.arange(128) % 17 == 3, np.nan, 0.0), {}),
    "p4_nsub3_fulldir": lambda: (fs.P4_2v().model_copy(update={"baffle": "rigid"}), fs.MediumParams(attenuation=0.2),
                                 *speckle(7, 37, zlim=(0.01, 0.08)),
                                 fs.focused(_pos(fs.P4_2v()), np.array([0.005, 0.04]), speed_of_sound=1540.0),
                                 {"element_splitting": 3, "full_frequency_directivity": True}),
}


@pytest.mark.parametrize("name", list(CASES))
def test_fp32_matches_oracle(name, cuda):
    params, medium, x, z, rc, D, kw = CASES[name]()
    print(name, flush=True)
    rf, _ = _run("hip", x, z, rc, D, params, medium, **kw)
    want = _oracle_rf(x, z, rc, D, params, medium, **kw)
    assert rel_l2(rf, want) < FP32_TOL


@pytest.mark.parametrize("name", list(CASES))
def test_mfma_matches_oracle(name, cuda):
    params, medium, x, z, rc, D, kw = CASES[name]()
    print(name, flush=True)
    rf, _ = _run("hip_mfma", x, z, rc, D, params, medium, **kw)
    want = _oracle_rf(x, z, rc, D, params, medium, **kw)
    assert rel_l2(rf, want) < MFMA_TOL


def test_multi_transmit_both_methods(cuda):
    p = fs.L11_5v()
    m = fs.MediumParams(attenuation=0.5)
    x, z, rc = speckle(7, 50)
    D = np.stack([fs.plane_wave(_pos(p), np.deg2rad(a), speed_of_sound=1540.0) for a in (-12, -3, 0, 7, 15)])
    want = _oracle_rf(x, z, rc, D, p, m)
    rf32, res = _run("hip", x, z, rc, D, p, m)
    rfmm, _ = _run("hip_mfma", x, z, rc, D, p, m)
    assert res.rf.shape[0] == 5 and res.spectrum.shape[0] == 5
    assert rel_l2(rf32, want) < FP32_TOL
    assert rel_l2(rfmm, want) < MFMA_TOL


def test_rf_stack_of_shared_plan_firings(cuda):
    """The multi-firing RF stack (samples, channels, firings) of the reference harness
    (tests/_picmus_phantom_shared.py:319-367: one plan from the widest firing, then one simus_compute per
    firing): simus_compute with 2-D delays and rf_stack give the same stack as per-firing calls."""
    p = fs.L11_5v()
    m = fs.MediumParams(attenuation=0.5)
    x, z, rc = speckle(9, 40)
    sc = np.stack([x, z], -1)
    D = np.stack([fs.plane_wave(_pos(p), np.deg2rad(a), speed_of_sound=1540.0) for a in (-6, 0, 9)])
    plan = fs.simus_precompute(sc, rc, D, p, m)
    per = [fs.simus_compute(sc, rc, D[i], plan, p, m, strategy="hip").rf for i in range(3)]
    stack = fs.rf_stack(fs.simus_compute(sc, rc, D, plan, p, m, strategy="hip").rf)
    assert stack.shape == (per[0].shape[0], 128, 3) and stack.dtype == torch.float32
    ref = fs.rf_stack(per)
    assert torch.equal(stack, ref)
    padded = fs.rf_stack(per, n_samples=per[0].shape[0] + 7)
    assert padded.shape[0] == per[0].shape[0] + 7 and torch.all(padded[-7:] == 0)


def test_phased_array_focused_beams_mfma(cuda):
    p = fs.P4_2v()
    x, z, rc = speckle(8, 60, xlim=(-0.04, 0.04), zlim=(0.01, 0.1))
    ang = np.deg2rad(np.linspace(-40, 40, 9))
    D = np.stack([fs.focused(_pos(p), np.array([0.06 * np.sin(a), 0.06 * np.cos(a)]), speed_of_sound=1540.0) for a in ang])
    want = _oracle_rf(x, z, rc, D, p, fs.MediumParams())
    rf, _ = _run("hip_mfma", x, z, rc, D, p, fs.MediumParams())
    assert rel_l2(rf, want) < MFMA_TOL


# ----------------------------------------------------------------- edge cases of the reference tests
def test_single_scatterer_and_zero_rc(cuda):
    """tests/test_simus.py:390-411."""
    p = fs.P4_2v()
    rf, _ = _run("hip", f32([0.0]), f32([3e-2]), f32([1.0]), np.zeros(64), p, fs.MediumParams())
    assert rf.shape[-1] == 64 and np.abs(rf).max() > 0
    x, z = f32(np.zeros(3)), f32(np.linspace(1e-2, 5e-2, 3))
    for strat in ("hip", "hip_mfma"):
        rf, _ = _run(strat, x, z, f32(np.zeros(3)), np.zeros(64), p, fs.MediumParams())
        assert np.abs(rf).max() < 1e-10


def test_attenuation_lowers_power(cuda):
    """tests/test_simus.py:413-427."""
    p = fs.P4_2v()
    x, z, rc = f32(np.zeros(3)), f32(np.linspace(1e-2, 5e-2, 3)), f32(np.ones(3))
    rf0, _ = _run("hip", x, z, rc, np.zeros(64), p, fs.MediumParams(attenuation=0.0))
    rf1, _ = _run("hip", x, z, rc, np.zeros(64), p, fs.MediumParams(attenuation=0.5))
    assert np.sum(rf1**2) < np.sum(rf0**2)


def test_out_of_field_and_empty(cuda):
    p = fs.L11_5v()
    x, z, rc = speckle(9, 20)
    a, _ = _run("hip", x, z, rc, np.zeros(128), p, fs.MediumParams())
    b, _ = _run("hip", np.append(x, 0.0), np.append(z, -0.01), np.append(rc, 3.0), np.zeros(128), p, fs.MediumParams())
    assert rel_l2(b, a) < 1e-6
    plan = fs.simus_precompute(np.stack([x, z], -1), rc, np.zeros(128), p)
    for strat in ("hip", "hip_mfma"):
        r = fs.simus_compute(np.zeros((0, 2)), np.zeros(0), np.zeros(128), plan, p, strategy=strat)
        assert r.rf.shape == (plan.n_fft // 2, 128) and float(r.rf.abs().max()) == 0.0


def test_scatterer_on_element_is_clipped(cuda):
    p = fs.L11_5v()
    x = f32([float(_pos(p)[40, 0]), 0.0])
    z = f32([0.0, 0.02])
    rc = f32([1.0, 1.0])
    rf, _ = _run("hip", x, z, rc, np.zeros(128), p, fs.MediumParams())
    want = _oracle_rf(x, z, rc, np.zeros(128), p, fs.MediumParams())
    assert np.all(np.isfinite(rf)) and rel_l2(rf, want) < FP32_TOL


def test_precompute_compute_equals_simus_and_reproducible(cuda):
    """tests/test_simus.py:302-319, plus bitwise run-to-run reproducibility (no atomics in the sums)."""
    import torch

    p = fs.P4_2v()
    sc = np.stack([np.zeros(3), np.linspace(1e-2, 5e-2, 3)], -1)
    a = fs.simus(sc, np.ones(3), np.zeros(64), p)
    plan = fs.simus_precompute(sc, np.ones(3), np.zeros(64), p)
    b = fs.simus_compute(sc, np.ones(3), np.zeros(64), plan, p)
    assert torch.equal(a.rf, b.rf) and torch.equal(a.spectrum, b.spectrum)
    x, z, rc = speckle(10, 3000)
    for strat in ("hip", "hip_mfma"):
        r1 = fs.simus(np.stack([x, z], -1), rc, np.zeros(64), p, strategy=strat)
        r2 = fs.simus(np.stack([x, z], -1), rc, np.zeros(64), p, strategy=strat)
        assert torch.equal(r1.rf, r2.rf)


def test_torch_inputs_match_numpy(cuda):
    import torch

    p = fs.L11_5v()
    x, z, rc = speckle(11, 40)
    sc = np.stack([x, z], -1)
    a = fs.simus(sc, rc, np.zeros(128), p)
    b = fs.simus(torch.as_tensor(sc, device=cuda), torch.as_tensor(rc, device=cuda),
                 torch.zeros(128, dtype=torch.float64, device=cuda), p)
    assert torch.equal(a.rf, b.rf)
    assert b.rf.device.type == "cuda"


def test_spectrum_layout_and_rf_tail(cuda):
    p = fs.L11_5v()
    m = fs.MediumParams()
    x, z, rc = speckle(12, 10)
    plan_o = O.simus_plan(x, z, np.zeros(128), p, m)
    rf_o, full_o, _ = O.simus_oracle(x, z, rc, np.zeros(128), p, m, plan_o)
    _, res = _run("hip", x, z, rc, np.zeros(128), p, m)
    full = res.spectrum.cpu().numpy()
    assert full.shape == full_o.shape
    assert np.abs(full[: plan_o["idx0"]]).max() == 0.0
    assert rel_l2(full, full_o) < FP32_TOL


# ----------------------------------------------------------------- full-size, size-independent properties
def _cfg1():
    rng = np.random.default_rng(2026)
    p = fs.TransducerParams(freq_center=7.6e6, pitch=3.0e-4, width=2.7e-4, n_elements=128, bandwidth=0.77)
    x = np.concatenate([rng.uniform(-0.02, 0.02, 50000), np.zeros(10)])
    z = np.concatenate([rng.uniform(0.005, 0.05, 50000), np.arange(1, 11) * 5e-3])
    rc = np.concatenate([rng.standard_normal(50000), np.full(10, 10.0)])
    return p, fs.MediumParams(attenuation=0.5), f32(x), f32(z), f32(rc)


def test_full_size_linearity_and_symmetry(cuda):
    import torch

    p, m, x, z, rc = _cfg1()
    sc = np.stack([x, z], -1)
    plan = fs.simus_precompute(sc, rc, np.zeros(128), p, m)
    half = rc.copy()
    half[::2] = 0.0
    rest = rc - half
    s_all = fs.simus_compute(sc, rc, np.zeros(128), plan, p, m).spectrum
    s_a = fs.simus_compute(sc, half, np.zeros(128), plan, p, m).spectrum
    s_b = fs.simus_compute(sc, rest, np.zeros(128), plan, p, m).spectrum
    assert float(torch.linalg.norm(s_a + s_b - s_all) / torch.linalg.norm(s_all)) < 1e-5
    mir = fs.simus_compute(np.stack([-x, z], -1), rc, np.zeros(128), plan, p, m).spectrum
    assert float(torch.linalg.norm(mir.flip(-1) - s_all) / torch.linalg.norm(s_all)) < 1e-5
    mm = fs.simus_compute(sc, rc, np.zeros(128), plan, p, m, strategy="hip_mfma").rf
    r32 = fs.simus_compute(sc, rc, np.zeros(128), plan, p, m).rf
    assert float(torch.linalg.norm(mm - r32) / torch.linalg.norm(r32)) < MFMA_TOL


@pytest.mark.parametrize("n_sub", [1, 3])
def test_two_block_tx_path_other_sub_element_counts(n_sub, cuda):
    """40 000 scatterers x 52 frequency blocks: the TX kernel takes two 32-frequency blocks per thread
    (at least 2048 workgroups). The plain spectrum against the split-BF16 kernel, which shares no code
    with it, and superposition over two halves of the reflectivity."""
    import torch

    p, m = fs.L11_5v(), fs.MediumParams(attenuation=0.3)
    x, z, rc = speckle(21, 40000)
    sc = np.stack([x, z], -1)
    plan = fs.simus_precompute(sc, rc, np.zeros(128), p, m, element_splitting=n_sub)
    assert plan.n_sub == n_sub
    n_blocks = -(-plan.selected_freqs.shape[0] // 32)
    assert -(-x.size // 256) * -(-n_blocks // 2) >= 2048
    r32 = fs.simus_compute(sc, rc, np.zeros(128), plan, p, m)
    rmm = fs.simus_compute(sc, rc, np.zeros(128), plan, p, m, strategy="hip_mfma")
    assert float(torch.linalg.norm(rmm.rf - r32.rf) / torch.linalg.norm(r32.rf)) < MFMA_TOL
    half = rc.copy()
    half[1::2] = 0.0
    s_a = fs.simus_compute(sc, half, np.zeros(128), plan, p, m).spectrum
    s_b = fs.simus_compute(sc, rc - half, np.zeros(128), plan, p, m).spectrum
    assert float(torch.linalg.norm(s_a + s_b - r32.spectrum) / torch.linalg.norm(r32.spectrum)) < 1e-5


# ----------------------------------------------------------------- pfield
def test_pfield_matches_oracle(cuda):
    p = fs.L11_5v()
    g = f32(np.stack(np.meshgrid(np.linspace(-0.012, 0.012, 13), np.linspace(0.004, 0.035, 11)), -1))
    d = fs.focused(_pos(p), np.array([0.003, 0.02]), speed_of_sound=1540.0)
    m = fs.MediumParams(attenuation=0.3)
    pf = fs.pfield(g, d, p, m).cpu().numpy()
    assert pf.shape == g.shape[:-1]
    pl = O.pfield_plan(g[..., 0].ravel(), g[..., 1].ravel(), d, p, m)
    want = O.pfield_oracle(g[..., 0].ravel(), g[..., 1].ravel(), d, p, m, pl).reshape(g.shape[:-1])
    assert rel_l2(pf, want) < FP32_TOL


def test_pfield_large_grid_matches_subset_and_oracle(cuda):
    """A grid large enough for the TX kernel to take two 32-frequency blocks per thread (at least
    2048 workgroups): 200 of its points against the same points computed alone under the same plan
    (one block per thread, the same per-point arithmetic), and 16 of them against the FP64 oracle."""
    p = fs.L11_5v()
    m = fs.MediumParams(attenuation=0.5)
    g = f32(np.stack(np.meshgrid(np.linspace(-0.02, 0.02, 300), np.linspace(0.002, 0.05, 256), indexing="ij"), -1))
    d = fs.focused(_pos(p), np.array([0.0, 0.02]), speed_of_sound=1540.0)
    plan = fs.pfield_precompute(g, d, p, m)
    n_blocks = -(-plan.selected_freqs.shape[0] // 32)
    assert -(-g[..., 0].size // 256) * -(-n_blocks // 2) >= 2048
    full = fs.pfield_compute(g, d, plan, p, m).cpu().numpy().ravel()
    pts = g.reshape(-1, 2)
    idx = np.sort(np.random.default_rng(11).choice(full.size, 200, replace=False))
    sub = fs.pfield_compute(pts[idx], d, plan, p, m).cpu().numpy()
    np.testing.assert_allclose(full[idx], sub, rtol=1e-6, atol=0)
    pl = O.pfield_plan(pts[:, 0].astype(np.float64), pts[:, 1].astype(np.float64), d, p, m)
    assert pl["freqs"].size == plan.selected_freqs.shape[0]
    o = idx[:16]
    want = O.pfield_oracle(pts[o, 0].astype(np.float64), pts[o, 1].astype(np.float64), d, p, m, pl)
    assert rel_l2(full[o], want) < FP32_TOL


def test_pfield_convex_out_of_field(cuda):
    p = fs.C5_2v()
    g = f32(np.array([[0.0, 0.03], [0.0, 0.001], [0.01, 0.04], [0.0, -0.01]]))
    pf = fs.pfield(g, np.zeros(128), p).cpu().numpy()
    pl = O.pfield_plan(g[:, 0], g[:, 1], np.zeros(128), p, fs.MediumParams())
    want = O.pfield_oracle(g[:, 0], g[:, 1], np.zeros(128), p, fs.MediumParams(), pl)
    assert pf[3] == 0.0 and pf[1] == 0.0  # behind the array / inside the arc
    assert rel_l2(pf, want) < FP32_TOL


# ----------------------------------------------------------------- element counts that do not fill 16-element tiles
@pytest.mark.parametrize("n_elem", [1, 7, 33, 100, 192, 1100])
def test_ragged_element_counts(n_elem, cuda):
    """Partial element tiles of both methods; the batched method runs up to 1024 elements (NotImplementedError
    above)."""
    p = fs.TransducerParams(freq_center=5e6, pitch=3.0e-4, width=2.5e-4, n_elements=n_elem, bandwidth=0.7)
    m = fs.MediumParams(attenuation=0.3)
    x, z, rc = speckle(90 + n_elem, 23, xlim=(-0.01, 0.01), zlim=(0.005, 0.03))
    D = np.stack([fs.plane_wave(_pos(p), np.deg2rad(a), speed_of_sound=1540.0) for a in (-4.0, 9.0)])
    want = _oracle_rf(x, z, rc, D, p, m)
    print("n_elem", n_elem, flush=True)
    rf32, _ = _run("hip", x, z, rc, D, p, m)
    assert rel_l2(rf32, want) < FP32_TOL
    if n_elem > 1024:
        with pytest.raises(NotImplementedError):
            _run("hip_mfma", x, z, rc, D, p, m)
    else:
        rfm, _ = _run("hip_mfma", x, z, rc, D, p, m)
        assert rel_l2(rfm, want) < MFMA_TOL


def test_device_plan_matches_host_plan(cuda):
    """simus_precompute / pfield_precompute on device tensors (one fs_max_distance pass) give the
    plans the host (NumPy FP64) computation gives; the reduction itself matches NumPy in FP64."""
    import torch

    from fastsimus_amd.plan import _max_distance

    x, z, rc = speckle(11, 5003)
    params, m = fs.L11_5v(), fs.MediumParams(attenuation=0.5)
    pos = _pos(params)
    D = np.stack([fs.plane_wave(pos, np.deg2rad(a), speed_of_sound=1540.0) for a in (-8.0, 8.0)])
    sc = np.stack([x, z], -1)
    host = fs.simus_precompute(sc, rc, D, params, m)
    dev = fs.simus_precompute(torch.as_tensor(sc, device=cuda), rc, D, params, m)
    assert (host.n_freq_full, host.freq_idx_start, host.n_fft, host.n_sub) == \
        (dev.n_freq_full, dev.freq_idx_start, dev.n_fft, dev.n_sub)
    np.testing.assert_allclose(dev.selected_freqs, host.selected_freqs, rtol=1e-12)
    xd = torch.as_tensor(x, dtype=torch.float64, device=cuda)
    zd = torch.as_tensor(z, dtype=torch.float64, device=cuda)
    want = float(np.sqrt((x[:, None] - pos[:, 0]) ** 2 + (z[:, None] - pos[:, 1]) ** 2).max())
    assert abs(_max_distance(xd, zd, pos[:, 0], pos[:, 1]) - want) <= 1e-14 * want
    assert _max_distance(xd[:0], zd[:0], pos[:, 0], pos[:, 1]) == 0.0
    grid = np.stack(np.meshgrid(np.linspace(-0.02, 0.02, 41), np.linspace(0.0, 0.04, 37)), -1).reshape(-1, 2)
    ph = fs.pfield_precompute(grid, D[0], params, m)
    pd = fs.pfield_precompute(torch.as_tensor(grid, device=cuda), D[0], params, m)
    np.testing.assert_allclose(pd.selected_freqs, ph.selected_freqs, rtol=1e-12)


def test_hip_path_never_calls_the_oracle(cuda, monkeypatch):
    """With every function of the oracle replaced by one that raises, both HIP methods still run
    (no CPU fallback through the oracle); cf. tests/backend/test_cupy.py:66-95 of the reference."""
    import types

    import oracle.simus_oracle as OM

    def _boom(*a, **k):
        raise AssertionError("the oracle was called by the product path")

    for name, obj in list(vars(OM).items()):
        if isinstance(obj, types.FunctionType):
            monkeypatch.setattr(OM, name, _boom)
    params, medium, x, z, rc, D, kw = CASES["l11_speckle_atten"]()
    for strategy in ("hip", "hip_mfma"):
        rf, _ = _run(strategy, x, z, rc, D, params, medium, **kw)
        assert np.isfinite(rf).all() and np.abs(rf).max() > 0


@pytest.mark.parametrize("n_elem,n_tx", [(48, 17), (64, 170), (100, 40), (128, 90)])
def test_mfma_many_transmits_against_plain_and_fp64(n_elem, n_tx, cuda):
    """Batched method with more than 16 transmits, where the multistatic matrix is applied to the transmits on
    the matrix cores (64- and 128-element tiles; 170 and 90 transmits take two passes of transmit tiles):
    against the plain FP32 method on every transmit and the FP64 oracle on two of them."""
    p = fs.TransducerParams(freq_center=7.6e6, pitch=3.0e-4, width=2.7e-4, n_elements=n_elem, bandwidth=0.77)
    m = fs.MediumParams(attenuation=0.5)
    half = 0.5 * p.pitch * n_elem
    x, z, rc = speckle(40 + n_elem, 300, xlim=(-half, half), zlim=(0.005, 0.03))
    pos = _pos(p)
    rng = np.random.default_rng(n_tx)
    D = np.stack([fs.plane_wave(pos, float(a), speed_of_sound=1540.0) for a in np.deg2rad(np.linspace(-15, 15, n_tx))])
    apod = rng.uniform(0.5, 1.0, D.shape)
    sc = np.stack([x, z], -1)
    rf_m = fs.simus(sc, rc, D, p, m, tx_apodization=apod, strategy="hip_mfma").rf.cpu().numpy()
    rf_p = fs.simus(sc, rc, D, p, m, tx_apodization=apod, strategy="hip").rf.cpu().numpy()
    e = rel_l2(rf_m, rf_p)
    print(f"E={n_elem} n_tx={n_tx}: split-BF16 vs plain FP32 over all transmits rel-L2 {e:.2e}", flush=True)
    assert e < 1e-4
    plan = O.simus_plan(x, z, D, p, m)
    for t in (0, n_tx - 1):
        want, _, _ = O.simus_oracle(x, z, rc, D[t], p, m, plan, apod=apod[t])
        err = rel_l2(rf_m[t], want)
        print(f"  tx {t}: split-BF16 vs FP64 rel-L2 {err:.2e}", flush=True)
        assert err < MFMA_TOL
