"""Host-side logic of the product (CPU only): parameter models, geometry, spectra,
delay laws and frequency plans, checked against the oracle's independent
restatement and the reference's known answers; plus the C-ABI library's
exported symbols and argument checks (no kernel launches)."""

from __future__ import annotations

import ctypes as C
import re
from math import inf, pi
from pathlib import Path

import numpy as np
import pytest

import fastsimus_amd as fs
from fastsimus_amd import _lib
from fastsimus_amd.plan import pfield_precompute, simus_precompute, two_way_pulse_duration
from oracle import simus_oracle as O
from tests.conftest import ROOT, speckle


# ----------------------------------------------------------------- parameter models (transducer_params.py:135-160)
def test_params_width_or_kerf():
    with pytest.raises(ValueError):
        fs.TransducerParams(freq_center=1e6, pitch=3e-4, n_elements=8)
    with pytest.raises(ValueError):
        fs.TransducerParams(freq_center=1e6, pitch=3e-4, n_elements=8, width=2e-4, kerf=1e-4)
    with pytest.raises(ValueError):
        fs.TransducerParams(freq_center=1e6, pitch=3e-4, n_elements=8, width=4e-4)
    with pytest.raises(ValueError):
        fs.TransducerParams(freq_center=1e6, pitch=3e-4, n_elements=8, kerf=3e-4)
    p = fs.TransducerParams(freq_center=2.5e6, pitch=300e-6, n_elements=128, width=250e-6)
    assert p.element_width == 250e-6 and p.kerf_width == pytest.approx(5e-5)
    q = fs.TransducerParams(freq_center=3.5e6, pitch=400e-6, n_elements=64, kerf=50e-6, radius=50e-3)
    assert q.element_width == pytest.approx(350e-6)


def test_params_frozen_and_baffle():
    p = fs.L11_5v()
    with pytest.raises(Exception):
        p.pitch = 1.0
    assert fs.TransducerParams(freq_center=1e6, pitch=3e-4, n_elements=2, width=1e-4, baffle="rigid").baffle == fs.BaffleType.RIGID
    assert fs.TransducerParams(freq_center=1e6, pitch=3e-4, n_elements=2, width=1e-4, baffle=1.75).baffle == 1.75
    with pytest.raises(ValueError):
        fs.MediumParams(speed_of_sound=0)
    with pytest.raises(ValueError):
        fs.MediumParams(attenuation=-1)


def test_presets():
    assert fs.P4_2v().n_elements == 64 and fs.P4_2v().element_width == pytest.approx(250e-6)
    assert fs.L11_5v().element_width == pytest.approx(270e-6)
    assert fs.L12_3v().n_elements == 192
    assert fs.C5_2v().radius == pytest.approx(49.57e-3)


# ----------------------------------------------------------------- geometry / spectra / delays vs oracle
@pytest.mark.parametrize("preset", [fs.P4_2v, fs.L11_5v, fs.C5_2v])
def test_element_positions_match_oracle(preset):
    p = preset()
    pos, theta, apex = fs.element_positions(p.n_elements, p.pitch, p.radius)
    ex, ez, th, ap = O.element_layout(p)
    np.testing.assert_array_equal(pos[:, 0], ex)
    np.testing.assert_allclose(pos[:, 1], ez, atol=1e-15)
    assert apex == pytest.approx(ap)
    if p.radius == inf:
        assert theta is None
    else:
        np.testing.assert_allclose(theta, th)
        # convex chord endpoints sit at z = 0 (geometry.py:70-84)
        assert abs(pos[0, 1]) < 1e-12 and abs(pos[-1, 1]) < 1e-12


def test_spectra_match_oracle():
    w = 2 * pi * np.linspace(0, 15e6, 333)
    np.testing.assert_allclose(fs.spectrum.pulse_spectrum(w, 7.6e6, 2), O.pulse_spectrum(w, 7.6e6, 2), atol=1e-15)
    np.testing.assert_allclose(fs.spectrum.probe_spectrum(w, 7.6e6, 0.77), O.probe_spectrum(w, 7.6e6, 0.77), rtol=1e-12)
    with pytest.raises(ValueError):
        fs.spectrum.probe_spectrum(w, 7.6e6, 0.0)


def test_plane_wave_and_focused_delays():
    p = fs.L11_5v()
    pos, _, _ = fs.element_positions(128, p.pitch, p.radius)
    d0 = fs.plane_wave(pos, 0.0, speed_of_sound=1540.0)
    assert np.all(d0 == 0)  # tests/test_tx_delay.py:120-133
    d = fs.plane_wave(pos, np.deg2rad(10), speed_of_sound=1540.0)
    assert d.min() == 0 and np.all(np.diff(d) > 0)
    np.testing.assert_allclose(np.diff(d), p.pitch * np.sin(np.deg2rad(10)) / 1540.0, rtol=1e-5)
    with pytest.raises(ValueError):
        fs.plane_wave(pos, pi / 2, speed_of_sound=1540.0)
    f = fs.focused(pos, np.array([0.0, 0.03]), speed_of_sound=1540.0)
    # converging wave: edge elements fire first (zero delay), centre last
    assert f[0] == pytest.approx(0.0) and np.argmax(f) in (63, 64)
    dv = fs.focused(pos, np.array([0.0, -0.01]), speed_of_sound=1540.0)
    assert np.argmin(dv) in (63, 64)
    dw = fs.diverging_wave(pos, 0.0, np.deg2rad(60), aperture_length=127 * p.pitch, speed_of_sound=1540.0)
    assert dw.min() == 0 and np.argmin(dw) in (63, 64)


def test_delays_keep_torch_type():
    torch = pytest.importorskip("torch")
    p = fs.L11_5v()
    pos, _, _ = fs.element_positions(128, p.pitch, p.radius)
    d = fs.plane_wave(torch.as_tensor(pos), 0.1, speed_of_sound=1540.0)
    assert isinstance(d, torch.Tensor) and d.shape == (128,)


# ----------------------------------------------------------------- plans vs oracle
@pytest.mark.parametrize("preset,seed", [(fs.L11_5v, 0), (fs.P4_2v, 1), (fs.C5_2v, 2)])
def test_simus_plan_matches_oracle(preset, seed):
    p = preset()
    x, z, rc = speckle(seed, 40, zlim=(0.005, 0.06))
    pos, _, _ = fs.element_positions(p.n_elements, p.pitch, p.radius)
    d = fs.plane_wave(pos, 0.05, speed_of_sound=1540.0) if p.radius == inf else np.zeros(p.n_elements)
    m = fs.MediumParams(attenuation=0.3)
    plan = simus_precompute(np.stack([x, z], -1), rc, d, p, m)
    o = O.simus_plan(x, z, d, p, m)
    np.testing.assert_array_equal(plan.selected_freqs, o["freqs"])
    np.testing.assert_allclose(plan.pulse_spectrum, o["pulse"], atol=1e-15)
    np.testing.assert_allclose(plan.probe_spectrum, o["probe"], rtol=1e-12)
    assert (plan.n_sub, plan.n_freq_full, plan.freq_idx_start, plan.n_fft) == (o["n_sub"], o["n_ff"], o["idx0"], o["n_fft"])
    assert plan.correction_factor == pytest.approx(o["corr"], rel=1e-15)
    assert plan.seg_length == pytest.approx(o["seg"])


def test_two_way_duration_matches_oracle():
    for fc, bw, n in [(2.72e6, 0.74, 1), (7.6e6, 0.77, 1), (5e6, 0.6, 2.5)]:
        assert two_way_pulse_duration(fc, bw, n) == pytest.approx(O.two_way_duration(fc, bw, n), abs=1e-15)


def test_plan_known_answer_p4_focused():
    """tests/test_simus.py:355-386."""
    p = fs.P4_2v()
    sc = np.stack([np.zeros(6), np.linspace(1e-2, 8e-2, 6)], -1)
    pos, _, _ = fs.element_positions(64, p.pitch, p.radius)
    d = fs.focused(pos, np.array([0.0, 0.03]), speed_of_sound=1540.0)
    plan = simus_precompute(sc, np.ones(6), d, p, fs=4 * p.freq_center)
    assert abs(plan.selected_freqs.shape[0] - 812) < 20


def test_batched_plan_uses_largest_delay():
    p = fs.L11_5v()
    pos, _, _ = fs.element_positions(128, p.pitch, p.radius)
    D = np.stack([fs.plane_wave(pos, np.deg2rad(a), speed_of_sound=1540.0) for a in (-16, 0, 10)])
    x, z, rc = speckle(3, 30)
    sc = np.stack([x, z], -1)
    plan_b = simus_precompute(sc, rc, D, p)
    plan_0 = simus_precompute(sc, rc, D[0], p)  # the -16 deg firing has the largest span
    np.testing.assert_array_equal(plan_b.selected_freqs, plan_0.selected_freqs)


def test_nan_delays_ignored_in_plan():
    p = fs.L11_5v()
    x, z, rc = speckle(4, 10)
    d = np.zeros(128)
    d[3] = np.nan
    a = simus_precompute(np.stack([x, z], -1), rc, d, p)
    b = simus_precompute(np.stack([x, z], -1), rc, np.zeros(128), p)
    np.testing.assert_array_equal(a.selected_freqs, b.selected_freqs)


def test_pfield_plan_matches_oracle():
    p = fs.L11_5v()
    x, z, _ = speckle(5, 30)
    d = np.zeros(128)
    plan = pfield_precompute(np.stack([x, z], -1), d, p)
    o = O.pfield_plan(x, z, d, p, fs.MediumParams())
    np.testing.assert_array_equal(plan.selected_freqs, o["freqs"])
    assert plan.correction_factor == pytest.approx(o["corr"], rel=1e-15)
    with pytest.raises(ValueError):
        pfield_precompute(np.zeros((0, 2)), d, p)


def test_strategy_names():
    from fastsimus_amd.simus import _method

    assert _method(None) == "fp32" and _method("cuda") == "fp32" and _method("hip_mfma") == "mfma"
    # default: the matrix-core path from 3 transmits on, for arrays it takes
    assert _method(None, 3, 128) == "mfma" and _method(None, 75, 1024) == "mfma"
    assert _method(None, 2, 128) == "fp32" and _method(None, 75, 1025) == "fp32"
    assert _method("hip", 75, 128) == "fp32"
    for s in ("python", "scan", "metal"):
        with pytest.raises(NotImplementedError):
            _method(s)


# ----------------------------------------------------------------- the C-ABI library (loads, exports, rejects; no GPU needed)
def _declared_symbols():
    hdr = (ROOT / "include" / "fastsimus.h").read_text()
    return sorted(set(re.findall(r"^\s*(?:int|size_t|double|const char\*)\s+(fs_\w+)\s*\(", hdr, flags=re.M)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    declared = _declared_symbols()
    assert set(declared) == set(_lib.EXPORTS)
    for name in declared:
        assert hasattr(lib, name), name
    assert lib.fs_version() >= 100
    assert lib.fs_status_string(0) == b"ok"


def test_library_rejects_bad_descriptors():
    lib = _lib.load()
    d = _lib.SimusDesc()
    assert lib.fs_simus_workspace_bytes(C.byref(d), 0) == 0  # empty descriptor
    assert lib.fs_simus_spectrum(None, 0, None, None, 0, None) == _lib.FS_ERR_INVALID
    assert lib.fs_spectrum_to_rf(None, 1, 1, 1, 1, 0, 4, None, None, None, None, 0, None) == _lib.FS_ERR_INVALID
    assert lib.fs_pfield(None, None, None, 0, None) == _lib.FS_ERR_INVALID
    assert lib.fs_simus_adjoint(None, 0, None, None, None, None, None, None, 0, None) == _lib.FS_ERR_INVALID
    assert lib.fs_simus_spectrum_batch(None, 1, 1, None, None, None, 0, None) == _lib.FS_ERR_INVALID
    assert lib.fs_tuning_get(None, 1, None) == _lib.FS_ERR_INVALID
    with pytest.raises(ValueError):
        _lib.check(_lib.FS_ERR_INVALID, "x")
    with pytest.raises(NotImplementedError):
        _lib.check(_lib.FS_ERR_UNSUPPORTED, "x")
    with pytest.raises(RuntimeError):
        _lib.check(_lib.FS_ERR_WORKSPACE, "x")


def test_compute_fails_loudly_without_gpu():
    torch = pytest.importorskip("torch")
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    p = fs.L11_5v()
    with pytest.raises(RuntimeError, match="GPU"):
        fs.simus(np.array([[0.0, 0.02]]), np.ones(1), np.zeros(128), p)


def test_missing_library_raises_instead_of_falling_back(monkeypatch, tmp_path):
    """The product path has no CPU route: without libfastsimus.so every compute call raises."""
    monkeypatch.setattr(_lib, "LIB_PATH", tmp_path / "libfastsimus.so")
    monkeypatch.setattr(_lib, "_lib", None)
    with pytest.raises(RuntimeError, match="not built"):
        _lib.load()


def test_product_code_never_imports_the_oracle():
    """The package (the product path) has no route to the FP64 oracle: no module of it imports
    `oracle` (the reference's tests check the same of its CUDA path, tests/backend/test_cupy.py:66-95)."""
    import ast
    from pathlib import Path

    import fastsimus_amd

    pkg = Path(fastsimus_amd.__file__).resolve().parent
    for f in sorted(pkg.rglob("*.py")):
        tree = ast.parse(f.read_text())
        for node in ast.walk(tree):
            if isinstance(node, ast.Import):
                names = [a.name for a in node.names]
            elif isinstance(node, ast.ImportFrom):
                names = [node.module or ""]
            else:
                continue
            assert not any(n == "oracle" or n.startswith("oracle.") for n in names), f"{f.name} imports the oracle"


def test_rf_stack_layout_and_padding():
    """(samples, channels, firings), ragged firings zero-padded at the end of the sample axis, padded
    (never cut) to n_samples, float32 or wider (tests/_picmus_phantom_shared.py:280-295)."""
    import torch

    rng = np.random.default_rng(0)
    a = rng.standard_normal((5, 4)).astype(np.float32)
    b = rng.standard_normal((3, 4)).astype(np.float32)
    st = fs.rf_stack([a, b])
    assert st.shape == (5, 4, 2) and st.dtype == torch.float32
    assert np.array_equal(st[:, :, 0].numpy(), a)
    assert np.array_equal(st[:3, :, 1].numpy(), b) and np.all(st[3:, :, 1].numpy() == 0)
    assert fs.rf_stack([a, b], n_samples=8).shape == (8, 4, 2)
    assert fs.rf_stack([a, b], n_samples=2).shape == (5, 4, 2)
    assert fs.rf_stack([a.astype(np.float64)]).dtype == torch.float64
    t = torch.as_tensor(rng.standard_normal((3, 6, 2)).astype(np.float32))
    assert torch.equal(fs.rf_stack(t), t.permute(1, 2, 0))
    with pytest.raises(ValueError):
        fs.rf_stack([a, np.zeros((5, 3), np.float32)])
    with pytest.raises(ValueError):
        fs.rf_stack([])


def _host_desc(n_scat=10, n_elem=8, n_sub=1, n_freq=40, n_tx=2):
    """A descriptor whose pointers are non-null placeholders: valid for size queries and argument checks that
    return before any device work, never for a launch."""
    d = _lib.SimusDesc()
    d.n_scat, d.n_elem, d.n_sub, d.n_freq, d.n_tx = n_scat, n_elem, n_sub, n_freq, n_tx
    for f in ("scat_x", "scat_z", "rc", "sub_x", "sub_z", "elem_cos", "elem_sin", "tx_delay", "tx_apod", "pulse_re",
              "pulse_im", "probe"):
        setattr(d, f, 8)
    d.f0, d.df, d.speed_of_sound, d.freq_center, d.seg_length = 2e6, 1e4, 1540.0, 5e6, 2e-4
    d.radius, d.correction = inf, 1.0
    return d


def test_adjoint_method_is_a_per_call_argument():
    """The adjoint kernels are chosen per call (0 auto, 1 matrix cores, 2 vector kernels): the workspace query
    returns 0 and the call FS_ERR_INVALID for any other value, before any GPU work; there is no process-wide
    switch left to race on."""
    lib = _lib.load()
    assert not hasattr(lib, "fs_set_adjoint_method") or "fs_set_adjoint_method" not in _lib.EXPORTS
    d = _host_desc()
    for m in (-1, 3):
        assert lib.fs_simus_adjoint_workspace_bytes(C.byref(d), m) == 0
        assert lib.fs_simus_adjoint(C.byref(d), m, 1, 1, 1, 1, 1, 1, 1 << 30, None) == _lib.FS_ERR_INVALID
    sizes = [lib.fs_simus_adjoint_workspace_bytes(C.byref(d), m) for m in (0, 1, 2)]
    assert all(s > 0 for s in sizes)
    assert sizes[0] == sizes[1] != sizes[2]  # auto takes the matrix cores at 8 elements, 1 sub-element


@pytest.mark.parametrize("n_elem,n_sub", [(64, 1), (128, 2), (192, 2), (200, 3), (130, 6)])
def test_batched_workspace_layout(n_elem, n_sub):
    """Workspace of the batched method for a 40-scatterer problem (one scatterer chunk): partial multistatic
    matrices of 16 ET x 16 ET per unit and padded frequency (one unit, or one per pair of 64-element groups
    above 128 elements), the scatterer table, the geometry table over elements padded to the 64-element
    groups (at most 4 sub-elements) and the queue word. Computed on the host, no device needed."""
    lib = _lib.load()
    S, NF = 40, 101
    d = _host_desc(n_scat=S, n_elem=n_elem, n_sub=n_sub, n_freq=NF, n_tx=3)
    got = lib.fs_simus_workspace_bytes(C.byref(d), _lib.FS_METHOD_MFMA_BF16X3)
    al = lambda v: (v + 255) // 256 * 256  # noqa: E731
    groups = -(-n_elem // 64)
    units = 1 if groups <= 2 else groups * (groups - 1) // 2
    ep = 64 if n_elem <= 64 else 128
    fbk = 2 if n_sub > 2 else 4
    nfpad = -(-NF // fbk) * fbk
    kpart = units * nfpad * ep * ep * 8
    terms = (S + 1) * 64 * groups * n_sub
    table = al(terms * 8) + al(terms * 8) + al(terms * 8) if n_sub <= 4 else 0
    assert got == al(kpart) + al(S * 16) + table + 256


def test_bench_traffic_profiles_match_their_workloads():
    """The bench lines' HBM traffic comes from the committed rocprofv3 passes of exactly their workload
    (FETCH_SIZE x fetch_scale + WRITE_SIZE x write_scale), and from none for other sizes."""
    import json

    import bench

    for path, key in ((bench.TRAFFIC_PROFILE, (200_000, 75, 1892, 128)),
                      (bench.ADJM_TRAFFIC_PROFILE, (65_536, 15, 1711, 128))):
        d = json.loads(path.read_text())
        assert (d["workload"]["n_scat"], d["workload"]["n_tx"], d["workload"]["n_freq"], d["workload"]["n_elem"]) == key
        assert abs(d["fetch_scale"] - 2.0) < 1e-3 and d["write_scale"] == 1.0
        assert bench.measured_traffic(*key, profile=path) == d["fetch_bytes"] + d["write_bytes"]
        assert bench.measured_traffic(key[0] + 1, *key[1:], profile=path) is None


@pytest.mark.parametrize("tilt_deg", [0.0, 10.0, -25.0])
def test_convex_plane_wave_is_the_projection_on_the_direction(tilt_deg):
    """A plane wave travelling along u = (sin tilt, cos tilt) reaches element p at u . p / c: the convex-array
    delays (the reference's point-to-line distance, tx_delay.py:50-66) equal (u . p - min u . p) / c."""
    p = fs.C5_2v()
    pos, _, _ = fs.element_positions(p.n_elements, p.pitch, p.radius)
    th = np.deg2rad(tilt_deg)
    d = fs.plane_wave(pos, th, speed_of_sound=1540.0, radius=p.radius)
    proj = pos @ np.array([np.sin(th), np.cos(th)])
    np.testing.assert_allclose(d, (proj - proj.min()) / 1540.0, rtol=0, atol=1e-15)


@pytest.mark.parametrize("tilt_deg,width_deg", [(0.0, 60.0), (10.0, 40.0), (-20.0, 90.0)])
def test_diverging_wave_virtual_source_spans_the_width(tilt_deg, width_deg):
    """The virtual source of a diverging wave sees the two ends of the aperture at tilt -+ width / 2 from the
    array normal, and the delays are those of that source (a focus behind the array)."""
    from fastsimus_amd.tx_delay import _angles_to_virtual_source

    q = fs.L11_5v()
    pos, _, _ = fs.element_positions(128, q.pitch, q.radius)
    aperture = 127 * q.pitch
    th, w = np.deg2rad(tilt_deg), np.deg2rad(width_deg)
    x0, z0 = _angles_to_virtual_source(aperture, th, w)
    assert z0 < 0
    a_left = np.arctan2(-aperture / 2 - x0, -z0)
    a_right = np.arctan2(aperture / 2 - x0, -z0)
    assert a_right - a_left == pytest.approx(w, abs=1e-12)
    assert (a_left + a_right) / 2 == pytest.approx(th, abs=1e-12)
    d = fs.diverging_wave(pos, th, w, aperture_length=aperture, speed_of_sound=1540.0)
    r = np.hypot(pos[:, 0] - x0, pos[:, 1] - z0)
    np.testing.assert_allclose(d, (r - r.min()) / 1540.0, rtol=0, atol=1e-15)
