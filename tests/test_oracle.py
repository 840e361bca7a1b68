"""Pins the FP64 oracle (CPU) against the reference's known answers, closed forms and brute force.

Every check here is chosen so that a plausible restatement mistake (dropped
term, wrong sign, wrong index, transposed operand) fails at least one of them.
Known answers come from the reference's own tests (cited per test).
"""

from __future__ import annotations

from math import cos, pi, sqrt

import numpy as np
import pytest

import fastsimus_amd as fs
from oracle import simus_oracle as O
from tests.conftest import f32, rel_l2, speckle


# ----------------------------------------------------------------- spectra (closed forms)
@pytest.mark.parametrize("fc,bw", [(2.72e6, 0.74), (7.6e6, 0.77), (3.57e6, 0.79), (5e6, 1.5)])
def test_probe_closed_forms(fc, bw):
    wc = 2 * pi * fc
    assert O.probe_spectrum(np.array([wc]), fc, bw)[0] == pytest.approx(1.0, abs=1e-15)
    edges = O.probe_spectrum(np.array([wc * (1 - bw / 2), wc * (1 + bw / 2)]), fc, bw)
    # two-way -6 dB at fc (1 +- B/2): one-way response 1/sqrt(2) (spectrum.py:92-107)
    np.testing.assert_allclose(edges, 1 / sqrt(2), rtol=1e-12)


def test_probe_rejects_bandwidth():
    with pytest.raises(ValueError):
        O.probe_spectrum(np.array([1.0]), 1e6, 2.0)


@pytest.mark.parametrize("n_w", [1, 2, 3])
def test_pulse_closed_forms(n_w):
    fc = 5e6
    w = 2 * pi * np.linspace(0.1e6, 9.9e6, 101)
    p = O.pulse_spectrum(w, fc, n_w)
    assert np.all(p.real == 0)  # purely imaginary
    np.testing.assert_allclose(O.pulse_spectrum(-w, fc, n_w), -p, atol=1e-15)  # odd in w
    # |pulse(wc)| = |1 - sinc(2 n_w)| = 1 for integer cycle counts
    assert abs(O.pulse_spectrum(np.array([2 * pi * fc]), fc, n_w)[0]) == pytest.approx(1.0, abs=1e-12)


@pytest.mark.parametrize("n_w", [1, 2, 3])
def test_pulse_sign_at_centre(n_w):
    """pulse(wc) = i (sinc(0) - sinc(2 n_w)) = +i exactly: a flipped sign would negate every RF."""
    fc = 5e6
    for fn in (O.pulse_spectrum, fs.spectrum.pulse_spectrum):
        v = fn(np.array([2 * pi * fc]), fc, n_w)[0]
        assert abs(v - 1j) < 1e-15, (fn.__module__, v)


# ----------------------------------------------------------------- convex geometry (utils/geometry.py:60-84,
# _pfield_math.py:41-67), oracle and product alike
def _convex_layouts(p):
    ex, ez, th, apex = O.element_layout(p)
    pos, theta, ap = fs.element_positions(p.n_elements, p.pitch, p.radius)
    return [("oracle", ex, ez, th, apex), ("product", pos[:, 0], pos[:, 1], theta, ap)]


@pytest.mark.parametrize("n,pitch,radius", [(128, 3e-4, 49.57e-3), (64, 4.7e-4, 20e-3), (5, 1e-3, 6e-3)])
def test_convex_elements_one_pitch_apart_on_the_arc(n, pitch, radius):
    p = fs.TransducerParams(freq_center=3.5e6, pitch=pitch, n_elements=n, kerf=pitch / 8, radius=radius)
    for who, ex, ez, th, apex in _convex_layouts(p):
        # adjacent centres are one pitch apart along the chord (chord = 2 R sin(asin(pitch / 2 R)))
        np.testing.assert_allclose(np.hypot(np.diff(ex), np.diff(ez)), pitch, rtol=1e-12, err_msg=who)
        # every centre on the circle of the radius around the arc centre (0, -apex), at angle theta
        np.testing.assert_allclose(np.hypot(ex, ez + apex), radius, rtol=1e-13, err_msg=who)
        np.testing.assert_allclose(ex, radius * np.sin(th), atol=1e-15, err_msg=who)
        # symmetric about the axis, chord endpoints at z = 0
        np.testing.assert_allclose(ex, -ex[::-1], atol=1e-15, err_msg=who)
        assert abs(ez[0]) < 1e-13 and abs(ez[-1]) < 1e-13, who


@pytest.mark.parametrize("n_sub", [1, 2, 3, 5])
def test_convex_subelements_on_the_tangent(n_sub):
    """Each sub-element centroid lies on its element's tangent at the offset o_j, oriented along increasing
    element angle: (sub - centre) . (centre - arc centre) = 0, |sub - centre| = |o_j|, and (sub - centre) has the
    direction of the chord from the previous to the next element (the tangent of equally spaced points)."""
    p = fs.C5_2v()
    E = p.n_elements
    W = p.element_width
    seg = W / n_sub
    o = -W / 2 + seg / 2 + seg * np.arange(n_sub)
    ex, ez, th, apex = O.element_layout(p)
    pos, theta, _ = fs.element_positions(E, p.pitch, p.radius)
    sub_o = np.stack(O.sub_elements(p, n_sub)[:2], -1)                    # (E, n_sub, 2)
    sx, sz, _, _ = fs.geometry.subelement_positions(p, n_sub)
    sub_p = np.stack([sx, sz], -1)
    for who, sub, cx, cz in (("oracle", sub_o, ex, ez), ("product", sub_p, pos[:, 0], pos[:, 1])):
        d = sub - np.stack([cx, cz], -1)[:, None, :]
        radial = np.stack([cx, cz + apex], -1)[:, None, :]
        np.testing.assert_allclose((d * radial).sum(-1), 0.0, atol=1e-15, err_msg=who)
        np.testing.assert_allclose(np.hypot(d[..., 0], d[..., 1]), np.abs(o)[None, :] * np.ones((E, 1)), rtol=1e-12,
                                   atol=1e-16, err_msg=who)
        chord = np.stack([cx[2:] - cx[:-2], cz[2:] - cz[:-2]], -1)
        chord /= np.hypot(chord[:, 0], chord[:, 1])[:, None]
        np.testing.assert_allclose(d[1:-1], o[None, :, None] * chord[:, None, :], atol=1e-15, err_msg=who)


# ----------------------------------------------------------------- plan (reference known answers)
def test_plan_p4_focused_frequency_count():
    """tests/test_simus.py:355-386: ~812 +- 20 selected frequencies."""
    p = fs.P4_2v()
    x = np.zeros(6)
    z = np.linspace(1e-2, 8e-2, 6)
    pos, _, _ = fs.element_positions(p.n_elements, p.pitch, p.radius)
    d = fs.focused(pos, np.array([0.0, 0.03]), speed_of_sound=1540.0)
    plan = O.simus_plan(x, z, d, p, fs.MediumParams())
    assert abs(plan["freqs"].size - 812) < 20


def test_plan_threshold_monotonic():
    """tests/test_simus.py:325-353: -60 dB keeps fewer bins than -100 dB."""
    p = fs.L11_5v()
    x, z = np.linspace(-1e-2, 1e-2, 6), np.linspace(0.5e-2, 3e-2, 6)
    a = O.simus_plan(x, z, np.zeros(128), p, fs.MediumParams(), db=-60)
    b = O.simus_plan(x, z, np.zeros(128), p, fs.MediumParams(), db=-100)
    assert 0 < a["freqs"].size < b["freqs"].size


def test_plan_nfft_identity():
    p = fs.L11_5v()
    x, z, _ = speckle(0, 50)
    plan = O.simus_plan(x, z, np.zeros(128), p, fs.MediumParams())
    assert plan["n_fft"] == 2 * (plan["n_ff"] - 1)  # fs = 4 fc
    np.testing.assert_allclose(plan["freqs"][0] / plan["df"], plan["idx0"], atol=1e-9)


def test_two_way_duration_is_full_width_above_threshold():
    # the reference slice length equals the full width of the centred pulse above 1/1023 (-60 dB)
    for fc, bw in [(2.72e6, 0.74), (7.6e6, 0.77)]:
        dt = 1e-9
        n = 2 ** int(np.ceil(np.log2(1 / dt / 2 / (fc / 32))))
        w = 2 * pi * np.linspace(0, 1 / dt / 2, n)
        y = np.fft.fftshift(np.fft.irfft(O.pulse_spectrum(w, fc) * O.probe_spectrum(w, fc, bw) ** 2))
        y /= np.abs(y).max()
        hit = np.flatnonzero(y > 1 / 1023)
        half = max(n - 1 - hit[0], hit[-1] - (n - 1))
        t = O.two_way_duration(fc, bw, 1.0)
        assert t == pytest.approx(2 * half * dt, abs=2.5e-9)
        assert 1.0 / fc < t < 20.0 / fc


# ----------------------------------------------------------------- geometry (test_pfield_math.py known answers)
def _one_element(width=1e-4, pitch=2e-4):
    return fs.TransducerParams(freq_center=5e6, pitch=pitch, n_elements=1, width=width)


def test_distance_345_and_clip():
    p = _one_element()
    m = fs.MediumParams(speed_of_sound=1540.0)
    g = O.green_terms(np.array([3e-3]), np.array([4e-3]), p, m, 1)
    assert g["r"][0, 0, 0] == pytest.approx(5e-3, rel=1e-12)
    lam2 = 1540.0 / 5e6 / 2
    g = O.green_terms(np.array([lam2 * 0.01]), np.array([lam2 * 0.01]), p, m, 1)
    assert g["r"][0, 0, 0] == pytest.approx(lam2)


def test_obliquity_per_baffle_and_horizon():
    m = fs.MediumParams()
    x, z = np.array([3e-3, 1e-3]), np.array([4e-3, 0.0])
    th = np.arcsin(3 / 5)
    soft = O.green_terms(x, z, _one_element(), m, 1)
    rigid = O.green_terms(x, z, _one_element().model_copy(update={"baffle": "rigid"}), m, 1)
    imp = O.green_terms(x, z, _one_element().model_copy(update={"baffle": 1.75}), m, 1)
    r = 5e-3
    kappa = pi * 5e6 * 1e-4 / 1540.0
    d = np.sin(kappa * 0.6) / (kappa * 0.6)
    base = d / sqrt(r)
    assert soft["amp"][0, 0, 0] == pytest.approx(cos(th) * base, rel=1e-12)
    assert rigid["amp"][0, 0, 0] == pytest.approx(base, rel=1e-12)
    assert imp["amp"][0, 0, 0] == pytest.approx(cos(th) / (cos(th) + 1.75) * base, rel=1e-12)
    # z = 0 beside the element: |theta| = pi/2 -> 1e-16 floor for every baffle
    for g in (soft, rigid, imp):
        assert abs(g["amp"][1, 0, 0]) < 1e-14


def test_green_step_unit_modulus_without_attenuation():
    """tests/test_pfield_helpers.py known answer: |step| = 1 with no attenuation; with attenuation it is below 1
    and falls with distance."""
    r = np.array([1e-4, 3e-3, 0.02, 0.11])
    np.testing.assert_allclose(np.abs(O.green_step(r, 4321.0, 1540.0, 0.0)), 1.0, atol=1e-15)
    a = np.abs(O.green_step(r, 4321.0, 1540.0, 0.5))
    assert np.all(a < 1.0) and np.all(np.diff(a) < 0)


def test_subelement_offsets():
    p = fs.TransducerParams(freq_center=5e6, pitch=1.2e-3, n_elements=1, width=1e-3)
    sx, sz, _, _ = O.sub_elements(p, 3)
    np.testing.assert_allclose(sx[0], [-1e-3 / 3, 0, 1e-3 / 3], atol=1e-15)
    np.testing.assert_allclose(sz[0], 0.0, atol=1e-15)
    sx, _, _, _ = O.sub_elements(p, 2)
    np.testing.assert_allclose(sx[0], [-2.5e-4, 2.5e-4], atol=1e-15)


def test_linear_positions_are_float32():
    ex, _, _, _ = O.element_layout(fs.L11_5v())
    assert np.array_equal(ex, ex.astype(np.float32).astype(np.float64))
    assert ex[0] == pytest.approx(-63.5 * 3e-4, rel=1e-6)


# ----------------------------------------------------------------- spectrum: recurrence vs brute force vs closed form
def _tiny_case(seed, n=5, params=None, medium=None, delays=None):
    params = params or fs.L11_5v()
    medium = medium or fs.MediumParams(attenuation=0.5)
    x, z, rc = speckle(seed, n)
    d = np.zeros(params.n_elements) if delays is None else delays
    plan = O.simus_plan(x, z, d, params, medium)
    return x, z, rc, d, params, medium, plan


@pytest.mark.parametrize("fulldir", [False, True])
def test_recurrence_matches_brute_force(fulldir):
    p = fs.TransducerParams(freq_center=5e6, pitch=3e-4, n_elements=8, width=2.5e-4, bandwidth=0.6)
    pos, _, _ = fs.element_positions(8, p.pitch, p.radius)
    d = fs.plane_wave(pos, np.deg2rad(7), speed_of_sound=1540.0)
    x, z, rc, d, p, m, plan = _tiny_case(3, 6, p, fs.MediumParams(attenuation=0.3), d)
    a = O.spectrum_recurrence(x, z, rc, d, p, m, plan, full_frequency_directivity=fulldir)
    b = O.spectrum_direct(x, z, rc, d, p, m, plan, full_frequency_directivity=fulldir)
    assert rel_l2(a, b) < 1e-9


def test_single_element_closed_form():
    """1 element x 1 scatterer, rigid baffle, no attenuation, n_sub = 1:
    X(f) = probe^2 pulse rc G^2 with G = sinc(k_c W/2 sin(theta)) e^{ikr} / sqrt(r)."""
    p = fs.TransducerParams(freq_center=3e6, pitch=4e-4, n_elements=1, width=3e-4, baffle="rigid")
    m = fs.MediumParams()
    x, z, rc = np.array([2e-3]), np.array([2.5e-2]), np.array([0.7])
    plan = O.simus_plan(x, z, np.zeros(1), p, m, element_splitting=1)
    got = O.spectrum_recurrence(x, z, rc, np.zeros(1), p, m, plan)[:, 0]
    r = sqrt(x[0] ** 2 + z[0] ** 2)
    st = x[0] / r
    a = pi * 3e6 * 3e-4 / 1540.0 * st
    dirv = np.sin(a) / a
    f = plan["freqs"]
    G = dirv * np.exp(1j * 2 * pi * f * r / 1540.0) / sqrt(r)
    want = plan["probe"] ** 2 * plan["pulse"] * rc[0] * G**2
    assert rel_l2(got, want) < 1e-12


def test_invariants_linearity_superposition_zero():
    x, z, rc, d, p, m, plan = _tiny_case(4, 4)
    s_all = O.spectrum_recurrence(x, z, rc, d, p, m, plan)
    s_sum = sum(O.spectrum_recurrence(x[i:i + 1], z[i:i + 1], rc[i:i + 1], d, p, m, plan) for i in range(4))
    assert rel_l2(s_all, s_sum) < 1e-12
    assert rel_l2(O.spectrum_recurrence(x, z, 3 * rc, d, p, m, plan), 3 * s_all) < 1e-13
    assert np.abs(O.spectrum_recurrence(x, z, 0 * rc, d, p, m, plan)).max() == 0.0


def test_mirror_symmetry_swaps_channels():
    x, z, rc, d, p, m, plan = _tiny_case(6, 4)
    a = O.spectrum_recurrence(x, z, rc, d, p, m, plan)
    b = O.spectrum_recurrence(-x, z, rc, d, p, m, plan)
    assert rel_l2(b[:, ::-1], a) < 1e-12


def test_out_of_field_contributes_nothing():
    x, z, rc, d, p, m, plan = _tiny_case(7, 3)
    a = O.spectrum_recurrence(x, z, rc, d, p, m, plan)
    b = O.spectrum_recurrence(np.append(x, 0.001), np.append(z, -0.01), np.append(rc, 5.0), d, p, m, plan)
    assert rel_l2(b, a) < 1e-15


def test_nan_delay_zeroes_element():
    x, z, rc, _, p, m, plan = _tiny_case(8, 3)
    d = np.zeros(p.n_elements)
    d[5] = np.nan
    w = np.ones(p.n_elements)
    w[5] = 0.0
    a = O.spectrum_recurrence(x, z, rc, d, p, m, plan)
    b = O.spectrum_recurrence(x, z, rc, np.zeros(p.n_elements), p, m, plan, apod=w)
    assert rel_l2(a, b) < 1e-15


def test_attenuation_lowers_power():
    """tests/test_simus.py:413-427."""
    x, z, rc, d, p, _, _ = _tiny_case(9, 3)
    m0, m1 = fs.MediumParams(), fs.MediumParams(attenuation=0.5)
    plan = O.simus_plan(x, z, d, p, m0)
    rf0, _, _ = O.simus_oracle(x, z, rc, d, p, m0, plan)
    rf1, _, _ = O.simus_oracle(x, z, rc, d, p, m1, plan)
    assert np.sum(rf1**2) < np.sum(rf0**2)


# ----------------------------------------------------------------- RF tail
def test_rf_matches_idft_definition():
    x, z, rc, d, p, m, plan = _tiny_case(10, 2)
    sel = O.spectrum_recurrence(x, z, rc, d, p, m, plan) * plan["corr"]
    rf, full = O.rf_from_spectrum(sel, plan, p.n_elements)
    N = plan["n_fft"]
    k = np.arange(plan["idx0"], plan["idx0"] + sel.shape[0])
    wk = np.where((k == 0) | (2 * k == N), 1.0, 2.0) * (k <= N // 2)
    n = np.arange((N + 1) // 2)
    ang = 2 * pi * np.outer(n, k) / N
    u = (np.cos(ang) * wk) @ sel.real + (np.sin(ang) * wk) @ sel.imag
    u /= N
    gate = 0.5 * (1 + np.tanh((np.abs(u) / np.abs(u).max() - 1e-5) / 1e-6))
    assert rel_l2(rf, u * gate) < 1e-12
    assert np.abs(full[: plan["idx0"]]).max() == 0.0


def test_zero_rc_gives_zero_rf():
    """tests/test_simus.py:403-411."""
    x, z, rc, d, p, m, plan = _tiny_case(11, 3)
    rf, _, _ = O.simus_oracle(x, z, 0 * rc, d, p, m, plan)
    assert np.abs(rf).max() < 1e-10


def test_echo_arrival_time():
    """A single on-axis scatterer's RF peaks near the round-trip time plus the pulse delay."""
    p = fs.L11_5v()
    m = fs.MediumParams()
    x, z, rc = np.array([0.0]), np.array([0.02]), np.array([1.0])
    plan = O.simus_plan(x, z, np.zeros(128), p, m)
    rf, _, _ = O.simus_oracle(x, z, rc, np.zeros(128), p, m, plan)
    fs_hz = 4 * p.freq_center
    t_peak = np.argmax(np.abs(rf[:, 64])) / fs_hz
    t0 = 2 * 0.02 / 1540.0
    assert t0 - 0.2e-6 < t_peak < t0 + O.two_way_duration(p.freq_center, p.bandwidth, 1.0) + 0.5e-6


# ----------------------------------------------------------------- pfield
def test_pfield_on_axis_focus_peak():
    p = fs.L11_5v()
    pos, _, _ = fs.element_positions(128, p.pitch, p.radius)
    d = fs.focused(pos, np.array([0.0, 0.02]), speed_of_sound=1540.0)
    z = np.linspace(0.01, 0.03, 21)
    x = np.zeros_like(z)
    plan = O.pfield_plan(x, z, d, p, fs.MediumParams())
    pr = O.pfield_oracle(x, z, d, p, fs.MediumParams(), plan)
    assert 0.012 < z[np.argmax(pr)] < 0.028
    assert np.all(pr > 0)


def test_pfield_plan_step_hand_computed():
    """pfield.py:276-282: df = step / (max clipped sub-element distance / c + max delay), then the grid of
    2 ceil(fc / df) + 1 points on [0, 2 fc]. One on-axis point, four elements, two sub-elements: the largest
    distance is to the outermost sub-element, computed here by hand (float32 element centre + W/4)."""
    p = fs.TransducerParams(freq_center=5e6, pitch=3.0e-4, width=2.7e-4, n_elements=4, bandwidth=0.6)
    m = fs.MediumParams(speed_of_sound=1540.0)
    x, z = np.array([0.0]), np.array([0.03])
    d = np.array([0.0, 1e-7, 2e-7, 5e-8])
    xe = float(np.float32(1.5) * np.float32(3.0e-4))
    rmax = np.hypot(xe + 2.7e-4 / 4, 0.03)
    df_max = 1.0 / (rmax / 1540.0 + 2e-7)
    n = 2 * int(np.ceil(5e6 / df_max)) + 1
    want = 2 * 5e6 / (n - 1)
    o = O.pfield_plan(x, z, d, p, m, element_splitting=2)
    assert o["df"] == pytest.approx(want, rel=1e-14)
    assert o["corr"] == pytest.approx(want * 2.7e-4, rel=1e-14)
    prod = fs.pfield_precompute(np.stack([x, z], -1), d, p, m, element_splitting=2)
    assert prod.selected_freqs[1] - prod.selected_freqs[0] == pytest.approx(want, rel=1e-9)
    assert prod.correction_factor == pytest.approx(want * 2.7e-4, rel=1e-14)
    # the frequency_step option scales the largest admissible step
    o2 = O.pfield_plan(x, z, d, p, m, element_splitting=2, frequency_step=0.5)
    assert o2["df"] == pytest.approx(2 * 5e6 / (2 * int(np.ceil(5e6 / (0.5 * df_max)))), rel=1e-14)


def _pp_power(plan):
    return float(np.sum(np.abs(plan["pulse"] * plan["probe"]) ** 2))


def test_pfield_single_element_closed_form():
    """One element, one sub-element, rigid baffle, no attenuation, zero delay: |G A| = D / sqrt(r) at every
    frequency (D the centre-frequency sinc directivity), so the RMS field reduces to
    D / sqrt(r) * sqrt(corr * sum_f |pulse_f probe_f|^2). Geometry computed here by hand."""
    p = fs.TransducerParams(freq_center=5e6, pitch=3.0e-4, width=2.7e-4, n_elements=1, bandwidth=0.6,
                            baffle="rigid")
    m = fs.MediumParams(speed_of_sound=1540.0, attenuation=0.0)
    x, z = np.array([0.002, -0.001]), np.array([0.015, 0.004])
    d = np.zeros(1)
    plan = O.pfield_plan(x, z, d, p, m, element_splitting=1)
    got = O.pfield_oracle(x, z, d, p, m, plan)
    r = np.hypot(x, z)
    kc = 2 * np.pi * p.freq_center / m.speed_of_sound
    D = np.sinc(kc * p.element_width / 2 * (x / r) / np.pi)
    want = np.abs(D) / np.sqrt(r) * np.sqrt(plan["corr"] * _pp_power(plan))
    assert np.allclose(got, want, rtol=1e-12, atol=0)


def test_pfield_two_element_interference_closed_form():
    """Two elements at +-pitch/2 (float32 positions, as the reference lays them out), a point on the axis and
    delays (0, dt): both terms have the same amplitude a, so |G1 A1 + G2 A2|^2 = a^2 (2 + 2 cos(2 pi f dt)) and
    the RMS field is a * sqrt(corr * sum_f |PP_f|^2 (2 + 2 cos(2 pi f dt)))."""
    p = fs.TransducerParams(freq_center=5e6, pitch=3.0e-4, width=2.7e-4, n_elements=2, bandwidth=0.6,
                            baffle="rigid")
    m = fs.MediumParams(speed_of_sound=1540.0, attenuation=0.0)
    x, z = np.array([0.0]), np.array([0.012])
    dt = 50e-9
    d = np.array([0.0, dt])
    plan = O.pfield_plan(x, z, d, p, m, element_splitting=1)
    got = O.pfield_oracle(x, z, d, p, m, plan)
    xe = float(np.float32(0.5) * np.float32(p.pitch))
    r = np.hypot(xe, z[0])
    kc = 2 * np.pi * p.freq_center / m.speed_of_sound
    a = abs(np.sinc(kc * p.element_width / 2 * (xe / r) / np.pi)) / np.sqrt(r)
    f = plan["freqs"]
    pw = np.sum(np.abs(plan["pulse"] * plan["probe"]) ** 2 * (2 + 2 * np.cos(2 * np.pi * f * dt)))
    want = a * np.sqrt(plan["corr"] * pw)
    assert got[0] == pytest.approx(want, rel=1e-12)
