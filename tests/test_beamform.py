"""IQ demodulation and delay-and-sum over the RF stack (``fastsimus_amd.beamform``).

CPU: the product's Butterworth design against SciPy's ``butter`` / ``sosfilt_zi``; the f-number of a point
element; the checker ``oracle.beamform_oracle`` pinned by closed forms (a tone at fc demodulates to 1, a
synthetic point echo sums coherently to elements-in-aperture times its amplitude at the target pixel).
GPU (marked ``gpu``): ``fs_rf2iq`` against SciPy's ``filtfilt`` recipe on SIMUS RF, the tone and
linearity; ``fs_das`` against the checker, the synthetic point echo, a SIMUS point target imaged at its
position, and compounding = the sum of the per-firing images. Agreement with PyMUST (the reference
harness's ``rf2iq`` / ``dasmtx``) is unpinned: PyMUST is not installed.
"""

from __future__ import annotations

import numpy as np
import pytest

import fastsimus_amd as fs
from fastsimus_amd import beamform as BF
from oracle import beamform_oracle as BO
from tests.conftest import rel_l2, speckle

signal = pytest.importorskip("scipy.signal")


# ----------------------------------------------------------------------------- CPU
@pytest.mark.parametrize("order,wn", [(5, 0.5), (5, 0.3), (4, 0.2), (5, 0.05), (1, 0.4), (2, 0.7)])
def test_butterworth_design_matches_scipy(order, wn):
    sos = BF.butterworth_sos(order, wn)
    b, a = signal.sos2tf(sos)
    bs, as_ = signal.butter(order, wn)
    n = bs.size
    np.testing.assert_allclose(b[:n], bs, rtol=1e-12, atol=1e-15 * np.abs(bs).max())
    np.testing.assert_allclose(a[:n], as_, rtol=1e-12, atol=1e-14)
    assert np.all(b[n:] == 0) and np.all(a[n:] == 0)
    # unit-DC sections, so the per-section steady states equal SciPy's cascade-scaled ones
    np.testing.assert_allclose(BF.section_steady_states(sos), signal.sosfilt_zi(sos), atol=1e-13)


def test_rf2iq_filter_parameters():
    f = BF.rf2iq_filter(4 * 7.6e6, 7.6e6)
    assert f.n_sections == 3 and f.padlen == 18  # 5th order: two biquads and a first-order section
    sos = np.array([[f.sos[i][j] for j in range(6)] for i in range(3)])
    np.testing.assert_allclose(signal.sos2tf(sos)[0][:6], signal.butter(5, 0.5)[0], rtol=1e-12)
    f2 = BF.rf2iq_filter(50e6, 5e6)  # 2 fc / fs = 0.2
    sos2 = np.array([[f2.sos[i][j] for j in range(6)] for i in range(3)])
    np.testing.assert_allclose(signal.sos2tf(sos2)[1][:6], signal.butter(5, 0.2)[1], rtol=1e-12)


def test_f_number_closed_form():
    """A point-like element (directivity cos(alpha)) reaches 1/sqrt(2) at 45 degrees: F = 1 / (2 tan 45) = 0.5;
    a wider element has a narrower cone and a larger F."""
    tiny = fs.TransducerParams(freq_center=5e6, pitch=3e-4, n_elements=8, width=1e-9)
    assert BF.f_number_for(tiny) == pytest.approx(0.5, rel=1e-6)
    wide = fs.TransducerParams(freq_center=5e6, pitch=6e-4, n_elements=8, width=5e-4)
    assert BF.f_number_for(wide) > BF.f_number_for(fs.TransducerParams(freq_center=5e6, pitch=6e-4, n_elements=8,
                                                                       width=2e-4)) > 0.5


def test_oracle_rf2iq_tone_and_linearity():
    fs_hz, fc = 4 * 7.6e6, 7.6e6
    n = np.arange(1000)
    tone = np.cos(2 * np.pi * fc * n / fs_hz)[:, None] * np.ones((1, 3))
    iq = BO.rf2iq(tone, fs_hz, fc)
    np.testing.assert_allclose(iq[100:-100], 1.0, atol=1e-9)
    rng = np.random.default_rng(0)
    a, b = rng.standard_normal((200, 2)), rng.standard_normal((200, 2))
    np.testing.assert_allclose(BO.rf2iq(2 * a - 3 * b, fs_hz, fc), 2 * BO.rf2iq(a, fs_hz, fc) - 3 * BO.rf2iq(b, fs_hz, fc),
                               atol=1e-12)


def _point_echo(p, target, delays, fs_hz, fc, n=2400, amp=0.7, sigma=20):
    """IQ of a point echo at ``target``: element e sees amp g((n - tau_e fs) / sigma) exp(-i 2 pi fc tau_e), g a
    Gaussian, tau_e the DAS delay of the target pixel, so the coherent sum there is amp x elements in aperture."""
    pos = fs.element_positions(p.n_elements, p.pitch, p.radius)[0]
    r = np.hypot(target[0] - pos[:, 0], target[1] - pos[:, 1])
    iq = []
    for d in np.atleast_2d(delays):
        tau = np.min(d + r / 1540.0) + r / 1540.0
        t = np.arange(n)[:, None]
        iq.append(amp * np.exp(-0.5 * ((t - tau[None] * fs_hz) / sigma) ** 2) * np.exp(-1j * 2 * np.pi * fc * tau)[None])
    return np.stack(iq), pos


def test_oracle_das_point_echo_closed_form():
    p = fs.L11_5v()
    fs_hz, fc = 4 * p.freq_center, p.freq_center
    target = (0.003, 0.02)
    D = np.stack([fs.plane_wave(fs.element_positions(128, p.pitch, p.radius)[0], np.deg2rad(a), speed_of_sound=1540.0)
                  for a in (-4.0, 6.0)])
    iq, pos = _point_echo(p, target, D, fs_hz, fc)
    F = 1.2
    inside = np.sum(np.abs(target[0] - pos[:, 0]) <= target[1] / (2 * F))
    img = BO.das(iq, D, pos[:, 0], pos[:, 1], [target[0]], [target[1]], fs_hz, fc, 1540.0, F)
    assert img[0] == pytest.approx(2 * 0.7 * inside, rel=1e-3)  # two firings, coherent, interpolation error < 1e-3


# ----------------------------------------------------------------------------- GPU
def _simus_rf(cuda, n=100, seed=501, angles=(-6.0, 0.0, 8.0), scat=None):
    p = fs.L11_5v()
    m = fs.MediumParams(attenuation=0.5)
    if scat is None:
        x, z, rc = speckle(seed, n, zlim=(0.005, 0.03))
    else:
        x, z, rc = scat
    pos = fs.element_positions(128, p.pitch, p.radius)[0]
    D = np.stack([fs.plane_wave(pos, np.deg2rad(a), speed_of_sound=1540.0) for a in angles])
    res = fs.simus(np.stack([x, z], -1), rc, D, p, m, strategy="hip")
    return p, D, res.rf


@pytest.mark.gpu
def test_rf2iq_matches_scipy_filtfilt(cuda):
    p, D, rf = _simus_rf(cuda)
    fs_hz, fc = 4 * p.freq_center, p.freq_center
    iq = BF.rf2iq(rf, fs_hz, fc).cpu().numpy()
    rf_h = rf.cpu().numpy()
    for t in range(rf_h.shape[0]):
        want = BO.rf2iq(rf_h[t], fs_hz, fc)
        e = rel_l2(iq[t], want)
        print(f"rf2iq firing {t}: rel-L2 vs SciPy filtfilt {e:.2e}", flush=True)
        assert e <= 1e-5, (t, e)


@pytest.mark.gpu
def test_rf2iq_tone_and_linearity(cuda):
    import torch

    # fs = 4 fc (the SIMUS sampling): the mixed tone's image at 2 fc sits at Nyquist, where the bilinear
    # Butterworth has its zeros, so the IQ is exactly the constant 1 away from the ends
    fs_hz, fc = 20e6, 5e6
    n = np.arange(1500)
    tone = np.cos(2 * np.pi * fc * n / fs_hz)[:, None] * np.ones((1, 5))
    iq = BF.rf2iq(torch.as_tensor(tone, dtype=torch.float32, device=cuda), fs_hz, fc).cpu().numpy()
    print("tone interior max |iq - 1|", np.abs(iq[200:-200] - 1).max(), flush=True)
    np.testing.assert_allclose(iq[200:-200], 1.0, atol=2e-5)
    # fs = 8 fc: the image at half-Nyquist is only attenuated (by 82 per pass); GPU against SciPy's filtfilt
    fs8 = 40e6
    tone8 = np.cos(2 * np.pi * fc * n / fs8)[:, None] * np.ones((1, 5))
    iq8 = BF.rf2iq(torch.as_tensor(tone8, dtype=torch.float32, device=cuda), fs8, fc).cpu().numpy()
    e8 = rel_l2(iq8, BO.rf2iq(tone8.astype(np.float32), fs8, fc))
    print(f"tone at fs = 8 fc vs SciPy filtfilt rel-L2 {e8:.2e}", flush=True)
    assert e8 <= 1e-5
    rng = np.random.default_rng(1)
    a = torch.as_tensor(rng.standard_normal((2, 600, 7)), dtype=torch.float32, device=cuda)
    b = torch.as_tensor(rng.standard_normal((2, 600, 7)), dtype=torch.float32, device=cuda)
    lhs = BF.rf2iq(2 * a - 3 * b, fs_hz, fc)
    rhs = 2 * BF.rf2iq(a, fs_hz, fc) - 3 * BF.rf2iq(b, fs_hz, fc)
    e = float(torch.linalg.norm(lhs - rhs) / torch.linalg.norm(rhs))
    print(f"rf2iq linearity rel-L2 {e:.2e}", flush=True)
    assert e <= 1e-6


@pytest.mark.gpu
def test_das_matches_checker_and_compounds(cuda):
    p, D, rf = _simus_rf(cuda)
    fs_hz, fc = 4 * p.freq_center, p.freq_center
    iq = BF.rf2iq(rf, fs_hz, fc)
    gx, gz = np.meshgrid(np.linspace(-0.012, 0.012, 41), np.linspace(0.006, 0.028, 33))
    img = BF.das(iq, D, p, gx, gz, fs_hz, fc, f_number=1.5).cpu().numpy()
    pos = fs.element_positions(128, p.pitch, p.radius)[0]
    want = BO.das(iq.cpu().numpy(), D, pos[:, 0], pos[:, 1], gx.astype(np.float32), gz.astype(np.float32), fs_hz, fc,
                  1540.0, 1.5).reshape(gx.shape)
    e = rel_l2(img, want)
    print(f"das vs checker rel-L2 {e:.2e}", flush=True)
    assert img.shape == gx.shape and e <= 1e-5
    per = BF.das(iq, D, p, gx, gz, fs_hz, fc, f_number=1.5, compound=False).cpu().numpy()
    assert per.shape == (3,) + gx.shape
    e2 = rel_l2(img, per.sum(0))
    print(f"compound vs sum of per-firing images rel-L2 {e2:.2e}", flush=True)
    assert e2 <= 1e-6


@pytest.mark.gpu
@pytest.mark.parametrize("n", [19, 20, 37, 100, 1111, 4099])
@pytest.mark.parametrize("wn", [0.5, 0.15])
def test_rf2iq_lengths_and_cutoffs_match_scipy(cuda, n, wn):
    """Sequence lengths from padlen + 1 (most of the 32 time chunks of a pass empty) to long ones that
    are not multiples of the chunk count, and a narrower low-pass (poles nearer z = 1, so the chunk
    transitions carry more of the state), against SciPy's filtfilt recipe."""
    import torch

    fs_hz, fc = 4e7, 5e6
    rng = np.random.default_rng(n)
    rf = rng.standard_normal((2, n, 3))
    bw = None if wn == 0.5 else wn * fs_hz
    iq = BF.rf2iq(torch.as_tensor(rf, dtype=torch.float32, device=cuda), fs_hz, fc, bandwidth=bw).cpu().numpy()
    for t in range(2):
        want = BO.rf2iq(rf[t].astype(np.float32), fs_hz, fc, bandwidth=bw)
        e = rel_l2(iq[t], want)
        print(f"n={n} wn={wn} firing {t}: rel-L2 {e:.2e}", flush=True)
        assert e <= 1e-5, (n, wn, t, e)


@pytest.mark.gpu
def test_channel_major_layout_matches_sample_major(cuda):
    """rf2iq's channel-major output is its sample-major output transposed, bit for bit (same arithmetic per
    sequence), and das reads either layout to the same image (same terms; the two layouts only change
    where the interpolation nodes are read from). Includes firings whose delays have NaN entries."""
    p, D, rf = _simus_rf(cuda)
    fs_hz, fc = 4 * p.freq_center, p.freq_center
    iq = BF.rf2iq(rf, fs_hz, fc)
    iqc = BF.rf2iq(rf, fs_hz, fc, channel_major=True)
    assert iqc.shape == (iq.shape[0], iq.shape[2], iq.shape[1])
    assert bool((iqc.permute(0, 2, 1) == iq).all())
    D = D.copy()
    D[1, 60:70] = np.nan
    gx, gz = np.meshgrid(np.linspace(-0.012, 0.012, 37), np.linspace(0.002, 0.03, 29))
    for compound in (True, False):
        a = BF.das(iq, D, p, gx, gz, fs_hz, fc, f_number=1.5, compound=compound).cpu().numpy()
        b = BF.das(iqc, D, p, gx, gz, fs_hz, fc, f_number=1.5, compound=compound, channel_major=True).cpu().numpy()
        assert bool(np.array_equal(a, b)), compound


@pytest.mark.gpu
@pytest.mark.parametrize("f_number,t0", [(0.0, 0.0), (1.0, 2.5e-6), (2.0, -1e-6)])
def test_das_random_iq_options_match_checker(cuda, f_number, t0):
    """Random IQ, diverging-wave and focused firings (some elements silent: NaN delays), a grid that
    reaches past the end of the record and above the array, the full aperture (F = 0) and record start
    times t0 other than 0, in both IQ layouts, against the FP64 checker."""
    import torch

    p = fs.L11_5v()
    fs_hz, fc = 4 * p.freq_center, p.freq_center
    pos = fs.element_positions(128, p.pitch, p.radius)[0]
    D = np.stack([fs.focused(pos, np.array([0.004, -0.01]), speed_of_sound=1540.0),
                  fs.focused(pos, np.array([-0.003, 0.015]), speed_of_sound=1540.0),
                  fs.plane_wave(pos, np.deg2rad(5.0), speed_of_sound=1540.0)])
    D[2, :20] = np.nan
    rng = np.random.default_rng(7)
    iq_h = rng.standard_normal((3, 700, 128)) + 1j * rng.standard_normal((3, 700, 128))
    iq = torch.as_tensor(iq_h.astype(np.complex64), device=cuda)
    gx, gz = np.meshgrid(np.linspace(-0.015, 0.015, 23), np.linspace(-0.002, 0.03, 19))
    want = BO.das(iq_h.astype(np.complex64), D, pos[:, 0], pos[:, 1], gx.astype(np.float32), gz.astype(np.float32),
                  fs_hz, fc, 1540.0, f_number, t0=t0).reshape(gx.shape)
    img = BF.das(iq, D, p, gx, gz, fs_hz, fc, f_number=f_number, t0=t0).cpu().numpy()
    e = rel_l2(img, want)
    print(f"F={f_number} t0={t0}: das vs checker rel-L2 {e:.2e}", flush=True)
    assert e <= 1e-5
    img_c = BF.das(iq.permute(0, 2, 1).contiguous(), D, p, gx, gz, fs_hz, fc, f_number=f_number, t0=t0,
                   channel_major=True).cpu().numpy()
    assert bool(np.array_equal(img, img_c))


@pytest.mark.gpu
def test_das_point_echo_closed_form(cuda):
    p = fs.L11_5v()
    fs_hz, fc = 4 * p.freq_center, p.freq_center
    target = (0.003, 0.02)
    D = np.stack([fs.plane_wave(fs.element_positions(128, p.pitch, p.radius)[0], np.deg2rad(a), speed_of_sound=1540.0)
                  for a in (-4.0, 6.0)])
    iq, pos = _point_echo(p, target, D, fs_hz, fc)
    F = 1.2
    inside = np.sum(np.abs(target[0] - pos[:, 0]) <= target[1] / (2 * F))
    img = BF.das(iq, D, p, np.array([target[0]]), np.array([target[1]]), fs_hz, fc, f_number=F).cpu().numpy()
    print("point echo:", img[0], "expected", 2 * 0.7 * inside, flush=True)
    assert img[0] == pytest.approx(2 * 0.7 * inside, rel=1e-3)


@pytest.mark.gpu
def test_simus_point_target_images_at_its_position(cuda):
    """One scatterer simulated under three plane waves, demodulated and beamformed (the harness's pipeline,
    ``beamform``): the envelope peaks at the scatterer's position within one pixel (lambda / 4)."""
    x0, z0 = 0.0023, 0.0187
    scat = (np.array([x0]), np.array([z0]), np.array([1.0]))
    p, D, rf = _simus_rf(cuda, scat=scat)
    fs_hz = 4 * p.freq_center
    lam = 1540.0 / p.freq_center
    step = lam / 4
    xs = x0 + step * np.arange(-16, 17)
    zs = z0 + step * np.arange(-16, 17)
    gx, gz = np.meshgrid(xs, zs)
    img = np.abs(BF.beamform(rf, D, p, gx, gz, fs_hz).cpu().numpy())
    iz, ix = np.unravel_index(np.argmax(img), img.shape)
    print(f"peak at dx {xs[ix] - x0:+.2e} m, dz {zs[iz] - z0:+.2e} m (pixel {step:.1e} m)", flush=True)
    assert abs(xs[ix] - x0) <= step * 1.01 and abs(zs[iz] - z0) <= step * 1.01
    # the rf_stack layout gives the same image
    img2 = np.abs(BF.beamform(fs.rf_stack(rf), D, p, gx, gz, fs_hz).cpu().numpy())
    assert rel_l2(img2, img) <= 1e-6
