"""FP64 CPU oracle for the SIMUS / PFIELD hot path.

TEST INFRASTRUCTURE ONLY. This module is the checker: it is imported by
``tests/``, by ``__graft_entry__.smoke()`` and by the CPU-baseline leg of
``bench.py``, and by nothing in the product package. It is a NumPy
float64 / complex128 restatement of the reference algorithm, written from the
reference sources read as text; every function cites the lines it follows.

Two independent forms of the spectrum are provided:

* ``spectrum_recurrence`` follows the reference PYTHON strategy step by step
  (per-frequency geometric progression, ``_simus_strategies.py:23-140``);
* ``spectrum_direct`` evaluates every Green's function term with its own
  ``exp`` at every frequency (no recurrence), for tiny brute-force cases.

Pinning (what checks each function, see ``tests/test_oracle.py``):

* ``pulse_spectrum`` / ``probe_spectrum``: closed forms (probe = 1 at fc,
  1/sqrt(2) at the two-way -6 dB edges fc(1 +- B/2); pulse odd in w and purely
  imaginary; |pulse(wc)| = 1 for integer cycle counts).
* ``simus_plan``: the reference's known answers (n_f ~ 812 +- 20 for P4-2v
  focused at 3 cm with scatterers to 8 cm, ``tests/test_simus.py:355-386``;
  fewer bins at -60 dB than at -100 dB, ``test_simus.py:325-353``;
  n_fft = 2 (n_ff - 1) when fs = 4 fc).
* ``pulse_spectrum``: pulse(wc) = +i exactly for integer cycle counts (fixes
  the sign of every RF).
* ``element_layout`` / ``sub_elements`` (convex): adjacent element centres
  exactly one pitch apart, every centre at the radius from the arc centre,
  every sub-element centroid on its element's tangent at the offset o_j,
  oriented along the direction of increasing element angle.
* ``geometry`` / ``green_terms`` / ``green_step``: 3-4-5 distance,
  lambda/2 clip, on-axis angle, obliquity per baffle, horizon floor,
  |step| = 1 without attenuation (``tests/test_pfield_math.py:101-234``).
* ``pfield_plan``: the frequency step of a hand-computed case (one on-axis
  point, the outermost sub-element, a known largest delay).
* ``spectrum_recurrence`` against ``spectrum_direct`` on brute-force inputs,
  and against the single-element single-scatterer closed form.
* ``rf_from_spectrum``: the inverse-DFT definition, the zero-input case and
  the echo-arrival-time invariant.
* ``pfield_oracle`` / ``pfield_plan``: two special cases that reduce to closed
  forms with the geometry computed by hand in the test (one element, rigid
  baffle, no attenuation: D / sqrt(r) * sqrt(corr sum |PP|^2); two elements and
  a delay dt on the axis: the 2 + 2 cos(2 pi f dt) interference), and the
  on-axis focal peak.
* the adjoint oracle (``simus_loss_grad_fp64``): central finite differences
  of the oracle loss itself.
"""

from __future__ import annotations

from math import ceil, inf, log, log2, pi

import numpy as np

NEPER_TO_DB = 20.0 / log(10.0)  # _pfield_math.py:21


# --------------------------------------------------------------------------- spectra
def pulse_spectrum(w, fc, n_w=1.0):
    """``spectrum.py:24-50``: i (sinc(T(w-wc)/2pi) - sinc(T(w+wc)/2pi)), T = n_w/fc."""
    w = np.asarray(w, dtype=np.float64)
    t = n_w / fc
    wc = 2 * pi * fc
    return 1j * (np.sinc(t * (w - wc) / (2 * pi)) - np.sinc(t * (w + wc) / (2 * pi)))


def probe_spectrum(w, fc, bw=0.75):
    """``spectrum.py:53-107``: sqrt(exp(-(|w-wc|/sigma)^p))."""
    if not (0.0 < bw < 2.0):
        raise ValueError("bandwidth must be in (0, 2.0)")
    w = np.asarray(w, dtype=np.float64)
    wc = 2 * pi * fc
    wb = bw * wc
    p = log(126) / log(2 * wc / wb)
    sigma = wb / 2 / log(2) ** (1 / p)
    return np.exp(-0.5 * (np.abs(w - wc) / sigma) ** p)


# --------------------------------------------------------------------------- geometry
def element_layout(params):
    """``utils/geometry.py:18-88`` (linear x in float32, convex arc)."""
    n, pitch, radius = params.n_elements, params.pitch, params.radius
    if radius == inf:
        xs = (np.arange(n, dtype=np.float32) - np.float32((n - 1) / 2)) * np.float32(pitch)
        return xs.astype(np.float64), np.zeros(n), np.zeros(n), 0.0
    half = np.arcsin(pitch / 2 / radius)
    chord = 2 * radius * np.sin(half * (n - 1))
    apex = float(np.sqrt(radius**2 - (chord / 2) ** 2))
    th = np.linspace(np.arctan2(-chord / 2, apex), np.arctan2(chord / 2, apex), n)
    return radius * np.sin(th), radius * np.cos(th) - apex, th, apex


def sub_elements(params, n_sub):
    """Sub-element centroids (E, n_sub): ``_pfield_math.py:41-67``."""
    ex, ez, th, apex = element_layout(params)
    w = params.element_width
    seg = w / n_sub
    o = -w / 2 + seg / 2 + seg * np.arange(n_sub)
    return ex[:, None] + o[None, :] * np.cos(th)[:, None], ez[:, None] - o[None, :] * np.sin(th)[:, None], th, apex


def green_terms(x, z, params, medium, n_sub, full_frequency_directivity=False):
    """Per (scatterer, element, sub-element) geometry of the Green's function.

    ``_pfield_math.py:71-118`` (distances, lambda/2 clip, asin angle),
    ``166-199`` (obliquity), ``simus.py:310-313`` (centre-frequency sinc).
    Returns dict with clipped distance ``r``, ``sin_theta`` and the real
    amplitude ``amp = obliquity * directivity / sqrt(r) / n_sub``.
    """
    c, fc = medium.speed_of_sound, params.freq_center
    sx, sz, th, apex = sub_elements(params, n_sub)
    dx = x[:, None, None] - sx[None]
    dz = z[:, None, None] - sz[None]
    r = np.sqrt(dx * dx + dz * dz)
    rc = np.where(r < c / fc / 2, c / fc / 2, r)
    theta = np.arcsin((dx + 1e-16) / (r + 1e-16)) - th[None, :, None]
    b = params.baffle
    bname = getattr(b, "value", b)
    if bname == "rigid":
        obl = np.ones_like(theta)
    elif bname == "soft":
        obl = np.cos(theta)
    else:
        obl = np.cos(theta) / (np.cos(theta) + float(b))
    obl = np.where(np.abs(theta) >= pi / 2, 1e-16, obl)
    sin_t = np.sin(theta)
    seg = params.element_width / n_sub
    amp = obl / np.sqrt(rc) / n_sub
    if not full_frequency_directivity:
        amp = amp * np.sinc(2 * pi * fc / c * seg / 2 * sin_t / pi)
    is_out = z < 0
    if params.radius != inf:
        is_out = is_out | (x**2 + (z + apex) ** 2 <= params.radius**2)
    return {"r": rc, "sin_theta": sin_t, "amp": amp, "is_out": is_out, "seg": seg}


# --------------------------------------------------------------------------- plan
def two_way_duration(fc, bw, n_w):
    """``simus.py:49-104`` restated: length of the reference trim slice, in seconds."""
    dt = 1e-9
    n = 2 ** ceil(log2(1 / dt / 2 / (fc / n_w / 32)))
    w = 2 * pi * np.linspace(0, 1 / dt / 2, n)
    y = np.fft.fftshift(np.fft.irfft(pulse_spectrum(w, fc, n_w) * probe_spectrum(w, fc, bw) ** 2))
    y = y / np.abs(y).max()
    hit = np.flatnonzero(y > 1 / 1023)
    if hit.size == 0:
        return n_w / fc
    k = min(hit[0] + 1, 2 * n - 2 - hit[-1])
    m = y.shape[0]
    start = m - k
    stop = k - 2 if k - 2 >= 0 else m + (k - 2)
    return max(0, start - stop) * dt


def band(fc, bw, n_w, db, df_max):
    """``_pfield_math.py:121-162`` + ``250-261``."""
    n = int(2 * ceil(fc / df_max) + 1)
    f = np.linspace(0, 2 * fc, n)
    mag = np.abs(pulse_spectrum(2 * pi * f, fc, n_w) * probe_spectrum(2 * pi * f, fc, bw))
    keep = 20 * np.log10(1e-200 + mag / mag.max()) > db
    i0, i1 = (np.flatnonzero(keep)[[0, -1]] if keep.any() else (0, 0))
    return f[i0:i1 + 1], float(f[1])


def simus_plan(x, z, delays, params, medium, fs=None, n_w=1.0, db=-60.0, element_splitting=None, frequency_step=1.0):
    """``simus.py:165-260``. Returns a dict of plan quantities."""
    c, fc, bw = medium.speed_of_sound, params.freq_center, params.bandwidth
    fs = 4 * fc if fs is None else fs
    n_sub = element_splitting if element_splitting is not None else ceil(params.element_width / (c / (fc * (1 + bw / 2))))
    ex, ez, _, _ = element_layout(params)
    maxd = float(np.sqrt((x[:, None] - ex) ** 2 + (z[:, None] - ez) ** 2).max())
    if n_w != inf:
        maxd += two_way_duration(fc, bw, n_w) * c
    d = np.nan_to_num(np.asarray(delays, dtype=np.float64), nan=0.0)
    df = frequency_step * 0.5 / (2 * maxd / c + d.max())
    n_ff = int(2 * ceil(fc / df) + 1)
    f, step = band(fc, bw, n_w, db, df)
    return {
        "freqs": f, "df": step, "n_sub": n_sub, "seg": params.element_width / n_sub,
        "pulse": pulse_spectrum(2 * pi * f, fc, n_w), "probe": probe_spectrum(2 * pi * f, fc, bw),
        "n_ff": n_ff, "idx0": int(round(f[0] / step)),
        "corr": (1.0 if n_w == inf else step) * params.element_width,
        "n_fft": ceil(fs / 2 / fc * (n_ff - 1)),
    }


# --------------------------------------------------------------------------- spectrum
def green_step(r, fstep, c, alpha):
    """Per-frequency-step multiplier of every Green's term, ``_pfield_math.py:203-247``:
    exp((-alpha(df) + i k(df)) r) with alpha(f) = alpha_dB / NEPER_TO_DB * f / 1e6 * 1e2 (Np/m)."""
    return np.exp((-(alpha / NEPER_TO_DB * fstep / 1e6 * 1e2) + 1j * 2 * pi * fstep / c) * np.asarray(r))


def _tx_weights(delays, apod):
    d = np.asarray(delays, dtype=np.float64)
    a = np.ones_like(d) if apod is None else np.asarray(apod, dtype=np.float64) * np.ones_like(d)
    nan = np.isnan(d)
    return np.where(nan, 0.0, d), np.where(nan, 0.0, a)


def spectrum_recurrence(x, z, rc, delays, params, medium, plan, apod=None, full_frequency_directivity=False):
    """Reference PYTHON strategy (``simus.py:263-333`` + ``_simus_strategies.py:23-140``).

    One transmit; returns the selected spectrum (n_f, E) *before* the
    correction factor, complex128.
    """
    c, alpha = medium.speed_of_sound, medium.attenuation
    g = green_terms(x, z, params, medium, plan["n_sub"], full_frequency_directivity)
    f = plan["freqs"]
    f0 = f[0]
    fstep = f[1] - f[0] if f.size > 1 else 0.0
    r = g["r"]
    k0r = 2 * pi * f0 / c * r
    a0 = alpha / NEPER_TO_DB * f0 / 1e6 * 1e2
    phase = np.exp(-a0 * r + 1j * (k0r - 2 * pi * np.floor(k0r / (2 * pi))))
    step = green_step(r, fstep, c, alpha)
    phase = phase * g["amp"] * plan["n_sub"]  # amp carries 1/n_sub; the reference divides in mean()
    d, w = _tx_weights(delays, apod)
    da = np.exp(1j * 2 * pi * f0 * d) * w
    dstep = np.exp(1j * 2 * pi * fstep * d)
    pp = plan["pulse"] * plan["probe"]
    out = np.zeros((f.size, params.n_elements), dtype=np.complex128)
    seg = g["seg"]
    for k in range(f.size):
        if full_frequency_directivity:
            dk = np.sinc(2 * pi * f[k] / c * seg / 2 * g["sin_theta"] / pi)
            mono = np.mean(phase * dk, axis=-1)
        else:
            mono = np.mean(phase, axis=-1)
        p = pp[k] * (mono @ da)
        p = np.where(g["is_out"], 0.0, p)
        out[k] = plan["probe"][k] * ((rc * p) @ mono)
        phase = phase * step
        da = da * dstep
    return out


def green_direct(x, z, params, medium, plan, k, full_frequency_directivity=False):
    """G(f_k, e, s), shape (S, E), with a fresh exp per term (no recurrence)."""
    c, alpha = medium.speed_of_sound, medium.attenuation
    g = green_terms(x, z, params, medium, plan["n_sub"], full_frequency_directivity)
    fk = plan["freqs"][0] + k * (plan["freqs"][1] - plan["freqs"][0]) if plan["freqs"].size > 1 else plan["freqs"][0]
    r = g["r"]
    amp = g["amp"]
    if full_frequency_directivity:
        amp = amp * np.sinc(2 * pi * fk / c * g["seg"] / 2 * g["sin_theta"] / pi)
    t = amp * np.exp(-alpha / NEPER_TO_DB * fk * 1e-4 * r) * np.exp(1j * 2 * pi * np.mod(fk * r / c, 1.0))
    return t.sum(axis=-1), g["is_out"], fk


def spectrum_direct(x, z, rc, delays, params, medium, plan, apod=None, full_frequency_directivity=False):
    """Brute-force double sum per frequency (independent of the recurrence)."""
    d, w = _tx_weights(delays, apod)
    nf = plan["freqs"].size
    out = np.zeros((nf, params.n_elements), dtype=np.complex128)
    for k in range(nf):
        G, out_mask, fk = green_direct(x, z, params, medium, plan, k, full_frequency_directivity)
        a = w * np.exp(1j * 2 * pi * np.mod(fk * d, 1.0))
        p = plan["pulse"][k] * plan["probe"][k] * (G @ a)
        p = np.where(out_mask, 0.0, p)
        out[k] = plan["probe"][k] * ((rc * p) @ G)
    return out


# --------------------------------------------------------------------------- RF
def rf_from_spectrum(sel, plan, n_elements):
    """``simus.py:336-367``: place band, irfft(conj), keep first half, smooth -100 dB gate."""
    full = np.zeros((plan["n_ff"], n_elements), dtype=np.complex128)
    full[plan["idx0"]:plan["idx0"] + sel.shape[0]] = sel
    rf = np.fft.irfft(np.conj(full), n=plan["n_fft"], axis=0)[: (plan["n_fft"] + 1) // 2]
    rel = np.abs(rf) / (np.abs(rf).max() + 1e-30)
    return rf * 0.5 * (1 + np.tanh((rel - 1e-5) / 1e-6)), full


def simus_oracle(x, z, rc, delays, params, medium, plan, apod=None, full_frequency_directivity=False, direct=False):
    """Full oracle: (rf, full_spectrum) for one transmit (``simus.py:392-495``)."""
    fn = spectrum_direct if direct else spectrum_recurrence
    sel = fn(x, z, rc, delays, params, medium, plan, apod, full_frequency_directivity) * plan["corr"]
    rf, full = rf_from_spectrum(sel, plan, params.n_elements)
    return rf, full, sel


# --------------------------------------------------------------------------- pfield
def pfield_plan(x, z, delays, params, medium, n_w=1.0, db=-60.0, element_splitting=None, frequency_step=1.0):
    """``pfield.py:223-298``."""
    c, fc, bw = medium.speed_of_sound, params.freq_center, params.bandwidth
    n_sub = element_splitting if element_splitting is not None else ceil(params.element_width / (c / (fc * (1 + bw / 2))))
    sx, sz, _, _ = sub_elements(params, n_sub)
    r = np.sqrt((x[:, None, None] - sx) ** 2 + (z[:, None, None] - sz) ** 2)
    rmax = max(float(r.max()), c / fc / 2)
    d = np.nan_to_num(np.asarray(delays, dtype=np.float64), nan=0.0)
    df = frequency_step / (rmax / c + d.max())
    f, step = band(fc, bw, n_w, db, df)
    return {"freqs": f, "df": step, "n_sub": n_sub, "seg": params.element_width / n_sub,
            "pulse": pulse_spectrum(2 * pi * f, fc, n_w), "probe": probe_spectrum(2 * pi * f, fc, bw),
            "corr": (1.0 if n_w == inf else step) * params.element_width}


def pfield_oracle(x, z, delays, params, medium, plan, apod=None, full_frequency_directivity=False):
    """``pfield.py:102-184, 301-378`` + ``_pfield_strategies.py:53-88``: sqrt(corr sum_f |PP sum G A|^2)."""
    d, w = _tx_weights(delays, apod)
    acc = np.zeros(x.shape[0])
    for k in range(plan["freqs"].size):
        G, out_mask, fk = green_direct(x, z, params, medium, plan, k, full_frequency_directivity)
        a = w * np.exp(1j * 2 * pi * np.mod(fk * d, 1.0))
        p = plan["pulse"][k] * plan["probe"][k] * (G @ a)
        acc += np.where(out_mask, 0.0, np.abs(p) ** 2)
    return np.sqrt(acc * plan["corr"])
