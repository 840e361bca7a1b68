"""Benchmark of the batched SIMUS hot path on one MI355X.

Workload (default): 128-element 7.6 MHz linear array (pitch 0.30 mm, width
0.27 mm, 77 % bandwidth), c = 1540 m/s, 0.5 dB/cm/MHz, fs = 4 fc; 200 000
speckle scatterers (x ~ U[-2, 2] cm, z ~ U[0.5, 5] cm, rc ~ N(0, 1), float32,
seed 2026); 75 plane waves over [-16, 16] deg sharing one frequency plan.

One step = the whole compute path for all 75 transmits with inputs resident
in HBM: the split-BF16 MFMA spectrum (multistatic matrix + transmit apply),
the placed full spectrum (``SimusResult.spectrum``) and the spectrum -> RF
transform with the -100 dB gate. The plan is built once, outside the timed
region (as the reference's own benchmark does). By default the tile shapes are
autotuned first and the step is captured in a HIP graph whose replays are
timed; the dominant kernel's duration comes from HIP events around its eager
launches in a second timed pass of the same step.

Printed JSON (one line): ``value`` = whole-job evaluations/s in
G(transmit . scatterer . element . frequency)/s for the MFMA path; beside it
the plain FP32 kernel's rate on the same inputs (a subset of the transmits:
the plain method's cost is exactly linear in the transmit count), the RF
relative L2 errors (MFMA vs FP32 on the full workload, both vs the FP64 oracle
on a scatterer subsample), the roofline of the dominant kernel and the CPU
baseline (the FP64 NumPy oracle on a bounded sample, host cores, rank 0).
"""

from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "G(scatterer·element·freq)/s, batched MFMA vs plain FP32 at ≤1e-3 rel RF err"
UNIT = "G(scatterer·element·freq)/s"
PEAK_BF16 = 2.5e15    # MI355X dense BF16 MFMA, FLOP/s
PEAK_FP32 = 157.3e12  # MI355X FP32 vector, FLOP/s
# VALU instructions the plain kernels (k_tx + k_rx) issue per evaluation at n_sub = 2 on the
# benchmarked sizes (SQ_INSTS_VALU x 64 / evaluations: k_tx 15.1, k_rx 16.8 with the geometry table)
VALU_PER_EVAL_NSUB2 = 31.9


def workload(seed: int = 2026, n_scat: int = 200_000, n_tx: int = 75):
    import fastsimus_amd as fs

    params = fs.TransducerParams(freq_center=7.6e6, pitch=3.0e-4, width=2.7e-4, n_elements=128, bandwidth=0.77)
    medium = fs.MediumParams(speed_of_sound=1540.0, attenuation=0.5)
    rng = np.random.default_rng(seed)
    x = rng.uniform(-0.02, 0.02, n_scat).astype(np.float32)
    z = rng.uniform(0.005, 0.05, n_scat).astype(np.float32)
    rc = rng.standard_normal(n_scat).astype(np.float32)
    pos, _, _ = fs.element_positions(params.n_elements, params.pitch, params.radius)
    angles = np.deg2rad(np.linspace(-16.0, 16.0, n_tx))
    delays = np.stack([fs.plane_wave(pos, float(a), speed_of_sound=medium.speed_of_sound) for a in angles])
    return params, medium, x, z, rc, delays


def rel_l2(a, b) -> float:
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def batched_useful_flops(n_scat: int, n_elem: int, n_freq: int, n_sub: int, n_tx: int) -> float:
    """Algorithmic flops of the batched form: per (scatterer, element, frequency) 8 n_sub for the Green's
    function (recurrence multiply + sub-element sum) and 16 per transmit (one complex MAC each for the TX and
    the RX contraction); the Green's function is shared by every transmit."""
    return float(n_scat) * n_elem * n_freq * (8 * n_sub + 16 * n_tx)


def msm_flops_per_launch(n_scat: int, n_elem: int, n_freq: int, fbk: int = 4) -> float:
    """MFMA FLOPs issued by one k_msm launch (upper-triangular 16x16 tiles per 16 scatterers and frequency:
    6 MFMAs for each off-diagonal tile, 4 for each diagonal one; frequencies padded to a multiple of the
    block depth fbk)."""
    et = 4 if n_elem <= 64 else 8
    mfma = 6 * (et * (et - 1) // 2) + 4 * et
    nfpad = -(-n_freq // fbk) * fbk
    tiles = -(-n_scat // 16)
    return tiles * nfpad * mfma * (16 * 16 * 32 * 2)


TRAFFIC_PROFILE = ROOT / "profiles" / "k_msm_planewave75_200k_traffic.json"
ADJM_TRAFFIC_PROFILE = ROOT / "profiles" / "k_adjm_adjoint15_traffic.json"


def measured_traffic(n_scat: int, n_tx: int, n_freq: int, n_elem: int, profile=TRAFFIC_PROFILE):
    """HBM bytes per launch of the profiled kernel (k_msm, or k_adjm with ADJM_TRAFFIC_PROFILE) from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE passes (calibrated on a 1 GiB copy), when they were taken on
    this very workload; else None."""
    try:
        d = json.loads(profile.read_text())
    except (OSError, ValueError):
        return None
    if d.get("workload") != {"n_scat": n_scat, "n_tx": n_tx, "n_freq": n_freq, "n_elem": n_elem}:
        return None
    return d["fetch_bytes"] + d["write_bytes"]


def cpu_baseline(params, medium, x, z, rc, delays, plan_o, n_sample: int, seed: int, repeats: int = 3):
    """FP64 NumPy oracle (the reference PYTHON strategy restated) on a bounded sample, 1 thread, best of
    ``repeats`` runs (plan excluded)."""
    from oracle import simus_oracle as O

    from threadpoolctl import threadpool_limits

    rng = np.random.default_rng(seed)
    idx = np.sort(rng.choice(x.size, n_sample, replace=False))
    xs, zs, rs = (a[idx].astype(np.float64) for a in (x, z, rc))
    t_mid = delays.shape[0] // 2
    times = []
    with threadpool_limits(limits=1):
        for _ in range(repeats):
            t0 = time.perf_counter()
            rf, _, _ = O.simus_oracle(xs, zs, rs, delays[t_mid], params, medium, plan_o)
            times.append(time.perf_counter() - t0)
    dt = min(times)
    evals = n_sample * params.n_elements * plan_o["freqs"].size
    return {"value": evals / dt / 1e9, "unit": UNIT, "cores": 1, "kind": "port",
            "sample": f"{n_sample} of {x.size} scatterers, 1 of {delays.shape[0]} transmits (tx {t_mid}), "
                      f"all {plan_o['freqs'].size} frequencies, FP64 NumPy, best of {repeats}: {dt:.2f} s "
                      f"(runs {', '.join(f'{t:.2f}' for t in times)} s), host "
                      f"{platform.processor() or platform.machine()} ({os.cpu_count()} logical CPUs visible); "
                      "rate per (scatterer, element, frequency) of one transmit"}, idx, t_mid, rf


def phantom_workload(seed: int = 2027, n_phantoms: int = 64, n_beams: int = 128):
    """Uneven phantom batch: 64-element 2.7 MHz phased array (pitch 0.30 mm, width 0.25 mm, 74 % bandwidth),
    phantoms of 10^U[4, 6] scatterers uniform over a 90-degree sector to 12 cm, rc ~ N(0, 1) float32,
    beams focused at 6 cm steered over [-40, 40] degrees."""
    import fastsimus_amd as fs

    params = fs.TransducerParams(freq_center=2.7e6, pitch=3.0e-4, width=2.5e-4, n_elements=64, bandwidth=0.74)
    medium = fs.MediumParams(speed_of_sound=1540.0)
    rng = np.random.default_rng(seed)
    phantoms = []
    for _ in range(n_phantoms):
        n = int(round(10 ** rng.uniform(4, 6)))
        phi = rng.uniform(-np.pi / 4, np.pi / 4, n)
        rho = 0.12 * np.sqrt(rng.uniform(0, 1, n))
        pts = np.stack([rho * np.sin(phi), rho * np.cos(phi)], -1).astype(np.float32)
        phantoms.append((pts, rng.standard_normal(n).astype(np.float32)))
    pos, _, _ = fs.element_positions(params.n_elements, params.pitch, params.radius)
    psi = np.deg2rad(np.linspace(-40.0, 40.0, n_beams))
    delays = np.stack([fs.focused(pos, np.array([0.06 * np.sin(a), 0.06 * np.cos(a)]), speed_of_sound=1540.0)
                       for a in psi])
    return params, medium, phantoms, delays


def bench_phantoms(args) -> int:
    """One step = the spectra (one persistent split-BF16 launch) and gated RF of every phantom and beam."""
    import torch

    from fastsimus_amd import _lib
    from fastsimus_amd import batch as B
    from fastsimus_amd import kernels as K

    dev = torch.device("cuda", 0)
    params, medium, phantoms, delays = phantom_workload(n_phantoms=args.n_phantoms)
    dph = [(torch.as_tensor(p, device=dev), torch.as_tensor(r, device=dev)) for p, r in phantoms]
    plan = B.simus_batch_precompute(dph, delays, params, medium)
    pts, rcs, off = B._concat(dph)
    prob = B._problem(pts, rcs, delays, None, plan, params, medium, False, dev)
    from fastsimus_amd import tune

    tuning = None if not args.tune else tune.autotune(
        prob, off, reps=1, candidates=[{"msm_fbk": 4, "msm_item_div": d} for d in (0, 4, 16)])
    ws = K.batch_workspace(prob, off, "mfma")
    nf, E, T, P, S = prob.n_freq, params.n_elements, prob.n_tx, len(phantoms), int(off[-1])
    lib = _lib.load()

    def step():
        sp = K.run_spectrum_batch(prob, off, "mfma", workspace=ws)
        rf, _, _ = K.run_rf(sp.reshape(P * T, nf, E), plan.n_freq_full, plan.freq_idx_start, plan.n_fft,
                            want_full=False)
        return rf

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    lib.fs_set_kernel_timing(1)
    _lib.kernel_time_ms()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    t_step = (time.perf_counter() - t0) / args.steps
    k_ms, k_n = _lib.kernel_time_ms()
    lib.fs_set_kernel_timing(0)
    k_avg = k_ms / max(k_n, 1) / 1e3
    tiles = sum(-(-int(off[q + 1] - off[q]) // 16) for q in range(P))
    nfpad = -(-nf // 4) * 4
    issued = tiles * nfpad * 10 * 6 * (16 * 16 * 32 * 2)  # 4 element tiles: 10 upper tile pairs
    useful = batched_useful_flops(S, E, nf, plan.n_sub, T)
    value = T * S * E * nf / t_step / 1e9

    # plain FP32 kernels on one beam of every phantom (cost linear in beams)
    prob1 = B._problem(pts, rcs, delays[T // 2], None, plan, params, medium, False, dev)
    ws1 = K.batch_workspace(prob1, off, "fp32")
    K.run_spectrum_batch(prob1, off, "fp32", workspace=ws1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    K.run_spectrum_batch(prob1, off, "fp32", workspace=ws1)
    torch.cuda.synchronize()
    plain_value = S * E * nf / (time.perf_counter() - t0) / 1e9

    # FP64 check on a sub-batch: the first 150 scatterers of 6 phantoms, 3 beams, shared plan
    from oracle import simus_oracle as O

    sub = [(p[:150], r[:150]) for p, r in phantoms[:6]]
    beams = delays[[0, T // 2, T - 1]]
    rf_sub = B.simus_batch(sub, beams, params, medium, strategy="hip_mfma").rf.cpu().numpy()
    allp = np.concatenate([p for p, _ in sub]).astype(np.float64)
    plan_o = O.simus_plan(allp[:, 0], allp[:, 1], beams, params, medium)
    from threadpoolctl import threadpool_limits

    errs = []
    t_cpu = 0.0
    with threadpool_limits(limits=1):
        for q, (p_, r_) in enumerate(sub):
            t0 = time.perf_counter()
            want = np.stack([O.simus_oracle(p_[:, 0].astype(np.float64), p_[:, 1].astype(np.float64),
                                            r_.astype(np.float64), d, params, medium, plan_o)[0] for d in beams])
            t_cpu += time.perf_counter() - t0
            errs.append(rel_l2(rf_sub[q], want))
    ev_cpu = sum(len(p_) for p_, _ in sub) * E * plan_o["freqs"].size * len(beams)
    cpu = {"value": ev_cpu / t_cpu / 1e9, "unit": UNIT, "cores": 1, "kind": "port",
           "sample": f"first 150 scatterers of 6 phantoms x 3 beams, all {plan_o['freqs'].size} frequencies of the "
                     f"sub-batch plan, FP64 NumPy, {t_cpu:.2f} s in all, host "
                     f"{platform.processor() or platform.machine()} ({os.cpu_count()} logical CPUs visible)"}

    out = {
        "metric": METRIC, "value": value, "unit": UNIT, "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": t_step * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16", "precision": "split-BF16 (hi+lo, 3 products) MFMA, FP32 accumulate; FP64 r and phase reseed",
        "data": "synthetic",
        "config": {"workload": f"uneven phantom batch: {P} phantoms of 10^U[4,6] scatterers in a 90-degree sector "
                               f"to 12 cm, 64-element 2.7 MHz phased array, {T} beams focused at 6 cm over "
                               "[-40,40] deg, one persistent launch",
                   "n_phantoms": P, "n_scat_total": S, "n_scat_min": int(np.diff(off).min()),
                   "n_scat_max": int(np.diff(off).max()), "n_elem": E, "n_tx": T, "n_freq": nf,
                   "n_sub": plan.n_sub, "method": "hip_mfma"},
        "roofline": {"bound": "mfma", "achieved": useful / k_avg / 1e12, "peak": PEAK_BF16 / 1e12, "unit": "TFLOP/s",
                     "frac": useful / k_avg / PEAK_BF16, "traffic": None, "kernel": "k_msm (persistent, all phantoms)",
                     "kernel_ms": k_avg * 1e3, "launches": k_n, "flops_per_launch": useful,
                     "flops_definition": "S E n_f (8 n_sub + 16 n_tx) useful flops / k_msm HIP-event duration",
                     "mfma_issued_flops_per_launch": issued, "mfma_issue_frac": issued / k_avg / PEAK_BF16},
        "plain_fp32": {"value": plain_value, "unit": UNIT, "transmits_timed": [T // 2]},
        "speedup_vs_plain": value / plain_value,
        "tuning": tuning,
        "rf_rel_l2_fp64_per_phantom": {"max": max(errs), "values": errs,
                                       "sample": "first 150 scatterers of phantoms 0-5, beams 0, 64, 127"},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


def adjoint_workload(seed: int = 2028):
    """Differentiable form: 128-element 7.6 MHz linear array, reflectivity on a 256 x 256 grid over
    x in [-2, 2] cm, z in [0.5, 4.5] cm jittered by U[-0.05, 0.05] mm, rc ~ N(0, 1), 15 plane waves in
    [-7, 7] deg, 0.5 dB/cm/MHz; the target phantom is rc + N(0, 0.1) (its RF taken at c = 1560 m/s)."""
    import fastsimus_amd as fs

    params = fs.TransducerParams(freq_center=7.6e6, pitch=3.0e-4, width=2.7e-4, n_elements=128, bandwidth=0.77)
    rng = np.random.default_rng(seed)
    gx, gz = np.meshgrid(np.linspace(-0.02, 0.02, 256), np.linspace(0.005, 0.045, 256), indexing="ij")
    x = (gx.ravel() + rng.uniform(-5e-5, 5e-5, gx.size)).astype(np.float32)
    z = (gz.ravel() + rng.uniform(-5e-5, 5e-5, gz.size)).astype(np.float32)
    rc = rng.standard_normal(x.size).astype(np.float32)
    rc_t = (rc + 0.1 * rng.standard_normal(x.size)).astype(np.float32)
    pos, _, _ = fs.element_positions(params.n_elements, params.pitch, params.radius)
    D = np.stack([fs.plane_wave(pos, float(a), speed_of_sound=1540.0) for a in np.deg2rad(np.linspace(-7, 7, 15))])
    return params, x, z, rc, rc_t, D, 0.5


def bench_adjoint(args) -> int:
    """Differentiable form: forward RF + gradient w.r.t. rc, x, z and c of an L2 loss, 256 x 256 jittered grid,
    15 plane waves in [-7, 7] deg; target from rc + N(0, 0.1) at c = 1560 m/s."""
    import torch

    import fastsimus_amd as fs
    from fastsimus_amd.autograd import simus_rf

    dev = torch.device("cuda", 0)
    params, x, z, rc, rc_t, D, att = adjoint_workload()
    plan = fs.simus_precompute(np.stack([x, z], -1), rc, D, params, fs.MediumParams(attenuation=att))
    xt, zt = torch.tensor(x, device=dev), torch.tensor(z, device=dev)
    with torch.no_grad():
        target = simus_rf(xt, zt, torch.tensor(rc_t, device=dev), 1560.0, D, params, plan, attenuation=att)

    def step(adjoint):
        xg = xt.clone().requires_grad_(True)
        zg = zt.clone().requires_grad_(True)
        rg = torch.tensor(rc, device=dev, requires_grad=True)
        cg = torch.tensor(1540.0, dtype=torch.float64, requires_grad=True)
        rf = simus_rf(xg, zg, rg, cg, D, params, plan, attenuation=att, adjoint=adjoint)
        loss = 0.5 * ((rf - target) ** 2).sum()
        loss.backward()
        return loss

    from fastsimus_amd import _lib

    lib = _lib.load()

    def timed(adjoint, steps, warmup):
        """Seconds per step, and ms per step and launches of the backward's dominant kernels (k_adjm per
        transmit group, or k_adj_tx + k_adj_pos per transmit) by HIP events on the stream they run on."""
        for _ in range(warmup):
            step(adjoint)
        torch.cuda.synchronize()
        lib.fs_set_kernel_timing(2)  # FS_TIMING_ADJOINT
        _lib.kernel_time_ms()
        t0 = time.perf_counter()
        for _ in range(steps):
            step(adjoint)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / steps
        k_ms, launches = _lib.kernel_time_ms()
        lib.fs_set_kernel_timing(0)
        return dt, k_ms / steps, launches

    t_step, k_ms, k_launches = timed(args.adjoint, args.steps, args.warmup)
    # the other backward kernels on the same step, for comparison (FP32 vector kernels regenerate the
    # Green's functions per transmit; the matrix-core form once per group of 16 transmits)
    other = "vector" if args.adjoint != "vector" else "matrix"
    t_other, k_other, _ = timed(other, 2, 1)
    nf = int(plan.selected_freqs.shape[0])
    n_tx, S, E = int(D.shape[0]), int(x.size), params.n_elements
    value = n_tx * S * E * nf / t_step / 1e9
    roofline = None
    if args.adjoint != "vector":
        # k_adjm: per workgroup (16 scatterers, 8 element slices of 16) and padded frequency, 24 MFMA
        # 16x16x32 per slice (Y = G R and W^T = R L^T, three split-bf16 products each for the real and the
        # imaginary part); one launch per group of 16 transmits, frequencies padded to 32-frequency chunks
        nch = -(-nf // 32)
        slices = 4 if E <= 64 else 8
        groups = -(-n_tx // 16)
        issued = groups * (-(-S // 16)) * nch * 32 * slices * 24 * 16 * 16 * 32 * 2
        # algorithmic flops of the transmit contractions: P_t and B_t (two complex MACs per (t, s, e, f)) and
        # W = sum_t (two complex MACs per (t, s, e, f))
        useful = 32 * n_tx * S * E * nf
        roofline = {"bound": "mfma", "achieved": useful / (k_ms * 1e-3) / 1e12, "peak": 2500.0, "unit": "TFLOP/s",
                    "frac": useful / (k_ms * 1e-3) / 2.5e15,
                    "traffic": measured_traffic(S, n_tx, nf, E, ADJM_TRAFFIC_PROFILE),
                    "kernel": "k_adjm (split-BF16 v_mfma_f32_16x16x32_bf16 + FP32 VALU terms)",
                    "kernel_ms": k_ms, "launches": k_launches, "flops_per_launch": useful / groups,
                    "flops_definition": "32 n_tx S E n_f useful flops (P_t, B_t and the transmit sum of W: two "
                                        "complex MACs each per (t, s, e, f)) / k_adjm HIP-event duration",
                    "mfma_issued_flops_per_launch": issued / groups,
                    "mfma_issue_tflops": issued / (k_ms * 1e-3) / 1e12,
                    "mfma_issue_frac": issued / (k_ms * 1e-3) / 2.5e15}
    cpu = None if args.no_cpu else adjoint_cpu_baseline(params, x, z, rc, rc_t, D, att)
    out = {
        "metric": "G(scatterer·element·freq)/s, forward + adjoint (dL/d rc, x, z, c)", "value": value, "unit": UNIT,
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": t_step * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "bf16", "data": "synthetic",
        "precision": "split-BF16 (hi+lo, 3 products) MFMA, FP32 accumulate; FP64 r and phase reseed",
        "config": {"workload": "differentiable SIMUS: 256x256 jittered grid (65536 scatterers), 15 plane waves in "
                               "[-7,7] deg, L2 loss vs RF of rc + N(0,0.1) at c = 1560 m/s",
                   "n_scat": int(x.size), "n_elem": params.n_elements, "n_tx": int(D.shape[0]), "n_freq": nf,
                   "n_sub": plan.n_sub,
                   "method": "hip_mfma forward (multistatic matrix) + hand-written adjoint", "adjoint": args.adjoint},
        f"{other}_adjoint_ms_per_step": t_other * 1e3, f"{other}_adjoint_kernel_ms": k_other,
        "roofline": roofline, "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


def adjoint_cpu_baseline(params, x, z, rc, rc_t, D, att, n_sample: int = 32, repeats: int = 3):
    """The FP64 torch restatement of the differentiable reference path (oracle.simus_grad, the scan strategy's
    forward with autograd through it) on a bounded sample of the jittered grid: forward RF, L2 loss against the
    target of the perturbed phantom at c = 1560 m/s, and the gradient w.r.t. rc, x, z and c, all 15 transmits and
    all frequencies, one thread, best of ``repeats`` (target and plan outside the timing)."""
    import torch

    import fastsimus_amd as fs
    from oracle import simus_oracle as O
    from oracle.simus_grad import simus_rf_torch

    medium = fs.MediumParams(attenuation=att)
    plan_o = O.simus_plan(x.astype(np.float64), z.astype(np.float64), D, params, medium)
    idx = np.sort(np.random.default_rng(13).choice(x.size, n_sample, replace=False))
    xs, zs, rs, rts = (a[idx].astype(np.float64) for a in (x, z, rc, rc_t))
    threads = torch.get_num_threads()
    torch.set_num_threads(1)
    try:
        with torch.no_grad():
            target = simus_rf_torch(torch.as_tensor(xs), torch.as_tensor(zs), torch.as_tensor(rts), 1560.0, D,
                                    params, att, plan_o)
        times = []
        for _ in range(repeats):
            t0 = time.perf_counter()
            xt, zt, rt = (torch.tensor(a, requires_grad=True) for a in (xs, zs, rs))
            ct = torch.tensor(1540.0, dtype=torch.float64, requires_grad=True)
            loss = 0.5 * ((simus_rf_torch(xt, zt, rt, ct, D, params, att, plan_o) - target) ** 2).sum()
            loss.backward()
            times.append(time.perf_counter() - t0)
    finally:
        torch.set_num_threads(threads)
    dt = min(times)
    evals = n_sample * params.n_elements * plan_o["freqs"].size * D.shape[0]
    return {"value": evals / dt / 1e9, "unit": UNIT, "cores": 1, "kind": "port",
            "sample": f"{n_sample} of {x.size} scatterers, all {D.shape[0]} transmits and "
                      f"{plan_o['freqs'].size} frequencies, FP64 torch forward + autograd backward "
                      f"(oracle.simus_grad), best of {repeats}: {dt:.2f} s (runs {', '.join(f'{t:.2f}' for t in times)} "
                      f"s), host {platform.processor() or platform.machine()} ({os.cpu_count()} logical CPUs visible); "
                      "rate per (transmit, scatterer, element, frequency) of forward + gradient"}


def single_workloads(seed: int = 2025):
    """The two single-transmit workloads: 2 000 speckle scatterers, zero delays (the CPU-runnable case), and
    50 000 speckle scatterers plus 10 bright points (x = 0, z = 5 .. 50 mm, rc = 10) under a 0 degree plane
    wave; 128-element 7.6 MHz linear array, 0.5 dB/cm/MHz, inputs drawn in float64 and rounded once to float32."""
    import fastsimus_amd as fs

    params = fs.TransducerParams(freq_center=7.6e6, pitch=3.0e-4, width=2.7e-4, n_elements=128, bandwidth=0.77)
    medium = fs.MediumParams(speed_of_sound=1540.0, attenuation=0.5)
    pos, _, _ = fs.element_positions(params.n_elements, params.pitch, params.radius)
    rng = np.random.default_rng(seed)
    x0 = rng.uniform(-0.02, 0.02, 2000).astype(np.float32)
    z0 = rng.uniform(0.005, 0.045, 2000).astype(np.float32)
    r0 = rng.standard_normal(2000).astype(np.float32)
    x1 = np.concatenate([rng.uniform(-0.02, 0.02, 50_000), np.zeros(10)]).astype(np.float32)
    z1 = np.concatenate([rng.uniform(0.005, 0.05, 50_000), np.linspace(0.005, 0.05, 10)]).astype(np.float32)
    r1 = np.concatenate([rng.standard_normal(50_000), np.full(10, 10.0)]).astype(np.float32)
    d0 = np.zeros(params.n_elements)
    d1 = fs.plane_wave(pos, 0.0, speed_of_sound=medium.speed_of_sound)
    return params, medium, (x0, z0, r0, d0), (x1, z1, r1, d1)


def bench_single(args) -> int:
    """Single-transmit workloads (one step = spectrum + RF of one simus_compute, inputs in HBM): the
    2 000-scatterer case against the FP64 NumPy oracle in full, and the 50 010-scatterer case with the
    plain FP32 kernels (FP64 error on a 1 010-scatterer sample: 1 000 speckle + the 10 points)."""
    import torch
    from threadpoolctl import threadpool_limits

    import fastsimus_amd as fs
    from fastsimus_amd import _lib
    from fastsimus_amd import kernels as K
    from oracle import simus_oracle as O

    dev = torch.device("cuda", 0)
    params, medium, c0, c1 = single_workloads()

    def problem(x, z, rc, d, plan):
        sc = np.stack([x, z], -1)
        return K.Problem(sc, rc, d, None, params, medium, n_sub=plan.n_sub, seg_length=plan.seg_length,
                         freqs=plan.selected_freqs, pulse=plan.pulse_spectrum, probe=plan.probe_spectrum,
                         correction=plan.correction_factor, full_frequency_directivity=False, device=dev)

    lib = _lib.load()

    def timed(prob, plan, method):
        """Seconds per step, transmit 0's RF, and ms per step of the dominant kernels (k_tx + k_rx for
        fp32, k_msm for mfma) by HIP events on the stream they run on."""
        def step():
            sp = K.run_spectrum(prob, method)
            return K.run_rf(sp, plan.n_freq_full, plan.freq_idx_start, plan.n_fft, want_full=False)[0]
        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        lib.fs_set_kernel_timing(1)
        _lib.kernel_time_ms()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            rf = step()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / args.steps
        k_ms, _ = _lib.kernel_time_ms()
        lib.fs_set_kernel_timing(0)
        return dt, rf[0].cpu().numpy(), k_ms / args.steps

    out = {}
    # 2 000 scatterers: both device methods, FP64 oracle in full on one core
    x, z, rc, d = c0
    plan = fs.simus_precompute(np.stack([x, z], -1), rc, d, params, medium)
    plan_o = O.simus_plan(x.astype(np.float64), z.astype(np.float64), d, params, medium)
    prob = problem(x, z, rc, d, plan)
    evals = x.size * params.n_elements * plan.selected_freqs.shape[0]
    with threadpool_limits(limits=1):
        t0 = time.perf_counter()
        want, _, _ = O.simus_oracle(x.astype(np.float64), z.astype(np.float64), rc.astype(np.float64), d, params,
                                    medium, plan_o)
        t_cpu = time.perf_counter() - t0
    r = {"n_scat": int(x.size), "n_freq": int(plan.selected_freqs.shape[0]),
         "cpu_baseline": {"value": evals / t_cpu / 1e9, "unit": UNIT, "cores": 1, "kind": "port",
                          "sample": f"all {x.size} scatterers, FP64 NumPy, {t_cpu:.2f} s"}}
    for method in ("fp32", "mfma"):
        t, rf, _ = timed(prob, plan, method)
        r[method] = {"value": evals / t / 1e9, "ms_per_step": t * 1e3, "rf_rel_l2_vs_fp64": rel_l2(rf, want)}
    out["2000_scatterers_zero_delays"] = r
    # 50 010 scatterers: plain FP32 kernels; FP64 error on a sample of the same plan
    x, z, rc, d = c1
    plan = fs.simus_precompute(np.stack([x, z], -1), rc, d, params, medium)
    plan_o = O.simus_plan(x.astype(np.float64), z.astype(np.float64), d, params, medium)
    evals = x.size * params.n_elements * plan.selected_freqs.shape[0]
    r = {"n_scat": int(x.size), "n_freq": int(plan.selected_freqs.shape[0])}
    for method in ("fp32", "mfma"):
        t, _, k_ms = timed(problem(x, z, rc, d, plan), plan, method)
        r[method] = {"value": evals / t / 1e9, "ms_per_step": t * 1e3, "kernel_ms": k_ms}
    idx = np.concatenate([np.sort(np.random.default_rng(3).choice(50_000, 1000, replace=False)),
                          np.arange(50_000, 50_010)])
    xs, zs, rs = x[idx], z[idx], rc[idx]
    with threadpool_limits(limits=1):
        t0 = time.perf_counter()
        want, _, _ = O.simus_oracle(xs.astype(np.float64), zs.astype(np.float64), rs.astype(np.float64), d, params,
                                    medium, plan_o)
        t_cpu = time.perf_counter() - t0
    for method in ("fp32", "mfma"):
        _, rf, _ = timed(problem(xs, zs, rs, d, plan), plan, method)
        r[method]["rf_rel_l2_vs_fp64"] = rel_l2(rf, want)
    ev_s = idx.size * params.n_elements * plan_o["freqs"].size
    r["fp64_sample"] = "1 010 scatterers (1 000 speckle + the 10 points), same plan"
    r["cpu_baseline"] = {"value": ev_s / t_cpu / 1e9, "unit": UNIT, "cores": 1, "kind": "port",
                         "sample": f"the 1 010-scatterer sample, FP64 NumPy, {t_cpu:.2f} s"}
    out["50010_scatterers_plane_wave"] = r
    main_r = out["50010_scatterers_plane_wave"]["fp32"]
    # dominant kernels of the value: k_tx + k_rx; 8 n_sub + 16 flops per evaluation, and the VALU
    # instructions they issue for it (counted with SQ_INSTS_VALU at n_sub = 2)
    k_s = main_r["kernel_ms"] / 1e3
    fl = evals * (8 * plan.n_sub + 16)
    roof = {"bound": "valu", "achieved": fl / k_s / 1e12, "peak": PEAK_FP32 / 1e12, "unit": "TFLOP/s",
            "frac": fl / k_s / PEAK_FP32, "traffic": None, "kernel": "k_tx + k_rx (FP32 VALU)",
            "kernel_ms": main_r["kernel_ms"],
            "valu_instructions_per_eval": VALU_PER_EVAL_NSUB2 if plan.n_sub == 2 else None,
            "valu_issue_tlanes_per_s": VALU_PER_EVAL_NSUB2 * evals / k_s / 1e12 if plan.n_sub == 2 else None,
            "valu_issue_peak_tlanes_per_s": PEAK_FP32 / 2 / 1e12}
    res = {
        "metric": "G(scatterer·element·freq)/s, single transmit (plain FP32 and split-BF16 MFMA)",
        "value": main_r["value"], "unit": UNIT, "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": main_r["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32", "data": "synthetic",
        "config": {"workload": "single transmit, 128-element 7.6 MHz linear array, 0.5 dB/cm/MHz: 2 000 speckle "
                               "scatterers with zero delays, and 50 000 speckle + 10 bright points under a 0 degree "
                               "plane wave (value: the latter, plain FP32)", "n_elem": params.n_elements},
        "configs": out, "roofline": roof,
        "cpu_baseline": out["50010_scatterers_plane_wave"]["cpu_baseline"],
    }
    print(json.dumps(res), flush=True)
    return 0


def bench_pfield(args) -> int:
    """RMS pressure field of one focused transmit of the 128-element linear array on a 512 x 512 grid
    (x in [-2, 2] cm, z in [0.2, 5] cm); one step = fs_pfield over the grid, inputs in HBM."""
    import torch

    import fastsimus_amd as fs
    from fastsimus_amd import _lib
    from fastsimus_amd import kernels as K
    from oracle import simus_oracle as O

    dev = torch.device("cuda", 0)
    params = fs.TransducerParams(freq_center=7.6e6, pitch=3.0e-4, width=2.7e-4, n_elements=128, bandwidth=0.77)
    medium = fs.MediumParams(speed_of_sound=1540.0, attenuation=0.5)
    gx, gz = np.meshgrid(np.linspace(-0.02, 0.02, 512), np.linspace(0.002, 0.05, 512), indexing="ij")
    grid = np.stack([gx, gz], -1).astype(np.float32)
    pos, _, _ = fs.element_positions(params.n_elements, params.pitch, params.radius)
    delays = fs.focused(pos, np.array([0.0, 0.02]), speed_of_sound=1540.0)
    plan = fs.pfield_precompute(grid, delays, params, medium)
    prob = K.Problem(grid, None, delays, None, params, medium, n_sub=plan.n_sub, seg_length=plan.seg_length,
                     freqs=plan.selected_freqs, pulse=plan.pulse_spectrum, probe=plan.probe_spectrum,
                     correction=plan.correction_factor, full_frequency_directivity=False, device=dev)
    lib = _lib.load()
    ws = K._workspace(lib.fs_pfield_workspace_bytes(prob.dref), dev)
    out = torch.empty(prob.x.numel(), dtype=torch.float32, device=dev)

    def step():
        _lib.check(lib.fs_pfield(prob.dref, out.data_ptr(), ws.data_ptr(), ws.numel(), K._stream(dev)), "fs_pfield")

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    t_step = (time.perf_counter() - t0) / args.steps
    nf = int(plan.selected_freqs.shape[0])
    npts = grid.shape[0] * grid.shape[1]
    evals = npts * params.n_elements * nf
    # FP64 oracle on a sample of grid points (same plan)
    rng = np.random.default_rng(11)
    idx = rng.choice(npts, 400, replace=False)
    flat = grid.reshape(-1, 2).astype(np.float64)
    plan_o = O.pfield_plan(flat[:, 0], flat[:, 1], delays, params, medium)
    t0 = time.perf_counter()
    want = O.pfield_oracle(flat[idx, 0], flat[idx, 1], delays, params, medium, plan_o)
    t_cpu = time.perf_counter() - t0
    err = rel_l2(out.cpu().numpy()[idx], want)
    fl = evals * (8 * plan.n_sub + 8)
    res = {
        "metric": "G(point·element·freq)/s, pfield (RMS pressure field)", "value": evals / t_step / 1e9,
        "unit": "G(point·element·freq)/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": t_step * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32", "data": "synthetic",
        "config": {"workload": "pfield: 128-element 7.6 MHz linear array focused at 2 cm, 512x512 grid over "
                               "[-2,2] x [0.2,5] cm, 0.5 dB/cm/MHz", "n_points": npts, "n_elem": params.n_elements,
                   "n_freq": nf, "n_sub": plan.n_sub},
        "roofline": {"bound": "valu", "achieved": fl / t_step / 1e12, "peak": PEAK_FP32 / 1e12, "unit": "TFLOP/s",
                     "frac": fl / t_step / PEAK_FP32, "traffic": None, "kernel": "k_tx (pfield mode)"},
        "rf_rel_l2": {"pfield_vs_fp64": err, "fp64_sample": "400 grid points"},
        "cpu_baseline": {"value": 400 * params.n_elements * plan_o["freqs"].size / t_cpu / 1e9,
                         "unit": "G(point·element·freq)/s", "cores": 1, "kind": "port",
                         "sample": f"400 of {npts} grid points, FP64 NumPy, {t_cpu:.1f} s"},
    }
    print(json.dumps(res), flush=True)
    return 0


def bench_beamform(args) -> int:
    """Consumers of the RF stack: the RF of 75 plane waves over [-16, 16] deg (128-element 7.6 MHz array,
    20 000 speckle scatterers, simulated once outside the timing) demodulated to IQ and delay-and-sum
    beamformed onto a 256 x 384 grid (x in [-2, 2] cm, z in [0.5, 4.5] cm), summed over the firings. One step
    = fs_rf2iq over the stack + fs_das of all firings; both kernels timed with HIP events."""
    import torch

    import fastsimus_amd as fs
    from fastsimus_amd import beamform as BF
    from oracle import beamform_oracle as BO

    dev = torch.device("cuda", 0)
    params, medium, x, z, rc, delays = workload(n_scat=20_000)
    res = fs.simus(np.stack([x, z], -1), rc, delays, params, medium, strategy="hip_mfma")
    rf = res.rf.contiguous()
    T, N, E = rf.shape
    fs_hz, fc = 4 * params.freq_center, params.freq_center
    gx, gz = np.meshgrid(np.linspace(-0.02, 0.02, 256), np.linspace(0.005, 0.045, 384))
    F = BF.f_number_for(params)
    # inputs resident on the device before timing: pixel grid and delays
    gx_d = torch.as_tensor(gx, dtype=torch.float32, device=rf.device)
    gz_d = torch.as_tensor(gz, dtype=torch.float32, device=rf.device)
    delays_d = torch.as_tensor(delays, dtype=torch.float64, device=rf.device)

    def step():
        iq = BF.rf2iq(rf, fs_hz, fc, channel_major=True)
        return iq, BF.das(iq, delays_d, params, gx_d, gz_d, fs_hz, fc, f_number=F, channel_major=True)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        iq, img = step()
    torch.cuda.synchronize()
    t_step = (time.perf_counter() - t0) / args.steps
    # the two kernels apart (HIP events on torch's current stream, which the library launches on)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    ev[0].record()
    iq = BF.rf2iq(rf, fs_hz, fc, channel_major=True)
    ev[1].record()
    img = BF.das(iq, delays_d, params, gx_d, gz_d, fs_hz, fc, f_number=F, channel_major=True)
    ev[2].record()
    torch.cuda.synchronize()
    t_iq, t_das = ev[0].elapsed_time(ev[1]), ev[1].elapsed_time(ev[2])
    # FP64 / SciPy checkers on samples: 2 firings of IQ, 300 pixels of the image
    rf_h, iq_h = rf.cpu().numpy(), iq.permute(0, 2, 1).cpu().numpy()
    e_iq = max(rel_l2(np.abs(iq_h[t]), np.abs(BO.rf2iq(rf_h[t], fs_hz, fc))) for t in (0, T - 1))
    pos, _, _ = fs.element_positions(E, params.pitch, params.radius)
    idx = np.random.default_rng(4).choice(gx.size, 300, replace=False)
    t0 = time.perf_counter()
    want = BO.das(iq_h, delays, pos[:, 0], pos[:, 1], gx.reshape(-1)[idx].astype(np.float32),
                  gz.reshape(-1)[idx].astype(np.float32), fs_hz, fc, 1540.0, F)
    t_cpu = time.perf_counter() - t0
    e_das = rel_l2(img.cpu().numpy().reshape(-1)[idx], want)
    evals = gx.size * E * T
    out = {
        "metric": "G(pixel·element·firing)/s, IQ demodulation + delay-and-sum of the RF stack", "value": evals / t_step / 1e9,
        "unit": "G(pixel·element·firing)/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": t_step * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32", "precision": "FP32 IIR sections and interpolation; FP64 delays and phase reduction",
        "data": "synthetic",
        "config": {"workload": "rf2iq + DAS compound of 75 plane-wave firings (128-element 7.6 MHz array, 20 000 "
                               "speckle scatterers) onto a 256 x 384 grid", "n_firings": T, "n_samples": N,
                   "n_elem": E, "n_pixels": int(gx.size), "f_number": F,
                   "iq_layout": "channel-major (firing, element, sample), as beamform() runs it"},
        "kernel_ms": {"fs_rf2iq": t_iq, "fs_das": t_das},
        "roofline": None,
        "rel_l2": {"iq_vs_scipy_filtfilt": e_iq, "das_vs_fp64_checker_300_pixels": e_das},
        "cpu_baseline": {"value": 300 * E * T / t_cpu / 1e9, "unit": "G(pixel·element·firing)/s", "cores": 1,
                         "kind": "port", "sample": f"300 pixels, FP64 NumPy, {t_cpu:.2f} s"},
    }
    print(json.dumps(out), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n-scat", type=int, default=200_000)
    ap.add_argument("--n-tx", type=int, default=75)
    ap.add_argument("--plain-tx", type=int, default=3, help="transmits timed with the plain FP32 kernel")
    ap.add_argument("--cpu-sample", type=int, default=1000, help="scatterers in the CPU-baseline / FP64 check sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--adjoint", choices=("auto", "matrix", "vector"), default="auto",
                    help="adjoint15: backward kernels (matrix cores where they apply, or FP32 vector kernels)")
    ap.add_argument("--workload", choices=("planewave75", "phantoms64", "adjoint15", "pfield", "single", "beamform"),
                    default="planewave75",
                    help="planewave75 (default, the headline metric), phantoms64 (uneven phantom batch, phased "
                         "array, focused beams), adjoint15 (forward + gradient of the differentiable form) or "
                         "pfield (RMS pressure field on a grid) or single (the single-transmit 2 000- and "
                         "50 010-scatterer cases)")
    ap.add_argument("--n-phantoms", type=int, default=64)
    ap.add_argument("--tune", action="store_true", help="phantoms64: run the tile-shape autotuner before timing")
    ap.add_argument("--no-tune", action="store_true",
                    help="headline: time the library's default tile shapes instead of the autotuner's choice")
    ap.add_argument("--no-graph", action="store_true", help="headline: time eager launches instead of graph replays")
    ap.add_argument("--copy-calib-mib", type=int, default=0,
                    help="only run a device-to-device copy of this many MiB (calibrates the FETCH_SIZE / WRITE_SIZE "
                         "counters of a profiler against a known byte count) and exit")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the library has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        return 2

    import torch

    import fastsimus_amd as fs
    from fastsimus_amd import _lib
    from fastsimus_amd import kernels as K

    dev = torch.device("cuda", 0)
    if args.copy_calib_mib > 0:
        n = args.copy_calib_mib * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(2):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calib_bytes": n * 4, "launches": 2}), flush=True)
        return 0
    if args.workload == "phantoms64":
        return bench_phantoms(args)
    if args.workload == "adjoint15":
        return bench_adjoint(args)
    if args.workload == "single":
        return bench_single(args)
    if args.workload == "pfield":
        return bench_pfield(args)
    if args.workload == "beamform":
        return bench_beamform(args)
    params, medium, x, z, rc, delays = workload(n_scat=args.n_scat, n_tx=args.n_tx)
    sc = np.stack([x, z], -1)
    plan = fs.simus_precompute(sc, rc, delays, params, medium)
    nf = int(plan.selected_freqs.shape[0])
    S, E, T = x.size, params.n_elements, delays.shape[0]

    def problem(d):
        return K.Problem(sc, rc, d, None, params, medium, n_sub=plan.n_sub, seg_length=plan.seg_length,
                         freqs=plan.selected_freqs, pulse=plan.pulse_spectrum, probe=plan.probe_spectrum,
                         correction=plan.correction_factor, full_frequency_directivity=False, device=dev)

    prob = problem(delays)
    lib = _lib.load()
    from fastsimus_amd import tune

    tuning = None if args.no_tune else tune.autotune(prob, reps=1)
    # one step: spectrum (multistatic matrix + apply), placed full spectrum and gated RF of every transmit
    bufs = {}

    def step():
        sp = K.run_spectrum(prob, "mfma")
        rf, full, _ = K.run_rf(sp, plan.n_freq_full, plan.freq_idx_start, plan.n_fft, want_full=True)
        bufs["rf"], bufs["full"] = rf, full
        return rf

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    graph = None
    if not args.no_graph:
        # capture on a side stream after the eager warm-up (kernel attributes set, code objects loaded);
        # kernel timing is off while capturing
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            rf_m = step()
        graph.replay()
        torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        if graph is not None:
            graph.replay()
        else:
            rf_m = step()
    torch.cuda.synchronize()
    t_step = (time.perf_counter() - t0) / args.steps
    # the same step eager, its k_msm launches bracketed by HIP events on their stream
    lib.fs_set_kernel_timing(1)
    _lib.kernel_time_ms()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        rf_e = step()
    torch.cuda.synchronize()
    t_eager = (time.perf_counter() - t0) / args.steps
    k_ms, k_n = _lib.kernel_time_ms()
    lib.fs_set_kernel_timing(0)
    graph_vs_eager = float((rf_m - rf_e).abs().max()) if graph is not None else 0.0
    evals = T * S * E * nf
    value = evals / t_step / 1e9
    k_avg = k_ms / max(k_n, 1) / 1e3
    fbk = (tuning["chosen"].get("msm_fbk") or 4) if tuning else 4
    issued = msm_flops_per_launch(S, E, nf, fbk)
    useful = batched_useful_flops(S, E, nf, plan.n_sub, T)
    roof = {"bound": "mfma", "achieved": useful / k_avg / 1e12, "peak": PEAK_BF16 / 1e12, "unit": "TFLOP/s",
            "frac": useful / k_avg / PEAK_BF16, "traffic": measured_traffic(S, T, nf, E),
            "kernel": "k_msm (split-BF16 v_mfma_f32_16x16x32_bf16)", "kernel_ms": k_avg * 1e3, "launches": k_n,
            "flops_per_launch": useful,
            "flops_definition": "S E n_f (8 n_sub + 16 n_tx) useful flops per launch (Green's function shared by "
                                "all transmits) / k_msm HIP-event duration",
            "mfma_issued_flops_per_launch": issued,
            "mfma_issue_tflops": issued / k_avg / 1e12, "mfma_issue_frac": issued / k_avg / PEAK_BF16}

    # plain FP32 kernel on a transmit subset of the same inputs
    sub = np.linspace(0, T - 1, max(1, args.plain_tx)).round().astype(int)
    prob_p = problem(delays[sub])
    sp_p = K.run_spectrum(prob_p, "fp32")
    torch.cuda.synchronize()
    lib.fs_set_kernel_timing(1)
    _lib.kernel_time_ms()
    t0 = time.perf_counter()
    sp_p = K.run_spectrum(prob_p, "fp32")
    rf_p, _, _ = K.run_rf(sp_p, plan.n_freq_full, plan.freq_idx_start, plan.n_fft, want_full=False)
    torch.cuda.synchronize()
    t_plain = time.perf_counter() - t0
    kp_ms, kp_n = _lib.kernel_time_ms()
    lib.fs_set_kernel_timing(0)
    evals_p = len(sub) * S * E * nf
    plain_value = evals_p / t_plain / 1e9
    fl_plain = evals_p * (8 * plan.n_sub + 16)
    plain = {"value": plain_value, "unit": UNIT, "transmits_timed": [int(i) for i in sub],
             "roofline": {"bound": "valu", "achieved": fl_plain / (kp_ms / 1e3) / 1e12, "peak": PEAK_FP32 / 1e12,
                          "unit": "TFLOP/s", "frac": fl_plain / (kp_ms / 1e3) / PEAK_FP32,
                          "kernel": "k_tx + k_rx (FP32 VALU)", "kernel_ms_total": kp_ms, "launches": kp_n,
                          # 32 of the flops per evaluation are useful, but the kernels issue 31.9 VALU
                          # instructions per evaluation at n_sub = 2 (k_tx 15.1 with two frequency
                          # blocks per thread + k_rx 16.8, counted with SQ_INSTS_VALU), so their bound
                          # is the VALU issue rate: 256 CUs x 128 lanes per cycle (two wave64
                          # instructions every 4 cycles per SIMD, reached without packed FP32) =
                          # 78.6 T lane-instructions/s at 2.4 GHz
                          "valu_instructions_per_eval": VALU_PER_EVAL_NSUB2 if plan.n_sub == 2 else None,
                          "valu_issue_tlanes_per_s": (VALU_PER_EVAL_NSUB2 * evals_p / (kp_ms / 1e3) / 1e12
                                                      if plan.n_sub == 2 else None),
                          "valu_issue_peak_tlanes_per_s": PEAK_FP32 / 2 / 1e12}}
    err = {"mfma_vs_fp32_full": rel_l2(rf_m[sub].cpu().numpy(), rf_p.cpu().numpy())}

    # FP64 oracle on a scatterer subsample: CPU baseline and absolute error of both GPU paths
    cpu = None
    if not args.no_cpu:
        from oracle import simus_oracle as O

        plan_o = O.simus_plan(x.astype(np.float64), z.astype(np.float64), delays, params, medium)
        cpu, idx, t_mid, rf_o = cpu_baseline(params, medium, x, z, rc, delays, plan_o, args.cpu_sample, seed=7)
        sc_s = sc[idx]
        for strat, key in (("hip", "fp32_vs_fp64"), ("hip_mfma", "mfma_vs_fp64")):
            r = fs.simus_compute(sc_s, rc[idx], delays[t_mid], plan, params, medium, strategy=strat)
            err[key] = rel_l2(r.rf.cpu().numpy(), rf_o)
        err["fp64_sample"] = f"{args.cpu_sample} scatterers, transmit {t_mid}"

    out = {
        "metric": METRIC, "value": value, "unit": UNIT, "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": t_step * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16", "precision": "split-BF16 (hi+lo, 3 products) MFMA, FP32 accumulate; FP64 r and phase reseed",
        "data": "synthetic",
        "config": {"workload": f"batched plane-wave compounding: 128-element 7.6 MHz linear array, {T} plane waves "
                               f"in [-16,16] deg, {S} speckle scatterers, 0.5 dB/cm/MHz",
                   "n_scat": S, "n_elem": E, "n_tx": T, "n_freq": nf, "n_sub": plan.n_sub, "n_fft": plan.n_fft,
                   "method": "hip_mfma", "tiles": "autotuned" if tuning else "library defaults",
                   "launch": "hipGraph replay" if graph is not None else "eager",
                   "step": "spectrum + placed full spectrum + gated RF, all transmits"},
        "eager_ms_per_step": t_eager * 1e3, "graph_vs_eager_rf_max_abs_diff": graph_vs_eager,
        "roofline": roof, "plain_fp32": plain, "speedup_vs_plain": value / plain_value, "rf_rel_l2": err,
        "tuning": tuning,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
