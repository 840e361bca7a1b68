"""Benchmark of the batched SIMUS hot path on one MI355X.

Workload (default): 128-element 7.6 MHz linear array (pitch 0.30 mm, width
0.27 mm, 77 % bandwidth), c = 1540 m/s, 0.5 dB/cm/MHz, fs = 4 fc; 200 000
speckle scatterers (x ~ U[-2, 2] cm, z ~ U[0.5, 5] cm, rc ~ N(0, 1), float32,
seed 2026); 75 plane waves over [-16, 16] deg sharing one frequency plan.

One step = the whole compute path for all 75 transmits with inputs resident
in HBM: the split-BF16 MFMA spectrum (multistatic matrix + transmit apply)
and the spectrum -> RF transform with the -100 dB gate. The plan is built once,
outside the timed region (as the reference's own benchmark does).

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


def msm_flops_per_launch(n_scat: int, n_elem: int, n_freq: int) -> float:
    """MFMA FLOPs issued by one k_msm launch (upper-triangular 16x16 tiles, 6 MFMAs each per 16 scatterers)."""
    et = 4 if n_elem <= 64 else 8
    pairs = et * (et + 1) // 2
    fbk = 2
    nfpad = -(-n_freq // fbk) * fbk
    tiles = -(-n_scat // 16)
    return tiles * nfpad * pairs * 6 * (16 * 16 * 32 * 2)


TRAFFIC_PROFILE = ROOT / "profiles" / "k_msm_planewave75_200k_traffic.json"


def measured_traffic(n_scat: int, n_tx: int, n_freq: int, n_elem: int):
    """HBM bytes per k_msm launch from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE passes
    (calibrated on a 1 GiB copy), when they were taken on this very workload; else None."""
    try:
        d = json.loads(TRAFFIC_PROFILE.read_text())
    except (OSError, ValueError):
        return None
    if d.get("workload") != {"n_scat": n_scat, "n_tx": n_tx, "n_freq": n_freq, "n_elem": n_elem}:
        return None
    return d["fetch_bytes"] + d["write_bytes"]


def cpu_baseline(params, medium, x, z, rc, delays, plan_o, n_sample: int, seed: int):
    """FP64 NumPy oracle (the reference PYTHON strategy restated) on a bounded sample, 1 thread."""
    from oracle import simus_oracle as O

    from threadpoolctl import threadpool_limits

    rng = np.random.default_rng(seed)
    idx = np.sort(rng.choice(x.size, n_sample, replace=False))
    xs, zs, rs = (a[idx].astype(np.float64) for a in (x, z, rc))
    t_mid = delays.shape[0] // 2
    with threadpool_limits(limits=1):
        t0 = time.perf_counter()
        rf, _, _ = O.simus_oracle(xs, zs, rs, delays[t_mid], params, medium, plan_o)
        dt = time.perf_counter() - t0
    evals = n_sample * params.n_elements * plan_o["freqs"].size
    return {"value": evals / dt / 1e9, "unit": UNIT, "cores": 1, "kind": "port",
            "sample": f"{n_sample} of {x.size} scatterers, 1 of {delays.shape[0]} transmits (tx {t_mid}), "
                      f"all {plan_o['freqs'].size} frequencies, FP64 NumPy, {dt:.1f} s, host {platform.processor() or platform.machine()} "
                      f"({os.cpu_count()} logical CPUs visible)"}, idx, t_mid, rf


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
    ap.add_argument("--copy-calib-mib", type=int, default=0,
                    help="only run a device-to-device copy of this many MiB (calibrates the FETCH_SIZE / WRITE_SIZE "
                         "counters of a profiler against a known byte count) and exit")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
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

    def step():
        sp = K.run_spectrum(prob, "mfma")
        rf, _, _ = K.run_rf(sp, plan.n_freq_full, plan.freq_idx_start, plan.n_fft, want_full=False)
        return sp, rf

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    lib.fs_set_kernel_timing(1)
    _lib.kernel_time_ms()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        sp_m, rf_m = step()
    torch.cuda.synchronize()
    t_step = (time.perf_counter() - t0) / args.steps
    k_ms, k_n = _lib.kernel_time_ms()
    lib.fs_set_kernel_timing(0)
    evals = T * S * E * nf
    value = evals / t_step / 1e9
    k_avg = k_ms / max(k_n, 1) / 1e3
    flops = msm_flops_per_launch(S, E, nf)
    roof = {"bound": "mfma", "achieved": flops / k_avg / 1e12, "peak": PEAK_BF16 / 1e12, "unit": "TFLOP/s",
            "frac": flops / k_avg / PEAK_BF16, "traffic": measured_traffic(S, T, nf, E),
            "kernel": "k_msm (split-BF16 v_mfma_f32_16x16x32_bf16)", "kernel_ms": k_avg * 1e3, "launches": k_n}

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
                          "kernel": "k_tx + k_rx (FP32 VALU)", "kernel_ms_total": kp_ms, "launches": kp_n}}
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
        "config": {"workload": "batched plane-wave compounding: 128-element 7.6 MHz linear array, 75 plane waves "
                               "in [-16,16] deg, 200000 speckle scatterers, 0.5 dB/cm/MHz",
                   "n_scat": S, "n_elem": E, "n_tx": T, "n_freq": nf, "n_sub": plan.n_sub, "n_fft": plan.n_fft,
                   "method": "hip_mfma"},
        "roofline": roof, "plain_fp32": plain, "speedup_vs_plain": value / plain_value, "rf_rel_l2": err,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
