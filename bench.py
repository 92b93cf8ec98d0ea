"""Benchmark of the Limber C_ell path on one MI355X.

Default workload (one "step" = one batched pass):
  1024 cosmologies drawn with numpy default_rng(0) uniform in
  Omega_c [0.2, 0.35], Omega_b [0.04, 0.06], h [0.6, 0.8], n_s [0.9, 1.0],
  sigma8 [0.7, 0.9], w0 [-1.3, -0.7], wa [-0.5, 0.5] (Omega_k = 0);
  10 tracers: 5 Smail weak-lensing bins (a=1, b=2, z0 in 0.5..1.5) and 5
  Gaussian number-count bins (mean 0.3..1.5, sigma_z 0.1, b = 1.5),
  10 gal/arcmin^2, sigma_e 0.26; 100 log-spaced ell in [10, 5000];
  512 Simpson intervals over a.  Output [1024, 55, 100] float32 C_ell.
Inputs are resident in HBM (torch CUDA tensors) before the timed region.

Also reported: the plain single-request drop-in (angular_cl on one
cosmology, host arrays in and out) throughput and p50/p99 latency, the
batched/plain ratio, the relative error against the FP64 oracle on sampled
cosmologies, the roofline of the dominant kernel (timed with HIP events on
its launch stream) and a CPU baseline (the FP64 oracle on a bounded sample).
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

BOX_LO = np.array([0.2, 0.04, 0.6, 0.9, 0.7, -1.3, -0.5])
BOX_HI = np.array([0.35, 0.06, 0.8, 1.0, 0.9, -0.7, 0.5])

# MI355X: 256 CUs x 64 FP32 lanes x 2.4 GHz peak engine clock.  One VALU
# lane-op per lane per cycle (non-packed FP32; packed FP32 math is not used);
# the published 157.3 TFLOP/s FP32 figure is this rate x 2 (FMA) x 2 (packed).
PEAK_LANE_OPS = 256 * 64 * 2.4e9
# Algorithmic VALU lane-ops per (cosmology, ell, a) point of the Limber
# kernel: EH98 123 + linear 12 + halofit 85 + k and weight 4 = 224, plus one
# FMA per tracer pair for the contraction.
OPS_PER_POINT = 224


def survey_probes(jc):
    wl = jc.probes.WeakLensing(
        [jc.redshift.smail_nz(1.0, 2.0, z0, gals_per_arcmin2=10.0) for z0 in (0.5, 0.75, 1.0, 1.25, 1.5)],
        sigma_e=0.26)
    nc = jc.probes.NumberCounts(
        [jc.redshift.gaussian_nz(mu, 0.1, gals_per_arcmin2=10.0) for mu in np.linspace(0.3, 1.5, 5)],
        jc.bias.constant_linear_bias(1.5))
    return [wl, nc]


def oracle_probes(O):
    wl = O.WeakLensing([O.smail_nz(1.0, 2.0, z0, gals_per_arcmin2=10.0) for z0 in (0.5, 0.75, 1.0, 1.25, 1.5)],
                       sigma_e=0.26)
    nc = O.NumberCounts([O.gaussian_nz(mu, 0.1, gals_per_arcmin2=10.0)
                         for mu in np.linspace(0.3, 1.5, 5).tolist()], 1.5)
    return [wl, nc]


def draw_cosmologies(B, seed=0):
    rng = np.random.default_rng(seed)
    th = BOX_LO + (BOX_HI - BOX_LO) * rng.random((B, 7))
    params = np.zeros((B, 8))
    params[:, [0, 1, 2, 3, 4, 6, 7]] = th
    return params


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--batch", type=int, default=1024)
    ap.add_argument("--plain-calls", type=int, default=200)
    ap.add_argument("--oracle-samples", type=int, default=2)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (every box has one GPU); "
              "--gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch

    import jax_cosmo_amd as jc
    from jax_cosmo_amd import _native

    dev = torch.device("cuda:0")
    probes = survey_probes(jc)
    ell = np.geomspace(10.0, 5000.0, 100)
    n_z = 512
    B = args.batch
    params = draw_cosmologies(B, 0)
    T = 10
    P = T * (T + 1) // 2
    n_cl = B * P * len(ell)

    ctx = _native.default_context()
    cp_dev = torch.tensor(params, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()

    def step():
        return jc.cl.angular_cl_batch(cp_dev, ell, probes, n_z=n_z)

    for _ in range(args.warmup):
        out = step()
    torch.cuda.synchronize()
    ctx.profile(True)
    ctx.profile_read()  # reset
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out = step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    prof = ctx.profile_read()
    ctx.profile(False)
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    value = n_cl / ((t1 - t0) / args.steps)

    lim_ms, lim_n = prof["limber"]
    lim_avg_ms = lim_ms / max(lim_n, 1)
    Na = n_z + 1
    ops_launch = float(B) * Na * len(ell) * (OPS_PER_POINT + P)
    achieved = ops_launch / (lim_avg_ms * 1e-3)
    roofline = {
        "bound": "valu",
        "kernel": "k_limber<10,0> (EH98 + halofit + pair contraction)",
        "achieved": achieved / 1e12,
        "peak": PEAK_LANE_OPS / 1e12,
        "unit": "TFLOP/s",
        "unit_note": "10^12 FP32 VALU lane-op slots/s (transcendental = 4 slots)",
        "frac": achieved / PEAK_LANE_OPS,
        "traffic": None,
        "avg_launch_ms": lim_avg_ms,
        "ops_per_launch": ops_launch,
        "kernel_ms": {k: (v[0] / max(v[1], 1)) for k, v in prof.items()},
    }

    # ---------------------------------------------------------------- plain drop-in
    cosmo0 = jc.Cosmology(*params[0])
    jc.cl.angular_cl(cosmo0, ell, probes, n_z=n_z)
    lat = []
    for i in range(args.plain_calls):
        c = jc.Cosmology(*params[i % B])
        ts = time.perf_counter()
        jc.cl.angular_cl(c, ell, probes, n_z=n_z)
        lat.append(time.perf_counter() - ts)
    lat = np.array(lat)
    plain_value = P * len(ell) / lat.mean()
    plain = {"value": plain_value, "unit": "C_ell/s", "p50_ms": float(np.percentile(lat, 50) * 1e3),
             "p99_ms": float(np.percentile(lat, 99) * 1e3), "calls": int(args.plain_calls),
             "what": "angular_cl(cosmo, ell, probes) one cosmology per call, host arrays"}

    # ---------------------------------------------------------------- parity sample
    from oracle import jc_oracle as O

    ncpu = len(os.sched_getaffinity(0))
    torch.set_num_threads(ncpu)
    oprobes = oracle_probes(O)
    out_np = out.float().cpu().numpy()
    errs = []
    for b in range(min(args.oracle_samples, B)):
        ref = O.angular_cl(O.Cosmo(*params[b]), torch.tensor(ell), oprobes, N=n_z).numpy()
        errs.append(float(np.max(np.abs(out_np[b] - ref) / np.abs(ref))))

    # ---------------------------------------------------------------- cpu baseline
    cpu = None
    if not args.no_cpu_baseline:
        done = 0
        tc0 = time.perf_counter()
        while True:
            O.angular_cl(O.Cosmo(*params[done % B]), torch.tensor(ell), oprobes, N=n_z)
            done += 1
            if time.perf_counter() - tc0 > args.cpu_seconds:
                break
        tc = time.perf_counter() - tc0
        cpu = {"value": done * P * len(ell) / tc, "unit": "C_ell/s", "cores": torch.get_num_threads(),
               "kind": "port",
               "sample": "%d cosmologies of the same 1024-cosmology workload through the FP64 "
                         "reference-faithful oracle (torch CPU, %d threads), %.1f s" % (done, torch.get_num_threads(), tc)}

    line = {
        "metric": "C_ell/s",
        "value": value,
        "unit": "C_ell/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "batched C_ell: %d cosmologies x 10 tracers (5 WL Smail + 5 NC Gaussian) "
                               "x 100 ell x 512 z intervals, f32 with f64 background/growth/halofit root" % B,
                   "batch": B, "tracers": T, "pairs": P, "n_ell": len(ell), "n_z": n_z},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "plain": plain,
        "batched_vs_plain": value / plain_value,
        "rel_err_vs_fp64_oracle": {"max": max(errs) if errs else None, "samples": len(errs)},
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
