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
# Algorithmic work per (cosmology, ell, a) point of the Limber integrand, in
# FP32 VALU lane-ops (add/mul/fma/compare = 1; div, exp, log, sin, sqrt = 4;
# non-integer pow = 9) for the minimal direct formulation: EH98 with wiggles
# 123 + linear power 12 + halofit 85 + k and Limber weight 4 = 224, plus one
# FMA per tracer pair for the contraction.  The dominant kernel performs the
# linear power, halofit, k / weight and contraction pieces (101 + pairs); the
# EH98 transfer function is evaluated once per node of a per-cosmology table
# (1024 nodes, built at the end of the table kernel) and only interpolated per
# point, so its 123 are charged to the table build, not to this kernel.
# `roofline.achieved` = (101 + pairs) x points per launch / launch time.
OPS_PER_POINT = 224
OPS_PER_POINT_LIMBER = 12 + 85 + 4
# What the default kernel executes per point (per-cosmology transfer table in
# place of EH98, halofit per-a coefficients loaded), same weights:
#   k, 1/k, ln k                                   3
#   table position (FP64 fma, clamp, fraction)     8
#   cubic Hermite piece + P_lin = T^2 D^2 A k^ns   5
#   k^3, k^-3, log2 y, y, 1/y                      7
#   halofit: 4 exp2 + 1 log2 + 1 shared reciprocal (6 x 4 = 24) + 25 simple   49
#   quadrature weight                              1
# = 73, plus per tracer one multiply and per pair one FMA: 138 on the
# 10-tracer workload.  The compiled loop body holds 124 VALU instructions per
# point, 6 of them transcendental (142 slots).  `executed_frac` uses it.
EXEC_OPS_PER_POINT = 73
# HBM bytes per k_limber launch on this workload, from two rocprofv3 --pmc runs
# of `bench.py --batch-only` (FETCH_SIZE 48212 KiB x 2.0 + WRITE_SIZE 34913
# KiB x 1.0 per dispatch; the factors make a plain 256 MiB copy read the
# counters as 256 MiB).  A counter run cannot be live inside the timed region.
LIMBER_HBM_BYTES = (48212 * 2.0 + 34913 * 1.0) * 1024
# VALU busy of k_limber from the same counters: SQ_ACTIVE_INST_VALU x 4 /
# 1024 SIMDs / (GRBM_GUI_ACTIVE / 8 XCDs).
LIMBER_VALU_BUSY = 109.8064e6 * 4 / 1024 / (4.2190e6 / 8)


def progress(msg):
    """One line per finished leg on stderr (the JSON result stays the only stdout line)."""
    print("bench: %s" % msg, file=sys.stderr, flush=True)


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


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baselines(O, jc, args, params, ell, n_z):
    """The FP64 oracle (reference-faithful mode: naive [Na, 257, 256] sigma^2
    tensor, argmin interpolation, per-z lensing Simpson) on this host's cores,
    one bounded sample per workload: C_ell/s (gradients/s for the likelihood
    gradient) and the per-call latency."""
    import torch

    secs = args.cpu_config_seconds
    out = {"cpu_model": cpu_model(), "threads": torch.get_num_threads(),
           "what": "FP64 reference-faithful restatement (torch CPU), not JAX"}

    def timed(fn, n_cl):
        lat = []
        t0 = time.perf_counter()
        while True:
            ts = time.perf_counter()
            fn(len(lat))
            lat.append(time.perf_counter() - ts)
            if time.perf_counter() - t0 > secs:
                break
        span = time.perf_counter() - t0
        return {"value": len(lat) * n_cl / span, "calls": len(lat), "seconds": span,
                "p50_ms": float(np.percentile(np.array(lat) * 1e3, 50))}

    # C1: Planck15, one Smail bin (1, 2, 1; zmax 4), 50 ell in [10, 3000], N = 256
    p15 = [0.2589, 0.0486, 0.6774, 0.9667, 0.8159, 0.0, -1.0, 0.0]
    ell1 = torch.tensor(np.geomspace(10.0, 3000.0, 50))
    wl1 = [O.WeakLensing([O.smail_nz(1.0, 2.0, 1.0, zmax=4.0)])]
    out["C1"] = dict(timed(lambda i: O.angular_cl(O.Cosmo(*p15), ell1, wl1, N=256), 50), unit="C_ell/s")
    # C2: Planck15, the 10-tracer survey, 100 ell, N = 512
    oprobes = oracle_probes(O)
    ell_t = torch.tensor(ell)
    out["C2"] = dict(timed(lambda i: O.angular_cl(O.Cosmo(*p15), ell_t, oprobes, N=n_z), 5500), unit="C_ell/s")
    # C4: gradient of the Gaussian log-likelihood (fixed sparse covariance at
    # Planck15, data = mean + L eps) by torch autograd, per cosmology
    mu = O.angular_cl(O.Cosmo(*p15), ell_t, oprobes, N=n_z).numpy()
    cov = O.gaussian_cl_covariance(ell_t, oprobes, torch.tensor(mu), O.noise_cl(ell_t, oprobes), f_sky=0.25,
                                   sparse=True)
    L = np.linalg.cholesky(np.transpose(cov.numpy(), (2, 0, 1)))
    eps = np.random.default_rng(2).normal(size=(len(ell), 55))
    data = torch.tensor((mu + np.einsum("lij,lj->il", L, eps)).flatten())
    gp = draw_cosmologies(64, 1)

    def grad_one(i):
        pv = torch.tensor(gp[i % 64], dtype=torch.float64, requires_grad=True)
        m = O.angular_cl(O.Cosmo.from_vector(pv), ell_t, oprobes, N=n_z).flatten()
        O.gaussian_log_likelihood(data, m, cov, include_logdet=False).backward()

    out["C4"] = dict(timed(grad_one, 1), unit="gradients/s")
    # C5: the request stream served one request after another (oracle forward,
    # plus autograd of <cl_bar, C_ell> for gradient requests)
    from jax_cosmo_amd import serving

    reqs = serving.make_stream(64, seed=3)
    zs = (0.5, 0.75, 1.0, 1.25, 1.5)
    mus = np.linspace(0.3, 1.5, 5).tolist()

    def req_probes(idx):
        wl = [O.smail_nz(1.0, 2.0, zs[i], gals_per_arcmin2=10.0) for i in idx if i < 5]
        nc = [O.gaussian_nz(mus[i - 5], 0.1, gals_per_arcmin2=10.0) for i in idx if i >= 5]
        return ([O.WeakLensing(wl, sigma_e=0.26)] if wl else []) + ([O.NumberCounts(nc, 1.5)] if nc else [])

    n_done = [0, 0]

    def serve_one(i):
        r = reqs[i % len(reqs)]
        T = len(r["tracers"])
        pv = torch.tensor(r["cosmo"], dtype=torch.float64, requires_grad=r["cl_bar"] is not None)
        c = O.angular_cl(O.Cosmo.from_vector(pv), torch.tensor(r["ell"]), req_probes(r["tracers"]), N=r["n_z"])
        if r["cl_bar"] is not None:
            (c * torch.tensor(r["cl_bar"], dtype=torch.float64)).sum().backward()
        n_done[0] += T * (T + 1) // 2 * len(r["ell"])

    res = timed(serve_one, 1)
    res["value"] = n_done[0] / res["seconds"]
    out["C5"] = dict(res, unit="C_ell/s", note="requests of the seeded stream (cosmology, 1-10 bins, 20-200 ell, "
                                               "n_z 256/512/1024, 30 % with the 7-parameter gradient), serially")
    return out


def batch_shaped_run(srv, reps=4):
    """The batch workload's shape as served requests (all 10 bins, the same
    100 ell, n_z 512, forward only), 2048 queued at once: the server's cost
    against the batch path for the same work per cosmology.  The first of
    reps passes is a warm-up; the median pass of the others is reported."""
    import time as _t

    ell100 = np.geomspace(10.0, 5000.0, 100)
    cps = draw_cosmologies(2048, 0)
    full = [dict(cosmo=cps[i], tracers=list(range(10)), ell=ell100, n_z=512, cl_bar=None)
            for i in range(len(cps))]
    prep_full = srv.prepare(full)
    passes = []
    for rep in range(reps):
        t0 = _t.perf_counter()
        srv.submit_prepared(prep_full)
        t_sub = _t.perf_counter() - t0
        done = 0
        lat = []
        while done < len(full):
            c = srv.poll_array()
            if c.size and (c["status"] != 0).any():
                raise RuntimeError("served request failed: %d" % int(c["status"][c["status"] != 0][0]))
            lat.append(c["latency_us"])
            done += c.size
        span = _t.perf_counter() - t0
        lat = np.concatenate(lat) / 1e3
        if rep:
            passes.append((span, t_sub, float(lat.min()), float(lat.max())))
    passes.sort()
    span, t_sub, first, last = passes[len(passes) // 2]
    return {"C_ell_per_s": len(full) * 55 * 100 / span, "requests": len(full), "span_s": span,
            "spans_s": sorted(p[0] for p in passes),
            "submit_ms": t_sub * 1e3, "first_done_ms": first, "last_done_ms": last,
            "what": "2048 forward requests of the batch workload's shape (10 bins, 100 ell, n_z 512) "
                    "queued at once; the median of %d passes after a warm-up; submit_ms: its bulk "
                    "submission call; first/last_done_ms: its completion times after submission "
                    "(server clock)" % len(passes)}


def serve_server(probes, args):
    from jax_cosmo_amd import serving

    return serving.Server(probes, max_batch=512, max_grad=160, max_ell=200, slots=args.serve_slots,
                          n_z=(256, 512, 1024))


def serve_leg(jc, probes, args):
    """Mixed request stream (seeded): cosmology from the box, 1..10 bins of
    the 10-bin pool, 20..200 log-spaced ell, n_z in {256, 512, 1024}, 30 %
    with the 7-parameter adjoint.  Plain: each request alone through the
    drop-in entry points.  Served: the continuous-batching server, first with
    every request queued at once (capacity), then Poisson arrivals at
    fractions of that capacity."""
    import time as _t

    from jax_cosmo_amd import serving

    ctx = jc._native.default_context()
    reqs = serving.make_stream(args.serve_requests, seed=3)
    pool = jc.cl.tracer_entries(probes)
    n_cl = [len(r["tracers"]) * (len(r["tracers"]) + 1) // 2 * len(r["ell"]) for r in reqs]
    total_cl = float(sum(n_cl))

    def subset_set(idx):
        es = [dict(pool[i]) for i in idx]
        return ctx.tracer_set(es)

    # plain per-request launches on a bounded prefix of the stream.  Cold:
    # the first pass, which registers each request's tracer subset and builds
    # its resident n(z) / lensing tables for its Simpson size on the host
    # inside the timed loop.  Warm: a second pass over the same requests with
    # every table resident, the like-for-like baseline of the served rate.
    n_plain = min(len(reqs), 512)

    def plain_pass():
        lat = []
        t0 = _t.perf_counter()
        for r in reqs[:n_plain]:
            ts = _t.perf_counter()
            sid = subset_set(r["tracers"])
            P = len(r["tracers"]) * (len(r["tracers"]) + 1) // 2
            ctx.angular_cl(r["cosmo"], sid, r["ell"], r["n_z"], P)
            if r["cl_bar"] is not None:
                ctx.angular_cl_vjp(r["cosmo"], sid, r["ell"], r["n_z"], r["cl_bar"][None])
            lat.append(_t.perf_counter() - ts)
        span = _t.perf_counter() - t0
        lat = np.array(lat) * 1e3
        return {"C_ell_per_s": sum(n_cl[:n_plain]) / span, "p50_ms": float(np.percentile(lat, 50)),
                "p99_ms": float(np.percentile(lat, 99)), "requests_timed": n_plain}

    plain_cold = plain_pass()
    plain_warm = plain_pass()
    plain_rate = plain_warm["C_ell_per_s"]

    out = {"requests": len(reqs), "C_ell_total": total_cl, "plain": plain_warm, "plain_cold": plain_cold,
           "plain_note": "one request per call through jc_angular_cl (+ jc_angular_cl_vjp for gradient "
                         "requests), host arrays; warm = every tracer subset's tables already resident "
                         "(the like-for-like baseline); cold = first pass, building them on the host"}
    with serve_server(probes, args) as srv:
        # the request descriptors and their output arrays are built before the
        # timed runs; a run submits them in bulk (all at once for the capacity
        # run, every request whose arrival time has passed for the Poisson runs)
        prep = srv.prepare(reqs)
        is_grad = np.array([r["cl_bar"] is not None for r in reqs])

        def run(rate):
            sched = None
            if rate:
                gaps = np.random.default_rng(4).exponential(1.0 / rate, size=len(reqs))
                sched = np.cumsum(gaps)
            st0 = srv.stats()
            lat, grad = [], []
            off = None  # request index = completion id - off (this thread's ids are consecutive)
            i = 0
            done = 0
            t_start = _t.perf_counter()
            while done < len(reqs):
                now = _t.perf_counter() - t_start
                j = len(reqs) if sched is None else int(np.searchsorted(sched, now, side="right"))
                if j > i:
                    first = int(srv.submit_prepared(prep, i, j)[0])
                    if off is None:
                        off = first - i
                    assert first - i == off
                    i = j
                c = srv.poll_array()
                if c.size:
                    if (c["status"] != 0).any():
                        raise RuntimeError("served request failed: %d" % int(c["status"][c["status"] != 0][0]))
                    lat.append(c["latency_us"])
                    grad.append(is_grad[c["id"] - off])
                    done += c.size
            span = _t.perf_counter() - t_start
            lat = np.concatenate(lat)
            grad = np.concatenate(grad)
            la = lat / 1e3
            lf = lat[~grad] / 1e3
            lg = lat[grad] / 1e3
            st1 = srv.stats()
            nd = max(st1["done"] - st0["done"], 1.0)
            return {"mean_wait_ms": (st1["wait_ms"] - st0["wait_ms"]) / nd,
                    "mean_service_ms": (st1["service_ms"] - st0["service_ms"]) / nd,
                    "batches": st1["batches"] - st0["batches"],
                    "mean_batch_device_ms": (st1["busy_ms"] - st0["busy_ms"]) / max(st1["batches"] - st0["batches"], 1),
                    "mean_batch_launch_ms": (st1["launch_ms"] - st0["launch_ms"]) / max(st1["batches"] - st0["batches"], 1),
                    "C_ell_per_s": total_cl / span, "p50_ms": float(np.percentile(la, 50)),
                    "p99_ms": float(np.percentile(la, 99)), "span_s": span,
                    "forward_only_p50_ms": float(np.percentile(lf, 50)) if lf.size else None,
                    "gradient_p50_ms": float(np.percentile(lg, 50)) if lg.size else None}

        run(None)  # warm-up (grid tables, graphs, caches)
        srv.stats()
        cap = run(None)
        st = srv.stats()
        cap["mean_batch"] = st["mean_batch"]
        out["capacity"] = cap
        req_rate = len(reqs) / cap["span_s"]
        out["poisson"] = []
        for frac in args.serve_fractions:
            res = run(frac * req_rate)
            res["arrival_rate_req_per_s"] = frac * req_rate
            res["fraction_of_capacity"] = frac
            out["poisson"].append(res)
        out["server_stats"] = srv.stats()

        out["batch_shaped"] = batch_shaped_run(srv)
    out["served_vs_plain"] = cap["C_ell_per_s"] / plain_rate
    out["served_vs_plain_cold"] = cap["C_ell_per_s"] / plain_cold["C_ell_per_s"]
    return out


def gradient_leg(jc, probes, ell, n_z, args):
    """Gradient of the Gaussian log-likelihood over the 7 parameters for a
    batch of seeded cosmologies: mean from the device, fixed sparse Gaussian
    covariance at Planck15 (f_sky 0.25, sigma_e 0.26, 10 gal/arcmin^2),
    data = fiducial mean + seeded Gaussian noise; the block inverses, the
    quadratic forms, their cotangent C^-1 (data - mu) and the adjoint all
    run on the device."""
    import time as _t

    import torch

    ctx = jc._native.default_context()
    sid = ctx.tracer_set(jc.cl.tracer_entries(probes))
    P = 55
    fid = np.array([0.2589, 0.0486, 0.6774, 0.9667, 0.8159, 0.0, -1.0, 0.0])
    mu = ctx.angular_cl(fid, sid, ell, n_z, P)[0].astype(np.float64)
    cov = jc.cl.gaussian_cl_covariance(ell, probes, mu, jc.cl.noise_cl(ell, probes), f_sky=0.25, sparse=True)
    L = np.linalg.cholesky(np.transpose(cov, (2, 0, 1)))  # [n_ell, P, P]
    eps = np.random.default_rng(2).normal(size=(len(ell), P))
    data = mu + np.einsum("lij,lj->il", L, eps)  # [P, n_ell]
    B = args.grad_batch
    rng = np.random.default_rng(1)
    th = BOX_LO + (BOX_HI - BOX_LO) * rng.random((B, 7))
    params = np.zeros((B, 8))
    params[:, [0, 1, 2, 3, 4, 6, 7]] = th
    dev = torch.device("cuda:0")
    prec, _ = ctx.sparse_inv(cov)  # device block inverses (FP64)
    prec_t = torch.tensor(prec, dtype=torch.float64, device=dev)
    data_t = torch.tensor(data.flatten(), dtype=torch.float64, device=dev)
    cp = torch.tensor(params, dtype=torch.float64, device=dev)
    el = torch.tensor(ell, dtype=torch.float64, device=dev)
    ll_out = torch.empty(B, dtype=torch.float64, device=dev)
    th_out = torch.empty((B, 7), dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream(dev).cuda_stream

    def step():
        # forward C_ell, -1/2 r^T C^-1 r and its cotangent, adjoint: one C call
        ctx.gaussian_loglike_grad_dev(cp.data_ptr(), B, sid, el.data_ptr(), len(ell), n_z, data_t.data_ptr(),
                                      prec_t.data_ptr(), 0.0, ll_out.data_ptr(), th_out.data_ptr(), stream=stream)
        return th_out

    for _ in range(3):  # warm-up (clocks, caches)
        step()
    torch.cuda.synchronize()
    t0 = _t.perf_counter()
    steps = 20
    for _ in range(steps):
        g = step()
    torch.cuda.synchronize()
    dt = (_t.perf_counter() - t0) / steps
    g = g.cpu().numpy()
    out = {"batch": B, "ms_per_batch": dt * 1e3, "gradients_per_s": B / dt,
           "finite": bool(np.all(np.isfinite(g))),
           "what": "log-likelihood -1/2 r^T C^-1 r and its gradient w.r.t. theta = (Omega_c, Omega_b, h, n_s, "
                   "sigma8, w0, wa), one jc_gaussian_loglike_grad call (forward, likelihood kernels, adjoint)"}
    # the same covariance as one dense 5500 x 5500 matrix (the default form of
    # gaussian_cl_covariance_and_mean): FP64 Cholesky once, then the same
    # gradient through the factor
    Cd = jc.sparse.to_dense(cov)
    ctx.dense_cholesky(Cd)  # warm-up
    torch.cuda.synchronize()
    t0 = _t.perf_counter()
    L, _ = ctx.dense_cholesky(Cd)
    t_chol = _t.perf_counter() - t0
    data_d = data.flatten()
    ctx.gaussian_loglike_grad_dense(params, sid, ell, n_z, data_d, L)
    torch.cuda.synchronize()
    t0 = _t.perf_counter()
    for _ in range(5):
        ctx.gaussian_loglike_grad_dense(params, sid, ell, n_z, data_d, L)
    t_dense = (_t.perf_counter() - t0) / 5
    out["dense"] = {"N": int(Cd.shape[0]), "cholesky_ms": t_chol * 1e3, "ms_per_batch": t_dense * 1e3,
                    "what": "dense covariance: FP64 Cholesky of the N x N matrix (host input, copied in and the "
                            "factor left on the device), then jc_gaussian_loglike_grad_dense for the batch "
                            "(host cosmologies in, log-likelihoods and gradients out)"}
    return out


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
    ap.add_argument("--cpu-config-seconds", type=float, default=4.0,
                    help="CPU seconds per workload of the per-workload CPU baselines")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-grad", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=4096)
    ap.add_argument("--serve-slots", type=int, default=2, help="batches in flight of the served stream")
    ap.add_argument("--serve-fractions", type=float, nargs="*", default=[0.25, 0.5, 1.0],
                    help="Poisson arrival rates of the served stream, as fractions of its measured capacity")
    ap.add_argument("--grad-batch", type=int, default=256)
    ap.add_argument("--tune", action="store_true", help="re-run the tile autotuner before timing")
    ap.add_argument("--tune-copy", default="", metavar="PATH",
                    help="with --tune, also copy the tuned tile cache to PATH")
    ap.add_argument("--batch-only", action="store_true",
                    help="only the timed batched steps (for profiler runs)")
    ap.add_argument("--serve-only", action="store_true",
                    help="only the served request stream (for profiler runs)")
    ap.add_argument("--serve-batch-shaped-only", action="store_true",
                    help="only the served requests of the batch workload's shape (for profiler runs)")
    ap.add_argument("--grad-only", action="store_true",
                    help="only the timed likelihood-gradient calls (for profiler runs)")
    ap.add_argument("--matrix-cores", action="store_true",
                    help="run the forward Limber sums on the FP32 matrix cores (k_limber_mx, off by default; "
                         "for its counters)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB five times (scales the "
                         "FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (every box has one GPU); "
              "--gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch

    if args.copy_calibration:
        n = args.copy_calibration * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": n * 4, "copies": 5}))
        return

    import jax_cosmo_amd as jc
    from jax_cosmo_amd import _native

    dev = torch.device("cuda:0")
    probes = survey_probes(jc)
    ell = np.geomspace(10.0, 5000.0, 100)
    n_z = 512
    if args.grad_only:
        print(json.dumps({"likelihood_gradient": gradient_leg(jc, probes, ell, n_z, args)}))
        return
    if args.serve_batch_shaped_only:
        with serve_server(probes, args) as srv:
            print(json.dumps({"batch_shaped": batch_shaped_run(srv, reps=6)}))
        return
    if args.serve_only:
        print(json.dumps({"served": serve_leg(jc, probes, args)}))
        return
    B = args.batch
    params = draw_cosmologies(B, 0)
    T = 10
    P = T * (T + 1) // 2
    n_cl = B * P * len(ell)

    ctx = _native.default_context()
    if args.matrix_cores:
        ctx.set_forward_options(matrix_cores=1)
    tuned = None
    if args.tune:
        from jax_cosmo_amd import autotune

        tuned = {"batch": autotune.tune_batch(probes, B, len(ell), n_z)}
        if not args.no_serve:
            tuned["server"] = autotune.tune_server(probes)
        if args.tune_copy:
            import shutil

            shutil.copy(autotune.cache_path(), args.tune_copy)
    cp_dev = torch.tensor(params, dtype=torch.float64, device=dev)
    ell_dev = torch.tensor(ell, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()

    def step():
        return jc.cl.angular_cl_batch(cp_dev, ell_dev, probes, n_z=n_z)

    def timed_steps(n, kernels):
        """n steps with HIP events around the launches of `kernels`: (ms per
        step, {kernel: (total_ms, launches)})."""
        ctx.profile(True, kernels=kernels)
        ctx.profile_read()  # reset
        torch.cuda.synchronize()
        ta = time.perf_counter()
        for _ in range(n):
            step()
        torch.cuda.synchronize()
        tb = time.perf_counter()
        p = ctx.profile_read()
        ctx.profile(False)
        return (tb - ta) * 1e3 / n, p

    for _ in range(args.warmup):
        out = step()
    torch.cuda.synchronize()
    # the timed region: events around the Limber launches only (the roofline's
    # kernel); each event record adds to the stream's time
    ctx.profile(True, kernels=("limber",))
    ctx.profile_read()  # reset
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out = step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    prof = ctx.profile_read()
    ctx.profile(False)
    tiles_c3 = ctx.get_tiles()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    value = n_cl / ((t1 - t0) / args.steps)
    # after the timed region: every kernel's time over as many steps, then
    # the same measurement as the timed region over 400 more steps (the first
    # ~10 ms of a process's GPU work run slower: the Limber kernel takes
    # ~0.245 ms per launch there and ~0.212 ms after a few hundred steps)
    _, prof_all = timed_steps(args.steps, None)
    steady_ms, prof_steady = timed_steps(400, ("limber",))
    steady = {"value": n_cl / (steady_ms * 1e-3), "unit": "C_ell/s", "ms_per_step": steady_ms, "steps": 400,
              "limber_avg_ms": prof_steady["limber"][0] / max(prof_steady["limber"][1], 1),
              "what": "the headline step timed the same way over 400 further steps, after the timed "
                      "region; not the headline value"}

    lim_ms, lim_n = prof["limber"]
    lim_avg_ms = lim_ms / max(lim_n, 1)
    Na = n_z + 1
    ops_launch = float(B) * Na * len(ell) * (OPS_PER_POINT_LIMBER + P)
    achieved = ops_launch / (lim_avg_ms * 1e-3)
    direct_launch = float(B) * Na * len(ell) * (OPS_PER_POINT + P)
    exec_launch = float(B) * Na * len(ell) * (EXEC_OPS_PER_POINT + T + P)
    # whole step: the per-point work plus the per-cosmology tables (n_eff/C
    # 257 Na 14, sigma^2(R) 257 x 256 x 8, lensing 40 and clustering 20 per
    # bin and node, background tables 60 k), the same lane-op weights
    step_ops = float(B) * (Na * len(ell) * (OPS_PER_POINT + P) + 257 * Na * 14 + 257 * 256 * 8
                           + 5 * Na * 40 + 5 * Na * 20 + 60e3)
    roofline = {
        "bound": "valu",
        "kernel": "k_limber<10,0,table,prefetch> (transfer table + halofit + pair contraction)",
        "achieved": achieved / 1e12,
        "peak": PEAK_LANE_OPS / 1e12,
        "unit": "TFLOP/s",
        "unit_note": "10^12 FP32 VALU lane-ops/s (one per lane per clock; div and transcendentals "
                     "weigh 4); algorithmic ops per launch = B x (n_z+1) x n_ell x (101 + pairs): the "
                     "linear power (12), halofit (85), k and weight (4) and contraction (1 per pair) "
                     "pieces of the per-point count, which this kernel performs; the EH98 piece (123 "
                     "per point in the direct form) is done once per node of the per-cosmology "
                     "transfer table in the table kernel",
        "frac": achieved / PEAK_LANE_OPS,
        "direct_formulation_ops_per_launch": direct_launch,
        "direct_formulation_note": "B x (n_z+1) x n_ell x (224 + pairs): EH98 evaluated at every point; "
                                   "this over the launch time is what the kernel would have to sustain "
                                   "to match it without the transfer table",
        "executed_ops_per_launch": exec_launch,
        "executed_frac": exec_launch / (lim_avg_ms * 1e-3) / PEAK_LANE_OPS,
        "traffic": LIMBER_HBM_BYTES,
        "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, calibrated)",
        "traffic_rate_GBps": LIMBER_HBM_BYTES / (lim_avg_ms * 1e-3) / 1e9,
        "valu_busy_pmc": LIMBER_VALU_BUSY,
        "avg_launch_ms": lim_avg_ms,
        "ops_per_launch": ops_launch,
        "kernel_ms": {k: (v[0] / max(v[1], 1)) for k, v in prof_all.items()},
        "kernel_ms_note": "per launch, over a pass of the same number of steps after the timed region "
                          "with events around every kernel; avg_launch_ms is from the timed region",
        "step_ops": step_ops,
        "step_frac": step_ops / (ms_per_step * 1e-3) / PEAK_LANE_OPS,
    }

    # ---------------------------------------------------------------- one HIP graph per step
    # the same batch captured once (the kernels of jc_angular_cl_noalloc) and
    # replayed per step: one graph launch instead of three kernel launches
    graph = jc.cl.BatchGraph(cp_dev, ell_dev, probes, n_z=n_z)
    for _ in range(args.warmup):
        graph.run()
    torch.cuda.synchronize()
    tg0 = time.perf_counter()
    for _ in range(args.steps):
        gout = graph.run()
    torch.cuda.synchronize()
    tg = (time.perf_counter() - tg0) / args.steps
    graph_leg = {"value": n_cl / tg, "unit": "C_ell/s", "ms_per_step": tg * 1e3,
                 "equal_to_eager": bool(torch.equal(gout, out)),
                 "what": "the same 1024-cosmology batch replayed from one captured HIP graph per step"}
    del graph
    progress("batch leg done: %.3f ms per step, graph replay %.3f ms" % (ms_per_step, tg * 1e3))

    if args.batch_only:
        print(json.dumps({"metric": "C_ell/s", "value": value, "unit": "C_ell/s", "ms_per_step": ms_per_step,
                          "batch_only": True, "kernel_ms": roofline["kernel_ms"], "graph_replay": graph_leg,
                          "steady_state": steady}))
        return

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

    # ---------------------------------------------------------------- same batch, host buffers
    # the PCIe-inclusive rate: cosmologies from host memory, C_ell back into a
    # host array (not the headline value, which keeps inputs and outputs in HBM)
    jc.cl.angular_cl_batch(params, ell, probes, n_z=n_z)
    th0 = time.perf_counter()
    n_host = 5
    for _ in range(n_host):
        jc.cl.angular_cl_batch(params, ell, probes, n_z=n_z)
    th = (time.perf_counter() - th0) / n_host
    host_batch = {"value": n_cl / th, "unit": "C_ell/s", "ms_per_batch": th * 1e3, "calls": n_host,
                  "what": "the same batch with host (pageable numpy) input and output arrays"}

    progress("plain and host-buffer legs done")

    # ---------------------------------------------------------------- served stream
    served = None if args.no_serve else serve_leg(jc, probes, args)
    progress("served leg done")

    # ---------------------------------------------------------------- likelihood gradient
    grad = None if args.no_grad else gradient_leg(jc, probes, ell, n_z, args)
    progress("gradient leg done")

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

    progress("parity sample done")

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
               "kind": "port", "cpu_model": cpu_model(),
               "sample": "%d cosmologies of the same 1024-cosmology workload through the FP64 "
                         "reference-faithful oracle (torch CPU, %d threads), %.1f s" % (done, torch.get_num_threads(), tc)}
        cpu["per_workload"] = cpu_baselines(O, jc, args, params, ell, n_z)

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
        "graph_replay": graph_leg,
        "steady_state": steady,
        "plain": plain,
        "batched_vs_plain": value / plain_value,
        "host_buffers": host_batch,
        "rel_err_vs_fp64_oracle": {"max": max(errs) if errs else None, "samples": len(errs)},
        "tiles": {"waves_per_block": tiles_c3[0], "wl_ctile": tiles_c3[1], "tuned_now": tuned is not None,
                  "note": "0 = library default (waves 2 if >= 2048 blocks else 4; wl_ctile 4 if B >= 192 else 1)"},
        "served": served,
        "batch_vs_served_batch_shaped": (value / served["batch_shaped"]["C_ell_per_s"]) if served else None,
        "likelihood_gradient": grad,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
