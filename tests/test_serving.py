"""GPU tests of the continuous-batching request loop (jc_server_* C ABI).

Each served result must equal the same request computed alone through the
plain entry points (jc_angular_cl / jc_angular_cl_vjp) within float32
rounding, and the FP64 oracle within the path's 1e-4 target.  Failure
handling (refused and over-budget requests) must not disturb the others.
"""
import numpy as np
import pytest
import torch

import jax_cosmo_amd as jc
from jax_cosmo_amd import serving
from oracle import jc_oracle as O

pytestmark = pytest.mark.gpu

ZS = (0.5, 0.75, 1.0, 1.25, 1.5)
MUS = np.linspace(0.3, 1.5, 5)


def pool_probes():
    wl = jc.probes.WeakLensing([jc.redshift.smail_nz(1.0, 2.0, z0, gals_per_arcmin2=10.0) for z0 in ZS])
    nc = jc.probes.NumberCounts([jc.redshift.gaussian_nz(mu, 0.1, gals_per_arcmin2=10.0) for mu in MUS],
                                jc.bias.constant_linear_bias(1.5))
    return [wl, nc]


def subset_probes(idx):
    wl = [jc.redshift.smail_nz(1.0, 2.0, ZS[i], gals_per_arcmin2=10.0) for i in idx if i < 5]
    nc = [jc.redshift.gaussian_nz(MUS[i - 5], 0.1, gals_per_arcmin2=10.0) for i in idx if i >= 5]
    out = []
    if wl:
        out.append(jc.probes.WeakLensing(wl))
    if nc:
        out.append(jc.probes.NumberCounts(nc, jc.bias.constant_linear_bias(1.5)))
    return out


def oracle_subset(idx):
    wl = [O.smail_nz(1.0, 2.0, ZS[i], gals_per_arcmin2=10.0) for i in idx if i < 5]
    nc = [O.gaussian_nz(float(MUS[i - 5]), 0.1, gals_per_arcmin2=10.0) for i in idx if i >= 5]
    out = []
    if wl:
        out.append(O.WeakLensing(wl))
    if nc:
        out.append(O.NumberCounts(nc, 1.5))
    return out


def rel(a, b):
    return float(np.max(np.abs(np.asarray(a, np.float64) - b) / np.abs(b)))


def test_served_mix_equals_alone():
    ctx = jc._native.default_context()
    # batches above 32 requests replay the full-capacity graphs, smaller ones
    # the small-grid graphs: both run here
    reqs = serving.make_stream(48, seed=5, n_z_choices=(256, 512), ell_range=(5, 60))
    with serving.Server(pool_probes(), max_batch=40, max_grad=24, max_ell=64, slots=2, n_z=(256, 512)) as srv:
        handles = {}
        for r in reqs:
            rid, out, grad = srv.submit(**r)
            handles[rid] = (r, out, grad)
        srv.wait(0, timeout_ms=120000)
        comps = srv.poll()
        st = srv.stats()
    assert len(comps) == len(reqs)
    assert all(c["status"] == 0 for c in comps)
    assert st["graph_replays"] >= 1 and st["batches"] >= 2
    assert st["done"] == len(reqs) and st["wait_ms"] >= 0.0 and st["service_ms"] > 0.0
    for rid, (r, out, grad) in handles.items():
        probes = subset_probes(r["tracers"])
        alone = jc.cl.angular_cl(jc.Cosmology(*r["cosmo"]), r["ell"], probes, n_z=r["n_z"])
        assert rel(out, alone) < 1e-5, (rid, rel(out, alone))
        if r["cl_bar"] is not None:
            sid = ctx.tracer_set(jc.cl.tracer_entries(probes))
            th = ctx.angular_cl_vjp(r["cosmo"], sid, r["ell"], r["n_z"], r["cl_bar"][None])[0]
            np.testing.assert_allclose(grad, th, rtol=1e-3, atol=1e-5 * np.abs(th).max())


def test_served_against_oracle():
    rng = np.random.default_rng(9)
    with serving.Server(pool_probes(), max_batch=8, max_grad=4, max_ell=64, slots=1, n_z=(512,)) as srv:
        cases = []
        for idx in ([0, 3, 7], [9], [1, 2, 4, 5, 6, 8]):
            cosmo = np.array([0.27, 0.05, 0.7, 0.96, 0.8, 0.0, -0.9, 0.1])
            ell = np.sort(rng.uniform(10, 4000, 17))
            cases.append((idx, cosmo, ell) + srv.submit(cosmo, idx, ell, n_z=512)[1:2])
        srv.wait(0)
        assert all(c["status"] == 0 for c in srv.poll())
    for idx, cosmo, ell, out in cases:
        ref = O.angular_cl(O.Cosmo(*cosmo), torch.tensor(ell), oracle_subset(idx), N=512).numpy()
        assert rel(out, ref) < 1e-4


def test_served_gamma_growth_against_oracle():
    # a server created for gamma growth keeps it for all its requests, while
    # the context's next plain call is back on the growth ODE
    cosmo = np.array([0.3, 0.05, 0.68, 0.97, 0.8, 0.0, -1.0, 0.0])
    ell = np.geomspace(20.0, 3000.0, 13)
    with serving.Server(pool_probes(), max_batch=8, max_grad=4, max_ell=64, slots=1, n_z=(256,),
                        gamma=0.55) as srv:
        jc.cl.angular_cl(jc.Cosmology(*cosmo), ell, subset_probes([0, 6]), n_z=256)
        _, out, _ = srv.submit(cosmo, [0, 6], ell, n_z=256)
        srv.wait(0)
        assert all(c["status"] == 0 for c in srv.poll())
    ref = O.angular_cl(O.Cosmo(*cosmo, gamma=0.55), torch.tensor(ell), oracle_subset([0, 6]), N=256).numpy()
    assert rel(out, ref) < 1e-4


def test_bad_and_oversized_requests_do_not_disturb_others():
    reqs = serving.make_stream(10, seed=8, n_z_choices=(256,), ell_range=(5, 30), grad_frac=0.0)
    with serving.Server(pool_probes(), max_batch=16, max_grad=4, max_ell=200, slots=2, n_z=(256,),
                        mem_budget=300_000) as srv:
        good = {}
        for i, r in enumerate(reqs):
            rid, out, _ = srv.submit(**r)
            good[rid] = (r, out)
            if i == 3:
                with pytest.raises(ValueError):
                    srv.submit(r["cosmo"], r["tracers"], r["ell"], n_z=512)  # size not served
                bad = r["cosmo"].copy()
                bad[2] = float("nan")
                with pytest.raises(ValueError):
                    srv.submit(bad, r["tracers"], r["ell"], n_z=256)
                # footprint above the budget: completes with JC_E_NOMEM
                big_id, _, _ = srv.submit(r["cosmo"], list(range(10)), np.geomspace(10, 5000, 200), n_z=256,
                                          cl_bar=np.ones((55, 200), np.float32))
        srv.wait(0)
        comps = {c["id"]: c for c in srv.poll()}
    assert comps[big_id]["status"] == -2
    for rid, (r, out) in good.items():
        assert comps[rid]["status"] == 0
        alone = jc.cl.angular_cl(jc.Cosmology(*r["cosmo"]), r["ell"], subset_probes(r["tracers"]), n_z=256)
        assert rel(out, alone) < 1e-5


@pytest.mark.parametrize("use_graph", [True, False])
def test_graph_and_direct_paths_agree(use_graph):
    reqs = serving.make_stream(12, seed=21, n_z_choices=(256,), ell_range=(8, 24))
    outs = []
    with serving.Server(pool_probes(), max_batch=12, max_grad=6, max_ell=32, slots=1, n_z=(256,),
                        use_graph=use_graph) as srv:
        for r in reqs:
            outs.append(srv.submit(**r)[1:])
        srv.wait(0)
        assert all(c["status"] == 0 for c in srv.poll())
        st = srv.stats()
    assert (st["graph_replays"] > 0) == use_graph
    for r, (out, grad) in zip(reqs, outs):
        alone = jc.cl.angular_cl(jc.Cosmology(*r["cosmo"]), r["ell"], subset_probes(r["tracers"]), n_z=256)
        assert rel(out, alone) < 1e-5


def test_served_full_shapes_against_plain_and_oracle():
    # the request shapes of the 4096-request stream: Simpson sizes 256, 512
    # and 1024 served together, 100-200 ell, several with the 7-parameter
    # gradient, all 10 bins and subsets; each against the same request alone
    # through the plain entry points and, for a sample, the FP64 oracle
    # (C_ell at 1e-4, gradients against the oracle's autograd at 1e-3)
    ctx = jc._native.default_context()
    rng = np.random.default_rng(17)
    shapes = [(list(range(10)), 200, 1024, True), ([0, 2, 4, 6, 8], 150, 1024, False),
              ([1, 3, 5, 7, 9], 120, 512, True), (list(range(10)), 100, 256, True),
              ([4], 200, 1024, True), ([0, 9], 180, 512, False), ([2, 5, 6], 160, 1024, True),
              (list(range(8)), 130, 256, False)]
    reqs = []
    for idx, n_ell, n_z, grad in shapes:
        th = serving.BOX_LO + (serving.BOX_HI - serving.BOX_LO) * rng.random(7)
        cosmo = np.array([th[0], th[1], th[2], th[3], th[4], 0.0, th[5], th[6]])
        P = len(idx) * (len(idx) + 1) // 2
        cb = rng.normal(size=(P, n_ell)).astype(np.float32) if grad else None
        reqs.append(dict(cosmo=cosmo, tracers=idx, ell=np.geomspace(10.0, 5000.0, n_ell), n_z=n_z, cl_bar=cb))
    with serving.Server(pool_probes(), max_batch=16, max_grad=8, max_ell=200, slots=2, n_z=(256, 512, 1024)) as srv:
        outs = [srv.submit(**r)[1:] for r in reqs]
        srv.wait(0, timeout_ms=120000)
        comps = srv.poll()
    assert len(comps) == len(reqs) and all(c["status"] == 0 for c in comps)
    for k, (r, (out, grad)) in enumerate(zip(reqs, outs)):
        probes = subset_probes(r["tracers"])
        alone = jc.cl.angular_cl(jc.Cosmology(*r["cosmo"]), r["ell"], probes, n_z=r["n_z"])
        assert rel(out, alone) < 1e-5, (k, rel(out, alone))
        if r["cl_bar"] is not None:
            sid = ctx.tracer_set(jc.cl.tracer_entries(probes))
            th = ctx.angular_cl_vjp(r["cosmo"], sid, r["ell"], r["n_z"], r["cl_bar"][None])[0]
            np.testing.assert_allclose(grad, th, rtol=1e-3, atol=1e-5 * np.abs(th).max())
    for k in (0, 2, 4):
        r, (out, grad) = reqs[k], outs[k]
        pv = torch.tensor(r["cosmo"], dtype=torch.float64, requires_grad=True)
        ocl = O.angular_cl(O.Cosmo.from_vector(pv), torch.tensor(r["ell"]), oracle_subset(r["tracers"]), N=r["n_z"])
        err = rel(out, ocl.detach().numpy())
        (ocl * torch.tensor(r["cl_bar"], dtype=torch.float64)).sum().backward()
        gref = pv.grad.numpy()[[0, 1, 2, 3, 4, 6, 7]]
        print("request %d: C_ell max rel err %.2e, grad max rel err %.2e" %
              (k, err, np.max(np.abs(grad - gref) / np.abs(gref))))
        assert err < 1e-4
        np.testing.assert_allclose(grad, gref, rtol=1e-3)


def test_bulk_submission_and_refusal():
    # jc_server_submit_many: requests queued in one call equal the same
    # requests alone; one invalid request refuses the whole call, nothing runs
    reqs = serving.make_stream(20, seed=31, n_z_choices=(256, 512), ell_range=(10, 70))
    with serving.Server(pool_probes(), max_batch=16, max_grad=8, max_ell=80, slots=2, n_z=(256, 512)) as srv:
        prep = srv.prepare(reqs)
        ids = np.concatenate([srv.submit_prepared(prep, 0, 7), srv.submit_prepared(prep, 7)])
        assert len(set(ids.tolist())) == 20
        bad = [dict(r) for r in reqs[:3]]
        bad[1]["n_z"] = 1024  # not served
        with pytest.raises(ValueError):
            srv.submit_prepared(srv.prepare(bad))
        srv.wait(0, timeout_ms=120000)
        comps = srv.poll()
        assert sorted(c["id"] for c in comps) == sorted(ids.tolist())
        assert all(c["status"] == 0 for c in comps)
    ctx = jc._native.default_context()
    for r, (out, grad) in zip(reqs, prep.outs):
        probes = subset_probes(r["tracers"])
        alone = jc.cl.angular_cl(jc.Cosmology(*r["cosmo"]), r["ell"], probes, n_z=r["n_z"])
        assert rel(out, alone) < 1e-5
        if r["cl_bar"] is not None:
            sid = ctx.tracer_set(jc.cl.tracer_entries(probes))
            th = ctx.angular_cl_vjp(r["cosmo"], sid, r["ell"], r["n_z"], r["cl_bar"][None])[0]
            np.testing.assert_allclose(grad, th, rtol=1e-3, atol=1e-5 * np.abs(th).max())


@pytest.mark.parametrize("status", [-2, -4])
def test_injected_batch_failure_leaves_other_batches_intact(status):
    # a batch that fails at launch (as a device allocation or runtime
    # failure would) completes its requests with that status; the batches
    # after it run and equal the same requests alone
    reqs = serving.make_stream(24, seed=41, n_z_choices=(256,), ell_range=(8, 40))
    with serving.Server(pool_probes(), max_batch=8, max_grad=8, max_ell=64, slots=1, n_z=(256,)) as srv:
        srv.inject_failure(1, status)
        prep = srv.prepare(reqs)
        ids = srv.submit_prepared(prep)
        srv.wait(0, timeout_ms=120000)
        comps = {c["id"]: c for c in srv.poll()}
    st = [comps[i]["status"] for i in ids.tolist()]
    failed = [k for k, s in enumerate(st) if s != 0]
    assert failed == list(range(len(failed))) and 1 <= len(failed) <= 8  # the first batch, in arrival order
    assert all(s == status for s in st[:len(failed)])
    ctx = jc._native.default_context()
    for k in range(len(failed), len(reqs)):
        r, (out, grad) = reqs[k], prep.outs[k]
        probes = subset_probes(r["tracers"])
        alone = jc.cl.angular_cl(jc.Cosmology(*r["cosmo"]), r["ell"], probes, n_z=256)
        assert rel(out, alone) < 1e-5
        if r["cl_bar"] is not None:
            sid = ctx.tracer_set(jc.cl.tracer_entries(probes))
            th = ctx.angular_cl_vjp(r["cosmo"], sid, r["ell"], 256, r["cl_bar"][None])[0]
            np.testing.assert_allclose(grad, th, rtol=1e-3, atol=1e-5 * np.abs(th).max())


def test_large_batches_delivered_by_copier_threads():
    # Deliveries above 256 KB are copied to the callers' arrays by the
    # server's copier threads, from two alternating result buffers per slot,
    # while the slot runs its next batch.  Full-size requests (10 bins, 200
    # ell: 44 KB of C_ell each) in batches of 64 exercise that path for
    # forward results and for the gradient pass; every result must equal the
    # same request alone, and no result array may be left unwritten.
    rng = np.random.default_rng(77)
    lo = np.array([0.2, 0.04, 0.6, 0.9, 0.7, -1.3, -0.5])
    hi = np.array([0.35, 0.06, 0.8, 1.0, 0.9, -0.7, 0.5])
    n = 192
    th = lo + (hi - lo) * rng.random((n, 7))
    cosmos = np.zeros((n, 8))
    cosmos[:, [0, 1, 2, 3, 4, 6, 7]] = th
    ell = np.geomspace(10.0, 5000.0, 200)
    reqs = []
    for k in range(n):
        cb = rng.normal(size=(55, 200)).astype(np.float32) if k % 3 == 0 else None
        reqs.append(dict(cosmo=cosmos[k], tracers=list(range(10)), ell=ell, n_z=512, cl_bar=cb))
    with serving.Server(pool_probes(), max_batch=64, max_grad=32, max_ell=200, slots=2, n_z=(512,)) as srv:
        prep = srv.prepare(reqs)
        for out, grad in prep.outs:  # sentinel: a result array left unwritten would keep it
            out.fill(np.nan)
            if grad is not None:
                grad.fill(np.nan)
        ids = srv.submit_prepared(prep)
        srv.wait(0, timeout_ms=120000)
        comps = srv.poll_array(4096)
        st = srv.stats()
    assert comps.size == n and np.all(comps["status"] == 0)
    assert sorted(comps["id"].tolist()) == sorted(ids.tolist())
    assert st["max_batch"] > 32  # full batches, whose deliveries go to the copier threads
    ctx = jc._native.default_context()
    sid = ctx.tracer_set(jc.cl.tracer_entries(pool_probes()))
    for k in list(range(0, n, 7)) + [n - 1]:
        out, grad = prep.outs[k]
        assert np.all(np.isfinite(out))
        alone = ctx.angular_cl(cosmos[k], sid, ell, 512, 55)[0]
        assert rel(out, alone) < 1e-5, k
        if reqs[k]["cl_bar"] is not None:
            assert np.all(np.isfinite(grad))
            th_alone = ctx.angular_cl_vjp(cosmos[k], sid, ell, 512, reqs[k]["cl_bar"][None])[0]
            np.testing.assert_allclose(grad, th_alone, rtol=1e-3, atol=1e-5 * np.abs(th_alone).max())


def test_concurrent_bulk_submitters():
    # Two threads submit bulk arrays at once through one client (which holds
    # its lock across each call, so that a call's id range is recorded before
    # any of its completions can be polled; the library's own concurrent
    # submitters are exercised by the sanitizer check program).  Every call's
    # ids are one consecutive range and every request completes with the
    # result it would have alone.
    import threading

    reqs = serving.make_stream(480, seed=91, n_z_choices=(256, 512), ell_range=(20, 120))
    halves = (reqs[:240], reqs[240:])
    with serving.Server(pool_probes(), max_batch=64, max_grad=24, max_ell=128, slots=2, n_z=(256, 512)) as srv:
        preps = [srv.prepare(h) for h in halves]
        ids = [None, None]
        errors = []

        def go(k):
            try:
                for part in range(3):  # three calls per thread
                    got = srv.submit_prepared(preps[k], 80 * part, 80 * (part + 1))
                    ids[k] = got if ids[k] is None else np.concatenate([ids[k], got])
            except Exception as e:  # reported below
                errors.append(e)

        th = [threading.Thread(target=go, args=(k,)) for k in (0, 1)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        assert not errors
        srv.wait(0, timeout_ms=120000)
        comps = srv.poll_array(4096)
    assert comps.size == 480 and np.all(comps["status"] == 0)
    for k in (0, 1):
        for part in range(3):
            seg = ids[k][80 * part:80 * (part + 1)]
            assert np.array_equal(seg, np.arange(seg[0], seg[0] + 80))
    assert len(set(ids[0].tolist()) | set(ids[1].tolist())) == 480
    for k in (0, 1):
        for j in (0, 79, 160, 239):
            r, (out, grad) = halves[k][j], preps[k].outs[j]
            alone = jc.cl.angular_cl(jc.Cosmology(*r["cosmo"]), r["ell"], subset_probes(r["tracers"]), n_z=r["n_z"])
            assert rel(out, alone) < 1e-5
