"""GPU parity of the likelihood-gradient workload: the gradient of the
Gaussian log-likelihood over (Omega_c, Omega_b, h, n_s, sigma8, w0, wa) with
10 tracers, 100 ell in [10, 5000] and 512 Simpson intervals, for cosmologies
drawn with default_rng(1) from the box Omega_c [0.2, 0.35], Omega_b [0.04,
0.06], h [0.6, 0.8], n_s [0.9, 1.0], sigma8 [0.7, 0.9], w0 [-1.3, -0.7],
wa [-0.5, 0.5].

The covariance is fixed: gaussian_cl_covariance (angular_cl.py:120-163,
sparse, f_sky 0.25) at Planck15 with sigma_e 0.26 and 10 gal/arcmin^2 noise;
the data are the Planck15 mean plus L eps per ell block (L the Cholesky factor
of the 55 x 55 block, eps from default_rng(2)).  The loss is
gaussian_log_likelihood(data, mu(theta), C, include_logdet=False)
(likelihood.py:9-61) with mu = angular_cl(...).flatten() (angular_cl.py:196).

At 100 ell the adjoint reads the per-cosmology transfer table and the table of
its theta derivatives (the path the gradient benchmark times).  Checks: the
FP64 oracle's torch autograd (the stand-in for jax.grad) and central finite
differences of the oracle (step 1e-6; why not 1e-4 is stated at the test),
both at 1e-3 relative; the device's halofit-root discontinuity against the
oracle's; the
forward-mode JVP (closed-form EH98 in duals) as an independent check of the
table derivatives; a 256-cosmology run against per-cosmology calls.
"""
import os

import numpy as np
import pytest
import torch

import jax_cosmo_amd as jc
from jax_cosmo_amd import _native
from oracle import jc_oracle as O

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
BOX_LO = np.array([0.2, 0.04, 0.6, 0.9, 0.7, -1.3, -0.5])
BOX_HI = np.array([0.35, 0.06, 0.8, 1.0, 0.9, -0.7, 0.5])
GRAD_IDX = [0, 1, 2, 3, 4, 6, 7]
GRAD_RTOL = 1e-3


def c2_probes():
    wl = jc.probes.WeakLensing(
        [jc.redshift.smail_nz(1.0, 2.0, z0, gals_per_arcmin2=10.0) for z0 in (0.5, 0.75, 1.0, 1.25, 1.5)],
        sigma_e=0.26)
    nc = jc.probes.NumberCounts(
        [jc.redshift.gaussian_nz(mu, 0.1, gals_per_arcmin2=10.0) for mu in np.linspace(0.3, 1.5, 5)],
        jc.bias.constant_linear_bias(1.5))
    return [wl, nc]


def c2_oracle_probes():
    wl = O.WeakLensing([O.smail_nz(1.0, 2.0, z0, gals_per_arcmin2=10.0) for z0 in (0.5, 0.75, 1.0, 1.25, 1.5)],
                       sigma_e=0.26)
    nc = O.NumberCounts([O.gaussian_nz(mu, 0.1, gals_per_arcmin2=10.0)
                         for mu in np.linspace(0.3, 1.5, 5).tolist()], 1.5)
    return [wl, nc]


@pytest.fixture(scope="module")
def setup():
    assert torch.cuda.is_available()
    torch.set_num_threads(min(16, os.cpu_count() or 1))
    g = np.load(os.path.join(GOLD, "c2_planck15.npz"))  # FP64 oracle mean at Planck15
    ell, mu_fid = g["ell"], g["cl"]
    probes = c2_probes()
    cov = jc.cl.gaussian_cl_covariance(ell, probes, mu_fid, jc.cl.noise_cl(ell, probes), f_sky=0.25, sparse=True)
    L = np.linalg.cholesky(np.transpose(cov, (2, 0, 1)))
    eps = np.random.default_rng(2).normal(size=(len(ell), 55))
    data = (mu_fid + np.einsum("lij,lj->il", L, eps)).flatten()
    th = BOX_LO + (BOX_HI - BOX_LO) * np.random.default_rng(1).random((256, 7))
    params = np.zeros((256, 8))
    params[:, GRAD_IDX] = th
    ctx = _native.default_context()
    sid = ctx.tracer_set(jc.cl.tracer_entries(probes))
    return dict(ell=ell, probes=probes, cov=cov, data=data, params=params, ctx=ctx, sid=sid)


def oracle_loglike(p, s, need_grad=True):
    pv = torch.tensor(p, dtype=torch.float64, requires_grad=need_grad)
    mu = O.angular_cl(O.Cosmo.from_vector(pv), torch.tensor(s["ell"]), c2_oracle_probes(), N=512).flatten()
    lo = O.gaussian_log_likelihood(torch.tensor(s["data"]), mu, torch.tensor(s["cov"]), include_logdet=False)
    if not need_grad:
        return float(lo), None
    lo.backward()
    return float(lo.detach()), pv.grad.numpy()[GRAD_IDX]


def test_loglike_gradient_against_oracle_autograd(setup):
    s = setup
    params = s["params"][:4]
    ll, grad = jc.likelihood.log_likelihood_and_grad(params, s["ell"], s["probes"], s["data"], s["cov"],
                                                     include_logdet=False, n_z=512)
    for b in range(4):
        ref_ll, ref_g = oracle_loglike(params[b], s)
        err = np.abs(grad[b] - ref_g) / np.abs(ref_g)
        print("cosmology %d: loglike %.6e (oracle %.6e), max rel grad err %.2e" % (b, ll[b], ref_ll, err.max()))
        assert ll[b] == pytest.approx(ref_ll, rel=1e-4)
        np.testing.assert_allclose(grad[b], ref_g, rtol=GRAD_RTOL)


def test_loglike_gradient_against_central_fd(setup):
    # Central differences of the FP64 oracle on all 7 parameters.  The step is
    # 1e-6, not 1e-4: the reference's likelihood is piecewise smooth with jump
    # discontinuities.  Its halofit root (power.py:113-115) interpolates the
    # decreasing sigma^2(R) table with interpolate.py:12-37, whose clip on a
    # decreasing table always picks the segment left of the nearest node, so
    # ln k_nl jumps where the nearest node changes.  At this cosmology one
    # such switch (a-node 406, d ln k_nl = -9.1e-4) lies between Omega_c + 3e-5
    # and + 4e-5 and moves the log-likelihood by +6.8, so a central difference
    # with step 1e-4 reads 1.5755e6 against the derivative 1.5412e6.  Steps of
    # 3e-5 and below agree with the oracle's autograd to 1e-8.
    s = setup
    p = s["params"][0]
    _, grad = jc.likelihood.log_likelihood_and_grad(p[None], s["ell"], s["probes"], s["data"], s["cov"],
                                                    include_logdet=False, n_z=512)
    h = 1e-6
    for gi, pi in enumerate(GRAD_IDX):
        pp, pm = p.copy(), p.copy()
        pp[pi] += h
        pm[pi] -= h
        fd = (oracle_loglike(pp, s, False)[0] - oracle_loglike(pm, s, False)[0]) / (2 * h)
        print("param %d: device %.8e  fd %.8e" % (pi, grad[0, gi], fd))
        assert grad[0, gi] == pytest.approx(fd, rel=GRAD_RTOL)


def test_halofit_root_jump_matches_oracle(setup):
    # the discontinuity above, on the device: the FP64 nonlinear scale of the
    # table kernel switches segment at the same a-node between the same two
    # cosmologies, by the same amount
    s = setup
    ctx = s["ctx"]
    knl = []
    for dx in (3e-5, 4e-5):
        p = s["params"][0].copy()
        p[0] += dx
        t = ctx.debug_tables(p, s["sid"], 512)
        knl.append(np.array([t[1024 + 8 * m + 2] for m in range(513)]))
    d = np.log(knl[1]) - np.log(knl[0])
    assert int(np.argmax(np.abs(d))) == 406
    assert d[406] == pytest.approx(-9.09e-4, rel=2e-3)


def test_vjp_table_mode_against_oracle_and_jvp(setup):
    # angular_cl_vjp with a random cotangent: B = 1 and a 5-cosmology batch
    # (both split the a range over several blocks), against the oracle's
    # autograd; then <jvp(v), u> = <v, vjp(u)> with the closed-form JVP
    s = setup
    ctx, sid, ell = s["ctx"], s["sid"], s["ell"]
    rng = np.random.default_rng(11)
    u = rng.normal(size=(5, 55, 100))
    th = ctx.angular_cl_vjp(s["params"][:5], sid, ell, 512, u.astype(np.float32))
    th1 = ctx.angular_cl_vjp(s["params"][:1], sid, ell, 512, u[:1].astype(np.float32))
    np.testing.assert_allclose(th1[0], th[0], rtol=1e-5)
    for b in (0, 3):
        pv = torch.tensor(s["params"][b], dtype=torch.float64, requires_grad=True)
        (O.angular_cl(O.Cosmo.from_vector(pv), torch.tensor(ell), c2_oracle_probes(), N=512)
         * torch.tensor(u[b])).sum().backward()
        ref = pv.grad.numpy()[GRAD_IDX]
        print("cosmology %d: vjp max rel err %.2e" % (b, np.max(np.abs(th[b] - ref) / np.abs(ref))))
        np.testing.assert_allclose(th[b], ref, rtol=GRAD_RTOL)
    v = rng.normal(size=(1, 7))
    _, dot = ctx.angular_cl_jvp(s["params"][:1], v, sid, ell, 512, 55)
    lhs = float((dot[0].astype(np.float64) * u[0]).sum())
    rhs = float(th[0] @ v[0])
    print("dot-product test: <jvp v, u> %.8e  <v, vjp u> %.8e" % (lhs, rhs))
    assert lhs == pytest.approx(rhs, rel=GRAD_RTOL)


def test_full_batch_256_properties(setup):
    # the benchmark's batch: finite, equal to per-cosmology calls (different
    # slicing and block shapes), log-likelihood equal to the two-step path,
    # and one more cosmology against the oracle
    s = setup
    params = s["params"]
    ll, grad = jc.likelihood.log_likelihood_and_grad(params, s["ell"], s["probes"], s["data"], s["cov"],
                                                     include_logdet=False, n_z=512)
    assert ll.shape == (256,) and grad.shape == (256, 7)
    assert np.all(np.isfinite(ll)) and np.all(np.isfinite(grad))
    mu = jc.cl.angular_cl_batch(params, s["ell"], s["probes"], n_z=512)
    ll2 = jc.likelihood.log_likelihood_batch(s["data"], mu, s["cov"], include_logdet=False)
    np.testing.assert_allclose(ll, ll2, rtol=1e-6)
    for b in (0, 77, 255):
        l1, g1 = jc.likelihood.log_likelihood_and_grad(params[b:b + 1], s["ell"], s["probes"], s["data"], s["cov"],
                                                       include_logdet=False, n_z=512)
        assert ll[b] == pytest.approx(l1[0], rel=1e-6)
        np.testing.assert_allclose(grad[b], g1[0], rtol=1e-4)
    ref_ll, ref_g = oracle_loglike(params[255], s)
    assert ll[255] == pytest.approx(ref_ll, rel=1e-4)
    np.testing.assert_allclose(grad[255], ref_g, rtol=GRAD_RTOL)
