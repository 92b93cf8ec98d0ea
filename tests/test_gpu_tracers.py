"""GPU parity of the tracer, n(z) and bias families beyond the headline
probes: NLA intrinsic alignments (probes.py:102-129) with the three bias
models of bias.py:10-59, per-bin multiplicative bias and sigma_e
(probes.py:204-222), fu_nz, delta_nz and systematic_shift
(redshift.py:80-171), against the FP64 oracle at the path's 1e-4 target and,
for the adjoint, the oracle's autograd at 1e-3."""
import numpy as np
import pytest
import torch

import jax_cosmo_amd as jc
from jax_cosmo_amd import serving
from oracle import jc_oracle as O

pytestmark = pytest.mark.gpu

CL_RTOL = 1e-4
P15 = [0.2589, 0.0486, 0.6774, 0.9667, 0.8159, 0.0, -1.0, 0.0]
COSMO2 = [0.31, 0.045, 0.72, 0.95, 0.78, 0.0, -0.9, 0.2]


def rel(a, b):
    return float(np.max(np.abs(np.asarray(a, np.float64) - b) / np.abs(b)))


def oracle_cl(params, ell, probes, n_z):
    return O.angular_cl(O.Cosmo(*params), torch.tensor(ell), probes, N=n_z).numpy()


def ia_case():
    dev = [jc.probes.WeakLensing([jc.redshift.smail_nz(1.0, 2.0, 0.6), jc.redshift.smail_nz(1.0, 2.0, 1.1)],
                                 ia_bias=jc.bias.constant_linear_bias(1.2), multiplicative_bias=[0.01, -0.02],
                                 sigma_e=[0.26, 0.3]),
           jc.probes.WeakLensing([jc.redshift.smail_nz(1.2, 1.8, 0.8, zmax=5.0), jc.redshift.fu_nz(0.471, 5.1843, 0.7259, zmax=4.0)],
                                 ia_bias=[jc.bias.inverse_growth_linear_bias(0.8), jc.bias.des_y1_ia_bias(0.5, 2.0, 0.62)])]
    orc = [O.WeakLensing([O.smail_nz(1.0, 2.0, 0.6), O.smail_nz(1.0, 2.0, 1.1)], ia_bias=1.2,
                         multiplicative_bias=[0.01, -0.02], sigma_e=[0.26, 0.3]),
           O.WeakLensing([O.smail_nz(1.2, 1.8, 0.8, zmax=5.0), O.fu_nz(0.471, 5.1843, 0.7259, zmax=4.0)],
                         ia_bias=[O.inverse_growth_linear_bias(0.8), O.des_y1_ia_bias(0.5, 2.0, 0.62)])]
    return dev, orc


def nz_case():
    dev = [jc.probes.WeakLensing([jc.redshift.delta_nz(0.9), jc.redshift.smail_nz(1.0, 2.0, 0.7), jc.redshift.delta_nz(1.6)],
                                 multiplicative_bias=0.03),
           jc.probes.NumberCounts([jc.redshift.systematic_shift(jc.redshift.smail_nz(1.0, 2.0, 0.4), 0.05, zmax=5.0),
                                   jc.redshift.systematic_shift(jc.redshift.gaussian_nz(0.9, 0.15), -0.04),
                                   jc.redshift.fu_nz(0.471, 5.1843, 0.7259, zmax=4.0)],
                                  [jc.bias.inverse_growth_linear_bias(1.3), jc.bias.des_y1_ia_bias(1.1, 0.7, 0.5),
                                   jc.bias.constant_linear_bias(1.6)])]
    orc = [O.WeakLensing([O.delta_nz(0.9), O.smail_nz(1.0, 2.0, 0.7), O.delta_nz(1.6)], multiplicative_bias=0.03),
           O.NumberCounts([O.systematic_shift(O.smail_nz(1.0, 2.0, 0.4), 0.05, zmax=5.0),
                           O.systematic_shift(O.gaussian_nz(0.9, 0.15), -0.04),
                           O.fu_nz(0.471, 5.1843, 0.7259, zmax=4.0)],
                          [O.inverse_growth_linear_bias(1.3), O.des_y1_ia_bias(1.1, 0.7, 0.5), 1.6])]
    return dev, orc


@pytest.mark.parametrize("case", ["ia", "nz"])
@pytest.mark.parametrize("params", [P15, COSMO2])
def test_families_against_oracle(case, params):
    dev, orc = ia_case() if case == "ia" else nz_case()
    ell = np.geomspace(10.0, 3000.0, 23)
    cl = jc.cl.angular_cl(jc.Cosmology(*params), ell, dev, n_z=256)
    ref = oracle_cl(params, ell, orc, 256)
    err = rel(cl, ref)
    print(case, "max rel err %.3e" % err)
    assert err < CL_RTOL


def test_families_batched_table_path():
    # 100 ell: the forward kernels take the transfer-table path; 3 cosmologies
    dev, orc = ia_case()
    ell = np.geomspace(10.0, 5000.0, 100)
    params = np.array([P15, COSMO2, [0.25, 0.05, 0.65, 0.98, 0.85, 0.0, -1.1, -0.3]])
    cl = jc.cl.angular_cl_batch(params, ell, dev, n_z=512)
    for b in range(3):
        assert rel(cl[b], oracle_cl(params[b], ell, orc, 512)) < CL_RTOL


def test_noise_with_per_bin_sigma_e():
    dev, orc = ia_case()
    ell = np.geomspace(10.0, 3000.0, 7)
    np.testing.assert_allclose(jc.cl.noise_cl(ell, dev), O.noise_cl(torch.tensor(ell), orc).numpy(), rtol=1e-14)


def test_adjoint_with_intrinsic_alignments():
    dev, orc = ia_case()
    ell = np.geomspace(20.0, 2000.0, 9)
    ctx = jc._native.default_context()
    sid = ctx.tracer_set(jc.cl.tracer_entries(dev))
    P = 10
    w = np.random.default_rng(5).normal(size=(P, len(ell)))
    th = ctx.angular_cl_vjp(np.array(COSMO2), sid, ell, 128, w.astype(np.float32)[None])[0]
    t = torch.tensor([COSMO2[i] for i in (0, 1, 2, 3, 4, 6, 7)], dtype=torch.float64, requires_grad=True)
    c = O.Cosmo(t[0], t[1], t[2], t[3], t[4], 0.0, t[5], t[6])
    cl = O.angular_cl(c, torch.tensor(ell), orc, N=128)
    (cl * torch.tensor(w)).sum().backward()
    np.testing.assert_allclose(th, t.grad.numpy(), rtol=1e-3, atol=1e-3 * np.abs(t.grad.numpy()).max())


def test_served_requests_with_new_families():
    dev, _ = nz_case()
    dev_ia, _ = ia_case()
    pool = dev + dev_ia  # 3 + 3 + 2 + 2 bins
    reqs = [dict(cosmo=np.array(COSMO2), tracers=[0, 2, 5], ell=np.geomspace(15, 900, 11), n_z=256, cl_bar=None),
            dict(cosmo=np.array(P15), tracers=[6, 7, 1, 8, 9], ell=np.geomspace(30, 2000, 17), n_z=256, cl_bar=None)]
    with serving.Server(pool, max_batch=4, max_grad=2, max_ell=32, slots=1, n_z=(256,)) as srv:
        outs = [srv.submit(**r)[1] for r in reqs]
        srv.wait(0)
        assert all(c["status"] == 0 for c in srv.poll())
    entries = jc.cl.tracer_entries(pool)
    ctx = jc._native.default_context()
    for r, out in zip(reqs, outs):
        # the served result equals the plain call on the same bins (each
        # request takes whole probes' worth of zmax, so the grids agree)
        sid = ctx.tracer_set([entries[i] for i in r["tracers"]])
        n = len(r["tracers"])
        alone = ctx.angular_cl(r["cosmo"], sid, r["ell"], 256, n * (n + 1) // 2)[0]
        assert rel(out, alone.astype(np.float64)) < 1e-5


def test_kde_catalogs_against_oracle():
    # redshift.py:126-156 evaluated on the device from the catalog, for a
    # weak-lensing bin, a number-count bin and a shifted KDE
    rng = np.random.default_rng(7)
    zc1 = rng.gamma(3.0, 0.25, 800)
    w1 = rng.uniform(0.2, 1.0, 800)
    zc2 = rng.normal(0.8, 0.2, 500).clip(0.01)
    dev = [jc.probes.WeakLensing([jc.redshift.kde_nz(zc1, w1, bw=0.05, zmax=4.0), jc.redshift.smail_nz(1.0, 2.0, 0.9)]),
           jc.probes.NumberCounts([jc.redshift.kde_nz(zc2, 1.0, bw=0.08, zmax=3.0),
                                   jc.redshift.systematic_shift(jc.redshift.kde_nz(zc2, 1.0, bw=0.08), 0.03, zmax=3.0)],
                                  jc.bias.constant_linear_bias(1.4))]
    orc = [O.WeakLensing([O.kde_nz(zc1, w1, 0.05, zmax=4.0), O.smail_nz(1.0, 2.0, 0.9)]),
           O.NumberCounts([O.kde_nz(zc2, np.ones(500), 0.08, zmax=3.0),
                           O.systematic_shift(O.kde_nz(zc2, np.ones(500), 0.08), 0.03, zmax=3.0)], 1.4)]
    ell = np.geomspace(10.0, 3000.0, 19)
    for params in (P15, COSMO2):
        cl = jc.cl.angular_cl(jc.Cosmology(*params), ell, dev, n_z=256)
        err = rel(cl, oracle_cl(params, ell, orc, 256))
        print("kde max rel err %.3e" % err)
        assert err < CL_RTOL


def test_refused_combinations():
    with pytest.raises(NotImplementedError):
        jc.cl.angular_cl(jc.Planck15(), np.array([100.0]),
                         [jc.probes.NumberCounts([jc.redshift.delta_nz(1.0)], jc.bias.constant_linear_bias(1.0))])
    with pytest.raises(NotImplementedError):
        jc.cl.angular_cl(jc.Planck15(), np.array([100.0]),
                         [jc.probes.WeakLensing([jc.redshift.delta_nz(1.0)], ia_bias=jc.bias.constant_linear_bias(1.0))])


# ---------------------------------------------------------------- user-defined n(z)
class SmailLike(jc.redshift.redshift_distribution):
    """The Smail formula written as a user subclass (redshift.py:15-31 pattern)."""

    def pz_fn(self, z):
        a, b, z0 = self.params
        return z ** a * np.exp(-((z / z0) ** b))


class Bimodal(jc.redshift.redshift_distribution):
    """A shape no built-in family has: two bumps and a tail."""

    def pz_fn(self, z):
        return z * z * np.exp(-z / 0.3) + 0.4 * np.exp(-0.5 * ((z - 1.4) / 0.2) ** 2)


def bimodal_t(z):
    return z * z * torch.exp(-z / 0.3) + 0.4 * torch.exp(-0.5 * ((z - 1.4) / 0.2) ** 2)


def test_user_nz_equals_builtin_family():
    # the same n(z) through the host function and through the device's own
    # Smail code: same tables up to rounding, same C_ell
    ell = np.geomspace(10.0, 3000.0, 30)
    user = [jc.probes.WeakLensing([SmailLike(1.0, 2.0, 0.7, zmax=4.0)]),
            jc.probes.NumberCounts([SmailLike(1.2, 1.5, 0.5)], jc.bias.constant_linear_bias(1.4))]
    ref = [jc.probes.WeakLensing([jc.redshift.smail_nz(1.0, 2.0, 0.7, zmax=4.0)]),
           jc.probes.NumberCounts([jc.redshift.smail_nz(1.2, 1.5, 0.5)], jc.bias.constant_linear_bias(1.4))]
    for params in (P15, COSMO2):
        a = jc.cl.angular_cl(jc.Cosmology(*params), ell, user, n_z=256)
        b = jc.cl.angular_cl(jc.Cosmology(*params), ell, ref, n_z=256)
        print("user vs built-in smail: max rel %.2e" % rel(a, b))
        assert rel(a, b) < 1e-6


def test_user_nz_against_oracle():
    shifted = jc.redshift.systematic_shift(Bimodal(zmax=5.0), 0.1, zmax=5.0)
    dev = [jc.probes.WeakLensing([Bimodal(zmax=5.0), shifted], ia_bias=jc.bias.constant_linear_bias(0.9)),
           jc.probes.NumberCounts([Bimodal(zmax=4.0)], jc.bias.inverse_growth_linear_bias(1.2))]
    orc = [O.WeakLensing([O.user_nz(bimodal_t, zmax=5.0), O.systematic_shift(O.user_nz(bimodal_t), 0.1, zmax=5.0)],
                         ia_bias=0.9),
           O.NumberCounts([O.user_nz(bimodal_t, zmax=4.0)], [O.inverse_growth_linear_bias(1.2)])]
    for n_ell, n_z in ((23, 256), (100, 512)):  # closed-form and transfer-table forward paths
        ell = np.geomspace(10.0, 3000.0, n_ell)
        cl = jc.cl.angular_cl(jc.Cosmology(*COSMO2), ell, dev, n_z=n_z)
        err = rel(cl, oracle_cl(COSMO2, ell, orc, n_z))
        print("user n(z) n_ell=%d n_z=%d: max rel err %.3e" % (n_ell, n_z, err))
        assert err < CL_RTOL
    # the adjoint through the same tables
    ell = np.geomspace(20.0, 2000.0, 9)
    ctx = jc._native.default_context()
    sid = ctx.tracer_set(jc.cl.tracer_entries(dev))
    w = np.random.default_rng(8).normal(size=(6, len(ell)))
    th = ctx.angular_cl_vjp(np.array(COSMO2), sid, ell, 128, w.astype(np.float32)[None])[0]
    t = torch.tensor([COSMO2[i] for i in (0, 1, 2, 3, 4, 6, 7)], dtype=torch.float64, requires_grad=True)
    pv = torch.cat([t[:5], torch.zeros(1, dtype=torch.float64), t[5:]])
    (O.angular_cl(O.Cosmo.from_vector(pv), torch.tensor(ell), orc, N=128) * torch.tensor(w)).sum().backward()
    np.testing.assert_allclose(th, t.grad.numpy(), rtol=1e-3, atol=1e-6 * np.abs(t.grad.numpy()).max())


def test_user_nz_errors_and_served_pool():
    class Broken(jc.redshift.redshift_distribution):
        def pz_fn(self, z):
            raise ZeroDivisionError("no n(z) here")

    class NotFinite(jc.redshift.redshift_distribution):
        def pz_fn(self, z):
            return np.where(z > 1.0, np.nan, 1.0)

    ell = np.geomspace(10.0, 1000.0, 8)
    with pytest.raises(ZeroDivisionError):
        jc.cl.angular_cl(jc.Planck15(), ell, [jc.probes.WeakLensing([Broken()])], n_z=128)
    with pytest.raises(ValueError):
        jc.cl.angular_cl(jc.Planck15(), ell, [jc.probes.WeakLensing([NotFinite()])], n_z=128)
    # the context still serves other sets
    ok = [jc.probes.WeakLensing([Bimodal(zmax=5.0)])]
    jc.cl.angular_cl(jc.Planck15(), ell, ok, n_z=128)
    # a server pool with user n(z) bins: its tables are built when the server
    # is created; the pool's bins share one zmax (the served a-grid's end)
    pool = [jc.probes.WeakLensing([Bimodal(zmax=5.0), jc.redshift.smail_nz(1.0, 2.0, 0.8, zmax=5.0)]),
            jc.probes.NumberCounts([Bimodal(zmax=5.0)], jc.bias.constant_linear_bias(1.3))]
    ell2 = np.geomspace(10.0, 2000.0, 70)
    with serving.Server(pool, max_batch=4, max_grad=2, max_ell=80, slots=2, n_z=(256,)) as srv:
        rid, out, grad = srv.submit(cosmo=COSMO2, ell=ell2, tracers=[0, 2], n_z=256)
        srv.wait(0, timeout_ms=60000)
        comps = srv.poll()
    assert [c["status"] for c in comps] == [0]
    alone = jc.cl.angular_cl(jc.Cosmology(*COSMO2), ell2, [jc.probes.WeakLensing([Bimodal(zmax=5.0)]),
                                                          jc.probes.NumberCounts([Bimodal(zmax=5.0)],
                                                                                 jc.bias.constant_linear_bias(1.3))],
                             n_z=256)
    assert rel(out, alone) < 1e-5
    # a request whose bins end below the pool's zmax would be integrated on
    # another a-grid than angular_cl of its bins: refused
    mixed = [jc.probes.WeakLensing([Bimodal(zmax=5.0), jc.redshift.smail_nz(1.0, 2.0, 0.8)])]
    with serving.Server(mixed, max_batch=4, max_grad=2, max_ell=80, slots=2, n_z=(256,)) as srv:
        with pytest.raises(NotImplementedError):
            srv.submit(cosmo=COSMO2, ell=ell2, tracers=[0], n_z=256)
        srv.submit(cosmo=COSMO2, ell=ell2, tracers=[1], n_z=256)
        srv.wait(0, timeout_ms=60000)
        assert [c["status"] for c in srv.poll()] == [0]
