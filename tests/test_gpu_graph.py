"""GPU test of the batch path captured as one HIP graph (BatchGraph: the
kernels of jc_angular_cl_noalloc under torch.cuda.graph capture): every
replay, with eager work and synchronisations in between and with the
parameters changed in place, equals the eager angular_cl_batch of the same
inputs.
"""
import numpy as np
import pytest
import torch

import jax_cosmo_amd as jc

pytestmark = pytest.mark.gpu

BOX_LO = np.array([0.2, 0.04, 0.6, 0.9, 0.7, -1.3, -0.5])
BOX_HI = np.array([0.35, 0.06, 0.8, 1.0, 0.9, -0.7, 0.5])


def c2_probes():
    wl = jc.probes.WeakLensing(
        [jc.redshift.smail_nz(1.0, 2.0, z0, gals_per_arcmin2=10.0) for z0 in (0.5, 0.75, 1.0, 1.25, 1.5)],
        sigma_e=0.26)
    nc = jc.probes.NumberCounts(
        [jc.redshift.gaussian_nz(mu, 0.1, gals_per_arcmin2=10.0) for mu in np.linspace(0.3, 1.5, 5)],
        jc.bias.constant_linear_bias(1.5))
    return [wl, nc]


def draw(B, seed):
    th = BOX_LO + (BOX_HI - BOX_LO) * np.random.default_rng(seed).random((B, 7))
    p = np.zeros((B, 8))
    p[:, [0, 1, 2, 3, 4, 6, 7]] = th
    return p


@pytest.mark.parametrize("B,n_ell,ell_on_device", [(96, 100, False), (5, 40, False), (7, 64, True)])
def test_batch_graph_replays_equal_eager(B, n_ell, ell_on_device):
    probes = c2_probes()
    ell = np.geomspace(10.0, 5000.0, n_ell)
    dev = torch.device("cuda:0")
    params = torch.tensor(draw(B, 1), dtype=torch.float64, device=dev)
    # the multipoles as a host array or as a CUDA tensor (the batch path's inputs resident in HBM)
    g = jc.cl.BatchGraph(params, torch.tensor(ell, device=dev) if ell_on_device else ell, probes, n_z=512)
    for k in range(4):
        if k >= 2:
            params.copy_(torch.tensor(draw(B, 10 + k), dtype=torch.float64, device=dev))
        print("replay %d: B=%d n_ell=%d" % (k, B, n_ell), flush=True)
        out = g.run().clone()
        torch.cuda.synchronize()
        # eager work between replays
        ref = jc.cl.angular_cl_batch(params.clone(), ell, probes, n_z=512)
        torch.cuda.synchronize()
        assert torch.equal(out, ref)
