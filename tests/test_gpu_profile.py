"""GPU test of the per-kernel profiling hooks (jc_profile_enable /
jc_profile_select / jc_profile_read): only the selected kernels' launches are
timed, a selection changes no result, and invalid masks are refused."""
import numpy as np
import pytest
import torch

import jax_cosmo_amd as jc
from jax_cosmo_amd import _native

pytestmark = pytest.mark.gpu


def probes():
    wl = jc.probes.WeakLensing([jc.redshift.smail_nz(1.0, 2.0, z0) for z0 in (0.5, 1.0)], sigma_e=0.26)
    nc = jc.probes.NumberCounts([jc.redshift.gaussian_nz(0.6, 0.1)], jc.bias.constant_linear_bias(1.5))
    return [wl, nc]


def test_profile_times_only_selected_kernels():
    ctx = _native.default_context()
    pr = probes()
    ell = np.geomspace(10.0, 3000.0, 40)
    p = np.tile(np.array([0.25, 0.05, 0.7, 0.96, 0.8, 0.0, -1.0, 0.0]), (6, 1))
    p[:, 0] += 0.01 * np.arange(6)
    params = torch.tensor(p, dtype=torch.float64, device="cuda:0")
    print("profile select: B=6 n_ell=40", flush=True)
    ref = jc.cl.angular_cl_batch(params, ell, pr, n_z=256)
    torch.cuda.synchronize()

    ctx.profile(True, kernels=("limber",))
    ctx.profile_read()  # reset
    out = jc.cl.angular_cl_batch(params, ell, pr, n_z=256)
    got = ctx.profile_read()
    ctx.profile(False)
    assert got["limber"][1] >= 1 and got["limber"][0] > 0.0
    assert got["tables"] == (0.0, 0) and got["wl"] == (0.0, 0)
    assert torch.equal(out, ref)

    ctx.profile(True)  # every kernel again
    ctx.profile_read()
    jc.cl.angular_cl_batch(params, ell, pr, n_z=256)
    got = ctx.profile_read()
    ctx.profile(False)
    for k in ("tables", "wl", "limber"):
        assert got[k][1] >= 1 and got[k][0] > 0.0, k

    # nothing is recorded while profiling is off
    jc.cl.angular_cl_batch(params, ell, pr, n_z=256)
    assert all(v == (0.0, 0) for v in ctx.profile_read().values())

    lib = _native.load_library()
    for bad in (0, 1 << len(_native.KERNELS)):
        with pytest.raises(ValueError):
            _native.check(lib.jc_profile_select(ctx.handle, bad))
