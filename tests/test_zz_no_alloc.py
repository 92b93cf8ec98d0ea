"""GPU test of jc_angular_cl_noalloc, the forward entry point that allocates
no device memory (its buffers live in a static per-GPU scratch array), called
through the C ABI against jc_angular_cl on the same inputs.  The two run the
same kernels with the same launch shapes, so the results must be identical;
two calls in a row, the second smaller, check that the scratch left by the
first does not leak into the second.  Also: host and device pointers, the
error for tables that are not resident and for calls larger than the array.
"""
import ctypes

import numpy as np
import pytest
import torch

import jax_cosmo_amd as jc
from jax_cosmo_amd import _native

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


def cosmologies(B, seed):
    th = BOX_LO + (BOX_HI - BOX_LO) * np.random.default_rng(seed).random((B, 7))
    p = np.zeros((B, 8))
    p[:, [0, 1, 2, 3, 4, 6, 7]] = th
    return p


def P_(a):
    return ctypes.c_void_p(a.ctypes.data)


def call(lib, name, ctx, cosmo, sid, ell, n_z, P):
    B = cosmo.shape[0]
    out = np.full((B, P, len(ell)), np.nan, dtype=np.float32)
    rc = getattr(lib, name)(ctx.handle, P_(cosmo), B, sid, P_(ell), len(ell), n_z, P_(out), 0)
    return rc, out


def test_noalloc_equals_allocating_entry_point():
    ctx = _native.default_context()
    lib = _native.load_library()
    sid = ctx.tracer_set(jc.cl.tracer_entries(c2_probes()))
    ctx.reset_stream()
    ell = np.geomspace(10.0, 5000.0, 100)
    # the tracer set's n_z = 512 tables are built by one allocating call
    rc, _ = call(lib, "jc_angular_cl", ctx, cosmologies(1, 0), sid, ell, 512, 55)
    assert rc == 0
    # two calls in a row, the second smaller (fewer cosmologies and ell, no
    # transfer table below 64 ell, sliced a range): identical to the
    # allocating entry point on the same inputs
    for B, n_ell, seed in ((256, 100, 1), (3, 40, 2)):
        cosmo = cosmologies(B, seed)
        el = np.ascontiguousarray(np.geomspace(10.0, 5000.0, n_ell))
        print("noalloc call: B=%d n_ell=%d n_z=512" % (B, n_ell), flush=True)
        rc, got = call(lib, "jc_angular_cl_noalloc", ctx, cosmo, sid, el, 512, 55)
        assert rc == 0, lib.jc_last_error().decode()
        rc, ref = call(lib, "jc_angular_cl", ctx, cosmo, sid, el, 512, 55)
        assert rc == 0
        assert np.all(np.isfinite(got))
        np.testing.assert_array_equal(got, ref)


def test_noalloc_device_pointers_and_errors():
    ctx = _native.default_context()
    lib = _native.load_library()
    probes = c2_probes()
    sid = ctx.tracer_set(jc.cl.tracer_entries(probes))
    ctx.reset_stream()
    ell = np.geomspace(10.0, 5000.0, 100)
    rc, _ = call(lib, "jc_angular_cl", ctx, cosmologies(1, 0), sid, ell, 512, 55)
    assert rc == 0
    dev = torch.device("cuda:0")
    for B in (64, 5):
        cosmo = cosmologies(B, 10 + B)
        cp = torch.tensor(cosmo, dtype=torch.float64, device=dev)
        el = torch.tensor(ell, dtype=torch.float64, device=dev)
        out = torch.full((B, 55, 100), float("nan"), dtype=torch.float32, device=dev)
        torch.cuda.synchronize()
        print("noalloc device-pointer call: B=%d" % B, flush=True)
        rc = lib.jc_angular_cl_noalloc(ctx.handle, ctypes.c_void_p(cp.data_ptr()), B, sid,
                                       ctypes.c_void_p(el.data_ptr()), 100, 512, ctypes.c_void_p(out.data_ptr()),
                                       _native.JC_DEVICE_PTRS)
        assert rc == 0, lib.jc_last_error().decode()
        assert lib.jc_synchronize(ctx.handle) == 0
        _, ref = call(lib, "jc_angular_cl", ctx, cosmo, sid, ell, 512, 55)
        np.testing.assert_array_equal(out.cpu().numpy(), ref)
    # tables for n_z = 386 were never built for this set: refused, nothing runs
    rc, _ = call(lib, "jc_angular_cl_noalloc", ctx, cosmologies(2, 3), sid, ell, 386, 55)
    assert rc == -1 and "not resident" in lib.jc_last_error().decode()
    # larger than the scratch array (about 108 KB per cosmology with host
    # output): refused before any launch
    big = cosmologies(2000, 4)
    rc, _ = call(lib, "jc_angular_cl_noalloc", ctx, big, sid, ell, 512, 55)
    assert rc == -1 and "scratch" in lib.jc_last_error().decode()
    # the context still works afterwards
    rc, got = call(lib, "jc_angular_cl_noalloc", ctx, big[:2], sid, ell, 512, 55)
    assert rc == 0
    _, ref = call(lib, "jc_angular_cl", ctx, big[:2], sid, ell, 512, 55)
    np.testing.assert_array_equal(got, ref)
