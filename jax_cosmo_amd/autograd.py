"""torch.autograd binding of the device derivative kernels.

``AngularClFunction`` runs the forward C_ell kernel and, in backward, the
hand-written adjoint (jc_angular_cl_vjp): the cotangent of C_ell is
contracted with dC/dtheta inside the fused Limber kernel, for
theta = (Omega_c, Omega_b, h, n_s, sigma8, w0, wa).  This replaces
``jax.grad`` through the reference's pytrees (core.py:98-141).
"""
import numpy as np
import torch

from ._native import default_context

GRAD_INDEX = (0, 1, 2, 3, 4, 6, 7)  # positions of theta in the 8-parameter vector


def _assemble(theta, omega_k):
    B = theta.shape[0]
    params = torch.empty((B, 8), dtype=torch.float64, device=theta.device)
    params[:, list(GRAD_INDEX)] = theta.detach().to(torch.float64)
    params[:, 5] = omega_k
    return params.contiguous()


class AngularClFunction(torch.autograd.Function):
    """C[B, P, n_ell] = angular_cl(theta[B, 7]); float32 output."""

    @staticmethod
    def forward(ctx, theta, omega_k, ell, set_id, n_z, n_pairs):
        dctx = default_context()
        params = _assemble(theta, omega_k)
        ell64 = np.ascontiguousarray(np.asarray(ell, dtype=np.float64).reshape(-1))
        if theta.is_cuda:
            el = torch.as_tensor(ell64, device=theta.device)
            out = torch.empty((params.shape[0], n_pairs, el.shape[0]), dtype=torch.float32,
                              device=theta.device)
            dctx.angular_cl_dev(params.data_ptr(), params.shape[0], set_id, el.data_ptr(),
                                el.shape[0], n_z, out.data_ptr(),
                                stream=torch.cuda.current_stream(theta.device).cuda_stream)
        else:
            out = torch.from_numpy(dctx.angular_cl(params.numpy(), set_id, ell64, n_z, n_pairs))
        ctx.save_for_backward(params)
        ctx.cfg = (ell64, set_id, n_z, theta.dtype, dctx.models())
        return out

    @staticmethod
    def backward(ctx, grad):
        (params,) = ctx.saved_tensors
        ell64, set_id, n_z, dtype, models = ctx.cfg
        dctx = default_context()
        # the models of the forward call, whatever calls came in between
        dctx.set_models(models[0], models[1])
        dctx.set_growth(models[2])
        g32 = grad.detach().to(torch.float32).contiguous()
        if params.is_cuda:
            el = torch.as_tensor(ell64, device=params.device)
            th = torch.empty((params.shape[0], 7), dtype=torch.float64, device=params.device)
            dctx.angular_cl_vjp_dev(params.data_ptr(), params.shape[0], set_id, el.data_ptr(),
                                    el.shape[0], n_z, g32.data_ptr(), th.data_ptr(),
                                    stream=torch.cuda.current_stream(params.device).cuda_stream)
        else:
            th = torch.from_numpy(dctx.angular_cl_vjp(params.numpy(), set_id, ell64, n_z,
                                                      g32.numpy()))
        return th.to(dtype), None, None, None, None, None


def angular_cl_torch(params, ell, probes, n_z=512):
    """angular_cl for a Cosmology whose parameters are (some) torch scalars.
    Returns [B=1, n_cls, n_ell]; gradients reach the torch parameters."""
    from .angular_cl import _n_cls, tracer_entries

    dev = None
    for p in params:
        if isinstance(p, torch.Tensor):
            dev = p.device
            break
    vals = []
    for i in GRAD_INDEX:
        p = params[i]
        vals.append(p.to(torch.float64).reshape(()) if isinstance(p, torch.Tensor)
                    else torch.tensor(float(p), dtype=torch.float64, device=dev))
    theta = torch.stack(vals).reshape(1, 7)
    ok = params[5]
    omega_k = float(ok.detach().cpu()) if isinstance(ok, torch.Tensor) else float(ok)
    dctx = default_context()
    set_id = dctx.tracer_set(tracer_entries(probes))
    P = _n_cls(probes)
    return AngularClFunction.apply(theta, omega_k, ell, set_id, n_z, P)


def angular_cl_jacobian(cosmo_params, ell, probes, n_z=512, transfer_fn=None, nonlinear_fn=None, gamma=None):
    """dC/dtheta [B, P, n_ell, 7] by seven device JVPs (forward mode), the
    counterpart of jax.jacfwd of the mean (notebook jax-cosmo-intro, cell 46)."""
    from . import power, transfer
    from .angular_cl import _check_fns, _n_cls, tracer_entries

    _check_fns(transfer_fn or transfer.Eisenstein_Hu, nonlinear_fn or power.halofit, gamma)
    dctx = default_context()
    set_id = dctx.tracer_set(tracer_entries(probes))
    P = _n_cls(probes)
    cp = np.asarray(cosmo_params, dtype=np.float64).reshape(-1, 8)
    B = cp.shape[0]
    jac = np.empty((B, P, len(np.atleast_1d(ell)), 7), dtype=np.float32)
    for d in range(7):
        t = np.zeros((B, 7))
        t[:, d] = 1.0
        _, dot = dctx.angular_cl_jvp(cp, t, set_id, np.asarray(ell, dtype=np.float64), n_z, P)
        jac[..., d] = dot
    return jac


def fisher_matrix(cosmo_params, ell, probes, C, n_z=512, transfer_fn=None, nonlinear_fn=None, gamma=None):
    """Fisher matrix F = J^T C^-1 J [B, 7, 7] of the C_ell mean for a fixed
    block-diagonal covariance C (sparse format [P, P, n_ell], e.g.
    gaussian_cl_covariance(..., sparse=True)), with J the device Jacobian over
    (Omega_c, Omega_b, h, n_s, sigma8, w0, wa): the reference's
    jax.jacfwd-based Fisher forecast (README.md, notebook jax-cosmo-intro)."""
    jac = angular_cl_jacobian(cosmo_params, ell, probes, n_z=n_z, transfer_fn=transfer_fn,
                              nonlinear_fn=nonlinear_fn, gamma=gamma).astype(np.float64)  # [B, P, L, 7]
    prec, _ = default_context().sparse_inv(C)  # [P, P, L]
    return np.einsum("bpla,pql,bqlc->bac", jac, prec, jac)


def log_likelihood_hessian(cosmo_params, ell, probes, data, C, n_z=512, transfer_fn=None, nonlinear_fn=None,
                           gamma=None):
    """Hessian [B, 7, 7] of the Gaussian log-likelihood -1/2 r^T C^-1 r,
    r = data - mu(theta), over (Omega_c, Omega_b, h, n_s, sigma8, w0, wa)
    for a fixed covariance C (sparse [P, P, n_ell], dense [N, N] or a
    likelihood.DenseFactor), at cosmologies whose mean equals the data:
    jax.hessian of the notebook's likelihood at its fiducial cosmology
    (jax-cosmo-intro cells 35-39, F = -H), where

        H = -J^T C^-1 J,   J = dmu/dtheta (seven device JVPs).

    Away from such a point the full Hessian adds sum_k (C^-1 r)_k d2mu_k,
    which this function leaves out (it then returns the Gauss-Newton
    Hessian): the device has first derivatives only, and central differences
    of the adjoint do not recover the term, because mu(theta) of the
    reference jumps wherever the nearest node of its halofit root
    interpolation changes (about every 1e-4 in Omega_c)."""
    from . import likelihood as lk

    cp = np.asarray(cosmo_params, dtype=np.float64).reshape(-1, 8)
    B = cp.shape[0]
    ell = np.asarray(ell, dtype=np.float64).reshape(-1)
    jac = angular_cl_jacobian(cp, ell, probes, n_z=n_z, transfer_fn=transfer_fn, nonlinear_fn=nonlinear_fn,
                              gamma=gamma).astype(np.float64)  # [B, P, L, 7]
    dctx = default_context()
    if lk._is_dense(C):
        f = lk.cholesky(C)
        Jm = jac.reshape(B, -1, 7)
        cj = np.stack([dctx.dense_solve(f.L, Jm[b].T.copy(), mode=2) for b in range(B)])  # [B, 7, N]
        H = -np.einsum("bna,bcn->bac", Jm, cj)
    else:
        prec, _ = dctx.sparse_inv(C)
        H = -np.einsum("bpla,pql,bqlc->bac", jac, prec, jac)
    return 0.5 * (H + np.transpose(H, (0, 2, 1)))
