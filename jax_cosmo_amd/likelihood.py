"""Gaussian likelihood (jax_cosmo/likelihood.py:9-61).

``gaussian_log_likelihood`` is the reference's function (float64 host code,
any covariance form).  ``log_likelihood_batch`` and
``log_likelihood_and_grad`` run on the device for block-diagonal (sparse
format) covariances: the block inverses, the quadratic forms of a batch of
means and, through the adjoint of angular_cl, the gradient with respect to
the cosmology."""
import numpy as np

from . import sparse

__all__ = ["gaussian_log_likelihood", "log_likelihood_batch", "log_likelihood_and_grad", "precision"]


def gaussian_log_likelihood(data, mu, C, include_logdet=True, inverse_method="inverse"):
    r = np.asarray(mu, dtype=np.float64) - np.asarray(data, dtype=np.float64)
    if sparse.is_sparse(C):
        rT_Cinv_r = float(r @ sparse.sparse_dot_vec(sparse.inv(C), r))
    else:
        C = np.asarray(C, dtype=np.float64)
        if inverse_method == "inverse":
            y = np.linalg.inv(C) @ r
        elif inverse_method == "cholesky":
            L = np.linalg.cholesky(C)
            y = np.linalg.solve(L.T, np.linalg.solve(L, r))
        else:
            raise NotImplementedError
        rT_Cinv_r = float(r @ y)
    if not include_logdet:
        return -0.5 * rT_Cinv_r
    if sparse.is_sparse(C):
        _, logdet = sparse.slogdet(C)
    else:
        _, logdet = np.linalg.slogdet(C)
    return -0.5 * (rT_Cinv_r - logdet)


def precision(C):
    """(C^-1, log det C) of a sparse-format covariance [P][P][n_ell], on the
    device (FP64 Gauss-Jordan per ell block)."""
    from ._native import default_context

    return default_context().sparse_inv(C)


def log_likelihood_batch(data, mu, C, include_logdet=True):
    """gaussian_log_likelihood for a batch of means mu [B][P][n_ell] (or one
    [P][n_ell]) against data [P * n_ell] and a sparse covariance: [B] values."""
    from ._native import default_context

    ctx = default_context()
    prec, logdet = ctx.sparse_inv(C)
    # likelihood.py:61 returns -1/2 (r^T C^-1 r - log det C)
    c0 = 0.5 * logdet if include_logdet else 0.0
    mu = np.asarray(mu)
    return ctx.gaussian_loglike(mu, np.asarray(data, dtype=np.float64), prec, c0)


def log_likelihood_and_grad(cosmo_params, ell, probes, data, C, include_logdet=True, n_z=512, transfer_fn=None,
                            nonlinear_fn=None, gamma=None):
    """Log-likelihood of data given the means angular_cl(cosmo_b, ell, probes)
    and its gradient with respect to (Omega_c, Omega_b, h, n_s, sigma8, w0,
    wa) for a batch of cosmologies [B, 8]: ([B], [B, 7])."""
    from . import power, transfer
    from ._native import default_context
    from .angular_cl import _check_fns, tracer_entries

    _check_fns(transfer_fn or transfer.Eisenstein_Hu, nonlinear_fn or power.halofit, gamma)
    ctx = default_context()
    sid = ctx.tracer_set(tracer_entries(probes))
    prec, logdet = ctx.sparse_inv(C)
    c0 = 0.5 * logdet if include_logdet else 0.0  # likelihood.py:61
    return ctx.gaussian_loglike_grad(np.asarray(cosmo_params, dtype=np.float64), sid, ell, n_z,
                                     np.asarray(data, dtype=np.float64), prec, c0)
