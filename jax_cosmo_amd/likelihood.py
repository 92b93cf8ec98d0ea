"""Gaussian likelihood (jax_cosmo/likelihood.py:9-61) on the device.

``gaussian_log_likelihood`` keeps the reference's signature and evaluates on
the GPU for both covariance forms: block-diagonal (the sparse format
[P][P][n_ell], FP64 block inverses) and dense ([N][N], N = P * n_ell, an FP64
Cholesky factor).  ``log_likelihood_batch`` and ``log_likelihood_and_grad``
take a batch of means or of cosmologies, the latter with the gradient with
respect to the cosmology through the adjoint of angular_cl.  Dense factors
stay resident on the device (``DenseFactor``); the factor of the last dense
covariance seen is reused while the array is unchanged.
"""
import numpy as np

__all__ = ["gaussian_log_likelihood", "log_likelihood_batch", "log_likelihood_and_grad", "precision",
           "cholesky", "DenseFactor"]


class DenseFactor:
    """Cholesky factor L (C = L L^T) of a dense covariance, on the device:
    ``L`` a CUDA float64 tensor [N][N], ``logdet`` = log det C."""

    def __init__(self, L, logdet):
        self.L = L
        self.logdet = float(logdet)
        self.N = int(L.shape[0])


_last = {"key": None, "ref": None, "factor": None}


def _fingerprint(C):
    # identity, address and a few sums: cheap next to the factorisation
    d = np.diagonal(C)
    return (id(C), C.__array_interface__["data"][0], C.shape, float(d.sum()), float(C[0].sum()), float(C[-1].sum()),
            float(C[:, C.shape[1] // 2].sum()))


def cholesky(C):
    """DenseFactor of a dense covariance (FP64 Cholesky on the device)."""
    from ._native import default_context

    if isinstance(C, DenseFactor):
        return C
    C = np.asarray(C, dtype=np.float64)
    key = _fingerprint(C)
    if _last["key"] == key and _last["ref"] is C:
        return _last["factor"]
    L, logdet = default_context().dense_cholesky(C)
    f = DenseFactor(L, logdet)
    _last.update(key=key, ref=C, factor=f)
    return f


def _is_dense(C):
    return isinstance(C, DenseFactor) or np.asarray(C).ndim == 2


def gaussian_log_likelihood(data, mu, C, include_logdet=True, inverse_method="inverse"):
    """-1/2 (r^T C^-1 r - log det C), r = mu - data (likelihood.py:9-61).

    C is the reference's sparse form [P][P][n_ell], a dense [N][N] array or a
    DenseFactor.  Both inverse methods of the reference give the same value;
    on the device the dense form is always solved through its Cholesky
    factor.  The residual is formed here in FP64 and everything after it
    runs in FP64 on the device."""
    if inverse_method not in ("inverse", "cholesky"):
        raise NotImplementedError
    r = np.asarray(data, dtype=np.float64).reshape(-1) - np.asarray(mu, dtype=np.float64).reshape(-1)
    # the device forms data - mu with mu in float32: passing (r, 0) keeps r exact
    return float(log_likelihood_batch(r, np.zeros((1, r.size), np.float32), C, include_logdet=include_logdet)[0])


def precision(C):
    """(C^-1, log det C) of a sparse-format covariance [P][P][n_ell], on the
    device (FP64 Gauss-Jordan per ell block)."""
    from ._native import default_context

    return default_context().sparse_inv(C)


def log_likelihood_batch(data, mu, C, include_logdet=True):
    """gaussian_log_likelihood for a batch of float32 means (the device's
    C_ell) against data [P * n_ell]: mu [B][P][n_ell] (or [B][N]; one mean may
    drop B), C sparse, dense or a DenseFactor: [B] values."""
    from ._native import default_context

    ctx = default_context()
    data = np.asarray(data, dtype=np.float64)
    if _is_dense(C):
        f = cholesky(C)
        c0 = 0.5 * f.logdet if include_logdet else 0.0  # likelihood.py:61
        mu = np.asarray(mu).reshape(-1, f.N)
        return ctx.gaussian_loglike_dense(mu, data, f.L, c0)
    prec, logdet = ctx.sparse_inv(C)
    # likelihood.py:61 returns -1/2 (r^T C^-1 r - log det C)
    c0 = 0.5 * logdet if include_logdet else 0.0
    P, _, n_ell = np.shape(C)
    mu = np.asarray(mu).reshape(-1, P, n_ell)
    return ctx.gaussian_loglike(mu, data, prec, c0)


def log_likelihood_and_grad(cosmo_params, ell, probes, data, C, include_logdet=True, n_z=512, transfer_fn=None,
                            nonlinear_fn=None, gamma=None):
    """Log-likelihood of data given the means angular_cl(cosmo_b, ell, probes)
    and its gradient with respect to (Omega_c, Omega_b, h, n_s, sigma8, w0,
    wa) for a batch of cosmologies [B, 8]: ([B], [B, 7]).  C sparse, dense or
    a DenseFactor."""
    from . import power, transfer
    from ._native import default_context
    from .angular_cl import _check_fns, tracer_entries

    _check_fns(transfer_fn or transfer.Eisenstein_Hu, nonlinear_fn or power.halofit, gamma)
    ctx = default_context()
    sid = ctx.tracer_set(tracer_entries(probes))
    params = np.asarray(cosmo_params, dtype=np.float64)
    data = np.asarray(data, dtype=np.float64)
    if _is_dense(C):
        f = cholesky(C)
        c0 = 0.5 * f.logdet if include_logdet else 0.0  # likelihood.py:61
        return ctx.gaussian_loglike_grad_dense(params, sid, ell, n_z, data, f.L, c0)
    prec, logdet = ctx.sparse_inv(C)
    c0 = 0.5 * logdet if include_logdet else 0.0  # likelihood.py:61
    return ctx.gaussian_loglike_grad(params, sid, ell, n_z, data, prec, c0)
