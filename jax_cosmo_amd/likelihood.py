"""Gaussian likelihood (jax_cosmo/likelihood.py:9-61), float64 host."""
import numpy as np

from . import sparse

__all__ = ["gaussian_log_likelihood"]


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
