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


def _torch_args(*xs):
    try:
        import torch
    except ImportError:
        return False
    return any(isinstance(x, torch.Tensor) for x in xs)


def gaussian_log_likelihood(data, mu, C, include_logdet=True, inverse_method="inverse"):
    """-1/2 (r^T C^-1 r - log det C), r = mu - data (likelihood.py:9-61).

    C is the reference's sparse form [P][P][n_ell], a dense [N][N] array or a
    DenseFactor.  Both inverse methods of the reference give the same value;
    on the device the dense form is always solved through its Cholesky
    factor.  The residual is formed here in FP64 and everything after it
    runs in FP64 on the device.  With torch arguments (the mean and
    covariance of a cosmology with torch parameters) the result is a torch
    scalar whose gradient reaches data, mu and C (jax.grad of the
    reference's likelihood, README.md:17-24)."""
    if inverse_method not in ("inverse", "cholesky"):
        raise NotImplementedError
    if _torch_args(data, mu, C):
        return GaussianLogLikelihood.apply(data, mu, C, bool(include_logdet))
    return _loglike_value(data, mu, C, include_logdet)


def _loglike_value(data, mu, C, include_logdet):
    r = np.asarray(data, dtype=np.float64).reshape(-1) - np.asarray(mu, dtype=np.float64).reshape(-1)
    # the device forms data - mu with mu in float32: passing (r, 0) keeps r exact
    return float(log_likelihood_batch(r, np.zeros((1, r.size), np.float32), C, include_logdet=include_logdet)[0])


def _np64(x):
    try:
        import torch

        if isinstance(x, torch.Tensor):
            return x.detach().to(torch.float64).cpu().numpy()
    except ImportError:
        pass
    return np.asarray(x, dtype=np.float64)


def _make_function():
    import torch

    class _GaussianLogLikelihood(torch.autograd.Function):
        """The device log-likelihood as a torch function of (data, mu, C).

        With s = C^-1 (data - mu): dL/dmu = s, dL/ddata = -s and, for the
        reference's L = -1/2 (r^T C^-1 r - log det C), dL/dC = 1/2 (s s^T
        + C^-1) entry by entry (1/2 s s^T without the log det).  s comes from
        the device likelihood kernels, C^-1 from the block inverses (sparse
        form) or from Cholesky solves (dense form)."""

        @staticmethod
        def forward(ctx, data, mu, C, include_logdet):
            from ._native import default_context

            dctx = default_context()
            d = _np64(data).reshape(-1)
            m = _np64(mu).reshape(-1)
            r = d - m
            zero = np.zeros((1, r.size), np.float32)
            Cn = C if isinstance(C, DenseFactor) else _np64(C)
            if _is_dense(Cn):
                f = cholesky(Cn)
                c0 = 0.5 * f.logdet if include_logdet else 0.0
                ll, s = dctx.gaussian_loglike_dense(zero, r, f.L, c0, want_dmu=True)
                ctx.dense = f
                ctx.prec = None
            else:
                prec, logdet = dctx.sparse_inv(Cn)
                c0 = 0.5 * logdet if include_logdet else 0.0
                P, _, n_ell = Cn.shape
                ll, s = dctx.gaussian_loglike(zero.reshape(1, P, n_ell), r, prec, c0, want_dmu=True)
                ctx.dense = None
                ctx.prec = prec
            ctx.s = np.asarray(s, dtype=np.float64).reshape(-1)  # C^-1 (data - mu) from the device
            ctx.include_logdet = include_logdet
            ctx.shapes = [(x.shape, x.dtype, x.device) if isinstance(x, torch.Tensor) else None
                          for x in (data, mu, C)]
            like = next(x for x in (mu, C, data) if isinstance(x, torch.Tensor))
            return torch.tensor(float(ll[0]), dtype=torch.float64, device=like.device)

        @staticmethod
        def backward(ctx, g):
            from ._native import default_context

            gv = float(g)
            s = ctx.s
            out = [None, None, None, None]

            def as_like(arr, k):
                shp, dt, dev = ctx.shapes[k]
                return torch.as_tensor(arr, device=dev).to(dt).reshape(shp)

            if ctx.needs_input_grad[0] and ctx.shapes[0] is not None:
                out[0] = as_like(-gv * s, 0)
            if ctx.needs_input_grad[1] and ctx.shapes[1] is not None:
                out[1] = as_like(gv * s, 1)
            if ctx.needs_input_grad[2] and ctx.shapes[2] is not None:
                dev = ctx.shapes[2][2]
                st = torch.as_tensor(s, device=dev)
                if ctx.prec is not None:
                    P, _, n_ell = ctx.prec.shape
                    sm = st.reshape(P, n_ell)
                    dC = 0.5 * sm[:, None, :] * sm[None, :, :]
                    if ctx.include_logdet:
                        dC = dC + 0.5 * torch.as_tensor(ctx.prec, device=dev)
                else:
                    dC = 0.5 * torch.outer(st, st)
                    if ctx.include_logdet:
                        N = ctx.dense.N
                        eye = torch.eye(N, dtype=torch.float64, device=ctx.dense.L.device)
                        cinv = default_context().dense_solve(ctx.dense.L, eye, mode=2)
                        dC = dC + 0.5 * cinv.to(dev)
                out[2] = (gv * dC).to(ctx.shapes[2][1]).reshape(ctx.shapes[2][0])
            return tuple(out)

    return _GaussianLogLikelihood


class _LazyFunction:
    _fn = None

    def apply(self, *args):
        if _LazyFunction._fn is None:
            _LazyFunction._fn = _make_function()
        return _LazyFunction._fn.apply(*args)


GaussianLogLikelihood = _LazyFunction()


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
