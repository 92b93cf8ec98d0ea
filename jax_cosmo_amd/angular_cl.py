"""Limber angular power spectra (jax_cosmo/angular_cl.py:15-196).

``angular_cl`` is the drop-in for the reference's entry point: the probe
list is registered once with the device context (its n(z) tables stay
resident), and the C_ell are computed by the HIP kernels behind the C ABI.
The covariance helpers are float64: host arrays for host inputs, torch
tensors (differentiable) when the C_ell come from a cosmology with torch
parameters.
"""
import numpy as np

from . import power
from . import transfer as tklib
from ._native import default_context
from .core import PARAM_NAMES

__all__ = ["angular_cl", "angular_cl_batch", "BatchGraph", "noise_cl", "gaussian_cl_covariance",
           "gaussian_cl_covariance_and_mean"]

DEFAULT_N_Z = 512  # angular_cl.py:96


def _get_cl_ordering(probes):
    """angular_cl.py:15-25."""
    n_tracers = sum([p.n_tracers for p in probes])
    cl_index = []
    for i in range(n_tracers):
        for j in range(i, n_tracers):
            cl_index.append((i, j))
    return cl_index


def _n_cls(probes):
    """len(_get_cl_ordering(probes)): pairs i <= j of the tracers."""
    n = sum(p.n_tracers for p in probes)
    return n * (n + 1) // 2


def _get_cov_blocks_ordering(probes):
    """angular_cl.py:28-46."""
    cl_index = _get_cl_ordering(probes)
    pos = {p: q for q, p in enumerate(cl_index)}

    def find_index(a, b):
        return pos[(a, b)] if (a, b) in pos else pos[(b, a)]

    cov_blocks = []
    for i, j in cl_index:
        for m, n in cl_index:
            cov_blocks.append((find_index(i, m), find_index(j, n), find_index(i, n), find_index(j, m)))
    return cov_blocks


def _check_fns(transfer_fn, nonlinear_fn, gamma=None):
    """Select the device models of the reference's transfer_fn and
    nonlinear_fn arguments (NotImplementedError for others) and the growth
    model of a Cosmology's gamma (None: the growth ODE)."""
    tm = tklib.transfer_model(transfer_fn)
    hm = power.halofit_model(nonlinear_fn)
    ctx = default_context()
    ctx.set_models(tm, hm)
    ctx.set_growth(gamma)


def tracer_entries(probes):
    """Flatten a probe list into device tracer descriptors (probe order, then
    bin order: the tracer order of angular_cl.py:15-25)."""
    entries = []
    for pi, p in enumerate(probes):
        if not hasattr(p, "tracer_entries"):
            raise NotImplementedError("unsupported probe type %s" % type(p).__name__)
        for e in p.tracer_entries():
            e = dict(e)
            e["probe"] = pi
            entries.append(e)
    if not 1 <= len(entries) <= 10:
        raise NotImplementedError("between 1 and 10 tracers are supported, got %d" % len(entries))
    return entries


def _is_torch(x):
    try:
        import torch
    except ImportError:  # pragma: no cover
        return False
    return isinstance(x, torch.Tensor)


def _n_ell(ell):
    return int(ell.reshape(-1).shape[0]) if _is_torch(ell) else int(np.asarray(ell).size)


def _empty(shape, like):
    """A float32 array of ``shape`` (with a zero extent) in the array library,
    and on the device, of ``like`` (numpy when ``like`` is None)."""
    if like is None:
        return np.zeros(shape, dtype=np.float32)
    import torch

    return torch.zeros(shape, dtype=torch.float32, device=like.device)


def angular_cl(cosmo, ell, probes, transfer_fn=tklib.Eisenstein_Hu, nonlinear_fn=power.halofit,
               n_z=DEFAULT_N_Z):
    """Angular C_ell of all tracer pairs under Limber (angular_cl.py:49-98).

    Returns [n_cls, n_ell] (float32 numpy; a torch tensor on the GPU with
    autograd when the cosmology holds torch parameters)."""
    _check_fns(transfer_fn, nonlinear_fn, cosmo.gamma)
    if n_z % 2 == 1:
        raise ValueError("N must be an even integer.")
    ctx = default_context()
    set_id = ctx.tracer_set(tracer_entries(probes))
    P = _n_cls(probes)
    params = cosmo.params()
    if _n_ell(ell) == 0:
        # vmap over an empty ell axis (angular_cl.py:66) gives [n_cls, 0]
        return _empty((P, 0), next((p for p in params + [ell] if _is_torch(p)), None))
    if any(_is_torch(p) for p in params):
        from .autograd import angular_cl_torch

        return angular_cl_torch(params, ell, probes, n_z=n_z)[0]
    ell = np.atleast_1d(np.asarray(ell, dtype=np.float64))
    theta = np.array([float(p) for p in params], dtype=np.float64)
    return ctx.angular_cl(theta, set_id, ell, n_z, P)[0]


def angular_cl_batch(cosmo_params, ell, probes, n_z=DEFAULT_N_Z, transfer_fn=tklib.Eisenstein_Hu,
                     nonlinear_fn=power.halofit, gamma=None):
    """C_ell for a batch of cosmologies in one device pass.

    cosmo_params: [B, 8] in Cosmology(...) order (Omega_c, Omega_b, h, n_s,
    sigma8, Omega_k, w0, wa), numpy or a CUDA torch tensor (float64); gamma:
    the growth index shared by the batch (None: the growth ODE).
    Returns [B, n_cls, n_ell] float32 (same array library as the input)."""
    from . import autotune

    _check_fns(transfer_fn, nonlinear_fn, gamma)
    ctx = default_context()
    set_id = ctx.tracer_set(tracer_entries(probes))
    P = _n_cls(probes)
    T = sum(p.n_tracers for p in probes)
    shape = tuple(cosmo_params.shape) if hasattr(cosmo_params, "shape") else np.shape(cosmo_params)
    if len(shape) not in (1, 2) or shape[-1] != 8:
        raise ValueError("cosmo_params must be [B, 8] (Omega_c, Omega_b, h, n_s, sigma8, Omega_k, "
                         "w0, wa), got shape %s" % (shape,))
    B = int(shape[0]) if len(shape) == 2 else 1
    n_ell = _n_ell(ell)
    if B == 0 or n_ell == 0:
        return _empty((B, P, n_ell), cosmo_params if _is_torch(cosmo_params) else None)
    autotune.apply_batch(ctx, B, n_ell, n_z, T)
    if _is_torch(cosmo_params) and cosmo_params.is_cuda:
        import torch

        cp = cosmo_params.to(torch.float64).reshape(B, 8).contiguous()
        el = _device_ell(ell, cp.device).reshape(-1)
        out = torch.empty((B, P, el.shape[0]), dtype=torch.float32, device=cp.device)
        ctx.angular_cl_dev(cp.data_ptr(), B, set_id, el.data_ptr(), el.shape[0], n_z,
                           out.data_ptr(), stream=torch.cuda.current_stream(cp.device).cuda_stream)
        return out
    cp = np.asarray(cosmo_params, dtype=np.float64).reshape(-1, 8)
    if _is_torch(ell):
        ell = ell.detach().cpu().numpy()
    return ctx.angular_cl(cp, set_id, np.asarray(ell, dtype=np.float64), n_z, P)


class BatchGraph:
    """A batch of angular_cl captured once as a HIP graph and replayed.

    params ([B, 8] CUDA float64) and ell stay where they are: update params
    in place between runs to evaluate other cosmologies.  The kernels are
    those of jc_angular_cl_noalloc (the tracer set's tables for n_z are built
    here first, the per-call buffers live in the library's static scratch),
    so the graph holds kernel launches only: one graph launch per batch.
    run() returns the graph's output tensor [B, n_cls, n_ell] (float32),
    overwritten by the next run.  Graphs share the scratch of their GPU: run
    them one at a time."""

    def __init__(self, params, ell, probes, n_z=DEFAULT_N_Z, transfer_fn=tklib.Eisenstein_Hu,
                 nonlinear_fn=power.halofit, gamma=None):
        import torch

        from . import autotune

        if not (_is_torch(params) and params.is_cuda and params.dtype == torch.float64 and params.dim() == 2
                and params.shape[1] == 8 and params.is_contiguous()):
            raise ValueError("params must be a contiguous [B, 8] CUDA float64 tensor")
        _check_fns(transfer_fn, nonlinear_fn, gamma)
        ctx = default_context()
        self.set_id = ctx.tracer_set(tracer_entries(probes))
        P = _n_cls(probes)
        B = params.shape[0]
        self.params = params
        if _is_torch(ell):
            self.ell = ell.detach().to(device=params.device, dtype=torch.float64).reshape(-1).contiguous()
        else:
            self.ell = torch.as_tensor(np.asarray(ell, dtype=np.float64),
                                       device=params.device).reshape(-1).contiguous()
        n_ell = self.ell.shape[0]
        if B < 1 or n_ell < 1:
            raise ValueError("empty batch")
        self.out = torch.empty((B, P, n_ell), dtype=torch.float32, device=params.device)
        # the resident tables of this set and n_z, and a first (eager) run
        angular_cl_batch(params[:1], self.ell, probes, n_z=n_z, transfer_fn=transfer_fn,
                         nonlinear_fn=nonlinear_fn, gamma=gamma)
        # the tiles of the whole batch's shape (the call above applied those of one cosmology);
        # the captured launches keep them
        autotune.apply_batch(ctx, B, n_ell, n_z, sum(p.n_tracers for p in probes))
        args = (params.data_ptr(), B, self.set_id, self.ell.data_ptr(), n_ell, n_z, self.out.data_ptr())
        ctx.angular_cl_noalloc_dev(*args, stream=torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            ctx.angular_cl_noalloc_dev(*args, stream=torch.cuda.current_stream().cuda_stream)
        ctx.reset_stream()

    def run(self):
        # the physics models and tiles are those of the capture (kernel arguments)
        self.graph.replay()
        return self.out


_ELL_CACHE = {}
_ELL_CACHE_MAX = 16


def _device_ell(ell, device):
    """ell as a float64 tensor on ``device``.  A device tensor is used as it
    is; a host array is uploaded once and kept (keyed by its bytes), so that
    repeated batches over the same multipoles copy nothing to the device."""
    import torch

    if _is_torch(ell):
        return ell.to(device=device, dtype=torch.float64).contiguous()
    a = np.ascontiguousarray(np.atleast_1d(np.asarray(ell, dtype=np.float64)))
    key = (str(device), a.tobytes())
    el = _ELL_CACHE.get(key)
    if el is None:
        if len(_ELL_CACHE) >= _ELL_CACHE_MAX:
            _ELL_CACHE.pop(next(iter(_ELL_CACHE)))
        el = torch.as_tensor(a, device=device).contiguous()
        _ELL_CACHE[key] = el
    return el


def noise_cl(ell, probes):
    """angular_cl.py:101-117: noise on the auto-spectra, [n_cls, n_ell]."""
    n_ell = len(np.atleast_1d(ell))
    noise = np.concatenate([np.atleast_1d(p.noise()) for p in probes])
    cl_index = _get_cl_ordering(probes)
    out = np.zeros((len(cl_index), n_ell))
    for q, (i, j) in enumerate(cl_index):
        delta = 1.0 - np.clip(np.abs(i - j), 0.0, 1.0)
        out[q] = noise[i] * delta * np.ones(n_ell)
    return out


def gaussian_cl_covariance(ell, probes, cl_signal, cl_noise, f_sky=0.25, sparse=True):
    """angular_cl.py:120-163 (float64).  With a torch cl_signal (the C_ell of
    a cosmology with torch parameters) the covariance is a torch tensor on the
    same device, built with differentiable tensor operations, so the
    cosmology dependence of the covariance reaches autograd as in the
    reference (which differentiates through it with jax.grad)."""
    if _is_torch(cl_signal) or _is_torch(cl_noise):
        return _gaussian_cl_covariance_torch(ell, probes, cl_signal, cl_noise, f_sky, sparse)
    ell = np.atleast_1d(np.asarray(ell, dtype=np.float64))
    n_ell = len(ell)
    cl_obs = np.asarray(cl_signal, dtype=np.float64) + np.asarray(cl_noise, dtype=np.float64)
    n_cls = cl_obs.shape[0]
    norm = (2 * ell + 1) * np.gradient(ell) * f_sky
    blocks = np.array(_get_cov_blocks_ordering(probes))
    cov = (cl_obs[blocks[:, 0]] * cl_obs[blocks[:, 1]] + cl_obs[blocks[:, 2]] * cl_obs[blocks[:, 3]]) / norm
    if sparse:
        return cov.reshape((n_cls, n_cls, n_ell))
    cov = cov.reshape((n_cls, n_cls, n_ell, 1)) * np.eye(n_ell)
    return cov.transpose((0, 2, 1, 3)).reshape((n_ell * n_cls, n_ell * n_cls))


def _gaussian_cl_covariance_torch(ell, probes, cl_signal, cl_noise, f_sky, sparse):
    import torch

    like = cl_signal if _is_torch(cl_signal) else cl_noise
    dev = like.device
    ell = np.atleast_1d(np.asarray(ell, dtype=np.float64))
    n_ell = len(ell)
    sig = torch.as_tensor(cl_signal, device=dev).to(torch.float64)
    noi = torch.as_tensor(cl_noise, device=dev).to(torch.float64)
    cl_obs = sig + noi
    n_cls = cl_obs.shape[0]
    norm = torch.as_tensor((2 * ell + 1) * np.gradient(ell) * f_sky, device=dev)
    blocks = torch.as_tensor(np.array(_get_cov_blocks_ordering(probes)), device=dev)
    cov = (cl_obs[blocks[:, 0]] * cl_obs[blocks[:, 1]] + cl_obs[blocks[:, 2]] * cl_obs[blocks[:, 3]]) / norm
    if sparse:
        return cov.reshape(n_cls, n_cls, n_ell)
    # dense: entry (p * n_ell + l, q * n_ell + l) holds block (p, q) at ell_l
    p, q, l = np.meshgrid(np.arange(n_cls), np.arange(n_cls), np.arange(n_ell), indexing="ij")
    rows = torch.as_tensor((p * n_ell + l).reshape(-1), device=dev)
    cols = torch.as_tensor((q * n_ell + l).reshape(-1), device=dev)
    out = torch.zeros((n_cls * n_ell, n_cls * n_ell), dtype=torch.float64, device=dev)
    return out.index_put((rows, cols), cov.reshape(-1))


def gaussian_cl_covariance_and_mean(cosmo, ell, probes, transfer_fn=tklib.Eisenstein_Hu,
                                    nonlinear_fn=power.halofit, f_sky=0.25, sparse=False,
                                    n_z=DEFAULT_N_Z):
    """angular_cl.py:166-196: (cl_signal.flatten(), covariance).  For a
    cosmology with torch parameters both are torch tensors and both carry
    its dependence to autograd (the mean through the device adjoint, the
    covariance through the C_ell it is built from)."""
    ell = np.atleast_1d(np.asarray(ell, dtype=np.float64))
    cl_signal = angular_cl(cosmo, ell, probes, transfer_fn=transfer_fn, nonlinear_fn=nonlinear_fn,
                           n_z=n_z)
    cl_noise = noise_cl(ell, probes)
    cov = gaussian_cl_covariance(ell, probes, cl_signal, cl_noise, f_sky, sparse)
    return cl_signal.flatten() if not _is_torch(cl_signal) else cl_signal.reshape(-1), cov


PARAMS = PARAM_NAMES
