"""Transfer function (jax_cosmo/transfer.py:10-156).

``Eisenstein_Hu`` identifies the EH98 transfer function, with baryon wiggles
(type='eisenhu_osc', the reference default) or without (type='eisenhu'); the
device kernels evaluate either.  Called directly it returns T(k) from the
device.  A caller selects the no-wiggle form the reference's way, with
``functools.partial(Eisenstein_Hu, type="eisenhu")`` as ``transfer_fn``.
"""
import functools

import numpy as np

__all__ = ["Eisenstein_Hu"]

TYPES = {"eisenhu_osc": 0, "eisenhu": 1}


def Eisenstein_Hu(cosmo, k, type="eisenhu_osc"):
    """T(k), k in h/Mpc (transfer.py:10-156), recovered from the device
    linear spectrum as sqrt(P_lin(k, a=1) / (k^n_s pknorm)) (D(1) = 1)."""
    if type not in TYPES:
        raise NotImplementedError(type)
    from . import power

    k = np.atleast_1d(np.asarray(k, dtype=float))
    fn = Eisenstein_Hu if type == "eisenhu_osc" else functools.partial(Eisenstein_Hu, type=type)
    plin = power.linear_matter_power(cosmo, k, a=1.0, transfer_fn=fn)
    pknorm = power.sigma8_norm(cosmo, transfer_fn=fn)
    return np.sqrt(plin / (k ** float(cosmo.n_s) * pknorm))


def transfer_model(transfer_fn, **kwargs):
    """Device code of a transfer_fn: Eisenstein_Hu itself, or a
    functools.partial of it fixing ``type``; anything else is refused."""
    fn, kw = transfer_fn, dict(kwargs)
    if isinstance(fn, functools.partial):
        kw = {**fn.keywords, **kw}
        fn = fn.func
    if fn is not Eisenstein_Hu or set(kw) - {"type"}:
        raise NotImplementedError("only transfer.Eisenstein_Hu is supported on the device path")
    t = kw.get("type", "eisenhu_osc")
    if t not in TYPES:
        raise NotImplementedError("Eisenstein_Hu type %r" % (t,))
    return TYPES[t]
