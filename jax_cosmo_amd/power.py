"""Matter power spectra (jax_cosmo/power.py:14-272), evaluated on the device.

``halofit`` identifies the Takahashi et al. (2012) halofit
(prescription='takahashi2012', the reference default).
"""
import functools

import numpy as np

from . import transfer as tklib
from ._native import default_context

__all__ = ["primordial_matter_power", "linear_matter_power", "nonlinear_matter_power", "halofit"]


def _params(cosmo):
    return np.array([float(p) for p in cosmo.params()], dtype=np.float64)


def primordial_matter_power(cosmo, k):  # power.py:14-18
    return np.asarray(k, dtype=float) ** float(cosmo.n_s)


def _pairs(k, a):
    k = np.atleast_1d(np.asarray(k, dtype=float))
    a = np.atleast_1d(np.asarray(a, dtype=float))
    k, a = np.broadcast_arrays(k, a)
    return k, a


PRESCRIPTIONS = {"takahashi2012": 0, "smith2003": 1}


def linear_matter_power(cosmo, k, a=1.0, transfer_fn=tklib.Eisenstein_Hu, **kwargs):
    """power.py:21-53 at paired (k, a) (broadcast)."""
    tm = tklib.transfer_model(transfer_fn, **kwargs)
    k, a = _pairs(k, a)
    ctx = default_context()
    ctx.set_models(tm, -1)
    ctx.set_growth(cosmo.gamma)
    plin, _ = ctx.matter_power(_params(cosmo), k.ravel(), a.ravel())
    return plin.reshape(k.shape).squeeze()


def halofit(cosmo, k, a, transfer_fn=tklib.Eisenstein_Hu, prescription="takahashi2012"):
    """power.py:144-262 (Takahashi 2012 or Smith 2003)."""
    if prescription not in PRESCRIPTIONS:
        raise NotImplementedError(prescription)
    tm = tklib.transfer_model(transfer_fn)
    k, a = _pairs(k, a)
    ctx = default_context()
    ctx.set_models(tm, PRESCRIPTIONS[prescription])
    ctx.set_growth(cosmo.gamma)
    _, pnl = ctx.matter_power(_params(cosmo), k.ravel(), a.ravel())
    return pnl.reshape(k.shape).squeeze()


def halofit_model(nonlinear_fn):
    """Device code of a nonlinear_fn: halofit itself, or a functools.partial
    of it fixing ``prescription``; anything else is refused."""
    fn, kw = nonlinear_fn, {}
    if isinstance(fn, functools.partial):
        kw = dict(fn.keywords)
        fn = fn.func
    if fn is not halofit or set(kw) - {"prescription"}:
        raise NotImplementedError("only power.halofit is supported on the device path")
    p = kw.get("prescription", "takahashi2012")
    if p not in PRESCRIPTIONS:
        raise NotImplementedError("halofit prescription %r" % (p,))
    return PRESCRIPTIONS[p]


def nonlinear_matter_power(cosmo, k, a=1.0, transfer_fn=tklib.Eisenstein_Hu, nonlinear_fn=halofit):
    """power.py:265-272."""
    hm = halofit_model(nonlinear_fn)
    return halofit(cosmo, k, a, transfer_fn=transfer_fn, prescription=["takahashi2012", "smith2003"][hm])


def sigma8_norm(cosmo, transfer_fn=tklib.Eisenstein_Hu):
    """pknorm = sigma8^2 / sigma^2(R=8) of power.py:41, from the device
    Romberg table (integrate.py:67-159)."""
    ctx = default_context()
    ctx.set_models(tklib.transfer_model(transfer_fn), -1)
    ctx.set_growth(cosmo.gamma)
    return ctx.sigma8_norm(_params(cosmo))
