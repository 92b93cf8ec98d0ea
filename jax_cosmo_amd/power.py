"""Matter power spectra (jax_cosmo/power.py:14-272), evaluated on the device.

``halofit`` identifies the Takahashi et al. (2012) halofit
(prescription='takahashi2012', the reference default).
"""
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


def linear_matter_power(cosmo, k, a=1.0, transfer_fn=tklib.Eisenstein_Hu, **kwargs):
    """power.py:21-53 at paired (k, a) (broadcast)."""
    if transfer_fn is not tklib.Eisenstein_Hu or kwargs.get("type", "eisenhu_osc") != "eisenhu_osc":
        raise NotImplementedError("only Eisenstein_Hu (eisenhu_osc) is supported")
    k, a = _pairs(k, a)
    plin, _ = default_context().matter_power(_params(cosmo), k.ravel(), a.ravel())
    return plin.reshape(k.shape).squeeze()


def halofit(cosmo, k, a, transfer_fn=tklib.Eisenstein_Hu, prescription="takahashi2012"):
    """power.py:144-262 (Takahashi 2012)."""
    if prescription != "takahashi2012":
        raise NotImplementedError(prescription)
    if transfer_fn is not tklib.Eisenstein_Hu:
        raise NotImplementedError("only Eisenstein_Hu is supported")
    k, a = _pairs(k, a)
    _, pnl = default_context().matter_power(_params(cosmo), k.ravel(), a.ravel())
    return pnl.reshape(k.shape).squeeze()


def nonlinear_matter_power(cosmo, k, a=1.0, transfer_fn=tklib.Eisenstein_Hu, nonlinear_fn=halofit):
    """power.py:265-272."""
    if nonlinear_fn is not halofit:
        raise NotImplementedError("only halofit is supported")
    return halofit(cosmo, k, a, transfer_fn=transfer_fn)


def sigma8_norm(cosmo):
    """pknorm = sigma8^2 / sigma^2(R=8) of power.py:41, from the device
    Romberg table (integrate.py:67-159)."""
    ctx = default_context()
    sid = ctx.tracer_set([dict(probe=0, kind=1, nz_kind=1, nz_params=(0.5, 0.1, 0.0, 0.0), nz_shifted=0,
                               zmax=10.0, gals=1.0, sigma_e=0.0, m_bias=0.0, bias_kind=0,
                               bias_params=(1.0, 0.0, 0.0))])
    return float(ctx.debug_tables(_params(cosmo), sid, 2)[640])
