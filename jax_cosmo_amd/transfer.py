"""Transfer function (jax_cosmo/transfer.py:10-156).

``Eisenstein_Hu`` identifies the EH98 transfer function with baryon wiggles
(type='eisenhu_osc', the reference default), which is what the device
kernels evaluate.  Called directly it returns T(k) from the device.
"""
import numpy as np

__all__ = ["Eisenstein_Hu"]


def Eisenstein_Hu(cosmo, k, type="eisenhu_osc"):
    """T(k), k in h/Mpc (transfer.py:10-156), recovered from the device
    linear spectrum as sqrt(P_lin(k, a=1) / (k^n_s pknorm)) (D(1) = 1)."""
    if type != "eisenhu_osc":
        raise NotImplementedError(type)
    from . import power

    k = np.atleast_1d(np.asarray(k, dtype=float))
    plin = power.linear_matter_power(cosmo, k, a=1.0)
    pknorm = power.sigma8_norm(cosmo)
    return np.sqrt(plin / (k ** float(cosmo.n_s) * pknorm))
