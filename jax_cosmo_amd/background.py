"""Background cosmology (jax_cosmo/background.py:25-512).

Closed-form functions are evaluated on the host in float64; the tabulated
quantities (radial comoving distance, growth factor) come from the device
tables, with the reference's table sizes and interpolation rule.
"""
import numpy as np

from . import constants as const
from ._native import default_context

__all__ = ["w", "f_de", "Esqr", "H", "Omega_m_a", "Omega_de_a", "radial_comoving_distance",
           "dchioverda", "growth_factor"]


def _f(x):
    return float(x)


def w(cosmo, a):  # background.py:25-53
    return _f(cosmo.w0) + (1.0 - np.asarray(a, dtype=float)) * _f(cosmo.wa)


def f_de(cosmo, a):  # background.py:56-90
    a = np.asarray(a, dtype=float)
    return (-3.0 * (1.0 + _f(cosmo.w0) + _f(cosmo.wa)) * np.log(a)
            + 3.0 * _f(cosmo.wa) * (a - 1.0))


def Esqr(cosmo, a):  # background.py:93-126
    a = np.asarray(a, dtype=float)
    return (_f(cosmo.Omega_m) * a ** -3.0 + _f(cosmo.Omega_k) * a ** -2.0
            + _f(cosmo.Omega_de) * np.exp(f_de(cosmo, a)))


def H(cosmo, a):  # background.py:129-143
    return const.H0 * np.sqrt(Esqr(cosmo, a))


def Omega_m_a(cosmo, a):  # background.py:146-168
    a = np.asarray(a, dtype=float)
    return _f(cosmo.Omega_m) * a ** -3.0 / Esqr(cosmo, a)


def Omega_de_a(cosmo, a):  # background.py:171-196
    return _f(cosmo.Omega_de) * np.exp(f_de(cosmo, a)) / Esqr(cosmo, a)


def dchioverda(cosmo, a):  # background.py:270-294
    a = np.asarray(a, dtype=float)
    return const.rh / (a ** 2 * np.sqrt(Esqr(cosmo, a)))


def _params(cosmo):
    return np.array([_f(p) for p in cosmo.params()], dtype=np.float64)


def radial_comoving_distance(cosmo, a, log10_amin=-3, steps=256):
    """background.py:199-242 (device table of 256 log-a nodes)."""
    if log10_amin != -3 or steps != 256:
        raise NotImplementedError("only the default chi table is supported")
    a = np.atleast_1d(np.asarray(a, dtype=float))
    chi, _ = default_context().background(_params(cosmo), a)
    return chi


def growth_factor(cosmo, a):
    """background.py:371-398 (ODE growth, device table of 128 log-a nodes)."""
    a = np.atleast_1d(np.asarray(a, dtype=float))
    _, g = default_context().background(_params(cosmo), a)
    return g
