"""Background cosmology (jax_cosmo/background.py:25-512).

Closed-form functions are evaluated on the host in float64; the tabulated
quantities (radial and transverse comoving distance, growth factor and rate)
come from the device tables (FP64), with the reference's table sizes,
integration schemes and interpolation rule.  The growth follows the
cosmology's model: the growth ODE, or dlnD/dlna = Omega_m(a)^gamma for a
Cosmology with gamma (background.py:371-440).
"""
import numpy as np

from . import constants as const
from ._native import default_context

__all__ = ["w", "f_de", "Esqr", "H", "Omega_m_a", "Omega_de_a", "radial_comoving_distance",
           "dchioverda", "transverse_comoving_distance", "angular_diameter_distance", "growth_factor",
           "growth_rate"]


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


def _background(cosmo, a):
    """FP64 (chi, D, f, f_k) of the device tables at the scale factors a."""
    ctx = default_context()
    ctx.set_growth(cosmo.gamma)
    return ctx.background_ex(_params(cosmo), np.atleast_1d(np.asarray(a, dtype=float)))


def radial_comoving_distance(cosmo, a, log10_amin=-3, steps=256):
    """background.py:199-242 (device table of 256 log-a nodes)."""
    if log10_amin != -3 or steps != 256:
        raise NotImplementedError("only the default chi table is supported")
    return _background(cosmo, a)[0]


def transverse_comoving_distance(cosmo, a):
    """background.py:297-344: R_H / sqrt|Omega_k| sinh (Omega_k > 0) or sin
    (Omega_k < 0) of sqrt|Omega_k| chi / R_H, chi for a flat cosmology."""
    return _background(cosmo, a)[3]


def angular_diameter_distance(cosmo, a):
    """background.py:347-368: a f_k(a)."""
    a = np.atleast_1d(np.asarray(a, dtype=float))
    return a * transverse_comoving_distance(cosmo, a)


def growth_factor(cosmo, a):
    """background.py:371-398: the ODE table (background.py:443-487) or the
    gamma table (background.py:515-549) of 128 log-a nodes, clip(interp, 0, 1)."""
    return _background(cosmo, a)[1]


def growth_rate(cosmo, a):
    """background.py:401-440: dlnD/dlna from the ODE table
    (background.py:490-512) or Omega_m(a)^gamma (background.py:551-582)."""
    return _background(cosmo, a)[2]
