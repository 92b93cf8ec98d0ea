"""MI355X-native drop-in for jax_cosmo's Limber angular power spectrum path.

Mirrors the names of jax_cosmo/__init__.py:1-20 that the path needs.
"""
from . import angular_cl as cl  # noqa: F401
from . import background, bias, constants, likelihood, power, probes, redshift, sparse, transfer  # noqa: F401
from .core import Cosmology  # noqa: F401
from .parameters import Planck15  # noqa: F401

__version__ = "0.1.0"
