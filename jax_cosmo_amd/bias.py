"""Galaxy and intrinsic-alignment bias models (jax_cosmo/bias.py:10-59).

Each model carries its device form: ``device_bias()`` returns (kind, params)
with kind one of the BIAS_* codes of the C ABI."""
import numpy as np

__all__ = ["constant_linear_bias", "inverse_growth_linear_bias", "des_y1_ia_bias"]

BIAS_CONST = 0
BIAS_INV_GROWTH = 1
BIAS_DES_Y1 = 2
BIAS_NONE = 3


class _bias:
    def __init__(self, *params):
        self.params = params
        self.config = {}

    def __repr__(self):
        return str(self.params)


class constant_linear_bias(_bias):
    """Redshift-independent linear bias b (bias.py:10-22)."""

    @property
    def b(self):
        return self.params[0]

    def __call__(self, cosmo, z):
        return self.params[0] * np.ones_like(np.asarray(z, dtype=float))

    def device_bias(self):
        return BIAS_CONST, (float(self.params[0]), 0.0, 0.0)


class inverse_growth_linear_bias(_bias):
    """b / D(a) (bias.py:25-40); D is evaluated on the device."""

    def __call__(self, cosmo, z):
        from . import background

        a = 1.0 / (1.0 + np.asarray(z, dtype=float))
        return self.params[0] / background.growth_factor(cosmo, a)

    def device_bias(self):
        return BIAS_INV_GROWTH, (float(self.params[0]), 0.0, 0.0)


class des_y1_ia_bias(_bias):
    """A ((1 + z) / (1 + z0))^eta (bias.py:43-59)."""

    def __call__(self, cosmo, z):
        A, eta, z0 = self.params
        return A * ((1.0 + np.asarray(z, dtype=float)) / (1.0 + z0)) ** eta

    def device_bias(self):
        A, eta, z0 = (float(p) for p in self.params)
        return BIAS_DES_Y1, (A, eta, z0)


def device_bias(b):
    """(kind, params) of a bias model, NotImplementedError for other callables."""
    if not hasattr(b, "device_bias"):
        raise NotImplementedError("bias model %s is not supported by the device path" % type(b).__name__)
    return b.device_bias()
