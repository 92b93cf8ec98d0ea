"""Galaxy bias models (jax_cosmo/bias.py:10-22)."""

__all__ = ["constant_linear_bias"]


class constant_linear_bias:
    """Redshift-independent linear bias b (bias.py:10-22)."""

    def __init__(self, b):
        self.params = (b,)
        self.config = {}

    @property
    def b(self):
        return self.params[0]

    def __call__(self, cosmo, z):
        import numpy as np

        return self.params[0] * np.ones_like(np.asarray(z, dtype=float))

    def __repr__(self):
        return str(self.params)
