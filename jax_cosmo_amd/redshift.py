"""Redshift distributions (jax_cosmo/redshift.py:10-77).

Host-side parameter holders.  Their n(z) are tabulated once, on the grids
the device kernels use, when a probe list is first registered with the
device context; the normalisation is the reference's 256-interval Simpson
over [0, zmax] (redshift.py:27-31).
"""
import numpy as np

steradian_to_arcmin2 = 11818102.86004228

__all__ = ["smail_nz", "gaussian_nz", "redshift_distribution"]

NZ_SMAIL = 0
NZ_GAUSS = 1


def _simps(f, a, b, N):
    x = np.linspace(a, b, N + 1)
    y = f(x)
    return (b - a) / N / 3 * np.sum(y[0:-1:2] + 4 * y[1::2] + y[2::2], axis=0)


class redshift_distribution:
    kind = None

    def __init__(self, *args, gals_per_arcmin2=1.0, zmax=10.0, **kwargs):
        self._norm = None
        self._gals_per_arcmin2 = gals_per_arcmin2
        self.params = args
        self.config = dict(zmax=zmax, **kwargs)

    def pz_fn(self, z):
        raise NotImplementedError

    def __call__(self, z):
        """Normalised n(z), evaluated on the host in float64."""
        if self._norm is None:
            self._norm = _simps(self.pz_fn, 0.0, self.config["zmax"], 256)
        return self.pz_fn(np.asarray(z, dtype=float)) / self._norm

    @property
    def zmax(self):
        return self.config["zmax"]

    @property
    def gals_per_arcmin2(self):
        return self._gals_per_arcmin2

    @property
    def gals_per_steradian(self):
        return self._gals_per_arcmin2 * steradian_to_arcmin2

    def device_params(self):
        """(kind, params[4]) for the device descriptor."""
        raise NotImplementedError(
            "%s is not supported by the device path" % type(self).__name__)

    def __repr__(self):
        return str(self.params)


class smail_nz(redshift_distribution):
    """n(z) ~ z^a exp(-(z/z0)^b) (redshift.py:61-77)."""

    kind = NZ_SMAIL

    def pz_fn(self, z):
        a, b, z0 = self.params
        return z ** a * np.exp(-((z / z0) ** b))

    def device_params(self):
        a, b, z0 = (float(p) for p in self.params)
        return NZ_SMAIL, (a, b, z0, 0.0)


class gaussian_nz(redshift_distribution):
    """Gaussian redshift bin n(z) ~ exp(-(z - mu)^2 / (2 sigma^2)), normalised
    like every redshift_distribution.  The reference ships no Gaussian bin;
    this follows its subclassing pattern (pz_fn + normalisation)."""

    kind = NZ_GAUSS

    def pz_fn(self, z):
        mu, sigma = self.params
        return np.exp(-0.5 * ((z - mu) / sigma) ** 2)

    def device_params(self):
        mu, sigma = (float(p) for p in self.params)
        return NZ_GAUSS, (mu, sigma, 0.0, 0.0)
