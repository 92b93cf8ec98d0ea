"""Redshift distributions (jax_cosmo/redshift.py:10-77).

Host-side parameter holders.  Their n(z) are tabulated once, on the grids
the device kernels use, when a probe list is first registered with the
device context; the normalisation is the reference's 256-interval Simpson
over [0, zmax] (redshift.py:27-31).
"""
import numpy as np

steradian_to_arcmin2 = 11818102.86004228

__all__ = ["smail_nz", "gaussian_nz", "fu_nz", "delta_nz", "kde_nz", "systematic_shift",
           "redshift_distribution"]

NZ_SMAIL = 0
NZ_GAUSS = 1
NZ_FU = 2
NZ_DELTA = 3
NZ_KDE = 4
NZ_HOST = 5


def _simps(f, a, b, N):
    x = np.linspace(a, b, N + 1)
    y = f(x)
    return (b - a) / N / 3 * np.sum(y[0:-1:2] + 4 * y[1::2] + y[2::2], axis=0)


class redshift_distribution:
    """Base class of the redshift distributions (redshift.py:15-58).  A
    subclass that defines only ``pz_fn(z)`` (numpy in, numpy out; the
    un-normalised n(z)), as in the reference's subclassing pattern, runs on
    the device path too: the library calls its pz_fn on the host, once per
    grid of redshifts it needs (normalisation nodes, the Simpson a-grid and the
    weak-lensing inner grids of each Simpson size), and keeps the tables
    resident."""

    kind = NZ_HOST

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
        """(kind, params[4], shifted) for the device descriptor; a
        user-defined distribution is tabulated through its pz_fn."""
        if type(self).pz_fn is redshift_distribution.pz_fn:
            raise NotImplementedError("%s defines no pz_fn" % type(self).__name__)
        return NZ_HOST, (0.0, 0.0, 0.0, 0.0), 0

    def host_pz(self):
        """The distribution whose pz_fn the library calls for this bin, or
        None when the device evaluates the n(z) itself."""
        return self if self.device_params()[0] == NZ_HOST else None

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
        return NZ_SMAIL, (a, b, z0, 0.0), 0


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
        return NZ_GAUSS, (mu, sigma, 0.0, 0.0), 0


class fu_nz(redshift_distribution):
    """Fu et al. (2008): n(z) ~ (z^a + z^(a b)) / (z^b + c) (redshift.py:80-105)."""

    kind = NZ_FU

    def pz_fn(self, z):
        a, b, c = self.params
        return (z ** a + z ** (a * b)) / (z ** b + c)

    def device_params(self):
        a, b, c = (float(p) for p in self.params)
        return NZ_FU, (a, b, c, 0.0), 0


class delta_nz(redshift_distribution):
    """A single source plane at z0 (redshift.py:108-123): weak lensing only,
    its kernel is the clip(chi_s - chi, 0) / clip(chi_s, 1) branch of
    probes.py:53-64; its normalisation is 1."""

    kind = NZ_DELTA

    def __init__(self, *args, **kwargs):
        super().__init__(*args, **kwargs)
        self._norm = 1.0

    def pz_fn(self, z):
        z0 = self.params[0]
        return np.where(np.asarray(z) == z0, 1.0, 0.0)

    def device_params(self):
        return NZ_DELTA, (float(self.params[0]), 0.0, 0.0, 0.0), 0


class kde_nz(redshift_distribution):
    """Gaussian KDE of a weighted redshift catalog (redshift.py:126-156); the
    device evaluates it from a resident copy of the catalog."""

    kind = NZ_KDE

    def _kernel(self, bw, X, x):
        return (1.0 / np.sqrt(2 * np.pi) / bw) * np.exp(-((X - x) ** 2) / (bw ** 2 * 2.0))

    def pz_fn(self, z):
        zcat, weight = self.params[:2]
        w = np.atleast_1d(weight)
        X = np.expand_dims(np.asarray(zcat, dtype=float), axis=-1)
        k = self._kernel(self.config["bw"], X, np.asarray(z, dtype=float))
        return np.dot(k.T, w) / np.sum(w)

    def device_params(self):
        return NZ_KDE, (float(self.config["bw"]), 0.0, 0.0, 0.0), 0

    def catalog(self):
        """(redshifts, weights) as float64 arrays of equal length."""
        zcat, weight = self.params[:2]
        z = np.ascontiguousarray(np.atleast_1d(zcat), dtype=np.float64)
        w = np.ascontiguousarray(np.broadcast_to(np.atleast_1d(weight), z.shape), dtype=np.float64)
        return z, w


class systematic_shift(redshift_distribution):
    """A shifted redshift distribution (redshift.py:159-171): the parent's
    un-normalised n(z) at clip(z - bias, 0), normalised over this object's
    own zmax."""

    def pz_fn(self, z):
        parent, bias = self.params[:2]
        return parent.pz_fn(np.clip(np.asarray(z, dtype=float) - bias, 0.0, None))

    def device_params(self):
        parent, bias = self.params[:2]
        if isinstance(parent, (systematic_shift, delta_nz)):
            raise NotImplementedError("systematic_shift of %s is not supported by the device path"
                                      % type(parent).__name__)
        kind, par, _ = parent.device_params()
        return kind, (par[0], par[1], par[2], float(bias)), 1

    def host_pz(self):
        # the library shifts the redshifts, then calls the parent's pz_fn
        return self.params[0].host_pz()

    def catalog(self):
        parent = self.params[0]
        return parent.catalog() if hasattr(parent, "catalog") else None
