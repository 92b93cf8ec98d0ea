"""Cosmology parameter container (jax_cosmo/core.py:12-198).

A plain host object: the eight primary parameters may be Python floats or
torch scalars (torch scalars with ``requires_grad`` make ``angular_cl``
differentiable through the device adjoint).
"""

__all__ = ["Cosmology"]

PARAM_NAMES = ("Omega_c", "Omega_b", "h", "n_s", "sigma8", "Omega_k", "w0", "wa")
GRAD_NAMES = ("Omega_c", "Omega_b", "h", "n_s", "sigma8", "w0", "wa")


class Cosmology:
    def __init__(self, Omega_c, Omega_b, h, n_s, sigma8, Omega_k, w0, wa, gamma=None):
        # gamma (core.py:13, 59-60): the growth follows dlnD/dlna =
        # Omega_m(a)^gamma (background.py:515-582) instead of the growth ODE.
        # It is a fixed number of the cosmology here: derivatives are taken
        # with respect to the seven other non-curvature parameters only.
        self._Omega_c = Omega_c
        self._Omega_b = Omega_b
        self._h = h
        self._n_s = n_s
        self._sigma8 = sigma8
        self._Omega_k = Omega_k
        self._w0 = w0
        self._wa = wa
        self._gamma = None if gamma is None else float(gamma)
        self._flags = {"gamma_growth": gamma is not None}
        self._workspace = {}

    def __str__(self):
        return (
            "Cosmological parameters: \n"
            + "    h:        " + str(self.h) + " \n"
            + "    Omega_b:  " + str(self.Omega_b) + " \n"
            + "    Omega_c:  " + str(self.Omega_c) + " \n"
            + "    Omega_k:  " + str(self.Omega_k) + " \n"
            + "    w0:       " + str(self.w0) + " \n"
            + "    wa:       " + str(self.wa) + " \n"
            + "    n:        " + str(self.n_s) + " \n"
            + "    sigma8:   " + str(self.sigma8)
        )

    def __repr__(self):
        return self.__str__()

    def params(self):
        """The 8 primary parameters in Cosmology(...) order."""
        return [getattr(self, "_" + n) for n in PARAM_NAMES]

    @property
    def Omega(self):
        return 1.0 - self._Omega_k

    @property
    def Omega_b(self):
        return self._Omega_b

    @property
    def Omega_c(self):
        return self._Omega_c

    @property
    def Omega_m(self):
        return self._Omega_b + self._Omega_c

    @property
    def Omega_de(self):
        return self.Omega - self.Omega_m

    @property
    def Omega_k(self):
        return self._Omega_k

    @property
    def k(self):
        ok = float(self._Omega_k)
        return -1 if ok > 0 else (1 if ok < 0 else 0)

    @property
    def sqrtk(self):
        return abs(float(self._Omega_k)) ** 0.5

    @property
    def h(self):
        return self._h

    @property
    def w0(self):
        return self._w0

    @property
    def wa(self):
        return self._wa

    @property
    def n_s(self):
        return self._n_s

    @property
    def sigma8(self):
        return self._sigma8

    @property
    def gamma(self):
        return self._gamma
