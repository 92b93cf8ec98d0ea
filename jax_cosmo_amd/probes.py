"""Tracer probes (jax_cosmo/probes.py:132-281).

Host-side descriptions of WeakLensing and NumberCounts probes.  Their
radial kernels (probes.py:16-99) are computed on the device.
"""
import numpy as np

from . import redshift as rds
from .bias import BIAS_NONE, device_bias

__all__ = ["WeakLensing", "NumberCounts"]

TRACER_WL = 0
TRACER_NC = 1


def _catalog(pz):
    return pz.catalog() if hasattr(pz, "catalog") else None


class WeakLensing:
    """probes.py:132-223: a weak-lensing probe over a list of redshift bins,
    with optional NLA intrinsic alignments (ia_bias: one bias model or one
    per bin, probes.py:102-129), a multiplicative bias and sigma_e (one value
    or one per bin)."""

    def __init__(self, redshift_bins, ia_bias=None, multiplicative_bias=0.0, sigma_e=0.26, **kwargs):
        if ia_bias is None:
            self.params = (list(redshift_bins), multiplicative_bias)
        else:
            self.params = (list(redshift_bins), multiplicative_bias, ia_bias)
        kwargs.setdefault("ia_enabled", ia_bias is not None)
        self.config = dict(sigma_e=sigma_e, **kwargs)

    @property
    def n_tracers(self):
        return len(self.params[0])

    @property
    def zmax(self):
        return max(pz.zmax for pz in self.params[0])

    def noise(self):
        """probes.py:210-223: sigma_e^2 / n_gal per bin [n_bins]."""
        pzs = self.params[0]
        ngals = np.array([pz.gals_per_steradian for pz in pzs])
        sigma_e = self.config["sigma_e"]
        if isinstance(sigma_e, list):
            sigma_e = np.array(sigma_e)
        return sigma_e ** 2 / ngals

    def tracer_entries(self):
        pzs, m = self.params[:2]
        ms = m if isinstance(m, (list, tuple)) else [m] * len(pzs)
        sig = self.config["sigma_e"]
        sigs = sig if isinstance(sig, (list, tuple)) else [sig] * len(pzs)
        if self.config["ia_enabled"]:
            ia = self.params[2]
            ias = ia if isinstance(ia, (list, tuple)) else [ia] * len(pzs)
        else:
            ias = [None] * len(pzs)
        out = []
        for pz, mi, si, bi in zip(pzs, ms, sigs, ias):
            kind, par, shifted = pz.device_params()
            if bi is None:
                bk, bp = BIAS_NONE, (0.0, 0.0, 0.0)
            else:
                if kind == rds.NZ_DELTA:
                    raise NotImplementedError("NLA kernel not properly implemented for delta redshift distributions")
                bk, bp = device_bias(bi)
            out.append(dict(kind=TRACER_WL, nz_kind=kind, nz_params=par, nz_shifted=shifted,
                            zmax=float(pz.zmax), gals=float(pz.gals_per_arcmin2), sigma_e=float(si),
                            m_bias=float(mi), bias_kind=bk, bias_params=bp, catalog=_catalog(pz),
                            host_pz=pz.host_pz()))
        return out


class NumberCounts:
    """probes.py:226-281: galaxy clustering over a list of redshift bins with
    a bias model (bias.py:10-59), one for all bins or one per bin."""

    def __init__(self, redshift_bins, bias, has_rsd=False, **kwargs):
        if has_rsd:
            raise NotImplementedError("RSD is not supported")
        self.params = (list(redshift_bins), bias)
        self.config = dict(has_rsd=has_rsd, **kwargs)

    @property
    def zmax(self):
        return max(pz.zmax for pz in self.params[0])

    @property
    def n_tracers(self):
        return len(self.params[0])

    def noise(self):
        """probes.py:274-281: 1 / n_gal per bin [n_bins]."""
        pzs = self.params[0]
        ngals = np.array([pz.gals_per_steradian for pz in pzs])
        return 1.0 / ngals

    def tracer_entries(self):
        pzs, bias = self.params
        biases = bias if isinstance(bias, (list, tuple)) else [bias] * len(pzs)
        out = []
        for pz, b in zip(pzs, biases):
            kind, par, shifted = pz.device_params()
            if kind == rds.NZ_DELTA:
                raise NotImplementedError("Density kernel not properly implemented for delta redshift distributions")
            bk, bp = device_bias(b)
            out.append(dict(kind=TRACER_NC, nz_kind=kind, nz_params=par, nz_shifted=shifted,
                            zmax=float(pz.zmax), gals=float(pz.gals_per_arcmin2), sigma_e=0.0, m_bias=0.0,
                            bias_kind=bk, bias_params=bp, catalog=_catalog(pz), host_pz=pz.host_pz()))
        return out
