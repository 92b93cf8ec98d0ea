"""FP64 CPU restatement of jax_cosmo's Limber angular power spectrum path.

THIS IS TEST INFRASTRUCTURE, NOT PRODUCT CODE.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it, and only as the checker.  The shipped package (``jax_cosmo_amd``)
never imports, calls or falls back to anything in this directory.

What it is
----------
A torch.float64 (CPU) re-statement of the reference algorithm, written from
the reference's source and following its discretisation step by step,
including its quirks (nearest-node ``interp`` with the first-segment and
decreasing-table behaviour, the log10/ln mix-up of the sigma_8 Romberg bounds,
the per-z 256-interval weak-lensing Simpson, the RK4-with-h=0-first-step
tables).  It is "reference-faithful": the halofit sigma^2(R, a) tensor is
built naively as [Na, 257, 256] and every chi lookup is an argmin over the 256
table nodes, exactly as the reference does.  torch autograd through it stands
in for ``jax.grad`` of the reference (the reference has no custom VJPs, so its
gradient is the derivative of this same discrete algorithm).

Reference files (paths relative to the reference repository root):
  jax_cosmo/constants.py:9-27, utils.py:2-9, core.py:143-198,
  parameters.py:10-20, background.py:25-512, transfer.py:10-156,
  power.py:14-272, probes.py:16-99,132-281, redshift.py:10-77,
  bias.py:10-22, angular_cl.py:15-196, likelihood.py:9-61,
  sparse.py:37-389, scipy/integrate.py:21-199, scipy/ode.py:6-22,
  scipy/interpolate.py:12-37.

Pinning status (each function's independent check lives in
tests/test_oracle_pins.py):
  simps, romb, odeint, interp, chi table, growth table (ODE and gamma),
  growth rate, transverse / angular diameter distance, sigma8 integrand,
  Smail / Gaussian normalisation, user-defined n(z) (user_nz with the Smail
  formula reproduces smail_nz), WL kernel quadrature, Limber outer
  quadrature, covariance orderings, sparse algebra: pinned.
  eisenstein_hu: the zero-baryon limit (T -> T_0(k; 1, 1) of EH98 eqs. 3,
  10, 11 written from the paper) pins k_eq, q, T_0 and the combination of
  the CDM terms; the no-wiggle branch is pinned against EH98 eqs. 28-31
  written from the paper (with the eq. 6 sound horizon, as the reference
  uses); plus T->1 as k->0 and the <= 12 % envelope.  Parity unpinned: the
  coefficients that only act with baryons in the wiggle form (alpha_c,
  beta_c of eqs. 11-12, the f and T_b shapes of eqs. 17-21, alpha_b, beta_b,
  beta_node, k_silk).
  halofit Takahashi-2012 coefficients: pinned by invariants only
  (P_nl/P_lin -> 1 as k -> 0, sigma(R_nl) ~ 1, n_eff and C against finite
  differences of ln sigma^2); the coefficient values themselves are
  parity unpinned.
  End-to-end C_ell against the real reference at 1e-4: parity unpinned (the
  reference needs JAX, which is absent).
"""

import math

import torch

DT = torch.float64

# ---------------------------------------------------------------- constants
# jax_cosmo/constants.py:9-27
C_KMS = 299792.458
TCMB = 2.726
RH = 2997.92458
H0 = 100.0
RHOCRIT = 2.7750e11
# jax_cosmo/redshift.py:10
STERADIAN_TO_ARCMIN2 = 11818102.86004228

PARAM_NAMES = ("Omega_c", "Omega_b", "h", "n_s", "sigma8", "Omega_k", "w0", "wa")
GRAD_NAMES = ("Omega_c", "Omega_b", "h", "n_s", "sigma8", "w0", "wa")


def _t(x):
    if isinstance(x, torch.Tensor):
        return x.to(DT)
    return torch.tensor(x, dtype=DT)


# ---------------------------------------------------------------- cosmology
class Cosmo:
    """Parameter holder, jax_cosmo/core.py:12-198 (derived properties
    core.py:143-198).  Parameters are float64 tensors so autograd can
    differentiate through them."""

    def __init__(self, Omega_c, Omega_b, h, n_s, sigma8, Omega_k, w0, wa, gamma=None):
        self.gamma = None if gamma is None else _t(gamma)  # core.py:13, 59-60
        self.Omega_c = _t(Omega_c)
        self.Omega_b = _t(Omega_b)
        self.h = _t(h)
        self.n_s = _t(n_s)
        self.sigma8 = _t(sigma8)
        self.Omega_k = _t(Omega_k)
        self.w0 = _t(w0)
        self.wa = _t(wa)
        self.workspace = {}

    @classmethod
    def from_vector(cls, v):
        return cls(*[v[i] for i in range(8)])

    @property
    def Omega_m(self):  # core.py:157-159
        return self.Omega_b + self.Omega_c

    @property
    def Omega(self):  # core.py:145-147
        return 1.0 - self.Omega_k

    @property
    def Omega_de(self):  # core.py:161-163
        return self.Omega - self.Omega_m


def planck15(**kw):
    """jax_cosmo/parameters.py:10-20."""
    p = dict(Omega_c=0.2589, Omega_b=0.04860, Omega_k=0.0, h=0.6774,
             n_s=0.9667, sigma8=0.8159, w0=-1.0, wa=0.0)
    p.update(kw)
    return Cosmo(**p)


def z2a(z):  # utils.py:2-4
    return 1.0 / (1.0 + z)


def a2z(a):  # utils.py:7-9
    return 1.0 / a - 1.0


# ---------------------------------------------------------------- numerics
def linspace(start, stop, num):
    """numpy/JAX linspace semantics: start + i*step, last point = stop.
    Accepts a tensor ``start`` (shape S) -> result shape [num, *S]."""
    start = _t(start)
    stop = _t(stop)
    i = torch.arange(num, dtype=DT).reshape((num,) + (1,) * start.dim())
    step = (stop - start) / (num - 1)
    out = start + i * step
    out = torch.cat([out[:-1], (stop + 0.0 * start).unsqueeze(0)], dim=0)
    return out


def simps_samples(y, dx, dim=0):
    """Composite Simpson of already-sampled values, jax_cosmo/scipy/
    integrate.py:162-199 (``S = dx/3 * sum(y[0:-1:2] + 4 y[1::2] + y[2::2])``)."""
    y = y.movedim(dim, 0)
    if (y.shape[0] - 1) % 2 == 1:
        raise ValueError("N must be an even integer.")
    s = torch.sum(y[0:-1:2] + 4 * y[1::2] + y[2::2], dim=0)
    return dx / 3 * s


def simps(f, a, b, N=128):
    """jax_cosmo/scipy/integrate.py:162-199.  ``a`` may be a tensor (the
    weak-lensing kernel integrates from an array of lower bounds)."""
    if N % 2 == 1:
        raise ValueError("N must be an even integer.")
    a = _t(a)
    b = _t(b)
    dx = (b - a) / N
    x = linspace(a, b, N + 1)
    y = f(x)
    return dx / 3 * torch.sum(y[0:-1:2] + 4 * y[1::2] + y[2::2], dim=0)


def romb(function, a, b, divmax=6):
    """Function Romberg, jax_cosmo/scipy/integrate.py:67-159.  Returns
    state[divmax] of the Richardson table (the reference returns state[i]
    with i == divmax after its loop)."""
    a = _t(a)
    b = _t(b)
    intrange = b - a
    ordsum = 0.5 * (function(a.reshape(1)) + function(b.reshape(1))).sum()  # _difftrap1
    result = intrange * ordsum
    state = [result] * (divmax + 1)
    for i in range(1, divmax + 1):
        n = 2 ** i
        numtosum = n // 2  # _difftrapn
        h = intrange / numtosum
        lox = a + 0.5 * h
        points = lox + h * torch.arange(numtosum, dtype=DT)
        ordsum = ordsum + torch.sum(function(points))
        x = intrange * ordsum / n
        new_state = [x]
        for k in range(divmax):  # lax.scan over state[:-1]
            tmp = 4.0 ** (k + 1)
            x = (tmp * x - state[k]) / (tmp - 1.0)  # _romberg_diff
            new_state.append(x)
        state = new_state
    return state[divmax]


def odeint(fn, y0, t):
    """Fixed-grid RK4 with a zero-length first step, jax_cosmo/scipy/
    ode.py:6-22.  Returns y at every t (y[0] == y0)."""
    y = y0
    t_prev = t[0]
    out = []
    for i in range(t.shape[0]):
        ti = t[i]
        h = ti - t_prev
        k1 = fn(y, t_prev)
        k2 = fn(y + h * k1 / 2, t_prev + h / 2)
        k3 = fn(y + h * k2 / 2, t_prev + h / 2)
        k4 = fn(y + h * k3, ti)
        y = y + 1.0 / 6.0 * h * (k1 + 2 * k2 + 2 * k3 + k4)
        t_prev = ti
        out.append(y)
    return torch.stack(out, dim=0)


def interp(x, xp, fp, chunk=1 << 15):
    """Nearest-node linear interpolation, jax_cosmo/scipy/interpolate.py:
    12-37, vectorised over x (any shape).

    ind = argmin((x - xp)^2) (first index on ties), clipped to [1, n-2];
    side s = sign(clip(x, xp[1], xp[-2]) - xp[ind]) with 0 -> +1; line through
    (xp[ind], fp[ind]) and (xp[ind+s], fp[ind+s]) evaluated as a*x + b.
    clip is max-then-min (JAX), so for a decreasing xp it returns xp[-2]."""
    shape = x.shape
    xf = x.reshape(-1)
    n = xp.shape[0]
    inds = []
    for s in range(0, xf.shape[0], chunk):
        xc = xf[s:s + chunk].detach()
        inds.append(torch.argmin((xc[:, None] - xp.detach()[None, :]) ** 2, dim=1))
    ind = torch.cat(inds) if inds else torch.zeros(0, dtype=torch.long)
    ind = ind.clamp(1, n - 2)
    xi = xp[ind]
    xcl = torch.minimum(torch.maximum(xf, xp[1]), xp[-2])
    s = torch.sign(xcl - xi)
    step = torch.where(s < 0, torch.tensor(-1), torch.tensor(1))
    a = (fp[ind + step] - fp[ind]) / (xp[ind + step] - xp[ind])
    b = fp[ind] - a * xp[ind]
    return (a * xf + b).reshape(shape)


def clip_min(x, lo):
    """jnp.clip(x, lo) == maximum(x, lo)."""
    return torch.maximum(x, torch.as_tensor(lo, dtype=DT))


# ---------------------------------------------------------------- background
def w(cosmo, a):  # background.py:25-53
    return cosmo.w0 + (1.0 - a) * cosmo.wa


def f_de(cosmo, a):  # background.py:56-90
    return -3.0 * (1.0 + cosmo.w0 + cosmo.wa) * torch.log(a) + 3.0 * cosmo.wa * (a - 1.0)


def Esqr(cosmo, a):  # background.py:93-126
    return (cosmo.Omega_m * torch.pow(a, -3) + cosmo.Omega_k * torch.pow(a, -2)
            + cosmo.Omega_de * torch.exp(f_de(cosmo, a)))


def H(cosmo, a):  # background.py:129-143
    return H0 * torch.sqrt(Esqr(cosmo, a))


def Omega_m_a(cosmo, a):  # background.py:146-168
    return cosmo.Omega_m * torch.pow(a, -3) / Esqr(cosmo, a)


def Omega_de_a(cosmo, a):  # background.py:171-196
    return cosmo.Omega_de * torch.exp(f_de(cosmo, a)) / Esqr(cosmo, a)


def dchioverda(cosmo, a):  # background.py:270-294
    return RH / (a ** 2 * torch.sqrt(Esqr(cosmo, a)))


def logspace(start, stop, num):
    return torch.pow(torch.tensor(10.0, dtype=DT), linspace(start, stop, num))


def chi_table(cosmo, log10_amin=-3, steps=256):
    """background.py:223-233: RK4 in ln a of a y-independent RHS, i.e. a
    cumulative Simpson with midpoint evaluations; chi = C_last - C."""
    key = "chi"
    if key not in cosmo.workspace:
        atab = logspace(log10_amin, 0.0, steps)

        def dchioverdlna(y, x):
            xa = torch.exp(x)
            return dchioverda(cosmo, xa) * xa

        chitab = odeint(dchioverdlna, torch.zeros((), dtype=DT), torch.log(atab))
        chitab = chitab[-1] - chitab
        cosmo.workspace[key] = (atab, chitab)
    return cosmo.workspace[key]


def radial_comoving_distance(cosmo, a):
    """background.py:199-242: clip(interp(a, atab, chitab), 0)."""
    atab, chitab = chi_table(cosmo)
    a = torch.atleast_1d(_t(a)) if not isinstance(a, torch.Tensor) else torch.atleast_1d(a)
    return clip_min(interp(a, atab, chitab), 0.0)


def growth_table(cosmo, log10_amin=-3, steps=128):
    """background.py:443-487: RK4 over a (not ln a) of y=[D, dD/da] from
    y0=[a0, 1]; g = D/D[-1], f = (dD/da / D[-1]) a / g.  For a cosmology with
    gamma, background.py:515-549: RK4 in ln a of dlnD/dlna = Omega_m(a)^gamma
    from ln a0; g = exp(y) / exp(y[-1]) (no f table: the rate is direct)."""
    key = "growth"
    if key not in cosmo.workspace and cosmo.gamma is not None:
        atab = logspace(log10_amin, 0.0, steps)

        def integrand(y, loga):
            return growth_rate_gamma(cosmo, torch.exp(loga))

        gtab = torch.exp(odeint(integrand, torch.log(atab[0]), torch.log(atab)))
        gtab = gtab / gtab[-1]
        cosmo.workspace[key] = (atab, gtab, None)
    if key not in cosmo.workspace:
        atab = logspace(log10_amin, 0.0, steps)

        def D_derivs(y, x):
            q = (2.0 - 0.5 * (Omega_m_a(cosmo, x)
                              + (1.0 + 3.0 * w(cosmo, x)) * Omega_de_a(cosmo, x))) / x
            r = 1.5 * Omega_m_a(cosmo, x) / x / x
            return torch.stack([y[1], -q * y[1] + r * y[0]])

        y0 = torch.stack([atab[0], torch.ones((), dtype=DT)])
        y = odeint(D_derivs, y0, atab)
        y1 = y[:, 0]
        gtab = y1 / y1[-1]
        ftab = y[:, 1] / y1[-1] * atab / gtab
        cosmo.workspace[key] = (atab, gtab, ftab)
    return cosmo.workspace[key]


def growth_rate_gamma(cosmo, a):  # background.py:551-582
    return Omega_m_a(cosmo, a) ** cosmo.gamma


def growth_rate(cosmo, a):
    """background.py:401-440, 490-512: interp(a, atab, ftab) of the ODE, or
    Omega_m(a)^gamma."""
    a = torch.atleast_1d(_t(a)) if not isinstance(a, torch.Tensor) else torch.atleast_1d(a)
    if cosmo.gamma is not None:
        return growth_rate_gamma(cosmo, a)
    atab, _, ftab = growth_table(cosmo)
    return interp(a, atab, ftab)


def transverse_comoving_distance(cosmo, a):
    """background.py:297-344: the branch of k = -sign(Omega_k) (core.py:
    165-171): open (Omega_k > 0) R_H/sqrt(k) sinh(sqrt(k) chi/R_H), flat chi,
    closed R_H/sqrt(k) sin(sqrt(k) chi/R_H), sqrt(k) = sqrt|Omega_k|."""
    chi = radial_comoving_distance(cosmo, a)
    ok = float(cosmo.Omega_k)
    sk = torch.sqrt(torch.abs(cosmo.Omega_k))
    if ok > 0:
        return RH / sk * torch.sinh(sk * chi / RH)
    if ok < 0:
        return RH / sk * torch.sin(sk * chi / RH)
    return chi


def angular_diameter_distance(cosmo, a):  # background.py:347-368
    a = torch.atleast_1d(_t(a)) if not isinstance(a, torch.Tensor) else torch.atleast_1d(a)
    return a * transverse_comoving_distance(cosmo, a)


def growth_factor(cosmo, a):
    """background.py:371-398, 488: clip(interp(a, atab, gtab), 0, 1)."""
    atab, gtab, _ = growth_table(cosmo)
    a = torch.atleast_1d(a)
    v = interp(a, atab, gtab)
    return torch.minimum(torch.maximum(v, torch.zeros((), dtype=DT)), torch.ones((), dtype=DT))


# ---------------------------------------------------------------- transfer
def eisenstein_hu(cosmo, k):
    """Eisenstein & Hu 1998, transfer.py:10-156: with baryon wiggles
    (type='eisenhu_osc', the default) or the no-wiggle fit (type='eisenhu')
    when ``cosmo.transfer_type`` says so (the oracle carries the reference's
    transfer_fn choice on the cosmology object)."""
    T_2_7_sqr = (TCMB / 2.7) ** 2
    h = cosmo.h
    h2 = h ** 2
    w_m = cosmo.Omega_m * h2
    w_b = cosmo.Omega_b * h2
    fb = cosmo.Omega_b / cosmo.Omega_m
    fc = (cosmo.Omega_m - cosmo.Omega_b) / cosmo.Omega_m

    k_eq = 7.46e-2 * w_m / T_2_7_sqr / h  # transfer.py:53
    z_eq = 2.50e4 * w_m / (T_2_7_sqr) ** 2  # :54

    b1 = 0.313 * torch.pow(w_m, -0.419) * (1.0 + 0.607 * torch.pow(w_m, 0.674))  # :57
    b2 = 0.238 * torch.pow(w_m, 0.223)
    z_d = (1291.0 * torch.pow(w_m, 0.251) / (1.0 + 0.659 * torch.pow(w_m, 0.828))
           * (1.0 + b1 * torch.pow(w_b, b2)))

    R_d = 31.5 * w_b / (T_2_7_sqr) ** 2 * (1.0e3 / z_d)  # :67
    R_eq = 31.5 * w_b / (T_2_7_sqr) ** 2 * (1.0e3 / z_eq)
    sh_d = (2.0 / (3.0 * k_eq) * torch.sqrt(6.0 / R_eq)
            * torch.log((torch.sqrt(1.0 + R_d) + torch.sqrt(R_eq + R_d)) / (1.0 + torch.sqrt(R_eq))))
    k_silk = (1.6 * torch.pow(w_b, 0.52) * torch.pow(w_m, 0.73)
              * (1.0 + torch.pow(10.4 * w_m, -0.95)) / h)  # :79-85

    if getattr(cosmo, "transfer_type", "eisenhu_osc") == "eisenhu":
        # transfer.py:88-105: gamma_eff and the EH98 (29) zero-baryon-wiggle fit
        alpha_gamma = (1.0 - 0.328 * torch.log(431.0 * w_m) * w_b / w_m
                       + 0.38 * torch.log(22.3 * w_m) * (cosmo.Omega_b / cosmo.Omega_m) ** 2)
        gamma_eff = cosmo.Omega_m * h * (alpha_gamma + (1.0 - alpha_gamma) / (1.0 + (0.43 * k * sh_d) ** 4))
        q = k * T_2_7_sqr / gamma_eff
        L = torch.log(2.0 * math.e + 1.8 * q)
        C = 14.2 + 731.0 / (1.0 + 62.5 * q)
        return L / (L + C * q * q)

    # transfer.py:107-117 (EH98 eqs. 11, 12)
    a1 = torch.pow(46.9 * w_m, 0.670) * (1.0 + torch.pow(32.1 * w_m, -0.532))
    a2 = torch.pow(12.0 * w_m, 0.424) * (1.0 + torch.pow(45.0 * w_m, -0.582))
    alpha_c = torch.pow(a1, -fb) * torch.pow(a2, -(fb ** 3))
    bb1 = 0.944 / (1.0 + torch.pow(458.0 * w_m, -0.708))
    bb2 = torch.pow(0.395 * w_m, -0.0266)
    beta_c = 1.0 + bb1 * (torch.pow(fc, bb2) - 1.0)
    beta_c = 1.0 / beta_c

    def T_tilde(k1, alpha, beta):  # transfer.py:120-126
        q = k1 / (13.41 * k_eq)
        L = torch.log(math.e + 1.8 * beta * q)
        C = 14.2 / alpha + 386.0 / (1.0 + 69.9 * torch.pow(q, 1.08))
        return L / (L + C * q * q)

    f = 1.0 / (1.0 + (k * sh_d / 5.4) ** 4)  # :129
    Tc = f * T_tilde(k, 1.0, beta_c) + (1.0 - f) * T_tilde(k, alpha_c, beta_c)

    y = (1.0 + z_eq) / (1.0 + z_d)  # :134-138
    x = torch.sqrt(1.0 + y)
    G_EH98 = y * (-6.0 * x + (2.0 + 3.0 * y) * torch.log((x + 1.0) / (x - 1.0)))
    alpha_b = 2.07 * k_eq * sh_d * torch.pow(1.0 + R_d, -0.75) * G_EH98

    beta_node = 8.41 * torch.pow(w_m, 0.435)  # :140-141
    tilde_s = sh_d / torch.pow(1.0 + (beta_node / (k * sh_d)) ** 3, 1.0 / 3.0)
    beta_b = 0.5 + fb + (3.0 - 2.0 * fb) * torch.sqrt((17.2 * w_m) ** 2 + 1.0)

    Tb = (T_tilde(k, 1.0, 1.0) / (1.0 + (k * sh_d / 5.2) ** 2)
          + alpha_b / (1.0 + (beta_b / (k * sh_d)) ** 3)
          * torch.exp(-torch.pow(k / k_silk, 1.4))) * torch.sinc(k * tilde_s / math.pi)
    return fb * Tb + fc * Tc


def eh_scalars(cosmo):
    """The k-independent EH98 quantities (for pin tests)."""
    T_2_7_sqr = (TCMB / 2.7) ** 2
    h = cosmo.h
    w_m = cosmo.Omega_m * h ** 2
    w_b = cosmo.Omega_b * h ** 2
    k_eq = 7.46e-2 * w_m / T_2_7_sqr / h
    z_eq = 2.50e4 * w_m / (T_2_7_sqr) ** 2
    b1 = 0.313 * torch.pow(w_m, -0.419) * (1.0 + 0.607 * torch.pow(w_m, 0.674))
    b2 = 0.238 * torch.pow(w_m, 0.223)
    z_d = (1291.0 * torch.pow(w_m, 0.251) / (1.0 + 0.659 * torch.pow(w_m, 0.828))
           * (1.0 + b1 * torch.pow(w_b, b2)))
    R_d = 31.5 * w_b / (T_2_7_sqr) ** 2 * (1.0e3 / z_d)
    R_eq = 31.5 * w_b / (T_2_7_sqr) ** 2 * (1.0e3 / z_eq)
    sh_d = (2.0 / (3.0 * k_eq) * torch.sqrt(6.0 / R_eq)
            * torch.log((torch.sqrt(1.0 + R_d) + torch.sqrt(R_eq + R_d)) / (1.0 + torch.sqrt(R_eq))))
    return dict(k_eq=k_eq, z_eq=z_eq, z_d=z_d, R_d=R_d, R_eq=R_eq, sh_d=sh_d)


# ---------------------------------------------------------------- power
def primordial_matter_power(cosmo, k):  # power.py:14-18
    return k ** cosmo.n_s


def sigmasqr(cosmo, R, kmin=0.0001, kmax=1000.0):
    """power.py:56-78.  Note the bounds are log10(kmin), log10(kmax) while the
    integrand exponentiates with e: the reference integrates ln k in [-4, 3]."""

    def int_sigma(logk):
        k = torch.exp(logk)
        x = k * R
        wv = 3.0 * (torch.sin(x) - x * torch.cos(x)) / (x * x * x)
        pk = eisenstein_hu(cosmo, k) ** 2 * primordial_matter_power(cosmo, k)
        return k * (k * wv) ** 2 * pk

    y = romb(int_sigma, math.log10(kmin), math.log10(kmax), divmax=7)
    return 1.0 / (2.0 * math.pi ** 2.0) * y


def linear_matter_power(cosmo, k, a=1.0):
    """power.py:21-53 (k and a broadcast against each other)."""
    k = torch.atleast_1d(_t(k) if not isinstance(k, torch.Tensor) else k)
    a = torch.atleast_1d(_t(a) if not isinstance(a, torch.Tensor) else a)
    g = growth_factor(cosmo, a.reshape(-1)).reshape(a.shape)
    t = eisenstein_hu(cosmo, k)
    if "pknorm" not in cosmo.workspace:
        cosmo.workspace["pknorm"] = cosmo.sigma8 ** 2 / sigmasqr(cosmo, 8.0)
    pknorm = cosmo.workspace["pknorm"]
    pk = primordial_matter_power(cosmo, k) * t ** 2 * g ** 2
    return pk * pknorm


def halofit_parameters(cosmo, a, a_chunk=32):
    """power.py:86-141 (reference-faithful: the [Na, 257, 256] sigma^2 tensor
    is built naively, in chunks of a to bound memory)."""
    a = torch.atleast_1d(a)
    logr = linspace(math.log(1e-4), math.log(1e1), 256)
    lnk_lo, lnk_hi = math.log(1e-4), math.log(1e4)
    logk = linspace(lnk_lo, lnk_hi, 257)
    dx = (lnk_hi - lnk_lo) / 256
    k = torch.exp(logk)
    r = torch.exp(logr)
    y = torch.outer(k, r)  # [257, 256]
    pk = linear_matter_power(cosmo, k)  # a = 1
    base = (pk * k ** 3).unsqueeze(1) * torch.exp(-(y ** 2)) / (2.0 * math.pi ** 2)  # [257, 256]
    g_a = growth_factor(cosmo, a)
    roots = []
    for s in range(0, a.shape[0], a_chunk):
        g = g_a[s:s + a_chunk]
        # int_sigma for every a of the chunk: [257, 256] * g^2 per a
        vals = base.unsqueeze(0) * (g ** 2).reshape(-1, 1, 1)  # [na, 257, 256]
        sigma = simps_samples(vals, dx, dim=1)  # [na, 256]
        for j in range(sigma.shape[0]):
            root = interp(torch.ones(1, dtype=DT), sigma[j], logr)
            # power.py:113-115: exp(root).clip(1e-6)
            roots.append(torch.maximum(torch.exp(root), torch.tensor(1e-6, dtype=DT)))
    R_nl = torch.cat(roots)
    k_nl = 1.0 / R_nl

    yy = torch.outer(k, 1.0 / k_nl)  # [257, Na]
    g = growth_factor(cosmo, a).unsqueeze(0)
    res = (pk * k ** 3).unsqueeze(1) * torch.exp(-(yy ** 2)) * g ** 2 / (2.0 * math.pi ** 2)
    dneff = 2 * res * yy ** 2
    dC = 4 * res * (yy ** 2 - yy ** 4)
    resi = simps_samples(torch.stack([dneff, dC], dim=1), dx, dim=0)  # [2, Na]
    n_eff = resi[0] - 3.0
    C = resi[0] ** 2 + resi[1]
    return k_nl, n_eff, C


def halofit(cosmo, k, a, params=None, prescription="takahashi2012"):
    """Halofit, power.py:144-262: prescription 'takahashi2012' (default) or
    'smith2003'.  ``k`` broadcasts against ``a`` on its last axis."""
    a = torch.atleast_1d(a)
    pklin = linear_matter_power(cosmo, k, a)
    if params is None:
        params = halofit_parameters(cosmo, a)
    k_nl, n, C = params
    om_m = Omega_m_a(cosmo, a)
    om_de = Omega_de_a(cosmo, a)
    wa_ = w(cosmo, a)

    if prescription == "smith2003":  # power.py:182-198, eq. C9-C18
        a_n = 10 ** (1.4861 + 1.8369 * n + 1.6762 * n ** 2 + 0.7940 * n ** 3 + 0.1670 * n ** 4 - 0.6206 * C)
        b_n = 10 ** (0.9463 + 0.9466 * n + 0.3084 * n ** 2 - 0.9400 * C)
        c_n = 10 ** (-0.2807 + 0.6669 * n + 0.3214 * n ** 2 - 0.0793 * C)
        gamma_n = 0.8649 + 0.2989 * n + 0.1631 * C
        alpha_n = 1.3884 + 0.3700 * n - 0.1452 * n ** 2
        beta_n = 0.8291 + 0.9854 * n + 0.3401 * n ** 2
        mu_n = 10 ** (-3.5442 + 0.1908 * n)
        nu_n = 10 ** (0.9585 + 1.2857 * n)
    elif prescription == "takahashi2012":
        a_n = 10 ** (1.5222 + 2.8553 * n + 2.3706 * n ** 2 + 0.9903 * n ** 3
                     + 0.2250 * n ** 4 - 0.6038 * C + 0.1749 * om_de * (1 + wa_))
        b_n = 10 ** (-0.5642 + 0.5864 * n + 0.5716 * n ** 2 - 1.5474 * C
                     + 0.2279 * om_de * (1 + wa_))
        c_n = 10 ** (0.3698 + 2.0404 * n + 0.8161 * n ** 2 + 0.5869 * C)
        gamma_n = 0.1971 - 0.0843 * n + 0.8460 * C
        alpha_n = torch.abs(6.0835 + 1.3373 * n - 0.1959 * n ** 2 - 5.5274 * C)
        beta_n = (2.0379 - 0.7354 * n + 0.3157 * n ** 2 + 1.2490 * n ** 3
                  + 0.3980 * n ** 4 - 0.1682 * C)
        mu_n = 0.0
        nu_n = 10 ** (5.2105 + 3.6902 * n)
    else:
        raise NotImplementedError(prescription)

    f1b = om_m ** (-0.0307)  # power.py:231-243
    f2b = om_m ** (-0.0585)
    f3b = om_m ** (0.0743)
    if prescription == "smith2003":
        frac = om_de / (1.0 - om_m)
        f1 = frac * f1b + (1 - frac) * om_m ** (-0.0732)
        f2 = frac * f2b + (1 - frac) * om_m ** (-0.1423)
        f3 = frac * f3b + (1 - frac) * om_m ** 0.0725
    else:
        f1, f2, f3 = f1b, f2b, f3b

    def f(x):
        return x / 4.0 + x ** 2 / 8.0

    d2l = k ** 3 * pklin / (2.0 * math.pi ** 2)
    y = k / k_nl
    d2q = d2l * ((1.0 + d2l) ** beta_n / (1 + alpha_n * d2l)) * torch.exp(-f(y))
    d2hprime = a_n * y ** (3 * f1) / (1.0 + b_n * y ** f2 + (c_n * f3 * y) ** (3.0 - gamma_n))
    d2h = d2hprime / (1.0 + mu_n / y + nu_n / y ** 2)
    d2nl = d2q + d2h
    return 2.0 * math.pi ** 2 / k ** 3 * d2nl


# ---------------------------------------------------------------- redshift
class NZ:
    """redshift_distribution, redshift.py:15-58: normalised by a 256-interval
    Simpson over [0, zmax]."""

    def __init__(self, kind, params, gals_per_arcmin2=1.0, zmax=10.0):
        self.kind = kind
        self.params = tuple(float(p) for p in params)
        self.gals_per_arcmin2 = gals_per_arcmin2
        self.zmax = zmax
        self._norm = None

    def pz_fn(self, z):
        if self.kind == "smail":  # redshift.py:75-77
            a, b, z0 = self.params
            return z ** a * torch.exp(-((z / z0) ** b))
        if self.kind == "gauss":
            # Gaussian bin, a redshift_distribution subclass in the pattern
            # of design.md:76-86 (the reference does not ship one).
            mu, sig = self.params
            return torch.exp(-0.5 * ((z - mu) / sig) ** 2)
        if self.kind == "fu":  # redshift.py:102-105
            a, b, c = self.params
            return (z ** a + z ** (a * b)) / (z ** b + c)
        if self.kind == "delta":  # redshift.py:120-123
            return torch.where(z == self.params[0], 1.0, 0.0).to(DT)
        raise NotImplementedError(self.kind)

    def __call__(self, z):  # redshift.py:27-31
        if self._norm is None:
            self._norm = simps(lambda t: self.pz_fn(t), 0.0, self.zmax, 256)
        return self.pz_fn(z) / self._norm

    @property
    def gals_per_steradian(self):  # redshift.py:44-47
        return self.gals_per_arcmin2 * STERADIAN_TO_ARCMIN2


def smail_nz(a, b, z0, gals_per_arcmin2=1.0, zmax=10.0):
    return NZ("smail", (a, b, z0), gals_per_arcmin2, zmax)


def gaussian_nz(mu, sigma, gals_per_arcmin2=1.0, zmax=10.0):
    return NZ("gauss", (mu, sigma), gals_per_arcmin2, zmax)


def fu_nz(a, b, c, gals_per_arcmin2=1.0, zmax=10.0):
    return NZ("fu", (a, b, c), gals_per_arcmin2, zmax)


def delta_nz(z0, gals_per_arcmin2=1.0, zmax=10.0):
    """redshift.py:108-123: a single plane, normalisation 1."""
    nz = NZ("delta", (z0,), gals_per_arcmin2, zmax)
    nz._norm = 1.0
    return nz


class kde_nz(NZ):
    """redshift.py:126-156: Gaussian KDE of a weighted catalog."""

    def __init__(self, zcat, weights, bw, gals_per_arcmin2=1.0, zmax=10.0):
        super().__init__("kde", (), gals_per_arcmin2, zmax)
        self.zcat = torch.as_tensor(zcat, dtype=DT)
        self.w = torch.atleast_1d(torch.as_tensor(weights, dtype=DT))
        self.bw = float(bw)

    def pz_fn(self, z):
        X = self.zcat.unsqueeze(-1)
        zz = z.reshape(1, -1)
        k = (1.0 / math.sqrt(2 * math.pi) / self.bw) * torch.exp(-((X - zz) ** 2) / (self.bw ** 2 * 2.0))
        return (self.w @ k / torch.sum(self.w)).reshape(z.shape)


class user_nz(NZ):
    """A user-defined redshift_distribution (redshift.py:15-31: a subclass
    that supplies pz_fn): ``fn`` maps a float64 tensor of redshifts to the
    un-normalised n(z), normalised like every distribution."""

    def __init__(self, fn, gals_per_arcmin2=1.0, zmax=10.0):
        super().__init__("user", (), gals_per_arcmin2, zmax)
        self.fn = fn

    def pz_fn(self, z):
        return self.fn(z)


class systematic_shift(NZ):
    """redshift.py:159-171: the parent's un-normalised n(z) at clip(z - bias, 0),
    normalised over this object's own zmax."""

    def __init__(self, parent, bias, gals_per_arcmin2=1.0, zmax=10.0):
        super().__init__("shift", (), gals_per_arcmin2, zmax)
        self.parent = parent
        self.bias = float(bias)

    def pz_fn(self, z):
        return self.parent.pz_fn(clip_min(z - self.bias, 0.0))


# ---------------------------------------------------------------- bias
def _bias_fn(b):
    """bias.py:10-59 models as callables (cosmo, z); a float is a
    constant_linear_bias."""
    if callable(b):
        return b
    return lambda cosmo, z: float(b) * torch.ones_like(z)


def inverse_growth_linear_bias(b):
    """bias.py:25-40: b / D(a)."""
    return lambda cosmo, z: b / growth_factor(cosmo, z2a(z))


def des_y1_ia_bias(A, eta, z0):
    """bias.py:43-59: A ((1 + z) / (1 + z0))^eta."""
    return lambda cosmo, z: A * ((1.0 + z) / (1.0 + z0)) ** eta


# ---------------------------------------------------------------- probes
class WeakLensing:
    """probes.py:132-223: optional NLA intrinsic alignments (one bias model or
    one per bin), multiplicative bias and sigma_e scalars or per bin."""

    def __init__(self, nzs, multiplicative_bias=0.0, sigma_e=0.26, ia_bias=None):
        self.nzs = list(nzs)
        self.m = multiplicative_bias
        self.sigma_e = sigma_e
        self.ia_bias = ia_bias

    @property
    def n_tracers(self):
        return len(self.nzs)

    @property
    def zmax(self):
        return max(nz.zmax for nz in self.nzs)

    def radial(self, cosmo, z):
        """weak_lensing_kernel radial part, probes.py:16-64: for extended
        n(z) per outer z its own 256-interval Simpson over [z, zmax]; for a
        single plane z_s, clip(chi_s - chi, 0) / clip(chi_s, 1)."""
        zmax = self.zmax
        chi = radial_comoving_distance(cosmo, z2a(z))

        def integrand(zp, nzs):  # zp: [257, Na]
            chi_p = radial_comoving_distance(cosmo, z2a(zp))
            dndz = torch.stack([nz(zp) for nz in nzs], dim=0)  # [nb, 257, Na]
            return dndz * clip_min(chi_p - chi, 0) / clip_min(chi_p, 1.0)

        out = [None] * len(self.nzs)
        ext = [i for i, nz in enumerate(self.nzs) if nz.kind != "delta"]
        if ext:
            # simps over the 257 axis with per-column dx: integrand output is
            # [nb, 257, Na]; move the 257 axis first.
            N = 256
            dx = (zmax - z) / N
            zp = linspace(z, torch.as_tensor(zmax, dtype=DT), N + 1)  # [257, Na]
            y = integrand(zp, [self.nzs[i] for i in ext]).movedim(1, 0)  # [257, nb, Na]
            integral = dx / 3 * torch.sum(y[0:-1:2] + 4 * y[1::2] + y[2::2], dim=0)
            for q, i in enumerate(ext):
                out[i] = integral[q] * (1.0 + z) * chi
        for i, nz in enumerate(self.nzs):
            if nz.kind == "delta":
                chi_s = radial_comoving_distance(cosmo, z2a(torch.tensor([nz.params[0]], dtype=DT)))
                out[i] = clip_min(chi_s - chi, 0) / clip_min(chi_s, 1.0) * (1.0 + z) * chi
        return torch.stack(out, dim=0)  # [nb, Na]

    def nla(self, cosmo, z):
        """nla_kernel radial part, probes.py:102-129: n(z) b(z) H(a) times
        -(C_1 rho_crit) Omega_m / D(a)."""
        if any(nz.kind == "delta" for nz in self.nzs):
            raise NotImplementedError("NLA kernel not properly implemented for delta redshift distributions")
        dndz = torch.stack([nz(z) for nz in self.nzs], dim=0)
        if isinstance(self.ia_bias, (list, tuple)):
            b = torch.stack([_bias_fn(bi)(cosmo, z) for bi in self.ia_bias], dim=0)
        else:
            b = _bias_fn(self.ia_bias)(cosmo, z)
        radial = dndz * b * H(cosmo, z2a(z))
        return radial * (-(5e-14 * RHOCRIT) * cosmo.Omega_m / growth_factor(cosmo, z2a(z)))

    def kernel(self, cosmo, z, ell):
        """probes.py:188-208 with probes.py:66-74: ell factor times (constant
        * lensing radial + NLA radial), times (1 + m).  Returns [nb, n_ell, Na]."""
        radial = 3.0 * H0 ** 2 * cosmo.Omega_m / 2.0 / C_KMS * self.radial(cosmo, z)
        if self.ia_bias is not None:
            radial = radial + self.nla(cosmo, z)
        ell = ell.reshape(-1, 1)
        ell_factor = torch.sqrt((ell - 1) * ell * (ell + 1) * (ell + 2)) / (ell + 0.5) ** 2
        m = self.m
        if isinstance(m, (list, tuple)):
            m = torch.tensor(m, dtype=DT).reshape(-1, 1, 1)
        return (ell_factor.unsqueeze(0) * radial.unsqueeze(1)) * (1.0 + m)

    def noise(self):  # probes.py:210-223
        ngals = torch.tensor([nz.gals_per_steradian for nz in self.nzs], dtype=DT)
        sig = self.sigma_e
        if isinstance(sig, (list, tuple)):
            sig = torch.tensor(sig, dtype=DT)
        return sig ** 2 / ngals


class NumberCounts:
    """probes.py:226-281 with density_kernel probes.py:77-99; bias a float
    (constant_linear_bias), a bias.py model or a list of them."""

    def __init__(self, nzs, bias):
        self.nzs = list(nzs)
        self.bias = bias

    @property
    def n_tracers(self):
        return len(self.nzs)

    @property
    def zmax(self):
        return max(nz.zmax for nz in self.nzs)

    def kernel(self, cosmo, z, ell):
        dndz = torch.stack([nz(z) for nz in self.nzs], dim=0)  # [nb, Na]
        if isinstance(self.bias, (list, tuple)):
            b = torch.stack([_bias_fn(bi)(cosmo, z) for bi in self.bias], dim=0)
        else:
            b = _bias_fn(self.bias)(cosmo, z)
        radial = dndz * b * H(cosmo, z2a(z))
        return radial.unsqueeze(1).expand(-1, ell.shape[0], -1)

    def noise(self):  # probes.py:274-281
        ngals = torch.tensor([nz.gals_per_steradian for nz in self.nzs], dtype=DT)
        return 1.0 / ngals


# ---------------------------------------------------------------- angular cl
def cl_ordering(n_tracers):
    """angular_cl.py:15-25: pairs (i, j), i <= j, row-major."""
    return [(i, j) for i in range(n_tracers) for j in range(i, n_tracers)]


def cov_blocks_ordering(n_tracers):
    """angular_cl.py:28-46."""
    cl_index = cl_ordering(n_tracers)
    pos = {p: q for q, p in enumerate(cl_index)}

    def find(a, b):
        return pos[(a, b)] if (a, b) in pos else pos[(b, a)]

    return [(find(i, m), find(j, n), find(i, n), find(j, m))
            for i, j in cl_index for m, n in cl_index]


def angular_cl(cosmo, ell, probes, N=512, prescription="takahashi2012"):
    """angular_cl.py:49-98: C_ij(ell) = Simpson over a (N intervals, uniform
    in a on [1/(1+zmax), 1]) of P_nl(k=(ell+1/2)/max(chi,1), a) K_i K_j
    dchi/da / max(chi^2, 1), divided by c^2.  Returns [P, n_ell].  The halofit
    prescription is the nonlinear_fn choice; the transfer function follows
    ``cosmo.transfer_type``."""
    ell = torch.atleast_1d(_t(ell) if not isinstance(ell, torch.Tensor) else ell)
    zmax = max(p.zmax for p in probes)
    a = linspace(z2a(zmax), 1.0, N + 1)  # [Na]
    chi = radial_comoving_distance(cosmo, a)
    k = (ell.reshape(-1, 1) + 0.5) / clip_min(chi, 1.0)  # [n_ell, Na]
    pk = halofit(cosmo, k, a, prescription=prescription)  # [n_ell, Na]
    kernels = torch.cat([p.kernel(cosmo, a2z(a), ell) for p in probes], dim=0)  # [T, n_ell, Na]
    T = kernels.shape[0]
    idx = cl_ordering(T)
    ii = torch.tensor([p[0] for p in idx])
    jj = torch.tensor([p[1] for p in idx])
    kk = kernels[ii] * kernels[jj]  # [P, n_ell, Na]
    result = pk * kk * dchioverda(cosmo, a) / clip_min(chi ** 2, 1.0)
    dx = (1.0 - z2a(zmax)) / N
    return simps_samples(result, dx, dim=2) / C_KMS ** 2


def noise_cl(ell, probes):
    """angular_cl.py:101-117."""
    n_ell = len(ell)
    noise = torch.cat([p.noise() for p in probes])
    idx = cl_ordering(noise.shape[0])
    rows = []
    for i, j in idx:
        delta = 1.0 - min(abs(i - j), 1)
        rows.append(noise[i] * delta * torch.ones(n_ell, dtype=DT))
    return torch.stack(rows)


def np_gradient(x):
    """np.gradient of a 1-D array with unit spacing (edge_order=1)."""
    g = torch.empty_like(x)
    g[1:-1] = (x[2:] - x[:-2]) / 2.0
    g[0] = x[1] - x[0]
    g[-1] = x[-1] - x[-2]
    return g


def gaussian_cl_covariance(ell, probes, cl_signal, cl_noise, f_sky=0.25, sparse=True):
    """angular_cl.py:120-163."""
    ell = torch.atleast_1d(_t(ell) if not isinstance(ell, torch.Tensor) else ell)
    n_ell = ell.shape[0]
    cl_obs = cl_signal + cl_noise
    n_cls = cl_obs.shape[0]
    norm = (2 * ell + 1) * np_gradient(ell) * f_sky
    T = sum(p.n_tracers for p in probes)
    blocks = torch.tensor(cov_blocks_ordering(T))
    cov = (cl_obs[blocks[:, 0]] * cl_obs[blocks[:, 1]]
           + cl_obs[blocks[:, 2]] * cl_obs[blocks[:, 3]]) / norm  # [P*P, n_ell]
    if sparse:
        return cov.reshape(n_cls, n_cls, n_ell)
    eye = torch.eye(n_ell, dtype=DT)
    cov = cov.reshape(n_cls, n_cls, n_ell, 1) * eye
    return cov.permute(0, 2, 1, 3).reshape(n_ell * n_cls, n_ell * n_cls)


def gaussian_cl_covariance_and_mean(cosmo, ell, probes, f_sky=0.25, sparse=False, N=512):
    """angular_cl.py:166-196."""
    ell = torch.atleast_1d(_t(ell) if not isinstance(ell, torch.Tensor) else ell)
    cl_signal = angular_cl(cosmo, ell, probes, N=N)
    cl_noise = noise_cl(ell, probes)
    cov = gaussian_cl_covariance(ell, probes, cl_signal, cl_noise, f_sky, sparse)
    return cl_signal.flatten(), cov


# ---------------------------------------------------------------- sparse / likelihood
def sparse_to_dense(s):
    """sparse.py:52-68: (a, b, nd) block-diagonal -> (a*nd, b*nd)."""
    a, b, nd = s.shape
    out = torch.zeros(a * nd, b * nd, dtype=s.dtype)
    for i in range(a):
        for j in range(b):
            out[i * nd:(i + 1) * nd, j * nd:(j + 1) * nd] = torch.diag(s[i, j])
    return out


def sparse_dot_vec(s, v):
    """sparse.py:140-162."""
    a, b, nd = s.shape
    return torch.einsum("ijk,jk->ik", s, v.reshape(b, nd)).reshape(-1)


def sparse_inv(s):
    """sparse.py:295-315: n_ell independent (P x P) inverses."""
    return torch.linalg.inv(s.permute(2, 0, 1)).permute(1, 2, 0)


def sparse_det(s):
    """Determinant of the dense equivalent (sparse.py:369-389 computes it by
    Schur complements; the value is the same)."""
    return torch.linalg.det(sparse_to_dense(s))


def gaussian_log_likelihood(data, mu, C, include_logdet=True):
    """likelihood.py:9-61 (sparse C of shape [P, P, n_ell] or dense)."""
    r = mu - data
    if C.dim() == 3:
        rT_Cinv_r = torch.dot(r, sparse_dot_vec(sparse_inv(C), r))
        if include_logdet:
            logdet = torch.logdet(sparse_to_dense(C))
    else:
        y = torch.linalg.solve(C, r)
        rT_Cinv_r = torch.dot(r, y)
        if include_logdet:
            logdet = torch.logdet(C)
    if not include_logdet:
        return -0.5 * rT_Cinv_r
    return -0.5 * (rT_Cinv_r - logdet)
