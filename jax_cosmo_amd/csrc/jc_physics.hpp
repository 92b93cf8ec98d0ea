// Cosmology physics of the Limber C_ell path, written once as templates on
// the scalar type (plain or dual, float or double; see jc_math.hpp).
//
// Reference files cited below are relative to the jax_cosmo repository root.
#pragma once
#include "jc_math.hpp"

namespace jc {

// jax_cosmo/constants.py:9-27
constexpr double C_KMS = 299792.458;
constexpr double TCMB = 2.726;
constexpr double RH = 2997.92458;
constexpr double H0 = 100.0;
constexpr double PI = 3.141592653589793;
constexpr double EULER_E = 2.718281828459045;  // np.exp(1.0), transfer.py:123
constexpr double LN10 = 2.302585092994046;

// Cosmology parameters in the order of Cosmology(...) (core.py:13), with the
// derived densities of core.py:143-163.
template <class S>
struct CosmoP {
  S Omega_c, Omega_b, h, n_s, sigma8, Omega_k, w0, wa;
  S Omega_m, Omega_de;
};

template <class S>
JC_D CosmoP<S> make_cosmo(const S* p) {
  CosmoP<S> c;
  c.Omega_c = p[0];
  c.Omega_b = p[1];
  c.h = p[2];
  c.n_s = p[3];
  c.sigma8 = p[4];
  c.Omega_k = p[5];
  c.w0 = p[6];
  c.wa = p[7];
  c.Omega_m = c.Omega_b + c.Omega_c;           // core.py:157-159
  using B = typename Base<S>::type;
  c.Omega_de = (B(1) - c.Omega_k) - c.Omega_m;  // core.py:145-147,161-163
  return c;
}

// background.py:25-53 (a is a plain grid value, never differentiated)
template <class S, class A>
JC_D S w_de(const CosmoP<S>& c, A a) {
  return c.w0 + (A(1) - a) * c.wa;
}
// background.py:56-90, given ln a
template <class S, class A>
JC_D S f_de_ln(const CosmoP<S>& c, A a, A lna) {
  return A(-3) * (A(1) + c.w0 + c.wa) * lna + A(3) * c.wa * (a - A(1));
}
template <class S, class A>
JC_D S f_de(const CosmoP<S>& c, A a) {
  return f_de_ln(c, a, jm::log(a));
}
// background.py:93-126 (np.power(a, -3) is an integer power: 1/(a a a)),
// given 1/a and exp(f_de)
template <class S, class A>
JC_D S Esqr_parts(const CosmoP<S>& c, A ia, const S& efde) {
  return c.Omega_m * (ia * ia * ia) + c.Omega_k * (ia * ia) + c.Omega_de * efde;
}
template <class S, class A>
JC_D S Esqr(const CosmoP<S>& c, A a) {
  return Esqr_parts(c, A(1) / a, jm::exp(f_de(c, a)));
}

// ------------------------------------------------------------------ EH98
// Per-cosmology constants of the Eisenstein & Hu (1998) transfer function
// with baryon oscillations (transfer.py:47-117,134-142).
template <class S>
struct EHc {
  S fb, fc;
  S inv13keq;  // 1/(13.41 k_eq), so q = k * inv13keq   (transfer.py:122)
  S ln_inv13keq;
  S beta_c18;  // 1.8 * beta_c                           (transfer.py:123)
  S c14a;      // 14.2 / alpha_c                          (transfer.py:124)
  S sh_d;      // sound horizon at drag [Mpc/h]           (transfer.py:70-75)
  S alpha_b, beta_b, beta_node;
  S ln_ksilk;   // ln k_silk
  S bnode_shd;  // beta_node / sh_d, so beta_node/(k sh_d) = bnode_shd / k
  S bb_shd;     // beta_b / sh_d
  // no-wiggle form (type="eisenhu", transfer.py:88-105)
  S alpha_g;    // alpha_gamma
  S tq;         // (T_cmb/2.7)^2 / (Omega_m h)
  int nw;       // 1: evaluate the no-wiggle form
};

// the same constants as another scalar type (dual parts zero)
template <class S2, class S1>
JC_D EHc<S2> eh_cast(const EHc<S1>& e) {
  EHc<S2> r;
  r.fb = Make<S2>::c(val(e.fb));
  r.fc = Make<S2>::c(val(e.fc));
  r.inv13keq = Make<S2>::c(val(e.inv13keq));
  r.ln_inv13keq = Make<S2>::c(val(e.ln_inv13keq));
  r.beta_c18 = Make<S2>::c(val(e.beta_c18));
  r.c14a = Make<S2>::c(val(e.c14a));
  r.sh_d = Make<S2>::c(val(e.sh_d));
  r.alpha_b = Make<S2>::c(val(e.alpha_b));
  r.beta_b = Make<S2>::c(val(e.beta_b));
  r.beta_node = Make<S2>::c(val(e.beta_node));
  r.ln_ksilk = Make<S2>::c(val(e.ln_ksilk));
  r.bnode_shd = Make<S2>::c(val(e.bnode_shd));
  r.bb_shd = Make<S2>::c(val(e.bb_shd));
  r.alpha_g = Make<S2>::c(val(e.alpha_g));
  r.tq = Make<S2>::c(val(e.tq));
  r.nw = e.nw;
  return r;
}

template <class S>
JC_D EHc<S> eh_setup(const CosmoP<S>& c) {
  using B = typename Base<S>::type;
  const B T27sq = B((TCMB / 2.7) * (TCMB / 2.7));
  S h2 = c.h * c.h;
  S w_m = c.Omega_m * h2;
  S w_b = c.Omega_b * h2;
  EHc<S> e;
  e.fb = c.Omega_b / c.Omega_m;
  e.fc = (c.Omega_m - c.Omega_b) / c.Omega_m;
  S k_eq = B(7.46e-2) * w_m / T27sq / c.h;      // eq. (3)
  S z_eq = B(2.50e4) * w_m / (T27sq * T27sq);    // eq. (2)
  // powers of w_m and w_b as exp(e ln x) from one logarithm each (the same
  // values as pow up to FP64 rounding; ln of the constant factors below)
  const S lwm = jm::log(w_m), lwb = jm::log(w_b);
  auto pwm = [&](B ln_c, B ex) { return jm::exp(ex * (lwm + ln_c)); };  // (c w_m)^ex
  S b1 = B(0.313) * pwm(B(0), B(-0.419)) * (B(1) + B(0.607) * pwm(B(0), B(0.674)));
  S b2 = B(0.238) * pwm(B(0), B(0.223));
  S z_d = B(1291.0) * pwm(B(0), B(0.251)) / (B(1) + B(0.659) * pwm(B(0), B(0.828))) *
          (B(1) + b1 * jm::exp(b2 * lwb));       // eq. (4)
  S R_d = B(31.5) * w_b / (T27sq * T27sq) * (B(1.0e3) / z_d);   // eq. (5)
  S R_eq = B(31.5) * w_b / (T27sq * T27sq) * (B(1.0e3) / z_eq);
  e.sh_d = B(2) / (B(3) * k_eq) * jm::sqrt(B(6) / R_eq) *
           jm::log((jm::sqrt(B(1) + R_d) + jm::sqrt(R_eq + R_d)) / (B(1) + jm::sqrt(R_eq)));
  S k_silk = B(1.6) * jm::exp(B(0.52) * lwb) * pwm(B(0), B(0.73)) *
             (B(1) + pwm(B(2.341805806147327), B(-0.95))) / c.h;   // eq. (7), ln 10.4
  e.ln_ksilk = jm::log(k_silk);
  // EH98 (11, 12), transfer.py:109-117
  // ln 46.9, ln 32.1; ln 12, ln 45
  S a1 = pwm(B(3.8480176754522337), B(0.670)) * (B(1) + pwm(B(3.4688560301359703), B(-0.532)));
  S a2 = pwm(B(2.4849066497880004), B(0.424)) * (B(1) + pwm(B(3.8066624897703196), B(-0.582)));
  S alpha_c = jm::exp(-e.fb * jm::log(a1)) * jm::exp(-(e.fb * e.fb * e.fb) * jm::log(a2));
  S bb1 = B(0.944) / (B(1) + pwm(B(6.126869184114185), B(-0.708)));  // 458
  S bb2 = pwm(B(-0.9288695140810151), B(-0.0266));                   // 0.395
  S beta_c = B(1) / (B(1) + bb1 * (jm::exp(bb2 * jm::log(e.fc)) - B(1)));
  e.inv13keq = B(1) / (B(13.41) * k_eq);
  e.ln_inv13keq = jm::log(e.inv13keq);
  e.beta_c18 = B(1.8) * beta_c;
  e.c14a = B(14.2) / alpha_c;
  // EH98 (19, 14, 21), transfer.py:134-138
  S y = (B(1) + z_eq) / (B(1) + z_d);
  S x = jm::sqrt(B(1) + y);
  S G = y * (B(-6) * x + (B(2) + B(3) * y) * jm::log((x + B(1)) / (x - B(1))));
  e.alpha_b = B(2.07) * k_eq * e.sh_d * jm::exp(B(-0.75) * jm::log(B(1) + R_d)) * G;
  e.beta_node = B(8.41) * pwm(B(0), B(0.435));
  e.beta_b = B(0.5) + e.fb + (B(3) - B(2) * e.fb) * jm::sqrt((B(17.2) * w_m) * (B(17.2) * w_m) + B(1));
  e.bnode_shd = e.beta_node / e.sh_d;
  e.bb_shd = e.beta_b / e.sh_d;
  // transfer.py:88-97
  e.alpha_g = B(1) - B(0.328) * jm::log(B(431.0) * w_m) * w_b / w_m +
              B(0.38) * jm::log(B(22.3) * w_m) * (e.fb * e.fb);
  e.tq = T27sq / (c.Omega_m * c.h);
  e.nw = 0;
  return e;
}

// A scalar of a lane (plain or dual), for the wave-wide form below.
JC_D double shfl_s(double x, int src) { return __shfl(x, src, 64); }
template <typename T, int N>
JC_D Dual<T, N> shfl_s(const Dual<T, N>& x, int src) {
  Dual<T, N> r;
  r.v = __shfl(x.v, src, 64);
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = __shfl(x.d[k], src, 64);
  return r;
}

// Per-lane arguments of the first exp pass of eh_setup_wave: the power
// (c w)^ex = exp(ex (ln w + ln c)) with w = w_b for lane 5, w_m otherwise.
__constant__ const double kEHPowLnC[16] = {0, 0, 0, 0, 0, 0, 0, 2.341805806147327, 3.8480176754522337,
                                           3.4688560301359703, 2.4849066497880004, 3.8066624897703196,
                                           6.126869184114185, -0.9288695140810151, 0, 0};
__constant__ const double kEHPowEx[16] = {-0.419, 0.674, 0.223, 0.251, 0.828, 0.52, 0.73, -0.95,
                                          0.670, -0.532, 0.424, -0.582, -0.708, -0.0266, 0.435, 0};

// eh_setup by the 64 lanes of one wave, all calling it, each getting the
// result: the independent powers and logarithms go to separate lanes, so
// the wave evaluates nine exp/log/sqrt passes instead of one thread's chain
// of ~35.  Every value comes from the same expression (same operands, same
// order) as in eh_setup, so the constants are identical.
template <class S>
JC_D EHc<S> eh_setup_wave(const CosmoP<S>& c, const int lane) {
  using B = typename Base<S>::type;
  const B T27sq = B((TCMB / 2.7) * (TCMB / 2.7));
  S h2 = c.h * c.h;
  S w_m = c.Omega_m * h2;
  S w_b = c.Omega_b * h2;
  EHc<S> e;
  e.fb = c.Omega_b / c.Omega_m;
  e.fc = (c.Omega_m - c.Omega_b) / c.Omega_m;
  S k_eq = B(7.46e-2) * w_m / T27sq / c.h;      // eq. (3)
  S z_eq = B(2.50e4) * w_m / (T27sq * T27sq);    // eq. (2)
  // ln w_m (lane 0) and ln w_b (lane 1)
  const S lx = jm::log(lane == 1 ? w_b : w_m);
  const S lwm = shfl_s(lx, 0), lwb = shfl_s(lx, 1);
  // the powers: P[k] from lane k
  const int k16 = lane & 15;
  const S pw = jm::exp(B(kEHPowEx[k16]) * ((k16 == 5 ? lwb : lwm) + B(kEHPowLnC[k16])));
  S P[15];
#pragma unroll
  for (int k = 0; k < 15; ++k) P[k] = shfl_s(pw, k);
  // ln f_c, ln(431 w_m), ln(22.3 w_m)
  const S l1 = jm::log(lane == 0 ? e.fc : (lane == 1 ? B(431.0) * w_m : B(22.3) * w_m));
  const S l_fc = shfl_s(l1, 0), l_431 = shfl_s(l1, 1), l_223 = shfl_s(l1, 2);
  // eq. (4), (7), (11, 12)
  S b1 = B(0.313) * P[0] * (B(1) + B(0.607) * P[1]);
  S b2 = B(0.238) * P[2];
  S k_silk = B(1.6) * P[5] * P[6] * (B(1) + P[7]) / c.h;
  S a1 = P[8] * (B(1) + P[9]);
  S a2 = P[10] * (B(1) + P[11]);
  S bb1 = B(0.944) / (B(1) + P[12]);
  S bb2 = P[13];
  e.inv13keq = B(1) / (B(13.41) * k_eq);
  const S l2 = jm::log(lane == 0 ? a1 : (lane == 1 ? a2 : (lane == 2 ? k_silk : e.inv13keq)));
  const S la1 = shfl_s(l2, 0), la2 = shfl_s(l2, 1);
  e.ln_ksilk = shfl_s(l2, 2);
  e.ln_inv13keq = shfl_s(l2, 3);
  const S x2 = lane == 0 ? b2 * lwb : (lane == 1 ? -e.fb * la1 : (lane == 2 ? -(e.fb * e.fb * e.fb) * la2 : bb2 * l_fc));
  const S ex2 = jm::exp(x2);
  const S e_zd = shfl_s(ex2, 0);
  S alpha_c = shfl_s(ex2, 1) * shfl_s(ex2, 2);
  S beta_c = B(1) / (B(1) + bb1 * (shfl_s(ex2, 3) - B(1)));
  S z_d = B(1291.0) * P[3] / (B(1) + B(0.659) * P[4]) * (B(1) + b1 * e_zd);
  S R_d = B(31.5) * w_b / (T27sq * T27sq) * (B(1.0e3) / z_d);   // eq. (5)
  S R_eq = B(31.5) * w_b / (T27sq * T27sq) * (B(1.0e3) / z_eq);
  S y = (B(1) + z_eq) / (B(1) + z_d);
  // eq. (6), (14): the square roots, then the logarithms
  const S sq = jm::sqrt(lane == 0 ? B(6) / R_eq
                                  : (lane == 1 ? B(1) + R_d : (lane == 2 ? R_eq + R_d : (lane == 3 ? R_eq : B(1) + y))));
  const S s6 = shfl_s(sq, 0), s1d = shfl_s(sq, 1), sed = shfl_s(sq, 2), seq = shfl_s(sq, 3), x = shfl_s(sq, 4);
  const S l3 = jm::log(lane == 0 ? (s1d + sed) / (B(1) + seq) : (lane == 1 ? (x + B(1)) / (x - B(1)) : B(1) + R_d));
  e.sh_d = B(2) / (B(3) * k_eq) * s6 * shfl_s(l3, 0);
  S G = y * (B(-6) * x + (B(2) + B(3) * y) * shfl_s(l3, 1));
  const S e_b = jm::exp(B(-0.75) * shfl_s(l3, 2));
  e.beta_c18 = B(1.8) * beta_c;
  e.c14a = B(14.2) / alpha_c;
  e.alpha_b = B(2.07) * k_eq * e.sh_d * e_b * G;
  e.beta_node = B(8.41) * P[14];
  e.beta_b = B(0.5) + e.fb + (B(3) - B(2) * e.fb) * jm::sqrt((B(17.2) * w_m) * (B(17.2) * w_m) + B(1));
  e.bnode_shd = e.beta_node / e.sh_d;
  e.bb_shd = e.beta_b / e.sh_d;
  e.alpha_g = B(1) - B(0.328) * l_431 * w_b / w_m + B(0.38) * l_223 * (e.fb * e.fb);
  e.tq = T27sq / (c.Omega_m * c.h);
  e.nw = 0;
  return e;
}

// sin(x)/x accurate for all x >= 0 (np.sinc(x/pi))
template <class S>
JC_D S sinc_x(const S& x) {
  using B = typename Base<S>::type;
  if (val(x) < B(0.25)) {
    S x2 = x * x;
    return B(1) - x2 * (B(1.0 / 6.0) - x2 * (B(1.0 / 120.0) - x2 * (B(1.0 / 5040.0) - x2 * B(1.0 / 362880.0))));
  }
  return jm::fdiv(jm::sin(x), x);
}

// T(k): transfer.py:119-149.  The caller supplies k, 1/k and ln k; every
// k-independent factor comes from eh_setup, so a point costs 16
// transcendental instructions.
template <class S, class K>
JC_D S eh_transfer(const EHc<S>& e, const K& k, const K& invk, const K& lnk) {
  using B = typename Base<S>::type;
  if (e.nw) {
    // transfer.py:93-105 (type="eisenhu"): gamma_eff and the EH98 (29) fit
    S u = B(0.43) * k * e.sh_d;
    u = u * u;
    S q = k * e.tq / (e.alpha_g + (B(1) - e.alpha_g) / (B(1) + u * u));
    S L = jm::log(B(2.0 * EULER_E) + B(1.8) * q);
    S Cq = B(14.2) + B(731.0) / (B(1) + B(62.5) * q);
    return L / (L + Cq * q * q);
  }
  S q = k * e.inv13keq;
  S lnq = lnk + e.ln_inv13keq;
  S q2 = q * q;
  // 386/(1+69.9 q^1.08) is shared by the three T_tilde calls
  S ccom = jm::fdiv(B(386), B(1) + B(69.9) * jm::exp(B(1.08) * lnq));
  S L_bc = jm::log(B(EULER_E) + e.beta_c18 * q);   // beta = beta_c (twice)
  S L_11 = jm::log(B(EULER_E) + B(1.8) * q);       // beta = 1
  S C1q2 = (B(14.2) + ccom) * q2;
  S T1bc = jm::fdiv(L_bc, L_bc + C1q2);            // T_tilde(k, 1, beta_c)
  S Tabc = jm::fdiv(L_bc, L_bc + (e.c14a + ccom) * q2);  // T_tilde(k, alpha_c, beta_c)
  S T11 = jm::fdiv(L_11, L_11 + C1q2);             // T_tilde(k, 1, 1)
  S ksh = k * e.sh_d;
  S u = ksh * B(1.0 / 5.4);
  u = u * u;
  S f = jm::fdiv(B(1), B(1) + u * u);              // eq. (18)
  S Tc = Tabc + f * (T1bc - Tabc);                 // f T1 + (1 - f) T2
  S xn = e.bnode_shd * invk;                       // beta_node / (k sh_d)
  S ks = ksh * jm::exp(B(-1.0 / 3.0) * jm::log(B(1) + xn * xn * xn));  // k tilde_s
  S v = ksh * B(1.0 / 5.2);
  S xb = e.bb_shd * invk;                          // beta_b / (k sh_d)
  S silk = jm::exp(-jm::exp(B(1.4) * (lnk - e.ln_ksilk)));
  S Tb = (jm::fdiv(T11, B(1) + v * v) + jm::fdiv(e.alpha_b, B(1) + xb * xb * xb) * silk) * sinc_x(ks);
  return e.fb * Tb + e.fc * Tc;
}

// ------------------------------------------------------------------ halofit
// Per-scale-factor Takahashi (2012) coefficients, power.py:199-238.
template <class S>
struct HFa {
  S a_n, b_n, lg_cf3, gam3, alpha_n, beta_n, nu_n, f1x3, f2, mu_n;  // lg_cf3 = log2(c_n f3)
};

// smith: the smith2003 prescription (power.py:182-198, 226-241), else
// takahashi2012 (power.py:199-224).  The coefficients are stored and used as
// floats, so their exponentials and logarithms are taken to float accuracy
// (expf_acc, logf_acc) from FP64 arguments.
template <class S>
JC_D HFa<S> halofit_coefs(const S& n, const S& C, const S& om_m, const S& om_de, const S& w, int smith = 0) {
  using B = typename Base<S>::type;
  const B ln10 = B(LN10);
  S n2 = n * n, n3 = n2 * n, n4 = n3 * n;
  HFa<S> h;
  if (smith) {
    h.a_n = jm::expf_acc(ln10 * (B(1.4861) + B(1.8369) * n + B(1.6762) * n2 + B(0.7940) * n3 + B(0.1670) * n4 -
                            B(0.6206) * C));
    h.b_n = jm::expf_acc(ln10 * (B(0.9463) + B(0.9466) * n + B(0.3084) * n2 - B(0.9400) * C));
    S c_n = jm::expf_acc(ln10 * (B(-0.2807) + B(0.6669) * n + B(0.3214) * n2 - B(0.0793) * C));
    S gamma_n = B(0.8649) + B(0.2989) * n + B(0.1631) * C;
    h.alpha_n = B(1.3884) + B(0.3700) * n - B(0.1452) * n2;
    h.beta_n = B(0.8291) + B(0.9854) * n + B(0.3401) * n2;
    h.mu_n = jm::expf_acc(ln10 * (B(-3.5442) + B(0.1908) * n));
    h.nu_n = jm::expf_acc(ln10 * (B(0.9585) + B(1.2857) * n));
    // f_i = frac f_ib + (1 - frac) f_ia, frac = Omega_de / (1 - Omega_m)
    S lom = jm::logf_acc(om_m);
    S frac = om_de / (B(1) - om_m);
    S f1 = frac * jm::expf_acc(B(-0.0307) * lom) + (B(1) - frac) * jm::expf_acc(B(-0.0732) * lom);
    h.f2 = frac * jm::expf_acc(B(-0.0585) * lom) + (B(1) - frac) * jm::expf_acc(B(-0.1423) * lom);
    S f3 = frac * jm::expf_acc(B(0.0743) * lom) + (B(1) - frac) * jm::expf_acc(B(0.0725) * lom);
    h.lg_cf3 = jm::logf_acc(c_n * f3) * B(1.4426950408889634);
    h.gam3 = B(3) - gamma_n;
    h.f1x3 = B(3) * f1;
    return h;
  }
  S dew = om_de * (B(1) + w);
  h.mu_n = Make<S>::c(0.0);
  h.a_n = jm::expf_acc(ln10 * (B(1.5222) + B(2.8553) * n + B(2.3706) * n2 + B(0.9903) * n3 +
                          B(0.2250) * n4 - B(0.6038) * C + B(0.1749) * dew));
  h.b_n = jm::expf_acc(ln10 * (B(-0.5642) + B(0.5864) * n + B(0.5716) * n2 - B(1.5474) * C +
                          B(0.2279) * dew));
  S c_n = jm::expf_acc(ln10 * (B(0.3698) + B(2.0404) * n + B(0.8161) * n2 + B(0.5869) * C));
  S gamma_n = B(0.1971) - B(0.0843) * n + B(0.8460) * C;
  h.alpha_n = jm::fabs(B(6.0835) + B(1.3373) * n - B(0.1959) * n2 - B(5.5274) * C);
  h.beta_n = B(2.0379) - B(0.7354) * n + B(0.3157) * n2 + B(1.2490) * n3 + B(0.3980) * n4 -
             B(0.1682) * C;
  h.nu_n = jm::expf_acc(ln10 * (B(5.2105) + B(3.6902) * n));
  S lom = jm::logf_acc(om_m);
  S f1 = jm::expf_acc(B(-0.0307) * lom);
  h.f2 = jm::expf_acc(B(-0.0585) * lom);
  S f3 = jm::expf_acc(B(0.0743) * lom);
  h.lg_cf3 = jm::logf_acc(c_n * f3) * B(1.4426950408889634);
  h.gam3 = B(3) - gamma_n;
  h.f1x3 = B(3) * f1;
  return h;
}

// P_nl(k, a), power.py:244-262 (mu_n = 0): returns 2 pi^2/k^3 (Delta_Q^2 +
// Delta_H^2) given log2 y, y = k/k_nl, 1/y, k^3, 1/k^3 and P_lin.
template <class S>
JC_D S halofit_pnl(const HFa<S>& h, const S& lgy, const S& y, const S& iy, const S& k3, const S& ik3,
                   const S& plin) {
  using B = typename Base<S>::type;
  S d2l = k3 * plin * B(1.0 / (2.0 * PI * PI));
  // f(y) = y/4 + y^2/8, in base-2 units (times log2 e)
  S fy2 = y * (B(0.25 * 1.4426950408889634) + y * B(0.125 * 1.4426950408889634));
  // Delta_Q^2 = Delta_L^2 (1+d2l)^beta / (1 + alpha d2l) exp(-f(y)) and
  // Delta_H^2 = a y^(3 f1) / (1 + b y^f2 + (c f3 y)^(3-gamma)) / (1 + mu/y + nu/y^2);
  // the three quotients share one reciprocal:
  //   P = [P_lin e d1 d2 + 2 pi^2 k^-3 a y^(3 f1) d0] / (d0 d1 d2)
  // powers as 2^(e log2 x): base-2 exponent and logarithm are the hardware's
  S eq = jm::exp2(h.beta_n * jm::log2(B(1) + d2l) - fy2);
  S yf1 = jm::exp2(h.f1x3 * lgy);
  S yf2 = jm::exp2(h.f2 * lgy);
  S ycg = jm::exp2(h.gam3 * (h.lg_cf3 + lgy));
  S d0 = B(1) + h.alpha_n * d2l;
  S d12 = (B(1) + h.b_n * yf2 + ycg) * (B(1) + h.mu_n * iy + h.nu_n * (iy * iy));
  return jm::fdiv(plin * eq * d12 + B(2.0 * PI * PI) * ik3 * (h.a_n * yf1) * d0, d0 * d12);
}

}  // namespace jc
