// Per-cosmology tables (FP64) and the weak-lensing radial kernels.
//
// k_tables: one 256-thread workgroup per cosmology builds, in FP64,
//   * the chi(a) table: RK4 in ln a of a y-independent right-hand side, i.e.
//     a cumulative Simpson with midpoint nodes (background.py:223-233,
//     scipy/ode.py:6-22), as a parallel prefix sum;
//   * the growth table: RK4 over a of [D, dD/da] (background.py:443-487),
//     right-hand-side coefficients evaluated in parallel, the 127 steps on
//     one lane;
//   * sigma_8 normalisation by the 129-sample Romberg table over ln k in
//     [-4, 3] (power.py:56-78, integrate.py:67-159);
//   * the halofit sigma^2(R) table on 256 R nodes from a 257-node ln k Simpson
//     (power.py:96-111).  sigma^2(R, a) = D(a)^2 sigma^2(R, 1) exactly, so one
//     table per cosmology replaces the reference's [Na, 257, 256] tensor, and
//     the exp(-(k R)^2) factors are a cosmology-independent resident table;
//   * per a-node: chi, D, the nonlinear scale by the reference's
//     nearest-node interpolation of the decreasing sigma^2 table
//     (power.py:112-115), n_eff and C (power.py:118-139), the Takahashi
//     coefficients (power.py:199-238), the Simpson weight of the Limber
//     integral (angular_cl.py:91-96) and the number-count kernels
//     (probes.py:77-99).
// k_wl: the weak-lensing radial kernel (probes.py:16-51) with the
//   reference's own per-z 256-interval Simpson over [z, zmax]: the
//   cosmology-independent parts (n(z') and Simpson weights, the chi-table
//   segment and offset of every z') are resident tables, so per cosmology
//   only the chi lookup and the (chi' - chi)/chi' factor are evaluated.
#include "jc_kernels.hpp"
#include "jc_physics.hpp"

namespace jc {

template <int ND>
struct Seed {
  static JC_D typename Types<ND>::D load(const double* params, const double* seeds, int c, int i) {
    typename Types<ND>::D r;
    r.v = params[c * 8 + i];
#pragma unroll
    for (int d = 0; d < ND; ++d) r.d[d] = seeds[(c * 8 + i) * ND + d];
    return r;
  }
};
template <>
struct Seed<0> {
  static JC_D double load(const double* params, const double*, int c, int i) { return params[c * 8 + i]; }
};

template <class S>
JC_D S zero() { return Make<S>::c(0.0); }

// interpolate.py:12-37 on a table with log-uniform nodes xp[0..n-1]
// (logspace(-3, 0, n)): nearest node (first on ties) clipped to [1, n-2],
// side from clip(x, xp[1], xp[n-2]), line evaluated as a*x + b.
template <class S>
JC_D S interp_logtab(double x, const double* xp, const S* fp, int n) {
  double u = (::log10(x) + 3.0) * (double)(n - 1) / 3.0;
  int i0 = (int)::floor(u);
  i0 = i0 < 0 ? 0 : (i0 > n - 1 ? n - 1 : i0);
  int ind = i0;
  // neighbours (covers any rounding in the log index); scan upward so the
  // first index wins ties
  int lo = i0 - 1 < 0 ? 0 : i0 - 1;
  int hi = i0 + 2 > n - 1 ? n - 1 : i0 + 2;
  double best = 1e300;
  for (int j = lo; j <= hi; ++j) {
    double dd = (x - xp[j]) * (x - xp[j]);
    if (dd < best) { best = dd; ind = j; }
  }
  ind = ind < 1 ? 1 : (ind > n - 2 ? n - 2 : ind);
  double xc = x > xp[1] ? x : xp[1];
  xc = xc < xp[n - 2] ? xc : xp[n - 2];
  int nb = (xc - xp[ind] < 0.0) ? ind - 1 : ind + 1;
  S a = (fp[nb] - fp[ind]) / (xp[nb] - xp[ind]);
  S b = fp[ind] - a * xp[ind];
  return a * x + b;
}

template <int ND>
__global__ __launch_bounds__(256) void k_tables(TablesArgs A) {
  using Sd = typename Types<ND>::D;
  using Sf = typename Types<ND>::F;
  constexpr int W = ND + 1;
  extern __shared__ __align__(16) unsigned char smem[];
  Sd* s_chi = reinterpret_cast<Sd*>(smem);  // 256
  Sd* s_g = s_chi + NCHI;                    // 128 (raw D, then normalised)
  Sd* s_sig = s_g + NGROW;                   // 256 (Romberg samples first)
  Sd* s_p1 = s_sig + NHF_R;                  // 257 (T(k) first)
  Sd* s_tmp = s_p1 + NHF_K;                  // 512
  Sf* s_p1f = reinterpret_cast<Sf*>(s_tmp + 512);  // 257

  const int c = blockIdx.x;
  const int tid = threadIdx.x;
  if (c >= A.B) return;
  if (A.dB && c >= *A.dB) return;  // served batches: grid sized for capacity

  Sd p[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) p[i] = Seed<ND>::load(A.params, A.seeds, c, i);
  const CosmoP<Sd> cp = make_cosmo(p);

  // ---------------------------------------------------------- chi table
  {
    const double ti = ::log(A.atab_chi[tid]);
    const double ai = ::exp(ti);
    Sd fi = RH / (ai * ai * jm::sqrt(Esqr(cp, ai))) * ai;  // dchioverda(e^t) e^t
    s_tmp[256 + tid] = fi;
    Sd fm = zero<Sd>();
    double h = 0.0;
    if (tid > 0) {
      const double tp = ::log(A.atab_chi[tid - 1]);
      h = ti - tp;
      const double am = ::exp(tp + h / 2);
      fm = RH / (am * am * jm::sqrt(Esqr(cp, am))) * am;
    }
    __syncthreads();
    Sd delta = zero<Sd>();
    if (tid > 0) {
      Sd fp = s_tmp[256 + tid - 1];
      delta = (1.0 / 6.0) * h * (fp + 2.0 * fm + 2.0 * fm + fi);
    }
    s_tmp[tid] = delta;
    __syncthreads();
    // inclusive Hillis-Steele scan
    for (int off = 1; off < 256; off <<= 1) {
      Sd v = s_tmp[tid];
      if (tid >= off) v = v + s_tmp[tid - off];
      __syncthreads();
      s_tmp[tid] = v;
      __syncthreads();
    }
    s_chi[tid] = s_tmp[255] - s_tmp[tid];
    __syncthreads();
  }

  // ---------------------------------------------------------- growth RHS
  // q(x), r(x) of background.py:465-475 at nodes (s_tmp[0..255]) and at the
  // RK4 midpoints a_{i-1} + h/2 (s_tmp[256..511])
  if (tid < NGROW) {
    const double x = A.atab_g[tid];
    Sd E2 = Esqr(cp, x);
    Sd omm = cp.Omega_m * (1.0 / (x * x * x)) / E2;
    Sd omde = cp.Omega_de * jm::exp(f_de(cp, x)) / E2;
    Sd q = (2.0 - 0.5 * (omm + (1.0 + 3.0 * w_de(cp, x)) * omde)) / x;
    Sd r = 1.5 * omm / x / x;
    s_tmp[2 * tid] = q;
    s_tmp[2 * tid + 1] = r;
  } else if (tid >= 128 && tid < 128 + NGROW) {
    const int i = tid - 128;
    if (i > 0) {
      const double xp = A.atab_g[i - 1];
      const double h = A.atab_g[i] - xp;
      const double x = xp + h / 2;
      Sd E2 = Esqr(cp, x);
      Sd omm = cp.Omega_m * (1.0 / (x * x * x)) / E2;
      Sd omde = cp.Omega_de * jm::exp(f_de(cp, x)) / E2;
      Sd q = (2.0 - 0.5 * (omm + (1.0 + 3.0 * w_de(cp, x)) * omde)) / x;
      Sd r = 1.5 * omm / x / x;
      s_tmp[256 + 2 * i] = q;
      s_tmp[256 + 2 * i + 1] = r;
    }
  }
  __syncthreads();

  const EHc<Sd> eh = eh_setup(cp);
  if (tid == 0) {
    // ode.py:6-22: the first step has h = 0 and returns y0
    Sd y0 = Make<Sd>::c(A.atab_g[0]);
    Sd y1 = Make<Sd>::c(1.0);
    s_g[0] = y0;
    for (int i = 1; i < NGROW; ++i) {
      const double h = A.atab_g[i] - A.atab_g[i - 1];
      const Sd q0 = s_tmp[2 * (i - 1)], r0 = s_tmp[2 * (i - 1) + 1];
      const Sd qm = s_tmp[256 + 2 * i], rm = s_tmp[256 + 2 * i + 1];
      const Sd q1 = s_tmp[2 * i], r1 = s_tmp[2 * i + 1];
      Sd k1a = y1, k1b = -q0 * y1 + r0 * y0;
      Sd ya = y0 + h * k1a / 2.0, yb = y1 + h * k1b / 2.0;
      Sd k2a = yb, k2b = -qm * yb + rm * ya;
      ya = y0 + h * k2a / 2.0;
      yb = y1 + h * k2b / 2.0;
      Sd k3a = yb, k3b = -qm * yb + rm * ya;
      ya = y0 + h * k3a;
      yb = y1 + h * k3b;
      Sd k4a = yb, k4b = -q1 * yb + r1 * ya;
      y0 = y0 + 1.0 / 6.0 * h * (k1a + 2.0 * k2a + 2.0 * k3a + k4a);
      y1 = y1 + 1.0 / 6.0 * h * (k1b + 2.0 * k2b + 2.0 * k3b + k4b);
      s_g[i] = y0;
    }
  } else if (tid >= 64) {
    // EH98 at the Romberg samples (power.py:66-73) and on the halofit ln k grid
    for (int task = tid - 64; task < NROMB + NHF_K; task += 192) {
      if (task < NROMB) {
        const double lk = -4.0 + task * (7.0 / 128.0);
        const double k = ::exp(lk);
        const double x = k * 8.0;
        const double wv = 3.0 * (::sin(x) - x * ::cos(x)) / (x * x * x);
        Sd T = eh_transfer(eh, k, 1.0 / k, lk);
        Sd kns = jm::exp(cp.n_s * lk);
        s_sig[task] = k * ((k * wv) * (k * wv)) * (T * T * kns);
      } else {
        const int i = task - NROMB;
        const double lk = A.hf_lnk[i];
        const double k = ::exp(lk);
        s_p1[i] = eh_transfer(eh, k, 1.0 / k, lk);
      }
    }
  }
  __syncthreads();

  if (tid == 0) {
    // integrate.py:67-159, divmax = 7, over [log10(1e-4), log10(1e3)]
    const double intrange = 7.0;
    Sd ordsum = 0.5 * (s_sig[0] + s_sig[128]);
    Sd state[8];
    for (int k = 0; k < 8; ++k) state[k] = intrange * ordsum;
    for (int i = 1; i <= 7; ++i) {
      const int n = 1 << i;
      const int stride = 128 >> (i - 1);
      for (int m = 0; m < n / 2; ++m) ordsum = ordsum + s_sig[stride / 2 + m * stride];
      Sd x = intrange * ordsum / (double)n;
      Sd ns[8];
      ns[0] = x;
      double tmp = 1.0;
      for (int k = 0; k < 7; ++k) {
        tmp *= 4.0;
        x = (tmp * x - state[k]) / (tmp - 1.0);
        ns[k + 1] = x;
      }
      for (int k = 0; k < 8; ++k) state[k] = ns[k];
    }
    Sd sig2 = state[7] * (1.0 / (2.0 * PI * PI));
    s_tmp[0] = cp.sigma8 * cp.sigma8 / sig2;  // pknorm
    Sd g1raw = s_g[NGROW - 1];
    for (int i = 0; i < NGROW; ++i) s_g[i] = s_g[i] / g1raw;  // background.py:482
    Sd g1 = interp_logtab(1.0, A.atab_g, s_g, NGROW);
    g1 = minc(maxc(g1, 0.0), 1.0);
    s_tmp[1] = g1;
  }
  __syncthreads();
  const Sd pknorm = s_tmp[0];
  {
    const Sd g1 = s_tmp[1];
    for (int i = tid; i < NHF_K; i += 256) {
      const double lk = A.hf_lnk[i];
      const double k = ::exp(lk);
      Sd T = s_p1[i];
      // linear_matter_power at a = 1 (power.py:21-53) times k^3/(2 pi^2)
      Sd pk = jm::exp(cp.n_s * lk) * (T * T) * (g1 * g1) * pknorm;
      Sd v = A.hf_w[i] * (pk * (k * k * k) * (1.0 / (2.0 * PI * PI)));
      s_p1[i] = v;
      s_p1f[i] = conv<Sf>(v);
    }
  }
  __syncthreads();
  {
    // sigma^2(R_j, a=1) without D^2: sum_i w_i P1_i exp(-(k_i R_j)^2)
    Sd acc = zero<Sd>();
    const double* E = A.hf_E + tid;
    for (int i = 0; i < NHF_K; ++i) acc = acc + s_p1[i] * E[i * NHF_R];
    s_sig[tid] = acc;
  }
  __syncthreads();

  // ---------------------------------------------------------- per a-node
  const double c2 = C_KMS * C_KMS;
  for (int m = tid; m < A.Na; m += 256) {
    const double a = A.a_list[m];
    Sd chi = maxc(interp_logtab(a, A.atab_chi, s_chi, NCHI), 0.0);   // background.py:242
    Sd g = interp_logtab(a, A.atab_g, s_g, NGROW);
    g = minc(maxc(g, 0.0), 1.0);                                       // background.py:488
    Sd E2 = Esqr(cp, a);
    Sd sqE = jm::sqrt(E2);
    Sd dchida = RH / (a * a * sqE);                                    // background.py:294
    Sd g2 = g * g;
    // nonlinear scale: interp(1.0, sigma^2(R, a), ln R) with sigma^2 decreasing
    const double g2v = val(g2);
    int lo = 0, hi = NHF_R;  // first j with g2 * sig_j < 1
    while (lo < hi) {
      int mid = (lo + hi) >> 1;
      if (g2v * val(s_sig[mid]) < 1.0) hi = mid; else lo = mid + 1;
    }
    int ind;
    if (lo == 0) ind = 0;
    else if (lo == NHF_R) ind = NHF_R - 1;
    else {
      const double d0 = (1.0 - g2v * val(s_sig[lo - 1])) * (1.0 - g2v * val(s_sig[lo - 1]));
      const double d1 = (1.0 - g2v * val(s_sig[lo])) * (1.0 - g2v * val(s_sig[lo]));
      ind = (d1 < d0) ? lo : lo - 1;
    }
    ind = ind < 1 ? 1 : (ind > NHF_R - 2 ? NHF_R - 2 : ind);
    // clip(1, xp[1], xp[-2]) returns xp[-2] for a decreasing table, so the
    // side is sign(xp[254] - xp[ind]): left segment unless ind == 254
    const double xcl = g2v * val(s_sig[NHF_R - 2]);
    const int nb = (xcl - g2v * val(s_sig[ind]) < 0.0) ? ind - 1 : ind + 1;
    Sd Sind = g2 * s_sig[ind];
    Sd Snb = g2 * s_sig[nb];
    Sd slope = (A.logr[nb] - A.logr[ind]) / (Snb - Sind);
    Sd icpt = A.logr[ind] - slope * Sind;
    Sd root = slope * 1.0 + icpt;
    Sd Rnl = maxc(jm::exp(root), 1e-6);                                // power.py:113-115
    // n_eff and C (power.py:118-139), float terms, double sums
    Sf Rf = conv<Sf>(Rnl);
    Sf g2f = conv<Sf>(g2);
    Sd sA = zero<Sd>(), sB = zero<Sd>();
    for (int i = 0; i < NHF_K; ++i) {
      Sf y = A.hf_k32[i] * Rf;
      Sf y2 = y * y;
      if (val(y2) > 87.0f) break;  // exp underflows; k grows with i
      Sf res = s_p1f[i] * g2f * jm::exp(-y2);
      sA = sA + conv<Sd>(2.0f * res * y2);
      sB = sB + conv<Sd>(4.0f * res * (y2 - y2 * y2));
    }
    Sd n_eff = sA - 3.0;
    Sd Cc = sA * sA + sB;
    Sd omm = cp.Omega_m * (1.0 / (a * a * a)) / E2;
    Sd omde = cp.Omega_de * jm::exp(f_de(cp, a)) / E2;
    HFa<Sd> hf = halofit_coefs(n_eff, Cc, omm, omde, w_de(cp, a));
    Sd wgt = A.a_wgt[m] * dchida / maxc(chi * chi, 1.0) / c2;          // angular_cl.py:91,96
    Sd chi1 = maxc(chi, 1.0);                                          // angular_cl.py:73
    Sd invchi = 1.0 / chi1;

    float* cf = A.coef + ((size_t)c * A.Na + m) * NCOEF * W;
    Mem<Sf>::st(cf + CF_INV_CHI * W, conv<Sf>(invchi));
    Mem<Sf>::st(cf + CF_CHI1 * W, conv<Sf>(chi1));
    Mem<Sf>::st(cf + CF_LNINVCHI * W, conv<Sf>(-jm::log(chi1)));
    Mem<Sf>::st(cf + CF_KNL * W, conv<Sf>(1.0 / Rnl));
    Mem<Sf>::st(cf + CF_PAD * W, Make<Sf>::c(0.0));
    Mem<Sf>::st(cf + CF_WGT * W, conv<Sf>(wgt));
    Mem<Sf>::st(cf + CF_G2P * W, conv<Sf>(g2 * pknorm));
    Mem<Sf>::st(cf + CF_LNRNL * W, conv<Sf>(jm::log(Rnl)));
    Mem<Sf>::st(cf + CF_RNL * W, conv<Sf>(Rnl));
    Mem<Sf>::st(cf + CF_AN * W, conv<Sf>(hf.a_n));
    Mem<Sf>::st(cf + CF_BN * W, conv<Sf>(hf.b_n));
    Mem<Sf>::st(cf + CF_LNCF3 * W, conv<Sf>(hf.ln_cf3));
    Mem<Sf>::st(cf + CF_GAM3 * W, conv<Sf>(hf.gam3));
    Mem<Sf>::st(cf + CF_ALPHA * W, conv<Sf>(hf.alpha_n));
    Mem<Sf>::st(cf + CF_BETA * W, conv<Sf>(hf.beta_n));
    Mem<Sf>::st(cf + CF_NU * W, conv<Sf>(hf.nu_n));
    Mem<Sf>::st(cf + CF_F1X3 * W, conv<Sf>(hf.f1x3));
    Mem<Sf>::st(cf + CF_F2 * W, conv<Sf>(hf.f2));
    Mem<Sf>::st(cf + CF_CHI * W, conv<Sf>(chi));
    Mem<Sf>::st(cf + CF_GROWTH * W, conv<Sf>(g));

    // number counts: n(z) b H(a), probes.py:87-99 (H = H0 sqrt(E^2))
    for (int t = 0; t < A.nNC; ++t) {
      Sd r = A.nc_nb[(size_t)t * A.Na + m] * (H0 * sqE);
      Mem<Sf>::st(A.R + (((size_t)c * A.Na + m) * A.Tpad + A.nc_col[t]) * W, conv<Sf>(r));
    }
    if (A.dbg) {
      double* d = A.dbg + (size_t)c * A.dbg_stride + 1024 + (size_t)m * 8;
      d[0] = val(chi); d[1] = val(g); d[2] = 1.0 / val(Rnl); d[3] = val(n_eff); d[4] = val(Cc);
      d[5] = val(dchida);
    }
  }

  // chi table for the weak-lensing kernel
  {
    Sd chi_i = s_chi[tid];
    Sd slope = zero<Sd>();
    if (tid < NCHI - 1) slope = (s_chi[tid + 1] - chi_i) / (A.atab_chi[tid + 1] - A.atab_chi[tid]);
    float* ct = A.chitab + ((size_t)c * NCHI + tid) * 2 * W;
    Mem<Sf>::st(ct, conv<Sf>(chi_i));
    Mem<Sf>::st(ct + W, conv<Sf>(slope));
    if (A.dbg) {
      double* d = A.dbg + (size_t)c * A.dbg_stride;
      d[tid] = val(chi_i);
      if (tid < NGROW) d[256 + tid] = val(s_g[tid]);
      d[384 + tid] = val(s_sig[tid]);
      if (tid == 0) d[640] = val(pknorm);
    }
  }
  if (tid == 0) {
    float* cr = A.crec + (size_t)c * NCREC * W;
    Mem<Sf>::st(cr + CR_FB * W, conv<Sf>(eh.fb));
    Mem<Sf>::st(cr + CR_FC * W, conv<Sf>(eh.fc));
    Mem<Sf>::st(cr + CR_INV13KEQ * W, conv<Sf>(eh.inv13keq));
    Mem<Sf>::st(cr + CR_BETAC18 * W, conv<Sf>(eh.beta_c18));
    Mem<Sf>::st(cr + CR_C14A * W, conv<Sf>(eh.c14a));
    Mem<Sf>::st(cr + CR_SHD * W, conv<Sf>(eh.sh_d));
    Mem<Sf>::st(cr + CR_ALPHAB * W, conv<Sf>(eh.alpha_b));
    Mem<Sf>::st(cr + CR_BETAB * W, conv<Sf>(eh.beta_b));
    Mem<Sf>::st(cr + CR_BETANODE * W, conv<Sf>(eh.beta_node));
    Mem<Sf>::st(cr + CR_LNKSILK * W, conv<Sf>(eh.ln_ksilk));
    Mem<Sf>::st(cr + CR_NS * W, conv<Sf>(cp.n_s));
    // probes.py:71: 3 H0^2 Omega_m / (2 c)
    Mem<Sf>::st(cr + CR_WLCONST * W, conv<Sf>(3.0 * H0 * H0 * cp.Omega_m / 2.0 / C_KMS));
    Mem<Sf>::st(cr + CR_PKNORM * W, conv<Sf>(pknorm));
    Mem<Sf>::st(cr + CR_LNINV13 * W, conv<Sf>(eh.ln_inv13keq));
    Mem<Sf>::st(cr + CR_BNODE_SHD * W, conv<Sf>(eh.bnode_shd));
    Mem<Sf>::st(cr + CR_BB_SHD * W, conv<Sf>(eh.bb_shd));
  }
}

size_t tables_smem(int nd) {
  const size_t sd = sizeof(double) * (nd + 1);
  const size_t sf = sizeof(float) * (nd + 1);
  return sd * (NCHI + NGROW + NHF_R + NHF_K + 512) + sf * NHF_K + 16;
}

hipError_t launch_tables(int nd, const TablesArgs& a, hipStream_t s) {
  const size_t sm = tables_smem(nd);
  switch (nd) {
    case 0: hipLaunchKernelGGL(k_tables<0>, dim3(a.B), dim3(256), sm, s, a); break;
    case 1: hipLaunchKernelGGL(k_tables<1>, dim3(a.B), dim3(256), sm, s, a); break;
    case 7: hipLaunchKernelGGL(k_tables<7>, dim3(a.B), dim3(256), sm, s, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// ------------------------------------------------------------------ WL kernel
// Threads own outer nodes m; a block covers 256 consecutive m for CT
// cosmologies, so every resident-table element loaded for (m, j) serves CT
// cosmologies.  The last node (a = 1, z = 0) has chi = 0 and so a zero kernel.
template <int ND, int CT>
__global__ __launch_bounds__(256) void k_wl(WLArgs A) {
  using Sf = typename Types<ND>::F;
  constexpr int W = ND + 1;
  __shared__ float s_tab[CT][NCHI][2][W];
  const int tid = threadIdx.x;
  const int m = blockIdx.x * 256 + tid;
  const int c0 = blockIdx.y * CT;
  const int B = A.dB ? min(*A.dB, A.B) : A.B;
  if (c0 >= B) return;
  for (int ct = 0; ct < CT; ++ct) {
    const int c = c0 + ct < B ? c0 + ct : B - 1;
    const float* src = A.chitab + (size_t)c * NCHI * 2 * W;
    for (int e = tid; e < NCHI * 2 * W; e += 256) (&s_tab[ct][0][0][0])[e] = src[e];
  }
  __syncthreads();
  const bool active = m < A.Na - 1;
  const int mm = active ? m : 0;
  Sf chim[CT];
  Sf acc[CT][NWL_MAX];
#pragma unroll
  for (int ct = 0; ct < CT; ++ct) {
    const int c = c0 + ct < B ? c0 + ct : B - 1;
    chim[ct] = Mem<Sf>::ld(A.coef + (((size_t)c * A.Na + mm) * NCOEF + CF_CHI) * W);
#pragma unroll
    for (int i = 0; i < NWL_MAX; ++i) acc[ct][i] = Make<Sf>::c(0.0);
  }
  if (active) {
    // j = 0 is z' = z, where chi' == chi and the integrand is exactly zero
    for (int j = 1; j < NWL_J; ++j) {
      const size_t o = (size_t)j * A.Na + m;
      const int L = A.Lidx[o];
      const float dl = A.delta[o];
      float nw[NWL_MAX];
#pragma unroll
      for (int i = 0; i < NWL_MAX; ++i)
        nw[i] = (i < A.nw) ? A.Nw[((size_t)i * NWL_J + j) * A.Na + m] : 0.0f;
#pragma unroll
      for (int ct = 0; ct < CT; ++ct) {
        Sf chi_l = Mem<Sf>::ld(&s_tab[ct][L][0][0]);
        Sf sl = Mem<Sf>::ld(&s_tab[ct][L][1][0]);
        Sf chip = maxc(chi_l + sl * dl, 0.0f);            // background.py:242
        Sf num = maxc(chip - chim[ct], 0.0f);             // probes.py:51
        Sf f = num * jm::rcp(maxc(chip, 1.0f));
#pragma unroll
        for (int i = 0; i < NWL_MAX; ++i)
          if (i < A.nw) acc[ct][i] = acc[ct][i] + nw[i] * f;
      }
    }
  }
  if (m < A.Na) {
#pragma unroll
    for (int ct = 0; ct < CT; ++ct) {
      const int c = c0 + ct;
      if (c >= B) break;
      Sf wl = Mem<Sf>::ld(A.crec + ((size_t)c * NCREC + CR_WLCONST) * W);
      Sf pref = (A.zp1[m] * chim[ct]) * wl;
#pragma unroll
      for (int i = 0; i < NWL_MAX; ++i) {
        if (i < A.nw) {
          Sf r = active ? acc[ct][i] * pref : Make<Sf>::c(0.0);
          Mem<Sf>::st(A.R + (((size_t)c * A.Na + m) * A.Tpad + A.col[i]) * W, r);
        }
      }
    }
  }
}

hipError_t launch_wl(int nd, int ctile, const WLArgs& a, hipStream_t s) {
  const int mt = (a.Na + 255) / 256;
  switch (nd * 16 + ctile) {
    case 0 * 16 + 1: hipLaunchKernelGGL((k_wl<0, 1>), dim3(mt, a.B), dim3(256), 0, s, a); break;
    case 0 * 16 + 4: hipLaunchKernelGGL((k_wl<0, 4>), dim3(mt, (a.B + 3) / 4), dim3(256), 0, s, a); break;
    case 0 * 16 + 8: hipLaunchKernelGGL((k_wl<0, 8>), dim3(mt, (a.B + 7) / 8), dim3(256), 0, s, a); break;
    case 1 * 16 + 1: hipLaunchKernelGGL((k_wl<1, 1>), dim3(mt, a.B), dim3(256), 0, s, a); break;
    case 7 * 16 + 1: hipLaunchKernelGGL((k_wl<7, 1>), dim3(mt, a.B), dim3(256), 0, s, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace jc
