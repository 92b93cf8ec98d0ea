// Per-cosmology tables (FP64) and the weak-lensing radial kernels.
//
// k_tables: one 256-thread workgroup per cosmology builds, in FP64,
//   * the chi(a) table: RK4 in ln a of a y-independent right-hand side, i.e.
//     a cumulative Simpson with midpoint nodes (background.py:223-233,
//     scipy/ode.py:6-22), as a parallel prefix sum;
//   * the growth table: RK4 over a of [D, dD/da] (background.py:443-487),
//     the ODE is linear, so each RK4 step is a 2x2 propagator built in
//     parallel and the 127 steps are composed by a parallel prefix product;
//     or, for a Cosmology with gamma, RK4 in ln a of the y-independent
//     dlnD/dlna = Omega_m(a)^gamma (background.py:515-549) as a prefix sum;
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
//     (probes.py:77-99);
//   * for forward calls in table mode, last, the cosmology's transfer table
//     T(k) k^(n_s/2) over the call's ln k range (ehtab_build).
// The EH98 constants are computed once per cosmology by one thread and
// shared through LDS.
// k_wl: the weak-lensing radial kernel (probes.py:16-51) with the
//   reference's own per-z 256-interval Simpson over [z, zmax]: the
//   cosmology-independent parts (n(z') and Simpson weights, the chi-table
//   segment and offset of every z') are resident tables, so per cosmology
//   only the chi lookup and the (chi' - chi)/chi' factor are evaluated.
#include "jc_records.hpp"

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

// FP64 (possibly dual) scalar in memory: component i at p[i]
template <class S>
struct DMem {
  static JC_D S ld(const double* p) { return p[0]; }
  static JC_D void st(double* p, const S& x) { p[0] = x; }
};
template <int N>
struct DMem<Dual<double, N>> {
  static JC_D Dual<double, N> ld(const double* p) {
    Dual<double, N> r;
    r.v = p[0];
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = p[1 + i];
    return r;
  }
  static JC_D void st(double* p, const Dual<double, N>& x) {
    p[0] = x.v;
#pragma unroll
    for (int i = 0; i < N; ++i) p[1 + i] = x.d[i];
  }
};

// interpolate.py:12-37 on a table with log-uniform nodes xp[0..n-1]
// (logspace(-3, 0, n)): nearest node (first on ties) clipped to [1, n-2],
// side from clip(x, xp[1], xp[n-2]), line evaluated as a*x + b.
template <class S>
JC_D S interp_logtab(double x, double log10x, const double* xp, const S* fp, int n) {
  double u = (log10x + 3.0) * (double)(n - 1) / 3.0;
  int i0 = (int)::floor(u);
  i0 = i0 < 0 ? 0 : (i0 > n - 1 ? n - 1 : i0);
  int ind = i0;
  // neighbours (covers any rounding in the log index); scan upward so the
  // first index wins ties
  // (clamped candidates repeat an end node, which never wins over its first
  // occurrence; the four reads are independent)
  double best = 1e300;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    int j = i0 - 1 + k;
    j = j < 0 ? 0 : (j > n - 1 ? n - 1 : j);
    const double dd = (x - xp[j]) * (x - xp[j]);
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

// Growth ODE right-hand side coefficients at x (background.py:465-475):
// y' = [[0, 1], [r, -q]] y, given ln x and 1/x.
template <class Sd>
JC_D void growth_qr(const CosmoP<Sd>& cp, double x, double lnx, double ix, Sd& q, Sd& r) {
  const Sd efde = jm::exp(f_de_ln(cp, x, lnx));
  const Sd E2 = Esqr_parts(cp, ix, efde);
  const Sd iE2 = 1.0 / E2;
  Sd omm = cp.Omega_m * (ix * ix * ix) * iE2;
  Sd omde = cp.Omega_de * efde * iE2;
  q = (2.0 - 0.5 * (omm + (1.0 + 3.0 * w_de(cp, x)) * omde)) * ix;
  r = 1.5 * omm * ix * ix;
}

// 2x2 matrix helpers for the growth propagators (row-major m[0..3])
template <class Sd>
JC_D void mat_mul(const Sd* a, const Sd* b, Sd* c) {  // c = a b
  c[0] = a[0] * b[0] + a[1] * b[2];
  c[1] = a[0] * b[1] + a[1] * b[3];
  c[2] = a[2] * b[0] + a[3] * b[2];
  c[3] = a[2] * b[1] + a[3] * b[3];
}

// Block-wide min / max of one float per thread (scratch: 2 floats per wave
// of LDS); every thread gets the result.
JC_D float block_min(float v, float* scratch) {
  for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o));
  __syncthreads();
  if ((threadIdx.x & 63) == 0) scratch[threadIdx.x >> 6] = v;
  __syncthreads();
  const int nw = blockDim.x >> 6;
  float r = scratch[0];
  for (int w = 1; w < nw; ++w) r = fminf(r, scratch[w]);
  return r;
}
JC_D float block_max(float v, float* scratch) { return -block_min(-v, scratch); }

// Read-only data seen through the constant address space: wave-uniform loads
// from it become scalar loads whatever the kernel stores elsewhere (a global
// store of the same type would otherwise stop the compiler from proving the
// data unchanged, and uniform reads would go through the vector memory path).
typedef const __attribute__((address_space(4))) double ConstD;

// Raw buffer descriptor of `bytes` bytes at p (gfx9 word 3: 32-bit data,
// no swizzle); a load whose offset lies past `bytes` returns zero.
JC_D __amdgpu_buffer_rsrc_t buf_rsrc(const void* p, int bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, bytes, 0x00020000);
}
JC_D float buf_f32(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}

// Cubic Hermite interpolant on [x_q, x_q + h] at fraction t, from the end
// values f0, f1 and the end derivatives d0, d1 (per unit x).
template <class Sd, class Tt>
JC_D Sd hermite(const Tt& t, double h, const Sd& f0, const Sd& f1, const Sd& d0, const Sd& d1) {
  const Sd e0 = h * d0, e1 = h * d1;
  const Sd c2 = 3.0 * (f1 - f0) - 2.0 * e0 - e1;
  const Sd c3 = 2.0 * (f0 - f1) + e0 + e1;
  return ((c3 * t + c2) * t + e0) * t + f0;
}

template <int W>
JC_D void ehtab_build(int c, int Na, int n_ell, const double* ell, int ell_stride, const float* coef, float* crec,
                      float4* ehtab, float* ehtab7, float* s_red, float2* s_nd);

template <int ND, int NT, int MINW>
__global__ __launch_bounds__(NT, MINW) void k_tables(TablesArgs A) {
  using Sd = typename Types<ND>::D;
  using Sf = typename Types<ND>::F;
  constexpr int W = ND + 1;
  extern __shared__ __align__(16) unsigned char smem[];
  Sd* s_chi = reinterpret_cast<Sd*>(smem);  // 256
  Sd* s_g = s_chi + NCHI;                    // 128 (raw D, then normalised)
  Sd* s_sig = s_g + NGROW;                   // 256
  Sd* s_p1 = s_sig + NHF_R;                  // 257
  Sd* s_tmp = s_p1 + NHF_K;                  // 512
  Sd* s_eh = s_tmp + 512;                    // 16: EH98 constants
  Sf* s_p1f = reinterpret_cast<Sf*>(s_eh + 16);  // 257
  float* s_k32 = reinterpret_cast<float*>(s_p1f + NHF_K);  // 257

  const int c = blockIdx.x;
  const int tid = threadIdx.x;
  if (c >= A.B) return;
  if (A.dB && c >= *A.dB) return;  // served batches: grid sized for capacity

  Sd p[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) p[i] = Seed<ND>::load(A.params, A.seeds, c, i);
  const CosmoP<Sd> cp = make_cosmo(p);
  for (int i = tid; i < NHF_K; i += NT) s_k32[i] = A.hf_k32[i];
  // the chi- and growth-table nodes, read by every a-node's interpolation
  const size_t xo = ((size_t)(reinterpret_cast<unsigned char*>(s_k32 + NHF_K) - smem) + 7) & ~(size_t)7;
  double* s_xc = reinterpret_cast<double*>(smem + xo);
  double* s_xg = s_xc + NCHI;
  for (int i = tid; i < NCHI; i += NT) s_xc[i] = A.atab_chi[i];
  for (int i = tid; i < NGROW; i += NT) s_xg[i] = A.atab_g[i];

  // ---------------------------------------------------------- EH98 constants
  // The k-independent EH98 quantities (transfer.py:47-85, 88-97, 109-117,
  // 134-142) once per cosmology, in FP64; the samples read them from LDS and
  // the Limber kernels from the cosmology record.  Forward tables: by wave 0,
  // its independent powers and logarithms on separate lanes (eh_setup_wave;
  // as one thread's chain they took 7.7 us of the block's first phase).
  // Dual tables: by thread 0 (gathering 15 dual powers per lane would spill).
  if (ND == 0 ? tid < 64 : tid == 0) {
    EHc<Sd> eh;
    if constexpr (ND == 0) eh = eh_setup_wave(cp, tid);
    else eh = eh_setup(cp);
    if (tid == 0) {
    const Sd vals[15] = {eh.fb, eh.fc, eh.inv13keq, eh.ln_inv13keq, eh.beta_c18, eh.c14a, eh.sh_d,
                         eh.alpha_b, eh.beta_b, eh.beta_node, eh.ln_ksilk, eh.bnode_shd, eh.bb_shd,
                         eh.alpha_g, eh.tq};
#pragma unroll
    for (int i = 0; i < 15; ++i) s_eh[i] = vals[i];
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
    Mem<Sf>::st(cr + CR_LNINV13 * W, conv<Sf>(eh.ln_inv13keq));
    Mem<Sf>::st(cr + CR_BNODE_SHD * W, conv<Sf>(eh.bnode_shd));
    Mem<Sf>::st(cr + CR_BB_SHD * W, conv<Sf>(eh.bb_shd));
    Mem<Sf>::st(cr + CR_ALPHAG * W, conv<Sf>(eh.alpha_g));
    Mem<Sf>::st(cr + CR_TQ * W, conv<Sf>(eh.tq));
    Mem<Sf>::st(cr + CR_NW * W, Make<Sf>::c(A.transfer_model ? 1.0 : 0.0));
    }
  }

  // ---------------------------------------------------------- chi table
  // dchi/dln a at the 256 nodes (s_sig) and at the RK4 midpoints (s_p1)
  for (int t = tid; t < 2 * NCHI; t += NT) {
    const int i = t & (NCHI - 1);
    const double* nd = A.chi_nodes + (size_t)i * 8;  // a, ln a, 1/a at the node and the midpoint
    if (t < NCHI) {
      const double ai = nd[0];
      const Sd E2 = Esqr_parts(cp, nd[2], jm::exp(f_de_ln(cp, ai, nd[1])));
      s_sig[i] = RH / (ai * ai * jm::sqrt(E2)) * ai;  // dchioverda(e^t) e^t
    } else if (i > 0) {
      const double am = nd[3];
      const Sd E2 = Esqr_parts(cp, nd[5], jm::exp(f_de_ln(cp, am, nd[4])));
      s_p1[i] = RH / (am * am * jm::sqrt(E2)) * am;
    }
  }
  __syncthreads();
  if (tid < NCHI) {
    Sd delta = zero<Sd>();
    if (tid > 0) {
      const double h = A.chi_nodes[(size_t)tid * 8 + 6];
      const Sd fm = s_p1[tid];
      delta = (1.0 / 6.0) * h * (s_sig[tid - 1] + 2.0 * fm + 2.0 * fm + s_sig[tid]);
    }
    s_tmp[tid] = delta;
  }
  __syncthreads();
  for (int off = 1; off < NCHI; off <<= 1) {  // inclusive Hillis-Steele scan
    Sd v = zero<Sd>();
    if (tid < NCHI) {
      v = s_tmp[tid];
      if (tid >= off) v = v + s_tmp[tid - off];
    }
    __syncthreads();
    if (tid < NCHI) s_tmp[tid] = v;
    __syncthreads();
  }
  if (tid < NCHI) s_chi[tid] = s_tmp[NCHI - 1] - s_tmp[tid];
  __syncthreads();

  // ---------------------------------------------------------- growth table
  // RK4 over a of the linear ODE y' = A(a) y (background.py:443-487).  Each
  // step is y_i = G_i y_{i-1} with G_i = I + h/6 (K1 + 2 K2 + 2 K3 + K4),
  // K1 = A(a_{i-1}), K2 = A(mid)(I + h/2 K1), K3 = A(mid)(I + h/2 K2),
  // K4 = A(a_i)(I + h K3): the same products the reference's RK4 forms,
  // regrouped; the prefix products G_i...G_1 come from a parallel scan.
  if (A.growth_model == 0) {
  for (int t = tid; t < 2 * NGROW; t += NT) {
    Sd q, r;
    if (t < NGROW) {
      const double* nd = A.g_nodes + (size_t)t * 8;
      growth_qr(cp, nd[0], nd[1], nd[2], q, r);
      s_tmp[2 * t] = q;
      s_tmp[2 * t + 1] = r;
    } else if (t > NGROW) {
      const int i = t - NGROW;
      const double* nd = A.g_nodes + (size_t)i * 8;
      growth_qr(cp, nd[3], nd[4], nd[5], q, r);
      s_tmp[256 + 2 * i] = q;
      s_tmp[256 + 2 * i + 1] = r;
    }
  }
  __syncthreads();
  Sd G[4];
  if (tid < NGROW) {
    if (tid == 0) {  // ode.py:6-22: the first step has h = 0
      G[0] = Make<Sd>::c(1.0); G[1] = zero<Sd>(); G[2] = zero<Sd>(); G[3] = Make<Sd>::c(1.0);
    } else {
      const int i = tid;
      const double h = A.atab_g[i] - A.atab_g[i - 1];
      const Sd q0 = s_tmp[2 * (i - 1)], r0 = s_tmp[2 * (i - 1) + 1];
      const Sd qm = s_tmp[256 + 2 * i], rm = s_tmp[256 + 2 * i + 1];
      const Sd q1 = s_tmp[2 * i], r1 = s_tmp[2 * i + 1];
      const Sd one = Make<Sd>::c(1.0), zro = zero<Sd>();
      const Sd K1[4] = {zro, one, r0, -q0};
      const Sd Am[4] = {zro, one, rm, -qm};
      const Sd A1[4] = {zro, one, r1, -q1};
      Sd M[4], K2[4], K3[4], K4[4];
      M[0] = 1.0 + h / 2 * K1[0]; M[1] = h / 2 * K1[1]; M[2] = h / 2 * K1[2]; M[3] = 1.0 + h / 2 * K1[3];
      mat_mul(Am, M, K2);
      M[0] = 1.0 + h / 2 * K2[0]; M[1] = h / 2 * K2[1]; M[2] = h / 2 * K2[2]; M[3] = 1.0 + h / 2 * K2[3];
      mat_mul(Am, M, K3);
      M[0] = 1.0 + h * K3[0]; M[1] = h * K3[1]; M[2] = h * K3[2]; M[3] = 1.0 + h * K3[3];
      mat_mul(A1, M, K4);
#pragma unroll
      for (int e = 0; e < 4; ++e)
        G[e] = ((e == 0 || e == 3) ? 1.0 : 0.0) + 1.0 / 6.0 * h * (K1[e] + 2.0 * K2[e] + 2.0 * K3[e] + K4[e]);
    }
  }
  __syncthreads();
  if (tid < NGROW)
    for (int e = 0; e < 4; ++e) s_tmp[4 * tid + e] = G[e];
  __syncthreads();
  for (int off = 1; off < NGROW; off <<= 1) {  // prefix products, later x earlier
    Sd P[4];
    const bool act = tid < NGROW && tid >= off;
    if (act) mat_mul(G, s_tmp + 4 * (tid - off), P);
    __syncthreads();
    if (act) {
      for (int e = 0; e < 4; ++e) {
        G[e] = P[e];
        s_tmp[4 * tid + e] = P[e];
      }
    }
    __syncthreads();
  }
  Sd y1 = zero<Sd>();
  if (tid < NGROW) {
    s_g[tid] = G[0] * A.atab_g[0] + G[1];  // D_i = (P_i y_0)[0], y_0 = [a_0, 1]
    y1 = G[2] * A.atab_g[0] + G[3];        // dD/da
  }
  __syncthreads();
  {
    const Sd g_last = s_g[NGROW - 1];
    Sd gi = zero<Sd>();
    if (tid < NGROW) gi = s_g[tid] / g_last;  // background.py:482
    __syncthreads();
    if (tid < NGROW) s_g[tid] = gi;
    // background calls: the growth-rate table (background.py:483-484)
    if (A.dbg && tid < NGROW) A.dbg[(size_t)c * A.dbg_stride + 768 + tid] = val(y1 / g_last * A.atab_g[tid] / gi);
  }
  } else {
  // gamma growth (background.py:515-549): RK4 in t = ln a of the
  // y-independent f(t) = Omega_m(e^t)^gamma (background.py:551-582) from
  // y_0 = ln a_0 with a first step of h = 0, i.e. a cumulative Simpson with
  // midpoint nodes (a prefix sum); D = exp(y) / exp(y_last).
  for (int t = tid; t < 2 * NGROW; t += NT) {
    if (t == NGROW) continue;
    const int i = t < NGROW ? t : t - NGROW;
    const double* nd = A.gg_nodes + (size_t)i * 8 + (t < NGROW ? 0 : 3);
    const Sd efde = jm::exp(f_de_ln(cp, nd[0], nd[1]));
    const Sd omm = cp.Omega_m * (nd[2] * nd[2] * nd[2]) / Esqr_parts(cp, nd[2], efde);
    s_tmp[t] = jm::exp(A.gamma * jm::log(omm));
  }
  __syncthreads();
  Sd dy = zero<Sd>();
  if (tid > 0 && tid < NGROW) {
    const Sd fm = s_tmp[NGROW + tid];
    dy = (1.0 / 6.0) * A.gg_nodes[(size_t)tid * 8 + 6] * (s_tmp[tid - 1] + 2.0 * fm + 2.0 * fm + s_tmp[tid]);
  }
  if (tid == 0) dy = Make<Sd>::c(A.gg_nodes[1]);
  __syncthreads();
  if (tid < NGROW) s_tmp[tid] = dy;
  __syncthreads();
  for (int off = 1; off < NGROW; off <<= 1) {  // inclusive Hillis-Steele scan
    Sd v = zero<Sd>();
    if (tid < NGROW) {
      v = s_tmp[tid];
      if (tid >= off) v = v + s_tmp[tid - off];
    }
    __syncthreads();
    if (tid < NGROW) s_tmp[tid] = v;
    __syncthreads();
  }
  if (tid < NGROW) s_g[tid] = jm::exp(s_tmp[tid]) / jm::exp(s_tmp[NGROW - 1]);  // background.py:546
  }

  // ---------------------------------------------------------- EH98 samples
  {
    EHc<Sd> eh;
    eh.fb = s_eh[0];
    eh.fc = s_eh[1];
    eh.inv13keq = s_eh[2];
    eh.ln_inv13keq = s_eh[3];
    eh.beta_c18 = s_eh[4];
    eh.c14a = s_eh[5];
    eh.sh_d = s_eh[6];
    eh.alpha_b = s_eh[7];
    eh.beta_b = s_eh[8];
    eh.beta_node = s_eh[9];
    eh.ln_ksilk = s_eh[10];
    eh.bnode_shd = s_eh[11];
    eh.bb_shd = s_eh[12];
    eh.alpha_g = s_eh[13];
    eh.tq = s_eh[14];
    eh.nw = A.transfer_model;
    for (int task = tid; task < NROMB + NHF_K; task += NT) {
      if (task < NROMB) {  // sigma_8 Romberg integrand (power.py:66-73)
        const double lk = -4.0 + task * (7.0 / 128.0);
        const double k = A.romb[task * 3], ik = A.romb[task * 3 + 1];
        Sd T = eh_transfer(eh, k, ik, lk);
        Sd kns = jm::exp(cp.n_s * lk);
        s_tmp[task] = A.romb[task * 3 + 2] * (T * T * kns);  // k (k W)^2 T^2 k^n_s
      } else {
        const int i = task - NROMB;
        const double k = A.hf_kd[i];
        s_p1[i] = eh_transfer(eh, k, 1.0 / k, A.hf_lnk[i]);
      }
    }
  }
  __syncthreads();

  // integrate.py:67-159, divmax = 7, over [log10(1e-4), log10(1e3)]: the
  // sums of the samples each trapezoid level adds (threads 0..7), then the
  // Richardson table on one thread
  if (tid < 8) {
    Sd sl = zero<Sd>();
    if (tid == 0) {
      sl = 0.5 * (s_tmp[0] + s_tmp[128]);
    } else {
      const int stride = 128 >> (tid - 1);
      for (int m = 0; m < (1 << (tid - 1)); ++m) sl = sl + s_tmp[stride / 2 + m * stride];
    }
    s_tmp[160 + tid] = sl;
  }
  __syncthreads();
  if (tid == 0) {
    const double intrange = 7.0;
    Sd ordsum = s_tmp[160];
    Sd state[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) state[k] = intrange * ordsum;
#pragma unroll
    for (int i = 1; i <= 7; ++i) {
      ordsum = ordsum + s_tmp[160 + i];
      Sd x = intrange * ordsum * (1.0 / (double)(1 << i));
      Sd ns[8];
      ns[0] = x;
      double tmp = 1.0;
#pragma unroll
      for (int k = 0; k < 7; ++k) {
        tmp *= 4.0;
        x = (tmp * x - state[k]) * (1.0 / (tmp - 1.0));
        ns[k + 1] = x;
      }
#pragma unroll
      for (int k = 0; k < 8; ++k) state[k] = ns[k];
    }
    Sd sig2 = state[7] * (1.0 / (2.0 * PI * PI));
    s_tmp[256] = cp.sigma8 * cp.sigma8 / sig2;  // pknorm
    Sd g1 = interp_logtab(1.0, 0.0, s_xg, s_g, NGROW);  // the LDS copy of the growth-table nodes
    s_tmp[257] = minc(maxc(g1, 0.0), 1.0);
  }
  __syncthreads();
  const Sd pknorm = s_tmp[256];
  {
    const Sd g1 = s_tmp[257];
    for (int i = tid; i < NHF_K; i += NT) {
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
  // Gaussian-filtered moments at a = 1 (without D^2) on the 256 R nodes:
  //   S_p(R_j) = sum_i w_i P1_i y^{2p} exp(-y^2),  y = k_i R_j,  p = 0..3,
  // summed as R_j^{2p} sum_i (w_i P1_i exp(-y^2)) k_i^{2p}: four FMAs per
  // term with the k_i^{2p} a resident table.
  // S_0 is the halofit sigma^2(R) table (power.py:96-111); S_1 and S_2 are
  // the n_eff and C sums of power.py:118-139 and, with dS_p/dln R =
  // 2p S_p - 2 S_{p+1}, give a cubic Hermite interpolant in ln R that is
  // evaluated at each a-node's R_nl (absolute error ~1e-8 in n_eff and C).
  // With 512 threads the i range is split in two halves.  P1 is not read
  // after this, so S_1..S_3 reuse its space.
  Sd* s_S = s_p1;  // [3][256]: S_1, S_2, S_3
  if constexpr (ND == 0 && NT == 256) {
    // Forward tables: the sums as a small product S[j][p] = sum_i E_ij V_ip
    // with V_ip = w_i P1_i k_i^{2p} staged in LDS.  Wave w takes a quarter of
    // the k nodes, lane l the four R nodes l, l + 64, l + 128, l + 192, so
    // each broadcast read of a V row feeds 16 FMAs and each E row is read
    // once, coalesced; the four waves' partial sums are added in a fixed
    // order through LDS.
    // [257][4], then the partial sums: 32-byte aligned after the float tables
    const size_t vo = ((size_t)(reinterpret_cast<unsigned char*>(s_xg + NGROW) - smem) + 31) & ~(size_t)31;
    double* s_V = reinterpret_cast<double*>(smem + vo);
    ConstD* kp = (ConstD*)A.hf_kp;
    for (int i = tid; i < NHF_K; i += NT) {
      const double v = s_p1[i];
      s_V[4 * i] = v;
      s_V[4 * i + 1] = v * kp[4 * i];
      s_V[4 * i + 2] = v * kp[4 * i + 1];
      s_V[4 * i + 3] = v * kp[4 * i + 2];
    }
    __syncthreads();
    const int lane = tid & 63, wv = tid >> 6;
    const int i0 = wv * 64, i1 = wv == 3 ? NHF_K : i0 + 64;
    double acc[4][4];
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int q = 0; q < 4; ++q) acc[r][q] = 0.0;
    const double* E = A.hf_E + lane;
    const double4* V4 = reinterpret_cast<const double4*>(s_V);
    auto term = [&](const double4& v, const double* e) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        acc[r][0] = fma(v.x, e[r], acc[r][0]);
        acc[r][1] = fma(v.y, e[r], acc[r][1]);
        acc[r][2] = fma(v.z, e[r], acc[r][2]);
        acc[r][3] = fma(v.w, e[r], acc[r][3]);
      }
    };
    int i = i0;
    for (; i + 2 <= i1; i += 2) {  // two k nodes' E rows in flight
      double e0[4], e1[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        e0[r] = E[(size_t)i * NHF_R + 64 * r];
        e1[r] = E[(size_t)(i + 1) * NHF_R + 64 * r];
      }
      const double4 v0 = V4[i], v1 = V4[i + 1];
      term(v0, e0);
      term(v1, e1);
    }
    if (i < i1) {
      double e0[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) e0[r] = E[(size_t)i * NHF_R + 64 * r];
      term(V4[i], e0);
    }
    // ((w0 + w2) + (w1 + w3)): waves 2, 3 hand theirs to 0, 1, then 1 to 0
    double* red = s_V;  // [2][16][64]
    __syncthreads();    // every wave has read V
    if (wv >= 2) {
#pragma unroll
      for (int q = 0; q < 16; ++q) red[((wv - 2) * 16 + q) * 64 + lane] = acc[q >> 2][q & 3];
    }
    __syncthreads();
    if (wv < 2) {
#pragma unroll
      for (int q = 0; q < 16; ++q) acc[q >> 2][q & 3] += red[(wv * 16 + q) * 64 + lane];
    }
    __syncthreads();
    if (wv == 1) {
#pragma unroll
      for (int q = 0; q < 16; ++q) red[q * 64 + lane] = acc[q >> 2][q & 3];
    }
    __syncthreads();
    if (wv == 0) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int j = lane + 64 * r;
        const double Rj = ::exp(A.logr[j]);
        const double R2 = Rj * Rj;
        s_sig[j] = acc[r][0] + red[(4 * r) * 64 + lane];
        s_S[j] = (acc[r][1] + red[(4 * r + 1) * 64 + lane]) * R2;
        s_S[NHF_R + j] = (acc[r][2] + red[(4 * r + 2) * 64 + lane]) * (R2 * R2);
        s_S[2 * NHF_R + j] = (acc[r][3] + red[(4 * r + 3) * 64 + lane]) * (R2 * R2 * R2);
      }
    }
  } else {
    constexpr int NH = NT >= 512 ? 2 : 1;
    const int j = tid & (NHF_R - 1);
    const int half = tid / NHF_R;
    const int i0 = half * ((NHF_K + 1) / 2);
    const int i1 = NH == 1 ? NHF_K : (half ? NHF_K : (NHF_K + 1) / 2);
    Sd m0 = zero<Sd>(), m1 = zero<Sd>(), m2 = zero<Sd>(), m3 = zero<Sd>();
    if (half < NH) {
      const double* E = A.hf_E + j;
      ConstD* kp = (ConstD*)A.hf_kp;  // resident k^{2p} rows: scalar loads
      // eight E rows in flight per step (the loop is bound by L2 latency)
      constexpr int U = 8;
      int i = i0;
      for (; i + U <= i1; i += U) {
        double e[U];
#pragma unroll
        for (int u = 0; u < U; ++u) e[u] = E[(size_t)(i + u) * NHF_R];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const Sd t0 = s_p1[i + u] * e[u];
          m0 = m0 + t0;
          m1 = m1 + t0 * kp[4 * (i + u)];
          m2 = m2 + t0 * kp[4 * (i + u) + 1];
          m3 = m3 + t0 * kp[4 * (i + u) + 2];
        }
      }
      for (; i < i1; ++i) {
        const Sd t0 = s_p1[i] * E[(size_t)i * NHF_R];
        m0 = m0 + t0;
        m1 = m1 + t0 * kp[4 * i];
        m2 = m2 + t0 * kp[4 * i + 1];
        m3 = m3 + t0 * kp[4 * i + 2];
      }
      const double Rj = ::exp(A.logr[j]);
      const double R2 = Rj * Rj;
      m1 = m1 * R2;
      m2 = m2 * (R2 * R2);
      m3 = m3 * (R2 * R2 * R2);
    }
    __syncthreads();
    if (NH == 2 && half == 1) {
      s_sig[j] = m0;
      s_S[j] = m1;
      s_S[NHF_R + j] = m2;
      s_S[2 * NHF_R + j] = m3;
    }
    __syncthreads();
    if (half == 0) {
      if (NH == 2) {
        m0 = m0 + s_sig[j];
        m1 = m1 + s_S[j];
        m2 = m2 + s_S[NHF_R + j];
        m3 = m3 + s_S[2 * NHF_R + j];
      }
      s_sig[j] = m0;
      s_S[j] = m1;
      s_S[NHF_R + j] = m2;
      s_S[2 * NHF_R + j] = m3;
    }
  }
  __syncthreads();

  // ---------------------------------------------------------- per a-node
  const double c2 = C_KMS * C_KMS;
  const float dlnk = (float)(A.hf_lnk[1] - A.hf_lnk[0]);
  const float lnk0 = (float)A.hf_lnk[0];
  for (int m = tid; m < A.Na; m += NT) {
    const double a = A.a_list[m];
    const double l10a = A.a_l10 ? A.a_l10[m] : ::log10(a);
    const double lna = A.a_ln ? A.a_ln[m] : ::log(a);
    const double ia = 1.0 / a;
    Sd chi = maxc(interp_logtab(a, l10a, s_xc, s_chi, NCHI), 0.0);  // background.py:242
    Sd g = interp_logtab(a, l10a, s_xg, s_g, NGROW);
    g = minc(maxc(g, 0.0), 1.0);                                           // background.py:488
    const Sd efde = jm::exp(f_de_ln(cp, a, lna));
    Sd E2 = Esqr_parts(cp, ia, efde);
    Sd sqE = jm::sqrt(E2);
    Sd dchida = RH / (a * a * sqE);                                        // background.py:294
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
    Sd Rnl = maxc(jm::exp(root), 1e-6);                                   // power.py:113-115
    // n_eff and C (power.py:118-139): n_eff + 3 = 2 g^2 S1(R_nl),
    // C = (n_eff+3)^2 + 4 g^2 (S1 - S2), S1/S2 by the Hermite interpolant of
    // the moment tables.  Outside the R grid (R_nl < 1e-4 or > 10, only for
    // scale factors well before a = 1/11) the sums are taken directly: float
    // terms, float partial sums of 16 terms, double totals, skipping y < 1e-3
    // (relative size < 1e-12) and y^2 > 40.
    Sd S1, S2;
    const double hR = A.logr[1] - A.logr[0];
    const Sd uR = (root - A.logr[0]) * (1.0 / hR);
    const int qR = (int)::floor(val(uR));
    if (qR >= 0 && qR <= NHF_R - 2) {
      const Sd tR = uR - (double)qR;
      const Sd* S1t = s_S;
      const Sd* S2t = s_S + NHF_R;
      const Sd* S3t = s_S + 2 * NHF_R;
      S1 = hermite(tR, hR, S1t[qR], S1t[qR + 1], 2.0 * (S1t[qR] - S2t[qR]), 2.0 * (S1t[qR + 1] - S2t[qR + 1]));
      S2 = hermite(tR, hR, S2t[qR], S2t[qR + 1], 4.0 * S2t[qR] - 2.0 * S3t[qR],
                   4.0 * S2t[qR + 1] - 2.0 * S3t[qR + 1]);
    } else {
      Sf Rf = conv<Sf>(Rnl);
      int ibeg = (int)((logf(1e-3f / val(Rf)) - lnk0) / dlnk);
      ibeg = ibeg < 0 ? 0 : (ibeg > NHF_K ? NHF_K : ibeg);
      S1 = zero<Sd>();
      S2 = zero<Sd>();
      for (int i0 = ibeg; i0 < NHF_K; i0 += 16) {
        Sf p1 = Make<Sf>::c(0.0), p2 = Make<Sf>::c(0.0);
        const int iend = i0 + 16 < NHF_K ? i0 + 16 : NHF_K;
        bool stop = false;
        for (int i = i0; i < iend; ++i) {
          Sf y = s_k32[i] * Rf;
          Sf y2 = y * y;
          if (val(y2) > 40.0f) { stop = true; break; }
          Sf ty2 = s_p1f[i] * jm::exp(-y2) * y2;
          p1 += ty2;
          p2 += ty2 * y2;
        }
        S1 = S1 + conv<Sd>(p1);
        S2 = S2 + conv<Sd>(p2);
        if (stop) break;
      }
    }
    Sd n3 = 2.0 * g2 * S1;
    Sd n_eff = n3 - 3.0;
    Sd Cc = n3 * n3 + 4.0 * g2 * (S1 - S2);
    Sd omm = cp.Omega_m * (ia * ia * ia) / E2;
    Sd omde = cp.Omega_de * efde / E2;
    HFa<Sd> hf = halofit_coefs(n_eff, Cc, omm, omde, w_de(cp, a), A.halofit_model);
    Sd wgt = A.a_wgt[m] * dchida / maxc(chi * chi, 1.0) / c2;             // angular_cl.py:91,96
    Sd chi1 = maxc(chi, 1.0);                                             // angular_cl.py:73
    Sd invchi = 1.0 / chi1;

    float* cf = A.coef + ((size_t)c * A.Na + m) * NCOEF * W;
    Mem<Sf>::st(cf + CF_INV_CHI * W, conv<Sf>(invchi));
    Mem<Sf>::st(cf + CF_CHI1 * W, conv<Sf>(chi1));
    // the logarithms stored as floats are taken to float accuracy
    Mem<Sf>::st(cf + CF_LNINVCHI * W, conv<Sf>(-jm::logf_acc(chi1)));
    Mem<Sf>::st(cf + CF_KNL * W, conv<Sf>(1.0 / Rnl));
    Mem<Sf>::st(cf + CF_MU * W, conv<Sf>(hf.mu_n));
    Mem<Sf>::st(cf + CF_WGT * W, conv<Sf>(wgt));
    Mem<Sf>::st(cf + CF_G2P * W, conv<Sf>(g2 * pknorm));
    Mem<Sf>::st(cf + CF_LGRNL * W, conv<Sf>(jm::logf_acc(Rnl) * 1.4426950408889634));
    Mem<Sf>::st(cf + CF_RNL * W, conv<Sf>(Rnl));
    Mem<Sf>::st(cf + CF_AN * W, conv<Sf>(hf.a_n));
    Mem<Sf>::st(cf + CF_BN * W, conv<Sf>(hf.b_n));
    Mem<Sf>::st(cf + CF_LGCF3 * W, conv<Sf>(hf.lg_cf3));
    Mem<Sf>::st(cf + CF_GAM3 * W, conv<Sf>(hf.gam3));
    Mem<Sf>::st(cf + CF_ALPHA * W, conv<Sf>(hf.alpha_n));
    Mem<Sf>::st(cf + CF_BETA * W, conv<Sf>(hf.beta_n));
    Mem<Sf>::st(cf + CF_NU * W, conv<Sf>(hf.nu_n));
    Mem<Sf>::st(cf + CF_F1X3 * W, conv<Sf>(hf.f1x3));
    Mem<Sf>::st(cf + CF_F2 * W, conv<Sf>(hf.f2));
    Mem<Sf>::st(cf + CF_CHI * W, conv<Sf>(chi));
    Mem<Sf>::st(cf + CF_GROWTH * W, conv<Sf>(g));

    // z-local kernels, H = H0 sqrt(E^2) (background.py:129-148): number
    // counts n(z) b H (probes.py:87-99) and the NLA terms (probes.py:114-124)
    for (int t = 0; t < A.nLoc; ++t) {
      Sd r = A.loc_nb[(size_t)t * A.Na + m] * (H0 * sqE);
      const int mode = A.loc_mode[t];
      if (mode & LOC_INVD) r = r / g;             // bias.py:38-40: b / D(a)
      if (mode & LOC_IA) r = r * cp.Omega_m / g;  // probes.py:120-122
      // R columns: the pool column, or (served batches) each position of
      // the request that selects the tracer
      const int lc = A.loc_col[t];
      uint32_t cmask = 1u << lc;
      if (A.rjobs) {
        const RJob& jb = A.rjobs[c];
        const int T = jb.T;
        cmask = 0;
#pragma unroll
        for (int i = 0; i < MAXT; ++i) cmask |= (i < T && jb.sel[i] == lc ? 1u : 0u) << i;
      }
      const size_t rb = ((size_t)c * A.Na + m) * A.Tpad;
      for (uint32_t mk = cmask; mk; mk &= mk - 1) Mem<Sf>::st(A.R + (rb + __builtin_ctz(mk)) * W, conv<Sf>(r));
    }
    if (A.dbg) {
      double* d = A.dbg + (size_t)c * A.dbg_stride + 1024 + (size_t)m * 8;
      d[0] = val(chi); d[1] = val(g); d[2] = 1.0 / val(Rnl); d[3] = val(n_eff); d[4] = val(Cc);
      d[5] = val(dchida);
    }
  }

  // chi table for the weak-lensing kernel
  if (tid < NCHI) {
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
  if (tid == 0) Mem<Sf>::st(A.crec + ((size_t)c * NCREC + CR_PKNORM) * W, conv<Sf>(pknorm));
  // forward calls with the table-mode Limber kernel: the cosmology's transfer
  // table, from its coefficient records just written (the LDS is free now)
  if constexpr (ND == 0 || ND == 7) {
    if (A.ehtab) {
      __syncthreads();
      // served batches: each cosmology's own multipoles (its request's columns)
      const double* ell = A.ell;
      int n_ell = A.n_ell, stride = 1;
      if (A.rjobs) {
        const RJob& jb = A.rjobs[c];
        ell = &A.rcols[jb.col0].ell;
        n_ell = jb.n_ell;
        stride = (int)(sizeof(RCol) / sizeof(double));
      }
      ehtab_build<W>(c, A.Na, n_ell, ell, stride, A.coef, A.crec, A.ehtab, ND == 7 ? A.ehtab7 : nullptr,
                     reinterpret_cast<float*>(smem), reinterpret_cast<float2*>(smem + 64));
    }
  }
}

// Growth rate and transverse comoving distance at the n scale factors of a
// background call (slots 6 and 7 of each per-a block of dbg), from the table
// kernel's FP64 output for the same call: chi in slot 0 and, for the ODE
// model, the growth-rate table in slots 768..895 of the table block.
__global__ __launch_bounds__(256) void k_background_ex(const double* params, const double* a, int n,
                                                       const double* atab_g, int growth_model, double gamma,
                                                       double* dbg) {
  const int m = blockIdx.x * 256 + threadIdx.x;
  if (m >= n) return;
  double p[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) p[i] = params[i];
  const CosmoP<double> cp = make_cosmo(p);
  double* d = dbg + 1024 + (size_t)m * 8;
  const double x = a[m];
  if (growth_model == 0) {
    d[6] = interp_logtab(x, ::log10(x), atab_g, dbg + 768, NGROW);  // background.py:490-512
  } else {  // background.py:551-582: Omega_m(a)^gamma at a
    const double ix = 1.0 / x;
    const double omm = cp.Omega_m * (ix * ix * ix) / Esqr_parts(cp, ix, ::exp(f_de_ln(cp, x, ::log(x))));
    d[6] = ::exp(gamma * ::log(omm));
  }
  // background.py:297-344: the branch of k = -sign(Omega_k)
  const double ok = cp.Omega_k, sk = ::sqrt(::fabs(ok)), chi = d[0];
  d[7] = ok > 0.0 ? RH / sk * ::sinh(sk * chi / RH) : (ok < 0.0 ? RH / sk * ::sin(sk * chi / RH) : chi);
}

hipError_t launch_background_ex(const double* params, const double* a, int n, const double* atab_g,
                                int growth_model, double gamma, double* dbg, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_background_ex, dim3((n + 255) / 256), dim3(256), 0, s, params, a, n, atab_g, growth_model,
                     gamma, dbg);
  return hipGetLastError();
}

size_t tables_smem(int nd, int threads) {
  const size_t sd = sizeof(double) * (nd + 1);
  const size_t sf = sizeof(float) * (nd + 1);
  size_t b = sd * (NCHI + NGROW + NHF_R + NHF_K + 512 + 16) + sf * NHF_K + sizeof(float) * NHF_K;
  b = (b + 7) / 8 * 8 + sizeof(double) * (NCHI + NGROW);  // chi- and growth-table nodes
  // forward tables at 256 threads: the moment product's V rows and partial sums
  if (nd == 0 && threads == 256) b = (b + 31) / 32 * 32 + sizeof(double) * 2 * 16 * 64;
  return b + 16;
}

// variant (forward tables only): 1 = 256 threads at 4 waves/SIMD (register
// spills), 2 = 256 threads at 2 waves/SIMD, 3 = 512 threads at 2 waves/SIMD
hipError_t launch_tables(int nd, const TablesArgs& a, hipStream_t s, int variant) {
  const size_t sm = tables_smem(nd, nd == 0 && variant == 3 ? 512 : 256);
  if (nd == 0) {
    switch (variant) {
      case 1: hipLaunchKernelGGL((k_tables<0, 256, 4>), dim3(a.B), dim3(256), sm, s, a); break;
      case 3: hipLaunchKernelGGL((k_tables<0, 512, 2>), dim3(a.B), dim3(512), sm, s, a); break;
      default: hipLaunchKernelGGL((k_tables<0, 256, 2>), dim3(a.B), dim3(256), sm, s, a); break;
    }
  } else if (nd == 1) {
    hipLaunchKernelGGL((k_tables<1, 256, 1>), dim3(a.B), dim3(256), sm, s, a);
  } else if (nd == 7) {
    hipLaunchKernelGGL((k_tables<7, 256, 1>), dim3(a.B), dim3(256), sm, s, a);
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// ------------------------------------------------------------------ transfer table
// Cosmology c's transfer table, built by one whole workgroup of the table
// kernel once the coefficient records are written (s_red: 8 floats, s_nd:
// NEHQ + 1 float2 of LDS).  The table spans the ln k range of the call's
// Limber points, ln k = ln(ell + 1/2) - ln max(chi, 1), formed in float as
// the Limber kernel forms it; the nodes are x_q = lo + q / invh with invh a
// float rounded down, so NEHQ intervals cover [lo, hi] and the kernel's FP64
// table position is exact.  Node values and ln k derivatives of
// T(k) k^(n_s/2) (transfer.py:119-149, power.py:14-18) come from the same
// EH98 code as the closed form, in float with forward-mode dual numbers; each
// value is moved from the float-rounded node to the exact node with its
// derivative.  W is the record width (1, or 8 in the adjoint's table kernel,
// whose records hold 7 theta derivatives after each value); there, ehtab7
// also receives the theta derivatives of the same function (float duals at
// fixed k) at the same NEHQ + 1 nodes, for 4-point Lagrange interpolation in
// the adjoint kernel.
template <int W>
JC_D void ehtab_build(int c, int Na, int n_ell, const double* ell, int ell_stride, const float* coef, float* crec,
                      float4* ehtab, float* ehtab7, float* s_red, float2* s_nd) {
  const int tid = threadIdx.x;
  float ll = 3.0e38f, lh = -3.0e38f, cl = 3.0e38f, ch = -3.0e38f;
  for (int l = tid; l < n_ell; l += blockDim.x) {
    const float v = (float)::log(ell[(size_t)l * ell_stride] + 0.5);
    ll = fminf(ll, v);
    lh = fmaxf(lh, v);
  }
  for (int m = tid; m < Na; m += blockDim.x) {
    const float v = coef[(((size_t)c * Na + m) * NCOEF + CF_LNINVCHI) * W];
    cl = fminf(cl, v);
    ch = fmaxf(ch, v);
  }
  const float lo = block_min(ll, s_red) + block_min(cl, s_red);
  const float hi = block_max(lh, s_red) + block_max(ch, s_red);
  const double span = fmax((double)hi - (double)lo, 1e-3);
  float invh = (float)(NEHQ / span);
  if ((double)invh * span > (double)NEHQ) invh = nextafterf(invh, 0.0f);
  const double h = 1.0 / (double)invh;

  using D1 = Dual<float, 1>;
  float ns;
  const EHc<D1> e1 = eh_cast<D1>(load_eh<float, W>(crec + (size_t)c * NCREC * W, ns));
  const float hns = 0.5f * ns;
  auto node = [&](int q, float& f, float& d) {
    const double xn = (double)lo + (double)q * h;
    D1 x;
    x.v = (float)xn;
    x.d[0] = 1.0f;
    D1 k, ik;
    k.v = jm::exp(x.v);
    k.d[0] = k.v;
    ik.v = jm::exp(-x.v);
    ik.d[0] = -ik.v;
    const D1 tn = eh_transfer(e1, k, ik, x) * jm::exp(hns * x);
    d = tn.d[0];
    f = tn.v + tn.d[0] * (float)(xn - (double)x.v);
  };
  for (int q = tid; q <= NEHQ; q += blockDim.x) {
    float f, d;
    node(q, f, d);
    s_nd[q] = make_float2(f, d);
  }
  __syncthreads();
  const float hf = (float)h;
  float4* dst = ehtab + (size_t)c * NEHQ;
  for (int q = tid; q < NEHQ; q += blockDim.x) {
    const float2 n0 = s_nd[q], n1 = s_nd[q + 1];
    const float e0 = hf * n0.y, e1d = hf * n1.y;
    dst[q] = make_float4(n0.x, e0, 3.0f * (n1.x - n0.x) - 2.0f * e0 - e1d, 2.0f * (n0.x - n1.x) + e0 + e1d);
  }
  if (tid == 0) {
    crec[((size_t)c * NCREC + CR_EHLO) * W] = lo;
    crec[((size_t)c * NCREC + CR_EHINVH) * W] = invh;
  }
  if constexpr (W == 8) {
    using D7 = Dual<float, 7>;
    D7 ns7;
    const EHc<D7> e7 = load_eh<D7, 8>(crec + (size_t)c * NCREC * 8, ns7);
    for (int q = tid; q <= NEHQ; q += blockDim.x) {
      const float x = (float)((double)lo + (double)q * h);
      const float k = jm::exp(x), ik = jm::exp(-x);
      const D7 tn = eh_transfer(e7, k, ik, x) * jm::exp(ns7 * (0.5f * x));
      float* o = ehtab7 + ((size_t)c * (NEHQ + 1) + q) * NEHD_ROW;
#pragma unroll
      for (int d = 0; d < 7; ++d) o[d] = tn.d[d];
      o[7] = 0.0f;
    }
  }
}

// ------------------------------------------------------------------ KDE n(z)
// One thread per evaluation point.  The catalog is sorted by redshift, so a
// point only visits galaxies within 8.5 bandwidths (the rest weigh below
// exp(-36) ~ 2e-16 of the nearest ones); exponentials in float, sums in double.
__global__ void k_kde(const double* __restrict__ z, int64_t n, const double* __restrict__ cz,
                      const double* __restrict__ cw, int64_t ncat, double bw, double scale,
                      double* __restrict__ out) {
  const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p >= n) return;
  const double zp = z[p];
  const double lo = zp - 8.5 * bw;
  int64_t a = 0, b = ncat;  // first galaxy with cz >= lo
  while (a < b) {
    const int64_t mid = (a + b) >> 1;
    if (cz[mid] < lo) a = mid + 1; else b = mid;
  }
  const double hi = zp + 8.5 * bw;
  const float c2 = (float)(-0.5 / (bw * bw));
  double s = 0.0;
  for (int64_t k = a; k < ncat && cz[k] <= hi; ++k) {
    const float d = (float)(zp - cz[k]);
    s += cw[k] * (double)jm::exp(c2 * d * d);
  }
  out[p] = s * scale;
}

hipError_t launch_kde(const double* z, int64_t n, const double* cz, const double* cw, int64_t ncat, double bw,
                      double wsum, double* out, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const double scale = 1.0 / (sqrt(2.0 * PI) * bw) / wsum;
  hipLaunchKernelGGL(k_kde, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, z, n, cz, cw, ncat, bw, scale, out);
  return hipGetLastError();
}

// ------------------------------------------------------------------ WL kernel
// A block covers 64 consecutive outer nodes m (one per lane) for CT
// cosmologies; its 4 waves split the 256 inner nodes j of the per-z Simpson
// into contiguous quarters and their partial sums are added in LDS in a fixed
// order (deterministic).  Every resident-table element loaded for (m, j)
// serves CT cosmologies; NW = weak-lensing bins held in registers (>= the
// set's count).  The last node (a = 1, z = 0) has chi = 0 and so a zero
// kernel; the lane of the node before it writes that zero.
constexpr int WL_M = 64;      // outer nodes per block
constexpr int WL_SLICES = 4;  // waves per block, one quarter of j each

template <int ND, int CT, int NW>
__global__ __launch_bounds__(WL_M * WL_SLICES) void k_wl(WLArgs A) {
  using Sf = typename Types<ND>::F;
  constexpr int W = ND + 1;
  extern __shared__ __align__(16) float wl_smem[];  // [CT][NCHI][2][W] tables, then [3][NACC][64] partial sums
  float* s_tab = wl_smem;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int m = blockIdx.x * WL_M + lane;
  const int c0 = blockIdx.y * CT;
  const int B = A.dB ? min(*A.dB, A.B) : A.B;
  if (c0 >= B) return;
  for (int ct = 0; ct < CT; ++ct) {
    const int c = c0 + ct < B ? c0 + ct : B - 1;
    const float* src = A.chitab + (size_t)c * NCHI * 2 * W;
    for (int e = tid; e < NCHI * 2 * W; e += WL_M * WL_SLICES) s_tab[ct * NCHI * 2 * W + e] = src[e];
  }
  __syncthreads();
  const int N = A.Na - 1;  // nodes with a nonzero kernel
  const bool active = m < N;
  const int mm = active ? m : 0;
  Sf chim[CT];
  Sf acc[CT][NW];
#pragma unroll
  for (int ct = 0; ct < CT; ++ct) {
    const int c = c0 + ct < B ? c0 + ct : B - 1;
    chim[ct] = Mem<Sf>::ld(A.coef + (((size_t)c * A.Na + mm) * NCOEF + CF_CHI) * W);
#pragma unroll
    for (int i = 0; i < NW; ++i) acc[ct][i] = Make<Sf>::c(0.0);
  }
  if (active) {
    // j = 0 is z' = z, where chi' == chi and the integrand is exactly zero
    constexpr int JS = (NWL_J - 1) / WL_SLICES;
    const int j0 = 1 + wave * JS;
    // raw buffer loads: the table bases and the bin row offsets stay in
    // SGPRs, so one 32-bit byte offset per lane advances per j (no 64-bit
    // address arithmetic in the loop); loads past a table's end read zero
    const int rs = NWL_J * A.Na;  // elements per table row block (and per bin of the weights)
    const __amdgpu_buffer_rsrc_t rL = buf_rsrc(A.Lidx, rs);
    const __amdgpu_buffer_rsrc_t rD = buf_rsrc(A.delta, rs * 4);
    const __amdgpu_buffer_rsrc_t rN = buf_rsrc(A.Nw, rs * 4 * NW);
    int oL = j0 * A.Na + m;
    int o4 = 4 * oL;
    for (int j = j0; j < j0 + JS; ++j, oL += A.Na, o4 += 4 * A.Na) {
      const int L = __builtin_amdgcn_raw_buffer_load_b8(rL, oL, 0, 0);
      const float dl = buf_f32(rD, o4, 0);
      float nw[NW];
#pragma unroll
      for (int i = 0; i < NW; ++i) nw[i] = buf_f32(rN, o4, i * rs * 4);  // zero rows past nw
      // cosmology-major tables: lanes (consecutive m, close segments L) read
      // nearby words of one table
      const float* tL = s_tab + L * 2 * W;
#pragma unroll
      for (int ct = 0; ct < CT; ++ct) {
        const float* t = tL + ct * NCHI * 2 * W;
        Sf chi_l, sl;
        if constexpr (W == 1) {  // one 8-byte read at a constant offset per cosmology
          const float2 v = *reinterpret_cast<const float2*>(t);
          chi_l = v.x;
          sl = v.y;
        } else {
          chi_l = Mem<Sf>::ld(t);
          sl = Mem<Sf>::ld(t + W);
        }
        // chi' = clip(interp, 0) (background.py:242): with chi >= 0 the clip
        // cannot change clip(chi' - chi, 0) / clip(chi', 1) (probes.py:51),
        // so it is left out
        Sf chip = chi_l + sl * dl;
        Sf num = maxc(chip - chim[ct], 0.0f);
        Sf f = num * jm::rcp(maxc(chip, 1.0f));
        // bins past the set's count carry nw = 0 and f is finite: no guard
#pragma unroll
        for (int i = 0; i < NW; ++i) acc[ct][i] = acc[ct][i] + nw[i] * f;
      }
    }
    // single source planes (probes.py:53-64): clip(chi_s - chi, 0) / clip(chi_s, 1)
    if (wave == 0) {
#pragma unroll
      for (int i = 0; i < NW; ++i) {
        if (i < A.nw) {
          const WLSource sp = A.src[i];
          if (sp.on) {
#pragma unroll
            for (int ct = 0; ct < CT; ++ct) {
              const float* t = s_tab + ((ct * NCHI + sp.L) * 2) * W;
              Sf chis = maxc(Mem<Sf>::ld(t) + Mem<Sf>::ld(t + W) * sp.off, 0.0f);
              Sf f = maxc(chis - chim[ct], 0.0f) * jm::rcp(maxc(chis, 1.0f));
              acc[ct][i] = acc[ct][i] + sp.fac * f;
            }
          }
        }
      }
    }
  }
  __syncthreads();  // the tables are no longer read: reuse the space
  float* red = wl_smem;
  if (wave > 0) {
#pragma unroll
    for (int ct = 0; ct < CT; ++ct)
#pragma unroll
      for (int i = 0; i < NW; ++i)
        Mem<Sf>::st(&red[(((wave - 1) * CT + ct) * NW + i) * W * 64 + lane], acc[ct][i], 64);
  }
  __syncthreads();
  if (wave != 0 || m >= N) return;
  // R columns of each bin and cosmology, taken before any store: the pool
  // column, or (served batches) every position of the request that
  // selects the bin; all MAXT positions are tested so that the loads issue
  // together
  uint32_t cm[CT][NW];
#pragma unroll
  for (int ct = 0; ct < CT; ++ct) {
    const int c = c0 + ct < B ? c0 + ct : B - 1;
    int T = 0;
    int sel[MAXT];
    if (A.rjobs) {
      T = A.rjobs[c].T;
#pragma unroll
      for (int p = 0; p < MAXT; ++p) sel[p] = A.rjobs[c].sel[p];
    }
#pragma unroll
    for (int i = 0; i < NW; ++i) {
      uint32_t mk = 0;
      if (i < A.nw) {
        const int cl = A.col[i];
        if (A.rjobs) {
#pragma unroll
          for (int p = 0; p < MAXT; ++p) mk |= (p < T && sel[p] == cl ? 1u : 0u) << p;
        } else {
          mk = 1u << cl;
        }
      }
      cm[ct][i] = mk;
    }
  }
#pragma unroll
  for (int ct = 0; ct < CT; ++ct) {
    const int c = c0 + ct;
    if (c >= B) break;
    Sf wl = Mem<Sf>::ld(A.crec + ((size_t)c * NCREC + CR_WLCONST) * W);
    Sf pref = (A.zp1[m] * chim[ct]) * wl;
#pragma unroll
    for (int i = 0; i < NW; ++i) {
      if (i < A.nw) {
        Sf r = acc[ct][i];
        for (int w = 1; w < WL_SLICES; ++w)
          r = r + Mem<Sf>::ld(&red[(((w - 1) * CT + ct) * NW + i) * W * 64 + lane], 64);
        r = r * pref;
        for (uint32_t mk = cm[ct][i]; mk; mk &= mk - 1) {
          const int cl = __builtin_ctz(mk);
          float* dst = A.R + (((size_t)c * A.Na + m) * A.Tpad + cl) * W;
          Sf v = r;
          if (A.ia) v = v + Mem<Sf>::ld(dst);  // the NLA term written by the table kernel
          Mem<Sf>::st(dst, v);
          // the a = 1 node: zero lensing kernel (chi = 0), the NLA term stays
          if (m == N - 1 && !A.ia) Mem<Sf>::st(A.R + (((size_t)c * A.Na + N) * A.Tpad + cl) * W, Make<Sf>::c(0.0));
        }
      }
    }
  }
}

template <int ND, int CT, int NW>
static void wl_launch_nw(const WLArgs& a, hipStream_t s) {
  constexpr int W = ND + 1;
  const size_t tab = (size_t)CT * NCHI * 2 * W;
  const size_t red = (size_t)(WL_SLICES - 1) * CT * NW * W * 64;
  const size_t sm = (tab > red ? tab : red) * sizeof(float);
  const dim3 grid((a.Na - 1 + WL_M - 1) / WL_M, (a.B + CT - 1) / CT);
  hipLaunchKernelGGL((k_wl<ND, CT, NW>), grid, dim3(WL_M * WL_SLICES), sm, s, a);
}

template <int ND, int CT>
static void wl_launch_ct(const WLArgs& a, hipStream_t s) {
  if (a.nw <= 5) wl_launch_nw<ND, CT, 5>(a, s);
  else wl_launch_nw<ND, CT, NWL_MAX>(a, s);
}

hipError_t launch_wl(int nd, int ctile, const WLArgs& a, hipStream_t s) {
  if (a.nw < 1 || a.nw > NWL_MAX) return hipErrorInvalidValue;
  switch (nd * 32 + ctile) {
    case 0 * 32 + 1: wl_launch_ct<0, 1>(a, s); break;
    case 0 * 32 + 4: wl_launch_ct<0, 4>(a, s); break;
    case 0 * 32 + 8: wl_launch_ct<0, 8>(a, s); break;
    case 0 * 32 + 16: wl_launch_ct<0, 16>(a, s); break;
    case 1 * 32 + 1: wl_launch_ct<1, 1>(a, s); break;
    case 7 * 32 + 1: wl_launch_ct<7, 1>(a, s); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace jc
