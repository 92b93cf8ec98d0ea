// The fused Limber kernel: for every (cosmology, ell) column and every a-node
// it evaluates the Eisenstein-Hu transfer function and the Takahashi halofit
// at k = (ell + 1/2) / max(chi, 1) (transfer.py:10-156, power.py:21-53,
// 144-262), weights it with the Simpson weight and dchi/da / max(chi^2, 1)
// / c^2, and contracts it with the tracer kernel products K_i K_j
// (angular_cl.py:66-98):
//
//   C_ij(ell) = F_i(ell) F_j(ell) sum_a R_i(a) R_j(a) M(a, ell),
//
// where R are the radial kernels and F the weak-lensing ell factor
// (probes.py:73).  One lane owns one column; all lanes of a wave walk the
// same a-nodes, so the per-a records are wave-uniform loads and the
// contraction is P register FMAs per point.  The a range is split over the
// waves of a block and reduced in LDS in a fixed order (deterministic).
//
// Modes (ND = number of dual directions):
//   ND = 0  forward C_ell;
//   ND = 1  forward plus one directional derivative (JVP);
//   ND = 7  adjoint: contracts dC/dtheta with a cotangent Cbar on the fly
//           (theta = Omega_c, Omega_b, h, n_s, sigma8, w0, wa) using
//           sum_{i<=j} c_ij d(R_i R_j) = sum_i dR_i V_i, V_i = sum_j ct_ij R_j.
#include "jc_kernels.hpp"
#include "jc_physics.hpp"

namespace jc {

template <class Sf, int W>
struct PointCoefs {
  Sf invchi, wgt, g2p, lnRnl, Rnl;
  HFa<Sf> hf;
};

template <class Sf, int W>
JC_D PointCoefs<Sf, W> load_coefs(const float* cf) {
  PointCoefs<Sf, W> r;
  r.invchi = Mem<Sf>::ld(cf + CF_INV_CHI * W);
  r.wgt = Mem<Sf>::ld(cf + CF_WGT * W);
  r.g2p = Mem<Sf>::ld(cf + CF_G2P * W);
  r.lnRnl = Mem<Sf>::ld(cf + CF_LNRNL * W);
  r.Rnl = Mem<Sf>::ld(cf + CF_RNL * W);
  r.hf.a_n = Mem<Sf>::ld(cf + CF_AN * W);
  r.hf.b_n = Mem<Sf>::ld(cf + CF_BN * W);
  r.hf.ln_cf3 = Mem<Sf>::ld(cf + CF_LNCF3 * W);
  r.hf.gam3 = Mem<Sf>::ld(cf + CF_GAM3 * W);
  r.hf.alpha_n = Mem<Sf>::ld(cf + CF_ALPHA * W);
  r.hf.beta_n = Mem<Sf>::ld(cf + CF_BETA * W);
  r.hf.nu_n = Mem<Sf>::ld(cf + CF_NU * W);
  r.hf.f1x3 = Mem<Sf>::ld(cf + CF_F1X3 * W);
  r.hf.f2 = Mem<Sf>::ld(cf + CF_F2 * W);
  return r;
}

template <class Sf, int W>
JC_D EHc<Sf> load_eh(const float* cr, Sf& ns) {
  EHc<Sf> e;
  e.fb = Mem<Sf>::ld(cr + CR_FB * W);
  e.fc = Mem<Sf>::ld(cr + CR_FC * W);
  e.inv13keq = Mem<Sf>::ld(cr + CR_INV13KEQ * W);
  e.ln_inv13keq = Mem<Sf>::ld(cr + CR_LNINV13 * W);
  e.beta_c18 = Mem<Sf>::ld(cr + CR_BETAC18 * W);
  e.c14a = Mem<Sf>::ld(cr + CR_C14A * W);
  e.sh_d = Mem<Sf>::ld(cr + CR_SHD * W);
  e.alpha_b = Mem<Sf>::ld(cr + CR_ALPHAB * W);
  e.beta_b = Mem<Sf>::ld(cr + CR_BETAB * W);
  e.beta_node = Mem<Sf>::ld(cr + CR_BETANODE * W);
  e.ln_ksilk = Mem<Sf>::ld(cr + CR_LNKSILK * W);
  ns = Mem<Sf>::ld(cr + CR_NS * W);
  return e;
}

// M(a, ell) = weight * P_nl(k, a) at k = (ell + 1/2) / max(chi, 1)
template <class Sf, int W>
JC_D Sf limber_point(const EHc<Sf>& eh, const Sf& ns, const PointCoefs<Sf, W>& pc, float lph) {
  Sf k = lph * pc.invchi;
  Sf lnk = jm::log(k);
  Sf Tk = eh_transfer(eh, k, lnk);
  Sf plin = jm::exp(ns * lnk) * (Tk * Tk) * pc.g2p;   // power.py:48-51
  Sf k3 = k * k * k;
  Sf lny = lnk + pc.lnRnl;
  Sf y = k * pc.Rnl;                                   // power.py:252
  return pc.wgt * halofit_pnl(pc.hf, lny, y, k3, plin);
}

template <int T, int ND>
__global__ __launch_bounds__(256) void k_limber(LimberArgs A) {
  using Sf = typename Types<ND>::F;
  constexpr int W = ND + 1;
  constexpr int P = T * (T + 1) / 2;
  constexpr bool VJP = (ND == 7);
  constexpr int NACC = VJP ? 7 : P * W;
  extern __shared__ float red[];  // [waves-1][NACC][64]

  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int nwv = A.waves;
  const int g = blockIdx.x * 64 + lane;
  const bool valid = g < A.ncols;
  const int gc = valid ? g : A.ncols - 1;
  const int c = gc / A.n_ell;
  const int li = gc - c * A.n_ell;
  const double ell = A.ell[li];
  const float lph = (float)(ell + 0.5);
  const float lf = (float)(::sqrt((ell - 1.0) * ell * (ell + 1.0) * (ell + 2.0)) / ((ell + 0.5) * (ell + 0.5)));
  float F[T];
#pragma unroll
  for (int i = 0; i < T; ++i) F[i] = ((A.wl_mask >> i) & 1u) ? lf : 1.0f;

  Sf ns;
  const EHc<Sf> eh = load_eh<Sf, W>(A.crec + (size_t)c * NCREC * W, ns);

  const int chunk = (A.Na + nwv - 1) / nwv;
  const int m0 = wave * chunk;
  const int m1 = min(A.Na, m0 + chunk);

  float acc[NACC];
#pragma unroll
  for (int q = 0; q < NACC; ++q) acc[q] = 0.0f;

  // adjoint: symmetric cotangent matrix ct_ij (diagonal doubled), times F_i F_j
  float ct[VJP ? T : 1][VJP ? T : 1];
  if constexpr (VJP) {
    int p = 0;
#pragma unroll
    for (int i = 0; i < T; ++i) {
#pragma unroll
      for (int j = i; j < T; ++j) {
        const float cb = A.cbar[((size_t)c * P + p) * A.n_ell + li] * F[i] * F[j];
        ct[i][j] = (i == j) ? 2.0f * cb : cb;
        ct[j][i] = ct[i][j];
        ++p;
      }
    }
  }

  const float* cf = A.coef + ((size_t)c * A.Na + m0) * NCOEF * W;
  const float* rr = A.R + ((size_t)c * A.Na + m0) * A.Tpad * W;
  for (int m = m0; m < m1; ++m, cf += NCOEF * W, rr += A.Tpad * W) {
    const PointCoefs<Sf, W> pc = load_coefs<Sf, W>(cf);
    const Sf M = limber_point<Sf, W>(eh, ns, pc, lph);
    Sf Rv[T];
#pragma unroll
    for (int i = 0; i < T; ++i) Rv[i] = Mem<Sf>::ld(rr + i * W);
    if constexpr (!VJP) {
      int p = 0;
#pragma unroll
      for (int i = 0; i < T; ++i) {
        const Sf u = Rv[i] * M;
#pragma unroll
        for (int j = i; j < T; ++j) {
          Sf q = Mem<Sf>::ld(&acc[p * W]);
          q += Rv[j] * u;
          Mem<Sf>::st(&acc[p * W], q);
          ++p;
        }
      }
    } else {
      float V[T];
      float G = 0.0f;
#pragma unroll
      for (int i = 0; i < T; ++i) {
        float s = 0.0f;
#pragma unroll
        for (int j = 0; j < T; ++j) s += ct[i][j] * val(Rv[j]);
        V[i] = s;
        G += 0.5f * val(Rv[i]) * s;
      }
#pragma unroll
      for (int d = 0; d < 7; ++d) {
        float s = 0.0f;
#pragma unroll
        for (int i = 0; i < T; ++i) s += Rv[i].d[d] * V[i];
        acc[d] += val(M) * s + G * M.d[d];
      }
    }
  }

  // deterministic reduction over the waves of the block
  if (nwv > 1) {
    if (wave > 0) {
#pragma unroll
      for (int q = 0; q < NACC; ++q) red[((wave - 1) * NACC + q) * 64 + lane] = acc[q];
    }
    __syncthreads();
    if (wave == 0) {
      for (int w = 1; w < nwv; ++w) {
#pragma unroll
        for (int q = 0; q < NACC; ++q) acc[q] += red[((w - 1) * NACC + q) * 64 + lane];
      }
    }
  }
  if (wave != 0 || !valid) return;
  if constexpr (!VJP) {
    int p = 0;
#pragma unroll
    for (int i = 0; i < T; ++i) {
#pragma unroll
      for (int j = i; j < T; ++j) {
        const float f = F[i] * F[j];
        const size_t o = ((size_t)c * P + p) * A.n_ell + li;
        A.out[o] = acc[p * W] * f;
        if constexpr (ND == 1) A.out_t[o] = acc[p * W + 1] * f;
        ++p;
      }
    }
  } else {
#pragma unroll
    for (int d = 0; d < 7; ++d) A.theta_col[(size_t)g * 7 + d] = acc[d];
  }
}

template <int T, int ND>
static hipError_t launch_T(const LimberArgs& a, hipStream_t s) {
  constexpr int P = T * (T + 1) / 2;
  constexpr int NACC = (ND == 7) ? 7 : P * (ND + 1);
  const size_t sm = (size_t)(a.waves - 1) * NACC * 64 * sizeof(float);
  const int blocks = (a.ncols + 63) / 64;
  hipLaunchKernelGGL((k_limber<T, ND>), dim3(blocks), dim3(64 * a.waves), sm, s, a);
  return hipGetLastError();
}

template <int ND>
static hipError_t launch_nd(int T, const LimberArgs& a, hipStream_t s) {
  switch (T) {
    case 1: return launch_T<1, ND>(a, s);
    case 2: return launch_T<2, ND>(a, s);
    case 3: return launch_T<3, ND>(a, s);
    case 4: return launch_T<4, ND>(a, s);
    case 5: return launch_T<5, ND>(a, s);
    case 6: return launch_T<6, ND>(a, s);
    case 7: return launch_T<7, ND>(a, s);
    case 8: return launch_T<8, ND>(a, s);
    case 9: return launch_T<9, ND>(a, s);
    case 10: return launch_T<10, ND>(a, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_limber(int nd, int T, const LimberArgs& a, hipStream_t s) {
  if (a.waves < 1 || a.waves > 4) return hipErrorInvalidValue;
  switch (nd) {
    case 0: return launch_nd<0>(T, a, s);
    case 1: return launch_nd<1>(T, a, s);
    case 7: return launch_nd<7>(T, a, s);
    default: return hipErrorInvalidValue;
  }
}

// theta[c][d] = sum over the n_ell columns of cosmology c (FP64, fixed order)
__global__ void k_theta_reduce(const float* __restrict__ tc, int B, int n_ell, double* __restrict__ th) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= B * 7) return;
  const int c = t / 7, d = t - c * 7;
  double s = 0.0;
  for (int l = 0; l < n_ell; ++l) s += (double)tc[((size_t)c * n_ell + l) * 7 + d];
  th[t] = s;
}

hipError_t launch_theta_reduce(const float* theta_col, int B, int n_ell, double* theta, hipStream_t s) {
  const int n = B * 7;
  hipLaunchKernelGGL(k_theta_reduce, dim3((n + 255) / 256), dim3(256), 0, s, theta_col, B, n_ell, theta);
  return hipGetLastError();
}

// P_lin and P_nl at (k_i, a_i) pairs of one cosmology whose per-a records were
// built on the a_i list (power.py:21-53, 144-272).
__global__ void k_pk_eval(PkArgs A) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.n) return;
  float ns;
  const EHc<float> eh = load_eh<float, 1>(A.crec, ns);
  const PointCoefs<float, 1> pc = load_coefs<float, 1>(A.coef + (size_t)i * NCOEF);
  const float k = (float)A.k[i];
  const float lnk = jm::log(k);
  const float Tk = eh_transfer(eh, k, lnk);
  const float plin = jm::exp(ns * lnk) * (Tk * Tk) * pc.g2p;
  const float k3 = k * k * k;
  const float pnl = halofit_pnl(pc.hf, lnk + pc.lnRnl, k * pc.Rnl, k3, plin);
  A.plin[i] = plin;
  A.pnl[i] = pnl;
}

hipError_t launch_pk_eval(const PkArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_pk_eval, dim3((a.n + 255) / 256), dim3(256), 0, s, a);
  return hipGetLastError();
}

}  // namespace jc
