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
#include "jc_records.hpp"

namespace jc {

// Per-lane constants of a column: ell + 1/2, its inverse and logarithm.
struct EllC {
  float lph, ilph, lnlph;
};
JC_D EllC ell_consts(double ell) {
  EllC e;
  e.lph = (float)(ell + 0.5);
  e.ilph = (float)(1.0 / (ell + 0.5));
  e.lnlph = (float)::log(ell + 0.5);
  return e;
}

// M(a, ell) = weight * P_nl(k, a) at k = (ell + 1/2) / max(chi, 1)
template <class Sf, int W>
JC_D Sf limber_point(const EHc<Sf>& eh, const Sf& ns, const PointCoefs<Sf, W>& pc, const EllC& lc) {
  Sf k = lc.lph * pc.invchi;
  Sf invk = lc.ilph * pc.chi1;
  Sf lnk = lc.lnlph + pc.lninvchi;
  Sf Tk = eh_transfer(eh, k, invk, lnk);
  Sf plin = jm::exp(ns * lnk) * (Tk * Tk) * pc.g2p;   // power.py:48-51
  Sf k3 = k * k * k;
  Sf ik3 = invk * invk * invk;
  Sf lgy = lnk * 1.4426950408889634f + pc.lgRnl;  // log2 y
  Sf y = k * pc.Rnl;                                   // power.py:252
  Sf iy = invk * pc.knl;
  return pc.wgt * halofit_pnl(pc.hf, lgy, y, iy, k3, ik3, plin);
}

// The same with T(k) k^(n_s/2) read from the cosmology's transfer table in
// LDS (cubic Hermite pieces uniform in ln k).  The table position is formed
// in FP64 so that it matches the node placement exactly; the remaining
// uncertainty is that of the float ln k itself.
JC_D float limber_point_tab(const float4* tab, double invh, double lo_invh, const PointCoefs<float, 1>& pc,
                            const EllC& lc) {
  const float k = lc.lph * pc.invchi;
  const float invk = lc.ilph * pc.chi1;
  const float lnk = lc.lnlph + pc.lninvchi;
  const double u = __builtin_fma((double)lnk, invh, -lo_invh);
  int q = (int)u;
  q = q < 0 ? 0 : (q > NEHQ - 1 ? NEHQ - 1 : q);
  const float t = (float)(u - (double)q);
  const float4 cc = tab[q];
  const float tn = __builtin_fmaf(__builtin_fmaf(__builtin_fmaf(cc.w, t, cc.z), t, cc.y), t, cc.x);
  const float plin = (tn * tn) * pc.g2p;               // power.py:48-51
  const float k3 = k * k * k;
  const float ik3 = invk * invk * invk;
  const float lgy = __builtin_fmaf(lnk, 1.4426950408889634f, pc.lgRnl);  // log2 y
  const float y = k * pc.Rnl;                          // power.py:252
  const float iy = invk * pc.knl;
  return pc.wgt * halofit_pnl(pc.hf, lgy, y, iy, k3, ik3, plin);
}

// Adjoint kernel: M = weight * P_nl at one point and its 7 parameter
// derivatives.  T(k) stays a forward dual (the EH98 constants and k carry
// theta); everything after it is differentiated in reverse, once per point:
// a backward sweep gives dM/dx for every input x of limber_point (the record
// fields, n_s, T), and dM/dtheta_d = sum_x dM/dx dx/dtheta_d with dx/dtheta
// the dual parts the table kernel stored.  The forward order of operations
// is limber_point's (halofit_pnl in base 2 with one shared reciprocal).
// tn = T(k) k^(n_s/2) is a forward dual: from the EH98 closed form (duals
// through the constants and k), or read from the transfer table with its
// theta derivatives (adj_tn_tab); plin = tn^2 D^2 pknorm.
template <int W>
JC_D void limber_point_adj(const Dual<float, 7>& tn, const PointCoefs<Dual<float, 7>, W>& pc, const EllC& lc,
                           float& Mv, float* Md) {
  using D = Dual<float, 7>;
  constexpr float LN2 = 0.69314718055994531f, LOG2E = 1.4426950408889634f;
  constexpr float C0 = (float)(1.0 / (2.0 * PI * PI)), C2 = (float)(2.0 * PI * PI);
  constexpr float A1 = 0.25f * LOG2E, A2 = 0.125f * LOG2E;
  // forward values
  const float k = lc.lph * pc.invchi.v, invk = lc.ilph * pc.chi1.v, lnk = lc.lnlph + pc.lninvchi.v;
  const float T = tn.v;
  const float g2p = pc.g2p.v;
  const float plin = (T * T) * g2p;
  const float k3 = k * k * k, ik3 = invk * invk * invk;
  const float lgy = __builtin_fmaf(lnk, LOG2E, pc.lgRnl.v);
  const float y = k * pc.Rnl.v, iy = invk * pc.knl.v;
  const HFa<D>& h = pc.hf;
  const float d2l = k3 * plin * C0;
  const float L1 = jm::log2(1.0f + d2l);
  const float fy2 = y * (A1 + y * A2);
  const float eq = jm::exp2(h.beta_n.v * L1 - fy2);
  const float yf1 = jm::exp2(h.f1x3.v * lgy);
  const float yf2 = jm::exp2(h.f2.v * lgy);
  const float ycg = jm::exp2(h.gam3.v * (h.lg_cf3.v + lgy));
  const float d0 = 1.0f + h.alpha_n.v * d2l;
  const float Aq = 1.0f + h.b_n.v * yf2 + ycg;
  const float Bq = 1.0f + h.mu_n.v * iy + h.nu_n.v * (iy * iy);
  const float d12 = Aq * Bq;
  const float hq = C2 * ik3 * (h.a_n.v * yf1);
  const float num = plin * eq * d12 + hq * d0;
  const float rden = __builtin_amdgcn_rcpf(d0 * d12);
  const float P = num * rden;
  Mv = pc.wgt.v * P;
  // backward sweep, seed dM = 1
  const float P_b = pc.wgt.v;
  const float num_b = P_b * rden;
  const float den_b = -P_b * P * rden;
  float plin_b = num_b * eq * d12;
  const float eq_b = num_b * plin * d12;
  float d12_b = num_b * plin * eq;
  const float hq_b = num_b * d0;
  float d0_b = num_b * hq;
  d0_b += den_b * d12;
  d12_b += den_b * d0;
  const float ik3_b = hq_b * C2 * h.a_n.v * yf1;
  const float an_b = hq_b * C2 * ik3 * yf1;
  const float yf1_b = hq_b * C2 * ik3 * h.a_n.v;
  const float A_b = d12_b * Bq, B_b = d12_b * Aq;
  const float mu_b = B_b * iy, nu_b = B_b * (iy * iy);
  float iy_b = B_b * (h.mu_n.v + 2.0f * h.nu_n.v * iy);
  const float bn_b = A_b * yf2;
  const float yf2_b = A_b * h.b_n.v;
  const float alpha_b = d0_b * d2l;
  float d2l_b = d0_b * h.alpha_n.v;
  const float tc = A_b * ycg * LN2;  // ycg_b = A_b
  const float gam3_b = tc * (h.lg_cf3.v + lgy), lgcf3_b = tc * h.gam3.v;
  float lgy_b = tc * h.gam3.v;
  const float t2 = yf2_b * yf2 * LN2;
  const float f2_b = t2 * lgy;
  lgy_b += t2 * h.f2.v;
  const float t1 = yf1_b * yf1 * LN2;
  const float f1x3_b = t1 * lgy;
  lgy_b += t1 * h.f1x3.v;
  const float te = eq_b * eq * LN2;
  const float beta_b = te * L1;
  d2l_b += te * h.beta_n.v * __builtin_amdgcn_rcpf((1.0f + d2l) * LN2);
  const float y_b = -te * (A1 + 2.0f * A2 * y);
  const float k3_b = d2l_b * plin * C0;
  plin_b += d2l_b * k3 * C0;
  float lnk_b = lgy_b * LOG2E;
  const float lgRnl_b = lgy_b;
  float k_b = y_b * pc.Rnl.v + k3_b * 3.0f * (k * k);
  const float Rnl_b = y_b * k;
  float invk_b = iy_b * pc.knl.v + ik3_b * 3.0f * (invk * invk);
  const float knl_b = iy_b * invk;
  const float T_b = plin_b * 2.0f * T * g2p;
  const float g2p_b = plin_b * (T * T);
  // k, 1/k and ln k reach theta through the record fields (tn's own k
  // dependence is inside its dual part)
  const float invchi_b = k_b * lc.lph, chi1_b = invk_b * lc.ilph;
#pragma unroll
  for (int d = 0; d < 7; ++d) {
    float s = P * pc.wgt.d[d] + T_b * tn.d[d] + g2p_b * pc.g2p.d[d];
    s += invchi_b * pc.invchi.d[d] + chi1_b * pc.chi1.d[d] + lnk_b * pc.lninvchi.d[d];
    s += lgRnl_b * pc.lgRnl.d[d] + Rnl_b * pc.Rnl.d[d] + knl_b * pc.knl.d[d];
    s += an_b * h.a_n.d[d] + bn_b * h.b_n.d[d] + lgcf3_b * h.lg_cf3.d[d] + gam3_b * h.gam3.d[d];
    s += alpha_b * h.alpha_n.d[d] + beta_b * h.beta_n.d[d] + nu_b * h.nu_n.d[d] + f1x3_b * h.f1x3.d[d];
    s += f2_b * h.f2.d[d] + mu_b * h.mu_n.d[d];
    Md[d] = s;
  }
}

// tn = T(k) k^(n_s/2) of the adjoint from the EH98 closed form
template <int W>
JC_D Dual<float, 7> adj_tn_closed(const EHc<Dual<float, 7>>& eh, const Dual<float, 7>& ns,
                                  const PointCoefs<Dual<float, 7>, W>& pc, const EllC& lc) {
  using D = Dual<float, 7>;
  const D kd = lc.lph * pc.invchi;
  const D invkd = lc.ilph * pc.chi1;
  const D lnkd = lc.lnlph + pc.lninvchi;
  return eh_transfer(eh, kd, invkd, lnkd) * jm::exp(ns * (0.5f * lnkd));
}

// ... and from the cosmology's transfer table: the value and its ln k slope
// from the cubic piece (as the forward kernel), the theta derivatives at
// fixed k by 4-point Lagrange interpolation of their values at the table's
// nodes q-1 .. q+2 (clamped at the ends; rows of 8 floats read through the
// L2), plus slope x d ln k.  Linear interpolation on a 4x coarser grid was
// 1.3e-3 off in d/d Omega_b: its baryon-wiggle part oscillates in ln k with
// a period of a few coarse nodes near k ~ 0.3 h/Mpc.
template <int W>
JC_D Dual<float, 7> adj_tn_tab(const float4* tab, const float* __restrict__ dtab, double invh, double lo_invh,
                               const PointCoefs<Dual<float, 7>, W>& pc, const EllC& lc) {
  const float lnk = lc.lnlph + pc.lninvchi.v;
  const double u = __builtin_fma((double)lnk, invh, -lo_invh);
  int q = (int)u;
  q = q < 0 ? 0 : (q > NEHQ - 1 ? NEHQ - 1 : q);
  const float t = (float)(u - (double)q);
  const float4 cc = tab[q];
  Dual<float, 7> tn;
  tn.v = __builtin_fmaf(__builtin_fmaf(__builtin_fmaf(cc.w, t, cc.z), t, cc.y), t, cc.x);
  const float slope = __builtin_fmaf(__builtin_fmaf(3.0f * cc.w, t, 2.0f * cc.z), t, cc.y) * (float)invh;
  // nodes j0 .. j0 + 3 with j0 = clamp(q - 1, 0, NEHQ - 3); s = position from j0
  const int j0 = q < 1 ? 0 : (q > NEHQ - 2 ? NEHQ - 3 : q - 1);
  const float s = t + (float)(q - j0);
  const float s1 = s - 1.0f, s2 = s - 2.0f, s3 = s - 3.0f;
  const float w0 = -(1.0f / 6.0f) * s1 * s2 * s3, w1 = 0.5f * s * s2 * s3;
  const float w2 = -0.5f * s * s1 * s3, w3 = (1.0f / 6.0f) * s * s1 * s2;
  const float4* r = reinterpret_cast<const float4*>(dtab + (size_t)j0 * NEHD_ROW);
  const float4 a0 = r[0], b0 = r[1], a1 = r[2], b1 = r[3], a2 = r[4], b2 = r[5], a3 = r[6], b3 = r[7];
  const float n0[7] = {a0.x, a0.y, a0.z, a0.w, b0.x, b0.y, b0.z};
  const float n1[7] = {a1.x, a1.y, a1.z, a1.w, b1.x, b1.y, b1.z};
  const float n2[7] = {a2.x, a2.y, a2.z, a2.w, b2.x, b2.y, b2.z};
  const float n3[7] = {a3.x, a3.y, a3.z, a3.w, b3.x, b3.y, b3.z};
#pragma unroll
  for (int d = 0; d < 7; ++d)
    tn.d[d] = __builtin_fmaf(w3, n3[d], __builtin_fmaf(w2, n2[d], __builtin_fmaf(w1, n1[d], w0 * n0[d]))) +
              slope * pc.lninvchi.d[d];
  return tn;
}

// The dispatcher deals blocks to the 8 XCDs in turn (block b to XCD b % 8),
// and each XCD has its own L2.  Neighbouring column groups read the same
// cosmology's records, so blocks are renumbered to give each XCD one
// contiguous range of groups (a bijection of [0, n) for any n): a record is
// then fetched into one L2 instead of two or three.
__device__ inline int xcd_group(int b, int n) {
  const int q = n >> 3, r = n & 7, x = b & 7;
  return x * q + min(x, r) + (b >> 3);
}

// accumulators reduced through LDS per pass (bounds the LDS a block needs)
template <int NACC>
struct RedChunk { static constexpr int value = NACC < 32 ? NACC : 32; };

// PF: 0 = the a-node's records are loaded when needed, 1 = the per-a
// coefficients of the next node are loaded one node ahead (the tracer
// kernels are loaded at the top of each node and used after P_nl).
// (forward kernels are held to 128 VGPRs: 4 waves per SIMD)
template <int T, int ND, bool TAB, int PF>
__global__ __launch_bounds__(256, (ND == 0 ? 4 : 1)) void k_limber(LimberArgs A) {
  using Sf = typename Types<ND>::F;
  constexpr int W = ND + 1;
  constexpr int P = T * (T + 1) / 2;
  constexpr bool VJP = (ND == 7);
  constexpr int NACC = VJP ? 7 : P * W;
  constexpr int RCH = RedChunk<NACC>::value;
  static_assert(!TAB || ND == 0 || ND == 7, "the transfer table serves the forward and adjoint kernels");
  // TAB: [2][NEHQ] tables (the adjoint reads its theta-derivative rows from
  // global memory); after the a loop, [waves-1][RCH][64] partial sums
  extern __shared__ float4 lds4[];
  float* red = reinterpret_cast<float*>(lds4);

  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int nwv = A.waves;
  const int grp = xcd_group(blockIdx.x, gridDim.x);
  const int g = grp * 64 + lane;
  const bool valid = g < A.ncols;
  const int gc = valid ? g : A.ncols - 1;
  const int c = gc / A.n_ell;
  const int li = gc - c * A.n_ell;
  const double ell = A.ell[li];
  const EllC lc = ell_consts(ell);
  const float lf = (float)(::sqrt((ell - 1.0) * ell * (ell + 1.0) * (ell + 2.0)) / ((ell + 0.5) * (ell + 0.5)));
  float F[T];
#pragma unroll
  for (int i = 0; i < T; ++i) F[i] = ((A.wl_mask >> i) & 1u) ? lf : 1.0f;

  Sf ns;
  EHc<Sf> eh;
  const float4* tab = nullptr;
  const float* dtab = nullptr;
  double invh = 0.0, lo_invh = 0.0;
  if constexpr (TAB) {
    // the block's 64 columns span at most two cosmologies (n_ell >= 64)
    const int cb = (grp * 64) / A.n_ell;
    const int ce = min(A.ncols - 1, grp * 64 + 63) / A.n_ell;
    for (int e = threadIdx.x; e < (ce - cb + 1) * NEHQ; e += blockDim.x)
      lds4[e] = A.ehtab[(size_t)cb * NEHQ + e];
    tab = lds4 + (c - cb) * NEHQ;
    if constexpr (VJP) dtab = A.ehtab7 + (size_t)c * (NEHQ + 1) * NEHD_ROW;
    const float* cr = A.crec + (size_t)c * NCREC * W;
    invh = (double)cr[CR_EHINVH * W];
    lo_invh = (double)cr[CR_EHLO * W] * invh;  // exact in FP64
    __syncthreads();
  } else {
    eh = load_eh<Sf, W>(A.crec + (size_t)c * NCREC * W, ns);
  }

  // the a range of this wave: one of slices x waves contiguous chunks
  const int nsl = A.slices;
  const int chunk = (A.Na + nsl * nwv - 1) / (nsl * nwv);
  const int m0 = min(A.Na, ((int)blockIdx.y * nwv + wave) * chunk);
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
  // one (column, a-node) point: P_nl, then the contraction
  auto point = [&](const PointCoefs<Sf, W>& pc, const float* rp) {
    Sf Rv[T];
#pragma unroll
    for (int i = 0; i < T; ++i) Rv[i] = Mem<Sf>::ld(rp + i * W);
    if constexpr (!VJP) {
      Sf M;
      if constexpr (TAB) M = limber_point_tab(tab, invh, lo_invh, pc, lc);
      else M = limber_point<Sf, W>(eh, ns, pc, lc);
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
      float Mv, Md[7];
      if constexpr (TAB) limber_point_adj<W>(adj_tn_tab<W>(tab, dtab, invh, lo_invh, pc, lc), pc, lc, Mv, Md);
      else limber_point_adj<W>(adj_tn_closed<W>(eh, ns, pc, lc), pc, lc, Mv, Md);
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
        acc[d] += Mv * s + G * Md[d];
      }
    }
  };
  if constexpr (PF) {
    // two register sets in turn: the next node's coefficients load while the
    // current node computes, with no copies between the sets
    constexpr int CS = NCOEF * W;
    const int RS = A.Tpad * W;
    PointCoefs<Sf, W> pa, pb;
    if (m0 < m1) pa = load_coefs<Sf, W>(cf);
    for (int m = m0; m < m1; m += 2, cf += 2 * CS, rr += 2 * RS) {
      const bool has_b = m + 1 < m1;
      if (has_b) pb = load_coefs<Sf, W>(cf + CS);
      point(pa, rr);
      if (!has_b) break;
      if (m + 2 < m1) pa = load_coefs<Sf, W>(cf + 2 * CS);
      point(pb, rr + RS);
    }
  } else {
    for (int m = m0; m < m1; ++m, cf += NCOEF * W, rr += A.Tpad * W) point(load_coefs<Sf, W>(cf), rr);
  }

  // deterministic reduction over the waves of the block, RCH accumulators
  // per pass (the LDS space of the tables is reused)
  if (nwv > 1) {
#pragma unroll
    for (int q0 = 0; q0 < NACC; q0 += RCH) {
      __syncthreads();
      if (wave > 0) {
#pragma unroll
        for (int q = q0; q < q0 + RCH && q < NACC; ++q) red[((wave - 1) * RCH + (q - q0)) * 64 + lane] = acc[q];
      }
      __syncthreads();
      if (wave == 0) {
        for (int w = 1; w < nwv; ++w) {
#pragma unroll
          for (int q = q0; q < q0 + RCH && q < NACC; ++q) acc[q] += red[((w - 1) * RCH + (q - q0)) * 64 + lane];
        }
      }
    }
  }
  if (wave != 0 || !valid) return;
  if constexpr (!VJP) {
    // with slices, this block's share goes to part[slice] (then out_t's at
    // part[slices + slice]) in the layout of out
    const size_t nout = (size_t)A.ncols * P;
    float* out = nsl > 1 ? A.part + (size_t)blockIdx.y * nout : A.out;
    float* out_t = nsl > 1 ? A.part + (size_t)(nsl + blockIdx.y) * nout : A.out_t;
    int p = 0;
#pragma unroll
    for (int i = 0; i < T; ++i) {
#pragma unroll
      for (int j = i; j < T; ++j) {
        const float f = F[i] * F[j];
        const size_t o = ((size_t)c * P + p) * A.n_ell + li;
        out[o] = acc[p * W] * f;
        if constexpr (ND == 1) out_t[o] = acc[p * W + 1] * f;
        ++p;
      }
    }
  } else {
    float* tc = nsl > 1 ? A.part + (size_t)blockIdx.y * A.ncols * 7 : A.theta_col;
#pragma unroll
    for (int d = 0; d < 7; ++d) tc[(size_t)g * 7 + d] = acc[d];
  }
}

// out[o] = part[o] + part[n + o] + ... (slices terms, in order)
__global__ void k_slice_sum(const float* __restrict__ part, int slices, int64_t n, float* __restrict__ out) {
  const int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= n) return;
  float s = part[o];
  for (int k = 1; k < slices; ++k) s += part[(size_t)k * n + o];
  out[o] = s;
}

hipError_t launch_slice_sum(const float* part, int slices, int64_t n, float* out, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_slice_sum, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, part, slices, n, out);
  return hipGetLastError();
}

// ------------------------------------------------------------------ matrix-core contraction
// Forward kernel with the pair contraction on the matrix cores: for one
// cosmology C_p(ell) = sum_a K_p(a) M(a, ell) with K_p = R_i R_j is a GEMM
// [pairs x a] x [a x ell].  A wave holds 32 columns x 2 a-nodes: lane l owns
// column (l & 31) at node 2q + (l >> 5), which is exactly the B operand layout
// of v_mfma_f32_32x32x2_f32 (B[k][n] in lane n + 32k), so each lane's
// M(a, ell) feeds the matrix core with no data movement.  The A operand
// (K_p at the same two nodes, lane p + 32k) is formed from the tracer kernel
// records; pairs fill one or two 32-row blocks.  Columns of a 32-column group
// span at most two cosmologies (n_ell >= 32); where they do, each
// cosmology's K multiplies its own columns only (B masked), so a group costs
// two products instead of one.  A block holds two column groups (64 columns)
// times `waves` splits of the a range.
typedef float floatx16 __attribute__((ext_vector_type(16)));

// pair index p -> (i, j), i <= j, row-major over i (angular_cl.py:15-25)
JC_D void pair_ij(int p, int T, int& i, int& j) {
  i = 0;
  while (p >= T - i) {
    p -= T - i;
    ++i;
  }
  j = i + p;
}

template <int T, bool TAB>
__global__ __launch_bounds__(512, 4) void k_limber_mx(LimberArgs A) {
  constexpr int P = T * (T + 1) / 2;
  constexpr int NRB = (P + 31) / 32;  // 32-row blocks of pairs
  constexpr int NACC = NRB * 16;
  constexpr int RCH = RedChunk<NACC>::value;
  extern __shared__ float4 lds4[];  // TAB: [<=3][NEHQ] tables; then [splits-1][2][RCH][64] partial sums
  float* red = reinterpret_cast<float*>(lds4);

  const int lane = threadIdx.x & 63;
  const int grp = (threadIdx.x >> 6) & 1;  // column group of the wave
  const int wave = threadIdx.x >> 7;       // a-range split of the wave
  const int nwv = A.waves;
  const int col0 = blockIdx.x * 64 + grp * 32;
  const int kk = lane >> 5;  // node parity of this lane
  const int g = col0 + (lane & 31);
  const bool valid = g < A.ncols;
  const int gc = valid ? g : A.ncols - 1;
  const int c = gc / A.n_ell;
  const int li = gc - c * A.n_ell;
  const int cA = col0 / A.n_ell;
  const int cB = min(col0 + 31, A.ncols - 1) / A.n_ell;
  const bool mixed = cA != cB;  // uniform over the block
  const double ell = A.ell[li];
  const EllC lc = ell_consts(ell);
  const float lf = (float)(::sqrt((ell - 1.0) * ell * (ell + 1.0) * (ell + 2.0)) / ((ell + 0.5) * (ell + 0.5)));

  // tracer indices of this lane's A-operand rows
  int ia[NRB], ja[NRB];
  bool pa[NRB];
#pragma unroll
  for (int rb = 0; rb < NRB; ++rb) {
    const int p = rb * 32 + (lane & 31);
    pa[rb] = p < P;
    pair_ij(pa[rb] ? p : 0, T, ia[rb], ja[rb]);
  }

  float ns = 0.0f;
  EHc<float> eh;
  const float4* tab = nullptr;
  double invh = 0.0, lo_invh = 0.0;
  if constexpr (TAB) {
    const int cb0 = (blockIdx.x * 64) / A.n_ell;
    const int ce0 = min(blockIdx.x * 64 + 63, A.ncols - 1) / A.n_ell;
    for (int e = threadIdx.x; e < (ce0 - cb0 + 1) * NEHQ; e += blockDim.x) lds4[e] = A.ehtab[(size_t)cb0 * NEHQ + e];
    tab = lds4 + (c - cb0) * NEHQ;
    const float* cr = A.crec + (size_t)c * NCREC;
    invh = (double)cr[CR_EHINVH];
    lo_invh = (double)cr[CR_EHLO] * invh;
    __syncthreads();
  } else {
    eh = load_eh<float, 1>(A.crec + (size_t)c * NCREC, ns);
  }

  floatx16 acc[NRB];
#pragma unroll
  for (int rb = 0; rb < NRB; ++rb)
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[rb][r] = 0.0f;

  // per-lane pointers to the tracer kernels of its A-operand rows
  const float* rAi[NRB];
  const float* rAj[NRB];
  const float* rBi[NRB];
  const float* rBj[NRB];
#pragma unroll
  for (int rb = 0; rb < NRB; ++rb) {
    rAi[rb] = A.R + (size_t)cA * A.Na * A.Tpad + ia[rb];
    rAj[rb] = A.R + (size_t)cA * A.Na * A.Tpad + ja[rb];
    rBi[rb] = A.R + (size_t)cB * A.Na * A.Tpad + ia[rb];
    rBj[rb] = A.R + (size_t)cB * A.Na * A.Tpad + ja[rb];
  }

  const int np = (A.Na + 1) / 2;  // node pairs
  const int chunk = (np + nwv - 1) / nwv;
  const int q0 = wave * chunk;
  const int q1 = min(np, q0 + chunk);

  auto step = [&](int q, const PointCoefs<float, 1>& pc) {
    const int m = 2 * q + kk;
    const bool vm = m < A.Na;
    const int mm = vm ? m : A.Na - 1;
    float M;
    if constexpr (TAB) M = limber_point_tab(tab, invh, lo_invh, pc, lc);
    else M = limber_point<float, 1>(eh, ns, pc, lc);
    M = (vm && valid) ? M : 0.0f;
    // the loads are unconditional (in-bounds rows) and masked afterwards
    const int ro = mm * A.Tpad;
    const float bA = (c == cA) ? M : 0.0f;
#pragma unroll
    for (int rb = 0; rb < NRB; ++rb) {
      float a = rAi[rb][ro] * rAj[rb][ro];
      a = (pa[rb] && vm) ? a : 0.0f;
      acc[rb] = __builtin_amdgcn_mfma_f32_32x32x2f32(a, bA, acc[rb], 0, 0, 0);
    }
    if (mixed) {
      const float bB = (c == cB) ? M : 0.0f;
#pragma unroll
      for (int rb = 0; rb < NRB; ++rb) {
        float a = rBi[rb][ro] * rBj[rb][ro];
        a = (pa[rb] && vm) ? a : 0.0f;
        acc[rb] = __builtin_amdgcn_mfma_f32_32x32x2f32(a, bB, acc[rb], 0, 0, 0);
      }
    }
  };
  // the records of the lane's next node load while the current one computes
  const float* cbase = A.coef + (size_t)c * A.Na * NCOEF;
  auto rec = [&](int q) {
    const int m = min(2 * q + kk, A.Na - 1);
    return load_coefs<float, 1>(cbase + (size_t)m * NCOEF);
  };
  PointCoefs<float, 1> pa_, pb_;
  if (q0 < q1) pa_ = rec(q0);
  for (int q = q0; q < q1; q += 2) {
    const bool has_b = q + 1 < q1;
    if (has_b) pb_ = rec(q + 1);
    step(q, pa_);
    if (!has_b) break;
    if (q + 2 < q1) pa_ = rec(q + 2);
    step(q + 1, pb_);
  }

  // deterministic reduction over the waves of the block
  if (nwv > 1) {
#pragma unroll
    for (int x0 = 0; x0 < NACC; x0 += RCH) {
      __syncthreads();
      if (wave > 0) {
#pragma unroll
        for (int x = x0; x < x0 + RCH && x < NACC; ++x)
          red[(((wave - 1) * 2 + grp) * RCH + (x - x0)) * 64 + lane] = acc[x >> 4][x & 15];
      }
      __syncthreads();
      if (wave == 0) {
        for (int w = 1; w < nwv; ++w) {
#pragma unroll
          for (int x = x0; x < x0 + RCH && x < NACC; ++x)
            acc[x >> 4][x & 15] += red[(((w - 1) * 2 + grp) * RCH + (x - x0)) * 64 + lane];
        }
      }
    }
  }
  if (wave != 0 || !valid) return;
  // result layout of the 32x32 product: lane l, register r holds row
  // 8 (r / 4) + 4 (l / 32) + r % 4 of column l % 32
#pragma unroll
  for (int rb = 0; rb < NRB; ++rb) {
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int p = rb * 32 + 8 * (r >> 2) + 4 * kk + (r & 3);
      if (p < P) {
        int ti, tj;
        pair_ij(p, T, ti, tj);
        const float f = (((A.wl_mask >> ti) & 1u) ? lf : 1.0f) * (((A.wl_mask >> tj) & 1u) ? lf : 1.0f);
        A.out[((size_t)c * P + p) * A.n_ell + li] = acc[rb][r] * f;
      }
    }
  }
}

// record prefetch per mode (measured on MI355X: the table mode's shorter
// point leaves load latency exposed; the closed form hides it and is
// register-bound)
constexpr int PF_TAB = 1;
constexpr int PF_CLOSED = 0;

template <int T, int ND>
static hipError_t launch_T(const LimberArgs& a, hipStream_t s) {
  constexpr int P = T * (T + 1) / 2;
  constexpr int NACC = (ND == 7) ? 7 : P * (ND + 1);
  constexpr int RCH = RedChunk<NACC>::value;
  const size_t red = (size_t)(a.waves - 1) * RCH * 64 * sizeof(float);
  const int blocks = (a.ncols + 63) / 64;
  if constexpr (ND == 0) {
    if (a.mx && a.n_ell >= 32) {
      constexpr int NACCX = ((P + 31) / 32) * 16;
      const size_t redx = (size_t)(a.waves - 1) * 2 * RedChunk<NACCX>::value * 64 * sizeof(float);
      const int bx = (a.ncols + 63) / 64;
      const dim3 blk(128 * a.waves);
      if (a.ehtab) {
        // 64 columns span at most 2 cosmologies when n_ell >= 64, else 3
        const size_t tab = (size_t)(a.n_ell >= 64 ? 2 : 3) * NEHQ * sizeof(float4);
        hipLaunchKernelGGL((k_limber_mx<T, true>), dim3(bx), blk, tab > redx ? tab : redx, s, a);
      } else {
        hipLaunchKernelGGL((k_limber_mx<T, false>), dim3(bx), blk, redx, s, a);
      }
      return hipGetLastError();
    }
    if (a.ehtab) {
      if (a.n_ell < 64) return hipErrorInvalidValue;  // a block would span more than two cosmologies
      const size_t tab = (size_t)2 * NEHQ * sizeof(float4);
      hipLaunchKernelGGL((k_limber<T, 0, true, PF_TAB>), dim3(blocks, a.slices), dim3(64 * a.waves),
                         tab > red ? tab : red, s, a);
      return hipGetLastError();
    }
  }
  if constexpr (ND == 7) {
    if (a.ehtab) {
      if (a.n_ell < 64 || !a.ehtab7) return hipErrorInvalidValue;
      const size_t tab = (size_t)2 * NEHQ * sizeof(float4);
      hipLaunchKernelGGL((k_limber<T, 7, true, PF_CLOSED>), dim3(blocks, a.slices), dim3(64 * a.waves),
                         tab > red ? tab : red, s, a);
      return hipGetLastError();
    }
  }
  hipLaunchKernelGGL((k_limber<T, ND, false, PF_CLOSED>), dim3(blocks, a.slices), dim3(64 * a.waves), red, s, a);
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
  if (a.slices < 1 || a.slices > 64 || (a.slices > 1 && (!a.part || a.mx))) return hipErrorInvalidValue;
  switch (nd) {
    case 0: return launch_nd<0>(T, a, s);
    case 1: return launch_nd<1>(T, a, s);
    case 7: return launch_nd<7>(T, a, s);
    default: return hipErrorInvalidValue;
  }
}

// theta[c][d] = sum over the n_ell columns of cosmology c (FP64, fixed order)
// One 256-thread block per cosmology: thread t adds the elements t, t + 256,
// ... of the cosmology's [n_ell][7] column partials (coalesced; with nsl > 1
// the a-range slices' partials, sstride apart, are added first) into seven
// FP64 sums (element e belongs to parameter e % 7; 256 = 36 * 7 + 4 moves it
// by four), each wave adds its lanes' sums by a butterfly, and thread q < 7
// adds the four waves' sums of parameter q: a fixed order, so the result does
// not depend on scheduling.
__global__ __launch_bounds__(256) void k_theta_reduce(const float* __restrict__ tc, int nsl, int64_t sstride,
                                                       int n_ell, double* __restrict__ th) {
  __shared__ double wsum[4][8];
  const int c = blockIdx.x;
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const int n = n_ell * 7;
  const float* p = tc + (size_t)c * n;
  double acc[7];
#pragma unroll
  for (int q = 0; q < 7; ++q) acc[q] = 0.0;
  int d = t % 7;
  for (int e = t; e < n; e += 256) {
    double v = 0.0;
#pragma unroll 4
    for (int y = 0; y < nsl; ++y) v += (double)p[(size_t)y * sstride + e];
#pragma unroll
    for (int q = 0; q < 7; ++q) acc[q] += (q == d) ? v : 0.0;
    d = (d >= 3) ? d - 3 : d + 4;
  }
#pragma unroll
  for (int q = 0; q < 7; ++q) {
    double v = acc[q];
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    if (lane == q) wsum[wave][q] = v;
  }
  __syncthreads();
  if (t < 7) th[(size_t)c * 7 + t] = ((wsum[0][t] + wsum[1][t]) + wsum[2][t]) + wsum[3][t];
}

hipError_t launch_theta_reduce(const float* theta_col, int nsl, int B, int n_ell, double* theta, hipStream_t s) {
  const int64_t sstride = (int64_t)B * n_ell * 7;
  hipLaunchKernelGGL(k_theta_reduce, dim3(B), dim3(256), 0, s, theta_col, nsl, sstride, n_ell, theta);
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
  const float invk = (float)(1.0 / A.k[i]);
  const float lnk = (float)::log(A.k[i]);
  const float Tk = eh_transfer(eh, k, invk, lnk);
  const float plin = jm::exp(ns * lnk) * (Tk * Tk) * pc.g2p;
  const float pnl = halofit_pnl(pc.hf, lnk * 1.4426950408889634f + pc.lgRnl, k * pc.Rnl, invk * pc.knl, k * k * k,
                                invk * invk * invk, plin);
  A.plin[i] = plin;
  A.pnl[i] = pnl;
}

hipError_t launch_pk_eval(const PkArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_pk_eval, dim3((a.n + 255) / 256), dim3(256), 0, s, a);
  return hipGetLastError();
}

// ------------------------------------------------------------------ ragged
// Served batches: requests with different tracer subsets (T <= 10) and ell
// lists share one launch.  The host packs the (request, ell) columns in
// buckets of equal T, each bucket padded to whole 64-column tiles, so T is
// uniform across a wave.  Row c of R holds request c's tracers in its own
// order.  Every request's cosmology has its own transfer table (built over
// its own ell range by the table kernel).  The a range of a tile is split
// over the waves of its block and over ragged_slices(live tiles) blocks
// (grid y), so small batches still fill the GPU; the slices' partials are
// added in a fixed order by k_ragged_slice_sum (forward) or k_ragged_reduce
// (adjoint).

// The forward ragged kernel with the tile's tracer count T as a constant
// (one body per T, chosen by a block-uniform switch): fixed pair loops
// with no per-pair tests, the next a-node's coefficients loaded one node
// ahead, and, when the tile's columns belong to at most two requests, their
// transfer tables staged into LDS (otherwise read through the L1/L2).
template <int T, bool STAGED>
__device__ __forceinline__ void ragged_fwd_tile(const RaggedArgs& A, const int tile0, const int nsl, float4* lds4,
                                                const uint64_t starts) {
  constexpr int P = T * (T + 1) / 2;
  constexpr int RCH = RedChunk<P>::value;
  float* red = reinterpret_cast<float*>(lds4);
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int nwv = A.waves;
  const RCol col = A.cols[tile0 + lane];
  const int first = A.cols[tile0].job;  // a tile starts with a real column
  const bool valid = col.job >= 0;
  const int job = valid ? col.job : first;
  const RJob& jb = A.jobs[job];
  const int c = jb.c;
  const uint32_t wl_mask = jb.wl_mask;
  const double ell = valid ? col.ell : A.cols[tile0].ell;
  const EllC lc = ell_consts(ell);

  int slot = 0;
  if constexpr (STAGED) {
    const int j2 = starts ? __shfl(job, (int)__builtin_ctzll(starts), 64) : first;
    const int c1 = A.jobs[first].c, c2 = A.jobs[j2].c;
    const int n = (starts ? 2 : 1) * NEHQ;
    for (int e = threadIdx.x; e < n; e += blockDim.x)
      lds4[e] = A.ehtab[(size_t)(e < NEHQ ? c1 : c2) * NEHQ + (e < NEHQ ? e : e - NEHQ)];
    slot = job == first ? 0 : NEHQ;
    __syncthreads();
  }
  const float4* gtab = A.ehtab + (size_t)c * NEHQ;
  const float* cr = A.crec + (size_t)c * NCREC;
  const double invh = (double)cr[CR_EHINVH];
  const double lo_invh = (double)cr[CR_EHLO] * invh;  // exact in FP64
  // this wave's share of the a range
  const int chunk = (A.Na + nsl * nwv - 1) / (nsl * nwv);
  const int m0 = min(A.Na, ((int)blockIdx.y * nwv + wave) * chunk);
  const int m1 = min(A.Na, m0 + chunk);

  float acc[P];
#pragma unroll
  for (int q = 0; q < P; ++q) acc[q] = 0.0f;

  // one (column, a-node) point: limber_point_tab with the table row from
  // LDS or global memory, then the contraction
  auto point = [&](const PointCoefs<float, 1>& pc, const float* rp) {
    const float k = lc.lph * pc.invchi;
    const float invk = lc.ilph * pc.chi1;
    const float lnk = lc.lnlph + pc.lninvchi;
    const double u = __builtin_fma((double)lnk, invh, -lo_invh);
    int q = (int)u;
    q = q < 0 ? 0 : (q > NEHQ - 1 ? NEHQ - 1 : q);
    const float t = (float)(u - (double)q);
    // the table row first: the tracer kernels below are needed only after P_nl
    const float4 cc = STAGED ? lds4[slot + q] : gtab[q];
    float Rv[T];
#pragma unroll
    for (int i = 0; i < T; ++i) Rv[i] = rp[i];
    const float tn = __builtin_fmaf(__builtin_fmaf(__builtin_fmaf(cc.w, t, cc.z), t, cc.y), t, cc.x);
    const float plin = (tn * tn) * pc.g2p;
    const float k3 = k * k * k;
    const float ik3 = invk * invk * invk;
    const float lgy = __builtin_fmaf(lnk, 1.4426950408889634f, pc.lgRnl);
    const float y = k * pc.Rnl;
    const float iy = invk * pc.knl;
    const float M = pc.wgt * halofit_pnl(pc.hf, lgy, y, iy, k3, ik3, plin);
    int p = 0;
#pragma unroll
    for (int i = 0; i < T; ++i) {
      const float uu = Rv[i] * M;
#pragma unroll
      for (int j = i; j < T; ++j) {
        acc[p] += Rv[j] * uu;
        ++p;
      }
    }
  };
  const float* cf = A.coef + ((size_t)c * A.Na + m0) * NCOEF;
  const float* rr = A.R + ((size_t)c * A.Na + m0) * A.Tpad;
  {
    // two register sets in turn: the next node's coefficients load while
    // the current node computes
    const int RS = A.Tpad;
    PointCoefs<float, 1> pa, pb;
    if (m0 < m1) pa = load_coefs<float, 1>(cf);
    for (int m = m0; m < m1; m += 2, cf += 2 * NCOEF, rr += 2 * RS) {
      const bool has_b = m + 1 < m1;
      if (has_b) pb = load_coefs<float, 1>(cf + NCOEF);
      point(pa, rr);
      if (!has_b) break;
      if (m + 2 < m1) pa = load_coefs<float, 1>(cf + 2 * NCOEF);
      point(pb, rr + RS);
    }
  }
  // fixed-order reduction over the waves, RCH accumulators per pass (the
  // LDS of the tables is reused)
  if (nwv > 1) {
#pragma unroll
    for (int q0 = 0; q0 < P; q0 += RCH) {
      __syncthreads();
      if (wave > 0) {
#pragma unroll
        for (int q = q0; q < q0 + RCH && q < P; ++q) red[((wave - 1) * RCH + (q - q0)) * 64 + lane] = acc[q];
      }
      __syncthreads();
      if (wave == 0) {
        for (int w = 1; w < nwv; ++w) {
#pragma unroll
          for (int q = q0; q < q0 + RCH && q < P; ++q) acc[q] += red[((w - 1) * RCH + (q - q0)) * 64 + lane];
        }
      }
    }
  }
  if (wave != 0 || !valid) return;
  const float lf = (float)(::sqrt((ell - 1.0) * ell * (ell + 1.0) * (ell + 2.0)) / ((ell + 0.5) * (ell + 0.5)));
  float* out = nsl > 1 ? A.part + (size_t)blockIdx.y * A.part_stride : A.out;
  int p = 0;
#pragma unroll
  for (int i = 0; i < T; ++i) {
    const float Fi = ((wl_mask >> i) & 1u) ? lf : 1.0f;
#pragma unroll
    for (int j = i; j < T; ++j) {
      const float Fj = ((wl_mask >> j) & 1u) ? lf : 1.0f;
      out[jb.out_off + (int64_t)p * jb.n_ell + col.l] = acc[p] * Fi * Fj;
      ++p;
    }
  }
}

__global__ __launch_bounds__(256, 4) void k_limber_ragged_fwd(RaggedArgs A) {
  extern __shared__ float4 lds4[];
  const int ncols = *A.ncols;
  // live tiles dealt to the 8 XCDs in contiguous ranges, as k_limber_ragged_adj
  const int live = (ncols + 63) >> 6;
  const int nsl = ragged_slices(live, A.waves, A.Na, A.max_slices);
  if ((int)blockIdx.y >= nsl) return;
  const int xq = live >> 3, xr = live & 7, xx = blockIdx.x & 7, xi = blockIdx.x >> 3;
  if (xi >= xq + (xx < xr ? 1 : 0)) return;
  const int tile0 = (xx * xq + min(xx, xr) + xi) * 64;
  const int first = A.cols[tile0].job;
  const int T = __builtin_amdgcn_readfirstlane(A.jobs[first].T);
  // the requests of the tile: a new one starts at each lane whose job
  // differs from the previous lane's (the same in every wave of the block)
  const int lane = threadIdx.x & 63;
  const int cj = A.cols[tile0 + lane].job;
  const int job = cj >= 0 ? cj : first;
  const int prev = __shfl_up(job, 1, 64);
  const uint64_t starts = __ballot(lane > 0 && job != prev);
  if (__popcll(starts) <= 1) {
    switch (T) {
      case 1: ragged_fwd_tile<1, true>(A, tile0, nsl, lds4, starts); break;
      case 2: ragged_fwd_tile<2, true>(A, tile0, nsl, lds4, starts); break;
      case 3: ragged_fwd_tile<3, true>(A, tile0, nsl, lds4, starts); break;
      case 4: ragged_fwd_tile<4, true>(A, tile0, nsl, lds4, starts); break;
      case 5: ragged_fwd_tile<5, true>(A, tile0, nsl, lds4, starts); break;
      case 6: ragged_fwd_tile<6, true>(A, tile0, nsl, lds4, starts); break;
      case 7: ragged_fwd_tile<7, true>(A, tile0, nsl, lds4, starts); break;
      case 8: ragged_fwd_tile<8, true>(A, tile0, nsl, lds4, starts); break;
      case 9: ragged_fwd_tile<9, true>(A, tile0, nsl, lds4, starts); break;
      default: ragged_fwd_tile<10, true>(A, tile0, nsl, lds4, starts); break;
    }
  } else {
    switch (T) {
      case 1: ragged_fwd_tile<1, false>(A, tile0, nsl, lds4, starts); break;
      case 2: ragged_fwd_tile<2, false>(A, tile0, nsl, lds4, starts); break;
      case 3: ragged_fwd_tile<3, false>(A, tile0, nsl, lds4, starts); break;
      case 4: ragged_fwd_tile<4, false>(A, tile0, nsl, lds4, starts); break;
      case 5: ragged_fwd_tile<5, false>(A, tile0, nsl, lds4, starts); break;
      case 6: ragged_fwd_tile<6, false>(A, tile0, nsl, lds4, starts); break;
      case 7: ragged_fwd_tile<7, false>(A, tile0, nsl, lds4, starts); break;
      case 8: ragged_fwd_tile<8, false>(A, tile0, nsl, lds4, starts); break;
      case 9: ragged_fwd_tile<9, false>(A, tile0, nsl, lds4, starts); break;
      default: ragged_fwd_tile<10, false>(A, tile0, nsl, lds4, starts); break;
    }
  }
}

// The adjoint: per column, the 7 parameter cotangents of the request's
// C_ell against its cotangent block, with the tile's tracer count as a
// constant (one body per T, chosen by a block-uniform switch): cotangent and
// tracer-kernel arrays of the request's own size and no per-pair tests, at 2
// waves per SIMD (one body for every T with per-pair tests needed 1).  The
// transfer table and its theta derivatives are read through the L1/L2, since
// one tile may span several requests.
template <int T>
__device__ __forceinline__ void ragged_adj_tile(const RaggedArgs& A, const int tile0, const int nsl, float* red) {
  using Sf = typename Types<7>::F;
  constexpr int W = 8;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int nwv = A.waves;
  const RCol col = A.cols[tile0 + lane];
  const int first = A.cols[tile0].job;  // a tile starts with a real column
  const bool valid = col.job >= 0;
  const RJob& jb = A.jobs[valid ? col.job : first];
  const int c = jb.c;
  const double ell = valid ? col.ell : A.cols[tile0].ell;
  const EllC lc = ell_consts(ell);
  const float lf = (float)(::sqrt((ell - 1.0) * ell * (ell + 1.0) * (ell + 2.0)) / ((ell + 0.5) * (ell + 0.5)));
  float F[T];
#pragma unroll
  for (int i = 0; i < T; ++i) F[i] = ((jb.wl_mask >> i) & 1u) ? lf : 1.0f;
  const float4* tab = A.ehtab + (size_t)c * NEHQ;
  const float* dtab = A.ehtab7 + (size_t)c * (NEHQ + 1) * NEHD_ROW;
  const float* cr = A.crec + (size_t)c * NCREC * W;
  const double invh = (double)cr[CR_EHINVH * W];
  const double lo_invh = (double)cr[CR_EHLO * W] * invh;  // exact in FP64
  const int chunk = (A.Na + nsl * nwv - 1) / (nsl * nwv);
  const int m0 = min(A.Na, ((int)blockIdx.y * nwv + wave) * chunk);
  const int m1 = min(A.Na, m0 + chunk);

  float acc[7];
#pragma unroll
  for (int q = 0; q < 7; ++q) acc[q] = 0.0f;
  // symmetric cotangent matrix (diagonal doubled) times F_i F_j
  float ct[T][T];
  {
    int p = 0;
#pragma unroll
    for (int i = 0; i < T; ++i) {
#pragma unroll
      for (int j = i; j < T; ++j) {
        const float v = valid ? A.cbar[jb.cbar_off + (int64_t)p * jb.n_ell + col.l] * F[i] * F[j] : 0.0f;
        ct[i][j] = (i == j) ? 2.0f * v : v;
        ct[j][i] = ct[i][j];
        ++p;
      }
    }
  }
  const float* cf = A.coef + ((size_t)c * A.Na + m0) * NCOEF * W;
  const float* rr = A.R + ((size_t)c * A.Na + m0) * A.Tpad * W;
  for (int m = m0; m < m1; ++m, cf += NCOEF * W, rr += A.Tpad * W) {
    const PointCoefs<Sf, W> pc = load_coefs<Sf, W>(cf);
    Sf Rv[T];
#pragma unroll
    for (int i = 0; i < T; ++i) Rv[i] = Mem<Sf>::ld(rr + i * W);
    float Mv, Md[7];
    limber_point_adj<W>(adj_tn_tab<W>(tab, dtab, invh, lo_invh, pc, lc), pc, lc, Mv, Md);
    float V[T];
    float G = 0.0f;
#pragma unroll
    for (int i = 0; i < T; ++i) {
      float sv = 0.0f;
#pragma unroll
      for (int j = 0; j < T; ++j) sv += ct[i][j] * val(Rv[j]);
      V[i] = sv;
      G += 0.5f * val(Rv[i]) * sv;
    }
#pragma unroll
    for (int d = 0; d < 7; ++d) {
      float sd = 0.0f;
#pragma unroll
      for (int i = 0; i < T; ++i) sd += Rv[i].d[d] * V[i];
      acc[d] += Mv * sd + G * Md[d];
    }
  }
  if (nwv > 1) {
    if (wave > 0) {
#pragma unroll
      for (int q = 0; q < 7; ++q) red[((wave - 1) * 7 + q) * 64 + lane] = acc[q];
    }
    __syncthreads();
    if (wave == 0) {
      for (int w = 1; w < nwv; ++w) {
#pragma unroll
        for (int q = 0; q < 7; ++q) acc[q] += red[((w - 1) * 7 + q) * 64 + lane];
      }
    }
  }
  if (wave != 0) return;
  float* tc = A.theta_col + (size_t)blockIdx.y * A.part_stride;
#pragma unroll
  for (int d = 0; d < 7; ++d) tc[(size_t)(tile0 + lane) * 7 + d] = valid ? acc[d] : 0.0f;
}

__global__ __launch_bounds__(256, 2) void k_limber_ragged_adj(RaggedArgs A) {
  extern __shared__ float red[];
  const int ncols = *A.ncols;
  // The grid is sized for capacity: the live tiles are dealt to the 8 XCDs
  // (block b runs on XCD b % 8) in contiguous ranges, so neighbouring tiles
  // (one request's columns) share an L2, and blocks past each XCD's range exit.
  const int live = (ncols + 63) >> 6;
  const int nsl = ragged_slices(live, A.waves, A.Na, A.max_slices);
  if ((int)blockIdx.y >= nsl) return;
  const int xq = live >> 3, xr = live & 7, xx = blockIdx.x & 7, xi = blockIdx.x >> 3;
  if (xi >= xq + (xx < xr ? 1 : 0)) return;
  const int tile0 = (xx * xq + min(xx, xr) + xi) * 64;
  const int T = __builtin_amdgcn_readfirstlane(A.jobs[A.cols[tile0].job].T);
  switch (T) {
    case 1: ragged_adj_tile<1>(A, tile0, nsl, red); break;
    case 2: ragged_adj_tile<2>(A, tile0, nsl, red); break;
    case 3: ragged_adj_tile<3>(A, tile0, nsl, red); break;
    case 4: ragged_adj_tile<4>(A, tile0, nsl, red); break;
    case 5: ragged_adj_tile<5>(A, tile0, nsl, red); break;
    case 6: ragged_adj_tile<6>(A, tile0, nsl, red); break;
    case 7: ragged_adj_tile<7>(A, tile0, nsl, red); break;
    case 8: ragged_adj_tile<8>(A, tile0, nsl, red); break;
    case 9: ragged_adj_tile<9>(A, tile0, nsl, red); break;
    default: ragged_adj_tile<10>(A, tile0, nsl, red); break;
  }
}

hipError_t launch_limber_ragged(int nd, int max_cols, const RaggedArgs& a, hipStream_t s) {
  if (a.waves < 1 || a.waves > 4 || max_cols < 64 || (max_cols % 64) != 0) return hipErrorInvalidValue;
  if (a.max_slices < 1 || a.max_slices > 64 || !a.ehtab || (nd == 7 && !a.ehtab7)) return hipErrorInvalidValue;
  if (nd == 0 && a.max_slices > 1 && !a.part) return hipErrorInvalidValue;
  const dim3 grid(max_cols / 64, a.max_slices);
  if (nd == 0) {
    // two transfer tables, or the reduction's [waves - 1][32][64] partials
    const size_t sm = std::max((size_t)2 * NEHQ * sizeof(float4), (size_t)(a.waves - 1) * 32 * 64 * sizeof(float));
    hipLaunchKernelGGL(k_limber_ragged_fwd, grid, dim3(64 * a.waves), sm, s, a);
  } else if (nd == 7) {
    const size_t sm = (size_t)(a.waves - 1) * 7 * 64 * sizeof(float);
    hipLaunchKernelGGL(k_limber_ragged_adj, grid, dim3(64 * a.waves), sm, s, a);
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// out[o] = part[0][o] + part[1][o] + ... over the forward launch's live
// slices (in order), for o in [range[0], range[1]); nothing with one slice
// (the Limber kernel then wrote out itself)
__global__ void k_ragged_slice_sum(const int* __restrict__ ncols, int waves, int Na, int max_slices,
                                   const float* __restrict__ part, int64_t stride, const int* __restrict__ range,
                                   float* __restrict__ out) {
  const int nsl = ragged_slices((*ncols + 63) >> 6, waves, Na, max_slices);
  if (nsl <= 1) return;
  const int64_t o = range[0] + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= range[1]) return;
  float v = part[o];
  for (int k = 1; k < nsl; ++k) v += part[(size_t)k * stride + o];
  out[o] = v;
}

hipError_t launch_ragged_slice_sum(const int* ncols, int waves, int Na, int max_slices, const float* part,
                                   int64_t part_stride, const int* range, int64_t max_out, float* out,
                                   hipStream_t s) {
  if (max_out <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_ragged_slice_sum, dim3((unsigned)((max_out + 255) / 256)), dim3(256), 0, s, ncols, waves,
                     Na, max_slices, part, part_stride, range, out);
  return hipGetLastError();
}

// theta[row(j)][d] = sum over the columns of request j and the adjoint
// launch's live slices (FP64).  One 256-thread block per request (the grid
// is sized for capacity): thread t adds the elements t, t + 256, ... of the
// request's [n_ell][7] column partials over the slices (coalesced; element e
// belongs to parameter e % 7, and 256 = 36 * 7 + 4 moves it by four), each
// wave adds its lanes by a butterfly and thread q < 7 adds the four waves'
// sums of parameter q in a fixed order.
__global__ __launch_bounds__(256) void k_ragged_reduce(RReduceArgs A) {
  __shared__ double wsum[4][8];
  const int j = blockIdx.x;
  if (j >= *A.njobs) return;  // uniform across the block
  const int nsl = ragged_slices((*A.ncols + 63) >> 6, A.waves, A.Na, A.max_slices);
  const RJob jb = A.jobs[j];
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const int n = jb.n_ell * 7;
  const float* p = A.theta_col + (size_t)jb.col0 * 7;
  double acc[7];
#pragma unroll
  for (int q = 0; q < 7; ++q) acc[q] = 0.0;
  int d = t % 7;
  for (int e = t; e < n; e += 256) {
    double v = 0.0;
    for (int k = 0; k < nsl; ++k) v += (double)p[(size_t)k * A.part_stride + e];
#pragma unroll
    for (int q = 0; q < 7; ++q) acc[q] += (q == d) ? v : 0.0;
    d = (d >= 3) ? d - 3 : d + 4;
  }
#pragma unroll
  for (int q = 0; q < 7; ++q) {
    double v = acc[q];
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    if (lane == q) wsum[wave][q] = v;
  }
  __syncthreads();
  // out_off: the request's row in the batch
  if (t < 7) A.theta[jb.out_off * 7 + t] = ((wsum[0][t] + wsum[1][t]) + wsum[2][t]) + wsum[3][t];
}

hipError_t launch_ragged_reduce(int max_jobs, const RReduceArgs& a, hipStream_t s) {
  if (max_jobs <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_ragged_reduce, dim3(max_jobs), dim3(256), 0, s, a);
  return hipGetLastError();
}

}  // namespace jc
