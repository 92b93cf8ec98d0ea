// Device-side readers of the per-cosmology (crec) and per-(cosmology, a)
// (coef) records that the table kernel writes and the Limber and transfer
// table kernels consume.
#pragma once
#include "jc_kernels.hpp"
#include "jc_physics.hpp"

namespace jc {

template <class Sf, int W>
struct PointCoefs {
  Sf invchi, chi1, lninvchi, wgt, g2p, lgRnl, Rnl, knl;
  HFa<Sf> hf;
};

template <class Sf, int W>
JC_D PointCoefs<Sf, W> load_coefs(const float* cf) {
  PointCoefs<Sf, W> r;
  r.invchi = Mem<Sf>::ld(cf + CF_INV_CHI * W);
  r.chi1 = Mem<Sf>::ld(cf + CF_CHI1 * W);
  r.lninvchi = Mem<Sf>::ld(cf + CF_LNINVCHI * W);
  r.wgt = Mem<Sf>::ld(cf + CF_WGT * W);
  r.g2p = Mem<Sf>::ld(cf + CF_G2P * W);
  r.lgRnl = Mem<Sf>::ld(cf + CF_LGRNL * W);
  r.Rnl = Mem<Sf>::ld(cf + CF_RNL * W);
  r.knl = Mem<Sf>::ld(cf + CF_KNL * W);
  r.hf.a_n = Mem<Sf>::ld(cf + CF_AN * W);
  r.hf.b_n = Mem<Sf>::ld(cf + CF_BN * W);
  r.hf.lg_cf3 = Mem<Sf>::ld(cf + CF_LGCF3 * W);
  r.hf.gam3 = Mem<Sf>::ld(cf + CF_GAM3 * W);
  r.hf.alpha_n = Mem<Sf>::ld(cf + CF_ALPHA * W);
  r.hf.beta_n = Mem<Sf>::ld(cf + CF_BETA * W);
  r.hf.nu_n = Mem<Sf>::ld(cf + CF_NU * W);
  r.hf.f1x3 = Mem<Sf>::ld(cf + CF_F1X3 * W);
  r.hf.f2 = Mem<Sf>::ld(cf + CF_F2 * W);
  r.hf.mu_n = Mem<Sf>::ld(cf + CF_MU * W);
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
  e.bnode_shd = Mem<Sf>::ld(cr + CR_BNODE_SHD * W);
  e.bb_shd = Mem<Sf>::ld(cr + CR_BB_SHD * W);
  e.alpha_g = Mem<Sf>::ld(cr + CR_ALPHAG * W);
  e.tq = Mem<Sf>::ld(cr + CR_TQ * W);
  e.nw = cr[CR_NW * W] != 0.0f;
  ns = Mem<Sf>::ld(cr + CR_NS * W);
  return e;
}

}  // namespace jc
