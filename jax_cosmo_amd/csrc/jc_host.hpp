// Host-side internals shared by the C ABI translation units: device
// buffers, registered tracer sets with their per-grid resident tables, and
// the context.
#pragma once
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/jc_limber.h"
#include "jc_kernels.hpp"

namespace jch {

int fail(int code, const char* fmt, ...);
extern thread_local std::string g_err;  // message of the last failure on this thread

#define HIPCHK(x)                                                                   \
  do {                                                                              \
    hipError_t e_ = (x);                                                            \
    if (e_ != hipSuccess)                                                           \
      return fail(e_ == hipErrorOutOfMemory ? JC_E_NOMEM : JC_E_HIP, "%s: %s (%s:%d)", \
                  #x, hipGetErrorString(e_), __FILE__, __LINE__);                   \
  } while (0)

// Device buffer owned by the context.
template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  hipError_t reserve(size_t count) {
    if (count <= n) return hipSuccess;
    release();
    hipError_t e = hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T));
    if (e == hipSuccess) n = count;
    else p = nullptr;
    return e;
  }
  hipError_t upload(const std::vector<T>& v, hipStream_t s) {
    hipError_t e = reserve(v.size());
    if (e != hipSuccess) return e;
    e = hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(s);
  }
};

// ---------------------------------------------------------------- n(z)
// Un-normalised n(z): redshift.py:61-77 (smail), the Gaussian bin,
// redshift.py:80-105 (fu), with the systematic shift of redshift.py:159-171
// (parent evaluated at max(z - shift, 0)).  delta_nz has no density here.
inline double pz_fn(const jc_tracer_desc& t, double z) {
  if (t.nz_shifted) z = std::max(z - t.nz_params[3], 0.0);
  if (t.nz_kind == JC_NZ_SMAIL) {
    const double a = t.nz_params[0], b = t.nz_params[1], z0 = t.nz_params[2];
    return pow(z, a) * exp(-pow(z / z0, b));
  }
  if (t.nz_kind == JC_NZ_FU) {
    const double a = t.nz_params[0], b = t.nz_params[1], c = t.nz_params[2];
    return (pow(z, a) + pow(z, a * b)) / (pow(z, b) + c);
  }
  if (t.nz_kind == JC_NZ_GAUSS) {
    const double mu = t.nz_params[0], sg = t.nz_params[1];
    const double u = (z - mu) / sg;
    return exp(-0.5 * u * u);
  }
  return 0.0;
}

// Cosmology-independent part of a bias model at z: constant_linear_bias
// and inverse_growth_linear_bias give b (the 1/D(a) is applied on the
// device), des_y1_ia_bias gives A ((1+z)/(1+z0))^eta (bias.py:10-59).
inline double bias_host(const jc_tracer_desc& t, double z) {
  if (t.bias_kind == JC_BIAS_DES_Y1)
    return t.bias_params[0] * pow((1.0 + z) / (1.0 + t.bias_params[2]), t.bias_params[1]);
  return t.bias_params[0];
}

// composite Simpson, integrate.py:162-199
template <class F>
double simps(F f, double a, double b, int N) {
  const double dx = (b - a) / N;
  double s = 0.0;
  for (int i = 0; i <= N; ++i) {
    const double x = (i == N) ? b : a + i * dx;
    const double w = (i == 0 || i == N) ? 1.0 : ((i & 1) ? 4.0 : 2.0);
    s += w * f(x);
  }
  return dx / 3 * s;
}

inline double linspace_at(double a, double b, int num, int i) {
  if (i == num - 1) return b;
  return a + i * ((b - a) / (num - 1));
}

// nearest-node segment of interpolate.py:12-37 on logspace(-3, 0, n):
// returns the lower node L of the segment the reference evaluates.
inline int interp_segment(double x, const std::vector<double>& xp) {
  const int n = (int)xp.size();
  double u = (log10(x) + 3.0) * (n - 1) / 3.0;
  int i0 = (int)floor(u);
  i0 = std::min(std::max(i0, 0), n - 1);
  int lo = std::max(i0 - 1, 0), hi = std::min(i0 + 2, n - 1);
  int ind = i0;
  double best = 1e300;
  for (int j = lo; j <= hi; ++j) {
    const double d = (x - xp[j]) * (x - xp[j]);
    if (d < best) {
      best = d;
      ind = j;
    }
  }
  ind = std::min(std::max(ind, 1), n - 2);
  double xc = std::min(std::max(x, xp[1]), xp[n - 2]);
  const int nb = (xc - xp[ind] < 0.0) ? ind - 1 : ind + 1;
  return std::min(ind, nb);
}

struct WLGroup {
  int nw = 0;
  bool ia = false;  // some bin adds an intrinsic-alignment kernel
  DevBuf<uint8_t> L;
  DevBuf<float> delta, Nw;
  DevBuf<int> col;
  DevBuf<jc::WLSource> src;  // [nw] single-plane (delta_nz) bins
};

struct GridTables {
  int N = 0, Na = 0;
  DevBuf<double> a_list, a_wgt, loc_nb, a_ln, a_l10;
  DevBuf<int> loc_col, loc_mode;
  DevBuf<float> zp1;
  int nLoc = 0;
  std::vector<std::unique_ptr<WLGroup>> wl;
};

// a kde_nz catalog, sorted by redshift, resident on the device
struct Catalog {
  DevBuf<double> z, w;
  int64_t n = 0;
  double wsum = 0.0;
};

struct TracerSet {
  std::vector<jc_tracer_desc> tr;
  std::vector<double> norm;
  std::vector<std::unique_ptr<Catalog>> cat;  // per bin, null unless kde_nz
  std::vector<std::pair<jc_pz_fn, void*>> host;  // per bin, null unless JC_NZ_HOST
  int T = 0, Tpad = 0;
  uint32_t wl_mask = 0;
  std::map<int, std::unique_ptr<GridTables>> grids;
};

int build_grid(jc_ctx* ctx, TracerSet* ts, int N, GridTables** out);
// un-normalised n(z) of bin i at the points z (the KDE on the device)
int eval_pz(jc_ctx* ctx, TracerSet* ts, int i, const std::vector<double>& z, std::vector<double>& out);

}  // namespace jch

struct jc_ctx {
  using TracerSet = jch::TracerSet;
  template <class T> using DevBuf = jch::DevBuf<T>;
  int device = 0;
  unsigned char* scratch = nullptr;  // this device's scratch array (jc_angular_cl_noalloc)
  hipStream_t own = nullptr;
  hipStream_t stream = nullptr;
  std::vector<double> atab_chi_h, atab_g_h;
  DevBuf<double> atab_chi, atab_g, hf_lnk, hf_w, hf_E, logr;
  DevBuf<float> hf_k32;
  DevBuf<double> hf_kp;
  DevBuf<double> chi_nodes, g_nodes, gg_nodes, romb, hf_kd;
  std::map<int64_t, std::unique_ptr<TracerSet>> sets;
  int64_t next_set = 1;
  // workspace
  DevBuf<double> params, seeds, ell, theta, dbg, pk_k, pk_a, pk_out;
  // likelihood workspace: inputs staged from host, mean, cotangent, partials
  DevBuf<double> lk_cosmo, lk_ell, lk_data, lk_prec, lk_part, lk_out, lk_grad, lk_piv;
  DevBuf<float> lk_mu, lk_s;
  // dense likelihood workspace: staged matrix, log L_ii, non-SPD flag, [B][N] solves
  DevBuf<double> dn_mat, dn_dlog, dn_z;
  DevBuf<int> dn_bad;
  DevBuf<float> chitab, coef, crec, R, out, out_t, cbar, theta_col, part;
  DevBuf<float4> ehtab;
  DevBuf<float> ehtab7;
  int waves = 0, wl_ctile = 0, tables_variant = 0;
  int num_cu = 256;        // compute units of the device (set at creation)
  int transfer_table = 1;  // forward calls read T(k) from a per-cosmology table
  int transfer_model = 0;  // 0: eisenhu_osc (transfer.py default), 1: eisenhu
  int halofit_model = 0;   // 0: takahashi2012 (power.py default), 1: smith2003
  int growth_model = 0;    // 0: growth ODE (Cosmology without gamma), 1: gamma growth
  double gamma = 0.0;      // growth index when growth_model == 1
  int matrix_cores = 0;    // forward calls contract the pairs on the matrix cores (measured slower; off)
  // optional per-kernel timing with HIP events on the launch stream
  bool prof = false;
  unsigned prof_kernels = ~0u;  // bit k: launches of kernel k are timed
  std::vector<hipEvent_t> ev_pool;
  struct Rec { int kind; hipEvent_t a, b; };
  std::vector<Rec> recs;
};

namespace jch {
// the context-resident, cosmology-independent table pointers of a TablesArgs
inline void resident_tables(const jc_ctx* ctx, jc::TablesArgs& ta) {
  ta.atab_chi = ctx->atab_chi.p;
  ta.atab_g = ctx->atab_g.p;
  ta.hf_lnk = ctx->hf_lnk.p;
  ta.hf_w = ctx->hf_w.p;
  ta.hf_k32 = ctx->hf_k32.p;
  ta.hf_kp = ctx->hf_kp.p;
  ta.hf_E = ctx->hf_E.p;
  ta.logr = ctx->logr.p;
  ta.chi_nodes = ctx->chi_nodes.p;
  ta.g_nodes = ctx->g_nodes.p;
  ta.gg_nodes = ctx->gg_nodes.p;
  ta.romb = ctx->romb.p;
  ta.hf_kd = ctx->hf_kd.p;
  ta.transfer_model = ctx->transfer_model;
  ta.halofit_model = ctx->halofit_model;
  ta.growth_model = ctx->growth_model;
  ta.gamma = ctx->gamma;
}
}  // namespace jch
