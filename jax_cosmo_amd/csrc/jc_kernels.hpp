// Kernel argument blocks and host-side launchers shared by the translation
// units of libjc_limber.so.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace jc {

// Fixed table sizes of the reference discretisation.
constexpr int NCHI = 256;     // chi table nodes, background.py:199 (steps=256)
constexpr int NGROW = 128;    // growth table nodes, background.py:443 (steps=128)
constexpr int NHF_K = 257;    // halofit ln k Simpson nodes, power.py:111 (N=256)
constexpr int NHF_R = 256;    // halofit ln R grid, power.py:92
constexpr int NROMB = 129;    // sigma_8 Romberg samples, power.py:76 (divmax=7)
constexpr int NWL_J = 257;    // per-z weak-lensing Simpson nodes, probes.py:51
constexpr int MAXT = 10;      // tracers per request handled by the Limber kernel
constexpr int NWL_MAX = 10;   // weak-lensing bins per kernel launch

// constants.py:23-27: intrinsic-alignment normalisation C_1 and the critical
// density [h^2 M_sun Mpc^-3]
constexpr double C1_IA = 5.0e-14;
constexpr double RHOCRIT = 2.7750e11;

// Per (cosmology, a-node) record consumed by the Limber kernel, in floats
// (each slot holds 1 + ND floats for the derivative kernels).
enum CoefSlot {
  CF_INV_CHI = 0,  // 1 / max(chi, 1)     (k = (ell + 1/2) / max(chi, 1))
  CF_CHI1,         // max(chi, 1)         (1/k)
  CF_LNINVCHI,     // -ln max(chi, 1)     (ln k)
  CF_WGT,          // Simpson weight * dchi/da / max(chi^2, 1) / c^2
  CF_G2P,          // D(a)^2 * pknorm
  CF_LGRNL,        // log2 R_nl = -log2 k_nl
  CF_RNL,          // R_nl
  CF_KNL,          // k_nl
  CF_AN, CF_BN, CF_LGCF3, CF_GAM3, CF_ALPHA, CF_BETA, CF_NU, CF_F1X3, CF_F2,
  CF_CHI,          // chi(a)
  CF_GROWTH,       // D(a)
  CF_MU,           // halofit mu_n (0 for takahashi2012)
  NCOEF
};
// Per-cosmology record (EH98 constants and the lensing prefactor).
enum CrecSlot {
  CR_FB = 0, CR_FC, CR_INV13KEQ, CR_BETAC18, CR_C14A, CR_SHD, CR_ALPHAB, CR_BETAB,
  CR_BETANODE, CR_LNKSILK, CR_NS, CR_WLCONST, CR_PKNORM, CR_LNINV13, CR_BNODE_SHD, CR_BB_SHD,
  CR_EHLO,    // ln k at the first node of the transfer table
  CR_EHINVH,  // 1 / (its ln k spacing)
  CR_ALPHAG, CR_TQ, CR_NW,  // the no-wiggle transfer form
  NCREC_USED
};
constexpr int NCREC = 21;
// Intervals of the per-cosmology transfer table: cubic Hermite pieces of
// T(k) k^(n_s/2) uniform in ln k over the k range of one forward call (16 KB
// per cosmology; interpolation error below 1e-6 of T).
constexpr int NEHQ = 1023;
// The adjoint's table of theta derivatives has a row of 8 floats (7
// derivatives, one pad) at each of the NEHQ + 1 nodes of the same grid (32 KB
// per cosmology, read through the L2); 4-point Lagrange interpolation.
constexpr int NEHD_ROW = 8;

struct RJob;
struct RCol;

struct TablesArgs {
  const double* params;   // [B][8] Cosmology(...) parameters
  const double* seeds;    // [B][8][ND] d param / d direction (ND > 0)
  int B, Na;
  const double* a_list;   // [Na] scale factors
  const double* a_wgt;    // [Na] quadrature weight of each node (0 = evaluation only)
  const double* atab_chi; // [256]
  const double* atab_g;   // [128]
  const double* hf_lnk;   // [257]
  const double* hf_w;     // [257] Simpson weights incl. dx/3
  const float* hf_k32;    // [257] exp(hf_lnk) as float
  const double* hf_kp;    // [257][4] k^2, k^4, k^6, 0 at k = exp(hf_lnk)
  const double* hf_E;     // [257][256] exp(-(k_i R_j)^2)
  const double* logr;     // [256]
  // cosmology-independent node quantities (host-computed once per context)
  const double* chi_nodes;  // [256][8]: a_i, ln a_i, 1/a_i, a_mid, ln a_mid, 1/a_mid, h_i (chi RK4 in ln a)
  const double* g_nodes;    // [128][8]: x_i, ln x_i, 1/x_i, x_mid, ln x_mid, 1/x_mid (growth RK4 in a)
  const double* gg_nodes;   // [128][8]: as chi_nodes on the growth grid (gamma growth RK4 in ln a)
  const double* romb;       // [129][3]: k, 1/k, k (k W(8k))^2 at the sigma_8 Romberg nodes
  const double* hf_kd;      // [257] exp(hf_lnk)
  const double* a_ln;       // [Na] ln a of the a-grid, or null (computed in the kernel)
  const double* a_l10;      // [Na] log10 a of the a-grid, or null
  // kernels local in z: number counts n b H (probes.py:77-99) and the NLA
  // intrinsic-alignment terms of weak-lensing bins (probes.py:102-129)
  int nLoc;
  const double* loc_nb;   // [nLoc][Na] cosmology-independent factor at each node
  const int* loc_col;     // [nLoc] tracer column
  const int* loc_mode;    // [nLoc] LOC_* flags
  int Tpad;
  float* chitab;          // [B][256][2][W] (chi, slope of segment i)
  float* coef;            // [B][Na][NCOEF][W]
  float* crec;            // [B][NCREC][W]
  float* R;               // [B][Na][Tpad][W]
  double* dbg;            // optional [B][DBG_STRIDE] FP64 intermediates
  int dbg_stride;
  const int* dB;          // optional device count: blocks >= *dB exit (grid sized for B)
  int transfer_model;     // 0: eisenhu_osc, 1: eisenhu (no wiggles)
  int halofit_model;      // 0: takahashi2012, 1: smith2003
  int growth_model;       // 0: growth ODE (background.py:443-487), 1: f = Omega_m(a)^gamma (background.py:515-582)
  double gamma;           // growth index of model 1
  // optional (forward only): build the transfer table at the end of the
  // table kernel over the ln k range of these multipoles
  float4* ehtab;          // [B][NEHQ] or null
  float* ehtab7;          // adjoint calls: [B][NEHQ + 1][8] theta derivatives at the table's nodes, or null
  const double* ell;      // [n_ell]
  int n_ell;
  // served batches: cosmology c's multipoles are instead the ell of columns
  // rjobs[c].col0 .. rjobs[c].col0 + rjobs[c].n_ell - 1 of rcols, and row
  // c of R holds request c's tracers in its own order (column i = pool
  // tracer rjobs[c].sel[i]; pool tracers it does not select are not written)
  const RJob* rjobs;
  const RCol* rcols;
};


// flags of a z-local kernel: times H(a) always; LOC_INVD: / D(a)
// (inverse_growth_linear_bias); LOC_IA: * Omega_m / D(a) (NLA normalisation)
enum LocMode { LOC_INVD = 1, LOC_IA = 2 };

// a single source plane (delta_nz, probes.py:53-64): chi-table segment and
// offset of its scale factor, and (1 + m)
struct WLSource {
  int32_t on;
  int32_t L;
  float off;
  float fac;
};

struct WLArgs {
  int B, Na, nw;
  int ia;                 // add to the z-local (intrinsic-alignment) values already in R
  const WLSource* src;    // [nw]
  const uint8_t* Lidx;    // [257][Na] chi-table segment of z'_{m,j}
  const float* delta;     // [257][Na] a'_{m,j} - atab[L]
  const float* Nw;        // [5 or NWL_MAX][257][Na] (dx_m/3) w_j n_i(z'_{m,j}) (1 + m_i), zero rows past nw
  const float* zp1;       // [Na] 1 + z_m
  const int* col;         // [nw] tracer column
  const float* chitab;
  const float* coef;
  const float* crec;
  float* R;
  int Tpad;
  const int* dB;          // optional device count of cosmologies (grid sized for B)
  // optional (served batches): row c of R in request c's tracer order, as
  // TablesArgs::rjobs
  const RJob* rjobs;
};

struct LimberArgs {
  int B, Na, n_ell, ncols;  // ncols = B * n_ell
  const double* ell;        // [n_ell]
  const float* coef;
  const float* crec;
  const float* R;
  int Tpad;
  uint32_t wl_mask;         // bit i set: tracer i is weak lensing (ell factor)
  float* out;               // [B][P][n_ell]
  float* out_t;             // [B][P][n_ell] tangent (JVP)
  const float* cbar;        // [B][P][n_ell] cotangent (VJP)
  float* theta_col;         // [ncols][7] per-column parameter cotangent (VJP)
  int waves;                // waves per block splitting the a range
  const float4* ehtab;      // [B][NEHQ] transfer table, or null: evaluate EH98 per point
  const float* ehtab7;      // adjoint: [B][NEHQ + 1][8] d/dtheta of the table's function at its nodes
  int mx;                   // forward, n_ell >= 32: contract the pairs on the matrix cores
  // small batches: the a range is also split over `slices` blocks per column
  // group (grid y); each writes its partial outputs, already scaled, to
  // part[slice][...] in the layout of out (then out_t) or theta_col, and
  // launch_slice_sum adds the slices in order
  int slices;               // 1 = no split
  float* part;
};

struct PkArgs {            // P(k, a) evaluation at (k_i, a_i) pairs
  int n;                   // points for one cosmology
  const double* k;         // [n]
  const float* coef;       // [Na == n][NCOEF] for that cosmology
  const float* crec;       // [NCREC]
  double* plin;            // [n]
  double* pnl;             // [n]
};

// ---------------------------------------------------------------- ragged batches
// One request of a served batch (all requests of a launch share n_z).
struct RJob {
  int32_t c;          // cosmology row in the launch's tables
  int32_t T;          // tracers of the request (1..10)
  int32_t n_ell;
  uint32_t wl_mask;   // bit i: tracer i of the request is weak lensing
  int64_t out_off;    // float offset of the request's [P][n_ell] output block
  int64_t cbar_off;   // float offset of its cotangent block (adjoint)
  int32_t col0;       // first column of the request in the column list
  int32_t sel[MAXT];  // pool tracer column of each of its tracers
  int32_t pad;
};
// One (request, ell) column; job < 0 pads a column tile.
struct RCol {
  int32_t job;
  int32_t l;
  double ell;
};

struct RaggedArgs {
  const int* ncols;      // device: number of columns (a multiple of 64)
  const RCol* cols;
  const RJob* jobs;
  int Na;
  const float* coef;
  const float* crec;
  const float* R;        // row c: request c's tracers in its own order
  int Tpad;
  float* out;
  const float* cbar;
  float* theta_col;      // adjoint: [max_slices][colcap][7] per-slice column partials
  int waves;
  // per-cosmology transfer tables (row c of the launch's tables) and, for
  // the adjoint, their theta derivatives
  const float4* ehtab;   // [rows][NEHQ]
  const float* ehtab7;   // [rows][NEHQ + 1][NEHD_ROW]
  // the a range of each 64-column tile is split over ragged_slices(live
  // tiles) blocks (grid y, at most max_slices); forward partials of slice s
  // go to part + s * part_stride in the layout of out when there is more
  // than one, adjoint partials to theta_col + s * part_stride
  int max_slices;
  float* part;
  int64_t part_stride;
};

struct RReduceArgs {
  const int* njobs;      // device count
  const RJob* jobs;
  const float* theta_col;
  double* theta;         // [njobs][7]
  // slices of the adjoint launch (recomputed from its live column count)
  const int* ncols;
  int waves, Na, max_slices;
  int64_t part_stride;
};

// Slices of a ragged launch with `tiles` live 64-column tiles: enough for
// about 2048 waves, each wave keeping at least 8 a-nodes.
__host__ __device__ inline int ragged_slices(int tiles, int waves, int Na, int max_slices) {
  if (tiles < 1) return 1;
  int s = (2048 + tiles * waves - 1) / (tiles * waves);
  const int smax = Na / (8 * waves);
  s = s < smax ? s : smax;
  s = s < max_slices ? s : max_slices;
  return s < 1 ? 1 : s;
}

// launchers (defined next to their kernels)
hipError_t launch_tables(int nd, const TablesArgs& a, hipStream_t s, int variant = 0);
hipError_t launch_wl(int nd, int ctile, const WLArgs& a, hipStream_t s);
// Gaussian likelihood on block-diagonal covariances (jc_likelihood.hip)
hipError_t launch_block_inv(const double* cov, int P, int n_ell, double* inv, double* piv, hipStream_t s);
hipError_t launch_loglike(const float* mu, int B, int P, int n_ell, const double* data, const double* prec,
                          double c0, float* s_out, double* part, double* loglike, hipStream_t s);
// Gaussian KDE n(z) (redshift.py:126-156) at n points from a catalog sorted
// by redshift: out[p] = sum_k w_k N(z_p - z_k; bw) / sum_k w_k
// growth rate and transverse distance of a background call (FP64 slots 6, 7)
hipError_t launch_background_ex(const double* params, const double* a, int n, const double* atab_g,
                                int growth_model, double gamma, double* dbg, hipStream_t s);
// out[o] = sum over s of part[s * n + o], s = 0..slices-1 in order
hipError_t launch_slice_sum(const float* part, int slices, int64_t n, float* out, hipStream_t s);
hipError_t launch_kde(const double* z, int64_t n, const double* cz, const double* cw, int64_t ncat, double bw,
                      double wsum, double* out, hipStream_t s);
hipError_t launch_limber(int nd, int T, const LimberArgs& a, hipStream_t s);
// theta[B][7] = sum over slices and ell of the per-column partials
// [nsl][B][n_ell][7] (nsl = 1: the partials of an unsliced launch)
hipError_t launch_theta_reduce(const float* theta_col, int nsl, int B, int n_ell, double* theta, hipStream_t s);
hipError_t launch_pk_eval(const PkArgs& a, hipStream_t s);
// grid sized for max_cols columns; blocks beyond *ncols exit at once
hipError_t launch_limber_ragged(int nd, int max_cols, const RaggedArgs& a, hipStream_t s);
hipError_t launch_ragged_reduce(int max_jobs, const RReduceArgs& a, hipStream_t s);
// dense covariances (jc_dense.hip): in-place lower Cholesky factor of the
// row-major N x N matrix A (log L_ii into dlog[N], *bad set on a
// non-positive pivot; the strict upper triangle is zeroed); Z = L^-1 Z
// (transpose 0) or L^-T Z (transpose 1) for B right-hand sides as rows of Z;
// the log-likelihood -1/2 |L^-1 (data - mu_b)|^2 + c0 of B means, with Z as
// [B][N] workspace and dmu (may be null) = C^-1 (data - mu_b) as float
hipError_t launch_dense_cholesky(double* A, int64_t N, double* dlog, int* bad, hipStream_t s);
hipError_t launch_dense_solve(const double* L, int64_t N, double* Z, int64_t B, int transpose, hipStream_t s);
hipError_t launch_dense_loglike(const float* mu, int64_t B, int64_t N, const double* data, const double* L, double c0,
                                double* Z, double* loglike, float* dmu, hipStream_t s);
// out[o] = sum of the forward partials of the launch's live slices, for o in
// [range[0], range[1]) (device values), grid sized for max_out floats
hipError_t launch_ragged_slice_sum(const int* ncols, int waves, int Na, int max_slices, const float* part,
                                   int64_t part_stride, const int* range, int64_t max_out, float* out,
                                   hipStream_t s);

}  // namespace jc
