/* C ABI of libjc_limber.so: MI355X (gfx950) Limber angular power spectra.
 *
 * Drop-in for the hot path of jax_cosmo (references are file:line in the
 * jax_cosmo repository):
 *   angular_cl(cosmo, ell, probes)              jax_cosmo/angular_cl.py:49-98
 *   the C_ell part of gaussian_cl_covariance_and_mean
 *                                               jax_cosmo/angular_cl.py:166-196
 *   jax.grad / jax.jvp of angular_cl w.r.t. the cosmology (the reference's
 *   autodiff through its pytrees, core.py:98-141)
 * Plain C types only; every pointer is owned by the caller.  Every call
 * returns 0 on success or a negative JC_E* code, with a message available
 * from jc_last_error() (per thread).  Calls on one context must be
 * serialised by the caller; distinct contexts may be used concurrently.
 */
#ifndef JC_LIMBER_H
#define JC_LIMBER_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define JC_OK 0
#define JC_E_ARG -1         /* bad argument (shape, value, unknown id)     */
#define JC_E_NOMEM -2       /* device allocation failed                     */
#define JC_E_UNSUPPORTED -3 /* feature outside the supported path           */
#define JC_E_HIP -4         /* HIP runtime error                            */

/* tracer kinds: probes.py:132 (WeakLensing) and probes.py:226 (NumberCounts) */
#define JC_TRACER_WL 0
#define JC_TRACER_NC 1
/* redshift distributions (nz_params): redshift.py:61-77 smail_nz (a, b, z0),
 * a Gaussian bin exp(-(z-mu)^2 / (2 sigma^2)) (mu, sigma), redshift.py:80-105
 * fu_nz (a, b, c), redshift.py:108-123 delta_nz (z0; weak lensing only),
 * redshift.py:126-156 kde_nz (bw; the catalog is passed to
 * jc_tracers_register_catalogs).
 * nz_shifted = 1 evaluates systematic_shift (redshift.py:159-171):
 * n(z) = parent(max(z - nz_params[3], 0)), normalised on its own zmax. */
#define JC_NZ_SMAIL 0
#define JC_NZ_GAUSS 1
#define JC_NZ_FU 2
#define JC_NZ_DELTA 3
#define JC_NZ_KDE 4
/* n(z) of a user-defined redshift_distribution (redshift.py:15-31: a subclass
 * with its own pz_fn), tabulated by a host function (jc_pz_fn) wherever the
 * path needs n(z); registered through jc_tracers_register_host. */
#define JC_NZ_HOST 5
/* bias models (bias_params): bias.py:10-22 constant_linear_bias (b),
 * bias.py:25-40 inverse_growth_linear_bias (b / D(a)), bias.py:43-59
 * des_y1_ia_bias (A, eta, z0: A ((1+z)/(1+z0))^eta).  Number counts use it as
 * the galaxy bias (probes.py:77-99); a weak-lensing bin with a model other
 * than JC_BIAS_NONE adds the NLA intrinsic-alignment kernel with that bias
 * (probes.py:102-129, 188-208). */
#define JC_BIAS_CONST 0
#define JC_BIAS_INV_GROWTH 1
#define JC_BIAS_DES_Y1 2
#define JC_BIAS_NONE 3

/* flags */
#define JC_DEVICE_PTRS 1 /* array arguments are device pointers; the call is
                            asynchronous on the context stream */

typedef struct jc_ctx jc_ctx;

/* One redshift bin of one probe (112 bytes). */
typedef struct jc_tracer_desc {
  int32_t kind;         /* JC_TRACER_WL or JC_TRACER_NC                        */
  int32_t nz_kind;      /* JC_NZ_*                                              */
  int32_t probe;        /* index of the owning probe in the probe list          */
  int32_t bias_kind;    /* JC_BIAS_*: NC galaxy bias; WL intrinsic alignments   */
  double nz_params[4];  /* see JC_NZ_*; [3] = systematic shift                  */
  double zmax;          /* redshift_distribution zmax (redshift.py:16)          */
  double gals_per_arcmin2; /* redshift.py:16, used for the noise only          */
  double sigma_e;       /* WeakLensing sigma_e of this bin (probes.py:219-222)  */
  double m_bias;        /* multiplicative bias of this bin (probes.py:204-207)  */
  double bias_params[3]; /* see JC_BIAS_*                                       */
  int32_t nz_shifted;   /* 1: systematic_shift of the n(z) by nz_params[3]      */
  int32_t reserved;     /* 0                                                    */
} jc_tracer_desc;

/* Context: owns a HIP stream, the resident tables (cosmology-independent
 * quadrature tables, registered tracer sets) and a device workspace.
 * device: HIP device ordinal. */
int jc_ctx_create(int device, jc_ctx** out);
int jc_ctx_destroy(jc_ctx* ctx);
/* Issue work on a caller's stream (e.g. torch.cuda.current_stream()) instead
 * of the context's own; stream = NULL selects the HIP null stream.
 * jc_ctx_reset_stream returns to the context's own stream. */
int jc_ctx_set_stream(jc_ctx* ctx, void* stream);
int jc_ctx_reset_stream(jc_ctx* ctx);
int jc_synchronize(jc_ctx* ctx);

/* Register the redshift bins of a probe list (probes in order, bins in order
 * inside each probe: the tracer order of angular_cl.py:15-25).  The n(z)
 * normalisations (redshift.py:27-31) and the per-grid n(z) tables are
 * computed here once and stay resident.  n <= 10. */
int jc_tracers_register(jc_ctx* ctx, const jc_tracer_desc* tracers, int64_t n, int64_t* set_id);
/* The same with redshift catalogs for JC_NZ_KDE bins: catalogs[i] points to
 * [2][catalog_len[i]] doubles (redshifts, then weights) for bin i and may be
 * NULL for the other bins.  The catalogs are copied; the Gaussian KDE
 * (redshift.py:126-156) is evaluated on the device wherever the path needs
 * n(z). */
int jc_tracers_register_catalogs(jc_ctx* ctx, const jc_tracer_desc* tracers, int64_t n,
                                 const double* const* catalogs, const int64_t* catalog_len,
                                 int64_t* set_id);
/* Host n(z) function of a JC_NZ_HOST bin: pz[k] = the un-normalised n(z[k])
 * for k < n (the bin's pz_fn, redshift.py:23-25); returns 0 on success.  For
 * a bin with nz_shifted = 1, z is already max(z - nz_params[3], 0). */
typedef int (*jc_pz_fn)(void* user, const double* z, double* pz, int64_t n);
/* jc_tracers_register_catalogs plus host n(z) functions: fns[i] and user[i]
 * serve bin i when its nz_kind is JC_NZ_HOST (fns and user may be NULL when no
 * bin is).  The library calls them from the calling thread of this function
 * (the normalisation, redshift.py:27-31) and of any later call that first
 * needs the set's tables at a new Simpson size (jc_angular_cl*, the
 * derivative and likelihood calls, jc_server_create); they must stay callable
 * until jc_tracers_release.  A function that returns nonzero, or a
 * non-finite value, fails that call with JC_E_ARG. */
int jc_tracers_register_host(jc_ctx* ctx, const jc_tracer_desc* tracers, int64_t n,
                             const double* const* catalogs, const int64_t* catalog_len,
                             const jc_pz_fn* fns, void* const* user, int64_t* set_id);
int jc_tracers_release(jc_ctx* ctx, int64_t set_id);

/* C_ell for B cosmologies.
 *   cosmo: [B][8] Omega_c, Omega_b, h, n_s, sigma8, Omega_k, w0, wa
 *          (Cosmology(...), core.py:13)
 *   ell:   [n_ell] multipoles
 *   n_z:   Simpson intervals over a (512 in angular_cl.py:96; even)
 *   cl:    [B][P][n_ell] float32, P = n(n+1)/2 pairs i <= j row-major
 * replaces angular_cl.py:49-98 (vmapped over cosmologies). */
int jc_angular_cl(jc_ctx* ctx, const double* cosmo, int64_t B, int64_t set_id,
                  const double* ell, int64_t n_ell, int64_t n_z, float* cl, int32_t flags);

/* jc_angular_cl without device allocation.  No counterpart in the reference
 * (jax_cosmo allocates through XLA): it is for callers that cannot allocate
 * device memory, such as code under HIP graph capture, tight serving loops
 * and harnesses that forbid hipMalloc.  Same arguments, results and errors as
 * jc_angular_cl; every per-call buffer (staged inputs, the per-cosmology
 * tables and records, transfer tables, partial sums, the staged output) lives
 * in a static device array of JC_NOALLOC_SCRATCH_BYTES per GPU, taken at
 * jc_ctx_create, and the call makes no hipMalloc / hipFree.
 *   - A call whose buffers exceed the array fails with JC_E_ARG and says so
 *     (10 tracers, 100 ell, n_z = 512: about 86 KB per cosmology with device
 *     pointers, so up to ~1500 cosmologies).
 *   - The tracer set's tables for this n_z must be resident: any earlier
 *     jc_angular_cl (or _jvp / _vjp) call with the same set and n_z builds
 *     them; otherwise JC_E_ARG.
 *   - Every call on one GPU shares that GPU's array, across contexts, and
 *     leaves its contents behind: calls that overlap on one GPU race, so the
 *     caller serialises them (one stream, or a synchronisation between).
 *   - With JC_DEVICE_PTRS the call only launches kernels on the context
 *     stream; with host pointers it also copies inputs in and the output out
 *     and returns after the copy. */
#define JC_NOALLOC_SCRATCH_BYTES (128u << 20)
int jc_angular_cl_noalloc(jc_ctx* ctx, const double* cosmo, int64_t B, int64_t set_id,
                          const double* ell, int64_t n_ell, int64_t n_z, float* cl, int32_t flags);

/* Forward-mode derivative: cl and cl_dot = dC/dtheta . tangent, with
 * tangent [B][7] over (Omega_c, Omega_b, h, n_s, sigma8, w0, wa).
 * Replaces jax.jvp / jax.jacfwd of angular_cl (notebook jax-cosmo-intro
 * cells 45-49). */
int jc_angular_cl_jvp(jc_ctx* ctx, const double* cosmo, const double* tangent, int64_t B,
                      int64_t set_id, const double* ell, int64_t n_ell, int64_t n_z,
                      float* cl, float* cl_dot, int32_t flags);

/* Adjoint: theta_bar[B][7] = sum_{p,l} cl_bar[b][p][l] dC[b][p][l]/dtheta.
 * Replaces jax.grad / jax.vjp of angular_cl (core.py:98-141 pytrees). */
int jc_angular_cl_vjp(jc_ctx* ctx, const double* cosmo, int64_t B, int64_t set_id,
                      const double* ell, int64_t n_ell, int64_t n_z, const float* cl_bar,
                      double* theta_bar, int32_t flags);

/* ------------------------------------------------------------------------
 * Gaussian likelihood on block-diagonal covariances, the reference's sparse
 * format [P][P][n_ell] (sparse.py:1-389): entry (p, q, l) couples pair p and
 * pair q at ell_l; data and means are pair-major (index p * n_ell + l, the
 * cl.flatten() of angular_cl.py:196).  JC_DEVICE_PTRS: arrays are device
 * pointers and the call is asynchronous, except jc_sparse_inv, which always
 * returns after the pivots are checked. */

/* Inverse of every P x P block (sparse.py:295-315: inv) and the log of the
 * absolute determinant of the whole matrix (the logdet of sparse.py:322-366:
 * slogdet), FP64, Gauss-Jordan with partial pivoting.  The blocks must be
 * invertible (JC_E_ARG for a singular block). P <= 64.  logdet may be NULL. */
int jc_sparse_inv(jc_ctx* ctx, const double* cov, int64_t P, int64_t n_ell, double* inv, double* logdet,
                  int32_t flags);

/* loglike[b] = -1/2 (data - mu_b)^T prec (data - mu_b) + c0 for B means
 * (mu: [B][P][n_ell] float32, prec: the inverse in the sparse format), the
 * quadratic form of likelihood.py:9-61 (c0 = +1/2 log det C to include it:
 * likelihood.py:61 returns -1/2 (r^T C^-1 r - log det C));
 * dloglike_dmu (may be NULL) = prec (data - mu_b), [B][P][n_ell] float32. */
int jc_gaussian_loglike(jc_ctx* ctx, const float* mu, int64_t B, int64_t P, int64_t n_ell, const double* data,
                        const double* prec, double c0, double* loglike, float* dloglike_dmu, int32_t flags);

/* The same for the means angular_cl(cosmo_b) of a registered tracer set,
 * with grad[b][7] = d loglike_b / d(Omega_c, Omega_b, h, n_s, sigma8, w0, wa):
 * jax.grad of a Gaussian likelihood through angular_cl (the likelihood of
 * gaussian_cl_covariance_and_mean, angular_cl.py:166-196, with a fixed
 * covariance).  One forward pass, the likelihood kernels, one adjoint pass. */
int jc_gaussian_loglike_grad(jc_ctx* ctx, const double* cosmo, int64_t B, int64_t set_id, const double* ell,
                             int64_t n_ell, int64_t n_z, const double* data, const double* prec, double c0,
                             double* loglike, double* grad, int32_t flags);

/* Dense covariances (likelihood.py:43-61 with C an N x N array, N = P *
 * n_ell, the default of gaussian_cl_covariance_and_mean, angular_cl.py:173),
 * FP64, row-major.  jc_dense_cholesky: the lower Cholesky factor L of C
 * (C = L L^T; the strict upper triangle of L is zero) and log det C; L may be
 * the same buffer as cov; JC_E_ARG when C is not positive definite; always
 * returns after the factor is checked.  N <= 65535. */
int jc_dense_cholesky(jc_ctx* ctx, const double* cov, int64_t N, double* L, double* logdet, int32_t flags);

/* X = L^-1 R (mode 0), L^-T R (mode 1) or C^-1 R = L^-T L^-1 R (mode 2) for
 * nrhs right-hand sides stored as rows, R and X [nrhs][N] (X may be R),
 * with L from jc_dense_cholesky (jax.scipy cho_solve of likelihood.py:50). */
int jc_dense_solve(jc_ctx* ctx, const double* L, int64_t N, const double* R, int64_t nrhs, int32_t mode,
                   double* X, int32_t flags);

/* loglike[b] = -1/2 (data - mu_b)^T C^-1 (data - mu_b) + c0 for B means
 * mu [B][N] float32, with C given by its factor L (jc_dense_cholesky): the
 * quadratic form as |L^-1 (data - mu_b)|^2 (the value of both inverse
 * methods of likelihood.py:46-52); dloglike_dmu (may be NULL) = C^-1
 * (data - mu_b), [B][N] float32. */
int jc_gaussian_loglike_dense(jc_ctx* ctx, const float* mu, int64_t B, int64_t N, const double* data,
                              const double* L, double c0, double* loglike, float* dloglike_dmu, int32_t flags);

/* jc_gaussian_loglike_grad with a dense covariance given by its factor L
 * ([P n_ell]^2): log-likelihood and its gradient w.r.t. (Omega_c, Omega_b,
 * h, n_s, sigma8, w0, wa) for the means angular_cl(cosmo_b). */
int jc_gaussian_loglike_grad_dense(jc_ctx* ctx, const double* cosmo, int64_t B, int64_t set_id, const double* ell,
                                   int64_t n_ell, int64_t n_z, const double* data, const double* L, double c0,
                                   double* loglike, double* grad, int32_t flags);

/* Linear and nonlinear matter power at n (k, a) pairs for one cosmology
 * (k in h/Mpc).  Replaces power.py:21-53 (linear_matter_power) and
 * power.py:265-272 (nonlinear_matter_power with halofit, Eisenstein_Hu). */
int jc_matter_power(jc_ctx* ctx, const double* cosmo, const double* k, const double* a, int64_t n,
                    double* pk_lin, double* pk_nl);

/* Background at n scale factors for one cosmology: radial comoving distance
 * [Mpc/h] (background.py:199-242) and growth factor (background.py:371-398).
 * Either output may be NULL. */
int jc_background(jc_ctx* ctx, const double* cosmo, const double* a, int64_t n, double* chi,
                  double* growth);

/* Background at n <= 4194304 scale factors for one cosmology, in FP64:
 * radial comoving distance [Mpc/h] (background.py:199-242), growth factor
 * (background.py:371-398), growth rate dlnD/dlna (background.py:401-440) and
 * transverse comoving distance [Mpc/h] (background.py:297-344; the angular
 * diameter distance of background.py:347-368 is a * fk).  The growth
 * quantities follow the context's growth model.  Any output may be NULL. */
int jc_background_ex(jc_ctx* ctx, const double* cosmo, const double* a, int64_t n, double* chi,
                     double* growth, double* rate, double* fk);

/* The sigma_8 normalisation of the linear power of one cosmology,
 * pknorm = sigma8^2 / sigma^2(R = 8 Mpc/h) (power.py:37-41, with the Romberg
 * integral of power.py:56-78 and its log10/ln bounds), FP64, under the
 * context's transfer model. */
int jc_sigma8_norm(jc_ctx* ctx, const double* cosmo, double* pknorm);

/* Tile shape used by the Limber kernel for a batch shape (the autotuner
 * writes it; 0 = automatic).  waves_per_block in {1, 2, 4}. */
int jc_set_tiles(jc_ctx* ctx, int32_t waves_per_block, int32_t wl_ctile);
int jc_get_tiles(jc_ctx* ctx, int32_t* waves_per_block, int32_t* wl_ctile);
/* Shape of the per-cosmology table kernel (autotuned): 0 = default,
 * 1 = 256 threads at 4 waves/SIMD, 2 = 256 threads at 2 waves/SIMD,
 * 3 = 512 threads at 2 waves/SIMD. */
int jc_set_tables_variant(jc_ctx* ctx, int32_t variant);
/* Physics models of the context's subsequent calls (-1 = unchanged):
 * the transfer_fn and nonlinear_fn arguments of angular_cl.py:49-51.
 *   transfer: JC_TRANSFER_EH_OSC (Eisenstein_Hu type="eisenhu_osc", the
 *     default) or JC_TRANSFER_EH_NOWIGGLE (type="eisenhu"), transfer.py:10-156
 *   nonlinear: JC_HALOFIT_TAKAHASHI2012 (default) or JC_HALOFIT_SMITH2003,
 *     power.py:144-262 */
#define JC_TRANSFER_EH_OSC 0
#define JC_TRANSFER_EH_NOWIGGLE 1
#define JC_HALOFIT_TAKAHASHI2012 0
#define JC_HALOFIT_SMITH2003 1
int jc_set_models(jc_ctx* ctx, int32_t transfer, int32_t nonlinear);

/* Growth model of the context's subsequent calls (servers keep the model of
 * their creation).  Replaces the gamma argument of Cosmology (core.py:13,
 * 59-60) and the dispatch of background.py:371-440:
 *   JC_GROWTH_ODE: the linear growth ODE, RK4 over a (background.py:443-512),
 *     for a Cosmology without gamma;
 *   JC_GROWTH_GAMMA: D from dlnD/dlna = Omega_m(a)^gamma, RK4 in ln a on the
 *     same 128 log-a nodes (background.py:515-582), |gamma| <= 10.
 * The gamma applies to every cosmology of a call; the derivative entry
 * points treat it as a constant (no d/d gamma). */
#define JC_GROWTH_ODE 0
#define JC_GROWTH_GAMMA 1
int jc_set_growth(jc_ctx* ctx, int32_t model, double gamma);

/* How forward calls (jc_angular_cl) evaluate; each option 1 = on, 0 = off,
 * -1 = unchanged.  Defaults: transfer_table 1, matrix_cores 0 (on MI355X the
 * FP32 matrix-core products do not overlap the per-point VALU work and the
 * 55 pairs fill 64-row tiles, so the per-lane FMAs are faster).
 *   transfer_table: the EH98 transfer function (transfer.py:10-156) from a
 *     per-cosmology table of cubic Hermite pieces in ln k, built once per
 *     call (error below 1e-6 of T(k)), instead of the closed form at every
 *     (ell, a) point;
 *   matrix_cores: the sum over tracer pairs and a-nodes as FP32 matrix-core
 *     products instead of per-lane FMAs.
 * Both apply when n_ell >= 32 (the table without matrix cores: n_ell >= 64).
 * The adjoint (jc_angular_cl_vjp, jc_gaussian_loglike_grad) also reads the
 * transfer table when n_ell >= 64, with a second table of its theta
 * derivatives; JVP calls use the closed form; derivative calls always sum
 * the pairs with per-lane FMAs. */
int jc_set_forward_options(jc_ctx* ctx, int32_t transfer_table, int32_t matrix_cores);

/* Per-call FP64 intermediates of the table kernel for one cosmology (test
 * hook): chi table [256], growth table [128], sigma^2(R) table [256],
 * pknorm, then for each a-node 8 values (chi, D, k_nl, n_eff, C, dchi/da).
 * out must hold 1024 + 8 * (n_z + 1) doubles. */
int jc_debug_tables(jc_ctx* ctx, const double* cosmo, int64_t set_id, int64_t n_z, double* out);

/* Per-kernel device time, measured with HIP events recorded on the launch
 * stream around every launch while profiling is on.  jc_profile_read fills
 * total_ms[k] and launches[k] for k < JC_NUM_KERNELS (tables, weak-lensing,
 * Limber, transfer table) accumulated since the previous read, then resets.
 * The transfer table is built at the end of the forward table kernel, so its
 * slot counts no launches of its own.  No reference counterpart (the
 * reference has no profiling hooks); used by the benchmark. */
#define JC_KERNEL_TABLES 0
#define JC_KERNEL_WL 1
#define JC_KERNEL_LIMBER 2
#define JC_KERNEL_EHTAB 3
#define JC_NUM_KERNELS 4
int jc_profile_enable(jc_ctx* ctx, int32_t on);
/* Which kernels profiling times: bit k = JC_KERNEL_k (default: all).  Each
 * timed launch adds two event records to its stream, so a caller timing a
 * whole pass can restrict them to the kernel it studies.  JC_E_ARG for an
 * empty mask or bits past JC_NUM_KERNELS. */
int jc_profile_select(jc_ctx* ctx, uint32_t kernels);
int jc_profile_read(jc_ctx* ctx, double* total_ms, int64_t* launches);

/* ------------------------------------------------------------------------
 * Served batches (no reference counterpart: the reference evaluates one
 * request per call; this is the continuous-batching request loop).
 * A server owns a scheduler thread and `slots` batches in flight, each on its
 * own stream with its own fixed-capacity device state.  Requests name a
 * cosmology, a subset of the tracers of a registered pool set, their own ell
 * list and Simpson size, and optionally a cotangent for the 7-parameter
 * adjoint.  Requests are admitted into the next free slot in arrival order
 * (with two or more slots and at most max_batch requests pending, slot 0
 * takes only requests without a gradient, so that they never wait for a
 * gradient pass to free a slot) while the batch's
 * per-request device footprint stays within mem_budget;
 * uneven requests share one pass of persistent kernels, replayed from a
 * captured hipGraph per slot.  Every request is integrated on the pool's
 * a-grid, which ends at the pool's largest zmax, and on each weak-lensing
 * probe's inner grids, which end at that probe's largest zmax: a request is
 * accepted (and then equals angular_cl of its own bins) only where its bins
 * reach those same ends; others are refused with JC_E_UNSUPPORTED. */
typedef struct jc_server jc_server;

typedef struct jc_server_config {
  int32_t max_batch;  /* requests per batch (default 256)                         */
  int32_t max_grad;   /* gradient requests per batch (default 64)                 */
  int32_t max_ell;    /* ell values per request (default 256)                     */
  int32_t slots;      /* batches in flight (default 2)                            */
  int32_t waves;      /* Limber kernel waves per block, 1/2/4 (0 = 2)             */
  int32_t wl_ctile;   /* weak-lensing cosmologies per block, 1/4/8 (0 = 4)        */
  int32_t use_graph;  /* 1: replay a captured hipGraph per batch                  */
  int32_t n_nz;       /* number of Simpson sizes served (1..4)                    */
  int32_t n_z[4];     /* the sizes, e.g. 256, 512, 1024                           */
  int64_t mem_budget; /* per-batch bytes of per-request device state (0 = capacity) */
} jc_server_config;   /* 56 bytes */

typedef struct jc_request {
  double cosmo[8];      /* Omega_c, Omega_b, h, n_s, sigma8, Omega_k, w0, wa      */
  int32_t n_tracers;    /* 1..10                                                  */
  int32_t tracers[10];  /* indices into the pool set, in output order             */
  int32_t n_ell;        /* 1..max_ell                                             */
  int32_t n_z;          /* one of the served sizes                                */
  int32_t want_grad;    /* 1: also grad_out = sum cl_bar * dC/dtheta              */
  int32_t reserved[2];
  const double* ell;    /* [n_ell], copied at submit                              */
  const float* cl_bar;  /* [P][n_ell] cotangent when want_grad, copied at submit  */
  float* cl_out;        /* [P][n_ell]: written before the completion is reported  */
  double* grad_out;     /* [7] when want_grad                                     */
} jc_request;           /* 160 bytes */

typedef struct jc_completion {
  int64_t id;
  int32_t status;       /* 0 or a JC_E_* code                                     */
  int32_t batch_size;   /* requests in the batch that served it                   */
  double latency_us;    /* submit -> results written                              */
  double t_submit_us;   /* server clock                                           */
  double t_done_us;
} jc_completion;        /* 40 bytes */

int jc_server_create(jc_ctx* ctx, int64_t pool_set, const jc_server_config* cfg, jc_server** out);
/* Thread-safe, non-blocking.  Invalid requests are refused here (nothing is
 * queued); a request larger than mem_budget completes with JC_E_NOMEM. */
int jc_server_submit(jc_server* srv, const jc_request* req, int64_t* id);
/* n requests at once (ids[n]): all are validated before any is queued, so an
 * invalid one refuses the whole call; they queue in order (in chunks, so the
 * first batch forms while the rest are copied) with consecutive ids. */
int jc_server_submit_many(jc_server* srv, const jc_request* reqs, int64_t n, int64_t* ids);
/* Completions since the previous poll (up to max). */
int jc_server_poll(jc_server* srv, jc_completion* out, int64_t max, int64_t* n);
/* Block until at most `outstanding` requests are queued or in flight, or
 * timeout_ms elapses (JC_E_ARG on timeout). */
int jc_server_wait(jc_server* srv, int64_t outstanding, int32_t timeout_ms);
/* Test hook (no reference counterpart): the next `batches` batches fail at
 * launch, before any work is queued, as a device allocation (status
 * JC_E_NOMEM) or runtime failure (JC_E_HIP) would; their requests complete
 * with that status and the server carries on with the next batch. */
int jc_server_inject_failure(jc_server* srv, int32_t batches, int32_t status);
/* stats[0..10]: batches, requests done, rejected, max batch, mean batch,
 * device busy ms (sum of batch spans), graph replays, direct launches; over
 * the requests done, summed ms from submission to their batch's launch and
 * from that launch to completion; summed host ms of the batch launches
 * (packing and enqueueing). */
int jc_server_stats(jc_server* srv, double* stats, int64_t n);
int jc_server_destroy(jc_server* srv);

const char* jc_last_error(void);
int jc_version(void);

#ifdef __cplusplus
}
#endif
#endif /* JC_LIMBER_H */
