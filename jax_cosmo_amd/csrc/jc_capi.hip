// Host side of libjc_limber.so: context, resident tables, tracer sets and the
// C ABI declared in include/jc_limber.h.
#include "jc_host.hpp"

using namespace jc;
using namespace jch;

// Device scratch of the allocation-free forward entry point
// (jc_angular_cl_noalloc): each GPU the library runs on holds its own copy
// until the process ends.  In host code `scratch` is only a stand-in for the
// symbol: its device address comes from hipGetSymbolAddress, once per device
// (scratch_address), and only that address is used.
static __device__ __align__(256) unsigned char scratch[JC_NOALLOC_SCRATCH_BYTES];

namespace jch {

namespace {
std::mutex g_scratch_mu;
std::map<int, unsigned char*> g_scratch_ptr;  // device ordinal -> address of its scratch
}  // namespace

// Address of `scratch` on `device`, which must be current.  Called from
// jc_ctx_create (the library's set-up), so the entry point itself makes no
// runtime call beyond copies and launches.
int scratch_address(int device, unsigned char** out) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  auto it = g_scratch_ptr.find(device);
  if (it != g_scratch_ptr.end()) {
    *out = it->second;
    return JC_OK;
  }
  void* p = nullptr;
  HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  if (!p) return fail(JC_E_HIP, "no device address for the scratch array on device %d", device);
  g_scratch_ptr[device] = static_cast<unsigned char*>(p);
  *out = static_cast<unsigned char*>(p);
  return JC_OK;
}

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

std::vector<double> logspace(double a, double b, int n) {
  std::vector<double> v(n);
  for (int i = 0; i < n; ++i) v[i] = pow(10.0, linspace_at(a, b, n, i));
  return v;
}

int ctx_init_tables(jc_ctx* ctx) {
  hipStream_t s = ctx->stream;
  ctx->atab_chi_h = logspace(-3.0, 0.0, NCHI);  // background.py:225
  ctx->atab_g_h = logspace(-3.0, 0.0, NGROW);   // background.py:463
  HIPCHK(ctx->atab_chi.upload(ctx->atab_chi_h, s));
  HIPCHK(ctx->atab_g.upload(ctx->atab_g_h, s));
  // halofit grids, power.py:92 and power.py:111
  std::vector<double> lnk(NHF_K), w(NHF_K), logr(NHF_R), E((size_t)NHF_K * NHF_R);
  std::vector<float> k32(NHF_K);
  std::vector<double> kp((size_t)NHF_K * 4, 0.0);
  const double lo = log(1e-4), hi = log(1e4);
  const double dx = (hi - lo) / (NHF_K - 1);
  for (int i = 0; i < NHF_K; ++i) {
    lnk[i] = linspace_at(lo, hi, NHF_K, i);
    w[i] = dx / 3 * ((i == 0 || i == NHF_K - 1) ? 1.0 : ((i & 1) ? 4.0 : 2.0));
    k32[i] = (float)exp(lnk[i]);
    const double k2 = exp(lnk[i]) * exp(lnk[i]);
    kp[(size_t)i * 4] = k2;
    kp[(size_t)i * 4 + 1] = k2 * k2;
    kp[(size_t)i * 4 + 2] = k2 * k2 * k2;
  }
  for (int j = 0; j < NHF_R; ++j) logr[j] = linspace_at(log(1e-4), log(1e1), NHF_R, j);
  for (int i = 0; i < NHF_K; ++i) {
    const double k = exp(lnk[i]);
    for (int j = 0; j < NHF_R; ++j) {
      const double y = k * exp(logr[j]);
      E[(size_t)i * NHF_R + j] = exp(-(y * y));
    }
  }
  HIPCHK(ctx->hf_lnk.upload(lnk, s));
  HIPCHK(ctx->hf_w.upload(w, s));
  HIPCHK(ctx->logr.upload(logr, s));
  HIPCHK(ctx->hf_E.upload(E, s));
  HIPCHK(ctx->hf_k32.upload(k32, s));
  HIPCHK(ctx->hf_kp.upload(kp, s));
  std::vector<double> kd(NHF_K);
  for (int i = 0; i < NHF_K; ++i) kd[i] = exp(lnk[i]);
  HIPCHK(ctx->hf_kd.upload(kd, s));
  // chi table nodes: RK4 in t = ln a with midpoints t_{i-1} + h/2
  // (background.py:223-233, scipy/ode.py:6-22)
  std::vector<double> cn((size_t)NCHI * 8, 0.0);
  for (int i = 0; i < NCHI; ++i) {
    double* n = &cn[(size_t)i * 8];
    const double ti = log(ctx->atab_chi_h[i]);
    n[0] = exp(ti);
    n[1] = ti;
    n[2] = 1.0 / n[0];
    if (i > 0) {
      const double tp = log(ctx->atab_chi_h[i - 1]);
      const double tm = tp + (ti - tp) / 2;
      n[3] = exp(tm);
      n[4] = tm;
      n[5] = 1.0 / n[3];
      n[6] = ti - tp;
    }
  }
  HIPCHK(ctx->chi_nodes.upload(cn, s));
  // the same quantities on the growth grid, for the gamma growth: RK4 in
  // t = ln a of dlnD/dlna = Omega_m(a)^gamma (background.py:532-545)
  std::vector<double> ggn((size_t)NGROW * 8, 0.0);
  for (int i = 0; i < NGROW; ++i) {
    double* n = &ggn[(size_t)i * 8];
    const double ti = log(ctx->atab_g_h[i]);
    n[0] = exp(ti);
    n[1] = ti;
    n[2] = 1.0 / n[0];
    if (i > 0) {
      const double tp = log(ctx->atab_g_h[i - 1]);
      const double tm = tp + (ti - tp) / 2;
      n[3] = exp(tm);
      n[4] = tm;
      n[5] = 1.0 / n[3];
      n[6] = ti - tp;
    }
  }
  HIPCHK(ctx->gg_nodes.upload(ggn, s));
  // growth table nodes: RK4 in a with midpoints x_{i-1} + (x_i - x_{i-1})/2
  std::vector<double> gn((size_t)NGROW * 8, 0.0);
  for (int i = 0; i < NGROW; ++i) {
    double* n = &gn[(size_t)i * 8];
    const double x = ctx->atab_g_h[i];
    n[0] = x;
    n[1] = log(x);
    n[2] = 1.0 / x;
    if (i > 0) {
      const double xp = ctx->atab_g_h[i - 1];
      const double xm = xp + (x - xp) / 2;
      n[3] = xm;
      n[4] = log(xm);
      n[5] = 1.0 / xm;
    }
  }
  HIPCHK(ctx->g_nodes.upload(gn, s));
  // sigma_8 Romberg nodes: ln k = -4 + 7 j / 128 (power.py:66-76, the
  // log10/ln bounds as in the reference), top-hat window of R = 8
  std::vector<double> rb((size_t)NROMB * 3);
  for (int j = 0; j < NROMB; ++j) {
    const double lk = -4.0 + j * (7.0 / 128.0);
    const double k = exp(lk);
    const double x = k * 8.0;
    const double wv = 3.0 * (sin(x) - x * cos(x)) / (x * x * x);
    rb[(size_t)j * 3] = k;
    rb[(size_t)j * 3 + 1] = 1.0 / k;
    rb[(size_t)j * 3 + 2] = k * ((k * wv) * (k * wv));
  }
  HIPCHK(ctx->romb.upload(rb, s));
  return JC_OK;
}

int eval_pz(jc_ctx* ctx, TracerSet* ts, int i, const std::vector<double>& z, std::vector<double>& out) {
  const jc_tracer_desc& t = ts->tr[i];
  out.resize(z.size());
  const Catalog* cat = i < (int)ts->cat.size() ? ts->cat[i].get() : nullptr;
  if (t.nz_kind == JC_NZ_HOST) {
    // a user-defined redshift_distribution: its pz_fn on the host
    std::vector<double> zz(z);
    if (t.nz_shifted)
      for (auto& v : zz) v = std::max(v - t.nz_params[3], 0.0);  // redshift.py:169-171
    const auto& h = ts->host[i];
    if (!zz.empty()) {
      const int rc = h.first(h.second, zz.data(), out.data(), (int64_t)zz.size());
      if (rc != 0) return fail(JC_E_ARG, "tracer %d: the host n(z) function failed (%d)", i, rc);
    }
    for (size_t p = 0; p < out.size(); ++p)
      if (!std::isfinite(out[p]))
        return fail(JC_E_ARG, "tracer %d: the host n(z) function gave %g at z = %g", i, out[p], z[p]);
    return JC_OK;
  }
  if (!cat) {
    for (size_t p = 0; p < z.size(); ++p) out[p] = pz_fn(t, z[p]);
    return JC_OK;
  }
  std::vector<double> zz(z);
  if (t.nz_shifted)
    for (auto& v : zz) v = std::max(v - t.nz_params[3], 0.0);  // redshift.py:169-171
  DevBuf<double> dz, dout;
  HIPCHK(dz.upload(zz, ctx->stream));
  HIPCHK(dout.reserve(zz.size()));
  HIPCHK(launch_kde(dz.p, (int64_t)zz.size(), cat->z.p, cat->w.p, cat->n, t.nz_params[0], cat->wsum, dout.p,
                    ctx->stream));
  HIPCHK(hipMemcpyAsync(out.data(), dout.p, zz.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return JC_OK;
}

// Build the per-N tables of a tracer set: the Simpson a-grid of
// angular_cl.py:96, the number-count n(z) b on it, and for every
// weak-lensing probe the per-z inner Simpson grid of probes.py:51.
int build_grid(jc_ctx* ctx, TracerSet* ts, int N, GridTables** out) {
  auto it = ts->grids.find(N);
  if (it != ts->grids.end()) {
    *out = it->second.get();
    return JC_OK;
  }
  std::unique_ptr<GridTables> g(new GridTables);
  hipStream_t s = ctx->stream;
  const int Na = N + 1;
  g->N = N;
  g->Na = Na;
  double zmax_all = 0.0;
  for (auto& t : ts->tr) zmax_all = std::max(zmax_all, t.zmax);  // angular_cl.py:63
  const double a0 = 1.0 / (1.0 + zmax_all);
  const double da = (1.0 - a0) / N;
  std::vector<double> a(Na), aw(Na), z(Na);
  std::vector<float> zp1(Na);
  for (int m = 0; m < Na; ++m) {
    a[m] = linspace_at(a0, 1.0, Na, m);
    aw[m] = da / 3 * ((m == 0 || m == N) ? 1.0 : ((m & 1) ? 4.0 : 2.0));
    z[m] = 1.0 / a[m] - 1.0;  // a2z, utils.py:7-9
    zp1[m] = (float)(1.0 + z[m]);
  }
  HIPCHK(g->a_list.upload(a, s));
  HIPCHK(g->a_wgt.upload(aw, s));
  {
    std::vector<double> aln(a.size()), al10(a.size());
    for (size_t m = 0; m < a.size(); ++m) {
      aln[m] = log(a[m]);
      al10[m] = log10(a[m]);
    }
    HIPCHK(g->a_ln.upload(aln, s));
    HIPCHK(g->a_l10.upload(al10, s));
  }
  HIPCHK(g->zp1.upload(zp1, s));
  // weak-lensing bins, one group per probe
  std::map<int, std::vector<int>> probes;
  for (int i = 0; i < ts->T; ++i)
    if (ts->tr[i].kind == JC_TRACER_WL) probes[ts->tr[i].probe].push_back(i);
  std::map<int, bool> probe_ia;
  for (auto& pr : probes)
    for (int i : pr.second) probe_ia[pr.first] = probe_ia[pr.first] || ts->tr[i].bias_kind != JC_BIAS_NONE;
  // z-local kernels: number counts n(z) b(z) H(a) (probes.py:77-99) and, for
  // the weak-lensing bins of a probe with intrinsic alignments, the NLA term
  // -(C_1 rho_crit) Omega_m / D(a) n(z) b(z) H(a) (probes.py:102-129) times
  // (1 + m) (probes.py:205-207); bins without IA in such a probe get zero
  std::vector<double> nb, pzv;
  std::vector<int> lcol, lmode;
  for (int i = 0; i < ts->T; ++i) {
    const auto& t = ts->tr[i];
    const bool wl_ia = t.kind == JC_TRACER_WL && probe_ia[t.probe];
    if (t.kind != JC_TRACER_NC && !wl_ia) continue;
    int rc = eval_pz(ctx, ts, i, z, pzv);
    if (rc) return rc;
    const bool has_b = t.bias_kind != JC_BIAS_NONE;
    int mode = (has_b && t.bias_kind == JC_BIAS_INV_GROWTH) ? LOC_INVD : 0;
    double f = 1.0;
    if (t.kind == JC_TRACER_WL) {
      mode |= LOC_IA;
      f = -(C1_IA * RHOCRIT) * (1.0 + t.m_bias);
    }
    lcol.push_back(i);
    lmode.push_back(mode);
    for (int m = 0; m < Na; ++m)
      nb.push_back(has_b ? pzv[m] / ts->norm[i] * bias_host(t, z[m]) * f : 0.0);
  }
  g->nLoc = (int)lcol.size();
  if (g->nLoc) {
    HIPCHK(g->loc_nb.upload(nb, s));
    HIPCHK(g->loc_col.upload(lcol, s));
    HIPCHK(g->loc_mode.upload(lmode, s));
  }
  for (auto& pr : probes) {
    const std::vector<int>& idx = pr.second;
    double zmax = 0.0;
    for (int i : idx) zmax = std::max(zmax, ts->tr[i].zmax);  // probes.py:24
    std::unique_ptr<WLGroup> wg(new WLGroup);
    wg->nw = (int)idx.size();
    wg->ia = probe_ia[pr.first];
    // single source planes: chi at a_s = 1/(1 + z0) through the same table
    // segment rule as the extended bins' inner nodes
    std::vector<WLSource> src(idx.size());
    for (size_t q = 0; q < idx.size(); ++q) {
      const auto& t = ts->tr[idx[q]];
      src[q] = WLSource{0, 0, 0.0f, 0.0f};
      if (t.nz_kind == JC_NZ_DELTA) {
        const double as = 1.0 / (1.0 + t.nz_params[0]);
        const int seg = interp_segment(as, ctx->atab_chi_h);
        src[q] = WLSource{1, seg, (float)(as - ctx->atab_chi_h[seg]), (float)(1.0 + t.m_bias)};
      }
    }
    HIPCHK(wg->src.upload(src, s));
    std::vector<uint8_t> L((size_t)NWL_J * Na);
    std::vector<float> dl((size_t)NWL_J * Na);
    // rows padded with zeros to the kernel's bin count (5 or NWL_MAX)
    const int nw_rows = wg->nw <= 5 ? 5 : NWL_MAX;
    std::vector<float> Nw((size_t)nw_rows * NWL_J * Na, 0.0f);
    // n(z') of every bin at every inner node, [j][m]
    std::vector<double> zin((size_t)NWL_J * Na);
    for (int m = 0; m < Na; ++m)
      for (int j = 0; j < NWL_J; ++j) zin[(size_t)j * Na + m] = linspace_at(z[m], zmax, NWL_J, j);
    std::vector<std::vector<double>> pzin(wg->nw);
    for (int q = 0; q < wg->nw; ++q) {
      if (ts->tr[idx[q]].nz_kind == JC_NZ_DELTA) continue;
      int rc = eval_pz(ctx, ts, idx[q], zin, pzin[q]);
      if (rc) return rc;
    }
    for (int m = 0; m < Na; ++m) {
      const double dxm = (zmax - z[m]) / (NWL_J - 1);
      for (int j = 0; j < NWL_J; ++j) {
        const double zp = linspace_at(z[m], zmax, NWL_J, j);
        const double ap = 1.0 / (1.0 + zp);  // z2a, utils.py:2-4
        const int seg = interp_segment(ap, ctx->atab_chi_h);
        L[(size_t)j * Na + m] = (uint8_t)seg;
        dl[(size_t)j * Na + m] = (float)(ap - ctx->atab_chi_h[seg]);
        const double wj = (j == 0 || j == NWL_J - 1) ? 1.0 : ((j & 1) ? 4.0 : 2.0);
        for (int q = 0; q < wg->nw; ++q) {
          const auto& t = ts->tr[idx[q]];
          Nw[((size_t)q * NWL_J + j) * Na + m] =
              t.nz_kind == JC_NZ_DELTA
                  ? 0.0f
                  : (float)(dxm / 3 * wj * pzin[q][(size_t)j * Na + m] / ts->norm[idx[q]] * (1.0 + t.m_bias));
        }
      }
    }
    HIPCHK(wg->L.upload(L, s));
    HIPCHK(wg->delta.upload(dl, s));
    HIPCHK(wg->Nw.upload(Nw, s));
    HIPCHK(wg->col.upload(idx, s));
    g->wl.push_back(std::move(wg));
  }
  *out = g.get();
  ts->grids[N] = std::move(g);
  return JC_OK;
}

// Consecutive 256-byte aligned regions of a fixed device buffer (the scratch
// array of the allocation-free entry point).
struct Bump {
  unsigned char* base = nullptr;
  size_t cap = 0, off = 0;
  template <class T>
  T* take(size_t count) {
    const size_t at = (off + 255) & ~(size_t)255;
    const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    if (at > cap || bytes > cap - at) return nullptr;
    off = at + bytes;
    return reinterpret_cast<T*>(base + at);
  }
};

// Workspace of `count` elements: the context's buffer (grown with hipMalloc
// when too small) or, with a Bump, the next scratch region.
template <class T>
int workspace(DevBuf<T>& buf, size_t count, Bump* sc, T** out) {
  if (sc) {
    *out = sc->take<T>(count);
    if (!*out) return fail(JC_E_ARG, "the call needs more than the %zu bytes of scratch", sc->cap);
    return JC_OK;
  }
  HIPCHK(buf.reserve(count));
  *out = buf.p;
  return JC_OK;
}

template <class T>
int to_device(jc_ctx* ctx, DevBuf<T>& buf, const T* src, size_t count, bool dev, const T** dptr, Bump* sc = nullptr) {
  if (dev) {
    *dptr = src;
    return JC_OK;
  }
  T* p = nullptr;
  int rc = workspace(buf, count, sc, &p);
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(p, src, count * sizeof(T), hipMemcpyHostToDevice, ctx->stream));
  *dptr = p;
  return JC_OK;
}

int check_cosmo_host(const double* cosmo, int64_t B) {
  for (int64_t b = 0; b < B; ++b) {
    const double* p = cosmo + 8 * b;
    for (int i = 0; i < 8; ++i)
      if (!std::isfinite(p[i])) return fail(JC_E_ARG, "cosmology %lld parameter %d is not finite", (long long)b, i);
    if (!(p[0] + p[1] > 0.0) || !(p[1] > 0.0) || !(p[2] > 0.0) || !(p[4] > 0.0))
      return fail(JC_E_ARG, "cosmology %lld: Omega_b, Omega_m, h and sigma8 must be positive", (long long)b);
  }
  return JC_OK;
}

int auto_waves(jc_ctx* ctx, int64_t ncols) {
  if (ctx->waves > 0) return ctx->waves;
  // enough waves to fill 256 CUs x 4 SIMDs a few times over
  const int64_t blocks = (ncols + 63) / 64;
  if (blocks >= 2048) return 2;
  return 4;
}
// Splits of each column group's a range over blocks (grid y), each wave
// keeping at least 8 a-nodes: small batches take enough slices for about
// 1024 waves.  The adjoint kernel holds `occ` waves per SIMD (1 with the
// EH98 closed form, 2 with the transfer table), so its blocks run in rounds
// of (CUs x 4 x occ / waves): it takes the fewest slices (up to 8) whose last
// round is at least 90 % full (measured on the 256-cosmology gradient, 400
// blocks: 3 slices with the closed form, -10 %; 5 with the table, -8 %).
// The forward kernels keep 4 waves per SIMD, which hide their tail, and
// take 1 slice at large batches.
int auto_slices(const jc_ctx* ctx, int64_t ncols, int waves, int Na, int nd, int occ) {
  const int64_t blocks = (ncols + 63) / 64;
  const int64_t smax = std::max(1, Na / (8 * waves));
  int64_t s = std::min<int64_t>((1024 + blocks * waves - 1) / (blocks * waves), smax);
  if (nd == 7 && s <= 1) {
    const double per_round = (double)ctx->num_cu * 4.0 * occ / waves;
    for (int64_t k = 1; k <= std::min<int64_t>(8, smax); ++k) {
      const double r = (double)(blocks * k) / per_round;
      if (r / std::ceil(r) >= 0.9) {
        s = k;
        break;
      }
    }
  }
  return (int)std::max<int64_t>(1, std::min<int64_t>(s, 64));
}
int auto_ctile(jc_ctx* ctx, int64_t B, int nd) {
  if (nd != 0) return 1;
  if (ctx->wl_ctile > 0) return ctx->wl_ctile;
  // measured on MI355X (N = 512, 10 tracers): one cosmology per block up to
  // ~128 (more blocks), four from ~256 (table reuse); eight is slower
  if (B >= 192) return 4;
  return 1;
}

// Timing events only: jc_profile_read synchronizes the stream before reading
// them, so they need no system-scope release (without it each record cost
// the 1024-cosmology batch's stream ~1.5 us less).
hipEvent_t prof_event(jc_ctx* ctx) {
  if (!ctx->ev_pool.empty()) {
    hipEvent_t e = ctx->ev_pool.back();
    ctx->ev_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess) return nullptr;
  return e;
}

// Launch wrapper: brackets the launch with events when profiling is on.
template <class F>
hipError_t timed(jc_ctx* ctx, int kind, F&& launch) {
  if (!ctx->prof || !((ctx->prof_kernels >> kind) & 1u)) return launch();
  hipEvent_t a = prof_event(ctx), b = prof_event(ctx);
  if (!a || !b) return launch();
  hipError_t e = hipEventRecord(a, ctx->stream);
  if (e != hipSuccess) return e;
  e = launch();
  if (e != hipSuccess) return e;
  e = hipEventRecord(b, ctx->stream);
  ctx->recs.push_back({kind, a, b});
  return e;
}

// The whole forward / derivative pipeline.  nd = 0 (C_ell), 1 (JVP), 7 (VJP).
// sc != null (forward calls only): every per-call buffer comes from the
// scratch regions of sc, and the tracer set's tables for n_z must already be
// resident, so the call allocates nothing.
int run_pipeline(jc_ctx* ctx, int nd, const double* cosmo, const double* tangent, int64_t B,
                 int64_t set_id, const double* ell, int64_t n_ell, int64_t n_z, float* cl,
                 float* cl_dot, const float* cl_bar, double* theta_bar, int32_t flags, Bump* sc = nullptr) {
  if (!ctx) return fail(JC_E_ARG, "null context");
  if (B < 1 || n_ell < 1) return fail(JC_E_ARG, "B and n_ell must be >= 1");
  if (n_z < 2 || (n_z % 2) != 0) return fail(JC_E_ARG, "N must be an even integer >= 2 (integrate.py:193)");
  if (n_z > 8192) return fail(JC_E_ARG, "n_z above 8192 is not supported");
  auto it = ctx->sets.find(set_id);
  if (it == ctx->sets.end()) return fail(JC_E_ARG, "unknown tracer set %lld", (long long)set_id);
  TracerSet* ts = it->second.get();
  const bool dev = (flags & JC_DEVICE_PTRS) != 0;
  if (!dev) {
    int rc = check_cosmo_host(cosmo, B);
    if (rc) return rc;
  }
  const int64_t ncols = B * n_ell;
  if (ncols > (int64_t)1 << 30) return fail(JC_E_ARG, "too many columns");
  if (sc && nd != 0) return fail(JC_E_ARG, "scratch workspaces serve forward calls only");
  HIPCHK(hipSetDevice(ctx->device));
  GridTables* g = nullptr;
  int rc = JC_OK;
  if (sc) {
    auto gi = ts->grids.find((int)n_z);
    if (gi == ts->grids.end())
      return fail(JC_E_ARG, "the tables of tracer set %lld for n_z = %lld are not resident: one jc_angular_cl call "
                  "with this set and n_z builds them", (long long)set_id, (long long)n_z);
    g = gi->second.get();
  } else {
    rc = build_grid(ctx, ts, (int)n_z, &g);
    if (rc) return rc;
  }
  hipStream_t s = ctx->stream;
  const int W = nd + 1;
  const int Na = g->Na;
  const int T = ts->T, P = T * (T + 1) / 2;

  const double* d_cosmo = nullptr;
  const double* d_ell = nullptr;
  rc = to_device(ctx, ctx->params, cosmo, (size_t)B * 8, dev, &d_cosmo, sc);
  if (rc) return rc;
  rc = to_device(ctx, ctx->ell, ell, (size_t)n_ell, dev, &d_ell, sc);
  if (rc) return rc;
  const double* d_seeds = nullptr;
  std::vector<double> seeds;
  if (nd > 0) {
    // d param_i / d direction: the 7 differentiable parameters skip Omega_k
    static const int pidx[7] = {0, 1, 2, 3, 4, 6, 7};
    std::vector<double> tan_h;
    if (nd == 1) {
      tan_h.resize((size_t)B * 7);
      if (dev) HIPCHK(hipMemcpyAsync(tan_h.data(), tangent, tan_h.size() * 8, hipMemcpyDeviceToHost, s));
      else memcpy(tan_h.data(), tangent, tan_h.size() * 8);
      HIPCHK(hipStreamSynchronize(s));
    }
    seeds.assign((size_t)B * 8 * nd, 0.0);
    for (int64_t b = 0; b < B; ++b)
      for (int q = 0; q < 7; ++q) {
        if (nd == 7) seeds[((size_t)b * 8 + pidx[q]) * 7 + q] = 1.0;
        else seeds[((size_t)b * 8 + pidx[q])] = tan_h[(size_t)b * 7 + q];
      }
    HIPCHK(ctx->seeds.reserve(seeds.size()));
    HIPCHK(hipMemcpyAsync(ctx->seeds.p, seeds.data(), seeds.size() * 8, hipMemcpyHostToDevice, s));
    d_seeds = ctx->seeds.p;
  }
  float *w_chitab = nullptr, *w_coef = nullptr, *w_crec = nullptr, *w_R = nullptr;
  if ((rc = workspace(ctx->chitab, (size_t)B * NCHI * 2 * W, sc, &w_chitab))) return rc;
  if ((rc = workspace(ctx->coef, (size_t)B * Na * NCOEF * W, sc, &w_coef))) return rc;
  if ((rc = workspace(ctx->crec, (size_t)B * NCREC * W, sc, &w_crec))) return rc;
  if ((rc = workspace(ctx->R, (size_t)B * Na * ts->Tpad * W, sc, &w_R))) return rc;

  TablesArgs ta{};
  ta.params = d_cosmo;
  ta.seeds = d_seeds;
  ta.B = (int)B;
  ta.Na = Na;
  ta.a_list = g->a_list.p;
  ta.a_ln = g->a_ln.p;
  ta.a_l10 = g->a_l10.p;
  ta.a_wgt = g->a_wgt.p;
  resident_tables(ctx, ta);
  ta.nLoc = g->nLoc;
  ta.loc_nb = g->loc_nb.p;
  ta.loc_col = g->loc_col.p;
  ta.loc_mode = g->loc_mode.p;
  ta.Tpad = ts->Tpad;
  ta.chitab = w_chitab;
  ta.coef = w_coef;
  ta.crec = w_crec;
  ta.R = w_R;
  ta.dbg = nullptr;
  // forward kernels whose column groups (32 for the matrix-core kernel, 64
  // otherwise) span at most two cosmologies read the transfer function from
  // a per-cosmology table
  const bool use_mx = nd == 0 && ctx->matrix_cores && n_ell >= 32;
  // the adjoint reads the table too, with its theta derivatives at the nodes
  const bool use_tab = (nd == 0 || nd == 7) && ctx->transfer_table && n_ell >= (use_mx ? 32 : 64);
  float4* w_ehtab = nullptr;
  if (use_tab) {
    // the table kernel ends by building each cosmology's transfer table
    if ((rc = workspace(ctx->ehtab, (size_t)B * NEHQ, sc, &w_ehtab))) return rc;
    ta.ehtab = w_ehtab;
    if (nd == 7) {
      HIPCHK(ctx->ehtab7.reserve((size_t)B * (NEHQ + 1) * NEHD_ROW));
      ta.ehtab7 = ctx->ehtab7.p;
    }
    ta.ell = d_ell;
    ta.n_ell = (int)n_ell;
  }
  // the Limber launch shape, its slice partials and the staged output are
  // settled before the first launch, so a call whose buffers do not fit the
  // scratch fails before any work is queued
  const int l_waves = auto_waves(ctx, ncols);
  const int l_slices = use_mx ? 1 : auto_slices(ctx, ncols, l_waves, Na, nd, use_tab ? 2 : 1);
  const size_t nslice = nd == 7 ? (size_t)ncols * 7 : (size_t)B * P * n_ell * (nd == 1 ? 2 : 1);
  const size_t nout = (size_t)B * P * n_ell;
  float* w_part = nullptr;
  if (l_slices > 1 && (rc = workspace(ctx->part, (size_t)l_slices * nslice, sc, &w_part))) return rc;
  float* d_out = cl;
  float* d_out_t = cl_dot;
  if (!dev && nd != 7) {
    if ((rc = workspace(ctx->out, nout, sc, &d_out))) return rc;
    if (nd == 1) {
      HIPCHK(ctx->out_t.reserve(nout));
      d_out_t = ctx->out_t.p;
    }
  }
  HIPCHK(timed(ctx, JC_KERNEL_TABLES, [&] { return launch_tables(nd, ta, s, ctx->tables_variant); }));

  for (auto& wg : g->wl) {
    WLArgs wa{};
    wa.B = (int)B;
    wa.Na = Na;
    wa.nw = wg->nw;
    wa.ia = wg->ia ? 1 : 0;
    wa.src = wg->src.p;
    wa.Lidx = wg->L.p;
    wa.delta = wg->delta.p;
    wa.Nw = wg->Nw.p;
    wa.zp1 = g->zp1.p;
    wa.col = wg->col.p;
    wa.chitab = w_chitab;
    wa.coef = w_coef;
    wa.crec = w_crec;
    wa.R = w_R;
    wa.Tpad = ts->Tpad;
    const int ct = auto_ctile(ctx, B, nd);
    HIPCHK(timed(ctx, JC_KERNEL_WL, [&] { return launch_wl(nd, ct, wa, s); }));
  }

  LimberArgs la{};
  la.B = (int)B;
  la.Na = Na;
  la.n_ell = (int)n_ell;
  la.ncols = (int)ncols;
  la.ell = d_ell;
  la.coef = w_coef;
  la.crec = w_crec;
  la.R = w_R;
  la.Tpad = ts->Tpad;
  la.wl_mask = ts->wl_mask;
  la.waves = l_waves;
  la.slices = l_slices;
  la.part = w_part;
  la.ehtab = use_tab ? w_ehtab : nullptr;
  la.ehtab7 = use_tab && nd == 7 ? ctx->ehtab7.p : nullptr;
  la.mx = use_mx ? 1 : 0;
  if (nd == 7) {
    const float* d_cbar = nullptr;
    rc = to_device(ctx, ctx->cbar, cl_bar, nout, dev, &d_cbar);
    if (rc) return rc;
    HIPCHK(ctx->theta_col.reserve((size_t)ncols * 7));
    la.cbar = d_cbar;
    la.theta_col = ctx->theta_col.p;
    HIPCHK(timed(ctx, JC_KERNEL_LIMBER, [&] { return launch_limber(nd, T, la, s); }));
    double* d_th = theta_bar;
    if (!dev) {
      HIPCHK(ctx->theta.reserve((size_t)B * 7));
      d_th = ctx->theta.p;
    }
    // the slices' partials are added inside the reduction
    HIPCHK(launch_theta_reduce(la.slices > 1 ? la.part : ctx->theta_col.p, la.slices, (int)B, (int)n_ell, d_th, s));
    if (!dev) {
      HIPCHK(hipMemcpyAsync(theta_bar, d_th, (size_t)B * 7 * sizeof(double), hipMemcpyDeviceToHost, s));
      HIPCHK(hipStreamSynchronize(s));
    }
    return JC_OK;
  }
  la.out = d_out;
  la.out_t = d_out_t;
  HIPCHK(timed(ctx, JC_KERNEL_LIMBER, [&] { return launch_limber(nd, T, la, s); }));
  if (la.slices > 1) {
    HIPCHK(launch_slice_sum(la.part, la.slices, (int64_t)nout, d_out, s));
    if (nd == 1) HIPCHK(launch_slice_sum(la.part + (size_t)la.slices * nout, la.slices, (int64_t)nout, d_out_t, s));
  }
  if (!dev) {
    HIPCHK(hipMemcpyAsync(cl, d_out, nout * sizeof(float), hipMemcpyDeviceToHost, s));
    if (nd == 1) HIPCHK(hipMemcpyAsync(cl_dot, d_out_t, nout * sizeof(float), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
  }
  return JC_OK;
}

// Per-a tables of one cosmology on an arbitrary list of scale factors (for
// the background and matter-power entry points).
int tables_on_alist(jc_ctx* ctx, const double* cosmo, const double* a, int64_t n, bool fp64_out = false) {
  if (n < 1 || n > (1 << 24)) return fail(JC_E_ARG, "bad number of points");
  int rc = check_cosmo_host(cosmo, 1);
  if (rc) return rc;
  for (int64_t i = 0; i < n; ++i)
    if (!(a[i] > 0.0) || !(a[i] <= 1.0)) return fail(JC_E_ARG, "scale factor %g outside (0, 1]", a[i]);
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  HIPCHK(ctx->params.reserve(8));
  HIPCHK(hipMemcpyAsync(ctx->params.p, cosmo, 64, hipMemcpyHostToDevice, s));
  HIPCHK(ctx->pk_a.reserve((size_t)n * 2));
  std::vector<double> aw((size_t)n * 2, 0.0);
  memcpy(aw.data(), a, (size_t)n * 8);
  HIPCHK(hipMemcpyAsync(ctx->pk_a.p, aw.data(), aw.size() * 8, hipMemcpyHostToDevice, s));
  HIPCHK(ctx->chitab.reserve((size_t)NCHI * 2));
  HIPCHK(ctx->coef.reserve((size_t)n * NCOEF));
  HIPCHK(ctx->crec.reserve(NCREC));
  HIPCHK(ctx->R.reserve((size_t)n));
  TablesArgs ta{};
  ta.params = ctx->params.p;
  ta.B = 1;
  ta.Na = (int)n;
  ta.a_list = ctx->pk_a.p;
  ta.a_wgt = ctx->pk_a.p + n;
  resident_tables(ctx, ta);
  ta.Tpad = 1;
  ta.chitab = ctx->chitab.p;
  ta.coef = ctx->coef.p;
  ta.crec = ctx->crec.p;
  ta.R = ctx->R.p;
  if (fp64_out) {  // FP64 per-a values: [1024 table slots][n][8]
    const size_t stride = 1024 + (size_t)8 * n;
    HIPCHK(ctx->dbg.reserve(stride));
    ta.dbg = ctx->dbg.p;
    ta.dbg_stride = (int)stride;
  }
  HIPCHK(launch_tables(0, ta, s));
  return JC_OK;
}

}  // namespace jch

extern "C" {

const char* jc_last_error(void) { return g_err.c_str(); }
int jc_version(void) { return 2; }

int jc_ctx_create(int device, jc_ctx** out) {
  if (!out) return fail(JC_E_ARG, "null output pointer");
  int ndev = 0;
  HIPCHK(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) return fail(JC_E_ARG, "device %d not present (%d devices)", device, ndev);
  HIPCHK(hipSetDevice(device));
  std::unique_ptr<jc_ctx> ctx(new jc_ctx);
  ctx->device = device;
  int ncu = 0;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && ncu > 0)
    ctx->num_cu = ncu;
  HIPCHK(hipStreamCreateWithFlags(&ctx->own, hipStreamNonBlocking));
  ctx->stream = ctx->own;
  int rc = scratch_address(device, &ctx->scratch);
  if (rc == JC_OK) rc = ctx_init_tables(ctx.get());
  if (rc) {
    (void)hipStreamDestroy(ctx->own);
    return rc;
  }
  *out = ctx.release();
  return JC_OK;
}

int jc_ctx_destroy(jc_ctx* ctx) {
  if (!ctx) return JC_OK;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  hipStream_t own = ctx->own;
  for (auto& r : ctx->recs) {
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  for (auto e : ctx->ev_pool) (void)hipEventDestroy(e);
  delete ctx;
  if (own) (void)hipStreamDestroy(own);
  return JC_OK;
}

int jc_ctx_set_stream(jc_ctx* ctx, void* stream) {
  if (!ctx) return fail(JC_E_ARG, "null context");
  ctx->stream = (hipStream_t)stream;  // NULL is the HIP null stream
  return JC_OK;
}

int jc_ctx_reset_stream(jc_ctx* ctx) {
  if (!ctx) return fail(JC_E_ARG, "null context");
  ctx->stream = ctx->own;
  return JC_OK;
}

int jc_synchronize(jc_ctx* ctx) {
  if (!ctx) return fail(JC_E_ARG, "null context");
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return JC_OK;
}

int jc_tracers_register(jc_ctx* ctx, const jc_tracer_desc* tr, int64_t n, int64_t* set_id) {
  return jc_tracers_register_catalogs(ctx, tr, n, nullptr, nullptr, set_id);
}

int jc_tracers_register_catalogs(jc_ctx* ctx, const jc_tracer_desc* tr, int64_t n, const double* const* catalogs,
                                 const int64_t* catalog_len, int64_t* set_id) {
  return jc_tracers_register_host(ctx, tr, n, catalogs, catalog_len, nullptr, nullptr, set_id);
}

int jc_tracers_register_host(jc_ctx* ctx, const jc_tracer_desc* tr, int64_t n, const double* const* catalogs,
                             const int64_t* catalog_len, const jc_pz_fn* fns, void* const* user, int64_t* set_id) {
  if (!ctx || !tr || !set_id) return fail(JC_E_ARG, "null argument");
  if (n < 1 || n > MAXT) return fail(JC_E_UNSUPPORTED, "between 1 and %d tracers are supported, got %lld", MAXT, (long long)n);
  std::unique_ptr<TracerSet> ts(new TracerSet);
  ts->T = (int)n;
  ts->Tpad = (int)((n + 3) / 4 * 4);
  for (int64_t i = 0; i < n; ++i) {
    const jc_tracer_desc& t = tr[i];
    if (t.kind != JC_TRACER_WL && t.kind != JC_TRACER_NC) return fail(JC_E_UNSUPPORTED, "tracer %lld: unknown kind %d", (long long)i, t.kind);
    if (t.nz_kind < JC_NZ_SMAIL || t.nz_kind > JC_NZ_HOST) return fail(JC_E_UNSUPPORTED, "tracer %lld: unknown n(z) kind %d", (long long)i, t.nz_kind);
    if (t.bias_kind < JC_BIAS_CONST || t.bias_kind > JC_BIAS_NONE) return fail(JC_E_UNSUPPORTED, "tracer %lld: unknown bias kind %d", (long long)i, t.bias_kind);
    if (t.kind == JC_TRACER_NC && t.bias_kind == JC_BIAS_NONE) return fail(JC_E_ARG, "tracer %lld: number counts need a bias model", (long long)i);
    if (!(t.zmax > 0.0) || !std::isfinite(t.zmax)) return fail(JC_E_ARG, "tracer %lld: zmax must be positive", (long long)i);
    if (t.nz_kind == JC_NZ_SMAIL && !(t.nz_params[2] > 0.0)) return fail(JC_E_ARG, "tracer %lld: smail z0 must be positive", (long long)i);
    if (t.nz_kind == JC_NZ_GAUSS && !(t.nz_params[1] > 0.0)) return fail(JC_E_ARG, "tracer %lld: gaussian sigma must be positive", (long long)i);
    if (t.nz_kind == JC_NZ_DELTA) {
      // probes.py:83-86 and 108-111 refuse single planes for the density and NLA kernels
      if (t.kind != JC_TRACER_WL) return fail(JC_E_UNSUPPORTED, "tracer %lld: delta_nz is only supported for weak lensing", (long long)i);
      if (t.bias_kind != JC_BIAS_NONE) return fail(JC_E_UNSUPPORTED, "tracer %lld: the NLA kernel does not support delta_nz", (long long)i);
      if (t.nz_shifted) return fail(JC_E_UNSUPPORTED, "tracer %lld: shifted delta_nz is not supported", (long long)i);
      if (!(t.nz_params[0] >= 0.0) || !std::isfinite(t.nz_params[0])) return fail(JC_E_ARG, "tracer %lld: delta_nz z0 must be >= 0", (long long)i);
    }
    if (t.nz_kind == JC_NZ_HOST && (!fns || !fns[i]))
      return fail(JC_E_ARG, "tracer %lld: a JC_NZ_HOST bin needs a host n(z) function", (long long)i);
    ts->tr.push_back(t);
    ts->cat.emplace_back();
    ts->host.emplace_back(t.nz_kind == JC_NZ_HOST ? fns[i] : nullptr,
                          t.nz_kind == JC_NZ_HOST && user ? user[i] : nullptr);
    if (t.nz_kind == JC_NZ_KDE) {
      if (!catalogs || !catalog_len || !catalogs[i] || catalog_len[i] < 1)
        return fail(JC_E_ARG, "tracer %lld: kde_nz needs a catalog", (long long)i);
      if (!(t.nz_params[0] > 0.0) || !std::isfinite(t.nz_params[0]))
        return fail(JC_E_ARG, "tracer %lld: kde_nz bandwidth must be positive", (long long)i);
      const int64_t nc = catalog_len[i];
      std::vector<std::pair<double, double>> zw((size_t)nc);
      double ws = 0.0;
      for (int64_t k = 0; k < nc; ++k) {
        zw[k] = {catalogs[i][k], catalogs[i][nc + k]};
        if (!std::isfinite(zw[k].first) || !std::isfinite(zw[k].second))
          return fail(JC_E_ARG, "tracer %lld: catalog entry %lld is not finite", (long long)i, (long long)k);
        ws += zw[k].second;
      }
      if (!(ws != 0.0)) return fail(JC_E_ARG, "tracer %lld: catalog weights sum to zero", (long long)i);
      std::sort(zw.begin(), zw.end());
      std::vector<double> cz((size_t)nc), cw((size_t)nc);
      for (int64_t k = 0; k < nc; ++k) {
        cz[k] = zw[k].first;
        cw[k] = zw[k].second;
      }
      std::unique_ptr<Catalog> c(new Catalog);
      HIPCHK(c->z.upload(cz, ctx->stream));
      HIPCHK(c->w.upload(cw, ctx->stream));
      c->n = nc;
      c->wsum = ws;
      ts->cat.back() = std::move(c);
    }
    // redshift.py:27-31: norm = simps(pz_fn, 0, zmax, 256); delta_nz: 1 (redshift.py:118)
    double nrm = 1.0;
    if (t.nz_kind != JC_NZ_DELTA) {
      // the nodes and weights of simps(f, 0, zmax, 256) (integrate.py:162-199)
      const double dx = t.zmax / 256;
      std::vector<double> zs(257), pv;
      for (int q = 0; q <= 256; ++q) zs[q] = (q == 256) ? t.zmax : q * dx;
      int rc = eval_pz(ctx, ts.get(), (int)i, zs, pv);
      if (rc) return rc;
      nrm = 0.0;
      for (int q = 0; q <= 256; ++q) nrm += ((q == 0 || q == 256) ? 1.0 : ((q & 1) ? 4.0 : 2.0)) * pv[q];
      nrm *= dx / 3;
    }
    if (!(nrm > 0.0) || !std::isfinite(nrm)) return fail(JC_E_ARG, "tracer %lld: n(z) normalisation is %g", (long long)i, nrm);
    ts->norm.push_back(nrm);
    if (t.kind == JC_TRACER_WL) ts->wl_mask |= 1u << i;
  }
  const int64_t id = ctx->next_set++;
  ctx->sets[id] = std::move(ts);
  *set_id = id;
  return JC_OK;
}

int jc_tracers_release(jc_ctx* ctx, int64_t set_id) {
  if (!ctx) return fail(JC_E_ARG, "null context");
  auto it = ctx->sets.find(set_id);
  if (it == ctx->sets.end()) return fail(JC_E_ARG, "unknown tracer set %lld", (long long)set_id);
  (void)hipStreamSynchronize(ctx->stream);
  ctx->sets.erase(it);
  return JC_OK;
}

int jc_angular_cl(jc_ctx* ctx, const double* cosmo, int64_t B, int64_t set_id, const double* ell,
                  int64_t n_ell, int64_t n_z, float* cl, int32_t flags) {
  if (!cosmo || !ell || !cl) return fail(JC_E_ARG, "null array argument");
  return run_pipeline(ctx, 0, cosmo, nullptr, B, set_id, ell, n_ell, n_z, cl, nullptr, nullptr, nullptr, flags);
}

int jc_angular_cl_noalloc(jc_ctx* ctx, const double* cosmo, int64_t B, int64_t set_id, const double* ell,
                          int64_t n_ell, int64_t n_z, float* cl, int32_t flags) {
  if (!ctx) return fail(JC_E_ARG, "null context");
  if (!ctx->scratch) return fail(JC_E_HIP, "the context has no scratch address");
  Bump sc;
  sc.base = ctx->scratch;
  sc.cap = JC_NOALLOC_SCRATCH_BYTES;
  return run_pipeline(ctx, 0, cosmo, nullptr, B, set_id, ell, n_ell, n_z, cl, nullptr, nullptr, nullptr, flags, &sc);
}

int jc_angular_cl_jvp(jc_ctx* ctx, const double* cosmo, const double* tangent, int64_t B, int64_t set_id,
                      const double* ell, int64_t n_ell, int64_t n_z, float* cl, float* cl_dot, int32_t flags) {
  if (!cosmo || !tangent || !ell || !cl || !cl_dot) return fail(JC_E_ARG, "null array argument");
  return run_pipeline(ctx, 1, cosmo, tangent, B, set_id, ell, n_ell, n_z, cl, cl_dot, nullptr, nullptr, flags);
}

int jc_angular_cl_vjp(jc_ctx* ctx, const double* cosmo, int64_t B, int64_t set_id, const double* ell,
                      int64_t n_ell, int64_t n_z, const float* cl_bar, double* theta_bar, int32_t flags) {
  if (!cosmo || !ell || !cl_bar || !theta_bar) return fail(JC_E_ARG, "null array argument");
  return run_pipeline(ctx, 7, cosmo, nullptr, B, set_id, ell, n_ell, n_z, nullptr, nullptr, cl_bar, theta_bar, flags);
}

// ---------------------------------------------------------------- likelihood
int jc_sparse_inv(jc_ctx* ctx, const double* cov, int64_t P, int64_t n_ell, double* inv, double* logdet,
                  int32_t flags) {
  if (!ctx || !cov || !inv) return fail(JC_E_ARG, "null argument");
  if (P < 1 || P > 64 || n_ell < 1) return fail(JC_E_ARG, "P must be 1..64 and n_ell >= 1");
  HIPCHK(hipSetDevice(ctx->device));
  const bool dev = (flags & JC_DEVICE_PTRS) != 0;
  const size_t n = (size_t)P * P * n_ell;
  const double* d_cov = nullptr;
  int rc = to_device(ctx, ctx->lk_prec, cov, n, dev, &d_cov);
  if (rc) return rc;
  double* d_inv = inv;
  if (!dev) {
    HIPCHK(ctx->lk_part.reserve(n));
    d_inv = ctx->lk_part.p;
  }
  HIPCHK(ctx->lk_piv.reserve((size_t)P * n_ell));
  HIPCHK(launch_block_inv(d_cov, (int)P, (int)n_ell, d_inv, ctx->lk_piv.p, ctx->stream));
  std::vector<double> piv((size_t)P * n_ell);
  HIPCHK(hipMemcpyAsync(piv.data(), ctx->lk_piv.p, piv.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  if (!dev) HIPCHK(hipMemcpyAsync(inv, d_inv, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  double ld = 0.0;
  for (int64_t l = 0; l < n_ell; ++l)
    for (int64_t k = 0; k < P; ++k) {
      const double pk = piv[(size_t)l * P + k];
      if (!(fabs(pk) > 0.0) || !std::isfinite(pk))
        return fail(JC_E_ARG, "covariance block %lld is singular", (long long)l);
      ld += log(fabs(pk));
    }
  if (logdet) *logdet = ld;
  return JC_OK;
}

int jc_gaussian_loglike(jc_ctx* ctx, const float* mu, int64_t B, int64_t P, int64_t n_ell, const double* data,
                        const double* prec, double c0, double* loglike, float* dloglike_dmu, int32_t flags) {
  if (!ctx || !mu || !data || !prec || !loglike) return fail(JC_E_ARG, "null argument");
  if (B < 1 || P < 1 || P > 64 || n_ell < 1) return fail(JC_E_ARG, "B >= 1, P in 1..64 and n_ell >= 1 required");
  HIPCHK(hipSetDevice(ctx->device));
  const bool dev = (flags & JC_DEVICE_PTRS) != 0;
  hipStream_t s = ctx->stream;
  const size_t nmu = (size_t)B * P * n_ell;
  const float* d_mu = mu;
  if (!dev) {
    HIPCHK(ctx->lk_mu.reserve(nmu));
    HIPCHK(hipMemcpyAsync(ctx->lk_mu.p, mu, nmu * sizeof(float), hipMemcpyHostToDevice, s));
    d_mu = ctx->lk_mu.p;
  }
  const double *d_data = nullptr, *d_prec = nullptr;
  int rc = to_device(ctx, ctx->lk_data, data, (size_t)P * n_ell, dev, &d_data);
  if (rc) return rc;
  rc = to_device(ctx, ctx->lk_prec, prec, (size_t)P * P * n_ell, dev, &d_prec);
  if (rc) return rc;
  float* d_s = dloglike_dmu;
  if (!dev && dloglike_dmu) {
    HIPCHK(ctx->lk_s.reserve(nmu));
    d_s = ctx->lk_s.p;
  }
  double* d_ll = loglike;
  if (!dev) {
    HIPCHK(ctx->lk_out.reserve((size_t)B));
    d_ll = ctx->lk_out.p;
  }
  HIPCHK(ctx->lk_part.reserve((size_t)B * n_ell));
  HIPCHK(launch_loglike(d_mu, (int)B, (int)P, (int)n_ell, d_data, d_prec, c0, d_s, ctx->lk_part.p, d_ll, s));
  if (!dev) {
    HIPCHK(hipMemcpyAsync(loglike, d_ll, (size_t)B * sizeof(double), hipMemcpyDeviceToHost, s));
    if (dloglike_dmu) HIPCHK(hipMemcpyAsync(dloglike_dmu, d_s, nmu * sizeof(float), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
  }
  return JC_OK;
}

int jc_gaussian_loglike_grad(jc_ctx* ctx, const double* cosmo, int64_t B, int64_t set_id, const double* ell,
                             int64_t n_ell, int64_t n_z, const double* data, const double* prec, double c0,
                             double* loglike, double* grad, int32_t flags) {
  if (!ctx || !cosmo || !ell || !data || !prec || !loglike || !grad) return fail(JC_E_ARG, "null argument");
  auto it = ctx->sets.find(set_id);
  if (it == ctx->sets.end()) return fail(JC_E_ARG, "unknown tracer set %lld", (long long)set_id);
  if (B < 1 || n_ell < 1) return fail(JC_E_ARG, "B and n_ell must be >= 1");
  HIPCHK(hipSetDevice(ctx->device));
  const bool dev = (flags & JC_DEVICE_PTRS) != 0;
  if (!dev) {
    int rc = check_cosmo_host(cosmo, B);
    if (rc) return rc;
  }
  hipStream_t s = ctx->stream;
  const int T = it->second->T, P = T * (T + 1) / 2;
  const double *d_cosmo = nullptr, *d_ell = nullptr, *d_data = nullptr, *d_prec = nullptr;
  int rc = to_device(ctx, ctx->lk_cosmo, cosmo, (size_t)B * 8, dev, &d_cosmo);
  if (rc) return rc;
  rc = to_device(ctx, ctx->lk_ell, ell, (size_t)n_ell, dev, &d_ell);
  if (rc) return rc;
  rc = to_device(ctx, ctx->lk_data, data, (size_t)P * n_ell, dev, &d_data);
  if (rc) return rc;
  rc = to_device(ctx, ctx->lk_prec, prec, (size_t)P * P * n_ell, dev, &d_prec);
  if (rc) return rc;
  const size_t nmu = (size_t)B * P * n_ell;
  HIPCHK(ctx->lk_mu.reserve(nmu));
  HIPCHK(ctx->lk_s.reserve(nmu));
  HIPCHK(ctx->lk_part.reserve((size_t)B * n_ell));
  double* d_ll = loglike;
  double* d_g = grad;
  if (!dev) {
    HIPCHK(ctx->lk_out.reserve((size_t)B));
    HIPCHK(ctx->lk_grad.reserve((size_t)B * 7));
    d_ll = ctx->lk_out.p;
    d_g = ctx->lk_grad.p;
  }
  // the mean, the log-likelihood and its derivative w.r.t. the mean, then
  // the adjoint of the mean with that cotangent
  rc = run_pipeline(ctx, 0, d_cosmo, nullptr, B, set_id, d_ell, n_ell, n_z, ctx->lk_mu.p, nullptr, nullptr, nullptr,
                    JC_DEVICE_PTRS);
  if (rc) return rc;
  HIPCHK(launch_loglike(ctx->lk_mu.p, (int)B, P, (int)n_ell, d_data, d_prec, c0, ctx->lk_s.p, ctx->lk_part.p, d_ll, s));
  rc = run_pipeline(ctx, 7, d_cosmo, nullptr, B, set_id, d_ell, n_ell, n_z, nullptr, nullptr, ctx->lk_s.p, d_g,
                    JC_DEVICE_PTRS);
  if (rc) return rc;
  if (!dev) {
    HIPCHK(hipMemcpyAsync(loglike, d_ll, (size_t)B * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(grad, d_g, (size_t)B * 7 * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
  }
  return JC_OK;
}

int jc_dense_cholesky(jc_ctx* ctx, const double* cov, int64_t N, double* L, double* logdet, int32_t flags) {
  if (!ctx || !cov || !L) return fail(JC_E_ARG, "null argument");
  if (N < 1 || N > 65535) return fail(JC_E_ARG, "N must be 1..65535");
  HIPCHK(hipSetDevice(ctx->device));
  const bool dev = (flags & JC_DEVICE_PTRS) != 0;
  hipStream_t s = ctx->stream;
  const size_t n = (size_t)N * N;
  double* A = L;
  if (dev) {
    if (A != cov) HIPCHK(hipMemcpyAsync(A, cov, n * sizeof(double), hipMemcpyDeviceToDevice, s));
  } else {
    HIPCHK(ctx->dn_mat.reserve(n));
    A = ctx->dn_mat.p;
    HIPCHK(hipMemcpyAsync(A, cov, n * sizeof(double), hipMemcpyHostToDevice, s));
  }
  HIPCHK(ctx->dn_dlog.reserve((size_t)N));
  HIPCHK(ctx->dn_bad.reserve(1));
  HIPCHK(hipMemsetAsync(ctx->dn_bad.p, 0, sizeof(int), s));
  HIPCHK(launch_dense_cholesky(A, N, ctx->dn_dlog.p, ctx->dn_bad.p, s));
  std::vector<double> dl((size_t)N);
  int bad = 0;
  HIPCHK(hipMemcpyAsync(dl.data(), ctx->dn_dlog.p, (size_t)N * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(&bad, ctx->dn_bad.p, sizeof(int), hipMemcpyDeviceToHost, s));
  if (!dev) HIPCHK(hipMemcpyAsync(L, A, n * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  if (bad) return fail(JC_E_ARG, "the covariance is not positive definite");
  double ld = 0.0;
  for (int64_t i = 0; i < N; ++i) ld += dl[(size_t)i];
  if (logdet) *logdet = 2.0 * ld;
  return JC_OK;
}

int jc_dense_solve(jc_ctx* ctx, const double* L, int64_t N, const double* R, int64_t nrhs, int32_t mode,
                   double* X, int32_t flags) {
  if (!ctx || !L || !R || !X) return fail(JC_E_ARG, "null argument");
  if (N < 1 || N > 65535 || nrhs < 1) return fail(JC_E_ARG, "N in 1..65535 and nrhs >= 1 required");
  if (mode < 0 || mode > 2) return fail(JC_E_ARG, "mode must be 0 (L^-1), 1 (L^-T) or 2 (C^-1)");
  HIPCHK(hipSetDevice(ctx->device));
  const bool dev = (flags & JC_DEVICE_PTRS) != 0;
  hipStream_t s = ctx->stream;
  const size_t n = (size_t)nrhs * N;
  const double* d_L = nullptr;
  int rc = to_device(ctx, ctx->dn_mat, L, (size_t)N * N, dev, &d_L);
  if (rc) return rc;
  double* Z = X;
  if (!dev) {
    HIPCHK(ctx->dn_z.reserve(n));
    Z = ctx->dn_z.p;
  }
  if (Z != R) HIPCHK(hipMemcpyAsync(Z, R, n * sizeof(double), dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s));
  if (mode != 1) HIPCHK(launch_dense_solve(d_L, N, Z, nrhs, 0, s));
  if (mode != 0) HIPCHK(launch_dense_solve(d_L, N, Z, nrhs, 1, s));
  if (!dev) {
    HIPCHK(hipMemcpyAsync(X, Z, n * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
  }
  return JC_OK;
}

int jc_gaussian_loglike_dense(jc_ctx* ctx, const float* mu, int64_t B, int64_t N, const double* data,
                              const double* L, double c0, double* loglike, float* dloglike_dmu, int32_t flags) {
  if (!ctx || !mu || !data || !L || !loglike) return fail(JC_E_ARG, "null argument");
  if (B < 1 || N < 1 || N > 65535) return fail(JC_E_ARG, "B >= 1 and N in 1..65535 required");
  HIPCHK(hipSetDevice(ctx->device));
  const bool dev = (flags & JC_DEVICE_PTRS) != 0;
  hipStream_t s = ctx->stream;
  const size_t nmu = (size_t)B * N;
  const float* d_mu = mu;
  if (!dev) {
    HIPCHK(ctx->lk_mu.reserve(nmu));
    HIPCHK(hipMemcpyAsync(ctx->lk_mu.p, mu, nmu * sizeof(float), hipMemcpyHostToDevice, s));
    d_mu = ctx->lk_mu.p;
  }
  const double *d_data = nullptr, *d_L = nullptr;
  int rc = to_device(ctx, ctx->lk_data, data, (size_t)N, dev, &d_data);
  if (rc) return rc;
  rc = to_device(ctx, ctx->dn_mat, L, (size_t)N * N, dev, &d_L);
  if (rc) return rc;
  float* d_s = dloglike_dmu;
  if (!dev && dloglike_dmu) {
    HIPCHK(ctx->lk_s.reserve(nmu));
    d_s = ctx->lk_s.p;
  }
  double* d_ll = loglike;
  if (!dev) {
    HIPCHK(ctx->lk_out.reserve((size_t)B));
    d_ll = ctx->lk_out.p;
  }
  HIPCHK(ctx->dn_z.reserve(nmu));
  HIPCHK(launch_dense_loglike(d_mu, B, N, d_data, d_L, c0, ctx->dn_z.p, d_ll, d_s, s));
  if (!dev) {
    HIPCHK(hipMemcpyAsync(loglike, d_ll, (size_t)B * sizeof(double), hipMemcpyDeviceToHost, s));
    if (dloglike_dmu) HIPCHK(hipMemcpyAsync(dloglike_dmu, d_s, nmu * sizeof(float), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
  }
  return JC_OK;
}

int jc_gaussian_loglike_grad_dense(jc_ctx* ctx, const double* cosmo, int64_t B, int64_t set_id, const double* ell,
                                   int64_t n_ell, int64_t n_z, const double* data, const double* L, double c0,
                                   double* loglike, double* grad, int32_t flags) {
  if (!ctx || !cosmo || !ell || !data || !L || !loglike || !grad) return fail(JC_E_ARG, "null argument");
  auto it = ctx->sets.find(set_id);
  if (it == ctx->sets.end()) return fail(JC_E_ARG, "unknown tracer set %lld", (long long)set_id);
  if (B < 1 || n_ell < 1) return fail(JC_E_ARG, "B and n_ell must be >= 1");
  const int T = it->second->T, P = T * (T + 1) / 2;
  const int64_t N = (int64_t)P * n_ell;
  if (N > 65535) return fail(JC_E_ARG, "pairs x ell must be <= 65535");
  HIPCHK(hipSetDevice(ctx->device));
  const bool dev = (flags & JC_DEVICE_PTRS) != 0;
  if (!dev) {
    int rc = check_cosmo_host(cosmo, B);
    if (rc) return rc;
  }
  hipStream_t s = ctx->stream;
  const double *d_cosmo = nullptr, *d_ell = nullptr, *d_data = nullptr, *d_L = nullptr;
  int rc = to_device(ctx, ctx->lk_cosmo, cosmo, (size_t)B * 8, dev, &d_cosmo);
  if (rc) return rc;
  rc = to_device(ctx, ctx->lk_ell, ell, (size_t)n_ell, dev, &d_ell);
  if (rc) return rc;
  rc = to_device(ctx, ctx->lk_data, data, (size_t)N, dev, &d_data);
  if (rc) return rc;
  rc = to_device(ctx, ctx->dn_mat, L, (size_t)N * N, dev, &d_L);
  if (rc) return rc;
  const size_t nmu = (size_t)B * N;
  HIPCHK(ctx->lk_mu.reserve(nmu));
  HIPCHK(ctx->lk_s.reserve(nmu));
  HIPCHK(ctx->dn_z.reserve(nmu));
  double* d_ll = loglike;
  double* d_g = grad;
  if (!dev) {
    HIPCHK(ctx->lk_out.reserve((size_t)B));
    HIPCHK(ctx->lk_grad.reserve((size_t)B * 7));
    d_ll = ctx->lk_out.p;
    d_g = ctx->lk_grad.p;
  }
  // the mean, the log-likelihood and its derivative w.r.t. the mean through
  // the Cholesky factor, then the adjoint of the mean with that cotangent
  rc = run_pipeline(ctx, 0, d_cosmo, nullptr, B, set_id, d_ell, n_ell, n_z, ctx->lk_mu.p, nullptr, nullptr, nullptr,
                    JC_DEVICE_PTRS);
  if (rc) return rc;
  HIPCHK(launch_dense_loglike(ctx->lk_mu.p, B, N, d_data, d_L, c0, ctx->dn_z.p, d_ll, ctx->lk_s.p, s));
  rc = run_pipeline(ctx, 7, d_cosmo, nullptr, B, set_id, d_ell, n_ell, n_z, nullptr, nullptr, ctx->lk_s.p, d_g,
                    JC_DEVICE_PTRS);
  if (rc) return rc;
  if (!dev) {
    HIPCHK(hipMemcpyAsync(loglike, d_ll, (size_t)B * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(grad, d_g, (size_t)B * 7 * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
  }
  return JC_OK;
}

int jc_matter_power(jc_ctx* ctx, const double* cosmo, const double* k, const double* a, int64_t n,
                    double* pk_lin, double* pk_nl) {
  if (!ctx || !cosmo || !k || !a || !pk_lin || !pk_nl) return fail(JC_E_ARG, "null argument");
  int rc = tables_on_alist(ctx, cosmo, a, n);
  if (rc) return rc;
  for (int64_t i = 0; i < n; ++i)
    if (!(k[i] > 0.0)) return fail(JC_E_ARG, "k must be positive");
  hipStream_t s = ctx->stream;
  HIPCHK(ctx->pk_k.reserve((size_t)n));
  HIPCHK(ctx->pk_out.reserve((size_t)n * 2));
  HIPCHK(hipMemcpyAsync(ctx->pk_k.p, k, (size_t)n * 8, hipMemcpyHostToDevice, s));
  PkArgs pa{};
  pa.n = (int)n;
  pa.k = ctx->pk_k.p;
  pa.coef = ctx->coef.p;
  pa.crec = ctx->crec.p;
  pa.plin = ctx->pk_out.p;
  pa.pnl = ctx->pk_out.p + n;
  HIPCHK(launch_pk_eval(pa, s));
  HIPCHK(hipMemcpyAsync(pk_lin, pa.plin, (size_t)n * 8, hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(pk_nl, pa.pnl, (size_t)n * 8, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  return JC_OK;
}

int jc_background(jc_ctx* ctx, const double* cosmo, const double* a, int64_t n, double* chi, double* growth) {
  if (!ctx || !cosmo || !a) return fail(JC_E_ARG, "null argument");
  int rc = tables_on_alist(ctx, cosmo, a, n);
  if (rc) return rc;
  std::vector<float> cf((size_t)n * NCOEF);
  HIPCHK(hipMemcpyAsync(cf.data(), ctx->coef.p, cf.size() * 4, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  for (int64_t i = 0; i < n; ++i) {
    if (chi) chi[i] = cf[(size_t)i * NCOEF + CF_CHI];
    if (growth) growth[i] = cf[(size_t)i * NCOEF + CF_GROWTH];
  }
  return JC_OK;
}

int jc_profile_enable(jc_ctx* ctx, int32_t on) {
  if (!ctx) return fail(JC_E_ARG, "null context");
  ctx->prof = on != 0;
  return JC_OK;
}

int jc_profile_select(jc_ctx* ctx, uint32_t kernels) {
  if (!ctx) return fail(JC_E_ARG, "null context");
  if (kernels == 0 || (kernels >> JC_NUM_KERNELS) != 0)
    return fail(JC_E_ARG, "kernel mask must be a non-empty subset of the %d kernel bits", JC_NUM_KERNELS);
  ctx->prof_kernels = kernels;
  return JC_OK;
}

int jc_profile_read(jc_ctx* ctx, double* total_ms, int64_t* launches) {
  if (!ctx || !total_ms || !launches) return fail(JC_E_ARG, "null argument");
  for (int k = 0; k < JC_NUM_KERNELS; ++k) {
    total_ms[k] = 0.0;
    launches[k] = 0;
  }
  if (!ctx->recs.empty()) HIPCHK(hipStreamSynchronize(ctx->stream));
  for (auto& r : ctx->recs) {
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, r.a, r.b));
    if (r.kind >= 0 && r.kind < JC_NUM_KERNELS) {
      total_ms[r.kind] += ms;
      launches[r.kind] += 1;
    }
    ctx->ev_pool.push_back(r.a);
    ctx->ev_pool.push_back(r.b);
  }
  ctx->recs.clear();
  return JC_OK;
}

int jc_set_tiles(jc_ctx* ctx, int32_t waves, int32_t ctile) {
  if (!ctx) return fail(JC_E_ARG, "null context");
  if (waves != 0 && waves != 1 && waves != 2 && waves != 4) return fail(JC_E_ARG, "waves_per_block must be 0, 1, 2 or 4");
  if (ctile != 0 && ctile != 1 && ctile != 4 && ctile != 8 && ctile != 16) return fail(JC_E_ARG, "wl_ctile must be 0, 1, 4, 8 or 16");
  ctx->waves = waves;
  ctx->wl_ctile = ctile;
  return JC_OK;
}

int jc_set_tables_variant(jc_ctx* ctx, int32_t variant) {
  if (!ctx) return fail(JC_E_ARG, "null context");
  if (variant < 0 || variant > 3) return fail(JC_E_ARG, "tables variant must be 0..3");
  ctx->tables_variant = variant;
  return JC_OK;
}

int jc_set_models(jc_ctx* ctx, int32_t transfer, int32_t nonlinear) {
  if (!ctx) return fail(JC_E_ARG, "null context");
  if (transfer < -1 || transfer > JC_TRANSFER_EH_NOWIGGLE) return fail(JC_E_UNSUPPORTED, "unknown transfer model %d", transfer);
  if (nonlinear < -1 || nonlinear > JC_HALOFIT_SMITH2003) return fail(JC_E_UNSUPPORTED, "unknown halofit model %d", nonlinear);
  if (transfer >= 0) ctx->transfer_model = transfer;
  if (nonlinear >= 0) ctx->halofit_model = nonlinear;
  return JC_OK;
}

int jc_set_growth(jc_ctx* ctx, int32_t model, double gamma) {
  if (!ctx) return fail(JC_E_ARG, "null context");
  if (model != JC_GROWTH_ODE && model != JC_GROWTH_GAMMA) return fail(JC_E_UNSUPPORTED, "unknown growth model %d", model);
  if (model == JC_GROWTH_GAMMA && !(std::fabs(gamma) <= 10.0)) return fail(JC_E_ARG, "growth index %g out of range", gamma);
  ctx->growth_model = model;
  ctx->gamma = model == JC_GROWTH_GAMMA ? gamma : 0.0;
  return JC_OK;
}

int jc_background_ex(jc_ctx* ctx, const double* cosmo, const double* a, int64_t n, double* chi, double* growth,
                     double* rate, double* fk) {
  if (!ctx || !cosmo || !a) return fail(JC_E_ARG, "null argument");
  if (n > (1 << 22)) return fail(JC_E_ARG, "at most 4194304 scale factors per call");
  int rc = tables_on_alist(ctx, cosmo, a, n, true);
  if (rc) return rc;
  HIPCHK(launch_background_ex(ctx->params.p, ctx->pk_a.p, (int)n, ctx->atab_g.p, ctx->growth_model, ctx->gamma,
                              ctx->dbg.p, ctx->stream));
  std::vector<double> d((size_t)n * 8);
  HIPCHK(hipMemcpyAsync(d.data(), ctx->dbg.p + 1024, d.size() * 8, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  for (int64_t i = 0; i < n; ++i) {
    const double* r = &d[(size_t)i * 8];
    if (chi) chi[i] = r[0];
    if (growth) growth[i] = r[1];
    if (rate) rate[i] = r[6];
    if (fk) fk[i] = r[7];
  }
  return JC_OK;
}

int jc_sigma8_norm(jc_ctx* ctx, const double* cosmo, double* pknorm) {
  if (!ctx || !cosmo || !pknorm) return fail(JC_E_ARG, "null argument");
  const double a1 = 1.0;
  int rc = tables_on_alist(ctx, cosmo, &a1, 1, true);
  if (rc) return rc;
  // slot 640 of the FP64 table block: sigma8^2 / sigma^2(R = 8) (power.py:41)
  HIPCHK(hipMemcpyAsync(pknorm, ctx->dbg.p + 640, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return JC_OK;
}

int jc_set_forward_options(jc_ctx* ctx, int32_t transfer_table, int32_t matrix_cores) {
  if (!ctx) return fail(JC_E_ARG, "null context");
  if (transfer_table < -1 || transfer_table > 1 || matrix_cores < -1 || matrix_cores > 1)
    return fail(JC_E_ARG, "forward options must be -1, 0 or 1");
  if (transfer_table >= 0) ctx->transfer_table = transfer_table;
  if (matrix_cores >= 0) ctx->matrix_cores = matrix_cores;
  return JC_OK;
}

int jc_get_tiles(jc_ctx* ctx, int32_t* waves, int32_t* ctile) {
  if (!ctx || !waves || !ctile) return fail(JC_E_ARG, "null argument");
  *waves = ctx->waves;
  *ctile = ctx->wl_ctile;
  return JC_OK;
}

int jc_debug_tables(jc_ctx* ctx, const double* cosmo, int64_t set_id, int64_t n_z, double* out) {
  if (!ctx || !cosmo || !out) return fail(JC_E_ARG, "null argument");
  auto it = ctx->sets.find(set_id);
  if (it == ctx->sets.end()) return fail(JC_E_ARG, "unknown tracer set %lld", (long long)set_id);
  if (n_z < 2 || (n_z % 2)) return fail(JC_E_ARG, "N must be an even integer >= 2");
  int rc = check_cosmo_host(cosmo, 1);
  if (rc) return rc;
  HIPCHK(hipSetDevice(ctx->device));
  GridTables* g = nullptr;
  rc = build_grid(ctx, it->second.get(), (int)n_z, &g);
  if (rc) return rc;
  TracerSet* ts = it->second.get();
  hipStream_t s = ctx->stream;
  const int Na = g->Na;
  const size_t stride = 1024 + (size_t)8 * Na;
  HIPCHK(ctx->params.reserve(8));
  HIPCHK(hipMemcpyAsync(ctx->params.p, cosmo, 64, hipMemcpyHostToDevice, s));
  HIPCHK(ctx->dbg.reserve(stride));
  HIPCHK(hipMemsetAsync(ctx->dbg.p, 0, stride * 8, s));
  HIPCHK(ctx->chitab.reserve((size_t)NCHI * 2));
  HIPCHK(ctx->coef.reserve((size_t)Na * NCOEF));
  HIPCHK(ctx->crec.reserve(NCREC));
  HIPCHK(ctx->R.reserve((size_t)Na * ts->Tpad));
  TablesArgs ta{};
  ta.params = ctx->params.p;
  ta.B = 1;
  ta.Na = Na;
  ta.a_list = g->a_list.p;
  ta.a_ln = g->a_ln.p;
  ta.a_l10 = g->a_l10.p;
  ta.a_wgt = g->a_wgt.p;
  resident_tables(ctx, ta);
  ta.nLoc = g->nLoc;
  ta.loc_nb = g->loc_nb.p;
  ta.loc_col = g->loc_col.p;
  ta.loc_mode = g->loc_mode.p;
  ta.Tpad = ts->Tpad;
  ta.chitab = ctx->chitab.p;
  ta.coef = ctx->coef.p;
  ta.crec = ctx->crec.p;
  ta.R = ctx->R.p;
  ta.dbg = ctx->dbg.p;
  ta.dbg_stride = (int)stride;
  HIPCHK(launch_tables(0, ta, s));
  HIPCHK(hipMemcpyAsync(out, ctx->dbg.p, stride * 8, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  return JC_OK;
}

}  // extern "C"
