// Exercises the C ABI of libjc_limber.so from a host program, for the
// sanitizer build (host code with AddressSanitizer and UndefinedBehavior-
// Sanitizer; device code as usual): argument checks, every compute entry
// point on small inputs, and the request loop with two submitting threads,
// failure injection and teardown.  Exit status 0 when every check passed.
// Without a GPU only the checks that need no device run.
#include <unistd.h>
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <atomic>
#include <random>
#include <thread>
#include <vector>

#include "../../../include/jc_limber.h"

static int g_fail = 0;
#define CHECK(cond, ...)                                   \
  do {                                                     \
    if (!(cond)) {                                         \
      fprintf(stderr, "check failed (%s:%d): ", __FILE__, __LINE__); \
      fprintf(stderr, __VA_ARGS__);                        \
      fprintf(stderr, " [%s]\n", jc_last_error());         \
      ++g_fail;                                            \
    }                                                      \
  } while (0)

static jc_tracer_desc smail(int probe, double z0) {
  jc_tracer_desc t;
  memset(&t, 0, sizeof t);
  t.kind = JC_TRACER_WL;
  t.nz_kind = JC_NZ_SMAIL;
  t.probe = probe;
  t.bias_kind = JC_BIAS_NONE;
  t.nz_params[0] = 1.0;
  t.nz_params[1] = 2.0;
  t.nz_params[2] = z0;
  t.zmax = 10.0;
  t.gals_per_arcmin2 = 10.0;
  t.sigma_e = 0.26;
  return t;
}

static jc_tracer_desc gauss(int probe, double mu) {
  jc_tracer_desc t;
  memset(&t, 0, sizeof t);
  t.kind = JC_TRACER_NC;
  t.nz_kind = JC_NZ_GAUSS;
  t.probe = probe;
  t.bias_kind = JC_BIAS_CONST;
  t.nz_params[0] = mu;
  t.nz_params[1] = 0.1;
  t.zmax = 10.0;
  t.gals_per_arcmin2 = 10.0;
  t.bias_params[0] = 1.5;
  return t;
}

static void host_only_checks() {
  CHECK(jc_ctx_create(0, nullptr) == JC_E_ARG, "null output");
  CHECK(jc_angular_cl(nullptr, nullptr, 1, 1, nullptr, 1, 512, nullptr, 0) == JC_E_ARG, "null context");
  CHECK(jc_angular_cl_noalloc(nullptr, nullptr, 1, 1, nullptr, 1, 512, nullptr, 0) == JC_E_ARG, "null context");
  CHECK(jc_server_submit(nullptr, nullptr, nullptr) == JC_E_ARG, "null server");
  CHECK(jc_server_destroy(nullptr) == JC_OK, "destroying null");
  CHECK(jc_version() == 2, "version");
}

static void device_checks(jc_ctx* ctx) {
  std::vector<jc_tracer_desc> tr;
  const double z0s[5] = {0.5, 0.75, 1.0, 1.25, 1.5};
  for (int i = 0; i < 5; ++i) tr.push_back(smail(0, z0s[i]));
  for (int i = 0; i < 5; ++i) tr.push_back(gauss(1, 0.3 + 0.3 * i));
  int64_t sid = 0;
  CHECK(jc_tracers_register(ctx, tr.data(), 10, &sid) == JC_OK, "register");
  const int B = 3, L = 70, P = 55;
  std::vector<double> cosmo(B * 8), ell(L);
  for (int b = 0; b < B; ++b) {
    const double p[8] = {0.25 + 0.02 * b, 0.049, 0.68, 0.96, 0.8, 0.0, -1.0, 0.1 * b};
    for (int i = 0; i < 8; ++i) cosmo[b * 8 + i] = p[i];
  }
  for (int l = 0; l < L; ++l) ell[l] = 10.0 * pow(500.0, l / (L - 1.0));
  std::vector<float> cl(B * P * L), cl2(B * P * L), dot(B * P * L), cbar(B * P * L, 1.0f);
  CHECK(jc_angular_cl(ctx, cosmo.data(), B, sid, ell.data(), L, 512, cl.data(), 0) == JC_OK, "angular_cl");
  CHECK(jc_angular_cl_noalloc(ctx, cosmo.data(), B, sid, ell.data(), L, 512, cl2.data(), 0) == JC_OK, "noalloc");
  CHECK(memcmp(cl.data(), cl2.data(), cl.size() * 4) == 0, "noalloc equals angular_cl");
  CHECK(jc_angular_cl(ctx, cosmo.data(), B, sid, ell.data(), L, 511, cl.data(), 0) == JC_E_ARG, "odd n_z");
  CHECK(jc_angular_cl(ctx, cosmo.data(), B, 12345, ell.data(), L, 512, cl.data(), 0) == JC_E_ARG, "unknown set");
  std::vector<double> tan(B * 7, 0.5), th(B * 7);
  CHECK(jc_angular_cl_jvp(ctx, cosmo.data(), tan.data(), B, sid, ell.data(), L, 256, cl.data(), dot.data(), 0) ==
            JC_OK, "jvp");
  CHECK(jc_angular_cl_vjp(ctx, cosmo.data(), B, sid, ell.data(), L, 256, cbar.data(), th.data(), 0) == JC_OK, "vjp");
  for (double v : th) CHECK(std::isfinite(v), "finite vjp");
  // likelihood algebra on the C_ell of cosmology 0
  std::vector<double> cov((size_t)P * P * L, 0.0), inv(cov.size()), data((size_t)P * L);
  for (int p = 0; p < P; ++p)
    for (int l = 0; l < L; ++l) {
      cov[((size_t)p * P + p) * L + l] = 1e-14 + 1e-3 * (double)cl[(size_t)p * L + l] * cl[(size_t)p * L + l];
      data[(size_t)p * L + l] = 1.01 * cl[(size_t)p * L + l];
    }
  double logdet = 0.0, ll[B], g[B * 7];
  CHECK(jc_sparse_inv(ctx, cov.data(), P, L, inv.data(), &logdet, 0) == JC_OK, "sparse_inv");
  CHECK(jc_gaussian_loglike(ctx, cl.data(), B, P, L, data.data(), inv.data(), 0.0, ll, nullptr, 0) == JC_OK, "loglike");
  CHECK(jc_gaussian_loglike_grad(ctx, cosmo.data(), B, sid, ell.data(), L, 256, data.data(), inv.data(), 0.0, ll, g,
                                 0) == JC_OK, "loglike_grad");
  // dense algebra on a small matrix
  const int N = 70;
  std::vector<double> C((size_t)N * N), Lf((size_t)N * N), R(2 * N, 1.0), X(2 * N);
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j) C[(size_t)i * N + j] = (i == j ? N : 0.0) + 1.0 / (1.0 + i + j);
  CHECK(jc_dense_cholesky(ctx, C.data(), N, Lf.data(), &logdet, 0) == JC_OK, "cholesky");
  CHECK(jc_dense_solve(ctx, Lf.data(), N, R.data(), 2, 2, X.data(), 0) == JC_OK, "dense solve");
  // background and power
  std::vector<double> a(40), k(40), chi(40), gr(40), rate(40), fk(40), pl(40), pn(40);
  for (int i = 0; i < 40; ++i) {
    a[i] = 0.1 + 0.9 * i / 39.0;
    k[i] = 1e-3 * pow(1e4, i / 39.0);
  }
  CHECK(jc_background_ex(ctx, cosmo.data(), a.data(), 40, chi.data(), gr.data(), rate.data(), fk.data()) == JC_OK,
        "background_ex");
  CHECK(jc_matter_power(ctx, cosmo.data(), k.data(), a.data(), 40, pl.data(), pn.data()) == JC_OK, "matter_power");
  double pkn = 0.0;
  CHECK(jc_sigma8_norm(ctx, cosmo.data(), &pkn) == JC_OK && pkn > 0.0, "sigma8_norm");

  // the request loop: two submitting threads, a failing batch, teardown
  jc_server_config cfg;
  memset(&cfg, 0, sizeof cfg);
  cfg.max_batch = 16;
  cfg.max_grad = 8;
  cfg.max_ell = 64;
  cfg.slots = 2;
  cfg.use_graph = 1;
  cfg.n_nz = 2;
  cfg.n_z[0] = 256;
  cfg.n_z[1] = 512;
  jc_server* srv = nullptr;
  CHECK(jc_server_create(ctx, sid, &cfg, &srv) == JC_OK, "server");
  if (!srv) return;
  CHECK(jc_server_inject_failure(srv, 1, JC_E_NOMEM) == JC_OK, "inject");
  const int NR = 60;
  std::vector<std::vector<float>> outs(2 * NR), cbars(2 * NR);
  std::vector<std::vector<double>> ells(2 * NR), grads(2 * NR, std::vector<double>(7));
  std::atomic<int> submitted{0};
  auto submitter = [&](int tid) {
    std::mt19937 rng(100 + tid);
    for (int i = 0; i < NR; ++i) {
      const int q = tid * NR + i;
      jc_request r;
      memset(&r, 0, sizeof r);
      for (int j = 0; j < 8; ++j) r.cosmo[j] = cosmo[j];
      r.cosmo[0] += 0.001 * (rng() % 50);
      r.n_tracers = 1 + rng() % 10;
      for (int t = 0; t < r.n_tracers; ++t) r.tracers[t] = t;
      r.n_ell = 5 + rng() % 60;
      r.n_z = (rng() & 1) ? 256 : 512;
      const int Pq = r.n_tracers * (r.n_tracers + 1) / 2;
      ells[q].resize(r.n_ell);
      for (int l = 0; l < r.n_ell; ++l) ells[q][l] = 10.0 + 37.0 * l;
      outs[q].assign((size_t)Pq * r.n_ell, 0.0f);
      r.ell = ells[q].data();
      r.cl_out = outs[q].data();
      if (rng() % 3 == 0) {
        cbars[q].assign((size_t)Pq * r.n_ell, 1.0f);
        r.want_grad = 1;
        r.cl_bar = cbars[q].data();
        r.grad_out = grads[q].data();
      }
      int64_t id = 0;
      if (jc_server_submit(srv, &r, &id) == JC_OK) submitted++;
    }
  };
  std::thread t0(submitter, 0), t1(submitter, 1);
  t0.join();
  t1.join();
  CHECK(submitted.load() == 2 * NR, "all submitted");
  CHECK(jc_server_wait(srv, 0, 120000) == JC_OK, "wait");
  std::vector<jc_completion> comps(4 * NR);
  int64_t n = 0;
  CHECK(jc_server_poll(srv, comps.data(), (int64_t)comps.size(), &n) == JC_OK, "poll");
  int ok = 0, nomem = 0;
  for (int64_t i = 0; i < n; ++i) {
    if (comps[i].status == JC_OK) ++ok;
    else if (comps[i].status == JC_E_NOMEM) ++nomem;
  }
  CHECK(n == 2 * NR && nomem >= 1 && ok + nomem == n, "completions %lld ok %d nomem %d", (long long)n, ok, nomem);
  jc_request bad;
  memset(&bad, 0, sizeof bad);
  bad.n_tracers = 3;
  bad.n_ell = 5;
  bad.n_z = 128;  // not served
  int64_t id = 0;
  CHECK(jc_server_submit(srv, &bad, &id) == JC_E_ARG, "refused request");
  CHECK(jc_server_destroy(srv) == JC_OK, "destroy server");
  CHECK(jc_tracers_release(ctx, sid) == JC_OK, "release");
}

int main() {
  host_only_checks();
  jc_ctx* ctx = nullptr;
  const int rc = jc_ctx_create(0, &ctx);
  if (rc != JC_OK) {
    printf("jc_abi_check: no device (%s): host-only checks, %d failed\n", jc_last_error(), g_fail);
    return g_fail ? 1 : 0;
  }
  device_checks(ctx);
  CHECK(jc_ctx_destroy(ctx) == JC_OK, "destroy context");
  printf("jc_abi_check: %d failed\n", g_fail);
  // Everything this library allocated is released above.  Leave without the
  // HIP runtime's static destructors: under the host AddressSanitizer the
  // runtime's exit-time frees can recycle quarantined device allocations
  // after the device runtime has unloaded, which the sanitizer's device
  // allocator rejects with a CHECK failure of its own.
  fflush(stdout);
  fflush(stderr);
  _exit(g_fail ? 1 : 0);
}
