// Continuous-batching request loop for served C_ell / gradient requests.
//
// A scheduler thread owns `slots` batches in flight.  Each slot has its own
// stream, fixed-capacity device state (one table block per served Simpson
// size) and pinned staging.  Forming a batch packs the admitted requests'
// inputs (cosmologies, job and column descriptors, cotangents) into the
// pinned staging area; one host-to-device copy moves it, an unpack kernel
// scatters it into the fixed device layout, then the persistent kernels run
// (grids sized for capacity, early exit past the live counts), so the whole
// kernel sequence is captured once per slot as a hipGraph and replayed per
// batch.  One device-to-host copy returns the packed outputs.
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <cmath>
#include <thread>

#include "jc_host.hpp"

using namespace jc;
using namespace jch;

namespace {

constexpr int MAXG = 4;         // served Simpson sizes
constexpr int MAXSEC = 64;      // staging sections
constexpr int TPAD_POOL = 12;   // pool of up to 10 tracers, padded
// a-range slices of the ragged Limber kernels (the live count is chosen per
// batch on the device, ragged_slices)
constexpr int SLICES_FWD = 8;
constexpr int SLICES_ADJ = 16;

struct StageSec {
  int64_t dst, src, bytes;
};
struct StageHdr {
  int32_t nsec, pad;
  StageSec sec[MAXSEC];
};

// scatter the staging sections into the slot arena (16-byte words)
__global__ void k_unpack(const unsigned char* __restrict__ stage, unsigned char* __restrict__ arena) {
  const StageHdr* h = reinterpret_cast<const StageHdr*>(stage);
  const int ns = h->nsec;
  for (int s = 0; s < ns; ++s) {
    const StageSec sc = h->sec[s];
    const int4* src = reinterpret_cast<const int4*>(stage + sc.src);
    int4* dst = reinterpret_cast<int4*>(arena + sc.dst);
    const int64_t nw = sc.bytes / 16;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nw; i += (int64_t)gridDim.x * blockDim.x)
      dst[i] = src[i];
  }
}

double now_us() {
  using namespace std::chrono;
  static const steady_clock::time_point t0 = steady_clock::now();
  return duration<double, std::micro>(steady_clock::now() - t0).count();
}

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// Launch grid bounds of a batch (the arena is always laid out for the full
// capacity; small batches replay graphs captured with small grids, so they
// do not dispatch the capacity's worth of empty blocks).
struct Caps {
  int C = 0, Gd = 0, colcap = 0, gcolcap = 0;
  int64_t outcap = 0;
};
constexpr int SMALL_BATCH = 32;  // requests of the small-grid graphs

struct Req {
  int64_t id = 0;
  jc_request r{};
  std::vector<double> ell;
  std::vector<float> cbar;
  int g = 0;     // served-size group
  int P = 0;
  size_t footprint = 0;
  double t_submit = 0.0;
  double t_launch = 0.0;  // when its batch was launched
  int64_t out_off = 0;
  int64_t th_row = -1;
};

// Per served Simpson size: device sub-buffers of a slot.
struct GroupBufs {
  double* params = nullptr;   // [C][8]
  double* gparams = nullptr;  // [Gd][8]
  double* gseeds = nullptr;   // [Gd][8][7]
  RJob* jobs = nullptr;       // [C]
  RJob* gjobs = nullptr;      // [Gd]
  RCol* cols = nullptr;       // [colcap]
  RCol* gcols = nullptr;      // [gcolcap]
  float *chitab = nullptr, *coef = nullptr, *crec = nullptr, *R = nullptr;
  float *chitab_d = nullptr, *coef_d = nullptr, *crec_d = nullptr, *R_d = nullptr;
  float4* ehtab = nullptr;     // [C][NEHQ] transfer tables of the forward rows
  float4* ehtab_d = nullptr;   // [Gd][NEHQ] ... of the gradient rows
  float* ehtab7 = nullptr;     // [Gd][NEHQ + 1][NEHD_ROW] their theta derivatives
  float* theta_col = nullptr;  // [SLICES_ADJ][gcolcap][7]
};

struct Slot {
  // `stream` runs the unpack, the forward pass and the output copy, then
  // ev_fwd; `stream2` the gradient pass (after the unpack) and its copy, then
  // ev_grad; ev_done follows both.  Requests without a gradient complete at
  // ev_fwd, the slot at ev_done.
  hipStream_t stream = nullptr, stream2 = nullptr;
  hipEvent_t ev_start = nullptr, ev_done = nullptr, ev_fwd = nullptr, ev_grad = nullptr, ev_unpacked = nullptr;
  bool fwd_reported = false;
  // one branch per served size in each pass: the chains of the Simpson
  // sizes are independent, so they fork from the pass's stream and join it
  hipStream_t side[2 * MAXG] = {};
  hipEvent_t ev_fork[2] = {}, ev_join[2 * MAXG] = {};
  unsigned char* arena = nullptr;  // device
  size_t arena_bytes = 0;
  unsigned char* d_stage = nullptr;
  unsigned char* h_stage = nullptr;  // pinned
  size_t stage_bytes = 0;
  int* counts = nullptr;            // device [MAXG][8]
  float* out = nullptr;             // device [outcap]
  float* part = nullptr;            // device [SLICES_FWD][outcap] forward slice partials
  float* cbar = nullptr;            // device [gout cap]
  double* theta = nullptr;          // device [Gtot][7]
  // two pinned result buffers, alternated per batch: copier threads hand one
  // batch's results to the callers while the slot runs the next batch.
  // drain[b] counts the copy tasks still reading buffer b (guarded by the
  // server mutex); a batch launches into a buffer whose count is 0.
  float* h_outb[2] = {nullptr, nullptr};
  double* h_thetab[2] = {nullptr, nullptr};
  int drain[2] = {0, 0};
  int cur = 0;                      // buffer of the batch in flight
  float* h_out = nullptr;           // = h_outb[cur]
  double* h_theta = nullptr;        // = h_thetab[cur]
  GroupBufs g[MAXG];
  // captured graphs [grid size: full, small][pass][mask - 1]: one per
  // non-empty subset of served sizes (bit g = the batch has work for size g
  // in that pass), so a batch launches no kernels for absent sizes
  std::vector<hipGraphExec_t> gx[2][2];
  bool busy = false;
  std::vector<Req*> batch;
  int bsize = 0;  // requests of the batch (some may have completed early)
  int64_t out_used = 0, th_used = 0;
};

// Results of some requests of one slot buffer, to be copied to the callers'
// arrays and completed.
struct CopyTask {
  Slot* sl = nullptr;
  int buf = 0;
  std::vector<Req*> reqs;
  int status = 0;
  int bsize = 0;
};

}  // namespace

struct jc_server {
  jc_ctx* ctx = nullptr;
  TracerSet* pool = nullptr;
  jc_server_config cfg{};
  int ng = 0;
  int Nz[MAXG] = {0}, Na[MAXG] = {0};
  GridTables* grid[MAXG] = {nullptr};
  int C = 0, Gd = 0, colcap = 0, gcolcap = 0;
  int64_t outcap = 0, goutcap = 0;
  Caps full, small;
  int waves = 2, ctile = 4, tables_variant = 0;
  int transfer_model = 0, halofit_model = 0;  // fixed at creation
  int growth_model = 0;
  double gamma = 0.0;
  std::vector<std::unique_ptr<Slot>> slots;
  // the pool's largest zmax, and each weak-lensing probe's (request checks)
  double zall = 0.0;
  std::vector<double> zprobe;
  // queues
  std::mutex mu;
  std::condition_variable cv, cv_idle;
  std::deque<Req*> pending;
  std::vector<jc_completion> done;
  int64_t next_id = 1;
  int64_t outstanding = 0;
  std::atomic<bool> stop{false};
  std::thread worker;
  // result delivery
  std::mutex cmu;
  std::condition_variable ccv;
  std::deque<CopyTask> ctasks;
  bool cstop = false;
  std::vector<std::thread> copiers;
  // stats
  double st_batches = 0, st_done = 0, st_rejected = 0, st_maxb = 0, st_sumb = 0, st_busy_ms = 0,
         st_graph = 0, st_direct = 0;
  // completed requests: summed submit -> launch and launch -> completion times (us)
  double st_wait_us = 0, st_service_us = 0;
  double st_launch_us = 0;  // host time in launch_batch (packing and enqueueing)
  // test hook (jc_server_inject_failure): batches still to fail at launch
  int inject_batches = 0, inject_status = JC_E_HIP;
  std::string worker_err;
};

namespace {

size_t req_footprint(const jc_server* S, const jc_request& r, int g) {
  const int P = r.n_tracers * (r.n_tracers + 1) / 2;
  const size_t Na = (size_t)S->Na[g];
  size_t b = Na * (NCOEF + TPAD_POOL) * 4 + NCHI * 2 * 4 + NCREC * 4 + (size_t)P * r.n_ell * 4 * (1 + SLICES_FWD) +
             (size_t)r.n_ell * sizeof(RCol) + sizeof(RJob) + (size_t)NEHQ * 16 + 64;
  if (r.want_grad)
    b += 8 * (Na * (NCOEF + TPAD_POOL) * 4 + NCHI * 2 * 4 + NCREC * 4) + (size_t)P * r.n_ell * 4 +
         (size_t)r.n_ell * (sizeof(RCol) + 28 * SLICES_ADJ) + sizeof(RJob) + (size_t)NEHQ * 16 +
         (size_t)(NEHQ + 1) * NEHD_ROW * 4 + 64 + 7 * 8;
  return b;
}

// Carve a slot's device arena.  Returns total bytes (dry run when base = 0).
size_t layout_slot(jc_server* S, Slot* sl, unsigned char* base) {
  size_t off = 0;
  auto take = [&](size_t bytes) {
    size_t o = off;
    off = align_up(off + bytes, 256);
    return base ? base + o : nullptr;
  };
  sl->counts = (int*)take(MAXG * 8 * sizeof(int));
  sl->out = (float*)take((size_t)S->outcap * 4);
  sl->part = (float*)take((size_t)SLICES_FWD * S->outcap * 4);
  sl->cbar = (float*)take((size_t)S->goutcap * 4);
  sl->theta = (double*)take((size_t)S->Gd * MAXG * 7 * 8);
  for (int g = 0; g < S->ng; ++g) {
    GroupBufs& b = sl->g[g];
    const size_t Na = S->Na[g];
    b.params = (double*)take((size_t)S->C * 8 * 8);
    b.gparams = (double*)take((size_t)S->Gd * 8 * 8);
    b.gseeds = (double*)take((size_t)S->Gd * 8 * 7 * 8);
    b.jobs = (RJob*)take((size_t)S->C * sizeof(RJob));
    b.gjobs = (RJob*)take((size_t)S->Gd * sizeof(RJob));
    b.cols = (RCol*)take((size_t)S->colcap * sizeof(RCol));
    b.gcols = (RCol*)take((size_t)S->gcolcap * sizeof(RCol));
    b.chitab = (float*)take((size_t)S->C * NCHI * 2 * 4);
    b.coef = (float*)take((size_t)S->C * Na * NCOEF * 4);
    b.crec = (float*)take((size_t)S->C * NCREC * 4);
    b.R = (float*)take((size_t)S->C * Na * TPAD_POOL * 4);
    b.chitab_d = (float*)take((size_t)S->Gd * NCHI * 2 * 4 * 8);
    b.coef_d = (float*)take((size_t)S->Gd * Na * NCOEF * 4 * 8);
    b.crec_d = (float*)take((size_t)S->Gd * NCREC * 4 * 8);
    b.R_d = (float*)take((size_t)S->Gd * Na * TPAD_POOL * 4 * 8);
    b.ehtab = (float4*)take((size_t)S->C * NEHQ * sizeof(float4));
    b.ehtab_d = (float4*)take((size_t)S->Gd * NEHQ * sizeof(float4));
    b.ehtab7 = (float*)take((size_t)S->Gd * (NEHQ + 1) * NEHD_ROW * 4);
    b.theta_col = (float*)take((size_t)SLICES_ADJ * S->gcolcap * 7 * 4);
  }
  return off;
}

TablesArgs tables_args(jc_server* S, int g, const GroupBufs& b, int nd, const int* dcount, const Caps& k) {
  jc_ctx* ctx = S->ctx;
  GridTables* gt = S->grid[g];
  TablesArgs ta{};
  ta.params = nd ? b.gparams : b.params;
  ta.seeds = nd ? b.gseeds : nullptr;
  ta.B = nd ? k.Gd : k.C;  // grid bound (the live count is *dcount)
  ta.Na = S->Na[g];
  ta.a_list = gt->a_list.p;
  ta.a_wgt = gt->a_wgt.p;
  resident_tables(ctx, ta);
  ta.transfer_model = S->transfer_model;
  ta.halofit_model = S->halofit_model;
  ta.growth_model = S->growth_model;
  ta.gamma = S->gamma;
  ta.a_ln = gt->a_ln.p;
  ta.a_l10 = gt->a_l10.p;
  ta.nLoc = gt->nLoc;
  ta.loc_nb = gt->loc_nb.p;
  ta.loc_col = gt->loc_col.p;
  ta.loc_mode = gt->loc_mode.p;
  ta.Tpad = TPAD_POOL;
  ta.chitab = nd ? b.chitab_d : b.chitab;
  ta.coef = nd ? b.coef_d : b.coef;
  ta.crec = nd ? b.crec_d : b.crec;
  ta.R = nd ? b.R_d : b.R;
  ta.dB = dcount;
  // every row's transfer table over its own request's ell range
  ta.ehtab = nd ? b.ehtab_d : b.ehtab;
  ta.ehtab7 = nd ? b.ehtab7 : nullptr;
  ta.rjobs = nd ? b.gjobs : b.jobs;
  ta.rcols = nd ? b.gcols : b.cols;
  return ta;
}

// The kernels of one pass of a batch (0: forward, 1: gradient) on stream
// `ps` (captured or direct): one branch per served size on the side
// streams, forked from and joined back into `ps`.
hipError_t enqueue_pass(jc_server* S, Slot* sl, int pass, hipStream_t ps, const Caps& k, unsigned mask) {
  hipError_t e;
  // weak-lensing cosmologies per block: the tuned tile for full batches, one
  // for the small-batch graphs, whose few cosmologies would otherwise leave
  // most of the device idle (a single-request batch: 64 -> ~20 us)
  const int ctile = &k == &S->small ? 1 : S->ctile;
  const int nd = pass ? 7 : 0;
  if ((e = hipEventRecord(sl->ev_fork[pass], ps)) != hipSuccess) return e;
  // The branches are enqueued stage by stage (every size's table kernel,
  // then every size's lensing kernels, ...), largest Simpson size first: a
  // replay submits the nodes in this order, and the submission is slow
  // enough that a branch enqueued whole before the next would start its
  // chain one branch later.
  auto branch = [&](int g) { return sl->side[pass * MAXG + g]; };
  for (int stage = 0; stage < 4; ++stage) {
    for (int g = S->ng - 1; g >= 0; --g) {
      if (!((mask >> g) & 1u)) continue;
      hipStream_t st = branch(g);
      GroupBufs& b = sl->g[g];
      int* cnt = sl->counts + g * 8;
      TablesArgs ta = tables_args(S, g, b, nd, cnt + (pass ? 2 : 0), k);
      if (stage == 0) {
        if ((e = hipStreamWaitEvent(st, sl->ev_fork[pass], 0)) != hipSuccess) return e;
        if ((e = launch_tables(nd, ta, st, S->tables_variant)) != hipSuccess) return e;
      } else if (stage == 1) {
        for (auto& wg : S->grid[g]->wl) {
          WLArgs wa{};
          wa.B = ta.B;
          wa.Na = ta.Na;
          wa.nw = wg->nw;
          wa.ia = wg->ia ? 1 : 0;
          wa.src = wg->src.p;
          wa.Lidx = wg->L.p;
          wa.delta = wg->delta.p;
          wa.Nw = wg->Nw.p;
          wa.zp1 = S->grid[g]->zp1.p;
          wa.col = wg->col.p;
          wa.chitab = ta.chitab;
          wa.coef = ta.coef;
          wa.crec = ta.crec;
          wa.R = ta.R;
          wa.Tpad = TPAD_POOL;
          wa.dB = ta.dB;
          wa.rjobs = ta.rjobs;  // R rows in each request's tracer order
          if ((e = launch_wl(nd, pass ? 1 : ctile, wa, st)) != hipSuccess) return e;
        }
      } else {
        RaggedArgs ra{};
        ra.ncols = cnt + (pass ? 3 : 1);
        ra.cols = pass ? b.gcols : b.cols;
        ra.jobs = pass ? b.gjobs : b.jobs;
        ra.Na = ta.Na;
        ra.coef = ta.coef;
        ra.crec = ta.crec;
        ra.R = ta.R;
        ra.Tpad = TPAD_POOL;
        ra.out = sl->out;
        ra.cbar = sl->cbar;
        ra.theta_col = b.theta_col;
        ra.waves = S->waves;
        ra.ehtab = ta.ehtab;
        ra.ehtab7 = ta.ehtab7;
        ra.max_slices = pass ? SLICES_ADJ : SLICES_FWD;
        ra.part = pass ? nullptr : sl->part;
        ra.part_stride = pass ? (int64_t)S->gcolcap * 7 : S->outcap;
        if (stage == 2) {
          if ((e = launch_limber_ragged(nd, pass ? k.gcolcap : k.colcap, ra, st)) != hipSuccess) return e;
        } else if (pass) {
          RReduceArgs rr{};
          rr.njobs = cnt + 2;
          rr.jobs = b.gjobs;
          rr.theta_col = b.theta_col;
          rr.theta = sl->theta;
          rr.ncols = ra.ncols;
          rr.waves = ra.waves;
          rr.Na = ra.Na;
          rr.max_slices = ra.max_slices;
          rr.part_stride = ra.part_stride;
          if ((e = launch_ragged_reduce(k.Gd, rr, st)) != hipSuccess) return e;
        } else {
          // the group's output range [counts[4], counts[5]): add the slices
          if ((e = launch_ragged_slice_sum(ra.ncols, ra.waves, ra.Na, ra.max_slices, sl->part, S->outcap, cnt + 4,
                                           k.outcap, sl->out, st)) != hipSuccess)
            return e;
        }
      }
    }
  }
  for (int g = S->ng - 1; g >= 0; --g) {
    if (!((mask >> g) & 1u)) continue;
    const int br = pass * MAXG + g;
    if ((e = hipEventRecord(sl->ev_join[br], branch(g))) != hipSuccess) return e;
    if ((e = hipStreamWaitEvent(ps, sl->ev_join[br], 0)) != hipSuccess) return e;
  }
  return hipSuccess;
}
hipError_t enqueue_unpack(Slot* sl) {
  hipLaunchKernelGGL(k_unpack, dim3(64), dim3(256), 0, sl->stream, sl->d_stage, sl->arena);
  return hipGetLastError();
}

int slot_init(jc_server* S, Slot* sl) {
  HIPCHK(hipStreamCreateWithFlags(&sl->stream, hipStreamNonBlocking));
  HIPCHK(hipStreamCreateWithFlags(&sl->stream2, hipStreamNonBlocking));
  HIPCHK(hipEventCreate(&sl->ev_start));
  HIPCHK(hipEventCreate(&sl->ev_done));
  HIPCHK(hipEventCreateWithFlags(&sl->ev_fwd, hipEventDisableTiming));
  HIPCHK(hipEventCreateWithFlags(&sl->ev_grad, hipEventDisableTiming));
  HIPCHK(hipEventCreateWithFlags(&sl->ev_unpacked, hipEventDisableTiming));
  for (int p = 0; p < 2; ++p) HIPCHK(hipEventCreateWithFlags(&sl->ev_fork[p], hipEventDisableTiming));
  for (int i = 0; i < 2 * MAXG; ++i) {
    HIPCHK(hipStreamCreateWithFlags(&sl->side[i], hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&sl->ev_join[i], hipEventDisableTiming));
  }
  sl->arena_bytes = layout_slot(S, sl, nullptr);
  HIPCHK(hipMalloc(&sl->arena, sl->arena_bytes));
  layout_slot(S, sl, sl->arena);
  // staging: header + every section at capacity
  size_t st = align_up(sizeof(StageHdr), 256) + align_up(MAXG * 8 * 4, 256);
  for (int g = 0; g < S->ng; ++g)
    st += align_up((size_t)S->C * 64, 256) + align_up((size_t)S->Gd * 64, 256) +
          align_up((size_t)S->C * sizeof(RJob), 256) + align_up((size_t)S->Gd * sizeof(RJob), 256) +
          align_up((size_t)S->colcap * sizeof(RCol), 256) + align_up((size_t)S->gcolcap * sizeof(RCol), 256);
  st += align_up((size_t)S->goutcap * 4, 256);
  sl->stage_bytes = st;
  HIPCHK(hipMalloc(&sl->d_stage, st));
  HIPCHK(hipHostMalloc((void**)&sl->h_stage, st, hipHostMallocDefault));
  for (int b = 0; b < 2; ++b) {
    HIPCHK(hipHostMalloc((void**)&sl->h_outb[b], (size_t)S->outcap * 4, hipHostMallocDefault));
    HIPCHK(hipHostMalloc((void**)&sl->h_thetab[b], (size_t)S->Gd * MAXG * 7 * 8, hipHostMallocDefault));
  }
  sl->h_out = sl->h_outb[0];
  sl->h_theta = sl->h_thetab[0];
  // constant seeds of the adjoint (identity over the 7 parameters, Omega_k skipped)
  std::vector<double> seeds((size_t)S->Gd * 8 * 7, 0.0);
  static const int pidx[7] = {0, 1, 2, 3, 4, 6, 7};
  for (int c = 0; c < S->Gd; ++c)
    for (int q = 0; q < 7; ++q) seeds[((size_t)c * 8 + pidx[q]) * 7 + q] = 1.0;
  for (int g = 0; g < S->ng; ++g)
    HIPCHK(hipMemcpy(sl->g[g].gseeds, seeds.data(), seeds.size() * 8, hipMemcpyHostToDevice));
  // zero counts and an empty staging header so a replay before any batch is a no-op
  std::vector<int> z(MAXG * 8, 0);
  HIPCHK(hipMemcpy(sl->counts, z.data(), z.size() * 4, hipMemcpyHostToDevice));
  StageHdr h0{};
  HIPCHK(hipMemcpy(sl->d_stage, &h0, sizeof h0, hipMemcpyHostToDevice));
  if (S->cfg.use_graph) {
    // one graph per grid size, pass and subset of served sizes, each
    // captured on the stream that replays it
    const unsigned nmask = 1u << S->ng;
    for (int sz = 0; sz < 2; ++sz)
      for (int pass = 0; pass < 2; ++pass) {
        sl->gx[sz][pass].assign(nmask - 1, nullptr);
        for (unsigned mask = 1; mask < nmask; ++mask) {
          hipStream_t ps = pass ? sl->stream2 : sl->stream;
          hipGraph_t graph = nullptr;
          HIPCHK(hipStreamBeginCapture(ps, hipStreamCaptureModeThreadLocal));
          hipError_t e = enqueue_pass(S, sl, pass, ps, sz ? S->small : S->full, mask);
          hipError_t e2 = hipStreamEndCapture(ps, &graph);
          if (e != hipSuccess) return fail(JC_E_HIP, "graph capture: %s", hipGetErrorString(e));
          if (e2 != hipSuccess) return fail(JC_E_HIP, "graph capture end: %s", hipGetErrorString(e2));
          e = hipGraphInstantiate(&sl->gx[sz][pass][mask - 1], graph, nullptr, nullptr, 0);
          (void)hipGraphDestroy(graph);
          if (e != hipSuccess) return fail(JC_E_HIP, "graph instantiate: %s", hipGetErrorString(e));
        }
      }
  }
  return JC_OK;
}

void slot_free(Slot* sl) {
  if (sl->stream) (void)hipStreamSynchronize(sl->stream);
  if (sl->stream2) (void)hipStreamSynchronize(sl->stream2);
  for (auto& a : sl->gx)
    for (auto& v : a)
      for (hipGraphExec_t x : v)
        if (x) (void)hipGraphExecDestroy(x);
  if (sl->arena) (void)hipFree(sl->arena);
  if (sl->d_stage) (void)hipFree(sl->d_stage);
  if (sl->h_stage) (void)hipHostFree(sl->h_stage);
  for (int b = 0; b < 2; ++b) {
    if (sl->h_outb[b]) (void)hipHostFree(sl->h_outb[b]);
    if (sl->h_thetab[b]) (void)hipHostFree(sl->h_thetab[b]);
  }
  if (sl->ev_start) (void)hipEventDestroy(sl->ev_start);
  if (sl->ev_done) (void)hipEventDestroy(sl->ev_done);
  for (int p = 0; p < 2; ++p)
    if (sl->ev_fork[p]) (void)hipEventDestroy(sl->ev_fork[p]);
  if (sl->ev_fwd) (void)hipEventDestroy(sl->ev_fwd);
  if (sl->ev_grad) (void)hipEventDestroy(sl->ev_grad);
  if (sl->ev_unpacked) (void)hipEventDestroy(sl->ev_unpacked);
  for (int i = 0; i < 2 * MAXG; ++i) {
    if (sl->side[i]) {
      (void)hipStreamSynchronize(sl->side[i]);
      (void)hipStreamDestroy(sl->side[i]);
    }
    if (sl->ev_join[i]) (void)hipEventDestroy(sl->ev_join[i]);
  }
  if (sl->stream) (void)hipStreamDestroy(sl->stream);
  if (sl->stream2) (void)hipStreamDestroy(sl->stream2);
}

// Pack the admitted requests of a slot into its staging area and launch.
hipError_t launch_batch(jc_server* S, Slot* sl) {
  unsigned char* hs = sl->h_stage;
  StageHdr* hdr = reinterpret_cast<StageHdr*>(hs);
  memset(hdr, 0, sizeof(StageHdr));
  size_t so = align_up(sizeof(StageHdr), 256);
  auto section = [&](void* dst_dev, const void* src, size_t bytes) {
    if (bytes == 0) return;
    const size_t b16 = align_up(bytes, 16);
    memcpy(hs + so, src, bytes);
    if (b16 > bytes) memset(hs + so + bytes, 0, b16 - bytes);
    StageSec& s = hdr->sec[hdr->nsec++];
    s.dst = (int64_t)((unsigned char*)dst_dev - sl->arena);
    s.src = (int64_t)so;
    s.bytes = (int64_t)b16;
    so = align_up(so + b16, 256);
  };
  int counts[MAXG * 8] = {0};
  bool small = true;  // every group fits the small grids
  unsigned fmask = 0, gmask = 0;  // sizes with forward / gradient work
  sl->out_used = 0;
  sl->th_used = 0;
  std::vector<float> cbar_all;
  for (int g = 0; g < S->ng; ++g) {
    std::vector<double> params, gparams;
    std::vector<RJob> jobs, gjobs;
    std::vector<RCol> cols, gcols;
    std::vector<Req*> grp;
    for (Req* q : sl->batch)
      if (q->g == g) grp.push_back(q);
    if (grp.empty()) continue;
    // forward jobs, grad jobs
    std::vector<int> fjob(grp.size()), gjob(grp.size(), -1);
    {
      size_t nc = 0, ng = 0, ngc = 0;
      for (Req* q : grp) {
        nc += q->r.n_ell;
        if (q->r.want_grad) {
          ng += 1;
          ngc += q->r.n_ell;
        }
      }
      jobs.reserve(grp.size());
      params.reserve(grp.size() * 8);
      cols.reserve(nc + 64 * MAXT);
      gjobs.reserve(ng);
      gparams.reserve(ng * 8);
      gcols.reserve(ngc + 64 * MAXT);
    }
    for (size_t i = 0; i < grp.size(); ++i) {
      Req* q = grp[i];
      RJob jb{};
      jb.c = (int)jobs.size();
      jb.T = q->r.n_tracers;
      jb.n_ell = q->r.n_ell;
      jb.wl_mask = 0;
      for (int t = 0; t < jb.T; ++t) {
        jb.sel[t] = q->r.tracers[t];
        if (S->pool->tr[q->r.tracers[t]].kind == JC_TRACER_WL) jb.wl_mask |= 1u << t;
      }
      q->out_off = sl->out_used;
      jb.out_off = q->out_off;
      sl->out_used += (int64_t)q->P * q->r.n_ell;
      fjob[i] = (int)jobs.size();
      jobs.push_back(jb);
      params.insert(params.end(), q->r.cosmo, q->r.cosmo + 8);
      if (q->r.want_grad) {
        RJob gj = jb;
        gj.c = (int)gjobs.size();
        gj.cbar_off = (int64_t)cbar_all.size();
        cbar_all.insert(cbar_all.end(), q->cbar.begin(), q->cbar.end());
        q->th_row = sl->th_used++;
        gj.out_off = q->th_row;
        gjob[i] = (int)gjobs.size();
        gjobs.push_back(gj);
        gparams.insert(gparams.end(), q->r.cosmo, q->r.cosmo + 8);
      }
    }
    // columns, bucketed by tracer count so T is uniform per 64-column tile
    for (int pass = 0; pass < 2; ++pass) {
      std::vector<RCol>& cl = pass ? gcols : cols;
      std::vector<RJob>& js = pass ? gjobs : jobs;
      for (int T = 1; T <= MAXT; ++T) {
        bool any = false;
        for (size_t i = 0; i < grp.size(); ++i) {
          const int j = pass ? gjob[i] : fjob[i];
          if (j < 0 || grp[i]->r.n_tracers != T) continue;
          any = true;
          js[j].col0 = (int)cl.size();
          for (int l = 0; l < grp[i]->r.n_ell; ++l) cl.push_back(RCol{j, l, grp[i]->ell[l]});
        }
        if (any)
          while (cl.size() % 64) cl.push_back(RCol{-1, 0, 10.0});
      }
    }
    counts[g * 8 + 0] = (int)jobs.size();
    counts[g * 8 + 1] = (int)cols.size();
    counts[g * 8 + 2] = (int)gjobs.size();
    counts[g * 8 + 3] = (int)gcols.size();
    counts[g * 8 + 4] = (int)grp.front()->out_off;  // the group's outputs are contiguous
    counts[g * 8 + 5] = (int)sl->out_used;
    fmask |= 1u << g;
    if (!gjobs.empty()) gmask |= 1u << g;
    const Caps& k = S->small;
    small = small && (int)jobs.size() <= k.C && (int)gjobs.size() <= k.Gd && (int)cols.size() <= k.colcap &&
            (int)gcols.size() <= k.gcolcap && sl->out_used - grp.front()->out_off <= k.outcap;
    GroupBufs& b = sl->g[g];
    section(b.params, params.data(), params.size() * 8);
    section(b.jobs, jobs.data(), jobs.size() * sizeof(RJob));
    section(b.cols, cols.data(), cols.size() * sizeof(RCol));
    if (!gjobs.empty()) {
      section(b.gparams, gparams.data(), gparams.size() * 8);
      section(b.gjobs, gjobs.data(), gjobs.size() * sizeof(RJob));
      section(b.gcols, gcols.data(), gcols.size() * sizeof(RCol));
    }
  }
  section(sl->counts, counts, sizeof counts);
  if (!cbar_all.empty()) section(sl->cbar, cbar_all.data(), cbar_all.size() * 4);
  hipError_t e;
  hipStream_t s1 = sl->stream, s2 = sl->stream2;
  if ((e = hipMemcpyAsync(sl->d_stage, hs, so, hipMemcpyHostToDevice, s1)) != hipSuccess) return e;
  if ((e = hipEventRecord(sl->ev_start, s1)) != hipSuccess) return e;
  if ((e = enqueue_unpack(sl)) != hipSuccess) return e;
  if ((e = hipEventRecord(sl->ev_unpacked, s1)) != hipSuccess) return e;
  // a batch without gradient work leaves the second stream alone: streams
  // share the process's few hardware queues, so even an empty marker there
  // can wait behind another slot's gradient kernels
  if (gmask && (e = hipStreamWaitEvent(s2, sl->ev_unpacked, 0)) != hipSuccess) return e;
  // the forward pass first: forward-only requests complete when it ends, and
  // the gradient pass fills the device beside it (measured on the mixed
  // stream: capacity +6 %, p50 at half and full load -13 % and -19 %)
  const bool graphs = !sl->gx[0][0].empty();
  for (int pass = 0; pass <= 1; ++pass) {
    const unsigned mask = pass ? gmask : fmask;
    if (!mask) continue;
    hipStream_t ps = pass ? s2 : s1;
    if (graphs) {
      if ((e = hipGraphLaunch(sl->gx[small ? 1 : 0][pass][mask - 1], ps)) != hipSuccess) return e;
    } else if ((e = enqueue_pass(S, sl, pass, ps, small ? S->small : S->full, mask)) != hipSuccess) {
      return e;
    }
  }
  {
    std::lock_guard<std::mutex> lk(S->mu);
    if (graphs) S->st_graph += 1;
    else S->st_direct += 1;
  }
  if (sl->out_used)
    if ((e = hipMemcpyAsync(sl->h_out, sl->out, sl->out_used * 4, hipMemcpyDeviceToHost, s1)) != hipSuccess)
      return e;
  if ((e = hipEventRecord(sl->ev_fwd, s1)) != hipSuccess) return e;
  if (gmask) {
    if (sl->th_used)
      if ((e = hipMemcpyAsync(sl->h_theta, sl->theta, sl->th_used * 7 * 8, hipMemcpyDeviceToHost, s2)) !=
          hipSuccess)
        return e;
    if ((e = hipEventRecord(sl->ev_grad, s2)) != hipSuccess) return e;
    if ((e = hipStreamWaitEvent(s1, sl->ev_grad, 0)) != hipSuccess) return e;
  }
  sl->fwd_reported = false;
  return hipEventRecord(sl->ev_done, s1);
}

// Post the completions of requests whose outputs are in place (under one
// lock) and free them.
void complete_many(jc_server* S, const std::vector<Req*>& reqs, int status, int bsize) {
  if (reqs.empty()) return;
  const double t = now_us();
  {
    std::lock_guard<std::mutex> lk(S->mu);
    for (Req* q : reqs) {
      jc_completion c{};
      c.id = q->id;
      c.status = status;
      c.batch_size = bsize;
      c.t_submit_us = q->t_submit;
      c.t_done_us = t;
      c.latency_us = t - q->t_submit;
      S->done.push_back(c);
    }
    S->outstanding -= (int64_t)reqs.size();
    if (status == JC_OK) {
      S->st_done += (double)reqs.size();
      for (Req* q : reqs) {
        S->st_wait_us += q->t_launch - q->t_submit;
        S->st_service_us += t - q->t_launch;
      }
    } else {
      S->st_rejected += (double)reqs.size();
    }
  }
  S->cv_idle.notify_all();
  for (Req* q : reqs) delete q;
}

// Copy the results of reqs from slot buffer buf to the callers' arrays, then
// complete them.  Gradient requests get their C_ell and their gradient.
void copy_out(jc_server* S, Slot* sl, int buf, const std::vector<Req*>& reqs, int status, int bsize) {
  if (status == JC_OK)
    for (Req* r : reqs) {
      memcpy(r->r.cl_out, sl->h_outb[buf] + r->out_off, (size_t)r->P * r->r.n_ell * 4);
      if (r->r.want_grad && r->r.grad_out) memcpy(r->r.grad_out, sl->h_thetab[buf] + r->th_row * 7, 7 * 8);
    }
  complete_many(S, reqs, status, bsize);
}

// Results up to this size (the small batches of low load) are copied by the
// scheduler thread itself, which saves the hand-over's thread wake-up on
// their latency; larger ones are split over the copier threads, so the
// scheduler goes on admitting and launching batches meanwhile.
constexpr size_t INLINE_COPY_BYTES = 256 << 10;

void deliver(jc_server* S, Slot* sl, std::vector<Req*>& reqs, int status, int bsize) {
  if (reqs.empty()) return;
  size_t bytes = 0;
  for (Req* r : reqs) bytes += (size_t)r->P * r->r.n_ell * 4;
  const int buf = sl->cur;
  const size_t nthr = S->copiers.size();
  if (status != JC_OK || bytes <= INLINE_COPY_BYTES || nthr == 0) {
    copy_out(S, sl, buf, reqs, status, bsize);
    reqs.clear();
    return;
  }
  const size_t nchunk = std::min(nthr, (bytes + INLINE_COPY_BYTES - 1) / INLINE_COPY_BYTES);
  const size_t per = (bytes + nchunk - 1) / nchunk;
  std::vector<CopyTask> tasks;
  CopyTask t;
  size_t acc = 0;
  for (Req* r : reqs) {
    t.reqs.push_back(r);
    acc += (size_t)r->P * r->r.n_ell * 4;
    if (acc >= per) {
      tasks.push_back(std::move(t));
      t = CopyTask{};
      acc = 0;
    }
  }
  if (!t.reqs.empty()) tasks.push_back(std::move(t));
  {
    std::lock_guard<std::mutex> lk(S->mu);
    sl->drain[buf] += (int)tasks.size();
  }
  {
    std::lock_guard<std::mutex> lk(S->cmu);
    for (CopyTask& x : tasks) {
      x.sl = sl;
      x.buf = buf;
      x.status = status;
      x.bsize = bsize;
      S->ctasks.push_back(std::move(x));
    }
  }
  S->ccv.notify_all();
  reqs.clear();
}

void copier_main(jc_server* S) {
  while (true) {
    CopyTask t;
    {
      std::unique_lock<std::mutex> lk(S->cmu);
      S->ccv.wait(lk, [&] { return S->cstop || !S->ctasks.empty(); });
      if (S->ctasks.empty()) return;  // stopping, nothing left
      t = std::move(S->ctasks.front());
      S->ctasks.pop_front();
    }
    copy_out(S, t.sl, t.buf, t.reqs, t.status, t.bsize);
    {
      std::lock_guard<std::mutex> lk(S->mu);
      t.sl->drain[t.buf] -= 1;
    }
    S->cv.notify_all();  // the buffer may be free for the next batch
  }
}

void worker_main(jc_server* S) {
  (void)hipSetDevice(S->ctx->device);
  std::vector<Slot*> order;  // busy slots in launch order
  while (true) {
    bool progressed = false;
    // requests without a gradient complete when their slot's forward pass
    // and output copy are done
    for (Slot* sl : order) {
      if (sl->fwd_reported || hipEventQuery(sl->ev_fwd) != hipSuccess) continue;
      const int bs = (int)sl->batch.size();
      std::vector<Req*> rest, fwd;
      for (Req* r : sl->batch) (r->r.want_grad ? rest : fwd).push_back(r);
      deliver(S, sl, fwd, JC_OK, bs);
      sl->batch.swap(rest);
      sl->bsize = bs;
      sl->fwd_reported = true;
      progressed = true;
    }
    // every finished slot, whatever its launch order: a short forward-only
    // batch frees its slot while an older batch's gradient pass still runs
    for (auto it = order.begin(); it != order.end();) {
      Slot* sl = *it;
      hipError_t q = hipEventQuery(sl->ev_done);
      if (q == hipErrorNotReady) {
        ++it;
        continue;
      }
      const int status = (q == hipSuccess) ? JC_OK : JC_E_HIP;
      float ms = 0.0f;
      if (q == hipSuccess && hipEventElapsedTime(&ms, sl->ev_start, sl->ev_done) == hipSuccess) {
        std::lock_guard<std::mutex> lk(S->mu);
        S->st_busy_ms += ms;
      }
      const int bs = sl->fwd_reported ? sl->bsize : (int)sl->batch.size();
      deliver(S, sl, sl->batch, status, bs);
      sl->batch.clear();
      sl->busy = false;
      it = order.erase(it);
      progressed = true;
    }
    // admission into a free slot.  With two or more slots and no backlog (at
    // most one full batch pending), slot 0 takes only requests without a
    // gradient (in arrival order among them), so that they never wait for a
    // batch's gradient pass to free a slot; otherwise, and in the other slots,
    // requests of both kinds are taken in arrival order (under a backlog the
    // slots share the gradient work, which keeps the capacity).
    Slot* free_slot = nullptr;
    std::vector<Req*> batch, rejected;
    for (size_t si = 0; si < S->slots.size() && !free_slot; ++si) {
      Slot* cand = S->slots[si].get();
      if (cand->busy) continue;
      {
        std::lock_guard<std::mutex> lk(S->mu);
        // a result buffer no copier still reads (the other one first)
        const int nb = cand->drain[cand->cur ^ 1] == 0 ? (cand->cur ^ 1) : cand->drain[cand->cur] == 0 ? cand->cur : -1;
        if (nb < 0) continue;
        cand->cur = nb;
        cand->h_out = cand->h_outb[nb];
        cand->h_theta = cand->h_thetab[nb];
        const bool fwd_only =
            si == 0 && S->slots.size() >= 2 && S->pending.size() <= (size_t)S->cfg.max_batch;
        size_t fp = 0;
        int64_t outu = 0, goutu = 0;
        int nC[MAXG] = {0}, nG[MAXG] = {0}, nc[MAXG] = {0}, ngc[MAXG] = {0};
        const size_t budget = (size_t)S->cfg.mem_budget;
        std::deque<Req*> skipped;  // gradient requests passed over by the forward slot, in order
        while (!S->pending.empty() && (int)batch.size() < S->cfg.max_batch) {
          Req* q = S->pending.front();
          if (q->footprint > budget) {
            S->pending.pop_front();
            rejected.push_back(q);
            continue;
          }
          if (fwd_only && q->r.want_grad) {
            S->pending.pop_front();
            skipped.push_back(q);
            continue;
          }
          const int g = q->g;
          const int64_t o = (int64_t)q->P * q->r.n_ell;
          const bool fits = fp + q->footprint <= budget && nC[g] + 1 <= S->C && outu + o <= S->outcap &&
                            nc[g] + q->r.n_ell + 64 * MAXT <= S->colcap &&
                            (!q->r.want_grad || (nG[g] + 1 <= S->Gd && goutu + o <= S->goutcap &&
                                                 ngc[g] + q->r.n_ell + 64 * MAXT <= S->gcolcap));
          if (!fits) break;
          S->pending.pop_front();
          batch.push_back(q);
          fp += q->footprint;
          nC[g] += 1;
          nc[g] += q->r.n_ell;
          outu += o;
          if (q->r.want_grad) {
            nG[g] += 1;
            ngc[g] += q->r.n_ell;
            goutu += o;
          }
        }
        // the passed-over requests go back to the front, in their order
        while (!skipped.empty()) {
          S->pending.push_front(skipped.back());
          skipped.pop_back();
        }
      }
      if (!batch.empty() || !rejected.empty()) free_slot = cand;
    }
    if (free_slot || !rejected.empty()) {
      complete_many(S, rejected, JC_E_NOMEM, 0);
      if (!batch.empty()) {
        free_slot->batch = batch;
        // test hook: a batch that fails at launch as an allocation or
        // runtime failure would, before anything is queued
        hipError_t e = hipSuccess;
        {
          std::lock_guard<std::mutex> lk(S->mu);
          if (S->inject_batches > 0) {
            S->inject_batches -= 1;
            e = S->inject_status == JC_E_NOMEM ? hipErrorOutOfMemory : hipErrorLaunchFailure;
          }
        }
        {
          const double tl = now_us();
          for (Req* r : batch) r->t_launch = tl;
        }
        if (e == hipSuccess) {
          const double t0 = now_us();
          e = launch_batch(S, free_slot);
          std::lock_guard<std::mutex> lk(S->mu);
          S->st_launch_us += now_us() - t0;
        }
        if (e != hipSuccess) {
          S->worker_err = hipGetErrorString(e);
          (void)hipStreamSynchronize(free_slot->stream);
          (void)hipStreamSynchronize(free_slot->stream2);
          const int st = e == hipErrorOutOfMemory ? JC_E_NOMEM : JC_E_HIP;
          complete_many(S, batch, st, (int)batch.size());
          free_slot->batch.clear();
        } else {
          free_slot->busy = true;
          order.push_back(free_slot);
          std::lock_guard<std::mutex> lk(S->mu);
          S->st_batches += 1;
          S->st_sumb += batch.size();
          if (batch.size() > S->st_maxb) S->st_maxb = batch.size();
        }
        progressed = true;
      }
    }
    if (S->stop.load() && order.empty()) {
      std::lock_guard<std::mutex> lk(S->mu);
      if (S->pending.empty()) break;
    }
    if (!progressed) {
      std::unique_lock<std::mutex> lk(S->mu);
      if (order.empty() && S->pending.empty() && !S->stop.load())
        S->cv.wait_for(lk, std::chrono::milliseconds(50));
      else  // batches in flight, or pending requests waiting for a result buffer
        S->cv.wait_for(lk, std::chrono::microseconds(20));
    }
  }
}

}  // namespace

extern "C" {

int jc_server_create(jc_ctx* ctx, int64_t pool_set, const jc_server_config* cfg, jc_server** out) {
  if (!ctx || !out) return fail(JC_E_ARG, "null argument");
  auto it = ctx->sets.find(pool_set);
  if (it == ctx->sets.end()) return fail(JC_E_ARG, "unknown tracer set %lld", (long long)pool_set);
  std::unique_ptr<jc_server> S(new jc_server);
  S->ctx = ctx;
  S->pool = it->second.get();
  if (S->pool->T > TPAD_POOL) return fail(JC_E_UNSUPPORTED, "pool sets hold at most 10 tracers");
  jc_server_config c = cfg ? *cfg : jc_server_config{};
  if (c.max_batch <= 0) c.max_batch = 256;
  if (c.max_grad <= 0) c.max_grad = 64;
  if (c.max_ell <= 0) c.max_ell = 256;
  if (c.slots <= 0) c.slots = 2;
  if (c.n_nz <= 0) {
    c.n_nz = 1;
    c.n_z[0] = 512;
  }
  if (c.n_nz > MAXG) return fail(JC_E_ARG, "at most %d Simpson sizes", MAXG);
  if (c.slots > 8 || c.max_batch > 4096 || c.max_grad > 1024 || c.max_ell > 4096)
    return fail(JC_E_ARG, "server capacity out of range");
  for (const auto& t : S->pool->tr) {
    S->zall = std::max(S->zall, t.zmax);
    if (t.kind != JC_TRACER_WL) continue;
    if (t.probe < 0) return fail(JC_E_ARG, "negative probe index");
    if ((size_t)t.probe >= S->zprobe.size()) S->zprobe.resize((size_t)t.probe + 1, 0.0);
    S->zprobe[t.probe] = std::max(S->zprobe[t.probe], t.zmax);
  }
  S->tables_variant = ctx->tables_variant;
  S->transfer_model = ctx->transfer_model;
  S->halofit_model = ctx->halofit_model;
  S->growth_model = ctx->growth_model;
  S->gamma = ctx->gamma;
  S->waves = c.waves ? c.waves : 2;
  S->ctile = c.wl_ctile ? c.wl_ctile : 4;
  if (S->waves != 1 && S->waves != 2 && S->waves != 4) return fail(JC_E_ARG, "waves must be 1, 2 or 4");
  if (S->ctile != 1 && S->ctile != 4 && S->ctile != 8 && S->ctile != 16) return fail(JC_E_ARG, "wl_ctile must be 1, 4, 8 or 16");
  HIPCHK(hipSetDevice(ctx->device));
  S->ng = c.n_nz;
  for (int g = 0; g < S->ng; ++g) {
    if (c.n_z[g] < 2 || (c.n_z[g] % 2) || c.n_z[g] > 8192) return fail(JC_E_ARG, "bad n_z %d", c.n_z[g]);
    S->Nz[g] = c.n_z[g];
    S->Na[g] = c.n_z[g] + 1;
    int rc = build_grid(ctx, S->pool, c.n_z[g], &S->grid[g]);
    if (rc) return rc;
  }
  S->C = c.max_batch;
  S->Gd = std::min(c.max_grad, c.max_batch);
  S->colcap = (int)align_up((size_t)S->C * c.max_ell + 64 * MAXT, 64);
  S->gcolcap = (int)align_up((size_t)S->Gd * c.max_ell + 64 * MAXT, 64);
  S->outcap = (int64_t)S->C * 55 * c.max_ell;
  S->goutcap = (int64_t)S->Gd * 55 * c.max_ell;
  S->full = Caps{S->C, S->Gd, S->colcap, S->gcolcap, S->outcap};
  {
    const int cs = std::min(S->C, SMALL_BATCH), gs = std::min(S->Gd, SMALL_BATCH / 2);
    S->small = Caps{cs, gs, std::min(S->colcap, (int)align_up((size_t)cs * c.max_ell + 64 * MAXT, 64)),
                    std::min(S->gcolcap, (int)align_up((size_t)gs * c.max_ell + 64 * MAXT, 64)),
                    std::min(S->outcap, (int64_t)cs * 55 * c.max_ell)};
  }
  S->cfg = c;
  // the per-request device state a batch may hold (capacity of one slot)
  Slot probe;
  size_t cap = layout_slot(S.get(), &probe, nullptr);
  if (c.mem_budget <= 0) S->cfg.mem_budget = (int64_t)cap;
  for (int i = 0; i < c.slots; ++i) {
    std::unique_ptr<Slot> sl(new Slot);
    int rc = slot_init(S.get(), sl.get());
    if (rc) {
      slot_free(sl.get());
      for (auto& x : S->slots) slot_free(x.get());
      return rc;
    }
    S->slots.push_back(std::move(sl));
  }
  {
    const unsigned hw = std::thread::hardware_concurrency();
    const unsigned nc = std::max(1u, std::min(4u, hw / 4));
    for (unsigned i = 0; i < nc; ++i) S->copiers.emplace_back(copier_main, S.get());
  }
  S->worker = std::thread(worker_main, S.get());
  *out = S.release();
  return JC_OK;
}

}  // extern "C"

namespace {
// Validate one request (nothing is allocated); *g = its served-size group.
int check_req(jc_server* S, const jc_request* r, int* gout) {
  if (r->n_tracers < 1 || r->n_tracers > MAXT) return fail(JC_E_ARG, "n_tracers must be 1..10");
  if (r->n_ell < 1 || r->n_ell > S->cfg.max_ell) return fail(JC_E_ARG, "n_ell must be 1..%d", S->cfg.max_ell);
  if (!r->ell || !r->cl_out) return fail(JC_E_ARG, "ell and cl_out are required");
  if (r->want_grad && (!r->cl_bar || !r->grad_out)) return fail(JC_E_ARG, "gradient requests need cl_bar and grad_out");
  int g = -1;
  for (int i = 0; i < S->ng; ++i)
    if (S->Nz[i] == r->n_z) g = i;
  if (g < 0) return fail(JC_E_ARG, "n_z %d is not served", r->n_z);
  for (int t = 0; t < r->n_tracers; ++t)
    if (r->tracers[t] < 0 || r->tracers[t] >= S->pool->T) return fail(JC_E_ARG, "tracer index %d", r->tracers[t]);
  // Requests are integrated on the pool's a-grid, which ends at the pool's
  // largest zmax (angular_cl.py:63), and on each weak-lensing probe's inner
  // grids, which end at that probe's largest zmax (probes.py:24).  A request
  // equals angular_cl of its own bins only where those are its own largest
  // zmax too.
  {
    const auto& tr = S->pool->tr;
    double zreq = 0.0;
    // the request's largest zmax per lensing probe (pool probes are few)
    double zq[TPAD_POOL];
    int pr[TPAD_POOL], np = 0;
    for (int t = 0; t < r->n_tracers; ++t) {
      const jc_tracer_desc& d = tr[r->tracers[t]];
      zreq = std::max(zreq, d.zmax);
      if (d.kind != JC_TRACER_WL) continue;
      int k = 0;
      while (k < np && pr[k] != d.probe) ++k;
      if (k == np) {
        pr[np] = d.probe;
        zq[np++] = d.zmax;
      } else {
        zq[k] = std::max(zq[k], d.zmax);
      }
    }
    if (zreq != S->zall)
      return fail(JC_E_UNSUPPORTED, "the request's bins end at zmax %g, the pool's a-grid at %g", zreq, S->zall);
    for (int k = 0; k < np; ++k)
      if (zq[k] != S->zprobe[pr[k]])
        return fail(JC_E_UNSUPPORTED, "the request's bins of lensing probe %d end at zmax %g, the probe's at %g",
                    pr[k], zq[k], S->zprobe[pr[k]]);
  }
  for (int i = 0; i < 8; ++i)
    if (!std::isfinite(r->cosmo[i])) return fail(JC_E_ARG, "cosmology parameter %d is not finite", i);
  if (!(r->cosmo[0] + r->cosmo[1] > 0.0) || !(r->cosmo[1] > 0.0) || !(r->cosmo[2] > 0.0) || !(r->cosmo[4] > 0.0))
    return fail(JC_E_ARG, "Omega_b, Omega_m, h and sigma8 must be positive");
  for (int i = 0; i < r->n_ell; ++i)
    if (!std::isfinite(r->ell[i]) || r->ell[i] < 1.0) return fail(JC_E_ARG, "ell values must be finite and >= 1");
  *gout = g;
  return JC_OK;
}

// Copy a validated request.
Req* make_req(jc_server* S, const jc_request* r, int g) {
  Req* q = new Req;
  q->r = *r;
  q->ell.assign(r->ell, r->ell + r->n_ell);
  q->P = r->n_tracers * (r->n_tracers + 1) / 2;
  if (r->want_grad) q->cbar.assign(r->cl_bar, r->cl_bar + (size_t)q->P * r->n_ell);
  q->g = g;
  q->footprint = req_footprint(S, *r, g);
  return q;
}
}  // namespace

extern "C" {

int jc_server_submit(jc_server* S, const jc_request* r, int64_t* id) {
  if (!S || !r || !id) return fail(JC_E_ARG, "null argument");
  return jc_server_submit_many(S, r, 1, id);
}

int jc_server_submit_many(jc_server* S, const jc_request* reqs, int64_t n, int64_t* ids) {
  if (!S || (n > 0 && (!reqs || !ids))) return fail(JC_E_ARG, "null argument");
  if (n < 0) return fail(JC_E_ARG, "negative request count");
  // every request is validated before any is queued
  std::vector<int> grp((size_t)n);
  for (int64_t i = 0; i < n; ++i) {
    int rc = check_req(S, reqs + i, &grp[(size_t)i]);
    if (rc) return fail(rc, "request %lld: %s", (long long)i, g_err.c_str());
  }
  // the call's ids are one consecutive range; the requests are copied and
  // queued in chunks, so the scheduler forms the first batch while the
  // rest are copied
  int64_t id0 = 0;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    if (S->stop.load()) return fail(JC_E_ARG, "server is stopping");
    id0 = S->next_id;
    S->next_id += n;
  }
  constexpr int64_t CHUNK = 128;
  std::vector<Req*> qs;
  qs.reserve((size_t)std::min(n, CHUNK));
  for (int64_t i0 = 0; i0 < n; i0 += CHUNK) {
    const int64_t i1 = std::min(n, i0 + CHUNK);
    qs.clear();
    for (int64_t i = i0; i < i1; ++i) qs.push_back(make_req(S, reqs + i, grp[(size_t)i]));
    {
      std::lock_guard<std::mutex> lk(S->mu);
      if (S->stop.load()) {  // destroyed meanwhile (a caller error): drop the rest
        for (Req* q : qs) delete q;
        return fail(JC_E_ARG, "server is stopping");
      }
      const double t = now_us();
      for (int64_t i = i0; i < i1; ++i) {
        Req* q = qs[(size_t)(i - i0)];
        q->id = id0 + i;
        q->t_submit = t;
        S->pending.push_back(q);
        ids[i] = q->id;
      }
      S->outstanding += i1 - i0;
    }
    S->cv.notify_one();
  }
  return JC_OK;
}

int jc_server_poll(jc_server* S, jc_completion* out, int64_t max, int64_t* n) {
  if (!S || !n || (max > 0 && !out)) return fail(JC_E_ARG, "null argument");
  std::lock_guard<std::mutex> lk(S->mu);
  const int64_t k = std::min<int64_t>(max, (int64_t)S->done.size());
  for (int64_t i = 0; i < k; ++i) out[i] = S->done[i];
  S->done.erase(S->done.begin(), S->done.begin() + k);
  *n = k;
  return JC_OK;
}

int jc_server_wait(jc_server* S, int64_t outstanding, int32_t timeout_ms) {
  if (!S) return fail(JC_E_ARG, "null server");
  std::unique_lock<std::mutex> lk(S->mu);
  const bool ok = S->cv_idle.wait_for(lk, std::chrono::milliseconds(timeout_ms),
                                      [&] { return S->outstanding <= outstanding; });
  if (!ok) return fail(JC_E_ARG, "timeout with %lld requests outstanding", (long long)S->outstanding);
  return JC_OK;
}

int jc_server_inject_failure(jc_server* S, int32_t batches, int32_t status) {
  if (!S) return fail(JC_E_ARG, "null server");
  if (batches < 0 || (status != JC_E_NOMEM && status != JC_E_HIP))
    return fail(JC_E_ARG, "batches >= 0 and status JC_E_NOMEM or JC_E_HIP required");
  std::lock_guard<std::mutex> lk(S->mu);
  S->inject_batches = batches;
  S->inject_status = status;
  return JC_OK;
}

int jc_server_stats(jc_server* S, double* st, int64_t n) {
  if (!S || !st) return fail(JC_E_ARG, "null argument");
  std::lock_guard<std::mutex> lk(S->mu);
  const double v[11] = {S->st_batches, S->st_done, S->st_rejected, S->st_maxb,
                        S->st_batches ? S->st_sumb / S->st_batches : 0.0, S->st_busy_ms, S->st_graph,
                        S->st_direct, S->st_wait_us * 1e-3, S->st_service_us * 1e-3, S->st_launch_us * 1e-3};
  for (int64_t i = 0; i < n && i < 11; ++i) st[i] = v[i];
  return JC_OK;
}

int jc_server_destroy(jc_server* S) {
  if (!S) return JC_OK;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->stop.store(true);
  }
  S->cv.notify_all();
  if (S->worker.joinable()) S->worker.join();
  {
    std::lock_guard<std::mutex> lk(S->cmu);
    S->cstop = true;
  }
  S->ccv.notify_all();
  for (std::thread& t : S->copiers)
    if (t.joinable()) t.join();
  for (auto& sl : S->slots) slot_free(sl.get());
  for (Req* q : S->pending) delete q;
  delete S;
  return JC_OK;
}

}  // extern "C"
