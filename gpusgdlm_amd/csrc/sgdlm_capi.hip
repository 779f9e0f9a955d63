/*
 * sgdlm_capi.hip -- the C-ABI of libsgdlm_hip.so: context, state machine,
 * device buffers and stage sequencing.
 *
 * Plays the role of SGDLM::HostWrapperImpl<double>
 * (rSGDLM/src/SGDLM/HostWrapperImpl.cu) for one GPU: the same state flags
 * (memory_initialized, sim_memory_initialized, evo_memory_initialized,
 * is_prior, use_state_evolution_matrix; HostWrapperImpl.cuh:101-127) and the
 * same call-order rules (checkInitialized / checkSimInitialized / checkPrior
 * / checkUseStateEvolutionMatrix, HostWrapperImpl.cu:1239-1286), but errors
 * are returned as status codes instead of being logged and ignored.
 *
 * Device data are plain strided arrays (no pointer arrays); draws are kept
 * in a packed draw-major layout theta[i * P + off_j + k], P = sum_j p_j.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/sgdlm.h"
#include "lu_symbolic.h"
#include "sgdlm_internal.h"

using namespace sgdlm;

/* Device memory of sgdlm_compute_vb_posterior_noalloc: every buffer of a
 * VB call beyond the model state and the caller's outputs is carved from
 * this array instead of being allocated.  Each GPU the library runs on
 * holds its own copy until the process ends.  In host code the name is
 * only a symbol handle: kernels get the per-device address that
 * hipGetSymbolAddress returns (scratch_ptr below). */
constexpr size_t kScratchBytes = 64ull << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

std::mutex g_scratch_mu;
std::map<int, void *> g_scratch_ptr; /* device -> address of scratch */

/* Address of this device's scratch array (with that device current);
 * looked up once per device. */
hipError_t scratch_ptr(int device, void **out) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  auto it = g_scratch_ptr.find(device);
  if (it != g_scratch_ptr.end()) {
    *out = it->second;
    return hipSuccess;
  }
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) {
    (void)hipGetLastError();
    *out = nullptr;
    return e;
  }
  g_scratch_ptr[device] = p;
  *out = p;
  return hipSuccess;
}

struct DBuf {
  void *ptr = nullptr;
  size_t bytes = 0;
  template <typename T> T *as() const { return static_cast<T *>(ptr); }
};

struct PendingEvent {
  std::string stage;
  hipEvent_t a, b;
};

/* A tuned VB LU configuration: kernel kind, fast-kernel shape (0 for the
 * register-resident kind) and prefix front limit, with what finding it
 * cost. */
struct LuTuned {
  int kind = SGDLM_LU_AUTO, threads = 0, nb = 0, max_front = -1;
  double tune_ms = 0.0, best_ms = 0.0;
  /* time of one draw on one workgroup of the chosen kernel (the timed
   * launch's time x its workgroups / draws; 0 = not measured) */
  double draw_wg_ms = 0.0;
  int candidates = 0;
  bool from_cache = false;
  /* the faster of the kinds that pivot every draw (partial-pivoting kernel,
   * register kernel with interchanges) and its time: what a call runs when
   * the last call handed nearly every draw to the fallback */
  int pa_kind = SGDLM_LU_PIVOTING;
  double pa_ms = 0.0;
};

/* A captured VB step and what it was captured for: replayed only while
 * every argument it baked in is unchanged. */
struct VbGraph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<uintptr_t> sig;
};

/* Tuned configurations of the process, by structure key. */
std::mutex g_tune_mu;
std::map<uint64_t, LuTuned> g_tune;

std::atomic<uint64_t> g_ctx_ids{0};

/* Streams a batch's LU launches of one stage run on concurrently (the
 * process has four hardware queues by default). */
constexpr size_t kGroupSideStreams = 4;
/* Parts a batch is split into, each on a stream of its own. */
constexpr int kBatchParts = 4;

/* One LU launch over several models: the group form of a kernel, its
 * launch shape, a model's argument struct for the by-value parameter and
 * the device table / workgroup map it reads. */
struct GrpLaunch {
  const void *fn = nullptr;
  int threads = 0, grid = 0;
  size_t lds = 0;
  std::vector<unsigned char> a0;
  const void *tab = nullptr;
  GrpMap gm;
};

/* A VB step of several models run as one launch per stage (one per kernel
 * variant for the LU): the device tables it reads, built for one list of
 * contexts and reused while nothing they bake in changes (sig). */
struct VbGroup {
  int n = 0, n_series = 0, max_p = 0, max_chunks = 0;
  long max_N = 0;
  void *blob = nullptr; /* model table, series map, epochs, stats, LU tables */
  void *work = nullptr; /* LU workspace of every model */
  VbModel *d_tab = nullptr;
  int *d_ser = nullptr;
  uint32_t *d_epochs = nullptr;
  double *d_stats = nullptr;
  uint32_t *h_epochs = nullptr; /* pinned */
  double *h_stats = nullptr;    /* pinned, 5 per model (4 weight
                                 * statistics, the LU fallback count) */
  std::vector<GrpLaunch> lu1, lu2; /* longest first */
  /* the LU launches of a stage run concurrently on these streams, forked
   * from and joined back to the leader's stream through the events */
  std::vector<hipStream_t> side;
  std::vector<hipEvent_t> ev;
};

void drop_vb_group(VbGroup *g) {
  if (!g)
    return;
  if (g->blob)
    (void)hipFree(g->blob);
  if (g->work)
    (void)hipFree(g->work);
  if (g->h_epochs)
    (void)hipHostFree(g->h_epochs);
  if (g->h_stats)
    (void)hipHostFree(g->h_stats);
  for (hipStream_t st : g->side)
    (void)hipStreamDestroy(st);
  for (hipEvent_t e : g->ev)
    (void)hipEventDestroy(e);
  delete g;
}

/* A batch: the models split into parts of about equal work, each a VbGroup
 * whose stages run on a stream of its own, so that one part's draws and
 * moments overlap another part's LU; captured and replayed as one graph in
 * graph mode. */
struct VbBatch {
  std::vector<uintptr_t> sig;
  std::vector<VbGroup *> parts;
  std::vector<std::vector<int>> members; /* model indices of each part */
  std::vector<hipStream_t> streams;      /* one per part (more than one part) */
  std::vector<hipEvent_t> ev;            /* fork, then one join per part */
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

void drop_vb_batch(VbBatch *b) {
  if (!b)
    return;
  if (b->exec)
    (void)hipGraphExecDestroy(b->exec);
  if (b->graph)
    (void)hipGraphDestroy(b->graph);
  for (VbGroup *g : b->parts)
    drop_vb_group(g);
  for (hipStream_t st : b->streams)
    (void)hipStreamDestroy(st);
  for (hipEvent_t e : b->ev)
    (void)hipEventDestroy(e);
  delete b;
}

} // namespace

struct sgdlm_ctx {
  std::string err;
  int device = 0;
  int n_cu = 0;
  size_t lds_max = 65536;
  hipStream_t stream = nullptr;
  /* pinned host copy of the last VB call's weight statistics (VbBuffers
   * stats: max log-det, sum exp(l - max), sum w^2, zero weights) */
  double *h_stats = nullptr;
  /* this device's scratch array (sgdlm_compute_vb_posterior_noalloc) */
  void *scratch = nullptr;

  size_t m = 0, max_p = 0;
  bool memory_initialized = false;
  bool sim_memory_initialized = false;
  bool evo_memory_initialized = false;
  bool use_state_evolution_matrix = false;
  bool is_prior = false;
  bool parents_set = false;
  size_t nsim = 0, nsim_batch = 0;

  int prec = SGDLM_PREC_FP64;
  int max_front = 32; /* sparse LU prefix: largest pivot degree (0 = off) */
  bool max_front_fixed = false; /* set by sgdlm_set_lu_max_front          */
  int max_front_used = 0;       /* front limit of the current plan        */
  int lu_threads = 0, lu_nb = 0; /* fast LU kernel tuning (0 = automatic) */
  int lu_kind = SGDLM_LU_AUTO;  /* pinned LU kernel kind                     */
  int lu_solve_kind = SGDLM_LU_AUTO; /* pinned forecast-solve kernel kind      */
  int max_front_default = 0;    /* front limit the default rule chose        */
  bool autotune = true;
  uint64_t tuned_key = 0;       /* structure key of `tuned` (0: none)        */
  LuTuned tuned;
  /* graph mode: the VB stages captured once and replayed */
  bool graph_mode = false;
  uint64_t gen = 0;             /* bumped whenever device tables are rebuilt */
  VbGraph vbg;
  long graph_captures = 0, graph_replays = 0;
  uint64_t seed = 1234ull;
  uint32_t epoch = 0;

  std::vector<unsigned> h_p, h_sp;
  std::vector<GammaEntry> ents; /* entries of Gamma of the current sets     */
  Structure S;

  /* model memory */
  std::map<std::string, DBuf> mem;
  /* simulation memory */
  std::map<std::string, DBuf> sim;
  bool vb_valid = false;
  bool lu_flags_valid = false; /* lu_flags holds the last LU call's count */
  /* the adaptive LU of run_vb (kind not pinned): draws of the last call that
   * ran the no-interchange kernel and how many of them it flagged; calls
   * since it; whether the current call pivots every draw */
  long fb_n = 0, run_n = 0;
  int fb_count = 0, pa_streak = 0;
  bool ran_pa = false;

  bool timing = false;
  std::map<std::string, std::pair<double, long>> times;
  std::vector<PendingEvent> pending;
  /* process-unique identity (a batch's cached device tables are keyed by
   * it, not by the address, which a later context may reuse) */
  uint64_t id = 0;
  /* the device tables of the last batch this context led (first in the
   * list of sgdlm_batch_compute_vb_posterior) */
  VbBatch *grp = nullptr;
  long batch_group_calls = 0, batch_stream_calls = 0;
  bool batch_no_group = false; /* sgdlm_set_batch_mode: one stream per context */
};

namespace {

int fail(sgdlm_ctx *c, int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  c->err = buf;
  return code;
}

#define HIPCHK(ctx, call)                                                      \
  do {                                                                         \
    hipError_t e_ = (call);                                                    \
    if (e_ != hipSuccess)                                                      \
      return fail((ctx), SGDLM_ERR_DEVICE, "%s: %s", #call,                    \
                  hipGetErrorString(e_));                                      \
  } while (0)

int dalloc(sgdlm_ctx *c, std::map<std::string, DBuf> &pool,
           const std::string &name, size_t bytes) {
  DBuf &b = pool[name];
  if (b.ptr && b.bytes >= bytes)
    return SGDLM_OK;
  if (b.ptr) {
    (void)hipFree(b.ptr);
    b.ptr = nullptr;
    b.bytes = 0;
  }
  if (bytes == 0)
    bytes = 16;
  hipError_t e = hipMalloc(&b.ptr, bytes);
  if (e != hipSuccess) {
    b.ptr = nullptr;
    (void)hipGetLastError();
    return fail(c, SGDLM_ERR_ALLOC, "device allocation of %zu bytes for %s failed: %s",
                bytes, name.c_str(), hipGetErrorString(e));
  }
  b.bytes = bytes;
  return SGDLM_OK;
}

void free_pool(std::map<std::string, DBuf> &pool) {
  for (auto &kv : pool)
    if (kv.second.ptr)
      (void)hipFree(kv.second.ptr);
  pool.clear();
}

template <typename T> T *M_(sgdlm_ctx *c, const char *n) {
  return c->mem[n].as<T>();
}
template <typename T> T *SIM(sgdlm_ctx *c, const char *n) {
  return c->sim[n].as<T>();
}

struct StageTimer {
  sgdlm_ctx *c;
  std::string stage;
  hipEvent_t a{}, b{};
  bool on;
  StageTimer(sgdlm_ctx *ctx, const char *s, bool enabled = true)
      : c(ctx), stage(s), on(enabled && ctx->timing) {
    if (on) {
      (void)hipEventCreate(&a);
      (void)hipEventCreate(&b);
      (void)hipEventRecord(a, c->stream);
    }
  }
  ~StageTimer() {
    if (on) {
      (void)hipEventRecord(b, c->stream);
      c->pending.push_back({stage, a, b});
    }
  }
};

int sync_and_collect(sgdlm_ctx *c) {
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (auto &pe : c->pending) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, pe.a, pe.b) == hipSuccess) {
      auto &t = c->times[pe.stage];
      t.first += ms;
      t.second += 1;
    }
    (void)hipEventDestroy(pe.a);
    (void)hipEventDestroy(pe.b);
  }
  c->pending.clear();
  return SGDLM_OK;
}

#define CHECK_LAUNCH(ctx)                                                      \
  do {                                                                         \
    hipError_t e_ = hipGetLastError();                                         \
    if (e_ != hipSuccess)                                                      \
      return fail((ctx), SGDLM_ERR_DEVICE, "kernel launch: %s",                \
                  hipGetErrorString(e_));                                      \
  } while (0)

int need_init(const sgdlm_ctx *c) {
  return c->memory_initialized ? SGDLM_OK : SGDLM_ERR_STATE;
}

int check_init(sgdlm_ctx *c) {
  if (!c->memory_initialized)
    return fail(c, SGDLM_ERR_STATE, "memory not initialised (call init_memory / set parents first)");
  return SGDLM_OK;
}
int check_parents(sgdlm_ctx *c) {
  if (!c->parents_set)
    return fail(c, SGDLM_ERR_STATE, "parental sets not set");
  return SGDLM_OK;
}
int check_sim(sgdlm_ctx *c) {
  if (!c->sim_memory_initialized)
    return fail(c, SGDLM_ERR_STATE, "simulation memory not initialised");
  return SGDLM_OK;
}
int check_prior(sgdlm_ctx *c, bool want_prior) {
  if (want_prior && !c->is_prior)
    return fail(c, SGDLM_ERR_STATE, "parameters are posterior parameters; a prior is required");
  if (!want_prior && c->is_prior)
    return fail(c, SGDLM_ERR_STATE, "parameters are prior parameters; a posterior is required");
  return SGDLM_OK;
}
int check_evo(sgdlm_ctx *c) {
  if (!c->use_state_evolution_matrix)
    return fail(c, SGDLM_ERR_STATE, "state evolution matrix not enabled");
  return SGDLM_OK;
}

/* The Philox counter holds the epoch in 24 bits (epoch << 8 | purpose):
 * a call that would take it to 2^24 fails instead of silently repeating
 * the random streams of earlier calls (sgdlm_set_rng_state starts a new
 * sequence). */
int check_epochs(sgdlm_ctx *c, uint32_t n) {
  if ((uint64_t)c->epoch + n >= (1ull << 24))
    return fail(c, SGDLM_ERR_STATE,
                "Philox epoch counter exhausted (epoch %u + %u >= 2^24); reset it with "
                "sgdlm_set_rng_state", c->epoch, n);
  return SGDLM_OK;
}

/* Allocate the simulation buffers for the current N and P. */
int ensure_sim_buffers(sgdlm_ctx *c) {
  const size_t N = c->nsim, m = c->m, mp = c->max_p;
  const size_t P = (size_t)c->S.P;
  int rc;
  if ((rc = dalloc(c, c->sim, "lam", N * m * 8)) ||
      (rc = dalloc(c, c->sim, "theta", N * std::max<size_t>(P, 1) * 8)) ||
      (rc = dalloc(c, c->sim, "logdet", N * 8)) ||
      (rc = dalloc(c, c->sim, "w", N * 8)) ||
      (rc = dalloc(c, c->sim, "stats", 8 * 8)) ||
      (rc = dalloc(c, c->sim, "lu_flags", (N + 2) * 4)))
    return rc;
  const int chunks = vb_chunks((long)N);
  if ((rc = dalloc(c, c->sim, "part1", (size_t)chunks * m * 8)) ||
      (rc = dalloc(c, c->sim, "part2", (size_t)chunks * m * (mp + 1) * (mp + 1) * 8)) ||
      (rc = dalloc(c, c->sim, "mean_lambda", m * 8)) ||
      (rc = dalloc(c, c->sim, "mean_log_lambda", m * 8)) ||
      (rc = dalloc(c, c->sim, "mean_m", m * mp * 8)) ||
      (rc = dalloc(c, c->sim, "mean_C", m * mp * mp * 8)) ||
      (rc = dalloc(c, c->sim, "mean_n", m * 8)) ||
      (rc = dalloc(c, c->sim, "mean_s", m * 8)) ||
      (rc = dalloc(c, c->sim, "mean_Q", m * 8)) ||
      (rc = dalloc(c, c->sim, "newton_iter", m * 4)))
    return rc;
  return SGDLM_OK;
}

VbBuffers vb_buffers(sgdlm_ctx *c) {
  VbBuffers B;
  B.w = SIM<double>(c, "w");
  B.stats = SIM<double>(c, "stats");
  B.part1 = SIM<double>(c, "part1");
  B.part2 = SIM<double>(c, "part2");
  B.mean_lambda = SIM<double>(c, "mean_lambda");
  B.mean_log_lambda = SIM<double>(c, "mean_log_lambda");
  B.mean_m = SIM<double>(c, "mean_m");
  B.mean_C = SIM<double>(c, "mean_C");
  B.mean_n = SIM<double>(c, "mean_n");
  B.mean_s = SIM<double>(c, "mean_s");
  B.mean_Q = SIM<double>(c, "mean_Q");
  B.newton_iter = SIM<int>(c, "newton_iter");
  B.chunks = vb_chunks((long)c->nsim);
  return B;
}

int h2d(sgdlm_ctx *c, void *dst, const void *src, size_t bytes) {
  HIPCHK(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
  return SGDLM_OK;
}
int d2h(sgdlm_ctx *c, void *dst, const void *src, size_t bytes) {
  HIPCHK(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
  return SGDLM_OK;
}

template <typename V> int upload(sgdlm_ctx *c, const char *name, const V &v) {
  const size_t bytes = std::max<size_t>(v.size(), 1) * sizeof(typename V::value_type);
  int rc;
  if ((rc = dalloc(c, c->mem, name, bytes)))
    return rc;
  if (!v.empty() && (rc = h2d(c, c->mem[name].ptr, v.data(), v.size() * sizeof(v[0]))))
    return rc;
  return SGDLM_OK;
}

/* Banded Schur tables of the one-wavefront LU after the prefix
 * (k_lu_wave.hip k_lu_wsch): bands of 16 rows, row stride M + 4 for the
 * M x M block (M = the dense block rounded up to 32, 64 or 128).  Leaves
 * e_wb_M = 0 without a prefix or beyond 128 rows. */
int build_wave_band_tables(sgdlm_ctx *c, const ElimPlan &P) {
  Structure &S = c->S;
  S.e_wb_M = 0;
  const int M = lu_wave_dim(P.M2, SGDLM_PREC_FP32);
  if (P.n1 == 0 || P.M2 == 0 || M == 0)
    return SGDLM_OK;
  const int stride = M + 4;
  BandPlan B;
  build_band_plan(P, 16, stride, M, B);
  int rc;
  if ((rc = upload(c, "wb_blk_ptr", B.blk_ptr)) || (rc = upload(c, "wb_st_off", B.st_off)) ||
      (rc = upload(c, "wb_pairs", B.st_pairs)) || (rc = upload(c, "wb_brounds", B.brounds)) ||
      (rc = upload(c, "wb_bcbase", B.bcbase)) || (rc = upload(c, "wb_sg_ptr", B.sg_ptr)) ||
      (rc = upload(c, "wb_sg_off", B.sg_off)) || (rc = upload(c, "wb_sg_idx", B.sg_idx)))
    return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  S.e_wb_blk_ptr = M_<int>(c, "wb_blk_ptr");
  S.e_wb_st_off = M_<int>(c, "wb_st_off");
  S.e_wb_pairs = M_<uint32_t>(c, "wb_pairs");
  S.e_wb_brounds = M_<int>(c, "wb_brounds");
  S.e_wb_bcbase = M_<int>(c, "wb_bcbase");
  S.e_wb_sg_ptr = M_<int>(c, "wb_sg_ptr");
  S.e_wb_sg_off = M_<int>(c, "wb_sg_off");
  S.e_wb_sg_idx = M_<int>(c, "wb_sg_idx");
  S.e_wb_stride = stride;
  S.e_wb_nband = B.n_band;
  S.e_wb_M = M;
  return SGDLM_OK;
}

/* LDS layout and banded Schur tables of the register-resident LU
 * (k_lu_rs.hip) for plan P.  Dense blocks of up to 12 tiles: four
 * wavefronts, two workgroups per CU in half the LDS each: two flags and
 * the reduction words, then the prefix slots and tables followed by the
 * staging band of S, all aliased by the staged panel tiles of the dense
 * LU (the slots are dead once S is gathered; the tables are loaded per
 * draw).  Otherwise (or when that does not fit) eight wavefronts, one
 * workgroup per CU: the prefix slots and tables, the flags, then one region
 * shared by the staged panel tiles and the staging band.  Leaves
 * e_rb_rows = 0 when the dense block is outside 6..18 tiles or does not
 * fit. */
int build_band_tables(sgdlm_ctx *c, const ElimPlan &P) {
  Structure &S = c->S;
  S.e_rb_rows = 0;
  if (P.M2 / 16 < 6 || P.M2 / 16 > kRsMaxTiles)
    return SGDLM_OK;
  const size_t pre = (((size_t)P.ns * 4 + 15) & ~(size_t)15) + P.tab.size() * 4;
  const size_t pre_al = (pre + 15) & ~(size_t)15;
  int nw = 8, MT = rs_padded_tiles(P.M2 / 16), stride = rs_stage_stride(MT), BR = 0;
  int tab_persist = 1;
  size_t flag_base = 0, panel_base = 0, pre_base = 0, stage_base = 0, lds = 0;
  if (P.M2 / 16 <= 12) {
    const int MT4 = rs_padded_tiles(P.M2 / 16, 4), st4 = rs_stage_stride(MT4, 4);
    const size_t pb = 16 + 8 * sizeof(double) + 16, pnl = rs_panel_bytes(MT4);
    for (int br = 16; br <= 16 * MT4; br += 16)
      if (pb + std::max(pnl, pre_al + (size_t)br * st4 * 4) <= c->lds_max / 2)
        BR = br;
    if (BR) {
      nw = 4;
      MT = MT4;
      stride = st4;
      tab_persist = 0;
      flag_base = 0;
      panel_base = pb;
      pre_base = pb;
      stage_base = pb + pre_al;
      lds = pb + std::max(pnl, pre_al + (size_t)BR * st4 * 4);
    }
  }
  if (nw == 8) {
    flag_base = pre_al;
    panel_base = flag_base + 16 + 8 * sizeof(double);
    const size_t panel = rs_panel_bytes(MT);
    for (int br = 16; br <= 16 * MT; br += 16)
      if (panel_base + std::max(panel, (size_t)br * stride * 4) <= c->lds_max)
        BR = br;
    if (BR == 0)
      return SGDLM_OK;
    pre_base = 0;
    stage_base = panel_base;
    lds = panel_base + std::max(panel, ((size_t)BR * stride * 4 + 15) & ~(size_t)15);
  }
  BandPlan B;
  build_band_plan(P, BR, stride, 16 * MT, B);
  int rc;
  if ((rc = upload(c, "rb_blk_ptr", B.blk_ptr)) || (rc = upload(c, "rb_st_off", B.st_off)) ||
      (rc = upload(c, "rb_pairs", B.st_pairs)) || (rc = upload(c, "rb_brounds", B.brounds)) ||
      (rc = upload(c, "rb_bcbase", B.bcbase)) || (rc = upload(c, "rb_sg_ptr", B.sg_ptr)) ||
      (rc = upload(c, "rb_sg_off", B.sg_off)) || (rc = upload(c, "rb_sg_idx", B.sg_idx)))
    return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  S.e_rb_blk_ptr = M_<int>(c, "rb_blk_ptr");
  S.e_rb_st_off = M_<int>(c, "rb_st_off");
  S.e_rb_pairs = M_<uint32_t>(c, "rb_pairs");
  S.e_rb_brounds = M_<int>(c, "rb_brounds");
  S.e_rb_bcbase = M_<int>(c, "rb_bcbase");
  S.e_rb_sg_ptr = M_<int>(c, "rb_sg_ptr");
  S.e_rb_sg_off = M_<int>(c, "rb_sg_off");
  S.e_rb_sg_idx = M_<int>(c, "rb_sg_idx");
  S.e_rb_stride = stride;
  S.e_rb_nband = B.n_band;
  S.e_rb_flag_base = flag_base;
  S.e_rb_panel_base = panel_base;
  S.e_rb_lds = lds;
  S.e_rb_rows = BR;
  S.e_rb_tiles = MT;
  S.e_rb_nw = nw;
  S.e_rb_tab_persist = tab_persist;
  S.e_rb_pre_base = pre_base;
  S.e_rb_stage_base = stage_base;
  return SGDLM_OK;
}

/* Symbolic elimination plan of the VB log-determinant LU for the current
 * parental sets (lu_symbolic.h).  Its LDS slots (FP64) and tables must fit
 * the LDS of one workgroup; the front limit is halved until they do.
 * front >= 0 (the autotuner): that limit instead of the context's rule. */
int build_elimination(sgdlm_ctx *c, const std::vector<GammaEntry> &ents, int front = -1) {
  ElimPlan P;
  const size_t budget = c->lds_max > 1024 ? c->lds_max - 1024 : 0;
  const int max_slots = (int)std::min<size_t>(65536, budget / 8);
  auto fits = [&](const ElimPlan &Q) {
    return (size_t)Q.ns * 8 + 16 + Q.tab.size() * 4 <= budget;
  };
  int mf = front >= 0 ? front : c->max_front;
  for (;; mf /= 2) {
    build_elim_plan((int)c->m, ents, mf, max_slots, P);
    if (fits(P) || mf == 0)
      break;
  }
  /* Unless the front limit was set explicitly: when the dense block just
   * misses the size up to which the 64-wide fast LU fits its LDS budget, a
   * somewhat larger front (a few more prefix eliminations) usually brings
   * it under; that costs far less than the narrower panels would. */
  const int cap = lu_dense_cap(c->lds_max);
  if (front < 0 && !c->max_front_fixed && mf == c->max_front && P.M2 > cap && P.M2 <= cap + 64) {
    for (int mf2 = mf + 4; mf2 <= 64; mf2 += 4) {
      ElimPlan Q;
      build_elim_plan((int)c->m, ents, mf2, max_slots, Q);
      if (!fits(Q))
        break;
      if (Q.M2 <= cap) {
        P = std::move(Q);
        mf = mf2;
        break;
      }
    }
  }
  c->max_front_used = mf;
  int rc;
  if ((rc = upload(c, "e_slot_init", P.slot_init)) || (rc = upload(c, "e_tab", P.tab)) ||
      (rc = upload(c, "e_sg_off", P.sg_off)) || (rc = upload(c, "e_sg_idx", P.sg_idx)) ||
      (rc = upload(c, "e_st_off", P.st_off)) || (rc = upload(c, "e_st_pairs", P.st_pairs)) ||
      (rc = upload(c, "e_wave_q", P.wave_q)) || (rc = upload(c, "e_node_of", P.node_of)) ||
      (rc = upload(c, "e_fwd_ptr", P.fwd_ptr)) || (rc = upload(c, "e_fwd", P.fwd)) ||
      (rc = upload(c, "e_bwd_ptr", P.bwd_ptr)) || (rc = upload(c, "e_bwd", P.bwd)))
    return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  Structure &S = c->S;
  S.e_n1 = P.n1;
  S.e_n2 = P.n2;
  S.e_M2 = P.M2;
  S.e_ns = P.ns;
  S.e_waves = P.n_waves;
  S.e_nsg = (int)P.sg_off.size();
  S.e_nblk = P.n_blk;
  S.e_tab_words = (int)P.tab.size();
  S.e_o_piv = P.o_piv;
  S.e_o_mptr = P.o_mptr;
  S.e_o_mult = P.o_mult;
  S.e_o_tptr = P.o_tptr;
  S.e_o_tslot = P.o_tslot;
  S.e_o_cptr = P.o_cptr;
  S.e_o_contrib = P.o_contrib;
  S.e_o_brounds = P.o_brounds;
  S.e_o_bcbase = P.o_bcbase;
  S.e_zero_pair = P.zero_pair;
  S.e_fma = P.prefix_fma;
  S.e_slot_init = M_<int>(c, "e_slot_init");
  S.e_tab = M_<int>(c, "e_tab");
  S.e_sg_off = M_<int>(c, "e_sg_off");
  S.e_sg_idx = M_<int>(c, "e_sg_idx");
  S.e_st_off = M_<int>(c, "e_st_off");
  S.e_st_pairs = M_<uint32_t>(c, "e_st_pairs");
  S.e_wave_q = M_<int>(c, "e_wave_q");
  S.e_node_of = M_<int>(c, "e_node_of");
  S.e_fwd_ptr = M_<int>(c, "e_fwd_ptr");
  S.e_fwd = M_<uint32_t>(c, "e_fwd");
  S.e_bwd_ptr = M_<int>(c, "e_bwd_ptr");
  S.e_bwd = M_<uint32_t>(c, "e_bwd");
  ++c->gen;
  int rc2;
  if ((rc2 = build_wave_band_tables(c, P)))
    return rc2;
  return build_band_tables(c, P);
}

/* Rebuild the device-side parent structure from h_p / h_sp. */
int build_structure(sgdlm_ctx *c) {
  const int m = (int)c->m, mp = (int)c->max_p;
  const unsigned mm = (unsigned)m * (unsigned)m;
  std::vector<int> off(m + 1, 0);
  for (int j = 0; j < m; ++j)
    off[j + 1] = off[j] + (int)c->h_p[j];
  typedef GammaEntry E;
  std::vector<E> ents;
  for (int j = 0; j < m; ++j)
    for (int k = 0; k < (int)c->h_p[j]; ++k) {
      const unsigned v = c->h_sp[(size_t)j * mp + k];
      if (v < mm)
        ents.push_back({(int)(v % (unsigned)m), (int)(v / (unsigned)m), off[j] + k});
    }
  std::sort(ents.begin(), ents.end(), [](const E &a, const E &b) {
    return a.row != b.row ? a.row < b.row : a.col < b.col;
  });
  const int ng = (int)ents.size();
  std::vector<int> gr(std::max(ng, 1)), gc(std::max(ng, 1)), gi(std::max(ng, 1)),
      rp(m + 1, 0);
  for (int e = 0; e < ng; ++e) {
    gr[e] = ents[e].row;
    gc[e] = ents[e].col;
    gi[e] = ents[e].idx;
    rp[ents[e].row + 1]++;
  }
  for (int r = 0; r < m; ++r)
    rp[r + 1] += rp[r];
  int rc;
  if ((rc = dalloc(c, c->mem, "off", (m + 1) * 4)) ||
      (rc = dalloc(c, c->mem, "g_row", std::max(ng, 1) * 4)) ||
      (rc = dalloc(c, c->mem, "g_col", std::max(ng, 1) * 4)) ||
      (rc = dalloc(c, c->mem, "g_idx", std::max(ng, 1) * 4)) ||
      (rc = dalloc(c, c->mem, "g_rowptr", (m + 1) * 4)))
    return rc;
  if ((rc = h2d(c, M_<int>(c, "off"), off.data(), (m + 1) * 4)) ||
      (rc = h2d(c, M_<int>(c, "g_row"), gr.data(), std::max(ng, 1) * 4)) ||
      (rc = h2d(c, M_<int>(c, "g_col"), gc.data(), std::max(ng, 1) * 4)) ||
      (rc = h2d(c, M_<int>(c, "g_idx"), gi.data(), std::max(ng, 1) * 4)) ||
      (rc = h2d(c, M_<int>(c, "g_rowptr"), rp.data(), (m + 1) * 4)))
    return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  Structure &S = c->S;
  S.m = m;
  S.max_p = mp;
  S.M = (m + 15) / 16 * 16;
  S.P = off[m];
  S.n_gam = ng;
  S.p = M_<unsigned>(c, "p");
  S.sp = M_<unsigned>(c, "sp");
  S.off = M_<int>(c, "off");
  S.g_row = M_<int>(c, "g_row");
  S.g_col = M_<int>(c, "g_col");
  S.g_idx = M_<int>(c, "g_idx");
  S.g_rowptr = M_<int>(c, "g_rowptr");
  c->ents = ents;
  if ((rc = build_elimination(c, ents)))
    return rc;
  c->max_front_default = c->max_front_used;
  return SGDLM_OK;
}

/* Key of the tuned-configuration table: FNV-1a over what the LU plans
 * depend on (sizes, parental sets, precision, device limits, front rule). */
uint64_t structure_key(const sgdlm_ctx *c) {
  uint64_t h = 1469598103934665603ull;
  auto mix = [&h](uint64_t v) {
    for (int i = 0; i < 8; ++i) {
      h ^= (v >> (8 * i)) & 0xffu;
      h *= 1099511628211ull;
    }
  };
  mix(c->m);
  mix(c->max_p);
  mix((uint64_t)c->prec);
  mix((uint64_t)c->n_cu);
  mix(c->lds_max);
  mix((uint64_t)c->max_front);
  for (size_t j = 0; j < c->h_p.size(); ++j) {
    mix(c->h_p[j]);
    for (unsigned k = 0; k < c->h_p[j]; ++k)
      mix(c->h_sp[j * c->max_p + k]);
  }
  return h | 1u; /* never 0 (= no tuned configuration) */
}

/* A choice pinned by the caller turns the autotuner off. */
bool lu_pinned(const sgdlm_ctx *c) {
  return c->lu_threads || c->lu_nb || c->lu_kind != SGDLM_LU_AUTO || c->max_front_fixed;
}

/* The VB LU plan: pinned choices, else the tuned configuration of the
 * current structure (when its front limit is the one built), else the
 * plan's defaults. */
int vb_lu_plan(sgdlm_ctx *c, LuPlan *plan) {
  int th = c->lu_threads, nb = c->lu_nb, kind = c->lu_kind;
  const uint64_t key = structure_key(c);
  if (!lu_pinned(c) && c->autotune && c->tuned_key != key) {
    /* tuned by another context of the process: usable as is when its front
     * limit is the one built (no re-planning here) */
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune.find(key);
    if (it != g_tune.end() && it->second.max_front == c->max_front_used) {
      c->tuned = it->second;
      c->tuned.from_cache = true;
      c->tuned.tune_ms = 0.0;
      c->tuned_key = key;
    }
  }
  if (!lu_pinned(c) && c->autotune && c->tuned_key == key &&
      c->tuned.max_front == c->max_front_used) {
    th = c->tuned.threads;
    nb = c->tuned.nb;
    kind = c->tuned.kind;
  }
  return lu_plan(c->S, c->prec, 0, c->n_cu, c->lds_max, th, nb, plan, kind);
}

int lu_workspace(sgdlm_ctx *c, int prec, int solve, LuPlan *plan) {
  int rc = solve ? lu_plan(c->S, prec, 1, c->n_cu, c->lds_max, 0, 0, plan, c->lu_solve_kind)
                 : vb_lu_plan(c, plan);
  if (rc)
    return fail(c, rc, "no LU configuration fits m=%zu (M=%d) in LDS", c->m, c->S.M);
  return dalloc(c, c->sim, "lu_work", plan->work_bytes);
}

/* Device buffers of one VB call beyond the model state: the context's
 * simulation pool (run_vb) or the static scratch array (the allocation-free
 * entry point). */
struct VbDev {
  double *lam = nullptr, *theta = nullptr, *logdet = nullptr;
  int *lu_flags = nullptr;
  void *lu_work = nullptr;
  VbBuffers B;
};

/* Times every feasible VB LU configuration (kernel kind x workgroup size x
 * panel width, pivoting for every draw, for the default front limit and
 * 24 / 48) on the first
 * min(N, 4096) draws of D (two runs each, the faster counts; one for a
 * candidate 1.5x slower than the best so far) and returns
 * the fastest.  The draws' log-determinants are recomputed by the caller;
 * leaves the elimination plan of the last front tried. */
int tune_vb_lu(sgdlm_ctx *c, long N, VbDev &D, LuTuned *best) {
  const long nt = std::min<long>(N, 4096);
  const auto h0 = std::chrono::steady_clock::now();
  std::vector<int> fronts{c->max_front_default};
  if (c->max_front > 0)
    for (int f : {24, 48})
      if (f != c->max_front_default)
        fronts.push_back(f);
  struct Cand {
    int kind, threads, nb;
  };
  std::vector<Cand> cands;
  if (c->prec != SGDLM_PREC_FP64)
    cands.push_back({SGDLM_LU_REGISTER, 0, 0});
  /* the register kernel with partial pivoting: wins where most draws need
   * row interchanges (early in a series) */
  if (c->prec == SGDLM_PREC_FP32)
    cands.push_back({SGDLM_LU_REGISTER_PIVOT, 0, 0});
  for (int th : {256, 512})
    for (int nb : {64, 32, 16})
      cands.push_back({SGDLM_LU_BLOCKED, th, nb});
  /* always feasible: row interchanges for every draw (the only kind for
   * dense blocks beyond the fast kernels' LDS) */
  cands.push_back({SGDLM_LU_PIVOTING, 0, 0});
  /* small models: one wavefront per draw (no prefix: timed at the default
   * front only) */
  const bool wave_ok = lu_wave_dim(c->S.M, c->prec) > 0;
  hipEvent_t e0, e1;
  HIPCHK(c, hipEventCreate(&e0));
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return fail(c, SGDLM_ERR_DEVICE, "hipEventCreate failed");
  }
  int rc = SGDLM_OK, n = 0;
  double best_ms = 1e300, pa_ms = 1e300;
  int pa_kind = SGDLM_LU_PIVOTING;
  bool have_np = false; /* a kind that is not interchanges-for-every-draw */
  for (int mf : fronts) {
    if (mf != c->max_front_used && (rc = build_elimination(c, c->ents, mf)))
      break;
    std::vector<Cand> cs = cands;
    if (wave_ok && mf == c->max_front_default)
      cs.push_back({SGDLM_LU_WAVE, 0, 0});
    /* dense blocks up to 128 rows after this front's prefix */
    if (c->S.e_wb_M > 0)
      cs.push_back({SGDLM_LU_WAVE_PREFIX, 0, 0});
    for (const Cand &k : cs) {
      LuPlan p;
      if (lu_plan(c->S, c->prec, 0, c->n_cu, c->lds_max, k.threads, k.nb, &p, k.kind))
        continue;
      if ((rc = dalloc(c, c->sim, "lu_work", p.work_bytes)))
        break;
      D.lu_work = c->sim["lu_work"].ptr;
      float ms = 1e30f;
      for (int rep = 0; rep < 2 && !rc; ++rep) {
        float t = 0.f;
        if (hipEventRecord(e0, c->stream) != hipSuccess ||
            launch_lu_logdet(c->S, p, nt, D.theta, D.lu_work, D.logdet, D.lu_flags, c->stream) ||
            hipEventRecord(e1, c->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&t, e0, e1) != hipSuccess)
          rc = fail(c, SGDLM_ERR_DEVICE, "LU autotuning launch failed: %s",
                    hipGetErrorString(hipGetLastError()));
        ms = std::min(ms, t);
        if (t > 1.5 * best_ms) /* clearly slower: no second run */
          break;
      }
      if (rc)
        break;
      ++n;
      /* the kinds that pivot every draw are kept apart: timed on early
       * draws they win, and a structure keeps its tuned kind for the rest
       * of the series; a call that finds nearly every draw flagged runs
       * them instead (run_vb) */
      const bool pa = k.kind == SGDLM_LU_PIVOTING || k.kind == SGDLM_LU_REGISTER_PIVOT;
      if (pa && ms < pa_ms) {
        pa_ms = ms;
        pa_kind = k.kind;
      }
      if (pa && have_np)
        continue;
      if (!pa && !have_np) {
        have_np = true;
        best_ms = 1e30f;
      }
      if (ms < best_ms) {
        best_ms = ms;
        const long wgs = p.wave_M > 0   ? std::min<long>(p.wave_grid, (nt + 3) / 4)
                         : p.wsch_M > 0 ? std::min<long>(p.wsch_grid, (nt + 3) / 4)
                         : p.rs_MT > 0  ? p.rs_grid
                         : p.nb        ? p.fast_grid
                                       : p.pivot_grid;
        best->draw_wg_ms = (double)ms * (double)std::min<long>(wgs, nt) / (double)nt;
        best->kind = k.kind;
        best->threads = k.threads;
        best->nb = k.nb;
        /* the front limit the plan was built with: build_elimination halves
         * a requested limit whose tables do not fit LDS */
        best->max_front = c->max_front_used;
      }
    }
    if (rc)
      break;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc)
    return rc;
  if (n == 0)
    return fail(c, SGDLM_ERR_ARG, "no LU configuration fits m=%zu in LDS", c->m);
  best->best_ms = best_ms;
  best->pa_kind = pa_kind;
  best->pa_ms = pa_ms < 1e299 ? pa_ms : 0.0;
  best->candidates = n;
  best->from_cache = false;
  best->tune_ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - h0).count();
  return SGDLM_OK;
}

/* Before the LU of a VB call: the tuned configuration of the current
 * structure from the process table, or tuned now on this call's draws;
 * rebuilds the elimination plan for its front limit and re-plans. */
int apply_lu_tuning(sgdlm_ctx *c, long N, VbDev &D, LuPlan &plan) {
  if (lu_pinned(c) || !c->autotune)
    return SGDLM_OK;
  const uint64_t key = structure_key(c);
  if (c->tuned_key == key && c->tuned.max_front == c->max_front_used)
    return SGDLM_OK;
  LuTuned t;
  bool found = false;
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune.find(key);
    if (it != g_tune.end()) {
      t = it->second;
      found = true;
    }
  }
  int rc;
  if (found) {
    t.from_cache = true;
    t.tune_ms = 0.0;
  } else {
    if ((rc = tune_vb_lu(c, N, D, &t)))
      return rc;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune[key] = t;
  }
  if (t.max_front != c->max_front_used && (rc = build_elimination(c, c->ents, t.max_front)))
    return rc;
  c->tuned = t;
  c->tuned_key = key;
  if ((rc = vb_lu_plan(c, &plan)))
    return fail(c, rc, "the tuned LU configuration does not fit");
  if ((rc = dalloc(c, c->sim, "lu_work", plan.work_bytes)))
    return rc;
  D.lu_work = c->sim["lu_work"].ptr;
  return SGDLM_OK;
}

/* The VB stages on the device (no host copies) for N draws; ends with the
 * weight statistics copied to the pinned c->h_stats.  tune: the LU
 * configuration may be tuned on this call's draws (device buffers from the
 * context's pools only). */
/* Enqueues the VB stages for N draws at epoch ep (or the epoch in the
 * device word ep_dev) on the context's stream; no host synchronisation, no
 * allocation unless tune (so that a capture can record it).  Ends with the
 * weight statistics copied to the pinned c->h_stats. */
int enqueue_vb(sgdlm_ctx *c, long N, VbDev &D, LuPlan &plan, bool t_on, bool tune, uint32_t ep,
               const uint32_t *ep_dev) {
  int rc;
  {
    StageTimer t(c, "draw", t_on);
    launch_cholesky(c->S, M_<double>(c, "C"), M_<double>(c, "U"), c->stream);
    launch_draw(c->S, N, M_<double>(c, "m"), M_<double>(c, "U"), M_<double>(c, "n"),
                M_<double>(c, "s"), c->seed, ep, D.lam, D.theta, c->stream, ep_dev);
  }
  CHECK_LAUNCH(c);
  if (tune && (rc = apply_lu_tuning(c, N, D, plan)))
    return rc;
  {
    StageTimer t(c, "lu", t_on);
    rc = launch_lu_logdet(c->S, plan, N, D.theta, D.lu_work, D.logdet, D.lu_flags, c->stream);
  }
  if (rc)
    return fail(c, rc, "LU launch failed: %s", hipGetErrorString(hipGetLastError()));
  {
    StageTimer t(c, "weights", t_on);
    launch_weights(N, D.logdet, D.B.w, D.B.stats, c->stream);
  }
  CHECK_LAUNCH(c);
  {
    StageTimer t(c, "moments", t_on);
    launch_moments(c->S, N, D.lam, D.theta, M_<double>(c, "m"), M_<double>(c, "n"),
                   M_<double>(c, "s"), D.B, c->stream);
  }
  CHECK_LAUNCH(c);
  HIPCHK(c, hipMemcpyAsync(c->h_stats, D.B.stats, 4 * sizeof(double), hipMemcpyDeviceToHost,
                           c->stream));
  /* the LU's fallback count (draws handed to the partial-pivoting kernel) */
  HIPCHK(c, hipMemcpyAsync(c->h_stats + 4, D.lu_flags, sizeof(int), hipMemcpyDeviceToHost,
                           c->stream));
  return SGDLM_OK;
}

int run_vb_stages(sgdlm_ctx *c, long N, VbDev D, LuPlan plan, bool timed, bool tune) {
  int rc;
  if ((rc = check_epochs(c, 1)))
    return rc;
  const uint32_t ep = ++c->epoch;
  return enqueue_vb(c, N, D, plan, timed && c->timing, tune, ep, nullptr);
}

void drop_graph(sgdlm_ctx *c) {
  if (c->vbg.exec)
    (void)hipGraphExecDestroy(c->vbg.exec);
  if (c->vbg.graph)
    (void)hipGraphDestroy(c->vbg.graph);
  c->vbg = VbGraph();
}

/* Everything a captured VB step bakes into its kernel arguments. */
std::vector<uintptr_t> vb_signature(sgdlm_ctx *c, long N, const VbDev &D, const LuPlan &p) {
  const VbBuffers &B = D.B;
  return {(uintptr_t)c->gen, (uintptr_t)c->prec, (uintptr_t)N, (uintptr_t)c->seed,
          (uintptr_t)(c->seed >> 32), (uintptr_t)D.lam, (uintptr_t)D.theta,
          (uintptr_t)D.logdet, (uintptr_t)D.lu_flags, (uintptr_t)D.lu_work, (uintptr_t)B.w,
          (uintptr_t)B.stats, (uintptr_t)B.part1, (uintptr_t)B.part2, (uintptr_t)B.mean_m,
          (uintptr_t)B.mean_C, (uintptr_t)B.mean_n, (uintptr_t)B.mean_s, (uintptr_t)B.chunks,
          (uintptr_t)M_<double>(c, "m"), (uintptr_t)M_<double>(c, "C"),
          (uintptr_t)M_<double>(c, "U"), (uintptr_t)M_<double>(c, "n"),
          (uintptr_t)M_<double>(c, "s"), (uintptr_t)c->h_stats, (uintptr_t)p.nb,
          (uintptr_t)p.threads, (uintptr_t)p.fast_grid, (uintptr_t)p.lds, (uintptr_t)p.grid,
          (uintptr_t)p.rs_MT, (uintptr_t)p.rs_grid, (uintptr_t)p.rs_lds, (uintptr_t)p.pivot_nb,
          (uintptr_t)p.pivot_grid, (uintptr_t)p.pivot_lds, (uintptr_t)p.pivot_schur_M,
          (uintptr_t)p.work_bytes, (uintptr_t)p.wave_M, (uintptr_t)p.wave_grid,
          (uintptr_t)p.wsch_M, (uintptr_t)p.wsch_grid, (uintptr_t)p.wsch_lds,
          (uintptr_t)p.rs_piv, (uintptr_t)p.rs_swap_base, (uintptr_t)p.rs_nw};
}

/* Graph mode: the VB step of run_vb captured once as a hipGraph (re-captured
 * when anything it baked in changes) and replayed; the epoch reaches the
 * draw kernels through a device word set by a one-thread kernel first. */
int run_vb_graph(sgdlm_ctx *c, long N, VbDev D, LuPlan plan) {
  int rc;
  if ((rc = check_epochs(c, 1)))
    return rc;
  if ((rc = dalloc(c, c->mem, "epoch", 16)))
    return rc;
  uint32_t *ep_dev = M_<uint32_t>(c, "epoch");
  const uint32_t ep = ++c->epoch;
  if (set_device_int(reinterpret_cast<int *>(ep_dev), (int)ep, c->stream))
    return fail(c, SGDLM_ERR_DEVICE, "epoch store failed: %s",
                hipGetErrorString(hipGetLastError()));
  std::vector<uintptr_t> sig = vb_signature(c, N, D, plan);
  sig.push_back((uintptr_t)ep_dev);
  if (!c->vbg.exec || c->vbg.sig != sig) {
    drop_graph(c);
    HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
    const int erc = enqueue_vb(c, N, D, plan, false, false, ep, ep_dev);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (erc || e != hipSuccess) {
      if (g)
        (void)hipGraphDestroy(g);
      (void)hipGetLastError();
      return erc ? erc : fail(c, SGDLM_ERR_DEVICE, "graph capture failed: %s", hipGetErrorString(e));
    }
    hipGraphExec_t x = nullptr;
    const hipError_t ei = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (ei != hipSuccess) {
      (void)hipGraphDestroy(g);
      (void)hipGetLastError();
      return fail(c, SGDLM_ERR_DEVICE, "graph instantiation failed: %s", hipGetErrorString(ei));
    }
    c->vbg.graph = g;
    c->vbg.exec = x;
    c->vbg.sig = sig;
    ++c->graph_captures;
  }
  HIPCHK(c, hipGraphLaunch(c->vbg.exec, c->stream));
  ++c->graph_replays;
  return SGDLM_OK;
}

/* The autotuner would time configurations on this call's draws. */
bool tuning_pending(sgdlm_ctx *c) {
  return !lu_pinned(c) && c->autotune &&
         !(c->tuned_key == structure_key(c) && c->tuned.max_front == c->max_front_used);
}

/* Adaptive LU of a VB call (kind not pinned, structure tuned): after a call
 * whose no-interchange kernel flagged more than 90 % of the draws (early in
 * a series) the call pivots every draw with the tuned kind for that,
 * trying the no-interchange kernel again every 8th call.  Sets c->ran_pa;
 * the counts are taken by vb_status. */
void adaptive_vb_plan(sgdlm_ctx *c, LuPlan *plan) {
  c->ran_pa = false;
  c->run_n = 0;
  const int pk = lu_plan_kind(*plan);
  if (lu_pinned(c) || tuning_pending(c) || (pk != SGDLM_LU_BLOCKED && pk != SGDLM_LU_REGISTER))
    return;
  const bool high = c->fb_n > 0 && (double)c->fb_count > 0.9 * (double)c->fb_n;
  if (!high) {
    c->pa_streak = 0;
    return;
  }
  if (++c->pa_streak % 8 == 0)
    return;
  LuPlan pa;
  if (!lu_plan(c->S, c->prec, 0, c->n_cu, c->lds_max, 0, 0, &pa, c->tuned.pa_kind) ||
      !lu_plan(c->S, c->prec, 0, c->n_cu, c->lds_max, 0, 0, &pa, SGDLM_LU_PIVOTING)) {
    *plan = pa;
    c->ran_pa = true;
  }
}

/* VB on the context's simulation pool (results stay on the device). */
int run_vb(sgdlm_ctx *c) {
  int rc;
  if ((rc = ensure_sim_buffers(c)))
    return rc;
  LuPlan plan;
  if ((rc = lu_workspace(c, c->prec, 0, &plan)))
    return rc;
  VbDev D;
  D.lam = SIM<double>(c, "lam");
  D.theta = SIM<double>(c, "theta");
  D.logdet = SIM<double>(c, "logdet");
  D.lu_flags = SIM<int>(c, "lu_flags");
  adaptive_vb_plan(c, &plan);
  if (c->ran_pa && (rc = dalloc(c, c->sim, "lu_work", plan.work_bytes)))
    return rc;
  D.lu_work = c->sim["lu_work"].ptr;
  D.B = vb_buffers(c);
  c->run_n = (long)c->nsim;
  if (c->graph_mode && !c->timing && !tuning_pending(c))
    rc = run_vb_graph(c, (long)c->nsim, D, plan);
  else
    rc = run_vb_stages(c, (long)c->nsim, D, plan, true, true);
  if (rc)
    return rc;
  c->lu_flags_valid = true;
  c->vb_valid = true;
  return SGDLM_OK;
}

/* After the stream of a VB call has drained: every importance weight zero
 * (each |det(I - Gamma_i)| zero or non-finite) is an error, also when the
 * results stay on the device. */
int vb_status(sgdlm_ctx *c) {
  if (c->run_n > 0 && !c->ran_pa) {
    std::memcpy(&c->fb_count, c->h_stats + 4, sizeof(int));
    c->fb_n = c->run_n;
  }
  c->run_n = 0;
  if (!(c->h_stats[1] > 0.0))
    return fail(c, SGDLM_ERR_NUMERIC,
                "every importance weight is zero (all |det(I-Gamma)| zero or non-finite)");
  return SGDLM_OK;
}

/* ---- a VB step of several models, one launch per stage ---- */

/* The device buffers and LU plan run_vb would use for c (no allocation of
 * the per-context LU workspace: the group brings its own). */
int group_member(sgdlm_ctx *c, VbDev *D, LuPlan *plan) {
  int rc;
  if ((rc = ensure_sim_buffers(c)))
    return rc;
  if ((rc = vb_lu_plan(c, plan)))
    return fail(c, rc, "no LU configuration fits m=%zu (M=%d) in LDS", c->m, c->S.M);
  adaptive_vb_plan(c, plan);
  c->run_n = (long)c->nsim;
  D->lam = SIM<double>(c, "lam");
  D->theta = SIM<double>(c, "theta");
  D->logdet = SIM<double>(c, "logdet");
  D->lu_flags = SIM<int>(c, "lu_flags");
  D->lu_work = nullptr;
  D->B = vb_buffers(c);
  return SGDLM_OK;
}

/* Builds the device tables of a step over the n contexts (all on one
 * device, checked, with their buffers in place).  SGDLM_ERR_ARG: this list
 * does not fit the group kernels (the caller runs the contexts one by one). */
int build_vb_group(sgdlm_ctx *const *cs, int n, const std::vector<VbDev> &Ds,
                   const std::vector<LuPlan> &plans, bool side_streams, VbGroup **out) {
  sgdlm_ctx *c0 = cs[0];
  VbGroup *g = new VbGroup();
  g->n = n;
  std::vector<VbModel> tab(n);
  std::vector<int> ser;
  for (int i = 0; i < n; ++i) {
    sgdlm_ctx *c = cs[i];
    const VbDev &D = Ds[i];
    VbModel &v = tab[i];
    v.m = c->S.m;
    v.max_p = c->S.max_p;
    v.chunks = D.B.chunks;
    v.ser0 = (int)ser.size();
    v.N = (long)c->nsim;
    v.P = c->S.P;
    v.seed = c->seed;
    v.p = c->S.p;
    v.off = c->S.off;
    v.m_t = M_<double>(c, "m");
    v.C = M_<double>(c, "C");
    v.n_t = M_<double>(c, "n");
    v.s_t = M_<double>(c, "s");
    v.U = M_<double>(c, "U");
    v.lam = D.lam;
    v.theta = D.theta;
    v.logdet = D.logdet;
    v.lu_flags = D.lu_flags;
    v.w = D.B.w;
    v.stats = D.B.stats;
    v.part1 = D.B.part1;
    v.part2 = D.B.part2;
    v.mean_lambda = D.B.mean_lambda;
    v.mean_log_lambda = D.B.mean_log_lambda;
    v.mean_m = D.B.mean_m;
    v.mean_C = D.B.mean_C;
    v.mean_n = D.B.mean_n;
    v.mean_s = D.B.mean_s;
    v.mean_Q = D.B.mean_Q;
    v.newton_iter = D.B.newton_iter;
    for (int j = 0; j < v.m; ++j)
      ser.push_back(i);
    g->max_p = std::max(g->max_p, v.max_p);
    g->max_chunks = std::max(g->max_chunks, v.chunks);
    g->max_N = std::max(g->max_N, v.N);
  }
  g->n_series = (int)ser.size();
  auto give_up = [&](int rc) {
    drop_vb_group(g);
    return rc;
  };
  if (g->max_p > 32 || g->n_series > 65535)
    return give_up(SGDLM_ERR_ARG);

  /* LU stages: every model's kernels, grouped by kernel */
  struct Item {
    int model;
    LuGrpKernel k;
  };
  std::vector<std::vector<Item>> st1, st2;
  std::map<const void *, size_t> at1, at2;
  for (int i = 0; i < n; ++i) {
    LuGrpKernel k1, k2;
    int init = 0;
    /* the measured time of a draw, when the plan is the tuned one */
    sgdlm_ctx *c = cs[i];
    const bool tuned = !lu_pinned(c) && c->autotune && c->tuned_key == structure_key(c) &&
                       c->tuned.max_front == c->max_front_used;
    if (lu_grp_stages(c->S, plans[i], tab[i].N, tab[i].theta, tab[i].logdet, tab[i].lu_flags,
                      tuned ? c->tuned.draw_wg_ms : 0.0, &init, &k1, &k2))
      return give_up(SGDLM_ERR_ARG);
    tab[i].flag_init = init;
    tab[i].flag_final = plans[i].rs_MT > 0 && plans[i].rs_piv ? (int)tab[i].N : -1;
    auto add = [&](std::vector<std::vector<Item>> &st, std::map<const void *, size_t> &at,
                   LuGrpKernel &k) {
      auto it = at.find(k.fn);
      if (it == at.end()) {
        it = at.emplace(k.fn, st.size()).first;
        st.emplace_back();
      }
      st[it->second].push_back({i, std::move(k)});
    };
    add(st1, at1, k1);
    if (k2.fn)
      add(st2, at2, k2);
  }
  /* workgroups per model within each launch: in proportion to its work
   * (draws x time of a draw on one workgroup), so that every model's
   * workgroups finish together, at least one and at most one per unit of
   * work; the models with the slowest draws first in the grid.  The
   * workspace of a model serves both stages. */
  std::vector<size_t> ws(n, 0);
  struct Shape {
    int grid = 0, threads = 0;
    size_t lds = 0;
    double cost = 0.0;
    std::vector<int> wg;
  };
  auto shape = [&](std::vector<Item> &items) {
    std::stable_sort(items.begin(), items.end(), [&](const Item &a, const Item &b) {
      return a.k.cost / (double)std::max(1L, tab[a.model].N) >
             b.k.cost / (double)std::max(1L, tab[b.model].N);
    });
    Shape sh;
    int per_cu = 1 << 20;
    for (const Item &it : items) {
      per_cu = std::min(per_cu, it.k.per_cu);
      sh.cost += it.k.cost;
      sh.lds = std::max(sh.lds, it.k.lds);
      sh.threads = it.k.threads;
    }
    const double G = (double)c0->n_cu * per_cu;
    const double per_wg = sh.cost > 0.0 ? sh.cost / G : 1.0;
    for (const Item &it : items) {
      long w = sh.cost > 0.0 ? (long)std::ceil(it.k.cost / per_wg - 1e-9) : 1;
      w = std::max(1L, std::min(w, std::max(1L, it.k.units)));
      sh.wg.push_back((int)w);
      sh.grid += (int)w;
      ws[it.model] = std::max(ws[it.model], (size_t)w * it.k.ws_wg);
    }
    return sh;
  };
  std::vector<Shape> sh1, sh2;
  for (auto &items : st1)
    sh1.push_back(shape(items));
  for (auto &items : st2)
    sh2.push_back(shape(items));
  std::vector<size_t> ws_off(n, 0);
  size_t work_bytes = 0;
  for (int i = 0; i < n; ++i) {
    ws_off[i] = work_bytes;
    work_bytes += (ws[i] + 255) & ~(size_t)255;
  }
  if (work_bytes) {
    if (hipMalloc(&g->work, work_bytes) != hipSuccess) {
      (void)hipGetLastError();
      g->work = nullptr;
      drop_vb_group(g);
      return fail(c0, SGDLM_ERR_ALLOC, "device allocation of %zu bytes for the batch LU "
                                       "workspace failed", work_bytes);
    }
  }
  /* the blob: model table, series map, epochs, stats, then per launch its
   * argument table and workgroup map */
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t o = off;
    off += (bytes + 255) & ~(size_t)255;
    return o;
  };
  const size_t o_tab = take(sizeof(VbModel) * n), o_ser = take(sizeof(int) * ser.size()),
               o_ep = take(sizeof(uint32_t) * n), o_st = take(sizeof(double) * 5 * n);
  struct Pending {
    std::vector<unsigned char> args;
    std::vector<int> inst, base, count;
    size_t o_args, o_inst, o_base, o_count;
  };
  std::vector<Pending> pend;
  auto lay = [&](std::vector<Item> &items, const Shape &sh) {
    Pending p;
    const size_t asz = items[0].k.args.size();
    int b = 0;
    for (size_t q = 0; q < items.size(); ++q) {
      Item &it = items[q];
      if (it.k.args.size() != asz)
        return false;
      if (it.k.work_field >= 0) {
        void *wp = static_cast<char *>(g->work) + ws_off[it.model];
        std::memcpy(it.k.args.data() + it.k.work_field, &wp, sizeof(void *));
      }
      p.args.insert(p.args.end(), it.k.args.begin(), it.k.args.end());
      p.base.push_back(b);
      p.count.push_back(sh.wg[q]);
      for (int w = 0; w < sh.wg[q]; ++w)
        p.inst.push_back((int)q);
      b += sh.wg[q];
    }
    p.o_args = take(p.args.size());
    p.o_inst = take(sizeof(int) * p.inst.size());
    p.o_base = take(sizeof(int) * p.base.size());
    p.o_count = take(sizeof(int) * p.count.size());
    pend.push_back(std::move(p));
    return true;
  };
  for (size_t q = 0; q < st1.size(); ++q)
    if (!lay(st1[q], sh1[q]))
      return give_up(SGDLM_ERR_ARG);
  for (size_t q = 0; q < st2.size(); ++q)
    if (!lay(st2[q], sh2[q]))
      return give_up(SGDLM_ERR_ARG);
  std::vector<unsigned char> host(off, 0);
  std::memcpy(host.data() + o_tab, tab.data(), sizeof(VbModel) * n);
  std::memcpy(host.data() + o_ser, ser.data(), sizeof(int) * ser.size());
  for (const Pending &p : pend) {
    std::memcpy(host.data() + p.o_args, p.args.data(), p.args.size());
    std::memcpy(host.data() + p.o_inst, p.inst.data(), sizeof(int) * p.inst.size());
    std::memcpy(host.data() + p.o_base, p.base.data(), sizeof(int) * p.base.size());
    std::memcpy(host.data() + p.o_count, p.count.data(), sizeof(int) * p.count.size());
  }
  if (hipMalloc(&g->blob, off) != hipSuccess) {
    (void)hipGetLastError();
    g->blob = nullptr;
    drop_vb_group(g);
    return fail(c0, SGDLM_ERR_ALLOC, "device allocation of %zu bytes for the batch tables failed",
                off);
  }
  if (hipHostMalloc((void **)&g->h_epochs, sizeof(uint32_t) * n, hipHostMallocDefault) !=
          hipSuccess ||
      hipHostMalloc((void **)&g->h_stats, sizeof(double) * 5 * n, hipHostMallocDefault) !=
          hipSuccess) {
    (void)hipGetLastError();
    drop_vb_group(g);
    return fail(c0, SGDLM_ERR_ALLOC, "pinned host allocation for the batch failed");
  }
  if (hipMemcpy(g->blob, host.data(), off, hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipGetLastError();
    drop_vb_group(g);
    return fail(c0, SGDLM_ERR_DEVICE, "upload of the batch tables failed");
  }
  char *base = static_cast<char *>(g->blob);
  g->d_tab = reinterpret_cast<VbModel *>(base + o_tab);
  g->d_ser = reinterpret_cast<int *>(base + o_ser);
  g->d_epochs = reinterpret_cast<uint32_t *>(base + o_ep);
  g->d_stats = reinterpret_cast<double *>(base + o_st);
  size_t pi = 0;
  auto mk = [&](std::vector<Item> &items, const Shape &sh, std::vector<GrpLaunch> &dst) {
    const Pending &p = pend[pi++];
    GrpLaunch L;
    L.fn = items[0].k.fn;
    L.threads = sh.threads;
    L.grid = sh.grid;
    L.lds = sh.lds;
    L.a0 = items[0].k.args;
    L.tab = base + p.o_args;
    L.gm.inst = reinterpret_cast<const int *>(base + p.o_inst);
    L.gm.base = reinterpret_cast<const int *>(base + p.o_base);
    L.gm.count = reinterpret_cast<const int *>(base + p.o_count);
    dst.push_back(std::move(L));
  };
  std::vector<double> c1, c2;
  for (size_t q = 0; q < st1.size(); ++q) {
    mk(st1[q], sh1[q], g->lu1);
    c1.push_back(sh1[q].cost);
  }
  for (size_t q = 0; q < st2.size(); ++q) {
    mk(st2[q], sh2[q], g->lu2);
    c2.push_back(sh2[q].cost);
  }
  /* longest launch first */
  auto order = [](std::vector<GrpLaunch> &L, const std::vector<double> &cost) {
    std::vector<size_t> ix(L.size());
    for (size_t q = 0; q < ix.size(); ++q)
      ix[q] = q;
    std::stable_sort(ix.begin(), ix.end(), [&](size_t a, size_t b) { return cost[a] > cost[b]; });
    std::vector<GrpLaunch> out;
    for (size_t q : ix)
      out.push_back(std::move(L[q]));
    L = std::move(out);
  };
  order(g->lu1, c1);
  order(g->lu2, c2);
  /* side streams (default priority) and fork / join events */
  const size_t n_side =
      side_streams ? std::min<size_t>(kGroupSideStreams, std::max(g->lu1.size(), g->lu2.size()))
                   : 0;
  for (size_t q = 0; q < n_side; ++q) {
    hipStream_t st = nullptr;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
      (void)hipGetLastError();
      drop_vb_group(g);
      return fail(c0, SGDLM_ERR_DEVICE, "batch stream creation failed");
    }
    g->side.push_back(st);
  }
  for (size_t q = 0; q < 2 * n_side + 2; ++q) {
    hipEvent_t e = nullptr;
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
      (void)hipGetLastError();
      drop_vb_group(g);
      return fail(c0, SGDLM_ERR_DEVICE, "batch event creation failed");
    }
    g->ev.push_back(e);
  }
  *out = g;
  return SGDLM_OK;
}

int launch_grp_lu(sgdlm_ctx *c, const GrpLaunch &L, hipStream_t st) {
  if (allow_dynamic_lds(L.fn, L.lds) != hipSuccess)
    return fail(c, SGDLM_ERR_DEVICE, "batch LU: LDS attribute refused (%zu bytes)", L.lds);
  GrpMap gm = L.gm;
  const void *tab = L.tab;
  void *args[] = {const_cast<unsigned char *>(L.a0.data()), &tab, &gm};
  if (hipLaunchKernel(L.fn, dim3(L.grid), dim3(L.threads), args, L.lds, st) != hipSuccess)
    return fail(c, SGDLM_ERR_DEVICE, "batch LU launch failed: %s",
                hipGetErrorString(hipGetLastError()));
  return SGDLM_OK;
}

/* The launches of one LU stage, concurrently: forked from c's stream onto
 * the group's side streams (round robin, longest first) and joined back. */
int launch_grp_stage(sgdlm_ctx *c, VbGroup *g, const std::vector<GrpLaunch> &Ls,
                     hipStream_t st) {
  int rc;
  if (Ls.size() <= 1 || g->side.empty()) {
    for (const GrpLaunch &L : Ls)
      if ((rc = launch_grp_lu(c, L, st)))
        return rc;
    return SGDLM_OK;
  }
  const size_t ns = std::min(g->side.size(), Ls.size());
  HIPCHK(c, hipEventRecord(g->ev[0], st));
  for (size_t q = 0; q < ns; ++q)
    HIPCHK(c, hipStreamWaitEvent(g->side[q], g->ev[0], 0));
  for (size_t q = 0; q < Ls.size(); ++q)
    if ((rc = launch_grp_lu(c, Ls[q], g->side[q % ns])))
      return rc;
  for (size_t q = 0; q < ns; ++q) {
    HIPCHK(c, hipEventRecord(g->ev[1 + q], g->side[q]));
    HIPCHK(c, hipStreamWaitEvent(st, g->ev[1 + q], 0));
  }
  return SGDLM_OK;
}

/* The stages of a group step on stream st (capturable: kernels and one
 * device-to-host copy of the statistics); stage times on c's timers when
 * t_on (st is c's stream). */
int enqueue_vb_group(sgdlm_ctx *c, VbGroup *g, hipStream_t st, bool t_on) {
  int rc;
  {
    StageTimer t(c, "batch_draw", t_on);
    launch_cholesky_grp(g->d_tab, g->d_ser, g->n_series, g->max_p, st);
    if ((rc = launch_draw_grp(g->d_tab, g->d_ser, g->d_epochs, g->n_series, g->max_N, g->max_p,
                              st)))
      return fail(c, rc, "batch draw launch failed");
  }
  CHECK_LAUNCH(c);
  {
    StageTimer t(c, "batch_lu", t_on);
    /* the fallback counters were set by the Cholesky launch */
    if ((rc = launch_grp_stage(c, g, g->lu1, st)) || (rc = launch_grp_stage(c, g, g->lu2, st)))
      return rc;
  }
  {
    StageTimer t(c, "batch_weights", t_on);
    launch_weights_grp(g->d_tab, g->n, st);
  }
  CHECK_LAUNCH(c);
  {
    StageTimer t(c, "batch_moments", t_on);
    if ((rc = launch_moments_grp(g->d_tab, g->d_ser, g->n_series, g->max_chunks, g->max_p, st)))
      return fail(c, rc, "batch moment launch failed");
  }
  launch_grp_stats(g->d_tab, g->n, g->d_stats, st);
  CHECK_LAUNCH(c);
  HIPCHK(c, hipMemcpyAsync(g->h_stats, g->d_stats, sizeof(double) * 5 * g->n,
                           hipMemcpyDeviceToHost, st));
  return SGDLM_OK;
}

/* Every part of a batch: one part on c0's stream, or each part on its own
 * stream, forked from c0's stream and joined back. */
int enqueue_vb_batch(sgdlm_ctx *c0, VbBatch *b, bool t_on) {
  int rc;
  if (b->parts.size() == 1)
    return enqueue_vb_group(c0, b->parts[0], c0->stream, t_on);
  HIPCHK(c0, hipEventRecord(b->ev[0], c0->stream));
  for (size_t q = 0; q < b->parts.size(); ++q) {
    HIPCHK(c0, hipStreamWaitEvent(b->streams[q], b->ev[0], 0));
    if ((rc = enqueue_vb_group(c0, b->parts[q], b->streams[q], false)))
      return rc;
    HIPCHK(c0, hipEventRecord(b->ev[1 + q], b->streams[q]));
    HIPCHK(c0, hipStreamWaitEvent(c0->stream, b->ev[1 + q], 0));
  }
  return SGDLM_OK;
}

/* The parts of a batch of n models: up to kBatchParts (or
 * SGDLM_BATCH_PARTS from the environment), at least four models each, by a greedy split of the estimated work (the LU's measured
 * or modelled time plus the draws and moments, which scale with m x N). */
std::vector<std::vector<int>> split_batch(sgdlm_ctx *const *cs, int n,
                                          const std::vector<LuPlan> &plans) {
  int want = kBatchParts;
  if (const char *e = std::getenv("SGDLM_BATCH_PARTS"))
    want = std::max(1, std::min(16, std::atoi(e)));
  const int P = std::max(1, std::min(want, n / 4));
  std::vector<double> w(n);
  for (int i = 0; i < n; ++i) {
    sgdlm_ctx *c = cs[i];
    const LuPlan &p = plans[i];
    const double N = (double)c->nsim;
    const bool tuned = !lu_pinned(c) && c->autotune && c->tuned_key == structure_key(c) &&
                       c->tuned.max_front == c->max_front_used && c->tuned.draw_wg_ms > 0.0;
    const double wgs = p.wave_M > 0   ? p.wave_grid
                       : p.wsch_M > 0 ? p.wsch_grid
                       : p.rs_MT > 0  ? p.rs_grid
                       : p.nb         ? p.fast_grid
                                      : p.pivot_grid;
    const double Md = (double)std::max(p.fast_M, 16);
    const double lu = tuned ? N * c->tuned.draw_wg_ms / std::max(1.0, wgs)
                            : N * 1e-9 * Md * Md * Md / 6.0;
    w[i] = lu + 2e-7 * N * (double)c->S.m;
  }
  std::vector<int> ix(n);
  for (int i = 0; i < n; ++i)
    ix[i] = i;
  std::stable_sort(ix.begin(), ix.end(), [&](int a, int b) { return w[a] > w[b]; });
  std::vector<std::vector<int>> parts(P);
  std::vector<double> load(P, 0.0);
  for (int i : ix) {
    const int q = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    parts[q].push_back(i);
    load[q] += w[i];
  }
  for (auto &pt : parts)
    std::sort(pt.begin(), pt.end());
  return parts;
}

int build_vb_batch(sgdlm_ctx *const *cs, int n, const std::vector<VbDev> &Ds,
                   const std::vector<LuPlan> &plans, std::vector<uintptr_t> sig, VbBatch **out) {
  sgdlm_ctx *c0 = cs[0];
  VbBatch *b = new VbBatch();
  b->sig = std::move(sig);
  b->members = split_batch(cs, n, plans);
  const bool multi = b->members.size() > 1;
  for (const auto &mem : b->members) {
    std::vector<sgdlm_ctx *> pc;
    std::vector<VbDev> pd;
    std::vector<LuPlan> pp;
    for (int i : mem) {
      pc.push_back(cs[i]);
      pd.push_back(Ds[i]);
      pp.push_back(plans[i]);
    }
    VbGroup *g = nullptr;
    if (build_vb_group(pc.data(), (int)pc.size(), pd, pp, !multi, &g)) {
      drop_vb_batch(b);
      return SGDLM_ERR_ARG;
    }
    b->parts.push_back(g);
  }
  if (multi) {
    for (size_t q = 0; q < b->parts.size(); ++q) {
      hipStream_t st = nullptr;
      if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
        (void)hipGetLastError();
        drop_vb_batch(b);
        return fail(c0, SGDLM_ERR_DEVICE, "batch stream creation failed");
      }
      b->streams.push_back(st);
    }
    for (size_t q = 0; q < b->parts.size() + 1; ++q) {
      hipEvent_t e = nullptr;
      if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
        (void)hipGetLastError();
        drop_vb_batch(b);
        return fail(c0, SGDLM_ERR_DEVICE, "batch event creation failed");
      }
      b->ev.push_back(e);
    }
  }
  *out = b;
  return SGDLM_OK;
}

/* One VB step of every context in cs as one launch per stage (per part)
 * on the first context's stream.  SGDLM_ERR_ARG before anything ran: the
 * list does not fit the group path (different devices, a pending LU
 * autotuning, draw bucket or grid limits); the caller then runs the
 * contexts one by one.  Per-context statuses go to rcs. */
int run_vb_group(sgdlm_ctx *const *cs, int n, std::vector<int> &rcs) {
  sgdlm_ctx *c0 = cs[0];
  for (int i = 0; i < n; ++i)
    if (cs[i]->device != c0->device || tuning_pending(cs[i]) || cs[i]->S.max_p > 32)
      return SGDLM_ERR_ARG;
  std::vector<VbDev> Ds(n);
  std::vector<LuPlan> plans(n);
  std::vector<uintptr_t> sig;
  for (int i = 0; i < n; ++i) {
    sgdlm_ctx *c = cs[i];
    if (group_member(c, &Ds[i], &plans[i]) || check_epochs(c, 1))
      return SGDLM_ERR_ARG; /* the one-by-one path reports it per context */
    std::vector<uintptr_t> s1 = vb_signature(c, (long)c->nsim, Ds[i], plans[i]);
    sig.push_back((uintptr_t)c->id);
    sig.insert(sig.end(), s1.begin(), s1.end());
    sig.push_back((uintptr_t)plans[i].rs_lds);
  }
  VbBatch *b = c0->grp;
  if (!b || b->sig != sig) {
    drop_vb_batch(c0->grp);
    c0->grp = nullptr;
    if (build_vb_batch(cs, n, Ds, plans, sig, &b))
      return SGDLM_ERR_ARG;
    c0->grp = b;
  }
  for (size_t q = 0; q < b->parts.size(); ++q) {
    VbGroup *g = b->parts[q];
    for (size_t k = 0; k < b->members[q].size(); ++k)
      g->h_epochs[k] = ++cs[b->members[q][k]]->epoch;
    HIPCHK(c0, hipMemcpyAsync(g->d_epochs, g->h_epochs, sizeof(uint32_t) * g->n,
                              hipMemcpyHostToDevice, c0->stream));
  }
  /* the step runs as one captured graph (replayed while the tables stay),
   * so that its parts' streams run concurrently; eagerly only for stage
   * timing */
  int rc;
  if (!c0->timing) {
    if (!b->exec) {
      HIPCHK(c0, hipStreamBeginCapture(c0->stream, hipStreamCaptureModeRelaxed));
      const int erc = enqueue_vb_batch(c0, b, false);
      hipGraph_t gr = nullptr;
      const hipError_t e = hipStreamEndCapture(c0->stream, &gr);
      if (erc || e != hipSuccess) {
        if (gr)
          (void)hipGraphDestroy(gr);
        (void)hipGetLastError();
        return erc ? erc : fail(c0, SGDLM_ERR_DEVICE, "batch graph capture failed: %s",
                                hipGetErrorString(e));
      }
      hipGraphExec_t x = nullptr;
      const hipError_t ei = hipGraphInstantiate(&x, gr, nullptr, nullptr, 0);
      if (ei != hipSuccess) {
        (void)hipGraphDestroy(gr);
        (void)hipGetLastError();
        return fail(c0, SGDLM_ERR_DEVICE, "batch graph instantiation failed: %s",
                    hipGetErrorString(ei));
      }
      b->graph = gr;
      b->exec = x;
      ++c0->graph_captures;
    }
    HIPCHK(c0, hipGraphLaunch(b->exec, c0->stream));
    ++c0->graph_replays;
  } else if ((rc = enqueue_vb_batch(c0, b, c0->timing))) {
    return rc;
  }
  if ((rc = sync_and_collect(c0)))
    return rc;
  for (size_t q = 0; q < b->parts.size(); ++q) {
    for (size_t k = 0; k < b->members[q].size(); ++k) {
      sgdlm_ctx *c = cs[b->members[q][k]];
      const double *hs = b->parts[q]->h_stats + 5 * k;
      std::memcpy(c->h_stats, hs, 4 * sizeof(double));
      const int fb = (int)hs[4];
      std::memcpy(c->h_stats + 4, &fb, sizeof(int));
      c->lu_flags_valid = true;
      c->vb_valid = true;
    }
  }
  for (int i = 0; i < n; ++i)
    rcs[i] = vb_status(cs[i]);
  ++c0->batch_group_calls;
  return SGDLM_OK;
}

} // namespace

/* ======================================================================== */
extern "C" {

int sgdlm_create(sgdlm_ctx **out, size_t no_gpus) {
  if (!out)
    return SGDLM_ERR_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    (void)hipGetLastError();
    return SGDLM_ERR_DEVICE;
  }
  (void)no_gpus; /* one GPU per context */
  sgdlm_ctx *c = new sgdlm_ctx();
  if (hipGetDevice(&c->device) != hipSuccess) {
    delete c;
    return SGDLM_ERR_DEVICE;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, c->device) != hipSuccess) {
    delete c;
    return SGDLM_ERR_DEVICE;
  }
  c->n_cu = prop.multiProcessorCount;
  {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device) == hipSuccess &&
        v > 0)
      c->lds_max = (size_t)v;
    if (prop.maxSharedMemoryPerMultiProcessor > c->lds_max)
      c->lds_max = prop.maxSharedMemoryPerMultiProcessor;
  }
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return SGDLM_ERR_DEVICE;
  }
  if (scratch_ptr(c->device, &c->scratch) != hipSuccess)
    c->scratch = nullptr; /* the allocation-free entry point then fails */
  {
    /* SGDLM_LU_AUTOTUNE=0: new contexts start with the autotuner off */
    const char *e = std::getenv("SGDLM_LU_AUTOTUNE");
    if (e && e[0] == '0')
      c->autotune = false;
  }
  if (hipHostMalloc((void **)&c->h_stats, 8 * sizeof(double), hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    (void)hipStreamDestroy(c->stream);
    delete c;
    return SGDLM_ERR_ALLOC;
  }
  c->id = ++g_ctx_ids;
  *out = c;
  return SGDLM_OK;
}

int sgdlm_destroy(sgdlm_ctx *c) {
  if (!c)
    return SGDLM_OK;
  (void)hipStreamSynchronize(c->stream);
  drop_graph(c);
  drop_vb_batch(c->grp);
  c->grp = nullptr;
  free_pool(c->sim);
  free_pool(c->mem);
  for (auto &pe : c->pending) {
    (void)hipEventDestroy(pe.a);
    (void)hipEventDestroy(pe.b);
  }
  (void)hipStreamDestroy(c->stream);
  (void)hipHostFree(c->h_stats);
  delete c;
  return SGDLM_OK;
}

const char *sgdlm_last_error(const sgdlm_ctx *c) { return c ? c->err.c_str() : "null context"; }

int sgdlm_get_no_series(const sgdlm_ctx *c, size_t *m) {
  if (!c || !m)
    return SGDLM_ERR_ARG;
  *m = c->m;
  return SGDLM_OK;
}
int sgdlm_get_max_p(const sgdlm_ctx *c, size_t *mp) {
  if (!c || !mp)
    return SGDLM_ERR_ARG;
  *mp = c->max_p;
  return SGDLM_OK;
}
int sgdlm_get_evolution_matrix_configuration(const sgdlm_ctx *c, int *u) {
  if (!c || !u)
    return SGDLM_ERR_ARG;
  *u = c->use_state_evolution_matrix ? 1 : 0;
  return SGDLM_OK;
}

int sgdlm_clear_memory(sgdlm_ctx *c) {
  if (!c)
    return SGDLM_ERR_ARG;
  (void)hipStreamSynchronize(c->stream);
  drop_graph(c);
  free_pool(c->sim);
  free_pool(c->mem);
  ++c->gen;
  c->memory_initialized = c->sim_memory_initialized = c->evo_memory_initialized = false;
  c->parents_set = false;
  c->vb_valid = false;
  c->nsim = c->nsim_batch = 0;
  c->S = Structure();
  return SGDLM_OK;
}

int sgdlm_init_memory(sgdlm_ctx *c, size_t m, size_t max_p) {
  if (!c)
    return SGDLM_ERR_ARG;
  if (m < 1 || max_p < 1)
    return fail(c, SGDLM_ERR_ARG, "m and max_p must be positive");
  if (max_p > 64)
    return fail(c, SGDLM_ERR_ARG, "max_p=%zu exceeds the supported 64", max_p);
  if ((unsigned long long)m * m >= 0x80000000ull)
    return fail(c, SGDLM_ERR_ARG, "m=%zu too large (m*m must stay below 2^31)", m);
  sgdlm_clear_memory(c);
  c->m = m;
  c->max_p = max_p;
  const size_t mp2 = max_p * max_p;
  int rc;
  const char *mats[] = {"delta", "C", "W", "fC", "fW", "U"};
  const char *vecs_mp[] = {"m", "fm", "F", "x"};
  const char *vecs_m[] = {"beta", "n", "s", "fn", "fs", "y"};
  for (const char *n : mats)
    if ((rc = dalloc(c, c->mem, n, m * mp2 * 8)))
      return rc;
  for (const char *n : vecs_mp)
    if ((rc = dalloc(c, c->mem, n, m * max_p * 8)))
      return rc;
  for (const char *n : vecs_m)
    if ((rc = dalloc(c, c->mem, n, m * 8)))
      return rc;
  if ((rc = dalloc(c, c->mem, "p", m * 4)) ||
      (rc = dalloc(c, c->mem, "sp", m * max_p * 4)))
    return rc;
  /* zero everything so getters before setters return zeros */
  for (auto &kv : c->mem)
    HIPCHK(c, hipMemsetAsync(kv.second.ptr, 0, kv.second.bytes, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->h_p.assign(m, 0);
  c->h_sp.assign(m * max_p, 0x80000000u);
  c->memory_initialized = true;
  if (c->use_state_evolution_matrix) {
    c->use_state_evolution_matrix = false;
    return sgdlm_manage_evo_memory(c, 1);
  }
  return SGDLM_OK;
}

int sgdlm_manage_evo_memory(sgdlm_ctx *c, int use) {
  if (!c)
    return SGDLM_ERR_ARG;
  if (use) {
    if (c->memory_initialized && !c->evo_memory_initialized) {
      int rc = dalloc(c, c->mem, "G", c->m * c->max_p * c->max_p * 8);
      if (rc)
        return rc;
      /* on the context's stream (non-blocking: a null-stream memset could
       * land after a later G upload on it) */
      HIPCHK(c, hipMemsetAsync(c->mem["G"].ptr, 0, c->mem["G"].bytes, c->stream));
      HIPCHK(c, hipStreamSynchronize(c->stream));
    }
  } else {
    auto it = c->mem.find("G");
    if (it != c->mem.end()) {
      (void)hipFree(it->second.ptr);
      c->mem.erase(it);
    }
  }
  c->use_state_evolution_matrix = use != 0;
  c->evo_memory_initialized = use != 0;
  return SGDLM_OK;
}

int sgdlm_get_nsim(const sgdlm_ctx *c, size_t *n) {
  if (!c || !n)
    return SGDLM_ERR_ARG;
  *n = c->nsim;
  return SGDLM_OK;
}
int sgdlm_get_nsim_batch(const sgdlm_ctx *c, size_t *n) {
  if (!c || !n)
    return SGDLM_ERR_ARG;
  *n = c->nsim_batch;
  return SGDLM_OK;
}

int sgdlm_init_sim_memory(sgdlm_ctx *c, size_t nsim, size_t nsim_batch) {
  if (!c)
    return SGDLM_ERR_ARG;
  int rc = check_init(c);
  if (rc)
    return rc;
  if (nsim < 1)
    return fail(c, SGDLM_ERR_ARG, "nsim must be positive");
  if (nsim >= 0x80000000ull)
    return fail(c, SGDLM_ERR_ARG, "nsim=%zu too large", nsim);
  (void)hipStreamSynchronize(c->stream);
  free_pool(c->sim);
  c->nsim = nsim;
  c->nsim_batch = nsim_batch > nsim ? nsim : nsim_batch;
  c->vb_valid = false;
  c->lu_flags_valid = false;
  c->sim_memory_initialized = true;
  if (c->parents_set)
    return ensure_sim_buffers(c);
  return SGDLM_OK;
}

int sgdlm_clear_sim_memory(sgdlm_ctx *c) {
  if (!c)
    return SGDLM_ERR_ARG;
  int rc = check_init(c);
  if (rc)
    return rc;
  (void)hipStreamSynchronize(c->stream);
  drop_graph(c);
  free_pool(c->sim);
  ++c->gen;
  c->sim_memory_initialized = false;
  c->vb_valid = false;
  c->lu_flags_valid = false;
  c->nsim = c->nsim_batch = 0;
  return SGDLM_OK;
}

int sgdlm_init_forecast_priors(sgdlm_ctx *c) {
  if (!c)
    return SGDLM_ERR_ARG;
  int rc = check_init(c);
  if (rc)
    return rc;
  const size_t m = c->m, mp = c->max_p;
  HIPCHK(c, hipMemcpyAsync(M_<double>(c, "fn"), M_<double>(c, "n"), m * 8, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(M_<double>(c, "fs"), M_<double>(c, "s"), m * 8, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(M_<double>(c, "fm"), M_<double>(c, "m"), m * mp * 8, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(M_<double>(c, "fC"), M_<double>(c, "C"), m * mp * mp * 8, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(M_<double>(c, "fW"), M_<double>(c, "W"), m * mp * mp * 8, hipMemcpyDeviceToDevice, c->stream));
  return sync_and_collect(c);
}

int sgdlm_get_is_prior(const sgdlm_ctx *c, int *v) {
  if (!c || !v)
    return SGDLM_ERR_ARG;
  *v = c->is_prior ? 1 : 0;
  return SGDLM_OK;
}
int sgdlm_set_is_prior(sgdlm_ctx *c, int v) {
  if (!c)
    return SGDLM_ERR_ARG;
  c->is_prior = v != 0;
  return SGDLM_OK;
}

int sgdlm_get_parameters(const sgdlm_ctx *cc, double *m_t, double *C_t,
                         double *n_t, double *s_t) {
  sgdlm_ctx *c = const_cast<sgdlm_ctx *>(cc);
  if (!c)
    return SGDLM_ERR_ARG;
  int rc = check_init(c);
  if (rc)
    return rc;
  const size_t m = c->m, mp = c->max_p;
  if (m_t && (rc = d2h(c, m_t, M_<double>(c, "m"), m * mp * 8)))
    return rc;
  if (C_t && (rc = d2h(c, C_t, M_<double>(c, "C"), m * mp * mp * 8)))
    return rc;
  if (n_t && (rc = d2h(c, n_t, M_<double>(c, "n"), m * 8)))
    return rc;
  if (s_t && (rc = d2h(c, s_t, M_<double>(c, "s"), m * 8)))
    return rc;
  return sync_and_collect(c);
}

int sgdlm_set_parameters(sgdlm_ctx *c, const double *m_t, const double *C_t,
                         const double *n_t, const double *s_t) {
  if (!c)
    return SGDLM_ERR_ARG;
  int rc = check_init(c);
  if (rc)
    return rc;
  const size_t m = c->m, mp = c->max_p;
  if (m_t && (rc = h2d(c, M_<double>(c, "m"), m_t, m * mp * 8)))
    return rc;
  if (C_t && (rc = h2d(c, M_<double>(c, "C"), C_t, m * mp * mp * 8)))
    return rc;
  if (n_t && (rc = h2d(c, M_<double>(c, "n"), n_t, m * 8)))
    return rc;
  if (s_t && (rc = h2d(c, M_<double>(c, "s"), s_t, m * 8)))
    return rc;
  return sync_and_collect(c);
}

int sgdlm_get_discount_factors(const sgdlm_ctx *cc, double *beta, double *delta) {
  sgdlm_ctx *c = const_cast<sgdlm_ctx *>(cc);
  if (!c)
    return SGDLM_ERR_ARG;
  int rc = check_init(c);
  if (rc)
    return rc;
  if (beta && (rc = d2h(c, beta, M_<double>(c, "beta"), c->m * 8)))
    return rc;
  if (delta && (rc = d2h(c, delta, M_<double>(c, "delta"), c->m * c->max_p * c->max_p * 8)))
    return rc;
  return sync_and_collect(c);
}

int sgdlm_set_discount_factors(sgdlm_ctx *c, const double *beta, const double *delta) {
  if (!c)
    return SGDLM_ERR_ARG;
  int rc = check_init(c);
  if (rc)
    return rc;
  if (beta && (rc = h2d(c, M_<double>(c, "beta"), beta, c->m * 8)))
    return rc;
  if (delta && (rc = h2d(c, M_<double>(c, "delta"), delta, c->m * c->max_p * c->max_p * 8)))
    return rc;
  return sync_and_collect(c);
}

int sgdlm_get_evolution_matrix(const sgdlm_ctx *cc, double *G) {
  sgdlm_ctx *c = const_cast<sgdlm_ctx *>(cc);
  if (!c)
    return SGDLM_ERR_ARG;
  int rc;
  if ((rc = check_init(c)) || (rc = check_evo(c)))
    return rc;
  if (G && (rc = d2h(c, G, M_<double>(c, "G"), c->m * c->max_p * c->max_p * 8)))
    return rc;
  return sync_and_collect(c);
}

int sgdlm_set_evolution_matrix(sgdlm_ctx *c, const double *G) {
  if (!c)
    return SGDLM_ERR_ARG;
  int rc;
  if ((rc = check_init(c)) || (rc = check_evo(c)))
    return rc;
  if (G && (rc = h2d(c, M_<double>(c, "G"), G, c->m * c->max_p * c->max_p * 8)))
    return rc;
  return sync_and_collect(c);
}

int sgdlm_get_parental_sets(const sgdlm_ctx *c, unsigned int *p, unsigned int *sp) {
  if (!c)
    return SGDLM_ERR_ARG;
  if (!c->memory_initialized)
    return SGDLM_ERR_STATE;
  if (p)
    memcpy(p, c->h_p.data(), c->m * 4);
  if (sp)
    memcpy(sp, c->h_sp.data(), c->m * c->max_p * 4);
  return SGDLM_OK;
}

int sgdlm_set_parental_sets(sgdlm_ctx *c, const unsigned int *p, const unsigned int *sp) {
  if (!c)
    return SGDLM_ERR_ARG;
  int rc = check_init(c);
  if (rc)
    return rc;
  const size_t m = c->m, mp = c->max_p;
  std::vector<unsigned> np = c->h_p, nsp = c->h_sp;
  if (p)
    np.assign(p, p + m);
  if (sp)
    nsp.assign(sp, sp + m * mp);
  const unsigned long long mm = (unsigned long long)m * m;
  std::vector<char> used(mm, 0);
  for (size_t j = 0; j < m; ++j) {
    if (np[j] > mp)
      return fail(c, SGDLM_ERR_ARG, "p[%zu]=%u exceeds max_p=%zu", j, np[j], mp);
    for (size_t k = 0; k < mp; ++k) {
      const unsigned v = nsp[j * mp + k];
      if (v >= mm)
        continue;
      if (k >= np[j])
        return fail(c, SGDLM_ERR_ARG, "series %zu: simultaneous parent position at k=%zu >= p=%u", j, k, np[j]);
      if (v % m == v / m)
        return fail(c, SGDLM_ERR_ARG, "series %zu: position %u is on the diagonal of Gamma", j, v);
      if (used[v])
        return fail(c, SGDLM_ERR_ARG, "position %u of Gamma is used twice", v);
      used[v] = 1;
    }
  }
  c->h_p = np;
  c->h_sp = nsp;
  if ((rc = h2d(c, M_<unsigned>(c, "p"), np.data(), m * 4)) ||
      (rc = h2d(c, M_<unsigned>(c, "sp"), nsp.data(), m * mp * 4)))
    return rc;
  if ((rc = build_structure(c)))
    return rc;
  c->parents_set = true;
  c->vb_valid = false;
  if (c->sim_memory_initialized)
    return ensure_sim_buffers(c);
  return SGDLM_OK;
}

int sgdlm_compute_prior(sgdlm_ctx *c, const double *G, int evolve_forecast_variables) {
  if (!c)
    return SGDLM_ERR_ARG;
  int rc;
  if ((rc = check_init(c)) || (rc = check_parents(c)))
    return rc;
  if (G && (rc = check_evo(c)))
    return rc;
  if (!evolve_forecast_variables && (rc = check_prior(c, false)))
    return rc;
  if (G && (rc = h2d(c, M_<double>(c, "G"), G, c->m * c->max_p * c->max_p * 8)))
    return rc;
  const bool f = evolve_forecast_variables != 0;
  {
    StageTimer t(c, "filter");
    launch_prior(c->S, M_<double>(c, f ? "fm" : "m"), M_<double>(c, f ? "fC" : "C"),
                 M_<double>(c, f ? "fn" : "n"), M_<double>(c, f ? "fW" : "W"),
                 M_<double>(c, "beta"), M_<double>(c, "delta"),
                 G ? M_<double>(c, "G") : nullptr, c->stream);
  }
  CHECK_LAUNCH(c);
  if ((rc = sync_and_collect(c)))
    return rc;
  if (!f)
    c->is_prior = true;
  return SGDLM_OK;
}

int sgdlm_compute_posterior(sgdlm_ctx *c, const double *y, const double *F) {
  if (!c || !y || !F)
    return c ? fail(c, SGDLM_ERR_ARG, "null y or F") : SGDLM_ERR_ARG;
  int rc;
  if ((rc = check_init(c)) || (rc = check_parents(c)) || (rc = check_prior(c, true)))
    return rc;
  if ((rc = h2d(c, M_<double>(c, "y"), y, c->m * 8)) ||
      (rc = h2d(c, M_<double>(c, "F"), F, c->m * c->max_p * 8)))
    return rc;
  {
    StageTimer t(c, "filter");
    launch_posterior(c->S, M_<double>(c, "m"), M_<double>(c, "C"), M_<double>(c, "n"),
                     M_<double>(c, "s"), M_<double>(c, "y"), M_<double>(c, "F"), c->stream);
  }
  CHECK_LAUNCH(c);
  if ((rc = sync_and_collect(c)))
    return rc;
  c->is_prior = false;
  return SGDLM_OK;
}

int sgdlm_compute_vb_posterior(sgdlm_ctx *c, double *mean_m, double *mean_C,
                               double *mean_n, double *mean_s, double *w,
                               double *sum_det) {
  if (!c)
    return SGDLM_ERR_ARG;
  int rc;
  if ((rc = check_init(c)) || (rc = check_parents(c)) || (rc = check_sim(c)) ||
      (rc = check_prior(c, false)))
    return rc;
  if ((rc = run_vb(c)))
    return rc;
  const bool any = mean_m || mean_C || mean_n || mean_s || w || sum_det;
  if (any) {
    rc = sgdlm_get_vb_result(c, mean_m, mean_C, mean_n, mean_s, w, nullptr, nullptr, nullptr);
    if (rc)
      return rc;
  }
  if ((rc = sync_and_collect(c)))
    return rc;
  if (sum_det)
    *sum_det = std::exp(c->h_stats[0]) * c->h_stats[1];
  return vb_status(c);
}

int sgdlm_compute_vb_posterior_noalloc(sgdlm_ctx *c, size_t nsim, double *mean_m,
                                       double *mean_C, double *mean_n, double *mean_s,
                                       double *w, double *sum_det, size_t *scratch_bytes) {
  if (!c)
    return SGDLM_ERR_ARG;
  int rc;
  if ((rc = check_init(c)) || (rc = check_parents(c)) || (rc = check_prior(c, false)))
    return rc;
  if (nsim < 1 || nsim >= 0x80000000ull)
    return fail(c, SGDLM_ERR_ARG, "nsim=%zu out of range", nsim);
  if (!c->scratch)
    return fail(c, SGDLM_ERR_DEVICE, "the scratch array of this device has no address");
  const size_t N = nsim, m = c->m, mp = c->max_p, P = std::max<size_t>((size_t)c->S.P, 1);
  const size_t chunks = (size_t)vb_chunks((long)N);
  /* carve the buffers (256-byte aligned) in a fixed order */
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t o = off;
    off += (bytes + 255) & ~(size_t)255;
    return o;
  };
  const size_t o_lam = take(N * m * 8), o_theta = take(N * P * 8), o_ld = take(N * 8),
               o_w = take(N * 8), o_stats = take(8 * 8), o_flags = take((N + 1) * 4),
               o_p1 = take(chunks * m * 8), o_p2 = take(chunks * m * (mp + 1) * (mp + 1) * 8),
               o_ml = take(m * 8), o_mll = take(m * 8), o_mm = take(m * mp * 8),
               o_mC = take(m * mp * mp * 8), o_mn = take(m * 8), o_ms = take(m * 8),
               o_mQ = take(m * 8), o_it = take(m * 4);
  LuPlan plan;
  rc = vb_lu_plan(c, &plan);
  if (rc)
    return fail(c, rc, "no LU configuration fits m=%zu (M=%d) in LDS", c->m, c->S.M);
  const size_t lu_room = off < kScratchBytes ? kScratchBytes - off : 0;
  const bool fits = lu_room > 0 && lu_plan_cap(c->S, lu_room, &plan) == SGDLM_OK;
  const size_t o_work = take(plan.work_bytes);
  if (scratch_bytes)
    *scratch_bytes = off;
  if (!fits || off > kScratchBytes)
    return fail(c, SGDLM_ERR_ARG,
                "nsim=%zu at m=%zu, max_p=%zu needs more than the %zu bytes of scratch of "
                "the allocation-free entry point (draws and moments alone take %zu)",
                nsim, m, mp, kScratchBytes, o_work);
  char *base = static_cast<char *>(c->scratch);
  VbDev D;
  D.lam = reinterpret_cast<double *>(base + o_lam);
  D.theta = reinterpret_cast<double *>(base + o_theta);
  D.logdet = reinterpret_cast<double *>(base + o_ld);
  D.lu_flags = reinterpret_cast<int *>(base + o_flags);
  D.lu_work = base + o_work;
  VbBuffers &B = D.B;
  B.w = reinterpret_cast<double *>(base + o_w);
  B.stats = reinterpret_cast<double *>(base + o_stats);
  B.part1 = reinterpret_cast<double *>(base + o_p1);
  B.part2 = reinterpret_cast<double *>(base + o_p2);
  B.mean_lambda = reinterpret_cast<double *>(base + o_ml);
  B.mean_log_lambda = reinterpret_cast<double *>(base + o_mll);
  B.mean_m = reinterpret_cast<double *>(base + o_mm);
  B.mean_C = reinterpret_cast<double *>(base + o_mC);
  B.mean_n = reinterpret_cast<double *>(base + o_mn);
  B.mean_s = reinterpret_cast<double *>(base + o_ms);
  B.mean_Q = reinterpret_cast<double *>(base + o_mQ);
  B.newton_iter = reinterpret_cast<int *>(base + o_it);
  B.chunks = (int)chunks;
  if ((rc = run_vb_stages(c, (long)N, D, plan, false, false)))
    return rc;
  if ((mean_m && (rc = d2h(c, mean_m, B.mean_m, m * mp * 8))) ||
      (mean_C && (rc = d2h(c, mean_C, B.mean_C, m * mp * mp * 8))) ||
      (mean_n && (rc = d2h(c, mean_n, B.mean_n, m * 8))) ||
      (mean_s && (rc = d2h(c, mean_s, B.mean_s, m * 8))) ||
      (w && (rc = d2h(c, w, B.w, N * 8))))
    return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (sum_det)
    *sum_det = std::exp(c->h_stats[0]) * c->h_stats[1];
  return vb_status(c);
}

int sgdlm_batch_compute_vb_posterior(sgdlm_ctx *const *ctxs, size_t count, int *status) {
  if (!ctxs && count)
    return SGDLM_ERR_ARG;
  int first = SGDLM_OK;
  std::vector<int> rcs(count, SGDLM_OK);
  /* one launch per stage over every model, when they all fit it */
  bool ready = count > 1 && count < (1u << 20);
  for (size_t i = 0; i < count && ready; ++i) {
    sgdlm_ctx *c = ctxs[i];
    ready = c && !check_init(c) && !check_parents(c) && !check_sim(c) && !check_prior(c, false) &&
            !c->batch_no_group;
  }
  if (ready) {
    const int rc = run_vb_group(ctxs, (int)count, rcs);
    if (rc == SGDLM_OK) {
      for (size_t i = 0; i < count; ++i) {
        if (status)
          status[i] = rcs[i];
        if (rcs[i] && !first)
          first = rcs[i];
      }
      return first;
    }
    if (rc != SGDLM_ERR_ARG) {
      /* the group step failed after it started: every model reports it */
      for (size_t i = 0; i < count; ++i) {
        if (ctxs[i] != ctxs[0])
          ctxs[i]->err = ctxs[0]->err;
        if (status)
          status[i] = rc;
      }
      return rc;
    }
    std::fill(rcs.begin(), rcs.end(), SGDLM_OK);
  }
  if (count && ctxs[0])
    ++ctxs[0]->batch_stream_calls;
  /* enqueue every context's stages on its own stream, then wait once each */
  for (size_t i = 0; i < count; ++i) {
    sgdlm_ctx *c = ctxs[i];
    int rc = c ? SGDLM_OK : SGDLM_ERR_ARG;
    if (!rc) {
      if (!(rc = check_init(c)) && !(rc = check_parents(c)) && !(rc = check_sim(c)) &&
          !(rc = check_prior(c, false)))
        rc = run_vb(c);
    }
    rcs[i] = rc;
  }
  for (size_t i = 0; i < count; ++i) {
    if (!rcs[i])
      rcs[i] = sync_and_collect(ctxs[i]);
    if (!rcs[i])
      rcs[i] = vb_status(ctxs[i]);
    if (status)
      status[i] = rcs[i];
    if (rcs[i] && !first)
      first = rcs[i];
  }
  return first;
}

int sgdlm_get_vb_result(const sgdlm_ctx *cc, double *mean_m, double *mean_C,
                        double *mean_n, double *mean_s, double *w,
                        double *log_w, double *ess, double *log_sum_det) {
  sgdlm_ctx *c = const_cast<sgdlm_ctx *>(cc);
  if (!c)
    return SGDLM_ERR_ARG;
  if (!c->vb_valid)
    return fail(c, SGDLM_ERR_STATE, "no VB result on the device");
  const size_t m = c->m, mp = c->max_p, N = c->nsim;
  int rc;
  if (mean_m && (rc = d2h(c, mean_m, SIM<double>(c, "mean_m"), m * mp * 8)))
    return rc;
  if (mean_C && (rc = d2h(c, mean_C, SIM<double>(c, "mean_C"), m * mp * mp * 8)))
    return rc;
  if (mean_n && (rc = d2h(c, mean_n, SIM<double>(c, "mean_n"), m * 8)))
    return rc;
  if (mean_s && (rc = d2h(c, mean_s, SIM<double>(c, "mean_s"), m * 8)))
    return rc;
  if (w && (rc = d2h(c, w, SIM<double>(c, "w"), N * 8)))
    return rc;
  if (log_w && (rc = d2h(c, log_w, SIM<double>(c, "logdet"), N * 8)))
    return rc;
  double stats[8];
  if ((rc = d2h(c, stats, SIM<double>(c, "stats"), 4 * 8)) || (rc = sync_and_collect(c)))
    return rc;
  if (log_w)
    for (size_t i = 0; i < N; ++i)
      if (!std::isfinite(log_w[i]))
        log_w[i] = -INFINITY;
  if (ess)
    *ess = 1.0 / stats[2];
  if (log_sum_det)
    *log_sum_det = stats[0] + std::log(stats[1]);
  return SGDLM_OK;
}

int sgdlm_evolve_forecast_samples(sgdlm_ctx *c, const double *G) {
  if (!c)
    return SGDLM_ERR_ARG;
  int rc;
  if ((rc = check_init(c)) || (rc = check_parents(c)) || (rc = check_sim(c)))
    return rc;
  if (G && (rc = check_evo(c)))
    return rc;
  if ((rc = ensure_sim_buffers(c)))
    return rc;
  if ((rc = check_epochs(c, 1)))
    return rc;
  if (G && (rc = h2d(c, M_<double>(c, "G"), G, c->m * c->max_p * c->max_p * 8)))
    return rc;
  const uint32_t ep = ++c->epoch;
  {
    StageTimer t(c, "evolve");
    launch_cholesky(c->S, M_<double>(c, "fW"), M_<double>(c, "U"), c->stream);
    launch_evolve(c->S, (long)c->nsim, M_<double>(c, "U"), M_<double>(c, "fn"),
                  M_<double>(c, "fs"), M_<double>(c, "beta"),
                  G ? M_<double>(c, "G") : nullptr, M_<double>(c, "fm"), c->seed, ep,
                  SIM<double>(c, "lam"), SIM<double>(c, "theta"), c->stream);
  }
  CHECK_LAUNCH(c);
  c->vb_valid = false;
  return sync_and_collect(c);
}

int sgdlm_compute_forecast(sgdlm_ctx *c, double *y, const double *x,
                           double *lambdas, double *thetas, int use_existing) {
  if (!c)
    return SGDLM_ERR_ARG;
  int rc;
  if ((rc = check_init(c)) || (rc = check_parents(c)) || (rc = check_sim(c)) ||
      (rc = check_prior(c, true)))
    return rc;
  if (!x)
    return fail(c, SGDLM_ERR_ARG, "null x_tp1");
  if ((rc = check_epochs(c, use_existing ? 1 : 2)))
    return rc;
  if ((rc = ensure_sim_buffers(c)))
    return rc;
  const size_t m = c->m, mp = c->max_p, N = c->nsim;
  if ((rc = dalloc(c, c->sim, "rhs", N * m * 8)) || (rc = dalloc(c, c->sim, "yf", N * m * 8)))
    return rc;
  if ((rc = h2d(c, M_<double>(c, "x"), x, m * mp * 8)))
    return rc;
  if (!use_existing) {
    const uint32_t ep = ++c->epoch;
    StageTimer t(c, "draw");
    launch_cholesky(c->S, M_<double>(c, "fC"), M_<double>(c, "U"), c->stream);
    launch_draw(c->S, (long)N, M_<double>(c, "fm"), M_<double>(c, "U"),
                M_<double>(c, "fn"), M_<double>(c, "fs"), c->seed, ep,
                SIM<double>(c, "lam"), SIM<double>(c, "theta"), c->stream);
  }
  CHECK_LAUNCH(c);
  const uint32_t ep_nu = ++c->epoch;
  launch_forecast_rhs(c->S, (long)N, M_<double>(c, "x"), M_<double>(c, "fm"),
                      SIM<double>(c, "theta"), SIM<double>(c, "lam"), c->seed, ep_nu,
                      SIM<double>(c, "rhs"), c->stream);
  CHECK_LAUNCH(c);
  const int prec = (c->prec == SGDLM_PREC_FP64) ? SGDLM_PREC_FP64 : SGDLM_PREC_FP32;
  LuPlan plan;
  if ((rc = lu_workspace(c, prec, 1, &plan)))
    return rc;
  {
    StageTimer t(c, "forecast_lu");
    rc = launch_lu_solve(c->S, plan, (long)N, SIM<double>(c, "theta"), c->sim["lu_work"].ptr,
                         SIM<double>(c, "rhs"), SIM<double>(c, "yf"),
                         prec == SGDLM_PREC_FP64 ? 0 : 1, SIM<int>(c, "lu_flags"), c->stream);
  }
  if (rc)
    return fail(c, rc, "forecast LU launch failed");
  c->lu_flags_valid = true;
  c->vb_valid = false;
  if (y && (rc = d2h(c, y, SIM<double>(c, "yf"), N * m * 8)))
    return rc;
  if (lambdas && (rc = d2h(c, lambdas, SIM<double>(c, "lam"), N * m * 8)))
    return rc;
  if (thetas) {
    /* unpack in slices to bound the extra device memory */
    const long slice = std::max<long>(1, (long)((256ull << 20) / (m * mp * 8)));
    if ((rc = dalloc(c, c->sim, "theta_full", (size_t)std::min<long>(slice, (long)N) * m * mp * 8)))
      return rc;
    for (long i0 = 0; i0 < (long)N; i0 += slice) {
      const long n = std::min<long>(slice, (long)N - i0);
      launch_unpack_theta(c->S, n, SIM<double>(c, "theta") + (size_t)i0 * c->S.P,
                          M_<double>(c, "fm"), SIM<double>(c, "theta_full"), c->stream);
      CHECK_LAUNCH(c);
      if ((rc = d2h(c, thetas + (size_t)i0 * m * mp, SIM<double>(c, "theta_full"), n * m * mp * 8)))
        return rc;
      HIPCHK(c, hipStreamSynchronize(c->stream));
    }
  }
  return sync_and_collect(c);
}

int sgdlm_set_precision(sgdlm_ctx *c, int prec) {
  if (!c)
    return SGDLM_ERR_ARG;
  if (prec != SGDLM_PREC_FP64 && prec != SGDLM_PREC_FP32 && prec != SGDLM_PREC_BF16X3)
    return fail(c, SGDLM_ERR_ARG, "unknown precision %d", prec);
  c->prec = prec;
  return SGDLM_OK;
}
int sgdlm_get_precision(const sgdlm_ctx *c, int *prec) {
  if (!c || !prec)
    return SGDLM_ERR_ARG;
  *prec = c->prec;
  return SGDLM_OK;
}
int sgdlm_set_rng_state(sgdlm_ctx *c, uint64_t seed, uint32_t epoch) {
  if (!c)
    return SGDLM_ERR_ARG;
  if (epoch >= (1u << 24))
    return fail(c, SGDLM_ERR_ARG, "epoch must be below 2^24");
  c->seed = seed;
  c->epoch = epoch;
  return SGDLM_OK;
}
int sgdlm_get_rng_state(const sgdlm_ctx *c, uint64_t *seed, uint32_t *epoch) {
  if (!c)
    return SGDLM_ERR_ARG;
  if (seed)
    *seed = c->seed;
  if (epoch)
    *epoch = c->epoch;
  return SGDLM_OK;
}
int sgdlm_set_timing(sgdlm_ctx *c, int enable) {
  if (!c)
    return SGDLM_ERR_ARG;
  c->timing = enable != 0;
  c->times.clear();
  return SGDLM_OK;
}
int sgdlm_get_timing(const sgdlm_ctx *c, const char *stage, double *ms, long *launches) {
  if (!c || !stage)
    return SGDLM_ERR_ARG;
  auto it = c->times.find(stage);
  if (ms)
    *ms = it == c->times.end() ? 0.0 : it->second.first;
  if (launches)
    *launches = it == c->times.end() ? 0 : it->second.second;
  return SGDLM_OK;
}
int sgdlm_get_lu_fallback_count(const sgdlm_ctx *cc, long *count) {
  sgdlm_ctx *c = const_cast<sgdlm_ctx *>(cc);
  if (!c || !count)
    return SGDLM_ERR_ARG;
  if (!c->lu_flags_valid || !c->sim_memory_initialized)
    return fail(c, SGDLM_ERR_STATE, "no VB or forecast LU has run");
  int n = 0;
  int rc;
  if ((rc = d2h(c, &n, SIM<int>(c, "lu_flags"), sizeof(int))) || (rc = sync_and_collect(c)))
    return rc;
  *count = n;
  return SGDLM_OK;
}

int sgdlm_get_lu_config(const sgdlm_ctx *cc, int *grid, size_t *lds, int *nb, int *threads) {
  sgdlm_ctx *c = const_cast<sgdlm_ctx *>(cc);
  if (!c)
    return SGDLM_ERR_ARG;
  if (!c->parents_set)
    return SGDLM_ERR_STATE;
  LuPlan plan;
  int rc = vb_lu_plan(c, &plan);
  if (rc)
    return fail(c, rc, "no LU configuration for threads=%d nb=%d", c->lu_threads, c->lu_nb);
  const bool reg = plan.rs_MT > 0, wave = plan.wave_M > 0, wsch = plan.wsch_M > 0;
  if (grid)
    *grid = wsch   ? plan.wsch_grid
            : wave ? plan.wave_grid
            : reg  ? plan.rs_grid
                   : (plan.nb ? plan.fast_grid : plan.grid);
  if (lds)
    *lds = wsch ? plan.wsch_lds : wave ? 0 : reg ? plan.rs_lds : plan.lds;
  if (nb)
    *nb = (wave || wsch) ? 0 : reg ? 16 : plan.nb;
  if (threads)
    *threads = (wave || wsch) ? 256 : reg ? 64 * plan.rs_nw : (plan.nb ? plan.threads : 0);
  return SGDLM_OK;
}

int sgdlm_get_lu_kernel(const sgdlm_ctx *cc, int *kind, int *dense_dim, int *pivot_dim) {
  sgdlm_ctx *c = const_cast<sgdlm_ctx *>(cc);
  if (!c)
    return SGDLM_ERR_ARG;
  if (!c->parents_set)
    return SGDLM_ERR_STATE;
  LuPlan plan;
  int rc = vb_lu_plan(c, &plan);
  if (rc)
    return fail(c, rc, "no LU configuration fits");
  if (kind)
    *kind = lu_plan_kind(plan);
  if (dense_dim)
    *dense_dim = plan.wsch_M > 0   ? plan.wsch_M
                 : plan.wave_M > 0 ? plan.wave_M
                 : plan.rs_MT > 0  ? 16 * plan.rs_MT
                                   : (plan.nb ? plan.fast_M : c->S.M);
  if (pivot_dim)
    *pivot_dim = plan.wave_M > 0 ? plan.wave_M
                 : plan.pivot_schur_M > 0 ? plan.pivot_schur_M : c->S.M;
  return SGDLM_OK;
}

int sgdlm_set_lu_kernel(sgdlm_ctx *c, int kind) {
  if (!c)
    return SGDLM_ERR_ARG;
  if (kind < SGDLM_LU_AUTO || kind > SGDLM_LU_REGISTER_PIVOT)
    return fail(c, SGDLM_ERR_ARG, "unknown LU kernel kind %d", kind);
  c->lu_kind = kind;
  return SGDLM_OK;
}

int sgdlm_set_graph_mode(sgdlm_ctx *c, int enable) {
  if (!c)
    return SGDLM_ERR_ARG;
  c->graph_mode = enable != 0;
  if (!c->graph_mode) {
    (void)hipStreamSynchronize(c->stream);
    drop_graph(c);
  }
  return SGDLM_OK;
}

int sgdlm_get_graph_stats(const sgdlm_ctx *c, long *captures, long *replays) {
  if (!c)
    return SGDLM_ERR_ARG;
  if (captures)
    *captures = c->graph_captures;
  if (replays)
    *replays = c->graph_replays;
  return SGDLM_OK;
}

int sgdlm_set_batch_mode(sgdlm_ctx *c, int mode) {
  if (!c || (mode != SGDLM_BATCH_GROUP && mode != SGDLM_BATCH_STREAMS))
    return SGDLM_ERR_ARG;
  c->batch_no_group = mode == SGDLM_BATCH_STREAMS;
  return SGDLM_OK;
}

int sgdlm_get_batch_stats(const sgdlm_ctx *c, long *group_calls, long *stream_calls) {
  if (!c)
    return SGDLM_ERR_ARG;
  if (group_calls)
    *group_calls = c->batch_group_calls;
  if (stream_calls)
    *stream_calls = c->batch_stream_calls;
  return SGDLM_OK;
}

int sgdlm_set_lu_autotune(sgdlm_ctx *c, int enable) {
  if (!c)
    return SGDLM_ERR_ARG;
  c->autotune = enable != 0;
  return SGDLM_OK;
}

int sgdlm_get_lu_autotune(const sgdlm_ctx *c, int *kind, int *threads, int *nb, int *max_front,
                          double *tune_ms, int *candidates, int *from_cache) {
  if (!c)
    return SGDLM_ERR_ARG;
  const bool have = c->parents_set && c->tuned_key != 0 && c->tuned_key == structure_key(c);
  LuTuned t;
  if (have)
    t = c->tuned;
  if (kind)
    *kind = have ? t.kind : SGDLM_LU_AUTO;
  if (threads)
    *threads = t.threads;
  if (nb)
    *nb = t.nb;
  if (max_front)
    *max_front = have ? t.max_front : -1;
  if (tune_ms)
    *tune_ms = t.tune_ms;
  if (candidates)
    *candidates = t.candidates;
  if (from_cache)
    *from_cache = t.from_cache ? 1 : 0;
  return SGDLM_OK;
}

int sgdlm_get_lu_solve_config(const sgdlm_ctx *cc, int *grid, size_t *lds, int *nb, int *threads,
                              int *slots_in_workspace) {
  sgdlm_ctx *c = const_cast<sgdlm_ctx *>(cc);
  if (!c)
    return SGDLM_ERR_ARG;
  if (!c->parents_set)
    return SGDLM_ERR_STATE;
  const int prec = (c->prec == SGDLM_PREC_FP64) ? SGDLM_PREC_FP64 : SGDLM_PREC_FP32;
  LuPlan plan;
  int rc = lu_plan(c->S, prec, 1, c->n_cu, c->lds_max, 0, 0, &plan, c->lu_solve_kind);
  if (rc)
    return fail(c, rc, "no LU solve configuration for m=%zu", c->m);
  const bool reg = plan.rs_chunk > 0;
  if (grid)
    *grid = reg ? plan.rs_grid : plan.nb ? plan.fast_grid : plan.pivot_grid;
  if (lds)
    *lds = reg ? plan.rs_lds : plan.nb ? plan.lds : plan.pivot_lds;
  if (nb)
    *nb = reg ? 16 : plan.nb;
  if (threads)
    *threads = reg ? 64 * plan.rs_nw : plan.nb ? plan.threads : 0;
  if (slots_in_workspace)
    *slots_in_workspace = reg || plan.slot_stride != 0;
  return SGDLM_OK;
}

int sgdlm_get_lu_solve_kernel(const sgdlm_ctx *cc, int *kind) {
  sgdlm_ctx *c = const_cast<sgdlm_ctx *>(cc);
  if (!c || !kind)
    return SGDLM_ERR_ARG;
  if (!c->parents_set)
    return SGDLM_ERR_STATE;
  const int prec = (c->prec == SGDLM_PREC_FP64) ? SGDLM_PREC_FP64 : SGDLM_PREC_FP32;
  LuPlan plan;
  int rc = lu_plan(c->S, prec, 1, c->n_cu, c->lds_max, 0, 0, &plan, c->lu_solve_kind);
  if (rc)
    return fail(c, rc, "no LU solve configuration for m=%zu", c->m);
  *kind = plan.rs_chunk > 0 ? SGDLM_LU_REGISTER : plan.nb ? SGDLM_LU_BLOCKED : SGDLM_LU_PIVOTING;
  return SGDLM_OK;
}

int sgdlm_set_lu_solve_kernel(sgdlm_ctx *c, int kind) {
  if (!c)
    return SGDLM_ERR_ARG;
  if (kind != SGDLM_LU_AUTO && kind != SGDLM_LU_BLOCKED && kind != SGDLM_LU_REGISTER &&
      kind != SGDLM_LU_PIVOTING)
    return fail(c, SGDLM_ERR_ARG, "unknown LU solve kernel kind %d", kind);
  c->lu_solve_kind = kind;
  return SGDLM_OK;
}

int sgdlm_set_lu_tuning(sgdlm_ctx *c, int threads, int nb) {
  if (!c)
    return SGDLM_ERR_ARG;
  if ((threads != 0 && threads != 256 && threads != 512) ||
      (nb != 0 && nb != 16 && nb != 32 && nb != 64))
    return fail(c, SGDLM_ERR_ARG, "threads must be 0, 256 or 512 and nb 0, 16, 32 or 64");
  c->lu_threads = threads;
  c->lu_nb = nb;
  return SGDLM_OK;
}

int sgdlm_set_lu_max_front(sgdlm_ctx *c, int max_front) {
  if (!c)
    return SGDLM_ERR_ARG;
  if (max_front < 0)
    return fail(c, SGDLM_ERR_ARG, "max_front must be >= 0");
  c->max_front = max_front;
  c->max_front_fixed = true;
  c->vb_valid = false;
  if (c->parents_set)
    return build_structure(c);
  return SGDLM_OK;
}

int sgdlm_get_lu_elimination(const sgdlm_ctx *c, int *max_front, int *n_prefix, int *n_dense,
                             int *n_waves, long *prefix_fma) {
  if (!c)
    return SGDLM_ERR_ARG;
  if (max_front)
    *max_front = c->max_front_used;
  if (!c->parents_set)
    return SGDLM_ERR_STATE;
  if (n_prefix)
    *n_prefix = c->S.e_n1;
  if (n_dense)
    *n_dense = c->S.e_n2;
  if (n_waves)
    *n_waves = c->S.e_waves;
  if (prefix_fma)
    *prefix_fma = (long)c->S.e_fma;
  return SGDLM_OK;
}

} // extern "C"
