// pbf_engine.hip -- the engine behind the C ABI of include/pbf.h.
//
// One context = one HIP stream, device buffers sized at creation for
// `capacity` particles and every scene's cells, and a cache of hipGraphs,
// one per (dt, iterations, diagnostics) shape, each holding the launch
// sequence of one substep.  All particle state is kept in cell-sorted slot
// order; `ids` maps a slot to the caller's particle index.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <exception>
#include <new>
#include <string>
#include <vector>

#include "../../include/pbf.h"
#include "pbf_device.h"
#include "pbf_internal.h"

using namespace pbf;

namespace {

enum Pass {
  P_PREDICT, P_SCAN, P_SCATTER, P_SETTLE, P_LIST, P_LAMBDA, P_DP, P_DENSITY, P_VORT, P_XSPH, P_DIAG, P_MLP, P_NPASS
};
const char *kPassNames =
    "predict_bin\nscan\nscatter\nsettle\nbuild_list\nlambda\ndp\ndensity\nvorticity\nxsph\ndiag\nmlp_infer";

// Substep kinds: the plain Jacobi solve, the learned warm start (network
// lambda at iteration 0), training-pair collection (features at iteration
// 0, lambda summed over all iterations), and the temporal warm start
// (iteration 0 applies the previous substep's lambda sum).
enum SubstepMode : uint32_t { M_PLAIN = 0, M_LEARNED = 1, M_COLLECT = 2, M_PREVIOUS = 3, M_RATIO = 4 };
// OR'ed into a substep mode: iterations 1.. project the accumulated lambda
// (the learned / collection modes only).
constexpr uint32_t M_PROJECT = 16;
// Graph-key bit of a learned substep followed by its device-side check.
constexpr uint32_t M_FALLBACK = 32;
// Graph-key bit of the plain re-solve of a rejected learned substep.
constexpr uint32_t M_RESOLVE = 128;
// OR'ed into a warm mode: the warm lambda field is smoothed over neighbours.
constexpr uint32_t M_SMOOTH = 64;

struct SceneHost {
  double lo[3], hi[3];
  uint32_t count;
  int dims[3];
  uint32_t ncells, cell_base, pbase;
  int field_kind = PBF_FIELD_NONE;
  int fres[3] = {0, 0, 0};
  double flo[3] = {0, 0, 0}, fhi[3] = {0, 0, 0};
  std::vector<float> field;  // cubic: sdf then vol; trilinear: interleaved
};

struct GraphEntry {
  uint32_t dt_bits, iters, diag, mode, aux;
  hipGraphExec_t exec;
  uint64_t used;  // launch serial of its latest replay
};

constexpr size_t kGraphCache = 16;  // executable graphs kept per context

struct PendingEvent {
  int pass;
  hipEvent_t a, b;
};

}  // namespace

struct pbf_ctx {
  pbf_params params;
  pbf_config cfg;
  std::string err;
  int device = 0;
  hipStream_t stream = nullptr;
  std::vector<SceneHost> scenes;
  bool batch = false;
  uint32_t n = 0, ncap = 0, ld = 0, maxn = 0, ncells = 0;
  std::vector<float> staged;
  uint64_t staged_serial = 0;
  // device state (slot order)
  float4 *pos = nullptr, *vel = nullptr, *acc = nullptr, *om = nullptr, *nrm = nullptr;
  float4 *XA = nullptr, *XB = nullptr, *Xold = nullptr, *B = nullptr, *xs = nullptr, *stage4 = nullptr;
  uint32_t *ids = nullptr, *cellkey = nullptr, *rank = nullptr, *tmp_key = nullptr, *tmp_src = nullptr;
  uint32_t *ncount = nullptr, *nbr = nullptr;
  uint32_t *nbr16 = nullptr, *wfmt = nullptr;  // 16-bit table and per-wavefront format
  float *lam = nullptr, *lam_sorted = nullptr;
  uint32_t *counts = nullptr, *offsets = nullptr;
  uint32_t *cellsub = nullptr;  // per cell: z-slab boundaries and count (null with one slab per cell)
  uint32_t slab_log2 = 2;
  std::vector<uint32_t *> scan_sums, scan_prefix;
  uint32_t *flags = nullptr;  // [0] bit 0 neighbour overflow, bit 1 non-finite position; [1] max neighbour count
  // scenes on device
  Scene *d_scenes = nullptr;
  uint32_t *sid = nullptr, *ranges = nullptr;
  std::vector<Scene> hscenes;
  float *pool = nullptr;
  size_t pool_floats = 0;
  bool any_field = false;
  // diagnostics
  double *ci = nullptr, *per_scene = nullptr, *hist = nullptr;
  uint32_t *hist_ctr = nullptr;
  uint32_t hist_cap = 4096;
  bool record_diag = false;
  // debug search
  uint32_t *dbg_counts = nullptr, *dbg_ids = nullptr, *ids_save = nullptr;
  bool last_was_search = false, have_dbg_counts = false, have_lists = false;
  // graphs and timing
  std::vector<GraphEntry> graphs;
  uint64_t graph_serial = 0, graph_evictions = 0;
  std::vector<hipEvent_t> event_pool;
  std::vector<PendingEvent> pending;
  double pass_ms[P_NPASS] = {0};
  uint32_t pass_count[P_NPASS] = {0};
  // learned warm start
  void *mlp = nullptr;           // network weights on the device
  uint32_t warm_mode = PBF_WARM_PLAIN;  // PBF_WARM_*
  bool warm_project = false;            // PBF_WARM_PROJECT
  bool warm_smooth = false;             // PBF_WARM_SMOOTH
  float *lam_tmp = nullptr;             // smoothed warm lambda
  double fb_threshold = 0.0;     // mean density error that rejects a learned substep (0: never)
  uint32_t fb_iters = 0;         // iterations of the plain re-solve
  float relax_set = -1.0f;       // scale applied to the network's lambda
  float relax_prev = 1.0f;       // scale applied to the previous lambda sum
  float *lam_acc = nullptr, *acc_sorted = nullptr, *feats = nullptr;
  float4 *vel_sorted = nullptr;
  // state saved before a learned substep that may be rejected
  float4 *snap_pos = nullptr, *snap_vel = nullptr, *snap_acc = nullptr, *snap_om = nullptr, *snap_nrm = nullptr;
  uint32_t *snap_ids = nullptr;
  float *snap_lam = nullptr, *snap_lacc = nullptr;
  uint64_t learned_substeps = 0;
  // device-side fallback decision: the decision word (pinned host memory and
  // its device address), the rejection counter, the
  // threshold, and the per-scene speed reduction
  void *decision = nullptr;
  uint32_t *decision_dev = nullptr, *fb_count = nullptr;
  double *fb_thr = nullptr, *per_scene2 = nullptr, *energy0 = nullptr, *energy1 = nullptr;
  // per-scene reductions: blocks per scene (about 16 k slots each) and partials
  uint32_t red_k = 1;
  double *partial = nullptr;
  // launch tuning
  pbf_tuning tuning = {{256, 256, 256, 256, 256, 256, 256, 256}, 0, {0, 0, 0, 0, 0, 0, 0}};
};

namespace {

// The message of the last failed create on this thread (pbf_last_error(NULL)).
thread_local std::string g_create_error;

}  // namespace

void pbf::set_thread_error(const std::string &msg) { g_create_error = msg; }

namespace {

int fail(pbf_ctx *c, int code, const std::string &msg) {
  if (c) c->err = msg;
  return code;
}

#define HIPCHK(expr)                                                                              \
  do {                                                                                            \
    hipError_t e_ = (expr);                                                                       \
    if (e_ != hipSuccess) return fail(c, PBF_E_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <typename T>
hipError_t dalloc(T **p, size_t count) {
  *p = nullptr;
  if (count == 0) count = 1;
  return hipMalloc((void **)p, count * sizeof(T));
}

Phys make_phys(const pbf_ctx *c, float dt) {
  const double r = c->params.particle_radius, h = c->params.kernel_radius, rho0 = c->params.rest_density;
  const double PI = 3.141592653589793;
  const double d = 2.0 * r;
  const double V64 = ((0.9 * d) * d) * d;  // SPH.ts:48-50
  Phys P;
  P.h = (float)h;
  P.h2 = (float)(h * h);
  P.half_h = (float)(0.5 * h);
  const float rf = (float)r;
  P.half_r = 0.5f * rf;
  P.two_r = 2.0f * rf;
  P.V = (float)V64;
  P.V2 = (float)(V64 * V64);
  P.m = (float)(V64 * rho0);
  P.Kp = (float)(315.0 / (64.0 * PI * std::pow(h, 9)));
  P.Ks = (float)(-45.0 / (PI * std::pow(h, 6)));
  P.Kc = (float)(32.0 / (PI * std::pow(h, 9)));
  P.bias = (float)(-std::pow(h, 6) / 64.0);
  P.two_rho0 = (float)(2.0 * rho0);
  P.eps_lambda = 1e-6f;
  P.dt = dt;
  P.inv_dt = (float)(1.0 / (double)dt);
  P.gx = 0.0f;
  P.gy = -c->cfg.gravity;
  P.gz = 0.0f;
  P.xsph_k = (c->cfg.xsph * P.inv_dt) * P.m;
  P.vort = c->cfg.vorticity;
  P.tension = c->cfg.surface_tension;
  P.ld = c->ld;
  P.maxn = c->maxn;
  P.nbr16 = c->nbr16;
  P.wfmt = c->wfmt;
  P.cellsub = c->cellsub;
  P.slab_log2 = c->cellsub ? c->slab_log2 : 0u;
  P.Kp64 = 315.0 / (64.0 * PI * std::pow(h, 9));
  P.h2_64 = h * h;
  P.V64 = V64;
  return P;
}

// Destroying an executable graph that is still running on the stream frees
// what its in-flight launch uses (and what a profiler's completion callbacks
// read), so every destroy waits for the stream first.
void clear_graphs(pbf_ctx *c) {
  if (c->graphs.empty()) return;
  (void)hipStreamSynchronize(c->stream);
  for (auto &g : c->graphs) (void)hipGraphExecDestroy(g.exec);
  c->graphs.clear();
}

// A full cache drops its least recently used executable, after the stream
// has drained (the one just launched may be that one's last user).
void evict_graph(pbf_ctx *c) {
  if (c->graphs.empty()) return;
  size_t old = 0;
  for (size_t k = 1; k < c->graphs.size(); ++k)
    if (c->graphs[k].used < c->graphs[old].used) old = k;
  (void)hipStreamSynchronize(c->stream);
  (void)hipGraphExecDestroy(c->graphs[old].exec);
  c->graphs.erase(c->graphs.begin() + (long)old);
  c->graph_evictions += 1;
}

LaunchCtx make_launch(pbf_ctx *c, float dt) {
  LaunchCtx L;
  L.stream = c->stream;
  L.n = c->n;
  L.multi = c->scenes.size() > 1;
  L.hash = c->params.grid_mode == PBF_GRID_HASH;
  L.field = c->any_field;
  L.fmode = 0;
  for (auto &s : c->scenes) {
    if (s.field_kind == PBF_FIELD_CUBIC) L.fmode = 1;
    else if (s.field_kind == PBF_FIELD_TRILINEAR && L.fmode == 0) L.fmode = 2;
  }
  L.s0 = c->hscenes.empty() ? Scene{} : c->hscenes[0];
  L.scenes = c->d_scenes;
  L.sid = c->sid;
  L.pool = c->pool;
  L.P = make_phys(c, dt);
  L.red_k = c->red_k;
  L.partial = c->partial;
  return L;
}

hipEvent_t take_event(pbf_ctx *c) {
  if (!c->event_pool.empty()) {
    hipEvent_t e = c->event_pool.back();
    c->event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) e = nullptr;
  return e;
}

// Per-pass timing: begin() returns the launch context with a start/stop
// event pair attached to the pass's first/last dispatch packet.
struct PassTimer {
  pbf_ctx *c;
  bool on;
  int pass = -1;
  hipEvent_t a = nullptr, b = nullptr;
  LaunchCtx begin(int p, const LaunchCtx &L) {
    LaunchCtx t = L;
    const int walk = p == P_LAMBDA ? 0 : p == P_DP ? 1 : p == P_DENSITY ? 2 : p == P_VORT ? 3 : p == P_XSPH ? 4
                   : p == P_DIAG ? 5 : -1;
    if (walk >= 0) t.bs = c->tuning.walk_block[walk];
    t.lvar = c->tuning.list_variant;
    if (!on) return t;
    pass = p;
    a = take_event(c);
    b = take_event(c);
    if (!a || !b) {
      pass = -1;
      return t;
    }
    t.ev0 = a;
    t.ev1 = b;
    return t;
  }
  void end() {
    if (!on || pass < 0) return;
    c->pending.push_back({pass, a, b});
    pass = -1;
  }
};

void drain_events(pbf_ctx *c) {
  for (auto &pe : c->pending) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, pe.a, pe.b) == hipSuccess) {
      c->pass_ms[pe.pass] += ms;
      c->pass_count[pe.pass] += 1;
    }
    c->event_pool.push_back(pe.a);
    c->event_pool.push_back(pe.b);
  }
  c->pending.clear();
}

// Predict, bin, scan, reorder and neighbour lists.  dbg: keep cell counts.
void enqueue_search(pbf_ctx *c, const LaunchCtx &L, PassTimer &T, bool dbg, uint32_t mode = M_PLAIN) {
  LaunchCtx t = T.begin(P_PREDICT, L);
  launch_predict_bin(t, c->pos, c->vel, c->acc, c->XB, c->cellkey, c->rank, c->counts, c->flags);
  T.end();
  if (dbg) (void)hipMemcpyAsync(c->dbg_counts, c->counts, (size_t)c->ncells * 4, hipMemcpyDeviceToDevice, c->stream);
  t = T.begin(P_SCAN, L);
  launch_scan(c->stream, c->counts, c->ncells, c->offsets, 1, c->scan_sums.data(), c->scan_prefix.data(), 0, t.ev0,
              t.ev1, c->cellsub);
  T.end();
  t = T.begin(P_SCATTER, L);
  launch_scatter(t, c->cellkey, c->rank, c->offsets, c->ids, c->tmp_key, c->tmp_src);
  T.end();
  t = T.begin(P_SETTLE, L);
  // the binning ranks are consumed by the scatter: their buffer takes the
  // final slots
  // the previous lambdas and the slot map only feed the learned and
  // data-generation modes' extra state (k_settle_extra, the features)
  const bool extra = mode != M_PLAIN;
  launch_settle(t, c->tmp_key, c->tmp_src, c->cellkey, c->offsets, c->XB, c->pos, c->lam, c->XA, c->Xold, c->ids,
                extra ? c->lam_sorted : nullptr, extra ? c->rank : nullptr);
  T.end();
  if (mode != M_PLAIN)
    launch_settle_extra(L, c->tmp_src, c->rank, c->lam_acc, c->acc_sorted, c->vel, c->vel_sorted);
  t = T.begin(P_LIST, L);
  launch_build_list(t, c->XA, c->offsets, c->nbr, c->ncount, c->flags);
  T.end();
}

// The launch sequence of one substep (PBF.ts:398-448).
// diag: record the substep's density error (history); ci_out: leave the
// per-particle FP64 error in c->ci for a check -- from the density pass, or,
// with cubic fields, from the diagnostic's own walk, which then also reduces
// it into per_scene.
void enqueue_substep(pbf_ctx *c, const LaunchCtx &L, uint32_t iters, bool diag, PassTimer &T, uint32_t mode_bits,
                     bool ci_out = false) {
  const bool project = (mode_bits & M_PROJECT) != 0, smooth = (mode_bits & M_SMOOTH) != 0;
  const uint32_t mode = mode_bits & ~(M_PROJECT | M_SMOOTH);
  enqueue_search(c, L, T, false, mode);
  float4 *X = c->XA, *Xn = c->XB;
  LearnBufs lb{c->lam_acc, c->acc_sorted, c->lam_sorted, c->vel_sorted, c->feats};
  for (uint32_t it = 0; it < iters; ++it) {
    LaunchCtx t = T.begin(P_LAMBDA, L);
    if (mode != M_PLAIN && it == 0) launch_lambda_features(t, X, c->B, c->nbr, c->ncount, lb);
    else launch_lambda(t, X, c->B, c->nbr, c->ncount, mode != M_PLAIN ? c->lam_acc : nullptr, project);
    T.end();
    if ((mode == M_LEARNED || mode == M_RATIO) && it == 0) {
      t = T.begin(P_MLP, L);
      launch_mlp_infer(c->stream, c->mlp, c->n, c->feats, X, c->lam_acc,
                       mode == M_RATIO ? PBF_WARM_RATIO_MAX * c->relax_prev : 0.0f, t.ev0, t.ev1);
      T.end();
    } else if (mode == M_PREVIOUS && it == 0) {
      launch_warm_previous(L, c->acc_sorted, X, c->lam_acc, c->relax_prev);
    }
    if (smooth && it == 0 && (mode == M_LEARNED || mode == M_RATIO || mode == M_PREVIOUS))
      launch_smooth_warm(L, X, c->nbr, c->ncount, c->lam_tmp, c->lam_acc);
    t = T.begin(P_DP, L);
    launch_dp(t, X, c->B, c->nbr, c->ncount, Xn);
    T.end();
    std::swap(X, Xn);
  }
  LaunchCtx t = T.begin(P_DENSITY, L);
  const bool fused = (diag || ci_out) && L.fmode != 1;
  launch_density(t, X, c->Xold, c->nbr, c->ncount, c->pos, c->vel, c->lam, fused ? c->ci : nullptr);
  T.end();
  t = T.begin(P_VORT, L);
  // the XSPH velocity sum goes to the free Jacobi buffer
  launch_vorticity(t, c->pos, c->vel, c->nbr, c->ncount, c->om, c->nrm, Xn);
  T.end();
  t = T.begin(P_XSPH, L);
  launch_xsph(t, c->pos, c->om, c->nrm, Xn, c->nbr, c->ncount, c->acc);
  T.end();
  if (diag) {
    t = T.begin(P_DIAG, L);
    if (fused)
      launch_diag_reduce(t, c->ci, c->ranges, (uint32_t)c->scenes.size(), c->per_scene, c->hist, c->hist_cap,
                         c->hist_ctr);
    else
      launch_diag(t, c->pos, c->nbr, c->ncount, c->ci, c->ranges, (uint32_t)c->scenes.size(), c->per_scene, c->hist,
                  c->hist_cap, c->hist_ctr);
    T.end();
  } else if (ci_out && !fused) {
    launch_diag(L, c->pos, c->nbr, c->ncount, c->ci, c->ranges, (uint32_t)c->scenes.size(), c->per_scene, nullptr,
                0, nullptr);
  }
}

// Runs an enqueue sequence once: directly, or as a cached hipGraph per
// (dt, iterations, diagnostics, mode, aux) shape.
template <typename Enqueue>
int run_enqueued(pbf_ctx *c, float dt, uint32_t iters, bool diag, uint32_t mode, uint32_t aux, Enqueue &&enq) {
  LaunchCtx L = make_launch(c, dt);
  const bool timing = (c->params.flags & PBF_FLAG_TIMING) != 0;
  const bool use_graph = !(c->params.flags & PBF_FLAG_NO_GRAPH) && !timing;
  PassTimer T{c, timing};
  c->last_was_search = false;
  c->have_dbg_counts = false;
  c->have_lists = true;
  if (!use_graph) {
    enq(L, T);
    HIPCHK(hipGetLastError());
    return PBF_OK;
  }
  uint32_t bits;
  std::memcpy(&bits, &dt, 4);
  hipGraphExec_t exec = nullptr;
  GraphEntry *hit = nullptr;
  for (auto &g : c->graphs)
    if (g.dt_bits == bits && g.iters == iters && g.diag == (uint32_t)diag && g.mode == mode && g.aux == aux) hit = &g;
  if (hit) {
    exec = hit->exec;
    hit->used = ++c->graph_serial;
  } else {
    hipGraph_t graph = nullptr;
    HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    enq(L, T);
    hipError_t le = hipGetLastError();
    hipError_t ee = hipStreamEndCapture(c->stream, &graph);
    if (le != hipSuccess || ee != hipSuccess) {
      if (graph) (void)hipGraphDestroy(graph);
      if (le != hipSuccess) return fail(c, PBF_E_DEVICE, std::string("launch during capture: ") + hipGetErrorString(le));
      return fail(c, PBF_E_DEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(ee));
    }
    hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ie != hipSuccess) return fail(c, PBF_E_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(ie));
    if (c->graphs.size() >= kGraphCache) evict_graph(c);
    c->graphs.push_back({bits, iters, (uint32_t)diag, mode, aux, exec, ++c->graph_serial});
  }
  HIPCHK(hipGraphLaunch(exec, c->stream));
  return PBF_OK;
}

int run_substep(pbf_ctx *c, float dt, uint32_t iters, uint32_t mode, bool diag) {
  return run_enqueued(c, dt, iters, diag, mode, 0u,
                      [&](const LaunchCtx &L, PassTimer &T) { enqueue_substep(c, L, iters, diag, T, mode); });
}

// Buffers of the learned and collection modes (allocated on first use).
int ensure_learn_buffers(pbf_ctx *c) {
  if (c->lam_acc) return PBF_OK;
  const size_t N = c->ncap;
  HIPCHK(dalloc(&c->lam_acc, N));
  HIPCHK(dalloc(&c->acc_sorted, N));
  HIPCHK(dalloc(&c->feats, 16 * N));
  HIPCHK(dalloc(&c->vel_sorted, N));
  HIPCHK(dalloc(&c->lam_tmp, N));
  HIPCHK(hipMemsetAsync(c->lam_acc, 0, N * 4, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return PBF_OK;
}

int ensure_snapshot(pbf_ctx *c) {
  if (c->snap_pos) return PBF_OK;
  const size_t N = c->ncap;
  HIPCHK(dalloc(&c->snap_pos, N));
  HIPCHK(dalloc(&c->snap_vel, N));
  HIPCHK(dalloc(&c->snap_acc, N));
  HIPCHK(dalloc(&c->snap_om, N));
  HIPCHK(dalloc(&c->snap_nrm, N));
  HIPCHK(dalloc(&c->snap_ids, N));
  HIPCHK(dalloc(&c->snap_lam, N));
  HIPCHK(dalloc(&c->snap_lacc, N));
  return PBF_OK;
}

// Copies the state carried from one substep to the next (slot order), and
// with `readable` the omega / normal fields a caller can read back (a substep
// that follows the restore recomputes them).
void copy_state(pbf_ctx *c, bool save, bool readable = true) {
  const size_t n = c->n;
  struct { void *live, *snap; size_t bytes; } t[] = {
      {c->pos, c->snap_pos, 16 * n}, {c->vel, c->snap_vel, 16 * n},   {c->acc, c->snap_acc, 16 * n},
      {c->ids, c->snap_ids, 4 * n},  {c->lam, c->snap_lam, 4 * n},    {c->lam_acc, c->snap_lacc, 4 * n},
      {c->om, c->snap_om, 16 * n},   {c->nrm, c->snap_nrm, 16 * n}};
  for (size_t k = 0; k < (readable ? 8u : 6u); ++k)
    if (t[k].live && t[k].snap)
      launch_copy(c->stream, save ? t[k].snap : t[k].live, save ? t[k].live : t[k].snap, t[k].bytes);
}

// One learned substep.  With a rejection threshold set, the check runs on the
// device in the substep's own launch sequence (graph): the energy is summed,
// the state saved, the learned substep runs (its density pass also leaves
// every particle's FP64 density error), one reduction takes the mean error,
// the largest |v| dt / h and the energy, and k_fallback_gate writes its
// decision to pinned host memory.  The host reads that one word once the
// stream has drained and, for a rejected substep, launches the restore and
// the plain re-solve (the collection mode: the plain solve that also keeps
// the lambda sum the next learned substep reads).  Predicating the re-solve's
// kernels on a device word instead (no host wait) was measured slower at
// 1M particles: its ~100 empty launches over every slot cost 0.25 ms a
// substep, more than the wait.
int run_learned_substep(pbf_ctx *c, float dt, uint32_t iters) {
  c->learned_substeps += 1;
  const uint32_t mode = (c->warm_mode == PBF_WARM_PREVIOUS ? M_PREVIOUS
                         : c->warm_mode == PBF_WARM_RATIO  ? M_RATIO
                         : c->warm_mode == PBF_WARM_PLAIN  ? M_COLLECT
                                                           : M_LEARNED) |
                        (c->warm_project ? M_PROJECT : 0u) | (c->warm_smooth ? M_SMOOTH : 0u);
  if (!(c->fb_threshold > 0.0)) return run_substep(c, dt, iters, mode, c->record_diag);
  int rc = ensure_snapshot(c);
  if (rc != PBF_OK) return rc;
  const uint32_t fb_iters = c->fb_iters ? c->fb_iters : iters;
  const uint32_t ns = (uint32_t)c->scenes.size();
  rc = run_enqueued(c, dt, iters, false, mode | M_FALLBACK, 0u, [&](const LaunchCtx &L, PassTimer &T) {
    launch_energy(L, c->pos, c->vel, c->ranges, ns, c->energy0);
    copy_state(c, true, false);
    enqueue_substep(c, L, iters, false, T, mode, true);
    launch_check_reduce(L, c->ci, c->pos, c->vel, c->ranges, ns, c->per_scene, c->per_scene2, c->energy1);
    launch_fallback_gate(c->stream, c->per_scene, c->per_scene2, c->energy0, c->energy1, ns, c->n, c->fb_thr,
                         c->decision_dev, c->fb_count);
  });
  if (rc != PBF_OK) return rc;
  HIPCHK(hipStreamSynchronize(c->stream));
  const bool rejected = *static_cast<volatile uint32_t *>(c->decision) != 0u;
  if (rejected) {
    rc = run_enqueued(c, dt, fb_iters, false, M_COLLECT | M_RESOLVE, 0u, [&](const LaunchCtx &L, PassTimer &T) {
      copy_state(c, false, false);
      enqueue_substep(c, L, fb_iters, false, T, M_COLLECT);
    });
    if (rc != PBF_OK) return rc;
  }
  if (c->record_diag) {
    LaunchCtx L = make_launch(c, dt);
    // an accepted substep's density error is the check's own (per_scene,
    // left by its diagnostic walk); a re-solved one is measured again
    if (rejected)
      launch_diag(L, c->pos, c->nbr, c->ncount, c->ci, c->ranges, ns, c->per_scene, c->hist, c->hist_cap,
                  c->hist_ctr);
    else
      launch_diag_history(L, c->per_scene, ns, c->hist, c->hist_cap, c->hist_ctr);
    HIPCHK(hipGetLastError());
  }
  return PBF_OK;
}

int check_flags(pbf_ctx *c) {
  uint32_t f[2] = {0, 0};
  HIPCHK(hipMemcpy(f, c->flags, 8, hipMemcpyDeviceToHost));
  if (f[0] & 2u) {
    HIPCHK(hipMemset(c->flags, 0, 4));
    return fail(c, PBF_E_OVERFLOW, "non-finite particle position: the simulation diverged");
  }
  if (f[0]) {
    HIPCHK(hipMemset(c->flags, 0, 4));
    char buf[160];
    std::snprintf(buf, sizeof(buf), "neighbour overflow: a particle had %u neighbours, max_neighbors is %u", f[1],
                  c->maxn);
    return fail(c, PBF_E_OVERFLOW, buf);
  }
  return PBF_OK;
}

int sync_ctx(pbf_ctx *c) {
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipStreamSynchronize(c->stream));
  drain_events(c);
  return check_flags(c);
}

// Grid of cell size h over the scene box; false when it cannot be indexed
// (an axis of 2^30 cells or more, or a dense grid of 2^31 cells or more).
bool compute_grid(const pbf_params &p, SceneHost &s) {
  uint64_t dense = 1;
  for (int a = 0; a < 3; ++a) {
    const double L = s.hi[a] - s.lo[a];
    const double D = std::ceil(L / p.kernel_radius - 1e-7);
    if (!(D < 1073741824.0)) return false;
    s.dims[a] = D < 1.0 ? 1 : (int)D;
    dense *= (uint64_t)s.dims[a];
  }
  if (p.grid_mode == PBF_GRID_HASH) {
    s.ncells = 1u << p.hash_log2;
  } else {
    if (dense > 0x7fffffffull) return false;
    s.ncells = (uint32_t)dense;
  }
  return true;
}

Scene to_device_scene(const pbf_ctx *c, const SceneHost &s, uint64_t field_off) {
  Scene d;
  std::memset(&d, 0, sizeof(d));
  for (int a = 0; a < 3; ++a) {
    d.lo[a] = (float)s.lo[a];
    d.hi[a] = (float)s.hi[a];
    d.dims[a] = s.dims[a];
    d.fres[a] = s.fres[a];
    d.flo[a] = (float)s.flo[a];
    double L = s.fhi[a] - s.flo[a];
    d.finv[a] = s.field_kind ? (float)(s.fres[a] / L) : 0.0f;
    d.fscale[a] = s.field_kind == PBF_FIELD_CUBIC ? (float)(2.0 * s.fres[a] / L) : d.finv[a];
  }
  d.inv_cell = (float)(1.0 / c->params.kernel_radius);
  d.cell_base = s.cell_base;
  d.ncells = s.ncells;
  d.hash_mask = c->params.grid_mode == PBF_GRID_HASH ? s.ncells - 1u : 0u;
  // the list build's pruned column walk needs the 27 cells around every
  // cell in 27 different buckets (always true on a dense grid)
  d.distinct27 = !d.hash_mask || pbf_hash_buckets_distinct(s.dims, c->params.hash_log2);
  d.pbase = s.pbase;
  d.pcount = s.count;
  d.field_kind = s.field_kind;
  d.quirk = (c->params.flags & PBF_FLAG_CUBIC_FIXED) ? 0u : 1u;
  d.field_off = field_off;
  return d;
}

// Rebuild the device scene table and the boundary-field pool.
int upload_scenes(pbf_ctx *c) {
  uint32_t most = 0;
  for (auto &s : c->scenes) most = std::max(most, s.count);
  c->red_k = std::min<uint32_t>(kMaxReduceBlocks, std::max<uint32_t>(1u, (most + 16383u) / 16384u));
  size_t total = 0;
  for (auto &s : c->scenes) total += s.field.size();
  std::vector<uint64_t> offs;
  c->hscenes.clear();
  c->any_field = false;
  uint64_t o = 0;
  for (auto &s : c->scenes) {
    offs.push_back(o);
    o += s.field.size();
    // keep float2 loads of trilinear nodes 8-byte aligned
    if (o & 1) o += 1;
    if (s.field_kind) c->any_field = true;
  }
  total = o;
  if (total > c->pool_floats || !c->pool) {
    if (c->pool) HIPCHK(hipFree(c->pool));
    c->pool = nullptr;
    HIPCHK(dalloc(&c->pool, total ? total : 2));
    c->pool_floats = total;
  }
  for (size_t k = 0; k < c->scenes.size(); ++k) {
    auto &s = c->scenes[k];
    if (!s.field.empty())
      HIPCHK(hipMemcpy(c->pool + offs[k], s.field.data(), s.field.size() * 4, hipMemcpyHostToDevice));
    c->hscenes.push_back(to_device_scene(c, s, offs[k]));
  }
  HIPCHK(hipMemcpy(c->d_scenes, c->hscenes.data(), c->hscenes.size() * sizeof(Scene), hipMemcpyHostToDevice));
  clear_graphs(c);
  return PBF_OK;
}

constexpr uint32_t kMaxCapacity = (1u << 27) - 256u;
constexpr uint32_t kMaxIterations = 4096;  // Jacobi iterations per substep (one graph node pair each)

int create_ctx_body(const pbf_params *params, uint32_t nscenes, const pbf_scene_desc *descs, bool batch, pbf_ctx **out,
                    pbf_ctx **made) {
  pbf_ctx *c = nullptr;
  g_create_error.clear();
  if (!params || !out || nscenes == 0) {
    g_create_error = "params, out and at least one scene are required";
    return PBF_E_INVALID;
  }
  *out = nullptr;
  c = new pbf_ctx();
  *made = c;
  c->params = *params;
  pbf_params &p = c->params;
  auto reject = [&](int code, const std::string &msg) {
    g_create_error = msg;
    *made = nullptr;
    delete c;
    return code;
  };
  if (!(p.particle_radius > 0.0) || !std::isfinite(p.particle_radius))
    return reject(PBF_E_INVALID, "particle_radius must be finite and > 0");
  if (!std::isfinite(p.kernel_radius) || !std::isfinite(p.rest_density) || !std::isfinite(p.time_step))
    return reject(PBF_E_INVALID, "kernel_radius, rest_density and time_step must be finite");
  if (p.kernel_radius <= 0.0) p.kernel_radius = 4.0 * p.particle_radius;
  if (p.rest_density <= 0.0) p.rest_density = 1000.0;
  if (p.time_step <= 0.0) p.time_step = 1.0 / 300.0;
  if (p.max_neighbors == 0) p.max_neighbors = 96;
  if (p.max_neighbors > 65535u) return reject(PBF_E_INVALID, "max_neighbors must be below 65536");
  if (p.grid_mode != PBF_GRID_DENSE && p.grid_mode != PBF_GRID_HASH)
    return reject(PBF_E_INVALID, "grid_mode must be PBF_GRID_DENSE or PBF_GRID_HASH");
  if (p.grid_mode == PBF_GRID_HASH && (p.hash_log2 < 4 || p.hash_log2 > 28))
    return reject(PBF_E_INVALID, "hash_log2 must be in [4, 28]");
  if (p.capacity > kMaxCapacity)  // 32-bit gather offsets (slot * 16 bytes)
    return reject(PBF_E_CAPACITY, "capacity above 2^27 - 256 particles");
  if (p.cell_slabs == 0) p.cell_slabs = 4;
  if (p.cell_slabs != 1 && p.cell_slabs != 2 && p.cell_slabs != 4)
    return reject(PBF_E_INVALID, "cell_slabs must be 1, 2 or 4");
  c->slab_log2 = p.cell_slabs == 4 ? 2u : p.cell_slabs == 2 ? 1u : 0u;
  c->cfg = {5u, 0.01f, 0.1f, 0.5f, 9.8f};  // src/common/config.ts:22-28
  c->batch = batch;
  uint32_t cell_base = 0, pbase = 0;
  uint64_t total_cells = 0, total_particles = 0;
  for (uint32_t k = 0; k < nscenes; ++k) {
    SceneHost s;
    for (int a = 0; a < 3; ++a) {
      s.lo[a] = descs ? descs[k].domain_lo[a] : p.domain_lo[a];
      s.hi[a] = descs ? descs[k].domain_hi[a] : p.domain_hi[a];
      if (!(s.hi[a] > s.lo[a]) || !std::isfinite(s.lo[a]) || !std::isfinite(s.hi[a]))
        return reject(PBF_E_INVALID, "scene domain must be finite with hi > lo");
    }
    s.count = descs ? descs[k].particle_count : 0;
    if (!compute_grid(p, s)) return reject(PBF_E_CAPACITY, "too many grid cells");
    s.cell_base = cell_base;
    s.pbase = pbase;
    total_cells += s.ncells;
    total_particles += s.count;
    if (total_cells > 0x7fffffffull) return reject(PBF_E_CAPACITY, "too many grid cells");
    cell_base += s.ncells;
    pbase += s.count;
    c->scenes.push_back(s);
  }
  if (batch && total_particles > p.capacity) {
    return reject(PBF_E_CAPACITY, "To Many Particles!");
  }
  c->ncells = (uint32_t)total_cells;
  c->ncap = ((p.capacity + 255u) / 256u) * 256u;
  if (c->ncap == 0) c->ncap = 256;
  // slot-major neighbour table: entry k of slot i at nbr[k * ld + i]
  c->maxn = p.max_neighbors;
  c->ld = c->ncap;
  hipError_t e = hipGetDevice(&c->device);
  if (e != hipSuccess) return reject(PBF_E_DEVICE, std::string("hipGetDevice: ") + hipGetErrorString(e));
  int rc = PBF_OK;
  auto step = [&](hipError_t err, const char *what) {
    if (rc == PBF_OK && err != hipSuccess) rc = fail(c, PBF_E_DEVICE, std::string(what) + ": " + hipGetErrorString(err));
  };
  step(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "hipStreamCreate");
  const size_t N = c->ncap;
  step(dalloc(&c->pos, N), "alloc"); step(dalloc(&c->vel, N), "alloc"); step(dalloc(&c->acc, N), "alloc");
  step(dalloc(&c->om, N), "alloc"); step(dalloc(&c->nrm, N), "alloc"); step(dalloc(&c->XA, N), "alloc");
  step(dalloc(&c->XB, N), "alloc"); step(dalloc(&c->Xold, N), "alloc"); step(dalloc(&c->B, N), "alloc");
  step(dalloc(&c->stage4, N), "alloc");
  step(dalloc(&c->ids, N), "alloc"); step(dalloc(&c->cellkey, N), "alloc"); step(dalloc(&c->rank, N), "alloc");
  step(dalloc(&c->tmp_key, N), "alloc"); step(dalloc(&c->tmp_src, N), "alloc"); step(dalloc(&c->ncount, N), "alloc");
  step(dalloc(&c->nbr, (size_t)c->maxn * c->ld), "alloc");
  step(dalloc(&c->nbr16, (size_t)((c->maxn + 7) / 8) * 4 * c->ld), "alloc");
  step(dalloc(&c->wfmt, (size_t)c->ncap / 64), "alloc");
  step(dalloc(&c->lam, N), "alloc"); step(dalloc(&c->lam_sorted, N), "alloc");
  step(dalloc(&c->counts, (size_t)c->ncells), "alloc");
  step(dalloc(&c->offsets, (size_t)c->ncells + 1), "alloc");
  if (c->slab_log2) step(dalloc(&c->cellsub, (size_t)c->ncells), "alloc");
  step(dalloc(&c->dbg_counts, (size_t)c->ncells), "alloc");
  step(dalloc(&c->dbg_ids, N), "alloc"); step(dalloc(&c->ids_save, N), "alloc");
  step(dalloc(&c->flags, 2), "alloc");
  step(dalloc(&c->d_scenes, nscenes), "alloc");
  step(dalloc(&c->sid, N), "alloc");
  step(dalloc(&c->ranges, 2 * (size_t)nscenes), "alloc");
  step(dalloc(&c->ci, N), "alloc");
  step(dalloc(&c->per_scene, 2 * (size_t)nscenes), "alloc");
  step(dalloc(&c->per_scene2, 2 * (size_t)nscenes), "alloc");
  step(dalloc(&c->energy0, 2 * (size_t)nscenes), "alloc");
  step(dalloc(&c->partial, 6 * (size_t)nscenes * kMaxReduceBlocks), "alloc");
  step(hipHostMalloc(&c->decision, 4, hipHostMallocMapped), "hipHostMalloc");
  if (rc == PBF_OK) {
    *static_cast<uint32_t *>(c->decision) = 0u;
    step(hipHostGetDevicePointer((void **)&c->decision_dev, c->decision, 0), "hipHostGetDevicePointer");
  }
  step(dalloc(&c->energy1, 2 * (size_t)nscenes), "alloc");
  step(dalloc(&c->fb_count, 1), "alloc");
  step(dalloc(&c->fb_thr, 1), "alloc");
  step(dalloc(&c->hist, 2 * (size_t)c->hist_cap), "alloc");
  step(dalloc(&c->hist_ctr, 1), "alloc");
  // scan workspace: one level per factor of 2048
  {
    uint32_t len = c->ncells;
    while (rc == PBF_OK && len > 2048u) {
      uint32_t nb = (len + 2047u) / 2048u;
      uint32_t *s = nullptr, *q = nullptr;
      step(dalloc(&s, nb), "alloc");
      step(dalloc(&q, (size_t)nb + 1), "alloc");
      c->scan_sums.push_back(s);
      c->scan_prefix.push_back(q);
      len = nb;
    }
    c->scan_sums.push_back(nullptr);
    c->scan_prefix.push_back(nullptr);
  }
  if (rc == PBF_OK) {
    step(hipMemset(c->counts, 0, (size_t)c->ncells * 4), "hipMemset");
    step(hipMemset(c->offsets, 0, ((size_t)c->ncells + 1) * 4), "hipMemset");
    step(hipMemset(c->flags, 0, 8), "hipMemset");
    step(hipMemset(c->hist_ctr, 0, 4), "hipMemset");
    step(hipMemset(c->fb_count, 0, 4), "hipMemset");
    step(hipMemset(c->fb_thr, 0, 8), "hipMemset");
    step(hipMemset(c->ncount, 0, N * 4), "hipMemset");
  }
  if (rc == PBF_OK) rc = upload_scenes(c);
  if (rc != PBF_OK) {
    g_create_error = c->err;
    *made = nullptr;
    pbf_destroy(c);
    return rc;
  }
  *out = c;
  return PBF_OK;
}

// create_ctx_body with host-allocation failures turned into a status; a
// context already made is released with its device buffers.
int create_ctx(const pbf_params *params, uint32_t nscenes, const pbf_scene_desc *descs, bool batch, pbf_ctx **out) {
  pbf_ctx *made = nullptr;
  try {
    return create_ctx_body(params, nscenes, descs, batch, out, &made);
  } catch (const std::exception &) {
    if (made) pbf_destroy(made);
    if (out) *out = nullptr;
    g_create_error = "out of host memory";
    return PBF_E_CAPACITY;
  }
}

int upload_positions(pbf_ctx *c, const float *xyzw, uint32_t count) {
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (c->batch) {
    uint64_t want = 0;
    for (auto &s : c->scenes) want += s.count;
    if (want != count) return fail(c, PBF_E_INVALID, "batched context: particle count must equal the sum of scene counts");
  } else {
    c->scenes[0].count = count;
    c->scenes[0].pbase = 0;
  }
  c->n = count;
  const size_t N = count;
  std::vector<float> zero(4 * (N ? N : 1), 0.0f);
  std::vector<float> p4(4 * (N ? N : 1), 0.0f);
  for (size_t i = 0; i < N; ++i) {
    p4[4 * i] = xyzw[4 * i];
    p4[4 * i + 1] = xyzw[4 * i + 1];
    p4[4 * i + 2] = xyzw[4 * i + 2];
    p4[4 * i + 3] = 0.0f;
  }
  if (N) {
    HIPCHK(hipMemcpy(c->pos, p4.data(), N * 16, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->vel, zero.data(), N * 16, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->acc, zero.data(), N * 16, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->om, zero.data(), N * 16, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->nrm, zero.data(), N * 16, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->B, zero.data(), N * 16, hipMemcpyHostToDevice));
    HIPCHK(hipMemset(c->lam, 0, N * 4));
    HIPCHK(hipMemset(c->ncount, 0, N * 4));
    if (c->lam_acc) HIPCHK(hipMemset(c->lam_acc, 0, N * 4));
  }
  std::vector<uint32_t> idv(N ? N : 1), sidv(N ? N : 1);
  for (size_t i = 0; i < N; ++i) idv[i] = (uint32_t)i;
  std::vector<uint32_t> rng;
  for (size_t k = 0; k < c->scenes.size(); ++k) {
    auto &s = c->scenes[k];
    for (uint32_t q = 0; q < s.count; ++q) sidv[s.pbase + q] = (uint32_t)k;
    rng.push_back(s.pbase);
    rng.push_back(s.count);
  }
  if (N) {
    HIPCHK(hipMemcpy(c->ids, idv.data(), N * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->sid, sidv.data(), N * 4, hipMemcpyHostToDevice));
  }
  HIPCHK(hipMemcpy(c->ranges, rng.data(), rng.size() * 4, hipMemcpyHostToDevice));
  HIPCHK(hipMemset(c->counts, 0, (size_t)c->ncells * 4));
  c->have_lists = false;
  c->last_was_search = false;
  return upload_scenes(c);
}

int read_sorted4(pbf_ctx *c, const float4 *src, const uint32_t *ids, float *out, uint32_t count, int clear_w = 0) {
  if (count != c->n) return fail(c, PBF_E_LENGTH, "readback count must equal the particle count");
  int rc = sync_ctx(c);
  if (rc != PBF_OK) return rc;
  launch_unsort4(c->stream, c->n, src, ids, c->stage4, clear_w);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(out, c->stage4, (size_t)c->n * 16, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return PBF_OK;
}

}  // namespace

// ===========================================================================
// C ABI
// Host allocations inside an entry point (staging vectors, graphs, event
// pools) report failure as a status: nothing unwinds through the C ABI.
#define PBF_CATCH(ctx)                                                      \
  catch (const std::bad_alloc &) {                                          \
    return fail(ctx, PBF_E_CAPACITY, "out of host memory");                 \
  }                                                                         \
  catch (const std::exception &ex) {                                        \
    return fail(ctx, PBF_E_DEVICE, std::string("internal error: ") + ex.what()); \
  }

extern "C" {

const char *pbf_version(void) { return "pbf-mi355x 1.0 (gfx950)"; }

int pbf_create(const pbf_params *params, pbf_ctx **out) { return create_ctx(params, 1, nullptr, false, out); }

int pbf_batch_create(const pbf_params *params, uint32_t nscenes, const pbf_scene_desc *scenes, pbf_ctx **out) {
  if (!scenes) return PBF_E_INVALID;
  return create_ctx(params, nscenes, scenes, true, out);
}

int pbf_destroy(pbf_ctx *c) {
  if (!c) return PBF_OK;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  clear_graphs(c);
  drain_events(c);
  for (auto e : c->event_pool) (void)hipEventDestroy(e);
  void *bufs[] = {c->pos, c->vel, c->acc, c->om, c->nrm, c->XA, c->XB, c->Xold, c->B, c->xs, c->stage4, c->ids,
                  c->cellkey, c->rank, c->tmp_key, c->tmp_src, c->ncount, c->nbr, c->nbr16, c->wfmt, c->lam, c->lam_sorted, c->counts,
                  c->offsets, c->cellsub, c->flags, c->d_scenes, c->sid, c->ranges, c->pool, c->ci, c->per_scene, c->hist,
                  c->hist_ctr, c->dbg_counts, c->dbg_ids, c->ids_save, c->lam_acc, c->acc_sorted, c->feats,
                  c->vel_sorted, c->snap_pos, c->snap_vel, c->snap_acc, c->snap_om, c->snap_nrm, c->snap_ids,
                  c->snap_lam, c->snap_lacc, c->per_scene2, c->fb_count, c->fb_thr, c->lam_tmp,
                  c->energy0, c->energy1, c->partial};
  mlp_engine_free(&c->mlp);
  if (c->decision) (void)hipHostFree(c->decision);
  for (void *b : bufs)
    if (b) (void)hipFree(b);
  for (auto p : c->scan_sums)
    if (p) (void)hipFree(p);
  for (auto p : c->scan_prefix)
    if (p) (void)hipFree(p);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return PBF_OK;
}

int pbf_reset(pbf_ctx *c) try {
  if (!c) return PBF_E_INVALID;
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipStreamSynchronize(c->stream));
  c->staged.clear();
  c->staged_serial = 0;
  c->n = 0;
  if (!c->batch) c->scenes[0].count = 0;
  for (auto &s : c->scenes) {
    s.field_kind = PBF_FIELD_NONE;
    s.field.clear();
  }
  c->have_lists = false;
  c->last_was_search = false;
  HIPCHK(hipMemset(c->counts, 0, (size_t)c->ncells * 4));
  HIPCHK(hipMemset(c->flags, 0, 8));
  HIPCHK(hipMemset(c->hist_ctr, 0, 4));
  return upload_scenes(c);
} PBF_CATCH(c)

int pbf_set_config(pbf_ctx *c, const pbf_config *cfg) try {
  if (!c || !cfg) return PBF_E_INVALID;
  if (cfg->iteration > kMaxIterations) return fail(c, PBF_E_INVALID, "iteration must be <= 4096");
  c->cfg = *cfg;
  clear_graphs(c);
  return PBF_OK;
} PBF_CATCH(c)

// Room for `cnt` more staged particles: more than the largest capacity can
// never be uploaded ('To Many Particles!' at upload), so it fails here before
// allocating host memory for them.
static int stage_room(pbf_ctx *c, uint32_t cnt, size_t *at) {
  *at = c->staged.size();
  if (*at / 4 + (size_t)cnt > kMaxCapacity) return fail(c, PBF_E_CAPACITY, "To Many Particles!");
  try {
    c->staged.resize(*at + 4 * (size_t)cnt);
  } catch (const std::bad_alloc &) {
    return fail(c, PBF_E_CAPACITY, "out of host memory for staged particles");
  }
  return PBF_OK;
}

int pbf_voxelize_cube(pbf_ctx *c, const double mn[3], const double mx[3], double jitter, uint64_t seed,
                      uint32_t *added) {
  if (!c || !mn || !mx) return PBF_E_INVALID;
  uint32_t cnt = 0;
  int rc = pbf_voxelize_cube_host(c->params.particle_radius, mn, mx, jitter, seed, 0, nullptr, 0, &cnt);
  if (rc != PBF_OK) return fail(c, rc, "voxelizeCube: invalid arguments");
  size_t at = 0;
  if ((rc = stage_room(c, cnt, &at)) != PBF_OK) return rc;
  rc = pbf_voxelize_cube_host(c->params.particle_radius, mn, mx, jitter, seed, c->staged_serial, c->staged.data() + at,
                              cnt, &cnt);
  if (rc != PBF_OK) return fail(c, rc, "voxelizeCube failed");
  c->staged_serial += cnt;
  if (added) *added = cnt;
  return PBF_OK;
}

int pbf_voxelize_sphere(pbf_ctx *c, const double center[3], double radius, double jitter, uint64_t seed,
                        uint32_t *added) {
  if (!c || !center) return PBF_E_INVALID;
  uint32_t cnt = 0;
  int rc = pbf_voxelize_sphere_host(c->params.particle_radius, center, radius, jitter, seed, 0, nullptr, 0, &cnt);
  if (rc != PBF_OK) return fail(c, rc, "voxelizeSphere: invalid arguments");
  size_t at = 0;
  if ((rc = stage_room(c, cnt, &at)) != PBF_OK) return rc;
  rc = pbf_voxelize_sphere_host(c->params.particle_radius, center, radius, jitter, seed, c->staged_serial,
                                c->staged.data() + at, cnt, &cnt);
  if (rc != PBF_OK) return fail(c, rc, "voxelizeSphere failed");
  c->staged_serial += cnt;
  if (added) *added = cnt;
  return PBF_OK;
}

int pbf_voxelize_mesh(pbf_ctx *c, const float *vertices, uint32_t nverts, const uint32_t *indices, uint32_t ntris,
                      const double world[16], double jitter, uint64_t seed, uint32_t *added) try {
  if (!c) return PBF_E_INVALID;
  uint32_t cnt = 0;
  int rc = pbf_voxelize_mesh_host(c->params.particle_radius, vertices, nverts, indices, ntris, world, jitter, seed, 0,
                                  nullptr, 0, &cnt);
  if (rc != PBF_OK) return fail(c, rc, "voxelizeMesh: invalid mesh or matrix");
  size_t at = 0;
  if ((rc = stage_room(c, cnt, &at)) != PBF_OK) return rc;
  rc = pbf_voxelize_mesh_host(c->params.particle_radius, vertices, nverts, indices, ntris, world, jitter, seed,
                              c->staged_serial, c->staged.data() + at, cnt, &cnt);
  if (rc != PBF_OK) return fail(c, rc, "voxelizeMesh failed");
  c->staged_serial += cnt;
  if (added) *added = cnt;
  return PBF_OK;
} PBF_CATCH(c)

uint32_t pbf_staged_count(const pbf_ctx *c) { return c ? (uint32_t)(c->staged.size() / 4) : 0u; }

int pbf_set_particle_position(pbf_ctx *c) try {
  if (!c) return PBF_E_INVALID;
  uint32_t count = (uint32_t)(c->staged.size() / 4);
  return pbf_set_positions(c, c->staged.data(), c->staged.size(), count);
} PBF_CATCH(c)

int pbf_set_positions(pbf_ctx *c, const float *xyzw, size_t length, uint32_t count) try {
  if (!c) return PBF_E_INVALID;
  if (length != (size_t)count * 4) return fail(c, PBF_E_LENGTH, "Illegal length of Particle Position Buffer!");
  if (count > c->params.capacity) return fail(c, PBF_E_CAPACITY, "To Many Particles!");
  if (count && !xyzw) return PBF_E_INVALID;
  return upload_positions(c, xyzw, count);
} PBF_CATCH(c)

// A field box the samplers can use: finite, hi > lo on every axis.
static bool field_box_ok(const double lo[3], const double hi[3]) {
  for (int a = 0; a < 3; ++a)
    if (!std::isfinite(lo[a]) || !std::isfinite(hi[a]) || !(hi[a] > lo[a])) return false;
  return true;
}
// The device samplers index a field array with 32-bit unsigned arithmetic.
constexpr size_t kMaxFieldValues = (size_t)1 << 31;
constexpr int32_t kMaxFieldRes = 1 << 16;

int pbf_set_cubic_field(pbf_ctx *c, uint32_t scene, const int32_t res[3], const double lo[3], const double hi[3],
                        const float *sdf, const float *vol, size_t length) try {
  if (!c || !res || !lo || !hi || !sdf || !vol) return PBF_E_INVALID;
  if (scene >= c->scenes.size()) return fail(c, PBF_E_INVALID, "scene index out of range");
  int r3[3] = {res[0], res[1], res[2]};
  for (int a = 0; a < 3; ++a)
    if (r3[a] < 1 || r3[a] > kMaxFieldRes) return fail(c, PBF_E_FORMAT, "Invalid Boundary Model File!");
  if (!field_box_ok(lo, hi)) return fail(c, PBF_E_INVALID, "field box must be finite with hi > lo");
  size_t fl = cubic_field_length(r3);
  if (fl >= kMaxFieldValues) return fail(c, PBF_E_CAPACITY, "boundary field too large");
  if (length != fl) return fail(c, PBF_E_FORMAT, "Invalid Boundary Model File!");
  auto &s = c->scenes[scene];
  try {
    s.field.assign(sdf, sdf + fl);
    s.field.insert(s.field.end(), vol, vol + fl);
  } catch (const std::bad_alloc &) {
    s.field.clear();
    s.field_kind = PBF_FIELD_NONE;
    return fail(c, PBF_E_CAPACITY, "out of host memory for the boundary field");
  }
  s.field_kind = PBF_FIELD_CUBIC;
  for (int a = 0; a < 3; ++a) { s.fres[a] = res[a]; s.flo[a] = lo[a]; s.fhi[a] = hi[a]; }
  HIPCHK(hipStreamSynchronize(c->stream));
  return upload_scenes(c);
} PBF_CATCH(c)

int pbf_set_boundary_data(pbf_ctx *c, const char *text, size_t length) try {
  if (!c || !text) return PBF_E_INVALID;
  int res[3];
  float *sdf = nullptr, *vol = nullptr;
  size_t fl = 0;
  if (parse_cdm(text, length, res, &sdf, &vol, &fl) != 0) return fail(c, PBF_E_FORMAT, "Invalid Boundary Model File!");
  const double lo[3] = {0.0, 0.0, 0.0}, hi[3] = {1.0, 1.0, 1.0};
  const int32_t r3[3] = {res[0], res[1], res[2]};
  int rc = pbf_set_cubic_field(c, 0, r3, lo, hi, sdf, vol, fl);
  free(sdf);
  free(vol);
  return rc;
} PBF_CATCH(c)

int pbf_set_volume_map(pbf_ctx *c, uint32_t scene, const int32_t res[3], const double lo[3], const double hi[3],
                       const float *sdf_vol) try {
  if (!c || !res || !lo || !hi || !sdf_vol) return PBF_E_INVALID;
  if (scene >= c->scenes.size()) return fail(c, PBF_E_INVALID, "scene index out of range");
  for (int a = 0; a < 3; ++a)
    if (res[a] < 1 || res[a] > kMaxFieldRes) return fail(c, PBF_E_INVALID, "volume map resolution must be in [1, 65536]");
  if (!field_box_ok(lo, hi)) return fail(c, PBF_E_INVALID, "field box must be finite with hi > lo");
  const size_t nn = ((size_t)res[0] + 1) * ((size_t)res[1] + 1) * ((size_t)res[2] + 1);
  if (2 * nn >= kMaxFieldValues) return fail(c, PBF_E_CAPACITY, "boundary field too large");
  auto &s = c->scenes[scene];
  try {
    s.field.assign(sdf_vol, sdf_vol + 2 * nn);
  } catch (const std::bad_alloc &) {
    s.field.clear();
    s.field_kind = PBF_FIELD_NONE;
    return fail(c, PBF_E_CAPACITY, "out of host memory for the boundary field");
  }
  s.field_kind = PBF_FIELD_TRILINEAR;
  for (int a = 0; a < 3; ++a) { s.fres[a] = res[a]; s.flo[a] = lo[a]; s.fhi[a] = hi[a]; }
  HIPCHK(hipStreamSynchronize(c->stream));
  return upload_scenes(c);
} PBF_CATCH(c)

int pbf_clear_boundary(pbf_ctx *c) try {
  if (!c) return PBF_E_INVALID;
  HIPCHK(hipStreamSynchronize(c->stream));
  for (auto &s : c->scenes) {
    s.field_kind = PBF_FIELD_NONE;
    s.field.clear();
  }
  return upload_scenes(c);
} PBF_CATCH(c)

int pbf_step(pbf_ctx *c, float dt, uint32_t substeps, uint32_t iterations) try {
  if (!c) return PBF_E_INVALID;
  if (!(dt > 0.0f) || substeps == 0) return fail(c, PBF_E_INVALID, "dt must be > 0 and substeps >= 1");
  if (iterations > kMaxIterations) return fail(c, PBF_E_INVALID, "iterations must be <= 4096");
  HIPCHK(hipSetDevice(c->device));
  if (c->n == 0) return PBF_OK;
  const uint32_t iters = iterations ? iterations : c->cfg.iteration;
  const float sdt = (float)((double)dt / (double)substeps);
  for (uint32_t s = 0; s < substeps; ++s) {
    int rc = c->warm_mode != PBF_WARM_PLAIN || c->warm_project ? run_learned_substep(c, sdt, iters)
                                                                : run_substep(c, sdt, iters, M_PLAIN, c->record_diag);
    if (rc != PBF_OK) return rc;
  }
  return PBF_OK;
} PBF_CATCH(c)

int pbf_set_warm_start(pbf_ctx *c, const float *blob, size_t length, uint32_t mode, float relax,
                       double fallback_threshold, uint32_t fallback_iterations) try {
  if (!c) return PBF_E_INVALID;
  const bool project = (mode & PBF_WARM_PROJECT) != 0, smooth = (mode & PBF_WARM_SMOOTH) != 0;
  mode &= ~(uint32_t)(PBF_WARM_PROJECT | PBF_WARM_SMOOTH);
  if (mode > PBF_WARM_RATIO) return fail(c, PBF_E_INVALID, "unknown warm-start mode");
  if (!(relax > 0.0f && relax <= 1.0f)) return fail(c, PBF_E_INVALID, "relax must be in (0, 1]");
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (blob) {
    if (length != pbf_mlp_blob_size())
      return fail(c, PBF_E_LENGTH, "network blob must hold pbf_mlp_blob_size() floats");
    for (size_t q = 0; q < length; ++q)
      if (!std::isfinite(blob[q])) return fail(c, PBF_E_FORMAT, "network blob holds a non-finite value");
    int rc = mlp_engine_set(&c->mlp, blob, c->stream);
    c->relax_set = -1.0f;
    if (rc != PBF_OK) return fail(c, rc, "could not upload the network weights");
  }
  if ((mode == PBF_WARM_LEARNED || mode == PBF_WARM_RATIO) && !c->mlp)
    return fail(c, PBF_E_STATE, "learned warm start needs network weights");
  if (c->mlp && relax != c->relax_set) {
    int rc = mlp_engine_relax(c->mlp, relax, c->stream);
    if (rc != PBF_OK) return fail(c, rc, "could not set the warm-start relaxation");
    c->relax_set = relax;
  }
  if (fallback_threshold < 0.0 || !std::isfinite(fallback_threshold))
    return fail(c, PBF_E_INVALID, "fallback_threshold must be finite and >= 0");
  if (mode != PBF_WARM_PLAIN || project) {
    int rc = ensure_learn_buffers(c);
    if (rc != PBF_OK) return rc;
  }
  c->relax_prev = relax;
  if (mode != c->warm_mode && c->lam_acc) HIPCHK(hipMemsetAsync(c->lam_acc, 0, (size_t)c->ncap * 4, c->stream));
  c->warm_mode = mode;
  c->warm_project = project;
  c->warm_smooth = smooth;
  c->fb_threshold = fallback_threshold;
  HIPCHK(hipMemcpyAsync(c->fb_thr, &c->fb_threshold, 8, hipMemcpyHostToDevice, c->stream));
  c->fb_iters = fallback_iterations;
  clear_graphs(c);
  HIPCHK(hipStreamSynchronize(c->stream));
  return PBF_OK;
} PBF_CATCH(c)

int pbf_get_tuning(const pbf_ctx *c, pbf_tuning *out) {
  if (!c || !out) return PBF_E_INVALID;
  *out = c->tuning;
  return PBF_OK;
}

int pbf_set_tuning(pbf_ctx *c, const pbf_tuning *t) {
  if (!c || !t) return PBF_E_INVALID;
  for (int k = 0; k < 8; ++k)
    if (t->walk_block[k] < 64 || t->walk_block[k] > 256 || t->walk_block[k] % 64)
      return fail(c, PBF_E_INVALID, "walk_block must be 64, 128, 192 or 256");
  if (t->list_variant >= kListVariants) return fail(c, PBF_E_INVALID, "list_variant must be 0..3");
  HIPCHK(hipStreamSynchronize(c->stream));
  c->tuning = *t;
  clear_graphs(c);
  return PBF_OK;
}

int pbf_autotune(pbf_ctx *c, float dt, uint32_t iterations, uint32_t repeats, pbf_tuning *out) try {
  if (!c) return PBF_E_INVALID;
  if (!(dt > 0.0f) || repeats == 0) return fail(c, PBF_E_INVALID, "dt must be > 0 and repeats >= 1");
  if (iterations > kMaxIterations || repeats > 1000)
    return fail(c, PBF_E_INVALID, "iterations must be <= 4096 and repeats <= 1000");
  if (c->n == 0) return fail(c, PBF_E_STATE, "no particles to tune on");
  HIPCHK(hipSetDevice(c->device));
  int rc = sync_ctx(c);
  if (rc == PBF_OK) rc = ensure_snapshot(c);
  if (rc != PBF_OK) return rc;
  const uint32_t iters = iterations ? iterations : c->cfg.iteration;
  copy_state(c, true);
  const pbf_tuning keep = c->tuning;
  const uint32_t keep_flags = c->params.flags;
  const bool keep_lists = c->have_lists, keep_search = c->last_was_search;
  c->params.flags |= PBF_FLAG_TIMING;
  for (int p = 0; p < P_NPASS; ++p) { c->pass_ms[p] = 0.0; c->pass_count[p] = 0; }
  // candidate k: every walk at block[k], the list at variant k
  const uint32_t blocks[kListVariants] = {256, 128, 64, 192};
  double ms[kListVariants][P_NPASS] = {};
  for (uint32_t k = 0; k < kListVariants && rc == PBF_OK; ++k) {
    for (int w = 0; w < 8; ++w) c->tuning.walk_block[w] = blocks[k];
    c->tuning.list_variant = k;
    copy_state(c, false);
    for (uint32_t r = 0; r <= repeats && rc == PBF_OK; ++r) {
      rc = run_substep(c, dt, iters, M_PLAIN, false);
      if (rc == PBF_OK) rc = sync_ctx(c);
      if (r == 0)  // the first substep warms the code and caches up
        for (int p = 0; p < P_NPASS; ++p) { c->pass_ms[p] = 0.0; c->pass_count[p] = 0; }
    }
    for (int p = 0; p < P_NPASS; ++p) {
      ms[k][p] = c->pass_ms[p];
      c->pass_ms[p] = 0.0;
      c->pass_count[p] = 0;
    }
  }
  c->params.flags = keep_flags;
  copy_state(c, false);
  HIPCHK(hipStreamSynchronize(c->stream));
  c->have_lists = keep_lists;
  c->last_was_search = keep_search;
  if (rc != PBF_OK) {
    c->tuning = keep;
    return rc;
  }
  pbf_tuning best = keep;
  const int walk_pass[6] = {P_LAMBDA, P_DP, P_DENSITY, P_VORT, P_XSPH, P_DIAG};
  for (int w = 0; w < 6; ++w) {
    uint32_t bk = 0;
    for (uint32_t k = 1; k < kListVariants; ++k)
      if (ms[k][walk_pass[w]] < ms[bk][walk_pass[w]]) bk = k;
    best.walk_block[w] = ms[bk][walk_pass[w]] > 0.0 ? blocks[bk] : keep.walk_block[w];
  }
  uint32_t bl = 0;
  for (uint32_t k = 1; k < kListVariants; ++k)
    if (ms[k][P_LIST] < ms[bl][P_LIST]) bl = k;
  best.list_variant = bl;
  c->tuning = best;
  clear_graphs(c);
  if (out) *out = best;
  return PBF_OK;
} PBF_CATCH(c)

int pbf_warm_start_stats(const pbf_ctx *c, uint64_t *learned_substeps, uint64_t *fallbacks) {
  if (!c) return PBF_E_INVALID;
  if (learned_substeps) *learned_substeps = c->learned_substeps;
  if (fallbacks) {
    // counted on the device by the fallback decision
    uint32_t f = 0;
    if (hipSetDevice(c->device) != hipSuccess || hipStreamSynchronize(c->stream) != hipSuccess ||
        hipMemcpy(&f, c->fb_count, 4, hipMemcpyDeviceToHost) != hipSuccess)
      return PBF_E_DEVICE;
    *fallbacks = f;
  }
  return PBF_OK;
}

int pbf_collect_training(pbf_ctx *c, float dt, uint32_t long_iterations, uint32_t iterations, float *features,
                         float *targets, uint32_t capacity, uint32_t *count) try {
  if (!c) return PBF_E_INVALID;
  if (!(dt > 0.0f) || long_iterations == 0) return fail(c, PBF_E_INVALID, "dt must be > 0 and long_iterations >= 1");
  if (long_iterations > kMaxIterations || iterations > kMaxIterations)
    return fail(c, PBF_E_INVALID, "iterations must be <= 4096");
  HIPCHK(hipSetDevice(c->device));
  if (count) *count = c->n;
  if (c->n == 0) return PBF_OK;
  if (capacity < c->n) return fail(c, PBF_E_LENGTH, "capacity is below the particle count");
  int rc = ensure_learn_buffers(c);
  if (rc == PBF_OK) rc = ensure_snapshot(c);
  if (rc != PBF_OK) return rc;
  // the long solve from the current state gives the pair ...
  copy_state(c, true);
  rc = run_substep(c, dt, long_iterations, M_COLLECT, false);
  if (rc == PBF_OK) rc = sync_ctx(c);
  if (rc != PBF_OK) return rc;
  // device rows are in slot order; return them in the caller's particle order
  const size_t n = c->n;
  std::vector<uint32_t> ids(n);
  HIPCHK(hipMemcpy(ids.data(), c->ids, n * 4, hipMemcpyDeviceToHost));
  if (features) {
    std::vector<float> f(16 * n);
    HIPCHK(hipMemcpy(f.data(), c->feats, n * 64, hipMemcpyDeviceToHost));
    for (size_t s = 0; s < n; ++s) std::memcpy(features + 16 * (size_t)ids[s], &f[16 * s], 64);
  }
  if (targets) {
    std::vector<float> t(n);
    HIPCHK(hipMemcpy(t.data(), c->lam_acc, n * 4, hipMemcpyDeviceToHost));
    for (size_t s = 0; s < n; ++s) targets[ids[s]] = t[s];
  }
  // ... and the state then advances by one substep of the selected solver
  copy_state(c, false);
  const uint32_t iters = iterations ? iterations : c->cfg.iteration;
  if (c->warm_mode != PBF_WARM_PLAIN || c->warm_project) return run_learned_substep(c, dt, iters);
  return run_substep(c, dt, iters, M_COLLECT, c->record_diag);
} PBF_CATCH(c)

int pbf_run(pbf_ctx *c) try {
  if (!c) return PBF_E_INVALID;
  return pbf_step(c, (float)c->params.time_step, 1, c->cfg.iteration ? c->cfg.iteration : 1);
} PBF_CATCH(c)

int pbf_sync(pbf_ctx *c) {
  if (!c) return PBF_E_INVALID;
  return sync_ctx(c);
}

int pbf_read_positions(pbf_ctx *c, float *xyzw, uint32_t count) try {
  if (!c || (!xyzw && count)) return PBF_E_INVALID;
  return read_sorted4(c, c->pos, c->ids, xyzw, count);
} PBF_CATCH(c)

int pbf_read_velocities(pbf_ctx *c, float *xyzw, uint32_t count) try {
  if (!c || (!xyzw && count)) return PBF_E_INVALID;
  return read_sorted4(c, c->vel, c->ids, xyzw, count, 1);
} PBF_CATCH(c)

int pbf_read_field(pbf_ctx *c, uint32_t field, float *out4, uint32_t count) try {
  if (!c || (!out4 && count)) return PBF_E_INVALID;
  const uint32_t *ids = c->last_was_search ? c->dbg_ids : c->ids;
  switch (field) {
    case PBF_FIELD_POSITION: return read_sorted4(c, c->pos, c->ids, out4, count);
    case PBF_FIELD_VELOCITY: return read_sorted4(c, c->vel, c->ids, out4, count, 1);
    case PBF_FIELD_ACCEL: return read_sorted4(c, c->acc, c->ids, out4, count);
    case PBF_FIELD_OMEGA: return read_sorted4(c, c->om, c->ids, out4, count);
    case PBF_FIELD_NORMAL: return read_sorted4(c, c->nrm, c->ids, out4, count, 1);
    case PBF_FIELD_BOUNDARY: return read_sorted4(c, c->B, c->ids, out4, count);
    case PBF_FIELD_PREDICTED: return read_sorted4(c, c->XA, ids, out4, count);
    case PBF_FIELD_LAMBDA:
    case PBF_FIELD_LAMBDA_SUM: {
      if (count != c->n) return fail(c, PBF_E_LENGTH, "readback count must equal the particle count");
      int rc = sync_ctx(c);
      if (rc != PBF_OK) return rc;
      if (field == PBF_FIELD_LAMBDA_SUM && !c->lam_acc) {  // no warm-start state yet: a zero sum
        std::memset(out4, 0, (size_t)c->n * 16);
        return PBF_OK;
      }
      launch_unsort1(c->stream, c->n, field == PBF_FIELD_LAMBDA ? c->lam : c->lam_acc, c->ids, c->stage4);
      HIPCHK(hipMemcpyAsync(out4, c->stage4, (size_t)c->n * 16, hipMemcpyDeviceToHost, c->stream));
      HIPCHK(hipStreamSynchronize(c->stream));
      return PBF_OK;
    }
    default: return fail(c, PBF_E_INVALID, "unknown field");
  }
} PBF_CATCH(c)

int pbf_write_field(pbf_ctx *c, uint32_t field, const float *in4, uint32_t count) try {
  if (!c || (!in4 && count)) return PBF_E_INVALID;
  if (field != PBF_FIELD_VELOCITY && field != PBF_FIELD_ACCEL && field != PBF_FIELD_LAMBDA &&
      field != PBF_FIELD_LAMBDA_SUM)
    return fail(c, PBF_E_INVALID, "only VELOCITY, ACCEL, LAMBDA and LAMBDA_SUM can be written");
  if (count != c->n) return fail(c, PBF_E_LENGTH, "upload count must equal the particle count");
  HIPCHK(hipSetDevice(c->device));
  int rc = sync_ctx(c);
  if (rc != PBF_OK) return rc;
  if (c->n == 0) return PBF_OK;
  if (field == PBF_FIELD_LAMBDA_SUM && (rc = ensure_learn_buffers(c)) != PBF_OK) return rc;
  // the staging rows are written by the host copy, then gathered into slot
  // order through the id map (slot i holds particle ids[i])
  HIPCHK(hipMemcpyAsync(c->stage4, in4, (size_t)c->n * 16, hipMemcpyHostToDevice, c->stream));
  switch (field) {
    case PBF_FIELD_VELOCITY: launch_sort4(c->stream, c->n, c->stage4, c->ids, c->vel, 1); break;
    case PBF_FIELD_ACCEL: launch_sort4(c->stream, c->n, c->stage4, c->ids, c->acc, 1); break;
    case PBF_FIELD_LAMBDA: launch_sort1(c->stream, c->n, c->stage4, c->ids, c->lam); break;
    default: launch_sort1(c->stream, c->n, c->stage4, c->ids, c->lam_acc); break;
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  return PBF_OK;
} PBF_CATCH(c)

uint32_t pbf_num_cells(const pbf_ctx *c) { return c ? c->ncells : 0u; }

int pbf_search(pbf_ctx *c, float dt) try {
  if (!c) return PBF_E_INVALID;
  if (!(dt > 0.0f)) return fail(c, PBF_E_INVALID, "dt must be > 0");
  HIPCHK(hipSetDevice(c->device));
  if (c->n == 0) return PBF_OK;
  LaunchCtx L = make_launch(c, dt);
  PassTimer T{c, (c->params.flags & PBF_FLAG_TIMING) != 0};
  HIPCHK(hipMemcpyAsync(c->ids_save, c->ids, (size_t)c->n * 4, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->lam_sorted, c->lam, (size_t)c->n * 4, hipMemcpyDeviceToDevice, c->stream));
  enqueue_search(c, L, T, true);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(c->dbg_ids, c->ids, (size_t)c->n * 4, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->ids, c->ids_save, (size_t)c->n * 4, hipMemcpyDeviceToDevice, c->stream));
  c->last_was_search = true;
  c->have_dbg_counts = true;
  c->have_lists = true;
  return sync_ctx(c);
} PBF_CATCH(c)

int pbf_read_cells(pbf_ctx *c, uint32_t *counts, uint32_t *offsets) try {
  if (!c) return PBF_E_INVALID;
  int rc = sync_ctx(c);
  if (rc != PBF_OK) return rc;
  std::vector<uint32_t> off((size_t)c->ncells + 1);
  HIPCHK(hipMemcpy(off.data(), c->offsets, off.size() * 4, hipMemcpyDeviceToHost));
  if (offsets) std::memcpy(offsets, off.data(), off.size() * 4);
  if (counts) {
    if (c->have_dbg_counts) {
      HIPCHK(hipMemcpy(counts, c->dbg_counts, (size_t)c->ncells * 4, hipMemcpyDeviceToHost));
    } else {
      for (uint32_t k = 0; k < c->ncells; ++k) counts[k] = off[k + 1] - off[k];
    }
  }
  return PBF_OK;
} PBF_CATCH(c)

int pbf_read_neighbors(pbf_ctx *c, uint32_t *nbr_off, uint32_t *nbr_ids, uint64_t capacity, uint64_t *pairs) try {
  if (!c) return PBF_E_INVALID;
  int rc = sync_ctx(c);
  if (rc != PBF_OK) return rc;
  if (!c->have_lists) return fail(c, PBF_E_STATE, "no neighbour lists yet (run pbf_step or pbf_search first)");
  const uint32_t n = c->n;
  std::vector<uint32_t> cnt(n ? n : 1), ids(n ? n : 1);
  if (n) {
    HIPCHK(hipMemcpy(cnt.data(), c->ncount, (size_t)n * 4, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(ids.data(), c->last_was_search ? c->dbg_ids : c->ids, (size_t)n * 4, hipMemcpyDeviceToHost));
  }
  uint32_t mx = 0;
  for (uint32_t i = 0; i < n; ++i) mx = std::max(mx, cnt[i]);
  std::vector<uint32_t> table((size_t)mx * c->ld), table16((size_t)((mx + 7) / 8) * 4 * c->ld);
  std::vector<uint32_t> wfmt((size_t)c->ncap / 64);
  if (mx) {
    HIPCHK(hipMemcpy(table.data(), c->nbr, table.size() * 4, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(table16.data(), c->nbr16, table16.size() * 4, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(wfmt.data(), c->wfmt, wfmt.size() * 4, hipMemcpyDeviceToHost));
  }
  // entry k of slot s in the wavefront's format (16-bit offsets or slots)
  auto entry = [&](uint32_t s, uint32_t k) -> uint32_t {
    if (!wfmt[s >> 6]) return table[(size_t)k * c->ld + s];
    const uint32_t w = table16[((size_t)(k >> 3) * c->ld + s) * 4 + ((k & 7u) >> 1)];
    return s - 32768u + ((k & 1u) ? (w >> 16) : (w & 0xffffu));
  };
  std::vector<uint32_t> by_id(n ? n : 1), slot_of(n ? n : 1);
  for (uint32_t i = 0; i < n; ++i) { by_id[ids[i]] = cnt[i]; slot_of[ids[i]] = i; }
  uint64_t tot = 0;
  for (uint32_t q = 0; q < n; ++q) {
    if (nbr_off) nbr_off[q] = (uint32_t)tot;
    tot += by_id[q];
  }
  if (pairs) *pairs = tot;
  if (nbr_off && tot > 0xffffffffull) return fail(c, PBF_E_CAPACITY, "neighbour offsets exceed 32 bits");
  if (nbr_off) nbr_off[n] = (uint32_t)tot;
  if (nbr_ids) {
    if (capacity < tot) return fail(c, PBF_E_CAPACITY, "neighbour id buffer too small");
    uint64_t w = 0;
    for (uint32_t q = 0; q < n; ++q) {
      uint32_t s = slot_of[q];
      for (uint32_t k = 0; k < cnt[s]; ++k) {
        const uint32_t e = entry(s, k);
        if (e >= n) return fail(c, PBF_E_STATE, "neighbour table entry out of range");
        nbr_ids[w++] = ids[e];
      }
    }
  }
  return PBF_OK;
} PBF_CATCH(c)

int pbf_density_error(pbf_ctx *c, uint32_t scene, double *mean, double *maxv) try {
  if (!c) return PBF_E_INVALID;
  int rc = sync_ctx(c);
  if (rc != PBF_OK) return rc;
  if (!c->have_lists || c->last_was_search)
    return fail(c, PBF_E_STATE, "no completed substep to evaluate (run pbf_step first)");
  const uint32_t ns = (uint32_t)c->scenes.size();
  LaunchCtx L = make_launch(c, (float)c->params.time_step);
  launch_diag(L, c->pos, c->nbr, c->ncount, c->ci, c->ranges, ns, c->per_scene, nullptr, 0, nullptr);
  HIPCHK(hipGetLastError());
  std::vector<double> ps(2 * (size_t)ns);
  HIPCHK(hipMemcpyAsync(ps.data(), c->per_scene, ps.size() * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (scene < ns) {
    uint32_t cnt = c->scenes[scene].count;
    if (mean) *mean = cnt ? ps[2 * scene] / (double)cnt : 0.0;
    if (maxv) *maxv = ps[2 * scene + 1];
  } else {
    double t = 0.0, m = 0.0;
    for (uint32_t s = 0; s < ns; ++s) { t += ps[2 * s]; m = std::max(m, ps[2 * s + 1]); }
    if (mean) *mean = c->n ? t / (double)c->n : 0.0;
    if (maxv) *maxv = m;
  }
  return PBF_OK;
} PBF_CATCH(c)

int pbf_record_density_error(pbf_ctx *c, uint32_t enable) {
  if (!c) return PBF_E_INVALID;
  HIPCHK(hipStreamSynchronize(c->stream));
  c->record_diag = enable != 0;
  HIPCHK(hipMemset(c->hist_ctr, 0, 4));
  return PBF_OK;
}

int pbf_read_density_history(pbf_ctx *c, double *mean_max, uint32_t capacity, uint32_t *count) try {
  if (!c) return PBF_E_INVALID;
  int rc = sync_ctx(c);
  if (rc != PBF_OK) return rc;
  uint32_t ctr = 0;
  HIPCHK(hipMemcpy(&ctr, c->hist_ctr, 4, hipMemcpyDeviceToHost));
  uint32_t avail = std::min(ctr, c->hist_cap);
  uint32_t take = std::min(avail, capacity);
  std::vector<double> h(2 * (size_t)c->hist_cap);
  HIPCHK(hipMemcpy(h.data(), c->hist, h.size() * 8, hipMemcpyDeviceToHost));
  // oldest first
  uint32_t start = ctr - avail;
  for (uint32_t k = 0; k < take; ++k) {
    uint32_t slot = (start + k) % c->hist_cap;
    mean_max[2 * k] = h[2 * slot];
    mean_max[2 * k + 1] = h[2 * slot + 1];
  }
  if (count) *count = take;
  return PBF_OK;
} PBF_CATCH(c)

uint32_t pbf_particle_count(const pbf_ctx *c) { return c ? c->n : 0u; }
uint32_t pbf_scene_count(const pbf_ctx *c) { return c ? (uint32_t)c->scenes.size() : 0u; }
const char *pbf_last_error(const pbf_ctx *c) {
  if (c) return c->err.c_str();
  return g_create_error.empty() ? "null context" : g_create_error.c_str();
}

int pbf_pass_times(pbf_ctx *c, double *ms, uint32_t *launches, uint32_t capacity, uint32_t *count,
                   const char **names) try {
  if (!c) return PBF_E_INVALID;
  int rc = sync_ctx(c);
  if (rc != PBF_OK) return rc;
  uint32_t k = std::min<uint32_t>(capacity, P_NPASS);
  for (uint32_t q = 0; q < k; ++q) {
    if (ms) ms[q] = c->pass_ms[q];
    if (launches) launches[q] = c->pass_count[q];
    c->pass_ms[q] = 0.0;
    c->pass_count[q] = 0;
  }
  if (count) *count = P_NPASS;
  if (names) *names = kPassNames;
  return PBF_OK;
} PBF_CATCH(c)

int pbf_copy_bandwidth(size_t bytes, uint32_t iterations, double *gbps) {
  pbf_ctx *c = nullptr;
  if (bytes < 16 || iterations == 0 || !gbps) return PBF_E_INVALID;
  bytes &= ~(size_t)15;
  void *a = nullptr, *b = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  HIPCHK(hipMalloc(&a, bytes));
  HIPCHK(hipMalloc(&b, bytes));
  HIPCHK(hipMemsetAsync(a, 1, bytes, st));
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  launch_copy_probe(st, a, b, bytes);  // warm-up
  HIPCHK(hipEventRecord(e0, st));
  for (uint32_t k = 0; k < iterations; ++k) launch_copy_probe(st, a, b, bytes);
  HIPCHK(hipEventRecord(e1, st));
  HIPCHK(hipEventSynchronize(e1));
  float ms = 0.0f;
  HIPCHK(hipEventElapsedTime(&ms, e0, e1));
  *gbps = 2.0 * (double)bytes * iterations / (ms * 1e-3) / 1e9;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(a);
  (void)hipFree(b);
  (void)hipStreamDestroy(st);
  return PBF_OK;
}

int pbf_check_device_sqrt(uint64_t *failures, uint64_t *checked) {
  if (!failures) return PBF_E_INVALID;
  const uint32_t lo = 0x0d800000u, hi = 0x7f7fffffu;  // 2^-100 .. FLT_MAX
  unsigned long long *bad = nullptr;
  hipStream_t st = nullptr;
  int rc = PBF_OK;
  auto ok = [&](hipError_t e) {
    if (e != hipSuccess && rc == PBF_OK) rc = PBF_E_DEVICE;
    return rc == PBF_OK;
  };
  if (ok(hipStreamCreateWithFlags(&st, hipStreamNonBlocking)) && ok(hipMalloc(&bad, sizeof *bad)) &&
      ok(hipMemsetAsync(bad, 0, sizeof *bad, st))) {
    launch_check_sqrt(st, lo, hi, bad);
    unsigned long long h = 0;
    if (ok(hipGetLastError()) && ok(hipMemcpyAsync(&h, bad, sizeof h, hipMemcpyDeviceToHost, st)) &&
        ok(hipStreamSynchronize(st))) {
      *failures = h;
      if (checked) *checked = (uint64_t)hi - lo + 1;
    }
  }
  if (bad) (void)hipFree(bad);
  if (st) (void)hipStreamDestroy(st);
  return rc;
}

}  // extern "C"
