// Batched execution and serving of PD reconstructions (no reference
// counterpart: the reference runs one reconstruction per process).
//
// Requests are admitted against an HBM budget, their iterates stay resident,
// and every engine step advances all resident requests of one method with a
// single task-list launch sequence per iteration (optionally replayed from a
// captured hipGraph).  Each request keeps its own step sizes, iteration count
// and adaptation schedule on the device, so a request's result does not
// depend on what else is resident.
#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <vector>

#include "solver.hpp"
#include "tasks.hpp"

namespace avionic {
bool time_space_weights(double tsw, double &ds, double &dt);
}
void avionic_set_error(const char *msg);

using namespace avionic;

namespace {

struct Slot {
  long long id = 0;
  int method = 0;
  int ds = 0;  // dual storage
  int W = 0, H = 0, C = 0, T = 0;
  long N = 0, K = 0;
  std::unique_ptr<CartesianOperator> own_op;  // built from the request's mask
  Operator *op = nullptr;                     // the request's operator
  const CartesianOperator *cart = nullptr;    // op when it runs the task-list passes
  DevBuf x, ext, y, z, d, khz, kcol, zhat, sc, slog, tlog, preg, pk, b1;
  avionic_params params{};
  int maxit = 0, k_host = 0;
  bool done = false;
  bool admitted = false;  // in its group's launch set (else waiting in the queue)
  // in-loop PD gap (avionic_request_desc::verbose / debug, params.stopPDGap)
  bool verbose = false, debug = false;
  int debugstep = 1;
  DevBuf gap_a, gap_b;
  std::vector<double> gap_log, norm_log, fid_log;
  size_t bytes = 0;
  Task task{};

  bool adapts(int k) const {
    return params.adaptSchedule == 1 || (params.adaptSchedule == 0 && (k < 10 || (k % 50) == 0));
  }
  // ictgv2.cpp:462-465
  bool wants_gap(int k) const {
    return (verbose && (k < 10 || k % 50 == 0)) || (debug && k % debugstep == 0) ||
           (params.stopPDGap > 0 && k % 20 == 0);
  }
  bool gap_enabled() const { return verbose || debug || params.stopPDGap > 0; }
};

size_t slot_bytes(int method, int ds, long N, long K, long CHW, int maxit) {
  const int P = primal_components(method), D = dual_components(method);
  size_t b = 0;
  b += 2 * sizeof(float2) * (size_t)P * N;                        // x, ext
  b += dual_bytes(ds) * (size_t)D * N;                            // y
  b += 4 * sizeof(float2) * (size_t)std::max(1L, K);              // z, d, kcol, zhat
  b += sizeof(float2) * (size_t)N;                                // khz
  b += sizeof(float2) * (size_t)CHW;                              // b1 copy
  b += 2 * sizeof(double) * (size_t)std::max(1, maxit);           // logs
  b += sizeof(double) * 3 * (size_t)((N + kRedThreads - 1) / kRedThreads + 1);  // partials
  return b;
}

// Launch variants of one iteration (TaskGrid::lanes, TaskGrid::col_first);
// every variant gives the same bits.  v = 2 [lanes == 1] + [column pass
// after the dual update].
constexpr int kVariants = 4;
inline int variant_lanes(int v) { return (v & 2) ? 1 : 2; }
inline int variant_col_first(int v) { return (v & 1) ? 0 : 1; }
inline int variant_of(int lanes, bool col_first) { return (lanes == 1 ? 2 : 0) + (col_first ? 0 : 1); }
// Variant timing: a launch set is timed once it has run this many plain
// iterations unchanged, over this many plain iterations per candidate.
constexpr int kTuneAfter = 8, kTuneSamples = 3;

struct Group {
  int method = 0;
  int ds = 0;  // dual storage
  std::vector<Slot *> slots;  // active (not done)
  bool dirty = true;
  DevBuf table;
  TaskGrid grid;
  hipGraphExec_t g_step[kVariants] = {}, g_plain[kVariants] = {};
  int variant = 0;                 // variant launched
  bool tuning = false;             // candidates still being timed
  int stable = 0, trial = 0;       // plain iterations since the rebuild / timed
  std::vector<int> cands;          // candidate variants of the launch set
  std::vector<float> samples[kVariants];
  std::array<long, 8> sig{};       // launch-set signature (decision cache key)
  void drop_graphs(int keep = -1) {
    for (int v = 0; v < kVariants; ++v) {
      if (v == keep) continue;
      if (g_step[v]) (void)hipGraphExecDestroy(g_step[v]);
      if (g_plain[v]) (void)hipGraphExecDestroy(g_plain[v]);
      g_step[v] = g_plain[v] = nullptr;
    }
  }
  TaskGrid grid_of(int v) const {
    TaskGrid gr = grid;
    gr.lanes = variant_lanes(v);
    gr.col_first = variant_col_first(v);
    return gr;
  }
};

}  // namespace

struct avionic_engine {
  // Every entry point taking an engine holds this for its whole call: calls
  // from several host threads are serialised (submit while another thread
  // steps or polls is safe).
  mutable std::mutex mu;
  avionic_engine_config cfg{};
  hipStream_t stream = nullptr;       // compute
  hipStream_t side = nullptr;         // compute, overlapped regulariser kernels
  hipStream_t copy_stream = nullptr;  // admissions
  hipEvent_t ev_in = nullptr, ev_out = nullptr, ev_req = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  long long next_id = 1;
  size_t used = 0;           // bytes of every request held (queued, running, finished)
  size_t resident_bytes = 0; // bytes of the admitted, unfinished requests
  int resident = 0;          // admitted, unfinished requests
  std::deque<Slot *> queue;  // submitted, not yet admitted (FIFO)
  long long captures = 0;    // graph captures + instantiations
  double capture_ms = 0.0;   // host time spent in them
  long long tunings = 0;     // launch sets whose variants were timed
  hipEvent_t ev_t0 = nullptr, ev_t1 = nullptr;  // variant timing
  std::map<std::array<long, 8>, int> tuned;     // signature -> fastest variant
  std::map<long long, std::unique_ptr<Slot>> slots;
  std::map<int, Group> groups;  // key = method * 3 + dual storage

  // FIFO admission: the queue head joins the launch set when the resident
  // count stays within max_batch and the resident bytes within
  // resident_bytes_cap (a request always runs when nothing else does).
  bool fits(const Slot &s) const {
    if (resident == 0) return true;
    if (resident >= cfg.max_batch) return false;
    return cfg.resident_bytes_cap == 0 || resident_bytes + s.bytes <= cfg.resident_bytes_cap;
  }
  void admit(Slot *s) {
    Group &grp = groups[s->method * 3 + s->ds];
    grp.method = s->method;
    grp.ds = s->ds;
    grp.slots.push_back(s);
    grp.dirty = true;
    s->admitted = true;
    resident += 1;
    resident_bytes += s->bytes;
  }
  void admit_queued() {
    while (!queue.empty() && fits(*queue.front())) {
      admit(queue.front());
      queue.pop_front();
    }
  }
  // a request leaves the launch set (finished or cancelled)
  void retire(Slot *s) {
    if (!s->admitted) return;
    s->admitted = false;
    resident -= 1;
    resident_bytes -= s->bytes;
  }

  ~avionic_engine() {
    for (auto &kv : groups) kv.second.drop_graphs();
    slots.clear();
    if (ev_in) (void)hipEventDestroy(ev_in);
    if (ev_out) (void)hipEventDestroy(ev_out);
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    if (ev_t0) (void)hipEventDestroy(ev_t0);
    if (ev_t1) (void)hipEventDestroy(ev_t1);
    if (ev_join) (void)hipEventDestroy(ev_join);
    if (side) (void)hipStreamDestroy(side);
    if (ev_req) (void)hipEventDestroy(ev_req);
    if (stream) (void)hipStreamDestroy(stream);
    if (copy_stream) (void)hipStreamDestroy(copy_stream);
  }

  void rebuild(Group &g) {
    TraceRange tr("avionic_engine_rebuild");
    // The previous step's iterations may still run on the stream: its graphs
    // are destroyed, and the task table rewritten, only once they are done.
    AV_HIP(hipStreamSynchronize(stream));
    g.drop_graphs();
    // tasks ordered by column kind, then row kind: each kind's kernel is
    // launched over its own span of blocks only
    std::stable_sort(g.slots.begin(), g.slots.end(), [&](const Slot *a, const Slot *b) {
      const int ka = col_kind(a->task.cg), kb = col_kind(b->task.cg);
      return ka != kb ? ka < kb : row_kind(a->task.cg) < row_kind(b->task.cg);
    });
    std::vector<Task> tasks;
    int bv = 0, bd1 = 0, bd2 = 0, bc = 0, br = 0;
    g.grid = TaskGrid{};
    bool col_seen[kColKinds] = {}, row_seen[kRowKinds] = {};
    for (Slot *s : g.slots) {
      Task t = s->task;
      t.b_vox = bv;
      t.b_dp1 = bd1;
      t.b_dp2 = bd2;
      t.b_col = bc;
      t.b_row = br;
      const int ck = col_kind(t.cg), rk = row_kind(t.cg);
      if (!col_seen[ck]) g.grid.col_lo[ck] = bc;
      if (!row_seen[rk]) g.grid.row_lo[rk] = br;
      col_seen[ck] = row_seen[rk] = true;
      bv += t.nvox_blocks;
      bd1 += t.ndp1_blocks;
      bd2 += t.ndp2_blocks;
      bc += t.ncol_blocks;
      br += t.nrow_blocks;
      g.grid.col_hi[ck] = bc;
      g.grid.row_hi[rk] = br;
      tasks.push_back(t);
    }
    g.grid.vox = bv;
    g.grid.dp1 = bd1;
    g.grid.dp2 = bd2;
    // launch variants: pairs of voxels unless configured off or no task has
    // an even width (then both lane variants launch the same kernels); the
    // column pass before or after the dual update (two streams only)
    const int lanes0 = (cfg.voxels_per_lane == 1 || bd2 == 0) ? 1 : 2;
    const bool first0 = default_col_first(g.grid);
    g.cands.clear();
    for (int lanes : {2, 1}) {
      if (lanes == 2 && lanes0 == 1) continue;
      if (lanes == 1 && lanes0 == 2 && cfg.fixed_launch) continue;
      for (bool first : {true, false}) {
        if (first != first0 && (cfg.fixed_launch || !side)) continue;
        g.cands.push_back(variant_of(lanes, first));
      }
    }
    // signature: the launch set's method, dual storage and the binary orders
    // of magnitude of its grids (launch sets of similar sizes share a
    // decision; a serving mix re-times only when its sizes move a lot)
    auto lg = [](long v) { long k = 0; while (v > 1) { v >>= 1; ++k; } return k; };
    g.sig = {g.method, g.ds, lanes0, lg(g.grid.vox), lg(g.grid.col), lg(g.grid.row), lg(bd1),
             lg(bd2)};
    g.variant = variant_of(lanes0, first0);
    g.tuning = false;
    g.stable = g.trial = 0;
    for (auto &v : g.samples) v.clear();
    if (g.cands.size() > 1) {
      auto it = tuned.find(g.sig);
      if (it != tuned.end())
        g.variant = it->second;
      else
        g.tuning = true;
    }
    g.grid.col = bc;
    g.grid.row = br;
    if (!tasks.empty()) {
      g.table.ensure(sizeof(Task) * tasks.size());
      AV_HIP(hipMemcpyAsync(g.table.p, tasks.data(), sizeof(Task) * tasks.size(),
                            hipMemcpyHostToDevice, stream));
      // `tasks` (pageable) is freed on return: the copy completes first
      AV_HIP(hipStreamSynchronize(stream));
    }
    g.dirty = false;
  }

  // PD gap after iteration k of one request (its state is final for k once
  // the engine stream reaches this point); finishes the request early when
  // the gap falls below stopPDGap: its device iteration limit drops to k + 1,
  // so the replayed iterations leave it idle.
  void evaluate_gap(Slot &s, int k) {
    GapState gs;
    gs.method = s.method;
    gs.ds = s.ds;
    gs.g = s.task.g;
    gs.rp = s.task.rp;
    gs.alpha0 = s.params.alpha0;
    gs.alpha1 = s.params.alpha1;
    gs.alpha = s.params.alpha;
    gs.lambda = s.params.lambda;
    gs.op = s.op;
    gs.x = s.x.as<float2>();
    gs.y = s.y.p;
    gs.z = s.z.as<float2>();
    gs.d = s.d.as<float2>();
    gs.khz = s.khz.as<float2>();
    gs.b1 = s.b1.as<float2>();
    gs.part_a = s.gap_a.as<double>();
    gs.part_b = s.gap_b.as<double>();
    const GapValues gv = compact_gap(gs, stream);
    s.gap_log.push_back(gv.gap);
    s.norm_log.push_back(gv.norm);
    s.fid_log.push_back(gv.fidelity);
    if (s.verbose)
      std::printf("request %lld: Normalized Primal-Dual Gap after %d iterations: %.4e\n", s.id, k,
                  gv.gap);
    if (s.params.stopPDGap > 0 && gv.gap < s.params.stopPDGap) {
      s.maxit = k + 1;
      AV_HIP(hipMemcpyAsync(reinterpret_cast<char *>(s.sc.p) + offsetof(DevScalars, maxit),
                            &s.maxit, sizeof(int), hipMemcpyHostToDevice, stream));
      AV_HIP(hipStreamSynchronize(stream));  // s.maxit is read by the copy
    }
  }

  // Data-term work of the group's non-Cartesian requests, launched from
  // launch_iteration (and captured with it): their operator's dual update
  // (z prox and K^H z), and on step iterations K(ext - x) norm partials.
  static void nc_hook(void *ctx, int phase, hipStream_t st) {
    const Group &g = *static_cast<const Group *>(ctx);
    for (Slot *s : g.slots) {
      if (s->cart) continue;
      const Task &t = s->task;
      if (phase == 0)
        s->op->dual_update(t.ext, t.b1, t.z, t.d, t.khz, t.sc, st);
      else
        (void)s->op->norm2(t.ext, t.x, t.b1, t.part_k, st);
    }
  }

  void run_iteration(Group &g, bool with_step) {
    const int n = (int)g.slots.size();
    const IterStreams st{stream, side, ev_fork, ev_join};
    IterHooks hooks;
    for (Slot *s : g.slots)
      if (!s->cart) {
        hooks.fn = &avionic_engine::nc_hook;
        hooks.ctx = &g;
        break;
      }
    // Variant timing: once the launch set has run kTuneAfter plain
    // iterations unchanged, its next plain iterations cycle through the
    // candidates, each timed with events (a host wait per timed iteration),
    // and the one with the lowest median is kept for this signature.
    int v = g.variant;
    bool timed = false;
    if (g.tuning && !with_step) {
      if (g.stable < kTuneAfter) {
        g.stable += 1;
      } else {
        v = g.cands[g.trial % g.cands.size()];
        timed = true;
      }
    }
    const TaskGrid gr = g.grid_of(v);
    if (timed) AV_HIP(hipEventRecord(ev_t0, stream));
    if (!cfg.use_graphs) {
      launch_iteration(g.method, g.ds, g.table.as<Task>(), n, gr, with_step, st, hooks);
    } else {
      hipGraphExec_t &ex = with_step ? g.g_step[v] : g.g_plain[v];
      if (!ex) {
        TraceRange tr("avionic_engine_capture");
        const auto t0 = std::chrono::steady_clock::now();
        hipGraph_t graph = nullptr;
        AV_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        launch_iteration(g.method, g.ds, g.table.as<Task>(), n, gr, with_step, st, hooks);
        AV_HIP(hipStreamEndCapture(stream, &graph));
        AV_HIP(hipGraphInstantiate(&ex, graph, nullptr, nullptr, 0));
        (void)hipGraphDestroy(graph);
        captures += 1;
        capture_ms +=
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        if (timed) AV_HIP(hipEventRecord(ev_t0, stream));  // capture time is not launch time
      }
      AV_HIP(hipGraphLaunch(ex, stream));
    }
    if (!timed) return;
    AV_HIP(hipEventRecord(ev_t1, stream));
    AV_HIP(hipEventSynchronize(ev_t1));
    float ms = 0.f;
    AV_HIP(hipEventElapsedTime(&ms, ev_t0, ev_t1));
    g.samples[v].push_back(ms);
    g.trial += 1;
    if (g.trial < kTuneSamples * (int)g.cands.size()) return;
    float best_ms = 0.f;
    int best = g.cands[0];
    for (int c : g.cands) {
      std::vector<float> t = g.samples[c];
      std::sort(t.begin(), t.end());
      const float med = t[t.size() / 2];
      if (c == g.cands[0] || med < best_ms) {
        best_ms = med;
        best = c;
      }
    }
    g.variant = best;
    g.tuning = false;
    tuned[g.sig] = best;
    tunings += 1;
    g.drop_graphs(best);
  }
};

namespace {

template <class F>
int eguard(F &&f) {
  try {
    f();
    avionic_set_error("");
    return AVIONIC_OK;
  } catch (const Error &e) {
    avionic_set_error(e.what());
    return e.code;
  } catch (const std::bad_alloc &) {
    avionic_set_error("out of memory");
    return AVIONIC_E_NOMEM;
  } catch (const std::exception &e) {
    avionic_set_error(e.what());
    return AVIONIC_E_INVALID;
  }
}

// Validate and complete a request's parameters (defaults, lambda, weights).
avionic_params request_params(const avionic_request_desc *d, const Operator &op) {
  avionic_params p{};
  if (d->params) {
    p = *d->params;
    if (p.adaptLambda) p.lambda = op.adapt_lambda(p.adaptLambdaK, p.adaptLambdaD);
  } else {
    Recon::defaults(kernel_method((int)d->method), &p);
    p.lambda = op.adapt_lambda(p.adaptLambdaK, p.adaptLambdaD);
  }
  double ds, dt;
  if (d->method == AVIONIC_TVTEMP) {
    // temporal TV: zero spatial weights, dt as configured (tv_temp.cpp:98)
    AV_CHECK(p.dt > 0.0, "dt must be positive");
    p.dx = p.dy = INFINITY;
  } else {
    AV_CHECK(time_space_weights(p.timeSpaceWeight, ds, dt), "ComputeTimeSpaceWeights: Invalid ratio.");
    p.dx = p.dy = ds;
    p.dt = dt;
  }
  if (d->method == AVIONIC_ICTGV2 || d->method == AVIONIC_ICTV) {
    AV_CHECK(time_space_weights(p.timeSpaceWeight2, ds, dt), "ComputeTimeSpaceWeights: Invalid ratio.");
    p.dx2 = p.dy2 = ds;
    p.dt2 = dt;
  }
  AV_CHECK(p.lambda > 0.0, "lambda must be positive");
  AV_CHECK(p.dualStorage >= 0 && p.dualStorage <= 2, "dualStorage must be 0, 1 or 2");
  AV_CHECK(p.adaptSchedule >= 0 && p.adaptSchedule <= 2, "adaptSchedule must be 0, 1 or 2");
  return p;
}

void check_desc(const avionic_request_desc *d) {
  AV_CHECK(d != nullptr, "request descriptor is null");
  AV_CHECK(d->method >= 0 && d->method <= AVIONIC_ICTV, "unknown method");
  AV_CHECK(d->width > 0 && d->height > 0 && d->coils > 0 && d->frames > 0, "empty request");
  AV_CHECK(d->data != nullptr && d->b1 != nullptr, "request data / b1 missing");
  check_grid(Grid3{(int)d->width, (int)d->height, (int)d->frames,
                   (long)d->width * d->height * d->frames});
}

}  // namespace

extern "C" {

int avionic_engine_create(const avionic_engine_config *cfg, avionic_engine **engine) {
  return eguard([&] {
    AV_CHECK(cfg && engine, "null argument");
    *engine = nullptr;
    auto e = std::make_unique<avionic_engine>();
    e->cfg = *cfg;
    if (e->cfg.iterations_per_tick <= 0) e->cfg.iterations_per_tick = 1;
    if (e->cfg.max_batch <= 0) e->cfg.max_batch = 1 << 20;
    if (e->cfg.hbm_budget_bytes == 0) {
      // default: 90% of the device memory free now
      size_t free_b = 0, total_b = 0;
      AV_HIP(hipMemGetInfo(&free_b, &total_b));
      e->cfg.hbm_budget_bytes = free_b / 10 * 9;
    }
    AV_HIP(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
    AV_HIP(hipStreamCreateWithFlags(&e->copy_stream, hipStreamNonBlocking));
    AV_HIP(hipStreamCreateWithFlags(&e->side, hipStreamNonBlocking));
    AV_HIP(hipEventCreate(&e->ev_t0));
    AV_HIP(hipEventCreate(&e->ev_t1));
    AV_HIP(hipEventCreateWithFlags(&e->ev_fork, hipEventDisableTiming));
    AV_HIP(hipEventCreateWithFlags(&e->ev_join, hipEventDisableTiming));
    AV_HIP(hipEventCreateWithFlags(&e->ev_req, hipEventDisableTiming));
    AV_HIP(hipEventCreateWithFlags(&e->ev_in, hipEventDisableTiming));
    AV_HIP(hipEventCreateWithFlags(&e->ev_out, hipEventDisableTiming));
    *engine = e.release();
  });
}

int avionic_engine_destroy(avionic_engine *engine) {
  return eguard([&] {
    if (!engine) return;
    (void)hipStreamSynchronize(engine->stream);
    delete engine;
  });
}

size_t avionic_request_footprint(const avionic_request_desc *d) {
  try {
    check_desc(d);
    long K = 0;
    if (d->op && d->op->op) {
      K = d->op->op->compact_elems();
    } else {
      CartPlan plan;
      plan.init((int)d->width, (int)d->height, (int)d->coils, (int)d->frames, d->mask,
                d->centered, nullptr);
      K = plan.kspace_elems();
    }
    const long N = (long)d->width * d->height * d->frames;
    const int maxit = d->params ? (int)d->params->maxIt : 500;
    const int ds = d->params ? d->params->dualStorage : 0;
    return slot_bytes(kernel_method((int)d->method), ds, N, K,
                      (long)d->coils * d->height * d->width, maxit);
  } catch (...) {
    return 0;
  }
}

int avionic_engine_submit(avionic_engine *e, const avionic_request_desc *d, long long *id) {
  return eguard([&] {
    AV_CHECK(e && id, "null argument");
    std::lock_guard<std::mutex> lock(e->mu);
    check_desc(d);
    auto s = std::make_unique<Slot>();
    s->method = kernel_method((int)d->method);
    s->W = (int)d->width; s->H = (int)d->height; s->C = (int)d->coils; s->T = (int)d->frames;
    s->N = (long)s->W * s->H * s->T;
    if (d->op) {
      AV_CHECK(d->op->op, "request operator is null");
      s->op = d->op->op;
      AV_CHECK(s->op->W == s->W && s->op->H == s->H && s->op->C == s->C && s->op->T == s->T,
               "request operator dimensions differ from the request's");
      s->cart = dynamic_cast<const CartesianOperator *>(s->op);
      AV_CHECK(s->cart || is_noncartesian(*s->op),
               "engine requests take Cartesian 2D+t or non-Cartesian operators");
    } else {
      s->own_op = std::make_unique<CartesianOperator>(s->W, s->H, s->C, s->T, d->mask, d->centered);
      s->op = s->own_op.get();
      s->cart = s->own_op.get();
    }
    s->K = s->op->compact_elems();
    s->params = request_params(d, *s->op);
    s->ds = s->params.dualStorage;
    s->maxit = (int)s->params.maxIt;
    s->verbose = d->verbose != 0;
    s->debug = d->debug != 0;
    s->debugstep = d->debugstep > 0 ? d->debugstep : 1;
    AV_CHECK(s->params.stopPDGap >= 0.0, "stopPDGap must not be negative");
    s->bytes = slot_bytes(s->method, s->ds, s->N, s->K, (long)s->C * s->H * s->W, s->maxit);
    if (e->used + s->bytes > e->cfg.hbm_budget_bytes)
      throw Error(AVIONIC_E_CAPACITY, "request does not fit the HBM budget now");
    const bool runs_now = e->queue.empty() && e->fits(*s);
    if (!runs_now && e->cfg.max_queue >= 0 && (int)e->queue.size() >= e->cfg.max_queue)
      throw Error(AVIONIC_E_CAPACITY, "admission queue is full");

    const int P = primal_components(s->method), D = dual_components(s->method);
    // inputs produced on the caller's stream; admissions copy on their own
    // stream so that they do not wait for the batch in flight
    hipStream_t st = e->copy_stream;
    AV_HIP(hipEventRecord(e->ev_req, (hipStream_t)d->stream));
    AV_HIP(hipStreamWaitEvent(st, e->ev_req, 0));
    const long N = s->N, K = std::max(1L, s->K);
    s->x.ensure(sizeof(float2) * P * N);
    s->ext.ensure(sizeof(float2) * P * N);
    s->y.ensure(dual_bytes(s->ds) * D * N);
    s->z.ensure(sizeof(float2) * K);
    s->d.ensure(sizeof(float2) * K);
    s->kcol.ensure(sizeof(float2) * K);
    s->zhat.ensure(sizeof(float2) * K);
    s->khz.ensure(sizeof(float2) * N);
    s->b1.ensure(sizeof(float2) * (size_t)s->C * s->H * s->W);
    s->slog.ensure(sizeof(double) * std::max(1, s->maxit));
    s->tlog.ensure(sizeof(double) * std::max(1, s->maxit));
    const Grid3 g{s->W, s->H, s->T, N};
    const int nvb = step_blocks(g);
    const int nrb = s->cart ? s->cart->plan().fused_row_blocks() : 0;
    const int npk = s->cart ? nrb : s->op->max_partials();
    s->preg.ensure(sizeof(double) * 2 * nvb);
    s->pk.ensure(sizeof(double) * std::max(1, npk));
    if (s->gap_enabled()) {
      const size_t np = gap_partials(g, *s->op);
      s->gap_a.ensure(sizeof(double) * np);
      s->gap_b.ensure(sizeof(double) * np);
    }
    AV_HIP(hipMemsetAsync(s->x.p, 0, s->x.bytes, st));
    AV_HIP(hipMemsetAsync(s->y.p, 0, s->y.bytes, st));
    AV_HIP(hipMemsetAsync(s->z.p, 0, s->z.bytes, st));
    if (d->x0)
      AV_HIP(hipMemcpyAsync(s->x.p, d->x0, sizeof(float2) * N, hipMemcpyDeviceToDevice, st));
    AV_HIP(hipMemcpyAsync(s->ext.p, s->x.p, s->x.bytes, hipMemcpyDeviceToDevice, st));
    AV_HIP(hipMemcpyAsync(s->b1.p, d->b1, s->b1.bytes, hipMemcpyDeviceToDevice, st));
    s->op->gather(reinterpret_cast<const float2 *>(d->data), s->d.as<float2>(), st);
    if (s->params.adaptSchedule == 2) {
      // fixed steps sigma = tau = 0.99 / ||K_full|| (FP64 power iteration on
      // the request's own operator and coil maps, ordered after their copy)
      const double w6[6] = {s->params.dx, s->params.dy, s->params.dt,
                            s->params.dx2, s->params.dy2, s->params.dt2};
      const int its = s->params.opnormIterations ? (int)s->params.opnormIterations
                                                 : kOpnormDefaultIterations;
      const double nk = operator_norm_fp64(s->method, s->W, s->H, s->T, w6, *s->op,
                                           s->b1.as<float2>(), its, st);
      AV_CHECK(nk > 0.0, "operator norm is zero");
      s->params.sigma = s->params.tau = 0.99 / nk;
    }
    DevScalars hs{};
    hs.sigma = s->params.sigma;
    hs.tau = s->params.tau;
    hs.lambda = s->params.lambda;
    hs.ratio = s->params.sigmaTauRatio;
    hs.iter = 0;
    hs.maxit = s->maxit;
    hs.schedule = s->params.adaptSchedule;
    s->sc.ensure(sizeof(DevScalars));
    AV_HIP(hipMemcpyAsync(s->sc.p, &hs, sizeof(hs), hipMemcpyHostToDevice, st));
    AV_HIP(hipStreamSynchronize(st));  // the caller may release its inputs now

    const double d6[6] = {s->params.dx, s->params.dy, s->params.dt,
                          s->params.dx2, s->params.dy2, s->params.dt2};
    Task &t = s->task;
    t.g = g;
    t.rp = make_reg_params(s->method, d6, s->params.alpha0, s->params.alpha1, s->params.alpha);
    if (s->cart) t.cg = s->cart->plan().g;  // else zero: no column / row blocks
    t.x = s->x.as<float2>();
    t.ext = s->ext.as<float2>();
    t.y = s->y.p;
    t.khz = s->khz.as<float2>();
    t.z = s->z.as<float2>();
    t.d = s->d.as<float2>();
    t.kcol = s->kcol.as<float2>();
    t.zhat = s->zhat.as<float2>();
    t.b1 = s->b1.as<float2>();
    t.sc = s->sc.as<DevScalars>();
    t.sig_log = s->slog.as<double>();
    t.tau_log = s->tlog.as<double>();
    t.part_reg = s->preg.as<double>();
    t.part_k = s->pk.as<double>();
    t.nvox_blocks = nvb;
    const bool pairs = g.W % 2 == 0;
    t.ndp1_blocks = pairs ? 0 : nvb;
    t.ndp2_blocks = pairs ? pair_blocks(g) : 0;
    t.ncol_blocks = s->cart ? t.cg.nxbf * t.cg.T : 0;
    t.nrow_blocks = nrb;
    t.npart_k = npk;
    t.schedule = s->params.adaptSchedule;

    s->id = e->next_id++;
    *id = s->id;
    e->used += s->bytes;
    if (s->maxit == 0) {
      s->done = true;
    } else if (runs_now) {
      e->admit(s.get());
    } else {
      e->queue.push_back(s.get());
    }
    e->slots[s->id] = std::move(s);
  });
}

int avionic_engine_step(avionic_engine *e, void *stream) {
  int resident = 0;
  const int rc = eguard([&] {
    TraceRange tr("avionic_engine_step");
    AV_CHECK(e, "engine is null");
    std::lock_guard<std::mutex> lock(e->mu);
    hipStream_t us = (hipStream_t)stream;
    AV_HIP(hipEventRecord(e->ev_in, us));
    AV_HIP(hipStreamWaitEvent(e->stream, e->ev_in, 0));
    e->admit_queued();
    for (auto &kv : e->groups) {
      Group &g = kv.second;
      // drop finished requests from the launch set
      auto it = std::remove_if(g.slots.begin(), g.slots.end(), [](Slot *s) { return s->done; });
      if (it != g.slots.end()) {
        g.slots.erase(it, g.slots.end());
        g.dirty = true;
      }
      if (g.slots.empty()) continue;
      if (g.dirty) e->rebuild(g);
      for (int i = 0; i < e->cfg.iterations_per_tick; ++i) {
        bool any = false, step = false;
        for (Slot *s : g.slots)
          if (s->k_host < s->maxit) {
            any = true;
            step = step || s->adapts(s->k_host);
          }
        if (!any) break;
        e->run_iteration(g, step);
        for (Slot *s : g.slots)
          if (s->k_host < s->maxit) {
            const int k = s->k_host++;
            if (s->wants_gap(k)) e->evaluate_gap(*s, k);
          }
      }
      for (Slot *s : g.slots)
        if (s->k_host >= s->maxit && !s->done) {
          s->done = true;
          e->retire(s);
        }
    }
    AV_HIP(hipEventRecord(e->ev_out, e->stream));
    AV_HIP(hipStreamWaitEvent(us, e->ev_out, 0));
    for (auto &kv : e->slots) resident += 1;
  });
  return rc == AVIONIC_OK ? resident : rc;
}

int avionic_engine_poll(avionic_engine *e, long long id, float *x_out, void *stream) {
  int result = 0;
  const int rc = eguard([&] {
    AV_CHECK(e, "engine is null");
    std::lock_guard<std::mutex> lock(e->mu);
    auto it = e->slots.find(id);
    if (it == e->slots.end()) throw Error(AVIONIC_E_NOTFOUND, "unknown request id");
    Slot *s = it->second.get();
    if (!s->done) {
      result = 0;
      return;
    }
    AV_CHECK(x_out != nullptr, "x_out is null");
    hipStream_t us = (hipStream_t)stream;
    AV_HIP(hipStreamSynchronize(e->stream));
    DevScalars hs{};
    AV_HIP(hipMemcpy(&hs, s->sc.p, sizeof(hs), hipMemcpyDeviceToHost));
    AV_HIP(hipMemcpyAsync(x_out, s->x.p, sizeof(float2) * s->N, hipMemcpyDeviceToDevice, us));
    AV_HIP(hipStreamSynchronize(us));
    e->used -= s->bytes;
    for (auto &kv : e->groups) {
      auto &v = kv.second.slots;
      auto f = std::find(v.begin(), v.end(), s);
      if (f != v.end()) {
        v.erase(f);
        kv.second.dirty = true;
      }
    }
    e->slots.erase(it);
    if (hs.diverged) throw Error(AVIONIC_E_DIVERGED, "NaN/Inf in a step-size reduction");
    result = 1;
  });
  return rc == AVIONIC_OK ? result : rc;
}

int avionic_engine_cancel(avionic_engine *e, long long id) {
  return eguard([&] {
    AV_CHECK(e, "engine is null");
    std::lock_guard<std::mutex> lock(e->mu);
    auto it = e->slots.find(id);
    if (it == e->slots.end()) throw Error(AVIONIC_E_NOTFOUND, "unknown request id");
    Slot *s = it->second.get();
    AV_HIP(hipStreamSynchronize(e->stream));
    for (auto &kv : e->groups) {
      auto &v = kv.second.slots;
      auto f = std::find(v.begin(), v.end(), s);
      if (f != v.end()) {
        v.erase(f);
        kv.second.dirty = true;
      }
    }
    auto q = std::find(e->queue.begin(), e->queue.end(), s);
    if (q != e->queue.end()) e->queue.erase(q);
    e->retire(s);
    e->used -= s->bytes;
    e->slots.erase(it);
  });
}

int avionic_engine_counters(const avionic_engine *e, avionic_engine_counts *c) {
  return eguard([&] {
    AV_CHECK(e && c, "null argument");
    std::lock_guard<std::mutex> lock(e->mu);
    c->held = (int)e->slots.size();
    c->running = e->resident;
    c->queued = (int)e->queue.size();
    c->bytes_held = e->used;
    c->bytes_running = e->resident_bytes;
    c->hbm_budget_bytes = e->cfg.hbm_budget_bytes;
    c->graph_captures = e->captures;
    c->capture_ms = e->capture_ms;
    c->launch_tunings = e->tunings;
  });
}

int avionic_engine_stats(const avionic_engine *e, int *resident, size_t *bytes) {
  return eguard([&] {
    AV_CHECK(e, "engine is null");
    std::lock_guard<std::mutex> lock(e->mu);
    if (resident) *resident = (int)e->slots.size();
    if (bytes) *bytes = e->used;
  });
}

// Average duration (ms) of each kernel family of the first resident group,
// each launched `reps` times alone between HIP events on the engine stream.
// The launches advance the resident requests' state: time on a throw-away
// engine.  Kernels whose tasks skip work at the current iteration (step-rule
// kernels outside the schedule) are timed at that cost.
int avionic_engine_time_kernels(avionic_engine *e, int reps, double *ms, const char **names,
                                size_t *n) {
  return eguard([&] {
    AV_CHECK(e && ms && n && reps > 0, "null argument");
    std::lock_guard<std::mutex> lock(e->mu);
    Group *g = nullptr;
    for (auto &kv : e->groups)
      if (!kv.second.slots.empty()) {
        g = &kv.second;
        break;
      }
    AV_CHECK(g != nullptr, "no resident requests");
    if (g->dirty) e->rebuild(*g);
    const TaskGrid gr = g->grid_of(g->variant);  // the variant the group runs
    const size_t m = std::min(*n, (size_t)kNumKernels);
    hipEvent_t a, b;
    AV_HIP(hipEventCreate(&a));
    AV_HIP(hipEventCreate(&b));
    for (size_t k = 0; k < m; ++k) {
      launch_kernel((int)k, g->method, g->ds, g->table.as<Task>(), (int)g->slots.size(), gr,
                    e->stream);  // warm
      AV_HIP(hipEventRecord(a, e->stream));
      for (int r = 0; r < reps; ++r)
        launch_kernel((int)k, g->method, g->ds, g->table.as<Task>(), (int)g->slots.size(), gr,
                      e->stream);
      AV_HIP(hipEventRecord(b, e->stream));
      AV_HIP(hipEventSynchronize(b));
      float t = 0.f;
      AV_HIP(hipEventElapsedTime(&t, a, b));
      ms[k] = (double)t / reps;
      if (names) names[k] = kernel_name((int)k);
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    *n = m;
  });
}

int avionic_engine_get_step_log(const avionic_engine *e, long long id, double *sigma, double *tau,
                                size_t *n) {
  return eguard([&] {
    AV_CHECK(e && n, "null argument");
    std::lock_guard<std::mutex> lock(e->mu);
    auto it = e->slots.find(id);
    if (it == e->slots.end()) throw Error(AVIONIC_E_NOTFOUND, "unknown request id");
    const Slot *s = it->second.get();
    AV_HIP(hipStreamSynchronize(e->stream));
    const size_t m = std::min(*n, (size_t)s->k_host);
    std::vector<double> hs(std::max<size_t>(1, s->maxit)), ht(std::max<size_t>(1, s->maxit));
    AV_HIP(hipMemcpy(hs.data(), s->slog.p, sizeof(double) * hs.size(), hipMemcpyDeviceToHost));
    AV_HIP(hipMemcpy(ht.data(), s->tlog.p, sizeof(double) * ht.size(), hipMemcpyDeviceToHost));
    double cs = s->params.sigma, ct = s->params.tau;
    for (size_t k = 0; k < m; ++k) {
      if (s->adapts((int)k)) {
        cs = hs[k];
        ct = ht[k];
      }
      if (sigma) sigma[k] = cs;
      if (tau) tau[k] = ct;
    }
    *n = m;
  });
}

int avionic_engine_get_gap_log(const avionic_engine *e, long long id, double *gap, double *norm,
                               double *fidelity, size_t *n) {
  return eguard([&] {
    AV_CHECK(e && n, "null argument");
    std::lock_guard<std::mutex> lock(e->mu);
    auto it = e->slots.find(id);
    if (it == e->slots.end()) throw Error(AVIONIC_E_NOTFOUND, "unknown request id");
    const Slot *s = it->second.get();
    const size_t m = std::min(*n, s->gap_log.size());
    for (size_t k = 0; k < m; ++k) {
      if (gap) gap[k] = s->gap_log[k];
      if (norm) norm[k] = s->norm_log[k];
      if (fidelity) fidelity[k] = s->fid_log[k];
    }
    *n = m;
  });
}

}  // extern "C"
