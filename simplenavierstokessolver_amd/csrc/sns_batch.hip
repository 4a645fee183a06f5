// sns_batch.hip -- many instances on one GPU: batches (snsg_batch_*) and the
// serving engine (snsg_server_*).
//
// Both are built on lanes.  A lane holds the instances of one shape class
// (same itot, jtot, treatment and temperature equation): every kernel of a
// step is launched once for the whole lane, grid.z = number of members, each
// member's argument blocks in a device table (ArgBatch), on the lane's own
// stream; the lane's step is recorded once as a HIP graph and replayed.  The
// kernels are the single-instance ones, so every member advances bitwise as
// it would alone.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <tuple>

#include "sns_ctx.h"

namespace {

struct Member {
  snsg_ctx *c;
  snsg_state *st;
};

// Only the predictor and the y solve depend on which of the two temperature
// buffers holds T (the parity), so they have a table per parity; parity 0 is
// the buffer each member's T was in when the tables were built.
struct Lane {
  std::vector<Member> mem;
  std::vector<double *> temp0;
  int itot = 0, jtot = 0, impx = 0, impy = 0, with_temp = 0;
  hipStream_t stream = nullptr;
  StepArgs shape{};  // argument blocks of the first member (launch configuration)
  char *d_tab = nullptr;
  size_t tab_cap = 0;
  LuFactorArgs *lu = nullptr;
  YFactorArgs *yf = nullptr;
  DtArgs *dt = nullptr;
  CorrArgs *bound = nullptr;
  PredArgs *pred[2][3] = {};
  YArgs *y[2][3][3] = {};
  PoisArgs *pois[3] = {};
  CorrArgs *corr[3] = {};
  hipGraph_t graph[2] = {};
  hipGraphExec_t exec[2] = {};
};

// shape class: size, treatment, temperature equation and launch shapes
using LaneKey = std::tuple<int, int, int, int, int, int, int, int, int, int>;

LaneKey key_of(const snsg_ctx *c) {
  const LaunchKey l = launch_key(c);
  return LaneKey(c->geo.itot, c->geo.jtot, c->treat.implicitx, c->treat.implicity,
                 c->cfg.with_temperature != 0, l.n1, l.chunk, l.cols, l.recompute, l.split);
}

// the class of a job's instance: its shapes come from the table as snsg_create
// reads it
LaneKey key_of(const snsg_config &cfg) {
  const LaunchKey l = launch_key(cfg.itot, cfg.jtot, tuned_launch(cfg.itot, cfg.jtot));
  return LaneKey(cfg.itot, cfg.jtot, cfg.implicitx != 0, cfg.implicity != 0,
                 cfg.with_temperature != 0, l.n1, l.chunk, l.cols, l.recompute, l.split);
}

int lane_init(Lane &k, const snsg_ctx *c) {
  k.itot = c->geo.itot;
  k.jtot = c->geo.jtot;
  k.impx = c->treat.implicitx;
  k.impy = c->treat.implicity;
  k.with_temp = c->cfg.with_temperature != 0;
  HIPCHECK(hipStreamCreateWithFlags(&k.stream, hipStreamNonBlocking));
  return 0;
}

void lane_drop_graphs(Lane &k) {
  for (int p = 0; p < 2; p++) {
    if (k.exec[p]) (void)hipGraphExecDestroy(k.exec[p]);
    if (k.graph[p]) (void)hipGraphDestroy(k.graph[p]);
    k.exec[p] = nullptr;
    k.graph[p] = nullptr;
  }
}

void lane_free(Lane &k) {
  if (k.stream) (void)hipStreamSynchronize(k.stream);
  lane_drop_graphs(k);
  if (k.d_tab) (void)hipFree(k.d_tab);
  if (k.stream) (void)hipStreamDestroy(k.stream);
  k = Lane();
}

// (Re)builds the argument tables for the current members, whose T is taken
// to be in parity 0 as they are now.  The lane must be idle.
int lane_build(Lane &k) {
  lane_drop_graphs(k);
  const size_t nz = k.mem.size();
  k.temp0.resize(nz);
  if (nz == 0) return 0;
  std::vector<char> host;
  auto reserve = [&](size_t bytes) {
    const size_t off = (host.size() + 255) / 256 * 256;
    host.resize(off + bytes);
    return off;
  };
  std::vector<StepArgs> sa[2];
  for (int p = 0; p < 2; p++) {
    sa[p].resize(nz);
    for (size_t m = 0; m < nz; m++) {
      snsg_state s = *k.mem[m].st;
      if (p == 1 && k.with_temp) std::swap(s.temp, s.temp_alt);
      step_args(k.mem[m].c, &s, true, sa[p][m]);
    }
  }
  for (size_t m = 0; m < nz; m++) k.temp0[m] = k.mem[m].st->temp;
  k.shape = sa[0][0];
  const size_t o_lu = reserve(nz * sizeof(LuFactorArgs));
  const size_t o_yf = reserve(nz * sizeof(YFactorArgs));
  const size_t o_dt = reserve(nz * sizeof(DtArgs));
  const size_t o_bound = reserve(nz * sizeof(CorrArgs));
  size_t o_pred[2][3], o_y[2][3][3], o_pois[3], o_corr[3];
  for (int p = 0; p < 2; p++)
    for (int rk = 0; rk < 3; rk++) {
      o_pred[p][rk] = reserve(nz * sizeof(PredArgs));
      // the three fields' blocks contiguous: one launch covers them
      const size_t oy = reserve(3 * nz * sizeof(YArgs));
      for (int f = 0; f < 3; f++) o_y[p][rk][f] = oy + f * nz * sizeof(YArgs);
    }
  for (int rk = 0; rk < 3; rk++) {
    o_pois[rk] = reserve(nz * sizeof(PoisArgs));
    o_corr[rk] = reserve(nz * sizeof(CorrArgs));
  }
  auto put = [&](size_t off, size_t m, const void *src, size_t sz) {
    memcpy(host.data() + off + m * sz, src, sz);
  };
  for (size_t m = 0; m < nz; m++) {
    put(o_lu, m, &sa[0][m].lu, sizeof(LuFactorArgs));
    put(o_yf, m, &sa[0][m].yf, sizeof(YFactorArgs));
    put(o_dt, m, &sa[0][m].dt, sizeof(DtArgs));
    put(o_bound, m, &sa[0][m].bound, sizeof(CorrArgs));
    for (int p = 0; p < 2; p++)
      for (int rk = 0; rk < 3; rk++) {
        put(o_pred[p][rk], m, &sa[p][m].pred[rk], sizeof(PredArgs));
        for (int f = 0; f < 3; f++) put(o_y[p][rk][f], m, &sa[p][m].y[rk][f], sizeof(YArgs));
      }
    for (int rk = 0; rk < 3; rk++) {
      put(o_pois[rk], m, &sa[0][m].pois[rk], sizeof(PoisArgs));
      put(o_corr[rk], m, &sa[0][m].corr[rk], sizeof(CorrArgs));
    }
  }
  if (host.size() > k.tab_cap) {
    if (k.d_tab) HIPCHECK(hipFree(k.d_tab));
    k.d_tab = nullptr;
    k.tab_cap = 0;
    const size_t cap = 2 * host.size();
    int rc = ctx_alloc((void **)&k.d_tab, cap);
    if (rc) return rc;
    k.tab_cap = cap;
  }
  HIPCHECK(hipMemcpyAsync(k.d_tab, host.data(), host.size(), hipMemcpyHostToDevice, k.stream));
  HIPCHECK(hipStreamSynchronize(k.stream));
  char *d = k.d_tab;
  k.lu = (LuFactorArgs *)(d + o_lu);
  k.yf = (YFactorArgs *)(d + o_yf);
  k.dt = (DtArgs *)(d + o_dt);
  k.bound = (CorrArgs *)(d + o_bound);
  for (int p = 0; p < 2; p++)
    for (int rk = 0; rk < 3; rk++) {
      k.pred[p][rk] = (PredArgs *)(d + o_pred[p][rk]);
      for (int f = 0; f < 3; f++) k.y[p][rk][f] = (YArgs *)(d + o_y[p][rk][f]);
    }
  for (int rk = 0; rk < 3; rk++) {
    k.pois[rk] = (PoisArgs *)(d + o_pois[rk]);
    k.corr[rk] = (CorrArgs *)(d + o_corr[rk]);
  }
  return 0;
}

// One time-loop iteration (dt decision + RK3 step) of every member, T in the
// parity-p buffer, enqueued on the lane's stream.
void lane_launch(const Lane &k, int p) {
  const int nz = (int)k.mem.size();
  const hipStream_t s = k.stream;
  const StepArgs &sh = k.shape;
  launch_dt_batch(s, k.dt, nz);
  if (k.impx) launch_lu_factor_batch(s, sh.lu, k.lu, nz);
  if (k.impy && sh.yf.L) launch_yfactor_batch(s, k.yf, nz);
  for (int rk = 0; rk < 3; rk++) {
    launch_pred_batch(s, sh.pred[rk], k.impx, k.pred[p][rk], nz);
    if (k.impy) {
      const int f0 = k.with_temp ? 0 : 1;
      if (sh.y[rk][f0].L) {
        launch_ypart_fields_batch(s, sh.y[rk][f0], k.y[p][rk][f0], nz, 3 - f0);
      } else {
        for (int f = f0; f < 3; f++)
          launch_ysolve_batch(s, kYKinds[f], sh.y[rk][f], k.y[p][rk][f], nz);
      }
    }
    for (int stage = 0; stage < 3; stage++)
      launch_poisson_batch(s, sh.pois[rk], k.pois[rk], nz, stage);
    if (sh.corr_on[rk]) launch_corr_batch(s, sh.corr[rk], k.corr[rk], nz);
  }
}

int lane_capture(Lane &k, int p) {
  HIPCHECK(hipStreamBeginCapture(k.stream, hipStreamCaptureModeThreadLocal));
  lane_launch(k, p);
  const hipError_t le = hipGetLastError();
  hipGraph_t g = nullptr;
  const hipError_t ce = hipStreamEndCapture(k.stream, &g);
  if (le != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    return set_err((int)le, "lane: launch during graph capture");
  }
  if (ce != hipSuccess) return set_err((int)ce, "hipStreamEndCapture");
  k.graph[p] = g;
  HIPCHECK(hipGraphInstantiate(&k.exec[p], g, nullptr, nullptr, 0));
  return 0;
}

// The first advective-bound partials of the members (the first dt decision
// of a run starts from the state as it is).
int lane_seed(const Lane &k) {
  launch_adv_bound_batch(k.stream, k.shape.bound, k.bound, (int)k.mem.size());
  return check_launch("lane: advective bound");
}

// One step of parity p: a graph replay, or eager launches the first time
// (which also sets up the kernels) followed by recording the graph.  The
// members' T and T' trade places (host side) once the step is enqueued.
int lane_step(Lane &k, int p, bool graphs) {
  if (graphs && k.exec[p]) {
    HIPCHECK(hipGraphLaunch(k.exec[p], k.stream));
  } else {
    lane_launch(k, p);
    int rc = check_launch("lane step");
    if (rc) return rc;
    if (graphs && (rc = lane_capture(k, p))) return rc;
  }
  if (k.with_temp)
    for (Member &m : k.mem) std::swap(m.st->temp, m.st->temp_alt);
  return 0;
}

// Parity of the members now (0: T where it was at the last build), -1 if
// they differ.
int lane_parity(const Lane &k) {
  if (!k.with_temp) return 0;
  int par = -2;
  for (size_t m = 0; m < k.mem.size(); m++) {
    const int p = k.mem[m].st->temp == k.temp0[m] ? 0 : 1;
    if (par == -2) par = p;
    else if (p != par) return -1;
  }
  return par < 0 ? 0 : par;
}

}  // namespace

// ------------------------------------------------------------ batches

struct snsg_batch {
  int device = 0;
  int flags = 0;
  std::vector<snsg_ctx *> ctx;
  std::vector<hipEvent_t> ready;  // one per member
  std::vector<Lane> lanes;
  std::vector<hipEvent_t> done;   // one per lane
  std::vector<std::vector<int>> ids;  // member indices of each lane
  bool counted = false;               // members' in_batch counts raised
};

static void batch_free(snsg_batch *b) {
  if (!b) return;
  (void)hipSetDevice(b->device);
  if (b->counted)
    for (snsg_ctx *c : b->ctx) c->in_batch--;
  for (Lane &k : b->lanes) lane_free(k);
  for (hipEvent_t e : b->ready)
    if (e) (void)hipEventDestroy(e);
  for (hipEvent_t e : b->done)
    if (e) (void)hipEventDestroy(e);
  delete b;
}

extern "C" int snsg_batch_create(int n, snsg_ctx *const *ctx, snsg_state *const *st, int flags,
                                 snsg_batch **out) {
  if (n < 1 || !ctx || !st || !out) return set_err(SNSG_E_ARG, "snsg_batch_create: arguments");
  for (int m = 0; m < n; m++) {
    if (!ctx[m] || !st[m] || !st[m]->ux)
      return set_err(SNSG_E_ARG, "snsg_batch_create: null instance or unallocated state");
    if (ctx[m]->device != ctx[0]->device)
      return set_err(SNSG_E_ARG, "snsg_batch_create: instances on different devices");
    for (int q = 0; q < m; q++)
      if (st[q] == st[m] || st[q]->ux == st[m]->ux || ctx[q] == ctx[m])
        return set_err(SNSG_E_ARG, "snsg_batch_create: an instance or state appears twice");
  }
  HIPCHECK(hipSetDevice(ctx[0]->device));
  snsg_batch *b = new snsg_batch();
  b->device = ctx[0]->device;
  b->flags = flags;
  b->ctx.assign(ctx, ctx + n);
  int rc = 0;
  auto fail = [&](int code) {
    batch_free(b);
    return code;
  };
  std::map<LaneKey, int> index;
  for (int m = 0; m < n; m++) {
    snsg_ctx *c = ctx[m];
    if (c->geo_st.jtot < 2) return fail(set_err(SNSG_E_SIZE, "snsg_batch_create: jtot < 2"));
    if ((rc = step_prepare(c))) return fail(rc);
    const LaneKey key = key_of(c);
    auto it = index.find(key);
    if (it == index.end()) {
      b->lanes.emplace_back();
      b->ids.emplace_back();
      if ((rc = lane_init(b->lanes.back(), c))) return fail(rc);
      it = index.emplace(key, (int)b->lanes.size() - 1).first;
    }
    b->lanes[it->second].mem.push_back({c, st[m]});
    b->ids[it->second].push_back(m);
  }
  for (Lane &k : b->lanes)
    if ((rc = lane_build(k))) return fail(rc);
  b->ready.assign(n, nullptr);
  b->done.assign(b->lanes.size(), nullptr);
  for (auto *v : {&b->ready, &b->done})
    for (hipEvent_t &e : *v)
      if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess)
        return fail(set_err(SNSG_E_ARG, "snsg_batch_create: event"));
  for (snsg_ctx *c : b->ctx) c->in_batch++;
  b->counted = true;
  *out = b;
  return 0;
}

extern "C" int snsg_batch_advance(snsg_batch *b, int nsteps) {
  if (!b || nsteps < 0) return set_err(SNSG_E_ARG, __func__);
  if (nsteps == 0) return 0;
  HIPCHECK(hipSetDevice(b->device));
  std::vector<int> par(b->lanes.size());
  for (size_t q = 0; q < b->lanes.size(); q++) {
    par[q] = lane_parity(b->lanes[q]);
    if (par[q] < 0)
      return set_err(SNSG_E_ARG,
                     "snsg_batch_advance: members of a class were advanced apart by an odd "
                     "number of steps");
  }
  // the lanes start after the work already queued on the members' streams
  for (size_t m = 0; m < b->ctx.size(); m++)
    HIPCHECK(hipEventRecord(b->ready[m], b->ctx[m]->stream));
  int rc;
  for (size_t q = 0; q < b->lanes.size(); q++) {
    for (int id : b->ids[q]) HIPCHECK(hipStreamWaitEvent(b->lanes[q].stream, b->ready[id], 0));
    if ((rc = lane_seed(b->lanes[q]))) return rc;
  }
  const bool graphs = !(b->flags & SNSG_BATCH_EAGER);
  for (int n = 0; n < nsteps; n++)
    for (size_t q = 0; q < b->lanes.size(); q++) {
      Lane &k = b->lanes[q];
      if ((rc = lane_step(k, par[q], graphs))) return rc;
      if (k.with_temp) par[q] ^= 1;
    }
  // later work on a member's stream comes after the batch's
  for (size_t q = 0; q < b->lanes.size(); q++) {
    HIPCHECK(hipEventRecord(b->done[q], b->lanes[q].stream));
    for (int id : b->ids[q]) HIPCHECK(hipStreamWaitEvent(b->ctx[id]->stream, b->done[q], 0));
  }
  return 0;
}

extern "C" int snsg_batch_synchronize(snsg_batch *b) {
  if (!b) return set_err(SNSG_E_ARG, __func__);
  HIPCHECK(hipSetDevice(b->device));
  for (Lane &k : b->lanes) HIPCHECK(hipStreamSynchronize(k.stream));
  return 0;
}

extern "C" int snsg_batch_classes(const snsg_batch *b, int *nclasses) {
  if (!b || !nclasses) return set_err(SNSG_E_ARG, __func__);
  *nclasses = (int)b->lanes.size();
  return 0;
}

extern "C" int snsg_batch_destroy(snsg_batch *b) {
  batch_free(b);
  return 0;
}

// ------------------------------------------------------------ serving

namespace {

enum { JOB_QUEUED = 0, JOB_ACTIVE = 1, JOB_DONE = 2, JOB_RELEASED = 3 };

struct Job {
  int id = 0;
  snsg_config cfg{};
  std::vector<double> dxf, dxc;
  uint64_t seed = 0;
  double amplitude = 0.;
  int nsteps = 0, done = 0, keep = 0;
  size_t bytes = 0;  // device bytes of its instance (snsg_instance_bytes)
  snsg_ctx *c = nullptr;
  snsg_state st{};
  hipEvent_t ready = nullptr;  // its set-up work on its own stream
  int state = JOB_QUEUED;
};

struct Pooled {
  snsg_ctx *c;
  snsg_state st;
};

// A lane of the engine: its jobs (in the order of the lane's members), the
// step in flight, and the set-up events of jobs that joined since the last
// step.
struct ServeLane {
  Lane k;
  std::vector<Job *> jobs;
  bool busy = false, dirty = false;
  int parity = 0;
  hipEvent_t t0 = nullptr, t1 = nullptr;
  std::vector<hipEvent_t> waits;
};

using Clock = std::chrono::steady_clock;

}  // namespace

struct snsg_server {
  int device = 0;
  size_t budget = 0;
  int max_active = 0;
  std::thread engine;
  // shared with the API threads (under mu)
  std::mutex mu;
  std::condition_variable cv;
  std::deque<Job *> queue;
  std::map<int, Job *> jobs;
  std::vector<Job *> released;
  int next_id = 1;
  bool stop = false;
  int failed = 0;
  std::string error;
  long submitted = 0, admitted = 0, completed = 0, steps = 0, active_peak = 0;
  double cell_steps = 0.;
  std::vector<float> lat;  // per job-step latency (ms)
  Clock::time_point t_start, t_last;
  bool started = false;
  size_t mem_used = 0, mem_peak = 0;
  int nlanes = 0;
  // engine thread only
  std::map<LaneKey, ServeLane *> lanes;
  std::multimap<LaneKey, Pooled> pool;
};

static void destroy_instance(snsg_ctx *c, snsg_state *st) {
  snsg_state_free(c, st);
  snsg_destroy(c);
}

static void engine_fail(snsg_server *s, int code, const char *what) {
  std::lock_guard<std::mutex> g(s->mu);
  if (!s->failed) {
    s->failed = code ? code : SNSG_E_ARG;
    s->error = std::string(what) + ": " + snsg_last_error();
  }
  s->cv.notify_all();
}

// Instance for job j: a pooled one of the same shape class, reconfigured, or
// a new one.  Then its initial state and first advective-bound partials, on
// its own stream; j->ready marks the end of that work.
static int engine_setup(snsg_server *s, Job *j, bool reuse) {
  int rc;
  const LaneKey key = key_of(j->cfg);
  if (reuse) {
    auto it = s->pool.find(key);
    j->c = it->second.c;
    j->st = it->second.st;
    s->pool.erase(it);
    if ((rc = ctx_configure(j->c, &j->cfg, j->dxf.data(), j->dxc.data()))) return rc;
    if ((rc = state_reset(j->c, &j->st))) return rc;
  } else {
    if ((rc = snsg_create(s->device, &j->cfg, j->dxf.data(), j->dxc.data(), &j->c))) return rc;
    if ((rc = snsg_state_alloc(j->c, &j->st))) return rc;
    if ((rc = step_prepare(j->c))) return rc;
    std::lock_guard<std::mutex> g(s->mu);
    s->mem_used += j->c->mem;
    s->mem_peak = std::max(s->mem_peak, s->mem_used);
  }
  if (j->cfg.with_temperature &&
      (rc = snsg_state_init_temperature(j->c, &j->st, j->seed, j->amplitude)))
    return rc;
  StepArgs sa;
  step_args(j->c, &j->st, true, sa);
  launch_adv_bound(j->c->stream, sa.bound);
  if ((rc = check_launch("server: advective bound"))) return rc;
  if (!j->ready) HIPCHECK(hipEventCreateWithFlags(&j->ready, hipEventDisableTiming));
  HIPCHECK(hipEventRecord(j->ready, j->c->stream));
  return 0;
}

// Admits queued jobs in arrival order while their memory fits the budget
// (freeing pooled instances of other shapes when needed).
static int engine_admit(snsg_server *s, bool &progressed) {
  for (;;) {
    Job *j;
    {
      std::lock_guard<std::mutex> g(s->mu);
      if (s->queue.empty() || s->stop) return 0;
      // slots: at most max_active jobs on the GPU at once (0: no limit)
      if (s->max_active > 0 && s->admitted - s->completed >= s->max_active) return 0;
      j = s->queue.front();
    }
    const LaneKey key = key_of(j->cfg);
    const bool reuse = s->pool.count(key) > 0;
    if (!reuse) {
      for (;;) {
        size_t used;
        {
          std::lock_guard<std::mutex> g(s->mu);
          used = s->mem_used;
        }
        if (used + j->bytes <= s->budget || s->pool.empty()) break;
        auto it = s->pool.begin();
        const size_t b = it->second.c->mem;
        destroy_instance(it->second.c, &it->second.st);
        s->pool.erase(it);
        std::lock_guard<std::mutex> g(s->mu);
        s->mem_used -= b;
      }
      std::lock_guard<std::mutex> g(s->mu);
      if (s->mem_used + j->bytes > s->budget) return 0;  // waits for running jobs to finish
    }
    {
      std::lock_guard<std::mutex> g(s->mu);
      s->queue.pop_front();
    }
    int rc = engine_setup(s, j, reuse);
    if (rc) return rc;
    ServeLane *&L = s->lanes[key];
    if (!L) {
      L = new ServeLane();
      if ((rc = lane_init(L->k, j->c))) return rc;
      HIPCHECK(hipEventCreate(&L->t0));
      HIPCHECK(hipEventCreate(&L->t1));
    }
    L->k.mem.push_back({j->c, &j->st});
    L->jobs.push_back(j);
    L->waits.push_back(j->ready);
    L->dirty = true;
    progressed = true;
    std::lock_guard<std::mutex> g(s->mu);
    j->state = JOB_ACTIVE;
    s->admitted++;
    s->active_peak = std::max(s->active_peak, s->admitted - s->completed);
    if (!s->started) {
      s->started = true;
      s->t_start = s->t_last = Clock::now();
    }
    s->nlanes = (int)s->lanes.size();
  }
}

// Collects the lanes whose step has finished; jobs that reached their step
// count leave their lane (kept for download, or pooled for reuse).
static int engine_collect(snsg_server *s, bool &progressed) {
  for (auto &kv : s->lanes) {
    ServeLane &L = *kv.second;
    if (!L.busy) continue;
    const hipError_t q = hipEventQuery(L.t1);
    if (q == hipErrorNotReady) continue;
    if (q != hipSuccess) return set_err((int)q, "server: lane step");
    float ms = 0.f;
    HIPCHECK(hipEventElapsedTime(&ms, L.t0, L.t1));
    L.busy = false;
    progressed = true;
    std::vector<Job *> left;
    std::vector<Member> mem;
    std::vector<Job *> finished;
    {
      std::lock_guard<std::mutex> g(s->mu);
      const Clock::time_point now = Clock::now();
      for (size_t m = 0; m < L.jobs.size(); m++) {
        Job *j = L.jobs[m];
        j->done++;
        s->steps++;
        s->cell_steps += (double)j->cfg.itot * (double)j->cfg.jtot;
        s->lat.push_back(ms);
        if (j->done >= j->nsteps) {
          finished.push_back(j);
        } else {
          left.push_back(j);
          mem.push_back(L.k.mem[m]);
        }
      }
      s->t_last = now;
    }
    if (finished.empty()) continue;
    L.jobs.swap(left);
    L.k.mem.swap(mem);
    L.dirty = true;
    for (Job *j : finished) {
      std::lock_guard<std::mutex> g(s->mu);
      j->state = JOB_DONE;
      s->completed++;
      if (!j->keep) {
        s->pool.emplace(kv.first, Pooled{j->c, j->st});
        j->c = nullptr;
      }
    }
    s->cv.notify_all();
  }
  // kept jobs released by the client
  std::vector<Job *> rel;
  {
    std::lock_guard<std::mutex> g(s->mu);
    rel.swap(s->released);
  }
  for (Job *j : rel) {
    if (j->c) s->pool.emplace(key_of(j->c), Pooled{j->c, j->st});
    j->c = nullptr;
    progressed = true;
  }
  return 0;
}

// Starts the next step of every idle lane that has jobs.
static int engine_launch(snsg_server *s, bool &progressed) {
  for (auto &kv : s->lanes) {
    ServeLane &L = *kv.second;
    if (L.busy || L.jobs.empty()) continue;
    int rc;
    if (L.dirty) {
      if ((rc = lane_build(L.k))) return rc;
      L.parity = 0;
      L.dirty = false;
    }
    for (hipEvent_t e : L.waits) HIPCHECK(hipStreamWaitEvent(L.k.stream, e, 0));
    L.waits.clear();
    HIPCHECK(hipEventRecord(L.t0, L.k.stream));
    if ((rc = lane_step(L.k, L.parity, true))) return rc;
    HIPCHECK(hipEventRecord(L.t1, L.k.stream));
    if (L.k.with_temp) L.parity ^= 1;
    L.busy = true;
    progressed = true;
  }
  return 0;
}

static void engine_main(snsg_server *s) {
  if (hipSetDevice(s->device) != hipSuccess) {
    engine_fail(s, SNSG_E_ARG, "server: hipSetDevice");
    return;
  }
  for (;;) {
    bool progressed = false;
    int rc = engine_collect(s, progressed);
    if (!rc) rc = engine_admit(s, progressed);
    bool stopping;
    {
      std::lock_guard<std::mutex> g(s->mu);
      stopping = s->stop;
    }
    if (!rc && !stopping) rc = engine_launch(s, progressed);
    if (rc) {
      engine_fail(s, rc, "server engine");
      stopping = true;
    }
    if (stopping) {
      bool busy = false;
      for (auto &kv : s->lanes) busy = busy || kv.second->busy;
      if (!busy || rc) return;
    }
    if (!progressed) std::this_thread::sleep_for(std::chrono::microseconds(20));
  }
}

extern "C" int snsg_server_create(int device, size_t mem_budget, int max_active,
                                  snsg_server **out) {
  if (!out || mem_budget == 0 || max_active < 0) return set_err(SNSG_E_ARG, __func__);
  int n = 0;
  HIPCHECK(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) return set_err(SNSG_E_ARG, "snsg_server_create: device");
  snsg_server *s = new snsg_server();
  s->device = device;
  s->budget = mem_budget;
  s->max_active = max_active;
  s->engine = std::thread(engine_main, s);
  *out = s;
  return 0;
}

extern "C" int snsg_server_submit(snsg_server *s, const snsg_job *job, int *id) {
  if (!s || !job || !job->dxf || !job->dxc || job->nsteps < 1)
    return set_err(SNSG_E_ARG, __func__);
  if (job->cfg.implicitx && !pred_impx_supported(job->cfg.itot))
    return set_err(SNSG_E_SIZE, "snsg_server_submit: implicit x needs itot <= 8192");
  size_t bytes = 0;
  int rc = snsg_instance_bytes(&job->cfg, &bytes);
  if (rc) return rc;
  if (bytes > s->budget)
    return set_err(SNSG_E_NOMEM, "snsg_server_submit: the job alone exceeds the memory budget");
  Job *j = new Job();
  j->cfg = job->cfg;
  j->dxf.assign(job->dxf, job->dxf + job->cfg.itot);
  j->dxc.assign(job->dxc, job->dxc + job->cfg.itot + 1);
  j->seed = job->seed;
  j->amplitude = job->amplitude;
  j->nsteps = job->nsteps;
  j->keep = job->keep;
  j->bytes = bytes;
  std::lock_guard<std::mutex> g(s->mu);
  if (s->failed) {
    delete j;
    return set_err(s->failed, s->error.c_str());
  }
  j->id = s->next_id++;
  s->jobs[j->id] = j;
  s->queue.push_back(j);
  s->submitted++;
  if (id) *id = j->id;
  return 0;
}

static Job *find_job(snsg_server *s, int id) {
  auto it = s->jobs.find(id);
  return it == s->jobs.end() ? nullptr : it->second;
}

extern "C" int snsg_server_wait(snsg_server *s, int id, double timeout_s) {
  if (!s) return set_err(SNSG_E_ARG, __func__);
  std::unique_lock<std::mutex> g(s->mu);
  Job *j = find_job(s, id);
  if (!j) return set_err(SNSG_E_ARG, "snsg_server_wait: unknown job");
  const auto until = Clock::now() + std::chrono::duration<double>(timeout_s);
  while (j->state == JOB_QUEUED || j->state == JOB_ACTIVE) {
    if (s->failed) return set_err(s->failed, s->error.c_str());
    if (s->cv.wait_until(g, until) == std::cv_status::timeout &&
        (j->state == JOB_QUEUED || j->state == JOB_ACTIVE))
      return set_err(SNSG_E_TIMEOUT, "snsg_server_wait: timed out");
  }
  return 0;
}

extern "C" int snsg_server_drain(snsg_server *s, double timeout_s) {
  if (!s) return set_err(SNSG_E_ARG, __func__);
  std::unique_lock<std::mutex> g(s->mu);
  const auto until = Clock::now() + std::chrono::duration<double>(timeout_s);
  while (s->completed < s->submitted) {
    if (s->failed) return set_err(s->failed, s->error.c_str());
    if (s->cv.wait_until(g, until) == std::cv_status::timeout && s->completed < s->submitted)
      return set_err(SNSG_E_TIMEOUT, "snsg_server_drain: timed out");
  }
  return 0;
}

extern "C" int snsg_server_status(snsg_server *s, int id, int *state, int *steps_done) {
  if (!s) return set_err(SNSG_E_ARG, __func__);
  std::lock_guard<std::mutex> g(s->mu);
  Job *j = find_job(s, id);
  if (!j) return set_err(SNSG_E_ARG, "snsg_server_status: unknown job");
  if (state) *state = j->state;
  if (steps_done) *steps_done = j->done;
  return 0;
}

// A finished job kept for download (the engine no longer touches it).
static int kept_job(snsg_server *s, int id, Job **out) {
  std::lock_guard<std::mutex> g(s->mu);
  Job *j = find_job(s, id);
  if (!j || j->state != JOB_DONE || !j->c)
    return set_err(SNSG_E_ARG, "snsg_server: job not finished or not kept");
  *out = j;
  return 0;
}

extern "C" int snsg_server_download(snsg_server *s, int id, int field, double *host) {
  if (!s || !host) return set_err(SNSG_E_ARG, __func__);
  Job *j;
  int rc = kept_job(s, id, &j);
  if (rc) return rc;
  return snsg_state_download(j->c, &j->st, field, host);
}

extern "C" int snsg_server_job_time(snsg_server *s, int id, double *time, double *dt) {
  if (!s) return set_err(SNSG_E_ARG, __func__);
  Job *j;
  int rc = kept_job(s, id, &j);
  if (rc) return rc;
  return snsg_get_time(j->c, time, dt);
}

extern "C" int snsg_server_release(snsg_server *s, int id) {
  if (!s) return set_err(SNSG_E_ARG, __func__);
  std::lock_guard<std::mutex> g(s->mu);
  Job *j = find_job(s, id);
  if (!j || j->state != JOB_DONE) return set_err(SNSG_E_ARG, "snsg_server_release: job not done");
  j->state = JOB_RELEASED;
  s->released.push_back(j);
  return 0;
}

extern "C" int snsg_server_stats_get(snsg_server *s, snsg_server_stats *out) {
  if (!s || !out) return set_err(SNSG_E_ARG, __func__);
  std::vector<float> lat;
  {
    std::lock_guard<std::mutex> g(s->mu);
    memset(out, 0, sizeof *out);
    out->submitted = s->submitted;
    out->admitted = s->admitted;
    out->completed = s->completed;
    out->queued = (long)s->queue.size();
    out->active = s->admitted - s->completed;
    out->active_peak = s->active_peak;
    out->steps = s->steps;
    out->cell_steps = s->cell_steps;
    out->busy_s = s->started ? std::chrono::duration<double>(s->t_last - s->t_start).count() : 0.;
    out->mem_used = s->mem_used;
    out->mem_peak = s->mem_peak;
    out->lanes = s->nlanes;
    lat = s->lat;
  }
  if (!lat.empty()) {
    std::sort(lat.begin(), lat.end());
    auto pct = [&](double q) {
      const size_t k = (size_t)ceil(q * (double)lat.size()) - 1;
      return (double)lat[std::min(k, lat.size() - 1)];
    };
    out->p50_ms = pct(0.50);
    out->p99_ms = pct(0.99);
    out->max_ms = lat.back();
  }
  return 0;
}

extern "C" int snsg_server_reset_stats(snsg_server *s) {
  if (!s) return set_err(SNSG_E_ARG, __func__);
  std::lock_guard<std::mutex> g(s->mu);
  s->steps = 0;
  s->cell_steps = 0.;
  s->lat.clear();
  s->t_start = s->t_last = Clock::now();
  s->started = s->admitted > s->completed;
  return 0;
}

extern "C" int snsg_server_destroy(snsg_server *s) {
  if (!s) return 0;
  {
    std::lock_guard<std::mutex> g(s->mu);
    s->stop = true;
  }
  if (s->engine.joinable()) s->engine.join();
  (void)hipSetDevice(s->device);
  for (auto &kv : s->lanes) {
    lane_free(kv.second->k);
    if (kv.second->t0) (void)hipEventDestroy(kv.second->t0);
    if (kv.second->t1) (void)hipEventDestroy(kv.second->t1);
    delete kv.second;
  }
  for (auto &kv : s->pool) destroy_instance(kv.second.c, &kv.second.st);
  for (auto &kv : s->jobs) {
    Job *j = kv.second;
    if (j->c) destroy_instance(j->c, &j->st);
    if (j->ready) (void)hipEventDestroy(j->ready);
    delete j;
  }
  delete s;
  return 0;
}
