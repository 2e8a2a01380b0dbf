// clv_server.hip -- the serving loop: continuous batching of Langevin runs.
//
// A worker thread owns the device.  Clients submit runs into an inbox; the
// worker moves them to its own waiting list and admits them into free slots of
// an idle pool of their shape, creating a pool when the HBM budget allows (and
// releasing empty pools of other shapes when it does not).  Each pool is a
// bucket of equal-shape chain slots with its own HIP stream, so pools of
// different shapes run concurrently on the GPU.  Pools tick independently: an
// idle pool with active slots is given k steps (the smallest remaining step
// count of its runs, capped so that a tick lasts about kTickTargetMs), the
// worker polls the tick's completion event, and a pool whose tick ended is
// accounted, its finished and cancelled runs read back and freed, and it is
// refilled and relaunched without waiting for the other pools.  Every run stops
// exactly after its own nsteps.  Nothing a run computes depends on its slot or
// on its neighbours: noise is keyed by the run's seed and step, reductions are
// per chain and in a fixed order.
//
// Locking: the client mutex guards the inbox, the run records' results and the
// counters, and is never held across device work (pool creation, admission,
// tick launches, read-back): clients that submit, poll or wait never queue
// behind an allocation or a stream synchronisation.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/clv.h"
#include "clv_internal.h"

using namespace clvi;

namespace {

double now_s() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

struct Run {
  int64_t id = 0;
  clv_request req{};
  clv_result res{};             // guarded by the server mutex
  std::atomic<bool> cancel{false};
  // worker only
  int pool = -1, slot = -1;
  int64_t remaining = 0;
  int64_t failed_gen = -1;      // generation at which creating its pool last failed
};

struct Pool {
  clv_ctx ctx;                 // one bucket, own stream
  std::vector<Run *> slot_run;  // run per slot, nullptr = free
  int nactive = 0;
  bool alive = false;
  bool busy = false;            // a tick is in flight on the stream
  int64_t k_inflight = 0;       // steps of that tick
  double ms_per_step = 0.0;     // measured by the last tick (0: unknown)
  hipEvent_t ev_start = nullptr, ev_done = nullptr;
};

bool same_shape(const clv_shape &a, const clv_shape &b) {
  return a.nx == b.nx && a.ny == b.ny && a.nz == b.nz && a.nt == b.nt;
}

constexpr int64_t kTickMax = 100;          // steps per tick at most
constexpr double kTickTargetMs = 2.0;      // aimed tick duration
constexpr int64_t kTickFirst = 10;         // steps of a pool's first tick
constexpr int64_t kPoolTargetBytes = 4LL << 30;
constexpr int kPoolMaxSlots = 64;

}  // namespace

struct clv_server {
  int device = 0;
  int precision = CLV_FP32;
  int64_t budget = 0;
  mutable std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Run *> inbox;                       // guarded by mu
  std::map<int64_t, std::unique_ptr<Run>> runs;  // guarded by mu
  int64_t next_id = 1;
  bool stop = false;
  std::string err;                               // guarded by mu
  clv_server_info info{};                        // guarded by mu
  // worker only
  std::deque<Run *> waiting;
  std::vector<std::unique_ptr<Pool>> pools;
  int64_t bytes_in_use = 0;
  int64_t gen = 0;              // bumps whenever a slot or pool memory is freed
  double setup_s = 0.0;         // time spent creating pools
  std::thread worker;
};

namespace {

int64_t pool_bytes(const Pool &p) { return p.ctx.bytes; }

void set_error(clv_server *s, const std::string &msg) {
  std::lock_guard<std::mutex> g(s->mu);
  s->err = msg;
}

// Create a pool for `shape` with as many slots as the target size and the
// remaining budget allow.  nullptr with *err = 0 when the budget has no room
// (wait for runs to finish), with *err < 0 when the device refused.
Pool *make_pool(clv_server *s, const clv_shape &shape, int *err) {
  *err = 0;
  const int64_t per = chain_bytes(shape, s->precision);
  const int64_t room = s->budget - s->bytes_in_use;
  int cap = int(std::min<int64_t>(kPoolMaxSlots, std::max<int64_t>(1, kPoolTargetBytes / per)));
  cap = int(std::min<int64_t>(cap, room / per));
  if (cap < 1) return nullptr;
  const double t0 = now_s();
  std::unique_ptr<Pool> p(new Pool());
  p->ctx.device = s->device;
  p->ctx.precision = s->precision;
  if (hipStreamCreateWithFlags(&p->ctx.stream, hipStreamNonBlocking) != hipSuccess) {
    (void)hipGetLastError();
    *err = CLV_ERR_HIP;
    set_error(s, "clv_server: hipStreamCreate failed");
    return nullptr;
  }
  p->ctx.buckets.emplace_back();
  Bucket &b = p->ctx.buckets.back();
  b.shape = shape;
  b.nch = cap;
  const int rc = setup_bucket(&p->ctx, b);
  if (rc != CLV_OK) {
    free_bucket(&p->ctx, b);
    (void)hipStreamDestroy(p->ctx.stream);
    *err = rc;
    set_error(s, "clv_server: creating a pool of " + std::to_string(cap) + " slots: " + p->ctx.err);
    return nullptr;
  }
  // setup_bucket zeroes the parameter records, so every slot starts inactive
  b.schedule = 1;
  b.host_active.assign(b.nch, 0);
  p->slot_run.assign(cap, nullptr);
  if (hipEventCreate(&p->ev_start) != hipSuccess || hipEventCreate(&p->ev_done) != hipSuccess) {
    (void)hipGetLastError();
    free_bucket(&p->ctx, b);
    (void)hipStreamDestroy(p->ctx.stream);
    *err = CLV_ERR_HIP;
    set_error(s, "clv_server: hipEventCreate failed");
    return nullptr;
  }
  p->alive = true;
  s->bytes_in_use += pool_bytes(*p);
  s->pools.push_back(std::move(p));
  s->setup_s += now_s() - t0;
  return s->pools.back().get();
}

void free_pool(clv_server *s, Pool &p) {
  if (!p.alive) return;
  (void)hipStreamSynchronize(p.ctx.stream);
  s->bytes_in_use -= pool_bytes(p);
  for (Bucket &b : p.ctx.buckets) free_bucket(&p.ctx, b);
  (void)hipStreamDestroy(p.ctx.stream);
  if (p.ev_start) (void)hipEventDestroy(p.ev_start);
  if (p.ev_done) (void)hipEventDestroy(p.ev_done);
  p.ev_start = p.ev_done = nullptr;
  p.ctx.buckets.clear();
  p.alive = false;
  p.busy = false;
  s->gen += 1;
}

bool any_pool_running(const clv_server *s) {
  for (const auto &p : s->pools)
    if (p->alive && (p->busy || p->nactive > 0)) return true;
  return false;
}

constexpr int kPlaced = 1, kWait = 0;  // place(): negative = fail the run with that code

// Find a free slot for run r (idle pool of its shape, new pool, or after
// releasing empty pools of other shapes).  kWait if it has to wait: a pool
// with a tick in flight takes new runs when the tick has ended, and runs in
// flight free budget when they finish.  A device allocation failure fails the
// run when nothing in flight could free memory; otherwise the run waits and
// the allocation is retried only after some memory has been freed.
int place(clv_server *s, Run &r, int *pool, int *slot) {
  const clv_shape &shape = r.req.shape;
  bool shape_busy = false;
  for (size_t k = 0; k < s->pools.size(); ++k) {
    Pool &p = *s->pools[k];
    if (!p.alive || !same_shape(p.ctx.buckets[0].shape, shape)) continue;
    if (p.busy) {
      for (Run *q : p.slot_run) shape_busy |= (q == nullptr);
      continue;
    }
    for (size_t i = 0; i < p.slot_run.size(); ++i)
      if (p.slot_run[i] == nullptr) {
        *pool = int(k);
        *slot = int(i);
        return kPlaced;
      }
  }
  if (shape_busy) return kWait;        // a slot frees up at the end of that pool's tick
  if (r.failed_gen == s->gen) return kWait;  // nothing freed since the last refusal
  int err = 0;
  Pool *np = make_pool(s, shape, &err);
  if (!np) {
    // release idle pools of other shapes, then try once more
    bool freed = false;
    for (auto &pp : s->pools)
      if (pp->alive && !pp->busy && pp->nactive == 0 && !same_shape(pp->ctx.buckets[0].shape, shape)) {
        free_pool(s, *pp);
        freed = true;
      }
    if (freed) np = make_pool(s, shape, &err);
  }
  if (!np) {
    if (err == 0) return kWait;                 // budget: runs in flight will free it
    if (!any_pool_running(s)) return err;       // nothing will ever free device memory
    r.failed_gen = s->gen;
    return kWait;
  }
  for (size_t k = 0; k < s->pools.size(); ++k)
    if (s->pools[k].get() == np) {
      *pool = int(k);
      *slot = 0;
      return kPlaced;
    }
  return kWait;
}

// A run that never reached a slot ends with `status` (worker; takes the lock).
void conclude_unplaced(clv_server *s, Run &r, int status, const std::string &msg) {
  std::lock_guard<std::mutex> g(s->mu);
  r.res.status = status;
  r.res.t_done = now_s();
  if (!msg.empty()) s->err = msg;
  s->info.queued -= 1;
  if (status == CLV_ERR_CANCELLED) s->info.cancelled += 1;
  else s->info.failed += 1;
  s->cv_done.notify_all();
}

// Read back and free the slots of pool p (idle) whose runs have finished their
// steps or were cancelled: observables (a finished run that reached no
// evaluation step is evaluated once, on its final fields) and step
// statistics, then one batched deactivation.  Device work without the lock;
// the results are published under it.
void reap(clv_server *s, Pool &p) {
  if (!p.alive || p.busy) return;
  std::vector<Run *> out;
  for (Run *r : p.slot_run)
    if (r && (r->remaining <= 0 || r->cancel.load())) out.push_back(r);
  if (out.empty()) return;
  // a run that completed its steps by this boundary ends as done, cancelled or not
  std::vector<char> cancelled(out.size());
  for (size_t i = 0; i < out.size(); ++i) cancelled[i] = out[i]->remaining > 0;
  Bucket &b = p.ctx.buckets[0];
  std::vector<clv_observables> obs(out.size());
  std::vector<ChainState> st(out.size());
  std::vector<int> rcs(out.size(), CLV_OK), locals;
  for (size_t i = 0; i < out.size(); ++i) {
    Run &r = *out[i];
    int rc = CLV_OK;
    std::memset(&obs[i], 0, sizeof(obs[i]));
    if (!cancelled[i]) {
      rc = chain_read_observables(&p.ctx, b, r.slot, &obs[i], nullptr, nullptr, nullptr);
      if (rc == CLV_OK && obs[i].count == 0) {
        chain_evaluate(&p.ctx, b, r.slot);
        rc = chain_read_observables(&p.ctx, b, r.slot, &obs[i], nullptr, nullptr, nullptr);
      }
    }
    if (rc == CLV_OK) rc = read_state(&p.ctx, b, r.slot, &st[i]);
    rcs[i] = rc;
    locals.push_back(r.slot);
  }
  const int rcd = chains_deactivate(&p.ctx, b, locals);
  std::lock_guard<std::mutex> g(s->mu);
  for (size_t i = 0; i < out.size(); ++i) {
    Run &r = *out[i];
    const int rc = rcs[i] != CLV_OK ? rcs[i] : rcd;
    if (rc == CLV_OK) {
      r.res.obs = obs[i];
      stats_from_state(st[i], &r.res.stats);
    }
    if (rc != CLV_OK) {
      r.res.status = rc;
      s->info.failed += 1;
      s->err = p.ctx.err;
    } else if (cancelled[i]) {
      r.res.status = CLV_ERR_CANCELLED;
      s->info.cancelled += 1;
    } else {
      r.res.status = CLV_RUN_DONE;
      s->info.completed += 1;
    }
    r.res.t_done = now_s();
    p.slot_run[r.slot] = nullptr;
    p.nactive -= 1;
    s->info.running -= 1;
  }
  s->gen += 1;
  s->cv_done.notify_all();
}

void worker_main(clv_server *s) {
  (void)hipSetDevice(s->device);
  double busy_since = -1.0;
  for (;;) {
    // ---- new submissions (the only part under the lock) ----
    {
      std::unique_lock<std::mutex> lk(s->mu);
      const bool any_active = any_pool_running(s);
      if (!any_active && s->waiting.empty() && s->inbox.empty()) {
        if (busy_since >= 0) {
          s->info.busy_seconds += now_s() - busy_since;
          busy_since = -1.0;
        }
        s->cv_work.wait(lk, [&] { return s->stop || !s->inbox.empty(); });
      }
      if (s->stop) break;
      while (!s->inbox.empty()) {
        s->waiting.push_back(s->inbox.front());
        s->inbox.pop_front();
      }
    }
    // ---- finished and cancelled runs of idle pools free their slots ----
    for (auto &pp : s->pools) reap(s, *pp);
    // ---- admission (tick boundary) ----
    std::vector<Run *> admitted;
    for (auto it = s->waiting.begin(); it != s->waiting.end();) {
      Run &r = **it;
      if (r.cancel.load()) {
        conclude_unplaced(s, r, CLV_ERR_CANCELLED, "");
        it = s->waiting.erase(it);
        continue;
      }
      const int64_t per = chain_bytes(r.req.shape, s->precision);
      if (per > s->budget) {
        conclude_unplaced(s, r, CLV_ERR_BUDGET,
                          "clv_server: run " + std::to_string(r.id) + " needs " + std::to_string(per) +
                              " bytes, more than the HBM budget");
        it = s->waiting.erase(it);
        continue;
      }
      int pool = -1, slot = -1;
      const int pr = place(s, r, &pool, &slot);
      if (pr == kWait) {
        ++it;
        continue;
      }
      if (pr < 0) {
        std::string msg;
        {
          std::lock_guard<std::mutex> g(s->mu);
          msg = s->err;
        }
        conclude_unplaced(s, r, pr, msg);
        it = s->waiting.erase(it);
        continue;
      }
      r.pool = pool;
      r.slot = slot;
      r.remaining = r.req.nsteps;
      s->pools[pool]->slot_run[slot] = &r;
      s->pools[pool]->nactive += 1;
      {
        std::lock_guard<std::mutex> g(s->mu);
        r.res.status = CLV_RUN_RUNNING;
        r.res.t_admit = now_s();
        s->info.queued -= 1;
        s->info.running += 1;
      }
      admitted.push_back(&r);
      it = s->waiting.erase(it);
    }
    for (Run *rp : admitted) {
      Run &r = *rp;
      Pool &p = *s->pools[r.pool];
      const int rc = chain_admit(&p.ctx, p.ctx.buckets[0], r.slot, r.req.params, r.req.eval_every,
                                 r.req.eval_from, r.req.n0, r.req.var);
      if (rc != CLV_OK) {
        std::lock_guard<std::mutex> g(s->mu);
        r.res.status = rc;
        r.res.t_done = now_s();
        s->err = p.ctx.err;
        p.slot_run[r.slot] = nullptr;
        p.nactive -= 1;
        s->info.running -= 1;
        s->info.failed += 1;
        s->cv_done.notify_all();
      }
    }
    // zero-step runs and runs cancelled during admission end before any tick
    if (!admitted.empty())
      for (auto &pp : s->pools) reap(s, *pp);
    {
      std::lock_guard<std::mutex> g(s->mu);
      s->info.pools = 0;
      for (auto &p : s->pools) s->info.pools += p->alive ? 1 : 0;
      s->info.bytes_in_use = s->bytes_in_use;
      s->info.pool_setup_seconds = s->setup_s;
    }
    // ---- launch a tick on every idle pool with active slots ----
    bool launched = false;
    for (auto &pp : s->pools) {
      Pool &p = *pp;
      if (!p.alive || p.busy || p.nactive == 0) continue;
      int64_t k = p.ms_per_step > 0.0 ? int64_t(kTickTargetMs / p.ms_per_step) : kTickFirst;
      k = std::max<int64_t>(1, std::min(k, kTickMax));
      for (Run *r : p.slot_run)
        if (r) k = std::min(k, r->remaining);
      Bucket &b = p.ctx.buckets[0];
      auto_plan(b, p.nactive);
      (void)hipEventRecord(p.ev_start, p.ctx.stream);
      for (int64_t n = 0; n < k; ++n) step_bucket(&p.ctx, b, -1);
      (void)hipEventRecord(p.ev_done, p.ctx.stream);
      p.busy = true;
      p.k_inflight = k;
      launched = true;
    }
    if (launched && busy_since < 0) busy_since = now_s();
    // ---- account the pools whose tick has ended ----
    bool progressed = false;
    double su = 0.0;
    for (auto &pp : s->pools) {
      Pool &p = *pp;
      if (!p.alive || !p.busy) continue;
      const hipError_t q = hipEventQuery(p.ev_done);
      if (q == hipErrorNotReady) continue;
      float ms = 0.f;
      if (q == hipSuccess && hipEventElapsedTime(&ms, p.ev_start, p.ev_done) == hipSuccess)
        p.ms_per_step = double(ms) / double(p.k_inflight);
      p.busy = false;
      progressed = true;
      su += double(p.nactive) * double(p.ctx.buckets[0].S) * double(p.k_inflight);
      for (Run *r : p.slot_run)
        if (r) r->remaining -= p.k_inflight;
      reap(s, p);
    }
    if (progressed) {
      std::lock_guard<std::mutex> g(s->mu);
      s->info.site_updates += su;
    }
    if (!progressed && !launched && admitted.empty())
      std::this_thread::sleep_for(std::chrono::microseconds(50));
  }
  std::lock_guard<std::mutex> g(s->mu);
  if (busy_since >= 0) s->info.busy_seconds += now_s() - busy_since;
}

}  // namespace

extern "C" {

double clv_server_clock(void) { return now_s(); }

// The worker may set the message at any time: the caller gets a copy, valid
// until this thread's next call.
const char *clv_server_error(const clv_server *srv) {
  static thread_local std::string copy;
  if (!srv) return "";
  std::lock_guard<std::mutex> g(srv->mu);
  copy = srv->err;
  return copy.c_str();
}

int clv_server_create(int32_t device, int32_t precision, int64_t hbm_budget_bytes,
                      clv_server **out) {
  if (!out) return CLV_ERR_ARG;
  *out = nullptr;
  if (precision != CLV_FP32 && precision != CLV_FP64) return CLV_ERR_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0 || device < 0 || device >= ndev)
    return CLV_ERR_HIP;
  if (hipSetDevice(device) != hipSuccess) return CLV_ERR_HIP;
  if (hbm_budget_bytes <= 0) {
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) return CLV_ERR_HIP;
    hbm_budget_bytes = int64_t(double(fr) * 0.8);
  }
  clv_server *s = new clv_server();
  s->device = device;
  s->precision = precision;
  s->budget = hbm_budget_bytes;
  s->info.budget_bytes = hbm_budget_bytes;
  s->worker = std::thread(worker_main, s);
  *out = s;
  return CLV_OK;
}

int clv_server_destroy(clv_server *s) {
  if (!s) return CLV_OK;
  {
    std::lock_guard<std::mutex> g(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  if (s->worker.joinable()) s->worker.join();
  (void)hipSetDevice(s->device);
  for (auto &p : s->pools) free_pool(s, *p);
  delete s;
  return CLV_OK;
}

int clv_server_submit(clv_server *s, const clv_request *req, int64_t *id) {
  const double t_arrival = now_s();  // before any lock: blocked time counts as latency
  if (!s || !req) return CLV_ERR_ARG;
  const clv_shape &sh = req->shape;
  const clv_params &pr = req->params;
  if (sh.nx <= 0 || sh.ny <= 0 || sh.nz <= 0 || sh.nt < 2 || !supported_x(sh.nx) ||
      int64_t(sh.nx) * sh.ny * sh.nz * sh.nt >= (int64_t(1) << 31) || req->nsteps < 0 ||
      pr.kin_dims != 1 || !(pr.DT > 0.0) || !(pr.eps > 0.0) || !(pr.a_x > 0.0) ||
      (pr.conj_drift_mode != CLV_CONJ_DRIFT_REFERENCE && pr.conj_drift_mode != CLV_CONJ_DRIFT_ACTION) ||
      !(req->n0 >= 0) || !(req->var >= 0)) {
    std::lock_guard<std::mutex> g(s->mu);
    s->err = "clv_server_submit: invalid request (shape, nsteps, kin_dims, DT, eps, a_x, "
             "conj_drift_mode, n0, var)";
    return CLV_ERR_ARG;
  }
  std::unique_ptr<Run> r(new Run());
  r->req = *req;
  r->res.status = CLV_RUN_QUEUED;
  r->res.t_submit = t_arrival;
  Run *rp = r.get();
  std::lock_guard<std::mutex> g(s->mu);
  r->id = s->next_id++;
  r->res.id = r->id;
  s->runs[rp->id] = std::move(r);
  s->inbox.push_back(rp);
  s->info.submitted += 1;
  s->info.queued += 1;
  if (id) *id = rp->id;
  s->cv_work.notify_all();
  return CLV_OK;
}

int clv_server_poll(clv_server *s, int64_t id, clv_result *out) {
  if (!s || !out) return CLV_ERR_ARG;
  std::lock_guard<std::mutex> g(s->mu);
  auto it = s->runs.find(id);
  if (it == s->runs.end()) return CLV_ERR_NOTFOUND;
  *out = it->second->res;
  return CLV_OK;
}

int clv_server_wait(clv_server *s, int64_t id, double timeout_s, clv_result *out) {
  if (!s || !out) return CLV_ERR_ARG;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->runs.find(id);
  if (it == s->runs.end()) return CLV_ERR_NOTFOUND;
  const auto deadline = std::chrono::steady_clock::now() +
                        std::chrono::duration_cast<std::chrono::steady_clock::duration>(
                            std::chrono::duration<double>(timeout_s));
  Run *r = it->second.get();
  const bool ok = s->cv_done.wait_until(lk, deadline, [&] {
    const int st = r->res.status;
    return st == CLV_RUN_DONE || st < 0;
  });
  *out = r->res;
  return ok ? CLV_OK : CLV_ERR_STATE;
}

int clv_server_release(clv_server *s, int64_t id) {
  if (!s) return CLV_ERR_ARG;
  std::lock_guard<std::mutex> g(s->mu);
  auto it = s->runs.find(id);
  if (it == s->runs.end()) return CLV_ERR_NOTFOUND;
  const int st = it->second->res.status;
  if (st == CLV_RUN_QUEUED || st == CLV_RUN_RUNNING) {
    s->err = "clv_server_release: run " + std::to_string(id) + " is not finished";
    return CLV_ERR_STATE;
  }
  s->runs.erase(it);
  return CLV_OK;
}

int clv_server_cancel(clv_server *s, int64_t id) {
  if (!s) return CLV_ERR_ARG;
  std::lock_guard<std::mutex> g(s->mu);
  auto it = s->runs.find(id);
  if (it == s->runs.end()) return CLV_ERR_NOTFOUND;
  const int st = it->second->res.status;
  if (st != CLV_RUN_QUEUED && st != CLV_RUN_RUNNING) {
    s->err = "clv_server_cancel: run " + std::to_string(id) + " has already finished";
    return CLV_ERR_STATE;
  }
  it->second->cancel.store(true);
  s->cv_work.notify_all();
  return CLV_OK;
}

int clv_server_info_get(clv_server *s, clv_server_info *out) {
  if (!s || !out) return CLV_ERR_ARG;
  std::lock_guard<std::mutex> g(s->mu);
  *out = s->info;
  out->budget_bytes = s->budget;
  return CLV_OK;
}

}  // extern "C"
