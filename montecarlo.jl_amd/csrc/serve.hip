// Serving form of the sweep: a request loop that admits independent chains by their HBM
// footprint and runs them concurrently on one device (continuous batching).
//
// The reference runs one chain per Julia process through run! (DQMC.jl:144-394); this is
// the many-request form of that loop.  A request carries a chain's shape, its hopping
// matrix, initial field, counter-stream seed and sweep count.  A scheduler thread admits
// queued requests in arrival order while the number of running chains is below
// max_active and the sum of their footprints (dqmc_footprint_bytes: the UDT stack and
// every other device buffer of the chain) fits the memory budget.
//   - An admitted request is prepared on a thread of its own (create, set hopping / field /
//     seed, build_stack on the chain's own stream), then joins the executor.
//   - An executor thread per size class (N <= 256, <= 576, <= 1024, larger), each with its
//     own stream, runs lockstep slice visits (sweep_spatial + propagate) of every FP64 chain
//     of its class that has joined, one launch per kernel kind covering all of them
//     (ensemble.h).  A lockstep step lasts as long as its largest chain's, so the classes
//     keep small requests from waiting on the largest ones; the classes' streams overlap on
//     the device.  A chain leaves after its 2 M nsweeps visits and a finishing thread reads
//     back field, G and statistics and frees it.  Requests that arrive while others run
//     join at the next step of their class, so their work shares the running launches.
//   - Mixed-precision requests run on their own thread and stream (dqmc_sweep).
// A chain's result does not depend on what else runs: every kernel of a chain reads and
// writes only that chain's buffers, in the chain's own order.
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/dqmc_hip.h"
#include "ensemble.h"

namespace {

enum { kQueued = 0, kRunning = 1, kDone = 2, kFailed = 3 };

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Request {
  int64_t id = 0;
  dqmc_config_t cfg{};
  std::vector<double> T;
  std::vector<int8_t> conf;
  uint64_t seed = 0;
  int nsweeps = 0;
  int64_t footprint = 0;
  int state = kQueued;            // guarded by the server mutex, as are t_start and t_done
  std::string err;
  std::vector<int8_t> conf_out;
  std::vector<double> G_out;
  dqmc_stats_t stats{};
  int64_t accepted = 0;
  double t_submit = 0, t_start = 0, t_done = 0;
  // executor state of a running request (touched by its own threads and the executor only)
  dqmc_chain_t h = nullptr;
  int64_t acc0 = 0;
  long long target = 0;
};

}  // namespace

struct dqmc_server_s {
  dqmc_server_config_t cfg{};
  std::mutex mu;
  std::condition_variable cv;
  std::deque<std::shared_ptr<Request>> queue;
  std::unordered_map<int64_t, std::shared_ptr<Request>> reqs;
  int64_t next_id = 1;
  int64_t used_bytes = 0;
  int active = 0;
  int peak_active = 0;
  int waiters = 0;        // threads inside dqmc_server_wait (destroy waits for them to leave)
  bool stop = false;
  std::thread scheduler;
  // the executors (one per size class): built chains wait in `joining`, then run in
  // lockstep steps
  static constexpr int kClasses = 4;
  std::thread executor[kClasses];
  std::deque<std::shared_ptr<Request>> joining[kClasses];
  int helpers = 0;        // preparing / finishing threads still running
};

namespace {

// Errors are kept per calling thread, so client threads that call the server concurrently
// never write the same string.
thread_local std::string t_server_error;

// Prepares a request's chain on the calling thread: create, hopping, field, seed and the
// stack build; false on failure (r.err).
bool prepare(Request& r, int device) {
  dqmc_config_t c = r.cfg;
  c.device = device;
  auto fail = [&](const char* what) {
    r.err = std::string(what) + ": " + dqmc_last_error(r.h);
    return false;
  };
  if (dqmc_create(&c, &r.h) != DQMC_OK) return fail("dqmc_create");
  const int n = c.n_sites;
  if (dqmc_set_hopping(r.h, r.T.data(), n) != DQMC_OK) return fail("dqmc_set_hopping");
  if (dqmc_set_field(r.h, r.conf.data(), n) != DQMC_OK) return fail("dqmc_set_field");
  if (dqmc_set_seed(r.h, r.seed, 0) != DQMC_OK) return fail("dqmc_set_seed");
  if (dqmc_build_stack(r.h) != DQMC_OK) return fail("dqmc_build_stack");
  dqmc_stats_t st{};
  if (dqmc_get_stats(r.h, &st) != DQMC_OK) return fail("dqmc_get_stats");
  r.acc0 = st.accepted;
  return true;
}

// Reads back the results of a finished chain (on its own stream again) and frees it.
bool finish(Request& r) {
  dqmc_chain_t h = r.h;
  auto fail = [&](const char* what) {
    r.err = std::string(what) + ": " + dqmc_last_error(h);
    return false;
  };
  const int n = r.cfg.n_sites;
  bool ok = true;
  if (dqmc_get_stats(h, &r.stats) != DQMC_OK) ok = fail("dqmc_get_stats");
  else if (r.stats.device_errors > 0) ok = fail("device error");
  if (ok) {
    r.accepted = r.stats.accepted - r.acc0;
    r.conf_out.assign((size_t)n * r.cfg.slices, 0);
    if (dqmc_get_field(h, r.conf_out.data(), n) != DQMC_OK) ok = fail("dqmc_get_field");
  }
  if (ok) {
    r.G_out.assign((size_t)n * n * r.cfg.flavors, 0.0);
    if (dqmc_get_greens(h, r.G_out.data(), n, 0) != DQMC_OK) ok = fail("dqmc_get_greens");
  }
  dqmc_destroy(h);
  r.h = nullptr;
  return ok;
}

// A request on its own thread and stream, start to finish (mixed precision).
bool run_alone(Request& r, int device) {
  bool ok = prepare(r, device);
  if (ok && dqmc_sweep(r.h, r.nsweeps, nullptr) != DQMC_OK) {
    r.err = std::string("dqmc_sweep: ") + dqmc_last_error(r.h);
    ok = false;
  }
  if (ok) return finish(r);
  if (r.h) { dqmc_destroy(r.h); r.h = nullptr; }
  return false;
}

void complete(dqmc_server_s* s, const std::shared_ptr<Request>& r, bool ok) {
  std::lock_guard<std::mutex> g(s->mu);
  r->state = ok ? kDone : kFailed;
  r->t_done = now_s();
  s->active--;
  s->used_bytes -= r->footprint;
  s->helpers--;
  s->cv.notify_all();
}

int size_class(int n) { return n <= 256 ? 0 : n <= 576 ? 1 : n <= 1024 ? 2 : 3; }

void executor_loop(dqmc_server_s* s, int cls) {
  dqmc::Ensemble* ens = nullptr;
  const int rc = dqmc::ensemble_create(s->cfg.device, &ens);
  std::vector<std::shared_ptr<Request>> act;
  std::vector<dqmc_chain_t> hs;
  auto& joining = s->joining[cls];
  for (;;) {
    {
      std::unique_lock<std::mutex> lk(s->mu);
      s->cv.wait(lk, [&]() { return !act.empty() || !joining.empty() || (s->stop && s->helpers == 0); });
      if (act.empty() && joining.empty() && s->stop && s->helpers == 0) break;
      while (!joining.empty()) {
        act.push_back(joining.front());
        joining.pop_front();
      }
    }
    // newly joined chains attach (a failed attach fails the request)
    for (size_t i = 0; i < act.size();) {
      Request& r = *act[i];
      if (r.target == 0) {
        if (rc != DQMC_OK || dqmc::ensemble_attach(ens, r.h) != DQMC_OK) {
          r.err = rc != DQMC_OK ? "executor could not start" : dqmc::ensemble_error(ens);
          dqmc_destroy(r.h);
          r.h = nullptr;
          complete(s, act[i], false);
          act.erase(act.begin() + i);
          continue;
        }
        r.target = dqmc::chain_steps(r.h) + 2LL * dqmc::chain_slices(r.h) * r.nsweeps;
      }
      i++;
    }
    hs.clear();
    for (auto& r : act) hs.push_back(r->h);
    bool step_ok = true;
    std::string step_err;
    if (!hs.empty() && dqmc::ensemble_step(ens, hs.data(), (int)hs.size()) != DQMC_OK) {
      step_ok = false;
      step_err = dqmc::ensemble_error(ens);
    }
    // chains that made all their visits (or all, after a failed step) leave and finish on
    // threads of their own
    for (size_t i = 0; i < act.size();) {
      std::shared_ptr<Request> r = act[i];
      if (step_ok && dqmc::chain_steps(r->h) < r->target) { i++; continue; }
      act.erase(act.begin() + i);
      const bool left = dqmc::ensemble_detach(ens, r->h) == DQMC_OK;
      if (!step_ok) r->err = "lockstep step failed: " + step_err;
      const bool ok = step_ok && left;
      std::thread([s, r, ok]() {
        bool fine = ok;
        if (fine) fine = finish(*r);
        else if (r->h) { dqmc_destroy(r->h); r->h = nullptr; }
        complete(s, r, fine);
      }).detach();
    }
  }
  if (ens) dqmc::ensemble_destroy(ens);
}

void scheduler_loop(dqmc_server_s* s) {
  std::unique_lock<std::mutex> lk(s->mu);
  for (;;) {
    s->cv.wait(lk, [&]() {
      if (s->stop) return true;
      if (s->queue.empty()) return false;
      const auto& r = s->queue.front();
      return s->active < s->cfg.max_active && s->used_bytes + r->footprint <= s->cfg.memory_budget;
    });
    if (s->stop) return;
    std::shared_ptr<Request> r = s->queue.front();
    s->queue.pop_front();
    s->active++;
    if (s->active > s->peak_active) s->peak_active = s->active;
    s->used_bytes += r->footprint;
    s->helpers++;
    r->state = kRunning;
    r->t_start = now_s();
    const int device = s->cfg.device;
    std::thread([s, r, device]() {
      if (r->cfg.precision != DQMC_PRECISION_FP64) {
        complete(s, r, run_alone(*r, device));
        return;
      }
      if (!prepare(*r, device)) {
        if (r->h) { dqmc_destroy(r->h); r->h = nullptr; }
        complete(s, r, false);
        return;
      }
      if (r->nsweeps == 0) {
        complete(s, r, finish(*r));
        return;
      }
      std::lock_guard<std::mutex> g(s->mu);
      // the executor of its size class owns it now; `helpers` drops when it finishes
      s->joining[size_class(r->cfg.n_sites)].push_back(r);
      s->cv.notify_all();
    }).detach();
  }
}

}  // namespace

extern "C" {

const char* dqmc_server_last_error(dqmc_server_t s) {
  (void)s;
  return t_server_error.c_str();
}

int dqmc_server_create(const dqmc_server_config_t* cfg, dqmc_server_t* out) {
  if (!cfg || !out) { t_server_error = "dqmc_server_create: null argument"; return DQMC_E_INVALID; }
  if (cfg->max_active < 1 || cfg->memory_budget <= 0) {
    t_server_error = "dqmc_server_create: max_active >= 1 and memory_budget > 0 required";
    return DQMC_E_INVALID;
  }
  auto* s = new dqmc_server_s();
  s->cfg = *cfg;
  s->scheduler = std::thread(scheduler_loop, s);
  for (int c = 0; c < dqmc_server_s::kClasses; c++) s->executor[c] = std::thread(executor_loop, s, c);
  *out = s;
  return DQMC_OK;
}

int dqmc_server_submit(dqmc_server_t s, const dqmc_request_t* rq, int64_t* id) {
  if (!s || !rq || !id) return DQMC_E_INVALID;
  const dqmc_config_t& c = rq->chain;
  int64_t fp = 0;
  if (dqmc_footprint_bytes(&c, &fp) != DQMC_OK) {
    t_server_error = std::string("dqmc_server_submit: ") + dqmc_last_error(nullptr);
    return DQMC_E_INVALID;
  }
  const int n = c.n_sites;
  if (!rq->hopping || !rq->field || rq->ld_hopping < n || rq->ld_field < n || rq->nsweeps < 0) {
    t_server_error = "dqmc_server_submit: null hopping/field, ld < n_sites or nsweeps < 0";
    return DQMC_E_INVALID;
  }
  if (fp > s->cfg.memory_budget) {
    t_server_error = "dqmc_server_submit: the chain's footprint exceeds the server's memory budget";
    return DQMC_E_ALLOC;
  }
  auto r = std::make_shared<Request>();
  r->cfg = c;
  r->T.resize((size_t)n * n);
  for (int j = 0; j < n; j++)
    for (int i = 0; i < n; i++) r->T[(size_t)i + (size_t)j * n] = rq->hopping[(size_t)i + (size_t)j * rq->ld_hopping];
  r->conf.resize((size_t)n * c.slices);
  for (int l = 0; l < c.slices; l++)
    for (int i = 0; i < n; i++) r->conf[(size_t)i + (size_t)l * n] = rq->field[(size_t)i + (size_t)l * rq->ld_field];
  r->seed = rq->seed;
  r->nsweeps = rq->nsweeps;
  r->footprint = fp;
  std::lock_guard<std::mutex> g(s->mu);
  if (s->stop) { t_server_error = "dqmc_server_submit: server is shutting down"; return DQMC_E_STATE; }
  r->id = s->next_id++;
  r->t_submit = now_s();
  s->reqs[r->id] = r;
  s->queue.push_back(r);
  s->cv.notify_all();
  *id = r->id;
  return DQMC_OK;
}

int dqmc_server_poll(dqmc_server_t s, int64_t id, int* state) {
  if (!s || !state) return DQMC_E_INVALID;
  std::lock_guard<std::mutex> g(s->mu);
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) { t_server_error = "dqmc_server_poll: unknown request id"; return DQMC_E_INVALID; }
  *state = it->second->state;
  return DQMC_OK;
}

int dqmc_server_wait(dqmc_server_t s, int64_t id) {
  if (!s) return DQMC_E_INVALID;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) { t_server_error = "dqmc_server_wait: unknown request id"; return DQMC_E_INVALID; }
  std::shared_ptr<Request> r = it->second;
  s->waiters++;
  s->cv.wait(lk, [&]() { return r->t_done > 0; });
  s->waiters--;
  s->cv.notify_all();
  return DQMC_OK;
}

int dqmc_server_result(dqmc_server_t s, int64_t id, int8_t* conf, int ld_conf, double* G, int ld_G,
                       int64_t* accepted, dqmc_stats_t* stats, double* latency) {
  if (!s) return DQMC_E_INVALID;
  std::shared_ptr<Request> r;
  {
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) { t_server_error = "dqmc_server_result: unknown request id"; return DQMC_E_INVALID; }
    r = it->second;
    if (r->t_done <= 0) { t_server_error = "dqmc_server_result: request not finished"; return DQMC_E_STATE; }
    s->reqs.erase(it);
  }
  if (r->state == kFailed) { t_server_error = r->err; return DQMC_E_STATE; }
  const int n = r->cfg.n_sites;
  if (conf) {
    if (ld_conf < n) { t_server_error = "dqmc_server_result: ld_conf < n_sites"; return DQMC_E_INVALID; }
    for (int l = 0; l < r->cfg.slices; l++)
      for (int i = 0; i < n; i++) conf[(size_t)i + (size_t)l * ld_conf] = r->conf_out[(size_t)i + (size_t)l * n];
  }
  if (G) {
    if (ld_G < n) { t_server_error = "dqmc_server_result: ld_G < n_sites"; return DQMC_E_INVALID; }
    for (int j = 0; j < n * r->cfg.flavors; j++)
      for (int i = 0; i < n; i++) G[(size_t)i + (size_t)j * ld_G] = r->G_out[(size_t)i + (size_t)j * n];
  }
  if (accepted) *accepted = r->accepted;
  if (stats) *stats = r->stats;
  if (latency) {
    latency[0] = r->t_done - r->t_submit;   // submit -> result ready
    latency[1] = r->t_start - r->t_submit;  // time queued before admission
  }
  return DQMC_OK;
}

int dqmc_server_status(dqmc_server_t s, int* queued, int* active, int* peak_active, int64_t* used_bytes) {
  if (!s) return DQMC_E_INVALID;
  std::lock_guard<std::mutex> g(s->mu);
  if (queued) *queued = (int)s->queue.size();
  if (active) *active = s->active;
  if (peak_active) *peak_active = s->peak_active;
  if (used_bytes) *used_bytes = s->used_bytes;
  return DQMC_OK;
}

int dqmc_server_destroy(dqmc_server_t s) {
  if (!s) return DQMC_E_INVALID;
  {
    std::unique_lock<std::mutex> lk(s->mu);
    // queued requests are dropped (marked failed, so their waiters return); running ones
    // finish first, and every thread inside dqmc_server_wait leaves before the server is freed
    const double t = now_s();
    for (auto& r : s->queue) {
      r->state = kFailed;
      r->err = "dropped: the server was destroyed before the request was admitted";
      r->t_done = t;
    }
    s->queue.clear();
    s->stop = true;
    s->cv.notify_all();
    s->cv.wait(lk, [&]() { return s->active == 0 && s->waiters == 0 && s->helpers == 0; });
  }
  s->scheduler.join();
  for (auto& t : s->executor) t.join();
  delete s;
  return DQMC_OK;
}

}  // extern "C"
