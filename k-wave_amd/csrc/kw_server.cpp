// Serving loop of include/kwave_hip.h: many k-Wave simulations resident on
// one device, admitted by HBM footprint and advanced concurrently.
//
// The reference has no in-process form of this: every kspaceFirstOrder3DC
// call writes an input file and launches the binary once
// (kspaceFirstOrder3DC.m:544), and a batch of simulations is a loop of such
// calls (kWaveTester.m:1936-1950).  Here a request is the same dataset set
// (kw_sim_desc), deep-copied at submit.  One admission thread takes requests
// in arrival order and starts a runner thread for each once the request's
// footprint fits the remaining budget and fewer than max_active are running.
// A runner creates the simulation (its own stream), queues its steps in
// quanta with at most two quanta in flight, copies every recorded output to
// host memory, frees the device memory and returns the reservation.
// Simulations on different streams overlap on the device; each one's kernels
// and their order are those of a lone kw_sim, so its results are identical.
// In batched mode (kw_server_config.batching) a runner hands its simulation
// to one batch worker instead, which advances all simulations of a grid
// shape and physics branch together: one launch per kernel of the step for
// the whole group, a quantum of steps replayed from a captured hipGraph.
#include <hip/hip_runtime.h>

#include <algorithm>
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

#include "../../include/kwave_hip.h"

namespace kw {
void set_last_error(const char* msg);
struct SimBatch;
std::string sim_batch_key(const kw_sim* s);
SimBatch* batch_create();
void batch_destroy(SimBatch* b);
kw_status batch_run(SimBatch* b, const std::vector<kw_sim*>& sims, int nsteps, hipStream_t st, bool use_graph);
void sim_detach(kw_sim* s);
}  // namespace kw

namespace {

using Clock = std::chrono::steady_clock;

kw_status srv_fail(kw_status st, const std::string& msg) {
  kw::set_last_error(msg.c_str());
  return st;
}

// A kw_sim_desc with every array it points to owned by the request.
struct OwnedDesc {
  kw_sim_desc d{};
  std::vector<std::vector<float>> f;
  std::vector<std::vector<uint64_t>> u;

  const float* keep_f(const float* p, uint64_t n) {
    if (!p || n == 0) return nullptr;
    f.emplace_back(p, p + n);
    return f.back().data();
  }
  const uint64_t* keep_u(const uint64_t* p, uint64_t n) {
    if (!p || n == 0) return nullptr;
    u.emplace_back(p, p + n);
    return u.back().data();
  }
  kw_real_array keep(const kw_real_array& a) { return kw_real_array{keep_f(a.data, a.count), a.data ? a.count : 0}; }

  explicit OwnedDesc(const kw_sim_desc& s) : d(s) {
    const uint64_t N = s.Nx * s.Ny * s.Nz;
    d.c0 = keep(s.c0);
    d.rho0 = keep(s.rho0);
    d.rho0_sgx = keep(s.rho0_sgx);
    d.rho0_sgy = keep(s.rho0_sgy);
    d.rho0_sgz = keep(s.rho0_sgz);
    d.BonA = keep(s.BonA);
    d.alpha_coeff = keep(s.alpha_coeff);
    d.p0_source_input = keep_f(s.p0_source_input, N);
    d.p_source_index = keep_u(s.p_source_index, s.p_source_npoints);
    d.p_source_input = keep_f(s.p_source_input, s.p_source_flag * (s.p_source_many ? s.p_source_npoints : 1));
    d.u_source_index = keep_u(s.u_source_index, s.u_source_npoints);
    const uint64_t useries = s.u_source_many ? s.u_source_npoints : 1;
    d.ux_source_input = keep_f(s.ux_source_input, s.ux_source_flag * useries);
    d.uy_source_input = keep_f(s.uy_source_input, s.uy_source_flag * useries);
    d.uz_source_input = keep_f(s.uz_source_input, s.uz_source_flag * useries);
    d.sensor_mask_index = keep_u(s.sensor_mask_index, s.sensor_npoints);
    d.transducer_source_input = keep_f(s.transducer_source_input, s.transducer_source_length);
    d.delay_mask = keep_u(s.delay_mask, s.transducer_source_flag ? s.u_source_npoints : 0);
  }
};

// A simulation handed to the batch worker by its runner.
struct BatchEntry {
  kw_sim* sim = nullptr;
  std::string key;
  uint64_t nt = 0;
  bool done = false;
  kw_status status = KW_OK;
  std::string error;
};

// Batched mode: one worker advances every batchable simulation.  Members of
// one batch key (grid shape, physics branch, launch configuration) form a
// group with its own stream; each round launches one quantum of batched
// steps per group (a replayed hipGraph once the group's membership has run
// a quantum), then waits for the groups and hands finished simulations back
// to their runners.  New members join at the next round.
struct BatchWorker {
  int device = 0;
  unsigned quantum = 16;
  std::mutex mu;
  std::condition_variable cv;  // work arrived / an entry finished / stop
  std::vector<std::shared_ptr<BatchEntry>> pending;
  struct Group {
    kw::SimBatch* batch = nullptr;
    hipStream_t stream = nullptr;
    std::vector<std::shared_ptr<BatchEntry>> members;
  };
  std::map<std::string, Group> groups;
  bool stopping = false;
  std::thread th;

  void loop();
  // blocks until the worker has run the simulation to its Nt (or failed)
  kw_status run(kw_sim* sim, const std::string& key, uint64_t nt, std::string* msg) {
    auto e = std::make_shared<BatchEntry>();
    e->sim = sim;
    e->key = key;
    e->nt = nt;
    std::unique_lock<std::mutex> lk(mu);
    pending.push_back(e);
    cv.notify_all();
    cv.wait(lk, [&] { return e->done; });
    *msg = e->error;
    return e->status;
  }
};

void BatchWorker::loop() {
  (void)hipSetDevice(device);
  for (;;) {
    {
      std::unique_lock<std::mutex> lk(mu);
      auto busy = [&] {
        for (auto& g : groups)
          if (!g.second.members.empty()) return true;
        return false;
      };
      cv.wait(lk, [&] { return stopping || !pending.empty() || busy(); });
      if (stopping && pending.empty() && !busy()) break;
      for (auto& e : pending) {
        Group& g = groups[e->key];
        if (!g.batch) {
          g.batch = kw::batch_create();
          if (hipStreamCreateWithFlags(&g.stream, hipStreamNonBlocking) != hipSuccess) {
            (void)hipGetLastError();
            g.stream = nullptr;
          }
        }
        if (!g.stream) {
          e->status = KW_ERR_DEVICE;
          e->error = "hipStreamCreateWithFlags failed";
          e->done = true;
          continue;
        }
        g.members.push_back(e);
      }
      pending.clear();
      cv.notify_all();
    }
    // one quantum per group, all queued before any wait so the groups overlap
    std::vector<Group*> ran;
    for (auto& kv : groups) {
      Group& g = kv.second;
      if (g.members.empty()) continue;
      std::vector<kw_sim*> sims;
      for (auto& m : g.members) sims.push_back(m->sim);
      const kw_status st = kw::batch_run(g.batch, sims, (int)quantum, g.stream, true);
      if (st != KW_OK) {
        const std::string msg = kw_last_error();
        (void)hipStreamSynchronize(g.stream);
        std::lock_guard<std::mutex> lk(mu);
        for (auto& m : g.members) {
          kw::sim_detach(m->sim);
          m->status = st;
          m->error = msg;
          m->done = true;
        }
        g.members.clear();
        cv.notify_all();
        continue;
      }
      ran.push_back(&g);
    }
    for (Group* g : ran) {
      const hipError_t e = hipStreamSynchronize(g->stream);
      std::lock_guard<std::mutex> lk(mu);
      for (size_t i = 0; i < g->members.size();) {
        BatchEntry& m = *g->members[i];
        if (e != hipSuccess || kw_sim_steps_done(m.sim) >= m.nt) {
          kw::sim_detach(m.sim);
          if (e != hipSuccess) {
            m.status = KW_ERR_DEVICE;
            m.error = std::string("batched quantum failed: ") + hipGetErrorString(e);
          }
          m.done = true;
          g->members.erase(g->members.begin() + (long)i);
        } else {
          ++i;
        }
      }
      cv.notify_all();
    }
  }
  for (auto& kv : groups) {
    kw::batch_destroy(kv.second.batch);
    if (kv.second.stream) (void)hipStreamDestroy(kv.second.stream);
  }
  groups.clear();
}

struct Request {
  uint64_t id = 0;
  std::unique_ptr<OwnedDesc> in;  // dropped once the simulation is on the device
  uint64_t footprint = 0, updates = 0, nt = 0;
  int state = KW_REQ_QUEUED;
  kw_status status = KW_OK;
  std::string error;
  double t_submit = 0, t_admit = 0, t_done = 0;
  std::map<int, std::vector<unsigned char>> out;
};

}  // namespace

struct kw_server {
  int device = 0;
  uint64_t budget = 0, reserved = 0, peak = 0;
  unsigned max_active = 8, quantum = 16;
  bool autotune = false;
  std::unique_ptr<BatchWorker> batcher;  // batched mode
  Clock::time_point t0;

  std::mutex mu;
  std::condition_variable cv_admit;  // queue or capacity changed
  std::condition_variable cv_done;   // a request finished
  std::map<uint64_t, std::shared_ptr<Request>> reqs;
  std::deque<std::shared_ptr<Request>> queue;
  std::vector<std::pair<std::thread, std::shared_ptr<Request>>> runners;
  unsigned active = 0;
  uint64_t next_id = 1, submitted = 0, completed = 0, failed = 0, updates_done = 0;
  bool stopping = false;
  std::thread admitter;

  double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }

  void admission_loop();
  void run_request(std::shared_ptr<Request> r);
  void finish(const std::shared_ptr<Request>& r, kw_status st, const std::string& msg);
};

void kw_server::finish(const std::shared_ptr<Request>& r, kw_status st, const std::string& msg) {
  std::lock_guard<std::mutex> lk(mu);
  r->t_done = now();
  r->status = st;
  r->state = st == KW_OK ? KW_REQ_DONE : KW_REQ_FAILED;
  r->error = msg;
  r->in.reset();
  if (st == KW_OK) {
    ++completed;
    updates_done += r->updates;
  } else {
    ++failed;
  }
  reserved -= r->footprint;
  --active;
  cv_admit.notify_all();
  cv_done.notify_all();
}

void kw_server::run_request(std::shared_ptr<Request> r) {
  (void)hipSetDevice(device);
  kw_sim* sim = nullptr;
  kw_status st = kw_sim_create(&r->in->d, &sim);
  if (st != KW_OK) {
    finish(r, st, kw_last_error());
    return;
  }
  if (autotune && (st = kw_sim_autotune(sim, 4, nullptr, nullptr, nullptr)) != KW_OK) {
    const std::string m = kw_last_error();
    kw_sim_destroy(sim);
    finish(r, st, m);
    return;
  }
  // the host copies of the inputs are no longer needed
  {
    std::lock_guard<std::mutex> lk(mu);
    r->in.reset();
  }
  std::string msg;
  const std::string bkey = batcher ? kw::sim_batch_key(sim) : std::string();
  if (!bkey.empty()) {
    // batched: the worker advances it with the other simulations of its key
    st = batcher->run(sim, bkey, r->nt, &msg);
    std::map<int, std::vector<unsigned char>> out;
    for (int id = 0; id < KW_OUT_COUNT && st == KW_OK; ++id) {
      const size_t b = kw_sim_output_bytes(sim, id);
      if (!b) continue;
      std::vector<unsigned char> h(b);
      st = kw_sim_fetch(sim, id, h.data(), b);
      if (st != KW_OK) msg = kw_last_error();
      out.emplace(id, std::move(h));
    }
    kw_sim_destroy(sim);
    if (st == KW_OK) {
      std::lock_guard<std::mutex> lk(mu);
      r->out = std::move(out);
    }
    finish(r, st, msg);
    return;
  }
  hipEvent_t ev[2] = {nullptr, nullptr};
  for (auto& e : ev)
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming | hipEventBlockingSync) != hipSuccess) {
      st = KW_ERR_DEVICE;
      msg = "hipEventCreateWithFlags failed";
    }
  // queue quanta of steps, keeping at most two in flight on the stream
  uint64_t done = 0;
  int k = 0;
  while (st == KW_OK && done < r->nt) {
    const uint64_t q = std::min<uint64_t>(quantum, r->nt - done);
    if (k >= 2 && hipEventSynchronize(ev[k & 1]) != hipSuccess) {
      st = KW_ERR_DEVICE;
      msg = "hipEventSynchronize failed";
      break;
    }
    st = kw_sim_run(sim, q, nullptr);
    if (st != KW_OK) {
      msg = kw_last_error();
      break;
    }
    if (hipEventRecord(ev[k & 1], (hipStream_t)kw_sim_stream(sim)) != hipSuccess) {
      st = KW_ERR_DEVICE;
      msg = "hipEventRecord failed";
      break;
    }
    done += q;
    ++k;
  }
  std::map<int, std::vector<unsigned char>> out;
  if (st == KW_OK) {
    for (int id = 0; id < KW_OUT_COUNT && st == KW_OK; ++id) {
      const size_t b = kw_sim_output_bytes(sim, id);
      if (!b) continue;
      std::vector<unsigned char> h(b);
      st = kw_sim_fetch(sim, id, h.data(), b);
      if (st != KW_OK) msg = kw_last_error();
      out.emplace(id, std::move(h));
    }
  } else {
    (void)kw_sim_sync(sim);
  }
  for (auto& e : ev)
    if (e) (void)hipEventDestroy(e);
  kw_sim_destroy(sim);
  if (st == KW_OK) {
    std::lock_guard<std::mutex> lk(mu);
    r->out = std::move(out);
  }
  finish(r, st, msg);
}

void kw_server::admission_loop() {
  (void)hipSetDevice(device);
  std::unique_lock<std::mutex> lk(mu);
  for (;;) {
    cv_admit.wait(lk, [&] {
      if (stopping) return true;
      if (queue.empty() || active >= max_active) return false;
      return reserved + queue.front()->footprint <= budget;
    });
    if (stopping) break;
    // join runners whose request has finished (finish() is their last use of mu)
    for (size_t i = 0; i < runners.size();) {
      const int st = runners[i].second->state;
      if (st == KW_REQ_DONE || st == KW_REQ_FAILED) {
        runners[i].first.join();
        runners[i] = std::move(runners.back());
        runners.pop_back();
      } else {
        ++i;
      }
    }
    std::shared_ptr<Request> r = queue.front();
    queue.pop_front();
    reserved += r->footprint;
    peak = std::max(peak, reserved);
    ++active;
    r->state = KW_REQ_RUNNING;
    r->t_admit = now();
    runners.emplace_back(std::thread(&kw_server::run_request, this, r), r);
  }
  // fail whatever is still queued
  for (auto& r : queue) {
    r->state = KW_REQ_FAILED;
    r->status = KW_ERR_INVALID;
    r->error = "server shut down before the request was admitted";
    r->t_done = now();
    r->in.reset();
    ++failed;
  }
  queue.clear();
  cv_done.notify_all();
}

extern "C" {

kw_status kw_server_create(const kw_server_config* cfg, kw_server** out) {
  if (!cfg || !out) return srv_fail(KW_ERR_INVALID, "null argument");
  *out = nullptr;
  if (cfg->abi_version != KW_ABI_VERSION) return srv_fail(KW_ERR_INVALID, "abi_version mismatch");
  auto s = std::make_unique<kw_server>();
  s->device = cfg->device;
  s->max_active = cfg->max_active ? cfg->max_active : 8;
  s->quantum = cfg->steps_per_quantum ? cfg->steps_per_quantum : 16;
  s->autotune = cfg->autotune != 0;
  if (hipSetDevice(s->device) != hipSuccess) return srv_fail(KW_ERR_DEVICE, "hipSetDevice failed");
  if (cfg->batching) {
    s->batcher = std::make_unique<BatchWorker>();
    s->batcher->device = s->device;
    s->batcher->quantum = s->quantum;
  }
  s->budget = cfg->hbm_budget_bytes;
  if (!s->budget) {
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) return srv_fail(KW_ERR_DEVICE, "hipMemGetInfo failed");
    s->budget = (uint64_t)(0.9 * (double)fr);
  }
  s->t0 = Clock::now();
  kw_server* p = s.release();
  if (p->batcher) p->batcher->th = std::thread(&BatchWorker::loop, p->batcher.get());
  p->admitter = std::thread(&kw_server::admission_loop, p);
  *out = p;
  return KW_OK;
}

kw_status kw_server_submit(kw_server* s, const kw_sim_desc* d, uint64_t* id) {
  if (!s || !d || !id) return srv_fail(KW_ERR_INVALID, "null argument");
  if (d->abi_version != KW_ABI_VERSION) return srv_fail(KW_ERR_INVALID, "abi_version mismatch");
  const uint64_t fp = kw_sim_footprint_bytes(d);
  if (!fp) return srv_fail(KW_ERR_INVALID, "invalid simulation description");
  if (fp > s->budget)
    return srv_fail(KW_ERR_OOM, "footprint " + std::to_string(fp) + " B exceeds the HBM budget " +
                                    std::to_string(s->budget) + " B");
  auto r = std::make_shared<Request>();
  r->in = std::make_unique<OwnedDesc>(*d);
  r->in->d.device = s->device;
  r->footprint = fp;
  r->nt = d->Nt;
  r->updates = d->Nx * d->Ny * d->Nz * d->Nt;
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->stopping) return srv_fail(KW_ERR_INVALID, "server is shutting down");
  r->id = s->next_id++;
  r->t_submit = s->now();
  s->reqs.emplace(r->id, r);
  s->queue.push_back(r);
  ++s->submitted;
  *id = r->id;
  s->cv_admit.notify_all();
  return KW_OK;
}

int kw_server_poll(kw_server* s, uint64_t id) {
  if (!s) return KW_REQ_UNKNOWN;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(id);
  return it == s->reqs.end() ? KW_REQ_UNKNOWN : it->second->state;
}

kw_status kw_server_wait(kw_server* s, uint64_t id, double timeout_s) {
  if (!s) return srv_fail(KW_ERR_INVALID, "null handle");
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) return srv_fail(KW_ERR_INVALID, "unknown request " + std::to_string(id));
  std::shared_ptr<Request> r = it->second;
  auto finished = [&] { return r->state == KW_REQ_DONE || r->state == KW_REQ_FAILED; };
  if (timeout_s < 0) {
    s->cv_done.wait(lk, finished);
    return KW_OK;
  }
  if (s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), finished)) return KW_OK;
  return srv_fail(KW_ERR_TIMEOUT, "request " + std::to_string(id) + " not finished");
}

kw_status kw_server_request_info(kw_server* s, uint64_t id, kw_request_info* info) {
  if (!s || !info) return srv_fail(KW_ERR_INVALID, "null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) return srv_fail(KW_ERR_INVALID, "unknown request " + std::to_string(id));
  const Request& r = *it->second;
  info->state = r.state;
  info->status = r.status;
  info->footprint_bytes = r.footprint;
  info->point_updates = r.updates;
  info->t_submit = r.t_submit;
  info->t_admit = r.t_admit;
  info->t_done = r.t_done;
  return KW_OK;
}

const char* kw_server_request_error(kw_server* s, uint64_t id) {
  if (!s) return "";
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(id);
  return it == s->reqs.end() ? "" : it->second->error.c_str();
}

size_t kw_server_output_bytes(kw_server* s, uint64_t id, int output_id) {
  if (!s) return 0;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(id);
  if (it == s->reqs.end() || it->second->state != KW_REQ_DONE) return 0;
  auto o = it->second->out.find(output_id);
  return o == it->second->out.end() ? 0 : o->second.size();
}

kw_status kw_server_fetch(kw_server* s, uint64_t id, int output_id, void* dst, size_t bytes) {
  if (!s || !dst) return srv_fail(KW_ERR_INVALID, "null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) return srv_fail(KW_ERR_INVALID, "unknown request " + std::to_string(id));
  const Request& r = *it->second;
  if (r.state != KW_REQ_DONE) return srv_fail(KW_ERR_INVALID, "request " + std::to_string(id) + " is not done");
  auto o = r.out.find(output_id);
  if (o == r.out.end()) return srv_fail(KW_ERR_INVALID, "output " + std::to_string(output_id) + " was not recorded");
  if (o->second.size() != bytes)
    return srv_fail(KW_ERR_INVALID, "output " + std::to_string(output_id) + " needs " +
                                        std::to_string(o->second.size()) + " bytes");
  std::memcpy(dst, o->second.data(), bytes);
  return KW_OK;
}

kw_status kw_server_release(kw_server* s, uint64_t id) {
  if (!s) return srv_fail(KW_ERR_INVALID, "null handle");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) return srv_fail(KW_ERR_INVALID, "unknown request " + std::to_string(id));
  if (it->second->state != KW_REQ_DONE && it->second->state != KW_REQ_FAILED)
    return srv_fail(KW_ERR_INVALID, "request " + std::to_string(id) + " is still queued or running");
  s->reqs.erase(it);
  return KW_OK;
}

kw_status kw_server_stats_get(kw_server* s, kw_server_stats* st) {
  if (!s || !st) return srv_fail(KW_ERR_INVALID, "null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  st->submitted = s->submitted;
  st->completed = s->completed;
  st->failed = s->failed;
  st->queued = s->queue.size();
  st->active = s->active;
  st->hbm_budget_bytes = s->budget;
  st->hbm_reserved_bytes = s->reserved;
  st->hbm_peak_reserved_bytes = s->peak;
  st->point_updates_done = s->updates_done;
  st->elapsed_s = s->now();
  return KW_OK;
}

kw_status kw_server_destroy(kw_server* s) {
  if (!s) return KW_OK;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stopping = true;
    s->cv_admit.notify_all();
  }
  s->admitter.join();
  // runners are only started by the admission thread, which has ended
  for (auto& t : s->runners) t.first.join();
  if (s->batcher) {
    // every runner has returned, so no simulation is left with the worker
    {
      std::lock_guard<std::mutex> lk(s->batcher->mu);
      s->batcher->stopping = true;
      s->batcher->cv.notify_all();
    }
    s->batcher->th.join();
  }
  delete s;
  return KW_OK;
}

}  // extern "C"
