// hg_server.hip -- the scene server: the scheduling policy of hg_sched.h
// driving resident MPM / DEM / hybrid handles, one stream each, from a
// worker thread.
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>

#include "../host/hg_sched.h"
#include "hg_hybrid_handle.h"
#include "../../include/hg_server.h"

namespace {

using hg::sched::Job;
using hg::sched::State;

// Deep copies of the descriptors' arrays (the caller may free its own).
struct OwnedMpm {
  std::vector<double> px, pn, pl, q, v, m, vol, hl, be, J;
  std::vector<int32_t> pb;
  std::vector<uint8_t> ks;
  hg_mpm_desc d{};
  void take(const hg_mpm_desc &s) {
    d = s;
    const size_t N = s.npoints, P = s.nplanes;
    auto cp = [](std::vector<double> &dst, const double *src, size_t n) {
      if (src) dst.assign(src, src + n);
      else dst.clear();
      return src ? dst.data() : nullptr;
    };
    d.plane_x = cp(px, s.plane_x, 2 * P);
    d.plane_n = cp(pn, s.plane_n, 2 * P);
    d.plane_lower = cp(pl, s.plane_lower, P);
    if (s.plane_behavior) pb.assign(s.plane_behavior, s.plane_behavior + P);
    d.plane_behavior = s.plane_behavior ? pb.data() : nullptr;
    d.q = cp(q, s.q, 2 * N);
    d.v = cp(v, s.v, 2 * N);
    d.m = cp(m, s.m, N);
    d.volume = cp(vol, s.volume, N);
    d.hl = cp(hl, s.hl, N);
    d.be_bar = cp(be, s.be_bar, 4 * N);
    d.J = cp(J, s.J, N);
    if (s.kinematically_scripted) ks.assign(s.kinematically_scripted, s.kinematically_scripted + N);
    d.kinematically_scripted = s.kinematically_scripted ? ks.data() : nullptr;
  }
};

struct OwnedDem {
  std::vector<double> q, v, mass, r, px, pn;
  std::vector<uint8_t> fixed;
  hg_dem_desc d{};
  void take(const hg_dem_desc &s) {
    d = s;
    const size_t B = s.nbodies, P = s.nplanes;
    auto cp = [](std::vector<double> &dst, const double *src, size_t n) {
      if (src) dst.assign(src, src + n);
      else dst.clear();
      return src ? dst.data() : nullptr;
    };
    d.q = cp(q, s.q, 3 * B);
    d.v = cp(v, s.v, 3 * B);
    d.mass = cp(mass, s.mass, 3 * B);
    d.radius = cp(r, s.radius, B);
    d.plane_x = cp(px, s.plane_x, 2 * P);
    d.plane_n = cp(pn, s.plane_n, 2 * P);
    if (s.fixed) fixed.assign(s.fixed, s.fixed + B);
    d.fixed = s.fixed ? fixed.data() : nullptr;
  }
};

enum class Kind { Mpm, Dem, Hybrid };

struct Req {
  Kind kind = Kind::Mpm;
  double dt = 0.0;
  OwnedMpm mpm;
  OwnedDem dem;
  std::vector<double> w;
  hg_mpm *m = nullptr;
  hg_dem *d = nullptr;
  hg_hybrid *h = nullptr;
  hipStream_t stream = nullptr;
  std::deque<hipEvent_t> events;
  // pinned host copies of the handles' device error words (MPM scene 0, DEM
  // scene 0, hybrid coupling), refreshed after every slice
  int32_t *herr = nullptr;
  // results
  std::vector<double> pq, pv, pb, pJ, bq, bv;
  double latency = 0.0;
};

class GpuExecutor : public hg::sched::Executor {
 public:
  explicit GpuExecutor(hg_ctx *ctx) : ctx_(ctx) {}
  ~GpuExecutor() override {
    for (hipEvent_t e : pool_) hipEventDestroy(e);
    for (int32_t *p : pinned_) hipHostFree(p);
  }

  bool start(Job &j) override {
    Req &r = *static_cast<Req *>(j.payload.get());
    hg_status st = HG_OK;
    uint64_t bytes = 0;
    switch (r.kind) {
      case Kind::Mpm:
        st = hg_mpm_create(ctx_, &r.mpm.d, &r.m);
        if (st == HG_OK) {
          r.stream = r.m->stream;
          hg_mpm_device_bytes(r.m, &bytes);
        }
        break;
      case Kind::Dem:
        st = hg_dem_create(ctx_, &r.dem.d, &r.d);
        if (st == HG_OK && !r.w.empty()) st = hg_dem_set_weights(r.d, 0, r.w.data());
        if (st == HG_OK) {
          r.stream = r.d->stream;
          hg_dem_device_bytes(r.d, &bytes);
        }
        break;
      case Kind::Hybrid:
        st = hg_hybrid_create(ctx_, &r.mpm.d, &r.dem.d, r.w.empty() ? nullptr : r.w.data(), &r.h);
        if (st == HG_OK) {
          r.stream = r.h->stream;
          hg_hybrid_device_bytes(r.h, &bytes);
        }
        break;
    }
    if (st == HG_OK && !(r.herr = pinned())) {
      st = HG_ERR_NOMEM;
      j.error = "pinned host memory for the error words";
    } else if (st != HG_OK) {
      j.error = hg_last_error(ctx_);
    }
    if (st != HG_OK) {
      release(r);
      return false;
    }
    // the estimate bounded admission; account what the handle holds
    if (bytes) j.bytes = std::min<uint64_t>(j.bytes, bytes);
    return true;
  }

  bool launch(Job &j, uint32_t k) override {
    Req &r = *static_cast<Req *>(j.payload.get());
    hg_status st = HG_OK;
    switch (r.kind) {
      case Kind::Mpm: st = hg_mpm_step_graph(r.m, r.dt, k); break;
      case Kind::Dem: st = hg_dem_step_graph(r.d, r.dt, k); break;
      case Kind::Hybrid: st = hg_hybrid_step_graph(r.h, r.dt, k); break;
    }
    if (st != HG_OK) {
      j.error = std::string("slice not queued: ") + hg_last_error(ctx_);
      return false;
    }
    // the handles' error words as they stand after this slice, read by poll
    // once the slice's event has completed
    const int32_t *src[3] = {r.m ? r.m->err : (r.h ? r.h->mpm->err : nullptr),
                             r.d ? r.d->err : (r.h ? r.h->dem->err : nullptr),
                             r.h ? r.h->err : nullptr};
    for (int w = 0; w < 3; ++w)
      if (src[w] && hipMemcpyAsync(r.herr + w, src[w], 4, hipMemcpyDeviceToHost, r.stream) != hipSuccess) {
        j.error = "error-word copy not queued";
        return false;
      }
    hipEvent_t e = event();
    if (!e || hipEventRecord(e, r.stream) != hipSuccess) {
      j.error = "slice event not recorded";
      if (e) pool_.push_back(e);
      return false;
    }
    r.events.push_back(e);
    return true;
  }

  int poll(Job &j) override {
    Req &r = *static_cast<Req *>(j.payload.get());
    int n = 0;
    while (!r.events.empty()) {
      const hipError_t q = hipEventQuery(r.events.front());
      if (q == hipErrorNotReady) break;
      if (q != hipSuccess) {
        j.error = std::string("device error: ") + hipGetErrorString(q);
        return -1;
      }
      pool_.push_back(r.events.front());
      r.events.pop_front();
      ++n;
    }
    // a finished slice left an error: retire the request now rather than
    // running its remaining steps on a void state
    if (n > 0) {
      const std::string e = decode_errors(r.herr);
      if (!e.empty()) {
        j.error = e;
        return -1;
      }
    }
    return n;
  }

  bool finish(Job &j) override {
    Req &r = *static_cast<Req *>(j.payload.get());
    hg_status st = HG_OK;
    if (r.kind != Kind::Dem) {
      hg_mpm *m = r.kind == Kind::Mpm ? r.m : r.h->mpm;
      const size_t N = r.mpm.d.npoints;
      r.pq.resize(2 * N);
      r.pv.resize(2 * N);
      r.pb.resize(4 * N);
      r.pJ.resize(N);
      st = r.kind == Kind::Hybrid ? hg_hybrid_sync(r.h) : HG_OK;
      if (st == HG_OK) st = hg_mpm_download_points(m, r.pq.data(), r.pv.data(), r.pb.data(), r.pJ.data());
    }
    if (st == HG_OK && r.kind != Kind::Mpm) {
      hg_dem *d = r.kind == Kind::Dem ? r.d : r.h->dem;
      const size_t B = r.dem.d.nbodies;
      r.bq.resize(3 * B);
      r.bv.resize(3 * B);
      st = hg_dem_download_bodies(d, 0, r.bq.data(), r.bv.data());
    }
    if (st != HG_OK) j.error = hg_last_error(ctx_);
    release(r);
    // the request's inputs are no longer needed
    r.mpm = OwnedMpm();
    r.dem = OwnedDem();
    r.w.clear();
    return st == HG_OK;
  }

  void abort(Job &j) override { release(*static_cast<Req *>(j.payload.get())); }

  double now() override {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
  }

 private:
  hipEvent_t event() {
    if (!pool_.empty()) {
      hipEvent_t e = pool_.back();
      pool_.pop_back();
      return e;
    }
    hipEvent_t e = nullptr;
    hipEventCreateWithFlags(&e, hipEventDisableTiming);
    return e;
  }
  // The handles' error bits (hg_mpm_scene_error, hg_dem_sync, hg_hybrid_sync)
  // as a message; empty when none is set.
  static std::string decode_errors(const int32_t *w) {
    const int32_t m = w[0], d = w[1], c = w[2];
    if ((m | d) & 0x100) return "cell sort: block totals do not match the counts";
    if (m & 1) return "a material point left the grid";
    if (m & 2) return "non-finite material point state";
    if (d & 4) return "more contacts on one body than its capacity";
    if (d & 2) return "non-finite body state";
    if (c & 8) return "coupling node list overflow";
    return std::string();
  }
  int32_t *pinned() {
    int32_t *p = nullptr;
    if (!pinned_.empty()) {
      p = pinned_.back();
      pinned_.pop_back();
    } else if (hipHostMalloc(reinterpret_cast<void **>(&p), 16, hipHostMallocDefault) != hipSuccess) {
      return nullptr;
    }
    p[0] = p[1] = p[2] = p[3] = 0;
    return p;
  }
  void release(Req &r) {
    if (r.stream) hipStreamSynchronize(r.stream);
    for (hipEvent_t e : r.events) pool_.push_back(e);
    r.events.clear();
    if (r.herr) pinned_.push_back(r.herr);
    r.herr = nullptr;
    if (r.m) hg_mpm_destroy(r.m);
    if (r.d) hg_dem_destroy(r.d);
    if (r.h) hg_hybrid_destroy(r.h);
    r.m = nullptr;
    r.d = nullptr;
    r.h = nullptr;
    r.stream = nullptr;
  }
  hg_ctx *ctx_;
  std::vector<hipEvent_t> pool_;
  std::vector<int32_t *> pinned_;
};

}  // namespace

struct hg_server {
  hg_ctx *ctx = nullptr;
  GpuExecutor ex;
  hg::sched::Scheduler sched;
  std::mutex mu;
  std::condition_variable wake;   // new work or stop
  std::condition_variable ended;  // a request ended
  std::vector<uint64_t> cancels;
  bool stop = false;
  std::thread worker;

  hg_server(hg_ctx *c, const hg::sched::Config &cfg) : ctx(c), ex(c), sched(ex, cfg) {}

  void run() {
    hipSetDevice(ctx->device);
    std::unique_lock<std::mutex> lk(mu);
    while (!stop) {
      for (uint64_t id : cancels) sched.cancel(id);
      if (!cancels.empty()) ended.notify_all();
      cancels.clear();
      const uint64_t before = sched.ended_count();
      const bool busy = sched.tick();
      if (sched.ended_count() != before) ended.notify_all();
      if (!busy) {
        wake.wait(lk, [&] { return stop || !cancels.empty() || sched.queued_count() > 0; });
      } else {
        // let submitters in, and give queued slices time to run
        lk.unlock();
        std::this_thread::sleep_for(std::chrono::microseconds(50));
        lk.lock();
      }
    }
  }
};

extern "C" hg_status hg_server_create(hg_ctx *ctx, const hg_server_config *cfg, hg_server **out) {
  if (!ctx || !out) return HG_ERR_INVALID;
  hg::sched::Config c;
  if (cfg) {
    if (cfg->budget_bytes) c.budget_bytes = cfg->budget_bytes;
    if (cfg->max_active) c.max_active = cfg->max_active;
    if (cfg->slice_steps) c.slice_steps = cfg->slice_steps;
    if (cfg->max_inflight) c.max_inflight = cfg->max_inflight;
    if (cfg->max_bypass) c.max_bypass = cfg->max_bypass;
  }
  auto *s = new hg_server(ctx, c);
  s->worker = std::thread([s] { s->run(); });
  *out = s;
  return HG_OK;
}

extern "C" hg_status hg_server_destroy(hg_server *s) {
  if (!s) return HG_OK;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->wake.notify_all();
  if (s->worker.joinable()) s->worker.join();
  // cancel what is left (releases device memory)
  for (uint64_t id = 1;; ++id) {
    Job *j = s->sched.job(id);
    if (!j) break;
    if (j->state == State::Queued || j->state == State::Running) s->sched.cancel(id);
  }
  delete s;
  return HG_OK;
}

static hg_status submit(hg_server *s, std::shared_ptr<Req> r, uint64_t bytes, uint64_t work,
                        uint64_t nsteps, uint64_t *id) {
  Job j;
  j.bytes = bytes;
  j.steps_total = nsteps;
  j.work_per_step = work;
  j.payload = r;
  // the hand-over time, taken before waiting for the worker's lock
  const double t = s->ex.now();
  uint64_t got;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    got = s->sched.submit(std::move(j), t);
  }
  s->wake.notify_all();
  if (id) *id = got;
  return HG_OK;
}

extern "C" hg_status hg_server_submit_mpm(hg_server *s, const hg_mpm_desc *desc, double dt,
                                          uint64_t nsteps, uint64_t *id) {
  if (!s || !desc || !(dt > 0.0)) return HG_ERR_INVALID;
  auto r = std::make_shared<Req>();
  r->kind = Kind::Mpm;
  r->dt = dt;
  r->mpm.take(*desc);
  uint64_t bytes = 0;
  if (hg_mpm_estimate_bytes(desc, 1, &bytes) != HG_OK) return HG_ERR_INVALID;
  return submit(s, r, bytes, desc->npoints, nsteps, id);
}

extern "C" hg_status hg_server_submit_dem(hg_server *s, const hg_dem_desc *desc,
                                          const double *weights, double dt, uint64_t nsteps,
                                          uint64_t *id) {
  if (!s || !desc || !(dt > 0.0)) return HG_ERR_INVALID;
  auto r = std::make_shared<Req>();
  r->kind = Kind::Dem;
  r->dt = dt;
  r->dem.take(*desc);
  if (weights) r->w.assign(weights, weights + desc->nbodies);
  uint64_t bytes = 0;
  if (hg_dem_estimate_bytes(desc, 1, &bytes) != HG_OK) return HG_ERR_INVALID;
  return submit(s, r, bytes, desc->nbodies, nsteps, id);
}

extern "C" hg_status hg_server_submit_hybrid(hg_server *s, const hg_mpm_desc *mpm,
                                             const hg_dem_desc *dem, const double *weights,
                                             double dt, uint64_t nsteps, uint64_t *id) {
  if (!s || !mpm || !dem || !(dt > 0.0)) return HG_ERR_INVALID;
  auto r = std::make_shared<Req>();
  r->kind = Kind::Hybrid;
  r->dt = dt;
  r->mpm.take(*mpm);
  r->dem.take(*dem);
  if (weights) r->w.assign(weights, weights + dem->nbodies);
  uint64_t bytes = 0;
  if (hg_hybrid_estimate_bytes(mpm, dem, &bytes) != HG_OK) return HG_ERR_INVALID;
  return submit(s, r, bytes, uint64_t(mpm->npoints) + dem->nbodies, nsteps, id);
}

extern "C" hg_status hg_server_cancel(hg_server *s, uint64_t id) {
  if (!s) return HG_ERR_INVALID;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    Job *j = s->sched.job(id);
    if (!j) return HG_ERR_INVALID;
    if (j->state == State::Queued) {
      s->sched.cancel(id);  // no device work to undo
      s->ended.notify_all();
      return HG_OK;
    }
    if (j->state != State::Running) return HG_OK;
    s->cancels.push_back(id);
  }
  s->wake.notify_all();
  return HG_OK;
}

extern "C" hg_status hg_server_poll(hg_server *s, uint64_t id, int32_t *state, uint64_t *steps_done) {
  if (!s) return HG_ERR_INVALID;
  std::lock_guard<std::mutex> lk(s->mu);
  const Job *j = s->sched.job(id);
  if (!j) return HG_ERR_INVALID;
  if (state) *state = int32_t(j->state);
  if (steps_done) *steps_done = j->steps_done;
  return HG_OK;
}

static bool ended_state(State st) {
  return st == State::Done || st == State::Failed || st == State::Cancelled;
}

extern "C" hg_status hg_server_wait(hg_server *s, uint64_t id, double timeout_s, int32_t *state) {
  if (!s) return HG_ERR_INVALID;
  std::unique_lock<std::mutex> lk(s->mu);
  const Job *j = s->sched.job(id);
  if (!j) return HG_ERR_INVALID;
  s->ended.wait_for(lk, std::chrono::duration<double>(timeout_s),
                    [&] { return ended_state(s->sched.job(id)->state); });
  if (state) *state = int32_t(s->sched.job(id)->state);
  return HG_OK;
}

extern "C" hg_status hg_server_drain(hg_server *s, double timeout_s) {
  if (!s) return HG_ERR_INVALID;
  std::unique_lock<std::mutex> lk(s->mu);
  const bool ok = s->ended.wait_for(lk, std::chrono::duration<double>(timeout_s), [&] {
    return s->sched.queued_count() == 0 && s->sched.active_count() == 0 && s->cancels.empty();
  });
  return ok ? HG_OK : HG_ERR_STATE;
}

static Req *done_req(hg_server *s, uint64_t id) {
  Job *j = s->sched.job(id);
  if (!j || j->state != State::Done || !j->payload) return nullptr;
  return static_cast<Req *>(j->payload.get());
}

extern "C" hg_status hg_server_result_points(hg_server *s, uint64_t id, double *q, double *v,
                                             double *be_bar, double *J) {
  if (!s) return HG_ERR_INVALID;
  std::lock_guard<std::mutex> lk(s->mu);
  Req *r = done_req(s, id);
  if (!r || r->kind == Kind::Dem) return HG_ERR_INVALID;
  if (q) std::memcpy(q, r->pq.data(), r->pq.size() * 8);
  if (v) std::memcpy(v, r->pv.data(), r->pv.size() * 8);
  if (be_bar) std::memcpy(be_bar, r->pb.data(), r->pb.size() * 8);
  if (J) std::memcpy(J, r->pJ.data(), r->pJ.size() * 8);
  return HG_OK;
}

extern "C" hg_status hg_server_result_bodies(hg_server *s, uint64_t id, double *q, double *v) {
  if (!s) return HG_ERR_INVALID;
  std::lock_guard<std::mutex> lk(s->mu);
  Req *r = done_req(s, id);
  if (!r || r->kind == Kind::Mpm) return HG_ERR_INVALID;
  if (q) std::memcpy(q, r->bq.data(), r->bq.size() * 8);
  if (v) std::memcpy(v, r->bv.data(), r->bv.size() * 8);
  return HG_OK;
}

extern "C" hg_status hg_server_latency(hg_server *s, uint64_t id, double *seconds) {
  if (!s || !seconds) return HG_ERR_INVALID;
  std::lock_guard<std::mutex> lk(s->mu);
  const Job *j = s->sched.job(id);
  if (!j || j->state != State::Done) return HG_ERR_INVALID;
  *seconds = j->done_t - j->submit_t;
  return HG_OK;
}

extern "C" hg_status hg_server_error(hg_server *s, uint64_t id, char *buf, size_t len) {
  if (!s || !buf || !len) return HG_ERR_INVALID;
  std::lock_guard<std::mutex> lk(s->mu);
  const Job *j = s->sched.job(id);
  if (!j) return HG_ERR_INVALID;
  std::strncpy(buf, j->error.c_str(), len - 1);
  buf[len - 1] = '\0';
  return HG_OK;
}

extern "C" hg_status hg_server_release(hg_server *s, uint64_t id) {
  if (!s) return HG_ERR_INVALID;
  std::lock_guard<std::mutex> lk(s->mu);
  Job *j = s->sched.job(id);
  if (!j) return HG_ERR_INVALID;
  if (ended_state(j->state)) j->payload.reset();
  return HG_OK;
}

extern "C" hg_status hg_server_stats_get(hg_server *s, hg_server_stats *out) {
  if (!s || !out) return HG_ERR_INVALID;
  std::lock_guard<std::mutex> lk(s->mu);
  const auto st = s->sched.stats();
  out->submitted = st.submitted;
  out->completed = st.completed;
  out->failed = st.failed;
  out->cancelled = st.cancelled;
  out->active = st.active;
  out->queued = st.queued;
  out->bytes_in_use = st.bytes_in_use;
  out->p50_latency_s = st.p50_latency_s;
  out->p99_latency_s = st.p99_latency_s;
  out->mean_latency_s = st.mean_latency_s;
  out->max_latency_s = st.max_latency_s;
  out->work_steps = st.work_steps;
  out->span_s = st.last_done_t > st.first_submit_t ? st.last_done_t - st.first_submit_t : 0.0;
  return HG_OK;
}
