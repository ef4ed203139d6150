// hg_sched.h -- the request scheduler behind the scene server: admission by
// device-memory footprint, continuous batching of resident scenes in slices
// of steps, completion and retirement.  It drives an Executor (the GPU one in
// hg_server.hip; a host-only fake in the scheduler tests), so the policy is
// testable without a GPU.
//
// Policy
//  - Admission: queued requests are admitted oldest first while their
//    footprint fits the remaining budget and an active slot is free.  A
//    request that does not fit may be passed by smaller ones at most
//    max_bypass times; after that it blocks admission until it fits, so
//    large requests are not starved.
//  - Stepping: every active request keeps up to max_inflight slices of
//    slice_steps steps queued on the device; slices of different requests
//    run concurrently (one stream each).
//  - Completion: when all steps of a request have finished, its results are
//    copied out, its device memory is released and its latency (submit to
//    completion) is recorded.  A request whose slice fails to launch, or
//    whose finished slice reports an error (its state left the valid domain,
//    a contact overflow, a device error), is retired alone with the error as
//    soon as that is seen, without running its remaining steps.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <deque>
#include <map>
#include <memory>
#include <string>
#include <vector>

namespace hg {
namespace sched {

enum class State : int32_t { Queued = 0, Running = 1, Done = 2, Failed = 3, Cancelled = 4 };

struct Job {
  uint64_t id = 0;
  uint64_t bytes = 0;        // device footprint while active
  uint64_t steps_total = 0;
  uint64_t steps_launched = 0;
  uint64_t steps_done = 0;
  uint64_t work_per_step = 0;  // particles + bodies
  uint32_t bypassed = 0;
  uint32_t inflight = 0;
  double submit_t = 0, admit_t = 0, done_t = 0;
  State state = State::Queued;
  std::string error;
  std::shared_ptr<void> payload;  // executor-owned request data and results
};

class Executor {
 public:
  virtual ~Executor() = default;
  // Allocates and uploads the job; false (with job.error) on failure.
  virtual bool start(Job &job) = 0;
  // Queues k more steps of the job on the device; false (with job.error) if
  // they could not be queued.
  virtual bool launch(Job &job, uint32_t k) = 0;
  // Number of queued slices of the job that have finished since the last
  // call (in launch order); -1 with job.error if the job failed.
  virtual int poll(Job &job) = 0;
  // Copies the results out and releases the device memory.
  virtual bool finish(Job &job) = 0;
  // Releases the device memory of an active job without results.
  virtual void abort(Job &job) = 0;
  virtual double now() = 0;
};

struct Config {
  uint64_t budget_bytes = 240ull << 30;
  uint32_t max_active = 64;
  uint32_t slice_steps = 50;
  uint32_t max_inflight = 2;
  uint32_t max_bypass = 8;
};

struct Stats {
  uint64_t submitted = 0, completed = 0, failed = 0, cancelled = 0;
  uint64_t active = 0, queued = 0, bytes_in_use = 0;
  double p50_latency_s = 0, p99_latency_s = 0, mean_latency_s = 0, max_latency_s = 0;
  double work_steps = 0;    // particle-steps (+ body-steps) of completed requests
  double first_submit_t = 0, last_done_t = 0;
};

class Scheduler {
 public:
  Scheduler(Executor &ex, const Config &cfg) : ex_(ex), cfg_(cfg) {}

  uint64_t submit(Job job) { return submit(std::move(job), ex_.now()); }

  // submit_t: when the caller handed the request over (before any wait for
  // the scheduler), so that latencies include that wait.
  uint64_t submit(Job job, double submit_t) {
    job.id = ++next_id_;
    job.submit_t = submit_t;
    job.state = State::Queued;
    if (stats_.submitted == 0) stats_.first_submit_t = job.submit_t;
    ++stats_.submitted;
    if (job.bytes > cfg_.budget_bytes) {
      job.state = State::Failed;
      job.error = "request needs more device memory than the server's budget";
      job.done_t = job.submit_t;
      ++stats_.failed;
      jobs_[job.id] = std::move(job);
      return next_id_;
    }
    queue_.push_back(job.id);
    jobs_[job.id] = std::move(job);
    return next_id_;
  }

  // Cancels a queued or running job; false if it had already ended.
  bool cancel(uint64_t id) {
    auto it = jobs_.find(id);
    if (it == jobs_.end()) return false;
    Job &j = it->second;
    if (j.state == State::Queued) {
      queue_.erase(std::remove(queue_.begin(), queue_.end(), id), queue_.end());
    } else if (j.state == State::Running) {
      ex_.abort(j);
      active_.erase(std::remove(active_.begin(), active_.end(), id), active_.end());
      bytes_ -= j.bytes;
    } else {
      return false;
    }
    j.state = State::Cancelled;
    j.done_t = ex_.now();
    ++stats_.cancelled;
    return true;
  }

  // One scheduling pass; returns true while work remains.
  bool tick() {
    progress();
    admit();
    feed();
    return !queue_.empty() || !active_.empty();
  }

  const Job *job(uint64_t id) const {
    auto it = jobs_.find(id);
    return it == jobs_.end() ? nullptr : &it->second;
  }
  Job *job(uint64_t id) {
    auto it = jobs_.find(id);
    return it == jobs_.end() ? nullptr : &it->second;
  }

  Stats stats() const {
    Stats s = stats_;
    s.active = active_.size();
    s.queued = queue_.size();
    s.bytes_in_use = bytes_;
    std::vector<double> lat = latencies_;
    if (!lat.empty()) {
      std::sort(lat.begin(), lat.end());
      s.p50_latency_s = quantile(lat, 0.50);
      s.p99_latency_s = quantile(lat, 0.99);
      double sum = 0;
      for (double x : lat) sum += x;
      s.mean_latency_s = sum / double(lat.size());
      s.max_latency_s = lat.back();
    }
    return s;
  }

  // Nearest-rank quantile of sorted values.
  static double quantile(const std::vector<double> &sorted, double q) {
    if (sorted.empty()) return 0.0;
    size_t k = size_t(std::ceil(q * double(sorted.size())));
    if (k == 0) k = 1;
    return sorted[std::min(k, sorted.size()) - 1];
  }

  const Config &config() const { return cfg_; }
  size_t queued_count() const { return queue_.size(); }
  size_t active_count() const { return active_.size(); }
  uint64_t ended_count() const { return stats_.completed + stats_.failed + stats_.cancelled; }

 private:
  void admit() {
    for (auto it = queue_.begin(); it != queue_.end();) {
      if (active_.size() >= cfg_.max_active) return;
      Job &j = jobs_[*it];
      if (bytes_ + j.bytes > cfg_.budget_bytes) {
        // this one waits; let smaller ones pass only a bounded number of times
        if (j.bypassed >= cfg_.max_bypass) return;
        ++it;
        continue;
      }
      // passing older waiting requests: each may be passed max_bypass times
      bool held = false;
      for (auto o = queue_.begin(); o != it; ++o) held |= jobs_[*o].bypassed >= cfg_.max_bypass;
      if (held) return;
      for (auto o = queue_.begin(); o != it; ++o) ++jobs_[*o].bypassed;
      const uint64_t id = *it;
      it = queue_.erase(it);
      j.admit_t = ex_.now();
      if (!ex_.start(j)) {
        j.state = State::Failed;
        j.done_t = ex_.now();
        ++stats_.failed;
        continue;
      }
      j.state = State::Running;
      bytes_ += j.bytes;
      active_.push_back(id);
    }
  }

  void feed() {
    std::vector<uint64_t> failed;
    for (uint64_t id : active_) {
      Job &j = jobs_[id];
      while (j.inflight < cfg_.max_inflight && j.steps_launched < j.steps_total) {
        const uint64_t k = std::min<uint64_t>(cfg_.slice_steps, j.steps_total - j.steps_launched);
        if (!ex_.launch(j, uint32_t(k))) {
          failed.push_back(id);
          break;
        }
        slices_[id].push_back(k);
        j.steps_launched += k;
        ++j.inflight;
      }
    }
    for (uint64_t id : failed) {
      Job &j = jobs_[id];
      ex_.abort(j);
      j.state = State::Failed;
      j.done_t = ex_.now();
      ++stats_.failed;
      bytes_ -= j.bytes;
      slices_.erase(id);
      active_.erase(std::remove(active_.begin(), active_.end(), id), active_.end());
    }
  }

  void progress() {
    for (size_t a = 0; a < active_.size();) {
      const uint64_t id = active_[a];
      Job &j = jobs_[id];
      const int done = ex_.poll(j);
      bool retire = false;
      if (done < 0) {
        ex_.abort(j);
        j.state = State::Failed;
        retire = true;
        ++stats_.failed;
      } else {
        auto &q = slices_[id];
        for (int k = 0; k < done && !q.empty(); ++k) {
          j.steps_done += q.front();
          q.pop_front();
          --j.inflight;
        }
        if (j.steps_done == j.steps_total) {
          if (ex_.finish(j)) {
            j.state = State::Done;
            ++stats_.completed;
            stats_.work_steps += double(j.work_per_step) * double(j.steps_total);
          } else {
            j.state = State::Failed;
            ++stats_.failed;
          }
          retire = true;
        }
      }
      if (retire) {
        j.done_t = ex_.now();
        if (j.state == State::Done) {
          latencies_.push_back(j.done_t - j.submit_t);
          stats_.last_done_t = std::max(stats_.last_done_t, j.done_t);
        }
        bytes_ -= j.bytes;
        slices_.erase(id);
        active_.erase(active_.begin() + long(a));
      } else {
        ++a;
      }
    }
  }

  Executor &ex_;
  Config cfg_;
  uint64_t next_id_ = 0;
  uint64_t bytes_ = 0;
  std::map<uint64_t, Job> jobs_;
  std::deque<uint64_t> queue_;
  std::vector<uint64_t> active_;
  std::map<uint64_t, std::deque<uint64_t>> slices_;
  std::vector<double> latencies_;
  Stats stats_;
};

}  // namespace sched
}  // namespace hg
