// Scheduler policy tests with a host-only executor whose slices finish on a
// simulated clock (no GPU).  Built and run by tests/test_scheduler.py.
#include <cstdio>
#include <cstdlib>
#include <map>
#include <string>
#include <vector>

#include "hg_sched.h"

using hg::sched::Config;
using hg::sched::Executor;
using hg::sched::Job;
using hg::sched::Scheduler;
using hg::sched::State;

static int failures = 0;
#define CHECK(c)                                                         \
  do {                                                                   \
    if (!(c)) {                                                          \
      std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #c);           \
      ++failures;                                                        \
    }                                                                    \
  } while (0)

// Each step costs step_cost time units; a job's slices run one after the
// other on its own stream; different jobs run side by side.
struct FakeExecutor : Executor {
  double t = 0.0;
  double step_cost = 1.0;
  std::map<uint64_t, std::vector<double>> ends;  // finish time of each queued slice
  std::map<uint64_t, double> stream_free;
  std::map<uint64_t, bool> fail_start, fail_poll, fail_launch;
  // a slice whose end passes this time reports an error (a state that left
  // the valid domain), as the GPU executor's per-slice error words do
  std::map<uint64_t, double> error_after;
  std::map<uint64_t, uint64_t> steps_queued;
  std::map<uint64_t, int> max_inflight_seen;
  std::vector<uint64_t> started, finished, aborted;

  bool start(Job &j) override {
    if (fail_start[j.id]) {
      j.error = "start failed";
      return false;
    }
    started.push_back(j.id);
    stream_free[j.id] = t;
    return true;
  }
  bool launch(Job &j, uint32_t k) override {
    if (fail_launch[j.id]) {
      j.error = "slice not queued";
      return false;
    }
    steps_queued[j.id] += k;
    const double s = std::max(t, stream_free[j.id]);
    const double e = s + step_cost * k;
    stream_free[j.id] = e;
    ends[j.id].push_back(e);
    const int inflight = int(ends[j.id].size());
    max_inflight_seen[j.id] = std::max(max_inflight_seen[j.id], inflight);
    return true;
  }
  int poll(Job &j) override {
    if (fail_poll[j.id]) {
      j.error = "state left the grid";
      return -1;
    }
    int n = 0;
    auto &v = ends[j.id];
    while (!v.empty() && v.front() <= t) {
      const auto it = error_after.find(j.id);
      if (it != error_after.end() && v.front() > it->second) {
        j.error = "non-finite body state";
        return -1;
      }
      v.erase(v.begin());
      ++n;
    }
    return n;
  }
  bool finish(Job &j) override {
    finished.push_back(j.id);
    return true;
  }
  void abort(Job &j) override {
    aborted.push_back(j.id);
    ends.erase(j.id);
  }
  double now() override { return t; }
};

static Job make(uint64_t bytes, uint64_t steps, uint64_t work = 1) {
  Job j;
  j.bytes = bytes;
  j.steps_total = steps;
  j.work_per_step = work;
  return j;
}

static void run_until_idle(Scheduler &s, FakeExecutor &ex, double dt = 1.0, int max_ticks = 100000) {
  for (int k = 0; k < max_ticks; ++k) {
    if (!s.tick()) return;
    ex.t += dt;
  }
  CHECK(false && "scheduler did not drain");
}

static void test_admission_by_budget() {
  FakeExecutor ex;
  Config c;
  c.budget_bytes = 100;
  c.slice_steps = 10;
  Scheduler s(ex, c);
  const uint64_t a = s.submit(make(60, 100));
  const uint64_t b = s.submit(make(50, 10));
  const uint64_t d = s.submit(make(30, 10));
  s.tick();
  CHECK(s.job(a)->state == State::Running);
  CHECK(s.job(b)->state == State::Queued);   // 60 + 50 > 100
  CHECK(s.job(d)->state == State::Running);  // passes b: 60 + 30 fits
  CHECK(s.stats().bytes_in_use == 90);
  run_until_idle(s, ex);
  CHECK(s.job(a)->state == State::Done && s.job(b)->state == State::Done &&
        s.job(d)->state == State::Done);
  CHECK(s.stats().bytes_in_use == 0);
  CHECK(s.stats().completed == 3);
  // every job ran exactly its steps, never more than two slices queued
  for (uint64_t id : {a, b, d}) {
    CHECK(s.job(id)->steps_done == s.job(id)->steps_total);
    CHECK(ex.max_inflight_seen[id] <= 2);
  }
}

static void test_bypass_is_bounded() {
  FakeExecutor ex;
  Config c;
  c.budget_bytes = 100;
  c.max_bypass = 2;
  c.slice_steps = 5;
  Scheduler s(ex, c);
  const uint64_t first = s.submit(make(50, 1000));  // holds half the budget for long
  const uint64_t big = s.submit(make(80, 10));
  std::vector<uint64_t> small;
  for (int k = 0; k < 5; ++k) small.push_back(s.submit(make(10, 50)));
  s.tick();
  // big waits; the first two small ones pass it, then admission stops behind it
  CHECK(s.job(first)->state == State::Running);
  CHECK(s.job(big)->state == State::Queued);
  CHECK(s.job(small[0])->state == State::Running);
  CHECK(s.job(small[1])->state == State::Running);
  CHECK(s.job(small[2])->state == State::Queued);
  CHECK(s.job(big)->bypassed == 2);
  run_until_idle(s, ex);
  // big ran before the small ones that were held behind it
  CHECK(s.job(big)->admit_t <= s.job(small[2])->admit_t);
  CHECK(s.stats().completed == 7);
}

static void test_over_budget_request_fails_at_submit() {
  FakeExecutor ex;
  Config c;
  c.budget_bytes = 100;
  Scheduler s(ex, c);
  const uint64_t a = s.submit(make(101, 10));
  CHECK(s.job(a)->state == State::Failed);
  CHECK(!s.job(a)->error.empty());
  CHECK(!s.tick());
}

static void test_cancel_and_failure_are_isolated() {
  FakeExecutor ex;
  Config c;
  c.budget_bytes = 1000;
  c.slice_steps = 10;
  Scheduler s(ex, c);
  const uint64_t a = s.submit(make(100, 200));
  const uint64_t b = s.submit(make(100, 200));
  const uint64_t d = s.submit(make(100, 200));
  const uint64_t q = s.submit(make(2000 - 1800, 50));
  s.tick();
  ex.t += 5;
  CHECK(s.cancel(b));
  CHECK(s.job(b)->state == State::Cancelled);
  CHECK(!s.cancel(b));
  ex.fail_poll[d] = true;
  run_until_idle(s, ex);
  CHECK(s.job(a)->state == State::Done);
  CHECK(s.job(d)->state == State::Failed && s.job(d)->error == "state left the grid");
  CHECK(s.job(q)->state == State::Done);
  const auto st = s.stats();
  CHECK(st.completed == 2 && st.failed == 1 && st.cancelled == 1);
  CHECK(st.bytes_in_use == 0);
  // a start failure retires that request only
  const uint64_t e = s.submit(make(10, 10));
  ex.fail_start[e] = true;
  const uint64_t f = s.submit(make(10, 10));
  run_until_idle(s, ex);
  CHECK(s.job(e)->state == State::Failed && s.job(f)->state == State::Done);
}

static void test_latency_quantiles() {
  std::vector<double> v;
  for (int k = 1; k <= 100; ++k) v.push_back(double(k));
  CHECK(Scheduler::quantile(v, 0.50) == 50.0);
  CHECK(Scheduler::quantile(v, 0.99) == 99.0);
  CHECK(Scheduler::quantile({3.0}, 0.99) == 3.0);
  // completed-request latencies: equal jobs side by side finish together
  FakeExecutor ex;
  Config c;
  c.slice_steps = 10;
  Scheduler s(ex, c);
  for (int k = 0; k < 4; ++k) s.submit(make(1, 40, 1000));
  run_until_idle(s, ex);
  const auto st = s.stats();
  CHECK(st.completed == 4);
  CHECK(st.p50_latency_s == st.p99_latency_s);
  CHECK(st.p99_latency_s >= 40.0);
  CHECK(st.work_steps == 4 * 40 * 1000.0);
}

static void test_continuous_admission() {
  // requests arriving while others run start without waiting for them to end
  FakeExecutor ex;
  Config c;
  c.slice_steps = 10;
  c.max_active = 2;
  Scheduler s(ex, c);
  const uint64_t a = s.submit(make(1, 1000));
  s.tick();
  ex.t += 1;
  const uint64_t b = s.submit(make(1, 20));
  s.tick();
  CHECK(s.job(b)->state == State::Running);  // admitted beside a
  const uint64_t d = s.submit(make(1, 20));
  s.tick();
  CHECK(s.job(d)->state == State::Queued);  // both slots taken
  run_until_idle(s, ex);
  CHECK(s.job(b)->done_t < s.job(a)->done_t);
  CHECK(s.job(d)->admit_t >= s.job(b)->done_t);
  CHECK(s.job(d)->done_t < s.job(a)->done_t);
}

static void test_slice_errors_retire_early() {
  // an error seen after a slice retires the request without running the
  // rest of its steps; a slice that cannot be queued retires it at once
  FakeExecutor ex;
  Config c;
  c.slice_steps = 10;
  Scheduler s(ex, c);
  const uint64_t a = s.submit(make(1, 1000));
  const uint64_t b = s.submit(make(1, 1000));
  const uint64_t d = s.submit(make(1, 60));
  ex.error_after[a] = 25.0;  // the slice ending at 30 carries the error
  ex.fail_launch[b] = true;
  run_until_idle(s, ex);
  CHECK(s.job(a)->state == State::Failed && s.job(a)->error == "non-finite body state");
  CHECK(s.job(a)->steps_done == 20);
  CHECK(ex.steps_queued[a] <= 40);  // at most max_inflight slices past the error
  CHECK(s.job(b)->state == State::Failed && s.job(b)->error == "slice not queued");
  CHECK(ex.steps_queued[b] == 0);
  CHECK(s.job(d)->state == State::Done);
  const auto st = s.stats();
  CHECK(st.failed == 2 && st.completed == 1 && st.bytes_in_use == 0 && st.active == 0);
  CHECK(ex.aborted.size() == 2);
}

int main() {
  test_admission_by_budget();
  test_bypass_is_bounded();
  test_over_budget_request_fails_at_submit();
  test_cancel_and_failure_are_isolated();
  test_latency_quantiles();
  test_continuous_admission();
  test_slice_errors_retire_early();
  if (failures) {
    std::printf("%d checks failed\n", failures);
    return 1;
  }
  std::printf("all scheduler checks passed\n");
  return 0;
}
