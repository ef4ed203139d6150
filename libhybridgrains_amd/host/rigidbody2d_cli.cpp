// rigidbody2d_cli.cpp -- command-line driver for <rigidbody2d_scene> circle
// scenes with a penalty_impact_map, the counterpart of
// rigidbody2dcli/rigidbody2d_cli.cpp: read, step with RigidBody2DSim::flow,
// report.
#include <algorithm>
#include <cstdio>

#include "cli_common.h"
#include "hg_sim.hpp"

static void report(const hg::RigidBody2DState &s, long it, double dt) {
  double ke = 0.0, xmax = -1e300, ymax = -1e300;
  for (size_t b = 0; b < s.nbodies(); ++b) {
    for (int k = 0; k < 3; ++k) ke += 0.5 * s.M[3 * b + k] * s.v[3 * b + k] * s.v[3 * b + k];
    xmax = std::max(xmax, s.q[3 * b]);
    ymax = std::max(ymax, s.q[3 * b + 1]);
  }
  std::printf("step %ld t %.9g bodies %zu kinetic_energy %.12e max_x %.12e max_y %.12e\n", it,
              double(it) * dt, s.nbodies(), ke, xmax, ymax);
}

int main(int argc, char **argv) {
  hgcli::Options o;
  const int pr = hgcli::parse(argc, argv, "rigidbody2d", o);
  if (pr != 0) return pr > 0 ? 0 : 1;
  hg::RigidBody2DScene sc;
  std::string err;
  if (!hg::readRigidBody2DXML(o.scene, sc, &err)) {
    std::fprintf(stderr, "%s\n", err.c_str());
    return 1;
  }
  const double dt = sc.dt.value();
  const long n = hgcli::step_count(o, dt, sc.end_time);
  if (n < 0) {
    std::fprintf(stderr, "The scene has no end_time: pass -e or -n.\n");
    return 1;
  }
  hg::RigidBody2DSim sim(sc.state, sc.spatial_grid_scale, o.precision);
  hg::PenaltyImpactFrictionMap ifmap(sc.penalty);
  sim.setResidency(hg::Residency::Resident);
  long it0 = 0;
  if (!o.resume.empty()) {
    std::vector<std::uint8_t> blob;
    if (!hgcli::read_snapshot(o.resume, hgcli::kSnapDem, dt, it0, blob)) return 1;
    sim.restore(blob, ifmap);
  }
  report(sim.getState(), it0, dt);
  const auto t0 = std::chrono::steady_clock::now();
  const long every = o.report_every > 0 ? o.report_every : std::max(n, 1L);
  for (long it = it0; it < n; ++it) {
    sim.flow(static_cast<unsigned>(it), sc.dt, ifmap, sc.reduce_bandwidth);
    const bool mark = (it + 1) % every == 0 || it + 1 == n;
    if (mark && o.serialize >= 0 &&
        !hgcli::write_snapshot(hgcli::snapshot_path(o.serialize, it + 1), hgcli::kSnapDem,
                               it + 1, dt, sim.snapshot(ifmap)))
      return 1;
    if (o.report_every > 0 && (it + 1) % o.report_every == 0 && it + 1 < n) {
      sim.syncToHost();
      report(sim.getState(), it + 1, dt);
    }
  }
  sim.syncToHost();
  const double secs = hgcli::seconds_since(t0);
  const long done = std::max(n - it0, 0L);
  report(sim.getState(), std::max(n, it0), dt);
  std::printf("Simulation complete: %ld steps in %.3f s (%.4g body-steps/s)\n", done, secs,
              secs > 0 ? double(done) * double(sc.state.nbodies()) / secs : 0.0);
  return 0;
}
