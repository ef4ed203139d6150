// mpmgrains2d_cli.cpp -- command-line driver for <mpmgrains2d> scenes, the
// counterpart of mpmgrains2d_cli/mpmgrains2d_cli.cpp: read the scene, step it
// with ExplicitIntegrator::flow, report.
#include <algorithm>
#include <cstdio>

#include "cli_common.h"
#include "hg_sim.hpp"

static void report(const hg::SimulationState &st, long it, double dt) {
  const hg::MaterialPoints &mp = st.material_points;
  double ke = 0.0, xmax = -1e300, ymax = -1e300;
  for (size_t p = 0; p < mp.npoints(); ++p) {
    const double vx = mp.v[2 * p], vy = mp.v[2 * p + 1];
    ke += 0.5 * mp.m[p] * (vx * vx + vy * vy);
    xmax = std::max(xmax, mp.q[2 * p]);
    ymax = std::max(ymax, mp.q[2 * p + 1]);
  }
  std::printf("step %ld t %.9g points %zu kinetic_energy %.12e max_x %.12e max_y %.12e\n", it,
              double(it) * dt, mp.npoints(), ke, xmax, ymax);
}

int main(int argc, char **argv) {
  hgcli::Options o;
  const int pr = hgcli::parse(argc, argv, "mpmgrains2d", o);
  if (pr != 0) return pr > 0 ? 0 : 1;
  hg::SimulationState st;
  std::string err;
  if (!hg::readMPMGrains2DXML(o.scene, st, &err)) {
    std::fprintf(stderr, "%s\n", err.c_str());
    return 1;
  }
  const double dt = st.dt.value();
  const long n = hgcli::step_count(o, dt, st.end_time);
  if (n < 0) {
    std::fprintf(stderr, "The scene has no end_time: pass -e or -n.\n");
    return 1;
  }
  hg::setResidency(st, hg::Residency::Resident, o.precision);
  long it0 = 0;
  if (!o.resume.empty()) {
    std::vector<std::uint8_t> blob;
    if (!hgcli::read_snapshot(o.resume, hgcli::kSnapMpm, dt, it0, blob)) return 1;
    hg::restoreDeviceState(st, blob);
  }
  report(st, it0, dt);
  const auto t0 = std::chrono::steady_clock::now();
  const long chunk = o.report_every > 0 ? o.report_every : std::max(n, 1L);
  for (long it = it0; it < n;) {
    const long k = std::min(chunk, n - it);
    hg::ExplicitIntegrator::flow_n(dt, static_cast<unsigned>(k), st);
    it += k;
    if (o.serialize >= 0 &&
        !hgcli::write_snapshot(hgcli::snapshot_path(o.serialize, it), hgcli::kSnapMpm, it, dt,
                               hg::snapshotDeviceState(st)))
      return 1;
    if (o.report_every > 0 && it < n) {
      hg::syncToHost(st);
      report(st, it, dt);
    }
  }
  hg::syncToHost(st);
  const double secs = hgcli::seconds_since(t0);
  const long done = std::max(n - it0, 0L);
  report(st, std::max(n, it0), dt);
  std::printf("Simulation complete: %ld steps in %.3f s (%.4g point-steps/s)\n", done, secs,
              secs > 0 ? double(done) * double(st.material_points.npoints()) / secs : 0.0);
  return 0;
}
