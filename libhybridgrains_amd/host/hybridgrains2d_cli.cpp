// hybridgrains2d_cli.cpp -- command-line driver for <hybridgrains2d_scene>
// files, the counterpart of hybridgrains2dnewcli/hybridgrains2d_cli.cpp: read
// the hybrid scene and the two scenes it names, step with
// HybridGrains2DSim::stepSystem, report.
//
// Grains start with weight 1.  A scene with a coupling_force element runs
// the zone resampling as the reference does (initializeAvoidAVoid, then every
// avoidAVoidFreq + 1 steps), which re-seeds the continuum and re-weights both
// sides.
#include <algorithm>
#include <cstdio>

#include "cli_common.h"
#include "hg_sim.hpp"

int main(int argc, char **argv) {
  hgcli::Options o;
  const int pr = hgcli::parse(argc, argv, "hybridgrains2d", o);
  if (pr != 0) return pr > 0 ? 0 : 1;
  hg::HybridIntegratorSettings hs;
  std::string err;
  hg::RigidBody2DScene dem;
  hg::SimulationState mpm;
  if (!hg::readHybridGrains2DXML(o.scene, hs, &err) ||
      !hg::readRigidBody2DXML(hs.discrete_file_name, dem, &err) ||
      !hg::readMPMGrains2DXML(hs.continuum_file_name, mpm, &err)) {
    std::fprintf(stderr, "%s\n", err.c_str());
    return 1;
  }
  const double dt = hs.overall_dt.value();
  if (dt != mpm.dt.value() || dt != dem.dt.value()) {
    std::fprintf(stderr,
                 "Currently only support [discrete dt == continuum dt] in 2d hybrid sim.\n");
    return 1;
  }
  const long n = hgcli::step_count(o, dt, hs.end_time);
  std::vector<double> w(dem.state.nbodies(), 1.0);
  hg::HybridGrains2DSim sim(hs, dem, mpm, w, o.precision);
  long it0 = 0;
  if (!o.resume.empty()) {
    std::vector<std::uint8_t> blob;
    if (!hgcli::read_snapshot(o.resume, hgcli::kSnapHybrid, dt, it0, blob)) return 1;
    sim.restore(blob);
  }
  const auto t0 = std::chrono::steady_clock::now();
  const long chunk = o.report_every > 0 ? o.report_every : std::max(n, 1L);
  for (long it = it0; it < n;) {
    const long k = std::min(chunk, n - it);
    sim.stepSystem(static_cast<unsigned>(k));
    it += k;
    if (o.serialize >= 0 &&
        !hgcli::write_snapshot(hgcli::snapshot_path(o.serialize, it), hgcli::kSnapHybrid, it, dt,
                               sim.snapshot()))
      return 1;
    if (o.report_every > 0) {
      const auto c = sim.coupledGrains();
      const hg::RigidBody2DState &d = sim.discreteState();
      const hg::SimulationState &cs = sim.continuumState();
      double mass = 0.0;
      for (double m : cs.material_points.m) mass += m;
      std::printf("step %ld t %.9g coupled_grains %ld grains %zu points %zu resamplings %ld "
                  "point_mass %.9g\n",
                  it, double(it) * dt, static_cast<long>(std::count(c.begin(), c.end(), 1)),
                  d.nbodies(), cs.material_points.npoints(), sim.resampleCount(), mass);
    }
  }
  const hg::RigidBody2DState &d = sim.discreteState();
  const hg::SimulationState &c = sim.continuumState();
  const double secs = hgcli::seconds_since(t0);
  std::printf("Simulation complete: %ld steps in %.3f s; %zu grains, %zu points\n",
              std::max(n - it0, 0L), secs, d.nbodies(), c.material_points.npoints());
  return 0;
}
