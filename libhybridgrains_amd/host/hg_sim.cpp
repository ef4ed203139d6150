// hg_sim.cpp -- the reference-named stepping API over the C-ABI.
#include "hg_sim.hpp"

#include <cstdio>
#include <cstdlib>

namespace hg {

// ---- device -------------------------------------------------------------------

Device::Device(int device) {
  const hg_status st = hg_ctx_create(device, &ctx_);
  if (st != HG_OK) {
    std::fprintf(stderr, "hg: cannot open GPU %d (status %d)\n", device, st);
    std::exit(EXIT_FAILURE);
  }
}

Device::~Device() {
  if (ctx_) hg_ctx_destroy(ctx_);
}

Device &Device::defaultDevice() {
  static Device d(0);
  return d;
}

void Device::check(hg_status st, const char *where) const {
  if (st == HG_OK) return;
  const char *msg = hg_last_error(ctx_);
  std::fprintf(stderr, "hg: %s failed (status %d): %s\n", where, st, msg ? msg : "");
  std::exit(EXIT_FAILURE);
}

// ---- MPM ----------------------------------------------------------------------

MpmDevice::MpmDevice(Device &dev, const SimulationState &state, int precision) : dev_(dev) {
  const hg_mpm_desc d = state.desc(precision);
  dev_.check(hg_mpm_create(dev_.ctx(), &d, &h_), "hg_mpm_create");
}

MpmDevice::~MpmDevice() {
  if (h_) hg_mpm_destroy(h_);
}

void MpmDevice::download(SimulationState &state) const {
  MaterialPoints &mp = state.material_points;
  dev_.check(hg_mpm_download_points(h_, mp.q.data(), mp.v.data(), mp.be_bar.data(), mp.J.data()),
             "ExplicitIntegrator (state check)");
}

MpmBinding &MpmBinding::of(const SimulationState &state) {
  if (!state.gpu_binding.p) state.gpu_binding.p = std::make_shared<MpmBinding>();
  return *static_cast<MpmBinding *>(state.gpu_binding.p.get());
}

void markHostModified(SimulationState &state) { MpmBinding::of(state).host_modified = true; }

void setResidency(SimulationState &state, Residency r, int precision) {
  MpmBinding &b = MpmBinding::of(state);
  b.residency = r;
  if (b.precision != precision) {
    b.precision = precision;
    b.host_modified = true;
  }
}

void syncToHost(SimulationState &state) {
  MpmBinding &b = MpmBinding::of(state);
  if (b.gpu && !b.host_modified) b.gpu->download(state);
}

static hg_mpm *bound(SimulationState &state) {
  MpmBinding &b = MpmBinding::of(state);
  if (!b.gpu || b.host_modified) {
    b.gpu.reset();
    Device &dev = b.device ? *b.device : Device::defaultDevice();
    b.gpu = std::make_unique<MpmDevice>(dev, state, b.precision);
    b.host_modified = false;
  }
  return b.gpu->handle();
}

static Device &device_of(SimulationState &state) {
  MpmBinding &b = MpmBinding::of(state);
  return b.device ? *b.device : Device::defaultDevice();
}

template <typename H, typename F>
static std::vector<std::uint8_t> snapshot_of(Device &dev, H *h, F fn, const char *where) {
  std::uint64_t n = 0;
  dev.check(fn(h, nullptr, &n), where);
  std::vector<std::uint8_t> blob(n);
  dev.check(fn(h, blob.data(), &n), where);
  blob.resize(n);
  return blob;
}

std::vector<std::uint8_t> snapshotDeviceState(SimulationState &state) {
  return snapshot_of(device_of(state), bound(state), hg_mpm_snapshot, "hg_mpm_snapshot");
}

void restoreDeviceState(SimulationState &state, const std::vector<std::uint8_t> &blob) {
  hg_mpm *h = bound(state);
  device_of(state).check(hg_mpm_restore(h, blob.data(), blob.size()), "hg_mpm_restore");
  MpmBinding::of(state).gpu->download(state);
}

static void after_call(SimulationState &state) {
  MpmBinding &b = MpmBinding::of(state);
  if (b.residency == Residency::SyncEveryCall) b.gpu->download(state);
}

namespace ExplicitIntegrator {

void flow(const double &dt, SimulationState &state) { flow_n(dt, 1, state); }

void flow_n(const double &dt, unsigned nsteps, SimulationState &state) {
  hg_mpm *h = bound(state);
  device_of(state).check(hg_mpm_step(h, dt, nsteps), "ExplicitIntegrator::flow");
  after_call(state);
}

// flow_zero_phase takes no dt (ExplicitIntegrator.h:10); the engine's grid
// pass also prepares the force and momentum update, which phase one then
// reruns if its dt differs from the last one used.
void flow_zero_phase(SimulationState &state) {
  hg_mpm *h = bound(state);
  device_of(state).check(hg_mpm_phase(h, 0, 0.0), "ExplicitIntegrator::flow_zero_phase");
}

void flow_first_phase(const double &dt, SimulationState &state) {
  hg_mpm *h = bound(state);
  device_of(state).check(hg_mpm_phase(h, 1, dt), "ExplicitIntegrator::flow_first_phase");
}

void flow_second_phase(const double &dt, SimulationState &state) {
  hg_mpm *h = bound(state);
  device_of(state).check(hg_mpm_phase(h, 2, dt), "ExplicitIntegrator::flow_second_phase");
  after_call(state);
}

}  // namespace ExplicitIntegrator

namespace MPMGrains2DSim {
void flow(const double &dt, SimulationState &state) { ExplicitIntegrator::flow(dt, state); }
void flow_zero_phase(SimulationState &state) { ExplicitIntegrator::flow_zero_phase(state); }
void flow_first_phase(const double &dt, SimulationState &state) {
  ExplicitIntegrator::flow_first_phase(dt, state);
}
void flow_second_phase(const double &dt, SimulationState &state) {
  ExplicitIntegrator::flow_second_phase(dt, state);
}
}  // namespace MPMGrains2DSim

// ---- DEM ----------------------------------------------------------------------

static hg_dem_desc dem_desc(const RigidBody2DState &s, const PenaltyImpactFrictionMapSettings &p,
                            double scale, int precision, std::vector<double> &px,
                            std::vector<double> &pn) {
  RigidBody2DScene tmp;  // reuse the scene's cell-width rule
  hg_dem_desc d{};
  d.precision = precision;
  d.nbodies = static_cast<std::uint32_t>(s.nbodies());
  d.q = s.q.data();
  d.v = s.v.data();
  d.mass = s.M.data();
  d.radius = s.radius.data();
  d.fixed = s.fixed.data();
  d.gravity[0] = s.gravity[0];
  d.gravity[1] = s.gravity[1];
  px.clear();
  pn.clear();
  for (const RigidBody2DStaticPlane &pl : s.planes) {
    px.insert(px.end(), {pl.x[0], pl.x[1]});
    pn.insert(pn.end(), {pl.n[0], pl.n[1]});
  }
  d.nplanes = static_cast<std::uint32_t>(s.planes.size());
  d.plane_x = px.data();
  d.plane_n = pn.data();
  d.kn = p.kn;
  d.kt = p.kt;
  d.gamman = p.gamman;
  d.gammat = p.gammat;
  d.mu = p.mu;
  tmp.state.radius = s.radius;
  d.cell_width = tmp.cellWidth(scale);
  return d;
}

RigidBody2DSim::RigidBody2DSim(const RigidBody2DState &state, double spatial_grid_scale,
                               int precision, Device *device)
    : dev_(device ? *device : Device::defaultDevice()),
      state_(state),
      weights_(state.nbodies(), 0.0),  // RigidBody2DState.cpp:116 starts them at 0
      scale_(spatial_grid_scale),
      precision_(precision) {}

RigidBody2DSim::~RigidBody2DSim() {
  if (h_) hg_dem_destroy(h_);
}

void RigidBody2DSim::ensure(const PenaltyImpactFrictionMap &ifmap) {
  if (h_ && !host_modified_) return;
  if (h_) {
    hg_dem_destroy(h_);
    h_ = nullptr;
  }
  std::vector<double> px, pn;
  const hg_dem_desc d = dem_desc(state_, ifmap.settings(), scale_, precision_, px, pn);
  dev_.check(hg_dem_create(dev_.ctx(), &d, &h_), "hg_dem_create");
  host_modified_ = false;
  weighted_on_device_ = false;
}

void RigidBody2DSim::syncToHost() {
  if (!h_ || host_modified_) return;
  dev_.check(hg_dem_download_bodies(h_, 0, state_.q.data(), state_.v.data()),
             "RigidBody2DSim (state check)");
}

void RigidBody2DSim::step(double dt, bool weighted) {
  if (weighted) {
    if (!weighted_on_device_ || uploaded_weights_ != weights_) {
      dev_.check(hg_dem_set_weights(h_, 0, weights_.data()), "hg_dem_set_weights");
      uploaded_weights_ = weights_;
      weighted_on_device_ = true;
    }
  } else if (weighted_on_device_) {
    dev_.check(hg_dem_set_weights(h_, 0, nullptr), "hg_dem_set_weights");
    weighted_on_device_ = false;
  }
  dev_.check(hg_dem_step(h_, dt, 1), weighted ? "RigidBody2DSim::flowWithWeight"
                                              : "RigidBody2DSim::flow");
  if (residency_ == Residency::SyncEveryCall) syncToHost();
}

void RigidBody2DSim::flow(unsigned, const Rational &dt, const PenaltyImpactFrictionMap &ifmap,
                          bool) {
  ensure(ifmap);
  step(dt.value(), false);
}

void RigidBody2DSim::flowWithWeight(unsigned, const Rational &dt,
                                    const PenaltyImpactFrictionMap &ifmap, bool) {
  ensure(ifmap);
  step(dt.value(), true);
}

std::vector<std::uint8_t> RigidBody2DSim::snapshot(const PenaltyImpactFrictionMap &ifmap) {
  ensure(ifmap);
  return snapshot_of(dev_, h_, hg_dem_snapshot, "hg_dem_snapshot");
}

void RigidBody2DSim::restore(const std::vector<std::uint8_t> &blob,
                             const PenaltyImpactFrictionMap &ifmap) {
  ensure(ifmap);
  dev_.check(hg_dem_restore(h_, blob.data(), blob.size()), "hg_dem_restore");
  dev_.check(hg_dem_download_bodies(h_, 0, state_.q.data(), state_.v.data()), "hg_dem_restore");
}

// ---- hybrid -------------------------------------------------------------------

HybridGrains2DSim::HybridGrains2DSim(const HybridIntegratorSettings &settings,
                                     const RigidBody2DScene &discrete,
                                     const SimulationState &continuum,
                                     const std::vector<double> &dem_weights, int precision,
                                     Device *device)
    : dev_(device ? *device : Device::defaultDevice()),
      settings_(settings),
      discrete_(discrete.state),
      continuum_(continuum) {
  if (settings_.integrator_style != "prediction_correction_advection_node_node") {
    std::fprintf(stderr, "unsupported integrator style\n");
    std::exit(EXIT_FAILURE);
  }
  if (continuum_.material_points.npoints() == 0) {
    std::fprintf(stderr, "hg: the hybrid step needs continuum points (the grains use their hl)\n");
    std::exit(EXIT_FAILURE);
  }
  if (dem_weights.size() != discrete_.nbodies()) {
    std::fprintf(stderr, "hg: one mass weight per grain is needed\n");
    std::exit(EXIT_FAILURE);
  }
  // The grains' masses carry their weights (updateCurrentMassUsingMassWeights,
  // RigidBody2DState.cpp:1182-1200); the hybrid's broad phase ignores
  // cell_scale and uses 1 (RigidBody2DSim.cpp:41-42).
  RigidBody2DState weighted = discrete_;
  for (size_t b = 0; b < weighted.nbodies(); ++b)
    for (int k = 0; k < 3; ++k) weighted.M[3 * b + k] *= dem_weights[b];
  std::vector<double> px, pn;
  const hg_dem_desc dd = dem_desc(weighted, discrete.penalty, 1.0, precision, px, pn);
  const hg_mpm_desc md = continuum_.desc(precision);
  dev_.check(hg_hybrid_create(dev_.ctx(), &md, &dd, dem_weights.data(), &h_), "hg_hybrid_create");
  // updateHybridStates_poorman runs after every step once coupling_force is
  // present; its homogenisation depends on these flags
  const PoormanEnrichmentSettings &p = settings_.poorman_settings;
  const bool hs = p.enabled && p.homogenize_stress && p.grid_smoothing_homogenized_stress;
  const bool hv = p.enabled && p.homogenize_velocity && p.grid_smoothing_homogenized_velocity;
  if (hs || hv)
    dev_.check(hg_hybrid_set_homogenize(h_, hs ? 1 : 0, hv ? 1 : 0), "hg_hybrid_set_homogenize");
}

std::vector<double> HybridGrains2DSim::homogenizedStress() {
  std::vector<double> s(4 * continuum_.physics_grid.numGridPoints());
  dev_.check(hg_hybrid_download_homogenized(h_, s.data(), nullptr), "hg_hybrid_download_homogenized");
  return s;
}

std::vector<double> HybridGrains2DSim::homogenizedVelocity() {
  std::vector<double> v(2 * continuum_.physics_grid.numGridPoints());
  dev_.check(hg_hybrid_download_homogenized(h_, nullptr, v.data()), "hg_hybrid_download_homogenized");
  return v;
}

void HybridGrains2DSim::clearHomogenized() {
  dev_.check(hg_hybrid_clear_homogenized(h_), "hg_hybrid_clear_homogenized");
}

std::vector<std::uint8_t> HybridGrains2DSim::snapshot() {
  return snapshot_of(dev_, h_, hg_hybrid_snapshot, "hg_hybrid_snapshot");
}

void HybridGrains2DSim::restore(const std::vector<std::uint8_t> &blob) {
  dev_.check(hg_hybrid_restore(h_, blob.data(), blob.size()), "hg_hybrid_restore");
  host_stale_ = true;
}

HybridGrains2DSim::~HybridGrains2DSim() {
  if (h_) hg_hybrid_destroy(h_);
}

void HybridGrains2DSim::stepSystem() { stepSystem(1); }

void HybridGrains2DSim::stepSystem(unsigned nsteps) {
  dev_.check(hg_hybrid_step(h_, settings_.overall_dt.value(), nsteps),
             "HybridGrains2DSim::stepSystem");
  host_stale_ = true;
}

const RigidBody2DState &HybridGrains2DSim::discreteState() {
  if (host_stale_) {
    hg_mpm *m = nullptr;
    hg_dem *d = nullptr;
    dev_.check(hg_hybrid_mpm(h_, &m), "hg_hybrid_mpm");
    dev_.check(hg_hybrid_dem(h_, &d), "hg_hybrid_dem");
    dev_.check(hg_hybrid_sync(h_), "HybridGrains2DSim (state check)");
    dev_.check(hg_dem_download_bodies(d, 0, discrete_.q.data(), discrete_.v.data()), "download");
    MaterialPoints &mp = continuum_.material_points;
    dev_.check(hg_mpm_download_points(m, mp.q.data(), mp.v.data(), mp.be_bar.data(), mp.J.data()),
               "download");
    host_stale_ = false;
  }
  return discrete_;
}

const SimulationState &HybridGrains2DSim::continuumState() {
  discreteState();
  return continuum_;
}

std::vector<std::uint8_t> HybridGrains2DSim::coupledGrains() {
  std::vector<std::uint8_t> c(discrete_.nbodies());
  dev_.check(hg_hybrid_download_coupling(h_, c.data()), "hg_hybrid_download_coupling");
  return c;
}

}  // namespace hg
