// hg_sim.cpp -- the reference-named stepping API over the C-ABI.
#include "hg_sim.hpp"

#include <cstdio>
#include <cstdlib>
#include <cstring>

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

// The resampler's host state (HybridGrains2DSim.h:50-75's ZoneTools,
// ResamplingTools and counters) and the fields of the state it edits.
struct HybridGrains2DSim::Resampler {
  ZoneTools zones;
  ResamplingTools tools;
  HybridResampleState s;
  PoormanEnrichmentSettings p;
  int counter = 0;            // m_avoid_a_void_step_counter
  bool initial_done = false;  // m_initial_resampling_performed
  bool hs = false, hv = false, ss = false, sv = false;
  long passes = 0;
  ResampleReport last;
};

HybridGrains2DSim::HybridGrains2DSim(const HybridIntegratorSettings &settings,
                                     const RigidBody2DScene &discrete,
                                     const SimulationState &continuum,
                                     const std::vector<double> &dem_weights, int precision,
                                     Device *device)
    : dev_(device ? *device : Device::defaultDevice()),
      settings_(settings),
      discrete_(discrete.state),
      continuum_(continuum),
      penalty_(discrete.penalty),
      precision_(precision) {
  if (settings_.integrator_style != "prediction_correction_advection_node_node") {
    std::fprintf(stderr, "unsupported integrator style\n");
    std::exit(EXIT_FAILURE);
  }
  if (dem_weights.size() != discrete_.nbodies()) {
    std::fprintf(stderr, "hg: one mass weight per grain is needed\n");
    std::exit(EXIT_FAILURE);
  }
  const PoormanEnrichmentSettings &ps = settings_.poorman_settings;
  if (ps.enabled) {
    // HybridGrains2DSim.cpp:21-73: the continuum starts without points (the
    // resampling seeds it); zone and resampling tools set up from the scene
    rs_.reset(new Resampler);
    Resampler &r = *rs_;
    r.p = ps;
    r.hs = ps.homogenize_stress;
    r.hv = ps.homogenize_velocity;
    r.ss = ps.homogenize_stress && ps.grid_smoothing_homogenized_stress;
    r.sv = ps.homogenize_velocity && ps.grid_smoothing_homogenized_velocity;
    HybridResampleState &s = r.s;
    s.continuum = continuum;
    MaterialPoints &mp = s.continuum.material_points;
    mp = MaterialPoints{};
    s.initial_particle_size = continuum.initial_particle_size;
    s.material_density = continuum.material_density;
    s.discrete = discrete.state;
    s.dem_q0 = s.discrete.q;
    s.dem_M0 = s.discrete.M;
    s.dem_weight = dem_weights;
    for (std::size_t b = 0; b < s.discrete.nbodies(); ++b)
      for (int k = 0; k < 3; ++k) s.discrete.M[3 * b + std::size_t(k)] *= dem_weights[b];
    const std::size_t nn = continuum.physics_grid.numGridPoints();
    s.homog_stress.assign(4 * nn, 0.0);
    s.homog_velocity.assign(2 * nn, 0.0);
    try {
      r.zones.init(continuum.physics_grid, ps.level_set_cell_width,
                   ps.allow_direct_transitions_between_discrete_and_continuum);
    } catch (const std::exception &e) {
      std::fprintf(stderr, "hg: %s\n", e.what());
      std::exit(EXIT_FAILURE);
    }
    r.tools.init(ps.dem_r_mean, ps.dem_r_std);
    initializeAvoidAVoid();
    return;
  }
  if (continuum_.material_points.npoints() == 0) {
    std::fprintf(stderr, "hg: the hybrid step needs continuum points (the grains use their hl)\n");
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
  // the reference saves the coupling weights every step
  dev_.check(hg_hybrid_set_homog_factor(h_, 1), "hg_hybrid_set_homog_factor");
}

// New device handles from the resampled state: points and grains with their
// weighted masses, the relabelled contact history, the homogenisation flags.
void HybridGrains2DSim::rebuildHandle() {
  Resampler &r = *rs_;
  HybridResampleState &s = r.s;
  if (s.continuum.material_points.npoints() == 0) {
    std::fprintf(stderr, "hg: the zone resampling left no continuum points\n");
    std::exit(EXIT_FAILURE);
  }
  if (h_) {
    hg_hybrid_destroy(h_);
    h_ = nullptr;
  }
  continuum_ = s.continuum;
  discrete_ = s.discrete;
  std::vector<double> px, pn;
  const hg_dem_desc dd = dem_desc(discrete_, penalty_, 1.0, precision_, px, pn);
  const hg_mpm_desc md = continuum_.desc(precision_);
  dev_.check(hg_hybrid_create(dev_.ctx(), &md, &dd, s.dem_weight.data(), &h_), "hg_hybrid_create");
  hg_dem *d = nullptr;
  dev_.check(hg_hybrid_dem(h_, &d), "hg_hybrid_dem");
  dev_.check(hg_dem_upload_contacts(d, 0, std::uint32_t(s.c_type.size()), s.c_type.data(),
                                    s.c_i.data(), s.c_j.data(), s.c_ds.data()),
             "hg_dem_upload_contacts");
  if (r.ss || r.sv)
    dev_.check(hg_hybrid_set_homogenize(h_, r.ss ? 1 : 0, r.sv ? 1 : 0), "hg_hybrid_set_homogenize");
  dev_.check(hg_hybrid_set_homog_factor(h_, 1), "hg_hybrid_set_homog_factor");
  host_stale_ = false;
}

// Device state into the resampler's copy: point and grain states, the
// contact history with the last step's forces and contact points, the
// homogenised accumulators.
void HybridGrains2DSim::pullState() {
  Resampler &r = *rs_;
  HybridResampleState &s = r.s;
  hg_mpm *m = nullptr;
  hg_dem *d = nullptr;
  dev_.check(hg_hybrid_mpm(h_, &m), "hg_hybrid_mpm");
  dev_.check(hg_hybrid_dem(h_, &d), "hg_hybrid_dem");
  dev_.check(hg_hybrid_sync(h_), "HybridGrains2DSim (state check)");
  MaterialPoints &mp = s.continuum.material_points;
  dev_.check(hg_mpm_download_points(m, mp.q.data(), mp.v.data(), mp.be_bar.data(), mp.J.data()),
             "download");
  dev_.check(hg_dem_download_bodies(d, 0, s.discrete.q.data(), s.discrete.v.data()), "download");
  std::uint32_t n = 0;
  dev_.check(hg_dem_download_contacts(d, 0, 0, nullptr, nullptr, nullptr, nullptr, &n), "contacts");
  s.c_type.resize(n);
  s.c_i.resize(n);
  s.c_j.resize(n);
  s.c_ds.resize(2 * std::size_t(n));
  s.c_f.resize(2 * std::size_t(n));
  s.c_p.resize(2 * std::size_t(n));
  dev_.check(hg_dem_download_contacts(d, 0, n, s.c_type.data(), s.c_i.data(), s.c_j.data(),
                                      s.c_ds.data(), &n),
             "contacts");
  dev_.check(hg_dem_download_contact_forces(d, 0, n, s.c_f.data(), s.c_p.data(), &n), "contacts");
  if (r.ss || r.sv)
    dev_.check(hg_hybrid_download_homogenized(h_, s.homog_stress.data(), s.homog_velocity.data()),
               "hg_hybrid_download_homogenized");
  continuum_.material_points.q = mp.q;
  continuum_.material_points.v = mp.v;
  continuum_.material_points.be_bar = mp.be_bar;
  continuum_.material_points.J = mp.J;
  discrete_.q = s.discrete.q;
  discrete_.v = s.discrete.v;
  host_stale_ = false;
}

void HybridGrains2DSim::pushHomogenized() {
  Resampler &r = *rs_;
  if (r.ss || r.sv)
    dev_.check(hg_hybrid_upload_homogenized(h_, r.s.homog_stress.data(), r.s.homog_velocity.data()),
               "hg_hybrid_upload_homogenized");
}

// HybridGrains2DSim::updateHybridStates_poorman (HybridGrains2DSim.cpp:301-559).
// after_step: the call that ends a device step, whose own accumulation is the
// one the reference makes at that point.
void HybridGrains2DSim::poormanCall(bool after_step) {
  Resampler &r = *rs_;
  HybridResampleState &s = r.s;
  if (r.counter == 0) {
    if (h_) pullState();
    // the resampling block's own homogenize calls (:405-427), which a
    // device step with grid smoothing has made already
    if (r.hs && !(after_step && r.ss)) accumulateHomogenizedStress(s);
    if (r.hv && !(after_step && r.sv)) accumulateHomogenizedVelocity(s);
    try {
      r.last = r.tools.resample(r.zones, s, r.p, double(r.p.avoidAVoidFreq));
    } catch (const std::exception &e) {
      std::fprintf(stderr, "hg: %s\n", e.what());
      std::exit(EXIT_FAILURE);
    }
    ++r.passes;
    rebuildHandle();
    // the call's closing accumulation (:545-556), on the resampled state
    if (r.ss) accumulateHomogenizedStress(s);
    if (r.sv) accumulateHomogenizedVelocity(s);
    pushHomogenized();
  } else {
    if (r.counter >= r.p.avoidAVoidFreq) r.counter = -1;
    if (!after_step && (r.ss || r.sv)) {
      pullState();
      if (r.ss) accumulateHomogenizedStress(s);
      if (r.sv) accumulateHomogenizedVelocity(s);
      pushHomogenized();
    }
  }
  ++r.counter;
}

// HybridGrains2DSim::initializeAvoidAVoid (HybridGrains2DSim.cpp:75-88)
void HybridGrains2DSim::initializeAvoidAVoid() {
  Resampler &r = *rs_;
  std::fill(r.s.homog_stress.begin(), r.s.homog_stress.end(), 0.0);
  std::fill(r.s.homog_velocity.begin(), r.s.homog_velocity.end(), 0.0);
  poormanCall(false);
  r.counter = 0;
  poormanCall(false);
}

ResampleReport HybridGrains2DSim::lastResample() const { return rs_ ? rs_->last : ResampleReport{}; }
long HybridGrains2DSim::resampleCount() const { return rs_ ? rs_->passes : 0; }

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

std::vector<double> HybridGrains2DSim::homogFactor() {
  std::vector<double> v(continuum_.physics_grid.numGridPoints());
  dev_.check(hg_hybrid_download_homog_factor(h_, v.data()), "hg_hybrid_download_homog_factor");
  return v;
}

void HybridGrains2DSim::clearHomogenized() {
  dev_.check(hg_hybrid_clear_homogenized(h_), "hg_hybrid_clear_homogenized");
}

namespace {
// A flat little-endian writer and reader for the resampler's host state.
struct BlobOut {
  std::vector<std::uint8_t> b;
  void raw(const void *p, std::size_t n) {
    const std::uint8_t *c = static_cast<const std::uint8_t *>(p);
    b.insert(b.end(), c, c + n);
  }
  template <typename T>
  void put(T v) { raw(&v, sizeof(T)); }
  template <typename T>
  void vec(const std::vector<T> &v) {
    put<std::uint64_t>(v.size());
    if (!v.empty()) raw(v.data(), v.size() * sizeof(T));
  }
};
struct BlobIn {
  const std::uint8_t *p, *e;
  bool ok = true;
  void raw(void *d, std::size_t n) {
    if (std::size_t(e - p) < n) {
      ok = false;
      return;
    }
    std::memcpy(d, p, n);
    p += n;
  }
  template <typename T>
  T get() {
    T v{};
    raw(&v, sizeof(T));
    return v;
  }
  template <typename T>
  void vec(std::vector<T> &v) {
    const std::uint64_t n = get<std::uint64_t>();
    if (!ok || n > std::uint64_t(e - p) / sizeof(T)) {
      ok = false;
      return;
    }
    v.resize(std::size_t(n));
    if (n) raw(v.data(), std::size_t(n) * sizeof(T));
  }
};
const char kResampleMagic[8] = {'H', 'G', 'H', 'Y', 'R', 'S', '0', '1'};
}  // namespace

std::vector<std::uint8_t> HybridGrains2DSim::snapshot() {
  std::vector<std::uint8_t> dev = snapshot_of(dev_, h_, hg_hybrid_snapshot, "hg_hybrid_snapshot");
  if (!rs_) return dev;
  const Resampler &r = *rs_;
  const HybridResampleState &s = r.s;
  BlobOut o;
  o.raw(kResampleMagic, 8);
  o.put<std::int32_t>(r.counter);
  o.put<std::uint8_t>(r.initial_done ? 1 : 0);
  o.put<std::int64_t>(r.passes);
  const std::string rnd = r.tools.saveRandomState();
  o.vec(std::vector<char>(rnd.begin(), rnd.end()));
  o.vec(r.zones.mpmIndicators());
  o.vec(r.zones.previousMpmIndicators());
  const MaterialPoints &mp = continuum_.material_points;
  o.vec(mp.q);
  o.vec(mp.v);
  o.vec(mp.m);
  o.vec(mp.volume);
  o.vec(mp.hl);
  o.vec(mp.J);
  o.vec(mp.be_bar);
  o.vec(mp.kinematically_scripted);
  o.vec(s.mpm_q0);
  o.vec(s.mpm_m0);
  o.vec(s.mpm_volume0);
  o.vec(s.mpm_weight);
  o.vec(discrete_.q);
  o.vec(discrete_.v);
  o.vec(discrete_.M);
  o.vec(discrete_.radius);
  o.vec(discrete_.fixed);
  o.vec(discrete_.geometry_indices);
  o.vec(discrete_.geometry_radius);
  o.vec(s.dem_q0);
  o.vec(s.dem_M0);
  o.vec(s.dem_weight);
  o.vec(dev);
  return o.b;
}

void HybridGrains2DSim::restore(const std::vector<std::uint8_t> &blob) {
  if (!rs_) {
    dev_.check(hg_hybrid_restore(h_, blob.data(), blob.size()), "hg_hybrid_restore");
    host_stale_ = true;
    return;
  }
  if (blob.size() < 8 || std::memcmp(blob.data(), kResampleMagic, 8) != 0) {
    std::fprintf(stderr, "hg: not a snapshot of a resampling hybrid simulation\n");
    std::exit(EXIT_FAILURE);
  }
  Resampler &r = *rs_;
  HybridResampleState &s = r.s;
  BlobIn in{blob.data() + 8, blob.data() + blob.size()};
  const int counter = in.get<std::int32_t>();
  const bool initial_done = in.get<std::uint8_t>() != 0;
  const long passes = long(in.get<std::int64_t>());
  std::vector<char> rnd;
  in.vec(rnd);
  std::vector<int> zc, zp;
  in.vec(zc);
  in.vec(zp);
  MaterialPoints &mp = s.continuum.material_points;
  in.vec(mp.q);
  in.vec(mp.v);
  in.vec(mp.m);
  in.vec(mp.volume);
  in.vec(mp.hl);
  in.vec(mp.J);
  in.vec(mp.be_bar);
  in.vec(mp.kinematically_scripted);
  in.vec(s.mpm_q0);
  in.vec(s.mpm_m0);
  in.vec(s.mpm_volume0);
  in.vec(s.mpm_weight);
  in.vec(s.discrete.q);
  in.vec(s.discrete.v);
  in.vec(s.discrete.M);
  in.vec(s.discrete.radius);
  in.vec(s.discrete.fixed);
  in.vec(s.discrete.geometry_indices);
  in.vec(s.discrete.geometry_radius);
  in.vec(s.dem_q0);
  in.vec(s.dem_M0);
  in.vec(s.dem_weight);
  std::vector<std::uint8_t> dev;
  in.vec(dev);
  const std::size_t N = mp.m.size(), B = s.discrete.radius.size();
  const bool sizes = mp.q.size() == 2 * N && mp.v.size() == 2 * N && mp.volume.size() == N &&
                     mp.hl.size() == N && mp.J.size() == N && mp.be_bar.size() == 4 * N &&
                     mp.kinematically_scripted.size() == N && s.mpm_q0.size() == 2 * N &&
                     s.mpm_m0.size() == N && s.mpm_volume0.size() == N &&
                     s.mpm_weight.size() == N && s.discrete.q.size() == 3 * B &&
                     s.discrete.v.size() == 3 * B && s.discrete.M.size() == 3 * B &&
                     s.discrete.fixed.size() == B && s.discrete.geometry_indices.size() == B &&
                     s.dem_q0.size() == 3 * B && s.dem_M0.size() == 3 * B &&
                     s.dem_weight.size() == B;
  if (!in.ok || !sizes ||
      !r.tools.loadRandomState(std::string(rnd.begin(), rnd.end()))) {
    std::fprintf(stderr, "hg: damaged resampling snapshot\n");
    std::exit(EXIT_FAILURE);
  }
  r.counter = counter;
  r.initial_done = initial_done;
  r.passes = passes;
  r.zones.setIndicators(zc, zp);
  s.c_type.clear();
  s.c_i.clear();
  s.c_j.clear();
  s.c_ds.clear();
  s.c_f.clear();
  s.c_p.clear();
  rebuildHandle();
  dev_.check(hg_hybrid_restore(h_, dev.data(), dev.size()), "hg_hybrid_restore");
  host_stale_ = true;
}

HybridGrains2DSim::~HybridGrains2DSim() {
  if (h_) hg_hybrid_destroy(h_);
}

void HybridGrains2DSim::stepSystem() { stepSystem(1); }

void HybridGrains2DSim::stepSystem(unsigned nsteps) {
  const double dt = settings_.overall_dt.value();
  if (!rs_) {
    dev_.check(hg_hybrid_step(h_, dt, nsteps), "HybridGrains2DSim::stepSystem");
    host_stale_ = true;
    return;
  }
  Resampler &r = *rs_;
  // stepSystem_..._node_node (HybridGrains2DSim.cpp:121-129): the first step
  // starts with two more resampling calls
  if (!r.initial_done) {
    poormanCall(false);
    poormanCall(false);
    r.initial_done = true;
  }
  // every step ends with a call; runs of steps whose calls do not resample
  // go to the device together
  unsigned left = nsteps;
  while (left > 0) {
    unsigned quiet = 0;  // steps before the one whose call resamples
    for (int c = r.counter; c != 0 && quiet < left; ++quiet) {
      if (c >= r.p.avoidAVoidFreq) c = -1;
      ++c;
    }
    if (quiet) {
      dev_.check(hg_hybrid_step(h_, dt, quiet), "HybridGrains2DSim::stepSystem");
      host_stale_ = true;
      for (unsigned k = 0; k < quiet; ++k) {
        if (r.counter >= r.p.avoidAVoidFreq) r.counter = -1;
        ++r.counter;
      }
      left -= quiet;
    }
    if (left > 0) {
      dev_.check(hg_hybrid_step(h_, dt, 1), "HybridGrains2DSim::stepSystem");
      host_stale_ = true;
      poormanCall(true);
      --left;
    }
  }
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
