// hg_sim.hpp -- C++ stepping API with the reference's names, over the C-ABI
// in hg.h.  A caller of libHybridGrains swaps its includes for this header
// (and hg_scene.hpp) and keeps its calls:
//
//   ExplicitIntegrator::flow(dt, state)            mpmgrains2d/ExplicitIntegrator.h:8-13
//   ExplicitIntegrator::flow_{zero,first,second}_phase
//   RigidBody2DSim::flow / flowWithWeight          rigidbody2d/RigidBody2DSim.h:116-127
//   HybridGrains2DSim::stepSystem                  hybridgrains2dnew/HybridGrains2DSim.h:38
//
// State lives in GPU memory between calls.  After each call the host copy
// (SimulationState::material_points, RigidBody2DState q/v) is refreshed, as
// the reference's in-place update leaves it, unless the state's residency is
// set to Residency::Resident; then syncToHost() refreshes it on demand.  A
// host-side edit of a state that is already on the GPU must be followed by
// markHostModified(), which re-uploads it on the next call (contact history
// and all).
//
// Errors: where the reference would stop (assert on an inconsistent state,
// std::exit on an unsupported configuration), these calls print the engine's
// message to stderr and std::exit(EXIT_FAILURE).
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "hg.h"
#include "hg_resample.hpp"
#include "hg_scene.hpp"

namespace hg {

enum class Residency { SyncEveryCall, Resident };

// One GPU context (hg_ctx); the default binds device 0.
class Device {
 public:
  explicit Device(int device = 0);
  ~Device();
  Device(const Device &) = delete;
  Device &operator=(const Device &) = delete;
  hg_ctx *ctx() const { return ctx_; }
  static Device &defaultDevice();
  // Exits with the context's last error when st is not HG_OK.
  void check(hg_status st, const char *where) const;

 private:
  hg_ctx *ctx_ = nullptr;
};

// ---- MPM ----------------------------------------------------------------------

// The device copy of a SimulationState (created on the first call).
class MpmDevice {
 public:
  MpmDevice(Device &dev, const SimulationState &state, int precision);
  ~MpmDevice();
  hg_mpm *handle() const { return h_; }
  void download(SimulationState &state) const;

 private:
  Device &dev_;
  hg_mpm *h_ = nullptr;
};

// A state's GPU binding, kept in SimulationState::gpu_binding.
struct MpmBinding {
  int precision = HG_FP64;
  Residency residency = Residency::SyncEveryCall;
  Device *device = nullptr;  // nullptr: Device::defaultDevice()
  std::unique_ptr<MpmDevice> gpu;
  bool host_modified = true;
  // The binding of a state (created with defaults on first use).
  static MpmBinding &of(const SimulationState &state);
};

void markHostModified(SimulationState &state);
void setResidency(SimulationState &state, Residency r, int precision = HG_FP64);
void syncToHost(SimulationState &state);

// Resumable snapshots of the device state (hg_mpm_snapshot / hg_mpm_restore;
// the reference serialises its host state instead, mpmgrains2d_cli.cpp:154-247).
// restoreDeviceState loads a snapshot taken from a state read from the same
// scene; the next flow continues exactly where the snapshot was taken.
std::vector<std::uint8_t> snapshotDeviceState(SimulationState &state);
void restoreDeviceState(SimulationState &state, const std::vector<std::uint8_t> &blob);

namespace ExplicitIntegrator {
void flow(const double &dt, SimulationState &state);
void flow_zero_phase(SimulationState &state);
void flow_first_phase(const double &dt, SimulationState &state);
void flow_second_phase(const double &dt, SimulationState &state);
// nsteps of flow with one host refresh at the end (no reference counterpart).
void flow_n(const double &dt, unsigned nsteps, SimulationState &state);
}  // namespace ExplicitIntegrator

namespace MPMGrains2DSim {
void flow(const double &dt, SimulationState &state);
void flow_zero_phase(SimulationState &state);
void flow_first_phase(const double &dt, SimulationState &state);
void flow_second_phase(const double &dt, SimulationState &state);
}  // namespace MPMGrains2DSim

// ---- DEM ----------------------------------------------------------------------

// PenaltyImpactFrictionMap (rigidbody2d/PenaltyImpactFrictionMap.h:362-436):
// the contact law's constants.  Its contact history lives with the
// simulation's device state.
class PenaltyImpactFrictionMap {
 public:
  PenaltyImpactFrictionMap(double kn, double kt, double gamman, double gammat, double mu)
      : s_{kn, kt, gamman, gammat, mu} {}
  explicit PenaltyImpactFrictionMap(const PenaltyImpactFrictionMapSettings &s) : s_(s) {}
  const PenaltyImpactFrictionMapSettings &settings() const { return s_; }
  std::string name() const { return "penalty_impact_friction_map"; }

 private:
  PenaltyImpactFrictionMapSettings s_;
};

// RigidBody2DSim (rigidbody2d/RigidBody2DSim.h:16-190) for circle scenes.
class RigidBody2DSim {
 public:
  RigidBody2DSim(const RigidBody2DState &state, double spatial_grid_scale = 1.0,
                 int precision = HG_FP64, Device *device = nullptr);
  ~RigidBody2DSim();
  RigidBody2DState &getState() { return state_; }
  const RigidBody2DState &getState() const { return state_; }
  // Mass weights (RigidBody2DState::massWeights) for flowWithWeight.
  std::vector<double> &massWeights() { return weights_; }
  void setResidency(Residency r) { residency_ = r; }
  void markHostModified() { host_modified_ = true; }
  void syncToHost();

  // RigidBody2DSim::flow with the penalty map (RigidBody2DSim.cpp:1013-1039).
  // The scripting callback, unconstrained map, CoR and friction solver of
  // the reference signature are ignored by the penalty map
  // (PenaltyImpactFrictionMap.cpp:1084-1089) and are not taken here.
  void flow(unsigned iteration, const Rational &dt, const PenaltyImpactFrictionMap &ifmap,
            bool reduce_bandwidth = false);
  // RigidBody2DSim::flowWithWeight (RigidBody2DSim.cpp:1041-1069)
  void flowWithWeight(unsigned iteration, const Rational &dt,
                      const PenaltyImpactFrictionMap &ifmap, bool reduce_bandwidth = false);
  // Bodies and contact history (hg_dem_snapshot / hg_dem_restore).
  std::vector<std::uint8_t> snapshot(const PenaltyImpactFrictionMap &ifmap);
  void restore(const std::vector<std::uint8_t> &blob, const PenaltyImpactFrictionMap &ifmap);
  hg_dem *handle() const { return h_; }

 private:
  void ensure(const PenaltyImpactFrictionMap &ifmap);
  void step(double dt, bool weighted);
  Device &dev_;
  RigidBody2DState state_;
  std::vector<double> weights_;
  std::vector<double> uploaded_weights_;
  bool weighted_on_device_ = false;
  double scale_;
  int precision_;
  Residency residency_ = Residency::SyncEveryCall;
  bool host_modified_ = true;
  hg_dem *h_ = nullptr;
};

// ---- hybrid -------------------------------------------------------------------

// HybridGrains2DSim (hybridgrains2dnew/HybridGrains2DSim.h:15-76): one
// continuum and one grain scene.  Without a coupling_force element the
// grains' mass weights come from the caller.  With one, the zone resampling
// runs as the reference runs it (hg_resample.hpp): the continuum's points
// are dropped and re-seeded, the constructor performs initializeAvoidAVoid
// (two resampling passes), and every avoidAVoidFreq + 1 steps the state is
// downloaded, resampled on the host and stepped on from new device handles.
class HybridGrains2DSim {
 public:
  HybridGrains2DSim(const HybridIntegratorSettings &settings, const RigidBody2DScene &discrete,
                    const SimulationState &continuum, const std::vector<double> &dem_weights,
                    int precision = HG_FP64, Device *device = nullptr);
  ~HybridGrains2DSim();
  // HybridGrains2DSim::stepSystem (HybridGrains2DSim.cpp:229-260)
  void stepSystem();
  void stepSystem(unsigned nsteps);
  const RigidBody2DState &discreteState();
  const SimulationState &continuumState();
  std::vector<std::uint8_t> coupledGrains();
  // ResamplingTools' homogenised fields (ResamplingTools.cpp:744-884), which
  // every step adds to when the scene's coupling_force sets homogenize_* and
  // grid_smoothing_homogenized_* (HybridGrains2DSim.cpp:545-556): 4 entries
  // per node (column-major 2x2) and 2 per node, nodes in PhysicsGrid's flat
  // order.
  std::vector<double> homogenizedStress();
  std::vector<double> homogenizedVelocity();
  // ResamplingTools::clearHomogenizedStress / clearHomogenizedVelocity
  void clearHomogenized();
  // PhysicsGrid::homog_factor as the last step saved it
  // (HybridGrains2DSim.cpp:169-197): 1 at nodes with continuum mass, 0.5 at
  // nodes with grain mass, else 0; 1 everywhere before the first step.
  std::vector<double> homogFactor();
  // Continuum, grains with their contact history, and the homogenisation
  // accumulators (hg_hybrid_snapshot / hg_hybrid_restore).
  // With resampling on, the blob also carries the resampler's host state
  // (counter, random streams, zones, reference masses and weights).
  std::vector<std::uint8_t> snapshot();
  void restore(const std::vector<std::uint8_t> &blob);
  hg_hybrid *handle() const { return h_; }
  bool resamplingEnabled() const { return rs_ != nullptr; }
  // counts of the last resampling pass, and the number of passes so far
  ResampleReport lastResample() const;
  long resampleCount() const;

 private:
  struct Resampler;
  void rebuildHandle();
  void pullState();
  void pushHomogenized();
  void poormanCall(bool after_step);
  void initializeAvoidAVoid();

  Device &dev_;
  HybridIntegratorSettings settings_;
  RigidBody2DState discrete_;
  SimulationState continuum_;
  PenaltyImpactFrictionMapSettings penalty_;
  int precision_ = HG_FP64;
  bool host_stale_ = false;
  hg_hybrid *h_ = nullptr;
  std::unique_ptr<Resampler> rs_;
};

}  // namespace hg
