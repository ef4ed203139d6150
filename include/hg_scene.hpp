// hg_scene.hpp -- host-side scene state and the three XML scene readers, with
// the reference's type and field names so callers of libHybridGrains keep
// their code.  Plain std::vector storage in the reference's layouts (2-vectors
// interleaved, 2x2 matrices column-major, DEM (x, y, theta) per body); no
// Eigen.
//
// Readers:
//   readMPMGrains2DXML  -- MPMGrains2DParser::readXMLFile
//                          (mpmgrains2dutils/MPMGrains2DParser.cpp:658-683)
//                          followed by InitialSimulationState::
//                          generateSimulationState (InitialSimulationState.cpp:147-228)
//   readRigidBody2DXML  -- RigidBody2DSceneParser::parseXMLSceneFile, circle
//                          scenes with a penalty_impact_map
//                          (rigidbody2dutils/RigidBody2DSceneParser.cpp:1127-1297)
//   readHybridGrains2DXML -- HybridGrains2DSceneParser::parseXMLSceneFile
//                          (hybridgrains2dnewutils/HybridGrains2DSceneParser.cpp:437-522)
// Each returns false and fills *error where the reference prints to stderr and
// returns false.
#pragma once

#include <cstdint>
#include <limits>
#include <memory>
#include <string>
#include <vector>

#include "hg.h"

namespace hg {

// scisim Rational<intmax_t> as the scene files write dt ("1 / 1000080",
// "0.00001", "3"); value() is double(numerator) / double(denominator) as
// Rational::operator double (scisim/Math/Rational.h:71-73).
struct Rational {
  std::int64_t numerator = 0, denominator = 1;
  double value() const { return double(numerator) / double(denominator); }
  bool positive() const { return numerator > 0 && denominator > 0; }
};
// RationalTools::extractFromString (scisim/Math/Rational.h:208-291)
bool parseRational(const std::string &text, Rational &out);

// ---- mpmgrains2d ------------------------------------------------------------

// MPMStaticPlane (mpmgrains2d/StaticPlane.h:6-22); n is normalised.
struct MPMStaticPlane {
  enum BoundaryBehavior { SLIDING = HG_PLANE_SLIDING, STICKING = HG_PLANE_STICKING };
  double x[2] = {0, 0};
  double n[2] = {0, 1};
  int boundary_behavior = SLIDING;
  double lower_bound = -std::numeric_limits<double>::infinity();
  double distanceToPoint(const double p[2]) const {
    return n[0] * (p[0] - x[0]) + n[1] * (p[1] - x[1]);
  }
};

// RectangleRegion (mpmgrains2d/InitialSimulationState.h)
struct RectangleRegion {
  double min[2], max[2];
  double rho;
  unsigned npoints_per_dim;
  double v[2];
  double omega;
};

// MaterialPoints (mpmgrains2d/MaterialPoints.h:38-156): the fields the step
// reads and writes.
struct MaterialPoints {
  std::vector<double> q, v;      // 2N
  std::vector<double> m, volume, hl, J;  // N
  std::vector<double> be_bar;    // 4N, column-major 2x2
  std::vector<std::uint8_t> kinematically_scripted;  // N
  std::size_t npoints() const { return m.size(); }
};

// PhysicsGrid geometry as PhysicsGrid::setDimensions leaves it
// (mpmgrains2d/PhysicsGrid.cpp:14-55).
struct PhysicsGrid {
  double min[2] = {0, 0}, max[2] = {0, 0};
  double cell_width = 0;
  std::uint32_t cell_count[2] = {0, 0};
  void setDimensions(const double grid_min[2], const double grid_max[2], double cell_w);
  std::size_t numGridPoints() const {
    return std::size_t(cell_count[0] + 1) * std::size_t(cell_count[1] + 1);
  }
};

// SimulationState (mpmgrains2d/SimulationState.h:16-68) plus the integrator
// settings the scene file carries.
struct SimulationState {
  MaterialPoints material_points;
  PhysicsGrid physics_grid;
  std::vector<MPMStaticPlane> static_planes;
  double shear_modulus = 0, bulk_modulus = 0, Drucker_Prager_alpha = 0;
  double near_earth_gravity[2] = {0, 0};
  double alpha = 0;  // PIC/FLIP blend
  int basis = HG_BASIS_UGIMP_LINEAR;
  // the first rectangle's point spacing (cell width / samples per side) and
  // density (InitialSimulationState.cpp:203-224), for the hybrid resampling
  double initial_particle_size = 0, material_density = 0;
  // scene-file settings (InitialSimulationState: dt, dt_string, end_time)
  Rational dt;
  std::string dt_string;
  double end_time = std::numeric_limits<double>::infinity();

  // Descriptor over this state's arrays (valid while the state is unchanged).
  hg_mpm_desc desc(int precision) const;
  // Storage behind desc() for the plane tables.
  mutable std::vector<double> plane_x_, plane_n_, plane_lower_;
  mutable std::vector<std::int32_t> plane_behavior_;

  // The GPU copy of this state, managed by the stepping API (hg_sim.hpp); a
  // copied state starts without one, a moved state takes it along.
  struct BindingSlot {
    std::shared_ptr<void> p;
    BindingSlot() = default;
    BindingSlot(const BindingSlot &) {}
    BindingSlot &operator=(const BindingSlot &) {
      p.reset();
      return *this;
    }
    BindingSlot(BindingSlot &&) = default;
    BindingSlot &operator=(BindingSlot &&) = default;
  };
  mutable BindingSlot gpu_binding;
};

// Fills the points of each rectangle as InitialSimulationState.cpp:45-145
// (rasterizeRectangle) does; false if a rectangle leaves the grid.
bool rasterizeRectangles(const std::vector<RectangleRegion> &regions, SimulationState &state,
                         std::string *error);
bool readMPMGrains2DXML(const std::string &file_name, SimulationState &state,
                        std::string *error);
bool readMPMGrains2DXMLText(const std::string &xml, SimulationState &state, std::string *error);

// ---- rigidbody2d ------------------------------------------------------------

// RigidBody2DStaticPlane (rigidbody2d/RigidBody2DStaticPlane.h:12-45)
struct RigidBody2DStaticPlane {
  double x[2] = {0, 0};
  double n[2] = {0, 1};
};

// RigidBody2DState (rigidbody2d/RigidBody2DState.h:13-190) for circle
// geometry: q, v (3B), the diagonal of M (3B), per-body radius (through the
// geometry index), fixed flags, summed near-earth gravity and static planes.
struct RigidBody2DState {
  std::vector<double> q, v, M;     // 3B
  std::vector<double> radius;      // B
  std::vector<std::uint8_t> fixed; // B
  std::vector<std::uint32_t> geometry_indices;  // B
  std::vector<double> geometry_radius;          // circle geometry table
  double gravity[2] = {0, 0};
  std::vector<RigidBody2DStaticPlane> planes;
  std::size_t nbodies() const { return radius.size(); }
};

// penalty_impact_map settings (PenaltyImpactFrictionMap.h:362-367)
struct PenaltyImpactFrictionMapSettings {
  double kn = 0, kt = 0, gamman = 0, gammat = 0, mu = 0;
};

struct RigidBody2DScene {
  RigidBody2DState state;
  PenaltyImpactFrictionMapSettings penalty;
  Rational dt;
  std::string dt_string;
  double end_time = std::numeric_limits<double>::infinity();
  double spatial_grid_scale = 1.0;  // collision_detection cell_scale
  bool reduce_bandwidth = false;    // cuthill_mckee: constraint order only
  // Broad-phase cell of SpatialGrid (SpatialGrid.cpp:70-79): the mean AABB
  // extent times the scale, at least the largest diameter (the engine's
  // 3x3-cell search needs that; the pair set does not depend on it).
  double cellWidth(double scale) const;
  hg_dem_desc desc(int precision, double scale) const;
  mutable std::vector<double> plane_x_, plane_n_;
};

bool readRigidBody2DXML(const std::string &file_name, RigidBody2DScene &scene,
                        std::string *error);
bool readRigidBody2DXMLText(const std::string &xml, RigidBody2DScene &scene, std::string *error);

// ---- hybridgrains2dnew --------------------------------------------------------

// PoormanEnrichmentSettings (hybridgrains2dnew/HybridIntegratorState.h:8-60),
// with the parser's defaults (HybridGrains2DSceneParser.cpp:150-255).
struct PoormanEnrichmentSettings {
  bool enabled = false;
  double rzone_half_thickness = 0.0, dem_r_mean = -1.0, dem_r_std = 0.0;
  double phi_window_size = 0.0, level_set_cell_width = 0.0;
  int phi_samples_per_cell_side = 0;
  double target_dense_packing_fraction = 0.0;
  int avoidAVoidFreq = 20;
  double phi_threshold = 0.5, rzone_level_set = 0.0, rho_dem = 0.96;
  bool allow_direct_transitions_between_discrete_and_continuum = false;
  bool newly_converted_continuum_stress_free = false;
  bool homogenize_stress = false, grid_smoothing_homogenized_stress = false;
  bool homogenize_velocity = false, grid_smoothing_homogenized_velocity = false;
};

struct HybridIntegratorSettings {
  std::string discrete_file_name, continuum_file_name;  // resolved against the scene's folder
  Rational overall_dt;
  std::string time_step_string;
  double end_time = std::numeric_limits<double>::infinity();
  std::string integrator_style;  // only prediction_correction_advection_node_node steps
  std::string kinematic_script = "none";
  PoormanEnrichmentSettings poorman_settings;
};

bool readHybridGrains2DXML(const std::string &file_name, HybridIntegratorSettings &settings,
                           std::string *error);

}  // namespace hg
