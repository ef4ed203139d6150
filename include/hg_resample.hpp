// hg_resample.hpp -- the hybrid zone resampling ("avoid a void" poor-man
// enrichment) of hybridgrains2dnew, on host state: packing fraction, distance
// field and zones, Poisson-style resampling of grains and material points,
// their quantities, deletions and mass weights.  The reference runs it every
// avoidAVoidFreq + 1 steps outside the explicit step
// (HybridGrains2DSim::updateHybridStates_poorman, HybridGrains2DSim.cpp:
// 301-559); HybridGrains2DSim (hg_sim.hpp) runs it between device steps on
// downloaded state and rebuilds the device handles from the result.
//
// The zone and resampling classes keep the reference's names, ZoneTools
// (ZoneTools.cpp) and ResamplingTools (ResamplingTools.cpp), with DistGrid and
// redistanceGrid (distgrid.cpp, levelsetutil.cpp) and the MT19937 stream of
// random.c; the geometry and search structures behind them are this
// library's own.
#pragma once

#include <cstdint>
#include <functional>
#include <random>
#include <string>
#include <vector>

#include "hg_scene.hpp"

namespace hg {

// ---- random.c -------------------------------------------------------------------
// MT19937 as published in 1998 (Matsumoto and Nishimura): the state is seeded
// with Knuth's multiplier 69069, and randomMT() returns the tempered word times
// 2^-32 in single precision.  The reference keeps one such stream per process,
// seeded with 4357 on first use; each resampler here owns one.
class MersenneTwister1998 {
 public:
  explicit MersenneTwister1998(std::uint32_t seed = 4357u) { seedMT(seed); }
  void seedMT(std::uint32_t seed);
  std::uint32_t randomMT_i();
  float randomMT();
  // the 624 state words and the position (for checks against other
  // implementations of the same recurrence)
  const std::uint32_t *state() const { return mt_; }
  int position() const { return mti_; }
  void setState(const std::uint32_t *words, int position);

 private:
  void twist();
  std::uint32_t mt_[624];
  int mti_ = 625;
};

// ---- geometry ---------------------------------------------------------------------
// Area of [x0, x1] x [y0, y1] inside the circle of radius r at (cx, cy) (what
// the reference's CircleRectangleIntersection.cpp computes).
double boxCircleOverlapArea(double x0, double x1, double y0, double y1, double cx, double cy,
                            double r);

// ---- distgrid.cpp, levelsetutil.cpp -------------------------------------------------
struct DistGrid {
  std::uint32_t N[2] = {0, 0};
  double min_coords[2] = {0, 0}, max_coords[2] = {0, 0};
  double h = 0;
  std::vector<double> minDist;    // (N+1)^2 nodes
  std::vector<int> closestPoint;  // (N+1)^2 nodes
  void resize(const double mn[2], const std::uint32_t n[2], double cell_h);
  int num_points() const { return int((N[0] + 1) * (N[1] + 1)); }
  int num_cells() const { return int(N[0] * N[1]); }
};
// Unsigned distance of every node to the nearest sample (samples 2P,
// interleaved): exact within 3 particleSize of a sample, then propagated
// node to node from the neighbours' closest samples.
void redistanceGrid(DistGrid &grid, const std::vector<double> &samples, double particleSize);

// ---- ZoneTools.cpp ---------------------------------------------------------------
enum ZoneType { ZT_CONTINUUM = 0, ZT_DISCRETE = 1, ZT_HYBRID = 2 };

// An axis-aligned zone (min x, min y, max x, max y) and an MPM cell index.
struct Zone {
  double v[4];
};
struct CellIndex {
  std::uint32_t i, j;
};

class ZoneTools {
 public:
  // ZoneTools::init (:6-65); the continuum grid's geometry
  void init(const PhysicsGrid &grid, double level_set_cell_width,
            bool allow_direct_transitions_between_discrete_and_continuum);
  // :185-270: packing fraction of the non-fixed circles per level-set cell
  // (window boxes around phi_samples_per_cell_side^2 sample points), then
  // 1 wherever a material point's window reaches (points 2N)
  void updatePackingFraction(const std::vector<double> &circle_pos,
                             const std::vector<double> &circle_r,
                             const std::vector<double> &points, double phi_window_size,
                             int phi_samples_per_cell_side);
  // :283-444: signed distance to the phi_threshold contour (positive inside)
  void updateDistanceField(double phi_threshold);
  // :678-685 (estimateLevelSetLevelZoneIndicators + sampleMPMLevelZoneIndicators)
  void computeRZones(double rzone_level_set, double rzone_half_thickness);

  bool isParticleInDiscreteZone_MPMIndicator(const double x[2]) const;
  bool isParticleInContinuumZone_MPMIndicator(const double x[2]) const;
  bool isParticleInHybridZone_MPMIndicator(const double x[2]) const;
  // :123-147
  double demWeight(const double x[2]) const;
  double mpmWeight(const double x[2]) const;

  const std::vector<Zone> &getRZones() const { return rzones_; }
  const std::vector<CellIndex> &getRZoneIndices() const { return rzone_idx_; }
  const std::vector<Zone> &getDirectTransitionZonesFromDiscreteToContinuum() const {
    return direct_;
  }
  const std::vector<CellIndex> &getDirectTransitionZonesFromDiscreteToContinuumIndices() const {
    return direct_idx_;
  }
  // :703-758 (points 2N; bodies q 3B with fixed flags)
  void determineMPMParticlesToDelete(const std::vector<double> &points,
                                     std::vector<unsigned> &ids) const;
  void determineDEMGrainsToDelete(const std::vector<double> &q,
                                  const std::vector<std::uint8_t> &fixed,
                                  std::vector<unsigned> &ids) const;
  // :760-814
  void identifyInnerContinuumRegion(std::vector<Zone> &zones,
                                    std::vector<CellIndex> &idx) const;
  void identifyInnerContinuumRegionExcludingDirectTransitionZones(
      std::vector<Zone> &zones, std::vector<CellIndex> &idx) const;
  bool allowDirectTransitionsBetweenDiscreteAndContinuum() const { return allow_direct_; }
  // the MPM-cell indicators carried from one resampling to the next
  void setIndicators(const std::vector<int> &current, const std::vector<int> &previous);

  const std::vector<double> &packingFractions() const { return phi_; }
  const DistGrid &distGrid() const { return dist_; }
  const std::vector<int> &levelSetIndicators() const { return zone_ls_; }
  const std::vector<int> &mpmIndicators() const { return zone_mpm_; }
  const std::vector<int> &previousMpmIndicators() const { return prev_zone_mpm_; }

 private:
  int mpmCellOf(const double x[2]) const;
  void estimateLevelSetLevelZoneIndicators(double level_set, double half_thickness);
  void sampleMPMLevelZoneIndicators();

  DistGrid dist_;
  std::vector<double> phi_;
  std::vector<int> zone_ls_, zone_mpm_, prev_zone_mpm_;
  std::uint32_t mpm_cells_[2] = {0, 0};
  double mpm_min_[2] = {0, 0}, mpm_h_ = 0;
  bool allow_direct_ = false;
  std::vector<Zone> rzones_, direct_;
  std::vector<CellIndex> rzone_idx_, direct_idx_;
};

// ---- ResamplingTools.cpp and the state edits it makes ------------------------------
// What the resampling needs beyond the step's state: the material points'
// reference positions and unweighted mass and volume, the grains' reference
// configuration, unweighted mass matrix diagonal and mass weights.
struct HybridResampleState {
  SimulationState continuum;  // q, v, m, volume, hl, J, be_bar (weighted m, volume)
  std::vector<double> mpm_q0, mpm_m0, mpm_volume0, mpm_weight;
  double initial_particle_size = 0, material_density = 0;
  RigidBody2DState discrete;  // q, v, M (weighted), radius, fixed, geometry
  std::vector<double> dem_q0, dem_M0, dem_weight;
  // contact history in the engine's download order: type (0 plane, 1 pair,
  // 2 kinematic), owner, partner, tangential spring (2 per contact)
  std::vector<std::int32_t> c_type, c_i, c_j;
  std::vector<double> c_ds;
  // the last step's force on the owner and contact point of each contact
  // (2 each), which the collision cache keeps for homogenizeStress
  std::vector<double> c_f, c_p;
  // homogenised accumulators, 4 and 2 per grid node
  std::vector<double> homog_stress, homog_velocity;
};

struct ResampleReport {
  int dem_before = 0, dem_after_sampling = 0, dem_after_delete = 0;
  int mpm_before = 0, mpm_after_sampling = 0, mpm_after_delete = 0;
  std::size_t rzones = 0;
};

class ResamplingTools {
 public:
  ResamplingTools() = default;
  // ResamplingTools::init (:6-14)
  void init(double dem_r_mean, double dem_r_std);
  // One pass of updateHybridStates_poorman's resampling block
  // (HybridGrains2DSim.cpp:305-540) on s: zones, sampling, quantities,
  // deletions, mass weights; clears the homogenised accumulators.
  ResampleReport resample(ZoneTools &zones, HybridResampleState &s,
                          const PoormanEnrichmentSettings &p, double avoid_a_void_freq);
  MersenneTwister1998 &rng() { return mt_; }
  // both random streams (the MT19937 words and the radius sampler's engine
  // and distribution) as text, for a snapshot
  std::string saveRandomState() const;
  bool loadRandomState(const std::string &text);
  // randomRadiusSampler (:18-25): a normal draw inside mean +- 2 std
  double randomRadiusSampler();

 private:
  void avoidAVoidMPM(HybridResampleState &s, const std::vector<Zone> &zones, int &before,
                     int &after);
  void avoidAVoidDEM(HybridResampleState &s, const std::vector<Zone> &zones, double rho_dem,
                     int &before, int &after);
  MersenneTwister1998 mt_;
  std::default_random_engine generator_;
  std::normal_distribution<double> dist_;
  double r_mean_ = -1, r_std_ = 0;
};

// ResamplingTools::homogenizeStress (ResamplingTools.cpp:744-801): each
// circle-pair contact adds -(f r^T + r f^T)/2 / h^2, with r between the
// bodies' current positions, to the nodes around its contact point (basis
// half width h/4).
void accumulateHomogenizedStress(HybridResampleState &s);
// ResamplingTools::homogenizeVelocity (:830-884): the non-fixed grains'
// mass-weighted velocity at every node their mass reaches.
void accumulateHomogenizedVelocity(HybridResampleState &s);

// ConstitutiveModel::computeStrainFromCauchyStress (ConstitutiveModel.cpp:99-117):
// J and b_bar (column-major) of a stress-free-reference state with Cauchy
// stress sigma (column-major).
void computeStrainFromCauchyStress(const double sigma[4], double kappa, double mu, double &J,
                                   double be_bar[4]);

// The basis weight of node (ix, iy) for a point at x with half width hl
// (uGIMP, linear or cubic, as the continuum's basis), and its stencil.
double basisWeight(int basis, const double x[2], double hl, std::uint32_t ix, std::uint32_t iy,
                   const double grid_min[2], double h);
void basisStencil(int basis, const double x[2], double hl, const double grid_min[2], double h,
                  long lo[2], long hi[2]);

}  // namespace hg
