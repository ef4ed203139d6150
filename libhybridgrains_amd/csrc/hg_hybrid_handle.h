// hg_hybrid_handle.h -- the hybrid handle (host side), shared with the
// scene server.
#pragma once

#include "hg_dem_handle.h"
#include "hg_mpm_handle.h"

namespace hghyb {

struct GrainGrid {
  double *dmass;   // [G] grain mass per node
  double *dmom;    // [2G] grain momentum per node
  double *vafter;  // [2G] node velocity after the exchange
  double *aafter;  // [2G] velocity increment lambda / m_dem
  uint32_t *stamp; // [G] 1 while the node is on the list
  uint32_t *list;  // touched nodes
  uint32_t *count; // [2]: list length of the steps of each parity
  uint32_t max_list;
};

// Per-step homogenisation accumulators on the MPM node storage.
struct HomogGrid {
  double *stress;  // [3G] xx, xy, yy
  double *vel;     // [2G]
  double *mass;    // [G]  this step's grain mass per node
  double *mom;     // [2G] this step's grain momentum per node
  uint32_t *stamp; // [G]  1 while the node is on the list
  uint32_t *list;  // nodes the grains reached this step
  uint32_t *count; // [2]
  uint32_t max_list;
};

}  // namespace hghyb

struct hg_hybrid {
  hg_ctx *ctx = nullptr;
  hg_mpm *mpm = nullptr;
  hg_dem *dem = nullptr;
  hipStream_t stream = nullptr;
  // The grain side runs on a second stream beside the continuum (eager
  // steps): the grain transfer, positions, broad-phase keys, DEM sort and the
  // next step's DEM contacts and integration beside G2P, the MPM sort and
  // P2G; events order the two.
  hipStream_t side = nullptr;
  hipEvent_t fork_ev = nullptr, join_ev = nullptr;
  hipEvent_t g2p_ev = nullptr;  // the fused G2P's completion, for the next grain mass
  bool capturing = false;       // steps are being captured into a graph
  hghyb::GrainGrid g = {};
  double *prex = nullptr, *prey = nullptr;
  uint8_t *coupled = nullptr;
  int32_t *err = nullptr;
  int par = 0;  // parity of the next step (which list counter it uses)
  hghyb::HomogGrid hg = {};
  bool homog_stress = false, homog_vel = false;
  bool graph_homog_stress = false, graph_homog_vel = false;  // as captured
  // coupling weights per MPM node in halves (hyb_homog_factor), recorded
  // while hfac_on
  uint8_t *hfac = nullptr;
  bool hfac_on = false, graph_hfac_on = false;
  hipGraphExec_t graph = nullptr;
  double graph_dt = -1.0;
  int graph_aw = 0, graph_ah = 0;  // MPM tile shape the graph was captured with
  std::vector<void *> allocs;
  uint64_t bytes = 0;  // device bytes of the coupling buffers
};
