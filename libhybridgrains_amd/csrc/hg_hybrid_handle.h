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

}  // namespace hghyb

struct hg_hybrid {
  hg_ctx *ctx = nullptr;
  hg_mpm *mpm = nullptr;
  hg_dem *dem = nullptr;
  hipStream_t stream = nullptr;
  hghyb::GrainGrid g = {};
  double *prex = nullptr, *prey = nullptr;
  uint8_t *coupled = nullptr;
  int32_t *err = nullptr;
  int par = 0;  // parity of the next step (which list counter it uses)
  hipGraphExec_t graph = nullptr;
  double graph_dt = -1.0;
  int graph_aw = 0, graph_ah = 0;  // MPM tile shape the graph was captured with
  std::vector<void *> allocs;
  uint64_t bytes = 0;  // device bytes of the coupling buffers
};
