// hg_mpm_handle.h -- the MPM handle (host side), shared with the hybrid module.
#pragma once

#include "hg_mpm_kernels.h"

struct hg_mpm {
  using SceneDev = hgmpm::SceneDev;
  using PtsRaw = hgmpm::PtsRaw;
  hg_ctx *ctx = nullptr;
  hipStream_t stream = nullptr;
  bool own_stream = true;
  double hl_first = 0.0;  // first point's half width (the grains' stencil width)
  int precision = HG_FP64;
  int basis = HG_BASIS_UGIMP_LINEAR;
  int aw = 16, ah = 16;  // P2G tile of owned nodes (threads per workgroup = aw*ah)
  int ppass = 1;         // FP32 fused step: points per tile pass, in threads per cell (1 or 2)
  std::vector<SceneDev> scenes;
  uint32_t nscenes = 0;
  uint32_t npoints = 0;
  uint32_t ncell = 0, nnode = 0, ntiles = 0;
  bool has_hl = false, has_flags = false;
  // points store their mass and volume (else each scene's are all equal and
  // live in its SceneDev)
  bool has_mv = false;
  PtsRaw buf[2] = {};
  int cur = 0;
  uint32_t *perm = nullptr, *pinv = nullptr, *key = nullptr, *slot = nullptr, *counts = nullptr,
           *cell_start = nullptr;
  // per-scan-block totals of the cell keys of the points in buf[c], at
  // scan_totals + c * hg_scan_totals_words(nscan)
  uint32_t *scan_totals = nullptr;
  uint32_t nscan = 0;
  SceneDev *d_scenes = nullptr;
  uint32_t *d_p_off = nullptr;
  uint32_t *d_tile_scene = nullptr;
  uint8_t *tile_active = nullptr;
  uint32_t *tile_list = nullptr;
  uint32_t *tile_count = nullptr;  // [2]: active tiles of the list of parity cur
  void *gva = nullptr;
  double *gm = nullptr, *gp = nullptr, *gpn = nullptr, *gf = nullptr;
  // FP32 handles: node sums in whole quanta (kFxQuantities arrays of nnode
  // 64-bit integers) for the fused step.  acc_ready: they hold the next rasterisation
  // of the current points; grid_ready: the node (v, a) (and grid_store's
  // other fields) belong to the current points.
  unsigned long long *acc = nullptr;
  bool acc_ready = false, grid_ready = false;
  bool pinv_ok = false;  // pinv holds the inverse of perm
  hipEvent_t fused_done = nullptr;  // recorded by the next fused G2P's completion
  uint32_t grid_store = 0;
  int32_t *err = nullptr;
  bool keep_all = false;
  uint32_t last_store = 0;
  int phase_next = 0;  // 0: flow or phase 0 allowed; 1: phase 1; 2: phase 2
  double dt_dev = -1.0;  // dt currently in the device scene table
  KernelProfiler prof;
  hipGraphExec_t graph = nullptr;
  uint64_t *obs = nullptr;  // scratch of hg_*_observables (allocated on first use)
  size_t obs_words = 0;
  std::vector<void *> allocs;
  uint64_t bytes = 0;  // device bytes allocated

  size_t tsize() const { return precision == HG_FP64 ? 8 : 4; }
  int tw() const;
  int th() const;
};


// The two MPM halves of the hybrid step (defined in hg_mpm.hip): P2G with
// the grid update, storing node mass, both momenta and (v, a); then G2P and
// the sort.
hg_status hg_mpm_launch_coupled_p2g(hg_mpm *h, double dt);
void hg_mpm_launch_g2p(hg_mpm *h);
// Before capturing steps into a graph: FP32 handles rasterise the current
// points now if no step has (so every captured step finds the sums ready).
void hg_mpm_prime(hg_mpm *h);
// FP32 handles that run the fused step (the hybrid's continuum reorders the
// halves around the grain work):
bool hg_mpm_is_fused(const hg_mpm *h);
// dt on the device and the node sums of the current points ready (a
// rasterisation if no step has left them)
hg_status hg_mpm_prepare_fused(hg_mpm *h, double dt);
// the grid update of the current points from the node sums (rasterised
// first if no step has), storing (v, a) (and the node mass with keep_mass),
// with the grain exchange at every node that holds grain mass (g may be null);
// done (if not null) is recorded by the kernel's own completion
hg_status hg_mpm_launch_coupled_grid(hg_mpm *h, double dt, const hgmpm::GrainNodes *g,
                                     bool keep_mass, hipEvent_t done = nullptr);
// the fused G2P and rasterisation alone, then the sort of the moved points
void hg_mpm_launch_fused_g2p(hg_mpm *h, hipEvent_t done = nullptr);
void hg_mpm_launch_fused_sort(hg_mpm *h);
// the node mass sums (whole quanta) and the quantum of scene 0
const unsigned long long *hg_mpm_mass_sums(const hg_mpm *h);
double hg_mpm_mass_quantum(const hg_mpm *h);
// Re-keys and re-sorts the points of the current buffer, as at creation,
// and waits for it (used after the buffer was rewritten).
hg_status hg_mpm_rebin(hg_mpm *h);
