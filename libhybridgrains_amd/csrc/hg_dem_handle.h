// hg_dem_handle.h -- DEM device layout and the handle (host side), shared
// with the hybrid module.
#pragma once

#include "hg_internal.h"

namespace hgdem {

constexpr int kMaxPlanes = 8;
// Per-body contact capacity (owned pair contacts, and incoming ones): a
// handle picks 8, 16 or 32 from its scenes' radii (hg_dem_contact_capacity);
// a body with more contacts sets the overflow error bit.
constexpr int kCapMin = 8;
constexpr int kCapMax = 32;
constexpr uint32_t kKinematic = 0x80000000u;
// One wavefront per workgroup: scenes of tens of thousands of bodies then
// still spread over every CU.
constexpr int kThreads = 64;

struct SceneDev {
  double g0, g1;
  double kn, kt, gamman, gammat, mu;
  double dt;
  double cmin0, cmin1, cw;  // broad-phase cell origin and width
  double icw;               // 1 / cw
  int32_t ncx, ncy;         // wrapped bucket table: Wx x Wy, powers of two >= 4
  uint32_t b_off, b_cnt, cell_off;
  int32_t nplanes;
  int32_t weighted;
  double plane_x[2 * kMaxPlanes];
  double plane_n[2 * kMaxPlanes];
};

struct BodiesRaw {
  double *x, *y, *th;
  void *vx, *vy, *om;
  double *r;
  void *im, *iI, *m;
  uint8_t *fixed;
  void *w;
};

struct SlotsRaw {
  uint32_t *okey;  // [B][cap]
  void *ods;       // [B][cap][2]
  uint8_t *ocnt;   // [B]
  uint8_t *pact;   // [B][kMaxPlanes]
  void *pds;       // [B][kMaxPlanes][2]
};

// The step's records read after it (contact-force downloads, the hybrid's
// homogenised stress): the force of each owned contact on its owner.
struct ForceRaw {
  void *oF;  // [B][cap][2]
};

// The broad-phase table of DEM-only steps: each bucket keeps up to kslots
// body indices in the order the integration's atomics file them; a body that
// finds its bucket full goes to an overflow list that every body also scans
// (empty in practice), so no candidate is ever lost.  Three tables rotate: a
// step reads table t (filled by the step before), files the moved bodies in
// table t + 1 (cleared by the step before) and clears table t + 2 (read by
// the step before), all in its one launch.
struct TableRaw {
  uint32_t *count;  // [ncell] bodies filed in each bucket (may exceed kslots)
  uint32_t *items;  // [ncell][kslots]
  uint32_t *bkey;   // [B] the bucket each body was filed in
  uint32_t *novf;   // bodies that did not fit
  uint32_t *ovf;    // [B] their indices
};

// Files body i in bucket k of table t (and records where it went).
__device__ __forceinline__ void table_insert(const TableRaw &t, int kslots, uint32_t k,
                                             uint32_t i) {
  t.bkey[i] = k;
  const uint32_t s = atomicAdd(&t.count[k], 1u);
  if (s < uint32_t(kslots)) t.items[size_t(k) * uint32_t(kslots) + s] = i;
  else t.ovf[atomicAdd(t.novf, 1u)] = i;
}

// Broad-phase cell of a position (unbounded; clamped to +-2^30 so that any
// finite or non-finite coordinate gives a valid integer).
__device__ __forceinline__ void dem_cell_of(const SceneDev &sc, double x, double y, int &cx,
                                            int &cy) {
  const double lim = 1073741824.0;
  const double fx = fmin(fmax(floor((x - sc.cmin0) * sc.icw), -lim), lim);
  const double fy = fmin(fmax(floor((y - sc.cmin1) * sc.icw), -lim), lim);
  cx = int(fx);
  cy = int(fy);
}

// The cell's bucket in the scene's wrapped table (a key of the batch).
__device__ __forceinline__ uint32_t dem_bucket(const SceneDev &sc, double x, double y) {
  int cx, cy;
  dem_cell_of(sc, x, y, cx, cy);
  return sc.cell_off + (uint32_t(cx) & uint32_t(sc.ncx - 1)) +
         uint32_t(sc.ncx) * (uint32_t(cy) & uint32_t(sc.ncy - 1));
}

}  // namespace hgdem

struct hg_dem {
  using SceneDev = hgdem::SceneDev;
  using BodiesRaw = hgdem::BodiesRaw;
  using SlotsRaw = hgdem::SlotsRaw;
  using ForceRaw = hgdem::ForceRaw;
  hg_ctx *ctx = nullptr;
  hipStream_t stream = nullptr;
  bool own_stream = true;
  int precision = HG_FP64;
  int cap = hgdem::kCapMin;  // contact slots per body (owned, and incoming)
  std::vector<SceneDev> scenes;
  uint32_t nscenes = 0, nbodies = 0, ncell = 0;
  SceneDev *d_scenes = nullptr;
  uint32_t *d_b_off = nullptr;
  BodiesRaw b = {};
  // a step writes the new position, angle and velocities here, and the two
  // swap after it (b always holds the current state)
  BodiesRaw b_alt = {};
  SlotsRaw slots[2] = {};
  int cur = 0;  // slots[cur] holds the last step's contacts
  ForceRaw f = {};
  // the hybrid's grains: the step does not file the bodies, the coupling's
  // last kernel files them at their final positions (in tables[tph])
  bool caller_files = false;
  int kslots = 8;
  hgdem::TableRaw tables[3] = {};
  int tph = 0;  // the table the next step reads
  int32_t *err = nullptr;
  double *F_out = nullptr;
  double *cpoint = nullptr;  // contact points of the owned circle pairs (kept on request)
  bool keep_force = false;
  double dt_dev = -1.0;
  KernelProfiler prof;
  hipGraphExec_t graph = nullptr;
  uint64_t *obs = nullptr;  // scratch of hg_*_observables (allocated on first use)
  size_t obs_words = 0;
  std::vector<void *> allocs;
  uint64_t bytes = 0;  // device bytes allocated
  size_t tsize() const { return precision == HG_FP64 ? 8 : 4; }
};


// Contact slots per body for a batch of scenes: 8, 16 or 32 (see hg_dem.hip).
int hg_dem_contact_capacity(const hg_dem_desc *descs, uint32_t nscenes);

// Launches one DEM step (one launch: contacts, integration, filing in the
// next bucket table), and re-bins the current positions into the table the
// next step reads (after a caller edited them).
void hg_dem_launch_step(hg_dem *h);
void hg_dem_launch_rebin(hg_dem *h);
// The hybrid's grains: steps leave the filing of the moved bodies to the
// caller, which files each body in tables[tph] after the step.
void hg_dem_caller_files(hg_dem *h);
