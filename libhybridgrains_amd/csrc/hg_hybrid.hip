// hg_hybrid.hip -- the hybridgrains2dnew node-node coupled step on gfx950.
//
// Replaces HybridGrains2DSim::stepSystem_constraintSetUp_prediction_correction_
// advection_node_node (hybridgrains2dnew/HybridGrains2DSim.cpp:116-227) and
// HybridCoupling2D (Coupling.cpp:22-591); the periodic zone resampling
// (updateHybridState) is outside the explicit step and not part of it.
//
// One MPM scene and one DEM scene share a stream.  The grain-side grid
// buffers (grain mass, grain momentum, post-exchange velocity and velocity
// increment) live on the MPM node storage; every node a coupled grain touches
// is appended once to a list (a stamp per node), and the list's nodes are
// cleared again at the end of the step, so nothing is swept over the whole
// grid and the step replays correctly from a hipGraph.  The list length
// alternates between two counters: each step's last kernel clears the one
// the next step uses.  In FP32 without the coupling weights (which read the
// grain mass after the grid update) there is no list: the grid update, which
// visits every node that can hold grain mass, clears a node's grain mass and
// momentum once it has exchanged there.
//
// Kernels of one step:
//   continuum stream       | grain stream
//   P2G + grid update      | DEM step (contacts and integration, one launch)
//   grain_mass (coupling test, grain mass and momentum spread)
//   exchange (also refreshes the node velocity and acceleration P2G wrote
//   at the nodes it changes)
//   G2P, MPM sort, tiles   | transfer, finish (coupled grain positions, every
//                          | grain filed in the DEM's next bucket table, the
//                          | next step's pre-step positions, list clean-up),
//                          | homogenisation when on, then the next step's
//                          | DEM step
// with events where one stream needs the other's results (FP32, with the
// fused continuum step: grid update with the exchange, then the fused G2P and
// the MPM sort beside the grain branch, whose next step's grain mass waits
// for the fused G2P's completion).  The fork is the completion of the continuum
// stream's last kernel before it (launched with its completion event), not an
// event record of its own: a record between two kernels of a stream delays
// the second by a few microseconds.  The pre-step
// positions the coupling uses are saved by the previous step's finish (and
// at creation), so the DEM integration may move the grains while grain_mass
// still reads them.  A graph capture queues the same kernels on both streams,
// ordered by recorded events (the kernels' completion events are for eager
// steps only).
#include "hg_hybrid_handle.h"
#include "hg_sort.h"

#include <hip/hip_ext.h>

#include <memory>

namespace hghyb {

// Grain kernels run one wavefront per workgroup: the hybrid zone holds tens
// of thousands of grains, and small workgroups spread them over every CU.
constexpr int kThreads = 64;

struct HybArgs {
  const hgmpm::SceneDev *msc;  // MPM scene 0
  const uint8_t *tile_active;
  double *gm, *gpn;
  const double *gp;
  void *gva;
  int32_t tw, th;
  double hl;  // grains use the first point's half width (Coupling.cpp:31)
  hgdem::BodiesRaw b;
  uint32_t nbodies;
  const hgdem::SceneDev *dsc;  // DEM scene 0 (broad-phase buckets)
  hgdem::TableRaw dtab;        // the bucket table the next DEM step reads
  int32_t kslots;
  double *prex, *prey;
  uint8_t *coupled;
  GrainGrid g;
  uint32_t *count, *next_count;  // this step's list length; the next step's
  int32_t *err;
  uint8_t *hfac;  // coupling weights per node in halves (0, 1, 2), or null
  // FP32 fused continuum: the node mass sums in quanta, and the quantum
  // (null when grain_mass reads the grid update's mass)
  const unsigned long long *acc_mass;
  double mass_quantum;
  // the grid update clears the grain mass and momentum where it exchanges
  // (FP32, no coupling weights): the grain mass neither stamps nor lists
  // the nodes it touches, and the finish has none to clear
  int32_t listless;
};

template <typename T>
__device__ __forceinline__ T ld(const void *p, uint32_t i) {
  return static_cast<const T *>(p)[i];
}

// Stencil of a grain at (x, y) on the MPM grid, as node index ranges that may
// reach off the grid; false if the coordinates are not finite or absurd.
template <int BASIS>
__device__ __forceinline__ bool grain_stencil(const hgmpm::SceneDev &sc, double x, double y,
                                              double hl, int &lx, int &hx, int &ly, int &hy) {
  double l0, h0, l1, h1;
  hgmpm::stencil_1d<BASIS>(x, hl, sc.min0, sc.h, sc.inv_h, l0, h0);
  hgmpm::stencil_1d<BASIS>(y, hl, sc.min1, sc.h, sc.inv_h, l1, h1);
  const double lim = 1.0e9;
  if (!(fabs(l0) < lim && fabs(h0) < lim && fabs(l1) < lim && fabs(h1) < lim)) return false;
  lx = int(l0);
  hx = int(h0);
  ly = int(l1);
  hy = int(h1);
  return true;
}

template <int BASIS>
__device__ __forceinline__ double grain_weight(const hgmpm::SceneDev &sc, double x, double y,
                                               double hl, int ix, int iy, int lx, int hx, int ly,
                                               int hy) {
  double ax, dx, ay, dy;
  hgmpm::basis_factor<double, BASIS>(x, hl, sc.min0, sc.h, sc.inv_h, ix, ix - lx, hx - lx + 1, ax,
                                     dx);
  hgmpm::basis_factor<double, BASIS>(y, hl, sc.min1, sc.h, sc.inv_h, iy, iy - ly, hy - ly + 1, ay,
                                     dy);
  if constexpr (BASIS == HG_BASIS_UGIMP_LINEAR) return (ax * (1.0 / (4.0 * hl * hl))) * ay;
  return ax * ay;
}

// A grain's stencil with its 1-D weights: node (lx + i, ly + j) of the S x S
// box has weight wx[i] * wy[j] (the x factor carries uGIMP's 1/(4 hl^2), as
// grain_weight applies it) when bit j*S + i of `on` is set (inside the
// stencil and on the grid).
template <int BASIS>
struct GrainStencil {
  static constexpr int S = hgmpm::BasisTraits<BASIS>::S;
  int lx, ly;
  uint32_t on;
  double wx[S], wy[S];
};

// The stencil's nodes alone (lx, ly and on; the weights are not formed).
template <int BASIS>
__device__ __forceinline__ bool grain_stencil_nodes(const hgmpm::SceneDev &sc, double x, double y,
                                                    double hl, GrainStencil<BASIS> &g, int &hx,
                                                    int &hy) {
  constexpr int S = GrainStencil<BASIS>::S;
  int lx, ly;
  g.on = 0u;
  if (!grain_stencil<BASIS>(sc, x, y, hl, lx, hx, ly, hy)) return false;
  g.lx = lx;
  g.ly = ly;
#pragma unroll
  for (int j = 0; j < S; ++j)
#pragma unroll
    for (int i = 0; i < S; ++i) {
      const int ix = lx + i, iy = ly + j;
      if (ix <= hx && iy <= hy && ix >= 0 && iy >= 0 && ix < sc.nnx && iy < sc.nny)
        g.on |= 1u << (j * S + i);
    }
  return true;
}

// The 1-D weights of a stencil whose nodes grain_stencil_nodes found.
template <int BASIS>
__device__ __forceinline__ void grain_stencil_fill(const hgmpm::SceneDev &sc, double x, double y,
                                                   double hl, int hx, int hy,
                                                   GrainStencil<BASIS> &g) {
  constexpr int S = GrainStencil<BASIS>::S;
  const int lx = g.lx, ly = g.ly;
  const double vnorm = (BASIS == HG_BASIS_UGIMP_LINEAR) ? 1.0 / (4.0 * hl * hl) : 1.0;
#pragma unroll
  for (int i = 0; i < S; ++i) {
    double ax = 0.0, ay = 0.0, d;
    if (lx + i <= hx)
      hgmpm::basis_factor<double, BASIS>(x, hl, sc.min0, sc.h, sc.inv_h, lx + i, i, hx - lx + 1,
                                         ax, d);
    if (ly + i <= hy)
      hgmpm::basis_factor<double, BASIS>(y, hl, sc.min1, sc.h, sc.inv_h, ly + i, i, hy - ly + 1,
                                         ay, d);
    g.wx[i] = (BASIS == HG_BASIS_UGIMP_LINEAR) ? ax * vnorm : ax;
    g.wy[i] = ay;
  }
}

template <int BASIS>
__device__ __forceinline__ bool grain_stencil_weights(const hgmpm::SceneDev &sc, double x,
                                                      double y, double hl,
                                                      GrainStencil<BASIS> &g) {
  int hx, hy;
  if (!grain_stencil_nodes<BASIS>(sc, x, y, hl, g, hx, hy)) return false;
  grain_stencil_fill<BASIS>(sc, x, y, hl, hx, hy, g);
  return true;
}

__device__ __forceinline__ bool node_active(const HybArgs &a, int ix, int iy) {
  const hgmpm::SceneDev &sc = *a.msc;
  const uint32_t tile = sc.tile_off + uint32_t(iy / a.th) * uint32_t(sc.tiles_x) + uint32_t(ix / a.tw);
  return a.tile_active[tile] != 0;
}

// HybridCoupling2D::rasterizeGrainMass (Coupling.cpp:22-129): a grain is
// coupled when every on-grid node of its stencil (at its pre-step position)
// carries more than 1e-6 of continuum mass; its mass is then spread with the
// MPM weights.  The same pass spreads its momentum at the post-DEM-step
// velocity (rasterizeGrainMom, Coupling.cpp:287-466), so the step's DEM step
// must be complete when it runs.
template <typename T, int BASIS>
__global__ __launch_bounds__(kThreads) void hyb_grain_mass(HybArgs a) {
  constexpr int S = GrainStencil<BASIS>::S;
  const uint32_t b = blockIdx.x * kThreads + threadIdx.x;
  const int lane = int(threadIdx.x & 63u);
  // every lane of a wave reaches the list append below
  uint32_t fresh = 0u;  // stencil nodes this grain stamped first
  GrainStencil<BASIS> g;
  g.on = 0u;
  if (b < a.nbodies) {
    const double x = a.prex[b], y = a.prey[b];
    a.coupled[b] = 0;
    const hgmpm::SceneDev &sc = *a.msc;
    int hx, hy;
    // the coupling test first (the weights only for a grain that passes)
    if (!a.b.fixed[b] && grain_stencil_nodes<BASIS>(sc, x, y, a.hl, g, hx, hy)) {
      bool all_pos = true;
#pragma unroll
      for (int q = 0; q < S * S; ++q) {
        if (!((g.on >> q) & 1u)) continue;
        const int ix = g.lx + q % S, iy = g.ly + q / S;
        const uint32_t node = sc.node_off + uint32_t(iy) * uint32_t(sc.snx) + uint32_t(ix);
        // the continuum mass of this step's grid: from the node sums when
        // they are what the grid update will read (zero off the points'
        // stencils), else from the grid update's output on active tiles
        const double m = a.acc_mass ? double((long long)a.acc_mass[node]) * a.mass_quantum
                                    : (node_active(a, ix, iy) ? a.gm[node] : 0.0);
        if (m <= 1.0e-6) all_pos = false;
      }
      if (all_pos) {
        grain_stencil_fill<BASIS>(sc, x, y, a.hl, hx, hy, g);
        a.coupled[b] = 1;
        const double mb = double(ld<T>(a.b.m, b));
        const double pm0 = mb * double(ld<T>(a.b.vx, b)), pm1 = mb * double(ld<T>(a.b.vy, b));
        // mass and momentum spread and stamps, all nodes at once
#pragma unroll
        for (int q = 0; q < S * S; ++q) {
          if (!((g.on >> q) & 1u)) continue;
          const uint32_t n = uint32_t(g.ly + q / S) * uint32_t(sc.snx) + uint32_t(g.lx + q % S);
          const double w = g.wx[q % S] * g.wy[q / S];
          atomicAdd(&a.g.dmass[n], w * mb);
          atomicAdd(&a.g.dmom[2 * n], w * pm0);
          atomicAdd(&a.g.dmom[2 * n + 1], w * pm1);
          if (!a.listless && atomicExch(&a.g.stamp[n], 1u) == 0u) fresh |= 1u << q;
        }
      }
    }
  }
  if (a.listless) return;
  // the first grain to stamp a node appends it to the list: one counter
  // update per wave, consecutive entries per lane
  const uint32_t mine = uint32_t(__popc(fresh));
  const uint32_t incl = hg_wave_incl_scan(mine);
  const uint32_t total = uint32_t(__shfl(int(incl), 63, 64));
  if (total == 0u) return;
  uint32_t base = 0;
  if (lane == 0) base = atomicAdd(a.count, total);
  base = uint32_t(__shfl(int(base), 0, 64)) + (incl - mine);
  if (fresh) {
    const hgmpm::SceneDev &sc = *a.msc;
#pragma unroll
    for (int q = 0; q < S * S; ++q) {
      if (!((fresh >> q) & 1u)) continue;
      const uint32_t n = uint32_t(g.ly + q / S) * uint32_t(sc.snx) + uint32_t(g.lx + q % S);
      if (base < a.g.max_list) a.g.list[base] = n;
      else atomicOr(a.err, 8);
      ++base;
    }
  }
}

// HybridCoupling2D::updateVelocitiesOnGrid (Coupling.cpp:533-579): the
// multiplier that equalises the two velocities at a shared node while
// conserving its momentum.  The node's velocity and acceleration are then
// recomputed from its new momentum as the grid update computes them
// (PhysicsGrid::lumpedMassVelocityAndAccelerationUpdate,
// PhysicsGrid.cpp:588-609); nodes the exchange leaves alone keep the values
// P2G wrote.
template <typename T>
__global__ __launch_bounds__(kThreads) void hyb_exchange(HybArgs a) {
  const uint32_t cnt = min(*a.count, a.g.max_list);
  for (uint32_t k = blockIdx.x * kThreads + threadIdx.x; k < cnt; k += gridDim.x * kThreads) {
    const uint32_t n = a.g.list[k];
    if (!(a.g.dmass[n] > 0.0)) continue;
    const hgmpm::GrainNodes gn = {a.g.dmass, a.g.dmom, a.g.vafter, a.g.aafter};
    hgmpm::exchange_node<T>(a.msc->dt, a.msc->node_off + n, n, gn, a.gm, a.gp, a.gpn, a.gva);
  }
}

// The continuum grid's coupling weights as HybridGrains2DSim's node-node step
// saves them (HybridGrains2DSim.cpp:169-197), in halves: 1 (0.5) at nodes
// holding grain mass, else 2 (1.0) at nodes with continuum mass after this
// step's P2G, else 0.  Output only.  It runs on the continuum stream after
// the exchange and before the grain branch forks off, so it precedes both
// the finish (which clears the grain mass) and the sort after G2P (which
// rewrites the tile flags it reads for the next step).  Nodes of tiles
// without points carry no mass (their P2G values are not this step's).
__global__ __launch_bounds__(kThreads) void hyb_homog_factor(HybArgs a) {
  // four nodes of one row per thread (tile widths are multiples of four);
  // blockIdx.y is the node row
  const hgmpm::SceneDev &sc = *a.msc;
  const int iy = int(blockIdx.y);
  const int ix0 = int(blockIdx.x * kThreads + threadIdx.x) * 4;
  if (ix0 >= sc.nnx) return;
  const uint32_t row = uint32_t(iy) * uint32_t(sc.snx);
  const bool active = node_active(a, ix0, iy);
  uint8_t c[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const uint32_t n = row + uint32_t(ix0 + k);
    c[k] = 0u;
    // grain mass only reaches nodes with continuum mass, so only nodes of
    // tiles with points are read
    if (active && ix0 + k < sc.nnx) {
      if (a.g.dmass[n] > 0.0) c[k] = 1u;
      else if (a.gm[sc.node_off + n] > 0.0) c[k] = 2u;
    }
  }
#pragma unroll
  for (int k = 0; k < 4; ++k)
    if (ix0 + k < sc.nnx) a.hfac[row + uint32_t(ix0 + k)] = c[k];
}

// HybridCoupling2D::transferToGrain (Coupling.cpp:468-527), FLIP with alpha 1.
template <typename T, int BASIS>
__device__ __forceinline__ void transfer_grain(const HybArgs &a, uint32_t b) {
  constexpr int S = GrainStencil<BASIS>::S;
  const hgmpm::SceneDev &sc = *a.msc;
  GrainStencil<BASIS> g;
  grain_stencil_weights<BASIS>(sc, a.prex[b], a.prey[b], a.hl, g);
  double vp0 = 0.0, vp1 = 0.0, a0 = 0.0, a1 = 0.0;
#pragma unroll
  for (int q = 0; q < S * S; ++q) {
    if (!((g.on >> q) & 1u)) continue;
    const uint32_t n = uint32_t(g.ly + q / S) * uint32_t(sc.snx) + uint32_t(g.lx + q % S);
    const double w = g.wx[q % S] * g.wy[q / S];
    vp0 += w * a.g.vafter[2 * n];
    vp1 += w * a.g.vafter[2 * n + 1];
    a0 += w * a.g.aafter[2 * n];
    a1 += w * a.g.aafter[2 * n + 1];
  }
  const double vf0 = double(ld<T>(a.b.vx, b)) + a0, vf1 = double(ld<T>(a.b.vy, b)) + a1;
  static_cast<T *>(a.b.vx)[b] = T(vp0 * (1.0 - 1.0) + vf0 * 1.0);
  static_cast<T *>(a.b.vy)[b] = T(vp1 * (1.0 - 1.0) + vf1 * 1.0);
}

template <typename T, int BASIS>
__global__ __launch_bounds__(kThreads) void hyb_transfer(HybArgs a) {
  const uint32_t b = blockIdx.x * kThreads + threadIdx.x;
  if (b >= a.nbodies || !a.coupled[b]) return;
  transfer_grain<T, BASIS>(a, b);
}

// End of the step: HybridCoupling2D::updatePositions_node_node
// (Coupling.cpp:581-591) for the coupled grains, every grain filed in the
// bucket table of the next DEM step at its final position, its position for
// the next step's coupling, and the clean-up of the touched nodes and of the
// next step's list counter.
template <typename T>
__device__ __forceinline__ void finish_grain(const HybArgs &a, uint32_t b, double dt) {
  double x = a.b.x[b], y = a.b.y[b];
  if (a.coupled[b]) {
    x = a.prex[b] + dt * double(ld<T>(a.b.vx, b));
    y = a.prey[b] + dt * double(ld<T>(a.b.vy, b));
    a.b.x[b] = x;
    a.b.y[b] = y;
  }
  a.prex[b] = x;
  a.prey[b] = y;
  hgdem::table_insert(a.dtab, a.kslots, hgdem::dem_bucket(*a.dsc, x, y), b);
}

// Transfer and finish of each grain in one launch, where no node list is kept
// (listless): nothing then clears node values a transfer of another grain
// still reads.
template <typename T, int BASIS>
__global__ __launch_bounds__(kThreads) void hyb_transfer_finish(HybArgs a, double dt) {
  const uint32_t b = blockIdx.x * kThreads + threadIdx.x;
  if (b == 0) *a.next_count = 0u;
  if (b >= a.nbodies) return;
  if (a.coupled[b]) transfer_grain<T, BASIS>(a, b);
  finish_grain<T>(a, b, dt);
}

template <typename T>
__global__ __launch_bounds__(kThreads) void hyb_finish(HybArgs a, double dt) {
  const uint32_t tid = blockIdx.x * kThreads + threadIdx.x;
  if (tid == 0) *a.next_count = 0u;
  if (tid < a.nbodies) finish_grain<T>(a, tid, dt);
  const uint32_t cnt = min(*a.count, a.g.max_list);
  for (uint32_t k = tid; k < cnt; k += gridDim.x * kThreads) {
    const uint32_t n = a.g.list[k];
    a.g.dmass[n] = 0.0;
    a.g.dmom[2 * n] = 0.0;
    a.g.dmom[2 * n + 1] = 0.0;
    a.g.vafter[2 * n] = 0.0;
    a.g.vafter[2 * n + 1] = 0.0;
    a.g.aafter[2 * n] = 0.0;
    a.g.aafter[2 * n + 1] = 0.0;
    a.g.stamp[n] = 0u;
  }
}

// ---- per-step homogenisation (ResamplingTools.cpp:744-884) -------------------

struct HomogArgs {
  HomogGrid hg;
  hgdem::SlotsRaw s;   // the contacts this step wrote
  const void *oF;      // their forces on the owner
  const double *cpoint;
  uint32_t *count, *next_count;
  int32_t stress, vel;
  uint32_t cap;        // contact slots per body
};

// Spreads one quantity set from position (x, y) with the MPM weights of half
// width hl = h/4, as ResamplingTools does for both fields.
template <int BASIS>
__device__ __forceinline__ bool homog_stencil(const hgmpm::SceneDev &sc, double x, double y,
                                              int &lx, int &hx, int &ly, int &hy) {
  return grain_stencil<BASIS>(sc, x, y, 0.25 * sc.h, lx, hx, ly, hy);
}

// Per grain: ResamplingTools::homogenizeStress (ResamplingTools.cpp:744-801)
// for the circle contacts it owns -- stress -(f r^T + r f^T)/2 with r at the
// end-of-step positions, spread from the contact point and divided by the
// cell area -- and the mass and momentum half of homogenizeVelocity
// (:818-866) for the grain itself.
template <typename T, int BASIS>
__global__ __launch_bounds__(kThreads) void hyb_homog_grains(HybArgs a, HomogArgs g) {
  const uint32_t b = blockIdx.x * kThreads + threadIdx.x;
  if (b >= a.nbodies) return;
  const hgmpm::SceneDev &sc = *a.msc;
  const double hl = 0.25 * sc.h;
  const double x = a.b.x[b], y = a.b.y[b];
  if (g.stress) {
    const double inv_area = 1.0 / (sc.h * sc.h);
    const int n = g.s.ocnt[b];
    for (int k = 0; k < n; ++k) {
      const uint32_t sl = b * g.cap + uint32_t(k);
      const uint32_t key = g.s.okey[sl];
      if (key & hgdem::kKinematic) continue;
      const uint32_t j = key;  // one DEM scene: bodies are numbered from 0
      const double f0 = double(ld<T>(g.oF, 2 * sl)), f1 = double(ld<T>(g.oF, 2 * sl + 1));
      const double r0 = x - a.b.x[j], r1 = y - a.b.y[j];
      const double sxx = -0.5 * (f0 * r0 + r0 * f0);
      const double sxy = -0.5 * (f1 * r0 + r1 * f0);
      const double syy = -0.5 * (f1 * r1 + r1 * f1);
      const double px = g.cpoint[2 * sl], py = g.cpoint[2 * sl + 1];
      int lx, hx, ly, hy;
      if (!homog_stencil<BASIS>(sc, px, py, lx, hx, ly, hy)) continue;
      for (int iy = ly; iy <= hy; ++iy)
        for (int ix = lx; ix <= hx; ++ix) {
          if (ix < 0 || iy < 0 || ix >= sc.nnx || iy >= sc.nny) continue;
          const uint32_t node = uint32_t(iy) * uint32_t(sc.snx) + uint32_t(ix);
          const double w = grain_weight<BASIS>(sc, px, py, hl, ix, iy, lx, hx, ly, hy);
          atomicAdd(&g.hg.stress[3 * node], w * sxx * inv_area);
          atomicAdd(&g.hg.stress[3 * node + 1], w * sxy * inv_area);
          atomicAdd(&g.hg.stress[3 * node + 2], w * syy * inv_area);
        }
    }
  }
  if (g.vel && !a.b.fixed[b]) {
    const double m = double(ld<T>(a.b.m, b));
    const double v0 = double(ld<T>(a.b.vx, b)), v1 = double(ld<T>(a.b.vy, b));
    int lx, hx, ly, hy;
    if (homog_stencil<BASIS>(sc, x, y, lx, hx, ly, hy)) {
      for (int iy = ly; iy <= hy; ++iy)
        for (int ix = lx; ix <= hx; ++ix) {
          if (ix < 0 || iy < 0 || ix >= sc.nnx || iy >= sc.nny) continue;
          const uint32_t node = uint32_t(iy) * uint32_t(sc.snx) + uint32_t(ix);
          const double w = grain_weight<BASIS>(sc, x, y, hl, ix, iy, lx, hx, ly, hy);
          atomicAdd(&g.hg.mom[2 * node], w * m * v0);
          atomicAdd(&g.hg.mom[2 * node + 1], w * m * v1);
          atomicAdd(&g.hg.mass[node], w * m);
          if (g.hg.stamp[node] == 0u && atomicExch(&g.hg.stamp[node], 1u) == 0u) {
            const uint32_t e = atomicAdd(g.count, 1u);
            if (e < g.hg.max_list) g.hg.list[e] = node;
            else atomicOr(a.err, 8);
          }
        }
    }
  }
}

// The rest of homogenizeVelocity (ResamplingTools.cpp:868-873): each reached
// node with positive grain mass adds its velocity; then the step's mass and
// momentum are cleared, and the next step's list counter.
__global__ __launch_bounds__(kThreads) void hyb_homog_nodes(HomogArgs g) {
  const uint32_t tid = blockIdx.x * kThreads + threadIdx.x;
  if (tid == 0) *g.next_count = 0u;
  const uint32_t cnt = min(*g.count, g.hg.max_list);
  for (uint32_t k = tid; k < cnt; k += gridDim.x * kThreads) {
    const uint32_t n = g.hg.list[k];
    const double m = g.hg.mass[n];
    if (m > 0.0) {
      g.hg.vel[2 * n] += g.hg.mom[2 * n] / m;
      g.hg.vel[2 * n + 1] += g.hg.mom[2 * n + 1] / m;
    }
    g.hg.mass[n] = 0.0;
    g.hg.mom[2 * n] = 0.0;
    g.hg.mom[2 * n + 1] = 0.0;
    g.hg.stamp[n] = 0u;
  }
}

}  // namespace hghyb

using namespace hghyb;


template <typename P>
static hg_status halloc(hg_hybrid *h, P **out, size_t bytes) {
  void *p = nullptr;
  if (bytes == 0) bytes = 16;
  hipError_t e = hipMalloc(&p, bytes);
  if (e != hipSuccess)
    return hg_fail(h->ctx, HG_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  h->allocs.push_back(p);
  h->bytes += bytes;
  *out = static_cast<P *>(p);
  return HG_OK;
}

#define HG_TRY(expr)            \
  do {                          \
    hg_status s_ = (expr);      \
    if (s_ != HG_OK) return s_; \
  } while (0)

static HybArgs make_args(hg_hybrid *h) {
  HybArgs a;
  a.msc = h->mpm->d_scenes;
  a.tile_active = h->mpm->tile_active;
  a.gm = h->mpm->gm;
  a.gpn = h->mpm->gpn;
  a.gp = h->mpm->gp;
  a.gva = h->mpm->gva;
  a.tw = h->mpm->tw();
  a.th = h->mpm->th();
  a.hl = h->mpm->hl_first;
  a.b = h->dem->b;
  a.nbodies = h->dem->nbodies;
  a.dsc = h->dem->d_scenes;
  // the table the DEM step after this one reads (this step's DEM step has
  // advanced the rotation already)
  a.dtab = h->dem->tables[h->dem->tph];
  a.kslots = h->dem->kslots;
  a.prex = h->prex;
  a.prey = h->prey;
  a.coupled = h->coupled;
  a.g = h->g;
  a.count = h->g.count + h->par;
  a.next_count = h->g.count + (1 - h->par);
  a.err = h->err;
  a.hfac = h->hfac_on ? h->hfac : nullptr;
  a.listless = hg_mpm_is_fused(h->mpm) && !h->hfac_on ? 1 : 0;
  a.acc_mass = nullptr;
  a.mass_quantum = 0.0;
  return a;
}

// The events that order the two streams of one device: no timing, and no
// system-scope fence at the record (each kernel's own release makes its
// writes visible on the device, and the host reads results after the
// continuum stream's copies).  Measured with in-kernel clocks, the kernel
// after a record then starts about 4.7 us after its predecessor instead of
// 7.5 us, and a waiting stream's kernel 9.3 us instead of 11 us after the
// event's kernel.
constexpr unsigned kOrderEventFlags = hipEventDisableTiming | hipEventDisableSystemFence;

// n steps.  concurrent: the grain branch runs on the second stream (eager
// steps), and each step's DEM contacts and integration are queued on it as
// soon as the previous step's grain work is queued, so they overlap the
// continuum's G2P and sort (also when captured into a graph: replays of the
// two branches take 0.115 ms per step here against 0.125 ms on one stream,
// and eager steps 0.104 ms).
template <typename T, int BASIS>
static void launch_hybrid_steps(hg_hybrid *h, double dt, uint32_t n, bool concurrent) {
  if (n == 0) return;
  KernelProfiler &prof = h->mpm->prof;
  const hipStream_t s = h->stream, side = concurrent ? h->side : h->stream;
  const uint32_t nb = std::max(h->dem->nbodies, 1u);
  const dim3 gb((nb + kThreads - 1) / kThreads), tb(kThreads);
  // work queued on stream from before this point precedes what stream to
  // queues after it
  auto order = [&](hipEvent_t ev, hipStream_t from, hipStream_t to) {
    if (from == to) return;
    hipEventRecord(ev, from);
    hipStreamWaitEvent(to, ev, 0);
  };
  auto dem_step = [&]() {
    h->dem->stream = side;
    hg_dem_launch_step(h->dem);
    h->dem->stream = s;
  };
  auto homog = [&](const HybArgs &a) {
    if (!(h->homog_stress || h->homog_vel)) return;
    HomogArgs g;
    g.hg = h->hg;
    g.s = h->dem->slots[h->dem->cur];
    g.oF = h->dem->f.oF;
    g.cpoint = h->dem->cpoint;
    g.count = h->hg.count + h->par;
    g.next_count = h->hg.count + (1 - h->par);
    g.stress = h->homog_stress ? 1 : 0;
    g.vel = h->homog_vel ? 1 : 0;
    g.cap = uint32_t(h->dem->cap);
    {
      ProfScope ps(prof, "hyb_homog_grains", side);
      hipLaunchKernelGGL((hyb_homog_grains<T, BASIS>), gb, tb, 0, side, a, g);
    }
    if (h->homog_vel) {
      ProfScope ps(prof, "hyb_homog_nodes", side);
      hipLaunchKernelGGL(hyb_homog_nodes, gb, tb, 0, side, g);
    }
  };
  auto homog_factor = [&](const HybArgs &a) {
    if (!a.hfac) return;
    // on the continuum stream, before the fork: it reads the tile flags,
    // which the sort after G2P rewrites on this stream for the next step
    const hgmpm::SceneDev &msc = h->mpm->scenes[0];
    const uint32_t per_row = (uint32_t(msc.nnx) + 4 * kThreads - 1) / (4 * kThreads);
    ProfScope ps(prof, "hyb_homog_factor", s);
    hipLaunchKernelGGL(hyb_homog_factor, dim3(per_row, uint32_t(msc.nny)), tb, 0, s, a);
  };
  auto grain_end = [&](const HybArgs &a) {
    if (a.listless) {
      ProfScope ps(prof, "hyb_transfer_finish", side);
      hipLaunchKernelGGL((hyb_transfer_finish<T, BASIS>), gb, tb, 0, side, a, dt);
    } else {
      {
        ProfScope ps(prof, "hyb_transfer", side);
        hipLaunchKernelGGL((hyb_transfer<T, BASIS>), gb, tb, 0, side, a);
      }
      {
        ProfScope ps(prof, "hyb_finish", side);
        hipLaunchKernelGGL(hyb_finish<T>, gb, tb, 0, side, a, dt);
      }
    }
    homog(a);
  };
  if (hg_mpm_is_fused(h->mpm)) {
    // FP32 continuum.  The grid update applies the exchange at the nodes
    // holding grain mass.  The grain branch (end of the coupling, the next
    // step's DEM step) runs beside the fused G2P; the next step's grain mass,
    // which reads the continuum mass from the node sums the fused G2P
    // completes, follows on the grain stream once the fused G2P's completion
    // event has fired, beside the continuum's sort; the continuum stream
    // waits for it before the next grid update.  (On one stream the grain
    // mass follows the sort.)
    hg_mpm_prepare_fused(h->mpm, dt);
    const bool listless = !h->hfac_on;
    const hgmpm::GrainNodes gn = {h->g.dmass, h->g.dmom, h->g.vafter, h->g.aafter,
                                  listless ? h->g.dmass : nullptr,
                                  listless ? h->g.dmom : nullptr};
    auto grain_mass_on = [&](hipStream_t st, hipEvent_t done = nullptr) {
      HybArgs a = make_args(h);
      a.acc_mass = hg_mpm_mass_sums(h->mpm);
      a.mass_quantum = hg_mpm_mass_quantum(h->mpm);
      ProfScope ps(prof, "hyb_grain_mass", st);
      if (done) hipExtLaunchKernelGGL((hyb_grain_mass<T, BASIS>), gb, tb, 0, st, nullptr, done, 0, a);
      else hipLaunchKernelGGL((hyb_grain_mass<T, BASIS>), gb, tb, 0, st, a);
    };
    order(h->fork_ev, s, side);
    dem_step();
    order(h->join_ev, side, s);
    grain_mass_on(s);
    for (uint32_t k = 0; k < n; ++k) {
      const HybArgs a = make_args(h);
      // the node mass is kept for the coupling weights when they are recorded;
      // the fork is the grid update's own completion when nothing follows it
      // before the fork (no event record between it and the fused G2P)
      // (kernels' completion events are not captured: a graph orders its
      // branches by recorded events, which cost nothing in replays)
      const bool ext = s != side && !h->capturing;
      const bool fork_at_grid = ext && !a.hfac;
      hg_mpm_launch_coupled_grid(h->mpm, dt, &gn, h->hfac_on, fork_at_grid ? h->fork_ev : nullptr);
      homog_factor(a);
      if (fork_at_grid) hipStreamWaitEvent(side, h->fork_ev, 0);
      else order(h->fork_ev, s, side);
      grain_end(a);
      if (k + 1 < n) dem_step();  // the next step's DEM step, ahead
      // the next step's grain mass on the grain stream, beside the MPM sort:
      // it waits for the fused G2P's completion (its node sums)
      const bool gm_side = s != side && k + 1 < n;
      hg_mpm_launch_fused_g2p(h->mpm, (gm_side && ext) ? h->g2p_ev : nullptr);
      if (gm_side && !ext) hipEventRecord(h->g2p_ev, s);
      h->par = 1 - h->par;
      hg_mpm_launch_fused_sort(h->mpm);
      if (gm_side) {
        // the join is the grain mass's own completion
        hipStreamWaitEvent(side, h->g2p_ev, 0);
        grain_mass_on(side, ext ? h->join_ev : nullptr);
        if (ext) hipStreamWaitEvent(s, h->join_ev, 0);
        else order(h->join_ev, side, s);
      } else {
        order(h->join_ev, side, s);
      }
      if (k + 1 < n && !gm_side) grain_mass_on(s);
    }
    return;
  }
  order(h->fork_ev, s, side);
  dem_step();
  for (uint32_t k = 0; k < n; ++k) {
    HybArgs a = make_args(h);
    // grain_mass needs this step's DEM step (post-step velocities) and, queued
    // before it on the grain stream, the previous step's finish (pre-step
    // positions, cleared node lists).  The join comes before the P2G: the
    // P2G's tiles fill every CU in a single wave of workgroups, so a grain
    // kernel still holding a CU when it starts would delay some of its tiles
    // by a whole tile time.
    order(h->join_ev, side, s);
    hg_mpm_launch_coupled_p2g(h->mpm, dt);
    {
      ProfScope ps(prof, "hyb_grain_mass", s);
      hipLaunchKernelGGL((hyb_grain_mass<T, BASIS>), gb, tb, 0, s, a);
    }
    // the fork is the exchange's own completion when nothing follows it
    // before the fork (no event record between it and the G2P)
    const bool fork_at_exchange = s != side && !h->capturing && !a.hfac;
    {
      // the list holds at most S*S nodes per grain: a grid over the grains
      // walks it in strides
      ProfScope ps(prof, "hyb_exchange", s);
      if (fork_at_exchange)
        hipExtLaunchKernelGGL(hyb_exchange<T>, gb, tb, 0, s, nullptr, h->fork_ev, 0, a);
      else
        hipLaunchKernelGGL(hyb_exchange<T>, gb, tb, 0, s, a);
    }
    homog_factor(a);
    if (fork_at_exchange) hipStreamWaitEvent(side, h->fork_ev, 0);
    else order(h->fork_ev, s, side);
    grain_end(a);
    if (k + 1 < n) dem_step();  // the next step's DEM step, ahead
    hg_mpm_launch_g2p(h->mpm);
    h->par = 1 - h->par;
  }
  // the call ends when both streams have
  order(h->join_ev, side, s);
}

static void hybrid_steps(hg_hybrid *h, double dt, uint32_t n, bool concurrent) {
  const bool f64 = h->mpm->precision == HG_FP64;
  switch (h->mpm->basis) {
    case HG_BASIS_LINEAR:
      f64 ? launch_hybrid_steps<double, HG_BASIS_LINEAR>(h, dt, n, concurrent)
          : launch_hybrid_steps<float, HG_BASIS_LINEAR>(h, dt, n, concurrent);
      break;
    case HG_BASIS_CUBIC:
      f64 ? launch_hybrid_steps<double, HG_BASIS_CUBIC>(h, dt, n, concurrent)
          : launch_hybrid_steps<float, HG_BASIS_CUBIC>(h, dt, n, concurrent);
      break;
    default:
      f64 ? launch_hybrid_steps<double, HG_BASIS_UGIMP_LINEAR>(h, dt, n, concurrent)
          : launch_hybrid_steps<float, HG_BASIS_UGIMP_LINEAR>(h, dt, n, concurrent);
      break;
  }
}

extern "C" hg_status hg_hybrid_destroy(hg_hybrid *h);

extern "C" hg_status hg_hybrid_create(hg_ctx *ctx, const hg_mpm_desc *mpm_desc,
                                      const hg_dem_desc *dem_desc, const double *dem_weights,
                                      hg_hybrid **out) {
  if (!ctx || !mpm_desc || !dem_desc || !out) return hg_fail(ctx, HG_ERR_INVALID, "null argument");
  if (mpm_desc->precision != dem_desc->precision)
    return hg_fail(ctx, HG_ERR_INVALID, "MPM and DEM precision must match");
  std::unique_ptr<hg_hybrid> h(new hg_hybrid());
  h->ctx = ctx;
  HG_TRY(hg_mpm_create(ctx, mpm_desc, &h->mpm));
  hg_status st = hg_dem_create(ctx, dem_desc, &h->dem);
  if (st != HG_OK) {
    hg_mpm_destroy(h->mpm);
    return st;
  }
  // the grains are filed in the DEM's bucket tables by the finish kernel, at
  // their final positions
  hg_dem_caller_files(h->dem);
  // one stream for both: the phases interleave
  HG_HIP(ctx, hipStreamSynchronize(h->dem->stream));
  HG_HIP(ctx, hipStreamDestroy(h->dem->stream));
  h->dem->stream = h->mpm->stream;
  h->dem->own_stream = false;
  h->stream = h->mpm->stream;
  hg_hybrid *hp = h.get();
  if (dem_weights) {
    st = hg_dem_set_weights(h->dem, 0, dem_weights);
    if (st != HG_OK) {
      hg_hybrid_destroy(h.release());
      return st;
    }
  }
  const uint32_t G = h->mpm->nnode;
  const int S = h->mpm->basis == HG_BASIS_LINEAR ? 2 : (h->mpm->basis == HG_BASIS_CUBIC ? 4 : 3);
  const uint64_t ml = std::min<uint64_t>(uint64_t(G), uint64_t(S * S) * std::max(h->dem->nbodies, 1u));
  h->g.max_list = uint32_t(ml);
  const uint32_t B = std::max(h->dem->nbodies, 1u);
  auto fail = [&](hg_status e) {
    hg_hybrid_destroy(h.release());
    return e;
  };
  hg_status e;
  if ((e = halloc(hp, &h->g.dmass, size_t(G) * 8)) != HG_OK) return fail(e);
  if ((e = halloc(hp, &h->g.dmom, size_t(G) * 16)) != HG_OK) return fail(e);
  if ((e = halloc(hp, &h->g.vafter, size_t(G) * 16)) != HG_OK) return fail(e);
  if ((e = halloc(hp, &h->g.aafter, size_t(G) * 16)) != HG_OK) return fail(e);
  if ((e = halloc(hp, &h->g.stamp, size_t(G) * 4)) != HG_OK) return fail(e);
  if ((e = halloc(hp, &h->g.list, size_t(h->g.max_list) * 4)) != HG_OK) return fail(e);
  if ((e = halloc(hp, &h->g.count, 8)) != HG_OK) return fail(e);
  if ((e = halloc(hp, &h->prex, size_t(B) * 8)) != HG_OK) return fail(e);
  if ((e = halloc(hp, &h->prey, size_t(B) * 8)) != HG_OK) return fail(e);
  if ((e = halloc(hp, &h->coupled, size_t(B))) != HG_OK) return fail(e);
  if ((e = halloc(hp, &h->err, 4)) != HG_OK) return fail(e);
  hipMemset(h->g.dmass, 0, size_t(G) * 8);
  hipMemset(h->g.dmom, 0, size_t(G) * 16);
  hipMemset(h->g.vafter, 0, size_t(G) * 16);
  hipMemset(h->g.aafter, 0, size_t(G) * 16);
  hipMemset(h->g.stamp, 0, size_t(G) * 4);
  hipMemset(h->g.count, 0, 8);
  hipMemset(h->coupled, 0, B);
  hipMemset(h->err, 0, 4);
  // the first step's pre-step positions (later steps get them from finish)
  if (h->dem->nbodies) {
    hipMemcpy(h->prex, h->dem->b.x, size_t(h->dem->nbodies) * 8, hipMemcpyDeviceToDevice);
    hipMemcpy(h->prey, h->dem->b.y, size_t(h->dem->nbodies) * 8, hipMemcpyDeviceToDevice);
  }
  if (hipStreamCreateWithFlags(&h->side, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&h->fork_ev, kOrderEventFlags) != hipSuccess ||
      hipEventCreateWithFlags(&h->join_ev, kOrderEventFlags) != hipSuccess ||
      hipEventCreateWithFlags(&h->g2p_ev, kOrderEventFlags) != hipSuccess)
    return fail(hg_fail(ctx, HG_ERR_HIP, "hybrid streams"));
  // the memsets and copies above ran on the null stream (other handles' work
  // is not waited for)
  if (hipStreamSynchronize(nullptr) != hipSuccess) return fail(hg_fail(ctx, HG_ERR_HIP, "hybrid init"));
  *out = h.release();
  return HG_OK;
}

extern "C" hg_status hg_hybrid_destroy(hg_hybrid *h) {
  if (!h) return HG_OK;
  if (h->stream) hipStreamSynchronize(h->stream);
  if (h->side) {
    hipStreamSynchronize(h->side);
    hipStreamDestroy(h->side);
  }
  if (h->fork_ev) hipEventDestroy(h->fork_ev);
  if (h->join_ev) hipEventDestroy(h->join_ev);
  if (h->g2p_ev) hipEventDestroy(h->g2p_ev);
  if (h->graph) hipGraphExecDestroy(h->graph);
  for (void *p : h->allocs) hipFree(p);
  if (h->dem) hg_dem_destroy(h->dem);
  if (h->mpm) hg_mpm_destroy(h->mpm);
  delete h;
  return HG_OK;
}

extern "C" hg_status hg_hybrid_step(hg_hybrid *h, double dt, uint32_t nsteps) {
  if (!h) return HG_ERR_INVALID;
  if (!(dt > 0.0)) return hg_fail(h->ctx, HG_ERR_INVALID, "dt must be > 0");
  if (h->mpm->phase_next != 0) return hg_fail(h->ctx, HG_ERR_INVALID, "a phased MPM step is in progress");
  // the DEM reads its dt from its scene table
  HG_TRY(hg_dem_step(h->dem, dt, 0));
  // profiled steps keep one stream, so that each event pair times one kernel
  // running alone
  hybrid_steps(h, dt, nsteps, !(h->mpm->prof.on || h->dem->prof.on));
  HG_HIP(h->ctx, hipGetLastError());
  return HG_OK;
}

extern "C" hg_status hg_hybrid_step_graph(hg_hybrid *h, double dt, uint32_t nsteps) {
  if (!h) return HG_ERR_INVALID;
  if (!(dt > 0.0)) return hg_fail(h->ctx, HG_ERR_INVALID, "dt must be > 0");
  if (h->mpm->phase_next != 0) return hg_fail(h->ctx, HG_ERR_INVALID, "a phased MPM step is in progress");
  if (h->mpm->prof.on || h->dem->prof.on) return hg_hybrid_step(h, dt, nsteps);
  HG_TRY(hg_dem_step(h->dem, dt, 0));
  if (h->graph && (dt != h->graph_dt || h->mpm->aw != h->graph_aw || h->mpm->ah != h->graph_ah ||
                   h->homog_stress != h->graph_homog_stress ||
                   h->homog_vel != h->graph_homog_vel || h->hfac_on != h->graph_hfac_on)) {
    hipGraphExecDestroy(h->graph);
    h->graph = nullptr;
  }
  // The solvers' buffers and the list counters alternate every step and the
  // DEM's bucket tables rotate by three, so a graph of six steps captured
  // with all of them at phase 0 replays from any later phase 0.  (Stepping
  // the MPM handle alone in between breaks that alignment; the steps then
  // run without the graph.)
  constexpr uint32_t kPeriod = 6;
  auto phase0 = [&]() {
    return h->mpm->cur == 0 && h->par == 0 && h->dem->cur == 0 && h->dem->tph == 0;
  };
  for (uint32_t k = 0; k < kPeriod - 1 && !phase0() && nsteps; ++k) {
    hybrid_steps(h, dt, 1, true);
    --nsteps;
  }
  const bool aligned = phase0();
  const uint32_t reps = aligned ? nsteps / kPeriod : 0;
  if (!aligned) {
    hybrid_steps(h, dt, nsteps, true);
    nsteps = 0;
  }
  if (reps) {
    if (!h->graph) {
      // the MPM dt upload synchronises; do it before capturing
      HG_TRY(hg_mpm_step(h->mpm, dt, 0));
      // each captured step finds the continuum's node sums ready
      hg_mpm_prime(h->mpm);
      hipGraph_t g = nullptr;
      HG_HIP(h->ctx, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
      // the six steps as one call on both streams (the grain branch beside
      // the continuum's G2P, as eager steps run)
      h->capturing = true;
      hybrid_steps(h, dt, kPeriod, true);
      h->capturing = false;
      HG_HIP(h->ctx, hipStreamEndCapture(h->stream, &g));
      HG_HIP(h->ctx, hipGraphInstantiate(&h->graph, g, nullptr, nullptr, 0));
      HG_HIP(h->ctx, hipGraphDestroy(g));
      h->graph_dt = dt;
      h->graph_aw = h->mpm->aw;
      h->graph_ah = h->mpm->ah;
      h->graph_homog_stress = h->homog_stress;
      h->graph_homog_vel = h->homog_vel;
      h->graph_hfac_on = h->hfac_on;
    }
    for (uint32_t k = 0; k < reps; ++k) HG_HIP(h->ctx, hipGraphLaunch(h->graph, h->stream));
  }
  if (nsteps % kPeriod) hybrid_steps(h, dt, nsteps % kPeriod, true);
  HG_HIP(h->ctx, hipGetLastError());
  return HG_OK;
}

extern "C" hg_status hg_hybrid_sync(hg_hybrid *h) {
  if (!h) return HG_ERR_INVALID;
  HG_HIP(h->ctx, hipStreamSynchronize(h->stream));
  int32_t e = 0;
  HG_HIP(h->ctx, hipMemcpy(&e, h->err, 4, hipMemcpyDeviceToHost));
  if (e & 8) return hg_fail(h->ctx, HG_ERR_STATE, "coupling node list overflow");
  HG_TRY(hg_dem_sync(h->dem));
  return hg_mpm_sync(h->mpm);
}

extern "C" hg_status hg_hybrid_mpm(hg_hybrid *h, hg_mpm **out) {
  if (!h || !out) return HG_ERR_INVALID;
  *out = h->mpm;
  return HG_OK;
}

extern "C" hg_status hg_hybrid_dem(hg_hybrid *h, hg_dem **out) {
  if (!h || !out) return HG_ERR_INVALID;
  *out = h->dem;
  return HG_OK;
}

extern "C" hg_status hg_hybrid_download_coupling(hg_hybrid *h, uint8_t *coupled) {
  if (!h) return HG_ERR_INVALID;
  HG_TRY(hg_hybrid_sync(h));
  if (coupled && h->dem->nbodies)
    HG_HIP(h->ctx, hipMemcpy(coupled, h->coupled, h->dem->nbodies, hipMemcpyDeviceToHost));
  return HG_OK;
}

extern "C" hg_status hg_hybrid_set_homogenize(hg_hybrid *h, int32_t stress, int32_t velocity) {
  if (!h) return HG_ERR_INVALID;
  HG_HIP(h->ctx, hipStreamSynchronize(h->stream));
  if ((stress || velocity) && !h->hg.stress) {
    const uint32_t G = h->mpm->nnode;
    const uint32_t B = std::max(h->dem->nbodies, 1u);
    const int S = h->mpm->basis == HG_BASIS_LINEAR ? 2 : (h->mpm->basis == HG_BASIS_CUBIC ? 4 : 3);
    h->hg.max_list = uint32_t(std::min<uint64_t>(uint64_t(G), uint64_t(S * S) * B));
    HG_TRY(halloc(h, &h->hg.stress, size_t(G) * 24));
    HG_TRY(halloc(h, &h->hg.vel, size_t(G) * 16));
    HG_TRY(halloc(h, &h->hg.mass, size_t(G) * 8));
    HG_TRY(halloc(h, &h->hg.mom, size_t(G) * 16));
    HG_TRY(halloc(h, &h->hg.stamp, size_t(G) * 4));
    HG_TRY(halloc(h, &h->hg.list, size_t(h->hg.max_list) * 4));
    HG_TRY(halloc(h, &h->hg.count, 8));
    HG_TRY(halloc(h, &h->dem->cpoint, size_t(B) * size_t(h->dem->cap) * 16));
    HG_HIP(h->ctx, hipMemset(h->hg.stress, 0, size_t(G) * 24));
    HG_HIP(h->ctx, hipMemset(h->hg.vel, 0, size_t(G) * 16));
    HG_HIP(h->ctx, hipMemset(h->hg.mass, 0, size_t(G) * 8));
    HG_HIP(h->ctx, hipMemset(h->hg.mom, 0, size_t(G) * 16));
    HG_HIP(h->ctx, hipMemset(h->hg.stamp, 0, size_t(G) * 4));
    HG_HIP(h->ctx, hipMemset(h->hg.count, 0, 8));
    HG_HIP(h->ctx, hipStreamSynchronize(nullptr));
  }
  h->homog_stress = stress != 0;
  h->homog_vel = velocity != 0;
  return HG_OK;
}

extern "C" hg_status hg_hybrid_clear_homogenized(hg_hybrid *h) {
  if (!h) return HG_ERR_INVALID;
  HG_HIP(h->ctx, hipStreamSynchronize(h->stream));
  if (!h->hg.stress) return HG_OK;
  const uint32_t G = h->mpm->nnode;
  HG_HIP(h->ctx, hipMemset(h->hg.stress, 0, size_t(G) * 24));
  HG_HIP(h->ctx, hipMemset(h->hg.vel, 0, size_t(G) * 16));
  HG_HIP(h->ctx, hipStreamSynchronize(nullptr));
  return HG_OK;
}

extern "C" hg_status hg_hybrid_download_homogenized(hg_hybrid *h, double *stress,
                                                    double *velocity) {
  if (!h) return HG_ERR_INVALID;
  HG_TRY(hg_hybrid_sync(h));
  // one MPM scene: its nodes start at 0 with the padded pitch snx
  const hgmpm::SceneDev &sc = h->mpm->scenes[0];
  const size_t nx = size_t(sc.nnx), ny = size_t(sc.nny), G = h->mpm->nnode;
  std::vector<double> st(h->hg.stress ? 3 * G : 0), ve(h->hg.vel ? 2 * G : 0);
  if (h->hg.stress) {
    HG_HIP(h->ctx, hipMemcpy(st.data(), h->hg.stress, 3 * G * 8, hipMemcpyDeviceToHost));
    HG_HIP(h->ctx, hipMemcpy(ve.data(), h->hg.vel, 2 * G * 8, hipMemcpyDeviceToHost));
  }
  for (size_t iy = 0; iy < ny; ++iy)
    for (size_t ix = 0; ix < nx; ++ix) {
      const size_t r = iy * nx + ix, d = iy * size_t(sc.snx) + ix;
      if (stress) {
        const bool on = !st.empty();
        stress[4 * r] = on ? st[3 * d] : 0.0;
        stress[4 * r + 1] = on ? st[3 * d + 1] : 0.0;
        stress[4 * r + 2] = on ? st[3 * d + 1] : 0.0;
        stress[4 * r + 3] = on ? st[3 * d + 2] : 0.0;
      }
      if (velocity) {
        const bool on = !ve.empty();
        velocity[2 * r] = on ? ve[2 * d] : 0.0;
        velocity[2 * r + 1] = on ? ve[2 * d + 1] : 0.0;
      }
    }
  return HG_OK;
}

// Coupling weights (HybridGrains2DSim.cpp:169-197) recorded by every later
// step while on.  The buffer starts at 1.0 everywhere, as the reference's grid
// does (PhysicsGrid.cpp:52).
extern "C" hg_status hg_hybrid_set_homog_factor(hg_hybrid *h, int32_t on) {
  if (!h) return HG_ERR_INVALID;
  HG_TRY(hg_hybrid_sync(h));
  if (on && !h->hfac) {
    const size_t G = h->mpm->nnode;
    HG_TRY(halloc(h, &h->hfac, G));
    HG_HIP(h->ctx, hipMemset(h->hfac, 2, G));
    HG_HIP(h->ctx, hipStreamSynchronize(nullptr));
  }
  h->hfac_on = on != 0;
  return HG_OK;
}

// The last recorded coupling weights, nny x nnx row-major (0, 0.5 or 1).
extern "C" hg_status hg_hybrid_download_homog_factor(hg_hybrid *h, double *out) {
  if (!h || !out) return HG_ERR_INVALID;
  if (!h->hfac)
    return hg_fail(h->ctx, HG_ERR_INVALID, "coupling weights were never switched on");
  HG_TRY(hg_hybrid_sync(h));
  const hgmpm::SceneDev &sc = h->mpm->scenes[0];
  const size_t nx = size_t(sc.nnx), ny = size_t(sc.nny);
  std::vector<uint8_t> c(h->mpm->nnode);
  HG_HIP(h->ctx, hipMemcpy(c.data(), h->hfac, c.size(), hipMemcpyDeviceToHost));
  for (size_t iy = 0; iy < ny; ++iy)
    for (size_t ix = 0; ix < nx; ++ix) out[iy * nx + ix] = 0.5 * double(c[iy * size_t(sc.snx) + ix]);
  return HG_OK;
}

// Sets the homogenised accumulators (the layout hg_hybrid_download_homogenized
// returns; stress must be symmetric, its (1,0) entry is not kept).
// Homogenisation buffers are allocated if they are not yet.
extern "C" hg_status hg_hybrid_upload_homogenized(hg_hybrid *h, const double *stress,
                                                  const double *velocity) {
  if (!h) return HG_ERR_INVALID;
  HG_TRY(hg_hybrid_sync(h));
  if (!h->hg.stress) {
    const bool s = h->homog_stress, v = h->homog_vel;
    HG_TRY(hg_hybrid_set_homogenize(h, 1, 1));
    h->homog_stress = s;
    h->homog_vel = v;
  }
  const hgmpm::SceneDev &sc = h->mpm->scenes[0];
  const size_t nx = size_t(sc.nnx), ny = size_t(sc.nny), G = h->mpm->nnode;
  std::vector<double> st(3 * G, 0.0), ve(2 * G, 0.0);
  HG_HIP(h->ctx, hipMemcpy(st.data(), h->hg.stress, 3 * G * 8, hipMemcpyDeviceToHost));
  HG_HIP(h->ctx, hipMemcpy(ve.data(), h->hg.vel, 2 * G * 8, hipMemcpyDeviceToHost));
  for (size_t iy = 0; iy < ny; ++iy)
    for (size_t ix = 0; ix < nx; ++ix) {
      const size_t r = iy * nx + ix, d = iy * size_t(sc.snx) + ix;
      if (stress) {
        st[3 * d] = stress[4 * r];
        st[3 * d + 1] = stress[4 * r + 2];
        st[3 * d + 2] = stress[4 * r + 3];
      }
      if (velocity) {
        ve[2 * d] = velocity[2 * r];
        ve[2 * d + 1] = velocity[2 * r + 1];
      }
    }
  HG_HIP(h->ctx, hipMemcpy(h->hg.stress, st.data(), 3 * G * 8, hipMemcpyHostToDevice));
  HG_HIP(h->ctx, hipMemcpy(h->hg.vel, ve.data(), 2 * G * 8, hipMemcpyHostToDevice));
  return HG_OK;
}

extern "C" hg_status hg_hybrid_estimate_bytes(const hg_mpm_desc *mpm, const hg_dem_desc *dem,
                                              uint64_t *bytes) {
  if (!mpm || !dem || !bytes) return HG_ERR_INVALID;
  uint64_t a = 0, b = 0;
  HG_TRY(hg_mpm_estimate_bytes(mpm, 1, &a));
  HG_TRY(hg_dem_estimate_bytes(dem, 1, &b));
  const uint64_t G = (uint64_t(mpm->cell_count[0]) + 2) * (uint64_t(mpm->cell_count[1]) + 2);
  const uint64_t B = std::max<uint64_t>(dem->nbodies, 1);
  *bytes = a + b + G * (8 + 16 * 3 + 4 + 4) + B * (8 + 8 + 1) + 64;
  return HG_OK;
}

extern "C" hg_status hg_hybrid_device_bytes(hg_hybrid *h, uint64_t *bytes) {
  if (!h || !bytes) return HG_ERR_INVALID;
  *bytes = h->bytes + h->mpm->bytes + h->dem->bytes;
  return HG_OK;
}
