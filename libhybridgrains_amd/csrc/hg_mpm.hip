// hg_mpm.hip -- MPM explicit step on gfx950 and its C-ABI.
//
// Replaces mpmgrains2d/ExplicitIntegrator.cpp:5-67 and the PhysicsGrid /
// MaterialPoints / ConstitutiveModel routines it calls.  See
// hg_mpm_kernels.h for the HBM layout.
#include "hg_mpm_handle.h"
#include "hg_sort.h"

#include <hip/hip_ext.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <numeric>
#include <type_traits>

using namespace hgmpm;

namespace {

constexpr int kG2PThreads = 256;

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------

template <typename T>
__device__ __forceinline__ T ld(const void *p, uint32_t i) {
  return static_cast<const T *>(p)[i];
}
template <typename T>
__device__ __forceinline__ void st(void *p, uint32_t i, T v) {
  static_cast<T *>(p)[i] = v;
}

// a * b + c: one fused operation in FP32 mode; in FP64 a rounded product and
// a rounded sum, the operations the CPU FP64 restatement performs.
__device__ __forceinline__ float madd(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double madd(double a, double b, double c) { return a * b + c; }

__device__ __forceinline__ uint32_t scene_of_point(const uint32_t *p_off, uint32_t nscenes,
                                                   uint32_t i) {
  uint32_t lo = 0, hi = nscenes;  // p_off[lo] <= i < p_off[hi]
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (p_off[mid] <= i) lo = mid; else hi = mid;
  }
  return lo;
}

// Computes the stencil base key of a point and whether its whole stencil is
// on the grid (SimulationState::isConsistent's bound check, SimulationState.cpp:54-62).
template <int BASIS>
__device__ __forceinline__ bool point_key(const SceneDev &sc, double x, double y, double hl,
                                          int &bx, int &by, int &cx, int &cy) {
  double lx, hx, ly, hy;
  stencil_1d<BASIS>(x, hl, sc.min0, sc.h, sc.inv_h, lx, hx);
  stencil_1d<BASIS>(y, hl, sc.min1, sc.h, sc.inv_h, ly, hy);
  const bool ok = (lx >= 0.0) && (ly >= 0.0) && (hx <= double(sc.nnx - 1)) &&
                  (hy <= double(sc.nny - 1));
  constexpr int S = BasisTraits<BASIS>::S;
  if (ok) {
    bx = int(lx);
    by = int(ly);
    cx = int(hx) - bx + 1;
    cy = int(hy) - by + 1;
  } else {
    // keep memory safe; the scene is flagged and its results are void
    bx = (lx >= 0.0) ? int(fmin(lx, double(sc.nnx - S))) : 0;
    by = (ly >= 0.0) ? int(fmin(ly, double(sc.nny - S))) : 0;
    bx = max(0, min(bx, sc.nnx - S));
    by = max(0, min(by, sc.nny - S));
    cx = S;
    cy = S;
  }
  return ok;
}

// ---------------------------------------------------------------------------
// P2G + grid update.  A workgroup takes tiles of TW x TH owned nodes from the
// list of active tiles, one thread per node.  The points whose stencils
// reach the tile have their base cell in the apron [x0-S+1, x0+TW) x
// [y0-S+1, y0+TH); in the cell-sorted order they form one run per apron row,
// and P2G stages each cell's points in point-id order.
//  1. stage (point-parallel): the apron points are loaded through the
//     permutation; their Cauchy stress (FP64) and their 1-D stencil weights
//     and weight derivatives for the S nodes of each axis go to LDS, each at
//     its point-id rank within its cell (the sort leaves a cell's points in
//     the order its atomics handed out slots).  The apron's points are
//     concatenated row by row and staged in chunks of at most CAP points
//     that end on cell boundaries.
//  2. gather (node-parallel): for each of the node's S x S cells -- rows
//     ascending, cells ascending -- the node walks the cell's points in id
//     order and sums mass, momentum and force (FP64 mode: into FP64 sums;
//     FP32 mode: a FP32 sum per cell, added to FP64 node sums).  The sum
//     order does not depend on the chunking or the tile shape, so each
//     node's value is a fixed function of the scene state (bitwise
//     reproducible, identical alone or in a batch, for any tile shape).
//  3. grid update of the node (forces, momentum, plane conditions, velocity
//     and acceleration), written as one (v, a) vector per node.
//
// Follows PhysicsGrid::rasterizePointMasses (PhysicsGrid.cpp:168-207),
// rasterizePointMomentum (:274-311), MaterialPoints::computeHyperelasticCauchyStress
// (MaterialPoints.cpp:183-202) with ConstitutiveModel::computeTau
// (ConstitutiveModel.cpp:83-97), PhysicsGrid::computeForces (:413-473),
// updateMomentum (:565-586), resolvePlaneCollisions (:611-673) and
// lumpedMassVelocityAndAccelerationUpdate (:588-609).
// ---------------------------------------------------------------------------
struct P2GArgs {
  const SceneDev *scenes;
  const uint32_t *tile_scene;
  const uint32_t *tile_list;
  const uint32_t *tile_count;
  PtsRaw pts;
  const uint32_t *perm;
  const uint32_t *cell_start;
  void *gva;
  double *gm, *gp, *gpn, *gf;
  uint32_t store;
};

// The grid update of one node from its sums (mass m, momentum p, internal
// force f): gravity, PhysicsGrid::updateMomentum (PhysicsGrid.cpp:565-586),
// resolvePlaneCollisions (:611-673) and
// lumpedMassVelocityAndAccelerationUpdate (:588-609), writing what store
// asks for.
template <typename T>
__device__ __forceinline__ void grid_node_momentum(const SceneDev &sc, int nx, int ny, double m,
                                                   double p0, double p1, double &f0, double &f1,
                                                   double &pn0, double &pn1) {
  const double dt = sc.dt;
  f0 += m * sc.g0;
  f1 += m * sc.g1;
  pn0 = 0.0;
  pn1 = 0.0;
  if (m > 0.0) {
    pn0 = p0 + dt * f0;
    pn1 = p1 + dt * f1;
  }
  const double px0 = sc.min0 + sc.h * double(nx);
  const double px1 = sc.min1 + sc.h * double(ny);
  for (int k = 0; k < sc.nplanes; ++k) {
    if (px1 < sc.plane_lower[k]) continue;
    const double n0 = sc.plane_n[2 * k], n1 = sc.plane_n[2 * k + 1];
    const double d = n0 * (px0 - sc.plane_x[2 * k]) + n1 * (px1 - sc.plane_x[2 * k + 1]);
    if (d > 0.0) continue;
    if (sc.plane_beh[k] == HG_PLANE_SLIDING) {
      if (m <= 0.0) continue;
      double u0 = pn0 / m, u1 = pn1 / m;
      const double vn = n0 * u0 + n1 * u1;
      if (vn > 0.0) continue;
      u0 -= vn * n0;
      u1 -= vn * n1;
      pn0 = m * u0;
      pn1 = m * u1;
    } else {
      pn0 = 0.0;
      pn1 = 0.0;
    }
  }
}

// The node's stored fields from its mass m, momentum p, new momentum pn and
// force f: the (v, a) vector and what store asks for.
template <typename T>
__device__ __forceinline__ void grid_node_store(const SceneDev &sc, uint32_t node, double m,
                                                double p0, double p1, double pn0, double pn1,
                                                double f0, double f1, uint32_t store, void *gva,
                                                double *gm, double *gp, double *gpn, double *gf) {
  using T4 = typename Vec4T<T>::type;
  if (store & kStoreVelAcc) {
    T4 va;
    if (m > 0.0) {
      const double dm = sc.dt * m;
      va.x = T(pn0 / m);
      va.y = T(pn1 / m);
      va.z = T((pn0 - p0) / dm);
      va.w = T((pn1 - p1) / dm);
    } else {
      va.x = va.y = va.z = va.w = T(0);
    }
    static_cast<T4 *>(gva)[node] = va;
  }
  if (store & kStoreMass) gm[node] = m;
  if (store & kStoreMom) {
    gp[2 * node] = p0;
    gp[2 * node + 1] = p1;
  }
  if (store & kStoreMomNew) {
    gpn[2 * node] = pn0;
    gpn[2 * node + 1] = pn1;
  }
  if (store & kStoreForce) {
    gf[2 * node] = f0;
    gf[2 * node + 1] = f1;
  }
}

template <typename T>
__device__ __forceinline__ void grid_node_update(const SceneDev &sc, int nx, int ny, double m,
                                                 double p0, double p1, double f0, double f1,
                                                 uint32_t store, void *gva, double *gm,
                                                 double *gp, double *gpn, double *gf) {
  const uint32_t node = sc.node_off + uint32_t(ny) * uint32_t(sc.snx) + uint32_t(nx);
  double pn0, pn1;
  grid_node_momentum<T>(sc, nx, ny, m, p0, p1, f0, f1, pn0, pn1);
  grid_node_store<T>(sc, node, m, p0, p1, pn0, pn1, f0, f1, store, gva, gm, gp, gpn, gf);
}

// What P2G needs of one point besides its weights.
template <typename T>
struct P2GRec {
  double x, y;
  T m, mv0, mv1, c, s00, s01, s11, hl;
  uint32_t cnt;  // stencil node counts: x in bits 0-3, y in bits 4-7
};

// Stencil node counts of a point whose stencil base node (bx, by) is its
// cell key (as the sort binned it): only uGIMP's count varies, and its last
// node is ceil((x + hl - min) / h) (uGIMPLinearBasisFunctions.cpp:68-90).
template <int BASIS>
__device__ __forceinline__ uint32_t stencil_counts(double x, double y, double hl, int bx, int by,
                                                   const SceneDev &sc) {
  constexpr int S = BasisTraits<BASIS>::S;
  if constexpr (BASIS == HG_BASIS_UGIMP_LINEAR) {
    const double cxd = ceil(div_rn((x + hl) - sc.min0, sc.h, sc.inv_h)) - double(bx) + 1.0;
    const int cntx = (cxd >= 1.0 && cxd <= double(S)) ? int(cxd) : S;
    const double cyd = ceil(div_rn((y + hl) - sc.min1, sc.h, sc.inv_h)) - double(by) + 1.0;
    const int cnty = (cyd >= 1.0 && cyd <= double(S)) ? int(cyd) : S;
    return uint32_t(cntx) | (uint32_t(cnty) << 4);
  } else {
    return uint32_t(S) | (uint32_t(S) << 4);
  }
}

// A point's fields as P2G loads them.
template <typename T>
struct P2GRaw {
  double x, y;
  T m, v0, v1, V, J, b00, b01, b11, hl;
  uint32_t id;
};

template <typename T>
__device__ __forceinline__ P2GRaw<T> p2g_load(const PtsRaw &pts, uint32_t p, const SceneDev &sc) {
  P2GRaw<T> r;
  r.id = pts.id[p];
  r.x = pts.x[p];
  r.y = pts.y[p];
  // mass, volume and half width are indexed by point id (never moved)
  r.m = pts.m ? ld<T>(pts.m, r.id) : T(sc.m_uniform);
  r.v0 = ld<T>(pts.vx, p);
  r.v1 = ld<T>(pts.vy, p);
  r.V = pts.V ? ld<T>(pts.V, r.id) : T(sc.V_uniform);
  r.J = ld<T>(pts.J, p);
  r.b00 = ld<T>(pts.b00, p);
  r.b01 = ld<T>(pts.b01, p);
  r.b11 = ld<T>(pts.b11, p);
  r.hl = pts.hl ? ld<T>(pts.hl, r.id) : T(sc.hl_uniform);
  return r;
}

template <typename T, int BASIS>
__device__ __forceinline__ P2GRec<T> p2g_record(const P2GRaw<T> &w, int bx, int by,
                                                const SceneDev &sc) {
  P2GRec<T> r;
  r.x = w.x;
  r.y = w.y;
  r.m = w.m;
  r.hl = w.hl;
  const double J = strain_load<T>(w.J);
  const double b00 = strain_load<T>(w.b00), b01 = double(w.b01), b11 = strain_load<T>(w.b11);
  // Cauchy stress sigma = tau / J; tau = 0 when J > 1 (no tension).  FP64.
  double t00 = 0.0, t01 = 0.0, t11 = 0.0;
  if (J <= 1.0) {
    const double e00 = b00 * J, e01 = b01 * J, e11 = b11 * J;
    const double det = e00 * e11 - e01 * e01;
    const double Jt = sqrt(det);
    const double inrm = 1.0 / Jt;
    const double bb00 = e00 * inrm, bb01 = e01 * inrm, bb11 = e11 * inrm;
    const double half_tr = (bb00 + bb11) * 0.5;
    const double c = sc.kappa * 0.5 * (Jt + 1.0) * (Jt - 1.0);
    t00 = c + sc.mu * (bb00 - half_tr);
    t01 = sc.mu * bb01;
    t11 = c + sc.mu * (bb11 - half_tr);
  }
  const double iJ = 1.0 / J;
  r.s00 = T(t00 * iJ);
  r.s01 = T(t01 * iJ);
  r.s11 = T(t11 * iJ);
  r.c = T(-double(w.V) * J);
  r.mv0 = r.m * w.v0;
  r.mv1 = r.m * w.v1;
  r.cnt = stencil_counts<BASIS>(r.x, r.y, double(r.hl), bx, by, sc);
  return r;
}

// Apron point indices kept in LDS per tile (larger aprons read them from HBM
// chunk by chunk).
constexpr int kP2GPermCap = 2048;

// Largest power of two <= n (n >= 1; 0 for n < 1).
constexpr int top_bit(int n) {
  int b = 1;
  while (b * 2 <= n) b *= 2;
  return n >= 1 ? b : 0;
}

template <typename T, int BASIS, int TW, int TH, int CAP>
__global__ __launch_bounds__(TW *TH) void mpm_p2g_kernel(P2GArgs args) {
  constexpr int S = BasisTraits<BASIS>::S;
  constexpr int AW = TW + S - 1;  // apron cells per row
  constexpr int AH = TH + S - 1;  // apron rows
  constexpr int NT = TW * TH;
  // per staged point: x weights and x derivative factors of its S x nodes,
  // the same for y, and the mass, momentum and stress terms
  __shared__ T s_wx[S][CAP], s_gx[S][CAP], s_wy[S][CAP], s_gy[S][CAP];
  __shared__ T s_m[CAP], s_mv0[CAP], s_mv1[CAP], s_f00[CAP], s_f01[CAP], s_f11[CAP];
  __shared__ uint32_t s_id[CAP];               // point ids, in sort order
  __shared__ uint32_t s_perm[kP2GPermCap];     // the apron's point indices, when they fit
  __shared__ uint32_t s_cs[AH][AW + 1];        // cell starts of the apron rows
  __shared__ uint32_t s_ro[AH + 1];            // row offsets in the concatenated list

  const uint32_t ntile_list = *args.tile_count;
  const int t = threadIdx.x;
  const int lx = t % TW, ly = t / TW;
  // Workgroup b runs on XCD b % 8 (the grid is a multiple of 8).  Each XCD
  // takes one contiguous eighth of the list, whose tiles are mostly
  // neighbours in the grid, so the apron points two tiles share are fetched
  // from HBM into that XCD's L2 once.
  const uint32_t per_xcd = (ntile_list + 7u) >> 3, xcd = blockIdx.x & 7u;
  for (uint32_t j = blockIdx.x >> 3; j < per_xcd; j += gridDim.x >> 3) {
    const uint32_t li = xcd * per_xcd + j;
    if (li >= ntile_list) break;
    const uint32_t tile = args.tile_list[li];
    const uint32_t sid = args.tile_scene[tile];
    const SceneDev &sc = args.scenes[sid];
    const uint32_t ltile = tile - sc.tile_off;
    const int x0 = int(ltile % uint32_t(sc.tiles_x)) * TW;
    const int y0 = int(ltile / uint32_t(sc.tiles_x)) * TH;
    const uint32_t *cs = args.cell_start + sc.cell_off;
    const int ax0 = x0 - (S - 1), ay0 = y0 - (S - 1);

    // Apron cell starts, row by row (cells of one row are consecutive keys;
    // cells off the grid hold no points and read as empty).
    for (int q = t; q < AH * (AW + 1); q += NT) {
      const int r = q / (AW + 1), c = q % (AW + 1);
      const int ry = ay0 + r;
      uint32_t v = 0;
      if (ry >= 0 && ry < sc.nny) {
        const int cx = min(max(ax0 + c, 0), sc.nnx);
        v = cs[uint32_t(ry) * uint32_t(sc.snx) + uint32_t(cx)];
      }
      s_cs[r][c] = v;
    }
    __syncthreads();
    if (t == 0) {
      uint32_t o = 0;
      for (int r = 0; r < AH; ++r) {
        s_ro[r] = o;
        o += s_cs[r][AW] - s_cs[r][0];
      }
      s_ro[AH] = o;
    }
    __syncthreads();
    const uint32_t total = s_ro[AH];

    const int nx = x0 + lx, ny = y0 + ly;
    double am = 0.0, ap0 = 0.0, ap1 = 0.0, af0 = 0.0, af1 = 0.0;

    // position k of the concatenated list -> apron row r and column col
    // (binary searches over the row offsets and the row's cell starts)
    auto locate = [&](uint32_t k, int &r, int &col) {
      r = 0;
#pragma unroll
      for (int step = top_bit(AH - 1); step > 0; step >>= 1)
        if (r + step <= AH - 1 && s_ro[r + step] <= k) r += step;
      const uint32_t pos = s_cs[r][0] + (k - s_ro[r]);
      col = 0;
#pragma unroll
      for (int step = top_bit(AW - 1); step > 0; step >>= 1)
        if (col + step <= AW - 1 && s_cs[r][col + step] <= pos) col += step;
    };

    // the apron's point indices, in list order, when they fit in LDS: every
    // chunk then needs one round trip to HBM instead of two
    const bool perm_lds = total <= uint32_t(kP2GPermCap);
    if (perm_lds) {
      for (uint32_t k = uint32_t(t); k < total; k += NT) {
        int r, col;
        locate(k, r, col);
        s_perm[k] = args.perm[s_cs[r][0] + (k - s_ro[r])];
      }
      __syncthreads();
    }

    // end of the chunk that starts at k0 (every thread computes it): at most
    // CAP points, ending on a cell boundary unless one cell fills the chunk
    auto chunk_end = [&](uint32_t k0) {
      const uint32_t target = min(k0 + uint32_t(CAP), total);
      if (target >= total) return total;
      int r, col;
      locate(target, r, col);
      const uint32_t cstart = s_ro[r] + (s_cs[r][col] - s_cs[r][0]);
      return cstart > k0 ? cstart : target;
    };
    constexpr int PER = (CAP + NT - 1) / NT;
    // a chunk's points as loaded (through the permutation) for one thread
    auto load_chunk = [&](uint32_t k0, uint32_t k1, P2GRaw<T> (&raw)[PER], int (&rrow)[PER],
                          int (&rcol)[PER]) {
#pragma unroll
      for (int m = 0; m < PER; ++m) {
        const uint32_t k = k0 + uint32_t(t) + uint32_t(m * NT);
        if (k < k1) {
          locate(k, rrow[m], rcol[m]);
          const uint32_t p =
              perm_lds ? s_perm[k] : args.perm[s_cs[rrow[m]][0] + (k - s_ro[rrow[m]])];
          raw[m] = p2g_load<T>(args.pts, p, sc);
        }
      }
    };

    // Chunks are consecutive pieces of the concatenated list that end on cell
    // boundaries (a cell of more than CAP points is split).  Within a cell the
    // points are staged in id order; each node walks its runs in ascending
    // list order within every chunk, so its sum order is the list order
    // (rows, cells, ids) whatever the chunk boundaries.  The next chunk's
    // points are loaded while the current one is gathered.
    P2GRaw<T> raw[PER];
    int rrow[PER], rcol[PER];
    uint32_t k0 = 0, k1 = chunk_end(0);
    if (k0 < total) load_chunk(k0, k1, raw, rrow, rcol);
    while (k0 < total) {
      // 1a. publish the chunk's ids
#pragma unroll
      for (int m = 0; m < PER; ++m) {
        const uint32_t k = k0 + uint32_t(t) + uint32_t(m * NT);
        if (k < k1) s_id[k - k0] = raw[m].id;
      }
      __syncthreads();

      // 1b. stage each point at its id rank within its cell
#pragma unroll
      for (int m = 0; m < PER; ++m) {
        const uint32_t k = k0 + uint32_t(t) + uint32_t(m * NT);
        if (k < k1) {
          const int r = rrow[m], col = rcol[m];
          const uint32_t cb0 = s_ro[r] + (s_cs[r][col] - s_cs[r][0]);
          const uint32_t ce0 = s_ro[r] + (s_cs[r][col + 1] - s_cs[r][0]);
          const uint32_t cb = max(cb0, k0) - k0, ce = min(ce0, k1) - k0;
          uint32_t rank = 0;
          for (uint32_t v = cb; v < ce; ++v) rank += (s_id[v] < raw[m].id) ? 1u : 0u;
          const uint32_t d = cb + rank;
          const int bx = ax0 + col, by = ay0 + r;
          const P2GRec<T> rec = p2g_record<T, BASIS>(raw[m], bx, by, sc);
          const T inv_h = T(sc.inv_h);
          const int cntx = int(rec.cnt & 15u), cnty = int(rec.cnt >> 4);
          T inv_vp = T(1);
          if constexpr (BASIS == HG_BASIS_UGIMP_LINEAR) inv_vp = T(1) / (T(4) * rec.hl * rec.hl);
          // A node's force term counts only where its weight wx*wy is nonzero
          // (PhysicsGrid.cpp:413-473 skips zero-weight nodes): an axis's
          // derivative factor is staged as zero where that axis's weight is
          // zero, so both gradient components gxs*wy and wx*gys vanish
          // exactly when one factor of the weight does.
#pragma unroll
          for (int k2 = 0; k2 < S; ++k2) {
            T ax, dx, ay, dy;
            basis_factor<T, BASIS>(rec.x, rec.hl, sc.min0, sc.h, inv_h, bx + k2, k2, cntx, ax,
                                   dx);
            basis_factor<T, BASIS>(rec.y, rec.hl, sc.min1, sc.h, inv_h, by + k2, k2, cnty, ay,
                                   dy);
            T wx, gx, gy;
            if constexpr (BASIS == HG_BASIS_UGIMP_LINEAR) {
              wx = ax * inv_vp;
              gx = dx * inv_vp;
              gy = dy;
            } else {
              wx = ax;
              gx = dx * inv_h;
              gy = dy * inv_h;
            }
            s_wx[k2][d] = wx;
            s_gx[k2][d] = (wx != T(0)) ? gx : T(0);
            s_wy[k2][d] = ay;
            s_gy[k2][d] = (ay != T(0)) ? gy : T(0);
          }
          s_m[d] = rec.m;
          s_mv0[d] = rec.mv0;
          s_mv1[d] = rec.mv1;
          // -V J sigma: the point's force per unit weight gradient
          s_f00[d] = rec.c * rec.s00;
          s_f01[d] = rec.c * rec.s01;
          s_f11[d] = rec.c * rec.s11;
        }
      }
      __syncthreads();

      // the next chunk's loads run under this chunk's gather
      const uint32_t nk0 = k1, nk1 = nk0 < total ? chunk_end(nk0) : total;
      if (nk0 < total) load_chunk(nk0, nk1, raw, rrow, rcol);

      // 2. gather: rows ascending, cells ascending, ids ascending.  The cell
      //    fixes the node's offset in the point's stencil, so every LDS read of
      //    an iteration is independent of the others.  The staged derivative
      //    factors already vanish where the weight does.  FP64 mode sums every
      //    term into the FP64 node sums; FP32 mode sums each cell's terms in
      //    FP32 (fused multiply-adds) and adds that partial to the FP64 node
      //    sums (chunks end on cell boundaries, so the partials do not depend
      //    on the chunking unless one cell holds more than a chunk).
      if (nx < sc.nnx && ny < sc.nny) {
        constexpr bool f64 = std::is_same<T, double>::value;
#pragma unroll
        for (int j0 = 0; j0 < S; ++j0) {
          const int j = S - 1 - j0;  // cell row ny - j: weight index j
          const int r = ly + j0;
          const uint32_t rb = s_ro[r] - s_cs[r][0];  // list position of the row's cell starts
#pragma unroll
          for (int i0 = 0; i0 < S; ++i0) {
            const int i = S - 1 - i0;  // cell column nx - i: weight index i
            const uint32_t lo = max(rb + s_cs[r][lx + i0], k0);
            const uint32_t hi = min(rb + s_cs[r][lx + i0 + 1], k1);
            T cm = 0, cp0 = 0, cp1 = 0, cf0 = 0, cf1 = 0;
            if constexpr (f64) {
              cm = am;
              cp0 = ap0;
              cp1 = ap1;
              cf0 = af0;
              cf1 = af1;
            }
            for (uint32_t k = lo; k < hi; ++k) {
              const uint32_t q = k - k0;
              const T wx = s_wx[i][q], gxs = s_gx[i][q];
              const T wy = s_wy[j][q], gys = s_gy[j][q];
              const T pm = s_m[q], pv0 = s_mv0[q], pv1 = s_mv1[q];
              const T f00 = s_f00[q], f01 = s_f01[q], f11 = s_f11[q];
              const T w = wx * wy;
              const T gx = gxs * wy;
              const T gy = wx * gys;
              cm = madd(w, pm, cm);
              cp0 = madd(w, pv0, cp0);
              cp1 = madd(w, pv1, cp1);
              if constexpr (f64) {
                cf0 += f00 * gx + f01 * gy;
                cf1 += f01 * gx + f11 * gy;
              } else {
                cf0 = madd(f01, gy, madd(f00, gx, cf0));
                cf1 = madd(f11, gy, madd(f01, gx, cf1));
              }
            }
            if constexpr (f64) {
              am = cm;
              ap0 = cp0;
              ap1 = cp1;
              af0 = cf0;
              af1 = cf1;
            } else if (hi > lo) {
              am += double(cm);
              ap0 += double(cp0);
              ap1 += double(cp1);
              af0 += double(cf0);
              af1 += double(cf1);
            }
          }
        }
      }
      __syncthreads();
      k0 = nk0;
      k1 = nk1;
    }

    // 3. grid update of the node.
    if (nx < sc.nnx && ny < sc.nny)
      grid_node_update<T>(sc, nx, ny, am, ap0, ap1, af0, af1, args.store, args.gva, args.gm,
                          args.gp, args.gpn, args.gf);
    __syncthreads();  // LDS is reused by the next tile
  }
}

// Active tiles: a tile is active when any cell of its apron holds a point.
// Writes the per-tile flags and appends active tiles to the list (in
// ascending order within a workgroup; workgroups append in any order, which
// P2G does not depend on).  The count alternates between two counters with
// the point buffers: this list's counter was cleared by the previous list.
struct TileListArgs {
  const SceneDev *scenes;
  const uint32_t *tile_scene;
  const uint32_t *cell_start;
  uint8_t *tile_active;
  uint32_t *tile_list, *tile_count, *next_count;
  uint32_t ntiles;
  int32_t tw, th, S;
  // the sort's scatter, run by the workgroups past the tile list's
  const uint32_t *key, *slot;
  uint32_t *perm, *pinv;
  uint32_t npoints;
};

// Workgroups [0, ceil(ntiles / 256)) list the active tiles, one lane per
// tile: the lane checks the tile's apron rows (a row's cells are consecutive
// keys, so two cell starts give its point count), and the workgroup appends
// its active tiles in ascending order with one atomic.  The workgroups after
// them scatter the sorted points into perm and pinv, 256 points each (both
// parts read only the scan's cell_start, so one launch does both).
__global__ __launch_bounds__(256) void mpm_tile_list_kernel(TileListArgs a) {
  __shared__ uint32_t s_wave[4];
  __shared__ uint32_t s_base;
  const uint32_t ntb = (a.ntiles + 255u) / 256u;
  if (blockIdx.x >= ntb) {
    const uint32_t i = (blockIdx.x - ntb) * 256u + threadIdx.x;
    if (i < a.npoints) hg_scatter_item(a.key, a.slot, a.cell_start, a.npoints, a.perm, a.pinv, i);
    return;
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t t = blockIdx.x * 256u + threadIdx.x;
  bool active = false;
  if (t < a.ntiles) {
    const SceneDev &sc = a.scenes[a.tile_scene[t]];
    const uint32_t lt = t - sc.tile_off;
    const int x0 = int(lt % uint32_t(sc.tiles_x)) * a.tw;
    const int y0 = int(lt / uint32_t(sc.tiles_x)) * a.th;
    const int ax0 = x0 - (a.S - 1), ay0 = y0 - (a.S - 1);
    const int aw = a.tw + a.S - 1, ah = a.th + a.S - 1;
    const uint32_t *cs = a.cell_start + sc.cell_off;
    const uint32_t c0 = uint32_t(min(max(ax0, 0), sc.nnx));
    const uint32_t c1 = uint32_t(min(max(ax0 + aw, 0), sc.nnx));
    const int r0 = max(ay0, 0), r1 = min(ay0 + ah, sc.nny);
    // every row's loads issue together (no early exit)
    uint32_t nonempty = 0;
#pragma unroll 8
    for (int ry = r0; ry < r1; ++ry) {
      const uint32_t row = uint32_t(ry) * uint32_t(sc.snx);
      nonempty |= uint32_t(cs[row + c1] > cs[row + c0]);
    }
    active = nonempty != 0u;
    a.tile_active[t] = active ? 1 : 0;
  }
  // the other parity's counter, for the next list (nothing reads it now)
  if (blockIdx.x == 0 && threadIdx.x == 0) *a.next_count = 0u;
  const uint64_t bal = __ballot(active);
  if (lane == 0) s_wave[wave] = uint32_t(__popcll(bal));
  __syncthreads();
  if (threadIdx.x == 0) {
    const uint32_t tot = s_wave[0] + s_wave[1] + s_wave[2] + s_wave[3];
    s_base = tot ? atomicAdd(a.tile_count, tot) : 0u;
  }
  __syncthreads();
  if (active) {
    uint32_t o = s_base + uint32_t(__popcll(bal & ((1ull << lane) - 1ull)));
    for (int w = 0; w < wave; ++w) o += s_wave[w];
    a.tile_list[o] = t;
  }
}

// Node velocity and acceleration from stored mass and momenta, for the split
// (phase-by-phase) step: PhysicsGrid::lumpedMassVelocityAndAccelerationUpdate
// (PhysicsGrid.cpp:588-609).
struct VelAccArgs {
  const SceneDev *scenes;
  const uint32_t *tile_scene;
  const uint32_t *tile_list, *tile_count;
  const double *gm, *gp, *gpn;
  void *gva;
  int32_t tw, th;
};

// pinv as the inverse of perm (what the scatter writes when asked for it).
__global__ __launch_bounds__(256) void pinv_from_perm_kernel(const uint32_t *perm, uint32_t n,
                                                             uint32_t *pinv) {
  const uint32_t j = blockIdx.x * 256u + threadIdx.x;
  if (j < n) {
    const uint32_t i = perm[j];
    if (i < n) pinv[i] = j;
  }
}

template <typename T>
__global__ __launch_bounds__(256) void mpm_velacc_kernel(VelAccArgs a) {
  const uint32_t n = *a.tile_count;
  using T4 = typename Vec4T<T>::type;
  for (uint32_t li = blockIdx.x; li < n; li += gridDim.x) {
    const uint32_t tile = a.tile_list[li];
    const SceneDev &sc = a.scenes[a.tile_scene[tile]];
    const uint32_t lt = tile - sc.tile_off;
    const int x0 = int(lt % uint32_t(sc.tiles_x)) * a.tw, y0 = int(lt / uint32_t(sc.tiles_x)) * a.th;
    for (int q = threadIdx.x; q < a.tw * a.th; q += 256) {
      const int nx = x0 + q % a.tw, ny = y0 + q / a.tw;
      if (nx >= sc.nnx || ny >= sc.nny) continue;
      const uint32_t node = sc.node_off + uint32_t(ny) * uint32_t(sc.snx) + uint32_t(nx);
      const double m = a.gm[node];
      T4 va;
      if (m > 0.0) {
        const double dm = sc.dt * m;
        va.x = T(a.gpn[2 * node] / m);
        va.y = T(a.gpn[2 * node + 1] / m);
        va.z = T((a.gpn[2 * node] - a.gp[2 * node]) / dm);
        va.w = T((a.gpn[2 * node + 1] - a.gp[2 * node + 1]) / dm);
      } else {
        va.x = va.y = va.z = va.w = T(0);
      }
      static_cast<T4 *>(a.gva)[node] = va;
    }
  }
}

// ---------------------------------------------------------------------------
// G2P: one thread per point, in the buffer's physical order (which is the
// previous sort's order); each point is written to its slot of the current
// cell order in the other buffer (pinv), so the buffers stay cell-sorted.  MaterialPoints::computeVelGrad
// (MaterialPoints.cpp:296-322), elasticPrediction (:324-348) with
// ConstitutiveModel::elasticPrediction (ConstitutiveModel.cpp:19-22),
// plasticCorrection (:350-369) with the Drucker-Prager return
// (ConstitutiveModel.cpp:26-81), updateVelocities (:204-245),
// updatePositions (:247-288) and resolvePlaneCollisions (:371-423).  The
// return map runs in FP64 in every precision.  The point is written to slot i
// of the other buffer, and its new cell key is counted for the next sort.
// ---------------------------------------------------------------------------
struct G2PArgs {
  const SceneDev *scenes;
  const uint32_t *p_off;
  uint32_t nscenes;
  uint32_t npoints;
  uint32_t nblocks_padded;
  PtsRaw src, dst;
  const uint32_t *pinv;
  const void *gva;
  uint32_t *key, *slot, *counts, *totals;
  int32_t *err;
};

// FP64 reciprocal square root, square root and reciprocal for the FP32 mode's
// strain update, from the hardware estimates (about 23 correct bits) and two
// Newton-Raphson steps each, which take them to within a few ulps of the
// correctly rounded results: far below what the FP32 strain storage keeps,
// and without the range-reduction and rounding-fixup sequences of the
// correctly rounded forms.  Inputs are positive and normal in every use;
// anything else yields a non-finite result that the point checks flag.
__device__ __forceinline__ double rsq_nr(double a) {
  double y = __builtin_amdgcn_rsq(a);
  y = fma(0.5 * y, fma(-(a * y), y, 1.0), y);
  return fma(0.5 * y, fma(-(a * y), y, 1.0), y);
}
// sqrt(a) given y = rsq_nr(a): a y, corrected once by its residual.
__device__ __forceinline__ double sqrt_rsq(double a, double y) {
  const double s = a * y;
  return fma(fma(-s, s, a), 0.5 * y, s);
}
__device__ __forceinline__ double rcp_nr(double b) {
  double r = __builtin_amdgcn_rcp(b);
  r = fma(fma(-b, r, 1.0), r, r);
  return fma(fma(-b, r, 1.0), r, r);
}

// The particle half of the step for one point: gathers the node (v, a) of
// its stencil (fetch(j, i) gives node (bx + i, by + j)), then
// MaterialPoints::computeVelGrad (MaterialPoints.cpp:296-322),
// elasticPrediction (:324-348) with ConstitutiveModel::elasticPrediction
// (ConstitutiveModel.cpp:19-22), plasticCorrection (:350-369) with the
// Drucker-Prager return (ConstitutiveModel.cpp:26-81), updateVelocities
// (:204-245), updatePositions (:247-288) and resolvePlaneCollisions
// (:371-423).  J and b_bar come in and go out in FP64; the return map runs in
// FP64 in every precision.
template <typename T, int BASIS, typename Fetch>
__device__ __forceinline__ void g2p_point(const SceneDev &sc, const Fetch &fetch, int bx, int by,
                                          int cx, int cy, T hl, bool scripted, double &x,
                                          double &y, T &v0, T &v1, double &J1, double &c00,
                                          double &c01, double &c11) {
  constexpr int S = BasisTraits<BASIS>::S;
  using T4 = typename Vec4T<T>::type;
  const double J0 = J1;
  const double b00 = c00, b01 = c01, b11 = c11;
  const double dt = sc.dt;
  T axs[S], dxs[S], ays[S], dys[S];
  basis_2d<T, BASIS>(x, y, hl, sc.min0, sc.min1, sc.h, T(sc.inv_h), bx, by, cx, cy, axs, dxs, ays,
                     dys);

  T L00 = 0, L10 = 0, L01 = 0, L11 = 0;
  T vp0 = 0, vp1 = 0, ac0 = 0, ac1 = 0;
  if constexpr (std::is_same<T, float>::value) {
    // FP32: the weights are products ax[i] ay[j], so each node row's sums
    // over i are formed first and then weighted by the row's y factors
    // (fused multiply-adds throughout).
#pragma unroll
    for (int j = 0; j < S; ++j) {
      T dvx = 0, dvy = 0, wvx = 0, wvy = 0, wax = 0, way = 0;
#pragma unroll
      for (int ii = 0; ii < S; ++ii) {
        const T4 nv = fetch(j, ii);
        dvx = madd(nv.x, dxs[ii], dvx);
        dvy = madd(nv.y, dxs[ii], dvy);
        wvx = madd(nv.x, axs[ii], wvx);
        wvy = madd(nv.y, axs[ii], wvy);
        wax = madd(nv.z, axs[ii], wax);
        way = madd(nv.w, axs[ii], way);
      }
      L00 = madd(ays[j], dvx, L00);
      L10 = madd(ays[j], dvy, L10);
      L01 = madd(dys[j], wvx, L01);
      L11 = madd(dys[j], wvy, L11);
      vp0 = madd(ays[j], wvx, vp0);
      vp1 = madd(ays[j], wvy, vp1);
      ac0 = madd(ays[j], wax, ac0);
      ac1 = madd(ays[j], way, ac1);
    }
  } else {
#pragma unroll
    for (int j = 0; j < S; ++j) {
#pragma unroll
      for (int ii = 0; ii < S; ++ii) {
        const T4 nv = fetch(j, ii);
        const T w = axs[ii] * ays[j];
        const T gx = dxs[ii] * ays[j];
        const T gy = axs[ii] * dys[j];
        L00 += nv.x * gx;
        L10 += nv.y * gx;
        L01 += nv.x * gy;
        L11 += nv.y * gy;
        vp0 += w * nv.x;
        vp1 += w * nv.y;
        ac0 += w * nv.z;
        ac1 += w * nv.w;
      }
    }
  }

  // elastic predictor b* = b_n + dt (L b_n + b_n L^T), b_n = J b_bar
  const double l00 = double(L00), l10 = double(L10), l01 = double(L01), l11 = double(L11);
  const double n00 = b00 * J0, n01 = b01 * J0, n11 = b11 * J0;
  const double Lb00 = l00 * n00 + l01 * n01;
  const double Lb10 = l10 * n00 + l11 * n01;
  const double Lb11 = l10 * n01 + l11 * n11;
  const double bL00 = n00 * l00 + n01 * l01;
  const double bL10 = n01 * l00 + n11 * l01;
  const double bL11 = n01 * l10 + n11 * l11;
  const double e00 = n00 + dt * (Lb00 + bL00);
  const double e01 = n01 + dt * (Lb10 + bL10);
  const double e11 = n11 + dt * (Lb11 + bL11);
  constexpr bool f32 = std::is_same<T, float>::value;
  if constexpr (f32) {
    // FP32 mode (J and b_bar stored in FP32): the same return in FP64
    // arithmetic, on b* itself.  The FP64 mode's round trip through b_bar
    // and back (s = b_bar J, whose determinant is det b* again) is skipped,
    // the yield test is taken on squares (mu |q| > ys with q = dev b* / J
    // becomes mu^2 |dev b*|^2 > ys^2 det b*, both sides non-negative), and
    // square roots and quotients come from rsq_nr / sqrt_rsq / rcp_nr.
    const double detE = e00 * e11 - e01 * e01;
    const double yE = rsq_nr(detE);
    const double JE = sqrt_rsq(detE, yE);
    double ys = 0.5 * sc.dp_alpha * sc.kappa * (1.0 + JE) * (1.0 - JE);
    if (ys < 0.0) ys = 0.0;
    const double half_tr = 0.5 * (e00 + e11);
    const double d00 = e00 - half_tr, d01 = e01, d11 = e11 - half_tr;
    const double dd = d00 * d00 + d01 * d01 + d01 * d01 + d11 * d11;
    if (sc.mu * sc.mu * dd > ys * ys * detE) {
      const double fn = sqrt_rsq(dd, rsq_nr(dd)) * yE;
      const double lambda2 = ys * rcp_nr(sc.mu * fn);
      const double detD = d00 * d11 - d01 * d01;
      const double l1sq = detE - lambda2 * lambda2 * detD;
      const double lambda1 = sqrt_rsq(l1sq, rsq_nr(l1sq));
      const double r00 = lambda1 + lambda2 * d00;
      const double r01 = lambda2 * d01;
      const double r11 = lambda1 + lambda2 * d11;
      const double detR = r00 * r11 - r01 * r01;
      const double yR = rsq_nr(detR);
      J1 = sqrt_rsq(detR, yR);
      c00 = r00 * yR;
      c01 = r01 * yR;
      c11 = r11 * yR;
    } else {
      J1 = JE;
      c00 = e00 * yE;
      c01 = e01 * yE;
      c11 = e11 * yE;
    }
  } else {
    double detE = e00 * e11 - e01 * e01;
    J1 = sqrt(detE);
    double nrm = sqrt(detE);
    c00 = e00 / nrm;
    c01 = e01 / nrm;
    c11 = e11 / nrm;

    // Drucker-Prager return on b* = J b_bar
    const double s00 = c00 * J1, s01 = c01 * J1, s11 = c11 * J1;
    const double detS = s00 * s11 - s01 * s01;
    const double Jr = sqrt(detS);
    double ys = 0.5 * sc.dp_alpha * sc.kappa * (1.0 + Jr) * (1.0 - Jr);
    if (ys < 0.0) ys = 0.0;
    const double half_tr = 0.5 * (s00 + s11);
    const double d00 = s00 - half_tr, d01 = s01 - 0.0, d11 = s11 - half_tr;
    const double q00 = d00 / Jr;
    const double q01 = d01 / Jr;
    const double q11 = d11 / Jr;
    const double fn = sqrt(q00 * q00 + q01 * q01 + q01 * q01 + q11 * q11);
    double r00 = s00, r01 = s01, r11 = s11;
    if (sc.mu * fn > ys) {
      const double lambda2 = ys / (sc.mu * fn);
      const double detD = d00 * d11 - d01 * d01;
      const double l1sq = detS - lambda2 * lambda2 * detD;
      const double lambda1 = sqrt(l1sq);
      r00 = lambda1 + lambda2 * d00;
      r01 = 0.0 + lambda2 * d01;
      r11 = lambda1 + lambda2 * d11;
    }
    const double detR = r00 * r11 - r01 * r01;
    J1 = sqrt(detR);
    const double rn = sqrt(detR);
    c00 = r00 / rn;
    c01 = r01 / rn;
    c11 = r11 / rn;
  }

  // PIC/FLIP velocity and advection
  const T alpha = T(sc.pic_alpha);
  const T vf0 = v0 + ac0 * T(dt), vf1 = v1 + ac1 * T(dt);
  if (!scripted) {
    v0 = (T(1) - alpha) * vp0 + alpha * vf0;
    v1 = (T(1) - alpha) * vp1 + alpha * vf1;
    x += dt * double(vp0);
    y += dt * double(vp1);
  } else {
    x += dt * double(v0);
    y += dt * double(v1);
  }

  // point plane conditions
  for (int k = 0; k < sc.nplanes; ++k) {
    if (y < sc.plane_lower[k]) continue;
    const double pn0 = sc.plane_n[2 * k], pn1 = sc.plane_n[2 * k + 1];
    const double d = pn0 * (x - sc.plane_x[2 * k]) + pn1 * (y - sc.plane_x[2 * k + 1]);
    if (d > 0.0) continue;
    x += -d * pn0;
    y += -d * pn1;
    if (sc.plane_beh[k] == HG_PLANE_SLIDING) {
      const double vn = pn0 * double(v0) + pn1 * double(v1);
      if (vn > 0.0) continue;
      v0 = T(double(v0) - vn * pn0);
      v1 = T(double(v1) - vn * pn1);
    } else {
      v0 = T(0);
      v1 = T(0);
    }
  }
}

template <typename T, int BASIS>
__global__ __launch_bounds__(kG2PThreads) void mpm_g2p_kernel(G2PArgs a) {
  using T4 = typename Vec4T<T>::type;
  const uint32_t blk = hg_xcd_remap(blockIdx.x, a.nblocks_padded);
  const uint32_t i = blk * kG2PThreads + threadIdx.x;
  if (i >= a.npoints) return;
  const uint32_t sid = scene_of_point(a.p_off, a.nscenes, i);
  const SceneDev &sc = a.scenes[sid];
  const uint32_t p = i;
  const uint32_t o = a.pinv[i];

  double x = a.src.x[p], y = a.src.y[p];
  T v0 = ld<T>(a.src.vx, p), v1 = ld<T>(a.src.vy, p);
  double J1 = strain_load<T>(ld<T>(a.src.J, p));
  double c00 = strain_load<T>(ld<T>(a.src.b00, p));
  double c01 = double(ld<T>(a.src.b01, p));
  double c11 = strain_load<T>(ld<T>(a.src.b11, p));
  const uint32_t id = a.src.id[p];
  const T hl = a.src.hl ? ld<T>(a.src.hl, id) : T(sc.hl_uniform);
  const bool scripted = a.src.flags ? (a.src.flags[p] != 0) : false;

  int bx, by, cx, cy;
  point_key<BASIS>(sc, x, y, double(hl), bx, by, cx, cy);
  const T4 *g = static_cast<const T4 *>(a.gva) + sc.node_off;
  auto fetch = [&](int j, int ii) {
    return g[uint32_t(by + j) * uint32_t(sc.snx) + uint32_t(bx) + uint32_t(ii)];
  };
  g2p_point<T, BASIS>(sc, fetch, bx, by, cx, cy, hl, scripted, x, y, v0, v1, J1, c00, c01, c11);

  // write the point to its sorted slot o of the other buffer and count its
  // new cell
  a.dst.x[o] = x;
  a.dst.y[o] = y;
  st<T>(a.dst.vx, o, v0);
  st<T>(a.dst.vy, o, v1);
  st<T>(a.dst.b00, o, strain_store<T>(c00));
  st<T>(a.dst.b01, o, T(c01));
  st<T>(a.dst.b11, o, strain_store<T>(c11));
  st<T>(a.dst.J, o, strain_store<T>(J1));
  a.dst.id[o] = id;  // (mass, volume and half width stay where the id puts them)
  if (a.src.flags) a.dst.flags[o] = a.src.flags[p];

  int nbx, nby, ncx, ncy;
  const bool ok = point_key<BASIS>(sc, x, y, double(hl), nbx, nby, ncx, ncy);
  const bool finite = isfinite(x) && isfinite(y) && isfinite(double(v0)) &&
                      isfinite(double(v1)) && isfinite(J1);
  if (!ok || !finite) atomicOr(&a.err[sid], (ok ? 0 : 1) | (finite ? 0 : 2));
  const uint32_t k = sc.cell_off + uint32_t(nby) * uint32_t(sc.snx) + uint32_t(nbx);
  a.key[o] = k;
  a.slot[o] = hg_run_slot(a.counts, k);
  hg_scan_block_count(a.totals, k, blk);
}

// Initial binning of points in creation order.
struct BinArgs {
  const SceneDev *scenes;
  const uint32_t *p_off;
  uint32_t nscenes, npoints;
  PtsRaw pts;
  uint32_t *key, *slot, *counts, *totals;
  int32_t *err;
};

template <typename T, int BASIS>
__global__ __launch_bounds__(256) void mpm_bin_kernel(BinArgs a) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= a.npoints) return;
  const uint32_t sid = scene_of_point(a.p_off, a.nscenes, i);
  const SceneDev &sc = a.scenes[sid];
  const T hl = a.pts.hl ? ld<T>(a.pts.hl, a.pts.id[i]) : T(sc.hl_uniform);
  int bx, by, cx, cy;
  const bool ok = point_key<BASIS>(sc, a.pts.x[i], a.pts.y[i], double(hl), bx, by, cx, cy);
  if (!ok) atomicOr(&a.err[sid], 1);
  const uint32_t k = sc.cell_off + uint32_t(by) * uint32_t(sc.snx) + uint32_t(bx);
  a.key[i] = k;
  a.slot[i] = atomicAdd(&a.counts[k], 1u);
  hg_scan_block_count(a.totals, k, blockIdx.x);
}

// ---------------------------------------------------------------------------
// FP32 step, fused: G2P of step n and the rasterisation (P2G) of step n + 1
// in one pass over each active tile, then a grid kernel.
//
// mpm_fused_kernel, per tile of TW x TH owned cells (persistent workgroups
// over the active-tile list, one contiguous range of it per XCD):
//  * the tile's points are the TH runs of its owned cells in the current
//    cell order (perm); they are loaded through perm (the buffer is ordered
//    by the previous sort, so the loads are nearly contiguous);
//  * G2P reads the node (v, a) of its stencil from an LDS copy of the
//    tile's node window, runs the particle update (g2p_point) and writes the
//    point once, to its position in the current order in the other buffer,
//    with its new cell key for the next sort;
//  * from the updated point it forms its rasterisation terms for the next
//    step -- mass w m, momentum w m v and internal force -V J sigma grad w at
//    each node of its new stencil (PhysicsGrid::rasterizePointMasses
//    :168-207, rasterizePointMomentum :274-311, computeForces :413-473, with
//    MaterialPoints::computeHyperelasticCauchyStress, MaterialPoints.cpp:
//    183-202) -- from FP32 weight factors with exact FP64 products, rounds
//    each to the nearest whole number of the scene's quantum (fx_shift) in
//    the same fused multiply-add and adds it into LDS accumulators covering the
//    tile's nodes one cell beyond its points' reach; a point that moved
//    further adds straight into the node sums in HBM after the pass;
//  * the accumulators are then added into the node sums in HBM (atomics, at
//    the tile borders shared with the neighbour tiles).
// The sums are whole numbers of quanta in 64-bit integers, so every addition
// is exact and the sums do not depend on the order of the additions: each
// node's sum is a fixed function of the scene state -- bitwise reproducible,
// the same alone or in a batch and for any tile shape.  In prime mode the kernel only rasterises the current points (no
// particle update).
//
// mpm_grid_kernel, per node of the active tiles (the new sort's): converts
// the sums back (fx_quantum), clears them for the next step and runs the
// grid update (grid_node_update).
// ---------------------------------------------------------------------------
struct FusedArgs {
  const SceneDev *scenes;
  const uint32_t *tile_scene, *tile_list, *tile_count;
  PtsRaw src, dst;
  const uint32_t *perm, *cell_start;
  const float4 *gva;
  unsigned long long *acc;  // kFxQuantities arrays of nnode node sums (whole quanta)
  uint32_t nnode;
  uint32_t *key, *slot, *counts, *totals;
  int32_t *err;
};

// Node-sum terms in whole quanta.  A term t scaled to quanta with |t| < 2^51,
// added to 1.5 2^52, lands in [2^52, 2^53), where the spacing of doubles is 1:
// the sum is t rounded to the nearest whole number (ties to even), and its bit
// pattern is kFxBias plus that number as a two's-complement integer.  A fused
// multiply-add with the magic number as its addend therefore forms a product
// and rounds it to quanta in one instruction (the caller bounds the terms
// below 2^50).
constexpr double kFxMagic = 0x1.8p52;
constexpr unsigned long long kFxBias = 0x4338000000000000ull;
__device__ __forceinline__ unsigned long long fx_bits(double biased) {
  return static_cast<unsigned long long>(__double_as_longlong(biased));
}

// A point's rasterisation terms: per axis the S weights and masked
// weight-derivative factors of its stencil (base node bx, by; cntx, cnty
// stencil nodes per axis), formed in FP32 and held in FP64 (exactly), so that
// every node's weight and gradient products are exact; and its mass, momentum
// and -V J sigma.  The stress comes from the strain as stored: the deviations
// dJ = J - 1 and d = b_bar - I (off-diagonal b_bar01).
template <int BASIS>
struct P2GTerms {
  static constexpr int S = BasisTraits<BASIS>::S;
  double wx[S], gx[S], wy[S], gy[S];
  // mass, momentum and -V J sigma, each times its quantity's 2^fx_shift
  // (an exact power-of-two scaling)
  double m, mv0, mv1, f00, f01, f11;
};

// MaterialPoints::computeHyperelasticCauchyStress (MaterialPoints.cpp:183-202)
// gives tau = kappa/2 (Jt^2 - 1) I + mu dev(e / Jt) for e = J b_bar and
// Jt = sqrt(det e), with no stress in tension (J > 1).  Written in the
// deviations a = e00 - 1, c = e11 - 1 and e01, all small, it needs no FP64:
// Jt^2 - 1 = det e - 1 = a + c + a c - e01^2, and dev(e / Jt) has the entries
// +-(a - c) / (2 Jt) and e01 / Jt, so every quantity keeps FP32's relative
// precision (none is formed as a difference of numbers near 1).
__device__ __forceinline__ void stress_from_deviations(const SceneDev &sc, float dJ, float d00,
                                                       float d01, float d11, float &t00,
                                                       float &t01, float &t11) {
  t00 = t01 = t11 = 0.0f;
  if (dJ <= 0.0f) {
    const float a = __builtin_fmaf(d00, dJ, d00 + dJ);
    const float c = __builtin_fmaf(d11, dJ, d11 + dJ);
    const float e01 = __builtin_fmaf(d01, dJ, d01);
    const float del = __builtin_fmaf(-e01, e01, __builtin_fmaf(a, c, a + c));
    const float rj = __builtin_amdgcn_rsqf(1.0f + del);
    const float p = 0.5f * float(sc.kappa) * del;
    const float h = 0.5f * float(sc.mu) * (a - c) * rj;
    t00 = p + h;
    t11 = p - h;
    t01 = float(sc.mu) * e01 * rj;
  }
}

template <int BASIS>
__device__ __forceinline__ void p2g_terms(const SceneDev &sc, double x, double y, float v0,
                                          float v1, float m, float V, float hl, float dJ,
                                          float d00, float d01, float d11, int bx, int by,
                                          int cntx, int cnty, P2GTerms<BASIS> &o) {
  constexpr int S = BasisTraits<BASIS>::S;
  float t00, t01, t11;
  stress_from_deviations(sc, dJ, d00, d01, d11, t00, t01, t11);
  // -V J sigma with sigma = tau / J is -V tau (products of two FP32 values,
  // exact in FP64)
  const double nV = -double(V);
  const int sf = sc.fx_shift[kFxF0], sp = sc.fx_shift[kFxMom0];
  o.f00 = ldexp(nV * double(t00), sf);
  o.f01 = ldexp(nV * double(t01), sf);
  o.f11 = ldexp(nV * double(t11), sf);
  const double md = double(m);
  o.m = ldexp(md, sc.fx_shift[kFxMass]);
  // products of two FP32 values are exact in FP64
  o.mv0 = ldexp(md * double(v0), sp);
  o.mv1 = ldexp(md * double(v1), sp);
  const float inv_h = float(sc.inv_h);
  float inv_vp = 1.0f;
  if constexpr (BASIS == HG_BASIS_UGIMP_LINEAR) inv_vp = 1.0f / (4.0f * hl * hl);
  float fax[S], fdx[S], fay[S], fdy[S];
  if constexpr (BASIS == HG_BASIS_UGIMP_LINEAR) {
    ugimp_1d_f32(x, hl, sc.min0, sc.h, inv_h, bx, cntx, fax, fdx);
    ugimp_1d_f32(y, hl, sc.min1, sc.h, inv_h, by, cnty, fay, fdy);
  } else {
#pragma unroll
    for (int k = 0; k < S; ++k) {
      basis_factor<float, BASIS>(x, hl, sc.min0, sc.h, inv_h, bx + k, k, cntx, fax[k], fdx[k]);
      basis_factor<float, BASIS>(y, hl, sc.min1, sc.h, inv_h, by + k, k, cnty, fay[k], fdy[k]);
    }
  }
#pragma unroll
  for (int k = 0; k < S; ++k) {
    const float ax = fax[k], dx = fdx[k], ay = fay[k], dy = fdy[k];
    float wxk, gxk, gyk;
    if constexpr (BASIS == HG_BASIS_UGIMP_LINEAR) {
      wxk = ax * inv_vp;
      gxk = dx * inv_vp;
      gyk = dy;
    } else {
      wxk = ax;
      gxk = dx * inv_h;
      gyk = dy * inv_h;
    }
    // a node's terms count only where its weight's factors are nonzero
    // (PhysicsGrid.cpp:413-473 skips zero-weight nodes); with a zero factor
    // every term of the node is then exactly zero
    o.wx[k] = double(wxk);
    o.gx[k] = double((wxk != 0.0f) ? gxk : 0.0f);
    o.wy[k] = double(ay);
    o.gy[k] = double((ay != 0.0f) ? gyk : 0.0f);
  }
}

// False when a scaled term could reach 2^50 quanta (|w| <= 1, so the terms
// are bounded by m, |m v| and (|f00| + 2 |f01| + |f11|) times the largest
// weight-derivative factor), or is not finite.
template <int BASIS>
__device__ __forceinline__ bool p2g_terms_in_range(const P2GTerms<BASIS> &o) {
  constexpr int S = BasisTraits<BASIS>::S;
  double g = 0.0;
#pragma unroll
  for (int k = 0; k < S; ++k) g = fmax(g, fmax(fabs(o.gx[k]), fabs(o.gy[k])));
  const double fb = (fabs(o.f00) + 2.0 * fabs(o.f01) + fabs(o.f11)) * g;
  const double lim = 1.0e15;  // below 2^50
  return fabs(o.m) < lim && fmax(fabs(o.mv0), fabs(o.mv1)) < lim && fb < lim;
}

// Adds a point's terms at the nodes of its stencil: add(q, i, j, bits) for
// quantity q at node (bx + i, by + j), where bits is kFxBias plus the term in
// whole quanta (mod 2^64).  The weight and gradient products are exact; each
// mass and momentum term is rounded once, each force term (the sum of two
// products) once after each product.
template <int BASIS, typename Add>
__device__ __forceinline__ void p2g_scatter(const P2GTerms<BASIS> &o, const Add &add) {
  constexpr int S = BasisTraits<BASIS>::S;
#pragma unroll
  for (int j = 0; j < S; ++j) {
#pragma unroll
    for (int i = 0; i < S; ++i) {
      const double w = o.wx[i] * o.wy[j];
      const double gx = o.gx[i] * o.wy[j];
      const double gy = o.wx[i] * o.gy[j];
      add(kFxMass, i, j, fx_bits(__builtin_fma(w, o.m, kFxMagic)));
      add(kFxMom0, i, j, fx_bits(__builtin_fma(w, o.mv0, kFxMagic)));
      add(kFxMom1, i, j, fx_bits(__builtin_fma(w, o.mv1, kFxMagic)));
      add(kFxF0, i, j, fx_bits(__builtin_fma(o.f01, gy, __builtin_fma(o.f00, gx, kFxMagic))));
      add(kFxF1, i, j, fx_bits(__builtin_fma(o.f11, gy, __builtin_fma(o.f01, gx, kFxMagic))));
    }
  }
}

// A point's stored state, as the fused kernel loads it.
struct FusedPoint {
  double x, y, J, c00, c01, c11;
  float v0, v1, m, V, hl;
  float sJ, s00, s01, s11;  // the strain as stored (FP32 deviations)
};

// The strain of a point from its stored form.
__device__ __forceinline__ void fused_strain(FusedPoint &q) {
  q.J = strain_load<float>(q.sJ);
  q.c00 = strain_load<float>(q.s00);
  q.c01 = double(q.s01);
  q.c11 = strain_load<float>(q.s11);
}

// The stored fields only (the strain as stored), so that a load can be issued
// a pass ahead of its use without holding the FP64 strain in registers.
__device__ __forceinline__ FusedPoint fused_load_raw(const PtsRaw &p, uint32_t i, uint32_t id,
                                                     const SceneDev &sc) {
  FusedPoint q;
  q.x = p.x[i];
  q.y = p.y[i];
  q.v0 = ld<float>(p.vx, i);
  q.v1 = ld<float>(p.vy, i);
  q.sJ = ld<float>(p.J, i);
  q.s00 = ld<float>(p.b00, i);
  q.s01 = ld<float>(p.b01, i);
  q.s11 = ld<float>(p.b11, i);
  // mass, volume and half width are indexed by point id (never moved)
  q.m = p.m ? ld<float>(p.m, id) : float(sc.m_uniform);
  q.V = p.V ? ld<float>(p.V, id) : float(sc.V_uniform);
  q.hl = p.hl ? ld<float>(p.hl, id) : float(sc.hl_uniform);
  return q;
}

// A rasterised point's stencil in the tile's window, packed into the stash:
// base node offsets from the window origin (8 bits each; the window is at most
// 64 + 5 nodes wide) and the stencil's node counts less one (2 bits each).
// kStashNone marks a lane with no point, kStashFar a point whose stencil
// leaves the window.
constexpr uint32_t kStashNone = 0xffffffffu, kStashFar = 0xfffffffeu;
__device__ __forceinline__ uint32_t stash_key(int dx, int dy, int cx, int cy) {
  return uint32_t(dx) | (uint32_t(dy) << 8) | (uint32_t(cx - 1) << 16) | (uint32_t(cy - 1) << 18);
}

// Points of a tile that moved more than one cell since the sort, whose terms
// are added after the tile's pass straight into the node sums.
constexpr int kFarCap = 256;

// Threads of a fused tile: MULT per owned cell (MULT points per cell per
// pass), a knob the autotuner sets.
template <int TW, int TH, int MULT> constexpr int fused_threads() { return MULT * TW * TH; }

// A workgroup barrier that orders LDS only: the global stores before it may
// still be in flight (no thread reads them before a full barrier).
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

template <int BASIS, int TW, int TH, bool PRIME, int MULT>
__global__ __launch_bounds__((fused_threads<TW, TH, MULT>())) __attribute__((amdgpu_waves_per_eu(4))) void mpm_fused_kernel(FusedArgs a) {
  constexpr int S = BasisTraits<BASIS>::S;
  constexpr int NT = fused_threads<TW, TH, MULT>();
  constexpr int GW = TW + S - 1, GH = TH + S - 1;  // G2P node window
  constexpr int RW = TW + S + 1, RH = TH + S + 1;  // rasterisation window
  constexpr int RN = RW * RH;
  __shared__ float4 s_va[PRIME ? 1 : GW * GH];
  __shared__ unsigned long long s_acc[kFxQuantities][RN];
  __shared__ uint32_t s_rb[TH], s_ro[TH + 1];
  __shared__ uint32_t s_far[kFarCap];
  __shared__ uint32_t s_nfar;
  // one pass's points after G2P (or as loaded, priming), for the terms
  __shared__ double s_px[NT], s_py[NT];
  __shared__ float s_pf[9][NT];
  __shared__ uint32_t s_pi[NT], s_pk[NT];
  static_assert(TW + S + 1 <= 256 && TH + S + 1 <= 256 && S <= 4, "stash key fields");

  const uint32_t ntile_list = *a.tile_count;
  const int t = threadIdx.x;
  // Workgroup b runs on XCD b % 8 (the grid is a multiple of 8); each XCD
  // takes one contiguous eighth of the list, so neighbouring tiles' shared
  // nodes stay in one L2.
  const uint32_t per_xcd = (ntile_list + 7u) >> 3, xcd = blockIdx.x & 7u;
  for (uint32_t jt = blockIdx.x >> 3; jt < per_xcd; jt += gridDim.x >> 3) {
    const uint32_t li = xcd * per_xcd + jt;
    if (li >= ntile_list) break;
    const uint32_t tile = a.tile_list[li];
    const uint32_t sid = a.tile_scene[tile];
    const SceneDev &sc = a.scenes[sid];
    const uint32_t ltile = tile - sc.tile_off;
    const int x0 = int(ltile % uint32_t(sc.tiles_x)) * TW;
    const int y0 = int(ltile / uint32_t(sc.tiles_x)) * TH;
    const uint32_t *cs = a.cell_start + sc.cell_off;
    // the owned cells' points: one run of the current order per cell row
    if (t < TH) {
      const int ry = y0 + t;
      uint32_t b = 0, e = 0;
      if (ry < sc.nny) {
        const uint32_t row = uint32_t(ry) * uint32_t(sc.snx);
        b = cs[row + uint32_t(min(x0, sc.nnx))];
        e = cs[row + uint32_t(min(x0 + TW, sc.nnx))];
      }
      s_rb[t] = b;
      s_ro[t + 1] = e - b;
    }
    if (t == 0) s_nfar = 0;
    if constexpr (!PRIME) {
      for (int q = t; q < GW * GH; q += NT) {
        const int nx = x0 + q % GW, ny = y0 + q / GW;
        float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
        if (nx < sc.snx && ny < sc.sny)
          v = a.gva[sc.node_off + uint32_t(ny) * uint32_t(sc.snx) + uint32_t(nx)];
        s_va[q] = v;
      }
    }
    for (int q = t; q < kFxQuantities * RN; q += NT) (&s_acc[0][0])[q] = 0ull;
    __syncthreads();
    if (t == 0) {
      uint32_t o = 0;
      s_ro[0] = 0;
      for (int r = 0; r < TH; ++r) {
        o += s_ro[r + 1];
        s_ro[r + 1] = o;
      }
    }
    __syncthreads();
    const uint32_t total = s_ro[TH];
    const int wx0 = x0 - 1, wy0 = y0 - 1;  // origin of the rasterisation window

    // Each pass takes NT points of the tile in two phases.  G2P runs on
    // consecutive lanes (neighbouring points share node gathers and slot
    // atomics) and leaves each moved point in LDS; the terms then run with
    // lane l of wave w on point l * NW + w, so that the points of one cell
    // (consecutive in the order) go to different waves and the lanes of one
    // LDS atomic add to distinct nodes instead of queueing on one address.
    constexpr int NW = NT / 64;
    static_assert(NT % 64 == 0, "whole waves per tile");
    const uint32_t lane_pt = uint32_t((t & 63) * NW + (t >> 6));
    // position in the current order of the tile's k-th point
    auto order_pos = [&](uint32_t k) {
      int r = 0;
#pragma unroll
      for (int q = 1; q < TH; ++q) r += (s_ro[q] <= k) ? 1 : 0;
      return s_rb[r] + (k - s_ro[r]);
    };
    // Loads run ahead of their use: a pass's point fields are requested
    // before the raster half of the pass before it, and their buffer indices
    // (perm) one pass earlier still.
    uint32_t sp = 0, p = 0, sp_n = 0, p_n = 0, id = 0;
    bool scripted = false;
    FusedPoint q;
    if (uint32_t(t) < total) {
      sp = order_pos(uint32_t(t));
      p = a.perm[sp];
    }
    if (uint32_t(NT + t) < total) {
      sp_n = order_pos(uint32_t(NT + t));
      p_n = a.perm[sp_n];
    }
    if (uint32_t(t) < total) {
      id = a.src.id[p];
      q = fused_load_raw(a.src, p, id, sc);
      if constexpr (!PRIME) scripted = a.src.flags ? (a.src.flags[p] != 0) : false;
    }
    for (uint32_t k0 = 0; k0 < total; k0 += uint32_t(NT)) {
      const uint32_t k = k0 + uint32_t(t);
      uint32_t tag = kStashNone, pk = kStashNone;
      HgRunSlot slot_v{0u, 0u};
      uint32_t slot_at = 0xffffffffu;
      if (k < total) {
        fused_strain(q);
        int bx, by, cx, cy;
        if constexpr (!PRIME) {
          point_key<BASIS>(sc, q.x, q.y, double(q.hl), bx, by, cx, cy);
          const bool inwin = bx >= x0 && bx < x0 + TW && by >= y0 && by < y0 + TH;
          // the points of the tile's cells gather from the LDS window; a
          // wave with a point whose stencil is elsewhere (moved since its
          // sort) takes the general form
          if (__all(inwin)) {
            const float4 *w = s_va + ((by - y0) * GW + (bx - x0));
            auto fetch = [&](int j, int ii) { return w[j * GW + ii]; };
            g2p_point<float, BASIS>(sc, fetch, bx, by, cx, cy, q.hl, scripted, q.x, q.y, q.v0,
                                    q.v1, q.J, q.c00, q.c01, q.c11);
          } else {
            const float4 *g = a.gva + sc.node_off;
            auto fetch = [&](int j, int ii) {
              return inwin ? s_va[(by - y0 + j) * GW + (bx - x0 + ii)]
                           : g[uint32_t(by + j) * uint32_t(sc.snx) + uint32_t(bx + ii)];
            };
            g2p_point<float, BASIS>(sc, fetch, bx, by, cx, cy, q.hl, scripted, q.x, q.y, q.v0,
                                    q.v1, q.J, q.c00, q.c01, q.c11);
          }
          // the point's one write, at its position in the current order; the
          // next step's terms use the stored (rounded) strain, as a
          // rasterisation of the stored state (after a restore) does
          q.sJ = strain_store<float>(q.J);
          q.s00 = strain_store<float>(q.c00);
          q.s01 = float(q.c01);
          q.s11 = strain_store<float>(q.c11);
          a.dst.x[sp] = q.x;
          a.dst.y[sp] = q.y;
          st<float>(a.dst.vx, sp, q.v0);
          st<float>(a.dst.vy, sp, q.v1);
          st<float>(a.dst.b00, sp, q.s00);
          st<float>(a.dst.b01, sp, q.s01);
          st<float>(a.dst.b11, sp, q.s11);
          st<float>(a.dst.J, sp, q.sJ);
          a.dst.id[sp] = id;
          if (a.src.flags) a.dst.flags[sp] = scripted ? 1 : 0;
          const bool ok = point_key<BASIS>(sc, q.x, q.y, double(q.hl), bx, by, cx, cy);
          const bool finite = isfinite(q.x) && isfinite(q.y) && isfinite(double(q.v0)) &&
                              isfinite(double(q.v1)) && isfinite(q.J);
          if (!ok || !finite) atomicOr(&a.err[sid], (ok ? 0 : 1) | (finite ? 0 : 2));
          const uint32_t key = sc.cell_off + uint32_t(by) * uint32_t(sc.snx) + uint32_t(bx);
          a.key[sp] = key;
          // the slot's atomic returns while the raster half runs; the slot
          // is taken and stored after it (waiting for the return here would
          // also wait for every store above)
          slot_v = hg_run_slot_issue(a.counts, key);
          slot_at = sp;
          hg_scan_block_count(a.totals, key, blockIdx.x);
        } else {
          point_key<BASIS>(sc, q.x, q.y, double(q.hl), bx, by, cx, cy);
        }
        s_px[t] = q.x;
        s_py[t] = q.y;
        s_pf[0][t] = q.v0;
        s_pf[1][t] = q.v1;
        s_pf[2][t] = q.m;
        s_pf[3][t] = q.V;
        s_pf[4][t] = q.hl;
        s_pf[5][t] = q.sJ;
        s_pf[6][t] = q.s00;
        s_pf[7][t] = q.s01;
        s_pf[8][t] = q.s11;
        tag = PRIME ? p : sp;
        pk = (bx >= wx0 && bx + S <= wx0 + RW && by >= wy0 && by + S <= wy0 + RH)
                 ? stash_key(bx - wx0, by - wy0, cx, cy)
                 : kStashFar;
      }
      s_pi[t] = tag;
      s_pk[t] = pk;
      // the next pass's fields and the pass after it's buffer indices
      {
        const uint32_t k1 = k0 + uint32_t(NT + t), k2 = k0 + uint32_t(2 * NT + t);
        sp = sp_n;
        p = p_n;
        if (k1 < total) {
          id = a.src.id[p];
          q = fused_load_raw(a.src, p, id, sc);
          if constexpr (!PRIME) scripted = a.src.flags ? (a.src.flags[p] != 0) : false;
        }
        if (k2 < total) {
          sp_n = order_pos(k2);
          p_n = a.perm[sp_n];
        }
      }
      lds_barrier();
      // the moved points' terms for the next step, at their new stencils
      const uint32_t rk = s_pk[lane_pt];
      if (rk == kStashFar) {
        // moved more than a cell since the sort (or left the grid): its
        // terms go straight into the node sums after the tile
        const uint32_t slot = atomicAdd(&s_nfar, 1u);
        if (slot < uint32_t(kFarCap)) s_far[slot] = s_pi[lane_pt];
        else atomicOr(&a.err[sid], kMpmFixedRangeBit);
      } else if (rk != kStashNone) {
        FusedPoint r;
        r.x = s_px[lane_pt];
        r.y = s_py[lane_pt];
        r.v0 = s_pf[0][lane_pt];
        r.v1 = s_pf[1][lane_pt];
        r.m = s_pf[2][lane_pt];
        r.V = s_pf[3][lane_pt];
        r.hl = s_pf[4][lane_pt];
        r.sJ = s_pf[5][lane_pt];
        r.s00 = s_pf[6][lane_pt];
        r.s01 = s_pf[7][lane_pt];
        r.s11 = s_pf[8][lane_pt];
        const int dx = int(rk & 0xffu), dy = int((rk >> 8) & 0xffu);
        const int cx = int((rk >> 16) & 3u) + 1, cy = int((rk >> 18) & 3u) + 1;
        P2GTerms<BASIS> o;
        p2g_terms<BASIS>(sc, r.x, r.y, r.v0, r.v1, r.m, r.V, r.hl, r.sJ, r.s00, r.s01, r.s11,
                         wx0 + dx, wy0 + dy, cx, cy, o);
        if (!p2g_terms_in_range<BASIS>(o)) atomicOr(&a.err[sid], kMpmFixedRangeBit);
        unsigned long long *base = &s_acc[0][0] + dy * RW + dx;
        p2g_scatter<BASIS>(o, [&](int f, int i, int j, unsigned long long v) {
          atomicAdd(base + f * RN + j * RW + i, v - kFxBias);
        });
      }
      if constexpr (!PRIME) {
        if (slot_at != 0xffffffffu) a.slot[slot_at] = hg_run_slot_take(slot_v);
      }
      lds_barrier();  // the stash is refilled by the next pass
    }
    __syncthreads();
    // the window's sums into the node sums (nodes off the grid get nothing)
    for (int q = t; q < RN; q += NT) {
      const int nx = wx0 + q % RW, ny = wy0 + q / RW;
      if (nx < 0 || ny < 0 || nx >= sc.nnx || ny >= sc.nny) continue;
      const uint32_t node = sc.node_off + uint32_t(ny) * uint32_t(sc.snx) + uint32_t(nx);
#pragma unroll
      for (int f = 0; f < kFxQuantities; ++f) {
        const unsigned long long v = s_acc[f][q];
        if (v) atomicAdd(&a.acc[size_t(f) * a.nnode + node], v);
      }
    }
    // the far points, from the state this pass wrote (or, priming, read)
    const uint32_t nfar = min(s_nfar, uint32_t(kFarCap));
    for (uint32_t f = uint32_t(t); f < nfar; f += uint32_t(NT)) {
      const uint32_t i = s_far[f];
      const PtsRaw &fp = PRIME ? a.src : a.dst;
      const FusedPoint q = fused_load_raw(fp, i, fp.id[i], sc);
      int bx, by, cx, cy;
      point_key<BASIS>(sc, q.x, q.y, double(q.hl), bx, by, cx, cy);
      P2GTerms<BASIS> o;
      p2g_terms<BASIS>(sc, q.x, q.y, q.v0, q.v1, q.m, q.V, q.hl, q.sJ, q.s00, q.s01, q.s11, bx,
                       by, cx, cy, o);
      if (!p2g_terms_in_range<BASIS>(o)) atomicOr(&a.err[sid], kMpmFixedRangeBit);
      const uint32_t nb = sc.node_off + uint32_t(by) * uint32_t(sc.snx) + uint32_t(bx);
      p2g_scatter<BASIS>(o, [&](int fq, int i2, int j2, unsigned long long v) {
        atomicAdd(&a.acc[size_t(fq) * a.nnode + nb + uint32_t(j2) * uint32_t(sc.snx) + uint32_t(i2)],
                  v - kFxBias);
      });
    }
    __syncthreads();  // LDS is reused by the next tile
  }
}

struct GridArgs {
  const SceneDev *scenes;
  const uint32_t *tile_scene, *tile_list, *tile_count;
  unsigned long long *acc;
  uint32_t nnode;
  void *gva;
  double *gm, *gp, *gpn, *gf;
  uint32_t store;
  int32_t *err;
  int32_t tw, th;
  GrainNodes ex;     // the hybrid's grain side of each node
  int32_t exchange;  // apply the grain exchange where grain mass lies
};

// Node sums of 2^62 quanta or more are flagged: further additions could wrap
// around.  With exchange set (the hybrid's continuum, one scene, node mass
// and both momenta stored) a node holding grain mass then gets the
// grain-continuum exchange.
__global__ __launch_bounds__(256) void mpm_grid_kernel(GridArgs a) {
  const uint32_t ntile_list = *a.tile_count;
  const uint32_t per_xcd = (ntile_list + 7u) >> 3, xcd = blockIdx.x & 7u;
  const int nt = a.tw * a.th;
  for (uint32_t jt = blockIdx.x >> 3; jt < per_xcd; jt += gridDim.x >> 3) {
    const uint32_t li = xcd * per_xcd + jt;
    if (li >= ntile_list) break;
    const uint32_t tile = a.tile_list[li];
    const uint32_t sid = a.tile_scene[tile];
    const SceneDev &sc = a.scenes[sid];
    const uint32_t ltile = tile - sc.tile_off;
    const int x0 = int(ltile % uint32_t(sc.tiles_x)) * a.tw;
    const int y0 = int(ltile / uint32_t(sc.tiles_x)) * a.th;
    for (int q = threadIdx.x; q < nt; q += 256) {
      const int nx = x0 + q % a.tw, ny = y0 + q / a.tw;
      if (nx >= sc.nnx || ny >= sc.nny) continue;
      const uint32_t node = sc.node_off + uint32_t(ny) * uint32_t(sc.snx) + uint32_t(nx);
      double v[kFxQuantities];
      bool big = false;
#pragma unroll
      for (int f = 0; f < kFxQuantities; ++f) {
        unsigned long long &w = a.acc[size_t(f) * a.nnode + node];
        const long long iv = (long long)w;
        w = 0ull;
        big |= iv > (1ll << 62) || iv < -(1ll << 62);
        v[f] = double(iv) * sc.fx_quantum[f];
      }
      if (big) atomicOr(&a.err[sid], kMpmFixedRangeBit);
      const double m = v[kFxMass], p0 = v[kFxMom0], p1 = v[kFxMom1];
      double f0 = v[kFxF0], f1 = v[kFxF1], pn0, pn1;
      grid_node_momentum<float>(sc, nx, ny, m, p0, p1, f0, f1, pn0, pn1);
      if (a.exchange) {
        // the grain exchange where grain mass lies, on the values in hand
        const uint32_t n = node - sc.node_off;
        const double m_dem = a.ex.dmass[n];
        if (m_dem > 0.0) {
          exchange_in_place(n, a.ex, m_dem, m, pn0, pn1);
          if (a.ex.clear_mass) {
            a.ex.clear_mass[n] = 0.0;
            a.ex.clear_mom[2 * n] = 0.0;
            a.ex.clear_mom[2 * n + 1] = 0.0;
          }
        }
      }
      grid_node_store<float>(sc, node, m, p0, p1, pn0, pn1, f0, f1, a.store, a.gva, a.gm, a.gp,
                             a.gpn, a.gf);
    }
  }
}

}  // namespace

// ===========================================================================
// host side
// ===========================================================================

static int stencil_width(int basis) {
  return basis == HG_BASIS_LINEAR ? 2 : (basis == HG_BASIS_CUBIC ? 4 : 3);
}
int hg_mpm::tw() const { return aw; }
int hg_mpm::th() const { return ah; }

template <typename P>
static hg_status dalloc(hg_mpm *h, P **out, size_t bytes) {
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

#define HG_TRY(expr)                 \
  do {                               \
    hg_status s_ = (expr);           \
    if (s_ != HG_OK) return s_;      \
  } while (0)

// ---- kernel dispatch ----------------------------------------------------------

template <typename T> struct P2GCap;
template <> struct P2GCap<float> { static constexpr int value = 384; };
template <> struct P2GCap<double> { static constexpr int value = 192; };

// Workgroups of the P2G launch: enough to fill the chip several times over;
// each takes active tiles from the list until it is exhausted.
static uint32_t p2g_grid(const hg_mpm *h) { return hg_pad8(std::max(1u, std::min(h->ntiles, 2048u))); }

template <typename T, int BASIS, int TW, int TH>
static void launch_p2g_t(hg_mpm *h, uint32_t store) {
  P2GArgs a;
  a.scenes = h->d_scenes;
  a.tile_scene = h->d_tile_scene;
  a.tile_list = h->tile_list;
  a.tile_count = h->tile_count + h->cur;
  a.pts = h->buf[h->cur];
  a.perm = h->perm;
  a.cell_start = h->cell_start;
  a.gva = h->gva;
  a.gm = h->gm;
  a.gp = h->gp;
  a.gpn = h->gpn;
  a.gf = h->gf;
  a.store = store;
  ProfScope ps(h->prof, "mpm_p2g", h->stream);
  hipLaunchKernelGGL((mpm_p2g_kernel<T, BASIS, TW, TH, P2GCap<T>::value>), dim3(p2g_grid(h)),
                     dim3(TW * TH), 0, h->stream, a);
}

template <typename T, int BASIS>
static void launch_p2g_b(hg_mpm *h, uint32_t store) {
  if (h->tw() == 32 && h->th() == 8) launch_p2g_t<T, BASIS, 32, 8>(h, store);
  else if (h->tw() == 16 && h->th() == 8) launch_p2g_t<T, BASIS, 16, 8>(h, store);
  else if (h->tw() == 64 && h->th() == 4) launch_p2g_t<T, BASIS, 64, 4>(h, store);
  else launch_p2g_t<T, BASIS, 16, 16>(h, store);
}

template <typename T>
static void launch_p2g_p(hg_mpm *h, uint32_t store) {
  switch (h->basis) {
    case HG_BASIS_LINEAR: launch_p2g_b<T, HG_BASIS_LINEAR>(h, store); break;
    case HG_BASIS_CUBIC: launch_p2g_b<T, HG_BASIS_CUBIC>(h, store); break;
    default: launch_p2g_b<T, HG_BASIS_UGIMP_LINEAR>(h, store); break;
  }
}

static void launch_p2g(hg_mpm *h, uint32_t store) {
  h->last_store = store;
  if (h->precision == HG_FP64) launch_p2g_p<double>(h, store);
  else launch_p2g_p<float>(h, store);
}

template <typename T, int BASIS>
static void launch_g2p_t(hg_mpm *h) {
  // the separate G2P writes each point at its sorted position (pinv), which
  // sorts for the fused step leave unwritten
  if (!h->pinv_ok && h->npoints) {
    hipLaunchKernelGGL(pinv_from_perm_kernel, dim3((h->npoints + 255) / 256), dim3(256), 0,
                       h->stream, h->perm, h->npoints, h->pinv);
    h->pinv_ok = true;
  }
  G2PArgs a;
  a.scenes = h->d_scenes;
  a.p_off = h->d_p_off;
  a.nscenes = h->nscenes;
  a.npoints = h->npoints;
  const uint32_t nb = (h->npoints + kG2PThreads - 1) / kG2PThreads;
  a.nblocks_padded = hg_pad8(nb);
  a.src = h->buf[h->cur];
  a.dst = h->buf[1 - h->cur];
  a.pinv = h->pinv;
  a.gva = h->gva;
  a.key = h->key;
  a.slot = h->slot;
  a.counts = h->counts;
  a.totals = h->scan_totals + size_t(1 - h->cur) * hg_scan_totals_words(h->nscan);  // the keys of buf[1 - cur]
  a.err = h->err;
  ProfScope ps(h->prof, "mpm_g2p", h->stream);
  hipLaunchKernelGGL((mpm_g2p_kernel<T, BASIS>), dim3(a.nblocks_padded), dim3(kG2PThreads), 0,
                     h->stream, a);
}

template <typename T>
static void launch_g2p_p(hg_mpm *h) {
  switch (h->basis) {
    case HG_BASIS_LINEAR: launch_g2p_t<T, HG_BASIS_LINEAR>(h); break;
    case HG_BASIS_CUBIC: launch_g2p_t<T, HG_BASIS_CUBIC>(h); break;
    default: launch_g2p_t<T, HG_BASIS_UGIMP_LINEAR>(h); break;
  }
}

template <typename T, int BASIS>
static void launch_bin_t(hg_mpm *h) {
  BinArgs a;
  a.scenes = h->d_scenes;
  a.p_off = h->d_p_off;
  a.nscenes = h->nscenes;
  a.npoints = h->npoints;
  a.pts = h->buf[h->cur];
  a.key = h->key;
  a.slot = h->slot;
  a.counts = h->counts;
  a.totals = h->scan_totals + size_t(h->cur) * hg_scan_totals_words(h->nscan);
  a.err = h->err;
  const uint32_t nb = (h->npoints + 255) / 256;
  if (nb) hipLaunchKernelGGL((mpm_bin_kernel<T, BASIS>), dim3(nb), dim3(256), 0, h->stream, a);
}

template <typename T>
static void launch_bin_p(hg_mpm *h) {
  switch (h->basis) {
    case HG_BASIS_LINEAR: launch_bin_t<T, HG_BASIS_LINEAR>(h); break;
    case HG_BASIS_CUBIC: launch_bin_t<T, HG_BASIS_CUBIC>(h); break;
    default: launch_bin_t<T, HG_BASIS_UGIMP_LINEAR>(h); break;
  }
}

static bool fused_mode(const hg_mpm *h);

static void launch_tile_list(hg_mpm *h, bool scatter) {
  TileListArgs a;
  a.scenes = h->d_scenes;
  a.tile_scene = h->d_tile_scene;
  a.cell_start = h->cell_start;
  a.tile_active = h->tile_active;
  a.tile_list = h->tile_list;
  a.tile_count = h->tile_count + h->cur;
  a.next_count = h->tile_count + (1 - h->cur);
  a.ntiles = h->ntiles;
  a.tw = h->tw();
  a.th = h->th();
  a.S = stencil_width(h->basis);
  a.key = h->key;
  a.slot = h->slot;
  a.perm = h->perm;
  // pinv only for the separate G2P; the fused step reads through perm
  a.pinv = fused_mode(h) ? nullptr : h->pinv;
  a.npoints = scatter ? h->npoints : 0u;
  if (scatter) h->pinv_ok = a.pinv != nullptr;
  const uint32_t blocks = (h->ntiles + 255) / 256 + (a.npoints + 255) / 256;
  ProfScope ps(h->prof, "mpm_tile_list", h->stream);
  hipLaunchKernelGGL(mpm_tile_list_kernel, dim3(blocks), dim3(256), 0, h->stream, a);
}

// Sorts the points of the current buffer by cell (perm, pinv, cell_start)
// and lists the active tiles: the scan, then one launch for the scatter and
// the tile list.
static void launch_sort(hg_mpm *h) {
  hg_counting_scan(h->prof, h->stream, h->counts, h->ncell,
                   h->scan_totals + size_t(h->cur) * hg_scan_totals_words(h->nscan),
                   h->scan_totals + size_t(1 - h->cur) * hg_scan_totals_words(h->nscan), h->nscan,
                   h->cell_start, h->npoints, h->err);
  launch_tile_list(h, true);
}

static void launch_g2p_and_sort(hg_mpm *h) {
  if (h->npoints) {
    if (h->precision == HG_FP64) launch_g2p_p<double>(h);
    else launch_g2p_p<float>(h);
  }
  // the points now live in the other buffer, which the sort orders
  h->cur = 1 - h->cur;
  launch_sort(h);
}

static void launch_velacc(hg_mpm *h) {
  VelAccArgs a;
  a.scenes = h->d_scenes;
  a.tile_scene = h->d_tile_scene;
  a.tile_list = h->tile_list;
  a.tile_count = h->tile_count + h->cur;
  a.gm = h->gm;
  a.gp = h->gp;
  a.gpn = h->gpn;
  a.gva = h->gva;
  a.tw = h->tw();
  a.th = h->th();
  ProfScope ps(h->prof, "mpm_velacc", h->stream);
  if (h->precision == HG_FP64)
    hipLaunchKernelGGL(mpm_velacc_kernel<double>, dim3(p2g_grid(h)), dim3(256), 0, h->stream, a);
  else
    hipLaunchKernelGGL(mpm_velacc_kernel<float>, dim3(p2g_grid(h)), dim3(256), 0, h->stream, a);
}

// ---- FP32 fused step ------------------------------------------------------------

template <int BASIS, int TW, int TH>
static void launch_fused_t(hg_mpm *h, bool prime) {
  FusedArgs a;
  a.scenes = h->d_scenes;
  a.tile_scene = h->d_tile_scene;
  a.tile_list = h->tile_list;
  a.tile_count = h->tile_count + h->cur;
  a.src = h->buf[h->cur];
  a.dst = h->buf[1 - h->cur];
  a.perm = h->perm;
  a.cell_start = h->cell_start;
  a.gva = static_cast<const float4 *>(h->gva);
  a.acc = h->acc;
  a.nnode = h->nnode;
  a.key = h->key;
  a.slot = h->slot;
  a.counts = h->counts;
  a.totals = h->scan_totals + size_t(1 - h->cur) * hg_scan_totals_words(h->nscan);  // the keys of buf[1 - cur]
  a.err = h->err;
  ProfScope ps(h->prof, prime ? "mpm_raster" : "mpm_fused", h->stream);
  const dim3 g(p2g_grid(h));
  if (h->ppass == 2) {
    const dim3 b(fused_threads<TW, TH, 2>());
    if (prime) hipLaunchKernelGGL((mpm_fused_kernel<BASIS, TW, TH, true, 2>), g, b, 0, h->stream, a);
    else if (h->fused_done)
      hipExtLaunchKernelGGL((mpm_fused_kernel<BASIS, TW, TH, false, 2>), g, b, 0, h->stream,
                            nullptr, h->fused_done, 0, a);
    else hipLaunchKernelGGL((mpm_fused_kernel<BASIS, TW, TH, false, 2>), g, b, 0, h->stream, a);
  } else {
    const dim3 b(fused_threads<TW, TH, 1>());
    if (prime) hipLaunchKernelGGL((mpm_fused_kernel<BASIS, TW, TH, true, 1>), g, b, 0, h->stream, a);
    else if (h->fused_done)
      hipExtLaunchKernelGGL((mpm_fused_kernel<BASIS, TW, TH, false, 1>), g, b, 0, h->stream,
                            nullptr, h->fused_done, 0, a);
    else hipLaunchKernelGGL((mpm_fused_kernel<BASIS, TW, TH, false, 1>), g, b, 0, h->stream, a);
  }
}

template <int BASIS>
static void launch_fused_b(hg_mpm *h, bool prime) {
  if (h->tw() == 32 && h->th() == 8) launch_fused_t<BASIS, 32, 8>(h, prime);
  else if (h->tw() == 16 && h->th() == 8) launch_fused_t<BASIS, 16, 8>(h, prime);
  else if (h->tw() == 64 && h->th() == 4) launch_fused_t<BASIS, 64, 4>(h, prime);
  else launch_fused_t<BASIS, 16, 16>(h, prime);
}

// The fused kernel over the current sort: prime rasterises the points of
// buf[cur] into the node sums; otherwise G2P moves them into buf[1 - cur]
// (keyed for the next sort) and rasterises them there.
static void launch_fused(hg_mpm *h, bool prime) {
  switch (h->basis) {
    case HG_BASIS_LINEAR: launch_fused_b<HG_BASIS_LINEAR>(h, prime); break;
    case HG_BASIS_CUBIC: launch_fused_b<HG_BASIS_CUBIC>(h, prime); break;
    default: launch_fused_b<HG_BASIS_UGIMP_LINEAR>(h, prime); break;
  }
}

// The grid kernel over the current active tiles: node sums -> grid update,
// sums cleared.
static void launch_grid(hg_mpm *h, uint32_t store, const GrainNodes *ex = nullptr,
                        hipEvent_t done = nullptr) {
  GridArgs a;
  a.scenes = h->d_scenes;
  a.tile_scene = h->d_tile_scene;
  a.tile_list = h->tile_list;
  a.tile_count = h->tile_count + h->cur;
  a.acc = h->acc;
  a.nnode = h->nnode;
  a.gva = h->gva;
  a.gm = h->gm;
  a.gp = h->gp;
  a.gpn = h->gpn;
  a.gf = h->gf;
  a.store = store;
  a.err = h->err;
  a.tw = h->tw();
  a.th = h->th();
  a.exchange = ex ? 1 : 0;
  if (ex) a.ex = *ex;
  else a.ex = GrainNodes{nullptr, nullptr, nullptr, nullptr};
  ProfScope ps(h->prof, "mpm_grid", h->stream);
  if (done)
    hipExtLaunchKernelGGL(mpm_grid_kernel, dim3(p2g_grid(h)), dim3(256), 0, h->stream, nullptr,
                          done, 0, a);
  else
    hipLaunchKernelGGL(mpm_grid_kernel, dim3(p2g_grid(h)), dim3(256), 0, h->stream, a);
  h->acc_ready = false;
  h->grid_ready = true;
  h->grid_store = store;
  h->last_store = store;
}

static bool fused_mode(const hg_mpm *h) { return h->acc && !h->keep_all; }

// Consumes node sums still pending (they would otherwise leak into a later
// rasterisation) and marks the grid as no longer the current points'.  Called
// before anything that changes the points or the grid outside the fused step.
static void fused_drain(hg_mpm *h) {
  if (!h->acc) return;
  if (h->acc_ready) launch_grid(h, kStoreVelAcc);
  h->grid_ready = false;
}

// The node (v, a) of the current points, with at least the fields in store.
static void fused_ensure_grid(hg_mpm *h, uint32_t store) {
  if (h->grid_ready && (h->grid_store & store) == store) return;
  if (!h->acc_ready) launch_fused(h, true);
  launch_grid(h, store);
}

// One FP32 step: G2P (with the next step's rasterisation), the sort, the
// grid update of the next step.
static void fused_step(hg_mpm *h) {
  fused_ensure_grid(h, kStoreVelAcc);
  launch_fused(h, false);
  h->cur = 1 - h->cur;
  launch_sort(h);
  launch_grid(h, kStoreVelAcc);
}

// (The node sums are all zero whenever acc_ready is false: only the grid
// kernel clears acc_ready, and it clears every sum of the active tiles,
// which hold all nonzero sums.)
void hg_mpm_prime(hg_mpm *h) {
  if (fused_mode(h) && !h->acc_ready) {
    launch_fused(h, true);
    h->acc_ready = true;
  }
}

static hg_status upload_dt(hg_mpm *h, double dt) {
  if (dt == h->dt_dev) return HG_OK;
  // the table is read by queued kernels: wait for them before rewriting it
  HG_HIP(h->ctx, hipStreamSynchronize(h->stream));
  for (SceneDev &s : h->scenes) s.dt = dt;
  HG_HIP(h->ctx, hipMemcpy(h->d_scenes, h->scenes.data(), h->scenes.size() * sizeof(SceneDev),
                           hipMemcpyHostToDevice));
  h->dt_dev = dt;
  // the node (v, a) depend on dt; the node sums do not
  h->grid_ready = false;
  if (h->graph) {
    hipGraphExecDestroy(h->graph);
    h->graph = nullptr;
  }
  return HG_OK;
}

static uint32_t fused_store(const hg_mpm *h) {
  return h->keep_all ? (kStoreVelAcc | kStoreMass | kStoreMom | kStoreMomNew | kStoreForce)
                     : kStoreVelAcc;
}

hg_status hg_mpm_launch_coupled_p2g(hg_mpm *h, double dt) {
  HG_TRY(upload_dt(h, dt));
  const uint32_t store = kStoreVelAcc | kStoreMass | kStoreMom | kStoreMomNew |
                         (h->keep_all ? kStoreForce : 0u);
  if (fused_mode(h)) {
    // the sums of the previous step's fused pass, or a rasterisation now
    if (!h->acc_ready) launch_fused(h, true);
    launch_grid(h, store);
    return HG_OK;
  }
  fused_drain(h);
  launch_p2g(h, store);
  return HG_OK;
}

bool hg_mpm_is_fused(const hg_mpm *h) { return fused_mode(h); }

hg_status hg_mpm_prepare_fused(hg_mpm *h, double dt) {
  HG_TRY(upload_dt(h, dt));
  hg_mpm_prime(h);
  return HG_OK;
}

hg_status hg_mpm_launch_coupled_grid(hg_mpm *h, double dt, const GrainNodes *g, bool keep_mass,
                                     hipEvent_t done) {
  HG_TRY(upload_dt(h, dt));
  // the exchange works on the values in hand; the node mass is kept for a
  // reader after the step's grid update
  const uint32_t store = kStoreVelAcc | (keep_mass ? kStoreMass : 0u);
  if (!h->acc_ready) launch_fused(h, true);
  launch_grid(h, store, g, done);
  return HG_OK;
}

void hg_mpm_launch_fused_g2p(hg_mpm *h, hipEvent_t done) {
  h->fused_done = done;
  launch_fused(h, false);
  h->fused_done = nullptr;
  h->cur = 1 - h->cur;
  h->acc_ready = true;
  h->grid_ready = false;
}

void hg_mpm_launch_fused_sort(hg_mpm *h) { launch_sort(h); }

const unsigned long long *hg_mpm_mass_sums(const hg_mpm *h) {
  return h->acc + size_t(kFxMass) * h->nnode;
}
double hg_mpm_mass_quantum(const hg_mpm *h) { return h->scenes[0].fx_quantum[kFxMass]; }

void hg_mpm_launch_g2p(hg_mpm *h) {
  if (fused_mode(h)) {
    launch_fused(h, false);
    h->cur = 1 - h->cur;
    launch_sort(h);
    h->acc_ready = true;
    h->grid_ready = false;
    return;
  }
  launch_g2p_and_sort(h);
}

// Bins the points of buf[cur] afresh (their block totals start from zero)
// and sorts them.
static void launch_rebin(hg_mpm *h) {
  fused_drain(h);
  hg_zero_words(h->stream, h->scan_totals + size_t(h->cur) * hg_scan_totals_words(h->nscan),
                hg_scan_totals_words(h->nscan));
  // the list is rebuilt at the same parity: its counter holds the previous
  // list's length, which the tile-list kernel would append to
  hg_zero_words(h->stream, h->tile_count + h->cur, 1);
  if (h->npoints) {
    if (h->precision == HG_FP64) launch_bin_p<double>(h);
    else launch_bin_p<float>(h);
  }
  launch_sort(h);
}

hg_status hg_mpm_rebin(hg_mpm *h) {
  launch_rebin(h);
  HG_HIP(h->ctx, hipGetLastError());
  HG_HIP(h->ctx, hipStreamSynchronize(h->stream));
  return HG_OK;
}

static void launch_step(hg_mpm *h) {
  if (fused_mode(h)) {
    fused_step(h);
    return;
  }
  fused_drain(h);
  launch_p2g(h, fused_store(h));
  launch_g2p_and_sort(h);
}

// ---- creation -----------------------------------------------------------------

static hg_status validate_desc(hg_ctx *ctx, const hg_mpm_desc &d) {
  if (d.precision != HG_FP64 && d.precision != HG_FP32)
    return hg_fail(ctx, HG_ERR_INVALID, "precision must be HG_FP64 or HG_FP32");
  if (d.basis != HG_BASIS_UGIMP_LINEAR && d.basis != HG_BASIS_LINEAR && d.basis != HG_BASIS_CUBIC)
    return hg_fail(ctx, HG_ERR_INVALID, "unknown basis");
  if (!(d.cell_width > 0.0)) return hg_fail(ctx, HG_ERR_INVALID, "cell_width must be > 0");
  if (d.cell_count[0] < 4 || d.cell_count[1] < 4)
    return hg_fail(ctx, HG_ERR_INVALID, "grid needs at least 4 cells per axis");
  if (uint64_t(d.cell_count[0] + 2) * uint64_t(d.cell_count[1] + 2) > (1ull << 31))
    return hg_fail(ctx, HG_ERR_UNSUPPORTED, "grid too large");
  if (d.nplanes > uint32_t(kMaxPlanes))
    return hg_fail(ctx, HG_ERR_UNSUPPORTED, "at most 8 static planes per scene");
  if (d.nplanes && (!d.plane_x || !d.plane_n || !d.plane_behavior))
    return hg_fail(ctx, HG_ERR_INVALID, "plane arrays missing");
  if (d.npoints && (!d.q || !d.v || !d.m || !d.volume || !d.hl || !d.be_bar || !d.J))
    return hg_fail(ctx, HG_ERR_INVALID, "point arrays missing");
  for (uint32_t i = 0; i < d.npoints; ++i) {
    if (!(d.m[i] > 0.0) || !(d.volume[i] > 0.0) || !(d.hl[i] > 0.0))
      return hg_fail(ctx, HG_ERR_INVALID, "point mass, volume and hl must be > 0");
    if (d.basis == HG_BASIS_UGIMP_LINEAR && !(2.0 * d.hl[i] < d.cell_width))
      return hg_fail(ctx, HG_ERR_INVALID, "uGIMP needs hl < cell_width/2");
    if (!(d.J[i] != 0.0)) return hg_fail(ctx, HG_ERR_INVALID, "J must be non-zero");
  }
  return HG_OK;
}

template <typename T>
static void fill_points(const hg_mpm_desc &d, uint32_t id0, std::vector<double> &x,
                        std::vector<double> &y, std::vector<T> (&f)[9],
                        std::vector<uint32_t> &ids, std::vector<uint8_t> &flags) {
  for (uint32_t i = 0; i < d.npoints; ++i) {
    x.push_back(d.q[2 * i]);
    y.push_back(d.q[2 * i + 1]);
    f[0].push_back(T(d.v[2 * i]));
    f[1].push_back(T(d.v[2 * i + 1]));
    f[2].push_back(strain_store<T>(d.be_bar[4 * i]));
    f[3].push_back(T(0.5 * (d.be_bar[4 * i + 1] + d.be_bar[4 * i + 2])));
    f[4].push_back(strain_store<T>(d.be_bar[4 * i + 3]));
    f[5].push_back(strain_store<T>(d.J[i]));
    f[6].push_back(T(d.m[i]));
    f[7].push_back(T(d.volume[i]));
    f[8].push_back(T(d.hl[i]));
    ids.push_back(id0 + i);
    flags.push_back(d.kinematically_scripted ? d.kinematically_scripted[i] : 0);
  }
}

template <typename T>
static hg_status upload_points(hg_mpm *h, const hg_mpm_desc *descs) {
  std::vector<double> x, y;
  std::vector<T> f[9];
  std::vector<uint32_t> ids;
  std::vector<uint8_t> flags;
  for (uint32_t s = 0; s < h->nscenes; ++s) fill_points<T>(descs[s], h->scenes[s].p_off, x, y, f, ids, flags);
  const uint32_t N = h->npoints;
  for (int b = 0; b < 2; ++b) {
    PtsRaw &p = h->buf[b];
    HG_TRY(dalloc(h, &p.x, N * 8));
    HG_TRY(dalloc(h, &p.y, N * 8));
    void **fs[9] = {&p.vx, &p.vy, &p.b00, &p.b01, &p.b11, &p.J, &p.m, &p.V, &p.hl};
    for (int k = 0; k < 9; ++k) {
      if ((k == 8 && !h->has_hl) || ((k == 6 || k == 7) && !h->has_mv)) {
        *fs[k] = nullptr;
        continue;
      }
      // mass, volume and half width are indexed by point id: one copy that
      // both buffers share
      if (k >= 6 && b == 1) {
        void *const shared[3] = {h->buf[0].m, h->buf[0].V, h->buf[0].hl};
        *fs[k] = shared[k - 6];
        continue;
      }
      HG_TRY(dalloc(h, fs[k], N * sizeof(T)));
    }
    HG_TRY(dalloc(h, &p.id, N * 4));
    if (h->has_flags) HG_TRY(dalloc(h, &p.flags, N));
    else p.flags = nullptr;
  }
  PtsRaw &p = h->buf[0];
  if (N) {
    HG_HIP(h->ctx, hipMemcpy(p.x, x.data(), N * 8, hipMemcpyHostToDevice));
    HG_HIP(h->ctx, hipMemcpy(p.y, y.data(), N * 8, hipMemcpyHostToDevice));
    void *fs[9] = {p.vx, p.vy, p.b00, p.b01, p.b11, p.J, p.m, p.V, p.hl};
    for (int k = 0; k < 9; ++k)
      if (fs[k]) HG_HIP(h->ctx, hipMemcpy(fs[k], f[k].data(), N * sizeof(T), hipMemcpyHostToDevice));
    HG_HIP(h->ctx, hipMemcpy(p.id, ids.data(), N * 4, hipMemcpyHostToDevice));
    if (p.flags) HG_HIP(h->ctx, hipMemcpy(p.flags, flags.data(), N, hipMemcpyHostToDevice));
  }
  return HG_OK;
}

// Fixed-point scales of a scene's node sums (FP32 step).  Each quantity has
// a unit that bounds one point's contribution at a node: mass m_max;
// momentum m_max V with V the larger of the elastic wave speed
// sqrt((kappa + 4 mu / 3) / rho_min) and the speed of a fall across the
// domain; internal force V0_max (kappa + mu) / h (the return map keeps the
// Kirchhoff stress within about kappa, and |grad w| h stays below 2).  The
// quantum is 2^-40 of the unit rounded up to a power of two: a term stays
// below 2^50 quanta (2^10 units) and a node sum below 2^62 (2^22 units).
static void plan_fixed_point(const hg_mpm_desc &d, SceneDev &sd) {
  double mmax = 0.0, vmax = 0.0, rho_min = INFINITY;
  for (uint32_t i = 0; i < d.npoints; ++i) {
    mmax = std::max(mmax, d.m[i]);
    vmax = std::max(vmax, d.volume[i]);
    rho_min = std::min(rho_min, d.m[i] / d.volume[i]);
  }
  const double wave = (rho_min > 0.0 && std::isfinite(rho_min))
                          ? std::sqrt(std::max(0.0, d.bulk_modulus + 4.0 / 3.0 * d.shear_modulus) / rho_min)
                          : 0.0;
  const double L = d.cell_width * double(std::max(d.cell_count[0], d.cell_count[1]));
  const double fall = std::sqrt(2.0 * std::hypot(d.gravity[0], d.gravity[1]) * L);
  double V = std::max(wave, fall);
  if (!(V > 0.0) || !std::isfinite(V)) V = 1.0;
  double fu = vmax * (d.bulk_modulus + d.shear_modulus) / d.cell_width;
  if (!(fu > 0.0) || !std::isfinite(fu)) fu = mmax > 0.0 ? mmax : 1.0;
  const double unit[kFxQuantities] = {mmax > 0.0 ? mmax : 1.0, (mmax > 0.0 ? mmax : 1.0) * V,
                                      (mmax > 0.0 ? mmax : 1.0) * V, fu, fu};
  for (int q = 0; q < kFxQuantities; ++q) {
    int e = 0;
    std::frexp(unit[q], &e);  // unit <= 2^e
    sd.fx_shift[q] = 40 - e;
    sd.fx_quantum[q] = std::ldexp(1.0, e - 40);
  }
}

static hg_status mpm_create_impl(hg_ctx *ctx, const hg_mpm_desc *descs, uint32_t nscenes,
                                 hg_mpm **out) {
  if (!ctx || !descs || !out || nscenes == 0) return hg_fail(ctx, HG_ERR_INVALID, "null argument");
  for (uint32_t s = 0; s < nscenes; ++s) {
    HG_TRY(validate_desc(ctx, descs[s]));
    if (descs[s].precision != descs[0].precision || descs[s].basis != descs[0].basis)
      return hg_fail(ctx, HG_ERR_INVALID, "all scenes of a batch share precision and basis");
  }
  HG_HIP(ctx, hipSetDevice(ctx->device));
  std::unique_ptr<hg_mpm> h(new hg_mpm());
  h->ctx = ctx;
  h->precision = descs[0].precision;
  h->basis = descs[0].basis;
  h->nscenes = nscenes;
  h->hl_first = descs[0].npoints ? descs[0].hl[0] : 0.25 * descs[0].cell_width;
  HG_HIP(ctx, hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));

  // scene tables
  const int S = stencil_width(h->basis);
  const int TW = h->tw(), TH = h->th();
  uint64_t p_off = 0, cell_off = 0, node_off = 0, tile_off = 0;
  std::vector<uint32_t> h_p_off;
  std::vector<uint32_t> tile_scene;
  for (uint32_t s = 0; s < nscenes; ++s) {
    const hg_mpm_desc &d = descs[s];
    SceneDev sd;
    std::memset(&sd, 0, sizeof(sd));
    sd.min0 = d.grid_min[0];
    sd.min1 = d.grid_min[1];
    sd.h = d.cell_width;
    sd.inv_h = 1.0 / d.cell_width;
    sd.mu = d.shear_modulus;
    sd.kappa = d.bulk_modulus;
    sd.dp_alpha = d.dp_alpha;
    sd.pic_alpha = d.pic_alpha;
    sd.g0 = d.gravity[0];
    sd.g1 = d.gravity[1];
    sd.dt = 0.0;
    sd.hl_uniform = d.npoints ? d.hl[0] : 0.25 * d.cell_width;
    sd.m_uniform = d.npoints ? d.m[0] : 0.0;
    sd.V_uniform = d.npoints ? d.volume[0] : 0.0;
    for (uint32_t i = 0; i < d.npoints; ++i) {
      if (d.hl[i] != d.hl[0]) h->has_hl = true;
      if (d.m[i] != d.m[0] || d.volume[i] != d.volume[0]) h->has_mv = true;
      if (d.kinematically_scripted && d.kinematically_scripted[i]) h->has_flags = true;
    }
    sd.nnx = int32_t(d.cell_count[0] + 1);
    sd.nny = int32_t(d.cell_count[1] + 1);
    sd.snx = sd.nnx + 1;
    sd.sny = sd.nny + 1;
    sd.tiles_x = (sd.nnx + TW - 1) / TW;
    sd.tiles_y = (sd.nny + TH - 1) / TH;
    sd.p_off = uint32_t(p_off);
    sd.p_cnt = d.npoints;
    sd.cell_off = uint32_t(cell_off);
    sd.node_off = uint32_t(node_off);
    sd.tile_off = uint32_t(tile_off);
    sd.nplanes = int32_t(d.nplanes);
    plan_fixed_point(d, sd);
    for (uint32_t k = 0; k < d.nplanes; ++k) {
      const double n0 = d.plane_n[2 * k], n1 = d.plane_n[2 * k + 1];
      const double nn = std::sqrt(n0 * n0 + n1 * n1);
      if (!(nn > 0.0)) return hg_fail(ctx, HG_ERR_INVALID, "plane normal is zero");
      sd.plane_x[2 * k] = d.plane_x[2 * k];
      sd.plane_x[2 * k + 1] = d.plane_x[2 * k + 1];
      sd.plane_n[2 * k] = n0 / nn;
      sd.plane_n[2 * k + 1] = n1 / nn;
      sd.plane_lower[k] = d.plane_lower ? d.plane_lower[k] : -INFINITY;
      sd.plane_beh[k] = d.plane_behavior[k];
    }
    h_p_off.push_back(uint32_t(p_off));
    for (int t = 0; t < sd.tiles_x * sd.tiles_y; ++t) tile_scene.push_back(s);
    p_off += d.npoints;
    cell_off += uint64_t(sd.snx) * uint64_t(sd.sny);
    node_off += uint64_t(sd.snx) * uint64_t(sd.sny);
    tile_off += uint64_t(sd.tiles_x) * uint64_t(sd.tiles_y);
    h->scenes.push_back(sd);
  }
  (void)S;
  if (p_off >= (1ull << 32) || cell_off >= (1ull << 31) || node_off >= (1ull << 31))
    return hg_fail(ctx, HG_ERR_UNSUPPORTED, "batch too large for 32-bit indexing");
  h_p_off.push_back(uint32_t(p_off));
  h->npoints = uint32_t(p_off);
  h->ncell = uint32_t(cell_off);
  h->nnode = uint32_t(node_off);
  h->ntiles = uint32_t(tile_off);
  h->nscan = (h->ncell + kScanChunk - 1) / kScanChunk;
  if (h->nscan > uint32_t(kScanChunk))
    return hg_fail(ctx, HG_ERR_UNSUPPORTED, "too many cells for the two-level scan");

  hg_mpm *hp = h.get();
  HG_TRY(dalloc(hp, &h->d_scenes, nscenes * sizeof(SceneDev)));
  HG_TRY(dalloc(hp, &h->d_p_off, (nscenes + 1) * 4));
  HG_TRY(dalloc(hp, &h->d_tile_scene, h->ntiles * 4));
  HG_TRY(dalloc(hp, &h->tile_active, h->ntiles));
  HG_TRY(dalloc(hp, &h->tile_list, size_t(h->ntiles) * 4));
  HG_TRY(dalloc(hp, &h->tile_count, 8));
  HG_TRY(dalloc(hp, &h->perm, size_t(h->npoints) * 4));
  HG_TRY(dalloc(hp, &h->pinv, size_t(h->npoints) * 4));
  HG_TRY(dalloc(hp, &h->key, size_t(h->npoints) * 4));
  HG_TRY(dalloc(hp, &h->slot, size_t(h->npoints) * 4));
  HG_TRY(dalloc(hp, &h->counts, size_t(h->ncell) * 4));
  HG_TRY(dalloc(hp, &h->cell_start, size_t(h->ncell + 1) * 4));
  HG_TRY(dalloc(hp, &h->scan_totals, size_t(hg_scan_totals_words(h->nscan)) * 8));
  HG_TRY(dalloc(hp, &h->gva, size_t(h->nnode) * 4 * hp->tsize()));
  HG_TRY(dalloc(hp, &h->gm, size_t(h->nnode) * 8));
  HG_TRY(dalloc(hp, &h->gp, size_t(h->nnode) * 16));
  HG_TRY(dalloc(hp, &h->gpn, size_t(h->nnode) * 16));
  HG_TRY(dalloc(hp, &h->gf, size_t(h->nnode) * 16));
  HG_TRY(dalloc(hp, &h->err, nscenes * 4));
  if (h->precision == HG_FP32) {
    HG_TRY(dalloc(hp, &h->acc, size_t(h->nnode) * kFxQuantities * 8));
    HG_HIP(ctx, hipMemset(h->acc, 0, size_t(h->nnode) * kFxQuantities * 8));
  }
  HG_HIP(ctx, hipMemcpy(h->d_scenes, h->scenes.data(), nscenes * sizeof(SceneDev), hipMemcpyHostToDevice));
  HG_HIP(ctx, hipMemcpy(h->d_p_off, h_p_off.data(), (nscenes + 1) * 4, hipMemcpyHostToDevice));
  HG_HIP(ctx, hipMemcpy(h->d_tile_scene, tile_scene.data(), h->ntiles * 4, hipMemcpyHostToDevice));
  // Zero everything a kernel may read before writing: masked stencil nodes
  // are read with weight zero and must hold finite values.
  HG_HIP(ctx, hipMemset(h->tile_active, 0, h->ntiles));
  HG_HIP(ctx, hipMemset(h->tile_count, 0, 8));
  HG_HIP(ctx, hipMemset(h->counts, 0, size_t(h->ncell) * 4));
  HG_HIP(ctx, hipMemset(h->scan_totals, 0, size_t(hg_scan_totals_words(h->nscan)) * 8));
  HG_HIP(ctx, hipMemset(h->perm, 0, size_t(h->npoints) * 4));
  HG_HIP(ctx, hipMemset(h->gva, 0, size_t(h->nnode) * 4 * hp->tsize()));
  HG_HIP(ctx, hipMemset(h->gm, 0, size_t(h->nnode) * 8));
  HG_HIP(ctx, hipMemset(h->gp, 0, size_t(h->nnode) * 16));
  HG_HIP(ctx, hipMemset(h->gpn, 0, size_t(h->nnode) * 16));
  HG_HIP(ctx, hipMemset(h->gf, 0, size_t(h->nnode) * 16));
  HG_HIP(ctx, hipMemset(h->err, 0, nscenes * 4));
  if (h->precision == HG_FP64) HG_TRY(upload_points<double>(hp, descs));
  else HG_TRY(upload_points<float>(hp, descs));
  // the memsets and copies above run on the null stream; the engine's
  // non-blocking stream does not wait for it, so finish them before the first
  // kernel (without waiting for other handles' streams)
  HG_HIP(ctx, hipStreamSynchronize(nullptr));

  // initial sort
  launch_rebin(hp);
  HG_HIP(ctx, hipGetLastError());
  HG_HIP(ctx, hipStreamSynchronize(h->stream));
  std::vector<int32_t> e(nscenes);
  HG_HIP(ctx, hipMemcpy(e.data(), h->err, nscenes * 4, hipMemcpyDeviceToHost));
  for (uint32_t s = 0; s < nscenes; ++s)
    if (e[s]) return hg_fail(ctx, HG_ERR_INVALID, "scene " + std::to_string(s) + ": a point's stencil is off the grid");
  *out = h.release();
  return HG_OK;
}

static void mpm_free(hg_mpm *h) {
  if (!h) return;
  if (h->stream) hipStreamSynchronize(h->stream);
  if (h->graph) hipGraphExecDestroy(h->graph);
  for (void *p : h->allocs) hipFree(p);
  if (h->stream && h->own_stream) hipStreamDestroy(h->stream);
  delete h;
}

static hg_status check_err(hg_mpm *h) {
  HG_HIP(h->ctx, hipStreamSynchronize(h->stream));
  std::vector<int32_t> e(h->nscenes);
  HG_HIP(h->ctx, hipMemcpy(e.data(), h->err, h->nscenes * 4, hipMemcpyDeviceToHost));
  for (uint32_t s = 0; s < h->nscenes; ++s) {
    if (e[s] & kSortTotalsBit) return hg_fail(h->ctx, HG_ERR_STATE, "cell sort: block totals do not match the counts");
    if (e[s] & 1) return hg_fail(h->ctx, HG_ERR_STATE, "scene " + std::to_string(s) + ": a material point left the grid");
    if (e[s] & 2) return hg_fail(h->ctx, HG_ERR_STATE, "scene " + std::to_string(s) + ": non-finite material point state");
    if (e[s] & kMpmFixedRangeBit)
      return hg_fail(h->ctx, HG_ERR_STATE, "scene " + std::to_string(s) + ": node sums outside the fixed-point range");
  }
  return HG_OK;
}

// ---- C-ABI --------------------------------------------------------------------

extern "C" hg_status hg_mpm_create(hg_ctx *ctx, const hg_mpm_desc *desc, hg_mpm **out) {
  return mpm_create_impl(ctx, desc, 1, out);
}

extern "C" hg_status hg_mpm_create_batch(hg_ctx *ctx, const hg_mpm_desc *descs, uint32_t nscenes,
                                         hg_mpm **out) {
  return mpm_create_impl(ctx, descs, nscenes, out);
}

extern "C" hg_status hg_mpm_destroy(hg_mpm *h) {
  mpm_free(h);
  return HG_OK;
}

extern "C" hg_status hg_mpm_step(hg_mpm *h, double dt, uint32_t nsteps) {
  if (!h) return HG_ERR_INVALID;
  if (!(dt > 0.0)) return hg_fail(h->ctx, HG_ERR_INVALID, "dt must be > 0");
  if (h->phase_next != 0) return hg_fail(h->ctx, HG_ERR_INVALID, "a phased step is in progress");
  HG_TRY(upload_dt(h, dt));
  for (uint32_t k = 0; k < nsteps; ++k) launch_step(h);
  HG_HIP(h->ctx, hipGetLastError());
  return HG_OK;
}

extern "C" hg_status hg_mpm_phase(hg_mpm *h, int32_t phase, double dt) {
  if (!h) return HG_ERR_INVALID;
  if (phase != h->phase_next)
    return hg_fail(h->ctx, HG_ERR_INVALID, "phases must run in the order 0, 1, 2");
  if (!(dt > 0.0) && phase != 0) return hg_fail(h->ctx, HG_ERR_INVALID, "dt must be > 0");
  if (phase == 0) {
    // Mass and momentum; the forces and the grid update need only the point
    // state this phase leaves unchanged, so they are computed in the same pass
    // and kept until phase 1 hands them over.
    fused_drain(h);
    h->phase_next = 1;
    HG_TRY(upload_dt(h, dt > 0.0 ? dt : (h->dt_dev > 0.0 ? h->dt_dev : 1.0)));
    launch_p2g(h, kStoreMass | kStoreMom | kStoreMomNew | kStoreForce);
  } else if (phase == 1) {
    if (dt != h->dt_dev) {
      // dt enters the momentum update: redo the grid pass with this dt
      HG_TRY(upload_dt(h, dt));
      launch_p2g(h, kStoreMass | kStoreMom | kStoreMomNew | kStoreForce);
    }
    h->phase_next = 2;
  } else {
    if (dt != h->dt_dev) return hg_fail(h->ctx, HG_ERR_INVALID, "phase 2 dt differs from phase 1 dt");
    launch_velacc(h);
    h->last_store |= kStoreVelAcc;
    launch_g2p_and_sort(h);
    h->phase_next = 0;
  }
  HG_HIP(h->ctx, hipGetLastError());
  return HG_OK;
}

extern "C" hg_status hg_mpm_sync(hg_mpm *h) {
  if (!h) return HG_ERR_INVALID;
  return check_err(h);
}

extern "C" hg_status hg_mpm_scene_error(hg_mpm *h, uint32_t scene, int32_t *flags) {
  if (!h || scene >= h->nscenes || !flags) return HG_ERR_INVALID;
  HG_HIP(h->ctx, hipStreamSynchronize(h->stream));
  HG_HIP(h->ctx, hipMemcpy(flags, h->err + scene, 4, hipMemcpyDeviceToHost));
  return HG_OK;
}

template <typename T>
static hg_status download_points_t(hg_mpm *h, uint32_t scene, double *q, double *v,
                                   double *be_bar, double *J) {
  const SceneDev &sc = h->scenes[scene];
  const uint32_t n = sc.p_cnt, off = sc.p_off;
  if (n == 0) return HG_OK;
  const PtsRaw &p = h->buf[h->cur];
  std::vector<double> x(n), y(n);
  std::vector<T> f[6];
  std::vector<uint32_t> id(n);
  HG_HIP(h->ctx, hipMemcpy(x.data(), p.x + off, n * 8, hipMemcpyDeviceToHost));
  HG_HIP(h->ctx, hipMemcpy(y.data(), p.y + off, n * 8, hipMemcpyDeviceToHost));
  HG_HIP(h->ctx, hipMemcpy(id.data(), p.id + off, n * 4, hipMemcpyDeviceToHost));
  const void *fs[6] = {p.vx, p.vy, p.b00, p.b01, p.b11, p.J};
  for (int k = 0; k < 6; ++k) {
    f[k].resize(n);
    HG_HIP(h->ctx, hipMemcpy(f[k].data(), static_cast<const T *>(fs[k]) + off, n * sizeof(T),
                             hipMemcpyDeviceToHost));
  }
  for (uint32_t s = 0; s < n; ++s) {
    const uint32_t i = id[s] - off;
    if (i >= n) return hg_fail(h->ctx, HG_ERR_STATE, "corrupt point ids");
    if (q) { q[2 * i] = x[s]; q[2 * i + 1] = y[s]; }
    if (v) { v[2 * i] = double(f[0][s]); v[2 * i + 1] = double(f[1][s]); }
    if (be_bar) {
      be_bar[4 * i] = strain_load<T>(f[2][s]);
      be_bar[4 * i + 1] = double(f[3][s]);
      be_bar[4 * i + 2] = double(f[3][s]);
      be_bar[4 * i + 3] = strain_load<T>(f[4][s]);
    }
    if (J) J[i] = strain_load<T>(f[5][s]);
  }
  return HG_OK;
}

extern "C" hg_status hg_mpm_download_points_scene(hg_mpm *h, uint32_t scene, double *q, double *v,
                                                  double *be_bar, double *J) {
  if (!h || scene >= h->nscenes) return HG_ERR_INVALID;
  if (h->phase_next != 0) return hg_fail(h->ctx, HG_ERR_INVALID, "a phased step is in progress");
  HG_TRY(check_err(h));
  if (h->precision == HG_FP64) return download_points_t<double>(h, scene, q, v, be_bar, J);
  return download_points_t<float>(h, scene, q, v, be_bar, J);
}

extern "C" hg_status hg_mpm_download_points(hg_mpm *h, double *q, double *v, double *be_bar,
                                            double *J) {
  return hg_mpm_download_points_scene(h, 0, q, v, be_bar, J);
}

extern "C" hg_status hg_mpm_set_keep_grid(hg_mpm *h, int32_t keep_all) {
  if (!h) return HG_ERR_INVALID;
  if ((keep_all != 0) != h->keep_all) fused_drain(h);
  h->keep_all = keep_all != 0;
  // the separate G2P reads pinv, which sorts for the fused step leave
  // unwritten: rebuilt here rather than inside a graph the next step captures
  if (h->keep_all && !h->pinv_ok && h->npoints) {
    hipLaunchKernelGGL(pinv_from_perm_kernel, dim3((h->npoints + 255) / 256), dim3(256), 0,
                       h->stream, h->perm, h->npoints, h->pinv);
    h->pinv_ok = true;
  }
  if (h->graph) {
    hipGraphExecDestroy(h->graph);
    h->graph = nullptr;
  }
  return HG_OK;
}

template <typename T>
static hg_status download_grid_t(hg_mpm *h, uint32_t scene, double *mass, double *mom,
                                 double *mom_new, double *force, double *vel, double *acc) {
  const SceneDev &sc = h->scenes[scene];
  const size_t ns = size_t(sc.snx) * size_t(sc.sny);
  const int TW = h->tw(), TH = h->th();
  std::vector<uint8_t> act(size_t(sc.tiles_x) * size_t(sc.tiles_y));
  HG_HIP(h->ctx, hipMemcpy(act.data(), h->tile_active + sc.tile_off, act.size(), hipMemcpyDeviceToHost));
  std::vector<double> gm(ns), gp(2 * ns), gpn(2 * ns), gf(2 * ns);
  std::vector<T> gva(4 * ns);
  const uint32_t st = h->last_store;
  if (st & kStoreMass) HG_HIP(h->ctx, hipMemcpy(gm.data(), h->gm + sc.node_off, ns * 8, hipMemcpyDeviceToHost));
  if (st & kStoreMom) HG_HIP(h->ctx, hipMemcpy(gp.data(), h->gp + 2 * size_t(sc.node_off), ns * 16, hipMemcpyDeviceToHost));
  if (st & kStoreMomNew) HG_HIP(h->ctx, hipMemcpy(gpn.data(), h->gpn + 2 * size_t(sc.node_off), ns * 16, hipMemcpyDeviceToHost));
  if (st & kStoreForce) HG_HIP(h->ctx, hipMemcpy(gf.data(), h->gf + 2 * size_t(sc.node_off), ns * 16, hipMemcpyDeviceToHost));
  if (st & kStoreVelAcc)
    HG_HIP(h->ctx, hipMemcpy(gva.data(), static_cast<const T *>(h->gva) + 4 * size_t(sc.node_off),
                             ns * 4 * sizeof(T), hipMemcpyDeviceToHost));
  for (int y = 0; y < sc.nny; ++y)
    for (int x = 0; x < sc.nnx; ++x) {
      const size_t r = size_t(y) * size_t(sc.nnx) + size_t(x);
      const size_t s = size_t(y) * size_t(sc.snx) + size_t(x);
      const bool on = act[size_t(y / TH) * size_t(sc.tiles_x) + size_t(x / TW)] != 0;
      if (mass) mass[r] = (on && (st & kStoreMass)) ? gm[s] : 0.0;
      for (int c = 0; c < 2; ++c) {
        if (mom) mom[2 * r + c] = (on && (st & kStoreMom)) ? gp[2 * s + c] : 0.0;
        if (mom_new) mom_new[2 * r + c] = (on && (st & kStoreMomNew)) ? gpn[2 * s + c] : 0.0;
        if (force) force[2 * r + c] = (on && (st & kStoreForce)) ? gf[2 * s + c] : 0.0;
        if (vel) vel[2 * r + c] = (on && (st & kStoreVelAcc)) ? double(gva[4 * s + c]) : 0.0;
        if (acc) acc[2 * r + c] = (on && (st & kStoreVelAcc)) ? double(gva[4 * s + 2 + c]) : 0.0;
      }
    }
  return HG_OK;
}

extern "C" hg_status hg_mpm_download_grid_scene(hg_mpm *h, uint32_t scene, double *mass, double *mom,
                                                double *mom_new, double *force, double *vel,
                                                double *acc) {
  if (!h || scene >= h->nscenes) return HG_ERR_INVALID;
  HG_HIP(h->ctx, hipStreamSynchronize(h->stream));
  if (h->precision == HG_FP64) return download_grid_t<double>(h, scene, mass, mom, mom_new, force, vel, acc);
  return download_grid_t<float>(h, scene, mass, mom, mom_new, force, vel, acc);
}

extern "C" hg_status hg_mpm_download_grid(hg_mpm *h, double *mass, double *mom, double *mom_new,
                                          double *force, double *vel, double *acc) {
  return hg_mpm_download_grid_scene(h, 0, mass, mom, mom_new, force, vel, acc);
}

extern "C" hg_status hg_mpm_counts(const hg_mpm *h, uint32_t *npoints, uint32_t *nnodes) {
  if (!h) return HG_ERR_INVALID;
  if (npoints) *npoints = h->scenes[0].p_cnt;
  if (nnodes) *nnodes = uint32_t(h->scenes[0].nnx) * uint32_t(h->scenes[0].nny);
  return HG_OK;
}

extern "C" hg_status hg_mpm_set_profiling(hg_mpm *h, int32_t on) {
  if (!h) return HG_ERR_INVALID;
  HG_HIP(h->ctx, hipStreamSynchronize(h->stream));
  h->prof.reset();
  h->prof.on = on != 0;
  return HG_OK;
}

extern "C" hg_status hg_mpm_hold(hg_mpm *h, double us) {
  if (!h) return HG_ERR_INVALID;
  return hg_stream_hold(h->ctx, h->stream, us);
}

extern "C" hg_status hg_mpm_kernel_times(hg_mpm *h, int32_t max_kernels, char *names, double *ms,
                                         int64_t *launches, int32_t *nkernels) {
  if (!h) return HG_ERR_INVALID;
  return hg_profiler_report(h->prof, h->ctx, max_kernels, names, ms, launches, nkernels);
}

extern "C" hg_status hg_mpm_set_points_per_pass(hg_mpm *h, int32_t points_per_pass) {
  if (!h) return HG_ERR_INVALID;
  if (points_per_pass != 1 && points_per_pass != 2)
    return hg_fail(h->ctx, HG_ERR_UNSUPPORTED, "points per pass must be 1 or 2 per cell");
  if (h->precision != HG_FP32) return HG_OK;  // the FP64 step has no such knob
  if (points_per_pass == h->ppass) return HG_OK;
  h->ppass = points_per_pass;
  if (h->graph) {
    hipGraphExecDestroy(h->graph);
    h->graph = nullptr;
  }
  return HG_OK;
}

extern "C" hg_status hg_mpm_tuning(const hg_mpm *h, int32_t *aw, int32_t *ah,
                                   int32_t *points_per_pass) {
  if (!h) return HG_ERR_INVALID;
  if (aw) *aw = h->aw;
  if (ah) *ah = h->ah;
  if (points_per_pass) *points_per_pass = h->precision == HG_FP32 ? h->ppass : 1;
  return HG_OK;
}

extern "C" hg_status hg_mpm_set_tile(hg_mpm *h, int32_t aw, int32_t ah) {
  if (!h) return HG_ERR_INVALID;
  if (!((aw == 16 && ah == 16) || (aw == 16 && ah == 8) || (aw == 32 && ah == 8) ||
        (aw == 64 && ah == 4)))
    return hg_fail(h->ctx, HG_ERR_UNSUPPORTED, "tile shape must be 16x16, 16x8, 32x8 or 64x4");
  if (h->phase_next != 0) return hg_fail(h->ctx, HG_ERR_INVALID, "a phased step is in progress");
  if (aw == h->aw && ah == h->ah) return HG_OK;
  // Tiles index the grid; rebuild the tile tables for the new shape.
  HG_HIP(h->ctx, hipStreamSynchronize(h->stream));
  h->aw = aw;
  h->ah = ah;
  const int TW = h->tw(), TH = h->th();
  uint64_t tile_off = 0;
  std::vector<uint32_t> tile_scene;
  for (uint32_t s = 0; s < h->nscenes; ++s) {
    SceneDev &sd = h->scenes[s];
    sd.tiles_x = (sd.nnx + TW - 1) / TW;
    sd.tiles_y = (sd.nny + TH - 1) / TH;
    sd.tile_off = uint32_t(tile_off);
    for (int t = 0; t < sd.tiles_x * sd.tiles_y; ++t) tile_scene.push_back(s);
    tile_off += uint64_t(sd.tiles_x) * uint64_t(sd.tiles_y);
  }
  h->ntiles = uint32_t(tile_off);
  void *nts = nullptr, *nta = nullptr, *ntl = nullptr;
  HG_HIP(h->ctx, hipMalloc(&nts, h->ntiles * 4));
  HG_HIP(h->ctx, hipMalloc(&nta, h->ntiles));
  HG_HIP(h->ctx, hipMalloc(&ntl, h->ntiles * 4));
  h->allocs.push_back(nts);
  h->allocs.push_back(nta);
  h->allocs.push_back(ntl);
  h->bytes += uint64_t(h->ntiles) * 9;
  h->d_tile_scene = static_cast<uint32_t *>(nts);
  h->tile_active = static_cast<uint8_t *>(nta);
  h->tile_list = static_cast<uint32_t *>(ntl);
  HG_HIP(h->ctx, hipMemcpy(h->d_tile_scene, tile_scene.data(), h->ntiles * 4, hipMemcpyHostToDevice));
  HG_HIP(h->ctx, hipMemset(h->tile_active, 0, h->ntiles));
  HG_HIP(h->ctx, hipMemcpy(h->d_scenes, h->scenes.data(), h->scenes.size() * sizeof(SceneDev),
                           hipMemcpyHostToDevice));
  if (h->graph) {
    hipGraphExecDestroy(h->graph);
    h->graph = nullptr;
  }
  // the active-tile list of the current sort, for the new shape
  HG_HIP(h->ctx, hipMemsetAsync(h->tile_count, 0, 8, h->stream));
  launch_tile_list(h, false);
  HG_HIP(h->ctx, hipStreamSynchronize(h->stream));
  return HG_OK;
}

extern "C" hg_status hg_mpm_step_graph(hg_mpm *h, double dt, uint32_t nsteps) {
  if (!h) return HG_ERR_INVALID;
  if (!(dt > 0.0)) return hg_fail(h->ctx, HG_ERR_INVALID, "dt must be > 0");
  if (h->phase_next != 0) return hg_fail(h->ctx, HG_ERR_INVALID, "a phased step is in progress");
  if (h->prof.on) return hg_mpm_step(h, dt, nsteps);  // events are not captured
  HG_TRY(upload_dt(h, dt));
  // every captured fused step expects the grid of its points ready
  if (fused_mode(h)) fused_ensure_grid(h, kStoreVelAcc);
  // The point buffers alternate every step, so one graph holds a pair of
  // steps that starts and ends on buffer 0.
  if (h->cur != 0 && nsteps) {
    launch_step(h);
    --nsteps;
  }
  const uint32_t pairs = nsteps / 2;
  if (pairs) {
    if (!h->graph) {
      hipGraph_t g = nullptr;
      HG_HIP(h->ctx, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
      launch_step(h);
      launch_step(h);
      HG_HIP(h->ctx, hipStreamEndCapture(h->stream, &g));
      HG_HIP(h->ctx, hipGraphInstantiate(&h->graph, g, nullptr, nullptr, 0));
      HG_HIP(h->ctx, hipGraphDestroy(g));
    }
    for (uint32_t k = 0; k < pairs; ++k) HG_HIP(h->ctx, hipGraphLaunch(h->graph, h->stream));
  }
  if (nsteps & 1u) launch_step(h);
  HG_HIP(h->ctx, hipGetLastError());
  return HG_OK;
}

extern "C" hg_status hg_mpm_estimate_bytes(const hg_mpm_desc *descs, uint32_t nscenes,
                                           uint64_t *bytes) {
  if (!descs || !bytes || nscenes == 0) return HG_ERR_INVALID;
  const uint64_t T = descs[0].precision == HG_FP64 ? 8 : 4;
  uint64_t N = 0, nodes = 0, tiles = 0;
  for (uint32_t s = 0; s < nscenes; ++s) {
    const hg_mpm_desc &d = descs[s];
    const uint64_t snx = uint64_t(d.cell_count[0]) + 2, sny = uint64_t(d.cell_count[1]) + 2;
    N += d.npoints;
    nodes += snx * sny;
    // the smallest tile (16x8) gives the most tiles
    tiles += ((snx + 15) / 16) * ((sny + 7) / 8);
  }
  const uint64_t per_point = 2 * (16 + 9 * T + 4 + 1) + 16;
  // cell lists, grid fields, and in FP32 the fixed-point node sums
  const uint64_t per_node = 8 + 4 * T + 8 + 16 * 3 + (T == 4 ? 8 * kFxQuantities : 0);
  const uint64_t nscan = (nodes + kScanChunk - 1) / kScanChunk;
  *bytes = N * per_point + nodes * per_node + 4 + nscan * 8 * kScanSubTotals + tiles * 5 +
           uint64_t(nscenes) * (sizeof(SceneDev) + 16) + (1u << 20);
  return HG_OK;
}

extern "C" hg_status hg_mpm_device_bytes(hg_mpm *h, uint64_t *bytes) {
  if (!h || !bytes) return HG_ERR_INVALID;
  *bytes = h->bytes;
  return HG_OK;
}

// Tile-shape autotuner: times the P2G pass (which reads the point state and
// writes only grid fields that every step recomputes), or in FP32 whole fused
// steps with one and with two points per cell per tile pass, for each
// supported tile shape on the handle's own scenes and keeps the fastest.  The choice changes no result: a node's sum order does not
// depend on the tile shape, and FP32 node sums are exact.
extern "C" hg_status hg_mpm_autotune(hg_mpm *h, double dt, int32_t reps, int32_t *aw_out,
                                     int32_t *ah_out, double *us /* [4] or NULL */) {
  if (!h) return HG_ERR_INVALID;
  if (!(dt > 0.0)) return hg_fail(h->ctx, HG_ERR_INVALID, "dt must be > 0");
  if (h->phase_next != 0) return hg_fail(h->ctx, HG_ERR_INVALID, "a phased step is in progress");
  if (reps < 1) reps = 1;
  HG_TRY(upload_dt(h, dt));
  const int shapes[4][2] = {{16, 16}, {16, 8}, {32, 8}, {64, 4}};
  const bool was_on = h->prof.on;
  h->prof.on = false;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  HG_HIP(h->ctx, hipEventCreate(&e0));
  HG_HIP(h->ctx, hipEventCreate(&e1));
  double best = 1e300;
  int bi = 0;
  hg_status st = HG_OK;
  // Fused FP32 handles time whole steps (G2P and rasterisation favour
  // different tiles) and then return to the state they started from: the
  // points and the error words are saved first and restored after, and a
  // restored handle continues bitwise as the uninterrupted one would.
  const bool fused = fused_mode(h);
  std::vector<uint8_t> saved;
  std::vector<int32_t> saved_err(h->nscenes);
  if (fused) {
    uint64_t nb = 0;
    HG_TRY(hg_mpm_snapshot(h, nullptr, &nb));
    saved.resize(nb);
    HG_TRY(hg_mpm_snapshot(h, saved.data(), &nb));
    HG_HIP(h->ctx, hipMemcpy(saved_err.data(), h->err, h->nscenes * 4, hipMemcpyDeviceToHost));
  }
  auto pass = [&]() {
    if (fused) launch_step(h);
    else launch_p2g(h, kStoreVelAcc);
  };
  // fused handles also try two points per cell per tile pass
  int bp = 1;
  const int npp = fused ? 2 : 1;
  for (int s = 0; s < 4 && st == HG_OK; ++s) {
    st = hg_mpm_set_tile(h, shapes[s][0], shapes[s][1]);
    if (st != HG_OK) break;
    double shape_best = 1e300;
    for (int pp = 1; pp <= npp && st == HG_OK; ++pp) {
      h->ppass = pp;
      pass();  // warm-up
      hipEventRecord(e0, h->stream);
      for (int r = 0; r < reps; ++r) pass();
      hipEventRecord(e1, h->stream);
      if (hipEventSynchronize(e1) != hipSuccess) {
        st = hg_fail(h->ctx, HG_ERR_HIP, "autotune timing failed");
        break;
      }
      float ms = 0.f;
      hipEventElapsedTime(&ms, e0, e1);
      const double t = 1e3 * double(ms) / reps;
      shape_best = std::min(shape_best, t);
      if (t < best) {
        best = t;
        bi = s;
        bp = pp;
      }
    }
    if (us) us[s] = shape_best;
  }
  h->ppass = bp;
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  h->prof.on = was_on;
  if (fused) {
    HG_TRY(hg_mpm_restore(h, saved.data(), saved.size()));
    HG_HIP(h->ctx, hipMemcpy(h->err, saved_err.data(), h->nscenes * 4, hipMemcpyHostToDevice));
  }
  HG_TRY(st);
  HG_TRY(hg_mpm_set_tile(h, shapes[bi][0], shapes[bi][1]));
  if (aw_out) *aw_out = shapes[bi][0];
  if (ah_out) *ah_out = shapes[bi][1];
  return HG_OK;
}
