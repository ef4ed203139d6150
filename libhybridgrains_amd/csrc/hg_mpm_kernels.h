// hg_mpm_kernels.h -- device side of the MPM explicit step (gfx950).
//
// Data layout in HBM (one batch = one or more independent scenes):
//  * points: SoA arrays concatenated over scenes (scene s owns
//    [p_off, p_off + p_cnt)); x,y are FP64 in every precision, the rest is T
//    (double for HG_FP64, float for HG_FP32); b_bar is stored as its three
//    independent entries (b00, b01, b11) because the step keeps it exactly
//    symmetric.  Two buffers: each step's G2P reads buffer A in its physical
//    order and writes each point to its slot of the current cell order in
//    buffer B, so the points are physically re-sorted every step at no extra
//    traffic (and stay nearly sorted, which keeps the next sort's gathers
//    coalesced).
//  * cells: one key per grid node that can be a stencil base; key = base node
//    (bx, by) as by*snx + bx, where snx = nnx + 1 pads one node column and row
//    so the masked third uGIMP node of a two-node stencil is always in bounds.
//  * grid: node storage with the same padded pitch, per scene at node_off.
//    The fused step writes only (v, a) per node as one T4; split/keep modes
//    also write FP64 mass, momentum, new momentum and force.
//
// Kernels of one step: p2g (node gather over the active tiles, also the whole
// grid update), g2p (gather + return map + advection + re-key), then the
// counting sort (scan, scatter) and the active-tile list
// (tile_list).
#pragma once

#include "hg_internal.h"

#include <type_traits>

namespace hgmpm {

constexpr int kMaxPlanes = 8;

enum : uint32_t {
  kStoreVelAcc = 1u,
  kStoreMass = 2u,
  kStoreMom = 4u,
  kStoreMomNew = 8u,
  kStoreForce = 16u,
};

// The FP32 step's node sums are exact: each point-node contribution, computed
// in FP32 and scaled by 2^shift (a per-scene power of two), is rounded to the
// nearest whole number of quanta and summed as a 64-bit integer, so the sum
// does not depend on the order in which the contributions arrive.  Quantities:
// node mass, the two momentum components, the two internal-force
// components.
enum : int { kFxMass = 0, kFxMom0 = 1, kFxMom1 = 2, kFxF0 = 3, kFxF1 = 4, kFxQuantities = 5 };
// MPM error bit: a contribution or a node sum outside the fixed-point range
// (never expected for physical states; the step's results are then void).
constexpr int32_t kMpmFixedRangeBit = 4;

struct SceneDev {
  double min0, min1, h;
  double mu, kappa, dp_alpha, pic_alpha, g0, g1;
  double dt;
  double hl_uniform;
  // every point's mass and volume where the points store none (all equal)
  double m_uniform, V_uniform;
  double inv_h;            // 1 / h
  int32_t nnx, nny;        // reference node counts (cell_count + 1)
  int32_t snx, sny;        // padded storage pitch
  int32_t tiles_x, tiles_y;
  uint32_t p_off, p_cnt;   // points
  uint32_t cell_off;       // keys
  uint32_t node_off;       // grid storage
  uint32_t tile_off;       // tiles
  int32_t nplanes;
  double plane_x[2 * kMaxPlanes];
  double plane_n[2 * kMaxPlanes];
  double plane_lower[kMaxPlanes];
  int32_t plane_beh[kMaxPlanes];
  // node sums of the FP32 step: contribution * 2^fx_shift[q] is summed in
  // whole quanta; fx_quantum[q] = 2^-fx_shift[q] converts back
  int32_t fx_shift[kFxQuantities];
  double fx_quantum[kFxQuantities];
};

// Stored strain.  FP32 mode keeps J - 1 and the diagonal of b_bar minus 1
// (b01 as it is), so that FP32 resolves the strain relative to its own size
// rather than to 1; FP64 mode keeps J and b_bar.  The conversions are exact in
// FP64 for the stored FP32 values.
template <typename T>
__host__ __device__ __forceinline__ double strain_load(T v) {
  if constexpr (std::is_same<T, float>::value) return 1.0 + double(v);
  else return v;
}
template <typename T>
__host__ __device__ __forceinline__ T strain_store(double v) {
  if constexpr (std::is_same<T, float>::value) return float(v - 1.0);
  else return v;
}

struct PtsRaw {
  double *x, *y;
  void *vx, *vy, *b00, *b01, *b11, *J, *m, *V, *hl;
  uint32_t *id;
  uint8_t *flags;
};

template <int BASIS> struct BasisTraits;
template <> struct BasisTraits<HG_BASIS_UGIMP_LINEAR> { static constexpr int S = 3; };
template <> struct BasisTraits<HG_BASIS_LINEAR> { static constexpr int S = 2; };
template <> struct BasisTraits<HG_BASIS_CUBIC> { static constexpr int S = 4; };

template <typename T> struct Vec4T;
template <> struct Vec4T<double> { using type = double4; };
template <> struct Vec4T<float> { using type = float4; };

// a / h rounded to nearest (the IEEE quotient), given inv_h = 1 / h rounded
// to nearest, in five FP64 operations instead of a division sequence.
// q = a inv_h is within two ulps of a / h; each step q += (a - q h) inv_h,
// whose remainder the fused multiply-add computes exactly, brings it closer.
// After the first step q is within an ulp, and Markstein's theorem (P.
// Markstein, IBM J. Res. Dev. 34(1), 1990; Muller et al., Handbook of
// Floating-Point Arithmetic, the chapter on FMA division) makes the second
// the correctly rounded quotient.  A zero quotient may come out as +0 where
// the division gives -0, which floor, ceil and every comparison treat alike.
// Overflowed or non-finite operands give NaN where the division may give an
// infinity; both fail the callers' range checks.
__device__ __forceinline__ double div_rn(double a, double h, double inv_h) {
  double q = a * inv_h;
  q = fma(fma(-q, h, a), inv_h, q);
  return fma(fma(-q, h, a), inv_h, q);
}

// Stencil of one axis: first node lo and last node hi (inclusive), as double
// so an off-grid or NaN coordinate can be detected before any integer cast.
// uGIMPLinearBasisFunctions.cpp:68-90, LinearBasisFunctions.cpp:48-60,
// CubicBasisFunctions.cpp:61-75.  The divisions by h are div_rn's.
template <int BASIS>
__device__ __forceinline__ void stencil_1d(double x, double hl, double gmin, double h,
                                           double inv_h, double &lo, double &hi) {
  if constexpr (BASIS == HG_BASIS_UGIMP_LINEAR) {
    lo = floor(div_rn((x - hl) - gmin, h, inv_h));
    hi = ceil(div_rn((x + hl) - gmin, h, inv_h));
  } else if constexpr (BASIS == HG_BASIS_LINEAR) {
    lo = floor(div_rn(x - gmin, h, inv_h));
    hi = lo + 1.0;
  } else {
    lo = floor(div_rn(x - gmin, h, inv_h)) - 1.0;
    hi = lo + 3.0;
  }
}

// 1-D basis factors for nodes lo..lo+S-1 along one axis; nodes past the
// stencil's hi are masked to exactly zero, as the reference never visits them.
// uGIMP: LinearIntegral / LinearIntegralGrad (uGIMPLinearBasisFunctions.cpp:6-47);
// Linear: linearN / lineardN (LinearBasisFunctions.cpp:6-23);
// Cubic: cubicN / cubicdN (CubicBasisFunctions.cpp:6-46).
// Divisions by the cell width are taken as products with its reciprocal
// (inv_w), which moves results by at most an ulp or two.
template <typename T, int BASIS>
__device__ __forceinline__ void basis_1d(double x, T hl, double gmin, double h, T inv_w, int lo,
                                         int cnt, T (&a)[BasisTraits<BASIS>::S],
                                         T (&d)[BasisTraits<BASIS>::S]) {
  constexpr int S = BasisTraits<BASIS>::S;
#pragma unroll
  for (int k = 0; k < S; ++k) {
    if constexpr (BASIS == HG_BASIS_UGIMP_LINEAR) {
      const double xi = h * double(lo + k) + gmin;
      const T r = T(x - xi);
      const T w = T(h);
      const T diff = fabs(r);
      const T sgn = r >= T(0) ? T(1) : T(-1);
      T av, dv;
      if (diff >= w + hl) {
        av = T(0);
        dv = T(0);
      } else if (diff >= w - hl) {
        const T u = w + hl - diff;
        av = u * u * (T(0.5) * inv_w);
        dv = -sgn * u * inv_w;
      } else if (diff >= hl) {
        av = T(2) * hl * (T(1) - diff * inv_w);
        dv = -sgn * T(2) * hl * inv_w;
      } else {
        av = T(2) * hl - (hl * hl + diff * diff) * inv_w;
        dv = T(2) * (-r) * inv_w;
      }
      a[k] = (k < cnt) ? av : T(0);
      d[k] = (k < cnt) ? dv : T(0);
    } else if constexpr (BASIS == HG_BASIS_LINEAR) {
      const T t = T((x - h * double(lo + k) - gmin) * double(inv_w));
      const T at = fabs(t);
      a[k] = at < T(1) ? T(1) - at : T(0);
      d[k] = (at <= T(1)) ? ((t >= T(0)) ? T(-1) : T(1)) : T(0);
    } else {
      const T t = T((x - h * double(lo + k) - gmin) * double(inv_w));
      const T at = fabs(t);
      const T sg = (t >= T(0)) ? T(1) : T(-1);
      if (at < T(1)) {
        a[k] = (T(0.5) * at - T(1)) * at * at + T(2.0 / 3.0);
        d[k] = sg * (T(1.5) * at - T(2)) * at;
      } else if (at < T(2)) {
        a[k] = ((-at * T(1.0 / 6.0) + T(1)) * at - T(2)) * at + T(4.0 / 3.0);
        d[k] = sg * ((T(-0.5) * at + T(2)) * at - T(2));
      } else {
        a[k] = T(0);
        d[k] = T(0);
      }
    }
  }
}

// FP32 uGIMP factors of the three nodes lo, lo + 1, lo + 2 along one axis,
// for a point whose support [x - hl, x + hl] starts in cell lo
// (lo = floor((x - hl - gmin) / h), so r0 = x - x_lo lies in [hl, h + hl))
// and is shorter than a cell (hl < h / 2, which scene creation requires).
// Each node then has two pieces of basis_1d's four to choose from:
//   node 0, r0 in [hl, h + hl):        tail (|r| >= h - hl) or linear;
//   node 1, r1 = r0 - h in [hl - h, hl): inner (|r| < hl) or linear, r1 < 0;
//   node 2, r2 = r0 - 2h in [hl - 2h, hl - h): tail, or nothing past h + hl
//                                        or past the stencil's count.
// The pieces are basis_1d's expressions.  The factors are continuous across
// the piece boundaries, so where rounding puts a distance just across a
// boundary the other piece's value differs only at rounding level.
__device__ __forceinline__ void ugimp_1d_f32(double x, float hl, double gmin, double h,
                                             float inv_w, int lo, int cnt, float (&a)[3],
                                             float (&d)[3]) {
  const float w = float(h);
  const float r0 = float(x - (h * double(lo) + gmin));
  const float r1 = float(x - (h * double(lo + 1) + gmin));
  const float r2 = float(x - (h * double(lo + 2) + gmin));
  const float c05 = 0.5f * inv_w;
  const float hl2 = 2.0f * hl;
  const float lin = hl2 * inv_w;  // |d| on the linear piece
  const float wph = w + hl, wmh = w - hl;
  {
    const float u = wph - r0;
    const bool tail = r0 >= wmh;
    a[0] = tail ? u * u * c05 : hl2 * (1.0f - r0 * inv_w);
    d[0] = tail ? -u * inv_w : -lin;
  }
  {
    const float diff = fabsf(r1);
    const bool inner = diff < hl;
    a[1] = inner ? hl2 - (hl * hl + diff * diff) * inv_w : hl2 * (1.0f - diff * inv_w);
    d[1] = inner ? 2.0f * (-r1) * inv_w : lin;
  }
  {
    const float diff = -r2;
    const float u = wph - diff;
    const bool none = diff >= wph || cnt < 3;
    a[2] = none ? 0.0f : u * u * c05;
    d[2] = none ? 0.0f : u * inv_w;
  }
}

// Both axes' factors, pre-scaled so that a node's weight is ax[i]*ay[j] and
// its gradient (dx[i]*ay[j], ax[i]*dy[j]): uGIMP divides the product of the
// overlap integrals by the point volume 4 hl^2 (uGIMP :108-123/:141-156);
// the standard bases divide the gradient by the cell width (Linear
// :113-135, Cubic :129-146).
template <typename T, int BASIS>
__device__ __forceinline__ void basis_2d(double x, double y, T hl, double min0, double min1,
                                         double h, T inv_h, int bx, int by, int cntx, int cnty,
                                         T (&ax)[BasisTraits<BASIS>::S],
                                         T (&dx)[BasisTraits<BASIS>::S],
                                         T (&ay)[BasisTraits<BASIS>::S],
                                         T (&dy)[BasisTraits<BASIS>::S]) {
  constexpr int S = BasisTraits<BASIS>::S;
  if constexpr (BASIS == HG_BASIS_UGIMP_LINEAR && std::is_same<T, float>::value) {
    ugimp_1d_f32(x, hl, min0, h, inv_h, bx, cntx, ax, dx);
    ugimp_1d_f32(y, hl, min1, h, inv_h, by, cnty, ay, dy);
  } else {
    basis_1d<T, BASIS>(x, hl, min0, h, inv_h, bx, cntx, ax, dx);
    basis_1d<T, BASIS>(y, hl, min1, h, inv_h, by, cnty, ay, dy);
  }
  if constexpr (BASIS == HG_BASIS_UGIMP_LINEAR) {
    const T inv_vp = T(1) / (T(4) * hl * hl);
#pragma unroll
    for (int k = 0; k < S; ++k) {
      ax[k] *= inv_vp;
      dx[k] *= inv_vp;
    }
  } else {
#pragma unroll
    for (int k = 0; k < S; ++k) {
      dx[k] *= inv_h;
      dy[k] *= inv_h;
    }
  }
}

// 1-D basis factor and its derivative for node offset k (0..S-1) from the
// point's stencil base, masked past the stencil's count.
template <typename T, int BASIS>
__device__ __forceinline__ void basis_factor(double x, T hl, double gmin, double h, T inv_h,
                                             int node, int k, int cnt, T &a, T &d) {
  if constexpr (BASIS == HG_BASIS_UGIMP_LINEAR) {
    const double xi = h * double(node) + gmin;
    const T r = T(x - xi);
    const T w = T(h);
    const T diff = fabs(r);
    const T sgn = r >= T(0) ? T(1) : T(-1);
    if (diff >= w + hl || k >= cnt) {
      a = T(0);
      d = T(0);
    } else if (diff >= w - hl) {
      const T u = w + hl - diff;
      a = u * u * (T(0.5) * inv_h);
      d = -sgn * u * inv_h;
    } else if (diff >= hl) {
      a = T(2) * hl * (T(1) - diff * inv_h);
      d = -sgn * T(2) * hl * inv_h;
    } else {
      a = T(2) * hl - (hl * hl + diff * diff) * inv_h;
      d = T(2) * (-r) * inv_h;
    }
  } else if constexpr (BASIS == HG_BASIS_LINEAR) {
    const T t = T((x - h * double(node) - gmin) * double(inv_h));
    const T at = fabs(t);
    a = at < T(1) ? T(1) - at : T(0);
    d = (at <= T(1)) ? ((t >= T(0)) ? T(-1) : T(1)) : T(0);
  } else {
    const T t = T((x - h * double(node) - gmin) * double(inv_h));
    const T at = fabs(t);
    const T sg = (t >= T(0)) ? T(1) : T(-1);
    if (at < T(1)) {
      a = (T(0.5) * at - T(1)) * at * at + T(2.0 / 3.0);
      d = sg * (T(1.5) * at - T(2)) * at;
    } else if (at < T(2)) {
      a = ((-at * T(1.0 / 6.0) + T(1)) * at - T(2)) * at + T(4.0 / 3.0);
      d = sg * ((T(-0.5) * at + T(2)) * at - T(2));
    } else {
      a = T(0);
      d = T(0);
    }
  }
}

// The grain-continuum exchange at one node holding grain mass
// (HybridCoupling2D::updateVelocitiesOnGrid, Coupling.cpp:533-579): the
// multiplier that equalises the two velocities while conserving the node's
// momentum; the continuum's new momentum and its (v, a) (as the grid update
// computes them, PhysicsGrid.cpp:588-609), and the grain side's velocity and
// acceleration after the exchange.  node indexes the continuum's arrays, n
// the grain-side arrays (the node's index within its scene).
struct GrainNodes {
  const double *dmass, *dmom;
  double *vafter, *aafter;
  // where not null, the grid update clears the grain mass and momentum of a
  // node once it has exchanged there (the same arrays as dmass and dmom)
  double *clear_mass = nullptr, *clear_mom = nullptr;
};

// The exchange on values in hand: m_dem is the node's grain mass, m_mpm its
// continuum mass and (pn0, pn1) the continuum's new momentum, which becomes
// pn - l.
__device__ __forceinline__ void exchange_in_place(uint32_t n, const GrainNodes &g, double m_dem,
                                                  double m_mpm, double &pn0, double &pn1) {
  const double msum = m_mpm + m_dem;
  const double l0 = (m_dem * pn0 - m_mpm * g.dmom[2 * n]) / msum;
  const double l1 = (m_dem * pn1 - m_mpm * g.dmom[2 * n + 1]) / msum;
  pn0 = pn0 - l0;
  pn1 = pn1 - l1;
  g.vafter[2 * n] = (g.dmom[2 * n] + l0) / m_dem;
  g.vafter[2 * n + 1] = (g.dmom[2 * n + 1] + l1) / m_dem;
  g.aafter[2 * n] = l0 / m_dem;
  g.aafter[2 * n + 1] = l1 / m_dem;
}

// The same at a node whose grid update is stored (mass, momentum, new
// momentum): the new momentum and the (v, a) vector are rewritten.
template <typename T>
__device__ __forceinline__ void exchange_node(double dt, uint32_t node, uint32_t n,
                                              const GrainNodes &g, const double *gm,
                                              const double *gp, double *gpn, void *gva) {
  using T4 = typename Vec4T<T>::type;
  const double m_mpm = gm[node];
  double pn0 = gpn[2 * node], pn1 = gpn[2 * node + 1];
  exchange_in_place(n, g, g.dmass[n], m_mpm, pn0, pn1);
  gpn[2 * node] = pn0;
  gpn[2 * node + 1] = pn1;
  T4 va;
  if (m_mpm > 0.0) {
    const double dm = dt * m_mpm;
    va.x = T(pn0 / m_mpm);
    va.y = T(pn1 / m_mpm);
    va.z = T((pn0 - gp[2 * node]) / dm);
    va.w = T((pn1 - gp[2 * node + 1]) / dm);
  } else {
    va.x = va.y = va.z = va.w = T(0);
  }
  static_cast<T4 *>(gva)[node] = va;
}

}  // namespace hgmpm
