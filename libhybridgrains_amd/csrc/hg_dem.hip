// hg_dem.hip -- circle DEM step with the penalty contact map, on gfx950.
//
// Replaces rigidbody2d PenaltyImpactFrictionMap::flow / flowWithWeights
// (rigidbody2d/PenaltyImpactFrictionMap.cpp:1083-1149) together with the
// active set RigidBody2DSim::computeActiveSet builds for circles and static
// planes (RigidBody2DSim.cpp:270-413, :777-842, :886-912, :1140-1178).
//
// Layout in HBM (bodies keep their creation order; scenes are concatenated):
//  * bodies: x, y, theta (FP64), vx, vy, omega (T), radius (FP64: the pair
//    set is decided in FP64 in every precision), 1/m, 1/I, m (T), fixed, and
//    optional contact weights (T).
//  * broad phase: cells of width >= the largest diameter, folded onto a
//    wrapped table of Wx x Wy buckets per scene (powers of two, each >= 4):
//    bucket = (cx mod Wx) + Wx (cy mod Wy).  The table covers the bodies' box
//    with room to spread, so the buckets of a compact scene are its cells;
//    cells a table width apart share a bucket, which only adds candidates
//    the distance test rejects (the 3x3 neighbour buckets stay distinct).
//    Each step files the moved bodies in the next of three rotating bucket
//    tables (up to kslots per bucket, an overflow list beyond).
//  * contacts: each contact is owned by one body -- the smaller index of a
//    circle pair, or the simulated body of a pair with a fixed body, as the
//    reference keys its CollisionTypeCache.  A body keeps up to cap owned
//    slots (partner key, tangential spring delta_s, force, torques; cap is
//    8, 16 or 32 per handle, from the scene's radii),
//    sorted by partner key, double-buffered so that the next step's
//    contacts join their spring history by partner; per-plane slots likewise.
//
// A step is one launch, dem_step_kernel: a team of lanes per body does the
// narrow phase over the 3x3 cells, the history join and the penalty force
// with the Coulomb projection of the spring for the body's own contacts and
// for those a smaller partner owns, and the plane contacts; its first lane
// then integrates (symplectic Euler) with the forces gathered in the
// reference's summation order -- gravity, planes, pairs owned by smaller
// partners, own pairs, kinematic pairs -- into the other state buffer and
// files the body in the next table (the hybrid's coupling files its grains
// itself, at their coupled positions).
#include "hg_dem_handle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>

namespace hgdem {

template <typename T>
__device__ __forceinline__ T ld(const void *p, uint32_t i) {
  return static_cast<const T *>(p)[i];
}
template <typename T>
__device__ __forceinline__ void st(void *p, uint32_t i, T v) {
  static_cast<T *>(p)[i] = v;
}

__device__ __forceinline__ uint32_t scene_of(const uint32_t *b_off, uint32_t ns, uint32_t i) {
  uint32_t lo = 0, hi = ns;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (b_off[mid] <= i) lo = mid; else hi = mid;
  }
  return lo;
}


// One contact's penalty force with the Coulomb projection of the tangential
// spring: PenaltyImpactFrictionMap::updateAnchorsAndAccumulateCollisionForces
// (PenaltyImpactFrictionMap.cpp:616-700; identical for every contact type),
// with secondRootOfQuadratic (:601-614).  FP64.
__device__ __forceinline__ void penalty_force(const SceneDev &sc, double w, double depth,
                                              double n0, double n1, double vr0, double vr1,
                                              double &ds0, double &ds1, double &f0, double &f1) {
  const double vdn = n0 * vr0 + n1 * vr1;
  const double vt0 = vr0 - vdn * n0, vt1 = vr1 - vdn * n1;
  const double vn0 = vr0 - vt0, vn1 = vr1 - vt1;
  const double cn = -sc.kn * w * depth;
  const double dn = 0.5 * sc.gamman * w;
  const double nf0 = cn * n0 - dn * vn0, nf1 = cn * n1 - dn * vn1;
  const double ck = -sc.kt * w;
  const double dt_ = 0.5 * sc.gammat * w;
  double ff0 = ck * ds0 - dt_ * vt0, ff1 = ck * ds1 - dt_ * vt1;
  const double mu_fn = sc.mu * sqrt(nf0 * nf0 + nf1 * nf1);
  const double ft = sqrt(ff0 * ff0 + ff1 * ff1);
  if (0.5 * sc.gammat * w * sqrt(vt0 * vt0 + vt1 * vt1) > mu_fn) {
    ds0 = 0.0;
    ds1 = 0.0;
    const double g = -0.5 * sc.gammat * w;
    ff0 = g * vt0;
    ff1 = g * vt1;
    const double z = ff0 * ff0 + ff1 * ff1;
    if (z > 0.0) {
      const double nz = sqrt(z);
      ff0 = ff0 / nz;
      ff1 = ff1 / nz;
    }
    ff0 *= mu_fn;
    ff1 *= mu_fn;
  } else if (ft > mu_fn) {
    const double a = sc.kt * w * sc.kt * w * (ds0 * ds0 + ds1 * ds1);
    const double b = sc.kt * w * sc.gammat * w * (ds0 * vt0 + ds1 * vt1);
    const double c = 0.25 * sc.gammat * w * sc.gammat * w * (vt0 * vt0 + vt1 * vt1) -
                     sc.mu * sc.mu * (nf0 * nf0 + nf1 * nf1);
    const double dscr = b * b - 4 * a * c;
    const double dsq = sqrt(dscr);
    const double root1 = (b > 0.0) ? (2.0 * c) / (-b - dsq) : (-b + dsq) / (2.0 * a);
    const double nd0 = root1 * ds0, nd1 = root1 * ds1;
    ff0 = ck * nd0 - dt_ * vt0;
    ff1 = ck * nd1 - dt_ * vt1;
    ds0 = nd0;
    ds1 = nd1;
  }
  f0 = nf0 + ff0;
  f1 = nf1 + ff1;
}

// delta_s update of a new or continuing contact (PenaltyImpactFrictionMap.cpp:61-80)
__device__ __forceinline__ void spring(bool had, double old0, double old1, double dt, double vr0,
                                       double vr1, double n0, double n1, double &ds0,
                                       double &ds1) {
  if (!had) {
    ds0 = dt * vr0;
    ds1 = dt * vr1;
  } else {
    ds0 = old0;
    ds1 = old1;
    ds0 += dt * vr0;
    ds1 += dt * vr1;
  }
  const double d = n0 * ds0 + n1 * ds1;
  ds0 -= d * n0;
  ds1 -= d * n1;
}

struct ContactArgs {
  const SceneDev *scenes;
  const uint32_t *b_off;
  uint32_t nscenes, nbodies, nblocks_padded;
  BodiesRaw b;
  SlotsRaw old_s, new_s;
  ForceRaw f;
  int32_t *err;
  double *cpoint;  // optional: contact point per owned slot (circle pairs)
  // the table read, the one the moved bodies are filed in, and the one
  // cleared for the step after (read by the step before)
  TableRaw tcur, tnext, tclr;
  int32_t kslots;
  // the state after the step (position, angle and velocities; the other
  // fields are b's) and, when kept, the total force per body
  BodiesRaw bn;
  double *F_out;
  int32_t file;  // file the moved bodies in tnext (else the caller does)
};

// A team of kTeam lanes per body: the lanes split the body's candidates and
// then its contacts, so that a body's chain of dependent loads is a fraction
// of what one lane would walk.  Contacts are collected unordered in LDS and
// each takes the slot of its rank among the body's keys, which is the order
// of the partner keys, as a sorted list would have it.  The team then sums
// the body's terms (rounded to T as stored) in the reference's order in its
// first lane and integrates the body into bn (the state this launch does not
// read), filing it in the next table unless the caller does (the hybrid's
// grains, filed at their coupled positions).  Each body's team forms every
// contact term the body receives, so no body waits for another's: the step
// is one launch.  Eight lanes per body (timed against four on the 20k column
// and the 1M hybrid scene's grains).
constexpr int kTeam = 8;
constexpr int kTeamThreads = 256;
constexpr int kTeamBodies = kTeamThreads / kTeam;

template <typename T, int CAP>
__global__ __launch_bounds__(kTeamThreads) void dem_step_kernel(ContactArgs a) {
  __shared__ uint32_t s_own[kTeamBodies][CAP];
  __shared__ uint32_t s_in[kTeamBodies][CAP];
  __shared__ int s_nown[kTeamBodies], s_nin[kTeamBodies];
  // the terms the integration sums: own and incoming pairs by slot, planes
  // by index (with their activity), each rounded to T (the FP32 mode's
  // terms then take half the LDS, so that more workgroups fit on a CU)
  __shared__ T s_of[kTeamBodies][CAP][3], s_if[kTeamBodies][CAP][3];
  __shared__ T s_pf[kTeamBodies][kMaxPlanes][3];
  __shared__ uint8_t s_pon[kTeamBodies][kMaxPlanes];
  const int tb = int(threadIdx.x) / kTeam, lane = int(threadIdx.x) % kTeam;
  const uint32_t blk = hg_xcd_remap(blockIdx.x, a.nblocks_padded);
  const uint32_t s = blk * uint32_t(kTeamBodies) + uint32_t(tb);
  const bool active = s < a.nbodies;
  if (lane == 0) {
    s_nown[tb] = 0;
    s_nin[tb] = 0;
  }
  __syncthreads();
  // teams take the bodies in creation order
  const uint32_t i = active ? s : 0u;
  const uint32_t sid = active ? scene_of(a.b_off, a.nscenes, i) : 0u;
  const SceneDev &sc = a.scenes[sid];
  const double xi = active ? a.b.x[i] : 0.0, yi = active ? a.b.y[i] : 0.0;
  const double ri = active ? a.b.r[i] : 0.0;
  const bool fi = active ? a.b.fixed[i] != 0 : true;
  // the table the next step files into: every bucket it holds bodies in is
  // some body's recorded bucket
  if (active && lane == 0) a.tclr.count[a.tclr.bkey[i]] = 0u;
  if (blockIdx.x == 0 && threadIdx.x == 0) *a.tclr.novf = 0u;
  if (active) {
    const bool finite_i = isfinite(xi) && isfinite(yi);
    int cx, cy;
    dem_cell_of(sc, xi, yi, cx, cy);
    const uint32_t wx = uint32_t(sc.ncx), wy = uint32_t(sc.ncy);
    const uint32_t ks = uint32_t(a.kslots);
    // the buckets of the 3x3 cells, then the overflow list
    uint32_t rb[9], rn[9], tot = 0;
#pragma unroll
    for (int r = 0; r < 9; ++r) {
      const uint32_t bk = sc.cell_off + (uint32_t(cx + r % 3 - 1) & (wx - 1u)) +
                          wx * (uint32_t(cy + r / 3 - 1) & (wy - 1u));
      rb[r] = bk * ks;
      rn[r] = finite_i ? min(a.tcur.count[bk], ks) : 0u;
      tot += rn[r];
    }
    const uint32_t nov = finite_i ? *a.tcur.novf : 0u;
    for (uint32_t t = uint32_t(lane); t < tot + nov; t += uint32_t(kTeam)) {
      uint32_t j;
      if (t < tot) {
        uint32_t off = t;
        int r = 0;
        while (off >= rn[r]) {
          off -= rn[r];
          ++r;
        }
        j = a.tcur.items[rb[r] + off];
      } else {
        j = a.tcur.ovf[t - tot];
        if (j < sc.b_off || j >= sc.b_off + sc.b_cnt) continue;  // another scene's body
      }
      if (j == i) continue;
      const bool fj = a.b.fixed[j] != 0;
      if (fi && fj) continue;
      const double dx = xi - a.b.x[j], dy = yi - a.b.y[j];
      if (!isfinite(dx) || !isfinite(dy)) continue;
      const double rs = ri + a.b.r[j];
      const double c0 = (dx * dx + dy * dy) - rs * rs;
      if (!(c0 <= 0.0)) continue;
      if (!fi && (fj || i < j)) {
        const int q = atomicAdd(&s_nown[tb], 1);
        if (q < CAP) s_own[tb][q] = (fj ? kKinematic : 0u) | (j - sc.b_off);
      } else if (!fi && !fj) {
        const int q = atomicAdd(&s_nin[tb], 1);
        if (q < CAP) s_in[tb][q] = j - sc.b_off;
      }
    }
  }
  __syncthreads();
  const int cnt_own = active ? s_nown[tb] : 0, cnt_in = active ? s_nin[tb] : 0;
  if (lane == 0 && (cnt_own > CAP || cnt_in > CAP)) atomicOr(&a.err[sid], 4);
  const int nown = min(cnt_own, CAP), nin = min(cnt_in, CAP);

  const double dt = sc.dt;
  // (a lane without a body reads body 0 and writes nothing)
  const double vxi = double(ld<T>(a.b.vx, i)), vyi = double(ld<T>(a.b.vy, i));
  const double omi = double(ld<T>(a.b.om, i));
  const double wi = sc.weighted ? double(ld<T>(a.b.w, i)) : 1.0;
  const uint32_t obase = i * uint32_t(CAP);
  const int old_n = a.old_s.ocnt[i];
  if (active) {
    for (int k = lane; k < nown; k += kTeam) {
      const uint32_t key = s_own[tb][k];
      int rank = 0;
      for (int q = 0; q < nown; ++q) rank += (s_own[tb][q] < key) ? 1 : 0;
      const bool kin = (key & kKinematic) != 0;
      const uint32_t j = (key & ~kKinematic) + sc.b_off;
      const double xj = a.b.x[j], yj = a.b.y[j], rj = a.b.r[j];
      // n = normalized(q_i - q_j) (RigidBody2DSim.cpp:320)
      const double d0 = xi - xj, d1 = yi - yj;
      const double dz = d0 * d0 + d1 * d1;
      double n0 = d0, n1 = d1;
      if (dz > 0.0) {
        const double nz = sqrt(dz);
        n0 = d0 / nz;
        n1 = d1 / nz;
      }
      const double gap = sqrt(d0 * d0 + d1 * d1) - ri - rj;
      const double depth = gap < 0.0 ? gap : 0.0;
      const double vxj = double(ld<T>(a.b.vx, j)), vyj = double(ld<T>(a.b.vy, j));
      const double omj = double(ld<T>(a.b.om, j));
      const uint32_t sl = obase + uint32_t(rank);
      double vr0, vr1, r00, r01, r10 = 0.0, r11 = 0.0, w;
      if (!kin) {
        // CircleCircleConstraint (CircleCircleConstraint.cpp:171-193, :224-233)
        const double f = ri / (ri + rj);
        const double p0 = xi + f * (xj - xi), p1 = yi + f * (yj - yi);
        r00 = p0 - xi;
        r01 = p1 - yi;
        r10 = p0 - xj;
        r11 = p1 - yj;
        vr0 = vxi + omi * (-r01) - vxj - omj * (-r11);
        vr1 = vyi + omi * r00 - vyj - omj * r10;
        w = sc.weighted ? (wi + double(ld<T>(a.b.w, j))) * 0.5 : 1.0;
        if (a.cpoint) {
          a.cpoint[2 * sl] = p0;
          a.cpoint[2 * sl + 1] = p1;
        }
      } else {
        // KinematicCircleCircleConstraint (KinematicObjectCircleConstraint.cpp:121-199)
        r00 = -ri * n0;
        r01 = -ri * n1;
        const double cp0 = xi - ri * n0, cp1 = yi - ri * n1;
        const double arm0 = cp0 - xj, arm1 = cp1 - yj;
        const double kv0 = vxj + omj * (-arm1), kv1 = vyj + omj * arm0;
        vr0 = vxi + omi * (-r01) - kv0;
        vr1 = vyi + omi * r00 - kv1;
        w = wi;
      }
      // spring history of this pair: the body's previous keys in a few vector
      // loads, then the matching slot's spring
      bool had = false;
      double o0 = 0.0, o1 = 0.0;
      {
        const uint4 *kv = reinterpret_cast<const uint4 *>(a.old_s.okey + obase);
        int hit = -1;
  #pragma unroll
        for (int q4 = 0; q4 < CAP / 4; ++q4) {
          const uint4 v = kv[q4];
          const uint32_t kk[4] = {v.x, v.y, v.z, v.w};
  #pragma unroll
          for (int e = 0; e < 4; ++e)
            if (hit < 0 && 4 * q4 + e < old_n && kk[e] == key) hit = 4 * q4 + e;
        }
        if (hit >= 0) {
          had = true;
          o0 = double(ld<T>(a.old_s.ods, 2 * (obase + uint32_t(hit))));
          o1 = double(ld<T>(a.old_s.ods, 2 * (obase + uint32_t(hit)) + 1));
        }
      }
      double ds0, ds1, F0, F1;
      spring(had, o0, o1, dt, vr0, vr1, n0, n1, ds0, ds1);
      penalty_force(sc, w, depth, n0, n1, vr0, vr1, ds0, ds1, F0, F1);
      a.new_s.okey[sl] = key;
      st<T>(a.new_s.ods, 2 * sl, T(ds0));
      st<T>(a.new_s.ods, 2 * sl + 1, T(ds1));
      st<T>(a.f.oF, 2 * sl, T(F0));
      st<T>(a.f.oF, 2 * sl + 1, T(F1));
      s_of[tb][rank][0] = T(F0);
      s_of[tb][rank][1] = T(F1);
      s_of[tb][rank][2] = T(r00 * F1 - r01 * F0);
    }
    // The contacts owned by a smaller partner j: the owner's computation above
    // (j as the body, this body as its partner: geometry, spring history from
    // j's previous slots, force), so that the integration finds the force this
    // body receives among its own records.  Every operation is the owner's, in
    // its order, so the values equal the owner's bit for bit.
    const uint32_t lid = i - sc.b_off;
    for (int k = lane; k < nin; k += kTeam) {
      const uint32_t key = s_in[tb][k];
      int rank = 0;
      for (int q = 0; q < nin; ++q) rank += (s_in[tb][q] < key) ? 1 : 0;
      const uint32_t j = key + sc.b_off;
      const double xo = a.b.x[j], yo = a.b.y[j], ro = a.b.r[j];
      const double vxo = double(ld<T>(a.b.vx, j)), vyo = double(ld<T>(a.b.vy, j));
      const double omo = double(ld<T>(a.b.om, j));
      const double d0 = xo - xi, d1 = yo - yi;
      const double dz = d0 * d0 + d1 * d1;
      double n0 = d0, n1 = d1;
      if (dz > 0.0) {
        const double nz = sqrt(dz);
        n0 = d0 / nz;
        n1 = d1 / nz;
      }
      const double gap = sqrt(d0 * d0 + d1 * d1) - ro - ri;
      const double depth = gap < 0.0 ? gap : 0.0;
      const double f = ro / (ro + ri);
      const double p0 = xo + f * (xi - xo), p1 = yo + f * (yi - yo);
      const double r00 = p0 - xo, r01 = p1 - yo;
      const double r10 = p0 - xi, r11 = p1 - yi;
      const double vr0 = vxo + omo * (-r01) - vxi - omi * (-r11);
      const double vr1 = vyo + omo * r00 - vyi - omi * r10;
      const double w = sc.weighted ? (double(ld<T>(a.b.w, j)) + wi) * 0.5 : 1.0;
      bool had = false;
      double o0 = 0.0, o1 = 0.0;
      {
        const uint32_t jbase = j * uint32_t(CAP);
        const int jold = a.old_s.ocnt[j];
        const uint4 *kv = reinterpret_cast<const uint4 *>(a.old_s.okey + jbase);
        int hit = -1;
  #pragma unroll
        for (int q4 = 0; q4 < CAP / 4; ++q4) {
          const uint4 v = kv[q4];
          const uint32_t kk[4] = {v.x, v.y, v.z, v.w};
  #pragma unroll
          for (int e = 0; e < 4; ++e)
            if (hit < 0 && 4 * q4 + e < jold && kk[e] == lid) hit = 4 * q4 + e;
        }
        if (hit >= 0) {
          had = true;
          o0 = double(ld<T>(a.old_s.ods, 2 * (jbase + uint32_t(hit))));
          o1 = double(ld<T>(a.old_s.ods, 2 * (jbase + uint32_t(hit)) + 1));
        }
      }
      double ds0, ds1, F0, F1;
      spring(had, o0, o1, dt, vr0, vr1, n0, n1, ds0, ds1);
      penalty_force(sc, w, depth, n0, n1, vr0, vr1, ds0, ds1, F0, F1);
      s_if[tb][rank][0] = T(F0);
      s_if[tb][rank][1] = T(F1);
      s_if[tb][rank][2] = T(r10 * F1 - r11 * F0);
    }
    if (lane == 0) a.new_s.ocnt[i] = uint8_t(nown);

    // static planes (RigidBody2DSim.cpp:777-842; StaticPlaneCircleConstraint.cpp)
    for (int k = lane; k < sc.nplanes; k += kTeam) {
      const uint32_t ps = i * kMaxPlanes + uint32_t(k);
      const double pn0 = sc.plane_n[2 * k], pn1 = sc.plane_n[2 * k + 1];
      const double rel0 = xi - sc.plane_x[2 * k], rel1 = yi - sc.plane_x[2 * k + 1];
      const double dist = pn0 * rel0 + pn1 * rel1;
      if (fi || !(dist <= ri)) {
        a.new_s.pact[ps] = 0;
        s_pon[tb][k] = 0;
        continue;
      }
      const double gap = dist - ri;
      const double depth = gap < 0.0 ? gap : 0.0;
      const double r00 = -ri * pn0, r01 = -ri * pn1;
      const double vr0 = vxi + omi * (-r01) - 0.0;
      const double vr1 = vyi + omi * r00 - 0.0;
      const bool had = a.old_s.pact[ps] != 0;
      const double o0 = had ? double(ld<T>(a.old_s.pds, 2 * ps)) : 0.0;
      const double o1 = had ? double(ld<T>(a.old_s.pds, 2 * ps + 1)) : 0.0;
      double ds0, ds1, F0, F1;
      spring(had, o0, o1, dt, vr0, vr1, pn0, pn1, ds0, ds1);
      penalty_force(sc, wi, depth, pn0, pn1, vr0, vr1, ds0, ds1, F0, F1);
      a.new_s.pact[ps] = 1;
      st<T>(a.new_s.pds, 2 * ps, T(ds0));
      st<T>(a.new_s.pds, 2 * ps + 1, T(ds1));
      s_pon[tb][k] = 1;
      s_pf[tb][k][0] = T(F0);
      s_pf[tb][k][1] = T(F1);
      s_pf[tb][k][2] = T(r00 * F1 - r01 * F0);
    }
  }

  __syncthreads();  // the team's terms are in LDS
  if (!active || lane != 0) return;
  // PenaltyImpactFrictionMap::flow's update (PenaltyImpactFrictionMap.cpp:
  // 1107-1114) with the forces gathered in the reference's summation order:
  // gravity, planes, pairs owned by smaller partners, own pairs, kinematic
  // pairs (the last two in key order)
  const double m = double(ld<T>(a.b.m, i));
  const double im = double(ld<T>(a.b.im, i)), iI = double(ld<T>(a.b.iI, i));
  const double th0 = a.b.th[i];
  double F0 = 0.0, F1 = 0.0, F2 = 0.0;
  F0 += m * sc.g0;
  F1 += m * sc.g1;
  for (int k = 0; k < sc.nplanes; ++k) {
    if (!s_pon[tb][k]) continue;
    F0 += double(s_pf[tb][k][0]);
    F1 += double(s_pf[tb][k][1]);
    F2 += double(s_pf[tb][k][2]);
  }
  for (int k = 0; k < nin; ++k) {
    F0 -= double(s_if[tb][k][0]);
    F1 -= double(s_if[tb][k][1]);
    F2 -= double(s_if[tb][k][2]);
  }
  for (int q = 0; q < nown; ++q) {
    F0 += double(s_of[tb][q][0]);
    F1 += double(s_of[tb][q][1]);
    F2 += double(s_of[tb][q][2]);
  }
  if (fi) F0 = F1 = F2 = 0.0;
  if (a.F_out) {
    a.F_out[3 * i] = F0;
    a.F_out[3 * i + 1] = F1;
    a.F_out[3 * i + 2] = F2;
  }
  const double v0 = vxi + (dt * im) * F0;
  const double v1 = vyi + (dt * im) * F1;
  const double v2 = omi + (dt * iI) * F2;
  const T tv0 = T(v0), tv1 = T(v1), tv2 = T(v2);
  st<T>(a.bn.vx, i, tv0);
  st<T>(a.bn.vy, i, tv1);
  st<T>(a.bn.om, i, tv2);
  const double x = xi + dt * double(tv0);
  const double y = yi + dt * double(tv1);
  a.bn.x[i] = x;
  a.bn.y[i] = y;
  a.bn.th[i] = th0 + dt * double(tv2);
  if (!(isfinite(x) && isfinite(y) && isfinite(v2))) atomicOr(&a.err[sid], 2);
  if (a.file) table_insert(a.tnext, a.kslots, dem_bucket(sc, x, y), i);
}

struct BinArgs {
  const SceneDev *scenes;
  const uint32_t *b_off;
  uint32_t nscenes, nbodies;
  const double *x, *y;
  TableRaw t;
  int32_t kslots;
};

// Files every body at its position in table t.
__global__ __launch_bounds__(kThreads) void dem_bin_kernel(BinArgs a) {
  const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= a.nbodies) return;
  const SceneDev &sc = a.scenes[scene_of(a.b_off, a.nscenes, i)];
  table_insert(a.t, a.kslots, dem_bucket(sc, a.x[i], a.y[i]), i);
}

}  // namespace hgdem

using namespace hgdem;

// ===========================================================================
// host side
// ===========================================================================

template <typename P>
static hg_status dalloc(hg_dem *h, P **out, size_t bytes) {
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

#define HG_TRY(expr)              \
  do {                            \
    hg_status s_ = (expr);        \
    if (s_ != HG_OK) return s_;   \
  } while (0)

template <typename T>
static void launch_step_t(hg_dem *h) {
  ContactArgs a;
  a.scenes = h->d_scenes;
  a.b_off = h->d_b_off;
  a.nscenes = h->nscenes;
  a.nbodies = h->nbodies;
  a.nblocks_padded = hg_pad8((h->nbodies + kTeamBodies - 1) / kTeamBodies);
  a.b = h->b;
  a.old_s = h->slots[h->cur];
  a.new_s = h->slots[1 - h->cur];
  a.f = h->f;
  a.err = h->err;
  a.cpoint = h->cpoint;
  a.tcur = h->tables[h->tph];
  a.tnext = h->tables[(h->tph + 1) % 3];
  a.tclr = h->tables[(h->tph + 2) % 3];
  a.kslots = h->kslots;
  a.bn = h->b_alt;
  a.F_out = h->keep_force ? h->F_out : nullptr;
  a.file = h->caller_files ? 0 : 1;
  ProfScope ps(h->prof, "dem_step", h->stream);
  const dim3 g(a.nblocks_padded), b(kTeamThreads);
  if (h->cap <= 8) hipLaunchKernelGGL((dem_step_kernel<T, 8>), g, b, 0, h->stream, a);
  else if (h->cap <= 16) hipLaunchKernelGGL((dem_step_kernel<T, 16>), g, b, 0, h->stream, a);
  else hipLaunchKernelGGL((dem_step_kernel<T, 32>), g, b, 0, h->stream, a);
}

void hg_dem_launch_step(hg_dem *h) {
  // one launch: contacts, integration into b_alt, filing in the next table
  if (h->nbodies) {
    if (h->precision == HG_FP64) launch_step_t<double>(h);
    else launch_step_t<float>(h);
  }
  std::swap(h->b.x, h->b_alt.x);
  std::swap(h->b.y, h->b_alt.y);
  std::swap(h->b.th, h->b_alt.th);
  std::swap(h->b.vx, h->b_alt.vx);
  std::swap(h->b.vy, h->b_alt.vy);
  std::swap(h->b.om, h->b_alt.om);
  h->tph = (h->tph + 1) % 3;
  h->cur = 1 - h->cur;
}

void hg_dem_launch_rebin(hg_dem *h) {
  // every table emptied, the one the next step reads filled from the
  // positions
  for (int c = 0; c < 3; ++c) {
    hipMemsetAsync(h->tables[c].count, 0, size_t(h->ncell) * 4, h->stream);
    hipMemsetAsync(h->tables[c].novf, 0, 4, h->stream);
  }
  if (h->nbodies) {
    BinArgs a;
    a.scenes = h->d_scenes;
    a.b_off = h->d_b_off;
    a.nscenes = h->nscenes;
    a.nbodies = h->nbodies;
    a.x = h->b.x;
    a.y = h->b.y;
    a.t = h->tables[h->tph];
    a.kslots = h->kslots;
    ProfScope ps(h->prof, "dem_bin", h->stream);
    hipLaunchKernelGGL(dem_bin_kernel, dim3((h->nbodies + kThreads - 1) / kThreads),
                       dim3(kThreads), 0, h->stream, a);
  }
}

static void launch_step(hg_dem *h) { hg_dem_launch_step(h); }

void hg_dem_caller_files(hg_dem *h) {
  h->caller_files = true;
  if (h->graph) {
    hipGraphExecDestroy(h->graph);
    h->graph = nullptr;
  }
}

static hg_status dem_upload_dt(hg_dem *h, double dt) {
  if (dt == h->dt_dev) return HG_OK;
  HG_HIP(h->ctx, hipStreamSynchronize(h->stream));
  for (SceneDev &s : h->scenes) s.dt = dt;
  HG_HIP(h->ctx, hipMemcpy(h->d_scenes, h->scenes.data(), h->scenes.size() * sizeof(SceneDev),
                           hipMemcpyHostToDevice));
  h->dt_dev = dt;
  if (h->graph) {
    hipGraphExecDestroy(h->graph);
    h->graph = nullptr;
  }
  return HG_OK;
}

template <typename T>
static hg_status upload_bodies(hg_dem *h, const hg_dem_desc *descs) {
  const uint32_t B = h->nbodies;
  std::vector<double> x, y, th, r;
  std::vector<T> vx, vy, om, im, iI, m;
  std::vector<uint8_t> fixed;
  for (uint32_t s = 0; s < h->nscenes; ++s) {
    const hg_dem_desc &d = descs[s];
    for (uint32_t k = 0; k < d.nbodies; ++k) {
      x.push_back(d.q[3 * k]);
      y.push_back(d.q[3 * k + 1]);
      th.push_back(d.q[3 * k + 2]);
      vx.push_back(T(d.v[3 * k]));
      vy.push_back(T(d.v[3 * k + 1]));
      om.push_back(T(d.v[3 * k + 2]));
      r.push_back(d.radius[k]);
      m.push_back(T(d.mass[3 * k]));
      im.push_back(T(1.0 / d.mass[3 * k]));
      iI.push_back(T(1.0 / d.mass[3 * k + 2]));
      fixed.push_back(d.fixed ? d.fixed[k] : 0);
    }
  }
  BodiesRaw &b = h->b;
  HG_TRY(dalloc(h, &b.x, B * 8));
  HG_TRY(dalloc(h, &b.y, B * 8));
  HG_TRY(dalloc(h, &b.th, B * 8));
  HG_TRY(dalloc(h, &b.r, B * 8));
  HG_TRY(dalloc(h, &b.vx, B * sizeof(T)));
  HG_TRY(dalloc(h, &b.vy, B * sizeof(T)));
  HG_TRY(dalloc(h, &b.om, B * sizeof(T)));
  // the other state buffer of the DEM-only steps
  HG_TRY(dalloc(h, &h->b_alt.x, B * 8));
  HG_TRY(dalloc(h, &h->b_alt.y, B * 8));
  HG_TRY(dalloc(h, &h->b_alt.th, B * 8));
  HG_TRY(dalloc(h, &h->b_alt.vx, B * sizeof(T)));
  HG_TRY(dalloc(h, &h->b_alt.vy, B * sizeof(T)));
  HG_TRY(dalloc(h, &h->b_alt.om, B * sizeof(T)));
  HG_TRY(dalloc(h, &b.im, B * sizeof(T)));
  HG_TRY(dalloc(h, &b.iI, B * sizeof(T)));
  HG_TRY(dalloc(h, &b.m, B * sizeof(T)));
  HG_TRY(dalloc(h, &b.w, B * sizeof(T)));
  HG_TRY(dalloc(h, &b.fixed, B));
  if (B) {
    HG_HIP(h->ctx, hipMemcpy(b.x, x.data(), B * 8, hipMemcpyHostToDevice));
    HG_HIP(h->ctx, hipMemcpy(b.y, y.data(), B * 8, hipMemcpyHostToDevice));
    HG_HIP(h->ctx, hipMemcpy(b.th, th.data(), B * 8, hipMemcpyHostToDevice));
    HG_HIP(h->ctx, hipMemcpy(b.r, r.data(), B * 8, hipMemcpyHostToDevice));
    HG_HIP(h->ctx, hipMemcpy(b.vx, vx.data(), B * sizeof(T), hipMemcpyHostToDevice));
    HG_HIP(h->ctx, hipMemcpy(b.vy, vy.data(), B * sizeof(T), hipMemcpyHostToDevice));
    HG_HIP(h->ctx, hipMemcpy(b.om, om.data(), B * sizeof(T), hipMemcpyHostToDevice));
    HG_HIP(h->ctx, hipMemcpy(b.im, im.data(), B * sizeof(T), hipMemcpyHostToDevice));
    HG_HIP(h->ctx, hipMemcpy(b.iI, iI.data(), B * sizeof(T), hipMemcpyHostToDevice));
    HG_HIP(h->ctx, hipMemcpy(b.m, m.data(), B * sizeof(T), hipMemcpyHostToDevice));
    HG_HIP(h->ctx, hipMemcpy(b.fixed, fixed.data(), B, hipMemcpyHostToDevice));
    std::vector<T> ones(B, T(1));
    HG_HIP(h->ctx, hipMemcpy(b.w, ones.data(), B * sizeof(T), hipMemcpyHostToDevice));
  }
  // the fields the steps do not rewrite are shared by both buffers
  h->b_alt.r = b.r;
  h->b_alt.im = b.im;
  h->b_alt.iI = b.iI;
  h->b_alt.m = b.m;
  h->b_alt.fixed = b.fixed;
  h->b_alt.w = b.w;
  for (int k = 0; k < 2; ++k) {
    SlotsRaw &s = h->slots[k];
    HG_TRY(dalloc(h, &s.okey, size_t(B) * h->cap * 4));
    HG_TRY(dalloc(h, &s.ods, size_t(B) * h->cap * 2 * sizeof(T)));
    HG_TRY(dalloc(h, &s.ocnt, B));
    HG_TRY(dalloc(h, &s.pact, size_t(B) * kMaxPlanes));
    HG_TRY(dalloc(h, &s.pds, size_t(B) * kMaxPlanes * 2 * sizeof(T)));
    HG_HIP(h->ctx, hipMemset(s.ocnt, 0, B ? B : 1));
    HG_HIP(h->ctx, hipMemset(s.pact, 0, size_t(B ? B : 1) * kMaxPlanes));
  }
  HG_TRY(dalloc(h, &h->f.oF, size_t(B) * h->cap * 2 * sizeof(T)));
  return HG_OK;
}

// Contact slots per body.  A body of radius R touches partners whose centres
// lie within R + r_max of its own.  While no two partners overlap each other
// by more than half the smallest radius, their centres stay at least
// 1.5 r_min apart, so about 2 pi (R + r_max) / (1.5 r_min) of them fit
// around it.  The capacity is the first of 8, 16 and 32 at or above that
// bound for R = r_max (16 for equal radii); a body with more contacts than
// its capacity sets the overflow error bit.
int hg_dem_contact_capacity(const hg_dem_desc *descs, uint32_t nscenes) {
  double rmin = INFINITY, rmax = 0.0;
  for (uint32_t s = 0; s < nscenes; ++s)
    for (uint32_t k = 0; k < descs[s].nbodies; ++k) {
      rmin = std::min(rmin, descs[s].radius[k]);
      rmax = std::max(rmax, descs[s].radius[k]);
    }
  if (!(rmax > 0.0) || !(rmin > 0.0)) return kCapMin;
  const double bound = std::ceil(2.0 * M_PI * (2.0 * rmax) / (1.5 * rmin));
  return bound <= 8.0 ? 8 : (bound <= 16.0 ? 16 : kCapMax);
}

// Broad-phase table of one scene: the cell is at least the largest diameter
// (the 3x3-cell search needs that); the table spans the caller's domain, or
// the bodies' box with as much room again on every side to spread, rounded up
// to powers of two and kept within 8 buckets per body.
static uint32_t pow2_at_least(double v) {
  uint32_t w = 4;
  while (w < 65536u && double(w) < v) w <<= 1;
  return w;
}

static void dem_plan_grid(const hg_dem_desc &d, SceneDev &sd) {
  double rmax = 0.0, x0 = INFINITY, y0 = INFINITY, x1 = -INFINITY, y1 = -INFINITY;
  for (uint32_t k = 0; k < d.nbodies; ++k) {
    rmax = std::max(rmax, d.radius[k]);
    x0 = std::min(x0, d.q[3 * k]);
    x1 = std::max(x1, d.q[3 * k]);
    y0 = std::min(y0, d.q[3 * k + 1]);
    y1 = std::max(y1, d.q[3 * k + 1]);
  }
  if (d.nbodies == 0) x0 = y0 = x1 = y1 = 0.0, rmax = 1.0;
  const double cw = std::max(d.cell_width, 2.0 * rmax * (1.0 + 1e-9));
  double lo0, lo1, ex, ey;
  if (d.domain_max[0] > d.domain_min[0] && d.domain_max[1] > d.domain_min[1]) {
    lo0 = d.domain_min[0];
    lo1 = d.domain_min[1];
    ex = (d.domain_max[0] - lo0) / cw + 2.0;
    ey = (d.domain_max[1] - lo1) / cw + 2.0;
  } else {
    const double bx = (x1 - x0) / cw + 1.0, by = (y1 - y0) / cw + 1.0;
    lo0 = x0 - 0.5 * bx * cw;
    lo1 = y0 - 0.5 * by * cw;
    ex = 2.0 * bx + 2.0;
    ey = 2.0 * by + 2.0;
  }
  uint32_t wx = pow2_at_least(ex), wy = pow2_at_least(ey);
  const uint64_t cap = std::max<uint64_t>(8ull * std::max<uint32_t>(d.nbodies, 1u), 4096ull);
  while (uint64_t(wx) * wy > cap && (wx > 4 || wy > 4)) {
    if (wx >= wy && wx > 4) wx >>= 1;
    else wy >>= 1;
  }
  sd.cmin0 = lo0;
  sd.cmin1 = lo1;
  sd.cw = cw;
  sd.icw = 1.0 / cw;
  sd.ncx = int32_t(wx);
  sd.ncy = int32_t(wy);
}

// Item slots per bucket of the DEM-only tables: a cell at least one largest
// diameter wide holds at most about (cw / (2 r_min) + 1)^2 centres of
// non-overlapping circles; half as many again for overlaps, rounded up to a
// power of two in [8, 64].  A body that finds its bucket full goes to the
// overflow list, so the figure sets only speed, never results.
static int dem_table_slots(const hg_dem_desc *descs, uint32_t nscenes) {
  double worst = 1.0;
  for (uint32_t s = 0; s < nscenes; ++s) {
    const hg_dem_desc &d = descs[s];
    if (d.nbodies == 0) continue;
    SceneDev sd;
    std::memset(&sd, 0, sizeof(sd));
    dem_plan_grid(d, sd);
    double rmin = INFINITY;
    for (uint32_t k = 0; k < d.nbodies; ++k) rmin = std::min(rmin, d.radius[k]);
    if (!(rmin > 0.0)) continue;
    const double n = sd.cw / (2.0 * rmin) + 1.0;
    worst = std::max(worst, 1.5 * n * n);
  }
  int k = 8;
  while (k < 64 && double(k) < worst) k *= 2;
  return k;
}

static hg_status dem_create_impl(hg_ctx *ctx, const hg_dem_desc *descs, uint32_t nscenes,
                                 hg_dem **out) {
  if (!ctx || !descs || !out || nscenes == 0) return hg_fail(ctx, HG_ERR_INVALID, "null argument");
  for (uint32_t s = 0; s < nscenes; ++s) {
    const hg_dem_desc &d = descs[s];
    if (d.precision != HG_FP64 && d.precision != HG_FP32)
      return hg_fail(ctx, HG_ERR_INVALID, "precision must be HG_FP64 or HG_FP32");
    if (d.precision != descs[0].precision)
      return hg_fail(ctx, HG_ERR_INVALID, "all scenes of a batch share the precision");
    if (d.nplanes > uint32_t(kMaxPlanes))
      return hg_fail(ctx, HG_ERR_UNSUPPORTED, "at most 8 static planes per scene");
    if (d.nbodies && (!d.q || !d.v || !d.mass || !d.radius))
      return hg_fail(ctx, HG_ERR_INVALID, "body arrays missing");
    if (d.nplanes && (!d.plane_x || !d.plane_n))
      return hg_fail(ctx, HG_ERR_INVALID, "plane arrays missing");
    for (uint32_t k = 0; k < d.nbodies; ++k)
      if (!(d.radius[k] > 0.0) || !(d.mass[3 * k] > 0.0) || !(d.mass[3 * k + 2] > 0.0))
        return hg_fail(ctx, HG_ERR_INVALID, "radius, mass and inertia must be > 0");
  }
  HG_HIP(ctx, hipSetDevice(ctx->device));
  std::unique_ptr<hg_dem> h(new hg_dem());
  h->ctx = ctx;
  h->precision = descs[0].precision;
  h->nscenes = nscenes;
  h->cap = hg_dem_contact_capacity(descs, nscenes);
  HG_HIP(ctx, hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
  uint64_t b_off = 0, cell_off = 0;
  std::vector<uint32_t> h_b_off;
  for (uint32_t s = 0; s < nscenes; ++s) {
    const hg_dem_desc &d = descs[s];
    SceneDev sd;
    std::memset(&sd, 0, sizeof(sd));
    sd.g0 = d.gravity[0];
    sd.g1 = d.gravity[1];
    sd.kn = d.kn;
    sd.kt = d.kt;
    sd.gamman = d.gamman;
    sd.gammat = d.gammat;
    sd.mu = d.mu;
    dem_plan_grid(d, sd);
    sd.b_off = uint32_t(b_off);
    sd.b_cnt = d.nbodies;
    sd.cell_off = uint32_t(cell_off);
    sd.nplanes = int32_t(d.nplanes);
    for (uint32_t k = 0; k < d.nplanes; ++k) {
      const double n0 = d.plane_n[2 * k], n1 = d.plane_n[2 * k + 1];
      const double nn = std::sqrt(n0 * n0 + n1 * n1);
      if (!(std::fabs(nn - 1.0) <= 1e-6)) return hg_fail(ctx, HG_ERR_INVALID, "plane normals must be unit");
      sd.plane_x[2 * k] = d.plane_x[2 * k];
      sd.plane_x[2 * k + 1] = d.plane_x[2 * k + 1];
      sd.plane_n[2 * k] = n0;
      sd.plane_n[2 * k + 1] = n1;
    }
    h_b_off.push_back(uint32_t(b_off));
    b_off += d.nbodies;
    cell_off += uint64_t(sd.ncx) * uint64_t(sd.ncy);
    h->scenes.push_back(sd);
  }
  if (b_off >= (1ull << 31) || cell_off >= (1ull << 31))
    return hg_fail(ctx, HG_ERR_UNSUPPORTED, "batch too large for 32-bit indexing");
  h_b_off.push_back(uint32_t(b_off));
  h->nbodies = uint32_t(b_off);
  h->ncell = uint32_t(cell_off);
  hg_dem *hp = h.get();
  HG_TRY(dalloc(hp, &h->d_scenes, nscenes * sizeof(SceneDev)));
  HG_TRY(dalloc(hp, &h->d_b_off, (nscenes + 1) * 4));
  HG_TRY(dalloc(hp, &h->err, nscenes * 4));
  HG_TRY(dalloc(hp, &h->F_out, size_t(h->nbodies) * 24));
  h->kslots = dem_table_slots(descs, nscenes);
  for (int c = 0; c < 3; ++c) {
    TableRaw &t = h->tables[c];
    HG_TRY(dalloc(hp, &t.count, size_t(h->ncell) * 4));
    HG_TRY(dalloc(hp, &t.items, size_t(h->ncell) * size_t(h->kslots) * 4));
    HG_TRY(dalloc(hp, &t.bkey, size_t(h->nbodies) * 4));
    HG_TRY(dalloc(hp, &t.novf, 4));
    HG_TRY(dalloc(hp, &t.ovf, size_t(h->nbodies) * 4));
    HG_HIP(ctx, hipMemset(t.count, 0, size_t(h->ncell) * 4));
    HG_HIP(ctx, hipMemset(t.bkey, 0, size_t(h->nbodies ? h->nbodies : 1) * 4));
    HG_HIP(ctx, hipMemset(t.novf, 0, 4));
  }
  HG_HIP(ctx, hipMemcpy(h->d_scenes, h->scenes.data(), nscenes * sizeof(SceneDev), hipMemcpyHostToDevice));
  HG_HIP(ctx, hipMemcpy(h->d_b_off, h_b_off.data(), (nscenes + 1) * 4, hipMemcpyHostToDevice));
  HG_HIP(ctx, hipMemset(h->err, 0, nscenes * 4));
  HG_HIP(ctx, hipMemset(h->F_out, 0, size_t(h->nbodies ? h->nbodies : 1) * 24));
  if (h->precision == HG_FP64) HG_TRY(upload_bodies<double>(hp, descs));
  else HG_TRY(upload_bodies<float>(hp, descs));
  // the copies and memsets above ran on the null stream, which the handle's
  // non-blocking stream does not wait for; other handles' work is not waited on
  HG_HIP(ctx, hipStreamSynchronize(nullptr));
  hg_dem_launch_rebin(hp);
  HG_HIP(ctx, hipGetLastError());
  HG_HIP(ctx, hipStreamSynchronize(h->stream));
  *out = h.release();
  return HG_OK;
}

static hg_status dem_check(hg_dem *h) {
  HG_HIP(h->ctx, hipStreamSynchronize(h->stream));
  std::vector<int32_t> e(h->nscenes);
  HG_HIP(h->ctx, hipMemcpy(e.data(), h->err, h->nscenes * 4, hipMemcpyDeviceToHost));
  for (uint32_t s = 0; s < h->nscenes; ++s) {
    if (e[s] & 4)
      return hg_fail(h->ctx, HG_ERR_UNSUPPORTED, "scene " + std::to_string(s) + ": more than " +
                                                     std::to_string(h->cap) + " contacts on one body");
    if (e[s] & 2) return hg_fail(h->ctx, HG_ERR_STATE, "scene " + std::to_string(s) + ": non-finite body state");
  }
  return HG_OK;
}

extern "C" hg_status hg_dem_create(hg_ctx *ctx, const hg_dem_desc *desc, hg_dem **out) {
  return dem_create_impl(ctx, desc, 1, out);
}

extern "C" hg_status hg_dem_create_batch(hg_ctx *ctx, const hg_dem_desc *descs, uint32_t nscenes,
                                         hg_dem **out) {
  return dem_create_impl(ctx, descs, nscenes, out);
}

extern "C" hg_status hg_dem_destroy(hg_dem *h) {
  if (!h) return HG_OK;
  if (h->stream) hipStreamSynchronize(h->stream);
  if (h->graph) hipGraphExecDestroy(h->graph);
  for (void *p : h->allocs) hipFree(p);
  if (h->stream && h->own_stream) hipStreamDestroy(h->stream);
  delete h;
  return HG_OK;
}

extern "C" hg_status hg_dem_step(hg_dem *h, double dt, uint32_t nsteps) {
  if (!h) return HG_ERR_INVALID;
  if (!(dt > 0.0)) return hg_fail(h->ctx, HG_ERR_INVALID, "dt must be > 0");
  HG_TRY(dem_upload_dt(h, dt));
  for (uint32_t k = 0; k < nsteps; ++k) launch_step(h);
  HG_HIP(h->ctx, hipGetLastError());
  return HG_OK;
}

extern "C" hg_status hg_dem_step_graph(hg_dem *h, double dt, uint32_t nsteps) {
  if (!h) return HG_ERR_INVALID;
  if (!(dt > 0.0)) return hg_fail(h->ctx, HG_ERR_INVALID, "dt must be > 0");
  if (h->prof.on) return hg_dem_step(h, dt, nsteps);
  HG_TRY(dem_upload_dt(h, dt));
  // The buffers a step uses repeat every six steps: the contact slots and
  // the state buffers alternate, the tables rotate by three.  A graph of one
  // period, captured from phase 0, replays from any later phase 0; steps
  // before and after run without it.
  const uint32_t period = 6u;
  auto phase0 = [&]() { return h->cur == 0 && h->tph == 0; };
  while (!phase0() && nsteps) {
    launch_step(h);
    --nsteps;
  }
  const uint32_t reps = nsteps / period;
  if (reps) {
    if (!h->graph) {
      hipGraph_t g = nullptr;
      HG_HIP(h->ctx, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
      for (uint32_t k = 0; k < period; ++k) launch_step(h);
      HG_HIP(h->ctx, hipStreamEndCapture(h->stream, &g));
      HG_HIP(h->ctx, hipGraphInstantiate(&h->graph, g, nullptr, nullptr, 0));
      HG_HIP(h->ctx, hipGraphDestroy(g));
    }
    for (uint32_t k = 0; k < reps; ++k) HG_HIP(h->ctx, hipGraphLaunch(h->graph, h->stream));
  }
  for (uint32_t k = 0; k < nsteps % period; ++k) launch_step(h);
  HG_HIP(h->ctx, hipGetLastError());
  return HG_OK;
}

extern "C" hg_status hg_dem_sync(hg_dem *h) {
  if (!h) return HG_ERR_INVALID;
  return dem_check(h);
}

template <typename T>
static hg_status set_weights_t(hg_dem *h, uint32_t scene, const double *w) {
  const SceneDev &sc = h->scenes[scene];
  std::vector<T> tw(sc.b_cnt);
  for (uint32_t k = 0; k < sc.b_cnt; ++k) tw[k] = T(w[k]);
  HG_HIP(h->ctx, hipStreamSynchronize(h->stream));
  if (sc.b_cnt)
    HG_HIP(h->ctx, hipMemcpy(static_cast<T *>(h->b.w) + sc.b_off, tw.data(), sc.b_cnt * sizeof(T),
                             hipMemcpyHostToDevice));
  return HG_OK;
}

extern "C" hg_status hg_dem_set_weights(hg_dem *h, uint32_t scene, const double *weights) {
  if (!h || scene >= h->nscenes) return HG_ERR_INVALID;
  HG_HIP(h->ctx, hipStreamSynchronize(h->stream));
  h->scenes[scene].weighted = weights ? 1 : 0;
  if (weights) {
    if (h->precision == HG_FP64) HG_TRY(set_weights_t<double>(h, scene, weights));
    else HG_TRY(set_weights_t<float>(h, scene, weights));
  }
  HG_HIP(h->ctx, hipMemcpy(h->d_scenes, h->scenes.data(), h->scenes.size() * sizeof(SceneDev),
                           hipMemcpyHostToDevice));
  if (h->graph) {
    hipGraphExecDestroy(h->graph);
    h->graph = nullptr;
  }
  return HG_OK;
}

template <typename T>
static hg_status download_bodies_t(hg_dem *h, uint32_t scene, double *q, double *v) {
  const SceneDev &sc = h->scenes[scene];
  const uint32_t n = sc.b_cnt, off = sc.b_off;
  if (!n) return HG_OK;
  std::vector<double> x(n), y(n), th(n);
  std::vector<T> vx(n), vy(n), om(n);
  HG_HIP(h->ctx, hipMemcpy(x.data(), h->b.x + off, n * 8, hipMemcpyDeviceToHost));
  HG_HIP(h->ctx, hipMemcpy(y.data(), h->b.y + off, n * 8, hipMemcpyDeviceToHost));
  HG_HIP(h->ctx, hipMemcpy(th.data(), h->b.th + off, n * 8, hipMemcpyDeviceToHost));
  HG_HIP(h->ctx, hipMemcpy(vx.data(), static_cast<T *>(h->b.vx) + off, n * sizeof(T), hipMemcpyDeviceToHost));
  HG_HIP(h->ctx, hipMemcpy(vy.data(), static_cast<T *>(h->b.vy) + off, n * sizeof(T), hipMemcpyDeviceToHost));
  HG_HIP(h->ctx, hipMemcpy(om.data(), static_cast<T *>(h->b.om) + off, n * sizeof(T), hipMemcpyDeviceToHost));
  for (uint32_t k = 0; k < n; ++k) {
    if (q) {
      q[3 * k] = x[k];
      q[3 * k + 1] = y[k];
      q[3 * k + 2] = th[k];
    }
    if (v) {
      v[3 * k] = double(vx[k]);
      v[3 * k + 1] = double(vy[k]);
      v[3 * k + 2] = double(om[k]);
    }
  }
  return HG_OK;
}

extern "C" hg_status hg_dem_download_bodies(hg_dem *h, uint32_t scene, double *q, double *v) {
  if (!h || scene >= h->nscenes) return HG_ERR_INVALID;
  HG_TRY(dem_check(h));
  if (h->precision == HG_FP64) return download_bodies_t<double>(h, scene, q, v);
  return download_bodies_t<float>(h, scene, q, v);
}

template <typename T>
static hg_status download_contacts_t(hg_dem *h, uint32_t scene, uint32_t max, int32_t *type,
                                     int32_t *ci, int32_t *cj, double *delta_s, uint32_t *count) {
  const SceneDev &sc = h->scenes[scene];
  const uint32_t n = sc.b_cnt, off = sc.b_off;
  const size_t C = size_t(h->cap);
  const SlotsRaw &s = h->slots[h->cur];
  std::vector<uint32_t> okey(size_t(n) * C);
  std::vector<T> ods(size_t(n) * C * 2), pds(size_t(n) * kMaxPlanes * 2);
  std::vector<uint8_t> ocnt(n), pact(size_t(n) * kMaxPlanes);
  if (n) {
    HG_HIP(h->ctx, hipMemcpy(okey.data(), s.okey + size_t(off) * C, okey.size() * 4, hipMemcpyDeviceToHost));
    HG_HIP(h->ctx, hipMemcpy(ods.data(), static_cast<T *>(s.ods) + size_t(off) * C * 2, ods.size() * sizeof(T), hipMemcpyDeviceToHost));
    HG_HIP(h->ctx, hipMemcpy(ocnt.data(), s.ocnt + off, n, hipMemcpyDeviceToHost));
    HG_HIP(h->ctx, hipMemcpy(pact.data(), s.pact + size_t(off) * kMaxPlanes, pact.size(), hipMemcpyDeviceToHost));
    HG_HIP(h->ctx, hipMemcpy(pds.data(), static_cast<T *>(s.pds) + size_t(off) * kMaxPlanes * 2, pds.size() * sizeof(T), hipMemcpyDeviceToHost));
  }
  uint32_t c = 0;
  auto emit = [&](int t, uint32_t i, uint32_t j, double d0, double d1) {
    if (c < max) {
      if (type) type[c] = t;
      if (ci) ci[c] = int32_t(i);
      if (cj) cj[c] = int32_t(j);
      if (delta_s) {
        delta_s[2 * c] = d0;
        delta_s[2 * c + 1] = d1;
      }
    }
    ++c;
  };
  // reference iteration order: planes, circle pairs, kinematic pairs, each by owner
  for (uint32_t i = 0; i < n; ++i)
    for (int k = 0; k < sc.nplanes; ++k)
      if (pact[size_t(i) * kMaxPlanes + k])
        emit(0, i, uint32_t(k), double(pds[2 * (size_t(i) * kMaxPlanes + k)]),
             double(pds[2 * (size_t(i) * kMaxPlanes + k) + 1]));
  for (int kin = 0; kin < 2; ++kin)
    for (uint32_t i = 0; i < n; ++i)
      for (int q = 0; q < ocnt[i]; ++q) {
        const size_t sl = size_t(i) * C + q;
        const bool isk = (okey[sl] & kKinematic) != 0;
        if (isk != (kin == 1)) continue;
        emit(kin ? 2 : 1, i, okey[sl] & ~kKinematic, double(ods[2 * sl]), double(ods[2 * sl + 1]));
      }
  *count = c;
  return HG_OK;
}

extern "C" hg_status hg_dem_download_contacts(hg_dem *h, uint32_t scene, uint32_t max,
                                              int32_t *type, int32_t *i, int32_t *j,
                                              double *delta_s, uint32_t *count) {
  if (!h || scene >= h->nscenes || !count) return HG_ERR_INVALID;
  HG_TRY(dem_check(h));
  if (h->precision == HG_FP64) return download_contacts_t<double>(h, scene, max, type, i, j, delta_s, count);
  return download_contacts_t<float>(h, scene, max, type, i, j, delta_s, count);
}

template <typename T>
static hg_status contact_forces_t(hg_dem *h, uint32_t scene, uint32_t max, double *force,
                                  double *point, uint32_t *count) {
  const SceneDev &sc = h->scenes[scene];
  const uint32_t n = sc.b_cnt, off = sc.b_off;
  const size_t C = size_t(h->cap);
  const SlotsRaw &s = h->slots[h->cur];
  std::vector<uint32_t> okey(size_t(n) * C);
  std::vector<T> oF(size_t(n) * C * 2);
  std::vector<double> cp(h->cpoint ? size_t(n) * C * 2 : 0);
  std::vector<uint8_t> ocnt(n), pact(size_t(n) * kMaxPlanes);
  if (n) {
    HG_HIP(h->ctx, hipMemcpy(okey.data(), s.okey + size_t(off) * C, okey.size() * 4, hipMemcpyDeviceToHost));
    HG_HIP(h->ctx, hipMemcpy(ocnt.data(), s.ocnt + off, n, hipMemcpyDeviceToHost));
    HG_HIP(h->ctx, hipMemcpy(pact.data(), s.pact + size_t(off) * kMaxPlanes, pact.size(), hipMemcpyDeviceToHost));
    HG_HIP(h->ctx, hipMemcpy(oF.data(), static_cast<T *>(h->f.oF) + size_t(off) * C * 2, oF.size() * sizeof(T), hipMemcpyDeviceToHost));
    if (!cp.empty())
      HG_HIP(h->ctx, hipMemcpy(cp.data(), h->cpoint + size_t(off) * C * 2, cp.size() * 8, hipMemcpyDeviceToHost));
  }
  uint32_t c = 0;
  auto emit = [&](double f0, double f1, double p0, double p1) {
    if (c < max) {
      if (force) {
        force[2 * c] = f0;
        force[2 * c + 1] = f1;
      }
      if (point) {
        point[2 * c] = p0;
        point[2 * c + 1] = p1;
      }
    }
    ++c;
  };
  // the order of hg_dem_download_contacts
  for (uint32_t i = 0; i < n; ++i)
    for (int k = 0; k < sc.nplanes; ++k)
      if (pact[size_t(i) * kMaxPlanes + k]) emit(0.0, 0.0, 0.0, 0.0);
  for (int kin = 0; kin < 2; ++kin)
    for (uint32_t i = 0; i < n; ++i)
      for (int q = 0; q < ocnt[i]; ++q) {
        const size_t sl = size_t(i) * C + q;
        if (((okey[sl] & kKinematic) != 0) != (kin == 1)) continue;
        emit(double(oF[2 * sl]), double(oF[2 * sl + 1]), cp.empty() ? 0.0 : cp[2 * sl],
             cp.empty() ? 0.0 : cp[2 * sl + 1]);
      }
  *count = c;
  return HG_OK;
}

extern "C" hg_status hg_dem_download_contact_forces(hg_dem *h, uint32_t scene, uint32_t max,
                                                    double *force, double *point,
                                                    uint32_t *count) {
  if (!h || scene >= h->nscenes || !count) return HG_ERR_INVALID;
  HG_TRY(dem_check(h));
  if (h->precision == HG_FP64) return contact_forces_t<double>(h, scene, max, force, point, count);
  return contact_forces_t<float>(h, scene, max, force, point, count);
}

template <typename T>
static hg_status upload_contacts_t(hg_dem *h, uint32_t scene, uint32_t count, const int32_t *type,
                                   const int32_t *ci, const int32_t *cj, const double *ds) {
  const SceneDev &sc = h->scenes[scene];
  const uint32_t n = sc.b_cnt, off = sc.b_off;
  const size_t C = size_t(h->cap);
  std::vector<uint32_t> okey(size_t(n) * C, 0u);
  std::vector<T> ods(size_t(n) * C * 2, T(0)), pds(size_t(n) * kMaxPlanes * 2, T(0));
  std::vector<uint8_t> ocnt(n, 0), pact(size_t(n) * kMaxPlanes, 0);
  for (uint32_t c = 0; c < count; ++c) {
    const int32_t t = type[c], i = ci[c], j = cj[c];
    if (i < 0 || uint32_t(i) >= n)
      return hg_fail(h->ctx, HG_ERR_INVALID, "contact owner out of range");
    if (t == 0) {
      if (j < 0 || j >= sc.nplanes) return hg_fail(h->ctx, HG_ERR_INVALID, "contact plane out of range");
      const size_t sl = size_t(i) * kMaxPlanes + size_t(j);
      pact[sl] = 1;
      pds[2 * sl] = T(ds[2 * c]);
      pds[2 * sl + 1] = T(ds[2 * c + 1]);
    } else if (t == 1 || t == 2) {
      if (j < 0 || uint32_t(j) >= n || (t == 1 && j <= i))
        return hg_fail(h->ctx, HG_ERR_INVALID, "contact partner out of range");
      if (ocnt[i] >= C)
        return hg_fail(h->ctx, HG_ERR_UNSUPPORTED, "more contacts on one body than its capacity (" + std::to_string(C) + ")");
      const size_t sl = size_t(i) * C + ocnt[i]++;
      okey[sl] = uint32_t(j) | (t == 2 ? kKinematic : 0u);
      ods[2 * sl] = T(ds[2 * c]);
      ods[2 * sl + 1] = T(ds[2 * c + 1]);
    } else {
      return hg_fail(h->ctx, HG_ERR_INVALID, "unknown contact type");
    }
  }
  // each body's contacts sorted by partner key, as the step keeps them
  for (uint32_t i = 0; i < n; ++i) {
    const size_t b = size_t(i) * C;
    for (int a = 1; a < ocnt[i]; ++a)
      for (int k = a; k > 0 && okey[b + k - 1] > okey[b + k]; --k) {
        std::swap(okey[b + k - 1], okey[b + k]);
        std::swap(ods[2 * (b + k - 1)], ods[2 * (b + k)]);
        std::swap(ods[2 * (b + k - 1) + 1], ods[2 * (b + k) + 1]);
      }
  }
  const SlotsRaw &s = h->slots[h->cur];
  if (n) {
    HG_HIP(h->ctx, hipMemcpy(s.okey + size_t(off) * C, okey.data(), okey.size() * 4, hipMemcpyHostToDevice));
    HG_HIP(h->ctx, hipMemcpy(static_cast<T *>(s.ods) + size_t(off) * C * 2, ods.data(), ods.size() * sizeof(T), hipMemcpyHostToDevice));
    HG_HIP(h->ctx, hipMemcpy(s.ocnt + off, ocnt.data(), n, hipMemcpyHostToDevice));
    HG_HIP(h->ctx, hipMemcpy(s.pact + size_t(off) * kMaxPlanes, pact.data(), pact.size(), hipMemcpyHostToDevice));
    HG_HIP(h->ctx, hipMemcpy(static_cast<T *>(s.pds) + size_t(off) * kMaxPlanes * 2, pds.data(), pds.size() * sizeof(T), hipMemcpyHostToDevice));
  }
  return HG_OK;
}

extern "C" hg_status hg_dem_upload_contacts(hg_dem *h, uint32_t scene, uint32_t count,
                                            const int32_t *type, const int32_t *i,
                                            const int32_t *j, const double *delta_s) {
  if (!h || scene >= h->nscenes || (count && (!type || !i || !j || !delta_s))) return HG_ERR_INVALID;
  HG_HIP(h->ctx, hipStreamSynchronize(h->stream));
  if (h->precision == HG_FP64) return upload_contacts_t<double>(h, scene, count, type, i, j, delta_s);
  return upload_contacts_t<float>(h, scene, count, type, i, j, delta_s);
}

extern "C" hg_status hg_dem_keep_force(hg_dem *h, int32_t keep) {
  if (!h) return HG_ERR_INVALID;
  HG_HIP(h->ctx, hipStreamSynchronize(h->stream));
  h->keep_force = keep != 0;
  if (h->graph) {
    hipGraphExecDestroy(h->graph);
    h->graph = nullptr;
  }
  return HG_OK;
}

extern "C" hg_status hg_dem_download_force(hg_dem *h, uint32_t scene, double *F) {
  if (!h || scene >= h->nscenes || !F) return HG_ERR_INVALID;
  HG_TRY(dem_check(h));
  const SceneDev &sc = h->scenes[scene];
  if (sc.b_cnt)
    HG_HIP(h->ctx, hipMemcpy(F, h->F_out + 3 * size_t(sc.b_off), size_t(sc.b_cnt) * 24, hipMemcpyDeviceToHost));
  return HG_OK;
}

extern "C" hg_status hg_dem_set_profiling(hg_dem *h, int32_t on) {
  if (!h) return HG_ERR_INVALID;
  HG_HIP(h->ctx, hipStreamSynchronize(h->stream));
  h->prof.reset();
  h->prof.on = on != 0;
  return HG_OK;
}

extern "C" hg_status hg_dem_hold(hg_dem *h, double us) {
  if (!h) return HG_ERR_INVALID;
  return hg_stream_hold(h->ctx, h->stream, us);
}

extern "C" hg_status hg_dem_kernel_times(hg_dem *h, int32_t max_kernels, char *names, double *ms,
                                         int64_t *launches, int32_t *nkernels) {
  if (!h) return HG_ERR_INVALID;
  return hg_profiler_report(h->prof, h->ctx, max_kernels, names, ms, launches, nkernels);
}

// Device bytes per body of the arrays dalloc gives a handle (bodies and the
// second state buffer, two contact-slot sets, the owned contacts' forces, the
// kept total force), for the estimate below.
static uint64_t dem_bytes_per_body(size_t T, int cap) {
  const uint64_t C = uint64_t(cap);
  const uint64_t bodies = 4 * 8 + 7 * T + 1 + 3 * 8 + 3 * T;
  const uint64_t slots = C * 4 + C * 2 * T + 1 + uint64_t(kMaxPlanes) + uint64_t(kMaxPlanes) * 2 * T;
  const uint64_t forces = C * 2 * T;
  return bodies + 2 * slots + forces + 24;
}

extern "C" hg_status hg_dem_estimate_bytes(const hg_dem_desc *descs, uint32_t nscenes,
                                           uint64_t *bytes) {
  if (!descs || !bytes || nscenes == 0) return HG_ERR_INVALID;
  uint64_t B = 0, cells = 0;
  for (uint32_t s = 0; s < nscenes; ++s) {
    SceneDev sd;
    std::memset(&sd, 0, sizeof(sd));
    dem_plan_grid(descs[s], sd);
    B += descs[s].nbodies;
    cells += uint64_t(sd.ncx) * uint64_t(sd.ncy);
  }
  const size_t T = descs[0].precision == HG_FP64 ? 8 : 4;
  // the three bucket tables: counts and item slots per bucket, bucket and
  // overflow entry per body
  const uint64_t tables = 3 * (cells * (uint64_t(dem_table_slots(descs, nscenes)) + 1) * 4 + B * 8 + 4);
  *bytes = B * dem_bytes_per_body(T, hg_dem_contact_capacity(descs, nscenes)) + tables +
           uint64_t(nscenes) * (sizeof(SceneDev) + 16) + (1u << 20);
  return HG_OK;
}

extern "C" hg_status hg_dem_device_bytes(hg_dem *h, uint64_t *bytes) {
  if (!h || !bytes) return HG_ERR_INVALID;
  *bytes = h->bytes;
  return HG_OK;
}
