/*
 * mpm_oracle.c -- TEST INFRASTRUCTURE: serial FP64 CPU restatement of the
 * mpmgrains2d ExplicitIntegrator step (mpmgrains2d/ExplicitIntegrator.cpp:5-67).
 * It is the checker for the HIP path and the CPU baseline; the product never
 * links it.  Compiled with -ffp-contract=off so every product and sum rounds
 * as written, in the reference's operation order.
 *
 * Parity pins for this file (tests/test_oracle_mpm.py): the closed form and
 * brute-force quadrature of the uGIMP overlap integral, partition of unity and
 * linear reproduction of every basis, P2G mass/momentum/angular-momentum
 * conservation (PhysicsGrid.cpp:207, :316-327), zero net internal force
 * (PhysicsGrid.cpp:545), exact free fall, plane sticking/sliding closed forms,
 * the return-map invariants (det preserved, yield surface hit exactly,
 * ConstitutiveModel.cpp:62-77) and the no-tension branch.  In
 * tests/test_oracle_pins.py: the internal force's sign and size (a uniformly
 * compressed block's nodal forces are its pressure integrated over its edges:
 * zero inside, -tau H outward) and the side of the PIC/FLIP blend (alpha 0
 * and 1 on linear grid fields).
 */
#include "hg_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---- small 2x2 helpers, column-major (m00, m10, m01, m11) ---------------- */
static double det2(const double *m) { return m[0] * m[3] - m[1] * m[2]; }
static double trace2(const double *m) { return m[0] + m[3]; }

/* ---- basis functions ------------------------------------------------------ */

/* uGIMPLinearBasisFunctions.cpp:6-27 (LinearIntegral) */
double oracle_ugimp_linear_integral(double xp, double hl, double xi, double w) {
  const double diff = fabs(xp - xi);
  if (diff >= w + hl) {
    return 0.0;
  } else if (diff >= w - hl) {
    return (w + hl - diff) * (w + hl - diff) / (2.0 * w);
  } else if (diff >= hl) {
    return 2.0 * hl * (1.0 - diff / w);
  } else {
    return 2.0 * hl - (hl * hl + diff * diff) / w;
  }
}

/* uGIMPLinearBasisFunctions.cpp:29-47 (LinearIntegralGrad) */
double oracle_ugimp_linear_integral_grad(double xp, double hl, double xi, double w) {
  const double diff = fabs(xp - xi);
  const double sgn = (xp - xi) >= 0.0 ? 1.0 : -1.0;
  if (diff >= w + hl) {
    return 0.0;
  } else if (diff >= w - hl) {
    return -sgn * (w + hl - diff) / w;
  } else if (diff >= hl) {
    return -sgn * 2.0 * hl / w;
  } else {
    return 2.0 * (xi - xp) / w;
  }
}

/* LinearBasisFunctions.cpp:6-23 */
static double linear_n(double x) {
  const double absx = fabs(x);
  return absx < 1.0 ? 1.0 - absx : 0.0;
}
static double linear_dn(double x) {
  if (fabs(x) <= 1.0) return (x >= 0.0) ? -1.0 : 1.0;
  return 0.0;
}

/* CubicBasisFunctions.cpp:6-46 */
static double cubic_n(double x) {
  const double absx = fabs(x);
  if (absx < 1.0) return (0.5 * absx - 1.0) * absx * absx + 2.0 / 3.0;
  if (absx < 2.0) return ((-absx / 6.0 + 1.0) * absx - 2.0) * absx + 4.0 / 3.0;
  return 0.0;
}
static double cubic_dn(double x) {
  const double absx = fabs(x);
  const double sgnx = (x >= 0.0) ? 1.0 : -1.0;
  if (absx < 1.0) return sgnx * (1.5 * absx - 2.0) * absx;
  if (absx < 2.0) return sgnx * ((-0.5 * absx + 2.0) * absx - 2.0);
  return 0.0;
}

void oracle_stencil(const oracle_mpm *s, const double xp[2], double hl,
                    int64_t lo[2], int64_t hi[2]) {
  for (int d = 0; d < 2; ++d) {
    if (s->basis == ORACLE_BASIS_UGIMP_LINEAR) {
      /* uGIMPLinearBasisFunctions.cpp:68-90 */
      const double xm = xp[d] - hl;
      const double xM = xp[d] + hl;
      lo[d] = (int64_t)floor((xm - s->grid_min[d]) / s->cell_width);
      hi[d] = (int64_t)ceil((xM - s->grid_min[d]) / s->cell_width);
    } else if (s->basis == ORACLE_BASIS_LINEAR) {
      /* LinearBasisFunctions.cpp:48-60 */
      lo[d] = (int64_t)floor((xp[d] - s->grid_min[d]) / s->cell_width);
      hi[d] = lo[d] + 1;
    } else {
      /* CubicBasisFunctions.cpp:61-75 */
      lo[d] = (int64_t)floor((xp[d] - s->grid_min[d]) / s->cell_width) - 1;
      hi[d] = lo[d] + 3;
    }
  }
}

double oracle_weight(const oracle_mpm *s, const double xp[2], double hl,
                     int64_t ix, int64_t iy) {
  const double h = s->cell_width;
  if (s->basis == ORACLE_BASIS_UGIMP_LINEAR) {
    /* uGIMPLinearBasisFunctions.cpp:108-123 */
    const double xi0 = h * (double)ix + s->grid_min[0];
    const double xi1 = h * (double)iy + s->grid_min[1];
    const double Vp = 4.0 * hl * hl;
    return oracle_ugimp_linear_integral(xp[0], hl, xi0, h) *
           oracle_ugimp_linear_integral(xp[1], hl, xi1, h) / Vp;
  }
  const double d0 = (xp[0] - h * (double)ix - s->grid_min[0]) / h;
  const double d1 = (xp[1] - h * (double)iy - s->grid_min[1]) / h;
  if (s->basis == ORACLE_BASIS_LINEAR) return linear_n(d0) * linear_n(d1); /* Linear :79-94 */
  return cubic_n(d0) * cubic_n(d1);                                        /* Cubic :96-110 */
}

void oracle_weight_grad(const oracle_mpm *s, const double xp[2], double hl,
                        int64_t ix, int64_t iy, double g[2]) {
  const double h = s->cell_width;
  if (s->basis == ORACLE_BASIS_UGIMP_LINEAR) {
    /* uGIMPLinearBasisFunctions.cpp:141-156 */
    const double xi0 = h * (double)ix + s->grid_min[0];
    const double xi1 = h * (double)iy + s->grid_min[1];
    const double Vp = 4.0 * hl * hl;
    const double wx = oracle_ugimp_linear_integral(xp[0], hl, xi0, h);
    const double wy = oracle_ugimp_linear_integral(xp[1], hl, xi1, h);
    g[0] = wy * oracle_ugimp_linear_integral_grad(xp[0], hl, xi0, h) / Vp;
    g[1] = wx * oracle_ugimp_linear_integral_grad(xp[1], hl, xi1, h) / Vp;
    return;
  }
  const double d0 = (xp[0] - h * (double)ix - s->grid_min[0]) / h;
  const double d1 = (xp[1] - h * (double)iy - s->grid_min[1]) / h;
  if (s->basis == ORACLE_BASIS_LINEAR) {
    /* LinearBasisFunctions.cpp:113-135 */
    const double wx = linear_n(d0), wy = linear_n(d1);
    g[0] = wy * linear_dn(d0) / h;
    g[1] = wx * linear_dn(d1) / h;
  } else {
    /* CubicBasisFunctions.cpp:129-146 */
    const double wx = cubic_n(d0), wy = cubic_n(d1);
    g[0] = wy * cubic_dn(d0) / h;
    g[1] = wx * cubic_dn(d1) / h;
  }
}

static uint32_t nx_nodes(const oracle_mpm *s) { return s->cell_count[0] + 1; }
static uint64_t num_nodes(const oracle_mpm *s) {
  return (uint64_t)(s->cell_count[0] + 1) * (uint64_t)(s->cell_count[1] + 1);
}

/* ---- grid phase ----------------------------------------------------------- */

/* PhysicsGrid.cpp:107-114 */
void oracle_mpm_clear_rasterized_data(oracle_mpm *s) {
  const uint64_t G = num_nodes(s);
  memset(s->mass, 0, G * sizeof(double));
  memset(s->mom, 0, 2 * G * sizeof(double));
  memset(s->mom_new, 0, 2 * G * sizeof(double));
  memset(s->force, 0, 2 * G * sizeof(double));
  memset(s->vel, 0, 2 * G * sizeof(double));
  memset(s->acc, 0, 2 * G * sizeof(double));
}

/* PhysicsGrid.cpp:168-207 (serial branch) */
void oracle_mpm_rasterize_point_masses(oracle_mpm *s) {
  const uint32_t nx = nx_nodes(s);
  for (uint32_t p = 0; p < s->npoints; ++p) {
    const double *xp = &s->q[2 * p];
    int64_t lo[2], hi[2];
    oracle_stencil(s, xp, s->hl[p], lo, hi);
    for (int64_t iy = lo[1]; iy <= hi[1]; ++iy)
      for (int64_t ix = lo[0]; ix <= hi[0]; ++ix) {
        const double w = oracle_weight(s, xp, s->hl[p], ix, iy);
        s->mass[iy * nx + ix] += w * s->m[p];
      }
  }
}

/* PhysicsGrid.cpp:274-311 (serial branch) */
void oracle_mpm_rasterize_point_momentum(oracle_mpm *s) {
  const uint32_t nx = nx_nodes(s);
  for (uint32_t p = 0; p < s->npoints; ++p) {
    const double *xp = &s->q[2 * p];
    int64_t lo[2], hi[2];
    oracle_stencil(s, xp, s->hl[p], lo, hi);
    const double pm0 = s->m[p] * s->v[2 * p];
    const double pm1 = s->m[p] * s->v[2 * p + 1];
    for (int64_t iy = lo[1]; iy <= hi[1]; ++iy)
      for (int64_t ix = lo[0]; ix <= hi[0]; ++ix) {
        const double w = oracle_weight(s, xp, s->hl[p], ix, iy);
        const int64_t n = iy * nx + ix;
        s->mom[2 * n] += w * pm0;
        s->mom[2 * n + 1] += w * pm1;
      }
  }
}

/* ConstitutiveModel.cpp:83-97 */
void oracle_compute_tau(const double be[4], double kappa, double mu,
                        int32_t just_homogenized, double tau[4]) {
  if (just_homogenized != 0) {
    tau[0] = tau[1] = tau[2] = tau[3] = 0.0;
    return;
  }
  const double detb = det2(be);
  const double J = sqrt(detb);
  const double nrm = pow(detb, 1.0 / 2.0);
  double bb[4];
  for (int k = 0; k < 4; ++k) bb[k] = be[k] / nrm;
  const double half_tr = trace2(bb) * 1.0 / 2.0;
  const double dev[4] = {bb[0] - half_tr, bb[1] - 0.0, bb[2] - 0.0, bb[3] - half_tr};
  const double c = kappa * 0.5 * (J + 1.0) * (J - 1.0);
  tau[0] = c + mu * dev[0];
  tau[1] = 0.0 + mu * dev[1];
  tau[2] = 0.0 + mu * dev[2];
  tau[3] = c + mu * dev[3];
}

/* MaterialPoints.cpp:183-202 */
void oracle_mpm_compute_cauchy_stress(oracle_mpm *s) {
  for (uint32_t p = 0; p < s->npoints; ++p) {
    double tau[4] = {0.0, 0.0, 0.0, 0.0};
    if (s->J[p] <= 1.0) {
      double be[4];
      for (int k = 0; k < 4; ++k) be[k] = s->be_bar[4 * p + k] * s->J[p];
      oracle_compute_tau(be, s->bulk_modulus, s->shear_modulus,
                         s->just_homogenized ? s->just_homogenized[p] : 0, tau);
    }
    for (int k = 0; k < 4; ++k) s->sigma[4 * p + k] = tau[k] / s->J[p];
  }
}

/* PhysicsGrid.cpp:413-473 (serial branch) */
void oracle_mpm_compute_forces(oracle_mpm *s) {
  const uint32_t nx = nx_nodes(s);
  const uint64_t G = num_nodes(s);
  for (uint32_t p = 0; p < s->npoints; ++p) {
    const double *xp = &s->q[2 * p];
    const double *sg = &s->sigma[4 * p];
    int64_t lo[2], hi[2];
    oracle_stencil(s, xp, s->hl[p], lo, hi);
    const double c = -s->volume[p] * s->J[p];
    for (int64_t iy = lo[1]; iy <= hi[1]; ++iy)
      for (int64_t ix = lo[0]; ix <= hi[0]; ++ix) {
        const double w = oracle_weight(s, xp, s->hl[p], ix, iy);
        if (w == 0.0) continue;
        double g[2];
        oracle_weight_grad(s, xp, s->hl[p], ix, iy, g);
        const double sw0 = sg[0] * g[0] + sg[2] * g[1];
        const double sw1 = sg[1] * g[0] + sg[3] * g[1];
        const int64_t n = iy * nx + ix;
        s->force[2 * n] += c * sw0;
        s->force[2 * n + 1] += c * sw1;
      }
  }
  for (uint64_t i = 0; i < G; ++i) {
    const int hyb = s->hybridized ? s->hybridized[i] : 0;
    if (hyb > 0) {
      s->force[2 * i] += 0.5 * s->mass[i] * s->gravity[0];
      s->force[2 * i + 1] += 0.5 * s->mass[i] * s->gravity[1];
    } else {
      s->force[2 * i] += s->mass[i] * s->gravity[0];
      s->force[2 * i + 1] += s->mass[i] * s->gravity[1];
    }
  }
  if (s->coupling_force) {
    for (uint64_t i = 0; i < G; ++i) {
      s->force[2 * i] += s->coupling_force[2 * i];
      s->force[2 * i + 1] += s->coupling_force[2 * i + 1];
    }
  }
}

/* PhysicsGrid.cpp:565-586 */
void oracle_mpm_update_momentum(oracle_mpm *s, double dt) {
  const uint64_t G = num_nodes(s);
  for (uint64_t i = 0; i < G; ++i) {
    if (s->mass[i] > 0.0) {
      s->mom_new[2 * i] = s->mom[2 * i] + dt * s->force[2 * i];
      s->mom_new[2 * i + 1] = s->mom[2 * i + 1] + dt * s->force[2 * i + 1];
    }
  }
}

/* PhysicsGrid.cpp:611-673 */
void oracle_mpm_grid_plane_collisions(oracle_mpm *s) {
  const uint32_t nx = nx_nodes(s);
  const uint64_t G = num_nodes(s);
  for (uint64_t i = 0; i < G; ++i) {
    const double px0 = s->grid_min[0] + s->cell_width * (double)(i % nx);
    const double px1 = s->grid_min[1] + s->cell_width * (double)(i / nx);
    for (uint32_t k = 0; k < s->nplanes; ++k) {
      if (px1 < s->plane_lower[k]) continue;
      const double *n = &s->plane_n[2 * k];
      const double *x0 = &s->plane_x[2 * k];
      const double d = n[0] * (px0 - x0[0]) + n[1] * (px1 - x0[1]);
      if (d > 0.0) continue;
      if (s->plane_behavior[k] == ORACLE_PLANE_SLIDING) {
        if (s->mass[i] <= 0.0) continue;
        double v0 = s->mom_new[2 * i] / s->mass[i];
        double v1 = s->mom_new[2 * i + 1] / s->mass[i];
        const double vn = n[0] * v0 + n[1] * v1;
        if (vn > 0.0) continue;
        v0 -= vn * n[0];
        v1 -= vn * n[1];
        s->mom_new[2 * i] = s->mass[i] * v0;
        s->mom_new[2 * i + 1] = s->mass[i] * v1;
      } else {
        s->mom_new[2 * i] = 0.0;
        s->mom_new[2 * i + 1] = 0.0;
      }
    }
  }
}

/* PhysicsGrid.cpp:588-609 */
void oracle_mpm_velocity_acceleration_update(oracle_mpm *s, double dt) {
  const uint64_t G = num_nodes(s);
  for (uint64_t i = 0; i < G; ++i) {
    if (s->mass[i] > 0.0) {
      s->vel[2 * i] = s->mom_new[2 * i] / s->mass[i];
      s->vel[2 * i + 1] = s->mom_new[2 * i + 1] / s->mass[i];
      const double dm = dt * s->mass[i];
      s->acc[2 * i] = (s->mom_new[2 * i] - s->mom[2 * i]) / dm;
      s->acc[2 * i + 1] = (s->mom_new[2 * i + 1] - s->mom[2 * i + 1]) / dm;
    } else {
      s->vel[2 * i] = s->vel[2 * i + 1] = 0.0;
      s->acc[2 * i] = s->acc[2 * i + 1] = 0.0;
    }
  }
}

/* ---- particle phase ------------------------------------------------------- */

/* MaterialPoints.cpp:296-322 */
void oracle_mpm_compute_vel_grad(oracle_mpm *s) {
  const uint32_t nx = nx_nodes(s);
  for (uint32_t p = 0; p < s->npoints; ++p) {
    const double *xp = &s->q[2 * p];
    int64_t lo[2], hi[2];
    oracle_stencil(s, xp, s->hl[p], lo, hi);
    double L[4] = {0.0, 0.0, 0.0, 0.0};
    for (int64_t iy = lo[1]; iy <= hi[1]; ++iy)
      for (int64_t ix = lo[0]; ix <= hi[0]; ++ix) {
        double g[2];
        oracle_weight_grad(s, xp, s->hl[p], ix, iy, g);
        const int64_t n = iy * nx + ix;
        const double v0 = s->vel[2 * n], v1 = s->vel[2 * n + 1];
        L[0] += v0 * g[0];
        L[1] += v1 * g[0];
        L[2] += v0 * g[1];
        L[3] += v1 * g[1];
      }
    for (int k = 0; k < 4; ++k) s->vel_grad[4 * p + k] = L[k];
  }
}

/* ConstitutiveModel.cpp:19-22: be* = be_n + dt (L be_n + be_n L^T) */
void oracle_elastic_prediction(const double b[4], const double L[4], double dt,
                               double be_star[4]) {
  /* (L b)_ij = L_i0 b_0j + L_i1 b_1j ; (b L^T)_ij = b_i0 L_j0 + b_i1 L_j1 */
  double Lb[4], bLt[4];
  Lb[0] = L[0] * b[0] + L[2] * b[1];
  Lb[1] = L[1] * b[0] + L[3] * b[1];
  Lb[2] = L[0] * b[2] + L[2] * b[3];
  Lb[3] = L[1] * b[2] + L[3] * b[3];
  bLt[0] = b[0] * L[0] + b[2] * L[2];
  bLt[1] = b[1] * L[0] + b[3] * L[2];
  bLt[2] = b[0] * L[1] + b[2] * L[3];
  bLt[3] = b[1] * L[1] + b[3] * L[3];
  for (int k = 0; k < 4; ++k) be_star[k] = b[k] + dt * (Lb[k] + bLt[k]);
}

/* MaterialPoints.cpp:290-294 */
static void normalized_determinant(const double A[4], double out[4]) {
  const double nrm = pow(det2(A), 1.0 / 2.0);
  for (int k = 0; k < 4; ++k) out[k] = A[k] / nrm;
}

/* MaterialPoints.cpp:324-348 */
void oracle_mpm_elastic_prediction(oracle_mpm *s, double dt) {
  for (uint32_t p = 0; p < s->npoints; ++p) {
    double bn[4], bs[4];
    for (int k = 0; k < 4; ++k) bn[k] = s->be_bar[4 * p + k] * s->J[p];
    oracle_elastic_prediction(bn, &s->vel_grad[4 * p], dt, bs);
    s->J[p] = sqrt(det2(bs));
    normalized_determinant(bs, &s->be_bar[4 * p]);
  }
}

/* ConstitutiveModel.cpp:26-81 (Drucker-Prager return on b_e) */
void oracle_plastic_correction(const double be_star[4], double kappa, double mu,
                               double alpha, int32_t *just_homogenized,
                               double be_next[4]) {
  if (just_homogenized && *just_homogenized != 0) {
    const double J = sqrt(det2(be_star));
    if (J > 1.0) {
      be_next[0] = J; be_next[1] = 0.0; be_next[2] = 0.0; be_next[3] = J;
      return;
    }
    *just_homogenized = 0;
  }
  const double J = sqrt(det2(be_star));
  double ys = 0.5 * alpha * kappa * (1.0 + J) * (1.0 - J);
  if (ys < 0.0) ys = 0.0;
  const double half_tr = 0.5 * trace2(be_star);
  const double dev[4] = {be_star[0] - half_tr, be_star[1] - 0.0,
                         be_star[2] - 0.0, be_star[3] - half_tr};
  double bar[4];
  for (int k = 0; k < 4; ++k) bar[k] = dev[k] / J;
  const double nrm = sqrt(bar[0] * bar[0] + bar[1] * bar[1] + bar[2] * bar[2] +
                          bar[3] * bar[3]);
  if (mu * nrm > ys) {
    const double lambda2 = ys / (mu * nrm);
    const double l1sq = det2(be_star) - lambda2 * lambda2 * det2(dev);
    const double lambda1 = sqrt(l1sq);
    be_next[0] = lambda1 + lambda2 * dev[0];
    be_next[1] = 0.0 + lambda2 * dev[1];
    be_next[2] = 0.0 + lambda2 * dev[2];
    be_next[3] = lambda1 + lambda2 * dev[3];
  } else {
    for (int k = 0; k < 4; ++k) be_next[k] = be_star[k];
  }
}

/* MaterialPoints.cpp:350-369 */
void oracle_mpm_plastic_correction(oracle_mpm *s) {
  for (uint32_t p = 0; p < s->npoints; ++p) {
    double bs[4], bn[4];
    for (int k = 0; k < 4; ++k) bs[k] = s->be_bar[4 * p + k] * s->J[p];
    oracle_plastic_correction(bs, s->bulk_modulus, s->shear_modulus, s->dp_alpha,
                              s->just_homogenized ? &s->just_homogenized[p] : NULL,
                              bn);
    s->J[p] = sqrt(det2(bn));
    normalized_determinant(bn, &s->be_bar[4 * p]);
  }
}

/* MaterialPoints.cpp:204-245 */
void oracle_mpm_update_velocities(oracle_mpm *s, double dt) {
  const uint32_t nx = nx_nodes(s);
  const double alpha = s->pic_alpha;
  for (uint32_t p = 0; p < s->npoints; ++p) {
    const double *xp = &s->q[2 * p];
    int64_t lo[2], hi[2];
    oracle_stencil(s, xp, s->hl[p], lo, hi);
    double vpic0 = 0.0, vpic1 = 0.0, a0 = 0.0, a1 = 0.0;
    for (int64_t iy = lo[1]; iy <= hi[1]; ++iy)
      for (int64_t ix = lo[0]; ix <= hi[0]; ++ix) {
        const double w = oracle_weight(s, xp, s->hl[p], ix, iy);
        const int64_t n = iy * nx + ix;
        vpic0 += w * s->vel[2 * n];
        vpic1 += w * s->vel[2 * n + 1];
        a0 += w * s->acc[2 * n];
        a1 += w * s->acc[2 * n + 1];
      }
    const double vf0 = s->v[2 * p] + a0 * dt;
    const double vf1 = s->v[2 * p + 1] + a1 * dt;
    if (!(s->kinematically_scripted && s->kinematically_scripted[p])) {
      s->v[2 * p] = (1.0 - alpha) * vpic0 + alpha * vf0;
      s->v[2 * p + 1] = (1.0 - alpha) * vpic1 + alpha * vf1;
    }
  }
}

/* MaterialPoints.cpp:247-288 */
void oracle_mpm_update_positions(oracle_mpm *s, double dt) {
  const uint32_t nx = nx_nodes(s);
  for (uint32_t p = 0; p < s->npoints; ++p) {
    double *xp = &s->q[2 * p];
    int64_t lo[2], hi[2];
    oracle_stencil(s, xp, s->hl[p], lo, hi);
    double g0 = 0.0, g1 = 0.0;
    for (int64_t iy = lo[1]; iy <= hi[1]; ++iy)
      for (int64_t ix = lo[0]; ix <= hi[0]; ++ix) {
        const double w = oracle_weight(s, xp, s->hl[p], ix, iy);
        const int64_t n = iy * nx + ix;
        g0 += w * s->vel[2 * n];
        g1 += w * s->vel[2 * n + 1];
      }
    if (!(s->kinematically_scripted && s->kinematically_scripted[p])) {
      xp[0] += dt * g0;
      xp[1] += dt * g1;
    } else {
      xp[0] += dt * s->v[2 * p];
      xp[1] += dt * s->v[2 * p + 1];
    }
  }
}

/* MaterialPoints.cpp:371-423 */
void oracle_mpm_point_plane_collisions(oracle_mpm *s) {
  for (uint32_t p = 0; p < s->npoints; ++p) {
    double *q = &s->q[2 * p];
    double *v = &s->v[2 * p];
    for (uint32_t k = 0; k < s->nplanes; ++k) {
      if (q[1] < s->plane_lower[k]) continue;
      const double *n = &s->plane_n[2 * k];
      const double *x0 = &s->plane_x[2 * k];
      const double d = n[0] * (q[0] - x0[0]) + n[1] * (q[1] - x0[1]);
      if (d > 0.0) continue;
      q[0] += -d * n[0];
      q[1] += -d * n[1];
      if (s->plane_behavior[k] == ORACLE_PLANE_SLIDING) {
        const double vn = n[0] * v[0] + n[1] * v[1];
        if (vn > 0.0) continue;
        const double nv0 = v[0] - vn * n[0];
        const double nv1 = v[1] - vn * n[1];
        v[0] = nv0;
        v[1] = nv1;
      } else {
        v[0] = 0.0;
        v[1] = 0.0;
      }
    }
  }
}

/* ---- ExplicitIntegrator --------------------------------------------------- */

void oracle_mpm_flow_zero_phase(oracle_mpm *s) {
  oracle_mpm_clear_rasterized_data(s);
  oracle_mpm_rasterize_point_masses(s);
  oracle_mpm_rasterize_point_momentum(s);
}

void oracle_mpm_flow_first_phase(oracle_mpm *s, double dt) {
  oracle_mpm_compute_cauchy_stress(s);
  oracle_mpm_compute_forces(s);
  oracle_mpm_update_momentum(s, dt);
  oracle_mpm_grid_plane_collisions(s);
}

void oracle_mpm_flow_second_phase(oracle_mpm *s, double dt) {
  oracle_mpm_velocity_acceleration_update(s, dt);
  oracle_mpm_compute_vel_grad(s);
  oracle_mpm_elastic_prediction(s, dt);
  oracle_mpm_plastic_correction(s);
  oracle_mpm_update_velocities(s, dt);
  oracle_mpm_update_positions(s, dt);
  oracle_mpm_point_plane_collisions(s);
}

void oracle_mpm_flow(oracle_mpm *s, double dt) {
  oracle_mpm_flow_zero_phase(s);
  oracle_mpm_flow_first_phase(s, dt);
  oracle_mpm_flow_second_phase(s, dt);
}

void oracle_mpm_flow_n(oracle_mpm *s, double dt, uint32_t nsteps) {
  for (uint32_t k = 0; k < nsteps; ++k) oracle_mpm_flow(s, dt);
}

int oracle_mpm_is_consistent(const oracle_mpm *s) {
  const double gmax0 = s->grid_min[0] + s->cell_width * (double)s->cell_count[0];
  const double gmax1 = s->grid_min[1] + s->cell_width * (double)s->cell_count[1];
  for (uint32_t p = 0; p < s->npoints; ++p) {
    const double x = s->q[2 * p], y = s->q[2 * p + 1];
    if (!isfinite(x) || !isfinite(y) || !isfinite(s->v[2 * p]) ||
        !isfinite(s->v[2 * p + 1]) || !isfinite(s->J[p]))
      return 1;
    if (x < s->grid_min[0] || y < s->grid_min[1] || x > gmax0 || y > gmax1) return 2;
    if (fabs(det2(&s->be_bar[4 * p]) - 1.0) > 1.0e-6) return 3;
  }
  return 0;
}

/* ---- OpenMP CPU baseline -------------------------------------------------- */
/* The reference's OpenMP structure: per-thread full-grid scatter buffers that
 * are zeroed, filled and reduced every transfer (PhysicsGrid.cpp:214-265), and
 * parallel-for loops over points and nodes elsewhere.  A step is three
 * phases, which the hybrid baseline interleaves with the grain work. */

/* Cleared grid, stress, and the mass, momentum and internal-force scatters
 * through T per-thread buffers (scratch: T * 5 * G doubles) and their
 * reduction.  The stress comes first so that the three scatters share one
 * pass over the buffers; it depends only on b and J. */
void oracle_mpm_omp_scatter(oracle_mpm *s, int T, double *scratch) {
#ifdef _OPENMP
  const uint32_t nx = nx_nodes(s);
  const uint64_t G = num_nodes(s);
  oracle_mpm_clear_rasterized_data(s);
#pragma omp parallel for num_threads(T)
  for (int t = 0; t < T; ++t) memset(scratch + (size_t)t * 5 * G, 0, 5 * G * sizeof(double));
#pragma omp parallel for num_threads(T)
  for (uint32_t p = 0; p < s->npoints; ++p) {
    double tau[4] = {0.0, 0.0, 0.0, 0.0};
    if (s->J[p] <= 1.0) {
      double be[4];
      for (int k = 0; k < 4; ++k) be[k] = s->be_bar[4 * p + k] * s->J[p];
      oracle_compute_tau(be, s->bulk_modulus, s->shear_modulus, 0, tau);
    }
    for (int k = 0; k < 4; ++k) s->sigma[4 * p + k] = tau[k] / s->J[p];
  }
#pragma omp parallel num_threads(T)
  {
    double *buf = scratch + (size_t)omp_get_thread_num() * 5 * G;
#pragma omp for
    for (uint32_t p = 0; p < s->npoints; ++p) {
      const double *xp = &s->q[2 * p];
      const double *sg = &s->sigma[4 * p];
      int64_t lo[2], hi[2];
      oracle_stencil(s, xp, s->hl[p], lo, hi);
      const double pm0 = s->m[p] * s->v[2 * p], pm1 = s->m[p] * s->v[2 * p + 1];
      const double c = -s->volume[p] * s->J[p];
      for (int64_t iy = lo[1]; iy <= hi[1]; ++iy)
        for (int64_t ix = lo[0]; ix <= hi[0]; ++ix) {
          const double w = oracle_weight(s, xp, s->hl[p], ix, iy);
          const int64_t n = iy * nx + ix;
          buf[n] += w * s->m[p];
          buf[G + 2 * n] += w * pm0;
          buf[G + 2 * n + 1] += w * pm1;
          if (w == 0.0) continue;
          double g[2];
          oracle_weight_grad(s, xp, s->hl[p], ix, iy, g);
          buf[3 * G + 2 * n] += c * (sg[0] * g[0] + sg[2] * g[1]);
          buf[3 * G + 2 * n + 1] += c * (sg[1] * g[0] + sg[3] * g[1]);
        }
    }
  }
#pragma omp parallel for num_threads(T)
  for (uint64_t i = 0; i < G; ++i) {
    for (int t = 0; t < T; ++t) {
      const double *buf = scratch + (size_t)t * 5 * G;
      s->mass[i] += buf[i];
      s->mom[2 * i] += buf[G + 2 * i];
      s->mom[2 * i + 1] += buf[G + 2 * i + 1];
      s->force[2 * i] += buf[3 * G + 2 * i];
      s->force[2 * i + 1] += buf[3 * G + 2 * i + 1];
    }
  }
#else
  (void)T;
  (void)scratch;
  oracle_mpm_flow_zero_phase(s);
  oracle_mpm_compute_cauchy_stress(s);
  oracle_mpm_compute_forces(s);
#endif
}

/* Gravity (halved at hybridised nodes), the coupling force, the momentum
 * update and the grid plane conditions (PhysicsGrid.cpp:544-673). */
void oracle_mpm_omp_grid_update(oracle_mpm *s, double dt, int T) {
  const uint64_t G = num_nodes(s);
#pragma omp parallel for num_threads(T)
  for (uint64_t i = 0; i < G; ++i) {
    const double gs = (s->hybridized && s->hybridized[i] > 0) ? 0.5 : 1.0;
    s->force[2 * i] += gs * s->mass[i] * s->gravity[0];
    s->force[2 * i + 1] += gs * s->mass[i] * s->gravity[1];
    if (s->coupling_force) {
      s->force[2 * i] += s->coupling_force[2 * i];
      s->force[2 * i + 1] += s->coupling_force[2 * i + 1];
    }
    if (s->mass[i] > 0.0) {
      s->mom_new[2 * i] = s->mom[2 * i] + dt * s->force[2 * i];
      s->mom_new[2 * i + 1] = s->mom[2 * i + 1] + dt * s->force[2 * i + 1];
    }
  }
  oracle_mpm_grid_plane_collisions(s);
}

/* Nodal velocity and acceleration, then the per-point passes of
 * flow_second_phase, each a parallel loop in the reference
 * (MaterialPoints.cpp:299, :328, :352, :208, :250, :373). */
void oracle_mpm_omp_point_update(oracle_mpm *s, double dt, int T) {
  const uint64_t G = num_nodes(s);
#pragma omp parallel for num_threads(T)
  for (uint64_t i = 0; i < G; ++i) {
    if (s->mass[i] > 0.0) {
      s->vel[2 * i] = s->mom_new[2 * i] / s->mass[i];
      s->vel[2 * i + 1] = s->mom_new[2 * i + 1] / s->mass[i];
      const double dm = dt * s->mass[i];
      s->acc[2 * i] = (s->mom_new[2 * i] - s->mom[2 * i]) / dm;
      s->acc[2 * i + 1] = (s->mom_new[2 * i + 1] - s->mom[2 * i + 1]) / dm;
    }
  }
#pragma omp parallel for num_threads(T)
  for (uint32_t p = 0; p < s->npoints; ++p) {
    oracle_mpm one = *s;
    one.npoints = 1;
    one.q += 2 * p; one.v += 2 * p; one.m += p; one.volume += p; one.hl += p;
    one.be_bar += 4 * p; one.J += p; one.sigma += 4 * p; one.vel_grad += 4 * p;
    one.just_homogenized = NULL;
    one.kinematically_scripted =
        s->kinematically_scripted ? s->kinematically_scripted + p : NULL;
    oracle_mpm_compute_vel_grad(&one);
    oracle_mpm_elastic_prediction(&one, dt);
    oracle_mpm_plastic_correction(&one);
    oracle_mpm_update_velocities(&one, dt);
    oracle_mpm_update_positions(&one, dt);
    oracle_mpm_point_plane_collisions(&one);
  }
}

int oracle_omp_threads(int nthreads) {
#ifdef _OPENMP
  return nthreads > 0 ? nthreads : omp_get_max_threads();
#else
  (void)nthreads;
  return 1;
#endif
}

void oracle_mpm_flow_n_omp(oracle_mpm *s, double dt, uint32_t nsteps, int nthreads) {
  const int T = oracle_omp_threads(nthreads);
  double *scratch = (double *)malloc((size_t)T * 5 * num_nodes(s) * sizeof(double));
  for (uint32_t step = 0; step < nsteps; ++step) {
    oracle_mpm_omp_scatter(s, T, scratch);
    oracle_mpm_omp_grid_update(s, dt, T);
    oracle_mpm_omp_point_update(s, dt, T);
  }
  free(scratch);
}
