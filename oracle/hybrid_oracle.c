/*
 * hybrid_oracle.c -- TEST INFRASTRUCTURE: serial FP64 CPU restatement of the
 * hybridgrains2dnew node-node coupled step (HybridGrains2DSim.cpp:116-227)
 * and HybridCoupling2D (Coupling.cpp:22-591), without the periodic zone
 * resampling (updateHybridState), which is outside the explicit step.
 * Checker only.
 *
 * Also the per-step homogenisation of the grains (ResamplingTools.cpp:744-884).
 *
 * Parity pins (tests/test_oracle_hybrid.py): the exchange conserves each
 * node's momentum and leaves equal MPM and DEM node velocities (the
 * reference's own Debug check, Coupling.cpp:566-577); grains the MPM mass does
 * not surround stay uncoupled; a coupled free fall stays exact.  Homogenised
 * stress: its node sum times the cell area equals the sum of the contacts'
 * stress tensors computed independently from the contact list (partition of
 * unity), and a compressed head-on pair gives -|f||r| along the axis;
 * homogenised velocity: grains moving together give that velocity at every
 * node they reach, once per step.  In tests/test_oracle_pins.py: the FLIP
 * transfer gives the grains the exchange multipliers' sum (continuum plus grain
 * momentum unchanged from before the exchange to after the transfer), coupled
 * positions advance from the pre-step position with the transferred velocity,
 * and one stencil node at or below 1e-6 keeps a grain uncoupled.
 */
#include "hg_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static uint64_t nodes(const oracle_mpm *s) {
  return (uint64_t)(s->cell_count[0] + 1) * (uint64_t)(s->cell_count[1] + 1);
}

static int node_valid(const oracle_mpm *s, int64_t ix, int64_t iy) {
  return ix >= 0 && iy >= 0 && ix <= (int64_t)s->cell_count[0] && iy <= (int64_t)s->cell_count[1];
}

/* the grains use the MPM basis with the first point's half width
 * (Coupling.cpp:31) */
static double grain_hl(const oracle_hybrid *h) {
  return h->mpm->npoints ? h->mpm->hl[0] : 0.25 * h->mpm->cell_width;
}

void oracle_hybrid_rasterize_grain_mass(oracle_hybrid *h) {
  oracle_mpm *s = h->mpm;
  oracle_dem *d = h->dem;
  const uint32_t nx = s->cell_count[0] + 1;
  const double hl = grain_hl(h);
  memset(h->dem_mass_grid, 0, nodes(s) * sizeof(double));
  if (s->hybridized) memset(s->hybridized, 0, nodes(s) * sizeof(int32_t));
  for (uint32_t b = 0; b < d->nbodies; ++b) {
    h->coupled[b] = 0;
    if (d->fixed && d->fixed[b]) continue;
    const double *qb = &h->dem_pos_before[3 * b];
    int64_t lo[2], hi[2];
    oracle_stencil(s, qb, hl, lo, hi);
    int all_pos = 1;
    for (int64_t iy = lo[1]; iy <= hi[1]; ++iy)
      for (int64_t ix = lo[0]; ix <= hi[0]; ++ix) {
        if (!node_valid(s, ix, iy)) continue;
        if (s->mass[iy * nx + ix] <= 1.0e-6) all_pos = 0;
      }
    if (!all_pos) continue;
    h->coupled[b] = 1;
    const double m = d->mass[3 * b];
    for (int64_t iy = lo[1]; iy <= hi[1]; ++iy)
      for (int64_t ix = lo[0]; ix <= hi[0]; ++ix) {
        if (!node_valid(s, ix, iy)) continue;
        const double w = oracle_weight(s, qb, hl, ix, iy);
        h->dem_mass_grid[iy * nx + ix] += w * m;
      }
  }
}

void oracle_hybrid_rasterize_grain_mom(oracle_hybrid *h) {
  oracle_mpm *s = h->mpm;
  oracle_dem *d = h->dem;
  const uint32_t nx = s->cell_count[0] + 1;
  const double hl = grain_hl(h);
  memset(h->dem_mom_grid, 0, 2 * nodes(s) * sizeof(double));
  for (uint32_t b = 0; b < d->nbodies; ++b) {
    if (!h->coupled[b]) continue;
    const double *qb = &h->dem_pos_before[3 * b];
    int64_t lo[2], hi[2];
    oracle_stencil(s, qb, hl, lo, hi);
    const double m = d->mass[3 * b];
    const double pm0 = m * d->v[3 * b], pm1 = m * d->v[3 * b + 1];
    for (int64_t iy = lo[1]; iy <= hi[1]; ++iy)
      for (int64_t ix = lo[0]; ix <= hi[0]; ++ix) {
        if (!node_valid(s, ix, iy)) continue;
        const double w = oracle_weight(s, qb, hl, ix, iy);
        const int64_t n = iy * nx + ix;
        h->dem_mom_grid[2 * n] += w * pm0;
        h->dem_mom_grid[2 * n + 1] += w * pm1;
      }
  }
}

void oracle_hybrid_update_velocities_on_grid(oracle_hybrid *h) {
  oracle_mpm *s = h->mpm;
  const uint64_t G = nodes(s);
  memset(h->v_after, 0, 2 * G * sizeof(double));
  memset(h->at_after, 0, 2 * G * sizeof(double));
  for (uint64_t n = 0; n < G; ++n) {
    const double m_dem = h->dem_mass_grid[n];
    if (!(m_dem > 0.0)) continue;
    const double m_mpm = s->mass[n];
    const double msum = m_mpm + m_dem;
    const double l0 = (m_dem * s->mom_new[2 * n] - m_mpm * h->dem_mom_grid[2 * n]) / msum;
    const double l1 = (m_dem * s->mom_new[2 * n + 1] - m_mpm * h->dem_mom_grid[2 * n + 1]) / msum;
    s->mom_new[2 * n] -= l0;
    s->mom_new[2 * n + 1] -= l1;
    const double da0 = h->dem_mom_grid[2 * n] + l0, da1 = h->dem_mom_grid[2 * n + 1] + l1;
    h->v_after[2 * n] = da0 / m_dem;
    h->v_after[2 * n + 1] = da1 / m_dem;
    h->at_after[2 * n] = l0 / m_dem;
    h->at_after[2 * n + 1] = l1 / m_dem;
  }
}

void oracle_hybrid_transfer_to_grain(oracle_hybrid *h) {
  oracle_mpm *s = h->mpm;
  oracle_dem *d = h->dem;
  const uint32_t nx = s->cell_count[0] + 1;
  const double hl = grain_hl(h);
  for (uint32_t b = 0; b < d->nbodies; ++b) {
    if (!h->coupled[b]) continue;
    const double *qb = &h->dem_pos_before[3 * b];
    int64_t lo[2], hi[2];
    oracle_stencil(s, qb, hl, lo, hi);
    double vp0 = 0.0, vp1 = 0.0, a0 = 0.0, a1 = 0.0;
    for (int64_t iy = lo[1]; iy <= hi[1]; ++iy)
      for (int64_t ix = lo[0]; ix <= hi[0]; ++ix) {
        if (!node_valid(s, ix, iy)) continue;
        const double w = oracle_weight(s, qb, hl, ix, iy);
        const int64_t n = iy * nx + ix;
        vp0 += w * h->v_after[2 * n];
        vp1 += w * h->v_after[2 * n + 1];
        a0 += w * h->at_after[2 * n];
        a1 += w * h->at_after[2 * n + 1];
      }
    const double vf0 = d->v[3 * b] + a0, vf1 = d->v[3 * b + 1] + a1;
    d->v[3 * b] = vp0 * (1.0 - 1.0) + vf0 * 1.0;
    d->v[3 * b + 1] = vp1 * (1.0 - 1.0) + vf1 * 1.0;
  }
}

void oracle_hybrid_update_positions(oracle_hybrid *h, double dt) {
  oracle_dem *d = h->dem;
  for (uint32_t b = 0; b < d->nbodies; ++b) {
    if (!h->coupled[b]) continue;
    d->q[3 * b] = h->dem_pos_before[3 * b] + dt * d->v[3 * b];
    d->q[3 * b + 1] = h->dem_pos_before[3 * b + 1] + dt * d->v[3 * b + 1];
  }
}

/* ResamplingTools.cpp:744-801.  The cache's circle contacts in its order;
 * r is taken at the current (end-of-step) positions, as the reference calls
 * this after the step. */
void oracle_hybrid_homogenize_stress(oracle_hybrid *h) {
  oracle_mpm *s = h->mpm;
  oracle_dem *d = h->dem;
  const uint32_t nx = s->cell_count[0] + 1;
  const double cw = s->cell_width;
  const double cell_area = cw * cw;
  const double hl = cw * 0.25;
  const uint32_t nc = oracle_dem_circle_contacts(d, NULL, NULL, NULL, NULL, 0);
  if (nc == 0) return;
  int32_t *ci = (int32_t *)malloc(2 * (size_t)nc * sizeof(int32_t));
  double *pf = (double *)malloc(4 * (size_t)nc * sizeof(double));
  if (!ci || !pf) {
    free(ci);
    free(pf);
    return;
  }
  oracle_dem_circle_contacts(d, ci, ci + nc, pf, pf + 2 * nc, nc);
  for (uint32_t k = 0; k < nc; ++k) {
    const uint32_t a = (uint32_t)ci[k], b = (uint32_t)ci[nc + k];
    const double *p = &pf[2 * k], *f = &pf[2 * nc + 2 * k];
    const double r[2] = {d->q[3 * a] - d->q[3 * b], d->q[3 * a + 1] - d->q[3 * b + 1]};
    /* -0.5 (f r^T + r f^T), column-major */
    double st[4];
    st[0] = -0.5 * (f[0] * r[0] + r[0] * f[0]);
    st[1] = -0.5 * (f[1] * r[0] + r[1] * f[0]);
    st[2] = -0.5 * (f[0] * r[1] + r[0] * f[1]);
    st[3] = -0.5 * (f[1] * r[1] + r[1] * f[1]);
    int64_t lo[2], hi[2];
    oracle_stencil(s, p, hl, lo, hi);
    for (int64_t iy = lo[1]; iy <= hi[1]; ++iy)
      for (int64_t ix = lo[0]; ix <= hi[0]; ++ix) {
        if (!node_valid(s, ix, iy)) continue;
        const double w = oracle_weight(s, p, hl, ix, iy);
        const int64_t n = iy * nx + ix;
        for (int e = 0; e < 4; ++e) h->homog_stress[4 * n + e] += w * st[e] / cell_area;
      }
  }
  free(ci);
  free(pf);
}

/* ResamplingTools.cpp:818-874: per-step nodal mass and momentum of the
 * non-fixed circles (end-of-step positions and velocities), then the nodal
 * velocity added wherever the mass is positive. */
int oracle_hybrid_homogenize_velocity(oracle_hybrid *h) {
  oracle_mpm *s = h->mpm;
  oracle_dem *d = h->dem;
  const uint32_t nx = s->cell_count[0] + 1;
  const uint64_t G = nodes(s);
  const double hl = s->cell_width * 0.25;
  double *mom = (double *)calloc(2 * G, sizeof(double));
  double *mass = (double *)calloc(G, sizeof(double));
  if (!mom || !mass) {
    free(mom);
    free(mass);
    return -1;
  }
  for (uint32_t b = 0; b < d->nbodies; ++b) {
    if (d->fixed && d->fixed[b]) continue;
    const double m = d->mass[3 * b];
    const double v0 = d->v[3 * b], v1 = d->v[3 * b + 1];
    const double *q = &d->q[3 * b];
    int64_t lo[2], hi[2];
    oracle_stencil(s, q, hl, lo, hi);
    for (int64_t iy = lo[1]; iy <= hi[1]; ++iy)
      for (int64_t ix = lo[0]; ix <= hi[0]; ++ix) {
        if (!node_valid(s, ix, iy)) continue;
        const double w = oracle_weight(s, q, hl, ix, iy);
        const int64_t n = iy * nx + ix;
        mom[2 * n] += w * m * v0;
        mom[2 * n + 1] += w * m * v1;
        mass[n] += w * m;
      }
  }
  for (uint64_t n = 0; n < G; ++n)
    if (mass[n] > 0.0) {
      h->homog_velocity[2 * n] += mom[2 * n] / mass[n];
      h->homog_velocity[2 * n + 1] += mom[2 * n + 1] / mass[n];
    }
  free(mom);
  free(mass);
  return 0;
}

int oracle_hybrid_step(oracle_hybrid *h, double dt, uint32_t nsteps) {
  oracle_mpm *s = h->mpm;
  oracle_dem *d = h->dem;
  for (uint32_t k = 0; k < nsteps; ++k) {
    memcpy(h->dem_pos_before, d->q, 3 * (size_t)d->nbodies * sizeof(double));
    oracle_mpm_flow_zero_phase(s);
    oracle_hybrid_rasterize_grain_mass(h);
    if (oracle_dem_flow(d, dt, 1)) return -1;
    oracle_mpm_flow_first_phase(s, dt);
    oracle_hybrid_rasterize_grain_mom(h);
    oracle_hybrid_update_velocities_on_grid(h);
    const uint64_t G = nodes(s);
    if (h->homog_factor)
      for (uint64_t n = 0; n < G; ++n)
        h->homog_factor[n] = h->dem_mass_grid[n] > 0.0 ? 0.5 : (s->mass[n] > 0.0 ? 1.0 : 0.0);
    oracle_hybrid_transfer_to_grain(h);
    oracle_mpm_flow_second_phase(s, dt);
    oracle_hybrid_update_positions(h, dt);
    if (h->homog_stress) oracle_hybrid_homogenize_stress(h);
    if (h->homog_velocity && oracle_hybrid_homogenize_velocity(h)) return -1;
  }
  return 0;
}

int oracle_hybrid_step_omp(oracle_hybrid *h, double dt, uint32_t nsteps, int nthreads) {
  oracle_mpm *s = h->mpm;
  oracle_dem *d = h->dem;
  const int T = oracle_omp_threads(nthreads);
  const uint64_t G = nodes(s);
  double *scratch = (double *)malloc((size_t)T * 5 * G * sizeof(double));
  if (!scratch) return -1;
  int rc = 0;
  for (uint32_t k = 0; k < nsteps && rc == 0; ++k) {
    memcpy(h->dem_pos_before, d->q, 3 * (size_t)d->nbodies * sizeof(double));
    oracle_mpm_omp_scatter(s, T, scratch);
    oracle_hybrid_rasterize_grain_mass(h);
    if (oracle_dem_flow(d, dt, 1)) {
      rc = -1;
      break;
    }
    oracle_mpm_omp_grid_update(s, dt, T);
    oracle_hybrid_rasterize_grain_mom(h);
    oracle_hybrid_update_velocities_on_grid(h);
    if (h->homog_factor)
      for (uint64_t n = 0; n < G; ++n)
        h->homog_factor[n] = h->dem_mass_grid[n] > 0.0 ? 0.5 : (s->mass[n] > 0.0 ? 1.0 : 0.0);
    oracle_hybrid_transfer_to_grain(h);
    oracle_mpm_omp_point_update(s, dt, T);
    oracle_hybrid_update_positions(h, dt);
  }
  free(scratch);
  return rc;
}
