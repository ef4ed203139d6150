/*
 * hg_oracle.h -- CPU FP64 restatement of libHybridGrains' explicit-step hot
 * path.  TEST INFRASTRUCTURE ONLY: this library is the checker that the HIP
 * path is compared against (and the timed CPU baseline).  Nothing in the
 * product links it.
 *
 * Every function follows the reference's operation order; the file:line of the
 * reference function it restates is given next to each declaration.  Layouts
 * mirror the reference's Eigen storage:
 *   - 2-vectors per point/node are interleaved (x0,y0,x1,y1,...), as
 *     Matrix2Xsc columns (mpmgrains2d/MaterialPoints.h:119-123);
 *   - 2x2 matrices are column-major (m00,m10,m01,m11), as Matrix22sc;
 *   - DEM generalized coordinates are (x,y,theta) per body, as VectorXs q/v
 *     (rigidbody2d/RigidBody2DState.h:166-170).
 */
#ifndef HG_ORACLE_H
#define HG_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORACLE_BASIS_UGIMP_LINEAR = 0, ORACLE_BASIS_LINEAR = 1, ORACLE_BASIS_CUBIC = 3 };
enum { ORACLE_PLANE_SLIDING = 0, ORACLE_PLANE_STICKING = 1 };

/* SimulationState (mpmgrains2d/SimulationState.h:16-68) with the PhysicsGrid
 * (mpmgrains2d/PhysicsGrid.h:13-95) and MaterialPoints
 * (mpmgrains2d/MaterialPoints.h:211-156) fields the explicit step touches.
 * All arrays are owned by the caller. */
typedef struct oracle_mpm {
  /* grid */
  double grid_min[2];
  double cell_width;
  uint32_t cell_count[2];
  double *mass;           /* G */
  double *mom;            /* 2G */
  double *mom_new;        /* 2G */
  double *force;          /* 2G */
  double *vel;            /* 2G */
  double *acc;            /* 2G */
  double *coupling_force; /* 2G, held at zero as in the reference */
  int32_t *hybridized;    /* G */
  /* points */
  uint32_t npoints;
  double *q;        /* 2N */
  double *v;        /* 2N */
  double *m;        /* N */
  double *volume;   /* N */
  double *hl;       /* N */
  double *be_bar;   /* 4N */
  double *J;        /* N */
  double *sigma;    /* 4N */
  double *vel_grad; /* 4N */
  int32_t *just_homogenized;        /* N */
  uint8_t *kinematically_scripted;  /* N */
  /* material, loading, boundaries */
  double shear_modulus, bulk_modulus, dp_alpha, pic_alpha;
  double gravity[2];
  uint32_t nplanes;
  double *plane_x;        /* 2P */
  double *plane_n;        /* 2P, unit */
  double *plane_lower;    /* P */
  int32_t *plane_behavior; /* P */
  int32_t basis;          /* ORACLE_BASIS_* */
} oracle_mpm;

/* ---- basis functions ---------------------------------------------------- */
/* uGIMPLinearBasisFunctions.cpp:6-27 / :29-42 */
double oracle_ugimp_linear_integral(double xp, double hl, double xi, double w);
double oracle_ugimp_linear_integral_grad(double xp, double hl, double xi, double w);
/* computeStencil for every basis (uGIMPLinearBasisFunctions.cpp:68-90,
 * LinearBasisFunctions.cpp:48-60, CubicBasisFunctions.cpp:61-75); lo/hi are
 * inclusive node indices, stored as signed so an underflow is visible. */
void oracle_stencil(const oracle_mpm *s, const double xp[2], double hl,
                    int64_t lo[2], int64_t hi[2]);
/* weight / weightGrad for every basis (uGIMP :108-123/:141-156, Linear
 * :79-94/:113-135, Cubic :96-110/:129-146) */
double oracle_weight(const oracle_mpm *s, const double xp[2], double hl,
                     int64_t ix, int64_t iy);
void oracle_weight_grad(const oracle_mpm *s, const double xp[2], double hl,
                        int64_t ix, int64_t iy, double g[2]);

/* ---- ExplicitIntegrator (mpmgrains2d/ExplicitIntegrator.cpp:5-67) -------- */
void oracle_mpm_clear_rasterized_data(oracle_mpm *s);           /* PhysicsGrid.cpp:107-114 */
void oracle_mpm_rasterize_point_masses(oracle_mpm *s);          /* PhysicsGrid.cpp:168-207 */
void oracle_mpm_rasterize_point_momentum(oracle_mpm *s);        /* PhysicsGrid.cpp:274-311 */
void oracle_mpm_compute_cauchy_stress(oracle_mpm *s);           /* MaterialPoints.cpp:183-202 */
void oracle_mpm_compute_forces(oracle_mpm *s);                  /* PhysicsGrid.cpp:413-473 */
void oracle_mpm_update_momentum(oracle_mpm *s, double dt);      /* PhysicsGrid.cpp:565-586 */
void oracle_mpm_grid_plane_collisions(oracle_mpm *s);           /* PhysicsGrid.cpp:611-673 */
void oracle_mpm_velocity_acceleration_update(oracle_mpm *s, double dt); /* PhysicsGrid.cpp:588-609 */
void oracle_mpm_compute_vel_grad(oracle_mpm *s);                /* MaterialPoints.cpp:296-322 */
void oracle_mpm_elastic_prediction(oracle_mpm *s, double dt);   /* MaterialPoints.cpp:324-348 */
void oracle_mpm_plastic_correction(oracle_mpm *s);              /* MaterialPoints.cpp:350-369 */
void oracle_mpm_update_velocities(oracle_mpm *s, double dt);    /* MaterialPoints.cpp:204-245 */
void oracle_mpm_update_positions(oracle_mpm *s, double dt);     /* MaterialPoints.cpp:247-288 */
void oracle_mpm_point_plane_collisions(oracle_mpm *s);          /* MaterialPoints.cpp:371-423 */
void oracle_mpm_flow_zero_phase(oracle_mpm *s);                 /* ExplicitIntegrator.cpp:5-16 */
void oracle_mpm_flow_first_phase(oracle_mpm *s, double dt);     /* ExplicitIntegrator.cpp:18-34 */
void oracle_mpm_flow_second_phase(oracle_mpm *s, double dt);    /* ExplicitIntegrator.cpp:36-61 */
void oracle_mpm_flow(oracle_mpm *s, double dt);                 /* ExplicitIntegrator.cpp:63-67 */
void oracle_mpm_flow_n(oracle_mpm *s, double dt, uint32_t nsteps);
/* ConstitutiveModel.cpp:19-22, :26-81, :83-97 on a single column-major 2x2 */
void oracle_elastic_prediction(const double be_n[4], const double L[4], double dt,
                               double be_star[4]);
void oracle_plastic_correction(const double be_star[4], double kappa, double mu,
                               double alpha, int32_t *just_homogenized,
                               double be_next[4]);
void oracle_compute_tau(const double be[4], double kappa, double mu,
                        int32_t just_homogenized, double tau[4]);
/* SimulationState::isConsistent's particle checks (SimulationState.cpp:52-111):
 * returns 0 when every point is finite and strictly inside the grid. */
int oracle_mpm_is_consistent(const oracle_mpm *s);

/* OpenMP variant of the full step with the reference's per-thread full-grid
 * scatter copies (PhysicsGrid.cpp:208-272, :329-411, :474-562).  Used only as
 * the multi-core CPU baseline; results differ from the serial oracle by
 * summation order only. */
void oracle_mpm_flow_n_omp(oracle_mpm *s, double dt, uint32_t nsteps, int nthreads);
/* The OpenMP step's phases (CPU baseline only): scratch holds T * 5 * G
 * doubles; oracle_omp_threads(0) is OpenMP's default thread count. */
int oracle_omp_threads(int nthreads);
void oracle_mpm_omp_scatter(oracle_mpm *s, int T, double *scratch);
void oracle_mpm_omp_grid_update(oracle_mpm *s, double dt, int T);
void oracle_mpm_omp_point_update(oracle_mpm *s, double dt, int T);

/* ---- DEM: rigidbody2d circles with PenaltyImpactFrictionMap -------------- */

enum { ORACLE_CONTACT_PLANE = 0, ORACLE_CONTACT_CIRCLE = 1, ORACLE_CONTACT_KINEMATIC = 2 };

/* RigidBody2DState (rigidbody2d/RigidBody2DState.h:13-190) restricted to
 * circles, static planes (RigidBody2DStaticPlane.h:12-45), gravity
 * (NearEarthGravityForce.cpp:34-51) and the penalty map's parameters
 * (PenaltyImpactFrictionMap.h:362-436).  The contact history (the map's
 * CollisionCache) is owned by the oracle: oracle_dem_init / oracle_dem_free. */
typedef struct oracle_dem {
  uint32_t nbodies;
  double *q;        /* 3B: x, y, theta */
  double *v;        /* 3B */
  double *mass;     /* 3B: diagonal of M (m, m, I) */
  double *minv;     /* 3B: diagonal of M^-1 */
  double *radius;   /* B */
  uint8_t *fixed;   /* B: kinematically scripted */
  double *weights;  /* B mass weights for flowWithWeights, NULL for flow */
  double gravity[2];
  uint32_t nplanes;
  double *plane_x;  /* 2P */
  double *plane_n;  /* 2P */
  double kn, kt, gamman, gammat, mu;
  void *hist;       /* contact history (opaque) */
} oracle_dem;

int oracle_dem_init(oracle_dem *s);
void oracle_dem_free(oracle_dem *s);
/* PenaltyImpactFrictionMap::flow (PenaltyImpactFrictionMap.cpp:1083-1115) or
 * flowWithWeights (:1117-1149) when s->weights is set, nsteps times. */
int oracle_dem_flow(oracle_dem *s, double dt, uint32_t nsteps);
/* The current contact cache, in the reference's iteration order (type, then
 * owner body, then insertion order).  Returns the number of contacts (may
 * exceed max; only max are written). */
uint32_t oracle_dem_contacts(const oracle_dem *s, int32_t *type, int32_t *i, int32_t *j,
                             double *delta_s, uint32_t max);
/* Circle pairs (i<j) the narrow phase accepts, by brute force over all pairs
 * (RigidBody2DSim::dispatchNarrowPhaseCollision with q0 == q1,
 * CollisionDetectionUtilities.cpp:8-13, :81-90).  Returns the count. */
uint32_t oracle_dem_pairs_bruteforce(const oracle_dem *s, int32_t *i, int32_t *j, uint32_t max);
/* Total force per body of the last step (gravity + contacts, fixed zeroed). */
void oracle_dem_last_force(const oracle_dem *s, double *F);
/* The circle-circle entries of the contact cache in its order, with each
 * contact's point and its normal + friction force on body i, as the cache
 * holds them after a step (PenaltyImpactFrictionMap.h:59-87, :843-844). */
uint32_t oracle_dem_circle_contacts(const oracle_dem *s, int32_t *i, int32_t *j, double *p,
                                    double *f, uint32_t max);

/* ---- hybrid: hybridgrains2dnew coupling ------------------------------------ */

/* HybridGrains2DSim's node-node coupled step without zone resampling
 * (hybridgrains2dnew/HybridGrains2DSim.cpp:116-227): the MPM and DEM states
 * plus HybridCoupling2D's per-step buffers (Coupling.h:16-70), owned by the
 * caller. */
typedef struct oracle_hybrid {
  oracle_mpm *mpm;
  oracle_dem *dem;          /* stepped with its weights (flowWithWeights) */
  double *dem_pos_before;   /* 3B */
  double *dem_mass_grid;    /* G */
  double *dem_mom_grid;     /* 2G */
  double *v_after;          /* 2G */
  double *at_after;         /* 2G */
  double *homog_factor;     /* G */
  uint8_t *coupled;         /* B */
  double *homog_stress;     /* 4G column-major 2x2 accumulators, NULL: not homogenised */
  double *homog_velocity;   /* 2G accumulators, NULL: not homogenised */
} oracle_hybrid;

/* Coupling.cpp:22-285 (serial branch, use_pre_position) */
void oracle_hybrid_rasterize_grain_mass(oracle_hybrid *h);
/* Coupling.cpp:287-466 */
void oracle_hybrid_rasterize_grain_mom(oracle_hybrid *h);
/* Coupling.cpp:533-579 */
void oracle_hybrid_update_velocities_on_grid(oracle_hybrid *h);
/* Coupling.cpp:468-527 (FLIP, alpha = 1) */
void oracle_hybrid_transfer_to_grain(oracle_hybrid *h);
/* Coupling.cpp:581-591 */
void oracle_hybrid_update_positions(oracle_hybrid *h, double dt);
/* ResamplingTools::homogenizeStress (ResamplingTools.cpp:744-801): adds every
 * circle contact's stress -(f r^T + r f^T)/2, spread with the MPM weights
 * (hl = h/4) from its contact point, divided by the cell area. */
void oracle_hybrid_homogenize_stress(oracle_hybrid *h);
/* ResamplingTools::homogenizeVelocity (ResamplingTools.cpp:818-874): adds the
 * grains' mass-weighted nodal velocity at every node they reach. */
int oracle_hybrid_homogenize_velocity(oracle_hybrid *h);
/* HybridGrains2DSim.cpp:116-227 without the zone resampling, nsteps times;
 * with homog_stress / homog_velocity set, each step then homogenises as
 * updateHybridStates_poorman does every step (HybridGrains2DSim.cpp:545-556). */
int oracle_hybrid_step(oracle_hybrid *h, double dt, uint32_t nsteps);
/* The same step with the MPM part in the reference's OpenMP structure on
 * nthreads threads (CPU baseline only; the grain work stays serial). */
int oracle_hybrid_step_omp(oracle_hybrid *h, double dt, uint32_t nsteps, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
