/*
 * hg.h -- C-ABI of the MI355X (gfx950) explicit-step engine for
 * libHybridGrains scenes: MPM (mpmgrains2d ExplicitIntegrator), circle DEM
 * (rigidbody2d PenaltyImpactFrictionMap) and the hybridgrains2dnew coupling.
 *
 * Conventions
 *  - Every call returns hg_status: HG_OK (0) or a negative HG_ERR_* code;
 *    hg_last_error(ctx) returns the message of the last failure on a context.
 *  - Handles are opaque and own their device memory; state stays resident in
 *    HBM between calls.  Each handle launches on one HIP stream of its own.
 *  - Calls on one handle are not thread-safe; calls on different handles are.
 *  - Host arrays use the reference's Eigen layouts: 2-vectors interleaved
 *    (x0,y0,x1,y1,...), 2x2 matrices column-major (m00,m10,m01,m11), DEM
 *    coordinates (x,y,theta) per body.
 *  - Point and body arrays passed to *_download come back in creation order,
 *    whatever order the engine keeps them in on the device.
 */
#ifndef HG_H
#define HG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t hg_status;
#define HG_OK 0
#define HG_ERR_INVALID (-1)     /* bad argument or descriptor */
#define HG_ERR_HIP (-2)         /* HIP runtime failure */
#define HG_ERR_NOMEM (-3)       /* device allocation failed */
#define HG_ERR_STATE (-4)       /* simulation left its valid domain (point off the grid, NaN) */
#define HG_ERR_UNSUPPORTED (-5) /* configuration outside what the engine implements */

enum { HG_FP64 = 0, HG_FP32 = 1 };
enum { HG_BASIS_UGIMP_LINEAR = 0, HG_BASIS_LINEAR = 1, HG_BASIS_CUBIC = 3 };
enum { HG_PLANE_SLIDING = 0, HG_PLANE_STICKING = 1 };

typedef struct hg_ctx hg_ctx;
typedef struct hg_mpm hg_mpm;
typedef struct hg_dem hg_dem;

/* ---- context -------------------------------------------------------------- */
/* No reference counterpart: the reference is CPU-only.  Binds one GPU. */
hg_status hg_ctx_create(int32_t device, hg_ctx **out);
hg_status hg_ctx_destroy(hg_ctx *ctx);
const char *hg_last_error(const hg_ctx *ctx);
/* Library version string, e.g. "hg 0.1 gfx950". */
const char *hg_version(void);

/* ---- MPM: mpmgrains2d SimulationState + ExplicitIntegrator ----------------- */
/*
 * Replaces the state a caller builds with
 * InitialSimulationState::generateSimulationState
 * (mpmgrains2d/InitialSimulationState.cpp:147-228) or
 * SimulationState::deserialize (mpmgrains2d/SimulationState.cpp:131-143):
 * grid geometry as PhysicsGrid::setDimensions leaves it (min, cell_width,
 * cell_count; PhysicsGrid.cpp:14-55), the material of SimulationState.h:46-57,
 * the static planes of StaticPlane.h:6-22 and the MaterialPoints fields the
 * step reads (MaterialPoints.h:290-326).
 */
typedef struct hg_mpm_desc {
  int32_t precision;         /* HG_FP64: plain FP64 path; HG_FP32: FP32 bulk, FP64 positions and return map */
  int32_t basis;             /* HG_BASIS_* (BasisFunctions.h:13) */
  double grid_min[2];
  double cell_width;
  uint32_t cell_count[2];
  double shear_modulus;      /* mu */
  double bulk_modulus;       /* kappa */
  double dp_alpha;           /* Drucker_Prager_alpha */
  double pic_alpha;          /* PIC/FLIP blend 'alpha' */
  double gravity[2];         /* near_earth_gravity */
  uint32_t nplanes;
  const double *plane_x;       /* 2*nplanes */
  const double *plane_n;       /* 2*nplanes (normalised by the engine, as MPMStaticPlane's ctor does) */
  const double *plane_lower;   /* nplanes; -INFINITY when absent */
  const int32_t *plane_behavior; /* nplanes; HG_PLANE_* */
  uint32_t npoints;
  const double *q;           /* 2N */
  const double *v;           /* 2N */
  const double *m;           /* N */
  const double *volume;      /* N */
  const double *hl;          /* N, uGIMP half width */
  const double *be_bar;      /* 4N */
  const double *J;           /* N */
  const uint8_t *kinematically_scripted; /* N or NULL */
} hg_mpm_desc;

hg_status hg_mpm_create(hg_ctx *ctx, const hg_mpm_desc *desc, hg_mpm **out);
hg_status hg_mpm_destroy(hg_mpm *h);

/* ExplicitIntegrator::flow (mpmgrains2d/ExplicitIntegrator.cpp:63-67), nsteps
 * times.  Asynchronous on the handle's stream unless an error check is due:
 * the engine checks its device error flag (point off the grid, non-finite
 * state: the isConsistent criteria of SimulationState.cpp:52-111) on every
 * hg_mpm_sync / download and returns HG_ERR_STATE. */
hg_status hg_mpm_step(hg_mpm *h, double dt, uint32_t nsteps);

/* ExplicitIntegrator::flow_zero_phase / flow_first_phase / flow_second_phase
 * (ExplicitIntegrator.cpp:5-61) for callers that interleave other work, as
 * the hybrid integrator does.  phase 0 rasterises mass and momentum; phase 1
 * adds the internal and gravity forces, updates the node momentum and applies
 * the grid plane conditions; phase 2 forms node velocity and acceleration and
 * runs the particle update.  The hybrid step (hg_hybrid_step) edits the node
 * momenta between phases 1 and 2 on the device. */
hg_status hg_mpm_phase(hg_mpm *h, int32_t phase, double dt);

/* Waits for the handle's stream and reports HG_ERR_STATE if any step so far
 * left the valid domain. */
hg_status hg_mpm_sync(hg_mpm *h);

/* Copies the points back to the caller, in creation order (any pointer may be
 * NULL).  Mirrors the MaterialPoints fields (MaterialPoints.h:290-326). */
hg_status hg_mpm_download_points(hg_mpm *h, double *q, double *v, double *be_bar,
                                 double *J);

/* Copies the grid fields of the last step, in the reference node order
 * (flat index y*(cell_count[0]+1)+x, PhysicsGrid.cpp:76-80).  Fields not kept
 * by the last step's mode are returned as zeros; hg_mpm_set_keep_grid(h, 1)
 * keeps all of them (mass, momentum, momentum_new, force, velocity,
 * acceleration), as PhysicsGrid holds after flow().  Without it an FP32
 * handle runs the fused step, which ends each step with the next step's grid
 * update: its velocity and acceleration are then those of the grid of the
 * current points (the step that would come next), and the other fields are
 * zeros. */
hg_status hg_mpm_set_keep_grid(hg_mpm *h, int32_t keep_all);
hg_status hg_mpm_download_grid(hg_mpm *h, double *mass, double *mom,
                               double *mom_new, double *force, double *vel,
                               double *acc);

/* Number of points / grid nodes (reference indexing) of scene 0. */
hg_status hg_mpm_counts(const hg_mpm *h, uint32_t *npoints, uint32_t *nnodes);

/* ---- batches of independent scenes (no reference counterpart: the reference
 * steps one SimulationState per process) ------------------------------------
 * One handle may hold several scenes of uneven size; every step kernel covers
 * all of them in one launch through per-scene offset tables.  All scenes of a
 * batch share precision and basis.  Each scene evolves bit-identically to the
 * same scene stepped alone.  hg_mpm_create is hg_mpm_create_batch with one
 * scene; the un-suffixed download calls address scene 0. */
hg_status hg_mpm_create_batch(hg_ctx *ctx, const hg_mpm_desc *descs, uint32_t nscenes,
                              hg_mpm **out);
hg_status hg_mpm_download_points_scene(hg_mpm *h, uint32_t scene, double *q, double *v,
                                       double *be_bar, double *J);
hg_status hg_mpm_download_grid_scene(hg_mpm *h, uint32_t scene, double *mass, double *mom,
                                     double *mom_new, double *force, double *vel,
                                     double *acc);
/* Per-scene device error bits after the steps queued so far (waits for the
 * stream): 1 = a point's stencil left the grid, 2 = non-finite point state;
 * 0x100 (scene 0 only) = the cell sort's per-block totals disagreed with its
 * cell counts (never expected; the step's results are void). */
hg_status hg_mpm_scene_error(hg_mpm *h, uint32_t scene, int32_t *flags);

/* P2G tile: aw x ah grid nodes per workgroup, one thread per node.
 * Supported: 16x16 (the default), 16x8, 32x8, 64x4.  Results do not depend
 * on it (bitwise, unless one cell holds more points than a staging chunk). */
hg_status hg_mpm_set_tile(hg_mpm *h, int32_t aw, int32_t ah);
/* Autotuner (no reference counterpart): times the P2G pass for each tile
 * shape on the handle's own scenes (reps launches each; the point state is
 * not changed), keeps the fastest and returns it; us (4 entries, or NULL)
 * receives the microseconds per launch of 16x16, 16x8, 32x8 and 64x4.  FP32
 * handles time whole fused steps instead (the point state is saved first and
 * restored after), each shape with one and with two threads per cell, and us
 * gets the faster of the two. */
hg_status hg_mpm_autotune(hg_mpm *h, double dt, int32_t reps, int32_t *aw, int32_t *ah,
                          double *us);
/* The tuning in use (no reference counterpart): the tile shape and, for FP32
 * handles, the points each tile takes per pass in threads per cell (1 or 2;
 * the autotuner times both and keeps the faster; 1 otherwise). */
hg_status hg_mpm_tuning(const hg_mpm *h, int32_t *aw, int32_t *ah, int32_t *points_per_pass);
/* Sets the FP32 step's points per tile pass (1 or 2 threads per cell; no
 * reference counterpart; ignored in FP64).  Results do not depend on it. */
hg_status hg_mpm_set_points_per_pass(hg_mpm *h, int32_t points_per_pass);

/* ---- timing (no reference counterpart) --------------------------------------
 * With profiling on, every kernel launch is bracketed by HIP events on the
 * handle's stream; hg_mpm_kernel_times returns, per kernel name, the summed
 * device milliseconds and the launch count since profiling was switched on. */
hg_status hg_mpm_set_profiling(hg_mpm *h, int32_t on);
hg_status hg_mpm_kernel_times(hg_mpm *h, int32_t max_kernels, char *names /* max_kernels*32 */,
                              double *ms, int64_t *launches, int32_t *nkernels);
/* Occupies the handle's stream for about us microseconds (a one-wavefront
 * kernel that waits on the device clock; at most one second).  Steps
 * submitted right after it queue up behind it, so profiled kernels then run
 * back to back and their event times hold no host submission gaps. */
hg_status hg_mpm_hold(hg_mpm *h, double us);
/* Streams nbytes from one device buffer to another iters times and returns
 * the achieved read+write bandwidth; the reference point for HBM counters and
 * the measured copy roofline of the box. */
hg_status hg_bench_copy(hg_ctx *ctx, uint64_t nbytes, int32_t iters, double *gbps);
/* Captures one step into a hipGraph and replays it nsteps times. */
hg_status hg_mpm_step_graph(hg_mpm *h, double dt, uint32_t nsteps);

/* ---- device memory (no reference counterpart) --------------------------------
 * hg_*_estimate_bytes: an upper bound of the device memory a handle created
 * from these descriptors allocates (admission by footprint); hg_*_device_bytes:
 * what a handle holds. */
hg_status hg_mpm_estimate_bytes(const hg_mpm_desc *descs, uint32_t nscenes, uint64_t *bytes);
hg_status hg_mpm_device_bytes(hg_mpm *h, uint64_t *bytes);

/* ---- DEM: rigidbody2d circles with PenaltyImpactFrictionMap ----------------- */
/*
 * Replaces the circle state RigidBody2DSceneParser builds
 * (rigidbody2dutils/RigidBody2DSceneParser.cpp:715-902, :976-1125): bodies
 * (RigidBody2DState.h:13-190 -- q, v, the diagonal of M, fixed flags), circle
 * radii (CircleGeometry.h), near-earth gravity, static planes
 * (RigidBody2DStaticPlane.h:12-45) and the penalty map's kn, kt, gamman,
 * gammat, mu (PenaltyImpactFrictionMap.h:362-367).  The contact history the
 * map keeps between steps (its CollisionCache) lives on the device.
 */
typedef struct hg_dem_desc {
  int32_t precision;       /* HG_FP64, or HG_FP32 (FP64 positions, radii and pair tests) */
  uint32_t nbodies;
  const double *q;         /* 3B: x, y, theta */
  const double *v;         /* 3B */
  const double *mass;      /* 3B: diagonal of M (m, m, I) */
  const double *radius;    /* B */
  const uint8_t *fixed;    /* B or NULL: kinematically scripted bodies */
  double gravity[2];
  uint32_t nplanes;        /* <= 8 */
  const double *plane_x;   /* 2P */
  const double *plane_n;   /* 2P, unit */
  double kn, kt, gamman, gammat, mu;
  double cell_width;       /* broad-phase cell; at least 2 * max radius (0: that) */
  double domain_min[2];    /* extent the broad-phase bucket table spans; bodies */
  double domain_max[2];    /* outside it stay correct (the table wraps); max <= min:
                              the bodies' box with as much room again to spread */
} hg_dem_desc;

hg_status hg_dem_create(hg_ctx *ctx, const hg_dem_desc *desc, hg_dem **out);
/* Several independent scenes stepped by the same launches (no reference
 * counterpart); each evolves bit-identically to the scene alone. */
hg_status hg_dem_create_batch(hg_ctx *ctx, const hg_dem_desc *descs, uint32_t nscenes,
                              hg_dem **out);
hg_status hg_dem_destroy(hg_dem *h);
/* RigidBody2DSim::flow with the penalty map (RigidBody2DSim.cpp:1013-1039 ->
 * PenaltyImpactFrictionMap::flow, PenaltyImpactFrictionMap.cpp:1083-1115),
 * nsteps times; with weights set, flowWithWeight / flowWithWeights
 * (RigidBody2DSim.cpp:1041-1069, PenaltyImpactFrictionMap.cpp:1117-1149). */
hg_status hg_dem_step(hg_dem *h, double dt, uint32_t nsteps);
hg_status hg_dem_step_graph(hg_dem *h, double dt, uint32_t nsteps);
/* Mass weights of one scene's bodies (RigidBody2DState::massWeights), used as
 * contact weights: plane w_i, pair (w_i + w_j)/2, kinematic w_i
 * (PenaltyImpactFrictionMap.cpp:296-599).  NULL returns to unweighted flow. */
hg_status hg_dem_set_weights(hg_dem *h, uint32_t scene, const double *weights);
/* Waits; HG_ERR_STATE on non-finite bodies, HG_ERR_UNSUPPORTED when a body had
 * more pair contacts than the handle's per-body capacity (the step's results
 * are then void).  The capacity is 8, 16 or 32, sized at creation from the
 * radii: the number of partners that fit around the largest body while no two
 * of them overlap by more than half the smallest radius (16 for equal radii).
 * The reference keeps unbounded per-body lists
 * (rigidbody2d/PenaltyImpactFrictionMap.h:272). */
hg_status hg_dem_sync(hg_dem *h);
hg_status hg_dem_download_bodies(hg_dem *h, uint32_t scene, double *q, double *v);
/* The contact cache after the last step in the reference's iteration order
 * (planes, circle pairs, kinematic pairs; each by owner body, then partner):
 * type 0 plane (j = plane index), 1 circle pair (i < j), 2 body i against
 * fixed body j; delta_s is the tangential spring.  *count is the total; at
 * most max entries are written. */
hg_status hg_dem_download_contacts(hg_dem *h, uint32_t scene, uint32_t max, int32_t *type,
                                   int32_t *i, int32_t *j, double *delta_s, uint32_t *count);
/* The last step's force on the owner and contact point of each contact, in
 * hg_dem_download_contacts' order (planes give zeros; points are kept only
 * while a hybrid homogenises, else zeros): what the reference's collision
 * cache holds for ResamplingTools::homogenizeStress (ResamplingTools.cpp:
 * 744-801). */
hg_status hg_dem_download_contact_forces(hg_dem *h, uint32_t scene, uint32_t max, double *force,
                                         double *point, uint32_t *count);
/* Replaces the contact history with count contacts in hg_dem_download_contacts'
 * form (pairs with i < j), as PenaltyImpactFrictionMap's caches hold them after
 * RigidBody2DState::removeBodies relabels them (PenaltyImpactFrictionMap.cpp:
 * 1199-1274). */
hg_status hg_dem_upload_contacts(hg_dem *h, uint32_t scene, uint32_t count, const int32_t *type,
                                 const int32_t *i, const int32_t *j, const double *delta_s);
/* Keep the total per-body force (3B) of each step for download. */
hg_status hg_dem_keep_force(hg_dem *h, int32_t keep);
hg_status hg_dem_download_force(hg_dem *h, uint32_t scene, double *F);
hg_status hg_dem_estimate_bytes(const hg_dem_desc *descs, uint32_t nscenes, uint64_t *bytes);
hg_status hg_dem_device_bytes(hg_dem *h, uint64_t *bytes);
hg_status hg_dem_set_profiling(hg_dem *h, int32_t on);
hg_status hg_dem_hold(hg_dem *h, double us);  /* as hg_mpm_hold */
hg_status hg_dem_kernel_times(hg_dem *h, int32_t max_kernels, char *names, double *ms,
                              int64_t *launches, int32_t *nkernels);

/* ---- hybrid: hybridgrains2dnew node-node coupling ---------------------------- */
typedef struct hg_hybrid hg_hybrid;
/*
 * Replaces HybridGrains2DSim(HybridIntegratorState, RigidBody2DState,
 * SimulationState) (hybridgrains2dnew/HybridGrains2DSim.h:19-21): one
 * continuum scene and one grain scene sharing a stream.  dem_weights (B or
 * NULL) are the grains' mass weights (RigidBody2DState::massWeights), which
 * the reference sets from the hybrid zones.  Precisions must match.
 */
hg_status hg_hybrid_create(hg_ctx *ctx, const hg_mpm_desc *mpm, const hg_dem_desc *dem,
                           const double *dem_weights, hg_hybrid **out);
hg_status hg_hybrid_destroy(hg_hybrid *h);
/* HybridGrains2DSim::stepSystem (HybridGrains2DSim.cpp:229-252) ->
 * stepSystem_constraintSetUp_prediction_correction_advection_node_node
 * (:116-227), nsteps times: continuum grid pass, grain mass rasterisation and
 * coupling test, weighted DEM step, grain momentum rasterisation, node
 * velocity exchange, FLIP transfer to the grains, continuum G2P and the
 * coupled grains' position update (Coupling.cpp:22-591).  The periodic zone
 * resampling is not part of the step. */
hg_status hg_hybrid_step(hg_hybrid *h, double dt, uint32_t nsteps);
/* The same steps replayed from a hipGraph of two steps (no reference
 * counterpart). */
hg_status hg_hybrid_step_graph(hg_hybrid *h, double dt, uint32_t nsteps);
/* Waits; reports the MPM and DEM errors of hg_mpm_sync / hg_dem_sync. */
hg_status hg_hybrid_sync(hg_hybrid *h);
/* The continuum and grain handles (owned by the hybrid handle) for the
 * download, kernel-time and profiling calls; stepping them on their own
 * while they are part of a hybrid handle is not supported. */
hg_status hg_hybrid_mpm(hg_hybrid *h, hg_mpm **out);
hg_status hg_hybrid_dem(hg_hybrid *h, hg_dem **out);
/* Per grain, 1 if the last step coupled it to the continuum
 * (HybridCoupling2D's dem_particle_coupled). */
hg_status hg_hybrid_download_coupling(hg_hybrid *h, uint8_t *coupled);
/* Per-step homogenisation of the grains onto the continuum grid, which
 * HybridGrains2DSim::updateHybridStates_poorman runs after every step when a
 * scene sets homogenize_stress / homogenize_velocity with their
 * grid_smoothing flags (HybridGrains2DSim.cpp:545-556): each step adds
 * ResamplingTools::homogenizeStress (ResamplingTools.cpp:744-801, the circle
 * contacts' stress spread from their contact points) and homogenizeVelocity
 * (:818-874, the grains' nodal velocity) to node accumulators.  Off by
 * default; switching it on allocates the accumulators (zeroed). */
hg_status hg_hybrid_set_homogenize(hg_hybrid *h, int32_t stress, int32_t velocity);
/* The accumulators, nodes in the reference's flat order (PhysicsGrid
 * flatNodeIndex): stress 4 per node (column-major 2x2), velocity 2 per node;
 * NULL skips one.  Zeros while homogenisation was never switched on. */
hg_status hg_hybrid_download_homogenized(hg_hybrid *h, double *stress, double *velocity);
/* Sets the homogenised accumulators from hg_hybrid_download_homogenized's
 * layout (the zone resampling's clear and accumulate steps run on the host). */
hg_status hg_hybrid_upload_homogenized(hg_hybrid *h, const double *stress, const double *velocity);
/* ResamplingTools::clearHomogenizedStress / clearHomogenizedVelocity
 * (ResamplingTools.cpp:736-742, :810-816). */
hg_status hg_hybrid_clear_homogenized(hg_hybrid *h);
/* The continuum grid's coupling weights (PhysicsGrid::homog_factor), which
 * HybridGrains2DSim's node-node step saves every step (HybridGrains2DSim.cpp:
 * 169-197): 1 at nodes with continuum mass after the step's P2G, 0 elsewhere,
 * 0.5 at nodes that hold grain mass.  Output only; no step reads them.  While
 * on, every step records them (one pass over the nodes on the grain stream;
 * the 1M-point hybrid step takes about 5-6 us longer); they start at 1
 * everywhere, as the reference's grid does (PhysicsGrid.cpp:52).  The
 * download returns the last recorded values, nny x nnx row-major;
 * HG_ERR_INVALID if never switched on. */
hg_status hg_hybrid_set_homog_factor(hg_hybrid *h, int32_t on);
hg_status hg_hybrid_download_homog_factor(hg_hybrid *h, double *out);
hg_status hg_hybrid_estimate_bytes(const hg_mpm_desc *mpm, const hg_dem_desc *dem,
                                   uint64_t *bytes);
hg_status hg_hybrid_device_bytes(hg_hybrid *h, uint64_t *bytes);

/* ---- observables (no reference counterpart: the reference writes its state
 * to HDF5 and leaves measurement to scripts) ------------------------------------
 * What a long granular collapse is judged by, from the resident state of one
 * scene: out[0] the kinetic energy sum 1/2 m |v|^2 (translational, per unit
 * depth), out[1] the largest x of the points (bodies), out[2] the largest y,
 * and out[3 + k] for k < nbins the largest y of the points whose x lies in
 * [x0 + k bin_width, x0 + (k + 1) bin_width), NaN for an empty bin (the pile
 * profile).  out holds 3 + nbins doubles.  Waits for the handle's stream;
 * results are bitwise reproducible.  The handle keeps a device scratch buffer
 * for the per-workgroup partials and the bins, allocated on the first call
 * and grown when a call asks for more bins. */
hg_status hg_mpm_observables(hg_mpm *h, uint32_t scene, double x0, double bin_width,
                             uint32_t nbins, double *out);
hg_status hg_dem_observables(hg_dem *h, uint32_t scene, double x0, double bin_width,
                             uint32_t nbins, double *out);
/* hg_mpm_observables without allocating: no reference counterpart; for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same results,
 * bitwise.  Its scratch is a static device array of 90 KiB on each GPU, held
 * until the process ends; nbins above 8192 is refused with HG_ERR_INVALID.
 * Every call on one GPU shares that array: calls that overlap on one GPU
 * (from several threads or streams) race. */
hg_status hg_mpm_observables_noalloc(hg_mpm *h, uint32_t scene, double x0, double bin_width,
                                     uint32_t nbins, double *out);

/* ---- snapshot and resume -------------------------------------------------------
 * The reference's CLIs serialise their host state for a bit-identical resume
 * (mpmgrains2d_cli.cpp:154-247, the -s/-r options); the DEM contact history
 * and the hybrid's homogenisation accumulators are beyond what they save.
 * hg_*_snapshot writes a handle's evolving state into buf: *bytes holds the
 * buffer size on entry and the snapshot's size on return (buf NULL: size
 * only).  hg_*_restore loads it into a handle created from the same
 * descriptors (same scenes, sizes and precision; DEM mass weights are set as
 * for the original handle); stepping then continues as the snapshotted
 * handle would -- bitwise for MPM and DEM, to the hybrid's own run-to-run
 * spread (its grain-side node sums are FP64 atomics).  A blob of another
 * kind, precision or size is refused with HG_ERR_INVALID. */
hg_status hg_mpm_snapshot(hg_mpm *h, void *buf, uint64_t *bytes);
hg_status hg_mpm_restore(hg_mpm *h, const void *buf, uint64_t bytes);
hg_status hg_dem_snapshot(hg_dem *h, void *buf, uint64_t *bytes);
hg_status hg_dem_restore(hg_dem *h, const void *buf, uint64_t bytes);
hg_status hg_hybrid_snapshot(hg_hybrid *h, void *buf, uint64_t *bytes);
hg_status hg_hybrid_restore(hg_hybrid *h, const void *buf, uint64_t bytes);

#ifdef __cplusplus
}
#endif
#endif
