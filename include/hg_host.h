/*
 * hg_host.h -- C-ABI of the host library (libhg_host.so): the scene-file
 * readers, for callers that are not C++ (ctypes, cgo, JNI).  Each reader
 * returns an opaque scene whose hg_*_desc points into the scene's own arrays,
 * ready for hg_mpm_create / hg_dem_create / hg_hybrid_create.
 */
#ifndef HG_HOST_H
#define HG_HOST_H

#include "hg.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hgx_mpm_scene hgx_mpm_scene;
typedef struct hgx_dem_scene hgx_dem_scene;
typedef struct hgx_hybrid_scene hgx_hybrid_scene;

/* MPMGrains2DParser::readXMLFile + InitialSimulationState::
 * generateSimulationState (mpmgrains2dutils/MPMGrains2DParser.cpp:658-683,
 * mpmgrains2d/InitialSimulationState.cpp:147-228).  On failure returns
 * HG_ERR_INVALID and writes the reader's message into err (errlen bytes). */
hg_status hgx_mpm_scene_read(const char *path, hgx_mpm_scene **out, char *err, size_t errlen);
hg_status hgx_mpm_scene_read_text(const char *xml, hgx_mpm_scene **out, char *err, size_t errlen);
/* Descriptor over the scene's arrays; valid until hgx_mpm_scene_free. */
hg_status hgx_mpm_scene_desc(hgx_mpm_scene *s, int32_t precision, hg_mpm_desc *out);
/* dt (as the scene's rational evaluates) and end_time (INFINITY if absent). */
hg_status hgx_mpm_scene_times(const hgx_mpm_scene *s, double *dt, double *end_time);
void hgx_mpm_scene_free(hgx_mpm_scene *s);

/* RigidBody2DSceneParser::parseXMLSceneFile for circle scenes with a
 * penalty_impact_map (rigidbody2dutils/RigidBody2DSceneParser.cpp:1127-1297). */
hg_status hgx_dem_scene_read(const char *path, hgx_dem_scene **out, char *err, size_t errlen);
hg_status hgx_dem_scene_read_text(const char *xml, hgx_dem_scene **out, char *err, size_t errlen);
/* cell_scale <= 0 takes the scene's collision_detection cell_scale. */
hg_status hgx_dem_scene_desc(hgx_dem_scene *s, int32_t precision, double cell_scale,
                             hg_dem_desc *out);
hg_status hgx_dem_scene_times(const hgx_dem_scene *s, double *dt, double *end_time);
void hgx_dem_scene_free(hgx_dem_scene *s);

/* HybridGrains2DSceneParser::parseXMLSceneFile
 * (hybridgrains2dnewutils/HybridGrains2DSceneParser.cpp:437-522) with the
 * discrete and continuum scenes it names, read from the hybrid scene's folder. */
hg_status hgx_hybrid_scene_read(const char *path, hgx_hybrid_scene **out, char *err,
                                size_t errlen);
hg_status hgx_hybrid_scene_parts(hgx_hybrid_scene *s, hgx_mpm_scene **continuum,
                                 hgx_dem_scene **discrete);
hg_status hgx_hybrid_scene_times(const hgx_hybrid_scene *s, double *dt, double *end_time);
/* The integrator_style string and the coupling settings the step uses. */
const char *hgx_hybrid_scene_style(const hgx_hybrid_scene *s);
void hgx_hybrid_scene_free(hgx_hybrid_scene *s);

/* ---- zone resampling on the host (hg_resample.hpp) ----------------------------
 * The pieces of the zone resampling (what hybridgrains2dnew/ResamplingTools.cpp,
 * ZoneTools.cpp, AvoidAvoid.cpp, random.c and CircleRectangleIntersection.cpp
 * compute) as plain calls, and one resampling pass on a hybrid scene without
 * any device work. */

/* random.c: n tempered words of a stream seeded with seed (randomMT_i), and
 * n single-precision reals of another such stream (randomMT). */
hg_status hgx_mt1998(uint32_t seed, uint32_t n, uint32_t *words, float *reals);
/* CircleRectangleIntersection.cpp: area of [x0,x1]x[y0,y1] inside the circle. */
double hgx_circle_rect_area(double x0, double x1, double y0, double y1, double cx, double cy,
                            double r);
/* ZoneTools on given circles (2 + 1 per circle) and points (2 per point):
 * packing fraction (ls cells), signed distance (ls nodes), level-set and
 * MPM-cell zone indicators (0 continuum, 1 discrete, 2 hybrid) for a grid of
 * cells[0] x cells[1] MPM cells of width h at gmin and level-set cells of
 * width ls_h.  Output sizes: ls_cells = ceil(cells * h / ls_h) per axis. */
hg_status hgx_zone_tools(const double gmin[2], double h, const uint32_t cells[2], double ls_h,
                         uint32_t ncircles, const double *circle_pos, const double *circle_r,
                         uint32_t npoints, const double *points, double window, int32_t spc,
                         double phi_threshold, double level_set, double half_thickness,
                         int32_t allow_direct, double *phi, double *dist, int32_t *ls_indicator,
                         int32_t *mpm_indicator);

/* One pass of the resampling (HybridGrains2DSim::updateHybridStates_poorman's
 * resampling block) on a hybrid scene as HybridGrains2DSim starts it: the
 * continuum's points dropped, the grains as read, no contacts, zero
 * homogenised fields.  Passes can be repeated on the result. */
typedef struct hgx_resample hgx_resample;
hg_status hgx_resample_create(const hgx_hybrid_scene *s, hgx_resample **out);
/* report: dem before, after sampling, after deletion; mpm the same; zones */
hg_status hgx_resample_run(hgx_resample *r, int64_t report[7]);
hg_status hgx_resample_counts(const hgx_resample *r, uint32_t *npoints, uint32_t *nbodies);
/* any pointer may be NULL */
hg_status hgx_resample_points(const hgx_resample *r, double *q, double *v, double *m,
                              double *volume, double *m0, double *volume0, double *weight,
                              double *J, double *be_bar, double *hl);
hg_status hgx_resample_bodies(const hgx_resample *r, double *q, double *v, double *M,
                              double *M0, double *weight, double *radius, uint8_t *fixed);
/* MPM-cell zone indicators of the last pass (cell_count[0] x cell_count[1]) */
hg_status hgx_resample_indicators(const hgx_resample *r, int32_t *mpm_indicator);
void hgx_resample_free(hgx_resample *r);

#ifdef __cplusplus
}
#endif
#endif
