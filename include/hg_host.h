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

#ifdef __cplusplus
}
#endif
#endif
