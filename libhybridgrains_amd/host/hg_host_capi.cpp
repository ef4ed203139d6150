// hg_host_capi.cpp -- extern "C" wrappers of the scene readers.
#include <cstring>

#include "hg_host.h"
#include "hg_scene.hpp"

struct hgx_mpm_scene {
  hg::SimulationState st;
};
struct hgx_dem_scene {
  hg::RigidBody2DScene sc;
};
struct hgx_hybrid_scene {
  hg::HybridIntegratorSettings settings;
  hgx_mpm_scene continuum;
  hgx_dem_scene discrete;
};

static hg_status report(const std::string &msg, char *err, size_t errlen) {
  if (err && errlen) {
    std::strncpy(err, msg.c_str(), errlen - 1);
    err[errlen - 1] = '\0';
  }
  return HG_ERR_INVALID;
}

extern "C" {

hg_status hgx_mpm_scene_read(const char *path, hgx_mpm_scene **out, char *err, size_t errlen) {
  if (!path || !out) return report("null argument", err, errlen);
  auto *s = new hgx_mpm_scene();
  std::string e;
  if (!hg::readMPMGrains2DXML(path, s->st, &e)) {
    delete s;
    return report(e, err, errlen);
  }
  *out = s;
  return HG_OK;
}

hg_status hgx_mpm_scene_read_text(const char *xml, hgx_mpm_scene **out, char *err,
                                  size_t errlen) {
  if (!xml || !out) return report("null argument", err, errlen);
  auto *s = new hgx_mpm_scene();
  std::string e;
  if (!hg::readMPMGrains2DXMLText(xml, s->st, &e)) {
    delete s;
    return report(e, err, errlen);
  }
  *out = s;
  return HG_OK;
}

hg_status hgx_mpm_scene_desc(hgx_mpm_scene *s, int32_t precision, hg_mpm_desc *out) {
  if (!s || !out) return HG_ERR_INVALID;
  *out = s->st.desc(precision);
  return HG_OK;
}

hg_status hgx_mpm_scene_times(const hgx_mpm_scene *s, double *dt, double *end_time) {
  if (!s) return HG_ERR_INVALID;
  if (dt) *dt = s->st.dt.value();
  if (end_time) *end_time = s->st.end_time;
  return HG_OK;
}

void hgx_mpm_scene_free(hgx_mpm_scene *s) { delete s; }

hg_status hgx_dem_scene_read(const char *path, hgx_dem_scene **out, char *err, size_t errlen) {
  if (!path || !out) return report("null argument", err, errlen);
  auto *s = new hgx_dem_scene();
  std::string e;
  if (!hg::readRigidBody2DXML(path, s->sc, &e)) {
    delete s;
    return report(e, err, errlen);
  }
  *out = s;
  return HG_OK;
}

hg_status hgx_dem_scene_read_text(const char *xml, hgx_dem_scene **out, char *err,
                                  size_t errlen) {
  if (!xml || !out) return report("null argument", err, errlen);
  auto *s = new hgx_dem_scene();
  std::string e;
  if (!hg::readRigidBody2DXMLText(xml, s->sc, &e)) {
    delete s;
    return report(e, err, errlen);
  }
  *out = s;
  return HG_OK;
}

hg_status hgx_dem_scene_desc(hgx_dem_scene *s, int32_t precision, double cell_scale,
                             hg_dem_desc *out) {
  if (!s || !out) return HG_ERR_INVALID;
  *out = s->sc.desc(precision, cell_scale > 0.0 ? cell_scale : s->sc.spatial_grid_scale);
  return HG_OK;
}

hg_status hgx_dem_scene_times(const hgx_dem_scene *s, double *dt, double *end_time) {
  if (!s) return HG_ERR_INVALID;
  if (dt) *dt = s->sc.dt.value();
  if (end_time) *end_time = s->sc.end_time;
  return HG_OK;
}

void hgx_dem_scene_free(hgx_dem_scene *s) { delete s; }

hg_status hgx_hybrid_scene_read(const char *path, hgx_hybrid_scene **out, char *err,
                                size_t errlen) {
  if (!path || !out) return report("null argument", err, errlen);
  auto *s = new hgx_hybrid_scene();
  std::string e;
  if (!hg::readHybridGrains2DXML(path, s->settings, &e) ||
      !hg::readRigidBody2DXML(s->settings.discrete_file_name, s->discrete.sc, &e) ||
      !hg::readMPMGrains2DXML(s->settings.continuum_file_name, s->continuum.st, &e)) {
    delete s;
    return report(e, err, errlen);
  }
  *out = s;
  return HG_OK;
}

hg_status hgx_hybrid_scene_parts(hgx_hybrid_scene *s, hgx_mpm_scene **continuum,
                                 hgx_dem_scene **discrete) {
  if (!s) return HG_ERR_INVALID;
  if (continuum) *continuum = &s->continuum;
  if (discrete) *discrete = &s->discrete;
  return HG_OK;
}

hg_status hgx_hybrid_scene_times(const hgx_hybrid_scene *s, double *dt, double *end_time) {
  if (!s) return HG_ERR_INVALID;
  if (dt) *dt = s->settings.overall_dt.value();
  if (end_time) *end_time = s->settings.end_time;
  return HG_OK;
}

const char *hgx_hybrid_scene_style(const hgx_hybrid_scene *s) {
  return s ? s->settings.integrator_style.c_str() : "";
}

void hgx_hybrid_scene_free(hgx_hybrid_scene *s) { delete s; }

}  // extern "C"
