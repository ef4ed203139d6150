// hg_host_capi.cpp -- extern "C" wrappers of the scene readers.
#include <algorithm>
#include <cstring>
#include <vector>

#include "hg_host.h"
#include "hg_resample.hpp"
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

hg_status hgx_mt1998(uint32_t seed, uint32_t n, uint32_t *words, float *reals) {
  hg::MersenneTwister1998 a(seed), b(seed);
  for (uint32_t k = 0; k < n; ++k) {
    if (words) words[k] = a.randomMT_i();
    if (reals) reals[k] = b.randomMT();
  }
  return HG_OK;
}

double hgx_circle_rect_area(double x0, double x1, double y0, double y1, double cx, double cy,
                            double r) {
  return hg::boxCircleOverlapArea(x0, x1, y0, y1, cx, cy, r);
}

hg_status hgx_zone_tools(const double gmin[2], double h, const uint32_t cells[2], double ls_h,
                         uint32_t ncircles, const double *circle_pos, const double *circle_r,
                         uint32_t npoints, const double *points, double window, int32_t spc,
                         double phi_threshold, double level_set, double half_thickness,
                         int32_t allow_direct, double *phi, double *dist, int32_t *ls_indicator,
                         int32_t *mpm_indicator) {
  if (!gmin || !cells || !(h > 0.0) || !(ls_h > 0.0)) return HG_ERR_INVALID;
  hg::PhysicsGrid g;
  const double gmax[2] = {gmin[0] + h * double(cells[0]), gmin[1] + h * double(cells[1])};
  g.setDimensions(gmin, gmax, h);
  hg::ZoneTools z;
  try {
    z.init(g, ls_h, allow_direct != 0);
  } catch (...) {
    return HG_ERR_INVALID;
  }
  std::vector<double> cp(circle_pos, circle_pos + 2 * std::size_t(ncircles));
  std::vector<double> cr(circle_r, circle_r + ncircles);
  std::vector<double> pts(points, points + 2 * std::size_t(npoints));
  z.updatePackingFraction(cp, cr, pts, window, spc);
  z.updateDistanceField(phi_threshold);
  z.computeRZones(level_set, half_thickness);
  const std::vector<double> &f = z.packingFractions();
  if (phi) std::copy(f.begin(), f.end(), phi);
  const std::vector<double> &d = z.distGrid().minDist;
  if (dist) std::copy(d.begin(), d.end(), dist);
  const std::vector<int> &li = z.levelSetIndicators();
  if (ls_indicator) std::copy(li.begin(), li.end(), ls_indicator);
  const std::vector<int> &mi = z.mpmIndicators();
  if (mpm_indicator) std::copy(mi.begin(), mi.end(), mpm_indicator);
  return HG_OK;
}

}  // extern "C"

struct hgx_resample {
  hg::ZoneTools zones;
  hg::ResamplingTools tools;
  hg::HybridResampleState s;
  hg::PoormanEnrichmentSettings p;
};

extern "C" {

hg_status hgx_resample_create(const hgx_hybrid_scene *sc, hgx_resample **out) {
  if (!sc || !out) return HG_ERR_INVALID;
  const hg::PoormanEnrichmentSettings &p = sc->settings.poorman_settings;
  if (!p.enabled) return HG_ERR_INVALID;
  auto *r = new hgx_resample();
  r->p = p;
  const hg::SimulationState &c = sc->continuum.st;
  r->s.continuum = c;
  r->s.continuum.material_points = hg::MaterialPoints{};
  r->s.initial_particle_size = c.initial_particle_size;
  r->s.material_density = c.material_density;
  r->s.discrete = sc->discrete.sc.state;
  r->s.dem_q0 = r->s.discrete.q;
  r->s.dem_M0 = r->s.discrete.M;
  r->s.dem_weight.assign(r->s.discrete.nbodies(), 1.0);
  r->s.homog_stress.assign(4 * c.physics_grid.numGridPoints(), 0.0);
  r->s.homog_velocity.assign(2 * c.physics_grid.numGridPoints(), 0.0);
  try {
    r->zones.init(c.physics_grid, p.level_set_cell_width,
                  p.allow_direct_transitions_between_discrete_and_continuum);
  } catch (...) {
    delete r;
    return HG_ERR_INVALID;
  }
  r->tools.init(p.dem_r_mean, p.dem_r_std);
  *out = r;
  return HG_OK;
}

hg_status hgx_resample_run(hgx_resample *r, int64_t report[7]) {
  if (!r) return HG_ERR_INVALID;
  hg::ResampleReport rep;
  try {
    rep = r->tools.resample(r->zones, r->s, r->p, double(r->p.avoidAVoidFreq));
  } catch (...) {
    return HG_ERR_INVALID;
  }
  if (report) {
    report[0] = rep.dem_before;
    report[1] = rep.dem_after_sampling;
    report[2] = rep.dem_after_delete;
    report[3] = rep.mpm_before;
    report[4] = rep.mpm_after_sampling;
    report[5] = rep.mpm_after_delete;
    report[6] = int64_t(rep.rzones);
  }
  return HG_OK;
}

hg_status hgx_resample_counts(const hgx_resample *r, uint32_t *npoints, uint32_t *nbodies) {
  if (!r) return HG_ERR_INVALID;
  if (npoints) *npoints = uint32_t(r->s.continuum.material_points.npoints());
  if (nbodies) *nbodies = uint32_t(r->s.discrete.nbodies());
  return HG_OK;
}

hg_status hgx_resample_points(const hgx_resample *r, double *q, double *v, double *m,
                              double *volume, double *m0, double *volume0, double *weight,
                              double *J, double *be_bar, double *hl) {
  if (!r) return HG_ERR_INVALID;
  const hg::MaterialPoints &mp = r->s.continuum.material_points;
  auto put = [](const std::vector<double> &src, double *dst) {
    if (dst) std::copy(src.begin(), src.end(), dst);
  };
  put(mp.q, q);
  put(mp.v, v);
  put(mp.m, m);
  put(mp.volume, volume);
  put(r->s.mpm_m0, m0);
  put(r->s.mpm_volume0, volume0);
  put(r->s.mpm_weight, weight);
  put(mp.J, J);
  put(mp.be_bar, be_bar);
  put(mp.hl, hl);
  return HG_OK;
}

hg_status hgx_resample_bodies(const hgx_resample *r, double *q, double *v, double *M,
                              double *M0, double *weight, double *radius, uint8_t *fixed) {
  if (!r) return HG_ERR_INVALID;
  const hg::RigidBody2DState &d = r->s.discrete;
  auto put = [](const std::vector<double> &src, double *dst) {
    if (dst) std::copy(src.begin(), src.end(), dst);
  };
  put(d.q, q);
  put(d.v, v);
  put(d.M, M);
  put(r->s.dem_M0, M0);
  put(r->s.dem_weight, weight);
  put(d.radius, radius);
  if (fixed) std::copy(d.fixed.begin(), d.fixed.end(), fixed);
  return HG_OK;
}

hg_status hgx_resample_indicators(const hgx_resample *r, int32_t *mpm_indicator) {
  if (!r || !mpm_indicator) return HG_ERR_INVALID;
  const std::vector<int> &mi = r->zones.mpmIndicators();
  std::copy(mi.begin(), mi.end(), mpm_indicator);
  return HG_OK;
}

void hgx_resample_free(hgx_resample *r) { delete r; }

}  // extern "C"
