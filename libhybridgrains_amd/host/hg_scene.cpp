// hg_scene.cpp -- the three scene-file readers and the host state behind the
// C-ABI descriptors.
#include "hg_scene.hpp"

#include <algorithm>
#include <cmath>
#include <sstream>

#include "hg_xml.h"

namespace hg {

namespace {

bool set_err(std::string *error, const std::string &m) {
  if (error) *error = m;
  return false;
}

// StringUtilities::extractFromString: the first whitespace-separated token
// read by a stream (scisim/StringUtilities.h:68-72).
template <typename T>
bool read_value(const std::string &s, T &out) {
  std::stringstream ss(s);
  ss >> out;
  return !ss.fail();
}

// extractScalarFromString accepts "inf" and "-inf" as well
// (scisim/StringUtilities.h:74-91).
bool read_scalar(const std::string &s, double &out) {
  if (read_value(s, out)) return true;
  if (s == "inf") {
    out = std::numeric_limits<double>::infinity();
    return true;
  }
  if (s == "-inf") {
    out = -std::numeric_limits<double>::infinity();
    return true;
  }
  return false;
}

// StringUtilities::readScalarList: exactly n whitespace-separated scalars.
bool read_list(const std::string &s, int n, double *out) {
  std::stringstream ss(s);
  std::string tok;
  int k = 0;
  while (ss >> tok) {
    if (k >= n) return false;
    if (!read_scalar(tok, out[k])) return false;
    ++k;
  }
  return k == n;
}

bool req_attr(const xml::Node &n, const char *key, const std::string *&out, std::string *error) {
  out = n.attr(key);
  if (!out)
    return set_err(error, std::string("Failed to locate ") + key + " attribute for " + n.name +
                              " node.");
  return true;
}

bool req_list(const xml::Node &n, const char *key, int cnt, double *out, std::string *error) {
  const std::string *a;
  if (!req_attr(n, key, a, error)) return false;
  if (!read_list(*a, cnt, out))
    return set_err(error, std::string("Failed to load ") + key + " attribute for " + n.name +
                              ", must provide " + std::to_string(cnt) + " scalars.");
  return true;
}

bool req_scalar(const xml::Node &n, const char *key, double &out, std::string *error) {
  const std::string *a;
  if (!req_attr(n, key, a, error)) return false;
  if (!read_value(*a, out))
    return set_err(error, std::string("Failed to load ") + key + " attribute for " + n.name +
                              ", must provide a scalar.");
  return true;
}

std::string strip_ws(const std::string &s) {
  std::string o;
  for (char c : s)
    if (!std::isspace(static_cast<unsigned char>(c))) o.push_back(c);
  return o;
}

bool read_end_time(const xml::Node &root, double &end_time, std::string *error) {
  const xml::Node *nd = root.child("end_time");
  if (!nd) return true;
  double t;
  if (!req_scalar(*nd, "t", t, error)) return false;
  if (!(t > 0.0))
    return set_err(error, "Failed to parse t attribute for end_time. Value must be a positive scalar.");
  end_time = t;
  return true;
}

const xml::Node *root_of(const xml::Node &doc, const char *name) {
  for (const auto &c : doc.children)
    if (c->name == name) return c.get();
  return nullptr;
}

}  // namespace

bool parseRational(const std::string &in, Rational &out) {
  const auto slashes = std::count(in.begin(), in.end(), '/');
  const auto periods = std::count(in.begin(), in.end(), '.');
  if (slashes > 2 || periods > 2) return false;
  if (slashes == 1 && periods != 0) return false;
  if (periods == 1 && slashes != 0) return false;
  const std::string s = strip_ws(in);
  if (slashes != 1) {
    // decimal: characteristic and mantissa split at the last period
    const size_t p = s.rfind('.');
    std::string left = p == std::string::npos ? s : s.substr(0, p);
    std::string right = p == std::string::npos ? std::string() : s.substr(p + 1);
    while (!right.empty() && right.back() == '0') right.pop_back();
    if (left.empty()) left = "0";
    if (left == "+") left = "+0";
    if (left == "-") left = "-0";
    std::int64_t li = 0, ri = 0;
    if (!read_value(left, li)) return false;
    if (!right.empty() && !read_value(right, ri)) return false;
    std::int64_t den = 1;
    for (size_t k = 0; k < right.size(); ++k) den *= 10;
    out.numerator = li * den + (li < 0 ? -1 : 1) * ri;
    out.denominator = den;
  } else {
    const size_t p = s.rfind('/');
    std::int64_t a = 0, b = 0;
    if (!read_value(s.substr(0, p), a) || !read_value(s.substr(p + 1), b)) return false;
    if (b == 0) return false;
    if (b < 0) {
      a = -a;
      b = -b;
    }
    out.numerator = a;
    out.denominator = b;
  }
  // reduced form, as Rational::simplify keeps it
  std::int64_t g = std::llabs(out.numerator), h = out.denominator;
  while (h) {
    const std::int64_t t = g % h;
    g = h;
    h = t;
  }
  if (g > 1) {
    out.numerator /= g;
    out.denominator /= g;
  }
  return true;
}

// ---- mpmgrains2d ------------------------------------------------------------

void PhysicsGrid::setDimensions(const double grid_min[2], const double grid_max[2],
                                double cell_w) {
  for (int d = 0; d < 2; ++d) {
    const double center = 0.5 * (grid_min[d] + grid_max[d]);
    const double width = grid_max[d] - grid_min[d];
    cell_count[d] = static_cast<std::uint32_t>(std::ceil(width / cell_w));
    const double new_width = double(cell_count[d]) * cell_w;
    min[d] = center - 0.5 * new_width;
    max[d] = center + 0.5 * new_width;
  }
  cell_width = cell_w;
}

hg_mpm_desc SimulationState::desc(int precision) const {
  hg_mpm_desc d{};
  d.precision = precision;
  d.basis = basis;
  d.grid_min[0] = physics_grid.min[0];
  d.grid_min[1] = physics_grid.min[1];
  d.cell_width = physics_grid.cell_width;
  d.cell_count[0] = physics_grid.cell_count[0];
  d.cell_count[1] = physics_grid.cell_count[1];
  d.shear_modulus = shear_modulus;
  d.bulk_modulus = bulk_modulus;
  d.dp_alpha = Drucker_Prager_alpha;
  d.pic_alpha = alpha;
  d.gravity[0] = near_earth_gravity[0];
  d.gravity[1] = near_earth_gravity[1];
  plane_x_.clear();
  plane_n_.clear();
  plane_lower_.clear();
  plane_behavior_.clear();
  for (const MPMStaticPlane &p : static_planes) {
    plane_x_.insert(plane_x_.end(), {p.x[0], p.x[1]});
    plane_n_.insert(plane_n_.end(), {p.n[0], p.n[1]});
    plane_lower_.push_back(p.lower_bound);
    plane_behavior_.push_back(p.boundary_behavior);
  }
  d.nplanes = static_cast<std::uint32_t>(static_planes.size());
  d.plane_x = plane_x_.data();
  d.plane_n = plane_n_.data();
  d.plane_lower = plane_lower_.data();
  d.plane_behavior = plane_behavior_.data();
  const MaterialPoints &mp = material_points;
  d.npoints = static_cast<std::uint32_t>(mp.npoints());
  d.q = mp.q.data();
  d.v = mp.v.data();
  d.m = mp.m.data();
  d.volume = mp.volume.data();
  d.hl = mp.hl.data();
  d.be_bar = mp.be_bar.data();
  d.J = mp.J.data();
  d.kinematically_scripted =
      mp.kinematically_scripted.size() == mp.npoints() ? mp.kinematically_scripted.data() : nullptr;
  return d;
}

// pointInsidePlane (InitialSimulationState.cpp:31-43): the first plane whose
// lower bound the point is below ends the test with "outside".
static bool point_inside_plane(const std::vector<MPMStaticPlane> &planes, const double p[2]) {
  for (const MPMStaticPlane &pl : planes) {
    if (p[1] < pl.lower_bound) return false;
    if (pl.distanceToPoint(p) < 0.0) return true;
  }
  return false;
}

bool rasterizeRectangles(const std::vector<RectangleRegion> &regions, SimulationState &st,
                         std::string *error) {
  const PhysicsGrid &g = st.physics_grid;
  MaterialPoints &mp = st.material_points;
  for (const RectangleRegion &r : regions) {
    for (int d = 0; d < 2; ++d) {
      if (r.min[d] < g.min[d])
        return set_err(error, "Error, initial rectangle region is below the simulation grid.");
      if (r.max[d] > g.max[d])
        return set_err(error, "Error, initial rectangle region is above the simulation grid.");
    }
    unsigned start[2], end[2];
    for (int d = 0; d < 2; ++d) {
      start[d] = static_cast<unsigned>(std::floor((r.min[d] - g.min[d]) / g.cell_width));
      end[d] = static_cast<unsigned>(std::ceil((r.max[d] - g.min[d]) / g.cell_width));
    }
    const double cx = 0.5 * (r.min[0] + r.max[0]), cy = 0.5 * (r.min[1] + r.max[1]);
    const double hl = 0.5 * g.cell_width / double(r.npoints_per_dim);
    const double vol = hl * hl * 4.0;
    const double mass = r.rho * vol;
    const double npd = double(r.npoints_per_dim);
    for (unsigned gy = start[1]; gy < end[1]; ++gy)
      for (unsigned gx = start[0]; gx < end[0]; ++gx)
        for (unsigned sy = 0; sy < r.npoints_per_dim; ++sy)
          for (unsigned sx = 0; sx < r.npoints_per_dim; ++sx) {
            const double p[2] = {g.min[0] + g.cell_width * (double(gx) + (double(sx) + 0.5) / npd),
                                 g.min[1] + g.cell_width * (double(gy) + (double(sy) + 0.5) / npd)};
            if (p[0] < r.min[0] || p[1] < r.min[1] || p[0] > r.max[0] || p[1] > r.max[1]) continue;
            if (point_inside_plane(st.static_planes, p)) continue;
            // v + omega * (e_z x arm)
            const double ax = p[0] - cx, ay = p[1] - cy;
            mp.q.insert(mp.q.end(), {p[0], p[1]});
            mp.v.insert(mp.v.end(), {r.v[0] + r.omega * (-ay), r.v[1] + r.omega * ax});
            mp.be_bar.insert(mp.be_bar.end(), {1.0, 0.0, 0.0, 1.0});
            mp.J.push_back(1.0);
            mp.m.push_back(mass);
            mp.volume.push_back(vol);
            mp.hl.push_back(hl);
            mp.kinematically_scripted.push_back(0);
          }
  }
  return true;
}

static bool read_mpm_root(const xml::Node &root, SimulationState &st, std::string *error) {
  // rectangles (MPMGrains2DParser.cpp:226-351)
  std::vector<RectangleRegion> rects;
  for (const xml::Node *nd : root.children_named("rectangle")) {
    RectangleRegion r{};
    if (!req_list(*nd, "min", 2, r.min, error)) return false;
    if (!req_list(*nd, "max", 2, r.max, error)) return false;
    if (r.min[0] >= r.max[0] || r.min[1] >= r.max[1])
      return set_err(error, "Failed to load rectangle region. Region max must be greater than region min.");
    if (!req_scalar(*nd, "density", r.rho, error)) return false;
    if (!(r.rho > 0.0)) return set_err(error, "Failed to load density attribute for rectangle node.");
    const std::string *a;
    if (!req_attr(*nd, "cell_samples_per_dim", a, error)) return false;
    int spd;
    if (!read_value(*a, spd) || spd < 0)
      return set_err(error, "Failed to load cell_samples_per_dim attribute for rectangle node.");
    r.npoints_per_dim = unsigned(spd);
    if (!req_list(*nd, "vel", 2, r.v, error)) return false;
    if (!req_scalar(*nd, "omega", r.omega, error)) return false;
    rects.push_back(r);
  }
  // material (:436-497)
  const xml::Node *mat = root.child("material");
  if (!mat) return set_err(error, "Failed to locate material node.");
  if (!req_scalar(*mat, "shear_modulus", st.shear_modulus, error)) return false;
  if (!req_scalar(*mat, "bulk_modulus", st.bulk_modulus, error)) return false;
  if (!req_scalar(*mat, "alpha", st.Drucker_Prager_alpha, error)) return false;
  if (st.shear_modulus < 0 || st.bulk_modulus < 0 || st.Drucker_Prager_alpha < 0)
    return set_err(error, "Failed to load material: moduli and alpha must be non-negative.");
  // grid (:499-571)
  const xml::Node *gr = root.child("grid");
  if (!gr) return set_err(error, "Failed to locate grid node.");
  double gmin[2], gmax[2], cw;
  if (!req_list(*gr, "min", 2, gmin, error)) return false;
  if (!req_list(*gr, "max", 2, gmax, error)) return false;
  if (!req_scalar(*gr, "cell_width", cw, error)) return false;
  if (!(cw > 0.0)) return set_err(error, "Failed to load cell_width attribute for grid node.");
  if (gmax[0] <= gmin[0] || gmax[1] <= gmin[1])
    return set_err(error, "Grid minimum must be less than grid maximum.");
  // integrator (:80-208)
  const xml::Node *in = root.child("integrator");
  if (!in) return set_err(error, "Failed to locate integrator node.");
  const std::string *a;
  if (!req_attr(*in, "dt", a, error)) return false;
  if (!parseRational(*a, st.dt) || !st.dt.positive())
    return set_err(error, "Failed to load dt attribute for integrator. Must provide a positive number.");
  st.dt_string = *a;
  if (!req_scalar(*in, "alpha", st.alpha, error)) return false;
  if (st.alpha < 0.0 || st.alpha > 1.0)
    return set_err(error, "Failed to load alpha attribute for integrator. Must provide a scalar in [0, 1].");
  if (!req_attr(*in, "shape_function_order", a, error)) return false;
  int order;
  if (!read_value(*a, order) || order <= 0)
    return set_err(error, "Failed to load shape_function_order attribute for integrator.");
  if (order != 1 && order != 3)
    return set_err(error, "Supported shape_function_order values are 1 (linear) and 3 (cubic).");
  bool ugimp = false;
  if (const std::string *c = in->attr("shape_function_category")) {
    if (*c == "uGIMP") ugimp = true;
    else if (*c != "Standard")
      return set_err(error, "Supported shape_function_category values are 'uGIMP' and 'Standard'.");
  }
  // InitialSimulationState.cpp:163-187
  if (ugimp) {
    if (order != 1) return set_err(error, "uGIMP basis functions are not first order.");
    st.basis = HG_BASIS_UGIMP_LINEAR;
  } else {
    st.basis = order == 1 ? HG_BASIS_LINEAR : HG_BASIS_CUBIC;
  }
  // static planes (:353-434)
  for (const xml::Node *nd : root.children_named("static_plane")) {
    MPMStaticPlane p;
    if (!req_list(*nd, "x", 2, p.x, error)) return false;
    if (!req_list(*nd, "n", 2, p.n, error)) return false;
    if (!req_attr(*nd, "boundary_behavior", a, error)) return false;
    if (*a == "sticking") p.boundary_behavior = MPMStaticPlane::STICKING;
    else if (*a == "sliding") p.boundary_behavior = MPMStaticPlane::SLIDING;
    else return set_err(error, "Invalid boundary_behavior attribute specified for static_plane. Valid values are: sticking, sliding");
    if (const std::string *lb = nd->attr("lower_bound")) {
      if (!read_value(*lb, p.lower_bound))
        return set_err(error, "Failed to load lower_bound attribute for static_plane.");
    }
    const double nn = std::sqrt(p.n[0] * p.n[0] + p.n[1] * p.n[1]);
    p.n[0] /= nn;
    p.n[1] /= nn;
    st.static_planes.push_back(p);
  }
  // gravity (:58-78): the last near_earth_gravity wins
  st.near_earth_gravity[0] = st.near_earth_gravity[1] = 0.0;
  for (const xml::Node *nd : root.children_named("near_earth_gravity"))
    if (!req_list(*nd, "f", 2, st.near_earth_gravity, error)) return false;
  if (!read_end_time(root, st.end_time, error)) return false;

  st.physics_grid.setDimensions(gmin, gmax, cw);
  if (rects.empty()) return set_err(error, "Provide at least one rectangle region.");
  // InitialSimulationState::generateSimulationState (:203-224): the first
  // rectangle sets the particle size and density the resampling uses
  st.initial_particle_size = st.physics_grid.cell_width / double(rects[0].npoints_per_dim);
  st.material_density = rects[0].rho;
  return rasterizeRectangles(rects, st, error);
}

bool readMPMGrains2DXMLText(const std::string &text, SimulationState &state, std::string *error) {
  std::string perr;
  auto doc = xml::parse(text, &perr);
  if (!doc) return set_err(error, "Failed to parse scene file: " + perr);
  const xml::Node *root = root_of(*doc, "mpmgrains2d");
  if (!root) return set_err(error, "Failed to locate root node mpmgrains2d.");
  SimulationState st;
  if (!read_mpm_root(*root, st, error)) return false;
  state = std::move(st);
  return true;
}

bool readMPMGrains2DXML(const std::string &file_name, SimulationState &state, std::string *error) {
  std::string perr;
  auto doc = xml::parse_file(file_name, &perr);
  if (!doc) return set_err(error, "Failed to read scene file: " + file_name + ": " + perr);
  const xml::Node *root = root_of(*doc, "mpmgrains2d");
  if (!root) return set_err(error, "Failed to locate root node mpmgrains2d in " + file_name);
  SimulationState st;
  if (!read_mpm_root(*root, st, error)) return false;
  state = std::move(st);
  return true;
}

// ---- rigidbody2d ------------------------------------------------------------

double RigidBody2DScene::cellWidth(double scale) const {
  const RigidBody2DState &s = state;
  if (s.nbodies() == 0) return 1.0;
  // mean AABB extent: both axes have extent 2r for a circle at rest
  double sum = 0.0, rmax = 0.0;
  for (double r : s.radius) {
    sum += 2.0 * r;
    rmax = std::max(rmax, r);
  }
  const double h = sum / double(s.nbodies()) * scale;
  return std::max(h, 2.0 * rmax);
}

hg_dem_desc RigidBody2DScene::desc(int precision, double scale) const {
  hg_dem_desc d{};
  d.precision = precision;
  d.nbodies = static_cast<std::uint32_t>(state.nbodies());
  d.q = state.q.data();
  d.v = state.v.data();
  d.mass = state.M.data();
  d.radius = state.radius.data();
  d.fixed = state.fixed.data();
  d.gravity[0] = state.gravity[0];
  d.gravity[1] = state.gravity[1];
  plane_x_.clear();
  plane_n_.clear();
  for (const RigidBody2DStaticPlane &p : state.planes) {
    plane_x_.insert(plane_x_.end(), {p.x[0], p.x[1]});
    plane_n_.insert(plane_n_.end(), {p.n[0], p.n[1]});
  }
  d.nplanes = static_cast<std::uint32_t>(state.planes.size());
  d.plane_x = plane_x_.data();
  d.plane_n = plane_n_.data();
  d.kn = penalty.kn;
  d.kt = penalty.kt;
  d.gamman = penalty.gamman;
  d.gammat = penalty.gammat;
  d.mu = penalty.mu;
  d.cell_width = cellWidth(scale);
  return d;
}

static bool read_dem_root(const xml::Node &root, RigidBody2DScene &sc, std::string *error) {
  if (!read_end_time(root, sc.end_time, error)) return false;
  const xml::Node *pm = root.child("penalty_impact_map");
  if (!pm)
    return set_err(error, "Only penalty_impact_map scenes are supported (no penalty_impact_map node).");
  {
    const std::string *a;
    if (!req_attr(*pm, "dt", a, error)) return false;
    if (!parseRational(*a, sc.dt) || !sc.dt.positive())
      return set_err(error, "Error, failed to load dt's value for penalty_impact_map.");
    sc.dt_string = *a;
    if (!req_scalar(*pm, "kn", sc.penalty.kn, error)) return false;
    if (!req_scalar(*pm, "kt", sc.penalty.kt, error)) return false;
    if (!req_scalar(*pm, "gamman", sc.penalty.gamman, error)) return false;
    if (!req_scalar(*pm, "gammat", sc.penalty.gammat, error)) return false;
    if (!req_scalar(*pm, "mu", sc.penalty.mu, error)) return false;
  }
  // collision_detection (RigidBody2DSceneParser.cpp:904-974)
  const xml::Node *cd = root.child("collision_detection");
  if (!cd) return set_err(error, "Failed to locate collision_detection node.");
  {
    const std::string *a;
    if (!req_attr(*cd, "broadphase", a, error)) return false;
    if (*a != "spatial_grid")
      return set_err(error, "Invalid broadphase; valid options are: spatial_grid");
    if (!req_scalar(*cd, "cell_scale", sc.spatial_grid_scale, error)) return false;
    if (!(sc.spatial_grid_scale > 0.0))
      return set_err(error, "Failed to load cell_scale attribute, must provide a positive scalar.");
    if (!req_attr(*cd, "bandwidth_reduction", a, error)) return false;
    if (*a == "none") sc.reduce_bandwidth = false;
    else if (*a == "cuthill_mckee") sc.reduce_bandwidth = true;
    else return set_err(error, "Invalid bandwidth_reduction; valid options are: none, cuthill_mckee");
  }
  RigidBody2DState &s = sc.state;
  // gravity forces add up (NearEarthGravityForce per node, :587-609)
  for (const xml::Node *nd : root.children_named("near_earth_gravity")) {
    double f[2];
    if (!req_list(*nd, "f", 2, f, error)) return false;
    s.gravity[0] += f[0];
    s.gravity[1] += f[1];
  }
  for (const xml::Node *nd : root.children_named("static_plane")) {
    RigidBody2DStaticPlane p;
    if (!req_list(*nd, "x", 2, p.x, error)) return false;
    if (!req_list(*nd, "n", 2, p.n, error)) return false;
    const double nn = std::sqrt(p.n[0] * p.n[0] + p.n[1] * p.n[1]);
    if (nn == 0.0) return set_err(error, "Failed to load n attribute for static_plane, must provide a nonzero vector.");
    p.n[0] /= nn;
    p.n[1] /= nn;
    s.planes.push_back(p);
  }
  if (root.child("planar_portal") || root.child("lees_edwards_portal"))
    return set_err(error, "Planar portals are not supported by this engine.");
  if (root.child("static_drum")) return set_err(error, "Static drums are not supported by this engine.");
  for (const xml::Node *nd : root.children_named("geometry")) {
    const std::string *t;
    if (!req_attr(*nd, "type", t, error)) return false;
    if (*t != "circle")
      return set_err(error, "Geometry type '" + *t + "' is not supported by this engine (circles only).");
    double r;
    if (!req_scalar(*nd, "r", r, error)) return false;
    if (!(r > 0.0)) return set_err(error, "Failed to load r attribute for circle geometry, must provide a positive scalar.");
    s.geometry_radius.push_back(r);
  }
  // bodies (:715-902); CircleGeometry::computeMassAndInertia: m = rho pi r^2,
  // I = m r^2 / 2 (CircleGeometry.cpp:44-48)
  for (const xml::Node *nd : root.children_named("rigid_body")) {
    double x[2], v[2], theta, omega, rho;
    if (!req_list(*nd, "x", 2, x, error)) return false;
    if (!req_scalar(*nd, "theta", theta, error)) return false;
    if (!req_list(*nd, "v", 2, v, error)) return false;
    if (!req_scalar(*nd, "omega", omega, error)) return false;
    if (!req_scalar(*nd, "rho", rho, error)) return false;
    if (!(rho > 0.0)) return set_err(error, "Failed to load rho attribute for rigid_body, must provide a single positive scalar.");
    const std::string *a;
    if (!req_attr(*nd, "geo_idx", a, error)) return false;
    int gi;
    if (!read_value(*a, gi) || gi >= int(s.geometry_radius.size()))
      return set_err(error, "Failed to load geo_idx attribute for rigid_body, must provide an index below the number of geometry instances.");
    if (gi < 0) gi += int(s.geometry_radius.size());
    if (gi < 0) return set_err(error, "geo_idx out of range.");
    std::uint8_t fixed = 0;
    if (const std::string *f = nd->attr("fixed")) {
      bool fv;
      if (!read_value(*f, fv)) return set_err(error, "Failed to load fixed attribute for rigid_body, fixed must be a boolean.");
      fixed = fv ? 1 : 0;
    }
    const double r = s.geometry_radius[size_t(gi)];
    const double m = rho * M_PI * r * r;
    const double I = 0.5 * m * r * r;
    s.q.insert(s.q.end(), {x[0], x[1], theta});
    s.v.insert(s.v.end(), {v[0], v[1], omega});
    s.M.insert(s.M.end(), {m, m, I});
    s.radius.push_back(r);
    s.fixed.push_back(fixed);
    s.geometry_indices.push_back(std::uint32_t(gi));
  }
  return true;
}

bool readRigidBody2DXMLText(const std::string &text, RigidBody2DScene &scene, std::string *error) {
  std::string perr;
  auto doc = xml::parse(text, &perr);
  if (!doc) return set_err(error, "Failed to parse scene file: " + perr);
  const xml::Node *root = root_of(*doc, "rigidbody2d_scene");
  if (!root) return set_err(error, "Failed to locate root node rigidbody2d_scene.");
  RigidBody2DScene sc;
  if (!read_dem_root(*root, sc, error)) return false;
  scene = std::move(sc);
  return true;
}

bool readRigidBody2DXML(const std::string &file_name, RigidBody2DScene &scene, std::string *error) {
  std::string perr;
  auto doc = xml::parse_file(file_name, &perr);
  if (!doc) return set_err(error, "Failed to read scene file: " + file_name + ": " + perr);
  const xml::Node *root = root_of(*doc, "rigidbody2d_scene");
  if (!root) return set_err(error, "Failed to locate root node rigidbody2d_scene in " + file_name);
  RigidBody2DScene sc;
  if (!read_dem_root(*root, sc, error)) return false;
  scene = std::move(sc);
  return true;
}

// ---- hybridgrains2dnew --------------------------------------------------------

static void opt_scalar(const xml::Node &n, const char *key, double &out) {
  if (const std::string *a = n.attr(key)) read_value(*a, out);
}
static void opt_int(const xml::Node &n, const char *key, int &out) {
  if (const std::string *a = n.attr(key)) read_value(*a, out);
}
static void opt_flag(const xml::Node &n, const char *key, bool &out) {
  if (const std::string *a = n.attr(key)) {
    int v = 0;
    if (read_value(*a, v)) out = v != 0;
  }
}

bool readHybridGrains2DXML(const std::string &file_name, HybridIntegratorSettings &out,
                           std::string *error) {
  std::string perr;
  auto doc = xml::parse_file(file_name, &perr);
  if (!doc) return set_err(error, "Failed to read scene file: " + file_name + ": " + perr);
  const xml::Node *root = root_of(*doc, "hybridgrains2d_scene");
  if (!root)
    return set_err(error, "Failed to locate root node hybridgrains2d_scene in xml scene file: " + file_name);
  HybridIntegratorSettings s;
  const std::string *a;
  const xml::Node *nd = root->child("discrete_sim");
  if (!nd) return set_err(error, "Failed to locate discrete_sim node.");
  if (!req_attr(*nd, "file_name", a, error)) return false;
  s.discrete_file_name = *a;
  nd = root->child("continuum_sim");
  if (!nd) return set_err(error, "Failed to locate continuum_sim node.");
  if (!req_attr(*nd, "file_name", a, error)) return false;
  s.continuum_file_name = *a;
  nd = root->child("time_step");
  if (!nd) return set_err(error, "Failed to locate time_step node.");
  if (!req_attr(*nd, "dt", a, error)) return false;
  if (!parseRational(*a, s.overall_dt) || !s.overall_dt.positive())
    return set_err(error, "Failed to load dt attribute for time_step, must provide a positive rational.");
  s.time_step_string = *a;
  nd = root->child("end_time");
  if (!nd) return set_err(error, "Failed to locate end_time node.");
  if (!req_attr(*nd, "t", a, error)) return false;
  {
    Rational et;  // a rational here (HybridGrains2DSceneParser.cpp:126-149)
    if (!parseRational(*a, et) || !et.positive())
      return set_err(error, "Failed to load t attribute for end_time. Must provide a positive number.");
    s.end_time = et.value();
  }
  // coupling_force (HybridGrains2DSceneParser.cpp:150-255)
  if (const xml::Node *cf = root->child("coupling_force")) {
    PoormanEnrichmentSettings &p = s.poorman_settings;
    p.enabled = true;
    opt_scalar(*cf, "rzoneHalfThickness", p.rzone_half_thickness);
    opt_scalar(*cf, "demRMean", p.dem_r_mean);
    opt_scalar(*cf, "demRStd", p.dem_r_std);
    opt_scalar(*cf, "phi_window_size", p.phi_window_size);
    opt_scalar(*cf, "level_set_cell_width", p.level_set_cell_width);
    opt_int(*cf, "phi_samples_per_cell_side", p.phi_samples_per_cell_side);
    opt_scalar(*cf, "target_dense_packing_fraction", p.target_dense_packing_fraction);
    opt_int(*cf, "avoidAVoidFreq", p.avoidAVoidFreq);
    opt_scalar(*cf, "phiThreshold", p.phi_threshold);
    opt_scalar(*cf, "rzoneLevelSet", p.rzone_level_set);
    opt_scalar(*cf, "rho_dem", p.rho_dem);
    opt_flag(*cf, "allow_direct_transitions_between_discrete_and_continuum",
             p.allow_direct_transitions_between_discrete_and_continuum);
    opt_flag(*cf, "newly_converted_continuum_stress_free", p.newly_converted_continuum_stress_free);
    opt_flag(*cf, "homogenize_stress", p.homogenize_stress);
    opt_flag(*cf, "grid_smoothing_homogenized_stress", p.grid_smoothing_homogenized_stress);
    opt_flag(*cf, "homogenize_velocity", p.homogenize_velocity);
    opt_flag(*cf, "grid_smoothing_homogenized_velocity", p.grid_smoothing_homogenized_velocity);
  }
  nd = root->child("integrator_style");
  if (!nd) return set_err(error, "Failed to locate integrator_style node.");
  if (!req_attr(*nd, "style", a, error)) return false;
  static const char *styles[] = {"old_version", "old_version_node_node",
                                 "prediction_correction_advection",
                                 "prediction_correction_advection_node_node", "iterative"};
  if (std::find_if(std::begin(styles), std::end(styles),
                   [&](const char *x) { return *a == x; }) == std::end(styles))
    return set_err(error, "Invalid integrator_style: " + *a);
  s.integrator_style = *a;
  nd = root->child("hybrid_kinematic_script");
  if (!nd) return set_err(error, "Failed to locate hybrid_kinematic_script node.");
  if (!req_attr(*nd, "type", a, error)) return false;
  if (*a != "none" && *a != "script_hybrid_fronts")
    return set_err(error, "Invalid hybrid_kinematic_script type: " + *a);
  s.kinematic_script = *a;
  // scene names relative to the hybrid scene's folder (:505-518)
  const size_t slash = file_name.rfind('/');
  if (slash != std::string::npos) {
    const std::string dir = file_name.substr(0, slash);
    s.discrete_file_name = dir + '/' + s.discrete_file_name;
    s.continuum_file_name = dir + '/' + s.continuum_file_name;
  }
  out = std::move(s);
  return true;
}

}  // namespace hg
