// hg_resample.cpp -- the hybrid zone resampling on host state (hg_resample.hpp).
// Each function cites the reference routine whose behaviour it follows.
#include "hg_resample.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <functional>
#include <memory>
#include <queue>
#include <sstream>
#include <stdexcept>

namespace hg {

// ---- random.c ---------------------------------------------------------------------
// MT19937 (M. Matsumoto, T. Nishimura, ACM TOMACS 8(1), 1998): 624-word state,
// twist with matrix 0x9908b0df, tempering (11, 7 & 0x9d2c5680, 15 & 0xefc60000,
// 18); the 1998 seeding fills the state with x_{k+1} = 69069 x_k mod 2^32.
void MersenneTwister1998::seedMT(std::uint32_t seed) {
  mt_[0] = seed;
  for (int k = 1; k < 624; ++k) mt_[k] = 69069u * mt_[k - 1];
  mti_ = 624;
}

void MersenneTwister1998::twist() {
  constexpr std::uint32_t kA = 0x9908b0dfu, kUpper = 0x80000000u, kLower = 0x7fffffffu;
  for (int k = 0; k < 624; ++k) {
    const std::uint32_t y = (mt_[k] & kUpper) | (mt_[(k + 1) % 624] & kLower);
    mt_[k] = mt_[(k + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? kA : 0u);
  }
  mti_ = 0;
}

std::uint32_t MersenneTwister1998::randomMT_i() {
  if (mti_ >= 624) {
    if (mti_ == 625) seedMT(4357u);
    twist();
  }
  std::uint32_t y = mt_[mti_++];
  y ^= y >> 11;
  y ^= (y << 7) & 0x9d2c5680u;
  y ^= (y << 15) & 0xefc60000u;
  y ^= y >> 18;
  return y;
}

void MersenneTwister1998::setState(const std::uint32_t *words, int position) {
  for (int k = 0; k < 624; ++k) mt_[k] = words[k];
  mti_ = position;
}

float MersenneTwister1998::randomMT() {
  // random.c: (float)y * (float)2^-32, a single-precision product
  return float(randomMT_i()) * float(2.3283064365386963e-10);
}

// ---- area of a box inside a circle ------------------------------------------------
// With the circle at the origin, the box's area inside it is the integral over
// u of the length of [y0, y1] inside [-c(u), c(u)], c(u) = sqrt(r^2 - u^2).
// Between the abscissae where c(u) meets |y0| or |y1| that length is one of
// 2c, c - y0, c + y1 or y1 - y0 (or zero), so the integral is a sum of closed
// forms over at most five pieces, with C(u) = (u c(u) + r^2 asin(u / r)) / 2
// the antiderivative of c.
namespace {
double semichord_integral(double u, double r) {
  const double t = std::max(-1.0, std::min(1.0, u / r));
  return 0.5 * (u * std::sqrt(std::max(0.0, r * r - u * u)) + r * r * std::asin(t));
}
}  // namespace

double boxCircleOverlapArea(double x0, double x1, double y0, double y1, double cx, double cy,
                            double r) {
  double ax = x0 - cx, bx = x1 - cx, ay = y0 - cy, by = y1 - cy;
  if (ax > bx) std::swap(ax, bx);
  if (ay > by) std::swap(ay, by);
  const double lo = std::max(ax, -r), hi = std::min(bx, r);
  if (!(lo < hi) || !(ay < r) || !(by > -r)) return 0.0;
  double cut[6] = {lo, hi};
  int n = 2;
  for (double y : {ay, by})
    if (std::fabs(y) < r) {
      const double u = std::sqrt(r * r - y * y);
      if (lo < -u && -u < hi) cut[n++] = -u;
      if (lo < u && u < hi) cut[n++] = u;
    }
  std::sort(cut, cut + n);
  double area = 0.0;
  for (int k = 0; k + 1 < n; ++k) {
    const double a = cut[k], b = cut[k + 1];
    if (!(a < b)) continue;
    const double mid = 0.5 * (a + b), c = std::sqrt(std::max(0.0, r * r - mid * mid));
    // the piece's length is kc c(u) + k1; which bound is the circle's
    // depends only on the piece
    const bool circle_top = c <= by, circle_bottom = -c >= ay;
    const double kc = circle_top ? (circle_bottom ? 2.0 : 1.0) : (circle_bottom ? 1.0 : 0.0);
    const double k1 = circle_top ? (circle_bottom ? 0.0 : -ay) : (circle_bottom ? by : by - ay);
    if (kc * c + k1 <= 0.0) continue;  // the box misses the circle over this piece
    area += kc * (semichord_integral(b, r) - semichord_integral(a, r)) + k1 * (b - a);
  }
  return area;
}

// ---- items binned by cell (compressed rows) -------------------------------------------
// Items (points, or boxes over several cells) counted into the cells of a
// uniform grid and laid out cell by cell, each cell's items in the order they
// were given; positions outside the grid fall into its border cells.
namespace {
class CellBins {
 public:
  CellBins(const double mn[2], const std::uint32_t n[2], double h) : h_(h) {
    for (int d = 0; d < 2; ++d) {
      min_[d] = mn[d];
      n_[d] = n[d];
    }
  }
  std::uint32_t cellOf(double x, int d) const {
    const double c = std::floor((x - min_[d]) / h_);
    return std::uint32_t(std::max(0.0, std::min(double(n_[d]) - 1.0, c)));
  }
  // items given as (lo, hi) cell ranges, one per item
  template <typename Range>
  void build(std::size_t count, Range range) {
    start_.assign(std::size_t(n_[0]) * n_[1] + 1, 0);
    for (int pass = 0; pass < 2; ++pass) {
      std::vector<std::uint32_t> fill;
      if (pass == 1) {
        for (std::size_t c = 1; c < start_.size(); ++c) start_[c] += start_[c - 1];
        items_.resize(start_.back());
        fill.assign(start_.begin(), start_.end() - 1);
      }
      for (std::size_t k = 0; k < count; ++k) {
        std::uint32_t lo[2], hi[2];
        range(k, lo, hi);
        for (std::uint32_t j = lo[1]; j <= hi[1]; ++j)
          for (std::uint32_t i = lo[0]; i <= hi[0]; ++i) {
            const std::size_t c = std::size_t(j) * n_[0] + i;
            if (pass == 0) ++start_[c + 1];
            else items_[fill[c]++] = int(k);
          }
      }
    }
  }
  // the items of cell (i, j)
  const int *begin(std::uint32_t i, std::uint32_t j) const {
    return items_.data() + start_[std::size_t(j) * n_[0] + i];
  }
  const int *end(std::uint32_t i, std::uint32_t j) const {
    return items_.data() + start_[std::size_t(j) * n_[0] + i + 1];
  }
  std::uint32_t size(int d) const { return n_[d]; }

 private:
  double min_[2], h_;
  std::uint32_t n_[2];
  std::vector<std::uint32_t> start_;
  std::vector<int> items_;
};
}  // namespace

// ---- distgrid.cpp, levelsetutil.cpp -----------------------------------------------------
void DistGrid::resize(const double mn[2], const std::uint32_t n[2], double cell_h) {
  for (int d = 0; d < 2; ++d) {
    min_coords[d] = mn[d];
    N[d] = n[d];
    max_coords[d] = mn[d] + cell_h * double(n[d]);
  }
  h = cell_h;
  minDist.assign(std::size_t(num_points()), 0.0);
  closestPoint.assign(std::size_t(num_points()), 0);
}

namespace {
void dist_node(const DistGrid &g, int i, int j, double p[2]) {
  p[0] = (g.max_coords[0] - g.min_coords[0]) * double(i) / double(g.N[0]) + g.min_coords[0];
  p[1] = (g.max_coords[1] - g.min_coords[1]) * double(j) / double(g.N[1]) + g.min_coords[1];
}
}  // namespace

// levelsetutil.cpp:27-142: exact distances to the samples within 3
// particleSize of each, then a breadth-first sweep in which every node hands
// its closest sample to its four neighbours while that shortens theirs.
void redistanceGrid(DistGrid &g, const std::vector<double> &samples, double particleSize) {
  const int np = g.num_points();
  const int W = int(g.N[0]) + 1;
  std::vector<double> absd(std::size_t(np), 1.0e10);
  std::fill(g.closestPoint.begin(), g.closestPoint.end(), -1);
  const std::size_t ns = samples.size() / 2;
  for (std::size_t p = 0; p < ns; ++p) {
    const double r = 3.0 * particleSize;
    const double sx = samples[2 * p], sy = samples[2 * p + 1];
    int lo[2], hi[2];
    const double mn[2] = {sx - r, sy - r}, mx[2] = {sx + r, sy + r};
    for (int d = 0; d < 2; ++d) {
      const double span = g.max_coords[d] - g.min_coords[d];
      lo[d] = std::max(0, std::min(int(g.N[d]),
                                   int(std::floor(double(g.N[d]) * (mn[d] - g.min_coords[d]) / span))));
      hi[d] = std::max(0, std::min(int(g.N[d]),
                                   int(std::ceil(double(g.N[d]) * (mx[d] - g.min_coords[d]) / span))));
    }
    for (int j = lo[1]; j <= hi[1]; ++j)
      for (int i = lo[0]; i <= hi[0]; ++i) {
        const int f = j * W + i;
        double gp[2];
        dist_node(g, i, j, gp);
        const double d = std::hypot(gp[0] - sx, gp[1] - sy);
        if (d < absd[std::size_t(f)]) {
          absd[std::size_t(f)] = d;
          g.closestPoint[std::size_t(f)] = int(p);
        }
      }
  }
  std::queue<int> q;
  for (int k = 0; k < np; ++k)
    if (g.closestPoint[std::size_t(k)] >= 0) q.push(k);
  while (!q.empty()) {
    const int k = q.front();
    q.pop();
    const int cp = g.closestPoint[std::size_t(k)];
    const int ci = k % W, cj = k / W;
    const int nb[4] = {ci > 0 ? cj * W + ci - 1 : -1, cj > 0 ? (cj - 1) * W + ci : -1,
                       ci < int(g.N[0]) ? cj * W + ci + 1 : -1,
                       cj < int(g.N[1]) ? (cj + 1) * W + ci : -1};
    for (int l = 0; l < 4; ++l) {
      const int n = nb[l];
      if (n < 0) continue;
      double gp[2];
      dist_node(g, n % W, n / W, gp);
      const double d = std::hypot(gp[0] - samples[2 * std::size_t(cp)],
                                  gp[1] - samples[2 * std::size_t(cp) + 1]);
      if (d < absd[std::size_t(n)] || g.closestPoint[std::size_t(n)] < 0) {
        absd[std::size_t(n)] = d;
        g.closestPoint[std::size_t(n)] = cp;
        q.push(n);
      }
    }
  }
  for (int k = 0; k < np; ++k) g.minDist[std::size_t(k)] = absd[std::size_t(k)];
}

// ---- ZoneTools.cpp ---------------------------------------------------------------------
void ZoneTools::init(const PhysicsGrid &grid, double ls_h, bool allow_direct) {
  const double ratio = grid.cell_width / ls_h;
  if (ratio < 1.0)
    throw std::runtime_error(
        "the level set cell width is larger than the mpm cell width (not supported)");
  allow_direct_ = allow_direct;
  std::uint32_t n[2];
  for (int d = 0; d < 2; ++d) {
    n[d] = std::uint32_t(std::ceil(double(grid.cell_count[d]) * grid.cell_width / ls_h));
    mpm_cells_[d] = grid.cell_count[d];
    mpm_min_[d] = grid.min[d];
  }
  mpm_h_ = grid.cell_width;
  dist_.resize(grid.min, n, ls_h);
  phi_.assign(std::size_t(dist_.num_cells()), 0.0);
  zone_ls_.assign(std::size_t(dist_.num_cells()), ZT_DISCRETE);
  const std::size_t nc = std::size_t(mpm_cells_[0]) * mpm_cells_[1];
  zone_mpm_.assign(nc, ZT_DISCRETE);
  prev_zone_mpm_.assign(nc, ZT_DISCRETE);
}

// :149-183 and :185-270.  Each level-set cell's packing fraction is the mean,
// over its spc x spc sample points, of the fraction of the window box around
// the point that the circles cover.  The windows lie on a regular lattice, so
// each circle adds its overlap to the windows whose boxes reach its bounding
// box, found by index arithmetic, and the cover of every window is then
// divided by the window's area.
void ZoneTools::updatePackingFraction(const std::vector<double> &cpos,
                                      const std::vector<double> &cr,
                                      const std::vector<double> &points, double window,
                                      int spc) {
  const double *rmin = dist_.min_coords;
  const std::uint32_t nx = dist_.N[0], ny = dist_.N[1];
  const double h = dist_.h, half = 0.5 * window;
  const std::uint32_t sx = nx * std::uint32_t(spc), sy = ny * std::uint32_t(spc);
  // the sample point of lattice index k along one axis (k = cell * spc + sub)
  auto sample = [&](int d, std::uint32_t k) {
    const std::uint32_t cell = k / std::uint32_t(spc), sub = k % std::uint32_t(spc);
    return rmin[d] + (double(cell) + (double(sub) + 0.5) / double(spc)) * h;
  };
  std::vector<double> cover(std::size_t(sx) * sy, 0.0);
  const double pitch = h / double(spc);
  auto lattice_range = [&](int d, double a, double b, std::uint32_t n, std::uint32_t &lo,
                           std::uint32_t &hi) {
    // samples whose window [s - half, s + half] reaches [a, b], with a margin
    // of one sample (those add nothing)
    const double kl = std::floor((a - half - rmin[d]) / pitch - 0.5) - 1.0;
    const double kh = std::ceil((b + half - rmin[d]) / pitch - 0.5) + 1.0;
    if (kh < 0.0 || kl > double(n) - 1.0 || n == 0) return false;
    lo = std::uint32_t(std::max(0.0, kl));
    hi = std::uint32_t(std::min(double(n) - 1.0, kh));
    return lo <= hi;
  };
  for (std::size_t c = 0; c < cr.size(); ++c) {
    const double x = cpos[2 * c], y = cpos[2 * c + 1], r = cr[c];
    std::uint32_t il, ih, jl, jh;
    if (!lattice_range(0, x - r, x + r, sx, il, ih) || !lattice_range(1, y - r, y + r, sy, jl, jh))
      continue;
    for (std::uint32_t j = jl; j <= jh; ++j) {
      const double py = sample(1, j);
      for (std::uint32_t i = il; i <= ih; ++i) {
        const double px = sample(0, i);
        cover[std::size_t(j) * sx + i] +=
            boxCircleOverlapArea(px - half, px + half, py - half, py + half, x, y, r);
      }
    }
  }
  for (std::uint32_t j = 0; j < ny; ++j)
    for (std::uint32_t i = 0; i < nx; ++i) {
      double pf = 0.0;
      for (int v = 0; v < spc; ++v)
        for (int u = 0; u < spc; ++u) {
          const std::uint32_t ki = i * std::uint32_t(spc) + std::uint32_t(u);
          const std::uint32_t kj = j * std::uint32_t(spc) + std::uint32_t(v);
          const double px = sample(0, ki), py = sample(1, kj);
          const double area = ((px + half) - (px - half)) * ((py + half) - (py - half));
          pf += cover[std::size_t(kj) * sx + ki] / area;
        }
      phi_[std::size_t(j) * nx + i] = pf / double(spc * spc);
    }
  // A material point sets the fraction to 1 in every level-set cell its
  // window reaches, except cells of the grid's first row and column.
  const std::size_t np = points.size() / 2;
  for (std::size_t p = 0; p < np; ++p) {
    int lo[2], hi[2];
    const std::uint32_t n[2] = {nx, ny};
    for (int d = 0; d < 2; ++d) {
      const double rel = points[2 * p + std::size_t(d)] - rmin[d];
      lo[d] = std::max(1, int(std::floor((rel - half) / h)));
      hi[d] = std::min(int(n[d]) - 1, int(std::floor((rel + half) / h)));
    }
    for (int j = lo[1]; j <= hi[1]; ++j)
      for (int i = lo[0]; i <= hi[0]; ++i) phi_[std::size_t(j) * nx + std::size_t(i)] = 1.0;
  }
}

// :283-444.  Boundary samples: 4 per edge where the packing fraction crosses
// the threshold (and on the grid border next to dense cells); the edges
// between cell rows are placed as the reference places them, one cell width
// above the shared edge.
void ZoneTools::updateDistanceField(double T) {
  std::vector<double> s;
  const std::uint32_t Nx = dist_.N[0], Ny = dist_.N[1];
  const double h = dist_.h, *mn = dist_.min_coords;
  auto edge = [&](double ax, double ay, double bx, double by) {
    for (int k = 0; k < 4; ++k) {
      const double t = (double(k) + 0.5) / 4.0;
      s.push_back((1.0 - t) * ax + t * bx);
      s.push_back((1.0 - t) * ay + t * by);
    }
  };
  auto dense = [&](std::size_t c) { return phi_[c] > T; };
  for (std::uint32_t j = 0; j <= Ny; ++j)
    for (std::uint32_t i = 0; i < Nx; ++i) {
      const double cl = mn[0] + double(i) * h, cr = mn[0] + double(i + 1) * h;
      const double ct = mn[1] + double(j + 1) * h;
      const std::size_t c = std::size_t(j) * Nx + i;
      bool add;
      if (j == 0) add = dense(c);
      else if (j == Ny) add = dense(std::size_t(j - 1) * Nx + i);
      else add = dense(std::size_t(j - 1) * Nx + i) != dense(c);
      if (add) edge(cl, ct, cr, ct);
    }
  for (std::uint32_t j = 0; j < Ny; ++j)
    for (std::uint32_t i = 0; i <= Nx; ++i) {
      const double cl = mn[0] + double(i) * h;
      const double cb = mn[1] + double(j) * h, ct = mn[1] + double(j + 1) * h;
      bool add;
      if (i == 0) add = dense(std::size_t(j) * Nx + i);
      else if (i == Nx) add = dense(std::size_t(j) * Nx + i - 1);
      else add = dense(std::size_t(j) * Nx + i - 1) != dense(std::size_t(j) * Nx + i);
      if (add) edge(cl, cb, cl, ct);
    }
  redistanceGrid(dist_, s, dist_.h);
  const std::uint32_t W = Nx + 1;
  std::vector<std::uint8_t> pos(std::size_t(dist_.num_points()), 0);
  for (std::uint32_t j = 0; j < Ny; ++j)
    for (std::uint32_t i = 0; i < Nx; ++i)
      if (dense(std::size_t(j) * Nx + i)) {
        const std::size_t n[4] = {std::size_t(j) * W + i, std::size_t(j) * W + i + 1,
                                  std::size_t(j + 1) * W + i, std::size_t(j + 1) * W + i + 1};
        for (std::size_t k : n) {
          dist_.minDist[k] = std::fabs(dist_.minDist[k]);
          pos[k] = 1;
        }
      }
  for (std::size_t k = 0; k < pos.size(); ++k)
    if (!pos[k]) dist_.minDist[k] = -std::fabs(dist_.minDist[k]);
}

// :446-494
void ZoneTools::estimateLevelSetLevelZoneIndicators(double L, double ht) {
  const std::uint32_t Nx = dist_.N[0], Ny = dist_.N[1], W = Nx + 1;
  for (std::uint32_t j = 0; j < Ny; ++j)
    for (std::uint32_t i = 0; i < Nx; ++i) {
      const double d[4] = {dist_.minDist[std::size_t(j) * W + i], dist_.minDist[std::size_t(j) * W + i + 1],
                           dist_.minDist[std::size_t(j + 1) * W + i],
                           dist_.minDist[std::size_t(j + 1) * W + i + 1]};
      const double lo = std::min(std::min(d[0], d[1]), std::min(d[2], d[3]));
      const double hi = std::max(std::max(d[0], d[1]), std::max(d[2], d[3]));
      int z;
      if (hi <= L - ht) z = ZT_DISCRETE;
      else if (lo >= L + ht) z = ZT_CONTINUUM;
      else z = ZT_HYBRID;
      zone_ls_[std::size_t(j) * Nx + i] = z;
    }
}

// :529-676
void ZoneTools::sampleMPMLevelZoneIndicators() {
  prev_zone_mpm_ = zone_mpm_;
  rzones_.clear();
  rzone_idx_.clear();
  direct_.clear();
  direct_idx_.clear();
  const std::uint32_t Nx = dist_.N[0], Ny = dist_.N[1];
  for (std::uint32_t j = 0; j < mpm_cells_[1]; ++j)
    for (std::uint32_t i = 0; i < mpm_cells_[0]; ++i) {
      const double cmin[2] = {mpm_min_[0] + double(i) * mpm_h_, mpm_min_[1] + double(j) * mpm_h_};
      const double cmax[2] = {mpm_min_[0] + (double(i) + 1.0) * mpm_h_,
                              mpm_min_[1] + (double(j) + 1.0) * mpm_h_};
      int lo[2], hi[2];
      const std::uint32_t n[2] = {Nx, Ny};
      for (int d = 0; d < 2; ++d) {
        lo[d] = int(std::floor((cmin[d] - dist_.min_coords[d]) / dist_.h));
        hi[d] = int(std::floor((cmax[d] - dist_.min_coords[d]) / dist_.h));
        lo[d] = std::min(int(n[d]) - 1, std::max(0, lo[d]));
        hi[d] = std::min(int(n[d]) - 1, std::max(0, hi[d]));
      }
      int nh = 0, nc = 0, nd = 0;
      for (int v = lo[1]; v <= hi[1]; ++v)
        for (int u = lo[0]; u <= hi[0]; ++u) {
          const int z = zone_ls_[std::size_t(v) * Nx + std::size_t(u)];
          if (z == ZT_CONTINUUM) ++nc;
          else if (z == ZT_DISCRETE) ++nd;
          else if (z == ZT_HYBRID) ++nh;
        }
      const std::size_t f = std::size_t(j) * mpm_cells_[0] + i;
      const Zone zone{{cmin[0], cmin[1], cmax[0], cmax[1]}};
      auto hybrid = [&]() {
        zone_mpm_[f] = ZT_HYBRID;
        rzones_.push_back(zone);
        rzone_idx_.push_back({i, j});
      };
      if (nh > 0 || (nd > 0 && nc > 0)) {
        hybrid();
      } else if (nd > 0) {
        if (allow_direct_ || prev_zone_mpm_[f] == ZT_DISCRETE || prev_zone_mpm_[f] == ZT_HYBRID)
          zone_mpm_[f] = ZT_DISCRETE;
        else
          hybrid();
      } else if (nc > 0) {
        if (allow_direct_) {
          zone_mpm_[f] = ZT_CONTINUUM;
          if (prev_zone_mpm_[f] == ZT_DISCRETE) {
            direct_.push_back(zone);
            direct_idx_.push_back({i, j});
          }
        } else if (prev_zone_mpm_[f] == ZT_CONTINUUM || prev_zone_mpm_[f] == ZT_HYBRID) {
          zone_mpm_[f] = ZT_CONTINUUM;
        } else {
          hybrid();
        }
      }
    }
}

void ZoneTools::computeRZones(double L, double ht) {
  estimateLevelSetLevelZoneIndicators(L, ht);
  sampleMPMLevelZoneIndicators();
}

void ZoneTools::setIndicators(const std::vector<int> &cur, const std::vector<int> &prev) {
  if (cur.size() == zone_mpm_.size()) zone_mpm_ = cur;
  if (prev.size() == prev_zone_mpm_.size()) prev_zone_mpm_ = prev;
}

int ZoneTools::mpmCellOf(const double x[2]) const {
  int c[2];
  for (int d = 0; d < 2; ++d) c[d] = int(std::floor((x[d] - mpm_min_[d]) / mpm_h_));
  if (c[0] < 0 || c[1] < 0 || c[0] >= int(mpm_cells_[0]) || c[1] >= int(mpm_cells_[1])) return -1;
  return c[1] * int(mpm_cells_[0]) + c[0];
}

bool ZoneTools::isParticleInDiscreteZone_MPMIndicator(const double x[2]) const {
  const int f = mpmCellOf(x);
  return f >= 0 && zone_mpm_[std::size_t(f)] == ZT_DISCRETE;
}
bool ZoneTools::isParticleInContinuumZone_MPMIndicator(const double x[2]) const {
  const int f = mpmCellOf(x);
  return f >= 0 && zone_mpm_[std::size_t(f)] == ZT_CONTINUUM;
}
bool ZoneTools::isParticleInHybridZone_MPMIndicator(const double x[2]) const {
  const int f = mpmCellOf(x);
  return f >= 0 && zone_mpm_[std::size_t(f)] == ZT_HYBRID;
}

double ZoneTools::demWeight(const double x[2]) const {
  if (isParticleInHybridZone_MPMIndicator(x)) return 0.5;
  if (isParticleInDiscreteZone_MPMIndicator(x)) return 1.0;
  return 0.0;
}
double ZoneTools::mpmWeight(const double x[2]) const {
  if (isParticleInHybridZone_MPMIndicator(x)) return 0.5;
  if (isParticleInContinuumZone_MPMIndicator(x)) return 1.0;
  return 0.0;
}

void ZoneTools::determineMPMParticlesToDelete(const std::vector<double> &pts,
                                              std::vector<unsigned> &ids) const {
  ids.clear();
  for (std::size_t p = 0; p < pts.size() / 2; ++p) {
    const int f = mpmCellOf(&pts[2 * p]);
    if (f < 0 || zone_mpm_[std::size_t(f)] == ZT_DISCRETE) ids.push_back(unsigned(p));
  }
}

void ZoneTools::determineDEMGrainsToDelete(const std::vector<double> &q,
                                           const std::vector<std::uint8_t> &fixed,
                                           std::vector<unsigned> &ids) const {
  ids.clear();
  for (std::size_t b = 0; b < fixed.size(); ++b) {
    if (fixed[b]) continue;
    const int f = mpmCellOf(&q[3 * b]);
    if (f < 0 || zone_mpm_[std::size_t(f)] == ZT_CONTINUUM) ids.push_back(unsigned(b));
  }
}

void ZoneTools::identifyInnerContinuumRegion(std::vector<Zone> &zones,
                                             std::vector<CellIndex> &idx) const {
  zones.clear();
  idx.clear();
  for (std::uint32_t j = 0; j < mpm_cells_[1]; ++j)
    for (std::uint32_t i = 0; i < mpm_cells_[0]; ++i) {
      const std::size_t f = std::size_t(j) * mpm_cells_[0] + i;
      if (zone_mpm_[f] != ZT_CONTINUUM) continue;
      zones.push_back({{mpm_min_[0] + double(i) * mpm_h_, mpm_min_[1] + double(j) * mpm_h_,
                        mpm_min_[0] + double(i + 1) * mpm_h_, mpm_min_[1] + double(j + 1) * mpm_h_}});
      idx.push_back({i, j});
    }
}

void ZoneTools::identifyInnerContinuumRegionExcludingDirectTransitionZones(
    std::vector<Zone> &zones, std::vector<CellIndex> &idx) const {
  zones.clear();
  idx.clear();
  for (std::uint32_t j = 0; j < mpm_cells_[1]; ++j)
    for (std::uint32_t i = 0; i < mpm_cells_[0]; ++i) {
      const std::size_t f = std::size_t(j) * mpm_cells_[0] + i;
      if (zone_mpm_[f] != ZT_CONTINUUM || prev_zone_mpm_[f] == ZT_DISCRETE) continue;
      zones.push_back({{mpm_min_[0] + double(i) * mpm_h_, mpm_min_[1] + double(j) * mpm_h_,
                        mpm_min_[0] + double(i + 1) * mpm_h_, mpm_min_[1] + double(j + 1) * mpm_h_}});
      idx.push_back({i, j});
    }
}

// ---- filling zones with disks (AvoidAvoid.cpp:102-303) -----------------------------
// The rule the reference's avoid-a-void sampler follows, which fixes the
// results: zone by zone in order, 16 trials each; a trial draws the position
// (two randomMT() reals scaling the zone's extent) and then the radius, and
// keeps the disk unless it comes closer to an existing disk (old or kept
// before it) than 0.97 of their radii's sum.  The search for such a disk is
// this file's own: disks are binned by centre in singly linked lists over a
// grid of cells two mean radii wide, and a trial visits the cells within its
// radius plus the largest radius of the set searched.  A trial blocked by an
// old disk is blocked in any order, so all trials are first checked against
// the old disks in parallel; the few that pass are then checked in draw order
// against the disks kept before them, which gives the sequential rule's result.
namespace {
class DiskFiller {
 public:
  DiskFiller(const std::vector<double> &pos, const std::vector<double> &radii,
             const std::vector<Zone> &zones, double mean_radius)
      : pos_(pos), r_(radii) {
    double lo[2] = {INFINITY, INFINITY}, hi[2] = {-INFINITY, -INFINITY};
    auto grow = [&](double x, double y) {
      lo[0] = std::min(lo[0], x);
      lo[1] = std::min(lo[1], y);
      hi[0] = std::max(hi[0], x);
      hi[1] = std::max(hi[1], y);
    };
    for (const Zone &z : zones) {
      grow(z.v[0], z.v[1]);
      grow(z.v[2], z.v[3]);
    }
    for (std::size_t k = 0; k < r_.size(); ++k) {
      grow(pos_[2 * k], pos_[2 * k + 1]);
      rmax_old_ = std::max(rmax_old_, r_[k]);
    }
    w_ = 2.0 * mean_radius;
    if (!(w_ > 0.0) || !std::isfinite(w_)) w_ = 1.0;
    for (int d = 0; d < 2; ++d) {
      org_[d] = lo[d] - w_;
      // at most about 2^24 cells: a wider spread of disks gets wider cells
      n_[d] = std::uint32_t(std::min(4096.0, std::floor((hi[d] - lo[d]) / w_) + 3.0));
    }
    w_ = std::max({w_, (hi[0] - lo[0] + 2.0 * w_) / double(n_[0] - 1),
                   (hi[1] - lo[1] + 2.0 * w_) / double(n_[1] - 1)});
    head_old_.assign(std::size_t(n_[0]) * n_[1], -1);
    head_new_.assign(std::size_t(n_[0]) * n_[1], -1);
    next_.assign(r_.size(), -1);
    for (std::size_t k = 0; k < r_.size(); ++k) link(head_old_, int(k));
  }

  void fill(const std::vector<Zone> &zones, const std::function<double()> &radius,
            MersenneTwister1998 &rng) {
    const std::size_t nt = 16 * zones.size();
    std::vector<double> tx(nt), ty(nt), tr(nt);
    for (std::size_t z = 0; z < zones.size(); ++z) {
      const Zone &zn = zones[z];
      const double ex = zn.v[2] - zn.v[0], ey = zn.v[3] - zn.v[1];
      for (int trial = 0; trial < 16; ++trial) {
        const std::size_t t = 16 * z + std::size_t(trial);
        const double ux = double(rng.randomMT());
        const double uy = double(rng.randomMT());
        tx[t] = ux * ex + zn.v[0];
        ty[t] = uy * ey + zn.v[1];
        tr[t] = radius();
      }
    }
    std::vector<std::uint8_t> open(nt);
#pragma omp parallel for schedule(static, 1024)
    for (long t = 0; t < long(nt); ++t)
      open[std::size_t(t)] = blocked(head_old_, rmax_old_, tx[std::size_t(t)], ty[std::size_t(t)],
                                     tr[std::size_t(t)]) ? 0 : 1;
    for (std::size_t t = 0; t < nt; ++t) {
      if (!open[t] || blocked(head_new_, rmax_new_, tx[t], ty[t], tr[t])) continue;
      pos_.push_back(tx[t]);
      pos_.push_back(ty[t]);
      r_.push_back(tr[t]);
      next_.push_back(-1);
      rmax_new_ = std::max(rmax_new_, tr[t]);
      link(head_new_, int(r_.size()) - 1);
    }
  }
  const std::vector<double> &positions() const { return pos_; }
  const std::vector<double> &radii() const { return r_; }
  int count() const { return int(r_.size()); }

 private:
  int cell(double v, int d) const {
    const double c = std::floor((v - org_[d]) / w_);
    return int(std::max(0.0, std::min(double(n_[d]) - 1.0, c)));
  }
  void link(std::vector<int> &head, int k) {
    const std::size_t c = std::size_t(cell(pos_[2 * std::size_t(k) + 1], 1)) * n_[0] +
                          std::size_t(cell(pos_[2 * std::size_t(k)], 0));
    next_[std::size_t(k)] = head[c];
    head[c] = k;
  }
  bool blocked(const std::vector<int> &head, double rmax, double x, double y, double r) const {
    const double reach = r + rmax;
    const int i0 = cell(x - reach, 0), i1 = cell(x + reach, 0);
    const int j0 = cell(y - reach, 1), j1 = cell(y + reach, 1);
    for (int j = j0; j <= j1; ++j)
      for (int i = i0; i <= i1; ++i)
        for (int k = head[std::size_t(j) * n_[0] + std::size_t(i)]; k >= 0;
             k = next_[std::size_t(k)]) {
          const std::size_t K = std::size_t(k);
          if (std::hypot(pos_[2 * K] - x, pos_[2 * K + 1] - y) < (r_[K] + r) * 0.97) return true;
        }
    return false;
  }

  std::vector<double> pos_, r_;
  double rmax_old_ = 0.0, rmax_new_ = 0.0, w_ = 1.0, org_[2] = {0.0, 0.0};
  std::uint32_t n_[2] = {1, 1};
  std::vector<int> head_old_, head_new_, next_;
};
}  // namespace

// ---- bases (uGIMPLinearBasisFunctions.cpp, LinearBasisFunctions.cpp,
// CubicBasisFunctions.cpp) ------------------------------------------------------------
namespace {
double ugimp_integral(double xp, double hl, double xi, double w) {
  const double diff = std::fabs(xp - xi);
  if (diff >= w + hl) return 0.0;
  if (diff >= w - hl) return (w + hl - diff) * (w + hl - diff) / (2.0 * w);
  if (diff >= hl) return 2.0 * hl * (1.0 - diff / w);
  return 2.0 * hl - (hl * hl + diff * diff) / w;
}
double linear_n(double t) {
  const double a = std::fabs(t);
  return a < 1.0 ? 1.0 - a : 0.0;
}
double cubic_n(double t) {
  const double a = std::fabs(t);
  if (a < 1.0) return 0.5 * a * a * a - a * a + 2.0 / 3.0;
  if (a < 2.0) return -a * a * a / 6.0 + a * a - 2.0 * a + 4.0 / 3.0;
  return 0.0;
}
}  // namespace

void basisStencil(int basis, const double x[2], double hl, const double gmin[2], double h,
                  long lo[2], long hi[2]) {
  for (int d = 0; d < 2; ++d) {
    if (basis == HG_BASIS_UGIMP_LINEAR) {
      lo[d] = long(std::floor(((x[d] - hl) - gmin[d]) / h));
      hi[d] = long(std::ceil(((x[d] + hl) - gmin[d]) / h));
    } else if (basis == HG_BASIS_LINEAR) {
      lo[d] = long(std::floor((x[d] - gmin[d]) / h));
      hi[d] = lo[d] + 1;
    } else {
      lo[d] = long(std::floor((x[d] - gmin[d]) / h)) - 1;
      hi[d] = lo[d] + 3;
    }
  }
}

double basisWeight(int basis, const double x[2], double hl, std::uint32_t ix, std::uint32_t iy,
                   const double gmin[2], double h) {
  const double xi[2] = {h * double(ix) + gmin[0], h * double(iy) + gmin[1]};
  if (basis == HG_BASIS_UGIMP_LINEAR)
    return ugimp_integral(x[0], hl, xi[0], h) * ugimp_integral(x[1], hl, xi[1], h) /
           (4.0 * hl * hl);
  const double t0 = (x[0] - xi[0]) / h, t1 = (x[1] - xi[1]) / h;
  if (basis == HG_BASIS_LINEAR) return linear_n(t0) * linear_n(t1);
  return cubic_n(t0) * cubic_n(t1);
}

void accumulateHomogenizedStress(HybridResampleState &s) {
  const PhysicsGrid &g = s.continuum.physics_grid;
  const std::size_t nn = g.numGridPoints();
  s.homog_stress.resize(4 * nn, 0.0);
  if (s.c_f.size() != 2 * s.c_type.size() || s.c_p.size() != 2 * s.c_type.size()) return;
  const std::uint32_t W = g.cell_count[0] + 1;
  const double h = g.cell_width, hl = 0.25 * h, area = h * h;
  const std::vector<double> &q = s.discrete.q;
  for (std::size_t c = 0; c < s.c_type.size(); ++c) {
    if (s.c_type[c] != 1) continue;
    const std::size_t i = std::size_t(s.c_i[c]), j = std::size_t(s.c_j[c]);
    const double r[2] = {q[3 * i] - q[3 * j], q[3 * i + 1] - q[3 * j + 1]};
    const double f[2] = {s.c_f[2 * c], s.c_f[2 * c + 1]};
    // column-major -(f r^T + r f^T) / 2
    const double S[4] = {-0.5 * (f[0] * r[0] + r[0] * f[0]), -0.5 * (f[1] * r[0] + r[1] * f[0]),
                         -0.5 * (f[0] * r[1] + r[0] * f[1]), -0.5 * (f[1] * r[1] + r[1] * f[1])};
    const double *p = &s.c_p[2 * c];
    long lo[2], hi[2];
    basisStencil(s.continuum.basis, p, hl, g.min, h, lo, hi);
    for (long y = std::max(lo[1], 0L); y <= std::min(hi[1], long(g.cell_count[1])); ++y)
      for (long x = std::max(lo[0], 0L); x <= std::min(hi[0], long(g.cell_count[0])); ++x) {
        const double w = basisWeight(s.continuum.basis, p, hl, std::uint32_t(x), std::uint32_t(y), g.min, h);
        const std::size_t node = std::size_t(y) * W + std::size_t(x);
        for (int a = 0; a < 4; ++a) s.homog_stress[4 * node + std::size_t(a)] += w * S[a] / area;
      }
  }
}

void accumulateHomogenizedVelocity(HybridResampleState &s) {
  const PhysicsGrid &g = s.continuum.physics_grid;
  const std::size_t nn = g.numGridPoints();
  s.homog_velocity.resize(2 * nn, 0.0);
  std::vector<double> mom(2 * nn, 0.0), mass(nn, 0.0);
  const std::uint32_t W = g.cell_count[0] + 1;
  const double h = g.cell_width, hl = 0.25 * h;
  const RigidBody2DState &d = s.discrete;
  for (std::size_t b = 0; b < d.nbodies(); ++b) {
    if (d.fixed[b]) continue;
    const double m = d.M[3 * b];
    const double *x = &d.q[3 * b];
    long lo[2], hi[2];
    basisStencil(s.continuum.basis, x, hl, g.min, h, lo, hi);
    for (long y = std::max(lo[1], 0L); y <= std::min(hi[1], long(g.cell_count[1])); ++y)
      for (long xx = std::max(lo[0], 0L); xx <= std::min(hi[0], long(g.cell_count[0])); ++xx) {
        const double w = basisWeight(s.continuum.basis, x, hl, std::uint32_t(xx), std::uint32_t(y), g.min, h);
        const std::size_t node = std::size_t(y) * W + std::size_t(xx);
        mom[2 * node] += w * m * d.v[3 * b];
        mom[2 * node + 1] += w * m * d.v[3 * b + 1];
        mass[node] += w * m;
      }
  }
  for (std::size_t node = 0; node < nn; ++node)
    if (mass[node] > 0.0) {
      s.homog_velocity[2 * node] += mom[2 * node] / mass[node];
      s.homog_velocity[2 * node + 1] += mom[2 * node + 1] / mass[node];
    }
}

// ConstitutiveModel.cpp:99-117
void computeStrainFromCauchyStress(const double s[4], double kappa, double mu, double &J,
                                   double be[4]) {
  const double T = s[0] + s[3];
  J = 0.5 * (T / kappa + std::sqrt((T * T) / (kappa * kappa) + 4.0));
  const double d[4] = {J * (s[0] - 0.5 * T) / mu, J * s[1] / mu, J * s[2] / mu,
                       J * (s[3] - 0.5 * T) / mu};
  const double t = std::sqrt(1.0 - (d[0] * d[3] - d[2] * d[1]));
  be[0] = d[0] + t;
  be[1] = d[1];
  be[2] = d[2];
  be[3] = d[3] + t;
}

// ---- ResamplingTools.cpp ------------------------------------------------------------------
void ResamplingTools::init(double r_mean, double r_std) {
  r_mean_ = r_mean;
  r_std_ = r_std;
  dist_ = std::normal_distribution<double>(r_mean, r_std);
}

std::string ResamplingTools::saveRandomState() const {
  std::ostringstream o;
  for (int k = 0; k < 624; ++k) o << mt_.state()[k] << ' ';
  o << mt_.position() << ' ' << generator_ << ' ' << dist_;
  return o.str();
}

bool ResamplingTools::loadRandomState(const std::string &text) {
  std::istringstream in(text);
  std::uint32_t w[624];
  int pos = 0;
  for (int k = 0; k < 624; ++k) in >> w[k];
  // (the engine's extractor does not skip leading blanks)
  in >> pos >> std::ws >> generator_ >> std::ws >> dist_;
  if (!in) return false;
  mt_.setState(w, pos);
  return true;
}

double ResamplingTools::randomRadiusSampler() {
  double r = -1.0e33;
  long draws = 0;
  while (r < r_mean_ - 2.0 * r_std_ || r > r_mean_ + 2.0 * r_std_) {
    if (++draws > 1000000)
      throw std::runtime_error("grain radius sampler: no draw within mean +- 2 std");
    r = dist_(generator_);
  }
  return r;
}

namespace {

// The points (or grains) binned by continuum cell, as the reference's
// setupUniformGridElementsAsPoints registers them (ZoneTools.cpp:816-839): the
// Gaussian-weighted sums below visit the cells of a box row by row and each
// cell's items in index order, and their rounding follows that order.
std::unique_ptr<CellBins> point_bins(const PhysicsGrid &g, const std::vector<double> &xy,
                                     std::size_t stride) {
  std::unique_ptr<CellBins> b(new CellBins(g.min, g.cell_count, g.cell_width));
  b->build(xy.size() / stride, [&](std::size_t k, std::uint32_t lo[2], std::uint32_t hi[2]) {
    for (int d = 0; d < 2; ++d) lo[d] = hi[d] = b->cellOf(xy[stride * k + std::size_t(d)], d);
  });
  return b;
}

template <typename F>
void for_ids_in_box(const CellBins &b, const double x[2], double r, F f) {
  const std::uint32_t i0 = b.cellOf(x[0] - r, 0), i1 = b.cellOf(x[0] + r, 0);
  const std::uint32_t j0 = b.cellOf(x[1] - r, 1), j1 = b.cellOf(x[1] + r, 1);
  for (std::uint32_t j = j0; j <= j1; ++j)
    for (std::uint32_t i = i0; i <= i1; ++i)
      for (const int *id = b.begin(i, j); id != b.end(i, j); ++id) f(*id);
}

// Grid: mass-weighted nodal fields of the first n points, normalised by the
// nodal mass, with a default where the mass is zero (Grid.cpp:149-353).
// Stencil nodes past the grid (which the reference only asserts against) are
// skipped here and in the interpolations below.  Only the nodes that the
// points to be interpolated reach are computed: each of those sums its points
// in index order, as a pass over the whole grid would, and the other nodes
// are never read.  The fields (J, v and b_bar) share one pass.
struct NodalFields {
  std::vector<int> slot;  // node -> compact index, or -1
  std::vector<double> m, val;  // per compact node: mass, then `len` values
  int len = 0;
};

void node_range(const SimulationState &st, const double x[2], double hl, long lo[2], long hi[2]) {
  const PhysicsGrid &g = st.physics_grid;
  basisStencil(st.basis, x, hl, g.min, g.cell_width, lo, hi);
  lo[0] = std::max(lo[0], 0L);
  lo[1] = std::max(lo[1], 0L);
  hi[0] = std::min(hi[0], long(g.cell_count[0]));
  hi[1] = std::min(hi[1], long(g.cell_count[1]));
}

NodalFields nodal_fields(const SimulationState &st, std::size_t n, std::size_t q_begin,
                         std::size_t q_end, const std::vector<const std::vector<double> *> &values,
                         const std::vector<int> &widths, const std::vector<double> &defaults) {
  const PhysicsGrid &g = st.physics_grid;
  const std::uint32_t W = g.cell_count[0] + 1;
  const MaterialPoints &mp = st.material_points;
  const double h1 = mp.hl.empty() ? 0.0 : mp.hl[0];
  NodalFields f;
  for (int w : widths) f.len += w;
  f.slot.assign(g.numGridPoints(), -1);
  int nslot = 0;
  for (std::size_t p = q_begin; p < q_end; ++p) {
    long lo[2], hi[2];
    node_range(st, &mp.q[2 * p], h1, lo, hi);
    for (long y = lo[1]; y <= hi[1]; ++y)
      for (long x = lo[0]; x <= hi[0]; ++x) {
        int &sl = f.slot[std::size_t(y) * W + std::size_t(x)];
        if (sl < 0) sl = nslot++;
      }
  }
  f.m.assign(std::size_t(nslot), 0.0);
  f.val.assign(std::size_t(nslot) * std::size_t(f.len), 0.0);
  for (std::size_t p = 0; p < n; ++p) {
    long lo[2], hi[2];
    node_range(st, &mp.q[2 * p], h1, lo, hi);
    for (long y = lo[1]; y <= hi[1]; ++y)
      for (long x = lo[0]; x <= hi[0]; ++x) {
        const int sl = f.slot[std::size_t(y) * W + std::size_t(x)];
        if (sl < 0) continue;
        const double w = basisWeight(st.basis, &mp.q[2 * p], h1, std::uint32_t(x), std::uint32_t(y),
                                     g.min, g.cell_width);
        f.m[std::size_t(sl)] += w * mp.m[p];
        double *out = &f.val[std::size_t(sl) * std::size_t(f.len)];
        for (std::size_t v = 0; v < values.size(); ++v) {
          const int len = widths[v];
          for (int k = 0; k < len; ++k)
            out[k] += w * mp.m[p] * (*values[v])[p * std::size_t(len) + std::size_t(k)];
          out += len;
        }
      }
  }
  for (std::size_t sl = 0; sl < f.m.size(); ++sl)
    for (int k = 0; k < f.len; ++k) {
      double &v = f.val[sl * std::size_t(f.len) + std::size_t(k)];
      v = f.m[sl] > 0.0 ? v / f.m[sl] : defaults[std::size_t(k)];
    }
  return f;
}

// The fields at x: out gets f.len values.
void interpolate(const SimulationState &st, const NodalFields &f, const double x[2], double *out) {
  const PhysicsGrid &g = st.physics_grid;
  const std::uint32_t W = g.cell_count[0] + 1;
  const double h1 = st.material_points.hl[0];
  long lo[2], hi[2];
  node_range(st, x, h1, lo, hi);
  for (int k = 0; k < f.len; ++k) out[k] = 0.0;
  for (long y = lo[1]; y <= hi[1]; ++y)
    for (long xx = lo[0]; xx <= hi[0]; ++xx) {
      const double w = basisWeight(st.basis, x, h1, std::uint32_t(xx), std::uint32_t(y), g.min,
                                   g.cell_width);
      const int sl = f.slot[std::size_t(y) * W + std::size_t(xx)];
      for (int k = 0; k < f.len; ++k)
        out[k] += w * f.val[std::size_t(sl) * std::size_t(f.len) + std::size_t(k)];
    }
}

void add_material_point(HybridResampleState &s, const double x[2], double hl) {
  MaterialPoints &mp = s.continuum.material_points;
  mp.q.insert(mp.q.end(), {x[0], x[1]});
  mp.v.insert(mp.v.end(), {0.0, 0.0});
  mp.m.push_back(0.0);
  mp.volume.push_back(0.0);
  mp.hl.push_back(hl);
  mp.J.push_back(1.0);
  mp.be_bar.insert(mp.be_bar.end(), {1.0, 0.0, 0.0, 1.0});
  mp.kinematically_scripted.push_back(0);
  s.mpm_q0.insert(s.mpm_q0.end(), {x[0], x[1]});
  s.mpm_m0.push_back(0.0);
  s.mpm_volume0.push_back(0.0);
  s.mpm_weight.push_back(1.0);
}

}  // namespace

// :58-104
void ResamplingTools::avoidAVoidMPM(HybridResampleState &s, const std::vector<Zone> &zones,
                                    int &before, int &after) {
  const double r = 0.5 * s.initial_particle_size;
  MaterialPoints &mp = s.continuum.material_points;
  before = int(mp.npoints());
  std::vector<double> radii(mp.npoints(), r);
  DiskFiller filler(mp.q, radii, zones, r);
  filler.fill(zones, [r]() { return r; }, mt_);
  after = filler.count();
  const std::vector<double> &pos = filler.positions();
  for (int k = before; k < after; ++k)
    add_material_point(s, &pos[2 * std::size_t(k)], s.initial_particle_size * 0.5);
}

// :106-196 with RigidBody2DState::addBodies (RigidBody2DState.cpp:440-602)
void ResamplingTools::avoidAVoidDEM(HybridResampleState &s, const std::vector<Zone> &zones,
                                    double rho, int &before, int &after) {
  RigidBody2DState &d = s.discrete;
  before = int(d.nbodies());
  std::vector<double> pos(2 * d.nbodies());
  for (std::size_t b = 0; b < d.nbodies(); ++b) {
    pos[2 * b] = d.q[3 * b];
    pos[2 * b + 1] = d.q[3 * b + 1];
  }
  if (r_mean_ < 0.0)
    throw std::runtime_error("demRMean is not set: the grain radius sampler needs it");
  DiskFiller filler(pos, d.radius, zones, r_mean_);
  filler.fill(zones, [this]() { return randomRadiusSampler(); }, mt_);
  after = filler.count();
  const std::vector<double> &np = filler.positions();
  const std::vector<double> &nr = filler.radii();
  for (int k = before; k < after; ++k) {
    const double x = np[2 * std::size_t(k)], y = np[2 * std::size_t(k) + 1], r = nr[std::size_t(k)];
    const double m = rho * M_PI * r * r, I = 0.5 * m * r * r;
    d.q.insert(d.q.end(), {x, y, 0.0});
    d.v.insert(d.v.end(), {0.0, 0.0, 0.0});
    d.M.insert(d.M.end(), {m, m, I});
    d.radius.push_back(r);
    d.fixed.push_back(0);
    d.geometry_indices.push_back(std::uint32_t(d.geometry_radius.size()));
    d.geometry_radius.push_back(r);
    s.dem_q0.insert(s.dem_q0.end(), {x, y, 0.0});
    s.dem_M0.insert(s.dem_M0.end(), {m, m, I});
    s.dem_weight.push_back(1.0);
  }
}

ResampleReport ResamplingTools::resample(ZoneTools &zt, HybridResampleState &s,
                                         const PoormanEnrichmentSettings &p, double freq) {
  ResampleReport rep;
  SimulationState &cs = s.continuum;
  MaterialPoints &mp = cs.material_points;
  RigidBody2DState &ds = s.discrete;
  const PhysicsGrid &g = cs.physics_grid;

  // zones (HybridGrains2DSim.cpp:308-327)
  {
    std::vector<double> cpos, cr;
    for (std::size_t b = 0; b < ds.nbodies(); ++b) {
      if (ds.fixed[b]) continue;
      cpos.push_back(ds.q[3 * b]);
      cpos.push_back(ds.q[3 * b + 1]);
      cr.push_back(ds.radius[b]);
    }
    zt.updatePackingFraction(cpos, cr, mp.q, p.phi_window_size, p.phi_samples_per_cell_side);
  }
  zt.updateDistanceField(p.phi_threshold);
  zt.computeRZones(p.rzone_level_set, p.rzone_half_thickness);
  rep.rzones = zt.getRZones().size();
  rep.dem_before = int(ds.nbodies());
  rep.mpm_before = int(mp.npoints());

  // sampling (:331-388)
  int g_before = 0, g_after = 0;
  avoidAVoidDEM(s, zt.getRZones(), p.rho_dem, g_before, g_after);
  int p_before = 0, p_after = 0, p_mid_after = 0, p_mid_before = 0;
  if (zt.allowDirectTransitionsBetweenDiscreteAndContinuum()) {
    avoidAVoidMPM(s, zt.getRZones(), p_before, p_mid_after);
    avoidAVoidMPM(s, zt.getDirectTransitionZonesFromDiscreteToContinuum(), p_mid_before, p_after);
  } else {
    avoidAVoidMPM(s, zt.getRZones(), p_before, p_after);
  }
  int p2_before = 0, p2_after = 0;
  std::vector<Zone> inner;
  std::vector<CellIndex> inner_idx;
  if (zt.allowDirectTransitionsBetweenDiscreteAndContinuum())
    zt.identifyInnerContinuumRegionExcludingDirectTransitionZones(inner, inner_idx);
  else
    zt.identifyInnerContinuumRegion(inner, inner_idx);
  if (inner.empty()) {
    p2_before = p2_after = p_after;
  } else {
    avoidAVoidMPM(s, inner, p2_before, p2_after);
  }
  rep.dem_after_sampling = int(ds.nbodies());
  rep.mpm_after_sampling = int(mp.npoints());

  std::unique_ptr<CellBins> ug_c = point_bins(g, mp.q, 2);
  std::unique_ptr<CellBins> ug_d = point_bins(g, ds.q, 3);
  const double h = g.cell_width, cell_volume = h * h;
  const std::size_t nn = g.numGridPoints();
  const std::uint32_t W = g.cell_count[0] + 1;

  // the homogenised fields: every step's accumulation (the last one with the
  // state just sampled from), scaled by 1 / freq when grid smoothing is on
  // (:405-427)
  std::vector<double> hs = s.homog_stress, hv = s.homog_velocity;
  hs.resize(4 * nn, 0.0);
  hv.resize(2 * nn, 0.0);
  if (p.homogenize_stress && p.grid_smoothing_homogenized_stress)
    for (double &x : hs) x *= 1.0 / freq;
  if (p.homogenize_velocity && p.grid_smoothing_homogenized_velocity)
    for (double &x : hv) x *= 1.0 / freq;

  // determineMPMQuantitiesForHybridZones (ResamplingTools.cpp:616-645)
  auto hybrid_quantities = [&](const std::vector<CellIndex> &zidx, int before, int after) {
    const double mass_per_cell = s.material_density * cell_volume;
    for (const CellIndex &c : zidx) {
      const double n = double(ug_c->end(c.i, c.j) - ug_c->begin(c.i, c.j));
      for (const int *id = ug_c->begin(c.i, c.j); id != ug_c->end(c.i, c.j); ++id) {
        s.mpm_m0[std::size_t(*id)] = mass_per_cell / n;
        s.mpm_volume0[std::size_t(*id)] = cell_volume / n;
      }
    }
    const double sigma = 0.5 * s.initial_particle_size;
    const double max_dist = 3.0 * s.initial_particle_size;
    // each new point reads only the points and grains that were there before
    // and writes only its own fields
#pragma omp parallel for schedule(dynamic, 256)
    for (int pi = before; pi < after; ++pi) {
      const std::size_t P = std::size_t(pi);
      const double x[2] = {mp.q[2 * P], mp.q[2 * P + 1]};
      double v_sum[2] = {0, 0}, q0_sum[2] = {0, 0}, J_sum = 0, be_sum[4] = {0, 0, 0, 0};
      double wsum = 0;
      int nnb = 0;
      for_ids_in_box(*ug_c, x, max_dist, [&](int id) {
        if (id >= before) return;
        const std::size_t k = std::size_t(id);
        const double dist = std::hypot(mp.q[2 * k] - x[0], mp.q[2 * k + 1] - x[1]);
        if (dist > max_dist) return;
        const double w = std::exp(-(dist * dist) / (2.0 * sigma * sigma));
        for (int a = 0; a < 2; ++a) {
          v_sum[a] += mp.v[2 * k + std::size_t(a)] * w;
          q0_sum[a] += s.mpm_q0[2 * k + std::size_t(a)] * w;
        }
        J_sum += mp.J[k] * w;
        for (int a = 0; a < 4; ++a) be_sum[a] += mp.be_bar[4 * k + std::size_t(a)] * w;
        wsum += w;
        ++nnb;
      });
      double stress[4] = {0, 0, 0, 0}, vel[2] = {0, 0};
      if (p.homogenize_stress || p.homogenize_velocity) {
        const double hl = h * 0.25;
        long lo[2], hi[2];
        basisStencil(cs.basis, x, hl, g.min, h, lo, hi);
        for (long y = std::max(lo[1], 0L); y <= std::min(hi[1], long(g.cell_count[1])); ++y)
          for (long xx = std::max(lo[0], 0L); xx <= std::min(hi[0], long(g.cell_count[0])); ++xx) {
            const double w = basisWeight(cs.basis, x, hl, std::uint32_t(xx), std::uint32_t(y), g.min, h);
            const std::size_t node = std::size_t(y) * W + std::size_t(xx);
            for (int a = 0; a < 4; ++a) stress[a] += w * hs[4 * node + std::size_t(a)];
            for (int a = 0; a < 2; ++a) vel[a] += w * hv[2 * node + std::size_t(a)];
          }
      }
      double J = 1.0, be[4] = {1.0, 0.0, 0.0, 1.0};
      if (nnb > 0) {
        J = J_sum / wsum;
        for (int a = 0; a < 4; ++a) be[a] = be_sum[a] / wsum;
        const double sd = std::sqrt(be[0] * be[3] - be[2] * be[1]);
        for (double &b : be) b = b / sd;
      }
      for_ids_in_box(*ug_d, x, max_dist, [&](int id) {
        if (id >= g_before) return;
        const std::size_t k = std::size_t(id);
        const double dist = std::hypot(ds.q[3 * k] - x[0], ds.q[3 * k + 1] - x[1]);
        if (dist > max_dist) return;
        const double w = std::exp(-(dist * dist) / (2.0 * sigma * sigma));
        for (int a = 0; a < 2; ++a) {
          v_sum[a] += ds.v[3 * k + std::size_t(a)] * w;
          q0_sum[a] += s.dem_q0[3 * k + std::size_t(a)] * w;
        }
        wsum += w;
      });
      double v[2] = {0, 0}, q0[2] = {x[0], x[1]};
      if (wsum != 0.0) {
        for (int a = 0; a < 2; ++a) {
          v[a] = v_sum[a] / wsum;
          q0[a] = q0_sum[a] / wsum;
        }
      }
      s.mpm_q0[2 * P] = q0[0];
      s.mpm_q0[2 * P + 1] = q0[1];
      const double *vv = p.homogenize_velocity ? vel : v;
      mp.v[2 * P] = vv[0];
      mp.v[2 * P + 1] = vv[1];
      if (p.homogenize_stress) {
        computeStrainFromCauchyStress(stress, cs.bulk_modulus, cs.shear_modulus, J, be);
        mp.J[P] = J;
      } else {
        mp.J[P] = p.newly_converted_continuum_stress_free ? 1.0 : J;
      }
      for (int a = 0; a < 4; ++a) mp.be_bar[4 * P + std::size_t(a)] = be[a];
    }
  };
  if (zt.allowDirectTransitionsBetweenDiscreteAndContinuum()) {
    hybrid_quantities(zt.getRZoneIndices(), p_before, p_mid_after);
    hybrid_quantities(zt.getDirectTransitionZonesFromDiscreteToContinuumIndices(), p_mid_before,
                      p_after);
  } else {
    hybrid_quantities(zt.getRZoneIndices(), p_before, p_after);
  }

  // determineDEMQuantitiesForHybridZones (:647-734)
  {
    const double sigma = r_mean_, max_dist = 6.0 * r_mean_;
#pragma omp parallel for schedule(dynamic, 256)
    for (int b = g_before; b < g_after; ++b) {
      const std::size_t B = std::size_t(b);
      const double x[2] = {ds.q[3 * B], ds.q[3 * B + 1]};
      double v_sum[3] = {0, 0, 0}, q0_sum[3] = {0, 0, 0}, wsum = 0;
      for_ids_in_box(*ug_d, x, max_dist, [&](int id) {
        if (id >= g_before) return;
        const std::size_t k = std::size_t(id);
        const double dist = std::hypot(ds.q[3 * k] - x[0], ds.q[3 * k + 1] - x[1]);
        if (dist > max_dist) return;
        const double w = std::exp(-(dist * dist) / (2.0 * sigma * sigma));
        for (int a = 0; a < 3; ++a) {
          v_sum[a] += ds.v[3 * k + std::size_t(a)] * w;
          q0_sum[a] += s.dem_q0[3 * k + std::size_t(a)] * w;
        }
        wsum += w;
      });
      for_ids_in_box(*ug_c, x, max_dist, [&](int id) {
        if (id >= p_before) return;
        const std::size_t k = std::size_t(id);
        const double dist = std::hypot(mp.q[2 * k] - x[0], mp.q[2 * k + 1] - x[1]);
        if (dist > max_dist) return;
        const double w = std::exp(-(dist * dist) / (2.0 * sigma * sigma));
        for (int a = 0; a < 2; ++a) {
          v_sum[a] += mp.v[2 * k + std::size_t(a)] * w;
          q0_sum[a] += s.mpm_q0[2 * k + std::size_t(a)] * w;
        }
        wsum += w;
      });
      for (int a = 0; a < 3; ++a) {
        ds.v[3 * B + std::size_t(a)] = wsum == 0.0 ? 0.0 : v_sum[a] / wsum;
        s.dem_q0[3 * B + std::size_t(a)] =
            wsum == 0.0 ? (a < 2 ? x[a] : 0.0) : q0_sum[a] / wsum;
      }
    }
  }

  // determineMPMQuantitiesForInnerContinuumZones (:381-420)
  {
    const std::size_t n = std::size_t(p2_before);
    // J, v and b_bar of the first n points, at the nodes the new points reach
    const NodalFields fld = nodal_fields(cs, n, std::size_t(p2_before), std::size_t(p2_after),
                                         {&mp.J, &mp.v, &mp.be_bar}, {1, 2, 4},
                                         {1.0, 0.0, 0.0, 1.0, 0.0, 0.0, 1.0});
    // gatherMassAndVolume (:198-280)
    const double sigma = 0.5 * s.initial_particle_size, r = 2.0 * s.initial_particle_size;
    for (int pi = p2_before; pi < p2_after; ++pi) {
      const std::size_t P = std::size_t(pi);
      const double x[2] = {mp.q[2 * P], mp.q[2 * P + 1]};
      double tot_m = 0, tot_v = 0, wsum = 0, q0_sum[2] = {0, 0};
      std::vector<std::size_t> nb;
      for_ids_in_box(*ug_c, x, r, [&](int id) {
        if (id >= pi) return;
        const std::size_t k = std::size_t(id);
        const double dx = x[0] - mp.q[2 * k], dy = x[1] - mp.q[2 * k + 1];
        const double d2 = dx * dx + dy * dy;
        if (d2 >= r * r) return;
        const double w = std::exp(-(d2 * d2) / (2.0 * sigma * sigma));
        wsum += w;
        q0_sum[0] += w * s.mpm_q0[2 * k];
        q0_sum[1] += w * s.mpm_q0[2 * k + 1];
        tot_m += s.mpm_m0[k];
        tot_v += s.mpm_volume0[k];
        nb.push_back(k);
      });
      // with no neighbour the reference divides zero by zero here
      s.mpm_q0[2 * P] = q0_sum[0] / wsum;
      s.mpm_q0[2 * P + 1] = q0_sum[1] / wsum;
      const double c = double(nb.size());
      s.mpm_m0[P] = tot_m / (c + 1.0);
      s.mpm_volume0[P] = tot_v / (c + 1.0);
      const double f = c / (c + 1.0);
      for (std::size_t k : nb) {
        s.mpm_m0[k] *= f;
        s.mpm_volume0[k] *= f;
      }
    }
#pragma omp parallel for schedule(dynamic, 256)
    for (int pi = p2_before; pi < p2_after; ++pi) {
      const std::size_t P = std::size_t(pi);
      double f7[7];
      interpolate(cs, fld, &mp.q[2 * P], f7);
      const double J = f7[0];
      const double v[2] = {f7[1], f7[2]};
      double be[4] = {f7[3], f7[4], f7[5], f7[6]};
      const double sd = std::sqrt(be[0] * be[3] - be[2] * be[1]);
      for (double &b : be) b /= sd;
      mp.v[2 * P] = v[0];
      mp.v[2 * P + 1] = v[1];
      if (p.newly_converted_continuum_stress_free) {
        mp.J[P] = 1.0;
        const double I[4] = {1.0, 0.0, 0.0, 1.0};
        for (int a = 0; a < 4; ++a) mp.be_bar[4 * P + std::size_t(a)] = I[a];
      } else {
        mp.J[P] = J;
        for (int a = 0; a < 4; ++a) mp.be_bar[4 * P + std::size_t(a)] = be[a];
      }
    }
  }

  // deletions (ResamplingTools.cpp:27-56; MaterialPoints::deleteMaterialPoint,
  // MaterialPoints.cpp:469-520; RigidBody2DState::removeBodies with the
  // contact cache relabelled, RigidBody2DState.cpp:646-759,
  // PenaltyImpactFrictionMap.cpp:1199-1274)
  {
    std::vector<unsigned> del;
    zt.determineMPMParticlesToDelete(mp.q, del);
    std::vector<std::uint8_t> gone(mp.npoints(), 0);
    for (unsigned k : del) gone[k] = 1;
    std::size_t dst = 0;
    for (std::size_t k = 0; k < mp.npoints(); ++k) {
      if (gone[k]) continue;
      if (dst != k) {
        for (int a = 0; a < 2; ++a) {
          mp.q[2 * dst + std::size_t(a)] = mp.q[2 * k + std::size_t(a)];
          mp.v[2 * dst + std::size_t(a)] = mp.v[2 * k + std::size_t(a)];
          s.mpm_q0[2 * dst + std::size_t(a)] = s.mpm_q0[2 * k + std::size_t(a)];
        }
        for (int a = 0; a < 4; ++a) mp.be_bar[4 * dst + std::size_t(a)] = mp.be_bar[4 * k + std::size_t(a)];
        mp.m[dst] = mp.m[k];
        mp.volume[dst] = mp.volume[k];
        mp.hl[dst] = mp.hl[k];
        mp.J[dst] = mp.J[k];
        mp.kinematically_scripted[dst] = mp.kinematically_scripted[k];
        s.mpm_m0[dst] = s.mpm_m0[k];
        s.mpm_volume0[dst] = s.mpm_volume0[k];
        s.mpm_weight[dst] = s.mpm_weight[k];
      }
      ++dst;
    }
    mp.q.resize(2 * dst);
    mp.v.resize(2 * dst);
    s.mpm_q0.resize(2 * dst);
    mp.be_bar.resize(4 * dst);
    mp.m.resize(dst);
    mp.volume.resize(dst);
    mp.hl.resize(dst);
    mp.J.resize(dst);
    mp.kinematically_scripted.resize(dst);
    s.mpm_m0.resize(dst);
    s.mpm_volume0.resize(dst);
    s.mpm_weight.resize(dst);

    zt.determineDEMGrainsToDelete(ds.q, ds.fixed, del);
    const std::size_t nb = ds.nbodies();
    std::vector<int> map(nb, 0);
    std::vector<std::uint8_t> gd(nb, 0);
    for (unsigned k : del) gd[k] = 1;
    std::size_t to = 0;
    for (std::size_t k = 0; k < nb; ++k) {
      if (gd[k]) {
        map[k] = -1;
        continue;
      }
      map[k] = int(to);
      if (to != k) {
        for (int a = 0; a < 3; ++a) {
          ds.q[3 * to + std::size_t(a)] = ds.q[3 * k + std::size_t(a)];
          ds.v[3 * to + std::size_t(a)] = ds.v[3 * k + std::size_t(a)];
          ds.M[3 * to + std::size_t(a)] = ds.M[3 * k + std::size_t(a)];
          s.dem_q0[3 * to + std::size_t(a)] = s.dem_q0[3 * k + std::size_t(a)];
          s.dem_M0[3 * to + std::size_t(a)] = s.dem_M0[3 * k + std::size_t(a)];
        }
        ds.radius[to] = ds.radius[k];
        ds.fixed[to] = ds.fixed[k];
        ds.geometry_indices[to] = ds.geometry_indices[k];
        s.dem_weight[to] = s.dem_weight[k];
      }
      ++to;
    }
    ds.q.resize(3 * to);
    ds.v.resize(3 * to);
    ds.M.resize(3 * to);
    s.dem_q0.resize(3 * to);
    s.dem_M0.resize(3 * to);
    ds.radius.resize(to);
    ds.fixed.resize(to);
    ds.geometry_indices.resize(to);
    s.dem_weight.resize(to);
    std::vector<std::int32_t> t2, i2, j2;
    std::vector<double> d2, f2, p2;
    const bool with_f = s.c_f.size() == 2 * s.c_type.size();
    const bool with_p = s.c_p.size() == 2 * s.c_type.size();
    for (std::size_t c = 0; c < s.c_type.size(); ++c) {
      const int oi = map[std::size_t(s.c_i[c])];
      if (oi < 0) continue;
      int oj = s.c_j[c];
      if (s.c_type[c] != 0) {
        oj = map[std::size_t(s.c_j[c])];
        if (oj < 0) continue;
      }
      t2.push_back(s.c_type[c]);
      i2.push_back(oi);
      j2.push_back(oj);
      d2.push_back(s.c_ds[2 * c]);
      d2.push_back(s.c_ds[2 * c + 1]);
      if (with_f) {
        f2.push_back(s.c_f[2 * c]);
        f2.push_back(s.c_f[2 * c + 1]);
      }
      if (with_p) {
        p2.push_back(s.c_p[2 * c]);
        p2.push_back(s.c_p[2 * c + 1]);
      }
    }
    s.c_type.swap(t2);
    s.c_i.swap(i2);
    s.c_j.swap(j2);
    s.c_ds.swap(d2);
    s.c_f.swap(f2);
    s.c_p.swap(p2);
  }
  rep.dem_after_delete = int(ds.nbodies());
  rep.mpm_after_delete = int(mp.npoints());

  // mass weights (HybridGrains2DSim.cpp:514-531): M = M0 w, m = m0 w, V = V0 w
  for (std::size_t b = 0; b < ds.nbodies(); ++b) {
    const double w = zt.demWeight(&ds.q[3 * b]);
    s.dem_weight[b] = w;
    for (int a = 0; a < 3; ++a) ds.M[3 * b + std::size_t(a)] = s.dem_M0[3 * b + std::size_t(a)] * w;
  }
  for (std::size_t k = 0; k < mp.npoints(); ++k) {
    const double w = zt.mpmWeight(&mp.q[2 * k]);
    s.mpm_weight[k] = w;
    mp.m[k] = s.mpm_m0[k] * w;
    mp.volume[k] = s.mpm_volume0[k] * w;
  }
  // ResamplingTools::clearHomogenizedStress / Velocity (:736-742, :822-828)
  std::fill(s.homog_stress.begin(), s.homog_stress.end(), 0.0);
  std::fill(s.homog_velocity.begin(), s.homog_velocity.end(), 0.0);
  return rep;
}

}  // namespace hg
