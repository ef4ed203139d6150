/*
 * dem_oracle.c -- TEST INFRASTRUCTURE: serial FP64 CPU restatement of the
 * rigidbody2d circle step with PenaltyImpactFrictionMap
 * (rigidbody2d/PenaltyImpactFrictionMap.cpp:28-294, :296-599, :616-1081,
 * :1083-1149) and the active set it builds (RigidBody2DSim.cpp:270-413,
 * :777-842, :886-912, :1140-1178).  Checker and CPU baseline only.
 *
 * The contact set is the narrow phase's: with q0 == q1 the reference's CCD
 * test reduces to dx*dx + dy*dy - pow(ra + rb, 2) <= 0
 * (CollisionDetectionUtilities.cpp:8-13, :81-90).  The spatial grid only
 * produces candidates (a superset), so candidates are enumerated here from a
 * uniform cell grid of width 2*max radius; tests pin this against brute force.
 * Constraints are visited in the order of the reference's std::set of (i<j)
 * pairs, then planes plane-major, and cached per owner body in insertion
 * order; the force loops visit planes, then circle pairs, then kinematic
 * pairs, each by owner body.  Bandwidth reduction (Cuthill-McKee) is not
 * applied: it only permutes the summation order of the forces.
 *
 * Parity pins (tests/test_oracle_dem.py): brute-force pair sets, the
 * analytic linear spring-dashpot head-on collision (contact time and
 * restitution), a disk sliding on a plane (deceleration mu g) and sticking
 * below the friction threshold, the Coulomb cone, momentum conservation of
 * pair forces, and exact free fall.
 */
#include "hg_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct contact {
  int32_t type;
  int32_t i, j; /* owner body; partner body or plane index */
  double depth;
  double n[2];
  double ds[2];
  double vrel[2];
  double r0[2], r1[2];
  double w;
  double p[2]; /* circle pairs: the contact point (CircleCircleConstraint.cpp:214) */
  double f[2]; /* normal + friction force on body i (set by the force pass) */
} contact;

typedef struct hist {
  contact *c;
  uint32_t n, cap;
  double *F; /* last total force, 3B */
} hist;

static int cmp_contact(const void *a, const void *b) {
  const contact *x = (const contact *)a, *y = (const contact *)b;
  if (x->type != y->type) return x->type < y->type ? -1 : 1;
  if (x->i != y->i) return x->i < y->i ? -1 : 1;
  if (x->j != y->j) return x->j < y->j ? -1 : 1;
  return 0;
}

static const contact *find_old(const hist *h, int32_t type, int32_t i, int32_t j) {
  uint32_t lo = 0, hi = h->n;
  contact key;
  key.type = type;
  key.i = i;
  key.j = j;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) / 2;
    const int c = cmp_contact(&h->c[mid], &key);
    if (c == 0) return &h->c[mid];
    if (c < 0) lo = mid + 1; else hi = mid;
  }
  return NULL;
}

int oracle_dem_init(oracle_dem *s) {
  hist *h = (hist *)calloc(1, sizeof(hist));
  if (!h) return -1;
  h->F = (double *)calloc(3 * (size_t)s->nbodies + 3, sizeof(double));
  s->hist = h;
  return 0;
}

void oracle_dem_free(oracle_dem *s) {
  hist *h = (hist *)s->hist;
  if (!h) return;
  free(h->c);
  free(h->F);
  free(h);
  s->hist = NULL;
}

static int push(contact **arr, uint32_t *n, uint32_t *cap, const contact *c) {
  if (*n == *cap) {
    uint32_t nc = *cap ? 2 * *cap : 1024;
    contact *p = (contact *)realloc(*arr, nc * sizeof(contact));
    if (!p) return -1;
    *arr = p;
    *cap = nc;
  }
  (*arr)[(*n)++] = *c;
  return 0;
}

/* narrow phase predicate (CollisionDetectionUtilities.cpp:8-13, :81-90 with
 * q1 == q0: the quadratic and linear coefficients vanish) */
static int circles_touch(const double *q, const double *r, uint32_t a, uint32_t b) {
  const double dx = q[3 * a] - q[3 * b];
  const double dy = q[3 * a + 1] - q[3 * b + 1];
  const double c0 = (dx * dx + dy * dy) - pow(r[a] + r[b], 2);
  return c0 <= 0.0;
}

typedef struct pair { uint32_t i, j; } pair;

static int cmp_pair(const void *a, const void *b) {
  const pair *x = (const pair *)a, *y = (const pair *)b;
  if (x->i != y->i) return x->i < y->i ? -1 : 1;
  if (x->j != y->j) return x->j < y->j ? -1 : 1;
  return 0;
}

/* all (i<j) circle pairs the narrow phase accepts, in std::set order */
static pair *touching_pairs(const oracle_dem *s, uint32_t *npairs) {
  const uint32_t B = s->nbodies;
  *npairs = 0;
  if (B < 2) return NULL;
  double rmax = 0.0, xmin = INFINITY, ymin = INFINITY, xmax = -INFINITY, ymax = -INFINITY;
  for (uint32_t b = 0; b < B; ++b) {
    if (s->radius[b] > rmax) rmax = s->radius[b];
    const double x = s->q[3 * b], y = s->q[3 * b + 1];
    if (x < xmin) xmin = x;
    if (x > xmax) xmax = x;
    if (y < ymin) ymin = y;
    if (y > ymax) ymax = y;
  }
  const double cw = 2.0 * rmax * (1.0 + 1e-9) + 1e-300;
  long nx = (long)floor((xmax - xmin) / cw) + 1, ny = (long)floor((ymax - ymin) / cw) + 1;
  if (nx < 1) nx = 1;
  if (ny < 1) ny = 1;
  /* guard degenerate inputs (NaN, far-flung bodies): fall back to all pairs */
  const int brute = !(nx > 0 && ny > 0 && (double)nx * (double)ny < 4.0 * B + 1e7);
  pair *out = NULL;
  uint32_t n = 0, cap = 0;
  if (brute) {
    for (uint32_t a = 0; a < B; ++a)
      for (uint32_t b = a + 1; b < B; ++b)
        if (circles_touch(s->q, s->radius, a, b)) {
          if (n == cap) {
            cap = cap ? 2 * cap : 1024;
            out = (pair *)realloc(out, cap * sizeof(pair));
          }
          out[n].i = a;
          out[n].j = b;
          ++n;
        }
  } else {
    const size_t ncell = (size_t)nx * (size_t)ny;
    uint32_t *start = (uint32_t *)calloc(ncell + 1, sizeof(uint32_t));
    uint32_t *cell = (uint32_t *)malloc(B * sizeof(uint32_t));
    uint32_t *order = (uint32_t *)malloc(B * sizeof(uint32_t));
    for (uint32_t b = 0; b < B; ++b) {
      long cx = (long)floor((s->q[3 * b] - xmin) / cw), cy = (long)floor((s->q[3 * b + 1] - ymin) / cw);
      if (cx < 0) cx = 0;
      if (cx >= nx) cx = nx - 1;
      if (cy < 0) cy = 0;
      if (cy >= ny) cy = ny - 1;
      cell[b] = (uint32_t)(cy * nx + cx);
      start[cell[b] + 1]++;
    }
    for (size_t c = 0; c < ncell; ++c) start[c + 1] += start[c];
    uint32_t *fill = (uint32_t *)malloc((ncell + 1) * sizeof(uint32_t));
    memcpy(fill, start, (ncell + 1) * sizeof(uint32_t));
    for (uint32_t b = 0; b < B; ++b) order[fill[cell[b]]++] = b;
    free(fill);
    for (uint32_t a = 0; a < B; ++a) {
      const long cx = (long)(cell[a] % (uint32_t)nx), cy = (long)(cell[a] / (uint32_t)nx);
      for (long yy = cy - 1; yy <= cy + 1; ++yy)
        for (long xx = cx - 1; xx <= cx + 1; ++xx) {
          if (xx < 0 || yy < 0 || xx >= nx || yy >= ny) continue;
          const size_t c = (size_t)yy * (size_t)nx + (size_t)xx;
          for (uint32_t k = start[c]; k < start[c + 1]; ++k) {
            const uint32_t b = order[k];
            if (b <= a) continue;
            if (circles_touch(s->q, s->radius, a, b)) {
              if (n == cap) {
                cap = cap ? 2 * cap : 1024;
                out = (pair *)realloc(out, cap * sizeof(pair));
              }
              out[n].i = a;
              out[n].j = b;
              ++n;
            }
          }
        }
    }
    free(start);
    free(cell);
    free(order);
    qsort(out, n, sizeof(pair), cmp_pair);
  }
  *npairs = n;
  return out;
}

uint32_t oracle_dem_pairs_bruteforce(const oracle_dem *s, int32_t *pi, int32_t *pj, uint32_t max) {
  uint32_t n = 0;
  for (uint32_t a = 0; a < s->nbodies; ++a)
    for (uint32_t b = a + 1; b < s->nbodies; ++b) {
      if (s->fixed && s->fixed[a] && s->fixed[b]) continue;
      if (circles_touch(s->q, s->radius, a, b)) {
        if (n < max) {
          pi[n] = (int32_t)a;
          pj[n] = (int32_t)b;
        }
        ++n;
      }
    }
  return n;
}

static double norm2(const double *a) { return sqrt(a[0] * a[0] + a[1] * a[1]); }
static double dot2(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1]; }
static double cross2(const double *a, const double *b) { return a[0] * b[1] - a[1] * b[0]; }

/* secondRootOfQuadratic (PenaltyImpactFrictionMap.cpp:601-614) */
static double second_root(double a, double b, double c, double ds) {
  if (b > 0.0) return (2.0 * c) / (-b - ds);
  return (-b + ds) / (2.0 * a);
}

/* new delta_s: dt*vrel on a new contact, old + dt*vrel on a continuing one,
 * then the normal component projected out (:61-80 and siblings) */
static void spring(const contact *old, double dt, const double *vrel, const double *n,
                   double *ds) {
  if (!old) {
    ds[0] = dt * vrel[0];
    ds[1] = dt * vrel[1];
  } else {
    ds[0] = old->ds[0];
    ds[1] = old->ds[1];
    ds[0] += dt * vrel[0];
    ds[1] += dt * vrel[1];
  }
  const double d = dot2(n, ds);
  ds[0] -= d * n[0];
  ds[1] -= d * n[1];
}

/* One contact's penalty force with the Coulomb projection of the tangential
 * spring (PenaltyImpactFrictionMap.cpp:616-700, identical for every type). */
static void contact_force(const oracle_dem *s, contact *c, double total[2]) {
  const double w = c->w;
  const double *n = c->n;
  const double *vel = c->vrel;
  const double vdn = dot2(n, vel);
  const double vt[2] = {vel[0] - vdn * n[0], vel[1] - vdn * n[1]};
  const double vn[2] = {vel[0] - vt[0], vel[1] - vt[1]};
  const double cn = -s->kn * w * c->depth;
  const double dn = 0.5 * s->gamman * w;
  const double nf[2] = {cn * n[0] - dn * vn[0], cn * n[1] - dn * vn[1]};
  const double ck = -s->kt * w;
  const double dt_ = 0.5 * s->gammat * w;
  double ff[2] = {ck * c->ds[0] - dt_ * vt[0], ck * c->ds[1] - dt_ * vt[1]};
  const double mu_fn = s->mu * norm2(nf);
  const double ft = norm2(ff);
  if (0.5 * s->gammat * w * norm2(vt) > mu_fn) {
    c->ds[0] = 0.0;
    c->ds[1] = 0.0;
    const double g = -0.5 * s->gammat * w;
    ff[0] = g * vt[0];
    ff[1] = g * vt[1];
    const double z = ff[0] * ff[0] + ff[1] * ff[1];
    if (z > 0.0) {
      const double nz = sqrt(z);
      ff[0] = ff[0] / nz;
      ff[1] = ff[1] / nz;
    }
    ff[0] *= mu_fn;
    ff[1] *= mu_fn;
  } else if (ft > mu_fn) {
    const double *d = c->ds;
    const double a = s->kt * w * s->kt * w * (d[0] * d[0] + d[1] * d[1]);
    const double b = s->kt * w * s->gammat * w * dot2(d, vt);
    const double cc = 0.25 * s->gammat * w * s->gammat * w * dot2(vt, vt) -
                      s->mu * s->mu * (nf[0] * nf[0] + nf[1] * nf[1]);
    const double dscr = b * b - 4 * a * cc;
    const double root1 = second_root(a, b, cc, sqrt(dscr));
    const double nd[2] = {root1 * d[0], root1 * d[1]};
    ff[0] = ck * nd[0] - dt_ * vt[0];
    ff[1] = ck * nd[1] - dt_ * vt[1];
    c->ds[0] = nd[0];
    c->ds[1] = nd[1];
  }
  total[0] = nf[0] + ff[0];
  total[1] = nf[1] + ff[1];
}

static int dem_step(oracle_dem *s, double dt) {
  hist *h = (hist *)s->hist;
  const uint32_t B = s->nbodies;
  contact *nc = NULL;
  uint32_t nn = 0, ncap = 0;

  /* active set, then the cache update in active-set order */
  uint32_t np = 0;
  pair *pairs = touching_pairs(s, &np);
  for (uint32_t k = 0; k < np; ++k) {
    uint32_t a = pairs[k].i, b = pairs[k].j;
    const int fa = s->fixed && s->fixed[a], fb = s->fixed && s->fixed[b];
    if (fa && fb) continue;
    /* a kinematic body is always listed second (RigidBody2DSim.cpp:277-285) */
    const int kinematic = fa || fb;
    if (fa) {
      const uint32_t t = a;
      a = b;
      b = t;
    }
    const double *qa = &s->q[3 * a], *qb = &s->q[3 * b];
    const double ra = s->radius[a], rb = s->radius[b];
    contact c;
    memset(&c, 0, sizeof(c));
    c.i = (int32_t)a;
    c.j = (int32_t)b;
    /* n = normalized(q0a - q0b) (RigidBody2DSim.cpp:320) */
    double d[2] = {qa[0] - qb[0], qa[1] - qb[1]};
    const double dz = d[0] * d[0] + d[1] * d[1];
    if (dz > 0.0) {
      const double nz = sqrt(dz);
      c.n[0] = d[0] / nz;
      c.n[1] = d[1] / nz;
    } else {
      c.n[0] = d[0];
      c.n[1] = d[1];
    }
    const double gap = norm2(d) - ra - rb;
    c.depth = gap < 0.0 ? gap : 0.0;
    if (!kinematic) {
      /* CircleCircleConstraint (RigidBody2DSim.cpp:323; CircleCircleConstraint.cpp:171-233) */
      c.type = ORACLE_CONTACT_CIRCLE;
      const double f = ra / (ra + rb);
      const double p[2] = {qa[0] + f * (qb[0] - qa[0]), qa[1] + f * (qb[1] - qa[1])};
      c.r0[0] = p[0] - qa[0];
      c.r0[1] = p[1] - qa[1];
      c.r1[0] = p[0] - qb[0];
      c.r1[1] = p[1] - qb[1];
      c.p[0] = p[0];
      c.p[1] = p[1];
      const double t0[2] = {-c.r0[1], c.r0[0]};
      const double t1[2] = {-c.r1[1], c.r1[0]};
      const double *va = &s->v[3 * a], *vb = &s->v[3 * b];
      c.vrel[0] = va[0] + va[2] * t0[0] - vb[0] - vb[2] * t1[0];
      c.vrel[1] = va[1] + va[2] * t0[1] - vb[1] - vb[2] * t1[1];
      c.w = s->weights ? (s->weights[a] + s->weights[b]) * 0.5 : 1.0;
    } else {
      /* KinematicCircleCircleConstraint (KinematicObjectCircleConstraint.cpp:121-199) */
      c.type = ORACLE_CONTACT_KINEMATIC;
      c.r0[0] = -ra * c.n[0];
      c.r0[1] = -ra * c.n[1];
      const double t[2] = {-c.r0[1], c.r0[0]};
      const double cp[2] = {qa[0] - ra * c.n[0], qa[1] - ra * c.n[1]};
      const double arm[2] = {cp[0] - qb[0], cp[1] - qb[1]};
      const double tk[2] = {-arm[1], arm[0]};
      const double *va = &s->v[3 * a], *vb = &s->v[3 * b];
      const double kv[2] = {vb[0] + vb[2] * tk[0], vb[1] + vb[2] * tk[1]};
      c.vrel[0] = va[0] + va[2] * t[0] - kv[0];
      c.vrel[1] = va[1] + va[2] * t[1] - kv[1];
      c.w = s->weights ? s->weights[a] : 1.0;
    }
    spring(find_old(h, c.type, c.i, c.j), dt, c.vrel, c.n, c.ds);
    if (push(&nc, &nn, &ncap, &c)) return -1;
  }
  free(pairs);
  /* body-plane pairs, plane-major (RigidBody2DSim.cpp:777-842) */
  for (uint32_t k = 0; k < s->nplanes; ++k) {
    const double *pn = &s->plane_n[2 * k], *px = &s->plane_x[2 * k];
    for (uint32_t b = 0; b < B; ++b) {
      if (s->fixed && s->fixed[b]) continue;
      const double *qb = &s->q[3 * b];
      const double rel[2] = {qb[0] - px[0], qb[1] - px[1]};
      if (!(dot2(pn, rel) <= s->radius[b])) continue;
      contact c;
      memset(&c, 0, sizeof(c));
      c.type = ORACLE_CONTACT_PLANE;
      c.i = (int32_t)b;
      c.j = (int32_t)k;
      c.n[0] = pn[0];
      c.n[1] = pn[1];
      const double gap = dot2(pn, rel) - s->radius[b];
      c.depth = gap < 0.0 ? gap : 0.0;
      c.r0[0] = -s->radius[b] * pn[0];
      c.r0[1] = -s->radius[b] * pn[1];
      const double t0[2] = {-c.r0[1], c.r0[0]};
      const double *vb = &s->v[3 * b];
      /* the plane point's velocity is zero for a static plane */
      c.vrel[0] = vb[0] + vb[2] * t0[0] - 0.0;
      c.vrel[1] = vb[1] + vb[2] * t0[1] - 0.0;
      c.w = s->weights ? s->weights[b] : 1.0;
      spring(find_old(h, c.type, c.i, c.j), dt, c.vrel, c.n, c.ds);
      if (push(&nc, &nn, &ncap, &c)) return -1;
    }
  }
  /* the reference's per-owner insertion order is (type, owner, partner) */
  qsort(nc, nn, sizeof(contact), cmp_contact);
  free(h->c);
  h->c = nc;
  h->n = nn;
  h->cap = ncap;

  /* forces: gravity (NearEarthGravityForce.cpp:34-51), then contacts */
  double *F = h->F;
  memset(F, 0, 3 * (size_t)B * sizeof(double));
  for (uint32_t b = 0; b < B; ++b) {
    F[3 * b] += s->mass[3 * b] * s->gravity[0];
    F[3 * b + 1] += s->mass[3 * b] * s->gravity[1];
  }
  for (uint32_t k = 0; k < nn; ++k) {
    contact *c = &h->c[k];
    double tf[2];
    contact_force(s, c, tf);
    c->f[0] = tf[0];
    c->f[1] = tf[1];
    const uint32_t a = (uint32_t)c->i;
    F[3 * a] += tf[0];
    F[3 * a + 1] += tf[1];
    F[3 * a + 2] += cross2(c->r0, tf);
    if (c->type == ORACLE_CONTACT_CIRCLE) {
      const uint32_t b = (uint32_t)c->j;
      F[3 * b] -= tf[0];
      F[3 * b + 1] -= tf[1];
      F[3 * b + 2] -= cross2(c->r1, tf);
    }
  }
  /* PenaltyImpactFrictionMap.cpp:1107-1114 */
  for (uint32_t b = 0; b < B; ++b)
    if (s->fixed && s->fixed[b]) F[3 * b] = F[3 * b + 1] = F[3 * b + 2] = 0.0;
  for (uint32_t k = 0; k < 3 * B; ++k) {
    s->v[k] = s->v[k] + (dt * s->minv[k]) * F[k];
    s->q[k] = s->q[k] + dt * s->v[k];
  }
  return 0;
}

int oracle_dem_flow(oracle_dem *s, double dt, uint32_t nsteps) {
  if (!s->hist && oracle_dem_init(s)) return -1;
  for (uint32_t k = 0; k < nsteps; ++k)
    if (dem_step(s, dt)) return -1;
  return 0;
}

uint32_t oracle_dem_contacts(const oracle_dem *s, int32_t *type, int32_t *i, int32_t *j,
                             double *delta_s, uint32_t max) {
  const hist *h = (const hist *)s->hist;
  if (!h) return 0;
  for (uint32_t k = 0; k < h->n && k < max; ++k) {
    if (type) type[k] = h->c[k].type;
    if (i) i[k] = h->c[k].i;
    if (j) j[k] = h->c[k].j;
    if (delta_s) {
      delta_s[2 * k] = h->c[k].ds[0];
      delta_s[2 * k + 1] = h->c[k].ds[1];
    }
  }
  return h->n;
}

uint32_t oracle_dem_circle_contacts(const oracle_dem *s, int32_t *i, int32_t *j, double *p,
                                    double *f, uint32_t max) {
  const hist *h = (const hist *)s->hist;
  if (!h) return 0;
  uint32_t n = 0;
  for (uint32_t k = 0; k < h->n; ++k) {
    const contact *c = &h->c[k];
    if (c->type != ORACLE_CONTACT_CIRCLE) continue;
    if (n < max) {
      if (i) i[n] = c->i;
      if (j) j[n] = c->j;
      if (p) {
        p[2 * n] = c->p[0];
        p[2 * n + 1] = c->p[1];
      }
      if (f) {
        f[2 * n] = c->f[0];
        f[2 * n + 1] = c->f[1];
      }
    }
    ++n;
  }
  return n;
}

void oracle_dem_last_force(const oracle_dem *s, double *F) {
  const hist *h = (const hist *)s->hist;
  if (h) memcpy(F, h->F, 3 * (size_t)s->nbodies * sizeof(double));
}
