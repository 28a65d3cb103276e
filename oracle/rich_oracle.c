/*
 * rich_oracle.c -- CPU restatement of RICH's finite-volume step, the checker
 * of the GPU path (test infrastructure; see rich_oracle.h).
 *
 * Compiled with -ffp-contract=off.  Every expression keeps the evaluation
 * order of the reference routine cited above it (paths relative to the
 * OzW1234/huji-rich tree), including the order in which overloaded
 * Vector2D/ComputationalCell/Conserved operators combine their operands.
 */
#include "rich_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* State-dependent scalars are `real`: double here, a dual number when
 * oracle/tangent.cpp compiles this file as C++ for forward-mode derivatives
 * (RO_TANGENT).  Mesh geometry stays double in both builds. */
#ifndef RO_TANGENT
typedef double real;
#define RO_VAL(x) (x)
static double vabs_r(double x, double y) { return sqrt(x * x + y * y); }
#endif

#define RO_MAX_FACES 64

typedef struct { real d, p, vx, vy; } cc_t;          /* ComputationalCell */
typedef struct { real d, p, vx, vy, e, c; } prim_t;  /* Primitive         */
typedef struct { real m, mx, my, E; } cons_t;        /* Conserved         */

/* misc/utils.cpp:3-8 */
static int is_nan_ref(real x) {
  const int b1 = (x >= 0);
  const int b2 = (x < 0);
  return 1 != b1 + b2;
}
/* std::min / std::max: (b < a) ? b : a  and  (a < b) ? b : a */
static real smin(real a, real b) { return (b < a) ? b : a; }
static real smax(real a, real b) { return (a < b) ? b : a; }
/* tessellation/geometry.cpp:18-21  abs(Vector2D) */
static double vabs(double x, double y) { return sqrt(x * x + y * y); }

/* ----------------------------------------------------- fingerprint -- */
/* Optional record of every branch the step takes (shock switch, limiter
 * gates, guards and active faces, neighbour extrema, HLLC wave pattern and
 * side, CFL argmin), summed per step as 64-bit hashes: two runs took the
 * same branches in a step exactly when (barring hash collisions) their sums
 * agree.  Off unless ro_fingerprint() gave a buffer.  Test infrastructure:
 * it tells where central differences may straddle a discontinuity. */
static uint64_t* g_fp = NULL;
static int32_t g_fp_n = 0, g_fp_step = -1, g_fp_stage = 0;
static uint64_t mix64(uint64_t x) { /* splitmix64 finaliser */
  x ^= x >> 30;
  x *= 0xbf58476d1ce4e5b9ull;
  x ^= x >> 27;
  x *= 0x94d049bb133111ebull;
  return x ^ (x >> 31);
}
static uint64_t fp_fold(uint64_t h, uint64_t v) { return mix64(h ^ (v + 0x9e3779b97f4a7c15ull)); }
/* Kinds of branch.  Jumps (the value is discontinuous across the switch):
 * 6 the shock switch, 7 the limiters' gates and guards.  Kinks (continuous
 * value, derivative discontinuous): 2 HLLC's wave pattern and the sides
 * giving its outer wave speeds, 3 the CFL argmin, 4 and 5 the faces that
 * set psi, the clamp at 0 and the neighbours giving the extrema in
 * slope_limit and shocked_slope_limit. */
enum { RO_FP_KINDS = 8, RO_FP_THREADS = 256 };
/* per-thread partial sums of the current step, one 64-byte line each;
 * sums commute, so the record does not depend on the thread count */
static uint64_t g_fp_part[RO_FP_THREADS][RO_FP_KINDS];
static int fp_thread(void) {
#ifdef _OPENMP
  return omp_get_thread_num() % RO_FP_THREADS;
#else
  return 0;
#endif
}
static void fp_add(uint64_t kind, uint64_t item, uint64_t code) {
  if (!g_fp || g_fp_step < 0 || g_fp_step >= g_fp_n) return;
  const uint64_t h = fp_fold(fp_fold(fp_fold(kind, (uint64_t)g_fp_stage), item), code);
  /* atomic, though each thread has its own line (more than RO_FP_THREADS
   * threads would share rows) */
#pragma omp atomic
  g_fp_part[fp_thread()][kind] += h;
}
/* fold the partial sums into the step's record */
static void fp_flush(void) {
  if (!g_fp || g_fp_step < 0 || g_fp_step >= g_fp_n) return;
  for (int t = 0; t < RO_FP_THREADS; ++t)
    for (int k = 0; k < RO_FP_KINDS; ++k) {
      g_fp[(size_t)g_fp_step * RO_FP_KINDS + k] += g_fp_part[t][k];
      g_fp_part[t][k] = 0;
    }
}

/* ---------------------------------------------------------------- EOS -- */
/* newtonian/common/ideal_gas.cpp:13-16 */
static real dp2e(double g, real d, real p) { return p / d / (g - 1); }
/* ideal_gas.cpp:40-44 */
static real dp2c(double g, real d, real p) { return sqrt(g * p / d); }
/* ideal_gas.cpp:18-23 */
static int de2p(double g, real d, real e, real* p) {
  if (e < 0) return RO_NEGATIVE_THERMAL;
  *p = (g - 1) * e * d;
  return RO_OK;
}
#ifndef RO_TANGENT
double ro_dp2e(double g, double d, double p) { return dp2e(g, d, p); }
double ro_dp2c(double g, double d, double p) { return dp2c(g, d, p); }
int ro_de2p(double g, double d, double e, double* p) { return de2p(g, d, e, p); }
#endif

/* --------------------------------------------------------------- HLLC -- */
/* hydrodynamic_variables.cpp:261-265; pow(abs(v),2) is abs(v)*abs(v) */
static real total_energy_density(const prim_t* s) {
  const real a = vabs_r(s->vx, s->vy);
  return s->d * (0.5 * (a * a) + s->e);
}
/* hydrodynamic_variables.cpp:267-272 */
static cons_t primitive_to_conserved(const prim_t* s) {
  cons_t r;
  r.m = s->d;
  r.mx = s->vx * s->d;
  r.my = s->vy * s->d;
  r.E = total_energy_density(s);
  return r;
}
/* hydrodynamic_variables.cpp:281-290 with n = (1,0): n/abs(n) = (1,0)
 * exactly, ScalarProd(v, nn) = vx*1 + vy*0. */
static cons_t primitive_to_flux_x(const prim_t* s) {
  const real un = s->vx * 1.0 + s->vy * 0.0;
  const real dun = s->d * un;
  cons_t r;
  r.m = s->d * un;
  r.mx = 1.0 * s->p + s->vx * dun;
  r.my = 0.0 * s->p + s->vy * dun;
  r.E = (total_energy_density(s) + s->p) * un;
  return r;
}
/* hllc.cpp:27-43 */
static void estimate_wave_speeds(const prim_t* l, const prim_t* r, real* sl, real* ss,
                                 real* sr) {
  const real dl = l->d, pl = l->p, vl = l->vx, cl = l->c;
  const real dr = r->d, pr = r->p, vr = r->vx, cr = r->c;
  *sl = smin(vl - cl, vr - cr);
  *sr = smax(vl + cl, vr + cr);
  *ss = (pr - pl + dl * vl * (*sl - vl) - dr * vr * (*sr - vr)) /
        (dl * (*sl - vl) - dr * (*sr - vr));
}
/* hllc.cpp:76-92 */
static cons_t starred_state(const prim_t* s, real sk, real ss) {
  const real dk = s->d, pk = s->p, uk = s->vx, vk = s->vy;
  const real ds = dk * (sk - uk) / (sk - ss);
  const real ek = total_energy_density(s);
  cons_t r;
  r.m = ds;
  r.mx = ds * ss;
  r.my = ds * vk;
  r.E = ek * ds / dk + ds * (ss - uk) * (ss + pk / dk / (sk - uk));
  return r;
}
/* f + s*(us - u), hydrodynamic_variables.cpp:217-236 */
static cons_t add_scaled_diff(const cons_t* f, real s, const cons_t* us, const cons_t* u) {
  cons_t r;
  r.m = f->m + s * (us->m - u->m);
  r.mx = f->mx + (us->mx - u->mx) * s;
  r.my = f->my + (us->my - u->my) * s;
  r.E = f->E + s * (us->E - u->E);
  return r;
}
/* hllc.cpp:97-154 (massflux_ = true, the Hllc() default, hllc.hpp:22) */
static int hllc(const prim_t* left, const prim_t* right, double velocity, cons_t* out,
                int* branch) {
  if (is_nan_ref(right->vx)) return RO_HLLC_NAN;
  prim_t l = *left, r = *right;
  /* Velocity -= velocity*normaldir, normaldir = (1,0) */
  l.vx = l.vx - 1.0 * velocity;
  l.vy = l.vy - 0.0 * velocity;
  r.vx = r.vx - 1.0 * velocity;
  r.vy = r.vy - 0.0 * velocity;
  real sl, ss, sr;
  estimate_wave_speeds(&l, &r, &sl, &ss, &sr);
  const cons_t ul = primitive_to_conserved(&l);
  const cons_t ur = primitive_to_conserved(&r);
  const cons_t fl = primitive_to_flux_x(&l);
  const cons_t fr = primitive_to_flux_x(&r);
  const cons_t usl = starred_state(&l, sl, ss);
  const cons_t usr = starred_state(&r, sr, ss);
  cons_t f;
  int b;
  if (sl > 0)
    f = fl, b = 0;
  else if (sl <= 0 && ss >= 0)
    f = add_scaled_diff(&fl, sl, &usl, &ul), b = 1;
  else if (ss < 0 && sr >= 0)
    f = add_scaled_diff(&fr, sr, &usr, &ur), b = 2;
  else if (sr < 0)
    f = fr, b = 3;
  else
    return RO_INVALID_WAVE_SPEEDS;
  if (branch) /* the pattern, and which side gave each outer wave speed */
    *branch = b | ((r.vx - r.c < l.vx - l.c) << 2) | ((l.vx + l.c < r.vx + r.c) << 3);
  /* back to the lab frame */
  f.E = f.E + (f.mx * (1.0 * velocity) + f.my * (0.0 * velocity) + 0.5 * f.m * velocity * velocity);
  const real vm = velocity * f.m;
  f.mx = f.mx + 1.0 * vm;
  f.my = f.my + 0.0 * vm;
  *out = f;
  return RO_OK;
}

#ifndef RO_TANGENT
int ro_hllc(const double L[6], const double R[6], double velocity, double out[4]) {
  const prim_t l = {L[0], L[1], L[2], L[3], L[4], L[5]};
  const prim_t r = {R[0], R[1], R[2], R[3], R[4], R[5]};
  cons_t f;
  const int rc = hllc(&l, &r, velocity, &f, NULL);
  if (rc == RO_OK) {
    out[0] = f.m;
    out[1] = f.mx;
    out[2] = f.my;
    out[3] = f.E;
  }
  return rc;
}
#endif

/* simple_flux_calculator.cpp:9-15 */
static prim_t convert_to_primitive(const cc_t* c, double g) {
  prim_t r;
  r.d = c->d;
  r.p = c->p;
  r.vx = c->vx;
  r.vy = c->vy;
  r.e = dp2e(g, c->d, c->p);
  r.c = dp2c(g, c->d, c->p);
  return r;
}
/* simple_flux_calculator.cpp:130-140; Projection = ScalarProd/abs
 * (geometry.cpp:107-110) */
static prim_t rotate(const prim_t* s, double nx, double ny, double px, double py) {
  prim_t r = *s;
  r.vx = (s->vx * nx + s->vy * ny) / vabs(nx, ny);
  r.vy = (s->vx * px + s->vy * py) / vabs(px, py);
  return r;
}
/* simple_flux_calculator.cpp:142-151 */
static cons_t rotate_back(const cons_t* c, double nx, double ny, double px, double py) {
  const double an = vabs(nx, ny), ap = vabs(px, py);
  cons_t r = *c;
  r.mx = nx * c->mx / an + px * c->my / ap;
  r.my = ny * c->mx / an + py * c->my / ap;
  return r;
}
/* simple_flux_calculator.cpp:153-165 */
static int rotate_solve_rotate_back(const prim_t* l, const prim_t* r, double velocity, double nx,
                                    double ny, double px, double py, cons_t* out, int* branch) {
  const prim_t rl = rotate(l, nx, ny, px, py);
  const prim_t rr = rotate(r, nx, ny, px, py);
  cons_t f;
  const int rc = hllc(&rl, &rr, velocity, &f, branch);
  if (rc != RO_OK) return rc;
  *out = rotate_back(&f, nx, ny, px, py);
  return RO_OK;
}
/* simple_flux_calculator.cpp:27-31 */
static void remove_parallel_component(double vx, double vy, double px, double py, double* ox,
                                      double* oy) {
  const double s = vx * px + vy * py;
  const double pp = px * px + py * py;
  *ox = vx - px * s / pp;
  *oy = vy - py * s / pp;
}

/* ------------------------------------------------------- ghost cells -- */
static int32_t orig_index(const ro_mesh* m, int32_t id) {
  return id < m->n_cells ? id : m->ghost_orig[id - m->n_cells];
}
static cc_t get_cell(ro_cells c, int32_t i) {
  cc_t r = {c.d[i], c.p[i], c.vx[i], c.vy[i]};
  return r;
}
static void put_cell(ro_cells c, int32_t i, cc_t v) {
  c.d[i] = v.d;
  c.p[i] = v.p;
  c.vx[i] = v.vx;
  c.vy[i] = v.vy;
}

/* GhostPointGenerator::GetOuterEdgesIndeces (GhostPointGenerator.cpp:18-39),
 * RigidWallGenerator::operator() (RigidWallGenerator.cpp:5-37),
 * PeriodicGhostGenerator::operator() (PeriodicGhostGenerator.cpp:3-17).
 * The flat_map insert keeps the first value offered for each ghost. */
void ro_ghost_cells(const ro_mesh* m, const ro_params* prm, ro_cells c) {
  const int32_t N = m->n_cells, G = m->n_ghosts;
  char* filled = (char*)calloc((size_t)(G > 0 ? G : 1), 1);
  for (int32_t f = 0; f < m->n_faces; ++f) {
    const int32_t a = m->face_first[f], b = m->face_second[f];
    const int side = (a >= N) ? 1 : ((b >= N) ? 2 : 0);
    if (side == 0) continue;
    const int32_t g = side == 1 ? a : b;
    if (filled[g - N]) continue;
    filled[g - N] = 1;
    if (prm->bc == 0) {
      cc_t t = get_cell(c, side == 2 ? a : b);
      double nx, ny;
      if (side == 1) {
        nx = m->pt_x[a] - m->pt_x[b];
        ny = m->pt_y[a] - m->pt_y[b];
      } else {
        nx = m->pt_x[b] - m->pt_x[a];
        ny = m->pt_y[b] - m->pt_y[a];
      }
      const double an = vabs(nx, ny);
      nx = nx / an;
      ny = ny / an;
      const real s2 = 2 * (t.vx * nx + t.vy * ny);
      t.vx = t.vx - nx * s2;
      t.vy = t.vy - ny * s2;
      put_cell(c, g, t);
    } else {
      put_cell(c, g, get_cell(c, orig_index(m, g)));
    }
  }
  free(filled);
}

/* ---------------------------------------------- LinearGaussImproved -- */
typedef struct { cc_t x, y; } slope_t;

/* spatial_reconstruction.cpp:3-6  CalcCentroid */
static void face_centroid(const ro_mesh* m, int32_t f, double* cx, double* cy) {
  *cx = (m->v1x[f] + m->v2x[f]) * 0.5;
  *cy = (m->v1y[f] + m->v2y[f]) * 0.5;
}
/* computational_cell_2d.cpp:59-68  ComputationalCellAddMult */
static void add_mult(cc_t* r, const cc_t* o, double s) {
  r->d += o->d * s;
  r->p += o->p * s;
  r->vx += o->vx * s;
  r->vy += o->vy * s;
}
/* LinearGaussImproved.cpp:159-172  interp / interp2 */
static cc_t interp(const cc_t* cell, const slope_t* sl, double tx, double ty, double cmx,
                   double cmy) {
  cc_t r = *cell;
  add_mult(&r, &sl->x, tx - cmx);
  add_mult(&r, &sl->y, ty - cmy);
  return r;
}

/* LinearGaussImproved.cpp:67-134 */
static int calc_naive_slope(const ro_mesh* m, int32_t ci, const cc_t* cell, int n,
                            const int32_t* faces, const int32_t* nbr, const cc_t* ncell,
                            slope_t* res) {
  if (n > 20) return RO_TOO_MANY_NEIGHBORS;
  const double centx = m->pt_x[ci], centy = m->pt_y[ci];
  const double cmx = m->cm_x[ci], cmy = m->cm_y[ci];
  const double vol = m->tess_volume[ci];
  double M0 = 0, M1 = 0, M2 = 0, M3 = 0;
  slope_t vc;
  memset(&vc, 0, sizeof(vc));
  for (int i = 0; i < n; ++i) {
    const int32_t f = faces[i], j = nbr[i];
    double fx, fy;
    face_centroid(m, f, &fx, &fy);
    const double cijx = fx - (m->cm_x[j] + cmx) * 0.5;
    const double cijy = fy - (m->cm_y[j] + cmy) * 0.5;
    const double L = vabs(m->v2x[f] - m->v1x[f], m->v2y[f] - m->v1y[f]);
    const double dx = m->pt_x[j] - centx, dy = m->pt_y[j] - centy;
    const double ad = vabs(dx, dy);
    const double rx = dx / ad * L, ry = dy / ad * L;
    M0 -= cijx * rx;
    M1 -= cijy * rx;
    M2 -= cijx * ry;
    M3 -= cijy * ry;
    if (i == 0) {
      const double sx = rx * 0.5, sy = ry * 0.5;
      vc.x.d = cell->d * sx;
      vc.x.p = cell->p * sx;
      vc.x.vx = cell->vx * sx;
      vc.x.vy = cell->vy * sx;
      vc.y.d = cell->d * sy;
      vc.y.p = cell->p * sy;
      vc.y.vx = cell->vx * sy;
      vc.y.vy = cell->vy * sy;
    } else {
      add_mult(&vc.x, cell, rx * 0.5);
      add_mult(&vc.y, cell, ry * 0.5);
    }
    add_mult(&vc.y, &ncell[i], ry * 0.5);
    add_mult(&vc.x, &ncell[i], rx * 0.5);
  }
  M0 += vol;
  M3 += vol;
  const double det = M0 * M3 - M1 * M2;
  if (fabs(det) < 1e-10 * vol * vol) return RO_SINGULAR_MATRIX;
  const double det_inv = 1.0 / det;
  const double i0 = M3 * det_inv, i1 = -M1 * det_inv, i2 = -M2 * det_inv, i3 = M0 * det_inv;
  res->x.d = vc.x.d * i0;
  res->x.p = vc.x.p * i0;
  res->x.vx = vc.x.vx * i0;
  res->x.vy = vc.x.vy * i0;
  res->y.d = vc.y.d * i3;
  res->y.p = vc.y.p * i3;
  res->y.vx = vc.y.vx * i3;
  res->y.vy = vc.y.vy * i3;
  add_mult(&res->x, &vc.y, i1);
  add_mult(&res->y, &vc.x, i2);
  return RO_OK;
}

/* LinearGaussImproved.cpp:401-422 */
static void get_boundary_slope(const ro_mesh* m, int32_t ci, const cc_t* cell, int n,
                               const int32_t* nbr, const cc_t* ncell, slope_t* res) {
  const double cmx = m->cm_x[ci], cmy = m->cm_y[ci];
  cc_t sy_ = {0, 0, 0, 0}, sx_ = {0, 0, 0, 0};
  double SxSy = 0, Sy2 = 0, Sx2 = 0;
  for (int i = 0; i < n; ++i) {
    const int32_t j = nbr[i];
    const double ey = m->cm_y[j] - cmy, ex = m->cm_x[j] - cmx;
    const cc_t dd = {ncell[i].d - cell->d, ncell[i].p - cell->p, ncell[i].vx - cell->vx,
                     ncell[i].vy - cell->vy};
    sy_.d += dd.d * ey;
    sy_.p += dd.p * ey;
    sy_.vx += dd.vx * ey;
    sy_.vy += dd.vy * ey;
    sx_.d += dd.d * ex;
    sx_.p += dd.p * ex;
    sx_.vx += dd.vx * ex;
    sx_.vy += dd.vy * ex;
    SxSy += ey * ex;
    Sx2 += ex * ex;
    Sy2 += ey * ey;
  }
  const double den = SxSy * SxSy - Sx2 * Sy2;
  res->x.d = (sy_.d * SxSy - sx_.d * Sy2) / den;
  res->x.p = (sy_.p * SxSy - sx_.p * Sy2) / den;
  res->x.vx = (sy_.vx * SxSy - sx_.vx * Sy2) / den;
  res->x.vy = (sy_.vy * SxSy - sx_.vy * Sy2) / den;
  res->y.d = (sx_.d * SxSy - sy_.d * Sx2) / den;
  res->y.p = (sx_.p * SxSy - sy_.p * Sx2) / den;
  res->y.vx = (sx_.vx * SxSy - sy_.vx * Sx2) / den;
  res->y.vy = (sx_.vy * SxSy - sy_.vy * Sx2) / den;
}

/* LinearGaussImproved.cpp:136-148 */
static real pressure_ratio(const cc_t* cell, int n, const cc_t* ncell) {
  real res = 1;
  const real p = cell->p;
  for (int i = 0; i < n; ++i) {
    if (p > ncell[i].p)
      res = smin(res, ncell[i].p / p);
    else
      res = smin(res, p / ncell[i].p);
  }
  return res;
}

/* one field of slope_limit (LinearGaussImproved.cpp:232-267); absguard
 * selects the velocity form of the +-1e-9 guard */
static int limit_field(real dphi, real cval, real cell, real maxd, real mind,
                       int absguard, real* psi) {
  /* returned: the branches taken (fingerprint) */
  int code = 0;
  if (fabs(dphi) > 0.05 * smax(fabs(maxd), fabs(mind)) || cval * cell < 0) {
    const real hi = absguard ? fabs(1e-9 * cell) : 1e-9 * cell;
    const real lo = absguard ? -fabs(1e-9 * cell) : -1e-9 * cell;
    code = 1 | ((fabs(maxd) < fabs(mind)) << 4);
    if (dphi > hi) {
      code |= 2 | ((maxd / dphi < *psi) << 3);
      *psi = smin(*psi, maxd / dphi);
    } else if (dphi < lo) {
      code |= 4 | ((mind / dphi < *psi) << 3);
      *psi = smin(*psi, mind / dphi);
    }
  }
  return code;
}

/* LinearGaussImproved.cpp:174-300 */
static void slope_limit(const ro_mesh* m, int32_t ci, const cc_t* cell, int n,
                        const int32_t* faces, const int32_t* nbr, const cc_t* ncell,
                        slope_t* sl) {
  cc_t cmax = *cell, cmin = *cell;
  uint64_t h = 0, hj = 0; /* fingerprint: kinks (extrema, active faces), jumps */
  for (int i = 0; i < n; ++i) {
    const int32_t f = faces[i];
    if (orig_index(m, m->face_first[f]) == orig_index(m, m->face_second[f])) continue;
    const cc_t* t = &ncell[i];
    if (g_fp)
      h = fp_fold(h, (uint64_t)((cmax.d < t->d) | (cmax.p < t->p) << 1 | (cmax.vx < t->vx) << 2 |
                                (cmax.vy < t->vy) << 3 | (t->d < cmin.d) << 4 |
                                (t->p < cmin.p) << 5 | (t->vx < cmin.vx) << 6 |
                                (t->vy < cmin.vy) << 7 | i << 8));
    cmax.d = smax(cmax.d, t->d);
    cmax.p = smax(cmax.p, t->p);
    cmax.vx = smax(cmax.vx, t->vx);
    cmax.vy = smax(cmax.vy, t->vy);
    cmin.d = smin(cmin.d, t->d);
    cmin.p = smin(cmin.p, t->p);
    cmin.vx = smin(cmin.vx, t->vx);
    cmin.vy = smin(cmin.vy, t->vy);
  }
  (void)nbr;
  const cc_t maxd = {cmax.d - cell->d, cmax.p - cell->p, cmax.vx - cell->vx, cmax.vy - cell->vy};
  const cc_t mind = {cmin.d - cell->d, cmin.p - cell->p, cmin.vx - cell->vx, cmin.vy - cell->vy};
  const double cmx = m->cm_x[ci], cmy = m->cm_y[ci];
  real psi[4] = {1, 1, 1, 1};
  for (int i = 0; i < n; ++i) {
    double fx, fy;
    face_centroid(m, faces[i], &fx, &fy);
    const cc_t cv = interp(cell, sl, fx, fy, cmx, cmy);
    const cc_t dphi = {cv.d - cell->d, cv.p - cell->p, cv.vx - cell->vx, cv.vy - cell->vy};
    const int c0 = limit_field(dphi.d, cv.d, cell->d, maxd.d, mind.d, 0, &psi[0]);
    const int c1 = limit_field(dphi.p, cv.p, cell->p, maxd.p, mind.p, 0, &psi[1]);
    const int c2 = limit_field(dphi.vx, cv.vx, cell->vx, maxd.vx, mind.vx, 1, &psi[2]);
    const int c3 = limit_field(dphi.vy, cv.vy, cell->vy, maxd.vy, mind.vy, 1, &psi[3]);
    if (g_fp) { /* bits 0-2 gate and guard (jumps), 3-4 the rest (kinks) */
      hj = fp_fold(hj, (uint64_t)((c0 & 7) | (c1 & 7) << 3 | (c2 & 7) << 6 | (c3 & 7) << 9 | i << 12));
      h = fp_fold(h, (uint64_t)((c0 >> 3) | (c1 >> 3) << 2 | (c2 >> 3) << 4 | (c3 >> 3) << 6 | i << 8));
    }
  }
  fp_add(4, (uint64_t)ci, h);
  fp_add(7, (uint64_t)ci, hj);
  sl->x.d *= psi[0];
  sl->y.d *= psi[0];
  sl->x.p *= psi[1];
  sl->y.p *= psi[1];
  sl->x.vx *= psi[2];
  sl->y.vx *= psi[2];
  sl->x.vy *= psi[3];
  sl->y.vy *= psi[3];
}

/* one field of shocked_slope_limit (LinearGaussImproved.cpp:343-366); the
 * guard has no abs() for any field (:346,352,358,364) */
static int shocked_field(real dphi, real cval, real cell, real nval, real maxd,
                         real mind, double theta, real* psi) {
  int code = 0; /* the branches taken (fingerprint) */
  if (fabs(dphi) > 0.1 * smax(fabs(maxd), fabs(mind)) || cval * cell < 0) {
    code = 1 | ((fabs(maxd) < fabs(mind)) << 4);
    if (fabs(dphi) > 1e-9 * cell) {
      const real q = theta * (nval - cell) / dphi;
      code |= 2 | ((q < 0.0) << 2) | ((smax(q, 0.0) < *psi) << 3);
      *psi = smin(*psi, smax(theta * (nval - cell) / dphi, 0.0));
    }
  }
  return code;
}

/* LinearGaussImproved.cpp:302-399 */
static void shocked_slope_limit(const ro_mesh* m, int32_t ci, const cc_t* cell, int n,
                                const int32_t* faces, const cc_t* ncell, double theta,
                                slope_t* sl) {
  cc_t cmax = *cell, cmin = *cell;
  uint64_t h = 0, hj = 0; /* fingerprint: kinks (extrema, active faces), jumps */
  for (int i = 0; i < n; ++i) {
    const cc_t* t = &ncell[i];
    if (g_fp)
      h = fp_fold(h, (uint64_t)((cmax.d < t->d) | (cmax.p < t->p) << 1 | (cmax.vx < t->vx) << 2 |
                                (cmax.vy < t->vy) << 3 | (t->d < cmin.d) << 4 |
                                (t->p < cmin.p) << 5 | (t->vx < cmin.vx) << 6 |
                                (t->vy < cmin.vy) << 7 | i << 8));
    cmax.d = smax(cmax.d, t->d);
    cmax.p = smax(cmax.p, t->p);
    cmax.vx = smax(cmax.vx, t->vx);
    cmax.vy = smax(cmax.vy, t->vy);
    cmin.d = smin(cmin.d, t->d);
    cmin.p = smin(cmin.p, t->p);
    cmin.vx = smin(cmin.vx, t->vx);
    cmin.vy = smin(cmin.vy, t->vy);
  }
  const cc_t maxd = {cmax.d - cell->d, cmax.p - cell->p, cmax.vx - cell->vx, cmax.vy - cell->vy};
  const cc_t mind = {cmin.d - cell->d, cmin.p - cell->p, cmin.vx - cell->vx, cmin.vy - cell->vy};
  const double cmx = m->cm_x[ci], cmy = m->cm_y[ci];
  real psi[4] = {1, 1, 1, 1};
  for (int i = 0; i < n; ++i) {
    double fx, fy;
    face_centroid(m, faces[i], &fx, &fy);
    const cc_t cv = interp(cell, sl, fx, fy, cmx, cmy);
    const cc_t dphi = {cv.d - cell->d, cv.p - cell->p, cv.vx - cell->vx, cv.vy - cell->vy};
    const int c0 = shocked_field(dphi.d, cv.d, cell->d, ncell[i].d, maxd.d, mind.d, theta, &psi[0]);
    const int c1 = shocked_field(dphi.p, cv.p, cell->p, ncell[i].p, maxd.p, mind.p, theta, &psi[1]);
    const int c2 =
        shocked_field(dphi.vx, cv.vx, cell->vx, ncell[i].vx, maxd.vx, mind.vx, theta, &psi[2]);
    const int c3 =
        shocked_field(dphi.vy, cv.vy, cell->vy, ncell[i].vy, maxd.vy, mind.vy, theta, &psi[3]);
    if (g_fp) { /* bits 0-1 gate and guard (jumps), 2-4 the rest (kinks) */
      hj = fp_fold(hj, (uint64_t)((c0 & 3) | (c1 & 3) << 2 | (c2 & 3) << 4 | (c3 & 3) << 6 | i << 8));
      h = fp_fold(h, (uint64_t)((c0 >> 2) | (c1 >> 2) << 3 | (c2 >> 2) << 6 | (c3 >> 2) << 9 | i << 12));
    }
  }
  fp_add(5, (uint64_t)ci, h);
  fp_add(7, (uint64_t)ci, hj);
  sl->x.d *= psi[0];
  sl->y.d *= psi[0];
  sl->x.p *= psi[1];
  sl->y.p *= psi[1];
  sl->x.vx *= psi[2];
  sl->y.vx *= psi[2];
  sl->x.vy *= psi[3];
  sl->y.vy *= psi[3];
}

/* LinearGaussImproved.cpp:425-494  calc_slope */
static int calc_slope(const ro_mesh* m, const ro_params* prm, ro_cells cng, int32_t ci,
                      slope_t* res, slope_t* naive) {
  const int32_t b = m->csr_off[ci], e = m->csr_off[ci + 1];
  const int n = e - b;
  if (n > RO_MAX_FACES) return RO_TOO_MANY_NEIGHBORS;
  const int32_t* faces = m->csr_face + b;
  int32_t nbr[RO_MAX_FACES];
  cc_t ncell[RO_MAX_FACES];
  int boundary = 0;
  for (int i = 0; i < n; ++i) {
    const int32_t f = faces[i];
    const int32_t a = m->face_first[f], s = m->face_second[f];
    nbr[i] = (a == ci) ? s : a;
    ncell[i] = get_cell(cng, nbr[i]);
  }
  for (int i = 0; i < n; ++i) {
    const int32_t f = faces[i];
    if (orig_index(m, m->face_first[f]) == orig_index(m, m->face_second[f])) {
      boundary = 1;
      break;
    }
  }
  const cc_t cell = get_cell(cng, ci);
  if (boundary) {
    get_boundary_slope(m, ci, &cell, n, nbr, ncell, res);
  } else {
    const int rc = calc_naive_slope(m, ci, &cell, n, faces, nbr, ncell, res);
    if (rc != RO_OK) return rc;
  }
  if (naive) *naive = *res;
  if (prm->slf) {
    const real cs = dp2c(prm->gamma, cell.d, cell.p);
    const double width = sqrt(m->tess_volume[ci] / M_PI);
    /* is_shock, LinearGaussImproved.cpp:150-157 */
    const int cond1 = (res->x.vx + res->y.vy) * width < (-prm->delta_v) * cs;
    const int cond2 = pressure_ratio(&cell, n, ncell) < prm->delta_p;
    fp_add(6, (uint64_t)ci, (uint64_t)(cond1 | cond2 << 1)); /* the shock switch */
    if (!(cond1 || cond2))
      slope_limit(m, ci, &cell, n, faces, nbr, ncell, res);
    else
      shocked_slope_limit(m, ci, &cell, n, faces, ncell, prm->theta, res);
  }
  return RO_OK;
}

static void put_slope(real* const s[8], int32_t i, const slope_t* v) {
  s[0][i] = v->x.d;
  s[1][i] = v->x.p;
  s[2][i] = v->x.vx;
  s[3][i] = v->x.vy;
  s[4][i] = v->y.d;
  s[5][i] = v->y.p;
  s[6][i] = v->y.vx;
  s[7][i] = v->y.vy;
}
static slope_t get_slope(real* const s[8], int32_t i) {
  slope_t v;
  v.x.d = s[0][i];
  v.x.p = s[1][i];
  v.x.vx = s[2][i];
  v.x.vy = s[3][i];
  v.y.d = s[4][i];
  v.y.p = s[5][i];
  v.y.vx = s[6][i];
  v.y.vy = s[7][i];
  return v;
}

/* ghost gradient: RigidWallGenerator.cpp:39-54 (zero) or
 * PeriodicGhostGenerator.cpp:19-24 (slope of the original cell) */
static slope_t ghost_gradient(const ro_mesh* m, const ro_params* prm, real* const slopes[8],
                              int32_t g) {
  slope_t z;
  memset(&z, 0, sizeof(z));
  if (prm->bc == 0) return z;
  return get_slope(slopes, orig_index(m, g));
}

static ro_error err_make(int code, int32_t index) {
  ro_error e;
  e.code = code;
  e.index = index;
  return e;
}

/* LinearGaussImproved::operator() (LinearGaussImproved.cpp:536-635).
 * cng has N+G entries with the ghost slots filled by ro_ghost_cells. */
ro_error ro_lgi(const ro_mesh* m, const ro_params* prm, ro_cells cng, real* const slopes[8],
                real* const naive[8], ro_cells left, ro_cells right) {
  const int32_t N = m->n_cells;
  const int nt = prm->nthreads > 1 ? prm->nthreads : 1;
  ro_error err = err_make(RO_OK, -1);
  if (nt == 1) {
    int32_t* boundary = (int32_t*)malloc(sizeof(int32_t) * (size_t)(m->n_faces + 1));
    int32_t nb = 0;
    for (int32_t i = 0; i < N; ++i) {
      slope_t s, nv;
      const int rc = calc_slope(m, prm, cng, i, &s, &nv);
      if (rc != RO_OK) {
        free(boundary);
        return err_make(rc, i);
      }
      put_slope(slopes, i, &s);
      if (naive) put_slope(naive, i, &nv);
      const cc_t cell = get_cell(cng, i);
      for (int32_t k = m->csr_off[i]; k < m->csr_off[i + 1]; ++k) {
        const int32_t f = m->csr_face[k];
        double fx, fy;
        face_centroid(m, f, &fx, &fy);
        const cc_t st = interp(&cell, &s, fx, fy, m->cm_x[i], m->cm_y[i]);
        if (m->face_first[f] == i) {
          put_cell(left, f, st);
          if (m->face_second[f] >= N) boundary[nb++] = f;
        } else {
          put_cell(right, f, st);
          if (m->face_first[f] >= N) boundary[nb++] = f;
        }
      }
    }
    for (int32_t k = 0; k < nb; ++k) {
      const int32_t f = boundary[k];
      const int32_t a = m->face_first[f], b = m->face_second[f];
      const int32_t g = a >= N ? a : b;
      double fx, fy;
      face_centroid(m, f, &fx, &fy);
      const cc_t gc = get_cell(cng, g);
      const slope_t gg = ghost_gradient(m, prm, slopes, g);
      const cc_t st = interp(&gc, &gg, fx, fy, m->cm_x[g], m->cm_y[g]);
      put_cell(a >= N ? left : right, f, st);
    }
    free(boundary);
    return err;
  }
  /* OpenMP form: same per-cell and per-face arithmetic, independent items */
  int32_t first_bad = N;
  int bad_code = RO_OK;
#pragma omp parallel for num_threads(nt) schedule(static)
  for (int32_t i = 0; i < N; ++i) {
    slope_t s, nv;
    const int rc = calc_slope(m, prm, cng, i, &s, &nv);
    if (rc != RO_OK) {
#pragma omp critical
      {
        if (i < first_bad) {
          first_bad = i;
          bad_code = rc;
        }
      }
      continue;
    }
    put_slope(slopes, i, &s);
    if (naive) put_slope(naive, i, &nv);
  }
  if (bad_code != RO_OK) return err_make(bad_code, first_bad);
#pragma omp parallel for num_threads(nt) schedule(static)
  for (int32_t f = 0; f < m->n_faces; ++f) {
    double fx, fy;
    face_centroid(m, f, &fx, &fy);
    const int32_t sides[2] = {m->face_first[f], m->face_second[f]};
    for (int k = 0; k < 2; ++k) {
      const int32_t c = sides[k];
      const slope_t s = c < N ? get_slope(slopes, c) : ghost_gradient(m, prm, slopes, c);
      const cc_t cell = get_cell(cng, c);
      put_cell(k == 0 ? left : right, f, interp(&cell, &s, fx, fy, m->cm_x[c], m->cm_y[c]));
    }
  }
  return err;
}

/* PCM::operator() (interpolations/PCM.cpp:5-37): each side's state is the
 * cell itself, or for a ghost side the ghost generator's value. */
void ro_pcm(const ro_mesh* m, const ro_params* prm, ro_cells cng, ro_cells left, ro_cells right) {
  ro_ghost_cells(m, prm, cng);
  for (int32_t f = 0; f < m->n_faces; ++f) {
    put_cell(left, f, get_cell(cng, m->face_first[f]));
    put_cell(right, f, get_cell(cng, m->face_second[f]));
  }
}

/* ------------------------------------------------------ flux calculators */
static int modular_face_flux(const ro_mesh* m, const ro_params* prm, int32_t f, const cc_t* lf,
                             const cc_t* rt, double wx, double wy, cons_t* out, int* branch) {
  const int32_t N = m->n_cells;
  const int32_t a = m->face_first[f], b = m->face_second[f];
  /* calc_parallel_normal (modular_flux_calculator.cpp:12-22) */
  const double px = m->v2x[f] - m->v1x[f], py = m->v2y[f] - m->v1y[f];
  double nx, ny;
  if (a >= 0 && a < N)
    remove_parallel_component(m->v1x[f] - m->pt_x[a], m->v1y[f] - m->pt_y[a], px, py, &nx, &ny);
  else
    remove_parallel_component(m->pt_x[b] - m->v1x[f], m->pt_y[b] - m->v1y[f], px, py, &nx, &ny);
  /* modular_flux_calculator.cpp:56-67 */
  const double speed = (nx * wx + ny * wy) / vabs(nx, ny);
  const prim_t pl = convert_to_primitive(lf, prm->gamma);
  const prim_t pr = convert_to_primitive(rt, prm->gamma);
  return rotate_solve_rotate_back(&pl, &pr, speed, nx, ny, px, py, out, branch);
}

/* Reflect (geometry.cpp:117-120) applied by reflect()
 * (simple_flux_calculator.cpp:17-25) */
static prim_t reflect_prim(const prim_t* s, double ax, double ay) {
  prim_t r = *s;
  const real s2 = 2 * (s->vx * ax + s->vy * ay);
  const double a = vabs(ax, ay);
  const double a2 = a * a;
  r.vx = ax * s2 / a2 - s->vx;
  r.vy = ay * s2 / a2 - s->vy;
  return r;
}

/* SimpleFluxCalculator: riemann_reduce + rotate_solve_rotate_back
 * (simple_flux_calculator.cpp:80-126,167-189).  Where the reference would
 * read cells[ghost] (a non-mirror ghost face) the filled ghost slot is used. */
static int simple_face_flux(const ro_mesh* m, const ro_params* prm, ro_cells cng, int32_t f,
                            double wx, double wy, cons_t* out, int* branch) {
  const int32_t N = m->n_cells;
  const int32_t a = m->face_first[f], b = m->face_second[f];
  const double px = m->v2x[f] - m->v1x[f], py = m->v2y[f] - m->v1y[f];
  const double nx = m->pt_x[b] - m->pt_x[a], ny = m->pt_y[b] - m->pt_y[a];
  const int fa = a >= 0 && a < N, fb = b >= 0 && b < N;
  const int mirror = orig_index(m, a) == orig_index(m, b);
  prim_t l, r;
  if (!fa) {
    const cc_t cb = get_cell(cng, b);
    r = convert_to_primitive(&cb, prm->gamma);
    if (mirror)
      l = reflect_prim(&r, px, py);
    else {
      const cc_t ca = get_cell(cng, a);
      l = convert_to_primitive(&ca, prm->gamma);
    }
  } else if (!fb) {
    const cc_t ca = get_cell(cng, a);
    l = convert_to_primitive(&ca, prm->gamma);
    if (mirror)
      r = reflect_prim(&l, px, py);
    else {
      const cc_t cb = get_cell(cng, b);
      r = convert_to_primitive(&cb, prm->gamma);
    }
  } else {
    const cc_t ca = get_cell(cng, a), cb = get_cell(cng, b);
    l = convert_to_primitive(&ca, prm->gamma);
    r = convert_to_primitive(&cb, prm->gamma);
  }
  const double velocity = (nx * wx + ny * wy) / vabs(nx, ny);
  return rotate_solve_rotate_back(&l, &r, velocity, nx, ny, px, py, out, branch);
}

ro_error ro_flux(const ro_mesh* m, const ro_params* prm, ro_cells cng, const double* wx,
                 const double* wy, ro_ext flux) {
  const int32_t F = m->n_faces, N = m->n_cells, NG = m->n_cells + m->n_ghosts;
  const int nt = prm->nthreads > 1 ? prm->nthreads : 1;
  ro_ghost_cells(m, prm, cng);
  ro_cells left = {0, 0, 0, 0}, right = {0, 0, 0, 0};
  real* sbuf = NULL;
  real* lbuf = NULL;
  (void)N;
  (void)NG;
  if (prm->order == 2 || prm->order == 3) {
    lbuf = (real*)malloc(sizeof(real) * 8 * (size_t)F);
    left.d = lbuf;
    left.p = lbuf + F;
    left.vx = lbuf + 2 * (size_t)F;
    left.vy = lbuf + 3 * (size_t)F;
    right.d = lbuf + 4 * (size_t)F;
    right.p = lbuf + 5 * (size_t)F;
    right.vx = lbuf + 6 * (size_t)F;
    right.vy = lbuf + 7 * (size_t)F;
  }
  if (prm->order == 2) {
    sbuf = (real*)malloc(sizeof(real) * 8 * (size_t)N);
    real* slopes[8];
    for (int k = 0; k < 8; ++k) slopes[k] = sbuf + (size_t)k * N;
    const ro_error e = ro_lgi(m, prm, cng, slopes, NULL, left, right);
    if (e.code != RO_OK) {
      free(sbuf);
      free(lbuf);
      return e;
    }
  } else if (prm->order == 3) {
    /* ModularFluxCalculator over PCM (modular_flux_calculator.cpp:39-71) */
    for (int32_t f = 0; f < F; ++f) {
      put_cell(left, f, get_cell(cng, m->face_first[f]));
      put_cell(right, f, get_cell(cng, m->face_second[f]));
    }
  }
  int32_t first_bad = F;
  int bad_code = RO_OK;
#pragma omp parallel for num_threads(nt) schedule(static) if (nt > 1)
  for (int32_t f = 0; f < F; ++f) {
    if (nt == 1 && bad_code != RO_OK) continue;
    const double w_x = wx ? wx[f] : 0.0, w_y = wy ? wy[f] : 0.0;
    cons_t c;
    int rc, branch = 0;
    if (prm->order == 2 || prm->order == 3) {
      const cc_t lf = get_cell(left, f), rt = get_cell(right, f);
      rc = modular_face_flux(m, prm, f, &lf, &rt, w_x, w_y, &c, &branch);
    } else {
      rc = simple_face_flux(m, prm, cng, f, w_x, w_y, &c, &branch);
    }
    fp_add(2, (uint64_t)f, (uint64_t)branch);
    if (rc != RO_OK) {
#pragma omp critical
      {
        if (f < first_bad) {
          first_bad = f;
          bad_code = rc;
        }
      }
      continue;
    }
    flux.m[f] = c.m;
    flux.Mx[f] = c.mx;
    flux.My[f] = c.my;
    flux.E[f] = c.E;
  }
  free(sbuf);
  free(lbuf);
  return err_make(bad_code, bad_code == RO_OK ? -1 : first_bad);
}

/* ------------------------------------------------------------ updaters -- */
/* SimpleExtensiveUpdater::operator() (simple_extensive_updater.cpp:11-35);
 * areas are CacheData::areas = SlabSymmetry::calcArea
 * (physical_geometry.cpp:8-11).  The OpenMP form gathers each cell's
 * faces in ascending order, which applies the same operations in the same
 * order to every cell. */
void ro_update_extensives(const ro_mesh* m, const ro_params* prm, ro_ext flux, real dt,
                          ro_ext ext) {
  const int32_t N = m->n_cells;
  const int nt = prm->nthreads > 1 ? prm->nthreads : 1;
  if (nt == 1) {
    for (int32_t f = 0; f < m->n_faces; ++f) {
      const double area = vabs(m->v1x[f] - m->v2x[f], m->v1y[f] - m->v2y[f]);
      const real s = dt * area;
      const real dm = flux.m[f] * s, dE = flux.E[f] * s, dx = flux.Mx[f] * s,
                 dy = flux.My[f] * s;
      const int32_t a = m->face_first[f], b = m->face_second[f];
      if (a >= 0 && a < N) {
        ext.m[a] -= dm;
        ext.E[a] -= dE;
        ext.Mx[a] -= dx;
        ext.My[a] -= dy;
      }
      if (b >= 0 && b < N) {
        ext.m[b] += dm;
        ext.E[b] += dE;
        ext.Mx[b] += dx;
        ext.My[b] += dy;
      }
    }
    return;
  }
#pragma omp parallel for num_threads(nt) schedule(static)
  for (int32_t c = 0; c < N; ++c) {
    real em = ext.m[c], eE = ext.E[c], ex = ext.Mx[c], ey = ext.My[c];
    for (int32_t k = m->csr_off[c]; k < m->csr_off[c + 1]; ++k) {
      const int32_t f = m->csr_face[k];
      const double area = vabs(m->v1x[f] - m->v2x[f], m->v1y[f] - m->v2y[f]);
      const real s = dt * area;
      const real dm = flux.m[f] * s, dE = flux.E[f] * s, dx = flux.Mx[f] * s,
                 dy = flux.My[f] * s;
      if (m->face_first[f] == c) {
        em -= dm;
        eE -= dE;
        ex -= dx;
        ey -= dy;
      }
      if (m->face_second[f] == c) {
        em += dm;
        eE += dE;
        ex += dx;
        ey += dy;
      }
    }
    ext.m[c] = em;
    ext.E[c] = eE;
    ext.Mx[c] = ex;
    ext.My[c] = ey;
  }
}

/* regular_update (simple_cell_updater.cpp:22-59) */
static int cell_update_one(const ro_mesh* m, double g, ro_ext ext, ro_cells cells, int32_t i) {
  const double volume = m->volume[i];
  const real mass = ext.m[i];
  const real d = mass / volume;
  if (d < 0) return RO_NEGATIVE_DENSITY;
  const real vx = ext.Mx[i] / mass, vy = ext.My[i] / mass;
  const real energy = ext.E[i] / mass - 0.5 * (vx * vx + vy * vy);
  real p;
  const int rc = de2p(g, d, energy, &p);
  if (rc != RO_OK) return rc;
  cells.d[i] = d;
  cells.p[i] = p;
  cells.vx[i] = vx;
  cells.vy[i] = vy;
  return RO_OK;
}

ro_error ro_update_cells(const ro_mesh* m, const ro_params* prm, ro_ext ext, ro_cells cells) {
  const int32_t N = m->n_cells;
  const int nt = prm->nthreads > 1 ? prm->nthreads : 1;
  if (nt == 1) {
    for (int32_t i = 0; i < N; ++i) {
      const int rc = cell_update_one(m, prm->gamma, ext, cells, i);
      if (rc != RO_OK) return err_make(rc, i);
    }
    return err_make(RO_OK, -1);
  }
  int32_t first_bad = N;
  int bad_code = RO_OK;
#pragma omp parallel for num_threads(nt) schedule(static)
  for (int32_t i = 0; i < N; ++i) {
    const int rc = cell_update_one(m, prm->gamma, ext, cells, i);
    if (rc != RO_OK) {
#pragma omp critical
      {
        if (i < first_bad) {
          first_bad = i;
          bad_code = rc;
        }
      }
    }
  }
  return err_make(bad_code, bad_code == RO_OK ? -1 : first_bad);
}

/* TimeStepCalculator::operator[] (simple_cfl.cpp:28-47); GetWidth is
 * sqrt(GetVolume/pi) (VoronoiMesh.cpp:410-413) */
static real cfl_cell(const ro_mesh* m, double g, ro_cells cells, const double* wx,
                     const double* wy, int32_t i) {
  real res = 0;
  const double radius = sqrt(m->tess_volume[i] / M_PI);
  const real c = dp2c(g, cells.d[i], cells.p[i]);
  const real vx = cells.vx[i], vy = cells.vy[i];
  for (int32_t k = m->csr_off[i]; k < m->csr_off[i + 1]; ++k) {
    const int32_t f = m->csr_face[k];
    const double ex = wx ? wx[f] : 0.0, ey = wy ? wy[f] : 0.0;
    res = fmax(res, (c + vabs_r(vx - ex, vy - ey)) / radius);
  }
  return 1.0 / res;
}

/* SimpleCFL::operator() (simple_cfl.cpp:58-71) with lazy_min
 * (misc/lazy_list.hpp:185-191) */
real ro_cfl(const ro_mesh* m, const ro_params* prm, ro_cells cells, const double* wx,
            const double* wy) {
  const int32_t N = m->n_cells;
  const int nt = prm->nthreads > 1 ? prm->nthreads : 1;
  real res = cfl_cell(m, prm->gamma, cells, wx, wy, 0);
  if (g_fp) { /* the cell that sets dt (the first of equal minima) */
    int32_t arg = 0;
    real best = res;
    for (int32_t i = 1; i < N; ++i) {
      const real v = cfl_cell(m, prm->gamma, cells, wx, wy, i);
      if (v < best) best = v, arg = i;
    }
    fp_add(3, (uint64_t)arg, 0);
  }
  if (nt == 1) {
    for (int32_t i = 1; i < N; ++i) res = smin(res, cfl_cell(m, prm->gamma, cells, wx, wy, i));
  } else {
#pragma omp parallel num_threads(nt)
    {
      real loc = res;
#pragma omp for schedule(static)
      for (int32_t i = 1; i < N; ++i) loc = smin(loc, cfl_cell(m, prm->gamma, cells, wx, wy, i));
#pragma omp critical
      res = smin(res, loc);
    }
  }
  return prm->cfl * res;
}

/* init_extensives (hdsim2d.cpp:41-66) with the CacheData volume */
void ro_init_extensives(const ro_mesh* m, const ro_params* prm, ro_cells cells, ro_ext ext) {
  for (int32_t i = 0; i < m->n_cells; ++i) {
    const double volume = m->volume[i];
    const real mass = volume * cells.d[i];
    const real vx = cells.vx[i], vy = cells.vy[i];
    ext.m[i] = mass;
    ext.E[i] = dp2e(prm->gamma, cells.d[i], cells.p[i]) * mass +
               0.5 * mass * (vx * vx + vy * vy);
    ext.Mx[i] = vx * mass;
    ext.My[i] = vy * mass;
  }
}

/* ExternalForceContribution with ZeroForce (hydrodynamics_2d.cpp:348-372,
 * zero_force.cpp:3-18): adds dt*0 to every extensive */
static void zero_force(int32_t N, real dt, ro_ext ext) {
  const real z = dt * 0.0;
  for (int32_t i = 0; i < N; ++i) {
    ext.m[i] += z;
    ext.Mx[i] += 0.0 * dt;
    ext.My[i] += 0.0 * dt;
    ext.E[i] += z;
  }
}

static real* alloc_d(size_t n) { return (real*)malloc(sizeof(real) * (n ? n : 1)); }

/* hdsim::TimeAdvance (hdsim2d.cpp:149-241) and hdsim::TimeAdvance2Heun
 * (hdsim2d.cpp:311-449) on a static mesh (no rebuild, no Hilbert
 * reordering; point velocities zero, face velocities as given). */
ro_error ro_run(const ro_mesh* m, const ro_params* prm, int32_t integrator, int32_t nsteps,
                const double* wx, const double* wy, ro_cells cng, ro_ext ext, double* dt_out) {
  const int32_t N = m->n_cells, F = m->n_faces, NG = m->n_cells + m->n_ghosts;
  real* fb = alloc_d(4 * (size_t)F);
  ro_ext flux = {fb, fb + F, fb + 2 * (size_t)F, fb + 3 * (size_t)F};
  real* mb = alloc_d(4 * (size_t)N);
  ro_ext mid = {mb, mb + N, mb + 2 * (size_t)N, mb + 3 * (size_t)N};
  real* cb = alloc_d(4 * (size_t)NG);
  ro_cells midc = {cb, cb + NG, cb + 2 * (size_t)NG, cb + 3 * (size_t)NG};
  ro_error err = err_make(RO_OK, -1);
  for (int32_t step = 0; step < nsteps && err.code == RO_OK; ++step) {
    fp_flush();
    g_fp_step = step;
    g_fp_stage = 0;
    const real dt = ro_cfl(m, prm, cng, wx, wy);
    if (dt_out) dt_out[step] = RO_VAL(dt);
    if (integrator == 1) {
      err = ro_flux(m, prm, cng, wx, wy, flux);
      if (err.code != RO_OK) break;
      ro_update_extensives(m, prm, flux, dt, ext);
      zero_force(N, dt, ext);
      err = ro_update_cells(m, prm, ext, cng);
    } else {
      err = ro_flux(m, prm, cng, wx, wy, flux);
      if (err.code != RO_OK) break;
      memcpy(mid.m, ext.m, sizeof(real) * (size_t)N);
      memcpy(mid.E, ext.E, sizeof(real) * (size_t)N);
      memcpy(mid.Mx, ext.Mx, sizeof(real) * (size_t)N);
      memcpy(mid.My, ext.My, sizeof(real) * (size_t)N);
      zero_force(N, dt, mid);
      ro_update_extensives(m, prm, flux, dt, mid);
      err = ro_update_cells(m, prm, mid, midc);
      if (err.code != RO_OK) break;
      g_fp_stage = 1;
      err = ro_flux(m, prm, midc, wx, wy, flux);
      if (err.code != RO_OK) break;
      zero_force(N, dt, mid);
      ro_update_extensives(m, prm, flux, dt, mid);
      /* average_extensive (hdsim2d.cpp:311-320): 0.5*(mid + old) */
      for (int32_t i = 0; i < N; ++i) {
        ext.m[i] = 0.5 * (mid.m[i] + ext.m[i]);
        ext.E[i] = 0.5 * (mid.E[i] + ext.E[i]);
        ext.Mx[i] = (mid.Mx[i] + ext.Mx[i]) * 0.5;
        ext.My[i] = (mid.My[i] + ext.My[i]) * 0.5;
      }
      err = ro_update_cells(m, prm, ext, cng);
    }
  }
  fp_flush();
  g_fp_step = -1;
  free(fb);
  free(mb);
  free(cb);
  return err;
}

#ifndef RO_TANGENT
/* Record ro_run's branches into per_step[n][8] (zeroed here); NULL stops. */
void ro_fingerprint(uint64_t* per_step, int32_t n) {
  g_fp = per_step;
  g_fp_n = per_step ? n : 0;
  g_fp_step = -1;
  for (int32_t k = 0; k < g_fp_n * RO_FP_KINDS; ++k) per_step[k] = 0;
  memset(g_fp_part, 0, sizeof(g_fp_part));
}
#endif

#ifndef RO_TANGENT
/* ------------------------------------------------ moving-mesh inputs -- */
/* Lagrangian::operator() (point_motions/lagrangian.cpp:3-13) with
 * RoundCells::operator() adding calc_dw (round_cells.cpp:63-76, 107-120;
 * cold = false): r the mesh point, s the CM, d = abs(s - r),
 * R = GetWidth; zero if d < 0.9 eta R, else
 * chi * max(dp2c, abs(v)) * (s - r) / R. */
void ro_point_velocity(const ro_mesh* m, double gamma, const double* d, const double* p,
                       const double* vx, const double* vy, double chi, double eta,
                       double* wx, double* wy) {
  for (int32_t i = 0; i < m->n_cells; ++i) {
    const double sx = m->cm_x[i] - m->pt_x[i], sy = m->cm_y[i] - m->pt_y[i];
    const double dist = vabs(sx, sy);
    const double R = sqrt(m->tess_volume[i] / M_PI);
    double dwx = 0, dwy = 0;
    if (!(dist < 0.9 * eta * R)) {
      const double c = smax(dp2c(gamma, d[i], p[i]), vabs(vx[i], vy[i]));
      /* chi_*c*(s - r) / R: (chi*c) * vector, then each component / R */
      const double k = chi * c;
      dwx = k * sx / R;
      dwy = k * sy / R;
    }
    /* res[i] += calc_dw(...) */
    wx[i] = vx[i] + dwx;
    wy[i] = vy[i] + dwy;
  }
}

/* CorrectPointsOverShoot (round_cells.cpp:14-61) for a box that is not
 * half-periodic, in the order right, left, up, down; each slowed
 * velocity is the one the next test sees. */
void ro_point_fix(const ro_mesh* m, double dt, const double box[4], double* wx, double* wy) {
  const double x0 = box[0], x1 = box[1], y0 = box[2], y1 = box[3];
  const double inv_dt = 1.0 / dt;
  for (int32_t i = 0; i < m->n_cells; ++i) {
    const double px = m->pt_x[i], py = m->pt_y[i];
    const double R = sqrt(m->tess_volume[i] / M_PI);
    double vx = wx[i], vy = wy[i];
    const double dists[4] = {x1 - px, px - x0, y1 - py, py - y0};
    for (int k = 0; k < 4; ++k) {
      const int hit = k == 0 ? (vx * dt * 2 + px > x1)
                    : k == 1 ? (vx * dt * 2 + px < x0)
                    : k == 2 ? (vy * dt * 2 + py > y1)
                             : (vy * dt * 2 + py < y0);
      if (!hit) continue;
      double factor = 0.25 * dists[k] * inv_dt / vabs(vx, vy);
      if (R * 0.1 > dists[k]) factor *= -0.1;
      vx = vx * factor;
      vy = vy * factor;
    }
    wx[i] = vx;
    wy[i] = vy;
  }
}

/* StationaryBox::operator() (stationary_box.cpp:20-40): zero on a face
 * whose sides have the same GetOriginalIndex, else
 * VoronoiMesh::CalcFaceVelocity (VoronoiMesh.cpp:146-151):
 * w' = ScalarProd(wl - wr, f - (rR + rL)/2) * (rR - rL) / pow(abs(rR - rL), 2),
 * w = 0.5*(wl + wr) + w', f the face centroid. */
void ro_face_velocity(const ro_mesh* m, int32_t bc, const double* wx, const double* wy,
                      double* fwx, double* fwy) {
  for (int32_t f = 0; f < m->n_faces; ++f) {
    const int32_t a = m->face_first[f], b = m->face_second[f];
    const int32_t oa = orig_index(m, a), ob = orig_index(m, b);
    if (bc == 0 && oa == ob) {
      fwx[f] = 0.0;
      fwy[f] = 0.0;
      continue;
    }
    const double wlx = wx[oa], wly = wy[oa], wrx = wx[ob], wry = wy[ob];
    const double rLx = m->pt_x[a], rLy = m->pt_y[a], rRx = m->pt_x[b], rRy = m->pt_y[b];
    double fx, fy;
    face_centroid(m, f, &fx, &fy);
    const double s = (wlx - wrx) * (fx - (rRx + rLx) / 2) + (wly - wry) * (fy - (rRy + rLy) / 2);
    const double dx = rRx - rLx, dy = rRy - rLy;
    const double a2 = pow(vabs(dx, dy), 2);
    fwx[f] = 0.5 * (wlx + wrx) + s * dx / a2;
    fwy[f] = 0.5 * (wly + wry) + s * dy / a2;
  }
}
#endif
