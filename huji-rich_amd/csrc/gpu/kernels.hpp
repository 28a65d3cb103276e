// HIP kernels of the FP64 finite-volume step (gfx950).
//
// Work decomposition:
//   k_slopes      one thread per cell: LinearGaussImproved's slope with the
//                 shock detector and both limiters (LinearGaussImproved.cpp:67-494)
//   k_flux2       one thread per face: face states from the slopes, rotated
//                 HLLC, rotation back (modular_flux_calculator.cpp:39-71)
//   k_flux1       one thread per face: first-order SimpleFluxCalculator
//   k_update      one thread per cell: ascending-face segmented sum of
//                 -/+ dt*A*F (SimpleExtensiveUpdater), ZeroForce, the Heun
//                 average, IdealGas primitive recovery (SimpleCellUpdater)
//                 and the cell's CFL term, min-reduced by wavefront
//                 shuffles into one 64-bit atomic per wave.
// Each cell gathers its faces through a CSR in ascending face order, which
// is the order in which the reference scatters face contributions, so sums
// are deterministic and equal to the reference's without atomics.
#pragma once

#include "physics.hpp"

namespace rg {

constexpr int BC_RIGID = 0;
constexpr int BC_PERIODIC = 1;
constexpr unsigned INFO_SECOND = 0x80000000u;  // cell is the face's second side
constexpr unsigned INFO_MIRROR = 0x40000000u;  // GetOriginalIndex equal on both sides
constexpr unsigned INFO_FACE = 0x3FFFFFFFu;
constexpr double PI = 3.14159265358979323846;  // M_PI
constexpr unsigned long long DBL_INF_BITS = 0x7FF0000000000000ull;

struct DMesh {
  int N, F, G, bc;
  const int* ff;
  const int* fs;
  const double *v1x, *v1y, *v2x, *v2y;
  const double *px, *py, *cmx, *cmy;
  const double *vol, *tvol, *area;
  const int* off;
  const unsigned* info;
  const int* nbr;
  const int* gorig;  // GetOriginalIndex of each ghost
  const int* gsrc;   // real cell a ghost copies (its generator's choice)
  const double *gnx, *gny;
  const double *wx, *wy;  // face velocities; null = zero
};

struct DCells {
  double *d, *p, *vx, *vy;
};
struct DExt {
  double *m, *E, *Mx, *My;
};
struct DSlopes {
  double* s[8];  // x: d p vx vy, y: d p vx vy
};
struct DParams {
  double gamma, cfl, delta_v, theta, delta_p;
  int slf;
};
// first-error record: atomicMin of (step, phase, index, code)
struct DErr {
  unsigned long long* word;
  int step, phase;
};

__device__ __forceinline__ void report(const DErr& e, int index, int code) {
  const unsigned long long key = (static_cast<unsigned long long>(e.step) << 39) |
                                 (static_cast<unsigned long long>(e.phase) << 36) |
                                 (static_cast<unsigned long long>(static_cast<unsigned>(index)) << 4) |
                                 static_cast<unsigned long long>(code);
  atomicMin(e.word, key);
}

__device__ __forceinline__ Cell load_cell(const DCells& c, int i) {
  return Cell{c.d[i], c.p[i], c.vx[i], c.vy[i]};
}

// Cell value at any point id: real cells directly, ghosts through the ghost
// generator -- RigidWallGenerator (copy of the real cell of the ghost's
// first boundary face with the normal velocity reversed,
// RigidWallGenerator.cpp:5-37) or PeriodicGhostGenerator (copy of the
// original, PeriodicGhostGenerator.cpp:3-17).
__device__ __forceinline__ Cell fetch(const DMesh& m, const DCells& c, int id) {
  if (id < m.N) return load_cell(c, id);
  const int g = id - m.N;
  Cell r = load_cell(c, m.gsrc[g]);
  if (m.bc == BC_RIGID) {
    const double nx = m.gnx[g], ny = m.gny[g];
    const double s2 = 2 * (r.vx * nx + r.vy * ny);
    r.vx = r.vx - nx * s2;
    r.vy = r.vy - ny * s2;
  }
  return r;
}

__device__ __forceinline__ Cell sub(const Cell& a, const Cell& b) {
  return Cell{a.d - b.d, a.p - b.p, a.vx - b.vx, a.vy - b.vy};
}

// one field of slope_limit (LinearGaussImproved.cpp:232-267)
__device__ __forceinline__ void limit_field(double dphi, double cval, double cell, double maxd,
                                            double mind, bool absguard, double& psi) {
  if (fabs(dphi) > 0.05 * smax(fabs(maxd), fabs(mind)) || cval * cell < 0) {
    const double hi = absguard ? fabs(1e-9 * cell) : 1e-9 * cell;
    const double lo = absguard ? -fabs(1e-9 * cell) : -1e-9 * cell;
    if (dphi > hi)
      psi = smin(psi, maxd / dphi);
    else if (dphi < lo)
      psi = smin(psi, mind / dphi);
  }
}

// one field of shocked_slope_limit (LinearGaussImproved.cpp:343-366)
__device__ __forceinline__ void shocked_field(double dphi, double cval, double cell, double nval,
                                              double maxd, double mind, double theta,
                                              double& psi) {
  if (fabs(dphi) > 0.1 * smax(fabs(maxd), fabs(mind)) || cval * cell < 0) {
    if (fabs(dphi) > 1e-9 * cell) psi = smin(psi, smax(theta * (nval - cell) / dphi, 0.0));
  }
}

__device__ __forceinline__ void minmax(Cell& mx, Cell& mn, const Cell& t) {
  mx.d = smax(mx.d, t.d);
  mx.p = smax(mx.p, t.p);
  mx.vx = smax(mx.vx, t.vx);
  mx.vy = smax(mx.vy, t.vy);
  mn.d = smin(mn.d, t.d);
  mn.p = smin(mn.p, t.p);
  mn.vx = smin(mn.vx, t.vx);
  mn.vy = smin(mn.vy, t.vy);
}

__device__ __forceinline__ void face_centroid(const DMesh& m, int f, double& cx, double& cy) {
  cx = (m.v1x[f] + m.v2x[f]) * 0.5;
  cy = (m.v1y[f] + m.v2y[f]) * 0.5;
}

// LinearGaussImproved calc_slope (LinearGaussImproved.cpp:425-494) for one cell.
__global__ void __launch_bounds__(256) k_slopes(DMesh m, DCells c, DSlopes out, DParams prm,
                                                DErr err) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m.N) return;
  const int b = m.off[i], e = m.off[i + 1], n = e - b;
  const Cell cell = load_cell(c, i);
  bool boundary = false;
  for (int k = b; k < e; ++k)
    if (m.info[k] & INFO_MIRROR) boundary = true;
  Cell sx{0, 0, 0, 0}, sy{0, 0, 0, 0};
  const double cmx = m.cmx[i], cmy = m.cmy[i];
  double pr = 1;
  Cell amax = cell, amin = cell, bmax = cell, bmin = cell;
  if (boundary) {
    // GetBoundarySlope (LinearGaussImproved.cpp:401-422)
    Cell phy{0, 0, 0, 0}, phx{0, 0, 0, 0};
    double SxSy = 0, Sx2 = 0, Sy2 = 0;
    for (int k = b; k < e; ++k) {
      const int j = m.nbr[k];
      const Cell nc = fetch(m, c, j);
      const double ey = m.cmy[j] - cmy, ex = m.cmx[j] - cmx;
      const Cell dd = sub(nc, cell);
      phy.d += dd.d * ey;
      phy.p += dd.p * ey;
      phy.vx += dd.vx * ey;
      phy.vy += dd.vy * ey;
      phx.d += dd.d * ex;
      phx.p += dd.p * ex;
      phx.vx += dd.vx * ex;
      phx.vy += dd.vy * ex;
      SxSy += ey * ex;
      Sx2 += ex * ex;
      Sy2 += ey * ey;
      pr = (cell.p > nc.p) ? smin(pr, nc.p / cell.p) : smin(pr, cell.p / nc.p);
      minmax(amax, amin, nc);
      if (!(m.info[k] & INFO_MIRROR)) minmax(bmax, bmin, nc);
    }
    const double den = SxSy * SxSy - Sx2 * Sy2;
    sx.d = (phy.d * SxSy - phx.d * Sy2) / den;
    sx.p = (phy.p * SxSy - phx.p * Sy2) / den;
    sx.vx = (phy.vx * SxSy - phx.vx * Sy2) / den;
    sx.vy = (phy.vy * SxSy - phx.vy * Sy2) / den;
    sy.d = (phx.d * SxSy - phy.d * Sx2) / den;
    sy.p = (phx.p * SxSy - phy.p * Sx2) / den;
    sy.vx = (phx.vx * SxSy - phy.vx * Sx2) / den;
    sy.vy = (phx.vy * SxSy - phy.vy * Sx2) / den;
  } else {
    // calc_naive_slope (LinearGaussImproved.cpp:67-134)
    if (n > 20) {
      report(err, i, ERR_TOO_MANY_NEIGHBORS);
      for (int q = 0; q < 8; ++q) out.s[q][i] = 0;
      return;
    }
    const double ptx = m.px[i], pty = m.py[i], vol = m.tvol[i];
    double M0 = 0, M1 = 0, M2 = 0, M3 = 0;
    Cell vcx{0, 0, 0, 0}, vcy{0, 0, 0, 0};
    for (int k = b; k < e; ++k) {
      const int f = static_cast<int>(m.info[k] & INFO_FACE);
      const int j = m.nbr[k];
      const Cell nc = fetch(m, c, j);
      double fx, fy;
      face_centroid(m, f, fx, fy);
      const double cijx = fx - (m.cmx[j] + cmx) * 0.5;
      const double cijy = fy - (m.cmy[j] + cmy) * 0.5;
      const double L = vabs(m.v2x[f] - m.v1x[f], m.v2y[f] - m.v1y[f]);
      const double dx = m.px[j] - ptx, dy = m.py[j] - pty;
      const double ad = vabs(dx, dy);
      const double rx = dx / ad * L, ry = dy / ad * L;
      M0 -= cijx * rx;
      M1 -= cijy * rx;
      M2 -= cijx * ry;
      M3 -= cijy * ry;
      const double hx = rx * 0.5, hy = ry * 0.5;
      if (k == b) {
        vcx = Cell{cell.d * hx, cell.p * hx, cell.vx * hx, cell.vy * hx};
        vcy = Cell{cell.d * hy, cell.p * hy, cell.vx * hy, cell.vy * hy};
      } else {
        add_mult(vcx, cell, hx);
        add_mult(vcy, cell, hy);
      }
      add_mult(vcy, nc, hy);
      add_mult(vcx, nc, hx);
      pr = (cell.p > nc.p) ? smin(pr, nc.p / cell.p) : smin(pr, cell.p / nc.p);
      minmax(amax, amin, nc);
      if (!(m.info[k] & INFO_MIRROR)) minmax(bmax, bmin, nc);
    }
    M0 += vol;
    M3 += vol;
    const double det = M0 * M3 - M1 * M2;
    if (fabs(det) < 1e-10 * vol * vol) {
      report(err, i, ERR_SINGULAR_MATRIX);
      for (int q = 0; q < 8; ++q) out.s[q][i] = 0;
      return;
    }
    const double det_inv = 1.0 / det;
    const double i0 = M3 * det_inv, i1 = -M1 * det_inv, i2 = -M2 * det_inv, i3 = M0 * det_inv;
    sx = Cell{vcx.d * i0, vcx.p * i0, vcx.vx * i0, vcx.vy * i0};
    sy = Cell{vcy.d * i3, vcy.p * i3, vcy.vx * i3, vcy.vy * i3};
    add_mult(sx, vcy, i1);
    add_mult(sy, vcx, i2);
  }
  if (prm.slf) {
    // is_shock (LinearGaussImproved.cpp:150-157)
    const double cs = dp2c(prm.gamma, cell.d, cell.p);
    const double width = sqrt(m.tvol[i] / PI);
    const bool shock = ((sx.vx + sy.vy) * width < (-prm.delta_v) * cs) || (pr < prm.delta_p);
    double psi0 = 1, psi1 = 1, psi2 = 1, psi3 = 1;
    if (!shock) {
      // slope_limit (LinearGaussImproved.cpp:174-300)
      const Cell maxd = sub(bmax, cell), mind = sub(bmin, cell);
      for (int k = b; k < e; ++k) {
        const int f = static_cast<int>(m.info[k] & INFO_FACE);
        double fx, fy;
        face_centroid(m, f, fx, fy);
        const Cell cv = interp(cell, sx, sy, fx, fy, cmx, cmy);
        const Cell dphi = sub(cv, cell);
        limit_field(dphi.d, cv.d, cell.d, maxd.d, mind.d, false, psi0);
        limit_field(dphi.p, cv.p, cell.p, maxd.p, mind.p, false, psi1);
        limit_field(dphi.vx, cv.vx, cell.vx, maxd.vx, mind.vx, true, psi2);
        limit_field(dphi.vy, cv.vy, cell.vy, maxd.vy, mind.vy, true, psi3);
      }
    } else {
      // shocked_slope_limit (LinearGaussImproved.cpp:302-399)
      const Cell maxd = sub(amax, cell), mind = sub(amin, cell);
      for (int k = b; k < e; ++k) {
        const int f = static_cast<int>(m.info[k] & INFO_FACE);
        const Cell nc = fetch(m, c, m.nbr[k]);
        double fx, fy;
        face_centroid(m, f, fx, fy);
        const Cell cv = interp(cell, sx, sy, fx, fy, cmx, cmy);
        const Cell dphi = sub(cv, cell);
        shocked_field(dphi.d, cv.d, cell.d, nc.d, maxd.d, mind.d, prm.theta, psi0);
        shocked_field(dphi.p, cv.p, cell.p, nc.p, maxd.p, mind.p, prm.theta, psi1);
        shocked_field(dphi.vx, cv.vx, cell.vx, nc.vx, maxd.vx, mind.vx, prm.theta, psi2);
        shocked_field(dphi.vy, cv.vy, cell.vy, nc.vy, maxd.vy, mind.vy, prm.theta, psi3);
      }
    }
    sx.d *= psi0;
    sy.d *= psi0;
    sx.p *= psi1;
    sy.p *= psi1;
    sx.vx *= psi2;
    sy.vx *= psi2;
    sx.vy *= psi3;
    sy.vy *= psi3;
  }
  out.s[0][i] = sx.d;
  out.s[1][i] = sx.p;
  out.s[2][i] = sx.vx;
  out.s[3][i] = sx.vy;
  out.s[4][i] = sy.d;
  out.s[5][i] = sy.p;
  out.s[6][i] = sy.vx;
  out.s[7][i] = sy.vy;
}

__device__ __forceinline__ void load_slope(const DSlopes& s, int i, Cell& sx, Cell& sy) {
  sx = Cell{s.s[0][i], s.s[1][i], s.s[2][i], s.s[3][i]};
  sy = Cell{s.s[4][i], s.s[5][i], s.s[6][i], s.s[7][i]};
}

// Face state on one side (LinearGaussImproved.cpp:571-590 for real cells,
// :597-634 for ghosts: ghost value plus the ghost gradient -- zero for
// rigid walls, the original's slope for periodic boxes).
__device__ __forceinline__ Cell side_state(const DMesh& m, const DCells& c, const DSlopes& s,
                                           int id, double fx, double fy) {
  Cell v, sx, sy;
  if (id < m.N) {
    v = load_cell(c, id);
    load_slope(s, id, sx, sy);
  } else {
    v = fetch(m, c, id);
    if (m.bc == BC_PERIODIC) {
      load_slope(s, m.gorig[id - m.N], sx, sy);
    } else {
      sx = Cell{0, 0, 0, 0};
      sy = Cell{0, 0, 0, 0};
    }
  }
  return interp(v, sx, sy, fx, fy, m.cmx[id], m.cmy[id]);
}

// ModularFluxCalculator body for one face (modular_flux_calculator.cpp:49-69).
__device__ __forceinline__ int modular_face_flux(const DMesh& m, int f, const Cell& lf,
                                                 const Cell& rt, double gamma, Cons& out) {
  const int a = m.ff[f], b = m.fs[f];
  const double v1x = m.v1x[f], v1y = m.v1y[f];
  // calc_parallel_normal (modular_flux_calculator.cpp:12-22) with
  // remove_parallel_component (simple_flux_calculator.cpp:27-31)
  const double px = m.v2x[f] - v1x, py = m.v2y[f] - v1y;
  double qx, qy;
  if (a >= 0 && a < m.N) {
    qx = v1x - m.px[a];
    qy = v1y - m.py[a];
  } else {
    qx = m.px[b] - v1x;
    qy = m.py[b] - v1y;
  }
  const double s = qx * px + qy * py;
  const double pp = px * px + py * py;
  const double nx = qx - px * s / pp, ny = qy - py * s / pp;
  const double an = vabs(nx, ny), ap = vabs(px, py);
  const double wx = m.wx ? m.wx[f] : 0.0, wy = m.wy ? m.wy[f] : 0.0;
  const double speed = (nx * wx + ny * wy) / an;
  const Prim pl = rotate(to_prim(lf, gamma), nx, ny, an, px, py, ap);
  const Prim pr = rotate(to_prim(rt, gamma), nx, ny, an, px, py, ap);
  Cons h;
  const int rc = hllc(pl, pr, speed, h);
  if (rc) return rc;
  out = rotate_back(h, nx, ny, an, px, py, ap);
  return 0;
}

__global__ void __launch_bounds__(256) k_flux2(DMesh m, DCells c, DSlopes s, DExt flux,
                                               DParams prm, DErr err) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= m.F) return;
  double fx, fy;
  face_centroid(m, f, fx, fy);
  const Cell lf = side_state(m, c, s, m.ff[f], fx, fy);
  const Cell rt = side_state(m, c, s, m.fs[f], fx, fy);
  Cons o;
  const int rc = modular_face_flux(m, f, lf, rt, prm.gamma, o);
  if (rc) {
    report(err, f, rc);
    o = Cons{0, 0, 0, 0};
  }
  flux.m[f] = o.m;
  flux.E[f] = o.E;
  flux.Mx[f] = o.mx;
  flux.My[f] = o.my;
}

// SpatialReconstruction output: res[f].first / .second
__global__ void __launch_bounds__(256) k_face_states(DMesh m, DCells c, DSlopes s, DCells left,
                                                     DCells right) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= m.F) return;
  double fx, fy;
  face_centroid(m, f, fx, fy);
  const Cell lf = side_state(m, c, s, m.ff[f], fx, fy);
  const Cell rt = side_state(m, c, s, m.fs[f], fx, fy);
  left.d[f] = lf.d;
  left.p[f] = lf.p;
  left.vx[f] = lf.vx;
  left.vy[f] = lf.vy;
  right.d[f] = rt.d;
  right.p[f] = rt.p;
  right.vx[f] = rt.vx;
  right.vy[f] = rt.vy;
}

// SimpleFluxCalculator body (simple_flux_calculator.cpp:80-126,167-189):
// normal = point difference; a mirror face reflects the real state about the
// face (Reflect, geometry.cpp:117-120).
__global__ void __launch_bounds__(256) k_flux1(DMesh m, DCells c, DExt flux, DParams prm,
                                               DErr err) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= m.F) return;
  const int a = m.ff[f], b = m.fs[f];
  const double px = m.v2x[f] - m.v1x[f], py = m.v2y[f] - m.v1y[f];
  const double nx = m.px[b] - m.px[a], ny = m.py[b] - m.py[a];
  const bool fa = a >= 0 && a < m.N, fb = b >= 0 && b < m.N;
  const int oa = fa ? a : m.gorig[a - m.N];
  const int ob = fb ? b : m.gorig[b - m.N];
  const bool mirror = oa == ob;  // GetOriginalIndex equal on both sides
  Prim l, r;
  auto reflect = [&](const Prim& s) {
    Prim q = s;
    const double s2 = 2 * (s.vx * px + s.vy * py);
    const double ap = vabs(px, py);
    const double a2 = ap * ap;
    q.vx = px * s2 / a2 - s.vx;
    q.vy = py * s2 / a2 - s.vy;
    return q;
  };
  if (!fa) {
    r = to_prim(load_cell(c, b), prm.gamma);
    l = mirror ? reflect(r) : to_prim(fetch(m, c, a), prm.gamma);
  } else if (!fb) {
    l = to_prim(load_cell(c, a), prm.gamma);
    r = mirror ? reflect(l) : to_prim(fetch(m, c, b), prm.gamma);
  } else {
    l = to_prim(load_cell(c, a), prm.gamma);
    r = to_prim(load_cell(c, b), prm.gamma);
  }
  const double an = vabs(nx, ny), ap = vabs(px, py);
  const double wx = m.wx ? m.wx[f] : 0.0, wy = m.wy ? m.wy[f] : 0.0;
  const double velocity = (nx * wx + ny * wy) / an;
  Cons h, o;
  const int rc = hllc(rotate(l, nx, ny, an, px, py, ap), rotate(r, nx, ny, an, px, py, ap),
                      velocity, h);
  if (rc) {
    report(err, f, rc);
    o = Cons{0, 0, 0, 0};
  } else {
    o = rotate_back(h, nx, ny, an, px, py, ap);
  }
  flux.m[f] = o.m;
  flux.E[f] = o.E;
  flux.Mx[f] = o.mx;
  flux.My[f] = o.my;
}

// TimeStepCalculator::operator[] (simple_cfl.cpp:28-47) for one cell.
__device__ __forceinline__ double cfl_inv(const DMesh& m, int i, double gamma, const Cell& q) {
  const double radius = sqrt(m.tvol[i] / PI);
  const double cs = dp2c(gamma, q.d, q.p);
  double res = 0;
  if (m.wx) {
    for (int k = m.off[i]; k < m.off[i + 1]; ++k) {
      const int f = static_cast<int>(m.info[k] & INFO_FACE);
      res = fmax(res, (cs + vabs(q.vx - m.wx[f], q.vy - m.wy[f])) / radius);
    }
  } else if (m.off[i + 1] > m.off[i]) {
    // all face velocities zero: every face gives the same term
    res = fmax(res, (cs + vabs(q.vx - 0.0, q.vy - 0.0)) / radius);
  }
  return 1.0 / res;
}

// wavefront min, then one 64-bit atomic per wave (positive doubles order
// like their bit patterns)
__device__ __forceinline__ void wave_min_to(unsigned long long* slot, double v) {
  for (int o = 32; o > 0; o >>= 1) v = smin(v, __shfl_xor(v, o, 64));
  if ((threadIdx.x & 63) == 0) atomicMin(slot, static_cast<unsigned long long>(__double_as_longlong(v)));
}

__global__ void __launch_bounds__(256) k_cfl(DMesh m, DCells c, double gamma,
                                             unsigned long long* slot) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double v = __longlong_as_double(static_cast<long long>(DBL_INF_BITS));
  if (i < m.N) v = cfl_inv(m, i, gamma, load_cell(c, i));
  wave_min_to(slot, v);
}

__global__ void k_fill_u64(unsigned long long* p, unsigned long long v, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

enum : int {
  UPD_ZF_BEFORE = 1,   // ZeroForce before the face sum (TimeAdvance2Heun)
  UPD_ZF_AFTER = 2,    // ZeroForce after the face sum (TimeAdvance)
  UPD_AVERAGE = 4,     // 0.5*(mid + old) (average_extensive)
  UPD_CELLS = 8,       // SimpleCellUpdater
  UPD_CFL = 16,        // next step's CFL term
  UPD_FACES = 32,      // apply the face fluxes
};

struct UpdArgs {
  DExt in, out, avg;
  DExt flux;
  DCells cells;
  double dt;                          // used when dt_slot is null
  const unsigned long long* dt_slot;  // min_i 1/res_i from the CFL reduction
  unsigned long long* reset_slot;     // set to +inf (next step's slot)
  unsigned long long* cfl_slot;       // receives this step's min
  double* dt_hist;
  int step;
  int flags;
};

// SimpleExtensiveUpdater (simple_extensive_updater.cpp:11-35) as a per-cell
// gather in ascending face order, ExternalForceContribution with ZeroForce
// (hydrodynamics_2d.cpp:348-372), average_extensive (hdsim2d.cpp:311-320),
// regular_update (simple_cell_updater.cpp:22-59), TimeStepCalculator.
__global__ void __launch_bounds__(256) k_update(DMesh m, UpdArgs a, DParams prm, DErr err) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const double dt = a.dt_slot ? prm.cfl * __longlong_as_double(static_cast<long long>(*a.dt_slot))
                              : a.dt;
  if (i == 0) {
    if (a.dt_hist) a.dt_hist[a.step] = dt;
    if (a.reset_slot) *a.reset_slot = DBL_INF_BITS;
  }
  double cfl_v = __longlong_as_double(static_cast<long long>(DBL_INF_BITS));
  if (i < m.N) {
    double em = a.in.m[i], eE = a.in.E[i], ex = a.in.Mx[i], ey = a.in.My[i];
    if (a.flags & UPD_ZF_BEFORE) {
      em += dt * 0.0;
      ex += 0.0 * dt;
      ey += 0.0 * dt;
      eE += dt * 0.0;
    }
    if (a.flags & UPD_FACES) {
      for (int k = m.off[i]; k < m.off[i + 1]; ++k) {
        const unsigned inf = m.info[k];
        const int f = static_cast<int>(inf & INFO_FACE);
        const double s = dt * m.area[f];
        const double dm = a.flux.m[f] * s, dE = a.flux.E[f] * s, dx = a.flux.Mx[f] * s,
                     dy = a.flux.My[f] * s;
        if (inf & INFO_SECOND) {
          em += dm;
          eE += dE;
          ex += dx;
          ey += dy;
        } else {
          em -= dm;
          eE -= dE;
          ex -= dx;
          ey -= dy;
        }
      }
    }
    if (a.flags & UPD_ZF_AFTER) {
      em += dt * 0.0;
      ex += 0.0 * dt;
      ey += 0.0 * dt;
      eE += dt * 0.0;
    }
    if (a.flags & UPD_AVERAGE) {
      em = 0.5 * (em + a.avg.m[i]);
      eE = 0.5 * (eE + a.avg.E[i]);
      ex = (ex + a.avg.Mx[i]) * 0.5;
      ey = (ey + a.avg.My[i]) * 0.5;
    }
    a.out.m[i] = em;
    a.out.E[i] = eE;
    a.out.Mx[i] = ex;
    a.out.My[i] = ey;
    if (a.flags & UPD_CELLS) {
      const double d = em / m.vol[i];
      const double vx = ex / em, vy = ey / em;
      const double e = eE / em - 0.5 * (vx * vx + vy * vy);
      if (d < 0) report(err, i, ERR_NEGATIVE_DENSITY);
      else if (e < 0) report(err, i, ERR_NEGATIVE_THERMAL);
      const double p = (prm.gamma - 1) * e * d;
      a.cells.d[i] = d;
      a.cells.p[i] = p;
      a.cells.vx[i] = vx;
      a.cells.vy[i] = vy;
      if (a.flags & UPD_CFL) cfl_v = cfl_inv(m, i, prm.gamma, Cell{d, p, vx, vy});
    }
  }
  if (a.flags & UPD_CFL) wave_min_to(a.cfl_slot, cfl_v);
}

// init_extensives (hdsim2d.cpp:41-66)
__global__ void __launch_bounds__(256) k_init_ext(DMesh m, DCells c, DExt x, double gamma) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m.N) return;
  const double volume = m.vol[i];
  const double d = c.d[i], p = c.p[i], vx = c.vx[i], vy = c.vy[i];
  const double mass = volume * d;
  x.m[i] = mass;
  x.E[i] = dp2e(gamma, d, p) * mass + 0.5 * mass * (vx * vx + vy * vy);
  x.Mx[i] = vx * mass;
  x.My[i] = vy * mass;
}

}  // namespace rg
