// Device-resident inputs of a moving-mesh step: point velocities
// (Lagrangian + RoundCells), the overshoot fix at the walls, and face
// velocities from point velocities.  Each kernel evaluates the expressions
// of the reference routine it cites in the same order, so it gives the same
// values the host forms.
#pragma once

#include "kernels.hpp"

namespace rg {

// Lagrangian (lagrangian.cpp:3-13) + RoundCells::calc_dw without the cold
// switch (round_cells.cpp:63-76): w = v + chi max(c, |v|) (CM - r) / R where
// the point is at least 0.9 eta R from its CM, R = GetWidth.
__global__ void __launch_bounds__(256) k_point_velocity(DMesh m, const D4* __restrict__ q,
                                                        double gamma, double chi, double eta,
                                                        D2* __restrict__ pv) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m.N) return;
  const Cell c = ld_cell(q, i);
  const D2 s = m.cm[i];
  const double sx = s.x - m.px[i], sy = s.y - m.py[i];
  const double d = sqrt(sx * sx + sy * sy);
  const double R = m.width[i];
  double dwx = 0.0, dwy = 0.0;
  if (!(d < 0.9 * eta * R)) {
    const double cs = dp2c(gamma, c.d, c.p);
    const double va = sqrt(c.vx * c.vx + c.vy * c.vy);
    const double k = chi * ((cs < va) ? va : cs);
    dwx = k * sx / R;
    dwy = k * sy / R;
  }
  pv[i] = D2{c.vx + dwx, c.vy + dwy};  // res[i] += calc_dw (round_cells.cpp:117)
}

// CorrectPointsOverShoot of a rigid box (round_cells.cpp:14-61) with
// dt = cfl * min_i inv_i from the dt slot.
__global__ void __launch_bounds__(256) k_point_fix(DMesh m, D2* __restrict__ pv,
                                                   const unsigned long long* __restrict__ slot,
                                                   double cfl, double x0, double x1, double y0,
                                                   double y1) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m.N) return;
  const double dt = cfl * __longlong_as_double(static_cast<long long>(*slot));
  const double inv_dt = 1.0 / dt;
  const double px = m.px[i], py = m.py[i], R = m.width[i];
  D2 v = pv[i];
  auto fix = [&](double dist) {
    double f = 0.25 * dist * inv_dt / sqrt(v.x * v.x + v.y * v.y);
    if (R * 0.1 > dist) f = f * -0.1;
    v.x = v.x * f;
    v.y = v.y * f;
  };
  if (v.x * dt * 2 + px > x1) fix(x1 - px);
  if (v.x * dt * 2 + px < x0) fix(px - x0);
  if (v.y * dt * 2 + py > y1) fix(y1 - py);
  if (v.y * dt * 2 + py < y0) fix(py - y0);
  pv[i] = v;
}

// VoronoiMesh::CalcFaceVelocity (VoronoiMesh.cpp:146-151) of every face,
// zero on mirror faces of a rigid box (StationaryBox, stationary_box.cpp:
// 20-40); a ghost side moves with its original point.
__global__ void __launch_bounds__(256) k_face_velocity(DMesh m, const D2* __restrict__ pv,
                                                       double* __restrict__ fwx,
                                                       double* __restrict__ fwy) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= m.F) return;
  const int a = m.ff[f], b = m.fs[f];
  const int oa = a < m.N ? a : m.gorig[a - m.N], ob = b < m.N ? b : m.gorig[b - m.N];
  if (m.bc == BC_RIGID && oa == ob) {
    fwx[f] = 0.0;
    fwy[f] = 0.0;
    return;
  }
  const D2 wl = pv[oa], wr = pv[ob];
  const double rLx = m.px[a], rLy = m.py[a], rRx = m.px[b], rRy = m.py[b];
  const double fx = (m.v1x[f] + m.v2x[f]) * 0.5, fy = (m.v1y[f] + m.v2y[f]) * 0.5;
  const double mx = (rRx + rLx) / 2, my = (rRy + rLy) / 2;
  const double s = (wl.x - wr.x) * (fx - mx) + (wl.y - wr.y) * (fy - my);
  const double dxr = rRx - rLx, dyr = rRy - rLy;
  const double ab = sqrt(dxr * dxr + dyr * dyr);
  const double d2 = ab * ab;
  fwx[f] = (wl.x + wr.x) * 0.5 + (dxr * s) / d2;
  fwy[f] = (wl.y + wr.y) * 0.5 + (dyr * s) / d2;
}

// Cell geometry from the face vertices, in each cell's ascending face
// order: Tessellation::GetVolume, a fan about the generating point
// (VoronoiMesh.cpp:420-434); SlabSymmetry::calcVolume, a fan of
// calc_triangle_area anchored at the first face's first vertex
// (physical_geometry.cpp:13-21, triangle_area.cpp:4-9); and CalcCellCM on
// the same fan (VoronoiMesh.cpp:436-451).  zcross(v) = (v.y, -v.x).
__global__ void __launch_bounds__(256) k_cell_geometry(DMesh m, double* __restrict__ tvol,
                                                       double* __restrict__ vol,
                                                       D2* __restrict__ cm) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m.N) return;
  const int b = m.off[i], e = m.off[i + 1];
  const double cx = m.px[i], cy = m.py[i];
  double area = 0;
  for (int k = b; k < e; ++k) {
    const int f = static_cast<int>(m.inc[k].info & INFO_FACE);
    const double p1x = m.v1x[f] - cx, p1y = m.v1y[f] - cy;
    const double p2x = m.v2x[f] - cx, p2y = m.v2y[f] - cy;
    area += 0.5 * fabs(p1x * p2y + p1y * (-p2x));
  }
  double v = 0, a2 = 0, pcx = 0, pcy = 0;
  if (e > b) {
    const int f0 = static_cast<int>(m.inc[b].info & INFO_FACE);
    const double ax = m.v1x[f0], ay = m.v1y[f0];
    for (int k = b + 1; k < e; ++k) {
      const int f = static_cast<int>(m.inc[k].info & INFO_FACE);
      const double q1x = m.v1x[f] - ax, q1y = m.v1y[f] - ay;
      const double q2x = m.v2x[f] - ax, q2y = m.v2y[f] - ay;
      const double at = 0.5 * fabs(q1x * q2y + q1y * (-q2x));
      v += at;
      a2 += at;
      const double s = at / 3.;
      pcx += ((ax + m.v1x[f]) + m.v2x[f]) * s;
      pcy += ((ay + m.v1y[f]) + m.v2y[f]) * s;
    }
  }
  tvol[i] = area;
  vol[i] = v;
  cm[i] = D2{pcx / a2, pcy / a2};
}

// Centres of mass of the ghosts (VoronoiMesh.cpp:359-398): a periodic copy
// is its original's CM shifted with its point; a rigid-wall ghost is its
// original's CM reflected across the wall its point lies beyond
// (GetReflection, VoronoiMesh.cpp:316-330: point + (2 (wall - x), 0) or
// (0, 2 (wall - y))).
__global__ void __launch_bounds__(256) k_ghost_cm(DMesh m, D2* __restrict__ cm, double x0,
                                                  double x1, double y0, double y1) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= m.G) return;
  const int id = m.N + g, o = m.gorig[g];
  const D2 c = cm[o];
  if (m.bc == BC_PERIODIC) {
    cm[id] = D2{c.x + (m.px[id] - m.px[o]), c.y + (m.py[id] - m.py[o])};
    return;
  }
  const double gx = m.px[id], gy = m.py[id];
  if (gx > x1)
    cm[id] = D2{c.x + 2 * (x1 - c.x), c.y + 0.0};
  else if (gy > y1)
    cm[id] = D2{c.x + 0.0, c.y + 2 * (y1 - c.y)};
  else if (gx < x0)
    cm[id] = D2{c.x + 2 * (x0 - c.x), c.y + 0.0};
  else
    cm[id] = D2{c.x + 0.0, c.y + 2 * (y0 - c.y)};
}

// dst[i] = src[order[i]] (VectorValues, misc/utils.hpp): the state follows
// the cells into a new order
template <class T>
__global__ void __launch_bounds__(256) k_gather(T* __restrict__ dst, const T* __restrict__ src,
                                               const int* __restrict__ order, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[order[i]];
}

}  // namespace rg
