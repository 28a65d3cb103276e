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

}  // namespace rg
