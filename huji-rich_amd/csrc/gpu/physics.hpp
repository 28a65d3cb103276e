// Device-side physics of the RICH finite-volume step: ideal-gas EOS, HLLC in
// the rotated face frame, and the face-state interpolation of
// LinearGaussImproved.  Built with -ffp-contract=off; each function keeps the
// operation order of the reference routine it cites so that FP64 results
// are bit-identical to the reference's (IEEE division and square root are
// correctly rounded on both sides).  The operand order of the reference's
// overloaded Vector2D / ComputationalCell / Conserved operators is kept too,
// since it decides where each rounding happens.
#pragma once

#include <hip/hip_runtime.h>

namespace rg {

struct Cell {  // ComputationalCell without tracers (computational_cell_2d.hpp:22-86)
  double d, p, vx, vy;
};
struct Prim {  // Primitive (hydrodynamic_variables.hpp:11-60)
  double d, p, vx, vy, e, c;
};
struct Cons {  // Conserved (hydrodynamic_variables.hpp:90-144)
  double m, mx, my, E;
};

enum : int {
  ERR_NEGATIVE_DENSITY = 1,
  ERR_NEGATIVE_THERMAL = 2,
  ERR_HLLC_NAN = 3,
  ERR_INVALID_WAVE_SPEEDS = 4,
  ERR_SINGULAR_MATRIX = 5,
  ERR_TOO_MANY_NEIGHBORS = 6,
};

// std::min / std::max as the reference calls them
__host__ __device__ __forceinline__ double smin(double a, double b) { return (b < a) ? b : a; }
__host__ __device__ __forceinline__ double smax(double a, double b) { return (a < b) ? b : a; }
// abs(Vector2D), geometry.cpp:18-21
__host__ __device__ __forceinline__ double vabs(double x, double y) { return sqrt(x * x + y * y); }
// is_nan, misc/utils.cpp:3-8
__host__ __device__ __forceinline__ bool is_nan_ref(double x) {
  const int b1 = (x >= 0);
  const int b2 = (x < 0);
  return 1 != b1 + b2;
}

// The face speed along the normal, (n . w) / |n| (modular_flux_calculator.cpp
// :56-67, simple_flux_calculator.cpp:112-114), for w = 0 (no face
// velocities): the numerator is a signed zero (or NaN), and a signed zero
// divided by a positive |n| (finite or not) is that zero, so the quotient
// needs no division; |n| = 0 or NaN keeps it.  Bit for bit the quotient.
__host__ __device__ __forceinline__ double zero_face_speed(double nx, double ny, double an) {
  const double num = nx * 0.0 + ny * 0.0;
  return an > 0.0 ? num : num / an;
}

// IdealGas::dp2e / dp2c (ideal_gas.cpp:13-16, 40-44)
__host__ __device__ __forceinline__ double dp2e(double g, double d, double p) { return p / d / (g - 1); }
__host__ __device__ __forceinline__ double dp2c(double g, double d, double p) { return sqrt(g * p / d); }

// TotalEnergyDensity (hydrodynamic_variables.cpp:261-265): pow(abs(v),2) is
// evaluated as abs(v)*abs(v)
__host__ __device__ __forceinline__ double ted(const Prim& s) {
  const double a = vabs(s.vx, s.vy);
  return s.d * (0.5 * (a * a) + s.e);
}

// rotate_back (simple_flux_calculator.cpp:142-151)
__host__ __device__ __forceinline__ Cons rotate_back(const Cons& c, double nx, double ny, double an,
                                            double px, double py, double ap) {
  Cons r = c;
  r.mx = nx * c.mx / an + px * c.my / ap;
  r.my = ny * c.mx / an + py * c.my / ap;
  return r;
}

// ComputationalCellAddMult (computational_cell_2d.cpp:59-68)
__host__ __device__ __forceinline__ void add_mult(Cell& r, const Cell& o, double s) {
  r.d += o.d * s;
  r.p += o.p * s;
  r.vx += o.vx * s;
  r.vy += o.vy * s;
}

// interp / interp2 (LinearGaussImproved.cpp:159-172)
__host__ __device__ __forceinline__ Cell interp(const Cell& c, const Cell& sx, const Cell& sy, double tx,
                                       double ty, double cmx, double cmy) {
  Cell r = c;
  add_mult(r, sx, tx - cmx);
  add_mult(r, sy, ty - cmy);
  return r;
}

// interp with the centroid-minus-CM offsets already formed
__host__ __device__ __forceinline__ Cell interp_off(const Cell& c, const Cell& sx, const Cell& sy, double ox,
                                           double oy) {
  Cell r = c;
  add_mult(r, sx, ox);
  add_mult(r, sy, oy);
  return r;
}

// One face: convert_to_primitive on both sides, rotate into the face frame,
// Hllc, rotate_back (simple_flux_calculator.cpp:9-15,130-165; hllc.cpp:97-154).
// The HLLC branch uses only one side's energy and tangential velocity, so
// those two quantities (three IEEE divisions) are computed for that side
// only; every value that is computed is evaluated exactly as above.
__host__ __device__ __forceinline__ int hllc_face(const Cell& L, const Cell& R, double g, double nx,
                                         double ny, double an, double px, double py, double ap,
                                         double velocity, Cons& out) {
  const double cl = dp2c(g, L.d, L.p), cr = dp2c(g, R.d, R.p);
  const double rvx = (R.vx * nx + R.vy * ny) / an;
  if (is_nan_ref(rvx)) return ERR_HLLC_NAN;
  const double lvx = (L.vx * nx + L.vy * ny) / an;
  // Velocity -= velocity*normaldir, normaldir = (1,0)
  const double ul = lvx - 1.0 * velocity, ur = rvx - 1.0 * velocity;
  const double sl = smin(ul - cl, ur - cr);
  const double sr = smax(ul + cl, ur + cr);
  const double ss = (R.p - L.p + L.d * ul * (sl - ul) - R.d * ur * (sr - ur)) /
                    (L.d * (sl - ul) - R.d * (sr - ur));
  const int branch = (sl > 0) ? 0 : (sl <= 0 && ss >= 0) ? 1 : (ss < 0 && sr >= 0) ? 2
                                                            : (sr < 0) ? 3 : 4;
  if (branch == 4) return ERR_INVALID_WAVE_SPEEDS;
  const bool left = branch <= 1;
  const Cell& S = left ? L : R;
  const double u = left ? ul : ur;
  const double sk = left ? sl : sr;
  // the upwind side's tangential velocity (rotate) and specific energy (dp2e)
  const double v = (S.vx * px + S.vy * py) / ap - 0.0 * velocity;
  const Prim s{S.d, S.p, u, v, dp2e(g, S.d, S.p), 0.0};
  const double un = s.vx * 1.0 + s.vy * 0.0;
  const double dun = s.d * un;
  const double es = ted(s);
  Cons f;
  if (branch == 0 || branch == 3) {
    f.m = s.d * un;
    f.mx = 1.0 * s.p + s.vx * dun;
    f.my = 0.0 * s.p + s.vy * dun;
    f.E = (es + s.p) * un;
  } else {
    const double ds = s.d * (sk - s.vx) / (sk - ss);  // starred_state, hllc.cpp:76-92
    const double usE = es * ds / s.d + ds * (ss - s.vx) * (ss + s.p / s.d / (sk - s.vx));
    f.m = s.d * un + sk * (ds - s.d);
    f.mx = (1.0 * s.p + s.vx * dun) + (ds * ss - s.vx * s.d) * sk;
    f.my = (0.0 * s.p + s.vy * dun) + (ds * s.vy - s.vy * s.d) * sk;
    f.E = (es + s.p) * un + sk * (usE - es);
  }
  f.E = f.E + (f.mx * (1.0 * velocity) + f.my * (0.0 * velocity) + 0.5 * f.m * velocity * velocity);
  const double vm = velocity * f.m;
  f.mx = f.mx + 1.0 * vm;
  f.my = f.my + 0.0 * vm;
  out = rotate_back(f, nx, ny, an, px, py, ap);
  return 0;
}

// Hllc::operator() (hllc.cpp:97-154) on two Primitives already in the face
// frame (x along the normal), with the energy and sound speed they carry,
// for a face moving at `velocity` along the normal: the same operations as
// hllc_face after its rotation, the upwind side's flux only.
__host__ __device__ __forceinline__ int hllc_prim(const Prim& L, const Prim& R, double velocity,
                                                  Cons& out) {
  if (is_nan_ref(R.vx)) return ERR_HLLC_NAN;
  // Velocity -= velocity*normaldir, normaldir = (1,0)
  const double ul = L.vx - 1.0 * velocity, ur = R.vx - 1.0 * velocity;
  const double sl = smin(ul - L.c, ur - R.c);
  const double sr = smax(ul + L.c, ur + R.c);
  const double ss = (R.p - L.p + L.d * ul * (sl - ul) - R.d * ur * (sr - ur)) /
                    (L.d * (sl - ul) - R.d * (sr - ur));
  const int branch = (sl > 0) ? 0 : (sl <= 0 && ss >= 0) ? 1 : (ss < 0 && sr >= 0) ? 2
                                                            : (sr < 0) ? 3 : 4;
  if (branch == 4) return ERR_INVALID_WAVE_SPEEDS;
  const bool left = branch <= 1;
  const Prim& S = left ? L : R;
  const double sk = left ? sl : sr;
  const Prim s{S.d, S.p, left ? ul : ur, S.vy - 0.0 * velocity, S.e, S.c};
  const double un = s.vx * 1.0 + s.vy * 0.0;
  const double dun = s.d * un;
  const double es = ted(s);
  Cons f;
  if (branch == 0 || branch == 3) {
    f.m = s.d * un;
    f.mx = 1.0 * s.p + s.vx * dun;
    f.my = 0.0 * s.p + s.vy * dun;
    f.E = (es + s.p) * un;
  } else {
    const double ds = s.d * (sk - s.vx) / (sk - ss);  // starred_state, hllc.cpp:76-92
    const double usE = es * ds / s.d + ds * (ss - s.vx) * (ss + s.p / s.d / (sk - s.vx));
    f.m = s.d * un + sk * (ds - s.d);
    f.mx = (1.0 * s.p + s.vx * dun) + (ds * ss - s.vx * s.d) * sk;
    f.my = (0.0 * s.p + s.vy * dun) + (ds * s.vy - s.vy * s.d) * sk;
    f.E = (es + s.p) * un + sk * (usE - es);
  }
  f.E = f.E + (f.mx * (1.0 * velocity) + f.my * (0.0 * velocity) + 0.5 * f.m * velocity * velocity);
  const double vm = velocity * f.m;
  f.mx = f.mx + 1.0 * vm;
  f.my = f.my + 0.0 * vm;
  out = f;
  return 0;
}

}  // namespace rg
