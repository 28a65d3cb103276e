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
__device__ __forceinline__ double smin(double a, double b) { return (b < a) ? b : a; }
__device__ __forceinline__ double smax(double a, double b) { return (a < b) ? b : a; }
// abs(Vector2D), geometry.cpp:18-21
__device__ __forceinline__ double vabs(double x, double y) { return sqrt(x * x + y * y); }
// is_nan, misc/utils.cpp:3-8
__device__ __forceinline__ bool is_nan_ref(double x) {
  const int b1 = (x >= 0);
  const int b2 = (x < 0);
  return 1 != b1 + b2;
}

// IdealGas::dp2e / dp2c (ideal_gas.cpp:13-16, 40-44)
__device__ __forceinline__ double dp2e(double g, double d, double p) { return p / d / (g - 1); }
__device__ __forceinline__ double dp2c(double g, double d, double p) { return sqrt(g * p / d); }

// convert_to_primitive (simple_flux_calculator.cpp:9-15)
__device__ __forceinline__ Prim to_prim(const Cell& c, double g) {
  return Prim{c.d, c.p, c.vx, c.vy, dp2e(g, c.d, c.p), dp2c(g, c.d, c.p)};
}

// TotalEnergyDensity (hydrodynamic_variables.cpp:261-265): pow(abs(v),2) is
// evaluated as abs(v)*abs(v)
__device__ __forceinline__ double ted(const Prim& s) {
  const double a = vabs(s.vx, s.vy);
  return s.d * (0.5 * (a * a) + s.e);
}

// HLLC in the face frame, normal along x (hllc.cpp:27-154, massflux_=true).
// Returns 0 or an error code.
__device__ __forceinline__ int hllc(Prim l, Prim r, double velocity, Cons& out) {
  if (is_nan_ref(r.vx)) return ERR_HLLC_NAN;
  // Velocity -= velocity*normaldir with normaldir = (1,0)
  l.vx = l.vx - 1.0 * velocity;
  l.vy = l.vy - 0.0 * velocity;
  r.vx = r.vx - 1.0 * velocity;
  r.vy = r.vy - 0.0 * velocity;
  // estimate_wave_speeds (hllc.cpp:27-43)
  const double sl = smin(l.vx - l.c, r.vx - r.c);
  const double sr = smax(l.vx + l.c, r.vx + r.c);
  const double ss = (r.p - l.p + l.d * l.vx * (sl - l.vx) - r.d * r.vx * (sr - r.vx)) /
                    (l.d * (sl - l.vx) - r.d * (sr - r.vx));
  Cons f;
  // branch order of hllc.cpp:134-148
  const int branch = (sl > 0) ? 0 : (sl <= 0 && ss >= 0) ? 1 : (ss < 0 && sr >= 0) ? 2
                                                                : (sr < 0) ? 3 : 4;
  if (branch == 0 || branch == 3) {
    // supersonic: Primitive2Flux of the upwind side (hydrodynamic_variables.cpp:281-290)
    const Prim& s = branch == 0 ? l : r;
    const double un = s.vx * 1.0 + s.vy * 0.0;
    const double dun = s.d * un;
    f.m = s.d * un;
    f.mx = 1.0 * s.p + s.vx * dun;
    f.my = 0.0 * s.p + s.vy * dun;
    f.E = (ted(s) + s.p) * un;
  } else if (branch == 1) {
    // fl + sl*(usl - ul)
    const double un = l.vx * 1.0 + l.vy * 0.0;
    const double dun = l.d * un;
    const double e_l = ted(l);
    const double ds = l.d * (sl - l.vx) / (sl - ss);  // starred_state, hllc.cpp:76-92
    const double usE = e_l * ds / l.d + ds * (ss - l.vx) * (ss + l.p / l.d / (sl - l.vx));
    f.m = l.d * un + sl * (ds - l.d);
    f.mx = (1.0 * l.p + l.vx * dun) + (ds * ss - l.vx * l.d) * sl;
    f.my = (0.0 * l.p + l.vy * dun) + (ds * l.vy - l.vy * l.d) * sl;
    f.E = (e_l + l.p) * un + sl * (usE - e_l);
  } else if (branch == 2) {
    // fr + sr*(usr - ur)
    const double un = r.vx * 1.0 + r.vy * 0.0;
    const double dun = r.d * un;
    const double e_r = ted(r);
    const double ds = r.d * (sr - r.vx) / (sr - ss);
    const double usE = e_r * ds / r.d + ds * (ss - r.vx) * (ss + r.p / r.d / (sr - r.vx));
    f.m = r.d * un + sr * (ds - r.d);
    f.mx = (1.0 * r.p + r.vx * dun) + (ds * ss - r.vx * r.d) * sr;
    f.my = (0.0 * r.p + r.vy * dun) + (ds * r.vy - r.vy * r.d) * sr;
    f.E = (e_r + r.p) * un + sr * (usE - e_r);
  } else {
    return ERR_INVALID_WAVE_SPEEDS;
  }
  // back to the lab frame (hllc.cpp:150-152)
  f.E = f.E + (f.mx * (1.0 * velocity) + f.my * (0.0 * velocity) + 0.5 * f.m * velocity * velocity);
  const double vm = velocity * f.m;
  f.mx = f.mx + 1.0 * vm;
  f.my = f.my + 0.0 * vm;
  out = f;
  return 0;
}

// rotate (simple_flux_calculator.cpp:130-140); an = |n|, ap = |p|
__device__ __forceinline__ Prim rotate(const Prim& s, double nx, double ny, double an, double px,
                                       double py, double ap) {
  Prim r = s;
  r.vx = (s.vx * nx + s.vy * ny) / an;
  r.vy = (s.vx * px + s.vy * py) / ap;
  return r;
}

// rotate_back (simple_flux_calculator.cpp:142-151)
__device__ __forceinline__ Cons rotate_back(const Cons& c, double nx, double ny, double an,
                                            double px, double py, double ap) {
  Cons r = c;
  r.mx = nx * c.mx / an + px * c.my / ap;
  r.my = ny * c.mx / an + py * c.my / ap;
  return r;
}

// ComputationalCellAddMult (computational_cell_2d.cpp:59-68)
__device__ __forceinline__ void add_mult(Cell& r, const Cell& o, double s) {
  r.d += o.d * s;
  r.p += o.p * s;
  r.vx += o.vx * s;
  r.vy += o.vy * s;
}

// interp / interp2 (LinearGaussImproved.cpp:159-172)
__device__ __forceinline__ Cell interp(const Cell& c, const Cell& sx, const Cell& sy, double tx,
                                       double ty, double cmx, double cmy) {
  Cell r = c;
  add_mult(r, sx, tx - cmx);
  add_mult(r, sy, ty - cmy);
  return r;
}

}  // namespace rg
