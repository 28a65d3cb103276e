/*
 * tangent_real.hpp -- forward-mode dual numbers for the oracle's tangent
 * build (test infrastructure, like the rest of oracle/).
 *
 * A `real` carries a value and RO_K directional derivatives.  The value
 * part of every operation is the plain double operation, so a tangent run
 * reproduces the double oracle's values bit for bit (-ffp-contract=off).
 * Comparisons, min/max and branches act on values only: the derivative is
 * that of the branch the values select (the "active branch").
 */
#ifndef RICH_TANGENT_REAL_HPP
#define RICH_TANGENT_REAL_HPP

#include <math.h>

#ifndef RO_K
#define RO_K 8
#endif

struct real {
  double v;
  double t[RO_K];
  real() = default;
  real(double x) : v(x) {
    for (int k = 0; k < RO_K; ++k) t[k] = 0.0;
  }
  real& operator+=(const real& b) {
    v = v + b.v;
    for (int k = 0; k < RO_K; ++k) t[k] = t[k] + b.t[k];
    return *this;
  }
  real& operator-=(const real& b) {
    v = v - b.v;
    for (int k = 0; k < RO_K; ++k) t[k] = t[k] - b.t[k];
    return *this;
  }
  real& operator*=(const real& b) {
    for (int k = 0; k < RO_K; ++k) t[k] = t[k] * b.v + v * b.t[k];
    v = v * b.v;
    return *this;
  }
};

inline real operator-(const real& a) {
  real r;
  r.v = -a.v;
  for (int k = 0; k < RO_K; ++k) r.t[k] = -a.t[k];
  return r;
}
inline real operator+(const real& a, const real& b) {
  real r;
  r.v = a.v + b.v;
  for (int k = 0; k < RO_K; ++k) r.t[k] = a.t[k] + b.t[k];
  return r;
}
inline real operator-(const real& a, const real& b) {
  real r;
  r.v = a.v - b.v;
  for (int k = 0; k < RO_K; ++k) r.t[k] = a.t[k] - b.t[k];
  return r;
}
inline real operator*(const real& a, const real& b) {
  real r;
  r.v = a.v * b.v;
  for (int k = 0; k < RO_K; ++k) r.t[k] = a.t[k] * b.v + a.v * b.t[k];
  return r;
}
inline real operator/(const real& a, const real& b) {
  real r;
  r.v = a.v / b.v;
  for (int k = 0; k < RO_K; ++k) r.t[k] = (a.t[k] - r.v * b.t[k]) / b.v;
  return r;
}
// double on one side: the double is a constant
inline real operator+(const real& a, double b) {
  real r = a;
  r.v = a.v + b;
  return r;
}
inline real operator+(double a, const real& b) {
  real r = b;
  r.v = a + b.v;
  return r;
}
inline real operator-(const real& a, double b) {
  real r = a;
  r.v = a.v - b;
  return r;
}
inline real operator-(double a, const real& b) {
  real r;
  r.v = a - b.v;
  for (int k = 0; k < RO_K; ++k) r.t[k] = -b.t[k];
  return r;
}
inline real operator*(const real& a, double b) {
  real r;
  r.v = a.v * b;
  for (int k = 0; k < RO_K; ++k) r.t[k] = a.t[k] * b;
  return r;
}
inline real operator*(double a, const real& b) {
  real r;
  r.v = a * b.v;
  for (int k = 0; k < RO_K; ++k) r.t[k] = a * b.t[k];
  return r;
}
inline real operator/(const real& a, double b) {
  real r;
  r.v = a.v / b;
  for (int k = 0; k < RO_K; ++k) r.t[k] = a.t[k] / b;
  return r;
}
inline real operator/(double a, const real& b) {
  real r;
  r.v = a / b.v;
  for (int k = 0; k < RO_K; ++k) r.t[k] = -r.v * b.t[k] / b.v;
  return r;
}

inline bool operator<(const real& a, const real& b) { return a.v < b.v; }
inline bool operator>(const real& a, const real& b) { return a.v > b.v; }
inline bool operator<=(const real& a, const real& b) { return a.v <= b.v; }
inline bool operator>=(const real& a, const real& b) { return a.v >= b.v; }
inline bool operator<(const real& a, double b) { return a.v < b; }
inline bool operator>(const real& a, double b) { return a.v > b; }
inline bool operator<=(const real& a, double b) { return a.v <= b; }
inline bool operator>=(const real& a, double b) { return a.v >= b; }
inline bool operator<(double a, const real& b) { return a < b.v; }
inline bool operator>(double a, const real& b) { return a > b.v; }

// At 0 the state only reaches sqrt as |v| = sqrt(vx^2 + vy^2) with a zero
// tangent; the derivative is taken as 0 there (the central difference of
// |v| at v = 0, and the exact derivative of |v|^2).
inline real sqrt(const real& a) {
  real r;
  r.v = ::sqrt(a.v);
  for (int k = 0; k < RO_K; ++k) r.t[k] = a.t[k] == 0.0 ? 0.0 : a.t[k] / (2.0 * r.v);
  return r;
}
inline real fabs(const real& a) { return a.v < 0 ? -a : a; }
// fmax as the C library returns it for ordered operands
inline real fmax(const real& a, const real& b) { return (a.v < b.v) ? b : a; }

#define RO_VAL(x) ((x).v)
static real vabs_r(const real& x, const real& y) { return sqrt(x * x + y * y); }

#endif
