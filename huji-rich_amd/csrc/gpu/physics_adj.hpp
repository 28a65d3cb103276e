// Reverse-mode (adjoint) derivatives of the per-face and per-cell physics
// of the step, for the fixed-topology adjoint on a static mesh (zero face
// velocities).  Each function recomputes the forward quantities it needs,
// takes the branch the forward values select (HLLC wave pattern, min/max
// choices), and accumulates the vector-Jacobian product of that branch into
// the adjoint arguments.  Where the forward evaluates |v| at v = 0 the
// derivative of |v| is taken as 0 (the central difference of |v| there).
#pragma once

#include "physics.hpp"

namespace rg {

// d|v|/dv, 0 at v = 0
__host__ __device__ __forceinline__ void dabs(double vx, double vy, double a, double& gx,
                                              double& gy) {
  gx = a > 0 ? vx / a : 0.0;
  gy = a > 0 ? vy / a : 0.0;
}

// Reverse of hllc_face with velocity = 0 (physics.hpp).  fb is the adjoint
// of the lab-frame flux (m, mx, my, E); adds the adjoints of the two face
// states to Lb, Rb.
__host__ __device__ inline void hllc_face_vjp(const Cell& L, const Cell& R, double g, double nx,
                                              double ny, double an, double px, double py,
                                              double ap, const Cons& fb, Cell& Lb, Cell& Rb) {
  // ---- forward
  const double cl = dp2c(g, L.d, L.p), cr = dp2c(g, R.d, R.p);
  const double ur = (R.vx * nx + R.vy * ny) / an;
  const double ul = (L.vx * nx + L.vy * ny) / an;
  const double a1 = ul - cl, a2 = ur - cr, b1 = ul + cl, b2 = ur + cr;
  const bool sl_right = a2 < a1;  // smin(a1, a2) = (a2 < a1) ? a2 : a1
  const bool sr_right = b1 < b2;  // smax(b1, b2) = (b1 < b2) ? b2 : b1
  const double sl = sl_right ? a2 : a1;
  const double sr = sr_right ? b2 : b1;
  const double num = R.p - L.p + L.d * ul * (sl - ul) - R.d * ur * (sr - ur);
  const double den = L.d * (sl - ul) - R.d * (sr - ur);
  const double ss = num / den;
  const int branch = (sl > 0) ? 0 : (sl <= 0 && ss >= 0) ? 1 : (ss < 0 && sr >= 0) ? 2 : 3;
  const bool left = branch <= 1;
  const Cell& S = left ? L : R;
  const double u = left ? ul : ur;
  const double sk = left ? sl : sr;
  const double v = (S.vx * px + S.vy * py) / ap;
  const double e = dp2e(g, S.d, S.p);
  const double dS = S.d, pS = S.p;
  const double amag = vabs(u, v);
  const double es = dS * (0.5 * (amag * amag) + e);
  // ---- reverse: rotate_back
  const double fxb = fb.mx * nx / an + fb.my * ny / an;
  const double fyb = fb.mx * px / ap + fb.my * py / ap;
  const double fmb = fb.m, fEb = fb.E;
  double dSb = 0, pSb = 0, ub = 0, vb = 0, skb = 0, ssb = 0, esb = 0;
  if (branch == 0 || branch == 3) {
    // fm = d u; fx = p + u (d u); fy = v (d u); fE = (es + p) u
    dSb += fmb * u + fxb * u * u + fyb * v * u;
    ub += fmb * dS + fxb * 2.0 * u * dS + fyb * v * dS + fEb * (es + pS);
    pSb += fxb + fEb * u;
    vb += fyb * dS * u;
    esb += fEb * u;
  } else {
    const double A = sk - u, B = sk - ss;
    const double ds = dS * A / B;
    const double Q = ss + pS / dS / A;
    const double usE = es * ds / dS + ds * (ss - u) * Q;
    double dsb = 0, usEb = 0, Ab = 0, Bb = 0;
    // fm = d u + sk (ds - d)
    dSb += fmb * (u - sk);
    ub += fmb * dS;
    skb += fmb * (ds - dS);
    dsb += fmb * sk;
    // fx = (p + u d u) + (ds ss - u d) sk
    pSb += fxb;
    ub += fxb * (2.0 * u * dS - dS * sk);
    dSb += fxb * (u * u - u * sk);
    dsb += fxb * ss * sk;
    ssb += fxb * ds * sk;
    skb += fxb * (ds * ss - u * dS);
    // fy = v d u + (ds v - v d) sk
    vb += fyb * (dS * u + (ds - dS) * sk);
    dSb += fyb * (v * u - v * sk);
    ub += fyb * v * dS;
    dsb += fyb * v * sk;
    skb += fyb * (ds * v - v * dS);
    // fE = (es + p) u + sk (usE - es)
    esb += fEb * (u - sk);
    pSb += fEb * u;
    ub += fEb * (es + pS);
    skb += fEb * (usE - es);
    usEb += fEb * sk;
    // usE = es ds / d + ds (ss - u) Q
    esb += usEb * ds / dS;
    dsb += usEb * (es / dS + (ss - u) * Q);
    dSb -= usEb * es * ds / (dS * dS);
    ssb += usEb * ds * Q;
    ub -= usEb * ds * Q;
    const double Qb = usEb * ds * (ss - u);
    // Q = ss + p / d / A
    ssb += Qb;
    pSb += Qb / (dS * A);
    dSb -= Qb * pS / (dS * dS * A);
    Ab -= Qb * pS / (dS * A * A);
    // ds = d A / B
    dSb += dsb * A / B;
    Ab += dsb * dS / B;
    Bb -= dsb * ds / B;
    // A = sk - u, B = sk - ss
    skb += Ab + Bb;
    ub -= Ab;
    ssb -= Bb;
  }
  // es = d (0.5 |(u,v)|^2 + e), |(u,v)|^2 differentiated as u^2 + v^2
  dSb += esb * (0.5 * (amag * amag) + e);
  ub += esb * dS * u;
  vb += esb * dS * v;
  const double eb = esb * dS;
  // e = p / d / (g - 1)
  pSb += eb / (dS * (g - 1));
  dSb -= eb * e / dS;
  // side states
  Cell& Sb = left ? Lb : Rb;
  Sb.d += dSb;
  Sb.p += pSb;
  Sb.vx += vb * px / ap;
  Sb.vy += vb * py / ap;
  double ulb = left ? ub : 0.0, urb = left ? 0.0 : ub;
  double slb = left ? skb : 0.0, srb = left ? 0.0 : skb;
  // ss = num / den
  const double numb = ssb / den, denb = -ssb * ss / den;
  // num = pR - pL + dL ul (sl - ul) - dR ur (sr - ur)
  Rb.p += numb;
  Lb.p -= numb;
  Lb.d += numb * ul * (sl - ul);
  ulb += numb * L.d * (sl - 2.0 * ul);
  slb += numb * L.d * ul;
  Rb.d -= numb * ur * (sr - ur);
  urb -= numb * R.d * (sr - 2.0 * ur);
  srb -= numb * R.d * ur;
  // den = dL (sl - ul) - dR (sr - ur)
  Lb.d += denb * (sl - ul);
  slb += denb * L.d;
  ulb -= denb * L.d;
  Rb.d -= denb * (sr - ur);
  srb -= denb * R.d;
  urb += denb * R.d;
  // wave-speed bounds
  double clb = 0, crb = 0;
  if (sl_right) {
    urb += slb;
    crb -= slb;
  } else {
    ulb += slb;
    clb -= slb;
  }
  if (sr_right) {
    urb += srb;
    crb += srb;
  } else {
    ulb += srb;
    clb += srb;
  }
  // normal velocities
  Lb.vx += ulb * nx / an;
  Lb.vy += ulb * ny / an;
  Rb.vx += urb * nx / an;
  Rb.vy += urb * ny / an;
  // c = sqrt(g p / d)
  Lb.p += clb * cl / (2.0 * L.p);
  Lb.d -= clb * cl / (2.0 * L.d);
  Rb.p += crb * cr / (2.0 * R.p);
  Rb.d -= crb * cr / (2.0 * R.d);
}

// Reverse of the primitive recovery of k_update (SimpleCellUpdater,
// simple_cell_updater.cpp:22-59): x = (m, E, Mx, My), cb the adjoint of
// (d, p, vx, vy); returns the adjoint of x.
__host__ __device__ __forceinline__ void recover_vjp(double gamma, double vol, double m, double E,
                                                     double Mx, double My, const Cell& cb,
                                                     double& mb, double& Eb, double& Mxb,
                                                     double& Myb) {
  const double d = m / vol;
  const double vx = Mx / m, vy = My / m;
  const double e = E / m - 0.5 * (vx * vx + vy * vy);
  const double eb = cb.p * (gamma - 1) * d;
  const double db = cb.d + cb.p * (gamma - 1) * e;
  const double vxb = cb.vx - eb * vx, vyb = cb.vy - eb * vy;
  Eb = eb / m;
  mb = db / vol - eb * E / (m * m) - (vxb * vx + vyb * vy) / m;
  Mxb = vxb / m;
  Myb = vyb / m;
}

// Reverse of init_extensives (hdsim2d.cpp:41-66): cells (d, p, vx, vy) ->
// (m, E, Mx, My); adds the adjoint of the cell to cb.
__host__ __device__ __forceinline__ void init_vjp(double gamma, double vol, const Cell& c,
                                                  double mb, double Eb, double Mxb, double Myb,
                                                  Cell& cb) {
  const double mass = vol * c.d;
  const double e = dp2e(gamma, c.d, c.p);
  const double k = c.vx * c.vx + c.vy * c.vy;
  // E = e m + 0.5 m k; Mx = vx m; My = vy m
  const double massb = mb + Eb * (e + 0.5 * k) + Mxb * c.vx + Myb * c.vy;
  const double ebar = Eb * mass;
  cb.d += massb * vol - ebar * e / c.d;
  cb.p += ebar / (c.d * (gamma - 1));
  cb.vx += Eb * mass * c.vx + Mxb * mass;
  cb.vy += Eb * mass * c.vy + Myb * mass;
}

// Reverse of the CFL inverse time step of one cell with zero face
// velocities: inv = width / (c + |v|) as 1 / ((c + |v|) / width).  Adds
// invb * d(inv)/d(cell) to cb.
__host__ __device__ __forceinline__ void cfl_inv_vjp(double gamma, double width, const Cell& c,
                                                     double invb, Cell& cb) {
  const double cs = dp2c(gamma, c.d, c.p);
  const double a = vabs(c.vx, c.vy);
  const double res = (cs + a) / width;
  const double resb = -invb / (res * res);
  const double csb = resb / width, ab = resb / width;
  double gx, gy;
  dabs(c.vx, c.vy, a, gx, gy);
  cb.vx += ab * gx;
  cb.vy += ab * gy;
  cb.p += csb * cs / (2.0 * c.p);
  cb.d -= csb * cs / (2.0 * c.d);
}

// Rigid-wall ghost velocity (I - 2 n n^T) v is symmetric: the adjoint
// reflects the same way.
__host__ __device__ __forceinline__ void reflect_vec(double nx, double ny, double& vx, double& vy) {
  const double s2 = 2 * (vx * nx + vy * ny);
  vx = vx - nx * s2;
  vy = vy - ny * s2;
}

}  // namespace rg
