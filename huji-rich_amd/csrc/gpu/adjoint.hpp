// Adjoint (reverse mode) of the fixed-topology FP64 step on a static mesh
// with zero face velocities: hdsim::TimeAdvance2Heun with LinearGaussImproved
// (+ slope limiter), ModularFluxCalculator + Hllc, SimpleExtensiveUpdater,
// SimpleCellUpdater and SimpleCFL.
//
// Each kernel recomputes the forward quantities it needs from the state the
// forward pass saved, takes the branches the forward values select (HLLC
// wave pattern, limiter choices, shock switch, the CFL argmin) and
// accumulates the vector-Jacobian product of that branch.  Contributions to
// other cells are written per face or per incidence and gathered by their
// receiver in a fixed order, so the adjoint is deterministic and uses no
// atomics.
#pragma once

#include "kernels.hpp"
#include "physics_adj.hpp"

namespace rg {

// Adjoint-only incidence tables (built once per mesh).
struct AdjMesh {
  const int* rev;        // [ne] ELL index of the same face seen from the neighbour (-1: ghost)
  const int* gf_off;     // [N+1] faces whose ghost side is a copy of this cell
  const unsigned* gf;    // face | INFO_SECOND (the ghost is the second side)
  const int* ge_off;     // [N+1] ELL entries whose neighbour is a ghost copying this cell
  const int* ge;         // ELL index
};

__device__ __forceinline__ D4 cell_d4(const Cell& c) { return D4{c.d, c.p, c.vx, c.vy}; }
__device__ __forceinline__ Cell d4_cell(const D4& v) { return Cell{v.a, v.b, v.c, v.d}; }
__device__ __forceinline__ void acc(Cell& r, const Cell& a) {
  r.d += a.d;
  r.p += a.p;
  r.vx += a.vx;
  r.vy += a.vy;
}
__device__ __forceinline__ void acc_scaled(Cell& r, const Cell& a, double s) {
  r.d += a.d * s;
  r.p += a.p * s;
  r.vx += a.vx * s;
  r.vy += a.vy * s;
}

// Deterministic block sum (fixed tree) of one double per thread into
// out[blockIdx.x]; blockDim.x == 256.
__device__ __forceinline__ void block_sum(double v, double* out) {
  __shared__ double sh[256];
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (static_cast<int>(threadIdx.x) < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = sh[0];
}

// The dt of a step, dt = cfl * min_i inv_i(P), has its adjoint applied one
// kernel late, by the next kernel that reads the step's cell adjoint: the
// face shares of d(loss)/d(dt) (dtb, both stages) and the per-block argmin
// partials are reduced by every block in the same fixed order.
struct PendingDt {
  const D4* q;  // cells the dt was computed from (null: none pending)
  const double* dtb;
  int ndtb;
  const double* vmin;
  const int* imin;
  int nmin;
  double cfl;
};

// Block-cooperative (every thread of the block calls it): the argmin cell
// and the adjoint it receives.
__device__ __forceinline__ void pending_dt(const DMesh& m, const PendingDt& pd, double gamma,
                                           int& cell, Cell& cb) {
  __shared__ double ssum[256];
  __shared__ double sv[256];
  __shared__ int si[256];
  double acc_v = 0;
  for (int k = threadIdx.x; k < pd.ndtb; k += 256) acc_v += pd.dtb[k];
  double v = __longlong_as_double(static_cast<long long>(DBL_INF_BITS));
  int ix = 0x7fffffff;
  for (int k = threadIdx.x; k < pd.nmin; k += 256) {
    if (pd.vmin[k] < v || (pd.vmin[k] == v && pd.imin[k] < ix)) {
      v = pd.vmin[k];
      ix = pd.imin[k];
    }
  }
  ssum[threadIdx.x] = acc_v;
  sv[threadIdx.x] = v;
  si[threadIdx.x] = ix;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (static_cast<int>(threadIdx.x) < s) {
      ssum[threadIdx.x] += ssum[threadIdx.x + s];
      const double o = sv[threadIdx.x + s];
      const int oi = si[threadIdx.x + s];
      if (o < sv[threadIdx.x] || (o == sv[threadIdx.x] && oi < si[threadIdx.x])) {
        sv[threadIdx.x] = o;
        si[threadIdx.x] = oi;
      }
    }
    __syncthreads();
  }
  cell = si[0];
  cb = Cell{0, 0, 0, 0};
  if (cell >= 0 && cell < m.N)
    cfl_inv_vjp(gamma, m.width[cell], ld_cell(pd.q, cell), pd.cfl * ssum[0], cb);
}

// Reverse of P_{n+1} = recover(X_{n+1}) and of the Heun average
// X_{n+1} = 0.5 (X_tmp + X_n): ga = gb = 0.5 (gX + recover^T gP), after the
// pending dt adjoint of step n+1 is added to gP.
__global__ void __launch_bounds__(256) k_adj_recover(DMesh m, const D4* __restrict__ X,
                                                     const D4* __restrict__ gP,
                                                     const D4* __restrict__ gX,
                                                     D4* __restrict__ ga, D4* __restrict__ gb,
                                                     double gamma, PendingDt pd) {
  int dcell = -1;
  Cell dcb{0, 0, 0, 0};
  if (pd.q) pending_dt(m, pd, gamma, dcell, dcb);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m.N) return;
  const D4 x = X[i], g = gX[i];
  Cell cb = d4_cell(gP[i]);
  if (i == dcell) acc(cb, dcb);
  double mb, Eb, Mxb, Myb;
  recover_vjp(gamma, m.vol[i], x.a, x.b, x.c, x.d, cb, mb, Eb, Mxb, Myb);
  const D4 h{0.5 * (g.a + mb), 0.5 * (g.b + Eb), 0.5 * (g.c + Mxb), 0.5 * (g.d + Myb)};
  ga[i] = h;
  gb[i] = h;
}

// Per face: the flux adjoint from the update X_out = X_in -/+ dt A F (seed
// gs = adjoint of X_out, real sides only), the HLLC reverse into the two
// face states, and this face's share of d(loss)/d(dt).
__global__ void __launch_bounds__(256) k_adj_flux(DMesh m, const D4* __restrict__ q,
                                                  const Rec* __restrict__ rec,
                                                  const D4* __restrict__ gs,
                                                  const double* __restrict__ dtp, double cfl,
                                                  D4* __restrict__ lb_out, D4* __restrict__ rb_out,
                                                  double* __restrict__ dtb_partials, double gamma) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  double dtb = 0;
  if (f < m.F) {
    const double dt = *dtp;
    const FaceFrame fr = m.fframe[f];
    const int a = m.ff[f], b = m.fs[f];
    Cell L, R;
    face_states(m, q, rec, a, b, fr, L, R);
    const double an = vabs(fr.nx, fr.ny), ap = vabs(fr.px, fr.py);
    Cons c;
    (void)hllc_face(L, R, gamma, fr.nx, fr.ny, an, fr.px, fr.py, ap, 0.0, c);
    const double area = m.area[f];
    const D4 z{0, 0, 0, 0};
    const D4 ga = a < m.N ? gs[a] : z, gb = b < m.N ? gs[b] : z;
    const D4 dg{gb.a - ga.a, gb.b - ga.b, gb.c - ga.c, gb.d - ga.d};
    const double s = dt * area;
    // extensive order (m, E, Mx, My); Cons order (m, mx, my, E)
    const Cons fb{s * dg.a, s * dg.c, s * dg.d, s * dg.b};
    dtb = area * (c.m * dg.a + c.E * dg.b + c.mx * dg.c + c.my * dg.d);
    Cell Lb{0, 0, 0, 0}, Rb{0, 0, 0, 0};
    hllc_face_vjp(L, R, gamma, fr.nx, fr.ny, an, fr.px, fr.py, ap, fb, Lb, Rb);
    lb_out[f] = cell_d4(Lb);
    rb_out[f] = cell_d4(Rb);
  }
  (void)cfl;
  block_sum(dtb, dtb_partials);
}

// Adjoints of cell i's value (v) and slopes (sx, sy) from the face states
// it produced (its own faces, and the ghost sides that copy it).
__device__ __forceinline__ void cell_state_adj(const DMesh& m, const AdjMesh& am, const D4* lb,
                                               const D4* rb, int i, int deg, int eb, Cell& v,
                                               Cell& sx, Cell& sy) {
  v = sx = sy = Cell{0, 0, 0, 0};
  // CH incidences at a time: their face words and offsets, then their face
  // states' adjoints, each batch of loads issued together; sums in slot order
  for (int s0 = 0; s0 < deg; s0 += CH) {
    unsigned info[CH];
    D2 o[CH];
    D4 gv[CH];
#pragma unroll
    for (int u = 0; u < CH; ++u) info[u] = m.einfo[eb + 64 * min(s0 + u, deg - 1)];
#pragma unroll
    for (int u = 0; u < CH; ++u) o[u] = m.eo[eb + 64 * min(s0 + u, deg - 1)];
#pragma unroll
    for (int u = 0; u < CH; ++u) {
      const int f = static_cast<int>(info[u] & INFO_FACE);
      gv[u] = (info[u] & INFO_SECOND) ? rb[f] : lb[f];
    }
#pragma unroll
    for (int u = 0; u < CH; ++u) {
      if (s0 + u >= deg) break;
      const Cell g = d4_cell(gv[u]);
      acc(v, g);
      acc_scaled(sx, g, o[u].x);
      acc_scaled(sy, g, o[u].y);
    }
  }
  for (int k = am.gf_off[i]; k < am.gf_off[i + 1]; ++k) {
    const unsigned w = am.gf[k];
    const int f = static_cast<int>(w & INFO_FACE);
    const bool second = (w & INFO_SECOND) != 0;
    Cell g = d4_cell(second ? rb[f] : lb[f]);
    const int gid = second ? m.fs[f] : m.ff[f];
    if (m.bc == BC_PERIODIC) {
      // ghost state = the original's value and slope at the ghost's CM
      const FaceRec r = m.frec[f];
      const D2 cg = m.cm[gid];
      acc(v, g);
      acc_scaled(sx, g, r.cx - cg.x);
      acc_scaled(sy, g, r.cy - cg.y);
    } else {
      // ghost state = the reflected source value, zero slope
      const D2 n = m.gn[gid - m.N];
      reflect_vec(n.x, n.y, g.vx, g.vy);
      acc(v, g);
    }
  }
}

// one field of the limiter with the choice recorded: psi = min(psi, ratio)
struct LimPick {
  double psi = 1;
  int slot = -1;   // incidence of the active ratio (-1: psi is the initial 1)
  int kind = 0;    // 0: maxd/dphi, 1: mind/dphi, 2: theta (nval - cell)/dphi
  double dphi = 0, ratio = 0;
};
__device__ __forceinline__ void lim_field(double dphi, double cval, double cell, double maxd,
                                          double mind, bool absguard, int s, LimPick& lp) {
  if (fabs(dphi) > 0.05 * smax(fabs(maxd), fabs(mind)) || cval * cell < 0) {
    const double hi = absguard ? fabs(1e-9 * cell) : 1e-9 * cell;
    const double lo = absguard ? -fabs(1e-9 * cell) : -1e-9 * cell;
    if (dphi > hi) {
      const double r = maxd / dphi;
      if (r < lp.psi) lp = LimPick{r, s, 0, dphi, r};
    } else if (dphi < lo) {
      const double r = mind / dphi;
      if (r < lp.psi) lp = LimPick{r, s, 1, dphi, r};
    }
  }
}
__device__ __forceinline__ void shocked_lim_field(double dphi, double cval, double cell,
                                                  double nval, double maxd, double mind,
                                                  double theta, int s, LimPick& lp) {
  if (fabs(dphi) > 0.1 * smax(fabs(maxd), fabs(mind)) || cval * cell < 0) {
    if (fabs(dphi) > 1e-9 * cell) {
      const double t = theta * (nval - cell) / dphi;
      const double r = smax(t, 0.0);
      // the clipped value 0 is a constant: no derivative
      if (r < lp.psi) lp = LimPick{r, (t < 0.0) ? -1 : s, 2, dphi, r};
    }
  }
}

__device__ __forceinline__ double fld(const Cell& c, int k) {
  return k == 0 ? c.d : k == 1 ? c.p : k == 2 ? c.vx : c.vy;
}
__device__ __forceinline__ void fld_add(Cell& c, int k, double v) {
  if (k == 0) c.d += v;
  else if (k == 1) c.p += v;
  else if (k == 2) c.vx += v;
  else c.vy += v;
}

// Per cell: the adjoint of the face states it produced (cell_state_adj),
// then the reverse of calc_slope (LinearGaussImproved.cpp:425-494) for the
// slope part: writes the cell's own share to gps and each incidence's share
// for the neighbour to emit[ELL entry].
__global__ void __launch_bounds__(256) k_adj_slopes(DMesh m, AdjMesh am, const D4* __restrict__ q,
                                                    const D4* __restrict__ lbf,
                                                    const D4* __restrict__ rbf,
                                                    D4* __restrict__ gps, D4* __restrict__ emit,
                                                    DParams prm) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m.N) return;
  const int deg = m.off[i + 1] - m.off[i], eb = ell_base(m, i);
  Cell vstate, gx, gy;
  cell_state_adj(m, am, lbf, rbf, i, deg, eb, vstate, gx, gy);
  const Cell cell = ld_cell(q, i);
  const CellGeo g = m.cgeo[i];
  const int kind = g.kind;
  const D2 cmi = m.cm[i];
  if (kind == KIND_TOO_MANY || kind == KIND_SINGULAR) {
    for (int s = 0; s < deg; ++s) emit[eb + 64 * s] = D4{0, 0, 0, 0};
    gps[i] = cell_d4(vstate);
    return;
  }
  const bool boundary = kind == KIND_BOUNDARY;
  // ---- forward: unlimited slope and neighbour extrema (with their incidence)
  Cell sx{0, 0, 0, 0}, sy{0, 0, 0, 0};
  double pr = 1;
  Cell amax = cell, amin = cell, bmax = cell, bmin = cell;
  int iamax[4] = {-1, -1, -1, -1}, iamin[4] = {-1, -1, -1, -1};
  int ibmax[4] = {-1, -1, -1, -1}, ibmin[4] = {-1, -1, -1, -1};
  Cell phy{0, 0, 0, 0}, phx{0, 0, 0, 0}, vcx{0, 0, 0, 0}, vcy{0, 0, 0, 0};
  // the gathers of CH slots at a time, issued back to back (fetch_k)
  for (int s0 = 0; s0 < deg; s0 += CH) {
    int nbk[CH];
    Cell nck[CH];
#pragma unroll
    for (int u = 0; u < CH; ++u) nbk[u] = m.enbr[eb + 64 * min(s0 + u, deg - 1)];
    fetch_k(m, q, nbk, nck);
#pragma unroll
    for (int u = 0; u < CH; ++u) {
      const int s = s0 + u;
      if (s >= deg) break;
      const int nb = nbk[u];
      const Cell nc = nck[u];
      if (boundary) {
        const D2 cj = m.cm[nb];
        const double ey = cj.y - cmi.y, ex = cj.x - cmi.x;
        const Cell dd = sub(nc, cell);
        phy.d += dd.d * ey;
        phy.p += dd.p * ey;
        phy.vx += dd.vx * ey;
        phy.vy += dd.vy * ey;
        phx.d += dd.d * ex;
        phx.p += dd.p * ex;
        phx.vx += dd.vx * ex;
        phx.vy += dd.vy * ex;
      } else {
        const D2 h = m.eh[eb + 64 * s];
        if (s == 0) {
          vcx = Cell{cell.d * h.x, cell.p * h.x, cell.vx * h.x, cell.vy * h.x};
          vcy = Cell{cell.d * h.y, cell.p * h.y, cell.vx * h.y, cell.vy * h.y};
        } else {
          add_mult(vcx, cell, h.x);
          add_mult(vcy, cell, h.y);
        }
        add_mult(vcy, nc, h.y);
        add_mult(vcx, nc, h.x);
      }
      pr = (cell.p > nc.p) ? smin(pr, nc.p / cell.p) : smin(pr, cell.p / nc.p);
      const bool mirror = (m.einfo[eb + 64 * s] & INFO_MIRROR) != 0;
      for (int k = 0; k < 4; ++k) {
        const double t = fld(nc, k);
        double* ax = k == 0 ? &amax.d : k == 1 ? &amax.p : k == 2 ? &amax.vx : &amax.vy;
        double* an = k == 0 ? &amin.d : k == 1 ? &amin.p : k == 2 ? &amin.vx : &amin.vy;
        if (*ax < t) { *ax = t; iamax[k] = s; }
        if (t < *an) { *an = t; iamin[k] = s; }
        if (!(boundary && mirror)) {
          double* bx = k == 0 ? &bmax.d : k == 1 ? &bmax.p : k == 2 ? &bmax.vx : &bmax.vy;
          double* bn = k == 0 ? &bmin.d : k == 1 ? &bmin.p : k == 2 ? &bmin.vx : &bmin.vy;
          if (*bx < t) { *bx = t; ibmax[k] = s; }
          if (t < *bn) { *bn = t; ibmin[k] = s; }
        }
      }
    }
  }
  if (boundary) {
    const double SxSy = g.c0, Sx2 = g.c1, Sy2 = g.c2, den = g.c3;
    sx = Cell{(phy.d * SxSy - phx.d * Sy2) / den, (phy.p * SxSy - phx.p * Sy2) / den,
              (phy.vx * SxSy - phx.vx * Sy2) / den, (phy.vy * SxSy - phx.vy * Sy2) / den};
    sy = Cell{(phx.d * SxSy - phy.d * Sx2) / den, (phx.p * SxSy - phy.p * Sx2) / den,
              (phx.vx * SxSy - phy.vx * Sx2) / den, (phx.vy * SxSy - phy.vy * Sx2) / den};
  } else {
    sx = Cell{vcx.d * g.c0, vcx.p * g.c0, vcx.vx * g.c0, vcx.vy * g.c0};
    sy = Cell{vcy.d * g.c3, vcy.p * g.c3, vcy.vx * g.c3, vcy.vy * g.c3};
    add_mult(sx, vcy, g.c1);
    add_mult(sy, vcx, g.c2);
  }
  // ---- forward: limiter choices
  LimPick lp[4];
  bool shock = false;
  if (prm.slf) {
    const double cs = dp2c(prm.gamma, cell.d, cell.p);
    shock = ((sx.vx + sy.vy) * g.width < (-prm.delta_v) * cs) || (pr < prm.delta_p);
    const Cell maxd = sub(shock ? amax : bmax, cell), mind = sub(shock ? amin : bmin, cell);
    for (int s0 = 0; s0 < deg; s0 += CH) {
      Cell nck[CH];
      if (shock) {
        int nbk[CH];
#pragma unroll
        for (int u = 0; u < CH; ++u) nbk[u] = m.enbr[eb + 64 * min(s0 + u, deg - 1)];
        fetch_k(m, q, nbk, nck);
      }
#pragma unroll
      for (int u = 0; u < CH; ++u) {
        const int s = s0 + u;
        if (s >= deg) break;
        const D2 o = m.eo[eb + 64 * s];
        const Cell cv = interp_off(cell, sx, sy, o.x, o.y);
        const Cell dphi = sub(cv, cell);
        if (!shock) {
          lim_field(dphi.d, cv.d, cell.d, maxd.d, mind.d, false, s, lp[0]);
          lim_field(dphi.p, cv.p, cell.p, maxd.p, mind.p, false, s, lp[1]);
          lim_field(dphi.vx, cv.vx, cell.vx, maxd.vx, mind.vx, true, s, lp[2]);
          lim_field(dphi.vy, cv.vy, cell.vy, maxd.vy, mind.vy, true, s, lp[3]);
        } else {
          const Cell nc = nck[u];
          shocked_lim_field(dphi.d, cv.d, cell.d, nc.d, maxd.d, mind.d, prm.theta, s, lp[0]);
          shocked_lim_field(dphi.p, cv.p, cell.p, nc.p, maxd.p, mind.p, prm.theta, s, lp[1]);
          shocked_lim_field(dphi.vx, cv.vx, cell.vx, nc.vx, maxd.vx, mind.vx, prm.theta, s, lp[2]);
          shocked_lim_field(dphi.vy, cv.vy, cell.vy, nc.vy, maxd.vy, mind.vy, prm.theta, s, lp[3]);
        }
      }
    }
  }
  // ---- reverse: limited slope = psi * unlimited slope (per field)
  Cell gsx, gsy, self = vstate;
  int tgt[4] = {-1, -1, -1, -1};
  double tval[4] = {0, 0, 0, 0};
  for (int k = 0; k < 4; ++k) {
    const double psi = lp[k].psi;
    double bx = fld(gx, k) * psi, by = fld(gy, k) * psi;
    if (lp[k].slot >= 0) {
      const double psib = fld(gx, k) * fld(sx, k) + fld(gy, k) * fld(sy, k);
      const double dphi = lp[k].dphi;
      const double numb = psib / dphi, dphib = -psib * lp[k].ratio / dphi;
      const D2 o = m.eo[eb + 64 * lp[k].slot];
      bx += dphib * o.x;
      by += dphib * o.y;
      int src;
      double w;
      if (lp[k].kind == 2) {
        src = lp[k].slot;  // theta (nval - cell)
        w = prm.theta * numb;
      } else {
        const bool mx = lp[k].kind == 0;
        src = shock ? (mx ? iamax[k] : iamin[k]) : (mx ? ibmax[k] : ibmin[k]);
        w = numb;  // (extremum - cell)
      }
      if (src >= 0) {
        tgt[k] = src;
        tval[k] = w;
        fld_add(self, k, -w);
      }
    }
    if (k == 0) { gsx.d = bx; gsy.d = by; }
    else if (k == 1) { gsx.p = bx; gsy.p = by; }
    else if (k == 2) { gsx.vx = bx; gsy.vx = by; }
    else { gsx.vy = bx; gsy.vy = by; }
  }
  // ---- reverse: unlimited slope
  Cell wa, wb;  // naive: per-slot weight = wa * hx + wb * hy; boundary: wa * ex + wb * ey
  if (boundary) {
    const double SxSy = g.c0, Sx2 = g.c1, Sy2 = g.c2, den = g.c3;
    wb = Cell{(gsx.d * SxSy - gsy.d * Sx2) / den, (gsx.p * SxSy - gsy.p * Sx2) / den,
              (gsx.vx * SxSy - gsy.vx * Sx2) / den, (gsx.vy * SxSy - gsy.vy * Sx2) / den};
    wa = Cell{(gsy.d * SxSy - gsx.d * Sy2) / den, (gsy.p * SxSy - gsx.p * Sy2) / den,
              (gsy.vx * SxSy - gsx.vx * Sy2) / den, (gsy.vy * SxSy - gsx.vy * Sy2) / den};
  } else {
    wa = Cell{gsx.d * g.c0 + gsy.d * g.c2, gsx.p * g.c0 + gsy.p * g.c2,
              gsx.vx * g.c0 + gsy.vx * g.c2, gsx.vy * g.c0 + gsy.vy * g.c2};
    wb = Cell{gsx.d * g.c1 + gsy.d * g.c3, gsx.p * g.c1 + gsy.p * g.c3,
              gsx.vx * g.c1 + gsy.vx * g.c3, gsx.vy * g.c1 + gsy.vy * g.c3};
  }
  for (int s = 0; s < deg; ++s) {
    double ax, ay;
    if (boundary) {
      const D2 cj = m.cm[m.enbr[eb + 64 * s]];
      ax = cj.x - cmi.x;
      ay = cj.y - cmi.y;
    } else {
      const D2 h = m.eh[eb + 64 * s];
      ax = h.x;
      ay = h.y;
    }
    Cell e{wa.d * ax + wb.d * ay, wa.p * ax + wb.p * ay, wa.vx * ax + wb.vx * ay,
           wa.vy * ax + wb.vy * ay};
    if (boundary) {
      self.d -= e.d;
      self.p -= e.p;
      self.vx -= e.vx;
      self.vy -= e.vy;
    } else {
      acc(self, e);
    }
    for (int k = 0; k < 4; ++k)
      if (tgt[k] == s) fld_add(e, k, tval[k]);
    emit[eb + 64 * s] = cell_d4(e);
  }
  gps[i] = cell_d4(self);
}

// Per cell: total adjoint of the stage's cell value (own share, the shares
// its neighbours emitted for it, and those emitted for ghosts that copy it),
// then either the recovery reverse (mode 2: ga += recover^T gp, X = the
// stage's extensives) or the step-start outputs (mode 1: gP = gp,
// gX = gb + ga) with the per-block CFL argmin of the step-start cells q.
__global__ void __launch_bounds__(256) k_adj_gather(DMesh m, AdjMesh am,
                                                    const D4* __restrict__ gps,
                                                    const D4* __restrict__ emit, int mode,
                                                    const D4* __restrict__ X, D4* ga,
                                                    const D4* __restrict__ gb, D4* gP, D4* gX,
                                                    const D4* __restrict__ q, double* vmin,
                                                    int* imin, double gamma) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double inv = __longlong_as_double(static_cast<long long>(DBL_INF_BITS));
  int ix = 0x7fffffff;
  if (i < m.N) {
    Cell v = d4_cell(gps[i]);
    const int deg = m.off[i + 1] - m.off[i], eb = ell_base(m, i);
    // CH incidences at a time: the ELL entries of their twins, then the
    // shares emitted there (entry 0 stands in for a ghost side, unused)
    for (int s0 = 0; s0 < deg; s0 += CH) {
      int r[CH];
      D4 e[CH];
#pragma unroll
      for (int u = 0; u < CH; ++u) r[u] = am.rev[eb + 64 * min(s0 + u, deg - 1)];
#pragma unroll
      for (int u = 0; u < CH; ++u) e[u] = emit[r[u] >= 0 ? r[u] : 0];
#pragma unroll
      for (int u = 0; u < CH; ++u) {
        if (s0 + u >= deg) break;
        if (r[u] >= 0) acc(v, d4_cell(e[u]));
      }
    }
    for (int k = am.ge_off[i]; k < am.ge_off[i + 1]; ++k) {
      const int idx = am.ge[k];
      Cell e = d4_cell(emit[idx]);
      if (m.bc == BC_RIGID) {
        const D2 n = m.gn[m.enbr[idx] - m.N];
        reflect_vec(n.x, n.y, e.vx, e.vy);
      }
      acc(v, e);
    }
    if (mode == 2) {
      const D4 x = X[i];
      double mb, Eb, Mxb, Myb;
      recover_vjp(gamma, m.vol[i], x.a, x.b, x.c, x.d, v, mb, Eb, Mxb, Myb);
      const D4 s = ga[i];
      ga[i] = D4{s.a + mb, s.b + Eb, s.c + Mxb, s.d + Myb};
    } else {
      gP[i] = cell_d4(v);
      const D4 a = ga[i], b = gb[i];
      gX[i] = D4{a.a + b.a, a.b + b.b, a.c + b.c, a.d + b.d};
      inv = cfl_inv(m, i, gamma, m.width[i], ld_cell(q, i));
      ix = i;
    }
  }
  if (mode != 1) return;  // uniform across the grid
  __shared__ double sv[256];
  __shared__ int si[256];
  sv[threadIdx.x] = inv;
  si[threadIdx.x] = ix;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (static_cast<int>(threadIdx.x) < s) {
      const double o = sv[threadIdx.x + s];
      const int oi = si[threadIdx.x + s];
      if (o < sv[threadIdx.x] || (o == sv[threadIdx.x] && oi < si[threadIdx.x])) {
        sv[threadIdx.x] = o;
        si[threadIdx.x] = oi;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    vmin[blockIdx.x] = sv[0];
    imin[blockIdx.x] = si[0];
  }
}

// Reverse of init_extensives, gP += init^T gX, after the pending dt adjoint
// of step 0 is added.
__global__ void __launch_bounds__(256) k_adj_init(DMesh m, const D4* __restrict__ q,
                                                  const D4* __restrict__ gX, D4* gP,
                                                  double gamma, PendingDt pd) {
  int dcell = -1;
  Cell dcb{0, 0, 0, 0};
  if (pd.q) pending_dt(m, pd, gamma, dcell, dcb);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m.N) return;
  const D4 x = gX[i];
  Cell cb = d4_cell(gP[i]);
  if (i == dcell) acc(cb, dcb);
  init_vjp(gamma, m.vol[i], ld_cell(q, i), x.a, x.b, x.c, x.d, cb);
  gP[i] = cell_d4(cb);
}

}  // namespace rg
