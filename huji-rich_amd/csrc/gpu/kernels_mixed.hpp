// Mixed-precision step: FP32 reconstruction (LinearGaussImproved) and FP32
// rotated HLLC; FP64 extensives, accumulated over faces with Neumaier
// compensation, and FP64 primitive recovery (pressure from total energy).
//
// FP32 operands are rounded from FP64 quantities that are formed before any
// cancellation: face-centroid offsets are taken relative to each side's
// centre of mass in FP64, face normals and tangents are unit vectors, and
// face area is folded into the flux, so no FP32 subtraction of absolute
// coordinates occurs.  The FP32 kernels use the GPU's fast reciprocal and
// square root.
#pragma once

#include "kernels.hpp"

namespace rg {

struct alignas(16) F4 {
  float a, b, c, d;
};
// per face: unit normal n/|n|, area, and the face centroid minus each side's
// CM.  The unit tangent is the normal turned by +90 degrees: a Voronoi face
// is perpendicular to the segment joining its two mesh points, and the sign
// of the tangent cancels between rotate and rotate_back.
struct alignas(32) FaceM {
  float nx, ny, area, pad, oxa, oya, oxb, oyb;
};
// per cell: kind, width, gradient matrix (naive) or least-squares sums (boundary)
struct alignas(32) CellGeoM {
  float kind, width, c0, c1, c2, c3, pad0, pad1;
};
// per cell record: cell, x-slope, y-slope
struct alignas(32) RecM {  // a cell's limited slopes in FP32
  float xd, xp, xvx, xvy, yd, yp, yvx, yvy;
};

// Incidences in wavefront-sliced ELL order: the cells of wavefront w are
// 64w..64w+63, its slot s for lane l is entry wofs[w] + 64 s + l, and it has
// (wofs[w+1] - wofs[w]) / 64 slots (the largest face count among its cells).
// A load of one slot across the wavefront reads contiguous memory.
struct MTables {
  const FaceM* face;
  const CellGeoM* cgeo;
  const int* wofs;
  const unsigned char* lanepos;  // ELL column of each cell (as in DMesh)
  const int* enbr;       // neighbour id (ghost ids as in the mesh)
  const float2* eh;      // r_ij/2 seen from the cell (gradient pass)
  const float2* eo;      // centroid - cell CM (limiter pass)
  const unsigned* einfo; // flux to read: face | INFO_SECOND, periodic copies shared
};

__device__ __forceinline__ int ell_base(const MTables& t, int i) {
  return t.wofs[i >> 6] + t.lanepos[i];
}

struct CellF {
  float d, p, vx, vy;
};

// The FP32 reconstruction and Riemann solver owe no bit parity to anything,
// so their multiply-adds may fuse (FMA: fewer instructions, one rounding).
// The FP64 update below keeps them separate, so that the two cells of a
// face add the same rounded products and the totals are conserved.
#pragma clang fp contract(fast)

__device__ __forceinline__ float fdiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float fminr(float a, float b) { return (b < a) ? b : a; }
__device__ __forceinline__ float fmaxr(float a, float b) { return (a < b) ? b : a; }

__device__ __forceinline__ CellF ld_cellf(const F4* q, int i) {
  const F4 v = q[i];
  return CellF{v.a, v.b, v.c, v.d};
}

__device__ __forceinline__ CellF fetchf(const DMesh& m, const F4* q, int id) {
  if (id < m.N) return ld_cellf(q, id);
  const int g = id - m.N;
  CellF r = ld_cellf(q, m.gsrc[g]);
  if (m.bc == BC_RIGID) {
    const D2 n = m.gn[g];
    const float nx = static_cast<float>(n.x), ny = static_cast<float>(n.y);
    const float s2 = 2.f * (r.vx * nx + r.vy * ny);
    r.vx = r.vx - nx * s2;
    r.vy = r.vy - ny * s2;
  }
  return r;
}

// fetchf() of K point ids with the cell loads issued back to back (ghost ids
// are resolved first, in a branch only boundary waves take).
template <int K>
__device__ __forceinline__ void fetchf_k(const DMesh& m, const F4* q, const int (&id)[K],
                                         CellF (&out)[K]) {
  int src[K];
  bool ghost = false;
#pragma unroll
  for (int u = 0; u < K; ++u) {
    src[u] = id[u];
    ghost |= id[u] >= m.N;
  }
  if (ghost) {
#pragma unroll
    for (int u = 0; u < K; ++u)
      if (id[u] >= m.N) src[u] = m.gsrc[id[u] - m.N];
  }
#pragma unroll
  for (int u = 0; u < K; ++u) out[u] = ld_cellf(q, src[u]);
  if (ghost && m.bc == BC_RIGID) {
#pragma unroll
    for (int u = 0; u < K; ++u)
      if (id[u] >= m.N) {
        const D2 n = m.gn[id[u] - m.N];
        const float nx = static_cast<float>(n.x), ny = static_cast<float>(n.y);
        CellF& r = out[u];
        const float s2 = 2.f * (r.vx * nx + r.vy * ny);
        r.vx = r.vx - nx * s2;
        r.vy = r.vy - ny * s2;
      }
  }
}

__device__ __forceinline__ void add_multf(CellF& r, const CellF& o, float s) {
  r.d += o.d * s;
  r.p += o.p * s;
  r.vx += o.vx * s;
  r.vy += o.vy * s;
}
__device__ __forceinline__ CellF subf(const CellF& a, const CellF& b) {
  return CellF{a.d - b.d, a.p - b.p, a.vx - b.vx, a.vy - b.vy};
}
__device__ __forceinline__ void minmaxf(CellF& mx, CellF& mn, const CellF& t) {
  mx.d = fmaxr(mx.d, t.d);
  mx.p = fmaxr(mx.p, t.p);
  mx.vx = fmaxr(mx.vx, t.vx);
  mx.vy = fmaxr(mx.vy, t.vy);
  mn.d = fminr(mn.d, t.d);
  mn.p = fminr(mn.p, t.p);
  mn.vx = fminr(mn.vx, t.vx);
  mn.vy = fminr(mn.vy, t.vy);
}
__device__ __forceinline__ void limit_fieldf(float dphi, float cval, float cell, float maxd,
                                             float mind, bool absguard, float& psi) {
  if (fabsf(dphi) > 0.05f * fmaxr(fabsf(maxd), fabsf(mind)) || cval * cell < 0) {
    const float hi = absguard ? fabsf(1e-9f * cell) : 1e-9f * cell;
    const float lo = -hi;
    if (dphi > hi)
      psi = fminr(psi, fdiv(maxd, dphi));
    else if (dphi < lo)
      psi = fminr(psi, fdiv(mind, dphi));
  }
}
__device__ __forceinline__ void shocked_fieldf(float dphi, float cval, float cell, float nval,
                                               float maxd, float mind, float theta, float& psi) {
  if (fabsf(dphi) > 0.1f * fmaxr(fabsf(maxd), fabsf(mind)) || cval * cell < 0) {
    if (fabsf(dphi) > 1e-9f * cell) psi = fminr(psi, fmaxr(fdiv(theta * (nval - cell), dphi), 0.f));
  }
}

// Static FP32 tables, formed in FP64 and rounded once.
RG_GLOBAL void __launch_bounds__(256) k_geo_mixed_faces(DMesh m, FaceM* face) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= m.F) return;
  const FaceRec r = m.frec[f];
  const FaceFrame fr = m.fframe[f];
  const D2 ca = m.cm[m.ff[f]], cb = m.cm[m.fs[f]];
  const double an = vabs(fr.nx, fr.ny);
  face[f] = FaceM{static_cast<float>(fr.nx / an), static_cast<float>(fr.ny / an),
                  static_cast<float>(m.area[f]), 0.f,
                  static_cast<float>(r.cx - ca.x), static_cast<float>(r.cy - ca.y),
                  static_cast<float>(r.cx - cb.x), static_cast<float>(r.cy - cb.y)};
}

// ELL geometry of every incidence: one thread per cell over its faces.
RG_GLOBAL void __launch_bounds__(256) k_geo_mixed_ell(DMesh m, float2* __restrict__ eh,
                                                       float2* __restrict__ eo) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m.N) return;
  const D2 ci = m.cm[i];
  const int base = ell_base(m, i);
  for (int k = m.off[i]; k < m.off[i + 1]; ++k) {
    const unsigned info = m.inc[k].info;
    const FaceRec r = m.frec[info & INFO_FACE];
    const bool second = (info & INFO_SECOND) != 0;
    const int at = base + 64 * (k - m.off[i]);
    eh[at] = make_float2(static_cast<float>(second ? -r.hx : r.hx),
                         static_cast<float>(second ? -r.hy : r.hy));
    eo[at] = make_float2(static_cast<float>(r.cx - ci.x), static_cast<float>(r.cy - ci.y));
  }
}

RG_GLOBAL void __launch_bounds__(256) k_geo_mixed_cells(DMesh m, CellGeoM* cg) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m.N) return;
  const CellGeo g = m.cgeo[i];
  cg[i] = CellGeoM{static_cast<float>(g.kind), static_cast<float>(g.width),
                   static_cast<float>(g.c0), static_cast<float>(g.c1), static_cast<float>(g.c2),
                   static_cast<float>(g.c3), 0.f, 0.f};
}

// Incidences per chunk of the FP32 slope loops: a cell's FP32 values take
// half the registers of FP64 ones, so all of a typical cell's faces fit in
// one chunk (measured 3 % faster than 4 on a 2M-cell mesh).
constexpr int CH_M = 8;

// LinearGaussImproved in FP32.
RG_GLOBAL_DEFAULT_SCHED void __launch_bounds__(256) k_slopes_mixed(DMesh m, MTables t,
                                                                   const F4* __restrict__ q,
                                                                   RecM* __restrict__ out,
                                                                   DParams prm, DErr err) {
  const int i = walk_block(prm) * blockDim.x + threadIdx.x;
  if (i >= m.N) return;
  const int b = m.off[i], e = m.off[i + 1], deg = e - b;
  const int eb = ell_base(t, i);
  const CellF cell = ld_cellf(q, i);
  const CellGeoM g = t.cgeo[i];
  const int kind = static_cast<int>(g.kind);
  CellF sx{0, 0, 0, 0}, sy{0, 0, 0, 0};
  float pr = 1.f;
  CellF amax = cell, amin = cell, bmax = cell, bmin = cell;
  if (kind == KIND_TOO_MANY || kind == KIND_SINGULAR) {
    report(err, i, kind == KIND_TOO_MANY ? ERR_TOO_MANY_NEIGHBORS : ERR_SINGULAR_MATRIX);
    out[i] = RecM{0, 0, 0, 0, 0, 0, 0, 0};
    return;
  }
  if (kind == KIND_BOUNDARY) {
    CellF phy{0, 0, 0, 0}, phx{0, 0, 0, 0};
    const D2 ci = m.cm[i];
    for (int k = b; k < e; ++k) {
      const Inc ic = m.inc[k];
      const CellF nc = fetchf(m, q, ic.nbr);
      const D2 cj = m.cm[ic.nbr];
      const float ey = static_cast<float>(cj.y - ci.y), ex = static_cast<float>(cj.x - ci.x);
      const CellF dd = subf(nc, cell);
      add_multf(phy, dd, ey);
      add_multf(phx, dd, ex);
      pr = (cell.p > nc.p) ? fminr(pr, fdiv(nc.p, cell.p)) : fminr(pr, fdiv(cell.p, nc.p));
      minmaxf(amax, amin, nc);
      if (!(ic.info & INFO_MIRROR)) minmaxf(bmax, bmin, nc);
    }
    const float SxSy = g.c0, Sx2 = g.c1, Sy2 = g.c2, rden = __builtin_amdgcn_rcpf(g.c3);
    sx = CellF{(phy.d * SxSy - phx.d * Sy2) * rden, (phy.p * SxSy - phx.p * Sy2) * rden,
               (phy.vx * SxSy - phx.vx * Sy2) * rden, (phy.vy * SxSy - phx.vy * Sy2) * rden};
    sy = CellF{(phx.d * SxSy - phy.d * Sx2) * rden, (phx.p * SxSy - phy.p * Sx2) * rden,
               (phx.vx * SxSy - phy.vx * Sx2) * rden, (phx.vy * SxSy - phy.vy * Sx2) * rden};
  } else {
    CellF vcx{0, 0, 0, 0}, vcy{0, 0, 0, 0};
    for (int s0 = 0; s0 < deg; s0 += CH_M) {
      float2 hh[CH_M];
      CellF nc[CH_M];
      int nb[CH_M];
      // slots past the last face re-read the last slot: no load behind a branch
#pragma unroll
      for (int u = 0; u < CH_M; ++u) nb[u] = t.enbr[eb + 64 * min(s0 + u, deg - 1)];
#pragma unroll
      for (int u = 0; u < CH_M; ++u) hh[u] = t.eh[eb + 64 * min(s0 + u, deg - 1)];
      fetchf_k(m, q, nb, nc);
#pragma unroll
      for (int u = 0; u < CH_M; ++u) {
        if (s0 + u >= deg) break;
        add_multf(vcx, cell, hh[u].x);
        add_multf(vcy, cell, hh[u].y);
        add_multf(vcy, nc[u], hh[u].y);
        add_multf(vcx, nc[u], hh[u].x);
        pr = (cell.p > nc[u].p) ? fminr(pr, fdiv(nc[u].p, cell.p)) : fminr(pr, fdiv(cell.p, nc[u].p));
        minmaxf(amax, amin, nc[u]);
      }
    }
    bmax = amax;
    bmin = amin;
    sx = CellF{vcx.d * g.c0, vcx.p * g.c0, vcx.vx * g.c0, vcx.vy * g.c0};
    sy = CellF{vcy.d * g.c3, vcy.p * g.c3, vcy.vx * g.c3, vcy.vy * g.c3};
    add_multf(sx, vcy, g.c1);
    add_multf(sy, vcx, g.c2);
  }
  if (prm.slf) {
    const float cs = fsqrt(fdiv(static_cast<float>(prm.gamma) * cell.p, cell.d));
    const bool shock = ((sx.vx + sy.vy) * g.width < static_cast<float>(-prm.delta_v) * cs) ||
                       (pr < static_cast<float>(prm.delta_p));
    float psi0 = 1, psi1 = 1, psi2 = 1, psi3 = 1;
    const CellF maxd = subf(shock ? amax : bmax, cell), mind = subf(shock ? amin : bmin, cell);
    const float theta = static_cast<float>(prm.theta);
    for (int s0 = 0; s0 < deg; s0 += CH_M) {
      float2 oo[CH_M];
      CellF nc[CH_M];
#pragma unroll
      for (int u = 0; u < CH_M; ++u) oo[u] = t.eo[eb + 64 * min(s0 + u, deg - 1)];
      if (shock) {
        int nb[CH_M];
#pragma unroll
        for (int u = 0; u < CH_M; ++u) nb[u] = t.enbr[eb + 64 * min(s0 + u, deg - 1)];
        fetchf_k(m, q, nb, nc);
      }
#pragma unroll
      for (int u = 0; u < CH_M; ++u) {
        if (s0 + u >= deg) break;
        CellF cv = cell;
        add_multf(cv, sx, oo[u].x);
        add_multf(cv, sy, oo[u].y);
        const CellF dphi = subf(cv, cell);
        if (!shock) {
          limit_fieldf(dphi.d, cv.d, cell.d, maxd.d, mind.d, false, psi0);
          limit_fieldf(dphi.p, cv.p, cell.p, maxd.p, mind.p, false, psi1);
          limit_fieldf(dphi.vx, cv.vx, cell.vx, maxd.vx, mind.vx, true, psi2);
          limit_fieldf(dphi.vy, cv.vy, cell.vy, maxd.vy, mind.vy, true, psi3);
        } else {
          shocked_fieldf(dphi.d, cv.d, cell.d, nc[u].d, maxd.d, mind.d, theta, psi0);
          shocked_fieldf(dphi.p, cv.p, cell.p, nc[u].p, maxd.p, mind.p, theta, psi1);
          shocked_fieldf(dphi.vx, cv.vx, cell.vx, nc[u].vx, maxd.vx, mind.vx, theta, psi2);
          shocked_fieldf(dphi.vy, cv.vy, cell.vy, nc[u].vy, maxd.vy, mind.vy, theta, psi3);
        }
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
  out[i] = RecM{sx.d, sx.p, sx.vx, sx.vy, sy.d, sy.p, sy.vx, sy.vy};
}

// Face states of both sides in FP32 ((ox, oy) = centroid - CM of that side;
// a rigid-wall ghost has zero gradient), with all loads issued together
__device__ __forceinline__ void face_states_m(const DMesh& m, const F4* q, const RecM* rec, int a,
                                              int b, const FaceM& fr, CellF& lt, CellF& rt) {
  const int id[2] = {a, b};
  int src[2] = {a, b};
  const bool ghost = a >= m.N || b >= m.N;
  if (ghost) {
#pragma unroll
    for (int u = 0; u < 2; ++u)
      if (id[u] >= m.N)
        src[u] = m.bc == BC_PERIODIC ? m.gorig[id[u] - m.N] : m.gsrc[id[u] - m.N];
  }
  RecM r[2];
  CellF v[2];
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    r[u] = rec[src[u]];
    v[u] = ld_cellf(q, src[u]);
  }
  if (ghost && m.bc == BC_RIGID) {
    // reflected copy with zero gradient, as fetchf() forms it
#pragma unroll
    for (int u = 0; u < 2; ++u)
      if (id[u] >= m.N) {
        const D2 n = m.gn[id[u] - m.N];
        const float nx = static_cast<float>(n.x), ny = static_cast<float>(n.y);
        const float s2 = 2.f * (v[u].vx * nx + v[u].vy * ny);
        v[u].vx = v[u].vx - nx * s2;
        v[u].vy = v[u].vy - ny * s2;
      }
  }
  const bool slope[2] = {!(ghost && m.bc == BC_RIGID && a >= m.N),
                         !(ghost && m.bc == BC_RIGID && b >= m.N)};
  const float ox[2] = {fr.oxa, fr.oxb}, oy[2] = {fr.oya, fr.oyb};
#pragma unroll
  for (int u = 0; u < 2; ++u)
    if (slope[u]) {
      add_multf(v[u], CellF{r[u].xd, r[u].xp, r[u].xvx, r[u].xvy}, ox[u]);
      add_multf(v[u], CellF{r[u].yd, r[u].yp, r[u].yvx, r[u].yvy}, oy[u]);
    }
  lt = v[0];
  rt = v[1];
}

// rotated HLLC in FP32 (same wave-speed estimate and star states as the FP64
// path); returns the lab-frame flux times the face area
__device__ __forceinline__ int hllc_face_m(const CellF& L, const CellF& R, float g, float nx,
                                           float ny, float px, float py, float area,
                                           float velocity, F4& out) {
  const float cl = fsqrt(fdiv(g * L.p, L.d)), cr = fsqrt(fdiv(g * R.p, R.d));
  const float rvx = R.vx * nx + R.vy * ny;
  if (!(rvx == rvx)) return ERR_HLLC_NAN;
  const float lvx = L.vx * nx + L.vy * ny;
  const float ul = lvx - velocity, ur = rvx - velocity;
  const float sl = fminr(ul - cl, ur - cr);
  const float sr = fmaxr(ul + cl, ur + cr);
  const float ss = fdiv(R.p - L.p + L.d * ul * (sl - ul) - R.d * ur * (sr - ur),
                        L.d * (sl - ul) - R.d * (sr - ur));
  const int branch = (sl > 0) ? 0 : (sl <= 0 && ss >= 0) ? 1 : (ss < 0 && sr >= 0) ? 2
                                                            : (sr < 0) ? 3 : 4;
  if (branch == 4) return ERR_INVALID_WAVE_SPEEDS;
  const bool left = branch <= 1;
  const CellF& S = left ? L : R;
  const float u = left ? ul : ur;
  const float sk = left ? sl : sr;
  const float v = S.vx * px + S.vy * py;
  const float rd = __builtin_amdgcn_rcpf(S.d);
  const float e = S.p * rd * __builtin_amdgcn_rcpf(g - 1.f);
  const float es = S.d * (0.5f * (u * u + v * v) + e);
  const float dun = S.d * u;
  float fm, fx, fy, fE;
  if (branch == 0 || branch == 3) {
    fm = dun;
    fx = S.p + u * dun;
    fy = v * dun;
    fE = (es + S.p) * u;
  } else {
    const float ds = fdiv(S.d * (sk - u), sk - ss);
    const float usE = es * ds * rd + ds * (ss - u) * (ss + fdiv(S.p * rd, sk - u));
    fm = dun + sk * (ds - S.d);
    fx = (S.p + u * dun) + (ds * ss - u * S.d) * sk;
    fy = v * dun + (ds * v - v * S.d) * sk;
    fE = (es + S.p) * u + sk * (usE - es);
  }
  fE = fE + (fx * velocity + 0.5f * fm * velocity * velocity);
  fx = fx + velocity * fm;
  // rotate back with unit vectors, times the face area
  out = F4{fm * area, fE * area, (fx * nx + fy * px) * area, (fx * ny + fy * py) * area};
  return 0;
}

// the face's speed along its unit normal
__device__ __forceinline__ float face_speed_m(const FaceM& fr, float wx, float wy) {
  return fr.nx * wx + fr.ny * wy;
}

RG_GLOBAL void __launch_bounds__(256) k_flux_mixed(DMesh m, MTables t, const F4* __restrict__ q,
                                                    const RecM* __restrict__ rec,
                                                    F4* __restrict__ flux, DParams prm, DErr err) {
  const int f = walk_block(prm) * blockDim.x + threadIdx.x;
  if (f >= m.F) return;
  const FaceM fr = t.face[f];
  CellF lt, rt;
  face_states_m(m, q, rec, m.ff[f], m.fs[f], fr, lt, rt);
  const float wx = m.wx ? static_cast<float>(m.wx[f]) : 0.f;
  const float wy = m.wy ? static_cast<float>(m.wy[f]) : 0.f;
  const float speed = face_speed_m(fr, wx, wy);
  F4 o{0, 0, 0, 0};
  const int rc = hllc_face_m(lt, rt, static_cast<float>(prm.gamma), fr.nx, fr.ny, -fr.ny, fr.nx,
                             fr.area, speed, o);
  if (rc) {
    report(err, f, rc);
    o = F4{0, 0, 0, 0};
  }
  flux[f] = o;
}

#pragma clang fp contract(off)

// Neumaier-compensated FP64 sum
struct KSum {
  double s, c;
  __device__ __forceinline__ void add(double x) {
    const double t = s + x;
    c += (fabs(s) >= fabs(x)) ? (s - t) + x : (x - t) + s;
    s = t;
  }
  __device__ __forceinline__ double value() const { return s + c; }
};

struct UpdArgsM {
  const D4* in;
  D4* out;
  const D4* avg;
  const F4* flux;  // flux times area
  F4* cells;       // FP32 copy for the next reconstruction
  const unsigned long long* dt_slot;
  unsigned long long* reset_slot;
  double* partials;
  double* dt_hist;
  int step;
  int flags;
};

RG_GLOBAL void __launch_bounds__(256) k_update_mixed(DMesh m, MTables t, UpdArgsM a, DParams prm,
                                                      DErr err) {
  const int i = walk_block(prm) * blockDim.x + threadIdx.x;
  const double dt = prm.cfl * __longlong_as_double(static_cast<long long>(*a.dt_slot));
  if (i == 0) {
    if (a.dt_hist) a.dt_hist[a.step] = dt;
    if (a.reset_slot) *a.reset_slot = DBL_INF_BITS;
  }
  double cfl_v = __longlong_as_double(static_cast<long long>(DBL_INF_BITS));
  if (i < m.N) {
    const D4 x = a.in[i];
    KSum sm{x.a, 0}, sE{x.b, 0}, sx{x.c, 0}, sy{x.d, 0};
    const int deg = m.off[i + 1] - m.off[i], eb = ell_base(t, i);
    for (int s0 = 0; s0 < deg; s0 += CH_UPD) {
      unsigned inf[CH_UPD];
      F4 fl[CH_UPD];
#pragma unroll
      for (int u = 0; u < CH_UPD; ++u)
        if (s0 + u < deg) inf[u] = t.einfo[eb + 64 * (s0 + u)];
#pragma unroll
      for (int u = 0; u < CH_UPD; ++u)
        if (s0 + u < deg) fl[u] = a.flux[inf[u] & INFO_FACE];
#pragma unroll
      for (int u = 0; u < CH_UPD; ++u) {
        if (s0 + u >= deg) break;
        const double s = (inf[u] & INFO_SECOND) ? dt : -dt;
        sm.add(s * static_cast<double>(fl[u].a));
        sE.add(s * static_cast<double>(fl[u].b));
        sx.add(s * static_cast<double>(fl[u].c));
        sy.add(s * static_cast<double>(fl[u].d));
      }
    }
    double em = sm.value(), eE = sE.value(), ex = sx.value(), ey = sy.value();
    if (a.flags & UPD_AVERAGE) {
      const D4 o = a.avg[i];
      em = 0.5 * (em + o.a);
      eE = 0.5 * (eE + o.b);
      ex = (ex + o.c) * 0.5;
      ey = (ey + o.d) * 0.5;
    }
    a.out[i] = D4{em, eE, ex, ey};
    // FP64 primitive recovery
    const double d = em / m.vol[i];
    const double vx = ex / em, vy = ey / em;
    const double e = eE / em - 0.5 * (vx * vx + vy * vy);
    if (d < 0)
      report(err, i, ERR_NEGATIVE_DENSITY);
    else if (e < 0)
      report(err, i, ERR_NEGATIVE_THERMAL);
    const double p = (prm.gamma - 1) * e * d;
    a.cells[i] = F4{static_cast<float>(d), static_cast<float>(p), static_cast<float>(vx),
                    static_cast<float>(vy)};
    if (a.flags & UPD_CFL) cfl_v = cfl_inv(m, i, prm.gamma, m.cgeo[i].width, Cell{d, p, vx, vy});
  }
  if (a.flags & UPD_CFL) block_min(cfl_v, a.partials);
}

// k_face_update in mixed precision: k_flux_mixed on every face the tile's
// cells read (a periodic boundary face through its lower-numbered copy, as
// k_update_mixed reads it) into LDS, then k_update_mixed's update of the
// tile's cells from those fluxes.  The update's outputs are other buffers
// than the cells and slopes the faces read.
RG_GLOBAL_DEFAULT_SCHED void __launch_bounds__(kTile) k_face_update_mixed(
    DMesh m, MTables t, DTiles tl, const F4* __restrict__ q, const RecM* __restrict__ rec,
    UpdArgsM a, DParams prm, DErr errf, DErr erru) {
  extern __shared__ F4 tfluxm[];
  // slots per chunk of the update: the fluxes come from LDS, so a short
  // chunk hides their latency and keeps the registers down
  constexpr int CHT = 4;
  const int tile = walk_block(prm);
  const double dt = prm.cfl * __longlong_as_double(static_cast<long long>(*a.dt_slot));
  const int e0 = tl.tofs[tile], ne = tl.tofs[tile + 1] - e0;
  const float g = static_cast<float>(prm.gamma);
  // each thread's next entry is loaded while the current one is solved
  FaceM frn{0, 0, 0, 0, 0, 0, 0, 0};
  FaceSides sdn{0, 0};
  if (static_cast<int>(threadIdx.x) < ne) {
    frn = tl.tfm[e0 + threadIdx.x];
    sdn = tl.tside[e0 + threadIdx.x];
  }
  for (int k = threadIdx.x; k < ne; k += blockDim.x) {
    const int e = e0 + k;
    const FaceM fr = frn;
    const FaceSides sd = sdn;
    if (k + static_cast<int>(blockDim.x) < ne) {
      frn = tl.tfm[e + blockDim.x];
      sdn = tl.tside[e + blockDim.x];
    }
    CellF lt, rt;
    face_states_m(m, q, rec, sd.a, sd.b, fr, lt, rt);
    float wx = 0.f, wy = 0.f;
    if (m.wx) {
      const int f = tl.tface[e];
      wx = static_cast<float>(m.wx[f]);
      wy = static_cast<float>(m.wy[f]);
    }
    const float speed = face_speed_m(fr, wx, wy);
    F4 o{0, 0, 0, 0};
    const int rc = hllc_face_m(lt, rt, g, fr.nx, fr.ny, -fr.ny, fr.nx, fr.area, speed, o);
    if (rc) {
      report(errf, tl.tface[e], rc);
      o = F4{0, 0, 0, 0};
    }
    tfluxm[k] = o;
  }
  const int i = tile * kTile + threadIdx.x;
  D4 x{0, 0, 0, 0}, av{0, 0, 0, 0};
  double vol = 0, wid = 0;
  int deg = 0, eb = 0;
  unsigned w0[CHT];  // the first chunk's slot words
  if (i < m.N) {
    x = a.in[i];
    if (a.flags & UPD_AVERAGE) av = a.avg[i];
    vol = m.vol[i];
    if (a.flags & UPD_CFL) wid = m.width[i];  // the FP64 width, as k_update_mixed reads it
    deg = m.off[i + 1] - m.off[i];
    eb = ell_base(t, i);
#pragma unroll
    for (int u = 0; u < CHT; ++u) w0[u] = u < deg ? tl.eloc[eb + 64 * u] : 0u;
  }
  __syncthreads();
  if (i == 0) {
    if (a.dt_hist) a.dt_hist[a.step] = dt;
    if (a.reset_slot) *a.reset_slot = DBL_INF_BITS;
  }
  double cfl_v = __longlong_as_double(static_cast<long long>(DBL_INF_BITS));
  if (i < m.N) {
    KSum sm{x.a, 0}, sE{x.b, 0}, sx{x.c, 0}, sy{x.d, 0};
    for (int s0 = 0; s0 < deg; s0 += CHT) {
      unsigned inf[CHT];
      F4 fl[CHT];
#pragma unroll
      for (int u = 0; u < CHT; ++u)
        if (s0 + u < deg) inf[u] = s0 == 0 ? w0[u] : tl.eloc[eb + 64 * (s0 + u)];
#pragma unroll
      for (int u = 0; u < CHT; ++u)
        if (s0 + u < deg) fl[u] = tfluxm[inf[u] & TILE_ENTRY];
#pragma unroll
      for (int u = 0; u < CHT; ++u) {
        if (s0 + u >= deg) break;
        const double s = (inf[u] & TILE_SECOND) ? dt : -dt;
        sm.add(s * static_cast<double>(fl[u].a));
        sE.add(s * static_cast<double>(fl[u].b));
        sx.add(s * static_cast<double>(fl[u].c));
        sy.add(s * static_cast<double>(fl[u].d));
      }
    }
    double em = sm.value(), eE = sE.value(), ex = sx.value(), ey = sy.value();
    if (a.flags & UPD_AVERAGE) {
      em = 0.5 * (em + av.a);
      eE = 0.5 * (eE + av.b);
      ex = (ex + av.c) * 0.5;
      ey = (ey + av.d) * 0.5;
    }
    a.out[i] = D4{em, eE, ex, ey};
    // FP64 primitive recovery
    const double d = em / vol;
    const double vx = ex / em, vy = ey / em;
    const double e = eE / em - 0.5 * (vx * vx + vy * vy);
    if (d < 0)
      report(erru, i, ERR_NEGATIVE_DENSITY);
    else if (e < 0)
      report(erru, i, ERR_NEGATIVE_THERMAL);
    const double p = (prm.gamma - 1) * e * d;
    a.cells[i] = F4{static_cast<float>(d), static_cast<float>(p), static_cast<float>(vx),
                    static_cast<float>(vy)};
    if (a.flags & UPD_CFL) cfl_v = cfl_inv(m, i, prm.gamma, wid, Cell{d, p, vx, vy});
  }
  if (a.flags & UPD_CFL) block_min(cfl_v, a.partials);
}

// The tiles' per-entry face data in entry order (set-up): the FP64 frame
// or the FP32 face record, whichever is requested.
RG_GLOBAL void __launch_bounds__(256) k_tile_frames(DMesh m, MTables t,
                                                     const int* __restrict__ tface,
                                                     FaceFrame* __restrict__ tframe,
                                                     FaceM* __restrict__ tfm, int n) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const int f = tface[e];
  if (tframe) tframe[e] = m.fframe[f];
  if (tfm) tfm[e] = t.face[f];
}

RG_GLOBAL void k_to_f32(const D4* __restrict__ s, F4* __restrict__ d, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const D4 v = s[i];
    d[i] = F4{static_cast<float>(v.a), static_cast<float>(v.b), static_cast<float>(v.c),
              static_cast<float>(v.d)};
  }
}

// Enqueue the kernels default_sched.hip builds.
hipError_t launch_slopes_mixed(int blocks, int threads, hipStream_t st, DMesh m, MTables t,
                               const F4* q, RecM* out, DParams prm, DErr err);
hipError_t launch_face_update_mixed(int blocks, int threads, size_t lds, hipStream_t st, DMesh m,
                                    MTables t, DTiles tl, const F4* q, const RecM* rec,
                                    UpdArgsM a, DParams prm, DErr errf, DErr erru);

}  // namespace rg
