// HIP kernels of the FP64 finite-volume step (gfx950).
//
// Per mesh (static geometry, computed once by k_geo_*):
//   per face  FaceRec   {h = r_ij/2 seen from the first side, centroid}
//             FaceFrame {centroid, rotated-frame normal n, face
//                        direction p}
//             area
//   per cell  CellGeo   {inverse gradient matrix, width, kind, face count}
//             -- the matrix of calc_naive_slope (or the least-squares sums
//             of GetBoundarySlope) depends on geometry only
//   per incidence (wavefront-sliced ELL) neighbour, face word, h, centroid
//             minus the cell's CM, area
// Per stage:
//   k_slopes  one thread per cell: the value-dependent half of
//             LinearGaussImproved (gradient right-hand side in face order,
//             shock detector, both limiters); writes a 64-byte record of
//             the limited x- and y-slopes read by the face kernel
//   k_flux2   one thread per face: face states, rotated HLLC, rotate back
//   k_flux_pcm one thread per face: ModularFluxCalculator over PCM
//   k_flux1   one thread per face: first-order SimpleFluxCalculator
//   k_update  one thread per cell: ascending-face segmented sum of
//             -/+ dt*A*F, ZeroForce, Heun average, primitive recovery and
//             the CFL term with per-block minima; k_reduce_min combines
//             them
// Precomputing a quantity once evaluates the same IEEE operations the
// reference evaluates on every call, so results stay bit-identical.
// Each cell gathers its faces in ascending face order, the order in which
// the reference scatters face contributions: sums are deterministic and
// equal to the reference's, with no atomics.
#pragma once

#include "physics.hpp"

// The library is built from two translation units: richgpu.hip, under the
// max-ILP instruction scheduler, and default_sched.hip, under the
// compiler's default one (the FP64 slope kernel and the mixed-precision
// stage kernels run faster there).  Each kernel of these headers is defined
// by one unit: RG_GLOBAL kernels by richgpu.hip, RG_GLOBAL_DEFAULT_SCHED
// ones by default_sched.hip.  In the other unit a kernel is a template that
// nothing instantiates, so that unit emits no code for it and the two link
// into one library.
#ifndef RG_GLOBAL
#define RG_GLOBAL __global__
#define RG_GLOBAL_DEFAULT_SCHED template <int = 0> __global__
#endif

namespace rg {

constexpr int BC_RIGID = 0;
constexpr int BC_PERIODIC = 1;
constexpr unsigned INFO_SECOND = 0x80000000u;  // cell is the face's second side
constexpr unsigned INFO_MIRROR = 0x40000000u;  // GetOriginalIndex equal on both sides
constexpr unsigned INFO_FACE = 0x3FFFFFFFu;
constexpr double PI = 3.14159265358979323846;  // M_PI
constexpr unsigned long long DBL_INF_BITS = 0x7FF0000000000000ull;

enum : int { KIND_NAIVE = 0, KIND_BOUNDARY = 1, KIND_SINGULAR = 2, KIND_TOO_MANY = 3 };

struct alignas(16) D2 {
  double x, y;
};
struct alignas(32) D4 {  // cell {d,p,vx,vy}; extensive / flux {m,E,Mx,My}
  double a, b, c, d;
};
struct alignas(16) CellGeo {
  double c0, c1, c2, c3;  // naive: inverse matrix i0..i3; boundary: SxSy, Sx2, Sy2, den
  double width;
  int kind, deg;          // slope kind, number of faces
};
struct alignas(32) FaceRec {
  double hx, hy, cx, cy;
};
// the face centroid, the normal n and the face direction p; |n| and |p| are
// recomputed with vabs, and interp's centroid minus CM of each side with the
// side's CM (the subtraction the reference evaluates)
struct alignas(16) FaceFrame {
  double cx, cy, nx, ny, px, py;
};
struct alignas(8) Inc {
  int nbr;
  unsigned info;
};
struct alignas(64) Rec {  // a cell's limited slopes
  double xd, xp, xvx, xvy;
  double yd, yp, yvx, yvy;
};

// Raw mesh arrays (SoA, as uploaded) and the static per-mesh tables.
struct DMesh {
  int N, F, G, bc;
  const int* ff;
  const int* fs;
  const double *v1x, *v1y, *v2x, *v2y;
  const double *px, *py;
  const D2* cm;  // [N+G]
  const double* tvol;
  const double* vol;
  const int* off;
  const Inc* inc;
  const int* gorig;  // GetOriginalIndex of each ghost
  const int* gsrc;   // real cell a ghost copies (its generator's choice)
  const D2* gn;      // rigid-wall reflection normal of each ghost
  const double *wx, *wy;  // face velocities; null = zero
  // static tables
  const CellGeo* cgeo;
  const FaceRec* frec;
  const FaceFrame* fframe;
  const double* area;
  const double* width;  // GetWidth = sqrt(GetVolume/pi)
  // Incidences in wavefront-sliced ELL order: the cells of wavefront w are
  // 64w..64w+63; slot s (the s-th face in ascending order) of the cell in
  // column l is entry wofs[w] + 64 s + l.  One slot load across a wavefront
  // reads contiguous memory, where a CSR walk reads one scattered line per
  // lane.  Within a wavefront the columns hold the cells in descending order
  // of face count (lanepos: a cell's column; lanecell: the cell of a column,
  // as an offset from 64w), so that the entries a slot load needs form a
  // prefix of the slot and no memory sector holds only padding.
  const int* wofs;
  const unsigned char* lanepos;   // [N]
  const unsigned char* lanecell;  // [64 * waves]
  const int* enbr;       // neighbour id
  const unsigned* einfo; // face | INFO_SECOND | INFO_MIRROR
  const D2* eh;          // r_ij/2 as seen from the cell (negated on the second side)
  const D2* eo;          // face centroid minus the cell's CM
  const double* earea;   // face area
};

__device__ __forceinline__ int ell_base(const DMesh& m, int i) {
  return m.wofs[i >> 6] + m.lanepos[i];
}

// Tiles of the fused face-and-update kernel (k_face_update).  Tile t holds
// the cells kTile*t .. kTile*t+kTile-1 and, as its entries tofs[t] ..
// tofs[t+1]-1, every face with a real side among them, once (a face between
// two tiles is an entry of both).  Per entry the face's frame and sides are
// stored in entry order, so the face phase streams them; the face id is read
// only for face velocities and errors.  Per ELL slot, eloc holds the entry of
// the slot's face within the cell's tile, with TILE_SECOND where the cell is
// the face's second side.
#ifndef RG_TILE
#define RG_TILE 256
#endif
constexpr int kTile = RG_TILE;  // cells per tile = threads per block of the tile kernels
constexpr unsigned TILE_SECOND = 0x8000u;
constexpr unsigned TILE_ENTRY = 0x7FFFu;
constexpr int kTileMaxEntries = 2040;  // fluxes and block_min within 64 KiB of LDS
struct FaceSides {
  int a, b;
};
struct FaceM;  // the mixed-precision face record
struct DTiles {
  const int* tofs;            // [tiles + 1]
  const int* tface;           // face id of each entry
  const FaceSides* tside;     // (first, second) of each entry
  const FaceFrame* tframe;    // frame of each entry (FP64)
  const FaceM* tfm;           // FP32 face record of each entry (mixed precision)
  const unsigned short* eloc; // per ELL slot
};

struct DParams {
  double gamma, cfl, delta_v, theta, delta_p;
  int slf;
  // Walk the blocks in reverse order.  The step loop alternates the walk
  // between consecutive kernels, so that each kernel starts on the data its
  // predecessor wrote last, which the caches still hold.
  int rev;
};
__device__ __forceinline__ int walk_block(const DParams& prm) {
  return prm.rev ? static_cast<int>(gridDim.x) - 1 - static_cast<int>(blockIdx.x)
                 : static_cast<int>(blockIdx.x);
}
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

__device__ __forceinline__ Cell ld_cell(const D4* q, int i) {
  const D4 v = q[i];
  return Cell{v.a, v.b, v.c, v.d};
}

// Cell value at any point id: real cells directly, ghosts through the ghost
// generator -- RigidWallGenerator (copy of the real cell of the ghost's
// first boundary face with the normal velocity reversed,
// RigidWallGenerator.cpp:5-37) or PeriodicGhostGenerator (copy of the
// original, PeriodicGhostGenerator.cpp:3-17).
__device__ __forceinline__ Cell fetch(const DMesh& m, const D4* q, int id) {
  if (id < m.N) return ld_cell(q, id);
  const int g = id - m.N;
  Cell r = ld_cell(q, m.gsrc[g]);
  if (m.bc == BC_RIGID) {
    const D2 n = m.gn[g];
    const double s2 = 2 * (r.vx * n.x + r.vy * n.y);
    r.vx = r.vx - n.x * s2;
    r.vy = r.vy - n.y * s2;
  }
  return r;
}

// fetch() of K point ids at once.  Ghost ids are resolved to their source
// cells first (rare: only waves that touch the boundary take that branch),
// then all K cell loads are issued back to back.  With fetch() called in a
// loop, the ghost branch between two gathers makes the compiler wait for
// the first gather before it issues the second.
template <int K>
__device__ __forceinline__ void fetch_k(const DMesh& m, const D4* q, const int (&id)[K],
                                        Cell (&out)[K]) {
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
  for (int u = 0; u < K; ++u) out[u] = ld_cell(q, src[u]);
  if (ghost && m.bc == BC_RIGID) {
#pragma unroll
    for (int u = 0; u < K; ++u)
      if (id[u] >= m.N) {
        const D2 n = m.gn[id[u] - m.N];
        Cell& r = out[u];
        const double s2 = 2 * (r.vx * n.x + r.vy * n.y);
        r.vx = r.vx - n.x * s2;
        r.vy = r.vy - n.y * s2;
      }
  }
}

__device__ __forceinline__ Cell sub(const Cell& a, const Cell& b) {
  return Cell{a.d - b.d, a.p - b.p, a.vx - b.vx, a.vy - b.vy};
}

__device__ __forceinline__ double wave_min(double v) {
  for (int o = 32; o > 0; o >>= 1) v = smin(v, __shfl_xor(v, o, 64));
  return v;
}

// Block minimum into partials[blockIdx.x]; k_reduce_min combines them in a
// separate launch (kernel boundaries order the memory traffic; no fences).
// Every thread of the block must call it.  Blocks of at most 1024 threads.
__device__ __forceinline__ void block_min(double v, double* partials) {
  __shared__ double wmin[16];
  v = wave_min(v);
  const int w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  if ((threadIdx.x & 63) == 0) wmin[w] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double mb = wmin[0];
    for (int k = 1; k < nw; ++k) mb = smin(mb, wmin[k]);
    partials[blockIdx.x] = mb;
  }
}

// One block: min of n partials into *slot (bits of a non-negative double).
RG_GLOBAL void __launch_bounds__(1024) k_reduce_min(const double* __restrict__ partials, int n,
                                                     unsigned long long* slot) {
  __shared__ double wmin[16];
  double v = __longlong_as_double(static_cast<long long>(DBL_INF_BITS));
  for (int k = threadIdx.x; k < n; k += blockDim.x) v = smin(v, partials[k]);
  v = wave_min(v);
  const int w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  if ((threadIdx.x & 63) == 0) wmin[w] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = wmin[0];
    for (int k = 1; k < nw; ++k) m = smin(m, wmin[k]);
    *slot = static_cast<unsigned long long>(__double_as_longlong(m));
  }
}

// RiemannSolver on n pairs of face-frame Primitives (SoA [6][n]: density,
// pressure, x and y velocity, energy, sound speed), face velocities w (null:
// zero), fluxes out [4][n] (mass, energy, x and y momentum).
RG_GLOBAL void __launch_bounds__(256) k_riemann(const double* __restrict__ l,
                                                 const double* __restrict__ r,
                                                 const double* __restrict__ w,
                                                 double* __restrict__ out, int n, DErr err) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const size_t N = static_cast<size_t>(n);
  const Prim L{l[i], l[N + i], l[2 * N + i], l[3 * N + i], l[4 * N + i], l[5 * N + i]};
  const Prim R{r[i], r[N + i], r[2 * N + i], r[3 * N + i], r[4 * N + i], r[5 * N + i]};
  Cons f{0, 0, 0, 0};
  const int rc = hllc_prim(L, R, w ? w[i] : 0.0, f);
  if (rc) report(err, i, rc);
  out[i] = f.m;
  out[N + i] = f.E;
  out[2 * N + i] = f.mx;
  out[3 * N + i] = f.my;
}

// ------------------------------------------------------------------ mesh --
// Per-face static tables.
RG_GLOBAL void __launch_bounds__(256) k_geo_faces(DMesh m, FaceRec* frec, FaceFrame* fframe,
                                                   double* area) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= m.F) return;
  const int a = m.ff[f], b = m.fs[f];
  const double v1x = m.v1x[f], v1y = m.v1y[f], v2x = m.v2x[f], v2y = m.v2y[f];
  // CalcCentroid (spatial_reconstruction.cpp:3-6)
  const double cx = (v1x + v2x) * 0.5, cy = (v1y + v2y) * 0.5;
  // calc_naive_slope's r_ij = normalize(r_j - r_i)*|e| for i = first
  // (LinearGaussImproved.cpp:85-86); the second side sees -r_ij exactly
  const double L = vabs(v2x - v1x, v2y - v1y);
  const double dx = m.px[b] - m.px[a], dy = m.py[b] - m.py[a];
  const double ad = vabs(dx, dy);
  const double rx = dx / ad * L, ry = dy / ad * L;
  frec[f] = FaceRec{rx * 0.5, ry * 0.5, cx, cy};
  // calc_parallel_normal (modular_flux_calculator.cpp:12-22) with
  // remove_parallel_component (simple_flux_calculator.cpp:27-31)
  const double ppx = v2x - v1x, ppy = v2y - v1y;
  double qx, qy;
  if (a >= 0 && a < m.N) {
    qx = v1x - m.px[a];
    qy = v1y - m.py[a];
  } else {
    qx = m.px[b] - v1x;
    qy = m.py[b] - v1y;
  }
  const double s = qx * ppx + qy * ppy;
  const double pp = ppx * ppx + ppy * ppy;
  const double nx = qx - ppx * s / pp, ny = qy - ppy * s / pp;
  fframe[f] = FaceFrame{cx, cy, nx, ny, ppx, ppy};
  // SlabSymmetry::calcArea (physical_geometry.cpp:8-11)
  area[f] = vabs(v1x - v2x, v1y - v2y);
}

// Per-cell static table: the geometric half of calc_naive_slope
// (LinearGaussImproved.cpp:67-134: matrix, singularity test, inverse) or of
// GetBoundarySlope (:401-422), and GetWidth (VoronoiMesh.cpp:410-413).
RG_GLOBAL void __launch_bounds__(256) k_geo_cells(DMesh m, CellGeo* cg, double* width) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m.N) return;
  const int b = m.off[i], e = m.off[i + 1], n = e - b;
  const D2 cmi = m.cm[i];
  CellGeo g{};
  g.width = sqrt(m.tvol[i] / PI);
  g.deg = n;
  bool boundary = false;
  for (int k = b; k < e; ++k)
    if (m.inc[k].info & INFO_MIRROR) boundary = true;
  if (boundary) {
    double SxSy = 0, Sx2 = 0, Sy2 = 0;
    for (int k = b; k < e; ++k) {
      const D2 cj = m.cm[m.inc[k].nbr];
      const double ey = cj.y - cmi.y, ex = cj.x - cmi.x;
      SxSy += ey * ex;
      Sx2 += ex * ex;
      Sy2 += ey * ey;
    }
    g.kind = KIND_BOUNDARY;
    g.c0 = SxSy;
    g.c1 = Sx2;
    g.c2 = Sy2;
    g.c3 = SxSy * SxSy - Sx2 * Sy2;
  } else if (n > 20) {
    g.kind = KIND_TOO_MANY;
  } else {
    const double ptx = m.px[i], pty = m.py[i], vol = m.tvol[i];
    double M0 = 0, M1 = 0, M2 = 0, M3 = 0;
    for (int k = b; k < e; ++k) {
      const int f = static_cast<int>(m.inc[k].info & INFO_FACE);
      const int j = m.inc[k].nbr;
      const double fx = (m.v1x[f] + m.v2x[f]) * 0.5, fy = (m.v1y[f] + m.v2y[f]) * 0.5;
      const D2 cj = m.cm[j];
      const double cijx = fx - (cj.x + cmi.x) * 0.5;
      const double cijy = fy - (cj.y + cmi.y) * 0.5;
      const double L = vabs(m.v2x[f] - m.v1x[f], m.v2y[f] - m.v1y[f]);
      const double dx = m.px[j] - ptx, dy = m.py[j] - pty;
      const double ad = vabs(dx, dy);
      const double rx = dx / ad * L, ry = dy / ad * L;
      M0 -= cijx * rx;
      M1 -= cijy * rx;
      M2 -= cijx * ry;
      M3 -= cijy * ry;
    }
    M0 += vol;
    M3 += vol;
    const double det = M0 * M3 - M1 * M2;
    if (fabs(det) < 1e-10 * vol * vol) {
      g.kind = KIND_SINGULAR;
      g.c0 = det;  // the UniversalError entry "Det was"
    } else {
      const double det_inv = 1.0 / det;
      g.kind = KIND_NAIVE;
      g.c0 = M3 * det_inv;
      g.c1 = -M1 * det_inv;
      g.c2 = -M2 * det_inv;
      g.c3 = M0 * det_inv;
    }
  }
  cg[i] = g;
  width[i] = g.width;
}

// ---------------------------------------------------------------- slopes --
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

// Incidences are processed in chunks: the chunk's loads are all issued
// before any of them is used (memory-level parallelism for the dependent
// index -> data gathers); accumulation stays in face order.  The slope
// loops keep many values per incidence and take 4 at a time; the update
// loops take 8 (one chunk for almost every cell; measured 15 % faster than 4
// on a 2M-cell mesh).
constexpr int CH = 4;
constexpr int CH_UPD = 8;

// The ELL arrays (one thread per cell over its faces): the incidences, and
// copies of the per-incidence geometry -- the same values the slope and
// update loops would read through the face index, so the results do not
// change.
RG_GLOBAL void __launch_bounds__(256) k_geo_ell(DMesh m, int* __restrict__ enbr,
                                                 unsigned* __restrict__ einfo, D2* __restrict__ eh,
                                                 D2* __restrict__ eo, double* __restrict__ earea) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m.N) return;
  const D2 ci = m.cm[i];
  const int base = ell_base(m, i);
  for (int k = m.off[i]; k < m.off[i + 1]; ++k) {
    const Inc ic = m.inc[k];
    const unsigned info = ic.info;
    const int at = base + 64 * (k - m.off[i]);
    enbr[at] = ic.nbr;
    einfo[at] = info;
    const int f = static_cast<int>(info & INFO_FACE);
    const FaceRec r = m.frec[f];
    const bool second = (info & INFO_SECOND) != 0;
    eh[at] = D2{second ? -r.hx : r.hx, second ? -r.hy : r.hy};
    eo[at] = D2{r.cx - ci.x, r.cy - ci.y};
    earea[at] = m.area[f];
  }
}

// calc_slope (LinearGaussImproved.cpp:425-494) with the geometric half
// taken from CellGeo / FaceRec.  Writes the cell's record.
__device__ __forceinline__ void slopes_cell(const DMesh& m, const D4* __restrict__ q,
                                            Rec* __restrict__ out, const DParams& prm,
                                            const DErr& err, int i) {
  if (i >= m.N) return;
  // A cell whose slope cannot be formed runs the loops with no faces and
  // writes a zero record at the end: an early return here would let the
  // compiler hold back every other load of the cell until its kind is known.
  const CellGeo g = m.cgeo[i];
  const int kind = g.kind;
  const bool bad = kind == KIND_TOO_MANY || kind == KIND_SINGULAR;
  const int deg = bad ? 0 : g.deg;
  const int eb = ell_base(m, i);
  const Cell cell = ld_cell(q, i);
  Cell sx{0, 0, 0, 0}, sy{0, 0, 0, 0};
  double pr = 1;
  Cell amax = cell, amin = cell, bmax = cell, bmin = cell;
  if (kind == KIND_BOUNDARY) {
    // GetBoundarySlope (LinearGaussImproved.cpp:401-422)
    Cell phy{0, 0, 0, 0}, phx{0, 0, 0, 0};
    const D2 cmi = m.cm[i];
    const int b = m.off[i], e = b + deg;
    for (int k = b; k < e; ++k) {
      const Inc ic = m.inc[k];
      const Cell nc = fetch(m, q, ic.nbr);
      const D2 cj = m.cm[ic.nbr];
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
      pr = (cell.p > nc.p) ? smin(pr, nc.p / cell.p) : smin(pr, cell.p / nc.p);
      minmax(amax, amin, nc);
      if (!(ic.info & INFO_MIRROR)) minmax(bmax, bmin, nc);
    }
    const double SxSy = g.c0, Sx2 = g.c1, Sy2 = g.c2, den = g.c3;
    sx = Cell{(phy.d * SxSy - phx.d * Sy2) / den, (phy.p * SxSy - phx.p * Sy2) / den,
              (phy.vx * SxSy - phx.vx * Sy2) / den, (phy.vy * SxSy - phx.vy * Sy2) / den};
    sy = Cell{(phx.d * SxSy - phy.d * Sx2) / den, (phx.p * SxSy - phy.p * Sx2) / den,
              (phx.vx * SxSy - phy.vx * Sx2) / den, (phx.vy * SxSy - phy.vy * Sx2) / den};
  } else {
    // right-hand side of calc_naive_slope, in face order with the cell's
    // own term first (LinearGaussImproved.cpp:91-104).  A cell on this path
    // has no mirror face, so one set of neighbour extrema serves both limiters.
    // Slots past the cell's last face re-read its last slot (in bounds and
    // already cached), so that no load sits behind a branch.
    Cell vcx{0, 0, 0, 0}, vcy{0, 0, 0, 0};
    for (int s0 = 0; s0 < deg; s0 += CH) {
      int nb[CH];
      D2 h[CH];
      Cell nc[CH];
#pragma unroll
      for (int u = 0; u < CH; ++u) nb[u] = m.enbr[eb + 64 * min(s0 + u, deg - 1)];
#pragma unroll
      for (int u = 0; u < CH; ++u) h[u] = m.eh[eb + 64 * min(s0 + u, deg - 1)];
      fetch_k(m, q, nb, nc);
#pragma unroll
      for (int u = 0; u < CH; ++u) {
        if (s0 + u >= deg) break;
        const double hx = h[u].x, hy = h[u].y;
        if (s0 + u == 0) {
          vcx = Cell{cell.d * hx, cell.p * hx, cell.vx * hx, cell.vy * hx};
          vcy = Cell{cell.d * hy, cell.p * hy, cell.vx * hy, cell.vy * hy};
        } else {
          add_mult(vcx, cell, hx);
          add_mult(vcy, cell, hy);
        }
        add_mult(vcy, nc[u], hy);
        add_mult(vcx, nc[u], hx);
        pr = (cell.p > nc[u].p) ? smin(pr, nc[u].p / cell.p) : smin(pr, cell.p / nc[u].p);
        minmax(amax, amin, nc[u]);
      }
    }
    bmax = amax;
    bmin = amin;
    const double i0 = g.c0, i1 = g.c1, i2 = g.c2, i3 = g.c3;
    sx = Cell{vcx.d * i0, vcx.p * i0, vcx.vx * i0, vcx.vy * i0};
    sy = Cell{vcy.d * i3, vcy.p * i3, vcy.vx * i3, vcy.vy * i3};
    add_mult(sx, vcy, i1);
    add_mult(sy, vcx, i2);
  }
  if (bad) {
    report(err, i, kind == KIND_TOO_MANY ? ERR_TOO_MANY_NEIGHBORS : ERR_SINGULAR_MATRIX);
    out[i] = Rec{0, 0, 0, 0, 0, 0, 0, 0};
    return;
  }
  if (prm.slf) {
    // is_shock (LinearGaussImproved.cpp:150-157)
    const double cs = dp2c(prm.gamma, cell.d, cell.p);
    const bool shock = ((sx.vx + sy.vy) * g.width < (-prm.delta_v) * cs) || (pr < prm.delta_p);
    double psi0 = 1, psi1 = 1, psi2 = 1, psi3 = 1;
    if (!shock) {
      // slope_limit (LinearGaussImproved.cpp:174-300); interp's centroid
      // minus CM is the precomputed offset
      const Cell maxd = sub(bmax, cell), mind = sub(bmin, cell);
      for (int s0 = 0; s0 < deg; s0 += CH) {
        D2 o[CH];
#pragma unroll
        for (int u = 0; u < CH; ++u) o[u] = m.eo[eb + 64 * min(s0 + u, deg - 1)];
#pragma unroll
        for (int u = 0; u < CH; ++u) {
          if (s0 + u >= deg) break;
          const Cell cv = interp_off(cell, sx, sy, o[u].x, o[u].y);
          const Cell dphi = sub(cv, cell);
          limit_field(dphi.d, cv.d, cell.d, maxd.d, mind.d, false, psi0);
          limit_field(dphi.p, cv.p, cell.p, maxd.p, mind.p, false, psi1);
          limit_field(dphi.vx, cv.vx, cell.vx, maxd.vx, mind.vx, true, psi2);
          limit_field(dphi.vy, cv.vy, cell.vy, maxd.vy, mind.vy, true, psi3);
        }
      }
    } else {
      // shocked_slope_limit (LinearGaussImproved.cpp:302-399)
      const Cell maxd = sub(amax, cell), mind = sub(amin, cell);
      for (int s0 = 0; s0 < deg; s0 += CH) {
        int nb[CH];
        D2 o[CH];
        Cell nc[CH];
#pragma unroll
        for (int u = 0; u < CH; ++u) nb[u] = m.enbr[eb + 64 * min(s0 + u, deg - 1)];
#pragma unroll
        for (int u = 0; u < CH; ++u) o[u] = m.eo[eb + 64 * min(s0 + u, deg - 1)];
        fetch_k(m, q, nb, nc);
#pragma unroll
        for (int u = 0; u < CH; ++u) {
          if (s0 + u >= deg) break;
          const Cell cv = interp_off(cell, sx, sy, o[u].x, o[u].y);
          const Cell dphi = sub(cv, cell);
          shocked_field(dphi.d, cv.d, cell.d, nc[u].d, maxd.d, mind.d, prm.theta, psi0);
          shocked_field(dphi.p, cv.p, cell.p, nc[u].p, maxd.p, mind.p, prm.theta, psi1);
          shocked_field(dphi.vx, cv.vx, cell.vx, nc[u].vx, maxd.vx, mind.vx, prm.theta, psi2);
          shocked_field(dphi.vy, cv.vy, cell.vy, nc[u].vy, maxd.vy, mind.vy, prm.theta, psi3);
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
  out[i] = Rec{sx.d, sx.p, sx.vx, sx.vy, sy.d, sy.p, sy.vx, sy.vy};
}

// Lane l of wavefront w takes the cell of ELL column l, so that the lanes
// that run out of faces first are the last ones.  Built by
// default_sched.hip; launch_slopes enqueues it.
RG_GLOBAL_DEFAULT_SCHED void __launch_bounds__(256) k_slopes(DMesh m, const D4* __restrict__ q,
                                                             Rec* __restrict__ out, DParams prm,
                                                             DErr err) {
  const int t = walk_block(prm) * blockDim.x + threadIdx.x;
  if (t >= m.N) return;
  slopes_cell(m, q, out, prm, err, (t & ~63) + m.lanecell[t]);
}
hipError_t launch_slopes(int blocks, int threads, hipStream_t st, DMesh m, const D4* q, Rec* out,
                         DParams prm, DErr err);
// Persistent form: a fixed grid walks the cells in tiles of blockDim.x.
RG_GLOBAL void __launch_bounds__(256) k_slopes_p(DMesh m, const D4* __restrict__ q,
                                                  Rec* __restrict__ out, DParams prm, DErr err) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < m.N; i += gridDim.x * blockDim.x)
    slopes_cell(m, q, out, prm, err, i);
}

// ---------------------------------------------------------------- fluxes --
// Face state on one side (LinearGaussImproved.cpp:571-590 for real cells,
// :597-634 for ghosts: the ghost value plus the ghost gradient -- zero for
// rigid walls, the original's slope for periodic boxes -- at the ghost's CM).
// (ox, oy) = face centroid minus that side's CM, as interp forms it.
__device__ __forceinline__ Cell side_state(const DMesh& m, const D4* q, const Rec* rec, int id,
                                           double ox, double oy) {
  if (id < m.N || m.bc == BC_PERIODIC) {
    const int o = id < m.N ? id : m.gorig[id - m.N];
    const Rec r = rec[o];
    const Cell v = ld_cell(q, o);
    const Cell sx{r.xd, r.xp, r.xvx, r.xvy}, sy{r.yd, r.yp, r.yvx, r.yvy};
    return interp_off(v, sx, sy, ox, oy);
  }
  const Cell z{0, 0, 0, 0};
  return interp_off(fetch(m, q, id), z, z, ox, oy);
}

// side_state of both sides of face f, with the loads of both sides issued
// together: ghost ids are resolved first (a branch only waves with a ghost
// face take), then both slope records, cells and CMs are loaded.  interp's
// centroid minus CM (LinearGaussImproved.cpp:159-172) takes a ghost side at
// the ghost's CM (:597-634).
__device__ __forceinline__ void face_states(const DMesh& m, const D4* q, const Rec* rec, int a,
                                            int b, const FaceFrame& fr, Cell& lf, Cell& rt) {
  const int id[2] = {a, b};
  int src[2] = {a, b};
  const bool ghost = a >= m.N || b >= m.N;
  if (ghost) {
#pragma unroll
    for (int u = 0; u < 2; ++u)
      if (id[u] >= m.N)
        src[u] = m.bc == BC_PERIODIC ? m.gorig[id[u] - m.N] : m.gsrc[id[u] - m.N];
  }
  Rec r[2];
  Cell v[2];
  D2 cm[2];
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    r[u] = rec[src[u]];
    v[u] = ld_cell(q, src[u]);
    cm[u] = m.cm[id[u]];
  }
  if (ghost && m.bc == BC_RIGID) {
    // RigidWallGenerator: reflected copy, zero gradient
#pragma unroll
    for (int u = 0; u < 2; ++u)
      if (id[u] >= m.N) {
        const D2 n = m.gn[id[u] - m.N];
        const double s2 = 2 * (v[u].vx * n.x + v[u].vy * n.y);
        v[u].vx = v[u].vx - n.x * s2;
        v[u].vy = v[u].vy - n.y * s2;
        r[u] = Rec{0, 0, 0, 0, 0, 0, 0, 0};
      }
  }
  const Cell sxa{r[0].xd, r[0].xp, r[0].xvx, r[0].xvy}, sya{r[0].yd, r[0].yp, r[0].yvx, r[0].yvy};
  const Cell sxb{r[1].xd, r[1].xp, r[1].xvx, r[1].xvy}, syb{r[1].yd, r[1].yp, r[1].yvx, r[1].yvy};
  lf = interp_off(v[0], sxa, sya, fr.cx - cm[0].x, fr.cy - cm[0].y);
  rt = interp_off(v[1], sxb, syb, fr.cx - cm[1].x, fr.cy - cm[1].y);
}

// ModularFluxCalculator (modular_flux_calculator.cpp:39-71) on one face,
// over LinearGaussImproved's face states or, with PCM, over PCM's
// (interpolations/PCM.cpp:5-37: the cell on each side, a ghost side as its
// ghost generator fills it).  With dt > 0 the face writes its extensive
// update dt*area*F (the products SimpleExtensiveUpdater forms,
// simple_extensive_updater.cpp:23-34) instead of F.
template <bool PCM = false>
__device__ __forceinline__ void flux2_face(const DMesh& m, const D4* __restrict__ q,
                                           const Rec* __restrict__ rec, D4* __restrict__ flux,
                                           const DParams& prm, const DErr& err, int f,
                                           double dt = 0.0) {
  if (f >= m.F) return;
  const FaceFrame fr = m.fframe[f];
  Cell lf, rt;
  if (PCM) {
    const int id[2] = {m.ff[f], m.fs[f]};
    Cell v[2];
    fetch_k(m, q, id, v);
    lf = v[0];
    rt = v[1];
  } else {
    face_states(m, q, rec, m.ff[f], m.fs[f], fr, lf, rt);
  }
  const double an = vabs(fr.nx, fr.ny), ap = vabs(fr.px, fr.py);
  // modular_flux_calculator.cpp:56-67
  const double speed = m.wx ? (fr.nx * m.wx[f] + fr.ny * m.wy[f]) / an
                            : zero_face_speed(fr.nx, fr.ny, an);
  Cons c;
  const int rc = hllc_face(lf, rt, prm.gamma, fr.nx, fr.ny, an, fr.px, fr.py, ap, speed, c);
  D4 o{0, 0, 0, 0};
  if (rc) {
    report(err, f, rc);
  } else if (dt > 0) {
    const double sc = dt * ap;  // the face area |v1 - v2| (physical_geometry.cpp:8-11)
    o = D4{c.m * sc, c.E * sc, c.mx * sc, c.my * sc};
  } else {
    o = D4{c.m, c.E, c.mx, c.my};
  }
  flux[f] = o;
}

// dt_slot non-null: write dt*area*F with dt = cfl * *dt_slot (the step loop)
RG_GLOBAL void __launch_bounds__(256) k_flux2(DMesh m, const D4* __restrict__ q,
                                               const Rec* __restrict__ rec, D4* __restrict__ flux,
                                               DParams prm, DErr err,
                                               const unsigned long long* __restrict__ dt_slot) {
  const double dt =
      dt_slot ? prm.cfl * __longlong_as_double(static_cast<long long>(*dt_slot)) : 0.0;
  flux2_face(m, q, rec, flux, prm, err, walk_block(prm) * blockDim.x + threadIdx.x, dt);
}
// ModularFluxCalculator over PCM; dt_slot as for k_flux2
RG_GLOBAL void __launch_bounds__(256) k_flux_pcm(DMesh m, const D4* __restrict__ q,
                                                  D4* __restrict__ flux, DParams prm, DErr err,
                                                  const unsigned long long* __restrict__ dt_slot) {
  const double dt =
      dt_slot ? prm.cfl * __longlong_as_double(static_cast<long long>(*dt_slot)) : 0.0;
  flux2_face<true>(m, q, nullptr, flux, prm, err, blockIdx.x * blockDim.x + threadIdx.x, dt);
}
// Persistent form: a fixed grid walks the faces in tiles of blockDim.x.
RG_GLOBAL void __launch_bounds__(256) k_flux2_p(DMesh m, const D4* __restrict__ q,
                                                 const Rec* __restrict__ rec,
                                                 D4* __restrict__ flux, DParams prm, DErr err) {
  for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < m.F; f += gridDim.x * blockDim.x)
    flux2_face(m, q, rec, flux, prm, err, f);
}

// SpatialReconstruction output res[f].first / .second
RG_GLOBAL void __launch_bounds__(256) k_face_states(DMesh m, const D4* __restrict__ q,
                                                     const Rec* __restrict__ rec, D4* left,
                                                     D4* right) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= m.F) return;
  const FaceFrame fr = m.fframe[f];
  const int a = m.ff[f], b = m.fs[f];
  const D2 ca = m.cm[a], cb = m.cm[b];
  const Cell lf = side_state(m, q, rec, a, fr.cx - ca.x, fr.cy - ca.y);
  const Cell rt = side_state(m, q, rec, b, fr.cx - cb.x, fr.cy - cb.y);
  left[f] = D4{lf.d, lf.p, lf.vx, lf.vy};
  right[f] = D4{rt.d, rt.p, rt.vx, rt.vy};
}

// PCM::operator() (interpolations/PCM.cpp:5-37): res[f].first / .second are
// the cells on the two sides, ghost sides as the ghost generator fills them
RG_GLOBAL void __launch_bounds__(256) k_face_states_pcm(DMesh m, const D4* __restrict__ q,
                                                         D4* left, D4* right) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= m.F) return;
  const Cell lf = fetch(m, q, m.ff[f]);
  const Cell rt = fetch(m, q, m.fs[f]);
  left[f] = D4{lf.d, lf.p, lf.vx, lf.vy};
  right[f] = D4{rt.d, rt.p, rt.vx, rt.vy};
}

// The two states SimpleFluxCalculator solves on face f, with its normal
// (the point difference) and face direction (simple_flux_calculator.cpp:
// 80-126): a mirror face reflects the real state about the face (Reflect,
// geometry.cpp:117-120).
__device__ __forceinline__ void simple_face_states(const DMesh& m, const D4* __restrict__ q, int f,
                                                   Cell& l, Cell& r, double& nx, double& ny,
                                                   double& px, double& py) {
  const int a = m.ff[f], b = m.fs[f];
  px = m.v2x[f] - m.v1x[f];
  py = m.v2y[f] - m.v1y[f];
  nx = m.px[b] - m.px[a];
  ny = m.py[b] - m.py[a];
  const bool fa = a >= 0 && a < m.N, fb = b >= 0 && b < m.N;
  const int oa = fa ? a : m.gorig[a - m.N];
  const int ob = fb ? b : m.gorig[b - m.N];
  const bool mirror = oa == ob;  // GetOriginalIndex equal on both sides
  // Reflect (geometry.cpp:117-120) of a side's velocity about the face
  auto reflect = [&](const Cell& s) {
    Cell o = s;
    const double s2 = 2 * (s.vx * px + s.vy * py);
    const double ap = vabs(px, py);
    const double a2 = ap * ap;
    o.vx = px * s2 / a2 - s.vx;
    o.vy = py * s2 / a2 - s.vy;
    return o;
  };
  if (!fa) {
    r = ld_cell(q, b);
    l = mirror ? reflect(r) : fetch(m, q, a);
  } else if (!fb) {
    l = ld_cell(q, a);
    r = mirror ? reflect(l) : fetch(m, q, b);
  } else {
    l = ld_cell(q, a);
    r = ld_cell(q, b);
  }
}

// SimpleFluxCalculator body (simple_flux_calculator.cpp:80-126,167-189).
RG_GLOBAL void __launch_bounds__(256) k_flux1(DMesh m, const D4* __restrict__ q,
                                               D4* __restrict__ flux, DParams prm, DErr err) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= m.F) return;
  Cell l, r;
  double nx, ny, px, py;
  simple_face_states(m, q, f, l, r, nx, ny, px, py);
  const double an = vabs(nx, ny), ap = vabs(px, py);
  const double wx = m.wx ? m.wx[f] : 0.0, wy = m.wy ? m.wy[f] : 0.0;
  const double velocity = (nx * wx + ny * wy) / an;
  Cons c;
  const int rc = hllc_face(l, r, prm.gamma, nx, ny, an, px, py, ap, velocity, c);
  D4 o{0, 0, 0, 0};
  if (rc)
    report(err, f, rc);
  else
    o = D4{c.m, c.E, c.mx, c.my};
  flux[f] = o;
}

// The entries of the reference's "Invalid wave speeds in hllc solver"
// error (hllc.cpp:46-71) for face f, recomputed from the inputs of the
// flux call that failed (cells q and, at second order, their slopes rec):
// the states Hllc received (rotated into the face frame, then shifted by
// the face speed), the face speed and the three wave speeds.  One thread.
RG_GLOBAL void k_face_entries(DMesh m, const D4* __restrict__ q, const Rec* __restrict__ rec,
                               int order, double gamma, int f, double* __restrict__ out) {
  if (blockIdx.x != 0 || threadIdx.x != 0 || f < 0 || f >= m.F) return;
  Cell L, R;
  double nx, ny, px, py;
  if (order == 1) {
    simple_face_states(m, q, f, L, R, nx, ny, px, py);
  } else {
    const FaceFrame fr = m.fframe[f];
    if (order == 3) {
      L = fetch(m, q, m.ff[f]);
      R = fetch(m, q, m.fs[f]);
    } else {
      face_states(m, q, rec, m.ff[f], m.fs[f], fr, L, R);
    }
    nx = fr.nx;
    ny = fr.ny;
    px = fr.px;
    py = fr.py;
  }
  const double an = vabs(nx, ny), ap = vabs(px, py);
  const double wx = m.wx ? m.wx[f] : 0.0, wy = m.wy ? m.wy[f] : 0.0;
  const double velocity = (nx * wx + ny * wy) / an;
  const Cell* side[2] = {&L, &R};
  double u[2], c[2];
  for (int k = 0; k < 2; ++k) {
    const Cell& S = *side[k];
    c[k] = dp2c(gamma, S.d, S.p);
    u[k] = (S.vx * nx + S.vy * ny) / an - 1.0 * velocity;
    out[6 * k + 0] = S.d;
    out[6 * k + 1] = S.p;
    out[6 * k + 2] = u[k];
    out[6 * k + 3] = (S.vx * px + S.vy * py) / ap - 0.0 * velocity;
    out[6 * k + 4] = c[k];
    out[6 * k + 5] = dp2e(gamma, S.d, S.p);
  }
  const double sl = smin(u[0] - c[0], u[1] - c[1]);
  const double sr = smax(u[0] + c[0], u[1] + c[1]);
  const double ss = (R.p - L.p + L.d * u[0] * (sl - u[0]) - R.d * u[1] * (sr - u[1])) /
                    (L.d * (sl - u[0]) - R.d * (sr - u[1]));
  out[12] = velocity;
  out[13] = sl;
  out[14] = ss;
  out[15] = sr;
}

// ------------------------------------------------------------------- CFL --
// TimeStepCalculator::operator[] (simple_cfl.cpp:28-47) for one cell.
__device__ __forceinline__ double cfl_inv(const DMesh& m, int i, double gamma, double width,
                                          const Cell& c) {
  const double cs = dp2c(gamma, c.d, c.p);
  double res = 0;
  if (m.wx) {
    for (int k = m.off[i]; k < m.off[i + 1]; ++k) {
      const int f = static_cast<int>(m.inc[k].info & INFO_FACE);
      res = fmax(res, (cs + vabs(c.vx - m.wx[f], c.vy - m.wy[f])) / width);
    }
  } else if (m.off[i + 1] > m.off[i]) {
    // all face velocities zero: every face gives the same term
    res = fmax(res, (cs + vabs(c.vx - 0.0, c.vy - 0.0)) / width);
  }
  return 1.0 / res;
}

RG_GLOBAL void __launch_bounds__(256) k_cfl(DMesh m, const D4* __restrict__ q, double gamma,
                                             double* partials) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double v = __longlong_as_double(static_cast<long long>(DBL_INF_BITS));
  if (i < m.N) v = cfl_inv(m, i, gamma, m.width[i], ld_cell(q, i));
  block_min(v, partials);
}

RG_GLOBAL void k_fill_u64(unsigned long long* p, unsigned long long v, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// ---------------------------------------------------------------- update --
enum : int {
  UPD_ZF_BEFORE = 1,  // ZeroForce before the face sum (TimeAdvance2Heun)
  UPD_ZF_AFTER = 2,   // ZeroForce after the face sum (TimeAdvance)
  UPD_AVERAGE = 4,    // 0.5*(mid + old) (average_extensive)
  UPD_CELLS = 8,      // SimpleCellUpdater
  UPD_CFL = 16,       // next step's CFL term
  UPD_FACES = 32,     // apply the face fluxes
  UPD_DELTA = 64,     // the face array holds dt*area*F already
};

struct UpdArgs {
  const D4* in;
  D4* out;
  const D4* avg;
  const D4* flux;
  D4* cells;
  double dt;                          // used when dt_slot is null
  const unsigned long long* dt_slot;  // min_i 1/res_i of the CFL reduction
  unsigned long long* reset_slot;     // set to +inf (next step's slot)
  double* partials;                   // per-block CFL minima
  double* dt_hist;
  int step;
  int flags;
};

// SimpleExtensiveUpdater (simple_extensive_updater.cpp:11-35) as a per-cell
// gather in ascending face order, ExternalForceContribution with ZeroForce
// (hydrodynamics_2d.cpp:348-372), average_extensive (hdsim2d.cpp:311-320),
// regular_update (simple_cell_updater.cpp:22-59), TimeStepCalculator.
// One cell with a given dt; returns its CFL inverse time step when
// UPD_CFL is set (else +inf).  kEarly issues the loads that do not depend on
// the face sum (Heun operand, volume, width, error word) before it, so that
// they do not wait behind the stores (more registers: the one-cell-per-thread
// kernels only).  With kTileFlux the face terms dt*area*F come from the
// tile's LDS array `tflux` through the per-slot words `eloc` (entry within
// the tile | TILE_SECOND) instead of the face array a.flux.
// Slots per chunk of the tile form's update (4 measured slower: 200.8 ->
// 206.5 us on a 2M-cell mesh, at the same register count).
constexpr int CH_TILE = CH_UPD;

// The loads of a cell's update that do not depend on the face sum (kEarly):
// extensives, Heun operand, volume, width, error word and, for the tile
// form, the first CH_TILE slot words.
struct UpdLoads {
  D4 x, o;
  double vol, wid;
  int deg, eb;
  bool frozen;
  unsigned w[CH_TILE];
};

template <bool kTileFlux = false>
__device__ __forceinline__ UpdLoads update_loads(const DMesh& m, const UpdArgs& a, const DErr& err,
                                                 int i, const unsigned short* __restrict__ eloc) {
  UpdLoads L;
  L.o = D4{0, 0, 0, 0};
  L.vol = L.wid = 0;
  if (a.flags & UPD_AVERAGE) L.o = a.avg[i];
  if (a.flags & UPD_CELLS) {
    L.vol = m.vol[i];
    if (a.flags & UPD_CFL) L.wid = m.width[i];
  }
  L.frozen = *err.word != ~0ull;
  L.x = a.in[i];
  L.deg = m.off[i + 1] - m.off[i];
  L.eb = ell_base(m, i);
  if (kTileFlux) {
#pragma unroll
    for (int u = 0; u < CH_TILE; ++u) L.w[u] = u < L.deg ? eloc[L.eb + 64 * u] : 0u;
  }
  return L;
}

template <bool kEarly = false, bool kTileFlux = false>
__device__ __forceinline__ double update_cell(const DMesh& m, const UpdArgs& a,
                                              const DParams& prm, const DErr& err, int i,
                                              double dt,
                                              const unsigned short* __restrict__ eloc = nullptr,
                                              const D4* tflux = nullptr,
                                              const UpdLoads* pre = nullptr) {
  double cfl_v = __longlong_as_double(static_cast<long long>(DBL_INF_BITS));
  // After a device-detected error the updates stop: the state stays as the
  // failing call left it (the reference throws there), so that the error's
  // entries can be recomputed from it.
  bool frozen = false;
  D4 o{0, 0, 0, 0};
  double vol = 0, wid = 0;
  UpdLoads L{};
  if (kEarly) {
    L = pre ? *pre : update_loads<kTileFlux>(m, a, err, i, eloc);
    o = L.o;
    vol = L.vol;
    wid = L.wid;
    frozen = L.frozen;
  } else if (*err.word != ~0ull) {
    return cfl_v;
  }
  {
    const D4 x = kEarly ? L.x : a.in[i];
    double em = x.a, eE = x.b, ex = x.c, ey = x.d;
    if (a.flags & UPD_ZF_BEFORE) {
      em += dt * 0.0;
      ex += 0.0 * dt;
      ey += 0.0 * dt;
      eE += dt * 0.0;
    }
    if (a.flags & UPD_FACES) {
      const int deg = kEarly ? L.deg : m.off[i + 1] - m.off[i];
      const int eb = kEarly ? L.eb : ell_base(m, i);
      constexpr int CH_UPD = kTileFlux ? CH_TILE : ::rg::CH_UPD;
      for (int s0 = 0; s0 < deg; s0 += CH_UPD) {
        unsigned inf[CH_UPD];
        D4 fl[CH_UPD];
        double ar[CH_UPD];
        const bool delta = kTileFlux || (a.flags & UPD_DELTA) != 0;
#pragma unroll
        for (int u = 0; u < CH_UPD; ++u)
          if (s0 + u < deg) {
            if (kTileFlux) {
              const unsigned w = (kEarly && s0 == 0) ? L.w[u] : eloc[eb + 64 * (s0 + u)];
              inf[u] = (w & TILE_ENTRY) | ((w & TILE_SECOND) ? INFO_SECOND : 0u);
            } else {
              inf[u] = m.einfo[eb + 64 * (s0 + u)];
            }
            ar[u] = delta ? 0.0 : m.earea[eb + 64 * (s0 + u)];
          }
#pragma unroll
        for (int u = 0; u < CH_UPD; ++u)
          if (s0 + u < deg) fl[u] = kTileFlux ? tflux[inf[u] & INFO_FACE] : a.flux[inf[u] & INFO_FACE];
#pragma unroll
        for (int u = 0; u < CH_UPD; ++u) {
          if (s0 + u >= deg) break;
          double dm, dE, dx, dy;
          if (delta) {
            dm = fl[u].a;
            dE = fl[u].b;
            dx = fl[u].c;
            dy = fl[u].d;
          } else {
            const double s = dt * ar[u];
            dm = fl[u].a * s;
            dE = fl[u].b * s;
            dx = fl[u].c * s;
            dy = fl[u].d * s;
          }
          if (inf[u] & INFO_SECOND) {
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
    }
    if (a.flags & UPD_ZF_AFTER) {
      em += dt * 0.0;
      ex += 0.0 * dt;
      ey += 0.0 * dt;
      eE += dt * 0.0;
    }
    if (a.flags & UPD_AVERAGE) {
      if (!kEarly) o = a.avg[i];
      em = 0.5 * (em + o.a);
      eE = 0.5 * (eE + o.b);
      ex = (ex + o.c) * 0.5;
      ey = (ey + o.d) * 0.5;
    }
    if (kEarly && frozen) return cfl_v;
    a.out[i] = D4{em, eE, ex, ey};
    if (a.flags & UPD_CELLS) {
      const double d = em / (kEarly ? vol : m.vol[i]);
      const double vx = ex / em, vy = ey / em;
      const double e = eE / em - 0.5 * (vx * vx + vy * vy);
      if (d < 0)
        report(err, i, ERR_NEGATIVE_DENSITY);
      else if (e < 0)
        report(err, i, ERR_NEGATIVE_THERMAL);
      const double p = (prm.gamma - 1) * e * d;
      a.cells[i] = D4{d, p, vx, vy};
      if (a.flags & UPD_CFL)
        cfl_v = cfl_inv(m, i, prm.gamma, kEarly ? wid : m.width[i], Cell{d, p, vx, vy});
    }
  }
  return cfl_v;
}

RG_GLOBAL void __launch_bounds__(256) k_update(DMesh m, UpdArgs a, DParams prm, DErr err) {
  const int i = walk_block(prm) * blockDim.x + threadIdx.x;
  const double dt = a.dt_slot ? prm.cfl * __longlong_as_double(static_cast<long long>(*a.dt_slot))
                              : a.dt;
  if (i == 0) {
    if (a.dt_hist) a.dt_hist[a.step] = dt;
    if (a.reset_slot) *a.reset_slot = DBL_INF_BITS;
  }
  double cfl_v = __longlong_as_double(static_cast<long long>(DBL_INF_BITS));
  if (i < m.N) cfl_v = update_cell<true>(m, a, prm, err, i, dt);
  if (a.flags & UPD_CFL) block_min(cfl_v, a.partials);
}

// One stage of the Heun step loop in one launch per tile of kTile cells
// (second order, dt*area*F): ModularFluxCalculator over LinearGaussImproved
// on every face of the tile (flux2_face's operations, in the same order) into
// LDS, then k_update's cell update of the tile's cells from those fluxes.
// The face kernel's output never leaves the chip.  A face on a tile boundary
// is computed by both tiles, with equal results.  The stage must not write
// what its faces read: the update's outputs (a.out, a.cells) are other
// buffers than the cells q and slopes rec.
RG_GLOBAL void __launch_bounds__(kTile) k_face_update(DMesh m, DTiles tl, const D4* __restrict__ q,
                                                     const Rec* __restrict__ rec, UpdArgs a,
                                                     DParams prm, DErr errf, DErr erru) {
  extern __shared__ D4 tflux[];
  const int t = walk_block(prm);
  const double dt = prm.cfl * __longlong_as_double(static_cast<long long>(*a.dt_slot));
  const int e0 = tl.tofs[t], ne = tl.tofs[t + 1] - e0;
  // each thread's next entry is loaded while the current one is solved
  FaceFrame frn{0, 0, 0, 0, 0, 0};
  FaceSides sdn{0, 0};
  if (static_cast<int>(threadIdx.x) < ne) {
    frn = tl.tframe[e0 + threadIdx.x];
    sdn = tl.tside[e0 + threadIdx.x];
  }
  for (int k = threadIdx.x; k < ne; k += blockDim.x) {
    const int e = e0 + k;
    const FaceFrame fr = frn;
    const FaceSides sd = sdn;
    if (k + static_cast<int>(blockDim.x) < ne) {
      frn = tl.tframe[e + blockDim.x];
      sdn = tl.tside[e + blockDim.x];
    }
    Cell lf, rt;
    face_states(m, q, rec, sd.a, sd.b, fr, lf, rt);
    const double an = vabs(fr.nx, fr.ny), ap = vabs(fr.px, fr.py);
    // modular_flux_calculator.cpp:56-67
    double speed;
    if (m.wx) {
      const int f = tl.tface[e];
      speed = (fr.nx * m.wx[f] + fr.ny * m.wy[f]) / an;
    } else {
      speed = zero_face_speed(fr.nx, fr.ny, an);
    }
    Cons c;
    const int rc = hllc_face(lf, rt, prm.gamma, fr.nx, fr.ny, an, fr.px, fr.py, ap, speed, c);
    D4 o{0, 0, 0, 0};
    if (rc) {
      report(errf, tl.tface[e], rc);
    } else {
      const double sc = dt * ap;  // the face area |v1 - v2| (physical_geometry.cpp:8-11)
      o = D4{c.m * sc, c.E * sc, c.mx * sc, c.my * sc};
    }
    tflux[k] = o;
  }
  // the update's loads that need no flux are in flight across the barrier
  const int i = t * kTile + threadIdx.x;
  UpdLoads L{};
  if (i < m.N) L = update_loads<true>(m, a, erru, i, tl.eloc);
  __syncthreads();
  if (i == 0) {
    if (a.dt_hist) a.dt_hist[a.step] = dt;
    if (a.reset_slot) *a.reset_slot = DBL_INF_BITS;
  }
  double cfl_v = __longlong_as_double(static_cast<long long>(DBL_INF_BITS));
  if (i < m.N) cfl_v = update_cell<true, true>(m, a, prm, erru, i, dt, tl.eloc, tflux, &L);
  if (a.flags & UPD_CFL) block_min(cfl_v, a.partials);
}

// init_extensives (hdsim2d.cpp:41-66)
RG_GLOBAL void __launch_bounds__(256) k_init_ext(DMesh m, const D4* __restrict__ q, D4* x,
                                                  double gamma) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m.N) return;
  const double volume = m.vol[i];
  const D4 c = q[i];
  const double mass = volume * c.a;
  x[i] = D4{mass, dp2e(gamma, c.a, c.b) * mass + 0.5 * mass * (c.c * c.c + c.d * c.d), c.c * mass,
            c.d * mass};
}

// SoA <-> AoS for the host API
RG_GLOBAL void k_pack4(const double* __restrict__ s, D4* __restrict__ d, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] = D4{s[i], s[n + i], s[2 * static_cast<size_t>(n) + i], s[3 * static_cast<size_t>(n) + i]};
}
RG_GLOBAL void k_unpack4(const D4* __restrict__ s, double* __restrict__ d, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const D4 v = s[i];
    d[i] = v.a;
    d[n + i] = v.b;
    d[2 * static_cast<size_t>(n) + i] = v.c;
    d[3 * static_cast<size_t>(n) + i] = v.d;
  }
}

}  // namespace rg
