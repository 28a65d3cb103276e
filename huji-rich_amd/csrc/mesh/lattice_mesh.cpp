// Jittered-lattice Voronoi mesh builder: produces the flat mesh arrays that a
// RICH VoronoiMesh would hand to the GPU path (see include/richmesh.h for the
// conventions followed and the reference lines they come from).
//
// The Delaunay triangulation of a jittered square lattice (jitter well below
// half the spacing) splits every lattice square along one diagonal, chosen by
// the in-circle test; ghost layers (mirror images for rigid walls, shifted
// copies for periodic boxes) extend the lattice by one row/column on each
// side, which is all a real cell's Voronoi polygon depends on.
#include "../../../include/richmesh.h"

#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <unordered_map>
#include <vector>
#include <algorithm>

namespace {

struct V2 {
  double x, y;
};

inline V2 sub(V2 a, V2 b) { return V2{a.x - b.x, a.y - b.y}; }

uint64_t splitmix64(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// uniform in [-1, 1)
double uniform_pm1(uint64_t& s) {
  const double u = static_cast<double>(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0);
  return 2.0 * u - 1.0;
}

// Hilbert curve index of (x, y) on an n x n grid, n a power of two.
uint64_t hilbert_d(uint64_t n, uint64_t x, uint64_t y) {
  uint64_t d = 0;
  for (uint64_t s = n / 2; s > 0; s /= 2) {
    const uint64_t rx = (x & s) > 0;
    const uint64_t ry = (y & s) > 0;
    d += s * s * ((3 * rx) ^ ry);
    if (ry == 0) {
      if (rx == 1) {
        x = s - 1 - x;
        y = s - 1 - y;
      }
      const uint64_t t = x;
      x = y;
      y = t;
    }
  }
  return d;
}

struct Tri {
  int32_t p[3];  // counter-clockwise point ids
  V2 cc;         // circumcentre
  double r;      // circumradius
};

}  // namespace

struct rm_mesh {
  int32_t N = 0, F = 0, G = 0;
  int32_t bc = RM_BC_RIGID;
  std::vector<int32_t> ff, fs;
  std::vector<double> v1x, v1y, v2x, v2y;
  std::vector<double> px, py, cmx, cmy;
  std::vector<double> vol, tvol;
  std::vector<int32_t> off, csr;
  std::vector<int32_t> gorig;
  std::vector<int32_t> li, lj;
};

namespace {

// point id of lattice site (i, j) with i in [-1, nx], j in [-1, ny]; -1 if absent
struct Sites {
  int nx, ny;
  std::vector<int32_t> id;  // (nx+2)*(ny+2)
  int32_t& at(int i, int j) { return id[static_cast<size_t>(j + 1) * (nx + 2) + (i + 1)]; }
};

}  // namespace

namespace {

struct EdgeRef {
  int32_t tlo, thi;  // triangles on both sides
  int32_t first, second;
  int32_t k;  // edge slot in tlo
};

// Faces, CSR and cell geometry from the Delaunay triangles and their
// shared edges (steps shared by both builders); m holds N, G, bc, the points
// and ghost_orig.
void finish_mesh(rm_mesh* m, const std::vector<Tri>& tris, std::vector<EdgeRef>& edges,
                 int flip_seed, const std::vector<int>& ghost_wall, double x0, double x1,
                 double y0, double y1) {
  const int32_t N = m->N, G = m->G;
  std::sort(edges.begin(), edges.end(), [](const EdgeRef& a, const EdgeRef& b) {
    return a.tlo != b.tlo ? a.tlo < b.tlo : a.k < b.k;
  });
  // 6. faces, dropping near-zero-length ones (VoronoiMesh.cpp:287-288)
  uint64_t fs = static_cast<uint64_t>(flip_seed) * 0x9E3779B97F4A7C15ull + 12345u;
  std::vector<std::vector<int32_t>> cell_faces(static_cast<size_t>(N));
  for (auto& cf : cell_faces) cf.reserve(8);
  for (const EdgeRef& e : edges) {
    const Tri &tl = tris[e.tlo], &th = tris[e.thi];
    const double lx = th.cc.x - tl.cc.x, ly = th.cc.y - tl.cc.y;
    const double len = std::sqrt(lx * lx + ly * ly);
    if (!(len > 1e-8 * std::sqrt(tl.r * th.r))) continue;
    int32_t first = e.first, second = e.second;
    if (flip_seed != 0 && (splitmix64(fs) & 1u)) std::swap(first, second);
    const int32_t f = static_cast<int32_t>(m->ff.size());
    m->ff.push_back(first);
    m->fs.push_back(second);
    m->v1x.push_back(tl.cc.x);
    m->v1y.push_back(tl.cc.y);
    m->v2x.push_back(th.cc.x);
    m->v2y.push_back(th.cc.y);
    if (first < N) cell_faces[first].push_back(f);
    if (second < N) cell_faces[second].push_back(f);
  }
  m->F = static_cast<int32_t>(m->ff.size());
  m->off.resize(static_cast<size_t>(N) + 1);
  m->off[0] = 0;
  for (int32_t c = 0; c < N; ++c)
    m->off[c + 1] = m->off[c] + static_cast<int32_t>(cell_faces[c].size());
  m->csr.resize(static_cast<size_t>(m->off[N]));
  for (int32_t c = 0; c < N; ++c)
    std::copy(cell_faces[c].begin(), cell_faces[c].end(), m->csr.begin() + m->off[c]);

  // 7. cell geometry, in each cell's face order
  m->vol.resize(N);
  m->tvol.resize(N);
  m->cmx.resize(static_cast<size_t>(N) + G);
  m->cmy.resize(static_cast<size_t>(N) + G);
  for (int32_t c = 0; c < N; ++c) {
    const int32_t b = m->off[c], e = m->off[c + 1];
    // Tessellation::GetVolume: fan about the generating point
    double area = 0;
    for (int32_t k = b; k < e; ++k) {
      const int32_t f = m->csr[k];
      const double p1x = m->v1x[f] - m->px[c], p1y = m->v1y[f] - m->py[c];
      const double p2x = m->v2x[f] - m->px[c], p2y = m->v2y[f] - m->py[c];
      area += 0.5 * std::fabs(p1x * p2y + p1y * (-p2x));
    }
    m->tvol[c] = area;
    // SlabSymmetry::calcVolume and CalcCellCM: fan about the first vertex
    // of the first face
    double vol = 0, a2 = 0, pcx = 0, pcy = 0;
    if (e > b) {
      const int32_t f0 = m->csr[b];
      const double ax = m->v1x[f0], ay = m->v1y[f0];
      for (int32_t k = b + 1; k < e; ++k) {
        const int32_t f = m->csr[k];
        const double q1x = m->v1x[f] - ax, q1y = m->v1y[f] - ay;
        const double q2x = m->v2x[f] - ax, q2y = m->v2y[f] - ay;
        vol += 0.5 * std::fabs(q1x * q2y + q1y * (-q2x));
        const double at = 0.5 * std::fabs(q1x * q2y + q1y * (-q2x));
        a2 += at;
        const double s = at / 3.;
        pcx += ((ax + m->v1x[f]) + m->v2x[f]) * s;
        pcy += ((ay + m->v1y[f]) + m->v2y[f]) * s;
      }
    }
    m->vol[c] = vol;
    m->cmx[c] = pcx / a2;
    m->cmy[c] = pcy / a2;
  }
  for (int32_t g = 0; g < G; ++g) {
    const int32_t o = m->gorig[g];
    const int32_t id = N + g;
    if (m->bc == RM_BC_RIGID) {
      const int w = ghost_wall[g];
      const double cx = m->cmx[o], cy = m->cmy[o];
      if (w == 0) {
        m->cmx[id] = cx + 2 * (x1 - cx);
        m->cmy[id] = cy + 0.0;
      } else if (w == 1) {
        m->cmx[id] = cx + 0.0;
        m->cmy[id] = cy + 2 * (y1 - cy);
      } else if (w == 2) {
        m->cmx[id] = cx + 2 * (x0 - cx);
        m->cmy[id] = cy + 0.0;
      } else {
        m->cmx[id] = cx + 0.0;
        m->cmy[id] = cy + 2 * (y0 - cy);
      }
    } else {
      m->cmx[id] = m->cmx[o] + (m->px[id] - m->px[o]);
      m->cmy[id] = m->cmy[o] + (m->py[id] - m->py[o]);
    }
  }
}

}  // namespace

extern "C" rm_mesh* rm_lattice_build(const rm_lattice_params* prm) {
  if (!prm || prm->nx < 2 || prm->ny < 2 || !(prm->x1 > prm->x0) || !(prm->y1 > prm->y0) ||
      !(prm->jitter >= 0.0) || prm->jitter > 0.05 ||
      (prm->bc != RM_BC_RIGID && prm->bc != RM_BC_PERIODIC))
    return nullptr;
  const int nx = prm->nx, ny = prm->ny;
  const double dx = (prm->x1 - prm->x0) / nx;
  const double dy = (prm->y1 - prm->y0) / ny;
  const double Lx = prm->x1 - prm->x0, Ly = prm->y1 - prm->y0;
  const int32_t N = nx * ny;
  rm_mesh* m = new rm_mesh();
  m->N = N;
  m->bc = prm->bc;

  // 1. jittered lattice, generated x-outer / y-inner like cartesian_mesh
  std::vector<V2> lat(static_cast<size_t>(N));
  {
    uint64_t s = prm->seed;
    for (int i = 0; i < nx; ++i)
      for (int j = 0; j < ny; ++j) {
        const double ux = uniform_pm1(s), uy = uniform_pm1(s);
        lat[static_cast<size_t>(i) * ny + j] =
            V2{prm->x0 + (i + 0.5) * dx + ux * prm->jitter * dx,
               prm->y0 + (j + 0.5) * dy + uy * prm->jitter * dy};
      }
  }
  // 2. Hilbert order of the lattice sites -> cell ids
  uint64_t hn = 1;
  while (hn < static_cast<uint64_t>(std::max(nx, ny))) hn *= 2;
  std::vector<std::pair<uint64_t, int32_t>> key(static_cast<size_t>(N));
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j)
      key[static_cast<size_t>(i) * ny + j] = {hilbert_d(hn, i, j), i * ny + j};
  std::sort(key.begin(), key.end());
  Sites site{nx, ny, std::vector<int32_t>(static_cast<size_t>(nx + 2) * (ny + 2), -1)};
  m->px.resize(N);
  m->py.resize(N);
  m->li.resize(N);
  m->lj.resize(N);
  for (int32_t c = 0; c < N; ++c) {
    const int32_t k = key[c].second;
    const int i = k / ny, j = k % ny;
    site.at(i, j) = c;
    m->px[c] = lat[k].x;
    m->py[c] = lat[k].y;
    m->li[c] = i;
    m->lj[c] = j;
  }
  // 3. ghost layers.  Groups as GetReflection numbers the walls
  //    (VoronoiMesh.cpp:316-330): 0 right, 1 up, 2 left, 3 down; periodic
  //    boxes add the four corner copies.
  std::vector<int> ghost_wall;  // wall group of each ghost (rigid)
  auto add_ghost = [&](int i, int j, int32_t orig, V2 p, int wall) {
    const int32_t id = N + static_cast<int32_t>(m->gorig.size());
    site.at(i, j) = id;
    m->gorig.push_back(orig);
    m->px.push_back(p.x);
    m->py.push_back(p.y);
    ghost_wall.push_back(wall);
  };
  if (prm->bc == RM_BC_RIGID) {
    const double R = prm->x1, U = prm->y1, L = prm->x0, D = prm->y0;
    for (int j = 0; j < ny; ++j) {
      const int32_t o = site.at(nx - 1, j);
      add_ghost(nx, j, o, V2{m->px[o] + 2 * (R - m->px[o]), m->py[o] + 0.0}, 0);
    }
    for (int i = 0; i < nx; ++i) {
      const int32_t o = site.at(i, ny - 1);
      add_ghost(i, ny, o, V2{m->px[o] + 0.0, m->py[o] + 2 * (U - m->py[o])}, 1);
    }
    for (int j = 0; j < ny; ++j) {
      const int32_t o = site.at(0, j);
      add_ghost(-1, j, o, V2{m->px[o] + 2 * (L - m->px[o]), m->py[o] + 0.0}, 2);
    }
    for (int i = 0; i < nx; ++i) {
      const int32_t o = site.at(i, 0);
      add_ghost(i, -1, o, V2{m->px[o] + 0.0, m->py[o] + 2 * (D - m->py[o])}, 3);
    }
  } else {
    auto shifted = [&](int i, int j, int wall) {
      const int ii = (i + nx) % nx, jj = (j + ny) % ny;
      const int32_t o = site.at(ii, jj);
      const double sx = i < 0 ? -Lx : (i >= nx ? Lx : 0.0);
      const double sy = j < 0 ? -Ly : (j >= ny ? Ly : 0.0);
      add_ghost(i, j, o, V2{m->px[o] + sx, m->py[o] + sy}, wall);
    };
    for (int j = 0; j < ny; ++j) shifted(nx, j, 0);
    for (int i = 0; i < nx; ++i) shifted(i, ny, 1);
    for (int j = 0; j < ny; ++j) shifted(-1, j, 2);
    for (int i = 0; i < nx; ++i) shifted(i, -1, 3);
    shifted(nx, ny, 4);
    shifted(-1, ny, 5);
    shifted(-1, -1, 6);
    shifted(nx, -1, 7);
  }
  const int32_t G = static_cast<int32_t>(m->gorig.size());
  m->G = G;
  auto P = [&](int32_t id) { return V2{m->px[id], m->py[id]}; };

  // 4. triangulate each lattice square of the extended lattice, squares in
  //    Hilbert order so that triangle (and hence face) ids follow the cells.
  const int sqx = nx + 1, sqy = ny + 1;  // squares (i,j), i in [-1,nx-1]
  uint64_t hs = 1;
  while (hs < static_cast<uint64_t>(std::max(sqx, sqy))) hs *= 2;
  std::vector<std::pair<uint64_t, int32_t>> sq_order(static_cast<size_t>(sqx) * sqy);
  for (int a = 0; a < sqx; ++a)
    for (int b = 0; b < sqy; ++b)
      sq_order[static_cast<size_t>(a) * sqy + b] = {hilbert_d(hs, a, b), a * sqy + b};
  std::sort(sq_order.begin(), sq_order.end());
  // per square: the triangle holding each side (0 bottom ab, 1 right bc,
  // 2 top cd, 3 left da) and the diagonal's two triangles
  std::vector<int32_t> side_tri(static_cast<size_t>(sqx) * sqy * 4, -1);
  std::vector<Tri> tris;
  tris.reserve(static_cast<size_t>(sqx) * sqy * 2);
  struct Diag {
    int32_t t0, t1;
  };
  std::vector<Diag> diag(static_cast<size_t>(sqx) * sqy, Diag{-1, -1});
  auto make_tri = [&](int32_t a, int32_t b, int32_t c) {
    Tri t;
    t.p[0] = a;
    t.p[1] = b;
    t.p[2] = c;
    const V2 p0 = P(a), B = sub(P(b), p0), C = sub(P(c), p0);
    const double d = 2.0 * (B.x * C.y - B.y * C.x);
    const double b2 = B.x * B.x + B.y * B.y, c2 = C.x * C.x + C.y * C.y;
    const double ux = (C.y * b2 - B.y * c2) / d;
    const double uy = (B.x * c2 - C.x * b2) / d;
    t.cc = V2{p0.x + ux, p0.y + uy};
    t.r = std::sqrt(ux * ux + uy * uy);
    tris.push_back(t);
    return static_cast<int32_t>(tris.size() - 1);
  };
  for (const auto& so : sq_order) {
    const int a_ = so.second / sqy, b_ = so.second % sqy;
    const int i = a_ - 1, j = b_ - 1;
    const int32_t A = site.at(i, j), B = site.at(i + 1, j), C = site.at(i + 1, j + 1),
                  D = site.at(i, j + 1);
    int32_t* st = &side_tri[static_cast<size_t>(so.second) * 4];
    const int present = (A >= 0) + (B >= 0) + (C >= 0) + (D >= 0);
    if (present == 4) {
      // in-circle test of D against (A,B,C), coordinates relative to D
      const V2 a = sub(P(A), P(D)), b = sub(P(B), P(D)), c = sub(P(C), P(D));
      const double det = (a.x * a.x + a.y * a.y) * (b.x * c.y - c.x * b.y) -
                         (b.x * b.x + b.y * b.y) * (a.x * c.y - c.x * a.y) +
                         (c.x * c.x + c.y * c.y) * (a.x * b.y - b.x * a.y);
      if (det > 0) {  // D inside circle(ABC): diagonal B-D
        const int32_t t0 = make_tri(A, B, D), t1 = make_tri(B, C, D);
        st[0] = t0;
        st[3] = t0;
        st[1] = t1;
        st[2] = t1;
        diag[so.second] = Diag{t0, t1};
      } else {  // diagonal A-C
        const int32_t t0 = make_tri(A, B, C), t1 = make_tri(A, C, D);
        st[0] = t0;
        st[1] = t0;
        st[2] = t1;
        st[3] = t1;
        diag[so.second] = Diag{t0, t1};
      }
    } else if (present == 3) {
      int32_t q[4] = {A, B, C, D};
      int32_t v[3];
      int k = 0;
      for (int s = 0; s < 4; ++s)
        if (q[s] >= 0) v[k++] = q[s];
      const int32_t t = make_tri(v[0], v[1], v[2]);
      if (A >= 0 && B >= 0) st[0] = t;
      if (B >= 0 && C >= 0) st[1] = t;
      if (C >= 0 && D >= 0) st[2] = t;
      if (D >= 0 && A >= 0) st[3] = t;
    }
  }
  // 5. Delaunay edges -> Voronoi faces.  Each edge is visited from its
  //    lower-numbered triangle, in triangle order and CCW edge order.
  std::vector<EdgeRef> edges;
  edges.reserve(tris.size() * 2);
  auto square_of = [&](int i, int j) -> int32_t {
    if (i < -1 || i > nx - 1 || j < -1 || j > ny - 1) return -1;
    return (i + 1) * sqy + (j + 1);
  };
  auto consider = [&](int32_t ta, int32_t tb, int32_t pa, int32_t pb) {
    if (ta < 0 || tb < 0) return;
    if (pa >= N && pb >= N) return;  // both ghosts: not a legal edge
    const int32_t tlo = std::min(ta, tb), thi = std::max(ta, tb);
    int k = -1;
    for (int s = 0; s < 3; ++s) {
      const int32_t u = tris[tlo].p[s], w = tris[tlo].p[(s + 1) % 3];
      if ((u == pa && w == pb) || (u == pb && w == pa)) k = s;
    }
    if (k < 0) return;
    edges.push_back(EdgeRef{tlo, thi, tris[tlo].p[k], tris[tlo].p[(k + 1) % 3], k});
  };
  for (int i = -1; i <= nx; ++i)
    for (int j = -1; j <= ny; ++j) {
      const int32_t A = site.at(i, j);
      if (A < 0) continue;
      if (i + 1 <= nx) {  // horizontal lattice edge (i,j)-(i+1,j)
        const int32_t B = site.at(i + 1, j);
        const int32_t sb = square_of(i, j - 1), sa = square_of(i, j);
        if (B >= 0 && sb >= 0 && sa >= 0)
          consider(side_tri[static_cast<size_t>(sb) * 4 + 2], side_tri[static_cast<size_t>(sa) * 4 + 0], A, B);
      }
      if (j + 1 <= ny) {  // vertical lattice edge (i,j)-(i,j+1)
        const int32_t B = site.at(i, j + 1);
        const int32_t sl = square_of(i - 1, j), sr = square_of(i, j);
        if (B >= 0 && sl >= 0 && sr >= 0)
          consider(side_tri[static_cast<size_t>(sl) * 4 + 1], side_tri[static_cast<size_t>(sr) * 4 + 3], A, B);
      }
    }
  for (size_t s = 0; s < diag.size(); ++s) {
    const Diag dg = diag[s];
    if (dg.t0 < 0) continue;
    // the shared edge of the two triangles
    int32_t shared[2];
    int n = 0;
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b)
        if (tris[dg.t0].p[a] == tris[dg.t1].p[b] && n < 2) shared[n++] = tris[dg.t0].p[a];
    if (n == 2) consider(dg.t0, dg.t1, shared[0], shared[1]);
  }
  finish_mesh(m, tris, edges, prm->flip_seed, ghost_wall, prm->x0, prm->x1, prm->y0, prm->y1);
  return m;
}

extern "C" rm_mesh* rm_points_build(const rm_points_params* prm) {
  if (!prm || prm->n_real < 3 || prm->n_ghost < 0 || prm->n_tri < 1 || !prm->x || !prm->y ||
      !prm->tri || (prm->n_ghost > 0 && (!prm->ghost_orig || !prm->ghost_wall)) ||
      !(prm->x1 > prm->x0) || !(prm->y1 > prm->y0))
    return nullptr;
  const int32_t N = prm->n_real, G = prm->n_ghost, NP = N + G;
  rm_mesh* m = new rm_mesh();
  m->N = N;
  m->G = G;
  m->bc = RM_BC_RIGID;
  m->px.assign(prm->x, prm->x + NP);
  m->py.assign(prm->y, prm->y + NP);
  std::vector<int> ghost_wall(static_cast<size_t>(G));
  for (int32_t g = 0; g < G; ++g) {
    if (prm->ghost_orig[g] < 0 || prm->ghost_orig[g] >= N || prm->ghost_wall[g] < 0 ||
        prm->ghost_wall[g] > 3) {
      delete m;
      return nullptr;
    }
    m->gorig.push_back(prm->ghost_orig[g]);
    ghost_wall[g] = prm->ghost_wall[g];
  }
  auto P = [&](int32_t id) { return V2{m->px[id], m->py[id]}; };
  // triangles counter-clockwise, numbered in Hilbert order of their centroids
  double bx0 = m->px[0], bx1 = bx0, by0 = m->py[0], by1 = by0;
  for (int32_t i = 1; i < NP; ++i) {
    bx0 = std::min(bx0, m->px[i]);
    bx1 = std::max(bx1, m->px[i]);
    by0 = std::min(by0, m->py[i]);
    by1 = std::max(by1, m->py[i]);
  }
  const uint64_t hn = 1u << 16;
  std::vector<std::pair<uint64_t, int32_t>> order(static_cast<size_t>(prm->n_tri));
  for (int32_t t = 0; t < prm->n_tri; ++t) {
    const int32_t* q = prm->tri + 3 * static_cast<size_t>(t);
    for (int k = 0; k < 3; ++k)
      if (q[k] < 0 || q[k] >= NP) {
        delete m;
        return nullptr;
      }
    const double cx = (m->px[q[0]] + m->px[q[1]] + m->px[q[2]]) / 3.0;
    const double cy = (m->py[q[0]] + m->py[q[1]] + m->py[q[2]]) / 3.0;
    const uint64_t ix = static_cast<uint64_t>((cx - bx0) / (bx1 - bx0 + 1e-300) * (hn - 1));
    const uint64_t iy = static_cast<uint64_t>((cy - by0) / (by1 - by0 + 1e-300) * (hn - 1));
    order[t] = {hilbert_d(hn, ix, iy), t};
  }
  std::sort(order.begin(), order.end());
  std::vector<Tri> tris;
  tris.reserve(order.size());
  for (const auto& o : order) {
    const int32_t* q = prm->tri + 3 * static_cast<size_t>(o.second);
    int32_t a = q[0], b = q[1], c = q[2];
    const V2 B0 = sub(P(b), P(a)), C0 = sub(P(c), P(a));
    if (B0.x * C0.y - B0.y * C0.x < 0) std::swap(b, c);
    Tri t;
    t.p[0] = a;
    t.p[1] = b;
    t.p[2] = c;
    const V2 p0 = P(a), B = sub(P(b), p0), C = sub(P(c), p0);
    const double d = 2.0 * (B.x * C.y - B.y * C.x);
    const double b2 = B.x * B.x + B.y * B.y, c2 = C.x * C.x + C.y * C.y;
    const double ux = (C.y * b2 - B.y * c2) / d;
    const double uy = (B.x * c2 - C.x * b2) / d;
    t.cc = V2{p0.x + ux, p0.y + uy};
    t.r = std::sqrt(ux * ux + uy * uy);
    tris.push_back(t);
  }
  // shared edges: the first triangle (in the order above) to list an edge
  // owns it; the second one closes it
  std::unordered_map<uint64_t, std::pair<int32_t, int32_t>> open;  // key -> (triangle, slot)
  open.reserve(tris.size() * 2);
  std::vector<EdgeRef> edges;
  edges.reserve(tris.size() * 3 / 2);
  for (int32_t t = 0; t < static_cast<int32_t>(tris.size()); ++t)
    for (int k = 0; k < 3; ++k) {
      const int32_t u = tris[t].p[k], w = tris[t].p[(k + 1) % 3];
      const uint64_t key = (static_cast<uint64_t>(std::min(u, w)) << 32) |
                           static_cast<uint32_t>(std::max(u, w));
      auto it = open.find(key);
      if (it == open.end()) {
        open.emplace(key, std::make_pair(t, k));
        continue;
      }
      const int32_t tlo = it->second.first, klo = it->second.second;
      open.erase(it);
      const int32_t a = tris[tlo].p[klo], b = tris[tlo].p[(klo + 1) % 3];
      if (a >= N && b >= N) continue;  // both ghosts: not a face of the mesh
      edges.push_back(EdgeRef{tlo, t, a, b, klo});
    }
  finish_mesh(m, tris, edges, 0, ghost_wall, prm->x0, prm->x1, prm->y0, prm->y1);
  return m;
}

extern "C" void rm_free(rm_mesh* m) { delete m; }

extern "C" void rm_sizes(const rm_mesh* m, int32_t* n_cells, int32_t* n_faces, int32_t* n_ghosts,
                         int32_t* nnz) {
  if (n_cells) *n_cells = m->N;
  if (n_faces) *n_faces = m->F;
  if (n_ghosts) *n_ghosts = m->G;
  if (nnz) *nnz = m->off.empty() ? 0 : m->off.back();
}

namespace {
template <class T>
void put(T* dst, const std::vector<T>& src) {
  if (dst && !src.empty()) std::copy(src.begin(), src.end(), dst);
}
}  // namespace

extern "C" void rm_export(const rm_mesh* m, int32_t* face_first, int32_t* face_second, double* v1x,
                          double* v1y, double* v2x, double* v2y, double* pt_x, double* pt_y,
                          double* cm_x, double* cm_y, double* volume, double* tess_volume,
                          int32_t* csr_off, int32_t* csr_face, int32_t* ghost_orig,
                          int32_t* lat_i, int32_t* lat_j) {
  put(face_first, m->ff);
  put(face_second, m->fs);
  put(v1x, m->v1x);
  put(v1y, m->v1y);
  put(v2x, m->v2x);
  put(v2y, m->v2y);
  put(pt_x, m->px);
  put(pt_y, m->py);
  put(cm_x, m->cmx);
  put(cm_y, m->cmy);
  put(volume, m->vol);
  put(tess_volume, m->tvol);
  put(csr_off, m->off);
  put(csr_face, m->csr);
  put(ghost_orig, m->gorig);
  put(lat_i, m->li);
  put(lat_j, m->lj);
}

extern "C" void rm_face_velocity(const rm_mesh* m, const double* wx, const double* wy, double* fwx,
                                 double* fwy) {
  const int32_t N = m->N;
  auto orig = [&](int32_t id) { return id < N ? id : m->gorig[id - N]; };
  for (int32_t f = 0; f < m->F; ++f) {
    const int32_t a = m->ff[f], b = m->fs[f];
    const int32_t oa = orig(a), ob = orig(b);
    if (m->bc == RM_BC_RIGID && oa == ob) {
      fwx[f] = 0.0;
      fwy[f] = 0.0;
      continue;
    }
    // CalcFaceVelocity(wl, wr, rL, rR, f):
    //   w' = ((wl-wr).(f-(rR+rL)/2)) * (rR-rL) / |rR-rL|^2 ; w = (wl+wr)/2 + w'
    const double wlx = wx[oa], wly = wy[oa], wrx = wx[ob], wry = wy[ob];
    const double rLx = m->px[a], rLy = m->py[a], rRx = m->px[b], rRy = m->py[b];
    const double fx = (m->v1x[f] + m->v2x[f]) * 0.5, fy = (m->v1y[f] + m->v2y[f]) * 0.5;
    const double mx = (rRx + rLx) / 2, my = (rRy + rLy) / 2;
    const double s = (wlx - wrx) * (fx - mx) + (wly - wry) * (fy - my);
    const double dxr = rRx - rLx, dyr = rRy - rLy;
    const double ab = std::sqrt(dxr * dxr + dyr * dyr);
    const double d2 = ab * ab;
    const double wpx = (dxr * s) / d2, wpy = (dyr * s) / d2;
    fwx[f] = (wlx + wrx) * 0.5 + wpx;
    fwy[f] = (wly + wry) * 0.5 + wpy;
  }
}
