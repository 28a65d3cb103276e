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
#include "delaunay.hpp"
#include "parallel.hpp"

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
  auto before = [](const EdgeRef& a, const EdgeRef& b) {
    return a.tlo != b.tlo ? a.tlo < b.tlo : a.k < b.k;
  };
  {
    const int64_t ne = static_cast<int64_t>(edges.size());
    const int nc = rmp::chunks(ne);
    std::vector<char> sorted(static_cast<size_t>(nc), 1);
    rmp::for_chunks(ne, nc, [&](int c, int64_t e0, int64_t e1) {
      for (int64_t i = std::max<int64_t>(e0, 1); i < e1; ++i)
        if (before(edges[i], edges[i - 1])) {
          sorted[c] = 0;
          break;
        }
    });
    if (std::find(sorted.begin(), sorted.end(), 0) != sorted.end())
      std::sort(edges.begin(), edges.end(), before);
  }
  // 6. faces, dropping near-zero-length ones (VoronoiMesh.cpp:287-288)
  uint64_t fs = static_cast<uint64_t>(flip_seed) * 0x9E3779B97F4A7C15ull + 12345u;
  auto long_enough = [&](const EdgeRef& e) {
    const Tri &tl = tris[e.tlo], &th = tris[e.thi];
    const double lx = th.cc.x - tl.cc.x, ly = th.cc.y - tl.cc.y;
    const double len = std::sqrt(lx * lx + ly * ly);
    return len > 1e-8 * std::sqrt(tl.r * th.r);
  };
  if (flip_seed == 0) {
    // kept faces counted per chunk, then written in place
    const int64_t ne = static_cast<int64_t>(edges.size());
    const int nc = rmp::chunks(ne);
    std::vector<int64_t> cnt(static_cast<size_t>(nc) + 1, 0);
    std::vector<char> kept(static_cast<size_t>(ne));
    rmp::for_chunks(ne, nc, [&](int c, int64_t e0, int64_t e1) {
      int64_t k = 0;
      for (int64_t i = e0; i < e1; ++i) {
        kept[i] = long_enough(edges[i]) ? 1 : 0;
        k += kept[i];
      }
      cnt[c + 1] = k;
    });
    for (int c = 0; c < nc; ++c) cnt[c + 1] += cnt[c];
    const size_t F = static_cast<size_t>(cnt[nc]);
    m->ff.resize(F);
    m->fs.resize(F);
    m->v1x.resize(F);
    m->v1y.resize(F);
    m->v2x.resize(F);
    m->v2y.resize(F);
    rmp::for_chunks(ne, nc, [&](int c, int64_t e0, int64_t e1) {
      int64_t f = cnt[c];
      for (int64_t i = e0; i < e1; ++i) {
        if (!kept[i]) continue;
        const EdgeRef& e = edges[i];
        const Tri &tl = tris[e.tlo], &th = tris[e.thi];
        m->ff[f] = e.first;
        m->fs[f] = e.second;
        m->v1x[f] = tl.cc.x;
        m->v1y[f] = tl.cc.y;
        m->v2x[f] = th.cc.x;
        m->v2y[f] = th.cc.y;
        ++f;
      }
    });
  }
  else {
    // a seeded face orientation draws one number per kept face, in order
    m->ff.reserve(edges.size());
    m->fs.reserve(edges.size());
    m->v1x.reserve(edges.size());
    m->v1y.reserve(edges.size());
    m->v2x.reserve(edges.size());
    m->v2y.reserve(edges.size());
    for (const EdgeRef& e : edges) {
      if (!long_enough(e)) continue;
      const Tri &tl = tris[e.tlo], &th = tris[e.thi];
      int32_t first = e.first, second = e.second;
      if (splitmix64(fs) & 1u) std::swap(first, second);
      m->ff.push_back(first);
      m->fs.push_back(second);
      m->v1x.push_back(tl.cc.x);
      m->v1y.push_back(tl.cc.y);
      m->v2x.push_back(th.cc.x);
      m->v2y.push_back(th.cc.y);
    }
  }
  m->F = static_cast<int32_t>(m->ff.size());
  // CSR: each cell's faces in ascending face order
  // (counts and slots taken with atomic adds; each cell's list is then
  // sorted, so the result does not depend on timing)
  m->off.assign(static_cast<size_t>(N) + 1, 0);
  rmp::parallel_for(m->F, [&](int64_t f) {
    if (m->ff[f] < N) __atomic_fetch_add(&m->off[m->ff[f] + 1], 1, __ATOMIC_RELAXED);
    if (m->fs[f] < N) __atomic_fetch_add(&m->off[m->fs[f] + 1], 1, __ATOMIC_RELAXED);
  });
  for (int32_t c = 0; c < N; ++c) m->off[c + 1] += m->off[c];
  m->csr.resize(static_cast<size_t>(m->off[N]));
  {
    std::vector<int32_t> fill(m->off.begin(), m->off.end() - 1);
    rmp::parallel_for(m->F, [&](int64_t f) {
      const int32_t fi = static_cast<int32_t>(f);
      if (m->ff[f] < N) m->csr[__atomic_fetch_add(&fill[m->ff[f]], 1, __ATOMIC_RELAXED)] = fi;
      if (m->fs[f] < N) m->csr[__atomic_fetch_add(&fill[m->fs[f]], 1, __ATOMIC_RELAXED)] = fi;
    });
    rmp::parallel_for(N, [&](int64_t c) {
      std::sort(m->csr.begin() + m->off[c], m->csr.begin() + m->off[c + 1]);
    });
  }

  // 7. cell geometry, in each cell's face order
  m->vol.resize(N);
  m->tvol.resize(N);
  m->cmx.resize(static_cast<size_t>(N) + G);
  m->cmy.resize(static_cast<size_t>(N) + G);
  rmp::parallel_for(N, [&](int64_t cc) {
    const int32_t c = static_cast<int32_t>(cc);
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
  });
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

namespace {

// Faces and geometry of a rigid-box mesh from Delaunay triangles of the
// real points [0, N) and their wall reflections [N, N+G).  With `half`
// (the opposite half-edge of each triangle edge, -1 where none) triangles
// must be counter-clockwise; without it they may have either orientation
// and shared edges are matched by point pair.
rm_mesh* points_mesh(int32_t N, int32_t G, const double* x, const double* y,
                     const int32_t* ghost_orig, const int32_t* gwall, int32_t n_tri,
                     const int32_t* tri, const int32_t* half, double x0, double x1, double y0,
                     double y1) {
  const int32_t NP = N + G;
  rm_mesh* m = new rm_mesh();
  m->N = N;
  m->G = G;
  m->bc = RM_BC_RIGID;
  m->px.assign(x, x + NP);
  m->py.assign(y, y + NP);
  std::vector<int> ghost_wall(static_cast<size_t>(G));
  for (int32_t g = 0; g < G; ++g) {
    if (ghost_orig[g] < 0 || ghost_orig[g] >= N || gwall[g] < 0 || gwall[g] > 3) {
      delete m;
      return nullptr;
    }
    m->gorig.push_back(ghost_orig[g]);
    ghost_wall[g] = gwall[g];
  }
  auto P = [&](int32_t id) { return V2{m->px[id], m->py[id]}; };
  // triangles numbered by their point ids (smallest first, then the next
  // two counter-clockwise), so they, and the faces after them, follow the
  // cell order and depend only on the set of triangles
  // bucketed by the smallest id: counts and slots taken with atomic adds,
  // so a bucket's initial order varies with timing; the sort below orders
  // it completely
  std::vector<int32_t> cnt(static_cast<size_t>(NP) + 1, 0);
  int bad_id = 0;
  rmp::parallel_for(n_tri, [&](int64_t t) {
    const int32_t* q = tri + 3 * static_cast<size_t>(t);
    for (int k = 0; k < 3; ++k)
      if (q[k] < 0 || q[k] >= NP) {
        __atomic_store_n(&bad_id, 1, __ATOMIC_RELAXED);
        return;
      }
    __atomic_fetch_add(&cnt[std::min(q[0], std::min(q[1], q[2])) + 1], 1, __ATOMIC_RELAXED);
  });
  if (bad_id) {
    delete m;
    return nullptr;
  }
  for (int32_t i = 0; i < NP; ++i) cnt[i + 1] += cnt[i];
  std::vector<std::pair<int32_t, int32_t>> order(static_cast<size_t>(n_tri));
  rmp::parallel_for(n_tri, [&](int64_t t) {
    const int32_t* q = tri + 3 * static_cast<size_t>(t);
    const int32_t lo = std::min(q[0], std::min(q[1], q[2]));
    order[__atomic_fetch_add(&cnt[lo], 1, __ATOMIC_RELAXED)] = {lo, static_cast<int32_t>(t)};
  });
  {
    // within one smallest id: by the ids that follow it (orientation as
    // given, which is counter-clockwise where `half` is)
    auto next2 = [&](int32_t t) {
      const int32_t* q = tri + 3 * static_cast<size_t>(t);
      const int k = q[0] <= q[1] && q[0] <= q[2] ? 0 : (q[1] <= q[2] ? 1 : 2);
      return std::make_pair(q[(k + 1) % 3], q[(k + 2) % 3]);
    };
    // each chunk sorts the buckets that start inside it
    const int64_t no = static_cast<int64_t>(order.size());
    rmp::for_chunks(no, rmp::chunks(no), [&](int, int64_t c0, int64_t c1) {
      int64_t b = c0;
      while (b > 0 && b < no && order[b].first == order[b - 1].first) ++b;
      while (b < c1) {
        int64_t e = b + 1;
        while (e < no && order[e].first == order[b].first) ++e;
        std::sort(order.begin() + b, order.begin() + e,
                  [&](const std::pair<int32_t, int32_t>& u, const std::pair<int32_t, int32_t>& v) {
                    const auto a = next2(u.second), b = next2(v.second);
                    return a != b ? a < b : u.second < v.second;  // ties: repeated triangles
                  });
        b = e;
      }
    });
  }
  std::vector<Tri> tris(order.size());
  // with `half`: each triangle's position in the order above, and hr, the
  // position of the triangle across each edge of the reordered triangle
  // (-1 on the hull), read in order by the edge pass below
  std::vector<int32_t> rank(half ? static_cast<size_t>(n_tri) : 0);
  std::vector<int32_t> hr(half ? 3 * order.size() : 0);
  if (half)
    rmp::parallel_for(n_tri, [&](int64_t r) { rank[order[r].second] = static_cast<int32_t>(r); });
  rmp::parallel_for(static_cast<int64_t>(order.size()), [&](int64_t r) {
    const int32_t* q = tri + 3 * static_cast<size_t>(order[r].second);
    int32_t a = q[0], b = q[1], c = q[2];
    if (!half) {
      const V2 B0 = sub(P(b), P(a)), C0 = sub(P(c), P(a));
      if (B0.x * C0.y - B0.y * C0.x < 0) std::swap(b, c);
    }
    // start at the smallest id, so the triangle (its circumcentre's
    // round-off, its edge slots) does not depend on the triangulator
    int k0 = 0;
    if (b < a && b < c) k0 = 1;
    if (c < a && c < b) k0 = 2;
    if (k0 == 1) {
      const int32_t t0 = a;
      a = b;
      b = c;
      c = t0;
    } else if (k0 == 2) {
      const int32_t t0 = c;
      c = b;
      b = a;
      a = t0;
    }
    if (half) {
      const int32_t* hq = half + 3 * static_cast<size_t>(order[r].second);
      for (int k = 0; k < 3; ++k) {
        const int32_t h = hq[(k + k0) % 3];
        hr[3 * static_cast<size_t>(r) + k] = h >= 0 ? rank[h / 3] : -1;
      }
    }
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
    tris[r] = t;
  });
  // shared edges: the first triangle (in the order above) to list an edge
  // owns it; the second one closes it
  std::vector<EdgeRef> edges;
  if (half) {
    // counted per chunk of triangles, then written in place (the
    // sequential order)
    auto owns = [&](int32_t r, int k) {
      const int32_t r2 = hr[3 * static_cast<size_t>(r) + k];
      if (r2 <= r) return false;  // hull edge (-1), or owned by an earlier triangle
      return tris[r].p[k] < N || tris[r].p[(k + 1) % 3] < N;  // two ghosts: not a face
    };
    const int nc = rmp::chunks(n_tri);
    std::vector<int64_t> ecnt(static_cast<size_t>(nc) + 1, 0);
    rmp::for_chunks(n_tri, nc, [&](int c, int64_t r0, int64_t r1) {
      int64_t k = 0;
      for (int64_t r = r0; r < r1; ++r)
        for (int j = 0; j < 3; ++j) k += owns(static_cast<int32_t>(r), j) ? 1 : 0;
      ecnt[c + 1] = k;
    });
    for (int c = 0; c < nc; ++c) ecnt[c + 1] += ecnt[c];
    edges.resize(static_cast<size_t>(ecnt[nc]));
    rmp::for_chunks(n_tri, nc, [&](int c, int64_t r0, int64_t r1) {
      int64_t o = ecnt[c];
      for (int64_t rr = r0; rr < r1; ++rr) {
        const int32_t r = static_cast<int32_t>(rr);
        for (int k = 0; k < 3; ++k)
          if (owns(r, k))
            edges[o++] = EdgeRef{r, hr[3 * static_cast<size_t>(r) + k], tris[r].p[k],
                                 tris[r].p[(k + 1) % 3], k};
      }
    });
  } else {
    edges.reserve(tris.size() * 3 / 2);
    std::unordered_map<uint64_t, std::pair<int32_t, int32_t>> open;  // key -> (triangle, slot)
    open.reserve(tris.size() * 2);
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
  }
  finish_mesh(m, tris, edges, 0, ghost_wall, x0, x1, y0, y1);
  return m;
}

}  // namespace

extern "C" rm_mesh* rm_points_build(const rm_points_params* prm) {
  if (!prm || prm->n_real < 3 || prm->n_ghost < 0 || prm->n_tri < 1 || !prm->x || !prm->y ||
      !prm->tri || (prm->n_ghost > 0 && (!prm->ghost_orig || !prm->ghost_wall)) ||
      !(prm->x1 > prm->x0) || !(prm->y1 > prm->y0))
    return nullptr;
  return points_mesh(prm->n_real, prm->n_ghost, prm->x, prm->y, prm->ghost_orig, prm->ghost_wall,
                     prm->n_tri, prm->tri, nullptr, prm->x0, prm->x1, prm->y0, prm->y1);
}

namespace {

// Reflection of point (px, py) across wall w of the box (0 right, 1 up,
// 2 left, 3 down).
inline void reflect_point(int w, double px, double py, double x0, double x1, double y0, double y1,
                          double& gx, double& gy) {
  gx = w == 0 ? px + 2 * (x1 - px) : w == 2 ? px + 2 * (x0 - px) : px + 0.0;
  gy = w == 1 ? py + 2 * (y1 - py) : w == 3 ? py + 2 * (y0 - py) : py + 0.0;
}

inline bool near_wall(int w, double px, double py, double x0, double x1, double y0, double y1,
                      double mg) {
  return w == 0 ? x1 - px < mg : w == 1 ? y1 - py < mg : w == 2 ? px - x0 < mg : py - y0 < mg;
}

// The points followed by the reflections across each wall of the points
// within the margin of it, wall by wall, points in index order.
void box_ghosts(int32_t N, const double* x, const double* y, double x0, double x1, double y0,
                double y1, double mg, std::vector<double>& X, std::vector<double>& Y,
                std::vector<int32_t>& go, std::vector<int32_t>& gw) {
  X.assign(x, x + N);
  Y.assign(y, y + N);
  go.clear();
  gw.clear();
  for (int w = 0; w < 4; ++w)
    for (int32_t i = 0; i < N; ++i) {
      if (!near_wall(w, x[i], y[i], x0, x1, y0, y1, mg)) continue;
      double gx, gy;
      reflect_point(w, x[i], y[i], x0, x1, y0, y1, gx, gy);
      X.push_back(gx);
      Y.push_back(gy);
      go.push_back(i);
      gw.push_back(w);
    }
}

// The mesh from a triangulation of points and ghosts: the triangles with a
// real vertex and the ghosts they use, ghosts renumbered in their order.
rm_mesh* mesh_of_triangulation(int32_t N, const std::vector<double>& X,
                               const std::vector<double>& Y, const std::vector<int32_t>& go,
                               const std::vector<int32_t>& gw, const std::vector<int32_t>& tri,
                               const std::vector<int32_t>& half, double x0, double x1, double y0,
                               double y1) {
  const int32_t T = static_cast<int32_t>(tri.size() / 3);
  const int32_t G0 = static_cast<int32_t>(go.size());
  // kept triangles (those with a real point) numbered in order: counted
  // per chunk, then numbered in place; the ghosts they use are marked
  std::vector<int32_t> keep(static_cast<size_t>(T), -1), gid(static_cast<size_t>(G0), -1);
  auto kept = [&](int64_t t) {
    const int32_t* q = &tri[3 * static_cast<size_t>(t)];
    return q[0] < N || q[1] < N || q[2] < N;
  };
  const int nc = rmp::chunks(T);
  std::vector<int32_t> kcnt(static_cast<size_t>(nc) + 1, 0);
  int frame_pt = 0;
  rmp::for_chunks(T, nc, [&](int c, int64_t t0, int64_t t1) {
    int32_t n = 0;
    for (int64_t t = t0; t < t1; ++t) {
      if (!kept(t)) continue;
      ++n;
      const int32_t* q = &tri[3 * static_cast<size_t>(t)];
      for (int k = 0; k < 3; ++k) {
        if (q[k] >= N + G0) __atomic_store_n(&frame_pt, 1, __ATOMIC_RELAXED);
        else if (q[k] >= N) __atomic_store_n(&gid[q[k] - N], 0, __ATOMIC_RELAXED);
      }
    }
    kcnt[c + 1] = n;
  });
  if (frame_pt) return nullptr;  // a real point on the hull of its reflections
  for (int c = 0; c < nc; ++c) kcnt[c + 1] += kcnt[c];
  const int32_t nt = kcnt[nc];
  rmp::for_chunks(T, nc, [&](int c, int64_t t0, int64_t t1) {
    int32_t n = kcnt[c];
    for (int64_t t = t0; t < t1; ++t)
      if (kept(t)) keep[t] = n++;
  });
  std::vector<double> px(X.begin(), X.begin() + N), py(Y.begin(), Y.begin() + N);
  std::vector<int32_t> gorig, gwall;
  int32_t G = 0;
  for (int32_t g = 0; g < G0; ++g)
    if (gid[g] == 0) {
      gid[g] = N + G++;
      px.push_back(X[N + g]);
      py.push_back(Y[N + g]);
      gorig.push_back(go[g]);
      gwall.push_back(gw[g]);
    }
  std::vector<int32_t> t2(3 * static_cast<size_t>(nt)), h2(3 * static_cast<size_t>(nt));
  rmp::parallel_for(T, [&](int64_t t) {
    if (keep[t] < 0) return;
    for (int k = 0; k < 3; ++k) {
      const size_t e = 3 * static_cast<size_t>(t) + k, e2 = 3 * static_cast<size_t>(keep[t]) + k;
      const int32_t v = tri[e];
      t2[e2] = v < N ? v : gid[v - N];
      const int32_t h = half[e];
      h2[e2] = (h >= 0 && keep[h / 3] >= 0) ? 3 * keep[h / 3] + h % 3 : -1;
    }
  });
  return points_mesh(N, G, px.data(), py.data(), gorig.data(), gwall.data(), nt, t2.data(),
                     h2.data(), x0, x1, y0, y1);
}

bool box_ok(const rm_box_points_params* p) {
  return p && p->n >= 3 && p->x1 > p->x0 && p->y1 > p->y0 && p->margin > 0;
}

}  // namespace

extern "C" rm_mesh* rm_box_points_build(const rm_box_points_params* prm) {
  if (!box_ok(prm) || !prm->x || !prm->y) return nullptr;
  std::vector<double> X, Y;
  std::vector<int32_t> go, gw, tri, half;
  box_ghosts(prm->n, prm->x, prm->y, prm->x0, prm->x1, prm->y0, prm->y1, prm->margin, X, Y, go,
             gw);
  if (rmd::triangulate(static_cast<int32_t>(X.size()), X.data(), Y.data(), tri, half) != 0)
    return nullptr;
  return mesh_of_triangulation(prm->n, X, Y, go, gw, tri, half, prm->x0, prm->x1, prm->y0,
                               prm->y1);
}

struct rm_mover {
  int32_t N = 0;
  double x0 = 0, x1 = 0, y0 = 0, y1 = 0, margin = 0;
  std::vector<int32_t> go, gw;    // ghosts of the last full build
  std::vector<uint8_t> has;       // [4 N]: point i has a ghost across wall w
  std::vector<double> X, Y;       // points, ghosts and a fixed frame
  std::vector<int32_t> tri, half; // their triangulation
  int32_t builds = 0, repairs = 0;
};

extern "C" rm_mover* rm_mover_create(const rm_box_points_params* prm) {
  if (!box_ok(prm)) return nullptr;
  rm_mover* mv = new rm_mover();
  mv->N = prm->n;
  mv->x0 = prm->x0;
  mv->x1 = prm->x1;
  mv->y0 = prm->y0;
  mv->y1 = prm->y1;
  mv->margin = prm->margin;
  return mv;
}

extern "C" rm_mesh* rm_mover_mesh(rm_mover* mv, const double* x, const double* y) {
  if (!mv || !x || !y) return nullptr;
  const int32_t N = mv->N;
  // finite points; with a triangulation kept, every point near a wall
  // still has its reflection
  const bool kept = !mv->tri.empty();
  int nonfinite = 0, unreflected = 0;
  rmp::parallel_for(N, [&](int64_t i) {
    if (!std::isfinite(x[i]) || !std::isfinite(y[i])) {
      __atomic_store_n(&nonfinite, 1, __ATOMIC_RELAXED);
      return;
    }
    if (!kept) return;
    for (int w = 0; w < 4; ++w)
      if (!mv->has[static_cast<size_t>(w) * N + i] &&
          near_wall(w, x[i], y[i], mv->x0, mv->x1, mv->y0, mv->y1, mv->margin))
        __atomic_store_n(&unreflected, 1, __ATOMIC_RELAXED);
  });
  if (nonfinite) return nullptr;
  bool ok = kept && !unreflected;
  if (ok) {  // the frame stays where it is
    std::copy(x, x + N, mv->X.begin());
    std::copy(y, y + N, mv->Y.begin());
    for (size_t g = 0; g < mv->go.size(); ++g)
      reflect_point(mv->gw[g], x[mv->go[g]], y[mv->go[g]], mv->x0, mv->x1, mv->y0, mv->y1,
                    mv->X[N + g], mv->Y[N + g]);
    ok = rmd::repair(static_cast<int32_t>(mv->X.size()), mv->X.data(), mv->Y.data(), mv->tri,
                     mv->half) == 0;
    if (ok) ++mv->repairs;
  }
  if (!ok) {
    box_ghosts(N, x, y, mv->x0, mv->x1, mv->y0, mv->y1, mv->margin, mv->X, mv->Y, mv->go, mv->gw);
    // a fixed frame of points around the ghosts (a square ring two mean
    // spacings beyond them, one spacing apart): the convex hull never
    // changes as points move, and the outermost ghosts are joined to the
    // frame by well-shaped triangles instead of nearly flat ones
    const double h = std::sqrt((mv->x1 - mv->x0) * (mv->y1 - mv->y0) / N);
    const double off = mv->margin + 2.0 * h;
    const double fx0 = mv->x0 - off, fx1 = mv->x1 + off, fy0 = mv->y0 - off, fy1 = mv->y1 + off;
    const int nx = static_cast<int>(std::ceil((fx1 - fx0) / h));
    const int ny = static_cast<int>(std::ceil((fy1 - fy0) / h));
    for (int i = 0; i < nx; ++i) {  // bottom and top, corners included
      const double fx = fx0 + (fx1 - fx0) * i / nx;
      mv->X.push_back(fx);
      mv->Y.push_back(fy0);
      mv->X.push_back(fx1 - (fx1 - fx0) * i / nx);
      mv->Y.push_back(fy1);
    }
    for (int j = 0; j < ny; ++j) {  // right and left
      mv->X.push_back(fx1);
      mv->Y.push_back(fy0 + (fy1 - fy0) * j / ny);
      mv->X.push_back(fx0);
      mv->Y.push_back(fy1 - (fy1 - fy0) * j / ny);
    }
    mv->has.assign(4 * static_cast<size_t>(N), 0);
    for (size_t g = 0; g < mv->go.size(); ++g)
      mv->has[static_cast<size_t>(mv->gw[g]) * N + mv->go[g]] = 1;
    if (rmd::triangulate(static_cast<int32_t>(mv->X.size()), mv->X.data(), mv->Y.data(), mv->tri,
                         mv->half) != 0) {
      mv->tri.clear();
      mv->half.clear();
      return nullptr;
    }
    ++mv->builds;
  }
  return mesh_of_triangulation(N, mv->X, mv->Y, mv->go, mv->gw, mv->tri, mv->half, mv->x0,
                               mv->x1, mv->y0, mv->y1);
}

extern "C" int rm_mover_permute(rm_mover* mv, const int32_t* order) {
  if (!mv || !order) return -1;
  const int32_t N = mv->N;
  std::vector<int32_t> inv(static_cast<size_t>(N), -1);
  for (int32_t i = 0; i < N; ++i) {
    const int32_t o = order[i];
    if (o < 0 || o >= N || inv[static_cast<size_t>(o)] >= 0) return -1;
    inv[static_cast<size_t>(o)] = i;
  }
  if (mv->tri.empty()) return 0;  // nothing kept: the next mesh is built afresh
  // the same triangles, points and ghosts under the new point numbers
  for (auto& t : mv->tri)
    if (t >= 0 && t < N) t = inv[static_cast<size_t>(t)];
  for (auto& g : mv->go) g = inv[static_cast<size_t>(g)];
  std::vector<uint8_t> has(mv->has.size());
  std::vector<double> X(mv->X.begin(), mv->X.begin() + N), Y(mv->Y.begin(), mv->Y.begin() + N);
  for (int32_t i = 0; i < N; ++i) {
    const size_t o = static_cast<size_t>(order[i]);
    for (size_t w = 0; w < 4; ++w) has[w * N + i] = mv->has[w * N + o];
    mv->X[static_cast<size_t>(i)] = X[o];
    mv->Y[static_cast<size_t>(i)] = Y[o];
  }
  mv->has.swap(has);
  return 0;
}

extern "C" void rm_mover_counts(const rm_mover* mv, int32_t* builds, int32_t* repairs) {
  if (builds) *builds = mv ? mv->builds : 0;
  if (repairs) *repairs = mv ? mv->repairs : 0;
}

extern "C" void rm_mover_free(rm_mover* mv) { delete mv; }

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
