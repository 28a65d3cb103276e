// Delaunay triangulation of points in the plane, for rebuilding a moving
// mesh on the host every step.
//
// Radial sweep, then edge flips:
//  1. seed triangle: the point nearest the bounding-box centre, its nearest
//     neighbour and the third point giving the smallest circumcircle;
//  2. the other points in order of distance from that circumcentre, each
//     joined to every hull edge it sees (a point farther than all earlier
//     ones lies outside their hull), the hull kept as a ring with an
//     angular hash for finding a visible edge; the edges facing the new
//     point are then flipped until locally Delaunay (Lawson);
//  3. a final pass of flips over all interior edges.
// Orientation is decided exactly (floating-point filter, exact expansion
// arithmetic when the filter cannot decide).  An in-circle test that the
// floating-point filter cannot decide keeps the current edge, so flipping
// always ends; such a quad is cocircular to round-off and either diagonal is
// a Delaunay triangulation.  Failures (duplicate points, a point the sweep
// cannot place, a flip budget exceeded) return an error so the caller can
// use another triangulator.
#include "delaunay.hpp"

#include "parallel.hpp"

#include "../../../include/richmesh.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <numeric>
#include <vector>

namespace {

constexpr double kEps = 1.1102230246251565e-16;  // 2^-53
constexpr double kCcwBound = (3.0 + 16.0 * kEps) * kEps;
constexpr double kIccBound = (10.0 + 96.0 * kEps) * kEps;

// a + b = x + y exactly, x = fl(a + b)
inline void two_sum(double a, double b, double& x, double& y) {
  x = a + b;
  const double bv = x - a;
  const double av = x - bv;
  y = (a - av) + (b - bv);
}

// a * b = x + y exactly
inline void two_prod(double a, double b, double& x, double& y) {
  x = a * b;
  y = std::fma(a, b, -x);
}

// Sign of the exact sum of n doubles: grow a nonoverlapping expansion one
// term at a time; its largest nonzero component carries the sign.
int exact_sign(const double* t, int n) {
  double e[32];
  int m = 0;
  for (int i = 0; i < n; ++i) {
    double q = t[i];
    int k = 0;
    for (int j = 0; j < m; ++j) {
      double s, r;
      two_sum(q, e[j], s, r);
      q = s;
      if (r != 0.0) e[k++] = r;
    }
    if (q != 0.0) e[k++] = q;
    m = k;
  }
  for (int j = m - 1; j >= 0; --j)
    if (e[j] != 0.0) return e[j] > 0 ? 1 : -1;
  return 0;
}

// > 0 if a, b, c turn counter-clockwise, < 0 clockwise, 0 collinear (exact)
int orient(double ax, double ay, double bx, double by, double cx, double cy) {
  const double l = (ax - cx) * (by - cy);
  const double r = (ay - cy) * (bx - cx);
  const double det = l - r;
  const double sum = std::fabs(l) + std::fabs(r);
  if (std::fabs(det) > kCcwBound * sum) return det > 0 ? 1 : -1;
  // ax by - ax cy - cx by - ay bx + ay cx + cy bx
  double t[12];
  two_prod(ax, by, t[0], t[1]);
  two_prod(-ax, cy, t[2], t[3]);
  two_prod(-cx, by, t[4], t[5]);
  two_prod(-ay, bx, t[6], t[7]);
  two_prod(ay, cx, t[8], t[9]);
  two_prod(cy, bx, t[10], t[11]);
  return exact_sign(t, 12);
}

// +1 if d is certainly inside the circumcircle of the counter-clockwise
// triangle a, b, c; 0 if floating point cannot tell or it is not inside.
int in_circle(double ax, double ay, double bx, double by, double cx, double cy, double dx,
              double dy) {
  const double adx = ax - dx, ady = ay - dy, bdx = bx - dx, bdy = by - dy;
  const double cdx = cx - dx, cdy = cy - dy;
  const double bc1 = bdx * cdy, cb1 = cdx * bdy, al = adx * adx + ady * ady;
  const double ca1 = cdx * ady, ac1 = adx * cdy, bl = bdx * bdx + bdy * bdy;
  const double ab1 = adx * bdy, ba1 = bdx * ady, cl = cdx * cdx + cdy * cdy;
  const double det = al * (bc1 - cb1) + bl * (ca1 - ac1) + cl * (ab1 - ba1);
  const double perm = (std::fabs(bc1) + std::fabs(cb1)) * al +
                      (std::fabs(ca1) + std::fabs(ac1)) * bl +
                      (std::fabs(ab1) + std::fabs(ba1)) * cl;
  return det > kIccBound * perm ? 1 : 0;
}

struct Sweep {
  const double *x, *y;
  int n;
  std::vector<int32_t> tri, half;  // 3 per triangle; half = opposite half-edge or -1
  std::vector<int32_t> hnext, hprev, htri;
  std::vector<char> on_hull;
  std::vector<int32_t> hash;
  double cx = 0, cy = 0;

  int key(double px, double py) const {
    const double dx = px - cx, dy = py - cy;
    const double s = std::fabs(dx) + std::fabs(dy);
    const double p = s > 0 ? dx / s : 0.0;
    const double a = (dy > 0 ? 3.0 - p : 1.0 + p) / 4.0;  // pseudo-angle in [0, 1]
    const int h = static_cast<int>(hash.size());
    return static_cast<int>(std::floor(a * h)) % h;
  }

  int add(int a, int b, int c, int ha, int hb, int hc) {
    const int t = static_cast<int>(tri.size());
    tri.push_back(a);
    tri.push_back(b);
    tri.push_back(c);
    half.push_back(ha);
    half.push_back(hb);
    half.push_back(hc);
    if (ha >= 0) half[ha] = t;
    if (hb >= 0) half[hb] = t + 1;
    if (hc >= 0) half[hc] = t + 2;
    return t;
  }

  bool vis(int p, int e, int q) const {  // hull edge e -> q seen from p
    return orient(x[e], y[e], x[q], y[q], x[p], y[p]) < 0;
  }

  std::vector<int32_t> stack;
  int64_t flips = 0, budget = 0;

  // The twin's opposite point lies in the circumcircle of e's triangle.
  bool illegal(int e) const {
    const int f = half[e];
    if (f < 0) return false;
    const int t = e - e % 3, k = e % 3;
    const int a = tri[e], b = tri[t + (k + 1) % 3], c = tri[t + (k + 2) % 3];
    const int d = tri[f - f % 3 + (f % 3 + 2) % 3];
    return in_circle(x[a], y[a], x[b], y[b], x[c], y[c], x[d], y[d]) != 0;
  }
  // Flip half-edge e (and its twin) if it is illegal; pushes the edges to
  // check next: the two facing the apex of e's triangle (insertion) or all
  // four outer edges.
  int flip(int e, bool all);
  int drain(bool all) {
    while (!stack.empty()) {
      const int e = stack.back();
      stack.pop_back();
      if (flip(e, all) != 0) return -4;
    }
    return 0;
  }
  int run();
};

int Sweep::flip(int e, bool all) {
  if (!illegal(e)) return 0;
  const int f = half[e];
  const int t = e - e % 3, k = e % 3;
  const int u = f - f % 3, j = f % 3;
  const int a = tri[e], b = tri[t + (k + 1) % 3], c = tri[t + (k + 2) % 3];
  const int d = tri[u + (j + 2) % 3];
  if (++flips > budget) return -4;
  const int Tbc = half[t + (k + 1) % 3], Tca = half[t + (k + 2) % 3];
  const int Uad = half[u + (j + 1) % 3], Udb = half[u + (j + 2) % 3];
  // t = (c, a, d), u = (c, d, b); d -> c and c -> d are the new edge
  tri[t] = c;
  tri[t + 1] = a;
  tri[t + 2] = d;
  tri[u] = c;
  tri[u + 1] = d;
  tri[u + 2] = b;
  half[t] = Tca;
  half[t + 1] = Uad;
  half[t + 2] = u;
  half[u] = t + 2;
  half[u + 1] = Udb;
  half[u + 2] = Tbc;
  // an outer edge on the hull moved: keep the ring's half-edge current
  // (while sweeping)
  const bool ring = !htri.empty();
  if (Tca >= 0) half[Tca] = t; else if (ring) htri[c] = t;
  if (Uad >= 0) half[Uad] = t + 1; else if (ring) htri[a] = t + 1;
  if (Udb >= 0) half[Udb] = u + 1; else if (ring) htri[d] = u + 1;
  if (Tbc >= 0) half[Tbc] = u + 2; else if (ring) htri[b] = u + 2;
  stack.push_back(t + 1);
  stack.push_back(u + 1);
  if (all) {
    stack.push_back(t);
    stack.push_back(u + 2);
  }
  return 0;
}

int Sweep::run() {
  double bx0 = x[0], bx1 = x[0], by0 = y[0], by1 = y[0];
  for (int i = 1; i < n; ++i) {
    bx0 = std::min(bx0, x[i]);
    bx1 = std::max(bx1, x[i]);
    by0 = std::min(by0, y[i]);
    by1 = std::max(by1, y[i]);
  }
  const double mx = 0.5 * (bx0 + bx1), my = 0.5 * (by0 + by1);
  auto d2 = [&](int i, double px, double py) {
    const double dx = x[i] - px, dy = y[i] - py;
    return dx * dx + dy * dy;
  };
  int i0 = 0;
  for (int i = 1; i < n; ++i)
    if (d2(i, mx, my) < d2(i0, mx, my)) i0 = i;
  int i1 = -1;
  for (int i = 0; i < n; ++i)
    if (i != i0 && d2(i, x[i0], y[i0]) > 0 &&
        (i1 < 0 || d2(i, x[i0], y[i0]) < d2(i1, x[i0], y[i0])))
      i1 = i;
  if (i1 < 0) return -1;
  auto circum = [&](int a, int b, int c, double& ux, double& uy) {
    const double bx = x[b] - x[a], by = y[b] - y[a], qx = x[c] - x[a], qy = y[c] - y[a];
    const double bl = bx * bx + by * by, ql = qx * qx + qy * qy;
    const double d = 0.5 / (bx * qy - by * qx);
    ux = (qy * bl - by * ql) * d;
    uy = (bx * ql - qx * bl) * d;
  };
  int i2 = -1;
  double best = INFINITY;
  for (int i = 0; i < n; ++i) {
    if (i == i0 || i == i1) continue;
    if (orient(x[i0], y[i0], x[i1], y[i1], x[i], y[i]) == 0) continue;
    double ux, uy;
    circum(i0, i1, i, ux, uy);
    const double r = ux * ux + uy * uy;
    if (r < best) {
      best = r;
      i2 = i;
    }
  }
  if (i2 < 0) return -1;  // all collinear
  if (orient(x[i0], y[i0], x[i1], y[i1], x[i2], y[i2]) < 0) std::swap(i1, i2);
  {
    double ux, uy;
    circum(i0, i1, i2, ux, uy);
    cx = x[i0] + ux;
    cy = y[i0] + uy;
  }
  std::vector<double> dist(static_cast<size_t>(n));
  for (int i = 0; i < n; ++i) dist[i] = d2(i, cx, cy);
  std::vector<int32_t> ids(static_cast<size_t>(n));
  std::iota(ids.begin(), ids.end(), 0);
  std::sort(ids.begin(), ids.end(), [&](int32_t a, int32_t b) {
    return dist[a] < dist[b] || (dist[a] == dist[b] && a < b);
  });

  const size_t ncap = 2 * static_cast<size_t>(n);
  tri.reserve(3 * ncap);
  half.reserve(3 * ncap);
  hnext.assign(static_cast<size_t>(n), -1);
  hprev.assign(static_cast<size_t>(n), -1);
  htri.assign(static_cast<size_t>(n), -1);
  on_hull.assign(static_cast<size_t>(n), 0);
  hash.assign(static_cast<size_t>(std::max(16.0, std::ceil(std::sqrt(static_cast<double>(n))))),
              -1);

  add(i0, i1, i2, -1, -1, -1);
  hnext[i0] = i1;
  hnext[i1] = i2;
  hnext[i2] = i0;
  hprev[i1] = i0;
  hprev[i2] = i1;
  hprev[i0] = i2;
  htri[i0] = 0;  // half-edge i0 -> i1
  htri[i1] = 1;
  htri[i2] = 2;
  on_hull[i0] = on_hull[i1] = on_hull[i2] = 1;
  hash[key(x[i0], y[i0])] = i0;
  hash[key(x[i1], y[i1])] = i1;
  hash[key(x[i2], y[i2])] = i2;

  const int H = static_cast<int>(hash.size());
  for (int k = 0; k < n; ++k) {
    const int p = ids[k];
    if (p == i0 || p == i1 || p == i2) continue;
    const double px = x[p], py = y[p];
    // a hull vertex near p's angle, then the first visible edge from just
    // before it
    int start = -1;
    const int hk = key(px, py);
    for (int j = 0; j < H; ++j) {
      const int s = hash[(hk + j) % H];
      if (s >= 0 && on_hull[s]) {
        start = s;
        break;
      }
    }
    if (start < 0) return -2;
    start = hprev[start];
    int e = start;
    while (!vis(p, e, hnext[e])) {
      e = hnext[e];
      if (e == start) return -3;  // inside the hull or a duplicate
    }
    // the first visible edge e -> q: triangle (e, p, q)
    int q = hnext[e];
    int t = add(e, p, q, -1, -1, htri[e]);
    stack.push_back(t + 2);
    const int first_e = e;
    int back_edge = t;    // half-edge e -> p of the first triangle
    int fwd_edge = t + 1; // half-edge p -> q of the last forward triangle
    // forward along the hull
    while (vis(p, q, hnext[q])) {
      const int r = hnext[q];
      t = add(q, p, r, fwd_edge, -1, htri[q]);
      stack.push_back(t + 2);
      fwd_edge = t + 1;
      on_hull[q] = 0;
      q = r;
    }
    // backward from e
    e = first_e;
    while (vis(p, hprev[e], e)) {
      const int s = hprev[e];
      t = add(s, p, e, -1, back_edge, htri[s]);
      stack.push_back(t + 2);
      back_edge = t;
      on_hull[e] = 0;
      e = s;
    }
    // hull ring: e -> p -> q
    hnext[e] = p;
    hprev[p] = e;
    hnext[p] = q;
    hprev[q] = p;
    htri[e] = back_edge;
    htri[p] = fwd_edge;
    on_hull[p] = 1;
    hash[hk] = p;
    hash[key(x[e], y[e])] = e;
    // edges facing p: flipped until locally Delaunay
    if (drain(false) != 0) return -4;
  }
  // every interior edge once more
  const int T = static_cast<int>(tri.size()) / 3;
  for (int e = 0; e < 3 * T; ++e)
    if (half[e] > e) stack.push_back(e);
  return drain(true);
}

}  // namespace

int rmd::triangulate(int32_t n, const double* x, const double* y, std::vector<int32_t>& tri,
                     std::vector<int32_t>& half) {
  if (n < 3 || !x || !y) return -1;
  for (int32_t i = 0; i < n; ++i)
    if (!std::isfinite(x[i]) || !std::isfinite(y[i])) return -1;
  Sweep s;
  s.x = x;
  s.y = y;
  s.n = n;
  s.budget = 64 * static_cast<int64_t>(n) + 1024;
  const int rc = s.run();
  if (rc != 0) return rc;
  tri.swap(s.tri);
  half.swap(s.half);
  return 0;
}

int rmd::repair(int32_t n, const double* x, const double* y, std::vector<int32_t>& tri,
                std::vector<int32_t>& half) {
  const int32_t T = static_cast<int32_t>(tri.size() / 3);
  if (n < 3 || !x || !y || half.size() != tri.size() || T < 1) return -1;
  // every triangle still counter-clockwise
  {
    const int nc = rmp::chunks(T);
    std::vector<char> bad(static_cast<size_t>(nc), 0);
    rmp::for_chunks(T, nc, [&](int k, int64_t b0, int64_t e0) {
      for (int64_t t = b0; t < e0; ++t) {
        const int32_t a = tri[3 * t], b = tri[3 * t + 1], c = tri[3 * t + 2];
        if (orient(x[a], y[a], x[b], y[b], x[c], y[c]) <= 0) {
          bad[k] = 1;
          return;
        }
      }
    });
    for (char v : bad)
      if (v) return -6;
  }
  // the boundary (edges with no twin, interior on their left) still convex
  // (the few edges with no twin are found in parallel and taken in edge order)
  std::vector<int32_t> hn(static_cast<size_t>(n), -1);
  {
    const int64_t ne = 3 * static_cast<int64_t>(T);
    const int nc = rmp::chunks(ne);
    std::vector<std::vector<int32_t>> open_e(static_cast<size_t>(nc));
    rmp::for_chunks(ne, nc, [&](int k, int64_t b0, int64_t e0) {
      for (int64_t e = b0; e < e0; ++e)
        if (half[e] < 0) open_e[k].push_back(static_cast<int32_t>(e));
    });
    for (const auto& v : open_e)
      for (const int32_t e : v) {
        const int32_t a = tri[e], b = tri[e - e % 3 + (e % 3 + 1) % 3];
        if (hn[a] >= 0) return -7;
        hn[a] = b;
      }
  }
  for (int32_t a = 0; a < n; ++a) {
    if (hn[a] < 0) continue;
    const int32_t b = hn[a], c = hn[b];
    if (c < 0 || orient(x[a], y[a], x[b], y[b], x[c], y[c]) < 0) return -7;
  }
  Sweep s;
  s.x = x;
  s.y = y;
  s.n = n;
  s.budget = 64 * static_cast<int64_t>(n) + 1024;
  s.tri.swap(tri);
  s.half.swap(half);
  // the edges no longer locally Delaunay (found in parallel, kept in edge
  // order), then flips from them
  {
    const int nc = rmp::chunks(3 * static_cast<int64_t>(T));
    std::vector<std::vector<int32_t>> bad(static_cast<size_t>(nc));
    rmp::for_chunks(3 * static_cast<int64_t>(T), nc, [&](int k, int64_t b0, int64_t e0) {
      for (int64_t e = b0; e < e0; ++e)
        if (s.half[e] > e && s.illegal(static_cast<int>(e))) bad[k].push_back(static_cast<int32_t>(e));
    });
    for (int k = 0; k < nc; ++k) s.stack.insert(s.stack.end(), bad[k].begin(), bad[k].end());
  }
  const int rc = s.drain(true);
  tri.swap(s.tri);
  half.swap(s.half);
  return rc;
}

extern "C" int rm_delaunay(int32_t n, const double* x, const double* y, int32_t* tri,
                           int32_t tri_cap, int32_t* n_tri) {
  if (!tri || !n_tri) return -1;
  std::vector<int32_t> t, h;
  const int rc = rmd::triangulate(n, x, y, t, h);
  if (rc != 0) return rc;
  const int32_t T = static_cast<int32_t>(t.size() / 3);
  if (T > tri_cap) return -5;
  std::copy(t.begin(), t.end(), tri);
  *n_tri = T;
  return 0;
}
