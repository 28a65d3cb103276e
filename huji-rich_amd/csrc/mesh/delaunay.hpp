// Delaunay triangulation shared by the point-mesh builders.
#pragma once

#include <cstdint>
#include <vector>

namespace rmd {

// Counter-clockwise triangles (3 point ids each) of n points and, for each
// triangle edge 3t+k (from tri[3t+k] to the next vertex), the opposite
// half-edge, or -1 on the convex hull.  Returns 0, or a negative code when
// the input has duplicate points or cannot be triangulated.
int triangulate(int32_t n, const double* x, const double* y, std::vector<int32_t>& tri,
                std::vector<int32_t>& half);

// Make a triangulation of moved points Delaunay again by edge flips, in
// place.  Returns 0, or a negative code (a triangle turned clockwise, the
// boundary no longer convex, the flip budget spent) when the triangulation
// must be built afresh.
int repair(int32_t n, const double* x, const double* y, std::vector<int32_t>& tri,
           std::vector<int32_t>& half);

}  // namespace rmd
