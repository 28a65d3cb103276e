/*
 * richmesh.h -- C ABI of the jittered-lattice Voronoi mesh builder.
 *
 * In RICH the Voronoi mesh is built on the host by VoronoiMesh
 * (source/tessellation/VoronoiMesh.cpp:333-399 Initialise, :250-302 build_v)
 * and read by the hydro plugins through the Tessellation accessors
 * (tessellation.hpp:21-224).  The GPU path keeps that split: a caller that
 * owns a VoronoiMesh flattens it into an rg_mesh (richgpu.h).  This library
 * produces the same flat arrays for the workloads used to test and measure
 * the path -- jittered square lattices in a rigid-wall box (SquareBox +
 * RigidWallGenerator) or a periodic box (PeriodicBox + PeriodicGhostGenerator)
 * -- following VoronoiMesh's conventions:
 *   - faces are Voronoi edges between two Delaunay circumcentres; faces shorter
 *     than 1e-8*sqrt(R1*R2) are dropped            (VoronoiMesh.cpp:287-288)
 *   - a face is created while visiting the lower-numbered of its two
 *     triangles; neighbours are that triangle's counter-clockwise edge
 *                                                  (VoronoiMesh.cpp:267-297)
 *   - each cell lists its faces in creation order, i.e. ascending face id
 *                                                  (VoronoiMesh.cpp:291-294)
 *   - centres of mass by CalcCellCM, Tessellation volume by GetVolume
 *                                                  (VoronoiMesh.cpp:420-451)
 *   - update volumes by SlabSymmetry::calcVolume   (physical_geometry.cpp:13-21)
 *   - ghost centres of mass mirrored (rigid) or shifted (periodic)
 *                                                  (VoronoiMesh.cpp:359-398)
 * Ghost points are numbered N..N+G-1 (RICH numbers them from N+3; the three
 * super-triangle points are not part of the flat mesh).
 * Cells are numbered in Hilbert order of the lattice (RICH Hilbert-orders the
 * generating points, VoronoiMesh.cpp:337-338).
 */
#ifndef RICHMESH_H
#define RICHMESH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RM_BC_RIGID 0
#define RM_BC_PERIODIC 1

typedef struct rm_lattice_params {
  int32_t nx, ny;          /* lattice points along x and y                 */
  double x0, x1, y0, y1;   /* box [x0,x1] x [y0,y1]                        */
  double jitter;           /* uniform jitter, fraction of the spacing      */
  uint64_t seed;           /* jitter seed                                  */
  int32_t bc;              /* RM_BC_RIGID or RM_BC_PERIODIC                */
  int32_t flip_seed;       /* 0: neighbours as built; else flip the
                              (first,second) order of a random half of the
                              faces, to exercise ghost-first faces         */
} rm_lattice_params;

typedef struct rm_mesh rm_mesh;

/* Build the mesh.  Returns NULL on bad parameters. */
rm_mesh* rm_lattice_build(const rm_lattice_params* p);
void rm_free(rm_mesh* m);

/* A mesh from arbitrary points in a rigid-wall box [x0,x1] x [y0,y1]: the
 * caller supplies the real points [0, n_real), ghost points reflected across
 * a wall (ghost_orig: the real point reflected; ghost_wall: 0 right, 1 up,
 * 2 left, 3 down) and the Delaunay triangles of all of them (any
 * orientation).  Triangles are numbered by their smallest point id;
 * faces, their order and the cell geometry then follow the same rules as
 * rm_lattice_build.  Returns NULL on bad input. */
typedef struct rm_points_params {
  int32_t n_real, n_ghost;
  const double *x, *y;            /* [n_real + n_ghost]                  */
  const int32_t *ghost_orig;      /* [n_ghost]                           */
  const int32_t *ghost_wall;      /* [n_ghost]                           */
  int32_t n_tri;
  const int32_t *tri;             /* [3 n_tri] point ids                 */
  double x0, x1, y0, y1;
} rm_points_params;

rm_mesh* rm_points_build(const rm_points_params* p);

/* The mesh of points in a rigid-wall box in one call (what
 * VoronoiMesh::Initialise, VoronoiMesh.cpp:333-456, builds for a SquareBox):
 * ghosts are the
 * reflections across each wall of the points within `margin` of it (walls
 * in the order 0 right, 1 up, 2 left, 3 down, points in index order), the
 * triangulation is rm_delaunay's, and the triangles with a real vertex and
 * the ghosts they use make the mesh as in rm_points_build.  Returns NULL on
 * bad input or when the points cannot be triangulated (e.g. duplicates). */
typedef struct rm_box_points_params {
  int32_t n;
  const double *x, *y;            /* [n]                                 */
  double x0, x1, y0, y1;
  double margin;
} rm_box_points_params;

rm_mesh* rm_box_points_build(const rm_box_points_params* p);

/* A moving mesh's rebuilds (VoronoiMesh::Update, VoronoiMesh.cpp:458-571,
 * which hdsim calls from MoveMeshPoints every step): rm_mover_mesh gives the
 * mesh of the points' current positions (n, box and margin of rm_mover_create; its x, y are not
 * read), the same mesh rm_box_points_build gives, up to the diagonal of
 * cocircular point quads.  It keeps the previous triangulation of points and
 * ghosts and, when the moved points leave every triangle counter-clockwise,
 * the boundary convex and every point near a wall with its reflection,
 * restores the Delaunay property by edge flips instead of triangulating
 * afresh.  The caller frees each mesh with rm_free. */
typedef struct rm_mover rm_mover;
rm_mover* rm_mover_create(const rm_box_points_params* p);
rm_mesh* rm_mover_mesh(rm_mover* mv, const double* x, const double* y);
/* The points renumbered, new point i = old point order[i] (a permutation of
 * [0, n), e.g. the Hilbert re-sort MoveMeshPoints applies every 25 cycles,
 * hdsim2d.cpp:222-228): the kept triangulation follows the new numbers, so
 * the next rm_mover_mesh repairs it as usual instead of triangulating
 * afresh.  Returns 0, or -1 for anything but a permutation (nothing
 * changed). */
int rm_mover_permute(rm_mover* mv, const int32_t* order);
/* Full triangulations and flip repairs so far. */
void rm_mover_counts(const rm_mover* mv, int32_t* builds, int32_t* repairs);
void rm_mover_free(rm_mover* mv);

/* Delaunay triangulation of n points (the host side of a moving-mesh
 * rebuild; the reference inserts points one at a time,
 * Delaunay::build_delaunay, Delaunay.cpp:296-402): radial sweep, then edge flips, exact orientation tests.
 * Writes *n_tri <= tri_cap counter-clockwise triangles [3 n_tri] and returns
 * 0, or a negative code (duplicate or unplaceable point, flip budget) for
 * the caller to fall back on another triangulator. */
int rm_delaunay(int32_t n, const double* x, const double* y, int32_t* tri,
                int32_t tri_cap, int32_t* n_tri);

/* Sizes: cells N, faces F, ghosts G, CSR entries nnz. */
void rm_sizes(const rm_mesh* m, int32_t* n_cells, int32_t* n_faces,
              int32_t* n_ghosts, int32_t* nnz);

/* Copy the flat arrays out; any pointer may be NULL.
 * face_first/face_second [F]; v1x,v1y,v2x,v2y [F];
 * pt_x,pt_y,cm_x,cm_y [N+G]; volume,tess_volume [N];
 * csr_off [N+1]; csr_face [nnz]; ghost_orig [G];
 * lat_i, lat_j [N] lattice coordinates of each cell. */
void rm_export(const rm_mesh* m, int32_t* face_first, int32_t* face_second,
               double* v1x, double* v1y, double* v2x, double* v2y,
               double* pt_x, double* pt_y, double* cm_x, double* cm_y,
               double* volume, double* tess_volume, int32_t* csr_off,
               int32_t* csr_face, int32_t* ghost_orig, int32_t* lat_i,
               int32_t* lat_j);

/* Face velocities of a static mesh with given point velocities [N]:
 * VoronoiMesh::CalcFaceVelocity (VoronoiMesh.cpp:146-151) applied as
 * StationaryBox (rigid; zero on mirror faces, stationary_box.cpp:20-40) or
 * PeriodicEdgeVelocities (periodic_edge_velocities.cpp:5-21). */
void rm_face_velocity(const rm_mesh* m, const double* wx, const double* wy,
                      double* fwx, double* fwy);

#ifdef __cplusplus
}
#endif

#endif
