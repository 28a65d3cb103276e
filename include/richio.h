/* HDF5 snapshots of a 2D run in the layout RICH writes and reads
 * (write_snapshot_to_hdf5 / read_hdf5_snapshot, hdf5_diagnostics.cpp:237-346
 * and :348-453; datasets as write_std_vector_to_hdf5 makes them,
 * hdf5_utils.hpp:29-51), so RICH's restart path, its test graders and its
 * plotting scripts read files written here and the other way round.
 *
 * File layout:
 *   /time [1] double, /cycle [1] int, /Volumes [N] double (CacheData::volumes)
 *   /geometry/{x_coordinate, y_coordinate} [N]   generating points
 *   /geometry/{x_vertices, y_vertices} [sum n]   each cell's polygon vertices,
 *       deduplicated and sorted by angle about the cell's CM (ConvexHull,
 *       ConvexHull.cpp:30-58)
 *   /geometry/n_vertices [N] double              vertices per cell
 *   /hydrodynamic/{density, pressure, x_velocity, y_velocity} [N]
 *   /tracers, /stickers, /appendices             groups (empty: no tracers)
 * Every dataset is 1D, little-endian, chunked by min(n, 100000) with deflate
 * level 6.
 *
 * Functions return 0, or a negative code (RIO_E_*). */
#ifndef RICHIO_H
#define RICHIO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { RIO_OK = 0, RIO_E_ARG = -1, RIO_E_HDF5 = -2, RIO_E_LAYOUT = -3, RIO_E_SIZE = -4 };

/* The cells' polygons and points, as in rg_mesh (real cells only). */
typedef struct rio_mesh {
  int32_t n_cells;
  const double *pt_x, *pt_y;         /* [N] generating points                  */
  const double *cm_x, *cm_y;         /* [N] centres of mass                    */
  const double *tess_volume;         /* [N] GetVolume (GetWidth for the
                                        vertex tolerance)                      */
  const double *volume;              /* [N] CacheData::volumes                 */
  const int32_t *csr_off, *csr_face; /* cell -> faces in ascending order       */
  const double *v1x, *v1y, *v2x, *v2y; /* face vertices                        */
} rio_mesh;

/* write_snapshot_to_hdf5 (hdf5_diagnostics.cpp:237-346): cells are the
 * primitive fields [N] each. */
int rio_write_snapshot(const char* path, const rio_mesh* mesh, const double* density,
                       const double* pressure, const double* x_velocity,
                       const double* y_velocity, double time, int32_t cycle);

/* Cell count of a snapshot (the length of /hydrodynamic/density). */
int rio_snapshot_cells(const char* path, int32_t* n_cells);

/* read_hdf5_snapshot (hdf5_diagnostics.cpp:348-453) without tracers and
 * stickers: points, cells, time and cycle into [n_cells] arrays (any may be
 * NULL).  RIO_E_SIZE if the file holds another cell count. */
int rio_read_snapshot(const char* path, int32_t n_cells, double* x, double* y, double* density,
                      double* pressure, double* x_velocity, double* y_velocity, double* time,
                      int32_t* cycle);

/* The polygon vertices written for each cell: *total = sum of counts; with
 * non-NULL arrays also n[N] counts and xv, yv [total] (capacity cap). */
int rio_cell_vertices(const rio_mesh* mesh, int64_t* total, double* n, double* xv, double* yv,
                      int64_t cap);

/* One line per dataset, "path class size length chunk deflate\n" (class f or
 * i, size in bytes, deflate level or -1), in name order, into buf (cap bytes,
 * NUL-terminated): for checking a file's layout. */
int rio_describe(const char* path, char* buf, int32_t cap);

#ifdef __cplusplus
}
#endif

#endif
