/*
 * richgpu.h -- C ABI of the MI355X (gfx950) finite-volume step of RICH.
 *
 * RICH's hdsim driver (source/newtonian/two_dimensional/hdsim2d.hpp:43-274)
 * calls five plugin objects per stage; each entry point below replaces the
 * body of one of them on a device-resident mesh and state:
 *
 *   TimeStepFunction::operator()       time_step_function.hpp:29-35
 *       -> rg_time_step                 (SimpleCFL, simple_cfl.cpp:58-71)
 *   SpatialReconstruction::operator()  spatial_reconstruction.hpp:37-42
 *       -> rg_reconstruct               (LinearGaussImproved.cpp:536-635)
 *   FluxCalculator::operator()         flux_calculator_2d.hpp:35-44
 *       -> rg_flux                      (ModularFluxCalculator
 *                                        modular_flux_calculator.cpp:39-71,
 *                                        SimpleFluxCalculator
 *                                        simple_flux_calculator.cpp:216-244,
 *                                        Hllc hllc.cpp:97-154)
 *   ExtensiveUpdater::operator()       extensive_updater.hpp:34-43
 *       -> rg_update_extensives         (simple_extensive_updater.cpp:11-35)
 *   CellUpdater::operator()            cell_updater_2d.hpp:34-41
 *       -> rg_update_cells              (simple_cell_updater.cpp:85-106)
 *
 * plus the state set-up hdsim does in its constructor (init_extensives,
 * hdsim2d.cpp:41-66) and a device-resident loop of whole steps
 * (hdsim::TimeAdvance hdsim2d.cpp:149-241, TimeAdvance2Heun :323-449) for a
 * static mesh, which is how throughput is measured.
 *
 * Conventions: the library owns device buffers, the caller owns host
 * buffers; host arrays are structure-of-arrays; every call is ordered on
 * the context's stream and returns 0 or an RG_E* code.  Device-detected
 * errors carry the reference's UniversalError text (rg_last_error).
 * A context is used from one host thread at a time (the reference plugins
 * hold mutable scratch and are not re-entrant either).
 */
#ifndef RICHGPU_H
#define RICHGPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rg_ctx rg_ctx;

/* status codes */
#define RG_OK 0
#define RG_E_ARG -1      /* bad argument / call order                */
#define RG_E_HIP -2      /* HIP runtime error                        */
#define RG_E_DEVICE 1    /* device-detected error: see rg_last_error */

/* device-detected error codes (UniversalError messages of the reference) */
#define RG_ERR_NEGATIVE_DENSITY 1     /* simple_cell_updater.cpp:33-34   */
#define RG_ERR_NEGATIVE_THERMAL 2     /* ideal_gas.cpp:20-21             */
#define RG_ERR_HLLC_NAN 3             /* hllc.cpp:102-103                */
#define RG_ERR_INVALID_WAVE_SPEEDS 4  /* hllc.cpp:143-148                */
#define RG_ERR_SINGULAR_MATRIX 5      /* LinearGaussImproved.cpp:111-119 */
#define RG_ERR_TOO_MANY_NEIGHBORS 6   /* LinearGaussImproved.cpp:73-79   */

#define RG_BC_RIGID 0     /* RigidWallGenerator + StationaryBox            */
#define RG_BC_PERIODIC 1  /* PeriodicGhostGenerator + PeriodicEdgeVelocities */

#define RG_FP64 0         /* every operation in FP64, reference order      */
#define RG_MIXED 1        /* FP32 reconstruction + HLLC, FP64 extensives and
                             primitive recovery (compensated sums)          */

typedef struct rg_params {
  double gamma;    /* IdealGas(gamma)                      ideal_gas.cpp:5-6 */
  double cfl;      /* SimpleCFL(cfl)                       simple_cfl.cpp:6   */
  int32_t order;   /* 1: SimpleFluxCalculator (piecewise constant)
                      2: ModularFluxCalculator + LinearGaussImproved
                      3: ModularFluxCalculator + PCM (PCM.cpp:5-37)           */
  int32_t bc;      /* RG_BC_RIGID / RG_BC_PERIODIC                            */
  int32_t slf;     /* LinearGaussImproved slf        LinearGaussImproved.hpp:31-40 */
  double delta_v;  /* shock detector |div v|*L/c       (default 0.2)         */
  double theta;    /* shocked-limiter coefficient      (default 0.5)         */
  double delta_p;  /* shock detector pressure ratio    (default 0.7)         */
  int32_t precision; /* RG_FP64 or RG_MIXED                                  */
} rg_params;

/* Flat mesh, as a caller holding a RICH Tessellation would flatten it:
 * faces = getAllEdges() (Edge.hpp:13-51), ghost ids remapped from
 * [N+3, TotalPointNumber) to [N, N+G), csr = GetCellEdges(i) in order. */
typedef struct rg_mesh {
  int32_t n_cells, n_faces, n_ghosts;
  const int32_t *face_first, *face_second; /* [F] Edge::neighbors            */
  const double *v1x, *v1y, *v2x, *v2y;     /* [F] Edge::vertices             */
  const double *pt_x, *pt_y;               /* [N+G] GetMeshPoint             */
  const double *cm_x, *cm_y;               /* [N+G] GetCellCM                */
  const double *volume;                    /* [N] CacheData::volumes         */
  const double *tess_volume;               /* [N] GetVolume (width, slopes)  */
  /* cm_x, cm_y, volume, tess_volume may all be NULL: the device then forms
   * them from the vertices in each cell's face order (CalcCellCM
   * VoronoiMesh.cpp:436-451, SlabSymmetry::calcVolume
   * physical_geometry.cpp:13-21, GetVolume VoronoiMesh.cpp:420-434) and
   * the ghosts' CMs from their originals' (shifted for a periodic box,
   * reflected across the wall each ghost lies beyond for a rigid one, which
   * needs rg_set_box; VoronoiMesh.cpp:359-398). */
  const int32_t *csr_off;                  /* [N+1]                          */
  const int32_t *csr_face;                 /* [csr_off[N]] GetCellEdges      */
  const int32_t *ghost_orig;               /* [G] GetOriginalIndex           */
} rg_mesh;

/* One UniversalError::AddEntry(name, value) of the reference. */
typedef struct rg_error_entry {
  char name[32];
  double value;
} rg_error_entry;

typedef struct rg_error {
  int32_t code;     /* RG_ERR_* or 0                                  */
  int32_t index;    /* cell or face index of the first failure        */
  int32_t step;     /* step of rg_advance (0 outside it)              */
  int32_t phase;    /* 0 cfl, 1 slopes, 2 fluxes, 3 cells (+3 stage 2) */
  char message[96]; /* reference message text                          */
  /* The entries the reference's exception carries, in its order:
   *   RG_ERR_NEGATIVE_THERMAL     "Cell index", "Cell mass", "Cell x momentum",
   *                               "Cell y momentum", "Cell energy"
   *                               (simple_cell_updater.cpp:50-57)
   *   RG_ERR_INVALID_WAVE_SPEEDS  "left density" .. "right wave speed", 16
   *                               entries (hllc.cpp:46-71)
   *   RG_ERR_SINGULAR_MATRIX      "Cell x cor", "Cell y cor", "Cell volume",
   *                               "Det was" (LinearGaussImproved.cpp:111-118)
   *   RG_ERR_TOO_MANY_NEIGHBORS   "Cell x cor", "Cell y cor" (:73-79)
   * none for the other codes (the reference adds none).  Filled by the
   * plugin entry points and by rg_advance / rg_advance_noalloc on an RG_FP64
   * context without instances (n_entries = 0 elsewhere): after a
   * device-detected error the step's updates stop, so the state the failing
   * call read is still there to recompute them from. */
  int32_t n_entries;
  rg_error_entry entries[16];
} rg_error;

/* lifecycle ---------------------------------------------------------- */
int rg_create(const rg_params* params, rg_ctx** out);
void rg_destroy(rg_ctx* ctx);
/* Run on a caller's hipStream_t (NULL: the library's own stream). */
int rg_set_stream(rg_ctx* ctx, void* hip_stream);

/* mesh and state (host arrays) ------------------------------------- */
/* Upload a mesh (hdsim holds the Tessellation, hdsim2d.hpp:54).  RG_E_ARG,
 * with nothing uploaded, unless: side ids are in [0, N+G), differ, and at
 * least one is a real cell; csr_off starts at 0 and never decreases; each
 * cell lists faces it is a side of, in ascending order; every face between
 * two real cells is listed by both; ghost_orig is in [0, N). */
int rg_set_mesh(rg_ctx* ctx, const rg_mesh* mesh);
/* The rigid box [x0, x1, y0, y1] (SquareBox, the OuterBoundary of a rigid
 * run): needed only to form a rigid mesh's geometry on the device. */
int rg_set_box(rg_ctx* ctx, const double* box);
/* The mesh geometry the device holds: volume [N], tess_volume [N],
 * cm_x/cm_y [N+G] (any may be NULL). */
int rg_get_geometry(rg_ctx* ctx, double* volume, double* tess_volume, double* cm_x,
                    double* cm_y);
/* Face velocities [F] as hdsim passes edge_velocities; NULL,NULL = all zero
 * (StationaryBox/PeriodicEdgeVelocities on a static mesh). */
int rg_set_face_velocity(rg_ctx* ctx, const double* wx, const double* wy);
/* moving mesh (rigid box, order 2, RG_FP64) ---------------------------
 * One hdsim::TimeAdvance2Heun step with Lagrangian + RoundCells points and
 * StationaryBox edges is rg_moving_begin, the host moving the points by
 * dt * velocity and rebuilding the Voronoi mesh, rg_replace_mesh, then
 * rg_moving_end; cells, extensives and point velocities stay on the device.
 *
 * Point velocities (point_motion_, hdsim2d.cpp:325; round_cells.cpp:78-134
 * with chi, eta), edge velocities, the SimpleCFL dt, ApplyFix of a rigid box
 * [x0, x1, y0, y1] (round_cells.cpp:14-61), edge velocities and dt again,
 * then stage 1 on the current mesh (hdsim2d.cpp:325-371).  Returns dt and,
 * where pvx/pvy are non-NULL, the point velocities [N]. */
int rg_moving_begin(rg_ctx* ctx, double chi, double eta, const double* box, double* dt_out,
                    double* pvx, double* pvy);
/* The tessellation rebuilt on the moved points (MoveMeshPoints,
 * hdsim2d.cpp:390; same cell count and order): replaces the mesh and keeps
 * cells, extensives, mid extensives and point velocities. */
int rg_replace_mesh(rg_ctx* ctx, const rg_mesh* mesh);
/* The cells in a new order, new cell i = old cell order[i] (a permutation
 * of [0, N)): cells, extensives, mid extensives and point velocities follow,
 * as hdsim permutes cells_, extensives_, mid_extensives and
 * point_velocities with the Hilbert order MoveMeshPoints returns every 25
 * cycles (hdsim2d.cpp:389-399).  The caller then gives the mesh of the
 * reordered points to rg_replace_mesh. */
int rg_permute_cells(rg_ctx* ctx, const int32_t* order);
/* Stage 2 on the replaced mesh: mid cells, edge velocities from the kept
 * point velocities, fluxes, the update, average_extensive and the cell
 * update (hdsim2d.cpp:401-441). */
int rg_moving_end(rg_ctx* ctx);
/* cells_ (ComputationalCell density, pressure, velocity), [N] each. */
int rg_set_cells(rg_ctx* ctx, const double* d, const double* p, const double* vx,
                 const double* vy);
int rg_get_cells(rg_ctx* ctx, double* d, double* p, double* vx, double* vy);
/* extensives_ (Extensive mass, energy, momentum), [N] each. */
int rg_set_extensives(rg_ctx* ctx, const double* m, const double* E, const double* Mx,
                      const double* My);
int rg_get_extensives(rg_ctx* ctx, double* m, double* E, double* Mx, double* My);
/* Fluxes [F] (per unit length) for rg_update_extensives, when they come
 * from a flux calculator outside this library. */
int rg_set_fluxes(rg_ctx* ctx, const double* fm, const double* fE, const double* fMx,
                  const double* fMy);

/* plugins ----------------------------------------------------------- */
/* init_extensives (hdsim2d.cpp:41-66) from the current cells. */
int rg_init_extensives(rg_ctx* ctx);
/* SimpleCFL::operator(): dt = cfl * min_i 1/max_faces((c+|v-w|)/R_i). */
int rg_time_step(rg_ctx* ctx, double* dt);
/* SpatialReconstruction: slopes [8][N] (d/dx of density, pressure, vx, vy,
 * then d/dy) and face states res[f].first/.second as left/right [4][F]
 * (density, pressure, vx, vy).  Any output may be NULL.  Order 2 is
 * LinearGaussImproved; orders 1 and 3 give PCM's states (PCM.cpp:5-37)
 * and zero slopes. */
int rg_reconstruct(rg_ctx* ctx, double* slopes, double* left, double* right);
/* FluxCalculator: fluxes per unit face length, kept on the device for
 * rg_update_extensives; copied out when the pointers are non-NULL [F]. */
int rg_flux(rg_ctx* ctx, double* fm, double* fE, double* fMx, double* fMy);
/* ExtensiveUpdater: extensives -= / += dt*area*flux over faces. */
int rg_update_extensives(rg_ctx* ctx, double dt);
/* RiemannSolver::operator()(left, right, velocity) (riemann_solver.hpp:22-23)
 * -> Hllc (hllc.cpp:97-154), for n pairs at once: Primitives already in the
 * face frame (x along the normal), left and right [6][n] (density,
 * pressure, x velocity, y velocity, energy, sound speed, as the caller's
 * EquationOfState filled them), the face velocity along the normal
 * velocity [n] (NULL: 0), into the conserved fluxes out [4][n] (mass,
 * energy, x momentum, y momentum).  Needs no mesh.  A failing pair gives the
 * reference's message, its index and, for invalid wave speeds, the
 * entries hllc.cpp:54-70 adds. */
int rg_riemann(rg_ctx* ctx, int32_t n, const double* left, const double* right,
               const double* velocity, double* out);
/* CellUpdater: cells from extensives and CacheData volumes. */
int rg_update_cells(rg_ctx* ctx);

/* device-resident steps on a static mesh ------------------------------ */
/* integrator 1: hdsim::TimeAdvance; 2: hdsim::TimeAdvance2Heun.  Each step
 * computes its own dt (SimpleCFL) and adds the ZeroForce source.  dt_out
 * [nsteps] receives the dts (may be NULL).  Synchronises once at the end. */
int rg_advance(rg_ctx* ctx, int32_t integrator, int32_t nsteps, double* dt_out);

/* No counterpart in the reference (hdsim allocates as it goes): rg_advance
 * for callers that cannot allocate device memory while they step -- code
 * under HIP graph capture, serving loops, harnesses that forbid hipMalloc.
 * Same arguments, results and errors as rg_advance for an RG_FP64 context
 * without instances (else RG_E_ARG); it calls no hipMalloc or hipFree.  The
 * dt slots, CFL block minima and dt history live in one fixed device array
 * of the library (1 MiB on each GPU it runs on, held until the process
 * ends), so nsteps <= 65536 and n_cells <= 16777216 per call (else
 * RG_E_ARG with a message in rg_last_error).  Every call on a GPU shares
 * that GPU's array: calls that overlap on one GPU (other host threads,
 * other streams) race. */
int rg_advance_noalloc(rg_ctx* ctx, int32_t integrator, int32_t nsteps, double* dt_out);

/* adjoint of the fixed-topology step ----------------------------------- */
/* No counterpart in the reference (RICH has no derivatives): the reverse
 * mode of rg_advance(ctx, 2, nsteps) for order 2, RG_FP64, zero face
 * velocities, for gradient-based callers (inverse problems, sensitivity
 * studies).  rg_adjoint_forward runs the same steps as rg_advance and keeps
 * a device checkpoint of (cells, extensives) every `every` steps; the
 * extensives must be rg_init_extensives of the cells (else RG_E_ARG). */
int rg_adjoint_forward(rg_ctx* ctx, int32_t nsteps, int32_t every, double* dt_out);
/* Given d(loss)/d(final cells) gcells [4][N] (density, pressure, vx, vy)
 * and optionally d(loss)/d(final extensives) gext [4][N] (NULL = 0),
 * writes d(loss)/d(initial cells) gcells0 [4][N], through init_extensives.
 * Each segment between checkpoints is recomputed once; the derivative is
 * that of the branches the forward values select (HLLC wave pattern,
 * limiter choices, shock switch, CFL argmin).  The final state of the
 * forward run is left in the context. */
int rg_adjoint_backward(rg_ctx* ctx, const double* gcells, const double* gext, double* gcells0);

/* batched ensemble -------------------------------------------------- */
/* No counterpart in the reference (it runs one hdsim per simulation): for
 * callers that advance many independent simulations together (parameter
 * scans, ensembles).  The context's mesh is the disjoint union of n_inst
 * meshes: the cells of instance k are cell_off[k]..cell_off[k+1]-1 and
 * every face and ghost belongs to one instance (checked; else RG_E_ARG).
 * rg_advance then gives each instance its own SimpleCFL dt (dt_out is
 * [nsteps][n_inst]) and runs each step as one hipGraph launch of persistent
 * kernels; integrator 2, order 2, RG_FP64, zero face velocities. */
int rg_set_instances(rg_ctx* ctx, int32_t n_inst, const int32_t* cell_off);
/* Launch shape of the step kernels: block (face/cell tile) 64, 128 or 256
 * threads and blocks_per_cu resident blocks per compute unit walking the
 * tiles (persistent), or blocks_per_cu = 0 for one thread per cell/face. */
int rg_ensemble_config(rg_ctx* ctx, int32_t block, int32_t blocks_per_cu);
/* Times trial_steps steps for every launch shape (blocks_per_cu 1..16 and
 * 0), keeps the fastest and
 * restores the state; table[3*r..] = (block, blocks_per_cu, ms per step)
 * for up to table_rows rows; *rows_out = rows measured. */
int rg_ensemble_autotune(rg_ctx* ctx, int32_t trial_steps, double* table, int32_t table_rows,
                         int32_t* rows_out);

/* diagnostics -------------------------------------------------------- */
int rg_last_error(rg_ctx* ctx, rg_error* err);
/* Per-kernel HIP-event timing of every launch on the context's stream. */
int rg_profile_enable(rg_ctx* ctx, int32_t on);
/* Kernel ids 0..rg_kernel_count()-1. */
int32_t rg_kernel_count(void);
const char* rg_kernel_name(int32_t id);
/* Synchronises, then returns launches and summed milliseconds of a kernel
 * since rg_profile_enable(ctx, 1). */
int rg_profile_read(rg_ctx* ctx, int32_t kernel_id, int64_t* launches, double* total_ms);
/* Bytes of device memory the context holds. */
int64_t rg_device_bytes(rg_ctx* ctx);
/* Whether the second-order Heun loops (rg_advance, the ensemble graph, the
 * adjoint) run each stage's faces and cell update fused per tile of 256
 * cells: 1 once the mesh's tiles are built, fit the kernel's shared memory
 * and hold at most 1.3 entries per face (cells numbered along a
 * space-filling curve give about 1.08; in random order about 2, where the
 * three-kernel stages are faster), else 0.  *entries and
 * *max_entries (either may be NULL) receive the faces solved per stage (a
 * face between two tiles counts twice) and the largest tile's count.  No
 * counterpart in the reference: a diagnostic of this library's layout. */
int32_t rg_tiles(rg_ctx* ctx, int64_t* entries, int32_t* max_entries);
/* on = 1 (the default): the second-order Heun loops run the fused tile
 * stages (built at their first call on a mesh); on = 0: they run the
 * three-kernel stages (slopes, faces into a face array, cell update), with
 * the same results bit for bit, and build no tiles (tiles already built stay
 * allocated until the next rg_set_mesh).  No counterpart in the reference:
 * for callers short of device memory, and for tests that compare the two
 * forms. */
int rg_use_tiles(rg_ctx* ctx, int32_t on);

#ifdef __cplusplus
}
#endif

#endif
