/*
 * rich_oracle.h -- CPU restatement of RICH's per-timestep finite-volume
 * update (OzW1234/huji-rich), used ONLY as the checker of the GPU path.
 *
 * TEST INFRASTRUCTURE.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only to check or
 * time against the GPU path.  The product never calls it.
 *
 * Every function restates one reference routine, keeping its operation
 * order so that results agree bit for bit under -ffp-contract=off; the
 * reference file:line is cited at each definition in rich_oracle.c.
 *
 * Layout: cells, extensives and face states are structure-of-arrays
 * (density, pressure, vx, vy) / (mass, energy, momentum x, momentum y);
 * slopes are 8 arrays (d/dx of the 4 fields, then d/dy of the 4 fields).
 * Ghost points are numbered N..N+G-1.
 */
#ifndef RICH_ORACLE_H
#define RICH_ORACLE_H

#include <stdint.h>

#if defined(__cplusplus) && !defined(RO_TANGENT)
extern "C" {
#endif

/* scalar type of the state arrays: double, or the dual numbers of the
 * forward-mode build (oracle/tangent.cpp) */
#ifndef RO_TANGENT
#define RO_REAL double
#else
#define RO_REAL real
#endif

typedef struct ro_mesh {
  int32_t n_cells, n_faces, n_ghosts;
  const int32_t *face_first, *face_second; /* [F]                     */
  const double *v1x, *v1y, *v2x, *v2y;     /* [F] face vertices        */
  const double *pt_x, *pt_y;               /* [N+G] generating points  */
  const double *cm_x, *cm_y;               /* [N+G] centres of mass    */
  const double *volume;                    /* [N] CacheData::volumes   */
  const double *tess_volume;               /* [N] Tessellation::GetVolume */
  const int32_t *csr_off, *csr_face;       /* cell -> faces, ascending */
  const int32_t *ghost_orig;               /* [G] GetOriginalIndex     */
} ro_mesh;

typedef struct ro_params {
  double gamma;   /* IdealGas adiabatic index                         */
  double cfl;     /* SimpleCFL                                        */
  int32_t order;  /* 1: SimpleFluxCalculator, 2: Modular + LGI,
                     3: Modular + PCM                                  */
  int32_t bc;     /* 0 RigidWallGenerator, 1 PeriodicGhostGenerator   */
  int32_t slf;    /* LinearGaussImproved slope limiter flag            */
  double delta_v, theta, delta_p;
  int32_t nthreads; /* 1 = reference loop order; >1 OpenMP (same results) */
} ro_params;

/* SoA views; arrays are read or written as the function states. */
typedef struct ro_cells { RO_REAL *d, *p, *vx, *vy; } ro_cells;
typedef struct ro_ext { RO_REAL *m, *E, *Mx, *My; } ro_ext;

/* Error codes mirror the UniversalError messages of the reference. */
enum {
  RO_OK = 0,
  RO_NEGATIVE_DENSITY = 1,      /* "Negative density"                 */
  RO_NEGATIVE_THERMAL = 2,      /* "Negative thermal energy"          */
  RO_HLLC_NAN = 3,              /* "Hllc::Solved entered with nan"    */
  RO_INVALID_WAVE_SPEEDS = 4,   /* "Invalid wave speeds in hllc solver" */
  RO_SINGULAR_MATRIX = 5,       /* "Singular matrix"                  */
  RO_TOO_MANY_NEIGHBORS = 6     /* "Cell has too many neighbors"      */
};
typedef struct ro_error { int32_t code; int32_t index; } ro_error;

#ifndef RO_TANGENT
/* --- single-face physics (for known-answer tests) --- */
/* prim = {density, pressure, vx, vy, energy, sound speed} (Primitive) */
int ro_hllc(const double left[6], const double right[6], double velocity,
            double out[4] /* mass, momentum x, momentum y, energy */);
double ro_dp2e(double gamma, double d, double p);
double ro_dp2c(double gamma, double d, double p);
int ro_de2p(double gamma, double d, double e, double* p);

/* --- plugins on a mesh --- */
/* Ghost cells [N+G] filled for every ghost that touches a face. */
void ro_ghost_cells(const ro_mesh* m, const ro_params* prm, ro_cells cells_ng);
/* LinearGaussImproved::operator(): slopes[8][N], naive[8][N] (may be
 * NULL), left/right face states [4][F]. */
ro_error ro_lgi(const ro_mesh* m, const ro_params* prm, ro_cells cells_ng,
                double* const slopes[8], double* const naive[8],
                ro_cells left, ro_cells right);
/* PCM::operator(): face states [4][F] are the cells on each side, ghost
 * sides from the ghost generator (cng's ghost slots, filled here). */
void ro_pcm(const ro_mesh* m, const ro_params* prm, ro_cells cells_ng, ro_cells left,
            ro_cells right);
/* FluxCalculator: ModularFluxCalculator over LinearGaussImproved (order 2)
 * or over PCM (order 3), or SimpleFluxCalculator (order 1).  cells_ng has
 * N+G entries (ghost slots are filled here).  w: face velocities [2][F]
 * (NULL = zero).  flux [4][F] per unit length. */
ro_error ro_flux(const ro_mesh* m, const ro_params* prm, ro_cells cells_ng,
                 const double* wx, const double* wy, ro_ext flux);
/* SimpleExtensiveUpdater. */
void ro_update_extensives(const ro_mesh* m, const ro_params* prm, ro_ext flux,
                          double dt, ro_ext ext);
/* SimpleCellUpdater. */
ro_error ro_update_cells(const ro_mesh* m, const ro_params* prm, ro_ext ext,
                         ro_cells cells);
/* SimpleCFL (dt, already multiplied by cfl). */
double ro_cfl(const ro_mesh* m, const ro_params* prm, ro_cells cells,
              const double* wx, const double* wy);
/* hdsim init_extensives. */
void ro_init_extensives(const ro_mesh* m, const ro_params* prm, ro_cells cells,
                        ro_ext ext);
/* Record the branches of each step of ro_run as 64-bit sums of hashes in
 * per_step[n][8], zeroed here, one sum per kind of branch.  Jumps (the
 * value is discontinuous there): [6] the shock switch, [7] the limiters'
 * gates and guards.  Kinks (the derivative is): [2] HLLC's wave pattern and
 * the sides giving its outer wave speeds, [3] the CFL argmin, [4] and [5]
 * slope_limit's and shocked_slope_limit's active faces, clamps and
 * extremum neighbours.  NULL switches the record off.  Equal sums: the
 * same branches. */
void ro_fingerprint(uint64_t* per_step, int32_t n);

/* --- moving-mesh inputs (rigid box) --- */
/* Lagrangian point velocities plus RoundCells::calc_dw without the cold
 * switch: w_i = v_i + chi max(c_i, |v_i|) (CM_i - r_i) / R_i where
 * |CM_i - r_i| >= 0.9 eta R_i, R = sqrt(GetVolume/pi).  [N] each. */
void ro_point_velocity(const ro_mesh* m, double gamma, const double* d, const double* p,
                       const double* vx, const double* vy, double chi, double eta,
                       double* wx, double* wy);
/* CorrectPointsOverShoot of the box [x0, x1, y0, y1] for step dt, in
 * place on the point velocities [N]. */
void ro_point_fix(const ro_mesh* m, double dt, const double box[4], double* wx, double* wy);
/* Face velocities [F] from point velocities [N] (StationaryBox with
 * CalcFaceVelocity; a ghost side moves with its original). */
void ro_face_velocity(const ro_mesh* m, int32_t bc, const double* wx, const double* wy,
                      double* fwx, double* fwy);

/* Whole steps on a static mesh: hdsim::TimeAdvance (integrator 1) or
 * hdsim::TimeAdvance2Heun (integrator 2).  cells_ng [N+G], ext [N].
 * dt_out[step] receives each step's dt (may be NULL). */
ro_error ro_run(const ro_mesh* m, const ro_params* prm, int32_t integrator,
                int32_t nsteps, const double* wx, const double* wy,
                ro_cells cells_ng, ro_ext ext, double* dt_out);
#endif /* RO_TANGENT */

#if defined(__cplusplus) && !defined(RO_TANGENT)
}
#endif

#endif
