// Batched ensemble: many independent simulations advanced together.  The
// context's mesh is the disjoint union of the instances' meshes (cells of
// instance k are coff[k]..coff[k+1]-1), so the cell and face kernels serve
// every instance in one launch whatever their sizes; only the time step is
// per instance.  A step is eight launches (six where the stages run fused
// per tile) captured once into a hipGraph and replayed per step; every quantity that changes between steps
// (the step index, the dt slots) lives in device memory, so the graph is the
// same for every step.
#pragma once

#include "kernels.hpp"

namespace rg {

struct Ens {
  int K;                      // instances
  const int* inst;            // [N] instance of each cell
  const int* coff;            // [K+1] first cell of each instance
  unsigned long long* dtA;    // [K] this step's min_i inv_i (double bits)
  unsigned long long* dtB;    // [K] the next step's
  double* inv;                // [N] CFL inverse time step of each cell's new state
  double* dt_hist;            // [hist_cap][K] dt of each step and instance (may be null)
  int* step;                  // step counter
  int hist_cap;
};

// SimpleExtensiveUpdater + SimpleCellUpdater + the CFL term per cell, each
// cell with its own instance's dt (update_cell is the single-instance body).
__global__ void __launch_bounds__(256) k_update_ens_p(DMesh m, UpdArgs a, Ens e, DParams prm,
                                                      DErr err) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < m.N; i += gridDim.x * blockDim.x) {
    const double dt =
        prm.cfl * __longlong_as_double(static_cast<long long>(e.dtA[e.inst[i]]));
    const double v = update_cell(m, a, prm, err, i, dt);
    if (a.flags & UPD_CFL) e.inv[i] = v;
  }
}

// One stage of every instance fused per tile of kTile cells (k_face_update
// with per-instance time steps): each face's dt*area*F with its instance's
// dt (both sides of a face belong to one instance) into LDS, then each cell's
// update with its own instance's dt and its CFL inverse into e.inv.
__global__ void __launch_bounds__(kTile) k_face_update_ens(DMesh m, DTiles tl,
                                                         const D4* __restrict__ q,
                                                         const Rec* __restrict__ rec, UpdArgs a,
                                                         Ens e, DParams prm, DErr errf, DErr erru) {
  extern __shared__ D4 tflux[];
  const int t = walk_block(prm);
  const int e0 = tl.tofs[t], ne = tl.tofs[t + 1] - e0;
  FaceFrame frn{0, 0, 0, 0, 0, 0};
  FaceSides sdn{0, 0};
  if (static_cast<int>(threadIdx.x) < ne) {
    frn = tl.tframe[e0 + threadIdx.x];
    sdn = tl.tside[e0 + threadIdx.x];
  }
  for (int k = threadIdx.x; k < ne; k += blockDim.x) {
    const int ent = e0 + k;
    const FaceFrame fr = frn;
    const FaceSides sd = sdn;
    if (k + static_cast<int>(blockDim.x) < ne) {
      frn = tl.tframe[ent + blockDim.x];
      sdn = tl.tside[ent + blockDim.x];
    }
    const int real = sd.a < m.N ? sd.a : sd.b;
    const double dt =
        prm.cfl * __longlong_as_double(static_cast<long long>(e.dtA[e.inst[real]]));
    Cell lf, rt;
    face_states(m, q, rec, sd.a, sd.b, fr, lf, rt);
    const double an = vabs(fr.nx, fr.ny), ap = vabs(fr.px, fr.py);
    // modular_flux_calculator.cpp:56-67 (an ensemble has no face velocities)
    const double speed = zero_face_speed(fr.nx, fr.ny, an);
    Cons c;
    const int rc = hllc_face(lf, rt, prm.gamma, fr.nx, fr.ny, an, fr.px, fr.py, ap, speed, c);
    D4 o{0, 0, 0, 0};
    if (rc) {
      report(errf, tl.tface[ent], rc);
    } else {
      const double sc = dt * ap;  // SimpleExtensiveUpdater's dt*area
      o = D4{c.m * sc, c.E * sc, c.mx * sc, c.my * sc};
    }
    tflux[k] = o;
  }
  const int i = t * kTile + threadIdx.x;
  UpdLoads L{};
  double dt = 0;
  if (i < m.N) {
    L = update_loads<true>(m, a, erru, i, tl.eloc);
    dt = prm.cfl * __longlong_as_double(static_cast<long long>(e.dtA[e.inst[i]]));
  }
  __syncthreads();
  if (i < m.N) {
    const double v = update_cell<true, true>(m, a, prm, erru, i, dt, tl.eloc, tflux, &L);
    if (a.flags & UPD_CFL) e.inv[i] = v;
  }
}

// CFL inverse of the current cells (the first step's dt).
__global__ void __launch_bounds__(256) k_inv_ens_p(DMesh m, const D4* __restrict__ q, Ens e,
                                                   double gamma) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < m.N; i += gridDim.x * blockDim.x)
    e.inv[i] = cfl_inv(m, i, gamma, m.width[i], ld_cell(q, i));
}

// Per-instance minimum of inv into dtB (lazy_min over the instance's cells,
// simple_cfl.cpp:58-71; the minimum is exact in any order).  A fixed grid of
// blocks of 256 walks the instances.
__global__ void __launch_bounds__(256) k_ens_min(Ens e) {
  __shared__ double sh[256];
  for (int k = blockIdx.x; k < e.K; k += gridDim.x) {
    double v = __longlong_as_double(static_cast<long long>(DBL_INF_BITS));
    for (int i = e.coff[k] + threadIdx.x; i < e.coff[k + 1]; i += 256) v = smin(v, e.inv[i]);
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
      if (static_cast<int>(threadIdx.x) < s) sh[threadIdx.x] = smin(sh[threadIdx.x], sh[threadIdx.x + s]);
      __syncthreads();
    }
    if (threadIdx.x == 0) e.dtB[k] = static_cast<unsigned long long>(__double_as_longlong(sh[0]));
    __syncthreads();
  }
}

// End of a step (one block): record each instance's dt, make the next
// step's minima current and advance the step counter.  With record = 0 only
// the minima are made current (the seed before the first step).
__global__ void __launch_bounds__(256) k_ens_finish(Ens e, double cfl, int record) {
  const int step = *e.step;
  for (int k = threadIdx.x; k < e.K; k += blockDim.x) {
    if (record && e.dt_hist && step < e.hist_cap)
      e.dt_hist[static_cast<size_t>(step) * e.K + k] =
          cfl * __longlong_as_double(static_cast<long long>(e.dtA[k]));
    e.dtA[k] = e.dtB[k];
  }
  __syncthreads();
  if (record && threadIdx.x == 0) *e.step = step + 1;
}

}  // namespace rg
