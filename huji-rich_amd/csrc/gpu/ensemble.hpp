// Batched ensemble: many independent simulations advanced together.  The
// context's mesh is the disjoint union of the instances' meshes (cells of
// instance k are coff[k]..coff[k+1]-1), so the cell and face kernels serve
// every instance in one launch whatever their sizes; only the time step is
// per instance.  A step is eight persistent launches captured once into a
// hipGraph and replayed per step; every quantity that changes between steps
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
