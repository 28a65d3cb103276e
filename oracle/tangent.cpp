/*
 * tangent.cpp -- forward-mode (dual-number) derivatives of the oracle's
 * step: the CPU reference gradient of the adjoint (test infrastructure, like
 * the rest of oracle/; never part of the product).
 *
 * rich_oracle.c is compiled a second time with `real` = dual numbers
 * (tangent_real.hpp), so the primal values are those of the double oracle
 * bit for bit and the RO_K tangents are the exact directional derivatives
 * of the same operations along the seeded directions (active branches).
 */
#define RO_TANGENT 1
#include "tangent_real.hpp"
#include "rich_oracle.c"

#include <vector>

extern "C" {

int rt_k(void) { return RO_K; }

/* Initial cells (values cells[4][N+G], tangents dcells[RO_K][4][N+G]; the
 * ghost slots are scratch) -> init_extensives -> ro_run(integrator, nsteps).
 * On return cells/dcells hold the final state, ext/dext (may be NULL) the
 * final extensives [4][N] / [RO_K][4][N], dt_out (may be NULL) each step's
 * dt.  Returns the error code; *err_index the failing index. */
int rt_run(const ro_mesh* m, const ro_params* prm, int32_t integrator, int32_t nsteps,
           const double* wx, const double* wy, double* cells, double* dcells, double* ext,
           double* dext, double* dt_out, int32_t* err_index) {
  const size_t N = (size_t)m->n_cells, NG = N + (size_t)m->n_ghosts;
  std::vector<real> c(4 * NG), x(4 * N);
  for (size_t j = 0; j < 4 * NG; ++j) {
    c[j].v = cells[j];
    for (int k = 0; k < RO_K; ++k) c[j].t[k] = dcells[(size_t)k * 4 * NG + j];
  }
  ro_cells cng = {c.data(), c.data() + NG, c.data() + 2 * NG, c.data() + 3 * NG};
  ro_ext e = {x.data(), x.data() + N, x.data() + 2 * N, x.data() + 3 * N};
  ro_init_extensives(m, prm, cng, e);
  const ro_error err = ro_run(m, prm, integrator, nsteps, wx, wy, cng, e, dt_out);
  for (size_t j = 0; j < 4 * NG; ++j) {
    cells[j] = c[j].v;
    for (int k = 0; k < RO_K; ++k) dcells[(size_t)k * 4 * NG + j] = c[j].t[k];
  }
  for (size_t j = 0; j < 4 * N; ++j) {
    if (ext) ext[j] = x[j].v;
    if (dext)
      for (int k = 0; k < RO_K; ++k) dext[(size_t)k * 4 * N + j] = x[j].t[k];
  }
  if (err_index) *err_index = err.index;
  return err.code;
}

}  // extern "C"
