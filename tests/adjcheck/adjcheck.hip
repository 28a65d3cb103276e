// Host-side view of the adjoint physics (physics_adj.hpp) for CPU unit
// tests: the same __host__ __device__ functions the GPU kernels call,
// exported with plain C signatures.  Test infrastructure only.
#include "../../huji-rich_amd/csrc/gpu/physics_adj.hpp"

using namespace rg;

extern "C" {

// the face speed for zero face velocities, as the face kernels form it
double ac_zero_face_speed(double nx, double ny, double an) { return zero_face_speed(nx, ny, an); }

// frame = {nx, ny, an, px, py, ap}; cells (d, p, vx, vy); flux (m, mx, my, E)
int ac_hllc(const double* L, const double* R, double g, const double* fr, double* out) {
  Cons c;
  const int rc = hllc_face(Cell{L[0], L[1], L[2], L[3]}, Cell{R[0], R[1], R[2], R[3]}, g, fr[0],
                           fr[1], fr[2], fr[3], fr[4], fr[5], 0.0, c);
  out[0] = c.m;
  out[1] = c.mx;
  out[2] = c.my;
  out[3] = c.E;
  return rc;
}

void ac_hllc_vjp(const double* L, const double* R, double g, const double* fr, const double* fb,
                 double* Lb, double* Rb) {
  Cell lb{0, 0, 0, 0}, rb{0, 0, 0, 0};
  hllc_face_vjp(Cell{L[0], L[1], L[2], L[3]}, Cell{R[0], R[1], R[2], R[3]}, g, fr[0], fr[1],
                fr[2], fr[3], fr[4], fr[5], Cons{fb[0], fb[1], fb[2], fb[3]}, lb, rb);
  Lb[0] = lb.d; Lb[1] = lb.p; Lb[2] = lb.vx; Lb[3] = lb.vy;
  Rb[0] = rb.d; Rb[1] = rb.p; Rb[2] = rb.vx; Rb[3] = rb.vy;
}

void ac_recover_vjp(double g, double vol, const double* x, const double* cb, double* xb) {
  recover_vjp(g, vol, x[0], x[1], x[2], x[3], Cell{cb[0], cb[1], cb[2], cb[3]}, xb[0], xb[1],
              xb[2], xb[3]);
}

void ac_init_vjp(double g, double vol, const double* c, const double* xb, double* cb) {
  Cell r{0, 0, 0, 0};
  init_vjp(g, vol, Cell{c[0], c[1], c[2], c[3]}, xb[0], xb[1], xb[2], xb[3], r);
  cb[0] = r.d; cb[1] = r.p; cb[2] = r.vx; cb[3] = r.vy;
}

void ac_cfl_inv_vjp(double g, double width, const double* c, double invb, double* cb) {
  Cell r{0, 0, 0, 0};
  cfl_inv_vjp(g, width, Cell{c[0], c[1], c[2], c[3]}, invb, r);
  cb[0] = r.d; cb[1] = r.p; cb[2] = r.vx; cb[3] = r.vy;
}

}  // extern "C"
