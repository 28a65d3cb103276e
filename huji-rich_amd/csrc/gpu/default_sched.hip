// The kernels built under the compiler's default instruction scheduler
// (richgpu.hip, the rest of the library, is built under the max-ILP one):
// the FP64 slope kernel of the step loops (k_slopes: LinearGaussImproved's
// slopes, LinearGaussImproved.cpp:536-570) and the mixed-precision stage
// kernels (k_slopes_mixed, k_face_update_mixed).  On the 2M-cell KH mesh
// they run at 126, 64 and 88 VGPRs and measured 169 / 175, 97.5 / 98.9 and
// 125.8 / 128.4 us against the max-ILP build, with the same results.  The
// launchers below are the library's only way in.
#include <hip/hip_runtime.h>

#define RG_GLOBAL template <int = 0> __global__
#define RG_GLOBAL_DEFAULT_SCHED __global__
#include "kernels_mixed.hpp"

namespace rg {

hipError_t launch_slopes(int blocks, int threads, hipStream_t st, DMesh m, const D4* q, Rec* out,
                         DParams prm, DErr err) {
  k_slopes<<<blocks, threads, 0, st>>>(m, q, out, prm, err);
  return hipPeekAtLastError();
}

hipError_t launch_slopes_mixed(int blocks, int threads, hipStream_t st, DMesh m, MTables t,
                               const F4* q, RecM* out, DParams prm, DErr err) {
  k_slopes_mixed<<<blocks, threads, 0, st>>>(m, t, q, out, prm, err);
  return hipPeekAtLastError();
}

hipError_t launch_face_update_mixed(int blocks, int threads, size_t lds, hipStream_t st, DMesh m,
                                    MTables t, DTiles tl, const F4* q, const RecM* rec,
                                    UpdArgsM a, DParams prm, DErr errf, DErr erru) {
  k_face_update_mixed<<<blocks, threads, lds, st>>>(m, t, tl, q, rec, a, prm, errf, erru);
  return hipPeekAtLastError();
}

}  // namespace rg
