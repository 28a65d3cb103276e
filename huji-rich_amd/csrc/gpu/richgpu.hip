// C ABI of the MI355X finite-volume step (include/richgpu.h): device buffers,
// mesh preprocessing, launches, error and timing plumbing.
#include "../../../include/richgpu.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <type_traits>
#include <vector>

#include "kernels.hpp"

using namespace rg;

namespace {

enum KernelId : int {
  KID_CFL = 0,
  KID_SLOPES,
  KID_FLUX2,
  KID_FLUX1,
  KID_UPDATE,
  KID_FACE_STATES,
  KID_INIT_EXT,
  KID_FILL,
  KID_COUNT
};
const char* const kKernelNames[KID_COUNT] = {"k_cfl",    "k_slopes",      "k_flux2",
                                             "k_flux1",  "k_update",      "k_face_states",
                                             "k_init_ext", "k_fill_u64"};

const char* err_message(int code) {
  switch (code) {
    case RG_ERR_NEGATIVE_DENSITY: return "Negative density";
    case RG_ERR_NEGATIVE_THERMAL: return "Negative thermal energy";
    case RG_ERR_HLLC_NAN: return "Hllc::Solved entered with nan";
    case RG_ERR_INVALID_WAVE_SPEEDS: return "Invalid wave speeds in hllc solver";
    case RG_ERR_SINGULAR_MATRIX: return "Singular matrix";
    case RG_ERR_TOO_MANY_NEIGHBORS: return "Cell has too many neighbors";
    default: return "";
  }
}

struct TimedLaunch {
  int kid;
  hipEvent_t a, b;
};

constexpr int kBlock = 256;
inline unsigned blocks_for(int n) { return static_cast<unsigned>((n + kBlock - 1) / kBlock); }

}  // namespace

struct rg_ctx {
  rg_params prm{};
  hipStream_t stream = nullptr;
  bool own_stream = false;
  int device = 0;

  int N = 0, F = 0, G = 0, nnz = 0;
  bool has_mesh = false, has_cells = false, has_ext = false, has_flux = false;
  bool has_w = false;

  std::vector<void*> mesh_bufs;   // freed on re-mesh
  std::vector<void*> state_bufs;  // freed on re-mesh
  int64_t bytes = 0;

  DMesh dm{};
  DCells cells{}, mid_cells{}, left{}, right{};
  DExt ext{}, mid_ext{}, flux{};
  DSlopes slopes{};
  double* wbuf = nullptr;
  unsigned long long* words = nullptr;  // [0],[1] dt slots, [2] error word
  unsigned long long* host_words = nullptr;
  double* dt_hist = nullptr;
  int dt_hist_cap = 0;

  bool prof = false;
  std::vector<TimedLaunch> pending;
  std::vector<hipEvent_t> event_pool;
  int64_t launches[KID_COUNT] = {};
  double total_ms[KID_COUNT] = {};

  rg_error last{};
};

namespace {

int fail_hip(rg_ctx* c, hipError_t e) {
  if (c) {
    c->last = rg_error{};
    c->last.code = -1;
    std::snprintf(c->last.message, sizeof(c->last.message), "HIP: %s", hipGetErrorString(e));
  }
  return RG_E_HIP;
}

#define HIPCHK(c, expr)                     \
  do {                                      \
    const hipError_t e_ = (expr);           \
    if (e_ != hipSuccess) return fail_hip((c), e_); \
  } while (0)

template <class T>
hipError_t dalloc(rg_ctx* c, std::vector<void*>& owner, T** p, size_t n) {
  void* q = nullptr;
  const size_t bytes = sizeof(T) * (n ? n : 1);
  const hipError_t e = hipMalloc(&q, bytes);
  if (e != hipSuccess) return e;
  owner.push_back(q);
  c->bytes += static_cast<int64_t>(bytes);
  *p = static_cast<T*>(q);
  return hipSuccess;
}

void free_all(rg_ctx* c, std::vector<void*>& owner) {
  for (void* p : owner) (void)hipFree(p);
  owner.clear();
}

hipEvent_t ev_get(rg_ctx* c) {
  if (!c->event_pool.empty()) {
    hipEvent_t e = c->event_pool.back();
    c->event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  (void)hipEventCreate(&e);
  return e;
}

// Launch with optional HIP-event timing on the context's stream.
template <class Fn>
hipError_t launch(rg_ctx* c, int kid, Fn&& fn) {
  hipEvent_t a = nullptr, b = nullptr;
  if (c->prof) {
    a = ev_get(c);
    (void)hipEventRecord(a, c->stream);
  }
  fn();
  const hipError_t e = hipGetLastError();
  if (c->prof) {
    b = ev_get(c);
    (void)hipEventRecord(b, c->stream);
    c->pending.push_back(TimedLaunch{kid, a, b});
  }
  return e;
}

void drain_timing(rg_ctx* c) {
  if (c->pending.empty()) return;
  (void)hipStreamSynchronize(c->stream);
  for (const TimedLaunch& t : c->pending) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, t.a, t.b) == hipSuccess) {
      c->launches[t.kid] += 1;
      c->total_ms[t.kid] += ms;
    }
    c->event_pool.push_back(t.a);
    c->event_pool.push_back(t.b);
  }
  c->pending.clear();
}

DParams dparams(const rg_ctx* c) {
  return DParams{c->prm.gamma, c->prm.cfl, c->prm.delta_v, c->prm.theta, c->prm.delta_p,
                 c->prm.slf};
}

DErr derr(rg_ctx* c, int step, int phase) { return DErr{c->words + 2, step, phase}; }

hipError_t fill_words(rg_ctx* c, int first, int n, unsigned long long v) {
  return launch(c, KID_FILL, [&] {
    k_fill_u64<<<1, 64, 0, c->stream>>>(c->words + first, v, n);
  });
}

// Read back the error word; returns RG_OK or RG_E_DEVICE and fills c->last.
int collect_error(rg_ctx* c) {
  HIPCHK(c, hipMemcpyAsync(c->host_words, c->words, 3 * sizeof(unsigned long long),
                           hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const unsigned long long w = c->host_words[2];
  if (w == ~0ull) {
    c->last = rg_error{};
    return RG_OK;
  }
  rg_error e{};
  e.code = static_cast<int32_t>(w & 0xF);
  e.index = static_cast<int32_t>((w >> 4) & 0xFFFFFFFFull);
  e.phase = static_cast<int32_t>((w >> 36) & 0x7);
  e.step = static_cast<int32_t>(w >> 39);
  std::snprintf(e.message, sizeof(e.message), "%s", err_message(e.code));
  c->last = e;
  // clear for the next call
  (void)fill_words(c, 2, 1, ~0ull);
  return RG_E_DEVICE;
}

bool ready(const rg_ctx* c, bool need_cells, bool need_ext) {
  return c && c->has_mesh && (!need_cells || c->has_cells) && (!need_ext || c->has_ext);
}

hipError_t run_slopes(rg_ctx* c, const DCells& cells, int step, int phase) {
  return launch(c, KID_SLOPES, [&] {
    k_slopes<<<blocks_for(c->N), kBlock, 0, c->stream>>>(c->dm, cells, c->slopes, dparams(c),
                                                         derr(c, step, phase));
  });
}

// FluxCalculator on `cells` into c->flux.
hipError_t run_flux(rg_ctx* c, const DCells& cells, int step, int phase_base) {
  if (c->prm.order == 2) {
    hipError_t e = run_slopes(c, cells, step, phase_base);
    if (e != hipSuccess) return e;
    return launch(c, KID_FLUX2, [&] {
      k_flux2<<<blocks_for(c->F), kBlock, 0, c->stream>>>(c->dm, cells, c->slopes, c->flux,
                                                           dparams(c), derr(c, step, phase_base + 1));
    });
  }
  return launch(c, KID_FLUX1, [&] {
    k_flux1<<<blocks_for(c->F), kBlock, 0, c->stream>>>(c->dm, cells, c->flux, dparams(c),
                                                         derr(c, step, phase_base + 1));
  });
}

hipError_t run_update(rg_ctx* c, const UpdArgs& a, int step, int phase) {
  return launch(c, KID_UPDATE, [&] {
    k_update<<<blocks_for(c->N), kBlock, 0, c->stream>>>(c->dm, a, dparams(c),
                                                         derr(c, step, phase));
  });
}

UpdArgs upd_args() {
  UpdArgs a{};
  return a;
}

}  // namespace

extern "C" {

int rg_create(const rg_params* params, rg_ctx** out) {
  if (!params || !out) return RG_E_ARG;
  if (!(params->gamma > 1.0) || !(params->cfl > 0.0) || (params->order != 1 && params->order != 2) ||
      (params->bc != RG_BC_RIGID && params->bc != RG_BC_PERIODIC) ||
      (params->precision != RG_FP64 && params->precision != RG_MIXED))
    return RG_E_ARG;
  rg_ctx* c = new (std::nothrow) rg_ctx();
  if (!c) return RG_E_ARG;
  c->prm = *params;
  hipError_t e = hipGetDevice(&c->device);
  if (e == hipSuccess) e = hipStreamCreate(&c->stream);
  if (e == hipSuccess) c->own_stream = true;
  if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&c->host_words),
                                         4 * sizeof(unsigned long long), hipHostMallocDefault);
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&c->words), 4 * sizeof(unsigned long long));
  if (e != hipSuccess) {
    rg_destroy(c);
    return RG_E_HIP;
  }
  if (fill_words(c, 0, 4, ~0ull) != hipSuccess) {
    rg_destroy(c);
    return RG_E_HIP;
  }
  *out = c;
  return RG_OK;
}

void rg_destroy(rg_ctx* c) {
  if (!c) return;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  free_all(c, c->mesh_bufs);
  free_all(c, c->state_bufs);
  if (c->dt_hist) (void)hipFree(c->dt_hist);
  if (c->words) (void)hipFree(c->words);
  if (c->host_words) (void)hipHostFree(c->host_words);
  for (const TimedLaunch& t : c->pending) {
    (void)hipEventDestroy(t.a);
    (void)hipEventDestroy(t.b);
  }
  for (hipEvent_t ev : c->event_pool) (void)hipEventDestroy(ev);
  if (c->own_stream && c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

int rg_set_stream(rg_ctx* c, void* s) {
  if (!c) return RG_E_ARG;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (c->own_stream) {
    (void)hipStreamDestroy(c->stream);
    c->own_stream = false;
  }
  if (s) {
    c->stream = static_cast<hipStream_t>(s);
  } else {
    HIPCHK(c, hipStreamCreate(&c->stream));
    c->own_stream = true;
  }
  return RG_OK;
}

int rg_set_mesh(rg_ctx* c, const rg_mesh* m) {
  if (!c || !m) return RG_E_ARG;
  const int N = m->n_cells, F = m->n_faces, G = m->n_ghosts;
  if (N <= 0 || F <= 0 || G < 0 || F > static_cast<int>(INFO_FACE)) return RG_E_ARG;
  if (!m->face_first || !m->face_second || !m->v1x || !m->v1y || !m->v2x || !m->v2y ||
      !m->pt_x || !m->pt_y || !m->cm_x || !m->cm_y || !m->volume || !m->tess_volume ||
      !m->csr_off || !m->csr_face || (G > 0 && !m->ghost_orig))
    return RG_E_ARG;
  // --- validate topology on the host
  if (m->csr_off[0] != 0) return RG_E_ARG;
  for (int i = 0; i < N; ++i)
    if (m->csr_off[i + 1] < m->csr_off[i]) return RG_E_ARG;
  const int nnz = m->csr_off[N];
  for (int f = 0; f < F; ++f) {
    const int a = m->face_first[f], b = m->face_second[f];
    if (a < 0 || b < 0 || a >= N + G || b >= N + G || a == b) return RG_E_ARG;
    if (a >= N && b >= N) return RG_E_ARG;
  }
  for (int g = 0; g < G; ++g)
    if (m->ghost_orig[g] < 0 || m->ghost_orig[g] >= N) return RG_E_ARG;
  auto orig = [&](int id) { return id < N ? id : m->ghost_orig[id - N]; };
  std::vector<unsigned> info(static_cast<size_t>(nnz));
  std::vector<int> nbr(static_cast<size_t>(nnz));
  for (int i = 0; i < N; ++i) {
    int prev = -1;
    for (int k = m->csr_off[i]; k < m->csr_off[i + 1]; ++k) {
      const int f = m->csr_face[k];
      if (f < 0 || f >= F || f <= prev) return RG_E_ARG;  // ascending face order
      prev = f;
      const int a = m->face_first[f], b = m->face_second[f];
      if (a != i && b != i) return RG_E_ARG;
      unsigned w = static_cast<unsigned>(f);
      if (b == i) w |= INFO_SECOND;
      if (orig(a) == orig(b)) w |= INFO_MIRROR;
      info[k] = w;
      nbr[k] = (a == i) ? b : a;
    }
  }
  // ghost sources: RigidWallGenerator copies the real cell of the ghost's
  // first boundary face (flat_map insert keeps the first) and reverses the
  // velocity along the ghost-minus-real direction; PeriodicGhostGenerator
  // copies the original.
  std::vector<int> gsrc(static_cast<size_t>(G));
  std::vector<double> gnx(static_cast<size_t>(G), 0.0), gny(static_cast<size_t>(G), 0.0);
  std::vector<char> seen(static_cast<size_t>(G), 0);
  for (int g = 0; g < G; ++g) gsrc[g] = m->ghost_orig[g];
  if (c->prm.bc == RG_BC_RIGID) {
    for (int f = 0; f < F; ++f) {
      const int a = m->face_first[f], b = m->face_second[f];
      const int side = (a >= N) ? 1 : ((b >= N) ? 2 : 0);
      if (!side) continue;
      const int g = (side == 1 ? a : b) - N;
      if (seen[g]) continue;
      seen[g] = 1;
      const int real = side == 1 ? b : a;
      gsrc[g] = real;
      double nx, ny;
      if (side == 1) {
        nx = m->pt_x[a] - m->pt_x[b];
        ny = m->pt_y[a] - m->pt_y[b];
      } else {
        nx = m->pt_x[b] - m->pt_x[a];
        ny = m->pt_y[b] - m->pt_y[a];
      }
      const double an = std::sqrt(nx * nx + ny * ny);
      gnx[g] = nx / an;
      gny[g] = ny / an;
    }
  }
  // SlabSymmetry::calcArea (physical_geometry.cpp:8-11)
  std::vector<double> area(static_cast<size_t>(F));
  for (int f = 0; f < F; ++f) {
    const double dx = m->v1x[f] - m->v2x[f], dy = m->v1y[f] - m->v2y[f];
    area[f] = std::sqrt(dx * dx + dy * dy);
  }
  // --- upload
  HIPCHK(c, hipStreamSynchronize(c->stream));
  free_all(c, c->mesh_bufs);
  free_all(c, c->state_bufs);
  c->bytes = 0;
  c->has_mesh = c->has_cells = c->has_ext = c->has_flux = c->has_w = false;
  c->wbuf = nullptr;
  c->left = DCells{};
  c->right = DCells{};
  const size_t NG = static_cast<size_t>(N) + G;
  DMesh d{};
  d.N = N;
  d.F = F;
  d.G = G;
  d.bc = c->prm.bc;
  auto up = [&](auto** dst, const auto* src, size_t n) -> hipError_t {
    using T = std::remove_const_t<std::remove_pointer_t<decltype(src)>>;
    T* p = nullptr;
    hipError_t e = dalloc(c, c->mesh_bufs, &p, n);
    if (e != hipSuccess) return e;
    if (n) e = hipMemcpyAsync(p, src, sizeof(T) * n, hipMemcpyHostToDevice, c->stream);
    *dst = p;
    return e;
  };
  int* p_ff = nullptr;
  int* p_fs = nullptr;
  double *p_v1x, *p_v1y, *p_v2x, *p_v2y, *p_px, *p_py, *p_cmx, *p_cmy, *p_vol, *p_tvol, *p_area;
  int *p_off, *p_nbr, *p_gorig, *p_gsrc;
  unsigned* p_info;
  double *p_gnx, *p_gny;
  const int zero_int = 0;
  HIPCHK(c, up(&p_ff, m->face_first, F));
  HIPCHK(c, up(&p_fs, m->face_second, F));
  HIPCHK(c, up(&p_v1x, m->v1x, F));
  HIPCHK(c, up(&p_v1y, m->v1y, F));
  HIPCHK(c, up(&p_v2x, m->v2x, F));
  HIPCHK(c, up(&p_v2y, m->v2y, F));
  HIPCHK(c, up(&p_px, m->pt_x, NG));
  HIPCHK(c, up(&p_py, m->pt_y, NG));
  HIPCHK(c, up(&p_cmx, m->cm_x, NG));
  HIPCHK(c, up(&p_cmy, m->cm_y, NG));
  HIPCHK(c, up(&p_vol, m->volume, N));
  HIPCHK(c, up(&p_tvol, m->tess_volume, N));
  HIPCHK(c, up(&p_area, area.data(), F));
  HIPCHK(c, up(&p_off, m->csr_off, static_cast<size_t>(N) + 1));
  HIPCHK(c, up(&p_info, info.data(), static_cast<size_t>(nnz)));
  HIPCHK(c, up(&p_nbr, nbr.data(), static_cast<size_t>(nnz)));
  HIPCHK(c, up(&p_gorig, G ? m->ghost_orig : &zero_int, G ? G : 1));
  HIPCHK(c, up(&p_gsrc, G ? gsrc.data() : &zero_int, G ? G : 1));
  const double zd = 0.0;
  HIPCHK(c, up(&p_gnx, G ? gnx.data() : &zd, G ? G : 1));
  HIPCHK(c, up(&p_gny, G ? gny.data() : &zd, G ? G : 1));
  d.ff = p_ff;
  d.fs = p_fs;
  d.v1x = p_v1x;
  d.v1y = p_v1y;
  d.v2x = p_v2x;
  d.v2y = p_v2y;
  d.px = p_px;
  d.py = p_py;
  d.cmx = p_cmx;
  d.cmy = p_cmy;
  d.vol = p_vol;
  d.tvol = p_tvol;
  d.area = p_area;
  d.off = p_off;
  d.info = p_info;
  d.nbr = p_nbr;
  d.gorig = p_gorig;
  d.gsrc = p_gsrc;
  d.gnx = p_gnx;
  d.gny = p_gny;
  d.wx = nullptr;
  d.wy = nullptr;
  // --- state buffers
  auto alloc4 = [&](double** q, size_t n) -> hipError_t {
    for (int k = 0; k < 4; ++k) {
      const hipError_t e = dalloc(c, c->state_bufs, &q[k], n);
      if (e != hipSuccess) return e;
    }
    return hipSuccess;
  };
  double* q[4];
  HIPCHK(c, alloc4(q, N));
  c->cells = DCells{q[0], q[1], q[2], q[3]};
  HIPCHK(c, alloc4(q, N));
  c->mid_cells = DCells{q[0], q[1], q[2], q[3]};
  HIPCHK(c, alloc4(q, N));
  c->ext = DExt{q[0], q[1], q[2], q[3]};
  HIPCHK(c, alloc4(q, N));
  c->mid_ext = DExt{q[0], q[1], q[2], q[3]};
  HIPCHK(c, alloc4(q, F));
  c->flux = DExt{q[0], q[1], q[2], q[3]};
  for (int k = 0; k < 8; ++k) HIPCHK(c, dalloc(c, c->state_bufs, &c->slopes.s[k], N));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->dm = d;
  c->N = N;
  c->F = F;
  c->G = G;
  c->nnz = nnz;
  c->has_mesh = true;
  return RG_OK;
}

int rg_set_face_velocity(rg_ctx* c, const double* wx, const double* wy) {
  if (!c || !c->has_mesh || (!wx) != (!wy)) return RG_E_ARG;
  if (!wx) {
    c->dm.wx = c->dm.wy = nullptr;
    c->has_w = false;
    return RG_OK;
  }
  if (!c->wbuf) HIPCHK(c, dalloc(c, c->mesh_bufs, &c->wbuf, 2 * static_cast<size_t>(c->F)));
  HIPCHK(c, hipMemcpyAsync(c->wbuf, wx, sizeof(double) * c->F, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->wbuf + c->F, wy, sizeof(double) * c->F, hipMemcpyHostToDevice,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->dm.wx = c->wbuf;
  c->dm.wy = c->wbuf + c->F;
  c->has_w = true;
  return RG_OK;
}

static int copy4_in(rg_ctx* c, double* const dst[4], const double* const src[4], size_t n) {
  for (int k = 0; k < 4; ++k)
    if (!src[k]) return RG_E_ARG;
  for (int k = 0; k < 4; ++k)
    HIPCHK(c, hipMemcpyAsync(dst[k], src[k], sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return RG_OK;
}

static int copy4_out(rg_ctx* c, double* const dst[4], double* const src[4], size_t n) {
  for (int k = 0; k < 4; ++k)
    if (dst[k])
      HIPCHK(c, hipMemcpyAsync(dst[k], src[k], sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return RG_OK;
}

int rg_set_cells(rg_ctx* c, const double* d, const double* p, const double* vx, const double* vy) {
  if (!ready(c, false, false)) return RG_E_ARG;
  double* const dst[4] = {c->cells.d, c->cells.p, c->cells.vx, c->cells.vy};
  const double* const src[4] = {d, p, vx, vy};
  const int rc = copy4_in(c, dst, src, c->N);
  if (rc == RG_OK) c->has_cells = true;
  return rc;
}

int rg_get_cells(rg_ctx* c, double* d, double* p, double* vx, double* vy) {
  if (!ready(c, true, false)) return RG_E_ARG;
  double* const dst[4] = {d, p, vx, vy};
  double* const src[4] = {c->cells.d, c->cells.p, c->cells.vx, c->cells.vy};
  return copy4_out(c, dst, src, c->N);
}

int rg_set_extensives(rg_ctx* c, const double* m, const double* E, const double* Mx,
                      const double* My) {
  if (!ready(c, false, false)) return RG_E_ARG;
  double* const dst[4] = {c->ext.m, c->ext.E, c->ext.Mx, c->ext.My};
  const double* const src[4] = {m, E, Mx, My};
  const int rc = copy4_in(c, dst, src, c->N);
  if (rc == RG_OK) c->has_ext = true;
  return rc;
}

int rg_get_extensives(rg_ctx* c, double* m, double* E, double* Mx, double* My) {
  if (!ready(c, false, true)) return RG_E_ARG;
  double* const dst[4] = {m, E, Mx, My};
  double* const src[4] = {c->ext.m, c->ext.E, c->ext.Mx, c->ext.My};
  return copy4_out(c, dst, src, c->N);
}

int rg_set_fluxes(rg_ctx* c, const double* fm, const double* fE, const double* fMx,
                  const double* fMy) {
  if (!ready(c, false, false)) return RG_E_ARG;
  double* const dst[4] = {c->flux.m, c->flux.E, c->flux.Mx, c->flux.My};
  const double* const src[4] = {fm, fE, fMx, fMy};
  const int rc = copy4_in(c, dst, src, c->F);
  if (rc == RG_OK) c->has_flux = true;
  return rc;
}

int rg_init_extensives(rg_ctx* c) {
  if (!ready(c, true, false)) return RG_E_ARG;
  HIPCHK(c, launch(c, KID_INIT_EXT, [&] {
    k_init_ext<<<blocks_for(c->N), kBlock, 0, c->stream>>>(c->dm, c->cells, c->ext, c->prm.gamma);
  }));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->has_ext = true;
  return RG_OK;
}

int rg_time_step(rg_ctx* c, double* dt) {
  if (!ready(c, true, false) || !dt) return RG_E_ARG;
  HIPCHK(c, fill_words(c, 0, 1, DBL_INF_BITS));
  HIPCHK(c, launch(c, KID_CFL, [&] {
    k_cfl<<<blocks_for(c->N), kBlock, 0, c->stream>>>(c->dm, c->cells, c->prm.gamma, c->words);
  }));
  const int rc = collect_error(c);
  if (rc != RG_OK) return rc;
  double v;
  std::memcpy(&v, &c->host_words[0], sizeof(double));
  *dt = c->prm.cfl * v;
  return RG_OK;
}

int rg_reconstruct(rg_ctx* c, double* slopes, double* left, double* right) {
  if (!ready(c, true, false) || c->prm.order != 2) return RG_E_ARG;
  HIPCHK(c, run_slopes(c, c->cells, 0, 1));
  if (left || right) {
    if (!c->left.d) {
      double* q[8];
      for (int k = 0; k < 8; ++k) HIPCHK(c, dalloc(c, c->state_bufs, &q[k], c->F));
      c->left = DCells{q[0], q[1], q[2], q[3]};
      c->right = DCells{q[4], q[5], q[6], q[7]};
    }
    HIPCHK(c, launch(c, KID_FACE_STATES, [&] {
      k_face_states<<<blocks_for(c->F), kBlock, 0, c->stream>>>(c->dm, c->cells, c->slopes, c->left,
                                                               c->right);
    }));
  }
  const int rc = collect_error(c);
  if (rc != RG_OK) return rc;
  if (slopes)
    for (int k = 0; k < 8; ++k)
      HIPCHK(c, hipMemcpyAsync(slopes + static_cast<size_t>(k) * c->N, c->slopes.s[k],
                               sizeof(double) * c->N, hipMemcpyDeviceToHost, c->stream));
  const DCells* lr[2] = {&c->left, &c->right};
  double* outs[2] = {left, right};
  for (int s = 0; s < 2; ++s) {
    if (!outs[s]) continue;
    double* src[4] = {lr[s]->d, lr[s]->p, lr[s]->vx, lr[s]->vy};
    for (int k = 0; k < 4; ++k)
      HIPCHK(c, hipMemcpyAsync(outs[s] + static_cast<size_t>(k) * c->F, src[k],
                               sizeof(double) * c->F, hipMemcpyDeviceToHost, c->stream));
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return RG_OK;
}

int rg_flux(rg_ctx* c, double* fm, double* fE, double* fMx, double* fMy) {
  if (!ready(c, true, false)) return RG_E_ARG;
  HIPCHK(c, run_flux(c, c->cells, 0, 1));
  const int rc = collect_error(c);
  if (rc != RG_OK) return rc;
  c->has_flux = true;
  double* const dst[4] = {fm, fE, fMx, fMy};
  double* const src[4] = {c->flux.m, c->flux.E, c->flux.Mx, c->flux.My};
  return copy4_out(c, dst, src, c->F);
}

int rg_update_extensives(rg_ctx* c, double dt) {
  if (!ready(c, false, true) || !c->has_flux) return RG_E_ARG;
  UpdArgs a = upd_args();
  a.in = c->ext;
  a.out = c->ext;
  a.flux = c->flux;
  a.dt = dt;
  a.flags = UPD_FACES;
  HIPCHK(c, run_update(c, a, 0, 3));
  return collect_error(c);
}

int rg_update_cells(rg_ctx* c) {
  if (!ready(c, false, true)) return RG_E_ARG;
  UpdArgs a = upd_args();
  a.in = c->ext;
  a.out = c->ext;
  a.cells = c->cells;
  a.flags = UPD_CELLS;
  HIPCHK(c, run_update(c, a, 0, 3));
  const int rc = collect_error(c);
  if (rc == RG_OK) c->has_cells = true;
  return rc;
}

int rg_advance(rg_ctx* c, int32_t integrator, int32_t nsteps, double* dt_out) {
  if (!ready(c, true, true) || (integrator != 1 && integrator != 2) || nsteps < 0 ||
      nsteps >= (1 << 24))
    return RG_E_ARG;
  if (nsteps == 0) return RG_OK;
  if (dt_out && c->dt_hist_cap < nsteps) {
    if (c->dt_hist) (void)hipFree(c->dt_hist);
    c->dt_hist = nullptr;
    c->dt_hist_cap = 0;
    HIPCHK(c, hipMalloc(reinterpret_cast<void**>(&c->dt_hist), sizeof(double) * nsteps));
    c->dt_hist_cap = nsteps;
  }
  // first dt: SimpleCFL on the current cells into slot 0
  HIPCHK(c, fill_words(c, 0, 1, DBL_INF_BITS));
  HIPCHK(c, launch(c, KID_CFL, [&] {
    k_cfl<<<blocks_for(c->N), kBlock, 0, c->stream>>>(c->dm, c->cells, c->prm.gamma, c->words);
  }));
  for (int s = 0; s < nsteps; ++s) {
    unsigned long long* cur = c->words + (s & 1);
    unsigned long long* nxt = c->words + ((s + 1) & 1);
    if (integrator == 1) {
      // hdsim::TimeAdvance: fluxes, extensive update, ZeroForce, cell update
      HIPCHK(c, run_flux(c, c->cells, s, 1));
      UpdArgs a = upd_args();
      a.in = c->ext;
      a.out = c->ext;
      a.flux = c->flux;
      a.cells = c->cells;
      a.dt_slot = cur;
      a.reset_slot = nullptr;
      a.cfl_slot = nxt;
      a.dt_hist = dt_out ? c->dt_hist : nullptr;
      a.step = s;
      a.flags = UPD_FACES | UPD_ZF_AFTER | UPD_CELLS | UPD_CFL;
      // the next slot must read +inf before this update min-reduces into it
      HIPCHK(c, fill_words(c, (s + 1) & 1, 1, DBL_INF_BITS));
      HIPCHK(c, run_update(c, a, s, 3));
    } else {
      // hdsim::TimeAdvance2Heun, stage 1
      HIPCHK(c, run_flux(c, c->cells, s, 1));
      UpdArgs a = upd_args();
      a.in = c->ext;
      a.out = c->mid_ext;
      a.flux = c->flux;
      a.cells = c->mid_cells;
      a.dt_slot = cur;
      a.reset_slot = nxt;
      a.dt_hist = dt_out ? c->dt_hist : nullptr;
      a.step = s;
      a.flags = UPD_FACES | UPD_ZF_BEFORE | UPD_CELLS;
      HIPCHK(c, run_update(c, a, s, 3));
      // stage 2 on the mid cells, then 0.5*(mid + old) and the cell update
      HIPCHK(c, run_flux(c, c->mid_cells, s, 4));
      UpdArgs b = upd_args();
      b.in = c->mid_ext;
      b.out = c->ext;
      b.avg = c->ext;
      b.flux = c->flux;
      b.cells = c->cells;
      b.dt_slot = cur;
      b.cfl_slot = nxt;
      b.step = s;
      b.flags = UPD_FACES | UPD_ZF_BEFORE | UPD_AVERAGE | UPD_CELLS | UPD_CFL;
      HIPCHK(c, run_update(c, b, s, 6));
    }
  }
  if (dt_out)
    HIPCHK(c, hipMemcpyAsync(dt_out, c->dt_hist, sizeof(double) * nsteps, hipMemcpyDeviceToHost,
                             c->stream));
  return collect_error(c);
}

int rg_last_error(rg_ctx* c, rg_error* err) {
  if (!c || !err) return RG_E_ARG;
  *err = c->last;
  return RG_OK;
}

int rg_profile_enable(rg_ctx* c, int32_t on) {
  if (!c) return RG_E_ARG;
  drain_timing(c);
  c->prof = on != 0;
  if (c->prof)
    for (int k = 0; k < KID_COUNT; ++k) {
      c->launches[k] = 0;
      c->total_ms[k] = 0;
    }
  return RG_OK;
}

int32_t rg_kernel_count(void) { return KID_COUNT; }

const char* rg_kernel_name(int32_t id) {
  return (id >= 0 && id < KID_COUNT) ? kKernelNames[id] : "";
}

int rg_profile_read(rg_ctx* c, int32_t id, int64_t* launches, double* total_ms) {
  if (!c || id < 0 || id >= KID_COUNT) return RG_E_ARG;
  drain_timing(c);
  if (launches) *launches = c->launches[id];
  if (total_ms) *total_ms = c->total_ms[id];
  return RG_OK;
}

int64_t rg_device_bytes(rg_ctx* c) { return c ? c->bytes : 0; }

}  // extern "C"
