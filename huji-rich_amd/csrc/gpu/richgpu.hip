// C ABI of the MI355X finite-volume step (include/richgpu.h): device buffers,
// mesh preprocessing, launches, error and timing plumbing.
#include "../../../include/richgpu.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <type_traits>
#include <unordered_map>
#include <vector>

#include "adjoint.hpp"
#include "ensemble.hpp"
#include "moving.hpp"
#include "../mesh/parallel.hpp"
#include "kernels_mixed.hpp"

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
  KID_GEO,
  KID_PACK,
  KID_REDUCE,
  KID_SLOPES_MIXED,
  KID_FLUX_MIXED,
  KID_UPDATE_MIXED,
  KID_ADJ_FLUX,
  KID_ADJ_SLOPES,
  KID_ADJ_CELL,
  KID_ENS,
  KID_MOVING,
  KID_FLUX_PCM,
  KID_RIEMANN,
  KID_FACE_UPDATE,
  KID_FACE_UPDATE_MIXED,
  KID_COUNT
};
const char* const kKernelNames[KID_COUNT] = {"k_cfl",         "k_slopes",   "k_flux2",
                                             "k_flux1",       "k_update",   "k_face_states",
                                             "k_init_ext",    "k_fill_u64", "k_geo",
                                             "k_pack",        "k_reduce_min", "k_slopes_mixed",
                                             "k_flux_mixed",  "k_update_mixed", "k_adj_flux",
                                             "k_adj_slopes",  "k_adj_cell",     "k_ens",
                                             "k_moving",      "k_flux_pcm",     "k_riemann",
                                             "k_face_update", "k_face_update_mixed"};

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

// Scratch of rg_advance_noalloc: [0], [1] the dt slots, [2] the error word,
// [3] unused, then the CFL block minima (one per block of kBlock cells),
// then the dt history.  1 MiB; each GPU holds its own copy until the
// process ends and every call on that GPU shares it.  Host code never
// touches `scratch` by name: it uses the device address that
// hipGetSymbolAddress returns for the current device (g_scratch_ptr).
constexpr int kScratchWords = 4;
constexpr int kScratchBlocks = 65536;  // cells: up to kScratchBlocks * kBlock
constexpr int kScratchSteps = 65536;   // dt history entries
static __device__ __align__(256) double scratch[kScratchWords + kScratchBlocks + kScratchSteps];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device, looked up once per
// device.
hipError_t scratch_on_current_device(int* dev_out, double** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  *dev_out = dev;
  *out = g_scratch_ptr[dev];
  return hipSuccess;
}

}  // namespace

struct rg_ctx {
  rg_params prm{};
  hipStream_t stream = nullptr;
  bool own_stream = false;

  int N = 0, F = 0, G = 0, nnz = 0;
  int dm_ne = 0;  // ELL entries
  bool has_mesh = false, has_cells = false, has_ext = false, has_flux = false;

  std::vector<std::pair<void*, size_t>> mesh_bufs;  // released on re-mesh
  // buffers of the previous mesh while a new one is set, reused where the
  // size fits (a moving mesh's tables keep nearly the same sizes)
  std::vector<std::pair<void*, size_t>> pool;
  int64_t bytes = 0;

  DMesh dm{};
  D4 *prim = nullptr, *mid_prim = nullptr, *ext = nullptr, *mid_ext = nullptr, *flux = nullptr;
  D4 *left = nullptr, *right = nullptr;
  Rec* rec = nullptr;
  double* staging = nullptr;   // 4*max(N,F) doubles for SoA host transfers
  double* partials = nullptr;  // per-block minima of the CFL reduction
  double* wbuf = nullptr;
  unsigned long long* words = nullptr;  // [0],[1] dt slots, [2] error word
  unsigned long long* host_words = nullptr;
  double* dt_hist = nullptr;
  int dt_hist_cap = 0;

  // mixed precision (rg_advance with RG_MIXED)
  MTables mt{};
  F4 *primM = nullptr, *mid_primM = nullptr, *fluxM = nullptr;
  RecM* recM = nullptr;

  // adjoint: tables, the recorded forward run and the reverse buffers
  bool ext_init = false;  // extensives are init_extensives(cells)
  AdjMesh am{};
  int adj_steps = 0, adj_every = 0, adj_nck = 0, adj_cap_every = 0, adj_cap_nck = 0;
  std::vector<void*> adj_bufs;  // checkpoints and tape (sized per run)
  D4 *ckP = nullptr, *ckX = nullptr, *finP = nullptr, *finX = nullptr;
  // two tapes of a segment's recompute (states, mid states, both stages'
  // slopes, dts): the recompute of one segment fills one while the reverse
  // sweep of the next reads the other
  struct Tape {
    D4 *P = nullptr, *X = nullptr, *Pm = nullptr, *Xm = nullptr;
    Rec *R1 = nullptr, *R2 = nullptr;
    double* dt = nullptr;
  } tape[2];
  // the recompute's own stream, fluxes, dt and error words and block minima
  hipStream_t rstream = nullptr;
  D4* rflux = nullptr;
  unsigned long long* rwords = nullptr;
  double* rpartials = nullptr;
  hipEvent_t tape_ready[2] = {nullptr, nullptr}, tape_free[2] = {nullptr, nullptr};
  D4 *gP = nullptr, *gX = nullptr, *ga = nullptr, *gb = nullptr, *gps = nullptr;
  D4 *lb = nullptr, *rb = nullptr, *emit = nullptr;
  double *dtb = nullptr, *vmin = nullptr;
  int* imin = nullptr;

  // batched ensemble (rg_set_instances): per-instance dt, graph per step
  int ens_K = 0, ens_cu = 0, ens_block = 256, ens_bpc = 4, ens_hist_cap = 0;
  int *ens_inst = nullptr, *ens_coff = nullptr, *ens_step = nullptr;
  unsigned long long *ens_dtA = nullptr, *ens_dtB = nullptr;
  double *ens_inv = nullptr, *ens_hist = nullptr;
  hipGraphExec_t ens_graph = nullptr;

  // what the adjoint tables are built from, kept on the host until the
  // first rg_adjoint_forward of the mesh needs them
  struct AdjSource {
    std::vector<Inc> inc;        // [nnz]
    std::vector<int> csr_off;    // [N+1]
    std::vector<int> ff, fs;     // [F]
    std::vector<int> gsrc;       // [G]
    std::vector<int> wofs;       // [waves+1]
    std::vector<unsigned char> lanepos;  // [N] ELL column of each cell
    std::vector<unsigned> minfo;  // [nnz] face words of the mixed update (RG_MIXED)
  } adj_src;

  // moving mesh: point velocities kept across rg_replace_mesh, and where
  // the state waits while the mesh is swapped
  D2* pv = nullptr;
  int pv_cap = 0;
  D4* keep = nullptr;
  size_t keep_cap = 0;
  // tiles of the fused face-and-update kernel, built at the first Heun step
  // loop of a mesh (tiles_ok: they fit the kernel's LDS array), and the
  // second state buffers that loop alternates with
  DTiles tiles{};
  int tiles_max = 0;
  int64_t tiles_total = 0;
  bool tiles_built = false, tiles_ok = false;
  bool tiles_on = true;  // rg_use_tiles: the fused stages are allowed
  std::vector<void*> tile_bufs;
  D4 *alt_prim = nullptr, *alt_ext = nullptr;
  double* riemann_buf = nullptr;  // rg_riemann's states and fluxes
  size_t riemann_cap = 0;         // doubles
  bool pv_valid = false;

  // the rigid box (x0, x1, y0, y1) of rg_set_box, for device geometry
  double box[4] = {0, 0, 0, 0};
  bool has_box = false;

  // device of the context's stream and its copy of the noalloc scratch
  int device = 0;
  double* scratch = nullptr;

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

#define HIPCHK(c, expr)                              \
  do {                                               \
    const hipError_t e_ = (expr);                    \
    if (e_ != hipSuccess) return fail_hip((c), e_);  \
  } while (0)

template <class T>
hipError_t dalloc(rg_ctx* c, T** p, size_t n) {
  void* q = nullptr;
  size_t bytes = sizeof(T) * (n ? n : 1);
  // the smallest pooled buffer that holds it without wasting over a quarter
  size_t best = c->pool.size();
  for (size_t k = 0; k < c->pool.size(); ++k) {
    const size_t b = c->pool[k].second;
    if (b >= bytes && b <= bytes + bytes / 4 + 4096 &&
        (best == c->pool.size() || b < c->pool[best].second))
      best = k;
  }
  if (best < c->pool.size()) {
    q = c->pool[best].first;
    bytes = c->pool[best].second;
    c->pool[best] = c->pool.back();
    c->pool.pop_back();
  } else {
    const hipError_t e = hipMalloc(&q, bytes);
    if (e != hipSuccess) return e;
  }
  c->mesh_bufs.emplace_back(q, bytes);
  c->bytes += static_cast<int64_t>(bytes);
  *p = static_cast<T*>(q);
  return hipSuccess;
}

void drain_pool(rg_ctx* c) {
  for (const auto& b : c->pool) (void)hipFree(b.first);
  c->pool.clear();
}

template <class T>
hipError_t upload(rg_ctx* c, T** p, const T* src, size_t n) {
  hipError_t e = dalloc(c, p, n);
  if (e == hipSuccess && n)
    e = hipMemcpyAsync(*p, src, sizeof(T) * n, hipMemcpyHostToDevice, c->stream);
  return e;
}

// Release the mesh's buffers: freed, or with `pool` kept for the next
// mesh's dalloc (the caller has synchronized the stream).
void free_mesh(rg_ctx* c, bool pool = false) {
  for (const auto& b : c->mesh_bufs) {
    if (pool)
      c->pool.push_back(b);
    else
      (void)hipFree(b.first);
  }
  c->mesh_bufs.clear();
  c->bytes = 0;
  c->prim = c->mid_prim = c->ext = c->mid_ext = c->flux = c->left = c->right = nullptr;
  c->rec = nullptr;
  c->staging = c->partials = c->wbuf = nullptr;
  c->mt = MTables{};
  c->primM = c->mid_primM = c->fluxM = nullptr;
  for (void* p : c->adj_bufs) (void)hipFree(p);
  c->adj_bufs.clear();
  for (void* p : c->tile_bufs) (void)hipFree(p);
  c->tile_bufs.clear();
  c->tiles = DTiles{};
  c->tiles_max = 0;
  c->tiles_total = 0;
  c->tiles_built = c->tiles_ok = false;
  c->alt_prim = c->alt_ext = nullptr;
  c->am = AdjMesh{};
  c->adj_src = rg_ctx::AdjSource{};
  c->adj_steps = c->adj_every = c->adj_nck = c->adj_cap_every = c->adj_cap_nck = 0;
  c->ckP = c->ckX = c->finP = c->finX = nullptr;
  c->tape[0] = c->tape[1] = rg_ctx::Tape{};
  c->rflux = nullptr;
  c->rwords = nullptr;
  c->rpartials = nullptr;
  c->gP = c->gX = c->ga = c->gb = c->gps = c->lb = c->rb = c->emit = nullptr;
  c->dtb = c->vmin = nullptr;
  c->imin = nullptr;
  c->ext_init = false;
  if (c->ens_graph) (void)hipGraphExecDestroy(c->ens_graph);
  c->ens_graph = nullptr;
  if (c->ens_hist) (void)hipFree(c->ens_hist);
  c->ens_hist = nullptr;
  c->ens_hist_cap = 0;
  c->ens_K = 0;
  c->ens_inst = c->ens_coff = c->ens_step = nullptr;
  c->ens_dtA = c->ens_dtB = nullptr;
  c->ens_inv = nullptr;
  c->recM = nullptr;
  c->has_mesh = c->has_cells = c->has_ext = c->has_flux = false;
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
                 c->prm.slf, 0};
}
DParams dparams_walk(const rg_ctx* c, int rev) {
  DParams p = dparams(c);
  p.rev = rev;
  return p;
}

DErr derr(rg_ctx* c, int step, int phase) { return DErr{c->words + 2, step, phase}; }
// The minimum of the CFL block minima into dt slot `slot`: one per block of
// kBlock cells, or per tile (`tiles`) after a fused stage.
inline unsigned tiles_for(int n) { return static_cast<unsigned>((n + kTile - 1) / kTile); }
hipError_t run_reduce_min(rg_ctx* c, int slot, bool tiles = false) {
  const int n = static_cast<int>(tiles ? tiles_for(c->N) : blocks_for(c->N));
  return launch(c, KID_REDUCE, [&] {
    k_reduce_min<<<1, 1024, 0, c->stream>>>(c->partials, n, c->words + slot);
  });
}

hipError_t fill_words(rg_ctx* c, int first, int n, unsigned long long v) {
  return launch(c, KID_FILL, [&] {
    k_fill_u64<<<1, 64, 0, c->stream>>>(c->words + first, v, n);
  });
}

// Where a call's device-detected error can be traced back to: the cells
// each flux stage read (phases 1-2 and 4-5) and the buffers the stage
// updates wrote (phases 3 and 6).  Null: not available.
struct ErrSrc {
  const D4* cells[2];
  const D4* ext[2];
};

// hllc.cpp:54-70 adds them in this order: left (density, pressure, x and y
// velocity, sound speed, energy), right, the interface velocity, then the
// three wave speeds
const char* const kWaveEntryNames[16] = {
    "left density",     "left pressure",     "left x velocity",  "left y velocity",
    "left sound speed", "left energy",       "right density",    "right pressure",
    "right x velocity", "right y velocity",  "right sound speed", "right energy",
    "interface velocity", "left wave speed", "center wave speed", "right wave speed"};

void add_entry(rg_error& e, const char* name, double v) {
  if (e.n_entries >= 16) return;
  rg_error_entry& x = e.entries[e.n_entries++];
  std::snprintf(x.name, sizeof(x.name), "%s", name);
  x.value = v;
}

template <class T>
hipError_t read_one(rg_ctx* c, const T* src, size_t i, T* out) {
  return hipMemcpyAsync(out, src + i, sizeof(T), hipMemcpyDeviceToHost, c->stream);
}

hipError_t run_slopes(rg_ctx* c, const D4* cells, int step, int phase, Rec* out,
                      int rev = 0);

// The UniversalError entries of error e (see rg_error), recomputed from the
// inputs the failing call read.
hipError_t fill_entries(rg_ctx* c, rg_error& e, const ErrSrc& src) {
  const int i = e.index;
  const DMesh& m = c->dm;
  if (e.code == RG_ERR_SINGULAR_MATRIX || e.code == RG_ERR_TOO_MANY_NEIGHBORS) {
    if (i < 0 || i >= c->N) return hipSuccess;
    double x = 0, y = 0, vol = 0;
    CellGeo g{};
    hipError_t r = read_one(c, m.px, i, &x);
    if (r == hipSuccess) r = read_one(c, m.py, i, &y);
    if (r == hipSuccess) r = read_one(c, m.tvol, i, &vol);
    if (r == hipSuccess) r = read_one(c, m.cgeo, i, &g);
    if (r == hipSuccess) r = hipStreamSynchronize(c->stream);
    if (r != hipSuccess) return r;
    add_entry(e, "Cell x cor", x);
    add_entry(e, "Cell y cor", y);
    if (e.code == RG_ERR_SINGULAR_MATRIX) {
      add_entry(e, "Cell volume", vol);
      add_entry(e, "Det was", g.c0);
    }
    return hipSuccess;
  }
  const int stage = e.phase >= 4 ? 1 : 0;
  if (e.code == RG_ERR_NEGATIVE_THERMAL) {
    const D4* x = src.ext[stage];
    if (!x || i < 0 || i >= c->N) return hipSuccess;
    D4 v{};
    hipError_t r = read_one(c, x, i, &v);
    if (r == hipSuccess) r = hipStreamSynchronize(c->stream);
    if (r != hipSuccess) return r;
    add_entry(e, "Cell index", static_cast<double>(i));
    add_entry(e, "Cell mass", v.a);
    add_entry(e, "Cell x momentum", v.c);
    add_entry(e, "Cell y momentum", v.d);
    add_entry(e, "Cell energy", v.b);
    return hipSuccess;
  }
  if (e.code == RG_ERR_INVALID_WAVE_SPEEDS) {
    const D4* q = src.cells[stage];
    if (!q || i < 0 || i >= c->F) return hipSuccess;
    hipError_t r = hipSuccess;
    if (c->prm.order == 2) r = run_slopes(c, q, 0, 1, c->rec);
    if (r != hipSuccess) return r;
    k_face_entries<<<1, 1, 0, c->stream>>>(m, q, c->rec, c->prm.order, c->prm.gamma, i,
                                            c->staging);
    r = hipGetLastError();
    double v[16];
    if (r == hipSuccess)
      r = hipMemcpyAsync(v, c->staging, sizeof(v), hipMemcpyDeviceToHost, c->stream);
    if (r == hipSuccess) r = hipStreamSynchronize(c->stream);
    if (r != hipSuccess) return r;
    for (int k = 0; k < 16; ++k) add_entry(e, kWaveEntryNames[k], v[k]);
  }
  return hipSuccess;
}

// Read back the error word; returns RG_OK or RG_E_DEVICE and fills c->last
// (with the reference's entries where `src` locates the failing call's
// inputs).
int collect_error(rg_ctx* c, const ErrSrc* src = nullptr) {
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
  if (src) {
    const hipError_t r = fill_entries(c, e, *src);
    if (r != hipSuccess) e.n_entries = 0;
  }
  c->last = e;
  (void)fill_words(c, 2, 1, ~0ull);  // clear for the next call
  return RG_E_DEVICE;
}

bool ready(const rg_ctx* c, bool need_cells, bool need_ext) {
  return c && c->has_mesh && (!need_cells || c->has_cells) && (!need_ext || c->has_ext);
}

hipError_t run_slopes(rg_ctx* c, const D4* cells, int step, int phase, Rec* out, int rev) {
  Rec* r = out ? out : c->rec;
  return launch(c, KID_SLOPES, [&] {
    (void)launch_slopes(blocks_for(c->N), kBlock, c->stream, c->dm, cells, r, dparams_walk(c, rev),
                        derr(c, step, phase));
  });
}

// FluxCalculator on `cells` into c->flux (slopes into `rec`, default c->rec).
// With a dt slot, the modular calculators write dt*area*F (for UPD_DELTA).
// `rev` walks the slope kernel's blocks in reverse and the face kernel's
// forward (0: the other way round).
hipError_t run_flux(rg_ctx* c, const D4* cells, int step, int phase_base, Rec* rec = nullptr,
                    const unsigned long long* dt_slot = nullptr, int rev = 0) {
  if (c->prm.order == 2) {
    Rec* r = rec ? rec : c->rec;
    const hipError_t e = run_slopes(c, cells, step, phase_base, r, rev);
    if (e != hipSuccess) return e;
    return launch(c, KID_FLUX2, [&] {
      k_flux2<<<blocks_for(c->F), kBlock, 0, c->stream>>>(c->dm, cells, r, c->flux,
                                                           dparams_walk(c, 1 - rev),
                                                           derr(c, step, phase_base + 1), dt_slot);
    });
  }
  if (c->prm.order == 3)
    return launch(c, KID_FLUX_PCM, [&] {
      k_flux_pcm<<<blocks_for(c->F), kBlock, 0, c->stream>>>(
          c->dm, cells, c->flux, dparams(c), derr(c, step, phase_base + 1), dt_slot);
    });
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

hipError_t run_cfl(rg_ctx* c, const D4* cells, int slot) {
  const hipError_t e = launch(c, KID_CFL, [&] {
    k_cfl<<<blocks_for(c->N), kBlock, 0, c->stream>>>(c->dm, cells, c->prm.gamma, c->partials);
  });
  if (e != hipSuccess) return e;
  return run_reduce_min(c, slot);
}

// host SoA [4][n] <-> device AoS
int put4(rg_ctx* c, D4* dst, const double* const src[4], int n) {
  for (int k = 0; k < 4; ++k)
    if (!src[k]) return RG_E_ARG;
  for (int k = 0; k < 4; ++k)
    HIPCHK(c, hipMemcpyAsync(c->staging + static_cast<size_t>(k) * n, src[k], sizeof(double) * n,
                             hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, launch(c, KID_PACK, [&] {
    k_pack4<<<blocks_for(n), kBlock, 0, c->stream>>>(c->staging, dst, n);
  }));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return RG_OK;
}

int get4(rg_ctx* c, double* const dst[4], const D4* src, int n) {
  HIPCHK(c, launch(c, KID_PACK, [&] {
    k_unpack4<<<blocks_for(n), kBlock, 0, c->stream>>>(src, c->staging, n);
  }));
  for (int k = 0; k < 4; ++k)
    if (dst[k])
      HIPCHK(c, hipMemcpyAsync(dst[k], c->staging + static_cast<size_t>(k) * n, sizeof(double) * n,
                               hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return RG_OK;
}


// State buffers of one Heun step: start (P, X), mid (Pm, Xm), end (Po, Xo),
// and where each stage's slopes go.  The default is the context's state
// with the end overwriting the start.
struct HeunBufs {
  const D4* P;
  const D4* X;
  D4* Pm;
  D4* Xm;
  D4* Po;
  D4* Xo;
  Rec* r1;
  Rec* r2;
};
HeunBufs ctx_bufs(rg_ctx* c) {
  return HeunBufs{c->prim, c->ext, c->mid_prim, c->mid_ext, c->prim, c->ext, c->rec, c->rec};
}

// One step of hdsim::TimeAdvance2Heun (hdsim2d.cpp:323-449) on a static
// mesh: the dt of slot s&1, fluxes, the update to the mid state, fluxes of
// the mid state, 0.5*(mid + old) and the cell update with the next dt's
// CFL terms into slot (s+1)&1.  dt_hist[hist] receives the step's dt.
// Consecutive kernels walk their blocks in opposite directions (slopes,
// faces, update: forward, reverse, forward in the first stage, reverse,
// forward, reverse in the second), so each starts where the previous one
// ended; the results do not depend on the order.
hipError_t heun_step(rg_ctx* c, int s, double* dt_hist, int hist, const HeunBufs& h) {
  const int cur = s & 1, nxt = (s + 1) & 1;
  // the modular calculators write dt*area*F straight from the flux kernel
  const int delta = c->prm.order >= 2 ? UPD_DELTA : 0;
  const unsigned long long* slot = delta ? c->words + cur : nullptr;
  hipError_t e = run_flux(c, h.P, s, 1, h.r1, slot, 0);
  if (e != hipSuccess) return e;
  UpdArgs a{};
  a.in = h.X;
  a.out = h.Xm;
  a.flux = c->flux;
  a.cells = h.Pm;
  a.dt_slot = c->words + cur;
  a.dt_hist = dt_hist;
  a.step = hist;
  a.flags = UPD_FACES | delta | UPD_ZF_BEFORE | UPD_CELLS;
  e = launch(c, KID_UPDATE, [&] {
    k_update<<<blocks_for(c->N), kBlock, 0, c->stream>>>(c->dm, a, dparams(c), derr(c, s, 3));
  });
  if (e != hipSuccess) return e;
  e = run_flux(c, h.Pm, s, 4, h.r2, slot, 1);
  if (e != hipSuccess) return e;
  UpdArgs b{};
  b.in = h.Xm;
  b.out = h.Xo;
  b.avg = h.X;
  b.flux = c->flux;
  b.cells = h.Po;
  b.dt_slot = c->words + cur;
  b.partials = c->partials;
  b.step = hist;
  b.flags = UPD_FACES | delta | UPD_ZF_BEFORE | UPD_AVERAGE | UPD_CELLS | UPD_CFL;
  e = launch(c, KID_UPDATE, [&] {
    k_update<<<blocks_for(c->N), kBlock, 0, c->stream>>>(c->dm, b, dparams_walk(c, 1),
                                                         derr(c, s, 6));
  });
  if (e != hipSuccess) return e;
  return run_reduce_min(c, nxt);
}

// Tile lists of the fused face-and-update kernels: for each tile of kTile
// cells, the distinct faces its cells' slots read (ref(k): face word of
// incidence k -- face | INFO_SECOND, the side whose sign the cell adds), in
// the order of their first use (cells, then slots), and per ELL slot the
// entry it reads.  Fails (false) where a tile would hold more than
// kTileMaxEntries faces.
struct TileLists {
  std::vector<int> tofs, tface;
  std::vector<unsigned short> eloc;
  int max = 0;
};

template <class Ref>
bool tile_lists(const rg_ctx::AdjSource& s, int N, int ne, Ref&& ref, TileLists& out) {
  const int nt = (N + kTile - 1) / kTile;
  std::vector<std::vector<int>> lists(static_cast<size_t>(nt));
  out.eloc.assign(static_cast<size_t>(std::max(ne, 1)), 0);
  std::atomic<bool> bad{false};
  rmp::parallel_for(nt, [&](int64_t tt) {
    const int t = static_cast<int>(tt), i0 = t * kTile, i1 = std::min(N, i0 + kTile);
    const int k0 = s.csr_off[i0], k1 = s.csr_off[i1];
    // (face, first use) pairs, distinct faces, then their entries by first use
    std::vector<std::pair<int, int>> fk;
    fk.reserve(static_cast<size_t>(k1 - k0));
    for (int k = k0; k < k1; ++k) fk.emplace_back(static_cast<int>(ref(k) & INFO_FACE), k);
    std::sort(fk.begin(), fk.end());
    std::vector<std::pair<int, int>> first;  // (first use, face)
    for (size_t u = 0; u < fk.size(); ++u)
      if (u == 0 || fk[u].first != fk[u - 1].first) first.emplace_back(fk[u].second, fk[u].first);
    std::sort(first.begin(), first.end());
    if (first.size() > static_cast<size_t>(kTileMaxEntries)) {
      bad.store(true, std::memory_order_relaxed);
      return;
    }
    std::vector<int>& faces = lists[static_cast<size_t>(t)];
    faces.resize(first.size());
    for (size_t u = 0; u < first.size(); ++u) faces[u] = first[u].second;
    // entry of each face: position in `faces` (binary search on a sorted copy)
    std::vector<std::pair<int, int>> pos(first.size());
    for (size_t u = 0; u < first.size(); ++u) pos[u] = {faces[u], static_cast<int>(u)};
    std::sort(pos.begin(), pos.end());
    for (int i = i0; i < i1; ++i) {
      const size_t base = static_cast<size_t>(s.wofs[i / 64] + s.lanepos[static_cast<size_t>(i)]);
      for (int k = s.csr_off[i]; k < s.csr_off[i + 1]; ++k) {
        const unsigned w = ref(k);
        const int f = static_cast<int>(w & INFO_FACE);
        const auto it = std::lower_bound(pos.begin(), pos.end(), std::make_pair(f, -1));
        out.eloc[base + 64 * static_cast<size_t>(k - s.csr_off[i])] = static_cast<unsigned short>(
            static_cast<unsigned>(it->second) | ((w & INFO_SECOND) ? TILE_SECOND : 0u));
      }
    }
  }, 64);
  if (bad.load()) return false;
  out.tofs.assign(static_cast<size_t>(nt) + 1, 0);
  out.max = 0;
  for (int t = 0; t < nt; ++t) {
    const int n = static_cast<int>(lists[static_cast<size_t>(t)].size());
    out.tofs[static_cast<size_t>(t) + 1] = out.tofs[t] + n;
    out.max = std::max(out.max, n);
  }
  out.tface.resize(static_cast<size_t>(std::max(out.tofs[nt], 1)));
  rmp::parallel_for(nt, [&](int64_t t) {
    std::copy(lists[static_cast<size_t>(t)].begin(), lists[static_cast<size_t>(t)].end(),
              out.tface.begin() + out.tofs[static_cast<size_t>(t)]);
  }, 256);
  return true;
}

template <class T>
hipError_t tile_alloc(rg_ctx* c, T** p, size_t n) {
  void* q = nullptr;
  const hipError_t e = hipMalloc(&q, sizeof(T) * std::max<size_t>(n, 1));
  if (e == hipSuccess) c->tile_bufs.push_back(q);
  *p = static_cast<T*>(q);
  return e;
}

template <class T>
hipError_t tile_upload(rg_ctx* c, T** p, const std::vector<T>& v) {
  hipError_t e = tile_alloc(c, p, v.size());
  if (e == hipSuccess && !v.empty())
    e = hipMemcpyAsync(*p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, c->stream);
  return e;
}

// The tiles of the fused face-and-update kernels and the second state
// buffers of the FP64 loop, built at the first Heun step loop of a mesh
// (second order): per entry the face's id, sides and frame (FP64) or FP32
// face record (mixed, whose slots read a periodic boundary face through its
// lower-numbered copy).  A mesh whose largest tile would hold more than
// kTileMaxEntries faces keeps the three-kernel stage (tiles_ok false), and so
// does one whose tiles hold more than kTileDupNum / kTileDupDen entries per
// face: the fused stage solves a face once per tile that reads it, and
// where cell ids have little spatial locality the repeated faces and the
// scattered reads cost more than the three-kernel stage's flux round trip.
// On a 2048 x 1024 Kelvin-Helmholtz lattice with part of its cells shuffled
// the two forms took the same time at about 1.39 entries per face (FP64)
// and 1.35 (mixed); the Hilbert-ordered lattice has 1.08.
// The fused stages run: allowed (rg_use_tiles) and the mesh's tiles fit.
constexpr int64_t kTileDupNum = 13, kTileDupDen = 10;
bool fused_ok(const rg_ctx* c) { return c->tiles_on && c->tiles_ok; }

hipError_t build_tiles(rg_ctx* c) {
  c->tiles_built = true;
  c->tiles_ok = false;
  const auto& s = c->adj_src;
  const int N = c->N;
  const bool mixed = c->prm.precision == RG_MIXED;
  if (s.inc.size() != static_cast<size_t>(c->nnz) || s.csr_off.size() != static_cast<size_t>(N) + 1 ||
      (mixed && s.minfo.size() != static_cast<size_t>(c->nnz)))
    return hipSuccess;
  TileLists L;
  const bool ok = mixed ? tile_lists(s, N, c->dm_ne, [&](int k) { return s.minfo[k]; }, L)
                        : tile_lists(s, N, c->dm_ne, [&](int k) { return s.inc[k].info; }, L);
  if (!ok) return hipSuccess;
  const int total = L.tofs.back();
  if (kTileDupDen * total > kTileDupNum * static_cast<int64_t>(c->F)) return hipSuccess;
  std::vector<FaceSides> tside(L.tface.size());
  rmp::parallel_for(static_cast<int64_t>(L.tface.size()), [&](int64_t e) {
    const int f = L.tface[static_cast<size_t>(e)];
    tside[static_cast<size_t>(e)] = FaceSides{s.ff[f], s.fs[f]};
  });
  int *tofs_d, *tface_d;
  FaceSides* tside_d;
  unsigned short* eloc_d;
  FaceFrame* tframe_d = nullptr;
  FaceM* tfm_d = nullptr;
  hipError_t e = tile_upload(c, &tofs_d, L.tofs);
  if (e == hipSuccess) e = tile_upload(c, &tface_d, L.tface);
  if (e == hipSuccess) e = tile_upload(c, &tside_d, tside);
  if (e == hipSuccess) e = tile_upload(c, &eloc_d, L.eloc);
  if (e == hipSuccess) {
    if (mixed)
      e = tile_alloc(c, &tfm_d, L.tface.size());
    else
      e = tile_alloc(c, &tframe_d, L.tface.size());
  }
  if (e == hipSuccess && !mixed) e = tile_alloc(c, &c->alt_prim, static_cast<size_t>(N));
  if (e == hipSuccess && !mixed) e = tile_alloc(c, &c->alt_ext, static_cast<size_t>(N));
  if (e != hipSuccess) return e;
  e = launch(c, KID_GEO, [&] {
    k_tile_frames<<<blocks_for(std::max(total, 1)), kBlock, 0, c->stream>>>(
        c->dm, c->mt, tface_d, tframe_d, tfm_d, total);
  });
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);  // the host vectors are uploaded
  if (e != hipSuccess) return e;
  c->tiles = DTiles{tofs_d, tface_d, tside_d, tframe_d, tfm_d, eloc_d};
  c->tiles_max = L.max;
  c->tiles_total = total;
  c->tiles_ok = true;
  return hipSuccess;
}

// One stage of the fused Heun step: the slope kernel, then k_face_update
// walking its tiles in reverse (so that it starts on the slopes written
// last).  dt from slot `cur` as in heun_step.
hipError_t fused_stage(rg_ctx* c, const D4* cells, Rec* r, const UpdArgs& a, int s,
                       int phase_base) {
  hipError_t e = run_slopes(c, cells, s, phase_base, r, 0);
  if (e != hipSuccess) return e;
  const size_t lds = sizeof(D4) * static_cast<size_t>(c->tiles_max);
  return launch(c, KID_FACE_UPDATE, [&] {
    k_face_update<<<tiles_for(c->N), kTile, lds, c->stream>>>(
        c->dm, c->tiles, cells, r, a, dparams_walk(c, 1), derr(c, s, phase_base + 1),
        derr(c, s, phase_base + 2));
  });
}

// heun_step with each stage's faces and update fused per tile
// (k_face_update).  The end state must not be the start state: the second
// stage's faces read the mid state only, but a failed step has to leave the
// start state as it was.
hipError_t heun_step_fused(rg_ctx* c, int s, double* dt_hist, int hist, const HeunBufs& h) {
  const int cur = s & 1, nxt = (s + 1) & 1;
  UpdArgs a{};
  a.in = h.X;
  a.out = h.Xm;
  a.cells = h.Pm;
  a.dt_slot = c->words + cur;
  a.dt_hist = dt_hist;
  a.step = hist;
  a.flags = UPD_FACES | UPD_DELTA | UPD_ZF_BEFORE | UPD_CELLS;
  hipError_t e = fused_stage(c, h.P, h.r1, a, s, 1);
  if (e != hipSuccess) return e;
  UpdArgs b{};
  b.in = h.Xm;
  b.out = h.Xo;
  b.avg = h.X;
  b.cells = h.Po;
  b.dt_slot = c->words + cur;
  b.partials = c->partials;
  b.step = hist;
  b.flags = UPD_FACES | UPD_DELTA | UPD_ZF_BEFORE | UPD_AVERAGE | UPD_CELLS | UPD_CFL;
  e = fused_stage(c, h.Pm, h.r2, b, s, 4);
  if (e != hipSuccess) return e;
  return run_reduce_min(c, nxt, true);
}

Ens ens_args(rg_ctx* c) {
  return Ens{c->ens_K,   c->ens_inst, c->ens_coff,     c->ens_dtA,  c->ens_dtB,
             c->ens_inv, c->ens_hist, c->ens_step,     c->ens_hist_cap};
}

// The first dt of every instance from the current cells.
// Grid of a launch over n items: bpc resident blocks per CU walking the
// items (persistent), or with bpc = 0 one thread per item.
unsigned ens_grid(const rg_ctx* c, int n, int block, int bpc) {
  return bpc > 0 ? static_cast<unsigned>(c->ens_cu * bpc)
                 : static_cast<unsigned>((n + block - 1) / block);
}

hipError_t ens_seed(rg_ctx* c) {
  const Ens e = ens_args(c);
  const unsigned grid = ens_grid(c, c->N, c->ens_block, c->ens_bpc);
  return launch(c, KID_ENS, [&] {
    k_inv_ens_p<<<grid, c->ens_block, 0, c->stream>>>(c->dm, c->prim, e, c->prm.gamma);
    k_ens_min<<<static_cast<unsigned>(std::min(e.K, 4096)), 256, 0, c->stream>>>(e);
    k_ens_finish<<<1, 256, 0, c->stream>>>(e, c->prm.cfl, 0);
  });
}

// The eight persistent launches of one Heun step of every instance
// (hdsim::TimeAdvance2Heun, hdsim2d.cpp:323-449), each instance with its
// own dt.  Everything that changes between steps is in device memory.
hipError_t ens_record_step(rg_ctx* c, int block, int bpc) {
  const Ens e = ens_args(c);
  const DParams prm = dparams(c);
  const unsigned grid = ens_grid(c, c->N, block, bpc), gridF = ens_grid(c, c->F, block, bpc);
  UpdArgs a{};
  a.in = c->ext;
  a.out = c->mid_ext;
  a.flux = c->flux;
  a.cells = c->mid_prim;
  a.flags = UPD_FACES | UPD_ZF_BEFORE | UPD_CELLS;
  UpdArgs b{};
  b.in = c->mid_ext;
  b.out = c->ext;
  b.avg = c->ext;
  b.flux = c->flux;
  b.cells = c->prim;
  b.flags = UPD_FACES | UPD_ZF_BEFORE | UPD_AVERAGE | UPD_CELLS | UPD_CFL;
  const hipStream_t st = c->stream;
  if (fused_ok(c)) {
    // each stage's faces and update fused per tile; the slope kernel keeps
    // the tuned launch shape, the tile kernel walks its tiles in reverse
    DParams pr = prm;
    pr.rev = 1;
    const size_t lds = sizeof(D4) * static_cast<size_t>(c->tiles_max);
    const unsigned nt = tiles_for(c->N);
    k_slopes_p<<<grid, block, 0, st>>>(c->dm, c->prim, c->rec, prm, derr(c, 0, 1));
    k_face_update_ens<<<nt, kTile, lds, st>>>(c->dm, c->tiles, c->prim, c->rec, a, e, pr,
                                              derr(c, 0, 2), derr(c, 0, 3));
    k_slopes_p<<<grid, block, 0, st>>>(c->dm, c->mid_prim, c->rec, prm, derr(c, 0, 4));
    k_face_update_ens<<<nt, kTile, lds, st>>>(c->dm, c->tiles, c->mid_prim, c->rec, b, e, pr,
                                              derr(c, 0, 5), derr(c, 0, 6));
  } else {
    k_slopes_p<<<grid, block, 0, st>>>(c->dm, c->prim, c->rec, prm, derr(c, 0, 1));
    k_flux2_p<<<gridF, block, 0, st>>>(c->dm, c->prim, c->rec, c->flux, prm, derr(c, 0, 2));
    k_update_ens_p<<<grid, block, 0, st>>>(c->dm, a, e, prm, derr(c, 0, 3));
    k_slopes_p<<<grid, block, 0, st>>>(c->dm, c->mid_prim, c->rec, prm, derr(c, 0, 4));
    k_flux2_p<<<gridF, block, 0, st>>>(c->dm, c->mid_prim, c->rec, c->flux, prm, derr(c, 0, 5));
    k_update_ens_p<<<grid, block, 0, st>>>(c->dm, b, e, prm, derr(c, 0, 6));
  }
  k_ens_min<<<static_cast<unsigned>(std::min(e.K, 4096)), 256, 0, st>>>(e);
  k_ens_finish<<<1, 256, 0, st>>>(e, c->prm.cfl, 1);
  return hipGetLastError();
}

// Capture one step into c->ens_graph with the given launch shape.
hipError_t ens_capture(rg_ctx* c, int block, int bpc) {
  if (c->ens_graph) (void)hipGraphExecDestroy(c->ens_graph);
  c->ens_graph = nullptr;
  // the tiles of the fused stages (allocated here, outside the capture)
  if (c->tiles_on && !c->tiles_built) {
    const hipError_t eb = build_tiles(c);
    if (eb != hipSuccess) return eb;
  }
  hipError_t e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) return e;
  const hipError_t e1 = ens_record_step(c, block, bpc);
  hipGraph_t g = nullptr;
  e = hipStreamEndCapture(c->stream, &g);
  if (e1 != hipSuccess) e = e1;
  if (e == hipSuccess) e = hipGraphInstantiate(&c->ens_graph, g, nullptr, nullptr, 0);
  if (g) (void)hipGraphDestroy(g);
  if (e == hipSuccess) {
    c->ens_block = block;
    c->ens_bpc = bpc;
  }
  return e;
}

// Adjoint tables of the current mesh.  rev: for each incidence (i, s) with a
// real neighbour j, the ELL entry of the same face in j's list.  gf: the
// faces whose ghost side copies cell b (b = the ghost's source), with the
// ghost's side.  ge: the incidences whose neighbour is a ghost copying b.
hipError_t build_adjoint_tables(rg_ctx* c) {
  const rg_ctx::AdjSource& a = c->adj_src;
  const int N = c->N, F = c->F;
  const size_t ne = static_cast<size_t>(c->dm_ne);
  auto ell_at = [&](int i, int k) {
    return static_cast<size_t>(a.wofs[i / 64] + a.lanepos[static_cast<size_t>(i)]) +
           64 * static_cast<size_t>(k - a.csr_off[i]);
  };
  // ELL entry of each face on its first [2f] and second [2f + 1] side
  std::vector<int> epos(2 * static_cast<size_t>(F), -1);
  rmp::parallel_for(N, [&](int64_t ii) {  // each (face, side) has one cell
    const int i = static_cast<int>(ii);
    for (int k = a.csr_off[i]; k < a.csr_off[i + 1]; ++k) {
      const unsigned w = a.inc[k].info;
      epos[2 * static_cast<size_t>(w & INFO_FACE) + ((w & INFO_SECOND) ? 1 : 0)] =
          static_cast<int>(ell_at(i, k));
    }
  });
  std::vector<int> rev(ne, -1);
  std::vector<int> gf_off(static_cast<size_t>(N) + 1, 0), ge_off(static_cast<size_t>(N) + 1, 0);
  for (int f = 0; f < F; ++f) {
    if (a.ff[f] >= N) ++gf_off[a.gsrc[a.ff[f] - N] + 1];
    if (a.fs[f] >= N) ++gf_off[a.gsrc[a.fs[f] - N] + 1];
  }
  for (int i = 0; i < N; ++i)
    for (int k = a.csr_off[i]; k < a.csr_off[i + 1]; ++k) {
      const int j = a.inc[k].nbr;
      if (j >= N) {
        ++ge_off[a.gsrc[j - N] + 1];
        continue;
      }
      const unsigned w = a.inc[k].info;
      rev[epos[2 * static_cast<size_t>(w & INFO_FACE) + ((w & INFO_SECOND) ? 0 : 1)]] =
          static_cast<int>(ell_at(i, k));
    }
  for (int i = 0; i < N; ++i) {
    gf_off[i + 1] += gf_off[i];
    ge_off[i + 1] += ge_off[i];
  }
  std::vector<unsigned> gf(static_cast<size_t>(std::max(gf_off[N], 1)), 0u);
  std::vector<int> ge(static_cast<size_t>(std::max(ge_off[N], 1)), 0);
  {
    std::vector<int> fill(gf_off.begin(), gf_off.end() - 1);
    for (int f = 0; f < F; ++f) {
      if (a.ff[f] >= N) gf[fill[a.gsrc[a.ff[f] - N]]++] = static_cast<unsigned>(f);
      if (a.fs[f] >= N) gf[fill[a.gsrc[a.fs[f] - N]]++] = static_cast<unsigned>(f) | INFO_SECOND;
    }
    std::vector<int> gfill(ge_off.begin(), ge_off.end() - 1);
    for (int i = 0; i < N; ++i)
      for (int k = a.csr_off[i]; k < a.csr_off[i + 1]; ++k)
        if (a.inc[k].nbr >= N) ge[gfill[a.gsrc[a.inc[k].nbr - N]]++] = static_cast<int>(ell_at(i, k));
  }
  int *rev_d, *gf_off_d, *ge_off_d, *ge_d;
  unsigned* gf_d;
  hipError_t e = upload(c, &rev_d, rev.data(), ne);
  if (e == hipSuccess) e = upload(c, &gf_off_d, gf_off.data(), gf_off.size());
  if (e == hipSuccess) e = upload(c, &gf_d, gf.data(), gf.size());
  if (e == hipSuccess) e = upload(c, &ge_off_d, ge_off.data(), ge_off.size());
  if (e == hipSuccess) e = upload(c, &ge_d, ge.data(), ge.size());
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);  // host vectors above
  if (e == hipSuccess) c->am = AdjMesh{rev_d, gf_off_d, gf_d, ge_off_d, ge_d};
  return e;
}

// hdsim::TimeAdvance2Heun with FP32 reconstruction and HLLC and FP64
// extensives (RG_MIXED).  The FP64 cells are recovered at the end.
int advance_mixed(rg_ctx* c, int nsteps, double* dt_out) {
  const int N = c->N, F = c->F;
  // each stage's faces and update fused per tile where the mesh has tiles
  if (c->tiles_on && !c->tiles_built) HIPCHK(c, build_tiles(c));
  const bool fused = fused_ok(c);
  const size_t lds = sizeof(F4) * static_cast<size_t>(c->tiles_max);
  HIPCHK(c, launch(c, KID_PACK, [&] {
    k_to_f32<<<blocks_for(N), kBlock, 0, c->stream>>>(c->prim, c->primM, N);
  }));
  HIPCHK(c, run_cfl(c, c->prim, 0));
  for (int s = 0; s < nsteps; ++s) {
    const int cur = s & 1, nxt = (s + 1) & 1;
    for (int stage = 0; stage < 2; ++stage) {
      // consecutive kernels walk their blocks in opposite directions, as in
      // the FP64 step (heun_step)
      const F4* q = stage == 0 ? c->primM : c->mid_primM;
      HIPCHK(c, launch(c, KID_SLOPES_MIXED, [&] {
        (void)launch_slopes_mixed(blocks_for(N), kBlock, c->stream, c->dm, c->mt, q, c->recM,
                                  dparams_walk(c, fused ? 0 : stage), derr(c, s, 1 + 3 * stage));
      }));
      if (!fused)
        HIPCHK(c, launch(c, KID_FLUX_MIXED, [&] {
          k_flux_mixed<<<blocks_for(F), kBlock, 0, c->stream>>>(
              c->dm, c->mt, q, c->recM, c->fluxM, dparams_walk(c, 1 - stage),
              derr(c, s, 2 + 3 * stage));
        }));
      UpdArgsM a{};
      a.flux = c->fluxM;
      a.dt_slot = c->words + cur;
      a.step = s;
      if (stage == 0) {
        a.in = c->ext;
        a.out = c->mid_ext;
        a.cells = c->mid_primM;
        a.reset_slot = c->words + nxt;
        a.dt_hist = dt_out ? c->dt_hist : nullptr;
        a.flags = 0;
      } else {
        a.in = c->mid_ext;
        a.out = c->ext;
        a.avg = c->ext;
        a.cells = c->primM;
        a.partials = c->partials;
        a.flags = UPD_AVERAGE | UPD_CFL;
      }
      if (fused)
        HIPCHK(c, launch(c, KID_FACE_UPDATE_MIXED, [&] {
          (void)launch_face_update_mixed(tiles_for(N), kTile, lds, c->stream, c->dm, c->mt,
                                         c->tiles, q, c->recM, a, dparams_walk(c, 1),
                                         derr(c, s, 2 + 3 * stage), derr(c, s, 3 + 3 * stage));
        }));
      else
        HIPCHK(c, launch(c, KID_UPDATE_MIXED, [&] {
          k_update_mixed<<<blocks_for(N), kBlock, 0, c->stream>>>(c->dm, c->mt, a,
                                                                  dparams_walk(c, stage),
                                                                  derr(c, s, 3 + 3 * stage));
        }));
    }
    HIPCHK(c, run_reduce_min(c, nxt, fused));
  }
  // FP64 cells from the FP64 extensives
  UpdArgs u{};
  u.in = c->ext;
  u.out = c->ext;
  u.cells = c->prim;
  u.flags = UPD_CELLS;
  HIPCHK(c, run_update(c, u, nsteps, 6));
  if (dt_out)
    HIPCHK(c, hipMemcpyAsync(dt_out, c->dt_hist, sizeof(double) * nsteps, hipMemcpyDeviceToHost,
                             c->stream));
  return collect_error(c);
}

// The FP64 step loop of rg_advance on a static mesh: the dt slots, block
// minima and dt history are the context's words, partials and dt_hist
// (record: write each step's dt to dt_hist).  Second-order Heun steps run
// fused (heun_step_fused) where the mesh has tiles -- built here at the
// first call of a mesh when may_build -- and the state then alternates
// between the context's two state buffers: step s reads one pair and writes
// the other, and c->prim / c->ext name the newest.  *fused tells which.
hipError_t advance_fp64(rg_ctx* c, int integrator, int nsteps, bool record, bool may_build,
                        bool* fused) {
  const bool can_fuse = integrator == 2 && c->prm.order == 2;
  if (can_fuse && may_build && c->tiles_on && !c->tiles_built) {
    const hipError_t eb = build_tiles(c);
    if (eb != hipSuccess) return eb;
  }
  *fused = can_fuse && fused_ok(c);
  // first dt: SimpleCFL on the current cells into slot 0
  hipError_t e = run_cfl(c, c->prim, 0);
  for (int s = 0; s < nsteps && e == hipSuccess; ++s) {
    const int cur = s & 1, nxt = (s + 1) & 1;
    if (integrator == 1) {
      // hdsim::TimeAdvance: fluxes, extensive update, ZeroForce, cell update
      e = run_flux(c, c->prim, s, 1);
      if (e != hipSuccess) break;
      UpdArgs a{};
      a.in = c->ext;
      a.out = c->ext;
      a.flux = c->flux;
      a.cells = c->prim;
      a.dt_slot = c->words + cur;
      a.partials = c->partials;
      a.dt_hist = record ? c->dt_hist : nullptr;
      a.step = s;
      a.flags = UPD_FACES | UPD_ZF_AFTER | UPD_CELLS | UPD_CFL;
      e = run_update(c, a, s, 3);
      if (e == hipSuccess) e = run_reduce_min(c, nxt);
    } else if (*fused) {
      const HeunBufs h{c->prim, c->ext, c->mid_prim, c->mid_ext,
                       c->alt_prim, c->alt_ext, c->rec, c->rec};
      e = heun_step_fused(c, s, record ? c->dt_hist : nullptr, s, h);
      if (e == hipSuccess) {
        std::swap(c->prim, c->alt_prim);
        std::swap(c->ext, c->alt_ext);
      }
    } else {
      e = heun_step(c, s, record ? c->dt_hist : nullptr, s, ctx_bufs(c));
    }
  }
  return e;
}

// The buffers the FP64 step loop's stages read and write (ErrSrc).
ErrSrc fp64_err_src(const rg_ctx* c, int integrator) {
  if (integrator == 1) return ErrSrc{{c->prim, nullptr}, {c->ext, nullptr}};
  const HeunBufs h = ctx_bufs(const_cast<rg_ctx*>(c));
  return ErrSrc{{h.P, h.Pm}, {h.Xm, h.Xo}};
}

// collect_error after advance_fp64.  A fused loop that failed at step s
// goes back to the buffers step s started from (the steps after a failure
// write no state), and the error's entries come from step s's buffers.
int collect_fp64(rg_ctx* c, int integrator, int nsteps, bool fused) {
  if (!fused) {
    const ErrSrc src = fp64_err_src(c, integrator);
    return collect_error(c, &src);
  }
  HIPCHK(c, hipMemcpyAsync(c->host_words, c->words, 3 * sizeof(unsigned long long),
                           hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const unsigned long long w = c->host_words[2];
  if (w != ~0ull) {
    const int step = static_cast<int>(w >> 39);
    if (step >= 0 && step < nsteps && ((nsteps - step) & 1)) {
      std::swap(c->prim, c->alt_prim);
      std::swap(c->ext, c->alt_ext);
    }
  }
  const ErrSrc src{{c->prim, c->mid_prim}, {c->mid_ext, c->alt_ext}};
  return collect_error(c, &src);
}

// For the duration of a rg_advance_noalloc call: the context's dt slots,
// error word, block minima and dt history are the device's scratch, and no
// per-launch timing events are recorded.
struct ScratchScope {
  rg_ctx* c;
  unsigned long long* words;
  double* partials;
  double* dt_hist;
  bool prof;
  explicit ScratchScope(rg_ctx* ctx)
      : c(ctx), words(ctx->words), partials(ctx->partials), dt_hist(ctx->dt_hist),
        prof(ctx->prof) {
    c->words = reinterpret_cast<unsigned long long*>(c->scratch);
    c->partials = c->scratch + kScratchWords;
    c->dt_hist = c->scratch + kScratchWords + kScratchBlocks;
    c->prof = false;
  }
  ~ScratchScope() {
    c->words = words;
    c->partials = partials;
    c->dt_hist = dt_hist;
    c->prof = prof;
  }
};

int arg_error(rg_ctx* c, const char* msg) {
  c->last = rg_error{};
  c->last.code = -1;
  std::snprintf(c->last.message, sizeof(c->last.message), "%s", msg);
  return RG_E_ARG;
}

}  // namespace

extern "C" {

int rg_create(const rg_params* params, rg_ctx** out) {
  if (!params || !out) return RG_E_ARG;
  if (!(params->gamma > 1.0) || !(params->cfl > 0.0) || (params->order < 1 || params->order > 3) ||
      (params->bc != RG_BC_RIGID && params->bc != RG_BC_PERIODIC) ||
      (params->precision != RG_FP64 && params->precision != RG_MIXED))
    return RG_E_ARG;
  rg_ctx* c = new (std::nothrow) rg_ctx();
  if (!c) return RG_E_ARG;
  c->prm = *params;
  hipError_t e = hipStreamCreate(&c->stream);
  if (e == hipSuccess) c->own_stream = true;
  if (e == hipSuccess)
    e = hipHostMalloc(reinterpret_cast<void**>(&c->host_words), 4 * sizeof(unsigned long long),
                      hipHostMallocDefault);
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&c->words), 4 * sizeof(unsigned long long));
  if (e == hipSuccess) e = fill_words(c, 0, 4, ~0ull);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  // rg_advance_noalloc's scratch on this device (a failure here only
  // disables that entry point)
  if (e == hipSuccess && scratch_on_current_device(&c->device, &c->scratch) != hipSuccess)
    c->scratch = nullptr;
  if (e != hipSuccess) {
    rg_destroy(c);
    return RG_E_HIP;
  }
  *out = c;
  return RG_OK;
}

void rg_destroy(rg_ctx* c) {
  if (!c) return;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  free_mesh(c);
  drain_pool(c);
  if (c->pv) (void)hipFree(c->pv);
  if (c->keep) (void)hipFree(c->keep);
  if (c->riemann_buf) (void)hipFree(c->riemann_buf);
  if (c->dt_hist) (void)hipFree(c->dt_hist);
  if (c->words) (void)hipFree(c->words);
  if (c->host_words) (void)hipHostFree(c->host_words);
  for (const TimedLaunch& t : c->pending) {
    (void)hipEventDestroy(t.a);
    (void)hipEventDestroy(t.b);
  }
  for (hipEvent_t ev : c->event_pool) (void)hipEventDestroy(ev);
  for (int b = 0; b < 2; ++b) {
    if (c->tape_ready[b]) (void)hipEventDestroy(c->tape_ready[b]);
    if (c->tape_free[b]) (void)hipEventDestroy(c->tape_free[b]);
  }
  if (c->rstream) {
    (void)hipStreamSynchronize(c->rstream);
    (void)hipStreamDestroy(c->rstream);
  }
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
      !m->pt_x || !m->pt_y || !m->csr_off || !m->csr_face || (G > 0 && !m->ghost_orig))
    return RG_E_ARG;
  // the cell geometry comes with the mesh, or (all four NULL) is computed on
  // the device; a rigid box's ghosts then need the box
  const int n_geo = !!m->cm_x + !!m->cm_y + !!m->volume + !!m->tess_volume;
  if (n_geo != 0 && n_geo != 4) return RG_E_ARG;
  const bool dev_geo = n_geo == 0;
  if (dev_geo && G > 0 && c->prm.bc == RG_BC_RIGID && !c->has_box) return RG_E_ARG;
  // --- validate topology on the host threads
  if (m->csr_off[0] != 0) return RG_E_ARG;
  for (int i = 0; i < N; ++i)
    if (m->csr_off[i + 1] < m->csr_off[i]) return RG_E_ARG;
  const int nnz = m->csr_off[N];
  int64_t faces_rr = 0;  // faces between two real cells
  {
    const int nc = rmp::chunks(F);
    std::vector<char> bad(static_cast<size_t>(nc), 0);
    std::vector<int64_t> rr(static_cast<size_t>(nc), 0);
    rmp::for_chunks(F, nc, [&](int ch, int64_t f0, int64_t f1) {
      int64_t n = 0;  // a local count: per-chunk counters share cache lines
      for (int64_t f = f0; f < f1; ++f) {
        const int a = m->face_first[f], b = m->face_second[f];
        if (a < 0 || b < 0 || a >= N + G || b >= N + G || a == b || (a >= N && b >= N)) {
          bad[ch] = 1;
          return;
        }
        if (a < N && b < N) ++n;
      }
      rr[ch] = n;
    });
    for (int ch = 0; ch < nc; ++ch) {
      if (bad[ch]) return RG_E_ARG;
      faces_rr += rr[ch];
    }
  }
  for (int g = 0; g < G; ++g)
    if (m->ghost_orig[g] < 0 || m->ghost_orig[g] >= N) return RG_E_ARG;
  auto orig = [&](int id) { return id < N ? id : m->ghost_orig[id - N]; };
  std::vector<Inc> inc(static_cast<size_t>(nnz));
  {
    // each cell's incidences on host threads; any bad entry fails the mesh
    const int nc = rmp::chunks(N);
    std::vector<char> bad(static_cast<size_t>(nc), 0);
    std::vector<int64_t> inc_rr(static_cast<size_t>(nc), 0);
    rmp::for_chunks(N, nc, [&](int ch, int64_t i0, int64_t i1) {
      int64_t n_rr = 0;
      for (int i = static_cast<int>(i0); i < static_cast<int>(i1); ++i) {
        int prev = -1;
        for (int k = m->csr_off[i]; k < m->csr_off[i + 1]; ++k) {
          const int f = m->csr_face[k];
          if (f < 0 || f >= F || f <= prev) {  // ascending face order
            bad[ch] = 1;
            return;
          }
          prev = f;
          const int a = m->face_first[f], b = m->face_second[f];
          if (a != i && b != i) {
            bad[ch] = 1;
            return;
          }
          unsigned w = static_cast<unsigned>(f);
          if (b == i) w |= INFO_SECOND;
          if (orig(a) == orig(b)) w |= INFO_MIRROR;
          inc[k] = Inc{(a == i) ? b : a, w};
          if (a < N && b < N) ++n_rr;
        }
      }
      inc_rr[ch] = n_rr;
    });
    int64_t listed_rr = 0;
    for (int ch = 0; ch < nc; ++ch) {
      if (bad[ch]) return RG_E_ARG;
      listed_rr += inc_rr[ch];
    }
    // Every face between two real cells must be in both cells' lists.  A
    // cell lists a face at most once (ascending order) and only as one of
    // its sides (checked above), so that holds exactly when such faces are
    // listed twice as often as there are of them.
    if (listed_rr != 2 * faces_rr) return RG_E_ARG;
  }
  // Ghost sources: RigidWallGenerator copies the real cell of the ghost's
  // first boundary face (its flat_map insert keeps the first) and reverses
  // the velocity along the ghost-minus-real direction
  // (RigidWallGenerator.cpp:5-37); PeriodicGhostGenerator copies the
  // original (PeriodicGhostGenerator.cpp:3-17).
  std::vector<int> gsrc(static_cast<size_t>(std::max(G, 1)), 0);
  std::vector<D2> gn(static_cast<size_t>(std::max(G, 1)), D2{0.0, 0.0});
  std::vector<char> seen(static_cast<size_t>(std::max(G, 1)), 0);
  for (int g = 0; g < G; ++g) gsrc[g] = m->ghost_orig[g];
  if (c->prm.bc == RG_BC_RIGID) {
    for (int f = 0; f < F; ++f) {
      const int a = m->face_first[f], b = m->face_second[f];
      const int side = (a >= N) ? 1 : ((b >= N) ? 2 : 0);
      if (!side) continue;
      const int g = (side == 1 ? a : b) - N;
      if (seen[g]) continue;
      seen[g] = 1;
      gsrc[g] = side == 1 ? b : a;
      double nx, ny;
      if (side == 1) {
        nx = m->pt_x[a] - m->pt_x[b];
        ny = m->pt_y[a] - m->pt_y[b];
      } else {
        nx = m->pt_x[b] - m->pt_x[a];
        ny = m->pt_y[b] - m->pt_y[a];
      }
      const double an = std::sqrt(nx * nx + ny * ny);
      gn[g] = D2{nx / an, ny / an};
    }
  }
  std::vector<D2> cm(dev_geo ? 0 : static_cast<size_t>(N) + G);
  if (!dev_geo)
    rmp::parallel_for(static_cast<int64_t>(N) + G,
                      [&](int64_t i) { cm[i] = D2{m->cm_x[i], m->cm_y[i]}; });
  // --- upload
  HIPCHK(c, hipStreamSynchronize(c->stream));
  drain_pool(c);
  free_mesh(c, true);
  const size_t NG = static_cast<size_t>(N) + G;
  DMesh d{};
  d.N = N;
  d.F = F;
  d.G = G;
  d.bc = c->prm.bc;
  int *ff, *fs, *off, *gorig, *gsrc_d;
  double *v1x, *v1y, *v2x, *v2y, *px, *py, *tvol, *vol, *area, *width;
  Inc* inc_d;
  D2 *cm_d, *gn_d;
  CellGeo* cgeo;
  FaceRec* frec;
  FaceFrame* fframe;
  const int zero = 0;
  HIPCHK(c, upload(c, &ff, m->face_first, F));
  HIPCHK(c, upload(c, &fs, m->face_second, F));
  HIPCHK(c, upload(c, &v1x, m->v1x, F));
  HIPCHK(c, upload(c, &v1y, m->v1y, F));
  HIPCHK(c, upload(c, &v2x, m->v2x, F));
  HIPCHK(c, upload(c, &v2y, m->v2y, F));
  HIPCHK(c, upload(c, &px, m->pt_x, NG));
  HIPCHK(c, upload(c, &py, m->pt_y, NG));
  if (dev_geo) {
    HIPCHK(c, dalloc(c, &tvol, N));
    HIPCHK(c, dalloc(c, &vol, N));
    HIPCHK(c, dalloc(c, &cm_d, NG));
  } else {
    HIPCHK(c, upload(c, &tvol, m->tess_volume, N));
    HIPCHK(c, upload(c, &vol, m->volume, N));
    HIPCHK(c, upload(c, &cm_d, cm.data(), NG));
  }
  HIPCHK(c, upload(c, &off, m->csr_off, static_cast<size_t>(N) + 1));
  HIPCHK(c, upload(c, &inc_d, inc.data(), static_cast<size_t>(nnz)));
  HIPCHK(c, upload(c, &gorig, G ? m->ghost_orig : &zero, G ? G : 1));
  HIPCHK(c, upload(c, &gsrc_d, gsrc.data(), gsrc.size()));
  HIPCHK(c, upload(c, &gn_d, gn.data(), gn.size()));
  // wavefront-sliced ELL of the incidences (slot s of a cell = its s-th
  // face in ascending order); within each wavefront the columns hold the
  // cells by descending face count (ties in cell order)
  const int nw = (N + 63) / 64;
  std::vector<int> wofs(static_cast<size_t>(nw) + 1, 0);
  std::vector<unsigned char> lanepos(static_cast<size_t>(std::max(N, 1)), 0);
  std::vector<unsigned char> lanecell(64 * static_cast<size_t>(std::max(nw, 1)), 0);
  rmp::parallel_for(nw, [&](int64_t ww) {  // a counting sort per wavefront
    const int w = static_cast<int>(ww), i0 = 64 * w, n = std::min(N, i0 + 64) - i0;
    int dg[64], wd = 0;
    for (int l = 0; l < n; ++l) {
      dg[l] = m->csr_off[i0 + l + 1] - m->csr_off[i0 + l];
      wd = std::max(wd, dg[l]);
    }
    int small[65] = {0};
    std::vector<int> large;  // face counts above 63 (not met in Voronoi meshes of interest)
    int* start = small;
    if (wd > 63) {
      large.assign(static_cast<size_t>(wd) + 1, 0);
      start = large.data();
    }
    for (int l = 0; l < n; ++l) ++start[wd - dg[l]];
    for (int k = 0, acc = 0; k <= wd; ++k) {
      const int cnt = start[k];
      start[k] = acc;
      acc += cnt;
    }
    unsigned char* cells = lanecell.data() + 64 * static_cast<size_t>(w);
    for (int l = 0; l < n; ++l) {
      const int col = start[wd - dg[l]]++;
      lanepos[static_cast<size_t>(i0 + l)] = static_cast<unsigned char>(col);
      cells[col] = static_cast<unsigned char>(l);
    }
    for (int l = n; l < 64; ++l) cells[l] = static_cast<unsigned char>(l);
    wofs[static_cast<size_t>(w) + 1] = 64 * wd;
  }, 256);
  for (int w = 0; w < nw; ++w) wofs[w + 1] += wofs[w];
  const size_t ne = static_cast<size_t>(std::max(wofs[nw], 1));
  c->dm_ne = static_cast<int>(ne);
  auto ell_at = [&](int i, int k) {
    return static_cast<size_t>(wofs[i / 64] + lanepos[static_cast<size_t>(i)]) +
           64 * static_cast<size_t>(k - m->csr_off[i]);
  };
  // the ELL arrays are filled on the device from the incidences (k_geo_ell);
  // padding entries stay zero
  int *wofs_d, *enbr_d;
  unsigned char *lanepos_d, *lanecell_d;
  unsigned* einfo_d;
  D2 *eh, *eo;
  double* earea;
  HIPCHK(c, upload(c, &wofs_d, wofs.data(), wofs.size()));
  HIPCHK(c, upload(c, &lanepos_d, lanepos.data(), lanepos.size()));
  HIPCHK(c, upload(c, &lanecell_d, lanecell.data(), lanecell.size()));
  HIPCHK(c, dalloc(c, &enbr_d, ne));
  HIPCHK(c, dalloc(c, &einfo_d, ne));
  HIPCHK(c, hipMemsetAsync(enbr_d, 0, sizeof(int) * ne, c->stream));
  HIPCHK(c, hipMemsetAsync(einfo_d, 0, sizeof(unsigned) * ne, c->stream));
  HIPCHK(c, dalloc(c, &eh, ne));
  HIPCHK(c, dalloc(c, &eo, ne));
  HIPCHK(c, dalloc(c, &earea, ne));
  HIPCHK(c, hipMemsetAsync(eh, 0, sizeof(D2) * ne, c->stream));
  HIPCHK(c, hipMemsetAsync(eo, 0, sizeof(D2) * ne, c->stream));
  HIPCHK(c, hipMemsetAsync(earea, 0, sizeof(double) * ne, c->stream));
  HIPCHK(c, dalloc(c, &cgeo, N));
  HIPCHK(c, dalloc(c, &frec, F));
  HIPCHK(c, dalloc(c, &fframe, F));
  HIPCHK(c, dalloc(c, &area, F));
  HIPCHK(c, dalloc(c, &width, N));
  d.ff = ff;
  d.fs = fs;
  d.v1x = v1x;
  d.v1y = v1y;
  d.v2x = v2x;
  d.v2y = v2y;
  d.px = px;
  d.py = py;
  d.cm = cm_d;
  d.tvol = tvol;
  d.vol = vol;
  d.off = off;
  d.inc = inc_d;
  d.gorig = gorig;
  d.gsrc = gsrc_d;
  d.gn = gn_d;
  d.wx = d.wy = nullptr;
  d.cgeo = cgeo;
  d.frec = frec;
  d.fframe = fframe;
  d.area = area;
  d.width = width;
  d.wofs = wofs_d;
  d.lanepos = lanepos_d;
  d.lanecell = lanecell_d;
  d.enbr = enbr_d;
  d.einfo = einfo_d;
  d.eh = eh;
  d.eo = eo;
  d.earea = earea;
  if (dev_geo) {
    // volumes and centres of mass from the vertices, then the ghosts' CMs
    HIPCHK(c, launch(c, KID_GEO, [&] {
      k_cell_geometry<<<blocks_for(N), kBlock, 0, c->stream>>>(d, tvol, vol, cm_d);
    }));
    if (G > 0)
      HIPCHK(c, launch(c, KID_GEO, [&] {
        k_ghost_cm<<<blocks_for(G), kBlock, 0, c->stream>>>(d, cm_d, c->box[0], c->box[1],
                                                           c->box[2], c->box[3]);
      }));
  }
  // static geometry tables
  HIPCHK(c, launch(c, KID_GEO, [&] {
    k_geo_faces<<<blocks_for(F), kBlock, 0, c->stream>>>(d, frec, fframe, area);
  }));
  HIPCHK(c, launch(c, KID_GEO, [&] {
    k_geo_cells<<<blocks_for(N), kBlock, 0, c->stream>>>(d, cgeo, width);
  }));
  HIPCHK(c, launch(c, KID_GEO, [&] {
    k_geo_ell<<<blocks_for(N), kBlock, 0, c->stream>>>(d, enbr_d, einfo_d, eh, eo, earea);
  }));
  // state
  HIPCHK(c, dalloc(c, &c->prim, N));
  HIPCHK(c, dalloc(c, &c->mid_prim, N));
  HIPCHK(c, dalloc(c, &c->ext, N));
  HIPCHK(c, dalloc(c, &c->mid_ext, N));
  HIPCHK(c, dalloc(c, &c->flux, F));
  HIPCHK(c, dalloc(c, &c->rec, N));
  HIPCHK(c, dalloc(c, &c->staging, 4 * static_cast<size_t>(std::max(N, F))));
  HIPCHK(c, dalloc(c, &c->partials, std::max(blocks_for(std::max(N, F)), tiles_for(N))));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (c->prm.precision == RG_MIXED) {
    // Incidence words of the FP32 update.  Periodic copies of a boundary
    // face: the real cell r of face f sees ghost g (a translate of cell b by
    // shift q); the copy of f on b's side is the face between b and a ghost
    // of r translated by -q.  Both cells read the lower-numbered copy, so
    // they add one FP32 flux with opposite signs and the totals telescope.
    std::vector<unsigned> minfo(static_cast<size_t>(std::max(nnz, 1)));
    for (int k = 0; k < nnz; ++k) minfo[k] = inc[k].info & (INFO_FACE | INFO_SECOND);
    if (c->prm.bc == RG_BC_PERIODIC && G > 0) {
      double x0 = m->pt_x[0], x1 = x0, y0 = m->pt_y[0], y1 = y0;
      for (int i = 1; i < N; ++i) {
        x0 = std::min(x0, m->pt_x[i]);
        x1 = std::max(x1, m->pt_x[i]);
        y0 = std::min(y0, m->pt_y[i]);
        y1 = std::max(y1, m->pt_y[i]);
      }
      const double hw = 0.5 * (x1 - x0), hh = 0.5 * (y1 - y0);
      auto key = [](int r, int b, int qx, int qy) {
        return (static_cast<unsigned long long>(r) << 32 | static_cast<unsigned>(b)) * 9u +
               static_cast<unsigned long long>((qx + 1) * 3 + (qy + 1));
      };
      std::unordered_map<unsigned long long, int> bface;
      std::vector<unsigned long long> twin_key(static_cast<size_t>(F), ~0ull);
      for (int f = 0; f < F; ++f) {
        const int a = m->face_first[f], b = m->face_second[f];
        const int g = a >= N ? a : (b >= N ? b : -1);
        if (g < 0) continue;
        const int r = g == a ? b : a, o = orig(g);
        if (o == r) continue;
        const double sx = m->pt_x[g] - m->pt_x[o], sy = m->pt_y[g] - m->pt_y[o];
        const int qx = sx > hw ? 1 : (sx < -hw ? -1 : 0);
        const int qy = sy > hh ? 1 : (sy < -hh ? -1 : 0);
        bface.emplace(key(r, o, qx, qy), f);
        twin_key[f] = key(o, r, -qx, -qy);
      }
      for (int k = 0; k < nnz; ++k) {
        const int f = static_cast<int>(inc[k].info & INFO_FACE);
        if (twin_key[f] == ~0ull) continue;
        const auto it = bface.find(twin_key[f]);
        if (it == bface.end() || it->second >= f) continue;
        const int t = it->second;  // this cell is the ghost side of face t
        minfo[k] = static_cast<unsigned>(t) | (m->face_second[t] >= N ? INFO_SECOND : 0u);
      }
    }
    std::vector<unsigned> einfoM(ne, 0u);
    for (int i = 0; i < N; ++i)
      for (int k = m->csr_off[i]; k < m->csr_off[i + 1]; ++k) einfoM[ell_at(i, k)] = minfo[k];
    c->adj_src.minfo = std::move(minfo);  // the mixed tiles' face words
    FaceM* faceM;
    CellGeoM* cgeoM;
    float2 *deh, *deo;
    unsigned* deinfo;
    HIPCHK(c, dalloc(c, &faceM, F));
    HIPCHK(c, dalloc(c, &cgeoM, N));
    HIPCHK(c, upload(c, &deinfo, einfoM.data(), ne));
    HIPCHK(c, dalloc(c, &deh, ne));
    HIPCHK(c, dalloc(c, &deo, ne));
    HIPCHK(c, hipMemsetAsync(deh, 0, sizeof(float2) * ne, c->stream));
    HIPCHK(c, hipMemsetAsync(deo, 0, sizeof(float2) * ne, c->stream));
    HIPCHK(c, dalloc(c, &c->primM, N));
    HIPCHK(c, dalloc(c, &c->mid_primM, N));
    HIPCHK(c, dalloc(c, &c->fluxM, F));
    HIPCHK(c, dalloc(c, &c->recM, N));
    HIPCHK(c, launch(c, KID_GEO, [&] {
      k_geo_mixed_faces<<<blocks_for(F), kBlock, 0, c->stream>>>(d, faceM);
    }));
    HIPCHK(c, launch(c, KID_GEO, [&] {
      k_geo_mixed_cells<<<blocks_for(N), kBlock, 0, c->stream>>>(d, cgeoM);
    }));
    HIPCHK(c, launch(c, KID_GEO, [&] {
      k_geo_mixed_ell<<<blocks_for(N), kBlock, 0, c->stream>>>(d, deh, deo);
    }));
    HIPCHK(c, hipStreamSynchronize(c->stream));  // host vectors above are uploaded
    c->mt = MTables{faceM, cgeoM, wofs_d, lanepos_d, enbr_d, deh, deo, deinfo};
  }
  c->dm = d;
  c->N = N;
  c->F = F;
  c->G = G;
  c->nnz = nnz;
  c->adj_src.inc = std::move(inc);
  c->adj_src.csr_off.assign(m->csr_off, m->csr_off + N + 1);
  c->adj_src.ff.assign(m->face_first, m->face_first + F);
  c->adj_src.fs.assign(m->face_second, m->face_second + F);
  c->adj_src.gsrc.assign(gsrc.begin(), gsrc.begin() + G);
  c->adj_src.wofs = std::move(wofs);
  c->adj_src.lanepos = std::move(lanepos);
  c->has_mesh = true;
  drain_pool(c);
  return RG_OK;
}

int rg_set_face_velocity(rg_ctx* c, const double* wx, const double* wy) {
  if (!c || !c->has_mesh || (!wx) != (!wy)) return RG_E_ARG;
  if (!wx) {
    c->dm.wx = c->dm.wy = nullptr;
    return RG_OK;
  }
  if (!c->wbuf) HIPCHK(c, dalloc(c, &c->wbuf, 2 * static_cast<size_t>(c->F)));
  HIPCHK(c, hipMemcpyAsync(c->wbuf, wx, sizeof(double) * c->F, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->wbuf + c->F, wy, sizeof(double) * c->F, hipMemcpyHostToDevice,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->dm.wx = c->wbuf;
  c->dm.wy = c->wbuf + c->F;
  return RG_OK;
}

int rg_set_cells(rg_ctx* c, const double* d, const double* p, const double* vx, const double* vy) {
  if (!ready(c, false, false)) return RG_E_ARG;
  const double* const src[4] = {d, p, vx, vy};
  const int rc = put4(c, c->prim, src, c->N);
  if (rc == RG_OK) c->has_cells = true;
  c->ext_init = false;
  return rc;
}

int rg_get_cells(rg_ctx* c, double* d, double* p, double* vx, double* vy) {
  if (!ready(c, true, false)) return RG_E_ARG;
  double* const dst[4] = {d, p, vx, vy};
  return get4(c, dst, c->prim, c->N);
}

int rg_set_extensives(rg_ctx* c, const double* m, const double* E, const double* Mx,
                      const double* My) {
  if (!ready(c, false, false)) return RG_E_ARG;
  const double* const src[4] = {m, E, Mx, My};
  const int rc = put4(c, c->ext, src, c->N);
  if (rc == RG_OK) c->has_ext = true;
  c->ext_init = false;
  return rc;
}

int rg_get_extensives(rg_ctx* c, double* m, double* E, double* Mx, double* My) {
  if (!ready(c, false, true)) return RG_E_ARG;
  double* const dst[4] = {m, E, Mx, My};
  return get4(c, dst, c->ext, c->N);
}

int rg_set_fluxes(rg_ctx* c, const double* fm, const double* fE, const double* fMx,
                  const double* fMy) {
  if (!ready(c, false, false)) return RG_E_ARG;
  const double* const src[4] = {fm, fE, fMx, fMy};
  const int rc = put4(c, c->flux, src, c->F);
  if (rc == RG_OK) c->has_flux = true;
  return rc;
}

int rg_init_extensives(rg_ctx* c) {
  if (!ready(c, true, false)) return RG_E_ARG;
  HIPCHK(c, launch(c, KID_INIT_EXT, [&] {
    k_init_ext<<<blocks_for(c->N), kBlock, 0, c->stream>>>(c->dm, c->prim, c->ext, c->prm.gamma);
  }));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->has_ext = true;
  c->ext_init = true;
  return RG_OK;
}

int rg_time_step(rg_ctx* c, double* dt) {
  if (!ready(c, true, false) || !dt) return RG_E_ARG;
  HIPCHK(c, run_cfl(c, c->prim, 0));
  const int rc = collect_error(c);
  if (rc != RG_OK) return rc;
  double v;
  std::memcpy(&v, &c->host_words[0], sizeof(double));
  *dt = c->prm.cfl * v;
  return RG_OK;
}

int rg_reconstruct(rg_ctx* c, double* slopes, double* left, double* right) {
  if (!ready(c, true, false)) return RG_E_ARG;
  const bool pcm = c->prm.order != 2;
  if (!pcm) HIPCHK(c, run_slopes(c, c->prim, 0, 1, c->rec));
  if (left || right) {
    if (!c->left) {
      HIPCHK(c, dalloc(c, &c->left, c->F));
      HIPCHK(c, dalloc(c, &c->right, c->F));
    }
    HIPCHK(c, launch(c, KID_FACE_STATES, [&] {
      if (pcm)
        k_face_states_pcm<<<blocks_for(c->F), kBlock, 0, c->stream>>>(c->dm, c->prim, c->left,
                                                                      c->right);
      else
        k_face_states<<<blocks_for(c->F), kBlock, 0, c->stream>>>(c->dm, c->prim, c->rec,
                                                                 c->left, c->right);
    }));
  }
  const ErrSrc src{{c->prim, nullptr}, {nullptr, nullptr}};
  const int rc = collect_error(c, &src);
  if (rc != RG_OK) return rc;
  if (slopes && pcm) {
    std::memset(slopes, 0, sizeof(double) * 8 * static_cast<size_t>(c->N));
  } else if (slopes) {
    std::vector<Rec> h(static_cast<size_t>(c->N));
    HIPCHK(c, hipMemcpyAsync(h.data(), c->rec, sizeof(Rec) * c->N, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const size_t n = static_cast<size_t>(c->N);
    for (size_t i = 0; i < n; ++i) {
      const Rec& r = h[i];
      const double v[8] = {r.xd, r.xp, r.xvx, r.xvy, r.yd, r.yp, r.yvx, r.yvy};
      for (int k = 0; k < 8; ++k) slopes[k * n + i] = v[k];
    }
  }
  const size_t F = static_cast<size_t>(c->F);
  if (left) {
    double* const dst[4] = {left, left + F, left + 2 * F, left + 3 * F};
    const int r2 = get4(c, dst, c->left, c->F);
    if (r2 != RG_OK) return r2;
  }
  if (right) {
    double* const dst[4] = {right, right + F, right + 2 * F, right + 3 * F};
    const int r2 = get4(c, dst, c->right, c->F);
    if (r2 != RG_OK) return r2;
  }
  return RG_OK;
}

int rg_flux(rg_ctx* c, double* fm, double* fE, double* fMx, double* fMy) {
  if (!ready(c, true, false)) return RG_E_ARG;
  HIPCHK(c, run_flux(c, c->prim, 0, 1));
  const ErrSrc src{{c->prim, nullptr}, {nullptr, nullptr}};
  const int rc = collect_error(c, &src);
  if (rc != RG_OK) return rc;
  c->has_flux = true;
  if (!fm && !fE && !fMx && !fMy) return RG_OK;
  double* const dst[4] = {fm, fE, fMx, fMy};
  return get4(c, dst, c->flux, c->F);
}

int rg_riemann(rg_ctx* c, int32_t n, const double* left, const double* right,
               const double* velocity, double* out) {
  if (!c || n < 0 || (n > 0 && (!left || !right || !out))) return RG_E_ARG;
  if (n == 0) return RG_OK;
  const size_t N = static_cast<size_t>(n), need = 17 * N;
  if (c->riemann_cap < need) {
    if (c->riemann_buf) (void)hipFree(c->riemann_buf);
    c->riemann_buf = nullptr;
    c->riemann_cap = 0;
    HIPCHK(c, hipMalloc(reinterpret_cast<void**>(&c->riemann_buf), sizeof(double) * need));
    c->riemann_cap = need;
  }
  double* const l = c->riemann_buf;
  double* const r = l + 6 * N;
  double* const w = r + 6 * N;
  double* const f = w + N;
  HIPCHK(c, hipMemcpyAsync(l, left, sizeof(double) * 6 * N, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(r, right, sizeof(double) * 6 * N, hipMemcpyHostToDevice, c->stream));
  if (velocity)
    HIPCHK(c, hipMemcpyAsync(w, velocity, sizeof(double) * N, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, launch(c, KID_RIEMANN, [&] {
    k_riemann<<<blocks_for(n), kBlock, 0, c->stream>>>(l, r, velocity ? w : nullptr, f, n,
                                                       derr(c, 0, 2));
  }));
  HIPCHK(c, hipMemcpyAsync(out, f, sizeof(double) * 4 * N, hipMemcpyDeviceToHost, c->stream));
  const int rc = collect_error(c);
  if (rc != RG_OK && c->last.code == RG_ERR_INVALID_WAVE_SPEEDS && c->last.index >= 0 &&
      c->last.index < n) {
    // the entries of hllc.cpp:54-70 from the failing pair's inputs
    const size_t i = static_cast<size_t>(c->last.index);
    const double v = velocity ? velocity[i] : 0.0;
    const double* const side[2] = {left, right};
    double ent[16], u[2], cs[2];
    for (int k = 0; k < 2; ++k) {
      const double* s = side[k];
      u[k] = s[2 * N + i] - 1.0 * v;
      cs[k] = s[5 * N + i];
      ent[6 * k + 0] = s[i];
      ent[6 * k + 1] = s[N + i];
      ent[6 * k + 2] = u[k];
      ent[6 * k + 3] = s[3 * N + i] - 0.0 * v;
      ent[6 * k + 4] = cs[k];
      ent[6 * k + 5] = s[4 * N + i];
    }
    const double sl = smin(u[0] - cs[0], u[1] - cs[1]);
    const double sr = smax(u[0] + cs[0], u[1] + cs[1]);
    const double dl = left[i], dr = right[i], pl = left[N + i], pr = right[N + i];
    ent[12] = v;
    ent[13] = sl;
    ent[14] = (pr - pl + dl * u[0] * (sl - u[0]) - dr * u[1] * (sr - u[1])) /
              (dl * (sl - u[0]) - dr * (sr - u[1]));
    ent[15] = sr;
    for (int k = 0; k < 16; ++k) add_entry(c->last, kWaveEntryNames[k], ent[k]);
  }
  return rc;
}

int rg_update_extensives(rg_ctx* c, double dt) {
  if (!ready(c, false, true) || !c->has_flux) return RG_E_ARG;
  UpdArgs a{};
  a.in = c->ext;
  a.out = c->ext;
  a.flux = c->flux;
  a.dt = dt;
  a.flags = UPD_FACES;
  c->ext_init = false;
  HIPCHK(c, run_update(c, a, 0, 3));
  return collect_error(c);
}

int rg_update_cells(rg_ctx* c) {
  if (!ready(c, false, true)) return RG_E_ARG;
  UpdArgs a{};
  a.in = c->ext;
  a.out = c->ext;
  a.cells = c->prim;
  a.flags = UPD_CELLS;
  c->ext_init = false;
  HIPCHK(c, run_update(c, a, 0, 3));
  const ErrSrc src{{nullptr, nullptr}, {c->ext, nullptr}};
  const int rc = collect_error(c, &src);
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
  c->ext_init = false;
  if (c->ens_K > 0) {
    if (integrator != 2 || c->prm.order != 2 || c->prm.precision != RG_FP64 || c->dm.wx)
      return RG_E_ARG;
    if (nsteps > c->ens_hist_cap) {
      if (c->ens_hist) (void)hipFree(c->ens_hist);
      c->ens_hist = nullptr;
      c->ens_hist_cap = 0;
      HIPCHK(c, hipMalloc(reinterpret_cast<void**>(&c->ens_hist),
                          sizeof(double) * static_cast<size_t>(nsteps) * c->ens_K));
      c->ens_hist_cap = nsteps;
      if (c->ens_graph) (void)hipGraphExecDestroy(c->ens_graph);  // holds the old pointer
      c->ens_graph = nullptr;
    }
    const bool prof = c->prof;
    c->prof = false;  // no per-launch events inside a graph
    HIPCHK(c, hipMemsetAsync(c->ens_step, 0, sizeof(int), c->stream));
    HIPCHK(c, ens_seed(c));
    if (!c->ens_graph) HIPCHK(c, ens_capture(c, c->ens_block, c->ens_bpc));
    for (int s = 0; s < nsteps; ++s) HIPCHK(c, hipGraphLaunch(c->ens_graph, c->stream));
    c->prof = prof;
    if (dt_out)
      HIPCHK(c, hipMemcpyAsync(dt_out, c->ens_hist,
                               sizeof(double) * static_cast<size_t>(nsteps) * c->ens_K,
                               hipMemcpyDeviceToHost, c->stream));
    return collect_error(c);
  }
  if (c->prm.precision == RG_MIXED && integrator == 2 && c->prm.order == 2)
    return advance_mixed(c, nsteps, dt_out);
  bool fused = false;
  HIPCHK(c, advance_fp64(c, integrator, nsteps, dt_out != nullptr, true, &fused));
  if (dt_out)
    HIPCHK(c, hipMemcpyAsync(dt_out, c->dt_hist, sizeof(double) * nsteps, hipMemcpyDeviceToHost,
                             c->stream));
  return collect_fp64(c, integrator, nsteps, fused);
}

int rg_advance_noalloc(rg_ctx* c, int32_t integrator, int32_t nsteps, double* dt_out) {
  if (!ready(c, true, true) || (integrator != 1 && integrator != 2) || nsteps < 0)
    return RG_E_ARG;
  if (c->ens_K > 0 || c->prm.precision != RG_FP64)
    return arg_error(c, "rg_advance_noalloc: FP64 contexts without instances only");
  if (nsteps > kScratchSteps)
    return arg_error(c, "rg_advance_noalloc: more than 65536 steps in one call");
  if (std::max(blocks_for(c->N), tiles_for(c->N)) > static_cast<unsigned>(kScratchBlocks))
    return arg_error(c, "rg_advance_noalloc: more than 16777216 cells");
  if (!c->scratch) {
    c->last = rg_error{};
    c->last.code = -1;
    std::snprintf(c->last.message, sizeof(c->last.message),
                  "rg_advance_noalloc: no scratch on this device");
    return RG_E_HIP;
  }
  if (nsteps == 0) return RG_OK;
  c->ext_init = false;
  ScratchScope scope(c);
  // the scratch keeps the last call's contents: start with clear dt slots
  // and error word (a fill kernel, no memset)
  HIPCHK(c, fill_words(c, 0, 3, ~0ull));
  // the fused stages only where an earlier rg_advance built the tiles: this
  // call allocates nothing
  bool fused = false;
  HIPCHK(c, advance_fp64(c, integrator, nsteps, dt_out != nullptr, false, &fused));
  if (dt_out)
    HIPCHK(c, hipMemcpyAsync(dt_out, c->dt_hist, sizeof(double) * nsteps, hipMemcpyDeviceToHost,
                             c->stream));
  return collect_fp64(c, integrator, nsteps, fused);
}

int rg_adjoint_forward(rg_ctx* c, int32_t nsteps, int32_t every, double* dt_out) {
  if (!ready(c, true, true) || c->prm.order != 2 || c->prm.precision != RG_FP64 || c->dm.wx ||
      nsteps <= 0 || nsteps >= (1 << 24) || every <= 0 || !c->ext_init)
    return RG_E_ARG;
  if (!c->am.rev) HIPCHK(c, build_adjoint_tables(c));
  const int N = c->N, F = c->F;
  const int nck = (nsteps + every - 1) / every, ev = std::min(every, nsteps);
  // buffers sized per run (reallocated when a run needs more)
  if (nck > c->adj_cap_nck || ev > c->adj_cap_every || !c->gP) {
    for (void* q : c->adj_bufs) (void)hipFree(q);
    c->adj_bufs.clear();
    auto get = [&](auto** ptr, size_t n) -> hipError_t {
      void* q = nullptr;
      const hipError_t e = hipMalloc(&q, sizeof(**ptr) * std::max<size_t>(n, 1));
      if (e == hipSuccess) {
        c->adj_bufs.push_back(q);
        *ptr = static_cast<std::remove_reference_t<decltype(*ptr)>>(q);
      }
      return e;
    };
    const size_t n = static_cast<size_t>(N), ne = static_cast<size_t>(c->dm_ne);
    HIPCHK(c, get(&c->ckP, n * nck));
    HIPCHK(c, get(&c->ckX, n * nck));
    HIPCHK(c, get(&c->finP, n));
    HIPCHK(c, get(&c->finX, n));
    for (rg_ctx::Tape& T : c->tape) {
      HIPCHK(c, get(&T.P, n * (ev + 1)));
      HIPCHK(c, get(&T.X, n * (ev + 1)));
      HIPCHK(c, get(&T.Pm, n * ev));
      HIPCHK(c, get(&T.Xm, n * ev));
      HIPCHK(c, get(&T.R1, n * ev));
      HIPCHK(c, get(&T.R2, n * ev));
      HIPCHK(c, get(&T.dt, static_cast<size_t>(ev)));
    }
    HIPCHK(c, get(&c->rflux, static_cast<size_t>(F)));
    HIPCHK(c, get(&c->rwords, 4));
    HIPCHK(c, get(&c->rpartials,
                  static_cast<size_t>(std::max(blocks_for(std::max(N, F)), tiles_for(N)))));
    HIPCHK(c, get(&c->gP, n));
    HIPCHK(c, get(&c->gX, n));
    HIPCHK(c, get(&c->ga, n));
    HIPCHK(c, get(&c->gb, n));
    HIPCHK(c, get(&c->gps, n));
    HIPCHK(c, get(&c->lb, static_cast<size_t>(F)));
    HIPCHK(c, get(&c->rb, static_cast<size_t>(F)));
    HIPCHK(c, get(&c->emit, ne));
    HIPCHK(c, get(&c->dtb, 2 * static_cast<size_t>(blocks_for(F))));
    HIPCHK(c, get(&c->vmin, static_cast<size_t>(blocks_for(N))));
    HIPCHK(c, get(&c->imin, static_cast<size_t>(blocks_for(N))));
    c->adj_cap_nck = nck;
    c->adj_cap_every = ev;
  }
  if (!c->rstream) {
    HIPCHK(c, hipStreamCreateWithFlags(&c->rstream, hipStreamNonBlocking));
    for (int b = 0; b < 2; ++b) {
      HIPCHK(c, hipEventCreateWithFlags(&c->tape_ready[b], hipEventDisableTiming));
      HIPCHK(c, hipEventCreateWithFlags(&c->tape_free[b], hipEventDisableTiming));
    }
  }
  if (dt_out && c->dt_hist_cap < nsteps) {
    if (c->dt_hist) (void)hipFree(c->dt_hist);
    c->dt_hist = nullptr;
    c->dt_hist_cap = 0;
    HIPCHK(c, hipMalloc(reinterpret_cast<void**>(&c->dt_hist), sizeof(double) * nsteps));
    c->dt_hist_cap = nsteps;
  }
  const size_t bytes = sizeof(D4) * static_cast<size_t>(N);
  c->ext_init = false;
  // the fused stages where the mesh has tiles, alternating the two state
  // buffers as rg_advance does
  if (c->tiles_on && !c->tiles_built) HIPCHK(c, build_tiles(c));
  const bool fused = fused_ok(c);
  HIPCHK(c, run_cfl(c, c->prim, 0));
  for (int s = 0; s < nsteps; ++s) {
    if (s % every == 0) {
      const size_t k = static_cast<size_t>(s / every) * N;
      HIPCHK(c, hipMemcpyAsync(c->ckP + k, c->prim, bytes, hipMemcpyDeviceToDevice, c->stream));
      HIPCHK(c, hipMemcpyAsync(c->ckX + k, c->ext, bytes, hipMemcpyDeviceToDevice, c->stream));
    }
    if (fused) {
      const HeunBufs h{c->prim, c->ext, c->mid_prim, c->mid_ext,
                       c->alt_prim, c->alt_ext, c->rec, c->rec};
      HIPCHK(c, heun_step_fused(c, s, dt_out ? c->dt_hist : nullptr, s, h));
      std::swap(c->prim, c->alt_prim);
      std::swap(c->ext, c->alt_ext);
    } else {
      HIPCHK(c, heun_step(c, s, dt_out ? c->dt_hist : nullptr, s, ctx_bufs(c)));
    }
  }
  HIPCHK(c, hipMemcpyAsync(c->finP, c->prim, bytes, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->finX, c->ext, bytes, hipMemcpyDeviceToDevice, c->stream));
  if (dt_out)
    HIPCHK(c, hipMemcpyAsync(dt_out, c->dt_hist, sizeof(double) * nsteps, hipMemcpyDeviceToHost,
                             c->stream));
  const int rc = collect_fp64(c, 2, nsteps, fused);
  if (rc != RG_OK) return rc;
  c->adj_steps = nsteps;
  c->adj_every = every;
  c->adj_nck = nck;
  return RG_OK;
}

namespace {

// Reverse of one Heun stage on the cells q with the taped slopes rec: seed
// c->ga = adjoint of the stage's output extensives.  mode 2 (second stage)
// leaves the adjoint of the mid extensives in c->ga; mode 1 (first stage)
// writes the step-start adjoints to c->gP and c->gX and the per-block CFL
// argmin of q.
hipError_t adj_stage(rg_ctx* c, const D4* q, const Rec* rec, const D4* X, const double* dtp,
                     double* dtb, int mode) {
  const int N = c->N, F = c->F;
  hipError_t e = launch(c, KID_ADJ_FLUX, [&] {
    k_adj_flux<<<blocks_for(F), kBlock, 0, c->stream>>>(c->dm, q, rec, c->ga, dtp, c->prm.cfl,
                                                         c->lb, c->rb, dtb, c->prm.gamma);
  });
  if (e != hipSuccess) return e;
  e = launch(c, KID_ADJ_SLOPES, [&] {
    k_adj_slopes<<<blocks_for(N), kBlock, 0, c->stream>>>(c->dm, c->am, q, c->lb, c->rb, c->gps,
                                                           c->emit, dparams(c));
  });
  if (e != hipSuccess) return e;
  return launch(c, KID_ADJ_CELL, [&] {
    k_adj_gather<<<blocks_for(N), kBlock, 0, c->stream>>>(c->dm, c->am, c->gps, c->emit, mode, X,
                                                           c->ga, c->gb, c->gP, c->gX, q, c->vmin,
                                                           c->imin, c->prm.gamma);
  });
}

}  // namespace

int rg_adjoint_backward(rg_ctx* c, const double* gcells, const double* gext, double* gcells0) {
  if (!c || !c->has_mesh || c->adj_steps <= 0 || !gcells || !gcells0) return RG_E_ARG;
  const int N = c->N, F = c->F, nsteps = c->adj_steps, every = c->adj_every;
  const size_t n = static_cast<size_t>(N), bytes = sizeof(D4) * n;
  const unsigned nbF = blocks_for(F), nbN = blocks_for(N);
  {
    const double* const src[4] = {gcells, gcells + n, gcells + 2 * n, gcells + 3 * n};
    const int rc = put4(c, c->gP, src, N);
    if (rc != RG_OK) return rc;
  }
  if (gext) {
    const double* const src[4] = {gext, gext + n, gext + 2 * n, gext + 3 * n};
    const int rc = put4(c, c->gX, src, N);
    if (rc != RG_OK) return rc;
  } else {
    HIPCHK(c, hipMemsetAsync(c->gX, 0, bytes, c->stream));
  }
  PendingDt pd{nullptr, c->dtb, static_cast<int>(2 * nbF), c->vmin, c->imin,
               static_cast<int>(nbN), c->prm.cfl};
  // The recompute runs on its own stream with its own fluxes, dt and error
  // words and block minima: while the enqueue below is inside a Recompute
  // scope, the context's stream and those buffers are the recompute's.
  struct Recompute {
    rg_ctx* c;
    explicit Recompute(rg_ctx* cc) : c(cc) { flip(); }
    ~Recompute() { flip(); }
    void flip() {
      std::swap(c->stream, c->rstream);
      std::swap(c->flux, c->rflux);
      std::swap(c->words, c->rwords);
      std::swap(c->partials, c->rpartials);
    }
  };
  {
    Recompute r(c);
    HIPCHK(c, fill_words(c, 2, 1, ~0ull));  // no error yet
  }
  for (int b = 0; b < 2; ++b) HIPCHK(c, hipEventRecord(c->tape_free[b], c->stream));
  // recompute a segment from its checkpoint into tape seg & 1: states
  // P/X[0..L], mid states, both stages' slopes and the dts
  auto recompute = [&](int seg) -> hipError_t {
    Recompute r(c);
    const rg_ctx::Tape& T = c->tape[seg & 1];
    const int k0 = seg * every, L = std::min(k0 + every, nsteps) - k0;
    hipError_t e = hipStreamWaitEvent(c->stream, c->tape_free[seg & 1], 0);
    if (e == hipSuccess)
      e = hipMemcpyAsync(T.P, c->ckP + static_cast<size_t>(seg) * n, bytes,
                         hipMemcpyDeviceToDevice, c->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(T.X, c->ckX + static_cast<size_t>(seg) * n, bytes,
                         hipMemcpyDeviceToDevice, c->stream);
    if (e == hipSuccess) e = run_cfl(c, T.P, k0 & 1);
    for (int t = 0; t < L && e == hipSuccess; ++t) {
      const size_t o = static_cast<size_t>(t) * n;
      const HeunBufs h{T.P + o,     T.X + o,     T.Pm + o, T.Xm + o,
                       T.P + o + n, T.X + o + n, T.R1 + o, T.R2 + o};
      // each step's end state goes to the next tape slot, as the fused
      // stages require
      e = fused_ok(c) ? heun_step_fused(c, k0 + t, T.dt, t, h) : heun_step(c, k0 + t, T.dt, t, h);
    }
    if (e == hipSuccess) e = hipEventRecord(c->tape_ready[seg & 1], c->stream);
    return e;
  };
  HIPCHK(c, recompute(c->adj_nck - 1));
  for (int seg = c->adj_nck - 1; seg >= 0; --seg) {
    const int k0 = seg * every, k1 = std::min(k0 + every, nsteps), L = k1 - k0;
    if (seg > 0) HIPCHK(c, recompute(seg - 1));  // beside this segment's reverse
    const rg_ctx::Tape& T = c->tape[seg & 1];
    HIPCHK(c, hipStreamWaitEvent(c->stream, c->tape_ready[seg & 1], 0));
    // reverse through the segment
    for (int t = L - 1; t >= 0; --t) {
      const size_t o = static_cast<size_t>(t) * n;
      // the dt adjoint of step k0+t+1 is pending except for the last step
      pd.q = (k0 + t + 1 < nsteps) ? T.P + o + n : nullptr;
      HIPCHK(c, launch(c, KID_ADJ_CELL, [&] {
        k_adj_recover<<<nbN, kBlock, 0, c->stream>>>(c->dm, T.X + o + n, c->gP, c->gX, c->ga,
                                                     c->gb, c->prm.gamma, pd);
      }));
      HIPCHK(c, adj_stage(c, T.Pm + o, T.R2 + o, T.Xm + o, T.dt + t, c->dtb, 2));
      HIPCHK(c, adj_stage(c, T.P + o, T.R1 + o, nullptr, T.dt + t, c->dtb + nbF, 1));
    }
    HIPCHK(c, hipEventRecord(c->tape_free[seg & 1], c->stream));
  }
  pd.q = c->ckP;  // the dt adjoint of step 0
  HIPCHK(c, launch(c, KID_ADJ_CELL, [&] {
    k_adj_init<<<nbN, kBlock, 0, c->stream>>>(c->dm, c->ckP, c->gX, c->gP, c->prm.gamma, pd);
  }));
  {
    double* const dst[4] = {gcells0, gcells0 + n, gcells0 + 2 * n, gcells0 + 3 * n};
    const int rc = get4(c, dst, c->gP, N);
    if (rc != RG_OK) return rc;
  }
  // leave the forward run's final state in place
  HIPCHK(c, hipMemcpyAsync(c->prim, c->finP, bytes, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->ext, c->finX, bytes, hipMemcpyDeviceToDevice, c->stream));
  int rc;
  {
    Recompute r(c);  // errors the recompute reported
    rc = collect_error(c);
  }
  const int rc0 = collect_error(c);
  return rc != RG_OK ? rc : rc0;
}

int rg_set_instances(rg_ctx* c, int32_t n_inst, const int32_t* cell_off) {
  if (!c || !c->has_mesh || n_inst <= 0 || !cell_off || cell_off[0] != 0 ||
      cell_off[n_inst] != c->N)
    return RG_E_ARG;
  for (int k = 0; k < n_inst; ++k)
    if (cell_off[k + 1] <= cell_off[k]) return RG_E_ARG;
  const int N = c->N, F = c->F, G = c->G;
  std::vector<int> inst(static_cast<size_t>(N));
  for (int k = 0; k < n_inst; ++k)
    for (int i = cell_off[k]; i < cell_off[k + 1]; ++i) inst[i] = k;
  // every face and ghost must belong to one instance
  std::vector<int> ff(static_cast<size_t>(F)), fs(static_cast<size_t>(F)),
      go(static_cast<size_t>(std::max(G, 1)));
  HIPCHK(c, hipMemcpy(ff.data(), c->dm.ff, sizeof(int) * F, hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(fs.data(), c->dm.fs, sizeof(int) * F, hipMemcpyDeviceToHost));
  if (G) HIPCHK(c, hipMemcpy(go.data(), c->dm.gorig, sizeof(int) * G, hipMemcpyDeviceToHost));
  auto owner = [&](int id) { return inst[id < N ? id : go[id - N]]; };
  for (int f = 0; f < F; ++f)
    if (owner(ff[f]) != owner(fs[f])) return RG_E_ARG;
  int dev = 0, cu = 0;
  HIPCHK(c, hipGetDevice(&dev));
  HIPCHK(c, hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev));
  HIPCHK(c, upload(c, &c->ens_inst, inst.data(), inst.size()));
  HIPCHK(c, upload(c, &c->ens_coff, cell_off, static_cast<size_t>(n_inst) + 1));
  HIPCHK(c, dalloc(c, &c->ens_dtA, n_inst));
  HIPCHK(c, dalloc(c, &c->ens_dtB, n_inst));
  HIPCHK(c, dalloc(c, &c->ens_inv, N));
  HIPCHK(c, dalloc(c, &c->ens_step, 1));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (c->ens_graph) (void)hipGraphExecDestroy(c->ens_graph);
  c->ens_graph = nullptr;
  c->ens_K = n_inst;
  c->ens_cu = cu;
  return RG_OK;
}

int rg_ensemble_config(rg_ctx* c, int32_t block, int32_t blocks_per_cu) {
  if (!c || c->ens_K <= 0 || (block != 64 && block != 128 && block != 256) || blocks_per_cu < 0 ||
      blocks_per_cu > 64)
    return RG_E_ARG;
  c->ens_block = block;
  c->ens_bpc = blocks_per_cu;
  if (c->ens_graph) (void)hipGraphExecDestroy(c->ens_graph);
  c->ens_graph = nullptr;
  return RG_OK;
}

int rg_ensemble_autotune(rg_ctx* c, int32_t trial_steps, double* table, int32_t table_rows,
                         int32_t* rows_out) {
  if (!ready(c, true, true) || c->ens_K <= 0 || trial_steps <= 0) return RG_E_ARG;
  const size_t bytes = sizeof(D4) * static_cast<size_t>(c->N);
  // released on every return: the saved state, the timing events, and the
  // profiling switch (off while the shapes are timed)
  struct Scope {
    rg_ctx* c;
    bool prof;
    D4 *saveP = nullptr, *saveX = nullptr;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    ~Scope() {
      if (saveP) (void)hipFree(saveP);
      if (saveX) (void)hipFree(saveX);
      if (t0) (void)hipEventDestroy(t0);
      if (t1) (void)hipEventDestroy(t1);
      c->prof = prof;
    }
  } sc{c, c->prof};
  D4*& saveP = sc.saveP;
  D4*& saveX = sc.saveX;
  HIPCHK(c, hipMalloc(reinterpret_cast<void**>(&saveP), bytes));
  HIPCHK(c, hipMalloc(reinterpret_cast<void**>(&saveX), bytes));
  HIPCHK(c, hipMemcpyAsync(saveP, c->prim, bytes, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(saveX, c->ext, bytes, hipMemcpyDeviceToDevice, c->stream));
  if (trial_steps + 1 > c->ens_hist_cap) {
    if (c->ens_hist) (void)hipFree(c->ens_hist);
    c->ens_hist = nullptr;
    c->ens_hist_cap = 0;
    HIPCHK(c, hipMalloc(reinterpret_cast<void**>(&c->ens_hist),
                        sizeof(double) * static_cast<size_t>(trial_steps + 1) * c->ens_K));
    c->ens_hist_cap = trial_steps + 1;
  }
  c->prof = false;
  hipEvent_t& t0 = sc.t0;
  hipEvent_t& t1 = sc.t1;
  HIPCHK(c, hipEventCreate(&t0));
  HIPCHK(c, hipEventCreate(&t1));
  const int blocks[3] = {64, 128, 256};
  const int bpcs[6] = {1, 2, 4, 8, 16, 0};  // 0: one thread per item
  double best = 1e300;
  int best_block = 256, best_bpc = 4, rows = 0;
  for (int bi = 0; bi < 3; ++bi)
    for (int pi = 0; pi < 6; ++pi) {
      const int block = blocks[bi], bpc = bpcs[pi];
      if (block * bpc > 2048) continue;  // at most 32 waves per CU
      HIPCHK(c, hipMemcpyAsync(c->prim, saveP, bytes, hipMemcpyDeviceToDevice, c->stream));
      HIPCHK(c, hipMemcpyAsync(c->ext, saveX, bytes, hipMemcpyDeviceToDevice, c->stream));
      HIPCHK(c, hipMemsetAsync(c->ens_step, 0, sizeof(int), c->stream));
      c->ens_block = block;
      c->ens_bpc = bpc;
      HIPCHK(c, ens_seed(c));
      HIPCHK(c, ens_capture(c, block, bpc));
      HIPCHK(c, hipGraphLaunch(c->ens_graph, c->stream));  // warm-up step
      HIPCHK(c, hipEventRecord(t0, c->stream));
      for (int s = 0; s < trial_steps; ++s) HIPCHK(c, hipGraphLaunch(c->ens_graph, c->stream));
      HIPCHK(c, hipEventRecord(t1, c->stream));
      HIPCHK(c, hipEventSynchronize(t1));
      float ms = 0.f;
      HIPCHK(c, hipEventElapsedTime(&ms, t0, t1));
      const double per = ms / trial_steps;
      if (table && rows < table_rows) {
        table[3 * rows] = block;
        table[3 * rows + 1] = bpc;
        table[3 * rows + 2] = per;
      }
      ++rows;
      if (per < best) {
        best = per;
        best_block = block;
        best_bpc = bpc;
      }
    }
  HIPCHK(c, hipMemcpyAsync(c->prim, saveP, bytes, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->ext, saveX, bytes, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->prof = sc.prof;  // the kept graph is captured as the caller profiles
  HIPCHK(c, ens_capture(c, best_block, best_bpc));
  if (rows_out) *rows_out = rows;
  return collect_error(c);
}

int rg_set_box(rg_ctx* c, const double* box) {
  if (!c || !box || !(box[0] < box[1]) || !(box[2] < box[3])) return RG_E_ARG;
  for (int k = 0; k < 4; ++k) c->box[k] = box[k];
  c->has_box = true;
  return RG_OK;
}

int rg_get_geometry(rg_ctx* c, double* volume, double* tess_volume, double* cm_x,
                    double* cm_y) {
  if (!c || !c->has_mesh) return RG_E_ARG;
  const size_t N = static_cast<size_t>(c->N), NG = N + static_cast<size_t>(c->G);
  if (volume)
    HIPCHK(c, hipMemcpyAsync(volume, c->dm.vol, sizeof(double) * N, hipMemcpyDeviceToHost,
                             c->stream));
  if (tess_volume)
    HIPCHK(c, hipMemcpyAsync(tess_volume, c->dm.tvol, sizeof(double) * N, hipMemcpyDeviceToHost,
                             c->stream));
  if (cm_x || cm_y) {
    std::vector<D2> cm(NG);
    HIPCHK(c, hipMemcpyAsync(cm.data(), c->dm.cm, sizeof(D2) * NG, hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (size_t i = 0; i < NG; ++i) {
      if (cm_x) cm_x[i] = cm[i].x;
      if (cm_y) cm_y[i] = cm[i].y;
    }
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return RG_OK;
}

int rg_permute_cells(rg_ctx* c, const int32_t* order) {
  if (!ready(c, true, true) || !order) return RG_E_ARG;
  const int N = c->N;
  {
    std::vector<char> seen(static_cast<size_t>(N), 0);
    for (int i = 0; i < N; ++i) {
      const int o = order[i];
      if (o < 0 || o >= N || seen[o]) return RG_E_ARG;
      seen[o] = 1;
    }
  }
  const size_t bytes = sizeof(D4) * static_cast<size_t>(N);
  if (c->keep_cap < 3 * static_cast<size_t>(N)) {
    if (c->keep) (void)hipFree(c->keep);
    c->keep = nullptr;
    c->keep_cap = 0;
    HIPCHK(c, hipMalloc(reinterpret_cast<void**>(&c->keep), 3 * bytes));
    c->keep_cap = 3 * static_cast<size_t>(N);
  }
  // the order goes through the staging buffer (4 max(N, F) doubles)
  int* ord = reinterpret_cast<int*>(c->staging);
  HIPCHK(c, hipMemcpyAsync(ord, order, sizeof(int) * N, hipMemcpyHostToDevice, c->stream));
  D4* const arrays[3] = {c->prim, c->ext, c->mid_ext};
  for (int k = 0; k < 3; ++k) {
    HIPCHK(c, launch(c, KID_MOVING, [&] {
      k_gather<D4><<<blocks_for(N), kBlock, 0, c->stream>>>(c->keep + static_cast<size_t>(k) * N,
                                                            arrays[k], ord, N);
    }));
  }
  for (int k = 0; k < 3; ++k)
    HIPCHK(c, hipMemcpyAsync(arrays[k], c->keep + static_cast<size_t>(k) * N, bytes,
                             hipMemcpyDeviceToDevice, c->stream));
  if (c->pv_valid && c->pv) {
    D2* tmp = reinterpret_cast<D2*>(c->keep);
    HIPCHK(c, launch(c, KID_MOVING, [&] {
      k_gather<D2><<<blocks_for(N), kBlock, 0, c->stream>>>(tmp, c->pv, ord, N);
    }));
    HIPCHK(c, hipMemcpyAsync(c->pv, tmp, sizeof(D2) * N, hipMemcpyDeviceToDevice, c->stream));
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->ext_init = false;
  return RG_OK;
}

int rg_replace_mesh(rg_ctx* c, const rg_mesh* m) {
  if (!c || !m || !c->has_mesh || !c->has_cells || !c->has_ext || m->n_cells != c->N)
    return RG_E_ARG;
  const size_t bytes = sizeof(D4) * static_cast<size_t>(c->N);
  if (c->keep_cap < 3 * static_cast<size_t>(c->N)) {  // cells, extensives, mid extensives
    if (c->keep) (void)hipFree(c->keep);
    c->keep = nullptr;
    c->keep_cap = 0;
    HIPCHK(c, hipMalloc(reinterpret_cast<void**>(&c->keep), 3 * bytes));
    c->keep_cap = 3 * static_cast<size_t>(c->N);
  }
  D4* keep = c->keep;
  hipError_t e = hipMemcpyAsync(keep, c->prim, bytes, hipMemcpyDeviceToDevice, c->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(keep + c->N, c->ext, bytes, hipMemcpyDeviceToDevice, c->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(keep + 2 * static_cast<size_t>(c->N), c->mid_ext, bytes,
                       hipMemcpyDeviceToDevice, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return fail_hip(c, e);
  const bool pv_valid = c->pv_valid;
  const int rc = rg_set_mesh(c, m);
  if (rc == RG_OK) {
    e = hipMemcpyAsync(c->prim, keep, bytes, hipMemcpyDeviceToDevice, c->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(c->ext, keep + c->N, bytes, hipMemcpyDeviceToDevice, c->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(c->mid_ext, keep + 2 * static_cast<size_t>(c->N), bytes,
                         hipMemcpyDeviceToDevice, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  }
  if (rc != RG_OK) return rc;
  if (e != hipSuccess) return fail_hip(c, e);
  c->has_cells = c->has_ext = true;
  c->pv_valid = pv_valid;
  return RG_OK;
}

namespace {

// StationaryBox face velocities of the current mesh from the device point
// velocities into the context's face-velocity buffer.
hipError_t moving_faces(rg_ctx* c) {
  if (!c->wbuf) {
    const hipError_t e = dalloc(c, &c->wbuf, 2 * static_cast<size_t>(c->F));
    if (e != hipSuccess) return e;
  }
  c->dm.wx = c->wbuf;
  c->dm.wy = c->wbuf + c->F;
  return launch(c, KID_MOVING, [&] {
    k_face_velocity<<<blocks_for(c->F), kBlock, 0, c->stream>>>(c->dm, c->pv, c->wbuf,
                                                                c->wbuf + c->F);
  });
}

}  // namespace

int rg_moving_begin(rg_ctx* c, double chi, double eta, const double* box, double* dt_out,
                    double* pvx, double* pvy) {
  if (!ready(c, true, true) || c->prm.order != 2 || c->prm.bc != RG_BC_RIGID ||
      c->prm.precision != RG_FP64 || c->ens_K > 0 || !box || !dt_out)
    return RG_E_ARG;
  const int N = c->N;
  if (c->pv_cap < N) {
    if (c->pv) (void)hipFree(c->pv);
    c->pv = nullptr;
    c->pv_cap = 0;
    HIPCHK(c, hipMalloc(reinterpret_cast<void**>(&c->pv), sizeof(D2) * N));
    c->pv_cap = N;
  }
  c->ext_init = false;
  // hdsim2d.cpp:325-344: point velocities, face velocities, dt; ApplyFix;
  // face velocities and dt again
  HIPCHK(c, launch(c, KID_MOVING, [&] {
    k_point_velocity<<<blocks_for(N), kBlock, 0, c->stream>>>(c->dm, c->prim, c->prm.gamma, chi,
                                                             eta, c->pv);
  }));
  HIPCHK(c, moving_faces(c));
  HIPCHK(c, run_cfl(c, c->prim, 0));
  HIPCHK(c, launch(c, KID_MOVING, [&] {
    k_point_fix<<<blocks_for(N), kBlock, 0, c->stream>>>(c->dm, c->pv, c->words, c->prm.cfl,
                                                        box[0], box[1], box[2], box[3]);
  }));
  HIPCHK(c, moving_faces(c));
  HIPCHK(c, run_cfl(c, c->prim, 0));
  // stage 1 on the old mesh (hdsim2d.cpp:346-371): mid = ext + ZeroForce + fluxes
  HIPCHK(c, run_flux(c, c->prim, 0, 1, c->rec, c->words));
  UpdArgs a{};
  a.in = c->ext;
  a.out = c->mid_ext;
  a.flux = c->flux;
  a.dt_slot = c->words;
  a.flags = UPD_FACES | UPD_DELTA | UPD_ZF_BEFORE;
  HIPCHK(c, run_update(c, a, 0, 3));
  unsigned long long w = 0;
  HIPCHK(c, hipMemcpyAsync(&w, c->words, sizeof(w), hipMemcpyDeviceToHost, c->stream));
  std::vector<D2> h;
  if (pvx || pvy) {
    h.resize(static_cast<size_t>(N));
    HIPCHK(c, hipMemcpyAsync(h.data(), c->pv, sizeof(D2) * N, hipMemcpyDeviceToHost, c->stream));
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  double inv;
  std::memcpy(&inv, &w, sizeof(inv));
  *dt_out = c->prm.cfl * inv;
  for (int i = 0; i < N && !h.empty(); ++i) {
    if (pvx) pvx[i] = h[i].x;
    if (pvy) pvy[i] = h[i].y;
  }
  const int rc = collect_error(c);
  c->pv_valid = rc == RG_OK;
  return rc;
}

int rg_moving_end(rg_ctx* c) {
  if (!ready(c, true, true) || !c->pv_valid || c->pv_cap < c->N) return RG_E_ARG;
  c->pv_valid = false;
  // mid cells with the new volumes (hdsim2d.cpp:401-404)
  UpdArgs u{};
  u.in = c->mid_ext;
  u.out = c->mid_ext;
  u.cells = c->mid_prim;
  u.flags = UPD_CELLS;
  HIPCHK(c, run_update(c, u, 0, 3));
  // stage 2 on the new mesh, the average and the cell update (:410-441)
  HIPCHK(c, moving_faces(c));
  HIPCHK(c, run_flux(c, c->mid_prim, 0, 4, c->rec, c->words));
  UpdArgs b{};
  b.in = c->mid_ext;
  b.out = c->ext;
  b.avg = c->ext;
  b.flux = c->flux;
  b.cells = c->prim;
  b.dt_slot = c->words;
  b.flags = UPD_FACES | UPD_DELTA | UPD_ZF_BEFORE | UPD_AVERAGE | UPD_CELLS;
  HIPCHK(c, run_update(c, b, 0, 6));
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

int rg_use_tiles(rg_ctx* c, int32_t on) {
  if (!c || (on != 0 && on != 1)) return RG_E_ARG;
  if (c->tiles_on != (on == 1)) {
    c->tiles_on = on == 1;
    // the ensemble's graph holds the other stage form
    if (c->ens_graph) (void)hipGraphExecDestroy(c->ens_graph);
    c->ens_graph = nullptr;
  }
  return RG_OK;
}

int32_t rg_tiles(rg_ctx* c, int64_t* entries, int32_t* max_entries) {
  const bool on = c && fused_ok(c);
  if (entries) *entries = on ? c->tiles_total : 0;
  if (max_entries) *max_entries = on ? c->tiles_max : 0;
  return on ? 1 : 0;
}

}  // extern "C"
