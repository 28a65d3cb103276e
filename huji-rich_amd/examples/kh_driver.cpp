// A C++ host that drives the library through its C ABI alone, the way an
// hdsim adapter does (no Python, no PyTorch):
//  1. builds a jittered periodic Kelvin-Helmholtz lattice (richmesh.h) and
//     uploads it as flat arrays (rg_set_mesh);
//  2. runs one hdsim::TimeAdvance step plugin by plugin (hdsim2d.cpp:149-241):
//     SimpleCFL, the flux calculator, ZeroForce, the extensive and cell
//     updaters, each a C-ABI call with host arrays in between;
//  3. runs the same step through the device-resident loop rg_advance on a
//     second context and checks that both give the same cells and
//     extensives, bit for bit;
//  4. times second-order Heun steps with rg_advance;
//  5. with an output file, runs K hdsim::TimeAdvance2Heun steps
//     (hdsim2d.cpp:323-449) in the call order of the plugin adapters -- the
//     time step, the flux calculator, ZeroForce and the extensive updater on
//     host copies of the mid extensives, the cell updater, the second stage,
//     the host average of old and mid extensives, the cell updater -- with
//     every plugin a C-ABI call and every state a host array in between, and
//     writes the initial cells, the dts, the cells and the extensives there
//     (raw doubles: cells [4][N], K dts, cells [4][N], extensives [4][N])
//     for a checker to compare.
// Prints one JSON line; exits non-zero on any error or mismatch.
//
//   kh_driver [nx ny heun_steps [K out_file]]      (default 512 256 100)
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "richgpu.h"
#include "richmesh.h"

namespace {

int fail(rg_ctx* c, const char* what, int rc) {
  rg_error e{};
  if (c) rg_last_error(c, &e);
  std::fprintf(stderr, "kh_driver: %s failed (%d): %s\n", what, rc, e.message);
  return 1;
}

#define CHECK(ctx, call)                       \
  do {                                         \
    const int rc_ = (call);                    \
    if (rc_ != RG_OK) return fail(ctx, #call, rc_); \
  } while (0)

}  // namespace

int main(int argc, char** argv) {
  const int nx = argc > 1 ? std::atoi(argv[1]) : 512;
  const int ny = argc > 2 ? std::atoi(argv[2]) : 256;
  const int steps = argc > 3 ? std::atoi(argv[3]) : 100;
  const int plugin_steps = argc > 5 ? std::atoi(argv[4]) : 0;
  const char* out_path = argc > 5 ? argv[5] : nullptr;
  if (nx < 4 || ny < 4 || steps < 1 || plugin_steps < 0) return 2;

  // 1. the mesh: periodic [0,2]x[0,1], jitter 1e-3 of the spacing
  const rm_lattice_params lp{nx, ny, 0.0, 2.0, 0.0, 1.0, 1e-3, 1, RM_BC_PERIODIC, 0};
  rm_mesh* rm = rm_lattice_build(&lp);
  if (!rm) return fail(nullptr, "rm_lattice_build", -1);
  int32_t N, F, G, nnz;
  rm_sizes(rm, &N, &F, &G, &nnz);
  std::vector<int32_t> ff(F), fs(F), off(N + 1), face(nnz), gorig(G > 0 ? G : 1);
  std::vector<double> v1x(F), v1y(F), v2x(F), v2y(F), px(N + G), py(N + G), cmx(N + G),
      cmy(N + G), vol(N), tvol(N);
  rm_export(rm, ff.data(), fs.data(), v1x.data(), v1y.data(), v2x.data(), v2y.data(), px.data(),
            py.data(), cmx.data(), cmy.data(), vol.data(), tvol.data(), off.data(), face.data(),
            gorig.data(), nullptr, nullptr);
  rm_free(rm);
  const rg_mesh mesh{N,          F,          G,          ff.data(),  fs.data(),  v1x.data(),
                     v1y.data(), v2x.data(), v2y.data(), px.data(),  py.data(),  cmx.data(),
                     cmy.data(), vol.data(), tvol.data(), off.data(), face.data(), gorig.data()};

  // the KH state: band |y-0.5|<0.25 rho=2, vx=+0.5; outside rho=1, vx=-0.5;
  // p=2.5; vy = 0.01 sin(4 pi x) exp(-(y-0.25)^2/0.005 - (y-0.75)^2/0.005)
  std::vector<double> d(N), p(N, 2.5), vx(N), vy(N);
  for (int i = 0; i < N; ++i) {
    const bool band = std::fabs(py[i] - 0.5) < 0.25;
    d[i] = band ? 2.0 : 1.0;
    vx[i] = band ? 0.5 : -0.5;
    vy[i] = 0.01 * std::sin(4 * M_PI * px[i]) *
            std::exp(-(py[i] - 0.25) * (py[i] - 0.25) / 0.005 -
                     (py[i] - 0.75) * (py[i] - 0.75) / 0.005);
  }

  const rg_params prm{5.0 / 3.0, 0.3, 2, RG_BC_PERIODIC, 1, 0.2, 0.5, 0.7, RG_FP64};
  rg_ctx *a = nullptr, *b = nullptr;
  CHECK(nullptr, rg_create(&prm, &a));
  CHECK(nullptr, rg_create(&prm, &b));
  for (rg_ctx* c : {a, b}) {
    CHECK(c, rg_set_mesh(c, &mesh));
    CHECK(c, rg_set_cells(c, d.data(), p.data(), vx.data(), vy.data()));
    CHECK(c, rg_init_extensives(c));
  }

  // 2. TimeAdvance plugin by plugin on context a
  double dt = 0;
  CHECK(a, rg_time_step(a, &dt));
  CHECK(a, rg_flux(a, nullptr, nullptr, nullptr, nullptr));  // fluxes stay on the device
  std::vector<double> em(N), eE(N), ex(N), ey(N);
  CHECK(a, rg_get_extensives(a, em.data(), eE.data(), ex.data(), ey.data()));
  for (int i = 0; i < N; ++i) {  // ZeroForce (zero_force.cpp:3-18): ext += dt * 0
    em[i] += dt * 0.0;
    ex[i] += 0.0 * dt;
    ey[i] += 0.0 * dt;
    eE[i] += dt * 0.0;
  }
  CHECK(a, rg_set_extensives(a, em.data(), eE.data(), ex.data(), ey.data()));
  CHECK(a, rg_update_extensives(a, dt));
  CHECK(a, rg_update_cells(a));

  // 3. the same step through the device loop on context b
  double dt_b = 0;
  CHECK(b, rg_advance(b, 1, 1, &dt_b));
  std::vector<double> ca(4 * static_cast<size_t>(N)), cb(4 * static_cast<size_t>(N)),
      xa(4 * static_cast<size_t>(N)), xb(4 * static_cast<size_t>(N));
  auto col = [&](std::vector<double>& v, int k) { return v.data() + static_cast<size_t>(k) * N; };
  CHECK(a, rg_get_cells(a, col(ca, 0), col(ca, 1), col(ca, 2), col(ca, 3)));
  CHECK(b, rg_get_cells(b, col(cb, 0), col(cb, 1), col(cb, 2), col(cb, 3)));
  CHECK(a, rg_get_extensives(a, col(xa, 0), col(xa, 1), col(xa, 2), col(xa, 3)));
  CHECK(b, rg_get_extensives(b, col(xb, 0), col(xb, 1), col(xb, 2), col(xb, 3)));
  const bool same = dt == dt_b &&
                    std::memcmp(ca.data(), cb.data(), ca.size() * sizeof(double)) == 0 &&
                    std::memcmp(xa.data(), xb.data(), xa.size() * sizeof(double)) == 0;

  // 4. Heun steps, device-resident (one warm-up step, then the timed ones)
  CHECK(b, rg_advance(b, 2, 1, nullptr));
  const auto t0 = std::chrono::steady_clock::now();
  CHECK(b, rg_advance(b, 2, steps, nullptr));
  const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  double mass = 0;
  CHECK(b, rg_get_extensives(b, col(xb, 0), nullptr, nullptr, nullptr));
  for (int i = 0; i < N; ++i) mass += xb[i];
  rg_destroy(a);
  rg_destroy(b);

  // 5. TimeAdvance2Heun in the adapters' call order, host arrays in between
  if (out_path && plugin_steps > 0) {
    rg_ctx* h = nullptr;
    CHECK(nullptr, rg_create(&prm, &h));
    CHECK(h, rg_set_mesh(h, &mesh));
    std::vector<double> cells(4 * static_cast<size_t>(N)), ext(4 * static_cast<size_t>(N)),
        mid(4 * static_cast<size_t>(N)), midc(4 * static_cast<size_t>(N)),
        flux(4 * static_cast<size_t>(F)), dts;
    auto row = [](std::vector<double>& v, int k, int n) { return v.data() + static_cast<size_t>(k) * n; };
    std::memcpy(row(cells, 0, N), d.data(), sizeof(double) * N);
    std::memcpy(row(cells, 1, N), p.data(), sizeof(double) * N);
    std::memcpy(row(cells, 2, N), vx.data(), sizeof(double) * N);
    std::memcpy(row(cells, 3, N), vy.data(), sizeof(double) * N);
    const std::vector<double> cells0 = cells;
    // hdsim's constructor: extensives from the cells (hdsim2d.cpp:41-66)
    CHECK(h, rg_set_cells(h, row(cells, 0, N), row(cells, 1, N), row(cells, 2, N), row(cells, 3, N)));
    CHECK(h, rg_init_extensives(h));
    CHECK(h, rg_get_extensives(h, row(ext, 0, N), row(ext, 1, N), row(ext, 2, N), row(ext, 3, N)));
    auto set_cells = [&](std::vector<double>& c) {
      return rg_set_cells(h, row(c, 0, N), row(c, 1, N), row(c, 2, N), row(c, 3, N));
    };
    auto zero_force = [&](std::vector<double>& x, double dt_) {  // zero_force.cpp:3-18
      for (int i = 0; i < N; ++i) {
        row(x, 0, N)[i] += dt_ * 0.0;
        row(x, 2, N)[i] += 0.0 * dt_;
        row(x, 3, N)[i] += 0.0 * dt_;
        row(x, 1, N)[i] += dt_ * 0.0;
      }
    };
    // eu_: fluxes and extensives in, extensives out
    auto update_extensives = [&](std::vector<double>& x, double dt_) {
      int rc = rg_set_extensives(h, row(x, 0, N), row(x, 1, N), row(x, 2, N), row(x, 3, N));
      if (rc == RG_OK)
        rc = rg_set_fluxes(h, row(flux, 0, F), row(flux, 1, F), row(flux, 2, F), row(flux, 3, F));
      if (rc == RG_OK) rc = rg_update_extensives(h, dt_);
      if (rc == RG_OK)
        rc = rg_get_extensives(h, row(x, 0, N), row(x, 1, N), row(x, 2, N), row(x, 3, N));
      return rc;
    };
    // cu_: extensives in, cells out
    auto update_cells = [&](std::vector<double>& x, std::vector<double>& c) {
      int rc = rg_set_extensives(h, row(x, 0, N), row(x, 1, N), row(x, 2, N), row(x, 3, N));
      if (rc == RG_OK) rc = rg_update_cells(h);
      if (rc == RG_OK) rc = rg_get_cells(h, row(c, 0, N), row(c, 1, N), row(c, 2, N), row(c, 3, N));
      return rc;
    };
    auto flux_calc = [&](std::vector<double>& c) {
      int rc = set_cells(c);
      if (rc == RG_OK) rc = rg_set_face_velocity(h, nullptr, nullptr);  // static mesh
      if (rc == RG_OK)
        rc = rg_flux(h, row(flux, 0, F), row(flux, 1, F), row(flux, 2, F), row(flux, 3, F));
      return rc;
    };
    for (int s = 0; s < plugin_steps; ++s) {
      double dt_h = 0;
      CHECK(h, set_cells(cells));
      CHECK(h, rg_time_step(h, &dt_h));  // tsf_, twice around ApplyFix (no motion here)
      CHECK(h, rg_time_step(h, &dt_h));
      CHECK(h, flux_calc(cells));        // mid_fluxes
      mid = ext;
      zero_force(mid, dt_h);             // ExternalForceContribution
      CHECK(h, update_extensives(mid, dt_h));
      CHECK(h, update_cells(mid, midc));  // mid_cells on the (unchanged) mesh
      CHECK(h, flux_calc(midc));          // fluxes of the mid cells
      zero_force(mid, dt_h);
      CHECK(h, update_extensives(mid, dt_h));
      for (int i = 0; i < N; ++i) {       // average_extensive (hdsim2d.cpp:311-320)
        row(ext, 0, N)[i] = 0.5 * (row(mid, 0, N)[i] + row(ext, 0, N)[i]);
        row(ext, 1, N)[i] = 0.5 * (row(mid, 1, N)[i] + row(ext, 1, N)[i]);
        row(ext, 2, N)[i] = (row(mid, 2, N)[i] + row(ext, 2, N)[i]) * 0.5;
        row(ext, 3, N)[i] = (row(mid, 3, N)[i] + row(ext, 3, N)[i]) * 0.5;
      }
      CHECK(h, update_cells(ext, cells));
      dts.push_back(dt_h);
    }
    rg_destroy(h);
    FILE* fo = std::fopen(out_path, "wb");
    if (!fo) return fail(nullptr, "fopen", -1);
    const bool ok = std::fwrite(cells0.data(), sizeof(double), cells0.size(), fo) == cells0.size() &&
                    std::fwrite(dts.data(), sizeof(double), dts.size(), fo) == dts.size() &&
                    std::fwrite(cells.data(), sizeof(double), cells.size(), fo) == cells.size() &&
                    std::fwrite(ext.data(), sizeof(double), ext.size(), fo) == ext.size();
    std::fclose(fo);
    if (!ok) return fail(nullptr, "fwrite", -1);
  }

  std::printf(
      "{\"cells\": %d, \"faces\": %d, \"plugin_step_equals_device_step\": %s, \"dt\": %.17g, "
      "\"heun_steps\": %d, \"ms_per_step\": %.4f, \"mcell_updates_per_s\": %.1f, "
      "\"mass\": %.17g}\n",
      N, F, same ? "true" : "false", dt, steps, 1e3 * s / steps, 1e-6 * N * steps / s, mass);
  return same ? 0 : 3;
}
