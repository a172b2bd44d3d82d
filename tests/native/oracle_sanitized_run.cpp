// Drives every entry point of the CPU oracle (oracle/mhd_oracle.h) on small boxes, for a
// build with AddressSanitizer and UndefinedBehaviorSanitizer (tests/test_oracle_sanitizers.py).
// Test infrastructure only: it checks the checker's memory accesses, index arithmetic and
// integer operations, over every boundary type, both energy fluxes and both step schemes.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "mhd_oracle.h"

static int fail(const char* what) {
  std::fprintf(stderr, "FAIL: %s\n", what);
  return 1;
}

static int run_box(int nx, int ny, int nz, const int bc[6], int eflux, int stages, int kh) {
  orc_config cfg;
  std::memset(&cfg, 0, sizeof cfg);
  cfg.nx = nx;
  cfg.ny = ny;
  cfg.nz = nz;
  for (int q = 0; q < 6; ++q) cfg.bc[q] = bc[q];
  for (int d = 0; d < 3; ++d) {
    cfg.xmin[d] = -0.5;
    cfg.xmax[d] = 0.5;
  }
  cfg.gam = kh ? 1.4 : 5.0 / 3.0;
  cfg.cfl = stages == 2 ? 0.3 : 0.05;
  cfg.alphabp = 0.1;
  cfg.energy_flux = eflux;
  cfg.grid_shift = kh;
  cfg.stages = stages;
  orc_ctx* c = orc_create(&cfg);
  if (!c) return fail("orc_create");
  orc_generate_grid(c);
  if (kh) {
    orc_generate_problem_kh(c, 1e-2, 1e-2);
  } else {
    orc_ic_params p = {1.0, 1.0, 0.5, 0.06, 0.12, 0.03, 0.1, 20261203ull};
    orc_generate_problem_shear(c, &p);
  }
  // the reference functions one by one, then whole steps
  const double dt = orc_control_timestep(c);
  orc_set_boundary(c);
  const double ch = orc_evaluate_ch(c);
  for (int d = 1; d <= 3; ++d) orc_numerical_flux(c, d, ch);
  orc_update_conserv(c, dt);
  orc_damp_psi(c, dt, ch);
  orc_update_primitive(c);
  orc_save_cons(c);
  orc_average_saved(c);
  orc_update_primitive(c);
  orc_prim_to_cons_all(c);
  std::vector<double> log(12);
  const double t = orc_run(c, 12, 0.0, log.data());
  double diag[4];
  orc_diagnostics(c, diag);
  int itot, jtot, ktot;
  orc_sizes(c, &itot, &jtot, &ktot);
  const double* P = orc_prim(c);
  const size_t n = (size_t)itot * jtot * ktot * 12;
  for (size_t q = 0; q < n; ++q)
    if (!std::isfinite(P[q])) return fail("non-finite primitive");
  orc_destroy(c);
  if (!(t > 0.0) || !(diag[0] > 0.0)) return fail("time or volume");
  return 0;
}

int main() {
  orc_set_threads(2);
  const int bcs[3][6] = {{1, 1, 2, 2, 1, 1}, {3, 3, 2, 2, 3, 3}, {2, 2, 3, 3, 1, 1}};
  int bad = 0;
  for (const auto& bc : bcs)
    for (int eflux = 0; eflux < 2; ++eflux)
      for (int stages : {1, 2}) bad |= run_box(13, 9, 7, bc, eflux, stages, eflux == 0);
  bad |= run_box(5, 4, 6, bcs[0], 1, 1, 0);
  // the single-point kernels
  double pm2[12], pm1[12], p0[12], pp1[12], r23[23], l23[23], f10[10];
  for (int v = 0; v < 12; ++v) {
    pm2[v] = 1.0 + 0.01 * v;
    pm1[v] = 1.1 + 0.02 * v;
    p0[v] = 0.9 + 0.015 * v;
    pp1[v] = 1.05 - 0.01 * v;
  }
  for (int d = 1; d <= 3; ++d) {
    orc_prim2riemann(p0, d, 1, l23);
    orc_prim2riemann(pp1, d, 1, r23);
    orc_hlld(l23, r23, f10);
    orc_face_flux(pm2, pm1, p0, pp1, d, 1, 1.5, f10);
  }
  if (!(std::isfinite(orc_vanleer(0.3, 0.1)) && orc_vanleer(0.3, -0.1) == 0.0)) bad |= fail("vanleer");
  (void)orc_splitmix64(20261203ull);
  if (bad) return 1;
  std::printf("oracle sanitized run ok\n");
  return 0;
}
