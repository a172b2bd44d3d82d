// mhd.cpp -- hydflux_mod over the MI355X library: each step function of
// src_cpp_kokkos_host/mhd.cpp forwards to its C-ABI counterpart.
#include "mhd.hpp"

#include <cstdio>
#include <cstdlib>

#include "boundary.hpp"

namespace hydflux_mod {

GridArray<double> G;
FieldArray<double> U;
FieldArray<double> Fx, Fy, Fz;
FieldArray<double> P;
double chg = 0.0;
double gam = 1.4;  // src_cpp_kokkos_host/mhd.cpp:27

namespace {
mhd_ctx* g_ctx = nullptr;
bool g_uploaded = false;

}  // namespace

// the host writes G, P and U before the loop (GenerateGrid / GenerateProblem); the first
// step function after allocation moves them to the device
void ensure_device_state() {
  if (g_uploaded) return;
  G.h2d();
  P.h2d();
  U.h2d();
  g_uploaded = true;
}

mhd_ctx* device_context() {
  if (!g_ctx) {
    std::fprintf(stderr, "hydflux_mod: AllocateHydroVariables has not been called\n");
    std::abort();
  }
  return g_ctx;
}

void check(int status, const char* what) {
  if (status != MHD_OK) {
    std::fprintf(stderr, "%s failed: %s (%d): %s\n", what, mhd_error_string(status), status, mhd_last_error());
    std::abort();
  }
}

template <>
void FieldArray<double>::h2d() {
  switch (kind) {
    case FieldKind::prim: check(mhd_upload_prim(device_context(), data), "mhd_upload_prim"); break;
    case FieldKind::cons: check(mhd_upload_cons(device_context(), data), "mhd_upload_cons"); break;
    default: break;  // fluxes are device scratch
  }
}

template <>
void FieldArray<double>::d2h() {
  switch (kind) {
    case FieldKind::prim: check(mhd_download_prim(device_context(), data), "mhd_download_prim"); break;
    case FieldKind::cons: check(mhd_download_cons(device_context(), data), "mhd_download_cons"); break;
    case FieldKind::flux1: check(mhd_download_flux(device_context(), 1, data), "mhd_download_flux"); break;
    case FieldKind::flux2: check(mhd_download_flux(device_context(), 2, data), "mhd_download_flux"); break;
    case FieldKind::flux3: check(mhd_download_flux(device_context(), 3, data), "mhd_download_flux"); break;
  }
}

void AllocateHydroVariables(GridArray<double>& G_, FieldArray<double>& U_, FieldArray<double>& Fx_,
                            FieldArray<double>& Fy_, FieldArray<double>& Fz_, FieldArray<double>& P_) {
  using namespace resolution_mod;
  mhd_config cfg{};
  cfg.nx = ngrid1; cfg.ny = ngrid2; cfg.nz = ngrid3;
  cfg.bc[0] = boundary_mod::boundary_xin;  cfg.bc[1] = boundary_mod::boundary_xout;
  cfg.bc[2] = boundary_mod::boundary_yin;  cfg.bc[3] = boundary_mod::boundary_yout;
  cfg.bc[4] = boundary_mod::boundary_zin;  cfg.bc[5] = boundary_mod::boundary_zout;
  cfg.xmin[0] = x1min; cfg.xmax[0] = x1max;
  cfg.xmin[1] = x2min; cfg.xmax[1] = x2max;
  cfg.xmin[2] = x3min; cfg.xmax[2] = x3max;
  cfg.gam = gam;
  cfg.cfl = 0.05;      // src_cpp_kokkos_host/mhd.cpp:1560
  cfg.alphabp = 0.1;   // src_cpp_kokkos_host/mhd.cpp:1614
  cfg.energy_flux = 0; // the energy flux as src_cpp_kokkos_host/mhd.cpp:1049-1050 computes it
  cfg.grid_shift = 1;
  cfg.precision = 0;
  cfg.device = 0;
  if (const char* d = std::getenv("MHD_DEVICE")) cfg.device = std::atoi(d);
  check(mhd_create(&cfg, &g_ctx), "mhd_create");
  g_uploaded = false;
  G_.allocate(ktot, jtot, itot);
  U_.allocate(mconsv, ktot, jtot, itot, FieldKind::cons);
  P_.allocate(nprim, ktot, jtot, itot, FieldKind::prim);
  Fx_.allocate(mconsv, ktot, jtot, itot, FieldKind::flux1);
  Fy_.allocate(mconsv, ktot, jtot, itot, FieldKind::flux2);
  Fz_.allocate(mconsv, ktot, jtot, itot, FieldKind::flux3);
}

void DeallocateHydroVariables(GridArray<double>& G_, FieldArray<double>& U_, FieldArray<double>& Fx_,
                              FieldArray<double>& Fy_, FieldArray<double>& Fz_, FieldArray<double>& P_) {
  G_.deallocate(); U_.deallocate(); Fx_.deallocate(); Fy_.deallocate(); Fz_.deallocate(); P_.deallocate();
  if (g_ctx) mhd_destroy(g_ctx);
  g_ctx = nullptr;
}

void ControlTimestep(const GridArray<double>&) {
  ensure_device_state();
  double dt = 0.0;
  check(mhd_control_timestep(device_context(), &dt), "mhd_control_timestep");
  resolution_mod::dt = dt;
}

void EvaluateCh() {
  ensure_device_state();
  check(mhd_evaluate_ch(device_context(), &chg), "mhd_evaluate_ch");
}

void GetNumericalFluxD(const int dir, const GridArray<double>&, const FieldArray<double>&, FieldArray<double>&) {
  ensure_device_state();
  check(mhd_numerical_flux(device_context(), dir, chg), "mhd_numerical_flux");
}

void UpdateConservU(const GridArray<double>&, const FieldArray<double>&, const FieldArray<double>&,
                    const FieldArray<double>&, FieldArray<double>&) {
  ensure_device_state();
  check(mhd_update_conserv(device_context(), resolution_mod::dt), "mhd_update_conserv");
}

void DampPsi(const GridArray<double>&, FieldArray<double>&) {
  ensure_device_state();
  check(mhd_damp_psi(device_context(), resolution_mod::dt, chg), "mhd_damp_psi");
}

void UpdatePrimitvP(const FieldArray<double>&, FieldArray<double>&) {
  ensure_device_state();
  check(mhd_update_primitive(device_context()), "mhd_update_primitive");
}

int AdvanceSteps(int n, double tmax) {
  ensure_device_state();
  std::vector<double> dts(n > 0 ? n : 1);
  int done = 0;
  check(mhd_run(device_context(), n, tmax, 0, &resolution_mod::time_sim, dts.data(), &done), "mhd_run");
  if (done > 0) resolution_mod::dt = dts[done - 1];
  return done;
}

}  // namespace hydflux_mod
