// mhd.cpp -- hydflux_mod over the MI355X library.  Each step function of
// src_cpp_kokkos_host/mhd.cpp forwards to its C-ABI counterpart (the staged path), or, while
// the reference loop's calls come in their order, is recorded and the whole loop body runs
// as the fused kernels when UpdatePrimitvP closes it (mhd_fused_begin_step / mhd_fused_step).
#include "mhd.hpp"

#include <cstdio>
#include <cstdlib>
#include <cstring>

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

StepPath initial_path() {
  const char* e = std::getenv("MHD_DROPIN_STAGED");
  return (e && std::strcmp(e, "0") != 0) ? StepPath::staged : StepPath::fused;
}
StepPath g_path = initial_path();

// The calls of the reference loop body in their order (src_cpp_kokkos_host/main.cpp:229-257).
enum Call { kNone, kCT, kBC, kEC, kF1, kF2, kF3, kUCU, kDP, kUPP };

// The step being recorded: ControlTimestep took dt and c_h from the fused loop, `last` is
// the last call recorded since.  dt / chg are the values handed out, which the later calls
// must still see in the globals for the fused step to be the step they describe.
struct Recording {
  bool active = false;
  int last = kNone;
  double dt = 0.0, chg = 0.0;
} g_rec;

// a recorded call continues the step: it is the next call of the loop body and the step
// scalars are unchanged (dt since ControlTimestep, chg since EvaluateCh)
bool continues(int call) {
  return g_rec.active && g_rec.last == call - 1 && resolution_mod::dt == g_rec.dt &&
         (call <= kEC || chg == g_rec.chg);
}

// aborts on a cell with rho <= 0 or a non-finite CFL bound (status bit 0)
void check_status(int st, const char* where) {
  if (st & 1) {
    std::fprintf(stderr, "%s: the state holds a cell with rho <= 0 or a non-finite CFL bound (status %d)\n", where,
                 st);
    std::abort();
  }
}

}  // namespace

void SetStepPath(StepPath p) {
  FlushRecordedStep();
  g_path = p;
}
StepPath GetStepPath() { return g_path; }

// the host writes G, P and U before the loop (GenerateGrid / GenerateProblem); the first
// step function after allocation moves them to the device
void ensure_device_state() {
  if (g_uploaded) return;
  g_uploaded = true;
  G.h2d();
  P.h2d();
  U.h2d();
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

// The calls recorded after ControlTimestep, run on the staged kernels in their order; the
// rest of the step then follows them there.  (The staged calls first bring P and U back
// from the fused loop's buffers.)
void FlushRecordedStep() {
  if (!g_rec.active) return;
  g_rec.active = false;
  mhd_ctx* c = device_context();
  for (int k = kBC; k <= g_rec.last; ++k) {
    switch (k) {
      case kBC: check(mhd_set_boundary(c), "mhd_set_boundary"); break;
      case kF1:
      case kF2:
      case kF3: check(mhd_numerical_flux(c, k - kF1 + 1, g_rec.chg), "mhd_numerical_flux"); break;
      case kUCU: check(mhd_update_conserv(c, g_rec.dt), "mhd_update_conserv"); break;
      case kDP: check(mhd_damp_psi(c, g_rec.dt, g_rec.chg), "mhd_damp_psi"); break;
      default: break;  // EvaluateCh: chg is already the value the staged call would return
    }
  }
}

template <>
void GridArray<double>::h2d() {
  FlushRecordedStep();
  check(mhd_set_grid(device_context(), a1.data(), b1.data(), a2.data(), b2.data(), a3.data(), b3.data()),
        "mhd_set_grid");
}

template <>
void FieldArray<double>::h2d() {
  if (kind == FieldKind::host) return;
  FlushRecordedStep();
  switch (kind) {
    case FieldKind::prim: check(mhd_upload_prim(device_context(), data), "mhd_upload_prim"); break;
    case FieldKind::cons: check(mhd_upload_cons(device_context(), data), "mhd_upload_cons"); break;
    default: break;  // fluxes are device scratch
  }
}

template <>
void FieldArray<double>::d2h() {
  if (kind == FieldKind::host) return;
  FlushRecordedStep();
  switch (kind) {
    case FieldKind::prim: check(mhd_download_prim(device_context(), data), "mhd_download_prim"); break;
    case FieldKind::cons: check(mhd_download_cons(device_context(), data), "mhd_download_cons"); break;
    case FieldKind::flux1: check(mhd_download_flux(device_context(), 1, data), "mhd_download_flux"); break;
    case FieldKind::flux2: check(mhd_download_flux(device_context(), 2, data), "mhd_download_flux"); break;
    case FieldKind::flux3: check(mhd_download_flux(device_context(), 3, data), "mhd_download_flux"); break;
    default: break;
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
  g_rec = Recording{};
  G_.allocate(ktot, jtot, itot);
  U_.allocate(mconsv, ktot, jtot, itot);
  P_.allocate(nprim, ktot, jtot, itot);
  Fx_.allocate(mconsv, ktot, jtot, itot);
  Fy_.allocate(mconsv, ktot, jtot, itot);
  Fz_.allocate(mconsv, ktot, jtot, itot);
  U_.kind = FieldKind::cons;
  P_.kind = FieldKind::prim;
  Fx_.kind = FieldKind::flux1;
  Fy_.kind = FieldKind::flux2;
  Fz_.kind = FieldKind::flux3;
}

void DeallocateHydroVariables(GridArray<double>& G_, FieldArray<double>& U_, FieldArray<double>& Fx_,
                              FieldArray<double>& Fy_, FieldArray<double>& Fz_, FieldArray<double>& P_) {
  g_rec = Recording{};
  G_.deallocate(); U_.deallocate(); Fx_.deallocate(); Fy_.deallocate(); Fz_.deallocate(); P_.deallocate();
  if (g_ctx) mhd_destroy(g_ctx);
  g_ctx = nullptr;
}

void ControlTimestep(const GridArray<double>&) {
  ensure_device_state();
  FlushRecordedStep();  // an unfinished step runs out staged first
  double dt = 0.0;
  int st = 0;
  if (g_path == StepPath::fused) {
    double ch = 0.0;
    check(mhd_fused_begin_step(device_context(), &dt, &ch, &st), "mhd_fused_begin_step");
    check_status(st, "ControlTimestep");
    g_rec = Recording{true, kCT, dt, ch};  // EvaluateCh will hand out ch
  } else {
    check(mhd_control_timestep(device_context(), &dt), "mhd_control_timestep");
    check(mhd_status(device_context(), &st), "mhd_status");
    check_status(st, "ControlTimestep");
  }
  resolution_mod::dt = dt;
}

bool RecordSetBoundaryCondition() {
  if (!continues(kBC)) return false;
  g_rec.last = kBC;
  return true;
}

void EvaluateCh() {
  ensure_device_state();
  if (continues(kEC)) {
    g_rec.last = kEC;
    chg = g_rec.chg;
    return;
  }
  FlushRecordedStep();
  check(mhd_evaluate_ch(device_context(), &chg), "mhd_evaluate_ch");
}

void GetNumericalFluxD(const int dir, const GridArray<double>&, const FieldArray<double>&, FieldArray<double>&) {
  ensure_device_state();
  if (dir >= 1 && dir <= 3 && continues(kF1 + dir - 1)) {
    g_rec.last = kF1 + dir - 1;
    return;
  }
  FlushRecordedStep();
  check(mhd_numerical_flux(device_context(), dir, chg), "mhd_numerical_flux");
}

void GetNumericalFlux1(const GridArray<double>& G_, const FieldArray<double>& P_, FieldArray<double>& Fx_) {
  GetNumericalFluxD(1, G_, P_, Fx_);
}
void GetNumericalFlux2(const GridArray<double>& G_, const FieldArray<double>& P_, FieldArray<double>& Fy_) {
  GetNumericalFluxD(2, G_, P_, Fy_);
}
void GetNumericalFlux3(const GridArray<double>& G_, const FieldArray<double>& P_, FieldArray<double>& Fz_) {
  GetNumericalFluxD(3, G_, P_, Fz_);
}

void UpdateConservU(const GridArray<double>&, const FieldArray<double>&, const FieldArray<double>&,
                    const FieldArray<double>&, FieldArray<double>&) {
  ensure_device_state();
  if (continues(kUCU)) {
    g_rec.last = kUCU;
    return;
  }
  FlushRecordedStep();
  check(mhd_update_conserv(device_context(), resolution_mod::dt), "mhd_update_conserv");
}

void DampPsi(const GridArray<double>&, FieldArray<double>&) {
  ensure_device_state();
  if (continues(kDP)) {
    g_rec.last = kDP;
    return;
  }
  FlushRecordedStep();
  check(mhd_damp_psi(device_context(), resolution_mod::dt, chg), "mhd_damp_psi");
}

void UpdatePrimitvP(const FieldArray<double>&, FieldArray<double>&) {
  ensure_device_state();
  if (continues(kUPP)) {
    // the recorded loop body, as the fused kernels with the dt and c_h ControlTimestep gave
    g_rec.active = false;
    check(mhd_fused_step(device_context()), "mhd_fused_step");
    return;
  }
  FlushRecordedStep();
  check(mhd_update_primitive(device_context()), "mhd_update_primitive");
}

int AdvanceSteps(int n, double tmax) {
  ensure_device_state();
  FlushRecordedStep();
  std::vector<double> dts(n > 0 ? n : 1);
  int done = 0;
  check(mhd_run(device_context(), n, tmax, 0, &resolution_mod::time_sim, dts.data(), &done), "mhd_run");
  if (done > 0) resolution_mod::dt = dts[done - 1];
  return done;
}

}  // namespace hydflux_mod
