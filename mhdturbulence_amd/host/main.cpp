// main.cpp -- Kelvin-Helmholtz driver with the call sequence of src_cpp_kokkos_host/main.cpp
// (GenerateGrid, GenerateProblem, then the step loop at main.cpp:228-267), linked against
// the drop-in hydflux_mod / boundary_mod layer.
//
//   mhd_kh_main [--steps N] [--fused | --staged] [--dump FILE] [--tprof]
//
// --steps N  loop-iteration cap (the reference breaks after step 150)
// --fused    run the loop body on the device through AdvanceSteps (dt stays in HBM)
// --staged   the reference loop with every step function on its own staged kernels
//            (default: the same loop, recorded and run as the fused kernels)
// --peek N   on step N download Fx right after GetNumericalFluxD(1) (an interruption of
//            the recorded step: that step runs out on the staged kernels)
// --dump F   write the final primitive array (host layout [k][j][n][i], f64) to F
// --tprof    append t-prof.csv rows (time, mix, sqrt(<vy^2>), A exp(Gamma t)) every step
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "boundary.hpp"
#include "config.hpp"
#include "mhd.hpp"

using namespace resolution_mod;
using namespace hydflux_mod;
using namespace boundary_mod;

// face and centre coordinates, KH/main.cpp:26-65 with one tile per axis
static void GenerateGrid(GridArray<double>& g) {
  const double d1 = (x1max - x1min) / double(ngrid1);
  const double d2 = (x2max - x2min) / double(ngrid2);
  const double d3 = (x3max - x3min) / double(ngrid3);
  for (int i = 0; i <= ie + ngh + 1; ++i) g.x1a(i) = d1 * (i - (ngh + 1)) + x1min;
  for (int i = 0; i <= ie + ngh; ++i) g.x1b(i) = 0.5 * (g.x1a(i + 1) + g.x1a(i));
  for (int j = 0; j <= je + ngh + 1; ++j) g.x2a(j) = d2 * (j - (ngh + 1)) + x2min;
  for (int j = 0; j <= je + ngh; ++j) g.x2b(j) = 0.5 * (g.x2a(j + 1) + g.x2a(j));
  for (int k = 0; k <= ke + ngh + 1; ++k) g.x3a(k) = d3 * (k - (ngh + 1)) + x3min;
  for (int k = 0; k <= ke + ngh; ++k) g.x3b(k) = 0.5 * (g.x3a(k + 1) + g.x3a(k));
}

// double shear layer of KH/main.cpp:66-184 (vy amplitude 1e-3, harmonic vx kick 1e-2)
static void GenerateProblem(const GridArray<double>& g, FieldArray<double>& p, FieldArray<double>& u) {
  const double pi = std::acos(-1.0);
  const double dv = 2.0, wid = 0.05, sig = 0.2, rrv = 1.0e-2;
  chg = 0.0;
  for (int k = 0; k <= ke + ngh; ++k)
    for (int j = 0; j <= je + ngh; ++j)
      for (int i = 0; i <= ie + ngh; ++i) {
        const double x = g.x1b(i), y = g.x2b(j);
        const double tp = std::tanh((y + 0.5) / wid), tm = std::tanh((y - 0.5) / wid);
        p(nden, k, j, i) = 1.0;
        p(npre, k, j, i) = 1.0;
        p(nve1, k, j, i) = 0.5 * dv * (tp - tm - 1.0);
        p(nve2, k, j, i) = 1.0e-3 * std::sin(2.0 * pi * x) *
                           (std::exp(-(y + 0.5) * (y + 0.5) / (sig * sig)) + std::exp(-(y - 0.5) * (y - 0.5) / (sig * sig)));
        p(nve3, k, j, i) = 0.0;
        p(nbm1, k, j, i) = p(nbm2, k, j, i) = p(nbm3, k, j, i) = p(nbps, k, j, i) = 0.0;
        p(nene, k, j, i) = p(npre, k, j, i) / (gam - 1.0);
        p(ncsp, k, j, i) = std::sqrt(gam * p(npre, k, j, i) / p(nden, k, j, i));
        p(nst, k, j, i) = 0.5 * (tp - tm);
      }
  for (int k = ks; k <= ke; ++k)
    for (int j = js; j <= je; ++j) {
      const double kick = dv * rrv * std::cos(6 * 2 * pi * (g.x2b(j) - x2min) / (x2max - x2min)) *
                          std::cos(6 * 2 * pi * (g.x3b(k) - x3min) / (x3max - x3min));
      for (int i = is; i <= ie; ++i) p(nve1, k, j, i) += kick;
    }
  for (int k = 0; k <= ke + ngh; ++k)
    for (int j = 0; j <= je + ngh; ++j)
      for (int i = 0; i <= ie + ngh; ++i) {
        const double rho = p(nden, k, j, i);
        const double v1 = p(nve1, k, j, i), v2 = p(nve2, k, j, i), v3 = p(nve3, k, j, i);
        const double b1 = p(nbm1, k, j, i), b2 = p(nbm2, k, j, i), b3 = p(nbm3, k, j, i);
        u(mden, k, j, i) = rho;
        u(mrv1, k, j, i) = rho * v1;
        u(mrv2, k, j, i) = rho * v2;
        u(mrv3, k, j, i) = rho * v3;
        u(meto, k, j, i) = rho * p(nene, k, j, i) + 0.5 * rho * (v1 * v1 + v2 * v2 + v3 * v3) +
                           0.5 * (b1 * b1 + b2 * b2 + b3 * b3);
        u(mbm1, k, j, i) = b1;
        u(mbm2, k, j, i) = b2;
        u(mbm3, k, j, i) = b3;
        u(mbps, k, j, i) = p(nbps, k, j, i);
        u(mst, k, j, i) = rho * p(nst, k, j, i);
      }
}

// t-prof row of RealTimeAnalysis (src_cpp_omp_device/main.cpp:23-69), sums on the device
static void RealTimeAnalysis(FILE* fp) {
  double d[4];
  check(mhd_diagnostics(device_context(), d), "mhd_diagnostics");
  const double amp = 1.2e-3, gamma = 1.49;
  std::fprintf(fp, " %24.16E %24.16E %24.16E %24.16E\n", time_sim, d[1], d[2], amp * std::exp(gamma * time_sim));
  std::fflush(fp);
}

int main(int argc, char** argv) {
  int steps = 150, peek = -1;
  bool fused = false, tprof = false;
  const char* dump = nullptr;
  for (int a = 1; a < argc; ++a) {
    if (!std::strcmp(argv[a], "--steps") && a + 1 < argc) steps = std::atoi(argv[++a]);
    else if (!std::strcmp(argv[a], "--fused")) fused = true;
    else if (!std::strcmp(argv[a], "--staged")) SetStepPath(StepPath::staged);
    else if (!std::strcmp(argv[a], "--tprof")) tprof = true;
    else if (!std::strcmp(argv[a], "--peek") && a + 1 < argc) peek = std::atoi(argv[++a]);
    else if (!std::strcmp(argv[a], "--dump") && a + 1 < argc) dump = argv[++a];
  }
  std::printf("setup grids and fields\n");
  AllocateHydroVariables(G, U, Fx, Fy, Fz, P);
  AllocateBoundaryVariables(Bs, Br);
  std::printf("grid size for x y z = %i %i %i\n", ngrid1, ngrid2, ngrid3);
  GenerateGrid(G);
  GenerateProblem(G, P, U);
  FILE* fpt = nullptr;
  if (tprof) {
    fpt = std::fopen("t-prof.csv", "w");
    std::fprintf(fpt, "# Gamma= %24.16E\n# 1:time 2:mix 3:v_y 4:exp(2Gamma*t)\n", 1.49);
  }
  std::printf("entering main loop\n");
  FILE* fpdt = std::fopen("dt.log", "w");
  const auto t0 = std::chrono::high_resolution_clock::now();
  int step = 0;
  if (fused) {
    ensure_device_state();
    ControlTimestep(G);  // the first dt, as the loop's first call computes it
    if (fpdt) std::fprintf(fpdt, "step=0, time=%.17g, dt=%.17g\n", time_sim, dt);
    step = AdvanceSteps(steps, time_max) - 1;
  } else {
    for (step = 0; step < steps; ++step) {
      ControlTimestep(G);
      if (step % 10 == 0) std::printf("step=%i/%i time=%e dt=%e\n", step, stepmax, time_sim, dt);
      if (fpdt && step < 100) std::fprintf(fpdt, "step=%d, time=%.17g, dt=%.17g\n", step, time_sim, dt);
      SetBoundaryCondition(P, Bs, Br);
      EvaluateCh();
      GetNumericalFluxD(1, G, P, Fx);
      if (step == peek) {
        Fx.d2h();
        double s1 = 0.0;
        for (int j = js; j <= je; ++j) s1 += Fx(mden, ks, j, is + 1);
        std::printf("peek step %d: sum of the x mass flux on a line %.17g\n", step, s1);
      }
      GetNumericalFluxD(2, G, P, Fy);
      GetNumericalFluxD(3, G, P, Fz);
      UpdateConservU(G, Fx, Fy, Fz, U);
      DampPsi(G, U);
      UpdatePrimitvP(U, P);
      time_sim += dt;
      if (fpt) RealTimeAnalysis(fpt);
      if (time_sim > time_max) break;
    }
    if (step == steps) step = steps - 1;
  }
  const auto t1 = std::chrono::high_resolution_clock::now();
  if (fpdt) std::fclose(fpdt);
  const double el = std::chrono::duration<double>(t1 - t0).count();
  std::printf("exiting main loop time=%e, step=%i\n", time_sim, step);
  std::printf("sim time [s]: %e\n", el);
  std::printf("time/count/cell : %e\n", el / (double(ngrid1) * ngrid2 * ngrid3) / (step + 1));
  if (fpt) {
    RealTimeAnalysis(fpt);
    std::fclose(fpt);
  }
  if (dump) {
    P.d2h();
    FILE* f = std::fopen(dump, "wb");
    if (!f || std::fwrite(P.data, sizeof(double), P.size, f) != P.size) {
      std::fprintf(stderr, "cannot write %s\n", dump);
      return 1;
    }
    std::fclose(f);
  }
  std::printf("program has been finished\n");
  DeallocateHydroVariables(G, U, Fx, Fy, Fz, P);
  DeallocateBoundaryVariables(Bs, Br);
  return 0;
}
