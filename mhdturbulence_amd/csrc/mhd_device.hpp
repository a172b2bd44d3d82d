// mhd_device.hpp -- device data layout of one box and the host-side launch wrappers.
//
// HBM layout (one box): structure of arrays, variable-major,
//   element (n, k, j, i) at  base + n*vstride + (k*jtot + j)*pitch + i + IOFF
// with the reference's index ranges (i in [0, itot), is = 2, itot = nx+5; same for j, k).
// IOFF = 14 puts the first interior cell (i = 2) on a 128-byte boundary, and pitch is a
// multiple of 16 doubles, so every interior row starts on a 128-byte line.  Rows of one
// variable are contiguous in i, so a wavefront reading 64 neighbouring i of one variable
// issues fully coalesced 512-byte loads.
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

namespace mhd {

constexpr int NGH = 2;     // ghost layers (KH/config.hpp:60)
constexpr int NPRIM = 12;  // rho v1 v2 v3 e p cs B1 B2 B3 psi X   (KH/mhd.hpp:249-253)
constexpr int NCONS = 10;  // rho m1 m2 m3 E B1 B2 B3 psi rhoX     (KH/mhd.hpp:264-268)
constexpr int IOFF = 14;

struct Geom {
  int nx, ny, nz;        // interior cells
  int itot, jtot, ktot;  // allocated extent per axis (n + 2*ngh + 1), as KH/config.hpp:56-58
  int pitch;             // doubles per (k, j) row
  size_t vstride;        // doubles per variable
  __host__ __device__ int cell(int k, int j, int i) const { return (k * jtot + j) * pitch + i + IOFF; }
};

// Boundary codes (KH/config.hpp:29-32)
enum { BC_PERIODIC = 1, BC_REFLECT = 2, BC_OUTFLOW = 3 };

// Per-box device scalars of the device-resident step loop.  red[] is a three-slot ring of
// (dt_min, c_h max) accumulators stored as IEEE bit patterns: for non-negative doubles the
// unsigned-integer order is the numeric order, so integer atomicMin/atomicMax reduce them
// exactly and independently of the order in which blocks arrive.  Step `iter` reads slot
// iter % 3.  Kernels take every per-step quantity from here, so the launch arguments of a
// step never change and a step can be captured once in a graph and replayed.
struct StepScalars {
  unsigned long long red[3][2];
  unsigned long long host_acc[2];  // accumulator pair of the one-call reference functions
  double time;
  double tmax;       // stop once time > tmax (<= 0: never)
  int iter;          // loop iterations issued since the last reset
  int steps_done;    // steps actually applied
  int stopped;
  int status;        // bit 0: a cell with rho <= 0 or a non-finite dt / c_h bound was met (sticky)
  int dt_log_cap;
  int halt;          // the step just applied left such a cell: the loop stops after it
};

constexpr unsigned long long DT_INIT_BITS = 0x7FF0000000000000ULL;  // +inf
constexpr unsigned long long CH_INIT_BITS = 0ULL;                   // +0.0

// A cell the step cannot continue from: rho <= 0, or a CFL bound / fast speed that is not
// finite and positive (a NaN anywhere in the cell, or e < 0 whose sound speed is a NaN).
// Every comparison is false for a NaN, so each test is written to fail on one.
__device__ __forceinline__ bool bad_cell(double rho, double cdt, double cch) {
  return !(rho > 0.0) || !(cdt > 0.0) || !(cch <= 1.7976931348623157e308);
}

// ---- launch wrappers (mhd_kernels.hip) ----
void launch_from_host_layout(const Geom& g, int nv, const double* src, double* dst, hipStream_t s);
void launch_to_host_layout(const Geom& g, int nv, const double* src, double* dst, hipStream_t s);
void launch_generate_grid(const Geom& g, const double* xmin, const double* xmax, int shift, double* x1a,
                          double* x1b, double* x2a, double* x2b, double* x3a, double* x3b, hipStream_t s);
void launch_problem_kh(const Geom& g, double gam, double vy_amp, double rrv, const double* xmin,
                       const double* xmax, const double* x1b, const double* x2b, const double* x3b, double* P,
                       double* U, hipStream_t s);
void launch_problem_shear(const Geom& g, double gam, const double* ic /* rho0,p0,v0,a,bx0,amp,sigma */,
                          uint64_t seed, const double* x2b, double* P, double* U, hipStream_t s);

// n doubles at p set to 0 by a kernel (a fill node, not a memset, if the stream is captured)
void launch_zero(double* p, size_t n, hipStream_t s);
// the interior cells of nv variables of src copied into dst, whose ghost cells are left as
// they are (the fused loop hands its state back this way: the reference's loop never
// writes the ghost cells of U)
void launch_copy_interior(const Geom& g, const double* src, double* dst, int nv, hipStream_t s);
// sc->host_acc := (+inf, 0); or, ring mode, all three slots
void launch_reset_scalars(StepScalars* sc, int ring, hipStream_t s);
// min-dt / max-c_h reduction over the interior cells of P.  ring=0: into sc->host_acc;
// ring=1: into sc->red[sc->iter % 3], the bounds of the step about to be taken; a bad_cell
// then also sets sc->stopped, so that step and every later one of the loop are skipped.
// Either way a bad_cell sets status bit 0.
void launch_reduce_prim(const Geom& g, const double* P, const double* x1a, const double* x2a, const double* x3a,
                        StepScalars* sc, int ring, hipStream_t s);
// ghost fill of an nv-variable array (P: nv = 12, U: nv = 10)
void launch_boundary(const Geom& g, const int* bc, double* A, int nv, hipStream_t s);
// chg: value, or (sc != nullptr) taken from sc->red[sc->iter % 3][1]
void launch_flux(const Geom& g, int dir, int eflux, int mixed, const double* P, double* F, double chg,
                 const StepScalars* sc, hipStream_t s);
// U -= dt*(dFx/dx + dFy/dy + dFz/dz).  sc != nullptr: dt = cfl*dtmin and c_h from the ring, the
// psi damping is applied in the same pass, and the update is skipped once the loop stopped.
// save != nullptr: the state before the update is copied there (the first Runge-Kutta stage);
// avg != nullptr: the result is averaged with avg, (avg + U) / 2 (the second stage).
void launch_update(const Geom& g, const double* x1a, const double* x2a, const double* x3a, const double* Fx,
                   const double* Fy, const double* Fz, double* U, double dt, double cfl, double alphabp,
                   const StepScalars* sc, hipStream_t s, double* save = nullptr, const double* avg = nullptr);
void launch_damp(const Geom& g, const double* x1a, const double* x2a, const double* x3a, double* U, double dt,
                 double chg, double alphabp, hipStream_t s);
void launch_update_prim(const Geom& g, double gam, const double* U, double* P, const StepScalars* sc,
                        hipStream_t s);
// time += dt, log dt, iter += 1, reset the ring slot the next iteration accumulates into
// (iter + ahead) % 3, stop once time > tmax
void launch_step_end(StepScalars* sc, double cfl, double* dt_log, int ahead, hipStream_t s);
// per-k-plane partial sums: partial[4*kk + {0,1,2,3}] = (vol, mix, vy^2, max|divB|)
void launch_diagnostics(const Geom& g, const double* P, const double* x1a, const double* x1b,
                        const double* x2a, const double* x2b, const double* x3a, const double* x3b,
                        double* partial, hipStream_t s);

}  // namespace mhd
