// mhd_fused.hpp -- the fused step: three kernels per loop iteration instead of eight.
//
//   xy   : x- and y-face fluxes and their divergence  acc = dFx/dx + dFy/dy
//          (a wavefront holds a 64/32/16-cell x-strip in its lanes; x neighbours come
//          through DPP wave shifts, the strip-edge faces from a per-chunk pre-pass parked
//          in LDS or, for FP32 full-wave strips, from the neighbouring wavefronts of the
//          block; it marches in y keeping a 3-row window, in LDS for FP64)
//   z    : z-face fluxes marching in z, then U_new = U - dt*(acc + dFz/dz), psi damping,
//          primitive recovery of U_new and the next iteration's dt / c_h reduction
//
// Only the 10 conserved variables stay resident (two ping-pong copies); primitives are
// recovered from U on load, as UpdatePrimitvP (src_cpp_kokkos_host/mhd.cpp:1447-1509)
// would have stored them.  No kernel writes a ghost cell: a load of a ghost row, column or
// plane reads the interior cell whose periodic / reflecting / outflow image it is (normal
// momentum negated at a reflecting wall), the values SetBoundaryCondition would store.
#pragma once

#include "mhd_device.hpp"

namespace mhd {

struct FusedArgs {
  Geom g;
  const double* x1a;
  const double* x2a;
  const double* x3a;
  double* U0;      // stage 0: iteration iter reads U[iter & 1] and writes the other copy
  double* U1;
  double* acc;     // NCONS * vstride, cell-major: acc[c * NCONS + m] (see acc32)
  StepScalars* sc;
  double gam, cfl, alphabp;
  int bc[6];
  int sw;          // x-strip width of the xy kernel (64, 32 or 16)
  int ns;          // strips per x-row
  int ly, lz;      // y-chunk and z-chunk lengths
  int eflux, mixed;
  // FP32 kernels only: 1 = the flux divergence is stored in FP32, variable-major
  // (((float*)acc)[m * vstride + c]); 0 = FP64, cell-major
  int acc32;
  // 0: one forward-Euler stage per iteration (the reference's step);
  // 1 / 2: the first / second stage of the two-stage Runge-Kutta step, U0 -> U1 and then
  // U1 -> U0 averaged with the U0 it replaces; the state is in U0 after every iteration, and
  // only the second stage reduces the next iteration's dt and c_h
  int stage;
  // 1: the FP64 xy march runs up to two wavefronts per SIMD (set by fused_configure for a box
  // run alone whose strips and boundaries allow it; batches keep 0)
  int duo;
};

// strip width and chunk lengths for a geometry (a.g, a.mixed and a.bc set), for a box run
// alone or in a batch; false if a variable column of the box spans 4 GiB or more (the fused kernels
// address columns with 32-bit offsets)
bool fused_configure(FusedArgs& a, bool alone = true);

// Dynamic LDS of the xy kernel (the pre-pass slots): per block (WAVES = 4 waves) for a box's
// tiling, the most a block may take (FUSED_XY_LDS_MAX, or for FP64 what the static window
// rings and edge exchange leave of the CU's XY_LDS_PER_CU bytes), and the longest y chunk a
// strip width allows within it
constexpr size_t FUSED_XY_LDS_MAX = 64 * 1024;
constexpr size_t XY_LDS_PER_CU = 160 * 1024;
size_t fused_xy_lds(const FusedArgs& a);
size_t fused_xy_lds_max(int mixed);
int fused_max_ly(int sw, int mixed);
// the longest y chunk the single-box xy launch takes for a box's strip width (a.sw, a.ns,
// a.duo set): fused_max_ly, or 64 for the two-wavefronts-per-SIMD FP64 march, which has no
// pre-pass slots
int fused_ly_cap(const FusedArgs& a);
// whether the FP64 xy march can run two wavefronts per SIMD at a box's strip width and boundaries,
// and whether the next single-box launch does (a.duo set and allowed)
bool fused_xy_two_blocks_ok(const FusedArgs& a);
bool fused_xy_two_blocks(const FusedArgs& a);

// blocks of 256 threads each kernel needs for one box
int fused_blocks_xy(const FusedArgs& a);
int fused_blocks_z(const FusedArgs& a);

void launch_fused_xy(const FusedArgs& a, hipStream_t s);
void launch_fused_z(const FusedArgs& a, hipStream_t s);

// Batched launches: boxes[] are device-resident descriptors, tab[block] = (box, block of
// that box); nblocks = tab length.  All boxes of a batch share eflux and precision.
// lds: dynamic LDS bytes of the xy launch, the largest fused_xy_lds over its boxes.
// stage: FusedArgs::stage of every box in the launch (the descriptors' own field is unused)
void launch_batch_xy(const FusedArgs* boxes, const int2* tab, int nblocks, size_t lds, int eflux, int mixed,
                     hipStream_t s, int stage = 0);
void launch_batch_z(const FusedArgs* boxes, const int2* tab, int nblocks, int eflux, int mixed, hipStream_t s,
                    int stage = 0);
// time += dt etc. for each listed box (live[] holds box indices); dt_last[box] gets its dt
void launch_batch_step_end(const FusedArgs* boxes, const int* live, int nlive, double* dt_last, hipStream_t s);

}  // namespace mhd
