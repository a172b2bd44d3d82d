// mhd_fused.hip -- fused step kernels for gfx950 (see mhd_fused.hpp for the data flow).
//
// Per loop iteration of src_cpp_kokkos_host/main.cpp:229-259 the fused path reads the
// conserved state once in each of the xy and z kernels (plus the x-strip edge cells),
// keeps primitives, slopes, interface states and fluxes in registers, and writes the flux
// divergence once and the new state once.  Every face flux is computed once (the x faces
// on strip edges once per adjacent strip), every limited slope once per cell and
// direction, as the reference's operation sequence allows: the same slope of cell f
// serves the right state of face f-1/2 and the left state of face f+1/2
// (KH/mhd.cpp:1134-1154).
#include "mhd_fused.hpp"
#include "mhd_physics.hpp"

#include <algorithm>

// Compilation units.  MHD_FUSED_UNIT=0 (default): the whole file.  The package builds it as
// two units: 1, the host side and the FP32 kernels, and 2, the FP64 kernels and their
// launches, which the build compiles with the instruction-level-parallelism scheduler
// (-mllvm -amdgpu-sched-strategy=max-ilp: the FP64 march 1 % faster at 256^3, while the
// FP32 kernels ran 4-5 % slower with it).
#ifndef MHD_FUSED_UNIT
#define MHD_FUSED_UNIT 0
#endif

namespace mhd {

// one batched launch: the device box descriptors, the work table, its length, the dynamic
// LDS of the xy kernel and the Runge-Kutta stage
struct BatchLaunch {
  const FusedArgs* boxes;
  const int2* tab;
  int nblocks;
  size_t lds;
  int stage;
};
// the FP64 launches (in unit 2, or in this unit when it is the whole file)
void launch_xy_f64(const FusedArgs& a, hipStream_t s);
void launch_z_f64(const FusedArgs& a, hipStream_t s);
void launch_bxy_f64(const BatchLaunch& b, int eflux, hipStream_t s);
void launch_bz_f64(const BatchLaunch& b, int eflux, hipStream_t s);

namespace {

constexpr int WAVES = 4;  // wavefronts per block; blocks carry independent waves

// FP64 march kernels need more than 256 registers per lane; at one wave per SIMD what does
// not fit the 256 architectural VGPRs goes to AGPRs instead of scratch.  The FP32 variants
// fit two waves per SIMD.
#ifndef MHD_F64_MINWAVES
#define MHD_F64_MINWAVES 1
#endif
template <typename R>
struct MinWaves {
  static constexpr int value = sizeof(R) == 8 ? MHD_F64_MINWAVES : 2;
};

__device__ __forceinline__ double bits2d(unsigned long long b) { return __longlong_as_double((long long)b); }
__device__ __forceinline__ unsigned long long d2bits(double d) { return (unsigned long long)__double_as_longlong(d); }

// A variable column is addressed as a wave-uniform base held in scalar registers plus a
// 32-bit byte offset, the scalar-base form of global loads and stores, so no load or store
// pays 64-bit vector address arithmetic.  The base goes through readfirstlane so the compiler
// keeps it scalar instead of folding the column offset into a per-lane 64-bit address.
// fused_configure rejects boxes with vstride * 8 >= 2^32.
__device__ __forceinline__ unsigned boff(int c) { return (unsigned)c * 8u; }
typedef __attribute__((address_space(1))) char gchar;  // global-memory bytes
__device__ __forceinline__ gchar* sbase(const double* p) {
  const unsigned long long v = reinterpret_cast<unsigned long long>(p);
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
  return (gchar*)(((unsigned long long)hi << 32) | lo);
}
__device__ __forceinline__ double ldc(const double* __restrict__ A, size_t vs, int m, unsigned ob) {
  return *(const __attribute__((address_space(1))) double*)(sbase(A + (size_t)m * vs) + ob);
}
__device__ __forceinline__ void stc(double* __restrict__ A, size_t vs, int m, unsigned ob, double v) {
  *(__attribute__((address_space(1))) double*)(sbase(A + (size_t)m * vs) + ob) = v;
}

// The NCONS column bases of one array, made scalar once before a march: readfirstlane is
// convergent, so the compiler does not hoist it out of a loop by itself.
struct Cols {
  gchar* p[NCONS];
};
__device__ __forceinline__ Cols make_cols(const double* A, size_t vs) {
  Cols c;
#pragma unroll
  for (int m = 0; m < NCONS; ++m) c.p[m] = sbase(A + (size_t)m * vs);
  return c;
}
__device__ __forceinline__ void load_raw(const Cols& U, int c, double* u) {
  const unsigned ob = boff(c);
#pragma unroll
  for (int m = 0; m < NCONS; ++m) u[m] = *(const __attribute__((address_space(1))) double*)(U.p[m] + ob);
}
__device__ __forceinline__ void store_cols(const Cols& A, unsigned ob, const double* u) {
#pragma unroll
  for (int m = 0; m < NCONS; ++m) *(__attribute__((address_space(1))) double*)(A.p[m] + ob) = u[m];
}

// The flux divergence acc passed from the xy to the z kernel is cell-major, NCONS doubles
// per cell at acc + c * NCONS: one 64-bit address per lane and five 16-byte accesses with
// immediate offsets (rows of a wavefront still cover one contiguous span).
__device__ __forceinline__ void store_acc(double* __restrict__ acc, int c, const double* v) {
  double2* p = reinterpret_cast<double2*>(acc + (size_t)c * NCONS);
#pragma unroll
  for (int m = 0; m < NCONS / 2; ++m) p[m] = make_double2(v[2 * m], v[2 * m + 1]);
}
__device__ __forceinline__ void load_acc(const double* __restrict__ acc, int c, double* v) {
  const double2* p = reinterpret_cast<const double2*>(acc + (size_t)c * NCONS);
#pragma unroll
  for (int m = 0; m < NCONS / 2; ++m) {
    const double2 t = p[m];
    v[2 * m] = t.x;
    v[2 * m + 1] = t.y;
  }
}

// The flux divergence in FP32 (precision 2): variable-major, so a wavefront's 64 cells of
// one variable are one 256-byte store or load
__device__ __forceinline__ void store_acc32(float* __restrict__ acc, size_t vs, int c, const double* v) {
#pragma unroll
  for (int m = 0; m < NCONS; ++m) acc[m * vs + c] = float(v[m]);
}
__device__ __forceinline__ void load_acc32(const float* __restrict__ acc, size_t vs, int c, double* v) {
#pragma unroll
  for (int m = 0; m < NCONS; ++m) v[m] = double(acc[m * vs + c]);
}

// conserved state of one cell -> primitive vector in precision R (UpdatePrimitvP arithmetic in FP64)
template <typename R>
__device__ __forceinline__ void load_prim(const double* __restrict__ U, size_t vs, int c, double gam, R* q) {
  double u[NCONS], p[NPRIM];
  const unsigned ob = boff(c);
#pragma unroll
  for (int m = 0; m < NCONS; ++m) u[m] = ldc(U, vs, m, ob);
  cons_to_prim(u, gam, p);
#pragma unroll
  for (int n = 0; n < NPRIM; ++n) q[n] = R(p[n]);
}

__device__ __forceinline__ void load_raw(const double* __restrict__ U, size_t vs, int c, double* u) {
  const unsigned ob = boff(c);
#pragma unroll
  for (int m = 0; m < NCONS; ++m) u[m] = ldc(U, vs, m, ob);
}

template <typename R>
__device__ __forceinline__ void recover(const double* u, double gam, R* q) {
  double p[NPRIM];
  cons_to_prim(u, gam, p);
#pragma unroll
  for (int n = 0; n < NPRIM; ++n) q[n] = R(p[n]);
}

template <typename R>
__device__ __forceinline__ R lane_get(R v, int src) {
  return __shfl(v, src, 64);
}

// Primitive recovery of the row/plane loaded at the start of a march pass, placed after
// the pass's Riemann solves and before its stores: the scheduling barrier keeps the
// arithmetic from being hoisted next to the loads (which would wait on them early), and the
// pinned results keep it from being sunk below the stores (a wait there drains the stores).
template <typename R>
__device__ __forceinline__ void recover_in_place(const double* u, double gam, R* q) {
  __builtin_amdgcn_sched_barrier(0);
  recover<R>(u, gam, q);
#pragma unroll
  for (int n = 0; n < NPRIM; ++n) asm volatile("" : "+v"(q[n]));
  __builtin_amdgcn_sched_barrier(0);
}

// lane r's value of v (r wave-uniform): a per-row constant kept one per lane.  Call it in
// wave-uniform control flow on a value from pin_lanes, so that every lane has computed v.
__device__ __forceinline__ double pin_lanes(double v) {
  asm volatile("" : "+v"(v));
  return v;
}
__device__ __forceinline__ double lane_row(double v, int r) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = __builtin_amdgcn_readlane((unsigned)b, r);
  const unsigned hi = __builtin_amdgcn_readlane((unsigned)(b >> 32), r);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// DPP wave shifts (the whole 64-lane wave): lane l receives v from lane l-1 (SHR) or l+1
// (SHL); lane 0 (SHR) / lane 63 (SHL) keep `old`.
#ifndef MHD_XY_DPP
#define MHD_XY_DPP 1
#endif
constexpr int DPP_WAVE_SHR1 = 0x138, DPP_WAVE_SHL1 = 0x130;
template <int CTRL>
__device__ __forceinline__ double dpp_shift(double v, double old) {
  const long long vb = __double_as_longlong(v), ob = __double_as_longlong(old);
  const int lo = __builtin_amdgcn_update_dpp((int)ob, (int)vb, CTRL, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp((int)(ob >> 32), (int)(vb >> 32), CTRL, 0xf, 0xf, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
template <int CTRL>
__device__ __forceinline__ float dpp_shift(float v, float old) {
  return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(v), CTRL, 0xf, 0xf, false));
}

// x neighbour values in a wavefront of strips: lane l takes lane l-1's (from_left) or l+1's
// (from_right) value.  Strip-edge lanes get a neighbour strip's value or their own and are
// overwritten by the caller's strip-edge select.  MHD_XY_DPP: one DPP wave shift per dword;
// otherwise a ds_bpermute from the clamped in-strip source lane.
template <typename R>
__device__ __forceinline__ R from_left(R v, int src_l) {
  if constexpr (MHD_XY_DPP) return dpp_shift<DPP_WAVE_SHR1>(v, v);
  else return lane_get(v, src_l);
}
template <typename R>
__device__ __forceinline__ R from_right(R v, int src_r) {
  if constexpr (MHD_XY_DPP) return dpp_shift<DPP_WAVE_SHL1>(v, v);
  else return lane_get(v, src_r);
}

// a value the compiler must hold in registers as is: an LDS read passed through here stays
// an unconditional read feeding a lane select, instead of becoming a branch per value
template <typename T>
__device__ __forceinline__ T in_reg(T v) {
  asm("" : "+v"(v));
  return v;
}

// source of ghost gg (0,1) on side (0 lo, 1 hi) of an axis with interior [lo, hi]
// (SendRecvBoundary single-tile rules, KH/boundary.cpp:152-220)
__device__ __forceinline__ int ghost_src(int bc_lo, int bc_hi, int side, int gg, int lo, int hi, bool& neg) {
  neg = false;
  const int bc = side ? bc_hi : bc_lo;
  if (bc == BC_PERIODIC) return side ? lo + gg : hi - NGH + 1 + gg;
  if (bc == BC_REFLECT) { neg = true; return side ? hi - gg : lo + NGH - 1 - gg; }
  return side ? hi : lo;
}

// The fused loop never writes ghost cells.  A load of a ghost row, column or plane reads the
// interior cell whose image that ghost is and, at a reflecting wall, negates the normal
// momentum: exactly the values the boundary fill would have stored there.  c in
// [lo - 2, hi + 3] along an axis of n interior cells; the index hi + 3 (n + 4, allocated but
// never a cell) maps like hi + 2, and no result depends on it.
__device__ __forceinline__ int image_of(int c, int n, int bc_lo, int bc_hi, bool& neg) {
  const int lo = NGH, hi = NGH + n - 1;
  neg = false;
  if (c >= lo && c <= hi) return c;
  const int side = c > hi;
  return ghost_src(bc_lo, bc_hi, side, side ? min(c - hi - 1, 1) : c, lo, hi, neg);
}

// v, negated when neg (a wave-uniform flag): one XOR of the sign bit, no branch
__device__ __forceinline__ double negate_if(double v, bool neg) {
  const unsigned long long m = neg ? 0x8000000000000000ull : 0ull;
  return __longlong_as_double((long long)(d2bits(v) ^ m));
}

// conserved state of the cell at (k, j, i) as seen through the boundary images of axis AX
// (0 x: i may be a ghost column; 1 y: j a ghost row; 2 z: k a ghost plane)
template <int AX>
__device__ __forceinline__ void load_image(const double* __restrict__ U, const Geom& g, const int* bc, int k, int j,
                                           int i, double* u) {
  bool neg;
  if constexpr (AX == 0) i = image_of(i, g.nx, bc[0], bc[1], neg);
  if constexpr (AX == 1) j = image_of(j, g.ny, bc[2], bc[3], neg);
  if constexpr (AX == 2) k = image_of(k, g.nz, bc[4], bc[5], neg);
  load_raw(U, g.vstride, g.cell(k, j, i), u);
  u[mrv1 + AX] = negate_if(u[mrv1 + AX], neg);
}

template <int AX, typename R>
__device__ __forceinline__ void load_prim_image(const double* __restrict__ U, const Geom& g, const int* bc, int k,
                                                int j, int i, double gam, R* q) {
  double u[NCONS];
  load_image<AX>(U, g, bc, k, j, i, u);
  recover<R>(u, gam, q);
}

// LDS window ring of one FP64 march wave (xy and z), in values of R: 4 slots of NPRIM x 64
constexpr int WIN_RING = 4 * NPRIM * 64;
#ifndef MHD_XY_RING
#define MHD_XY_RING 1
#endif
#ifndef MHD_XY_COOP
#define MHD_XY_COOP 1
#endif
#ifndef MHD_XY_SHARE
#define MHD_XY_SHARE 1
#endif
// MHD_XY_CARRY=1 (default): the FP64 march keeps row j of pass j (read from the ring for the
// y states) in registers for pass j+1's x faces instead of reading it from the ring again:
// 24 more VGPRs (224), 14 fewer LDS and 34 fewer VALU instructions per row, xy 1.4 % faster
#ifndef MHD_XY_CARRY
#define MHD_XY_CARRY 1
#endif
// block-shared strip-edge exchange of the xy kernels, per wave: edge cells of 4 rows (2 sides
// each) and the right interface state of the last cell, in values of R; one face flux
constexpr int XY_EDGE_R = 9 * NPRIM;

// ---------------------------------------------------------------- x + y divergence
// One wavefront: x-strip s of sw cells (64/sw strips stacked in k), rows [j0, j1) of
// chunk yc.  Lanes hold one cell each; x neighbours come through cross-lane shuffles.
// The two x faces on the strip's edges need cells of the neighbouring strips: a pre-pass
// computes them for every row of the chunk (one face per lane) and parks the face fluxes
// and the outside neighbour cells in this wave's LDS slots, so the row march itself only
// loads the conserved state of the row two ahead.  A face shared by two strips is computed
// by both with the same instructions on the same inputs, so both see the same bits.
//
// LDS of one wave: F[slot][NCONS] doubles, then N[slot][NPRIM] of R,
// slot = (sub * 2 + side) * ly + r  (side 0: left edge face, 1: right edge face; r = row - j0)
//
// Block-cooperative edges (WIDE FP32 strips, MHD_XY_COOP=1): a block's four waves are four
// neighbouring strips of one x-row, or whole x-rows of one or two strips, of the same row
// chunk, and march in step.  An edge between two of them needs no pre-pass: each wave
// publishes its two edge cells as it recovers a row, the right interface state of its last
// cell and the flux of its first face; two block barriers per row order them.  The wave on
// the right computes the shared face (lane 0's left face), the wave on the left reads it.
// With x periodic and the block holding whole x-rows, a row's two end faces pair up the
// same way (the ghost cells are copies of the cells at the other end).  Only edges without
// a neighbour in the block use the pre-pass, whose 4-cell neighbourhoods would otherwise be
// fetched from memory a second time when the march reaches those rows.
template <typename R>
__host__ __device__ constexpr size_t xy_slot_bytes() { return NCONS * sizeof(double) + NPRIM * sizeof(R); }

template <typename R>
__host__ __device__ inline size_t xy_wave_lds(int sw, int ly) { return (size_t)(64 / sw) * 2 * ly * xy_slot_bytes<R>(); }

// WIDE: every strip is one full wave (sw = 64, nx a multiple of 64); x neighbours then move
// by DPP wave shifts, and the strip-edge values from LDS enter as the shifts' end-lane values.
//
// The block-cooperative edges apply to FP32 strips: at two wavefronts per SIMD the other
// block on the CU runs while one waits at a barrier.  FP64 marches run one wavefront per
// SIMD and are bound by its instruction issue, where the barriers cost more than the
// pre-pass they save.
template <typename R, bool WIDE>
__host__ __device__ inline bool xy_coop(const FusedArgs& a) {
  // MHD_XY_COOP=2: FP64 strips too (a measurement switch; the default keeps FP64 on the
  // pre-pass, see above)
  if (!(WIDE && MHD_XY_COOP && (sizeof(R) == 4 || MHD_XY_COOP == 2))) return false;
  if (a.ns % WAVES == 0) return true;
  // one strip per x-row shares an edge only with itself, across a periodic x
  if (a.ns == 1 && !(a.bc[0] == BC_PERIODIC && a.bc[1] == BC_PERIODIC)) return false;
  return WAVES % a.ns == 0 && a.g.nz % (WAVES / a.ns) == 0;
}

// Shared pre-pass (WIDE FP64 strips, four strips of one x-row per block, MHD_XY_SHARE=1):
// the block computes each strip-edge face of its chunk once, 4 or 5 faces per row (4 when
// the block holds the whole periodic x-row), instead of every wave computing both of its
// own edges, and one block barrier publishes them.  Per face and row it keeps the flux and
// both neighbour cells, in the block's slots of the per-wave layout:
//   F[(fi * ly + r) * NCONS], L[(fi * ly + r) * NPRIM] (cell fc-1), R[...] (cell fc),
// fi = strip position of the face in the block (the left face of strip fi; fi = 4 the right
// face of the last strip).
template <typename R, bool WIDE>
__host__ __device__ inline bool xy_share(const FusedArgs& a) {
  return WIDE && MHD_XY_SHARE && MHD_XY_COOP != 2 && sizeof(R) == 8 && a.ns % WAVES == 0;
}

// Two wavefronts per SIMD for the FP64 march (MHD_XY_DUO=1): where every strip edge of a
// block has its neighbour in the block (full-wave strips, x periodic, a block holding one or
// more whole x-rows), the FP64 march takes the block-shared edges and no pre-pass, and its
// window ring drops the row that the carried registers already hold (3 slots: rows j, j+1
// and the upper state of row j-1; row j-1 is the carry).  A wavefront then takes 19.4 KB of
// LDS instead of 24 KiB plus pre-pass slots, so a CU holds up to eight of them: two per SIMD,
// each covering the other's barrier, LDS and memory waits.  Blocks hold 4 wavefronts for
// nx = 64, 128, 256 (4, 2 or 1 whole x-rows; two blocks per CU) and one whole x-row of
// nx / 64 wavefronts for nx = 192 (two blocks, six wavefronts per CU) and 512 (one block of
// eight).  At nx = 384 one block of six ran 3 % slower than the one-wavefront kernel (its
// SIMDs hold 2, 2, 1 and 1 wavefronts, all behind the same barriers), so it is not used.
#ifndef MHD_XY_DUO
#define MHD_XY_DUO 1
#endif
constexpr int WIN_RING3 = 3 * NPRIM * 64;

// Second Runge-Kutta stage in the W form (MHD_RK2_W=1, FP64 divergence storage): the stage's
// xy kernel, which is bound by instruction issue and has HBM bandwidth to spare, reads U^n
// of each cell and stores W = U^n - dt (dFx/dx + dFy/dy) in place of the divergence for the
// nine variables psi damping leaves alone, and the z kernel forms (W + U* - dt dFz/dz) / 2:
// (U^n + E(U*)) / 2 summed in another order.  psi keeps the plain form (its divergence stored,
// U^n of psi read by the z kernel), so the damping stays where it was.  The z kernel reads
// 248 instead of 320 bytes per cell.  Precision 2 keeps the plain form (its divergence is
// stored in FP32, W could not be).
#ifndef MHD_RK2_W
#define MHD_RK2_W 1
#endif
// wavefronts per block of the two-wavefronts-per-SIMD march for a box, 0 where it does not apply
__host__ __device__ inline int xy_duo_waves(const FusedArgs& a) {
  if (!(MHD_XY_DUO && MHD_XY_RING && MHD_XY_CARRY && !a.mixed && a.sw == 64 && a.g.nx % 64 == 0)) return 0;
  if (!(a.bc[0] == BC_PERIODIC && a.bc[1] == BC_PERIODIC)) return 0;
  // four strips of one x-row, or 4/ns whole x-rows of one y chunk
  if (a.ns == WAVES || ((a.ns == 1 || a.ns == 2) && a.g.nz % (WAVES / a.ns) == 0)) return WAVES;
  // one whole x-row
  if (a.ns == 3 || a.ns == 8) return a.ns;
  return 0;
}
__host__ __device__ inline bool xy_duo(const FusedArgs& a) { return xy_duo_waves(a) > 0; }
// static LDS of one block of the march: window rings and edge exchange
__host__ __device__ constexpr size_t xy_duo_lds(int bw) {
  return (size_t)bw * (WIN_RING3 + XY_EDGE_R + NCONS) * sizeof(double);
}
static_assert((WAVES + 1) * (NCONS + 2 * NPRIM) * sizeof(double) <= WAVES * 2 * (NCONS + NPRIM) * sizeof(double),
              "the shared pre-pass slots fit the block's per-wave slots");
// static LDS of the edge exchange (FP32 only)
template <typename R>
__host__ __device__ constexpr int xy_edge_len(int per_wave) {
  return (sizeof(R) == 4 || MHD_XY_COOP == 2) ? WAVES * per_wave : 1;
}
//
// FP64 (MHD_XY_RING=1): the window rows and the previous pass's upper y state live in this
// wave's LDS ring, ring[(slot * NPRIM + n) * 64 + lane]: slots 0..2 hold rows j-1, j, j+1 in
// rotation and slot 3 the upper state of row j-1.  A row is written once per pass and read
// where it is used instead of being carried (and rotated) in registers, which leaves the
// registers to the Riemann solves.  FP32 keeps the window in registers.
// ACC32: the divergence is stored in FP32 (FusedArgs::acc32), fixed at compile time so the
// other kernels carry nothing of it
// DUO: the two-wavefronts-per-SIMD FP64 march (xy_duo holds): block-shared edges on both sides of
// every strip, 3-slot ring
// BW: wavefronts per block; W2: the second Runge-Kutta stage in the W form (MHD_RK2_W)
template <int EFLUX, typename R, bool WIDE, bool ACC32 = false, bool DUO = false, int BW = WAVES, bool W2 = false>
__device__ __forceinline__ void xy_body(const FusedArgs& a, int stage, int w, int lane, char* lds, R* ring, R* edgeP,
                                        double* edgeF, int wv) {
  static_assert(!DUO || (WIDE && sizeof(R) == 8 && MHD_XY_RING && MHD_XY_CARRY), "DUO: FP64 full-wave strips");
  static_assert(!W2 || !ACC32, "the W form stores FP64");
  const StepScalars* sc = a.sc;
  if (sc->stopped) return;
  const Geom& g = a.g;
  const size_t vs = g.vstride;
  const int iter = sc->iter;
  const double* __restrict__ U = stage == 2 ? a.U1 : (stage == 0 && (iter & 1)) ? a.U1 : a.U0;
  const double chg = bits2d(sc->red[iter % 3][1]);
  const double hchg = 0.5 / chg;
  const int SW = a.sw;
  const int SPW = 64 / SW;
  const int sub = lane / SW, l = lane - sub * SW;
  const int nkg = (g.nz + SPW - 1) / SPW;
  const int nyc = (g.ny + a.ly - 1) / a.ly;
  const int s = w % a.ns;
  const int t = w / a.ns;
  const int kg = t % nkg;
  const int yc = t / nkg;
  if (yc >= nyc) return;  // uniform over the wavefront
  const int kraw = NGH + kg * SPW + sub;
  const bool kvalid = kraw < NGH + g.nz;
  const int k = kvalid ? kraw : NGH + g.nz - 1;
  const int i0 = NGH + s * SW;
  const int ilast = min(i0 + SW, NGH + g.nx) - 1;
  const int iraw = i0 + l;
  const bool ivalid = iraw <= ilast;
  const int i = ivalid ? iraw : ilast;
  const bool active = kvalid && ivalid;
  const bool is_left = (l == 0);
  const bool is_right = (iraw == ilast);
  const int src_l = sub * SW + (l > 0 ? l - 1 : l);
  const int src_r = sub * SW + (l < SW - 1 ? l + 1 : l);
  const int j0 = NGH + yc * a.ly;
  const int j1 = min(j0 + a.ly, NGH + g.ny);
  const int nrows = j1 - j0;
  const double xinv = 1.0 / (a.x1a[i + 1] - a.x1a[i]);
  const double dt2 = W2 ? a.cfl * bits2d(sc->red[iter % 3][0]) : 0.0;  // W2: dt of this step
  // 1/dy of the chunk's rows, row j0 + r in lane r (a.ly <= 64), read back per row by index
  const double yinv_l =
      pin_lanes(1.0 / (a.x2a[j0 + min(lane, nrows - 1) + 1] - a.x2a[j0 + min(lane, nrows - 1)]));
  double* ldsF = reinterpret_cast<double*>(lds);
  R* ldsN = reinterpret_cast<R*>(lds + (size_t)SPW * 2 * a.ly * NCONS * sizeof(double));
  // block-uniform: a block holds 4 neighbouring strips of one x-row (ns a multiple of 4), or
  // 4/ns whole x-rows of the same y chunk (ns = 1 or 2 and nz a multiple of 4/ns), so all of
  // its waves march the same rows and take the same barriers
  const bool coop = DUO || xy_coop<R, WIDE>(a);
  const int rw = min(a.ns, BW);  // strips of one x-row in the block
  const int pos = wv % rw, grp = wv - pos;
  const bool wrap = coop && a.ns <= WAVES && a.bc[0] == BC_PERIODIC && a.bc[1] == BC_PERIODIC;
  const bool coop_l = DUO || (coop && (pos > 0 || wrap));
  const bool coop_r = DUO || (coop && (pos < rw - 1 || wrap));
  const int nb_l = grp + (pos + rw - 1) % rw, nb_r = grp + (pos + 1) % rw;
  // the edge cell (side 0: first, 1: last) of row jr that wavefront `wave` published
  auto edge_cell = [&](int wave, int jr, int side) { return edgeP + wave * XY_EDGE_R + ((jr & 3) * 2 + side) * NPRIM; };
  R* const edge_rt = edgeP + wv * XY_EDGE_R + 8 * NPRIM;

  const bool share = !DUO && xy_share<R, WIDE>(a);
  // the pre-pass slots this wave's march reads, per row r: left / right face flux and the
  // outside neighbour cells
  const double* fL;
  const double* fR;
  const R* nL;
  const R* nR;
  if (share) {
    // the block's slots start at this wave's minus wv per-wave blocks
    char* blk = lds - (size_t)wv * xy_wave_lds<R>(SW, a.ly);
    double* shF = reinterpret_cast<double*>(blk);
    R* shL = reinterpret_cast<R*>(blk + (size_t)(WAVES + 1) * a.ly * NCONS * sizeof(double));
    R* shR = shL + (size_t)(WAVES + 1) * a.ly * NPRIM;
    const bool wrap4 = a.ns == WAVES && a.bc[0] == BC_PERIODIC && a.bc[1] == BC_PERIODIC;
    const int nface = wrap4 ? WAVES : WAVES + 1;
    const int items = nface * nrows;
    const int s_first = s - pos;
    for (int f = wv * 64 + lane; f < items; f += 64 * WAVES) {
      const int r = f % nrows, fi = f / nrows;
      const int fc = NGH + (s_first + fi) * 64;  // the face between cells fc-1 and fc
      R q[4][NPRIM];
#pragma unroll
      for (int c = 0; c < 4; ++c) load_prim_image<0, R>(U, g, a.bc, k, j0 + r, fc - 2 + c, a.gam, q[c]);
      R ql[NPRIM], qr[NPRIM], unused[NPRIM];
      cell_states<R>(q[0], q[1], q[2], ql, unused);
      cell_states<R>(q[1], q[2], q[3], unused, qr);
      double F[NCONS];
      riemann_flux<0, EFLUX, R>(ql, qr, chg, hchg, F);
      const int slot = fi * a.ly + r;
#pragma unroll
      for (int m = 0; m < NCONS; ++m) shF[slot * NCONS + m] = F[m];
#pragma unroll
      for (int n = 0; n < NPRIM; ++n) {
        shL[slot * NPRIM + n] = q[1][n];
        shR[slot * NPRIM + n] = q[2][n];
      }
    }
    const int rfi = (wrap4 && pos == WAVES - 1) ? 0 : pos + 1;
    fL = shF + (size_t)pos * a.ly * NCONS;
    fR = shF + (size_t)rfi * a.ly * NCONS;
    nL = shL + (size_t)pos * a.ly * NPRIM;
    nR = shR + (size_t)rfi * a.ly * NPRIM;
    __syncthreads();
  } else {
    fL = ldsF + (size_t)(sub * 2) * a.ly * NCONS;
    fR = fL + (size_t)a.ly * NCONS;
    nL = ldsN + (size_t)(sub * 2) * a.ly * NPRIM;
    nR = nL + (size_t)a.ly * NPRIM;
  }

  // pre-pass: the strip-edge x faces of every row of the chunk without a block neighbour,
  // one face per lane
  const int side0 = coop_l ? 1 : 0;
  const int nsides = share ? 0 : 2 - (coop_l ? 1 : 0) - (coop_r ? 1 : 0);
  const int nf = SPW * nsides * nrows;
  for (int f = lane; f < nf; f += 64) {
    const int r = f % nrows, side = side0 + (f / nrows) % max(nsides, 1), sb = f / (max(nsides, 1) * nrows);
    const int kr = min(NGH + kg * SPW + sb, NGH + g.nz - 1);
    const int fc = side ? ilast + 1 : i0;  // the face between cells fc-1 and fc
    R q[4][NPRIM];
#pragma unroll
    for (int c = 0; c < 4; ++c) load_prim_image<0, R>(U, g, a.bc, kr, j0 + r, fc - 2 + c, a.gam, q[c]);
    R ql[NPRIM], qr[NPRIM], unused[NPRIM];
    cell_states<R>(q[0], q[1], q[2], ql, unused);
    cell_states<R>(q[1], q[2], q[3], unused, qr);
    double F[NCONS];
    riemann_flux<0, EFLUX, R>(ql, qr, chg, hchg, F);
    const int slot = (sb * 2 + side) * a.ly + r;
#pragma unroll
    for (int m = 0; m < NCONS; ++m) ldsF[slot * NCONS + m] = F[m];
#pragma unroll
    for (int n = 0; n < NPRIM; ++n) ldsN[slot * NPRIM + n] = side ? q[2][n] : q[1][n];
  }
  // LDS traffic of one wavefront is ordered; the fence keeps the compiler from moving the
  // slot reads of the march above the writes
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");

  constexpr bool PRING = MHD_XY_RING && sizeof(R) == 8;
  auto rdr = [&](int slot, R* q) {
#pragma unroll
    for (int n = 0; n < NPRIM; ++n) q[n] = ring[(slot * NPRIM + n) * 64 + lane];
  };
  auto wr = [&](int slot, const R* q) {
#pragma unroll
    for (int n = 0; n < NPRIM; ++n) ring[(slot * NPRIM + n) * 64 + lane] = q[n];
  };
  R Pm[NPRIM], P0[NPRIM], Pp[NPRIM], upprev[NPRIM];
  constexpr bool CARRY = PRING && MHD_XY_CARRY;
  R Px[NPRIM];  // CARRY: row j-1 for this pass's x faces, kept from the previous pass
  double Fyprev[NCONS];
  load_prim_image<1, R>(U, g, a.bc, k, j0 - 2, i, a.gam, Pm);
  load_prim_image<1, R>(U, g, a.bc, k, j0 - 1, i, a.gam, P0);
  load_prim_image<1, R>(U, g, a.bc, k, j0, i, a.gam, Pp);
  const Cols Uc = make_cols(U, vs);
  const Cols U0c = make_cols(W2 ? a.U0 : U, vs);  // W2: U^n (stage 2 reads U* from U1)
  if (coop) {
    if (lane == 0 || lane == 63) {
      R* e = edge_cell(wv, j0, lane == 63);
#pragma unroll
      for (int n = 0; n < NPRIM; ++n) e[n] = Pp[n];
    }
    __syncthreads();
  }
#pragma unroll
  for (int n = 0; n < NPRIM; ++n) upprev[n] = R(0);
#pragma unroll
  for (int m = 0; m < NCONS; ++m) Fyprev[m] = 0.0;
  // DUO: slots 0 and 1 hold rows j and j+1 in rotation (s0, sp), slot 2 the upper state of
  // row j-1; row j-1 itself is the carry Px
  if constexpr (DUO) {
#pragma unroll
    for (int n = 0; n < NPRIM; ++n) Px[n] = Pm[n];
    wr(0, P0);
    wr(1, Pp);
    wr(2, upprev);
  } else if constexpr (PRING) {
    wr(0, Pm);
    wr(1, P0);
    wr(2, Pp);
    wr(3, upprev);
  }
  int sm = 0, s0 = DUO ? 0 : 1, sp = DUO ? 1 : 2;

  // Pass j: the window holds rows j-1, j, j+1.  The raw state of row j+2 is loaded at the
  // start and turned into primitives after both Riemann solves have covered the load
  // latency; only computed values cross the loop edge, so no load in flight is copied.
  for (int j = j0 - 1; j <= j1; ++j) {
    double un[NCONS];
    // row j+2, or the row it is the image of (made scalar: j is uniform over the wave)
    bool neg_n;
    const int jn = image_of(__builtin_amdgcn_readfirstlane(j + 2), g.ny, a.bc[2], a.bc[3], neg_n);
    load_raw(Uc, g.cell(k, jn, i), un);
    const int row = j - 1;  // the cell row completed in this pass
    const bool out_row = row >= j0;
    double ax[NCONS];
    if (out_row) {
      const int rr = row - j0;
      if constexpr (CARRY) {
#pragma unroll
        for (int n = 0; n < NPRIM; ++n) Pm[n] = Px[n];
      } else if constexpr (PRING) {
        rdr(sm, Pm);
      }
      R qm[NPRIM], qp[NPRIM];
      const R* srcl = coop_l ? edge_cell(nb_l, row, 1) : nL + rr * NPRIM;
      const R* srcr = coop_r ? edge_cell(nb_r, row, 0) : nR + rr * NPRIM;
#pragma unroll
      for (int n = 0; n < NPRIM; ++n) {
        const R nl = in_reg(srcl[n]);
        const R nr = in_reg(srcr[n]);
        if constexpr (WIDE) {
          qm[n] = dpp_shift<DPP_WAVE_SHR1>(Pm[n], nl);
          qp[n] = dpp_shift<DPP_WAVE_SHL1>(Pm[n], nr);
        } else {
          const R vm = from_left(Pm[n], src_l);
          const R vp = from_right(Pm[n], src_r);
          qm[n] = is_left ? nl : vm;
          qp[n] = is_right ? nr : vp;
        }
      }
      R rt[NPRIM], lt[NPRIM];
      cell_states<R>(qm, Pm, qp, rt, lt);
      if (coop) {
        if (lane == 63) {
#pragma unroll
          for (int n = 0; n < NPRIM; ++n) edge_rt[n] = rt[n];
        }
        __syncthreads();
      }
      R rtl[NPRIM];
      if (coop_l) {
        // lane 0 takes the right state of the left neighbour's last cell
        const R* ol = edgeP + nb_l * XY_EDGE_R + 8 * NPRIM;
#pragma unroll
        for (int n = 0; n < NPRIM; ++n) rtl[n] = dpp_shift<DPP_WAVE_SHR1>(rt[n], in_reg(ol[n]));
      } else {
#pragma unroll
        for (int n = 0; n < NPRIM; ++n) {
          if constexpr (WIDE) rtl[n] = dpp_shift<DPP_WAVE_SHR1>(rt[n], rt[n]);  // lane 0's face comes from LDS
          else rtl[n] = from_left(rt[n], src_l);
        }
      }
      double Fl[NCONS];
      riemann_flux<0, EFLUX, R>(rtl, lt, chg, hchg, Fl);
      if (!coop_l) {
#pragma unroll
        for (int m = 0; m < NCONS; ++m) {
          const double fl = in_reg(fL[rr * NCONS + m]);
          Fl[m] = is_left ? fl : Fl[m];
        }
      }
      if (coop) {
        if (lane == 0) {
#pragma unroll
          for (int m = 0; m < NCONS; ++m) edgeF[wv * NCONS + m] = Fl[m];
        }
        __syncthreads();
      }
      const double* srcf = coop_r ? edgeF + nb_r * NCONS : fR + rr * NCONS;
#pragma unroll
      for (int m = 0; m < NCONS; ++m) {
        const double fe = in_reg(srcf[m]);
        if constexpr (WIDE) {
          ax[m] = (dpp_shift<DPP_WAVE_SHL1>(Fl[m], fe) - Fl[m]) * xinv;
        } else {
          const double fr = from_right(Fl[m], src_r);
          ax[m] = ((is_right ? fe : fr) - Fl[m]) * xinv;
        }
      }
    }
    // W2: U^n of the cell completed in this pass, issued by every lane on every pass (row
    // >= j0 - 2 >= 0 is in the array; unused when not stored), after the x faces so that it
    // is held only through the y face
    double un0[W2 ? NCONS : 1];
    if constexpr (W2) {
      const unsigned ob = boff(g.cell(k, row, i));
#pragma unroll
      for (int m = 0; m < NCONS; ++m)
        if (m != mbps) un0[m] = *(const __attribute__((address_space(1))) double*)(U0c.p[m] + ob);
    }
    const double yinv = lane_row(yinv_l, max(row - j0, 0));
    // y states of row j and the y face j-1/2
    R up[NPRIM], dn[NPRIM];
    if constexpr (DUO) {
      rdr(s0, P0);
      rdr(sp, Pp);
      cell_states<R>(Px, P0, Pp, up, dn);
#pragma unroll
      for (int n = 0; n < NPRIM; ++n) Px[n] = P0[n];
      rdr(2, upprev);
      wr(2, up);
    } else if constexpr (PRING) {
      rdr(sm, Pm);
      rdr(s0, P0);
      rdr(sp, Pp);
      cell_states<R>(Pm, P0, Pp, up, dn);
      if constexpr (CARRY) {
#pragma unroll
        for (int n = 0; n < NPRIM; ++n) Px[n] = P0[n];
      }
      rdr(3, upprev);
      wr(3, up);
    } else {
      cell_states<R>(Pm, P0, Pp, up, dn);
    }
    // the first pass (j = j0 - 1) solves a face against the zero state upprev starts as; its
    // flux only becomes Fyprev of pass j0, which completes no row (its row j0 - 1 is a halo)
    double Fy[NCONS];
    riemann_flux<1, EFLUX, R>(upprev, dn, chg, hchg, Fy);
    // row j+2 enters before this pass's stores are issued: a wait for a load behind a
    // pending store would drain the store too (one counter for both on this architecture)
    R Pn[NPRIM];
    un[mrv2] = negate_if(un[mrv2], neg_n);
    recover_in_place<R>(un, a.gam, Pn);
    if constexpr (DUO) wr(s0, Pn);  // row j's slot: row j is the carry now
    else if constexpr (PRING) wr(sm, Pn);
    if (coop && (lane == 0 || lane == 63)) {
      R* e = edge_cell(wv, j + 2, lane == 63);
#pragma unroll
      for (int n = 0; n < NPRIM; ++n) e[n] = Pn[n];
    }
    if (out_row && active) {
      double d[NCONS];
#pragma unroll
      for (int m = 0; m < NCONS; ++m) d[m] = ax[m] + (Fy[m] - Fyprev[m]) * yinv;
      if constexpr (W2) {
#pragma unroll
        for (int m = 0; m < NCONS; ++m)
          if (m != mbps) d[m] = un0[m] - dt2 * d[m];
      }
      if constexpr (ACC32) store_acc32(reinterpret_cast<float*>(a.acc), vs, g.cell(k, row, i), d);
      else store_acc(a.acc, g.cell(k, row, i), d);
    }
    if constexpr (DUO) {
      const int t0 = s0;
      s0 = sp;
      sp = t0;
    } else if constexpr (PRING) {
      const int t0 = sm;
      sm = s0;
      s0 = sp;
      sp = t0;
    } else {
#pragma unroll
      for (int n = 0; n < NPRIM; ++n) {
        upprev[n] = up[n];
        Pm[n] = P0[n];
        P0[n] = Pp[n];
        Pp[n] = Pn[n];
      }
    }
#pragma unroll
    for (int m = 0; m < NCONS; ++m) Fyprev[m] = Fy[m];
  }
}

// strips fill whole waves (the DPP neighbour path applies)
__host__ __device__ inline bool xy_wide(const FusedArgs& a) { return MHD_XY_DPP && a.sw == 64 && a.g.nx % 64 == 0; }

// static LDS of the xy kernels: the FP64 window rings (96 KiB per block); the pre-pass
// slots are dynamic (at most FUSED_XY_LDS_MAX)
template <typename R>
__host__ __device__ constexpr int xy_ring_len() { return (MHD_XY_RING && sizeof(R) == 8) ? WAVES * WIN_RING : 1; }

template <int EFLUX, typename R, bool WIDE, bool ACC32, bool W2 = false>
__global__ __launch_bounds__(256, MinWaves<R>::value) void k_fused_xy(FusedArgs a) {
  extern __shared__ __align__(16) char smem[];
  __shared__ R xring[xy_ring_len<R>()];
  __shared__ R xedgeP[xy_edge_len<R>(XY_EDGE_R)];
  __shared__ double xedgeF[xy_edge_len<R>(NCONS)];
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform over the wave
  xy_body<EFLUX, R, WIDE, ACC32, false, WAVES, W2>(a, a.stage, blockIdx.x * WAVES + wv, threadIdx.x & 63,
                          smem + wv * xy_wave_lds<R>(a.sw, a.ly), xring + (xy_ring_len<R>() > 1 ? wv * WIN_RING : 0),
                          xedgeP, xedgeF, wv);
}

// FP64 march at two wavefronts per SIMD (xy_duo): blocks of BW wavefronts, static LDS only
// (BW x 18 KiB of window rings and the edge exchange: 77.5 KB for BW = 4, 155 KB for 8)
template <int EFLUX, int BW, bool W2 = false>
__global__ __launch_bounds__(64 * BW, 2) void k_fused_xy_duo(FusedArgs a) {
  __shared__ double xring[BW * WIN_RING3];
  __shared__ double xedgeP[BW * XY_EDGE_R];
  __shared__ double xedgeF[BW * NCONS];
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform over the wave
  xy_body<EFLUX, double, true, false, true, BW, W2>(a, a.stage, blockIdx.x * BW + wv, threadIdx.x & 63, nullptr,
                                                xring + wv * WIN_RING3, xedgeP, xedgeF, wv);
}
static_assert(xy_duo_lds(8) <= XY_LDS_PER_CU && 2 * xy_duo_lds(4) <= XY_LDS_PER_CU &&
                  2 * xy_duo_lds(3) <= XY_LDS_PER_CU,
              "two-wavefront march blocks fit the CU's LDS");

// ---------------------------------------------------------------- z + update + recovery + reduction
// The z kernels also keep the raw conserved state of the last three planes loaded: the old
// state of the cell a pass updates was loaded three passes earlier as the plane ahead, so it
// is read back from LDS instead of from memory.
constexpr int Z_RAW = 3 * NCONS * 64;
// one wavefront: an x-strip of sw cells (the xy kernel's strip width) on 64/sw consecutive
// rows, one column per lane, planes [k0, k1) of chunk zc.  The columns are independent, so
// narrow boxes stack rows in a wave instead of leaving lanes idle.
// STAGE: FusedArgs::stage, fixed at compile time so the one-stage kernel carries nothing of
// the other two.  The second Runge-Kutta stage (STAGE 2) averages each updated cell with the
// state it overwrites.
template <int EFLUX, typename R, int STAGE, bool ACC32 = false>
__device__ __forceinline__ void z_body(const FusedArgs& a, int w, int lane, R* ring, double* raw) {
  constexpr bool AVG = STAGE == 2;
  // the W form of the second stage (MHD_RK2_W): the divergence buffer holds W (the xy
  // kernel's U^n - dt dF_xy) for every variable but psi, so only U^n of psi is read here
  constexpr bool WFORM = AVG && !ACC32 && MHD_RK2_W;
  StepScalars* sc = a.sc;
  if (sc->stopped) return;
  const Geom& g = a.g;
  const size_t vs = g.vstride;
  const int iter = sc->iter;
  const bool odd = STAGE == 2 || (STAGE == 0 && (iter & 1));
  const double* __restrict__ U = odd ? a.U1 : a.U0;
  double* __restrict__ Uo = odd ? a.U0 : a.U1;
  constexpr bool reduce = STAGE != 1;
  const unsigned long long* rd = sc->red[iter % 3];
  const double dt = a.cfl * bits2d(rd[0]);
  const double chg = bits2d(rd[1]);
  const double hchg = 0.5 / chg;
  const int SW = a.sw;
  const int SPW = 64 / SW;
  const int sub = lane / SW, l = lane - sub * SW;
  const int nrg = (g.ny + SPW - 1) / SPW;  // row groups
  const int nzc = (g.nz + a.lz - 1) / a.lz;
  const int s = w % a.ns;
  const int t = w / a.ns;
  const int jg = t % nrg;
  const int zc = t / nrg;
  if (zc >= nzc) return;  // uniform over the wavefront
  const int iraw = NGH + s * SW + l;
  const int jraw = NGH + jg * SPW + sub;
  const bool active = iraw < NGH + g.nx && jraw < NGH + g.ny;
  const int i = min(iraw, NGH + g.nx - 1);
  const int j = min(jraw, NGH + g.ny - 1);
  const int k0 = NGH + zc * a.lz;
  const int k1 = min(k0 + a.lz, NGH + g.nz);
  const double dxa = a.x1a[i + 1] - a.x1a[i];
  const double dya = a.x2a[j + 1] - a.x2a[j];
  // dz and 1/dz of the chunk's planes, plane k0 + r in lane r (a.lz <= 64)
  const int kl = k0 + min(lane, k1 - k0 - 1);
  const double dz_l = pin_lanes(a.x3a[kl + 1] - a.x3a[kl]);
  const double zinv_l = pin_lanes(1.0 / (a.x3a[kl + 1] - a.x3a[kl]));

  // FP64: the window planes and the previous pass's upper interface state live in this
  // wave's LDS ring, ring[(slot * NPRIM + n) * 64 + lane]: slots 0..2 hold planes k-1, k, k+1
  // in rotation (sm, s0, sp) and slot 3 the upper state of plane k-1.  A plane is written
  // once and read once per pass instead of being copied through registers, which leaves the
  // registers to the Riemann solve.  FP32 (two blocks per CU, which leaves no LDS for a
  // window ring) keeps them in registers (wm, w0, wp, wu).  In both,
  // raw[((p % 3) * NCONS + m) * 64 + lane] holds the conserved state of plane p for the last
  // three planes loaded.
  constexpr bool PRING = sizeof(R) == 8;
  R wm[NPRIM], w0[NPRIM], wp[NPRIM], wu[NPRIM];
  auto rdr = [&](int slot, R* q) {
#pragma unroll
    for (int n = 0; n < NPRIM; ++n) q[n] = ring[(slot * NPRIM + n) * 64 + lane];
  };
  auto wr = [&](int slot, const R* q) {
#pragma unroll
    for (int n = 0; n < NPRIM; ++n) ring[(slot * NPRIM + n) * 64 + lane] = q[n];
  };
  load_prim_image<2, R>(U, g, a.bc, k0 - 2, j, i, a.gam, wm);
  load_prim_image<2, R>(U, g, a.bc, k0 - 1, j, i, a.gam, w0);
  const Cols Uc = make_cols(U, vs), Uoc = make_cols(Uo, vs);
  {
    double u0[NCONS];
    load_raw(U, vs, g.cell(k0, j, i), u0);
#pragma unroll
    for (int m = 0; m < NCONS; ++m) raw[((k0 % 3) * NCONS + m) * 64 + lane] = u0[m];
    recover<R>(u0, a.gam, wp);
  }
#pragma unroll
  for (int n = 0; n < NPRIM; ++n) wu[n] = R(0);
  if constexpr (PRING) {
    wr(0, wm);
    wr(1, w0);
    wr(2, wp);
    wr(3, wu);
  }
  int sm = 0, s0 = 1, sp = 2;
  double Fzprev[NCONS];
#pragma unroll
  for (int m = 0; m < NCONS; ++m) Fzprev[m] = 0.0;
  double dtl = __builtin_inf(), chl = 0.0;
  bool bad = false;  // an updated cell of this lane is a bad_cell

  // Pass k: cell k-1 is updated.  The raw state of plane k+2 and the old state and
  // divergence of cell k-1 are loaded at the start; plane k+2 becomes primitives after the
  // Riemann solve (only computed values cross the loop edge).
  for (int k = k0 - 1; k <= k1; ++k) {
    const int kk = k - 1;  // the cell updated in this pass
    const bool upd = kk >= k0 && active;
    double un[NCONS];
    // plane k+2, or the plane it is the image of (made scalar: k is uniform over the wave)
    bool neg_n;
    const int kn = image_of(__builtin_amdgcn_readfirstlane(k + 2), g.nz, a.bc[4], a.bc[5], neg_n);
    load_raw(Uc, g.cell(kn, j, i), un);
    double uo[NCONS], ac[NCONS];
    const unsigned oc = boff(g.cell(kk, j, i));
    const double dza = lane_row(dz_l, max(kk - k0, 0));
    const double zinv = lane_row(zinv_l, max(kk - k0, 0));
    const int rs = (k + 2) % 3;  // raw slot of plane k-1 (read now) and of plane k+2 (written below)
#pragma unroll
    for (int m = 0; m < NCONS; ++m) uo[m] = raw[(rs * NCONS + m) * 64 + lane];
    // issued by every lane on every pass (cell kk >= k0 - 2 >= 0 is in the array; unused
    // when not updated), so the number of loads in flight is the same on every path and a
    // wait for an earlier load need not drain these
    if constexpr (ACC32) load_acc32(reinterpret_cast<const float*>(a.acc), vs, g.cell(kk, j, i), ac);
    else load_acc(a.acc, g.cell(kk, j, i), ac);
    // the second stage's U^n of cell kk, loaded under the same rule as the divergence
    double un0[AVG ? NCONS : 1];
    if constexpr (WFORM) un0[mbps] = *(const __attribute__((address_space(1))) double*)(Uoc.p[mbps] + oc);
    else if constexpr (AVG) load_raw(Uoc, g.cell(kk, j, i), un0);
    R up[NPRIM], dn[NPRIM], upprev[NPRIM];
    if constexpr (PRING) {
      R Pm[NPRIM], P0[NPRIM], Pp[NPRIM];
      rdr(sm, Pm);
      rdr(s0, P0);
      rdr(sp, Pp);
      cell_states<R>(Pm, P0, Pp, up, dn);
      rdr(3, upprev);
      wr(3, up);
    } else {
      cell_states<R>(wm, w0, wp, up, dn);
#pragma unroll
      for (int n = 0; n < NPRIM; ++n) {
        upprev[n] = wu[n];
        wu[n] = up[n];
      }
    }
    double Fz[NCONS] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (k >= k0) riemann_flux<2, EFLUX, R>(upprev, dn, chg, hchg, Fz);
    // plane k+2 enters before this pass's stores are issued (see the xy march)
    R Pn[NPRIM];
    un[mrv3] = negate_if(un[mrv3], neg_n);
    recover_in_place<R>(un, a.gam, Pn);
    if constexpr (PRING) {
      wr(sm, Pn);
    } else {
#pragma unroll
      for (int n = 0; n < NPRIM; ++n) {
        wm[n] = w0[n];
        w0[n] = wp[n];
        wp[n] = Pn[n];
      }
    }
#pragma unroll
    for (int m = 0; m < NCONS; ++m) raw[(rs * NCONS + m) * 64 + lane] = un[m];
    if (upd) {
      double u[NCONS];
#pragma unroll
      for (int m = 0; m < NCONS; ++m) {
        if (WFORM && m != mbps) u[m] = uo[m] - dt * ((Fz[m] - Fzprev[m]) * zinv);
        else u[m] = uo[m] - dt * (ac[m] + (Fz[m] - Fzprev[m]) * zinv);
      }
      // DampPsi (KH/mhd.cpp:1625-1631)
      const double dhl = fmin(dxa, fmin(dya, dza));
      const double taui = a.alphabp * chg / dhl;
      u[mbps] = u[mbps] * (1.0 - dt * taui);
      if constexpr (WFORM) {
#pragma unroll
        for (int m = 0; m < NCONS; ++m) u[m] = 0.5 * ((m == mbps ? un0[m] : ac[m]) + u[m]);  // ac holds W
      } else if constexpr (AVG) {
#pragma unroll
        for (int m = 0; m < NCONS; ++m) u[m] = 0.5 * (un0[m] + u[m]);
      }
      store_cols(Uoc, oc, u);
      // the next iteration's ControlTimestep / EvaluateCh bounds of this cell
      if constexpr (reduce) {
        double q[NPRIM];
        cons_to_prim(u, a.gam, q);
        const double cdt = cell_dt(q, dxa, dya, dza), cch = cell_ch(q);
        bad |= bad_cell(q[nden], cdt, cch);
        dtl = fmin(dtl, cdt);
        chl = fmax(chl, cch);
      }
    }
    const int t0 = sm;
    sm = s0;
    s0 = sp;
    sp = t0;
#pragma unroll
    for (int m = 0; m < NCONS; ++m) Fzprev[m] = Fz[m];
  }
  if constexpr (!reduce) return;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    dtl = fmin(dtl, __shfl_xor(dtl, o, 64));
    chl = fmax(chl, __shfl_xor(chl, o, 64));
  }
  const unsigned long long anybad = __ballot(bad);
  if (lane == 0) {
    unsigned long long* nx = sc->red[(iter + 1) % 3];
    if (!(chl == chl)) chl = __builtin_inf();
    atomicMin(&nx[0], d2bits(dtl));
    atomicMax(&nx[1], d2bits(chl));
    // the step-end launch stops the loop after this step (the new state holds the cell)
    if (anybad) {
      atomicOr(&sc->status, 1);
      atomicExch(&sc->halt, 1);
    }
  }
}

template <int EFLUX, typename R, int STAGE, bool ACC32>
__global__ __launch_bounds__(256, MinWaves<R>::value) void k_fused_z(FusedArgs a) {
  // LDS: FP64 window ring (96 KiB) + raw ring (60 KiB) per block; FP32 raw ring only
  __shared__ R zring[sizeof(R) == 8 ? WAVES * WIN_RING : 1];
  __shared__ double zraw[WAVES * Z_RAW];
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform over the wave
  z_body<EFLUX, R, STAGE, ACC32>(a, blockIdx.x * WAVES + wv, threadIdx.x & 63,
                                 zring + (sizeof(R) == 8 ? wv * WIN_RING : 0),
                          zraw + wv * Z_RAW);
}

// ---------------------------------------------------------------- batched launches
// Many boxes in one launch: tab[block] = (box, block index within that box).  Every box's
// block count is a whole number of blocks, so a block never straddles two boxes and the
// box descriptor read is uniform over the block.  stage: FusedArgs::stage for every box.
template <int EFLUX, typename R, bool ACC32, bool W2 = false>
__global__ __launch_bounds__(256, MinWaves<R>::value) void k_batch_xy(const FusedArgs* __restrict__ boxes,
                                                                     const int2* __restrict__ tab, int stage) {
  extern __shared__ __align__(16) char smem[];
  __shared__ R xring[xy_ring_len<R>()];
  __shared__ R xedgeP[xy_edge_len<R>(XY_EDGE_R)];
  __shared__ double xedgeF[xy_edge_len<R>(NCONS)];
  const int2 e = tab[blockIdx.x];
  const FusedArgs& a = boxes[__builtin_amdgcn_readfirstlane(e.x)];
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform over the wave
  char* lds = smem + wv * xy_wave_lds<R>(a.sw, a.ly);
  R* ring = xring + (xy_ring_len<R>() > 1 ? wv * WIN_RING : 0);
  if (xy_wide(a))
    xy_body<EFLUX, R, true, ACC32, false, WAVES, W2>(a, stage, e.y * WAVES + wv, threadIdx.x & 63, lds, ring, xedgeP,
                                                     xedgeF, wv);
  else
    xy_body<EFLUX, R, false, ACC32, false, WAVES, W2>(a, stage, e.y * WAVES + wv, threadIdx.x & 63, lds, ring, xedgeP,
                                                      xedgeF, wv);
}

template <int EFLUX, typename R, int STAGE, bool ACC32>
__global__ __launch_bounds__(256, MinWaves<R>::value) void k_batch_z(const FusedArgs* __restrict__ boxes,
                                                                    const int2* __restrict__ tab) {
  const int2 e = tab[blockIdx.x];
  const FusedArgs& a = boxes[__builtin_amdgcn_readfirstlane(e.x)];
  __shared__ R zring[sizeof(R) == 8 ? WAVES * WIN_RING : 1];
  __shared__ double zraw[WAVES * Z_RAW];
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform over the wave
  z_body<EFLUX, R, STAGE, ACC32>(a, e.y * WAVES + wv, threadIdx.x & 63, zring + (sizeof(R) == 8 ? wv * WIN_RING : 0),
                          zraw + wv * Z_RAW);
}

#if MHD_FUSED_UNIT != 2
// step end of every live box (k_step_end with the fused ring offset), one thread per box
__global__ void k_batch_step_end(const FusedArgs* __restrict__ boxes, const int* __restrict__ live, int nlive,
                                 double* __restrict__ dt_last) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nlive) return;
  const int b = live[t];
  const FusedArgs& a = boxes[b];
  StepScalars* sc = a.sc;
  if (sc->stopped) return;
  const double dt = a.cfl * bits2d(sc->red[sc->iter % 3][0]);
  sc->time += dt;
  sc->steps_done += 1;
  sc->iter += 1;
  unsigned long long* nxt = sc->red[(sc->iter + 1) % 3];
  nxt[0] = DT_INIT_BITS;
  nxt[1] = CH_INIT_BITS;
  if (dt_last) dt_last[b] = dt;
  if (sc->tmax > 0.0 && sc->time > sc->tmax) sc->stopped = 1;
  if (sc->halt) sc->stopped = 1;  // the state this step left holds a bad cell
}
#endif

inline unsigned cdiv(long long a, long long b) { return (unsigned)((a + b - 1) / b); }

template <int EFLUX, typename R, bool ACC32, bool W2>
void launch_xy_w(const FusedArgs& a, hipStream_t s) {
  if constexpr (sizeof(R) == 8) {
    const int bw = a.duo ? xy_duo_waves(a) : 0;
    const dim3 grid(fused_blocks_xy(a)), block(64 * bw);
    switch (bw) {
      case 3: hipLaunchKernelGGL((k_fused_xy_duo<EFLUX, 3, W2>), grid, block, 0, s, a); return;
      case 4: hipLaunchKernelGGL((k_fused_xy_duo<EFLUX, 4, W2>), grid, block, 0, s, a); return;
      case 8: hipLaunchKernelGGL((k_fused_xy_duo<EFLUX, 8, W2>), grid, block, 0, s, a); return;
      default: break;
    }
  }
  if (xy_wide(a))
    hipLaunchKernelGGL((k_fused_xy<EFLUX, R, true, ACC32, W2>), dim3(fused_blocks_xy(a)), dim3(64 * WAVES),
                       fused_xy_lds(a), s, a);
  else
    hipLaunchKernelGGL((k_fused_xy<EFLUX, R, false, ACC32, W2>), dim3(fused_blocks_xy(a)), dim3(64 * WAVES),
                       fused_xy_lds(a), s, a);
}
// the second Runge-Kutta stage takes the W form where the divergence is stored in FP64
template <int EFLUX, typename R, bool ACC32>
void launch_xy_t(const FusedArgs& a, hipStream_t s) {
  if constexpr (!ACC32 && MHD_RK2_W) {
    if (a.stage == 2) {
      launch_xy_w<EFLUX, R, ACC32, true>(a, s);
      return;
    }
  }
  launch_xy_w<EFLUX, R, ACC32, false>(a, s);
}
template <int EFLUX, typename R, bool ACC32>
void launch_z_t(const FusedArgs& a, hipStream_t s) {
  const dim3 grid(fused_blocks_z(a)), block(64 * WAVES);
  if (a.stage == 1) hipLaunchKernelGGL((k_fused_z<EFLUX, R, 1, ACC32>), grid, block, 0, s, a);
  else if (a.stage == 2) hipLaunchKernelGGL((k_fused_z<EFLUX, R, 2, ACC32>), grid, block, 0, s, a);
  else hipLaunchKernelGGL((k_fused_z<EFLUX, R, 0, ACC32>), grid, block, 0, s, a);
}

template <int EFLUX, typename R, bool ACC32>
void launch_bxy_t(const BatchLaunch& b, hipStream_t s) {
  if constexpr (!ACC32 && MHD_RK2_W) {
    if (b.stage == 2) {
      hipLaunchKernelGGL((k_batch_xy<EFLUX, R, ACC32, true>), dim3(b.nblocks), dim3(64 * WAVES), b.lds, s, b.boxes,
                         b.tab, b.stage);
      return;
    }
  }
  hipLaunchKernelGGL((k_batch_xy<EFLUX, R, ACC32>), dim3(b.nblocks), dim3(64 * WAVES), b.lds, s, b.boxes, b.tab,
                     b.stage);
}
template <int EFLUX, typename R, bool ACC32>
void launch_bz_t(const BatchLaunch& b, hipStream_t s) {
  const dim3 grid(b.nblocks), block(64 * WAVES);
  if (b.stage == 1) hipLaunchKernelGGL((k_batch_z<EFLUX, R, 1, ACC32>), grid, block, 0, s, b.boxes, b.tab);
  else if (b.stage == 2) hipLaunchKernelGGL((k_batch_z<EFLUX, R, 2, ACC32>), grid, block, 0, s, b.boxes, b.tab);
  else hipLaunchKernelGGL((k_batch_z<EFLUX, R, 0, ACC32>), grid, block, 0, s, b.boxes, b.tab);
}

// FP32 launches: prec = mhd_config.precision (1 FP32 reconstruction + HLLD; 2 as 1 with the
// divergence stored in FP32)
#define MHD_DISPATCH_F32(fn, eflux, prec, a, s)                          \
  do {                                                                   \
    if ((eflux) == 0) {                                                  \
      if ((prec) == 2) fn<0, float, true>(a, s);                         \
      else fn<0, float, false>(a, s);                                    \
    } else {                                                             \
      if ((prec) == 2) fn<1, float, true>(a, s);                         \
      else fn<1, float, false>(a, s);                                    \
    }                                                                    \
  } while (0)

}  // namespace

#if MHD_FUSED_UNIT != 1
void launch_xy_f64(const FusedArgs& a, hipStream_t s) {
  if (a.eflux == 0) launch_xy_t<0, double, false>(a, s);
  else launch_xy_t<1, double, false>(a, s);
}
void launch_z_f64(const FusedArgs& a, hipStream_t s) {
  if (a.eflux == 0) launch_z_t<0, double, false>(a, s);
  else launch_z_t<1, double, false>(a, s);
}
void launch_bxy_f64(const BatchLaunch& b, int eflux, hipStream_t s) {
  if (eflux == 0) launch_bxy_t<0, double, false>(b, s);
  else launch_bxy_t<1, double, false>(b, s);
}
void launch_bz_f64(const BatchLaunch& b, int eflux, hipStream_t s) {
  if (eflux == 0) launch_bz_t<0, double, false>(b, s);
  else launch_bz_t<1, double, false>(b, s);
}
#endif

#if MHD_FUSED_UNIT != 2

size_t fused_xy_lds_max(int mixed) {
  if (mixed) return FUSED_XY_LDS_MAX;
  const size_t fixed = sizeof(double) * ((size_t)xy_ring_len<double>() + xy_edge_len<double>(XY_EDGE_R) +
                                         xy_edge_len<double>(NCONS));
  return std::min(FUSED_XY_LDS_MAX, XY_LDS_PER_CU - fixed);
}

bool fused_xy_two_blocks_ok(const FusedArgs& a) {
  FusedArgs d = a;
  d.mixed = 0;
  return xy_duo(d);
}
bool fused_xy_two_blocks(const FusedArgs& a) { return a.duo && xy_duo(a); }

int fused_ly_cap(const FusedArgs& a) {
  if (a.duo && xy_duo(a)) return 64;  // one row constant per lane, no pre-pass slots
  return std::min(64, fused_max_ly(a.sw, a.mixed));
}

int fused_max_ly(int sw, int mixed) {
  const size_t per_row = (size_t)WAVES * (mixed ? xy_wave_lds<float>(sw, 1) : xy_wave_lds<double>(sw, 1));
  return (int)(fused_xy_lds_max(mixed) / per_row);
}

size_t fused_xy_lds(const FusedArgs& a) {
  if (a.duo && xy_duo(a)) return 0;
  return (size_t)WAVES * (a.mixed ? xy_wave_lds<float>(a.sw, a.ly) : xy_wave_lds<double>(a.sw, a.ly));
}

bool fused_configure(FusedArgs& a, bool alone) {
  // strip width with the fewest idle lanes (ties: the wider strip).  In FP32, 64-cell strips
  // whenever at most a quarter of their lanes idle: there the narrower strips lose more to
  // their y chunks, which the pre-pass slots of 2 or 4 strips per wavefront cap at 32 or 16
  // rows, and to their edge faces than the wide strip loses to idle lanes (48- and 96-wide
  // boxes: 22 % and 5 % faster per cell in a batch of precision 1)
  int best = 64, best_lanes = 1 << 30;
  for (int sw : {64, 32, 16}) {
    const int lanes = (a.g.nx + sw - 1) / sw * sw;
    if (lanes < best_lanes) { best = sw; best_lanes = lanes; }
  }
  if (a.mixed) {
    const int lanes64 = (a.g.nx + 63) / 64 * 64;
    if (4 * (lanes64 - a.g.nx) <= lanes64) best = 64;
  } else {
    // FP64: 64-cell strips whenever at most an eighth of their lanes idle.  Their y chunks
    // may be twice as long (46 rows against 23) and they take the DPP shifts of full-width
    // strips: at 408^3, 600^3 and 798^3 (9, 6 and 4 % of the 64-cell lanes idle) the xy
    // kernel ran 12-17 % faster and the step 7-10 %; at 150^3 (22 % idle) 32-cell strips
    // stay faster
    const int lanes64 = (a.g.nx + 63) / 64 * 64;
    if (8 * (lanes64 - a.g.nx) <= lanes64) best = 64;
  }
  a.sw = best;
  a.ns = (a.g.nx + best - 1) / best;
  // Chunk lengths.  Each chunk of rows (planes) is one wavefront's march, and the chip holds
  // `slots` such wavefronts at once: one per SIMD (256 CUs x 4) in FP64, two in FP32, for
  // both kernels.  A box run alone splits ny (nz) into c chunks of at most 48 rows (64
  // planes) and no fewer than 8, with c minimising ceil(c * wavefronts per chunk / slots) *
  // (chunk + 4): the rounds of wavefronts the launch takes times the rows each marches, its
  // 4 halo rows (planes) included.  Longer chunks re-read fewer halo rows and run fewer
  // pre-passes; a last round that fills a fraction of the chip costs a whole one.  In a
  // batch (the other boxes fill the chip) 32 rows and 32 planes.
  auto split = [](int n, int cap, long long per_chunk, long long slots) {
    const int cmin = (n + cap - 1) / cap, cmax = std::max(cmin, (n + 7) / 8);
    long long best_cost = -1;
    int len = n;
    for (int c = cmin; c <= cmax; ++c) {
      const int l = (n + c - 1) / c;
      if (l < 8 && c > cmin) break;
      const long long cost = ((per_chunk * c + slots - 1) / slots) * (l + 4);
      if (best_cost < 0 || cost < best_cost) {
        best_cost = cost;
        len = l;
      }
    }
    return std::max(std::min(n, 8), len);
  };
  const long long slots = a.mixed ? 2048 : 1024;
  // the two-wavefronts-per-SIMD FP64 march (xy_duo) has no
  // pre-pass slots to cap its chunk (64 rows: one per lane of its row constants)
  const long long strips = (long long)a.ns * ((a.g.nz + 64 / best - 1) / (64 / best));
  bool duo = alone && xy_duo(a);
  if (duo) {
    // wavefronts a CU holds: whole blocks within its LDS, at most two per SIMD
    const int bw = xy_duo_waves(a);
    const long long per_cu = std::min<long long>(8, bw * (long long)(XY_LDS_PER_CU / xy_duo_lds(bw)));
    a.ly = split(a.g.ny, 64, strips, 256 * per_cu);
    // only where its chunks give more wavefronts than the chip has SIMDs: a smaller launch
    // leaves SIMDs idle either way and pays the block barriers for nothing (64^3: 7 % slower)
    duo = strips * ((a.g.ny + a.ly - 1) / a.ly) > 1024;
  }
  a.duo = duo ? 1 : 0;
  if (!duo) {
    const int ycap = std::min(alone ? 48 : 32, fused_max_ly(best, a.mixed));
    a.ly = alone ? split(a.g.ny, ycap, strips, slots) : std::min(ycap, a.g.ny);
  }
  const long long cols = (long long)a.ns * ((a.g.ny + 64 / best - 1) / (64 / best));
  a.lz = alone ? split(a.g.nz, 64, cols, slots) : std::min(32, a.g.nz);
  return a.g.vstride * sizeof(double) < (size_t(1) << 32);  // 32-bit column byte offsets
}

int fused_blocks_xy(const FusedArgs& a) {
  const int spw = 64 / a.sw;
  const long long waves = (long long)a.ns * ((a.g.nz + spw - 1) / spw) * ((a.g.ny + a.ly - 1) / a.ly);
  const int bw = (a.duo && xy_duo(a)) ? xy_duo_waves(a) : WAVES;
  return (int)cdiv(waves, bw);
}

int fused_blocks_z(const FusedArgs& a) {
  const int spw = 64 / a.sw;
  const long long waves = (long long)a.ns * ((a.g.ny + spw - 1) / spw) * ((a.g.nz + a.lz - 1) / a.lz);
  return (int)cdiv(waves, WAVES);
}

void launch_batch_xy(const FusedArgs* boxes, const int2* tab, int nblocks, size_t lds, int eflux, int mixed,
                     hipStream_t s, int stage) {
  if (nblocks <= 0) return;
  const BatchLaunch b{boxes, tab, nblocks, lds, stage};
  if (mixed) MHD_DISPATCH_F32(launch_bxy_t, eflux, mixed, b, s);
  else launch_bxy_f64(b, eflux, s);
}
void launch_batch_z(const FusedArgs* boxes, const int2* tab, int nblocks, int eflux, int mixed, hipStream_t s,
                    int stage) {
  if (nblocks <= 0) return;
  const BatchLaunch b{boxes, tab, nblocks, 0, stage};
  if (mixed) MHD_DISPATCH_F32(launch_bz_t, eflux, mixed, b, s);
  else launch_bz_f64(b, eflux, s);
}
void launch_batch_step_end(const FusedArgs* boxes, const int* live, int nlive, double* dt_last, hipStream_t s) {
  if (nlive > 0)
    hipLaunchKernelGGL(k_batch_step_end, dim3(cdiv(nlive, 256)), dim3(256), 0, s, boxes, live, nlive, dt_last);
}

void launch_fused_xy(const FusedArgs& a, hipStream_t s) {
  if (a.mixed) MHD_DISPATCH_F32(launch_xy_t, a.eflux, a.acc32 ? 2 : 1, a, s);
  else launch_xy_f64(a, s);
}
void launch_fused_z(const FusedArgs& a, hipStream_t s) {
  if (a.mixed) MHD_DISPATCH_F32(launch_z_t, a.eflux, a.acc32 ? 2 : 1, a, s);
  else launch_z_f64(a, s);
}
#endif  // MHD_FUSED_UNIT != 2

}  // namespace mhd
