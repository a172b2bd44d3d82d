// mhd_fused.hip -- fused step kernels for gfx950 (see mhd_fused.hpp for the data flow).
//
// Per loop iteration of src_cpp_kokkos_host/main.cpp:229-259 the fused path reads the
// conserved state once in each of the xy and z kernels (plus the x-strip edges), keeps
// primitives, slopes, interface states and fluxes in registers, and writes the flux
// divergence once and the new state once.  Every face flux is computed exactly once
// (the strip-edge x faces by the edge kernel), every limited slope once per cell and
// direction, as the reference's operation sequence allows: the same slope of cell f
// serves the right state of face f-1/2 and the left state of face f+1/2
// (KH/mhd.cpp:1134-1154).
#include "mhd_fused.hpp"
#include "mhd_physics.hpp"

#include <algorithm>

namespace mhd {

namespace {

constexpr int WAVES = 4;  // wavefronts per block; blocks carry independent waves

// FP64 march kernels need more than 256 VGPRs; one wave per SIMD lets them keep their
// window in AGPRs instead of scratch.  The FP32 variants fit two waves per SIMD.
#ifndef MHD_F64_MINWAVES
#define MHD_F64_MINWAVES 1
#endif
// MHD_PREFETCH=1: the march loads the raw state of the row two ahead, and the z kernel
// issues the loads of the cell it updates before that iteration's Riemann solve, so the
// load latency overlaps arithmetic even at one wavefront per SIMD.
#ifndef MHD_PREFETCH
#define MHD_PREFETCH 1
#endif
template <typename R>
struct MinWaves {
  static constexpr int value = sizeof(R) == 8 ? MHD_F64_MINWAVES : 2;
};

__device__ __forceinline__ double bits2d(unsigned long long b) { return __longlong_as_double((long long)b); }
__device__ __forceinline__ unsigned long long d2bits(double d) { return (unsigned long long)__double_as_longlong(d); }

// conserved state of one cell -> primitive vector in precision R (UpdatePrimitvP arithmetic in FP64)
template <typename R>
__device__ __forceinline__ void load_prim(const double* __restrict__ U, size_t vs, int c, double gam, R* q) {
  double u[NCONS], p[NPRIM];
#pragma unroll
  for (int m = 0; m < NCONS; ++m) u[m] = U[m * vs + c];
  cons_to_prim(u, gam, p);
#pragma unroll
  for (int n = 0; n < NPRIM; ++n) q[n] = R(p[n]);
}

__device__ __forceinline__ void load_raw(const double* __restrict__ U, size_t vs, int c, double* u) {
#pragma unroll
  for (int m = 0; m < NCONS; ++m) u[m] = U[m * vs + c];
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

// source of ghost gg (0,1) on side (0 lo, 1 hi) of an axis with interior [lo, hi]
// (SendRecvBoundary single-tile rules, KH/boundary.cpp:152-220)
__device__ __forceinline__ int ghost_src(int bc_lo, int bc_hi, int side, int gg, int lo, int hi, bool& neg) {
  neg = false;
  const int bc = side ? bc_hi : bc_lo;
  if (bc == BC_PERIODIC) return side ? lo + gg : hi - NGH + 1 + gg;
  if (bc == BC_REFLECT) { neg = true; return side ? hi - gg : lo + NGH - 1 - gg; }
  return side ? hi : lo;
}

// the ghost cells whose boundary image is interior cell (i,j,k) get its new state
__device__ __forceinline__ void write_images(double* __restrict__ Uo, const Geom& g, const int* bc, int i, int j,
                                             int k, const double* u) {
  const int n3[3] = {g.nx, g.ny, g.nz};
  const int idx[3] = {i, j, k};
#pragma unroll
  for (int ax = 0; ax < 3; ++ax) {
    const int c = idx[ax], lo = NGH, hi = NGH + n3[ax] - 1;
    if (c > lo + 1 && c < hi - 1) continue;
    for (int p = 0; p < 4; ++p) {
      const int side = p >> 1, gg = p & 1;
      bool neg;
      if (ghost_src(bc[2 * ax], bc[2 * ax + 1], side, gg, lo, hi, neg) != c) continue;
      int t[3] = {i, j, k};
      t[ax] = side ? hi + 1 + gg : gg;
      const int cc = g.cell(t[2], t[1], t[0]);
#pragma unroll
      for (int m = 0; m < NCONS; ++m) Uo[m * g.vstride + cc] = (neg && m == 1 + ax) ? -u[m] : u[m];
    }
  }
}

__device__ __forceinline__ size_t edge_index(const FusedArgs& a, int m, int k, int j, int e) {
  return (((size_t)m * a.g.nz + (k - NGH)) * a.g.ny + (j - NGH)) * (a.ns + 1) + e;
}

// ---------------------------------------------------------------- edge x faces
// x faces at the left edge of each strip and the right edge of the last one:
// face index f = NGH + min(e*sw, nx), e in [0, ns]; one thread per (e, j, k).
template <int EFLUX, typename R>
__device__ __forceinline__ void edge_body(const FusedArgs& a, long long t) {
  const StepScalars* sc = a.sc;
  if (sc->stopped) return;
  const Geom& g = a.g;
  const int iter = sc->iter;
  const double* __restrict__ U = (iter & 1) ? a.U1 : a.U0;
  const double chg = bits2d(sc->red[iter % 3][1]);
  const int ne = a.ns + 1;
  if (t >= (long long)ne * g.ny * g.nz) return;
  const int e = (int)(t % ne);
  const long long r = t / ne;
  const int j = NGH + (int)(r % g.ny);
  const int k = NGH + (int)(r / g.ny);
  const int f = NGH + min(e * a.sw, g.nx);
  R q[4][NPRIM];
#pragma unroll
  for (int s = 0; s < 4; ++s) load_prim<R>(U, g.vstride, g.cell(k, j, f - 2 + s), a.gam, q[s]);
  R ql[NPRIM], qr[NPRIM], tmp[NPRIM];
  cell_states<R>(q[0], q[1], q[2], ql, tmp);
  cell_states<R>(q[1], q[2], q[3], tmp, qr);
  double F[NCONS];
  riemann_flux<0, EFLUX, R>(ql, qr, chg, 0.5 / chg, F);
#pragma unroll
  for (int m = 0; m < NCONS; ++m) a.edge[edge_index(a, m, k, j, e)] = F[m];
}

template <int EFLUX, typename R>
__global__ __launch_bounds__(256) void k_fused_edge(FusedArgs a) {
  edge_body<EFLUX, R>(a, (long long)blockIdx.x * blockDim.x + threadIdx.x);
}

// ---------------------------------------------------------------- x + y divergence
// one wavefront: x-strip s of sw cells (64/sw strips stacked in k), rows [j0, j1) of chunk yc
template <int EFLUX, typename R>
__device__ __forceinline__ void xy_body(const FusedArgs& a, int w, int lane) {
  const StepScalars* sc = a.sc;
  if (sc->stopped) return;
  const Geom& g = a.g;
  const size_t vs = g.vstride;
  const int iter = sc->iter;
  const double* __restrict__ U = (iter & 1) ? a.U1 : a.U0;
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
  const double xinv = 1.0 / (a.x1a[i + 1] - a.x1a[i]);

  R Pm[NPRIM], P0[NPRIM], Pp[NPRIM], upprev[NPRIM];
  double Fyprev[NCONS];
  load_prim<R>(U, vs, g.cell(k, j0 - 2, i), a.gam, Pm);
  load_prim<R>(U, vs, g.cell(k, j0 - 1, i), a.gam, P0);
#pragma unroll
  for (int n = 0; n < NPRIM; ++n) upprev[n] = R(0);
#pragma unroll
  for (int m = 0; m < NCONS; ++m) Fyprev[m] = 0.0;
#if MHD_PREFETCH
  double un[NCONS];
  load_raw(U, vs, g.cell(k, j0, i), un);
#endif

  for (int j = j0 - 1; j <= j1; ++j) {
#if MHD_PREFETCH
    recover<R>(un, a.gam, Pp);
    load_raw(U, vs, g.cell(k, j + 2, i), un);  // row j+2 <= ny+4 < jtot
#else
    load_prim<R>(U, vs, g.cell(k, j + 1, i), a.gam, Pp);
#endif
    const int row = j - 1;  // the cell row completed in this pass
    const bool out_row = row >= j0;
    double ax[NCONS];
    if (out_row) {
      // x neighbours of row `row` (= Pm): shuffles inside the strip, loads at its edges
      R nb1[NPRIM], nb2[NPRIM];
      if (is_left || is_right) load_prim<R>(U, vs, g.cell(k, row, is_left ? i - 1 : i + 1), a.gam, nb1);
      if (is_left && is_right) load_prim<R>(U, vs, g.cell(k, row, i + 1), a.gam, nb2);
      R qm[NPRIM], qp[NPRIM];
#pragma unroll
      for (int n = 0; n < NPRIM; ++n) {
        const R vm = lane_get(Pm[n], src_l);
        const R vp = lane_get(Pm[n], src_r);
        qm[n] = is_left ? nb1[n] : vm;
        qp[n] = is_right ? (is_left ? nb2[n] : nb1[n]) : vp;
      }
      R rt[NPRIM], lt[NPRIM];
      cell_states<R>(qm, Pm, qp, rt, lt);
      R rtl[NPRIM];
#pragma unroll
      for (int n = 0; n < NPRIM; ++n) rtl[n] = lane_get(rt[n], src_l);
      double Fl[NCONS];
      riemann_flux<0, EFLUX, R>(rtl, lt, chg, hchg, Fl);
      if (is_left) {
#pragma unroll
        for (int m = 0; m < NCONS; ++m) Fl[m] = a.edge[edge_index(a, m, k, row, s)];
      }
      double Fr[NCONS];
#pragma unroll
      for (int m = 0; m < NCONS; ++m) Fr[m] = lane_get(Fl[m], src_r);
      if (is_right) {
#pragma unroll
        for (int m = 0; m < NCONS; ++m) Fr[m] = a.edge[edge_index(a, m, k, row, s + 1)];
      }
#pragma unroll
      for (int m = 0; m < NCONS; ++m) ax[m] = (Fr[m] - Fl[m]) * xinv;
    }
    // y states of row j and the y face j-1/2
    R up[NPRIM], dn[NPRIM];
    cell_states<R>(Pm, P0, Pp, up, dn);
    double Fy[NCONS] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (j >= j0) riemann_flux<1, EFLUX, R>(upprev, dn, chg, hchg, Fy);
    if (out_row && active) {
      const double yinv = 1.0 / (a.x2a[row + 1] - a.x2a[row]);
      const int c = g.cell(k, row, i);
#pragma unroll
      for (int m = 0; m < NCONS; ++m) a.acc[m * vs + c] = ax[m] + (Fy[m] - Fyprev[m]) * yinv;
    }
#pragma unroll
    for (int n = 0; n < NPRIM; ++n) {
      upprev[n] = up[n];
      Pm[n] = P0[n];
      P0[n] = Pp[n];
    }
#pragma unroll
    for (int m = 0; m < NCONS; ++m) Fyprev[m] = Fy[m];
  }
}

template <int EFLUX, typename R>
__global__ __launch_bounds__(256, MinWaves<R>::value) void k_fused_xy(FusedArgs a) {
  xy_body<EFLUX, R>(a, blockIdx.x * WAVES + (threadIdx.x >> 6), threadIdx.x & 63);
}

// ---------------------------------------------------------------- z + update + recovery + reduction
// one wavefront: 64 x-cells at row j, planes [k0, k1) of chunk zc
template <int EFLUX, typename R>
__device__ __forceinline__ void z_body(const FusedArgs& a, int w, int lane) {
  StepScalars* sc = a.sc;
  if (sc->stopped) return;
  const Geom& g = a.g;
  const size_t vs = g.vstride;
  const int iter = sc->iter;
  const double* __restrict__ U = (iter & 1) ? a.U1 : a.U0;
  double* __restrict__ Uo = (iter & 1) ? a.U0 : a.U1;
  const unsigned long long* rd = sc->red[iter % 3];
  const double dt = a.cfl * bits2d(rd[0]);
  const double chg = bits2d(rd[1]);
  const double hchg = 0.5 / chg;
  const int nxs = (g.nx + 63) / 64;
  const int nzc = (g.nz + a.lz - 1) / a.lz;
  const int s = w % nxs;
  const int t = w / nxs;
  const int jj = t % g.ny;
  const int zc = t / g.ny;
  if (zc >= nzc) return;  // uniform over the wavefront
  const int iraw = NGH + s * 64 + lane;
  const bool active = iraw < NGH + g.nx;
  const int i = active ? iraw : NGH + g.nx - 1;
  const int j = NGH + jj;
  const int k0 = NGH + zc * a.lz;
  const int k1 = min(k0 + a.lz, NGH + g.nz);
  const double dxa = a.x1a[i + 1] - a.x1a[i];
  const double dya = a.x2a[j + 1] - a.x2a[j];

  R Pm[NPRIM], P0[NPRIM], Pp[NPRIM], upprev[NPRIM];
  double Fzprev[NCONS];
  load_prim<R>(U, vs, g.cell(k0 - 2, j, i), a.gam, Pm);
  load_prim<R>(U, vs, g.cell(k0 - 1, j, i), a.gam, P0);
#pragma unroll
  for (int n = 0; n < NPRIM; ++n) upprev[n] = R(0);
#pragma unroll
  for (int m = 0; m < NCONS; ++m) Fzprev[m] = 0.0;
  double dtl = __builtin_inf(), chl = 0.0;
#if MHD_PREFETCH
  double un[NCONS];
  load_raw(U, vs, g.cell(k0, j, i), un);
#endif

  for (int k = k0 - 1; k <= k1; ++k) {
    const int kk = k - 1;  // the cell updated in this pass
    const bool upd = kk >= k0 && active;
#if MHD_PREFETCH
    recover<R>(un, a.gam, Pp);
    load_raw(U, vs, g.cell(k + 2, j, i), un);  // plane k+2 <= nz+4 < ktot
    double uo[NCONS], ac[NCONS];
    if (upd) {
      const int c = g.cell(kk, j, i);
#pragma unroll
      for (int m = 0; m < NCONS; ++m) {
        uo[m] = U[m * vs + c];
        ac[m] = a.acc[m * vs + c];
      }
    }
#else
    load_prim<R>(U, vs, g.cell(k + 1, j, i), a.gam, Pp);
#endif
    R up[NPRIM], dn[NPRIM];
    cell_states<R>(Pm, P0, Pp, up, dn);
    double Fz[NCONS] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (k >= k0) riemann_flux<2, EFLUX, R>(upprev, dn, chg, hchg, Fz);
    if (upd) {
      const double dza = a.x3a[kk + 1] - a.x3a[kk];
      const double zinv = 1.0 / dza;
      const int c = g.cell(kk, j, i);
      double u[NCONS];
#pragma unroll
      for (int m = 0; m < NCONS; ++m) {
#if MHD_PREFETCH
        const double div = ac[m] + (Fz[m] - Fzprev[m]) * zinv;
        u[m] = uo[m] - dt * div;
#else
        const double div = a.acc[m * vs + c] + (Fz[m] - Fzprev[m]) * zinv;
        u[m] = U[m * vs + c] - dt * div;
#endif
      }
      // DampPsi (KH/mhd.cpp:1625-1631)
      const double dhl = fmin(dxa, fmin(dya, dza));
      const double taui = a.alphabp * chg / dhl;
      u[mbps] = u[mbps] * (1.0 - dt * taui);
#pragma unroll
      for (int m = 0; m < NCONS; ++m) Uo[m * vs + c] = u[m];
      write_images(Uo, g, a.bc, i, j, kk, u);
      // the next iteration's ControlTimestep / EvaluateCh bounds of this cell
      double q[NPRIM];
      cons_to_prim(u, a.gam, q);
      dtl = fmin(dtl, cell_dt(q, dxa, dya, dza));
      chl = fmax(chl, cell_ch(q));
    }
#pragma unroll
    for (int n = 0; n < NPRIM; ++n) {
      upprev[n] = up[n];
      Pm[n] = P0[n];
      P0[n] = Pp[n];
    }
#pragma unroll
    for (int m = 0; m < NCONS; ++m) Fzprev[m] = Fz[m];
  }
  const bool bad = !(dtl > 0.0) || !(chl == chl) || isinf(chl);
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
    if (anybad) atomicOr(&sc->status, 1);
  }
}

template <int EFLUX, typename R>
__global__ __launch_bounds__(256, MinWaves<R>::value) void k_fused_z(FusedArgs a) {
  z_body<EFLUX, R>(a, blockIdx.x * WAVES + (threadIdx.x >> 6), threadIdx.x & 63);
}

// ---------------------------------------------------------------- batched launches
// Many boxes in one launch: tab[block] = (box, block index within that box).  Every box's
// block count is a whole number of blocks, so a block never straddles two boxes and the
// box descriptor read is uniform over the block.
template <int EFLUX, typename R>
__global__ __launch_bounds__(256) void k_batch_edge(const FusedArgs* __restrict__ boxes, const int2* __restrict__ tab) {
  const int2 e = tab[blockIdx.x];
  const FusedArgs& a = boxes[__builtin_amdgcn_readfirstlane(e.x)];
  edge_body<EFLUX, R>(a, (long long)e.y * blockDim.x + threadIdx.x);
}

template <int EFLUX, typename R>
__global__ __launch_bounds__(256, MinWaves<R>::value) void k_batch_xy(const FusedArgs* __restrict__ boxes,
                                                                     const int2* __restrict__ tab) {
  const int2 e = tab[blockIdx.x];
  const FusedArgs& a = boxes[__builtin_amdgcn_readfirstlane(e.x)];
  xy_body<EFLUX, R>(a, e.y * WAVES + (threadIdx.x >> 6), threadIdx.x & 63);
}

template <int EFLUX, typename R>
__global__ __launch_bounds__(256, MinWaves<R>::value) void k_batch_z(const FusedArgs* __restrict__ boxes,
                                                                    const int2* __restrict__ tab) {
  const int2 e = tab[blockIdx.x];
  const FusedArgs& a = boxes[__builtin_amdgcn_readfirstlane(e.x)];
  z_body<EFLUX, R>(a, e.y * WAVES + (threadIdx.x >> 6), threadIdx.x & 63);
}

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
}

inline unsigned cdiv(long long a, long long b) { return (unsigned)((a + b - 1) / b); }

template <int EFLUX, typename R>
void launch_xy_t(const FusedArgs& a, hipStream_t s) {
  hipLaunchKernelGGL((k_fused_xy<EFLUX, R>), dim3(fused_blocks_xy(a)), dim3(64 * WAVES), 0, s, a);
}

template <int EFLUX, typename R>
void launch_z_t(const FusedArgs& a, hipStream_t s) {
  hipLaunchKernelGGL((k_fused_z<EFLUX, R>), dim3(fused_blocks_z(a)), dim3(64 * WAVES), 0, s, a);
}

template <int EFLUX, typename R>
void launch_edge_t(const FusedArgs& a, hipStream_t s) {
  hipLaunchKernelGGL((k_fused_edge<EFLUX, R>), dim3(fused_blocks_edge(a)), dim3(256), 0, s, a);
}

struct BatchLaunch {
  const FusedArgs* boxes;
  const int2* tab;
  int nblocks;
};

template <int EFLUX, typename R>
void launch_bedge_t(const BatchLaunch& b, hipStream_t s) {
  hipLaunchKernelGGL((k_batch_edge<EFLUX, R>), dim3(b.nblocks), dim3(256), 0, s, b.boxes, b.tab);
}
template <int EFLUX, typename R>
void launch_bxy_t(const BatchLaunch& b, hipStream_t s) {
  hipLaunchKernelGGL((k_batch_xy<EFLUX, R>), dim3(b.nblocks), dim3(64 * WAVES), 0, s, b.boxes, b.tab);
}
template <int EFLUX, typename R>
void launch_bz_t(const BatchLaunch& b, hipStream_t s) {
  hipLaunchKernelGGL((k_batch_z<EFLUX, R>), dim3(b.nblocks), dim3(64 * WAVES), 0, s, b.boxes, b.tab);
}

#define MHD_DISPATCH2(fn, eflux, mixed, a, s)                 \
  do {                                                         \
    if ((eflux) == 0) {                                        \
      if (mixed) fn<0, float>(a, s); else fn<0, double>(a, s); \
    } else {                                                   \
      if (mixed) fn<1, float>(a, s); else fn<1, double>(a, s); \
    }                                                          \
  } while (0)
#define MHD_DISPATCH(fn, a, s) MHD_DISPATCH2(fn, (a).eflux, (a).mixed, a, s)

}  // namespace

void fused_configure(FusedArgs& a) {
  // strip width with the fewest idle lanes (ties: the wider strip)
  int best = 64, best_lanes = 1 << 30;
  for (int sw : {64, 32, 16}) {
    const int lanes = (a.g.nx + sw - 1) / sw * sw;
    if (lanes < best_lanes) { best = sw; best_lanes = lanes; }
  }
  a.sw = best;
  a.ns = (a.g.nx + best - 1) / best;
  a.ly = std::min(32, a.g.ny);
  a.lz = std::min(32, a.g.nz);
}

size_t fused_edge_elems(const FusedArgs& a) { return (size_t)NCONS * a.g.nz * a.g.ny * (a.ns + 1); }

int fused_blocks_edge(const FusedArgs& a) { return (int)cdiv((long long)(a.ns + 1) * a.g.ny * a.g.nz, 256); }

int fused_blocks_xy(const FusedArgs& a) {
  const int spw = 64 / a.sw;
  const long long waves = (long long)a.ns * ((a.g.nz + spw - 1) / spw) * ((a.g.ny + a.ly - 1) / a.ly);
  return (int)cdiv(waves, WAVES);
}

int fused_blocks_z(const FusedArgs& a) {
  const long long waves = (long long)((a.g.nx + 63) / 64) * a.g.ny * ((a.g.nz + a.lz - 1) / a.lz);
  return (int)cdiv(waves, WAVES);
}

void launch_batch_edge(const FusedArgs* boxes, const int2* tab, int nblocks, int eflux, int mixed, hipStream_t s) {
  if (nblocks > 0) MHD_DISPATCH2(launch_bedge_t, eflux, mixed, (BatchLaunch{boxes, tab, nblocks}), s);
}
void launch_batch_xy(const FusedArgs* boxes, const int2* tab, int nblocks, int eflux, int mixed, hipStream_t s) {
  if (nblocks > 0) MHD_DISPATCH2(launch_bxy_t, eflux, mixed, (BatchLaunch{boxes, tab, nblocks}), s);
}
void launch_batch_z(const FusedArgs* boxes, const int2* tab, int nblocks, int eflux, int mixed, hipStream_t s) {
  if (nblocks > 0) MHD_DISPATCH2(launch_bz_t, eflux, mixed, (BatchLaunch{boxes, tab, nblocks}), s);
}
void launch_batch_step_end(const FusedArgs* boxes, const int* live, int nlive, double* dt_last, hipStream_t s) {
  if (nlive > 0)
    hipLaunchKernelGGL(k_batch_step_end, dim3(cdiv(nlive, 256)), dim3(256), 0, s, boxes, live, nlive, dt_last);
}

void launch_fused_edge(const FusedArgs& a, hipStream_t s) { MHD_DISPATCH(launch_edge_t, a, s); }
void launch_fused_xy(const FusedArgs& a, hipStream_t s) { MHD_DISPATCH(launch_xy_t, a, s); }
void launch_fused_z(const FusedArgs& a, hipStream_t s) { MHD_DISPATCH(launch_z_t, a, s); }

}  // namespace mhd
