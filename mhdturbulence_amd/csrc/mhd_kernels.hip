// mhd_kernels.hip -- staged kernels of the MHD step (one launch per reference function),
// layout transposes, problem generators and diagnostics, for gfx950.
//
// The staged kernels are the device side of the reference-shaped C-ABI entry points
// (ControlTimestep, SetBoundaryCondition, EvaluateCh, GetNumericalFluxD, UpdateConservU,
// DampPsi, UpdatePrimitvP of src_cpp_kokkos_host).  The fused step kernel lives in
// mhd_fused.hip.
#include <algorithm>

#include "mhd_device.hpp"
#include "mhd_physics.hpp"

namespace mhd {

namespace {

constexpr int BLK = 256;

__device__ __forceinline__ double bits_to_double(unsigned long long b) { return __longlong_as_double((long long)b); }
__device__ __forceinline__ unsigned long long double_to_bits(double d) {
  return (unsigned long long)__double_as_longlong(d);
}

inline unsigned cdiv(long long a, long long b) { return (unsigned)((a + b - 1) / b); }

// ---------------------------------------------------------------- layout transposes
// host layout: ((k*jtot + j)*nv + n)*itot + i   (reference FieldArray, KH/mhd.hpp:164-175)
__global__ void k_from_host(Geom g, int nv, const double* __restrict__ src, double* __restrict__ dst) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int row = blockIdx.y;  // (k*jtot + j)*nv + n
  if (i >= g.itot) return;
  const int n = row % nv;
  const int kj = row / nv;
  const int j = kj % g.jtot, k = kj / g.jtot;
  dst[n * g.vstride + g.cell(k, j, i)] = src[(size_t)row * g.itot + i];
}

__global__ void k_to_host(Geom g, int nv, const double* __restrict__ src, double* __restrict__ dst) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int row = blockIdx.y;
  if (i >= g.itot) return;
  const int n = row % nv;
  const int kj = row / nv;
  const int j = kj % g.jtot, k = kj / g.jtot;
  dst[(size_t)row * g.itot + i] = src[n * g.vstride + g.cell(k, j, i)];
}

// ---------------------------------------------------------------- grid
// GenerateGrid (KH/main.cpp:26-65).  Contraction is off so each face is
// dx*(i-off) rounded, then + xmin rounded, exactly as the host computes it.
__global__ void k_grid(int n, int tot, double xmin, double xmax, int off, double* xa, double* xb) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= tot) return;
  const double dx = (xmax - xmin) / double(n);
  const double a0 = dx * (i - off) + xmin;
  xa[i] = a0;
  if (i < tot - 1) {
    const double a1 = dx * (i + 1 - off) + xmin;
    xb[i] = 0.5 * (a1 + a0);
  } else {
    xb[i] = 0.0;
  }
}

// ---------------------------------------------------------------- problem generators
__device__ __forceinline__ void store_state(const Geom& g, int c, const double* q, double* P, double* U) {
#pragma clang fp contract(off)
  for (int n = 0; n < NPRIM; ++n) P[n * g.vstride + c] = q[n];
  const double rho = q[nden];
  const double v1 = q[nve1], v2 = q[nve2], v3 = q[nve3];
  const double b1 = q[nbm1], b2 = q[nbm2], b3 = q[nbm3];
  const double ekin = 0.5 * rho * (v1 * v1 + v2 * v2 + v3 * v3);
  const double emag = 0.5 * (b1 * b1 + b2 * b2 + b3 * b3);
  U[mden * g.vstride + c] = rho;
  U[mrv1 * g.vstride + c] = rho * v1;
  U[mrv2 * g.vstride + c] = rho * v2;
  U[mrv3 * g.vstride + c] = rho * v3;
  U[meto * g.vstride + c] = rho * q[nene] + ekin + emag;
  U[mbm1 * g.vstride + c] = b1;
  U[mbm2 * g.vstride + c] = b2;
  U[mbm3 * g.vstride + c] = b3;
  U[mbps * g.vstride + c] = q[nbps];
  U[mst * g.vstride + c] = rho * q[nst];
}

// Kelvin-Helmholtz state of KH/main.cpp:66-184 (vy amplitude vy_amp, harmonic vx kick rrv)
__global__ void k_problem_kh(Geom g, double gam, double vy_amp, double rrv, double ymin, double ymax, double zmin,
                             double zmax, const double* __restrict__ x1b, const double* __restrict__ x2b,
                             const double* __restrict__ x3b, double* P, double* U) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  const int k = blockIdx.z;
  if (i >= g.itot - 1) return;
  const double pi = 3.141592653589793;  // std::acos(-1.0)
  const double dv = 2.0, wid = 0.05, sig = 0.2;
  const double x = x1b[i], y = x2b[j];
  const double tp = tanh((y + 0.5) / wid);
  const double tm = tanh((y - 0.5) / wid);
  double q[NPRIM];
  q[nden] = 1.0;
  q[npre] = 1.0;
  q[nve1] = 0.5 * dv * (tp - tm - 1.0);
  q[nve2] = vy_amp * sin(2.0 * pi * x) *
            (exp(-(y + 0.5) * (y + 0.5) / (sig * sig)) + exp(-(y - 0.5) * (y - 0.5) / (sig * sig)));
  q[nve3] = 0.0;
  q[nbm1] = q[nbm2] = q[nbm3] = q[nbps] = 0.0;
  q[nene] = q[npre] / (gam - 1.0);
  q[ncsp] = sqrt(gam * q[npre] / q[nden]);
  q[nst] = 0.5 * (tp - tm);
  if (i >= NGH && i < NGH + g.nx && j >= NGH && j < NGH + g.ny && k >= NGH && k < NGH + g.nz) {
    const double z = x3b[k];
    q[nve1] += dv * rrv * cos(6 * 2 * pi * (y - ymin) / (ymax - ymin)) * cos(6 * 2 * pi * (z - zmin) / (zmax - zmin));
  }
  store_state(g, g.cell(k, j, i), q, P, U);
}

__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ULL;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
  return x ^ (x >> 31);
}

// Single shear layer with a seeded per-cell vy kick; see mhd_generate_problem_shear.
__global__ void k_problem_shear(Geom g, double gam, double rho0, double p0, double v0, double a, double bx0,
                                double amp, double sigma, uint64_t seed, const double* __restrict__ x2b, double* P,
                                double* U) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  const int k = blockIdx.z;
  if (i >= g.itot - 1) return;
  const double y = x2b[j];
  const double th = tanh(y / a);
  double vy = 0.0;
  if (i >= NGH && i < NGH + g.nx && j >= NGH && j < NGH + g.ny && k >= NGH && k < NGH + g.nz) {
    const uint64_t id = ((uint64_t)(k - NGH) * g.ny + (j - NGH)) * g.nx + (i - NGH);
    const double u01 = (double)(splitmix64(seed ^ id) >> 11) * 0x1.0p-53;
    vy = amp * exp(-(y / sigma) * (y / sigma)) * (2.0 * u01 - 1.0);
  }
  double q[NPRIM];
  q[nden] = rho0;
  q[npre] = p0;
  q[nve1] = v0 * th;
  q[nve2] = vy;
  q[nve3] = 0.0;
  q[nbm1] = bx0;
  q[nbm2] = q[nbm3] = q[nbps] = 0.0;
  q[nene] = p0 / ((gam - 1.0) * rho0);
  q[ncsp] = sqrt(gam * p0 / rho0);
  q[nst] = 0.5 * (1.0 + th);
  store_state(g, g.cell(k, j, i), q, P, U);
}

// ---------------------------------------------------------------- reductions
__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

// Block-level (min dt, max ch) -> one atomic pair per block.  bad: this thread met a
// bad_cell; stop: also stop the device loop (the bounds are those of the step about to run).
__device__ __forceinline__ void block_commit(double dtl, double chl, bool bad, unsigned long long* acc,
                                             StepScalars* sc, bool stop) {
  __shared__ double s_dt[BLK / 64], s_ch[BLK / 64];
  __shared__ int s_bad;
  if (threadIdx.x == 0) s_bad = 0;
  __syncthreads();
  if (bad) s_bad = 1;
  dtl = wave_min(dtl);
  chl = wave_max(chl);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) { s_dt[w] = dtl; s_ch[w] = chl; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double d = s_dt[0], c = s_ch[0];
    for (int q = 1; q < BLK / 64; ++q) { d = fmin(d, s_dt[q]); c = fmax(c, s_ch[q]); }
    if (!(c == c)) c = __builtin_inf();
    atomicMin(&acc[0], double_to_bits(d));
    atomicMax(&acc[1], double_to_bits(c));
    if (s_bad) {
      atomicOr(&sc->status, 1);
      if (stop) atomicExch(&sc->stopped, 1);
    }
  }
}

__global__ __launch_bounds__(BLK) void k_reduce_prim(Geom g, const double* __restrict__ P,
                                                    const double* __restrict__ x1a, const double* __restrict__ x2a,
                                                    const double* __restrict__ x3a, StepScalars* sc, int ring) {
  const long long ncell = (long long)g.nx * g.ny * g.nz;
  double dtl = __builtin_inf(), chl = 0.0;
  bool bad = false;
  for (long long t = (long long)blockIdx.x * BLK + threadIdx.x; t < ncell; t += (long long)gridDim.x * BLK) {
    const int i = NGH + (int)(t % g.nx);
    const long long r = t / g.nx;
    const int j = NGH + (int)(r % g.ny);
    const int k = NGH + (int)(r / g.ny);
    const int c = g.cell(k, j, i);
    double q[NPRIM];
#pragma unroll
    for (int n = 0; n < NPRIM; ++n) q[n] = P[n * g.vstride + c];
    const double cdt = cell_dt(q, x1a[i + 1] - x1a[i], x2a[j + 1] - x2a[j], x3a[k + 1] - x3a[k]);
    const double cch = cell_ch(q);
    bad |= bad_cell(q[nden], cdt, cch);
    dtl = fmin(dtl, cdt);
    chl = fmax(chl, cch);
  }
  unsigned long long* acc = ring ? sc->red[sc->iter % 3] : sc->host_acc;
  block_commit(dtl, chl, bad, acc, sc, ring != 0);
}

__global__ __launch_bounds__(BLK) void k_zero(double* __restrict__ p, size_t n) {
  for (size_t t = (size_t)blockIdx.x * BLK + threadIdx.x; t < n; t += (size_t)gridDim.x * BLK) p[t] = 0.0;
}

__global__ void k_reset_scalars(StepScalars* sc, int ring) {
  if (threadIdx.x != 0) return;
  sc->host_acc[0] = DT_INIT_BITS;
  sc->host_acc[1] = CH_INIT_BITS;
  if (ring) {
    for (int q = 0; q < 3; ++q) { sc->red[q][0] = DT_INIT_BITS; sc->red[q][1] = CH_INIT_BITS; }
  }
}

// ---------------------------------------------------------------- boundary
// Ghost fill of SetBoundaryCondition (KH/boundary.cpp:750-893) for one tile per axis:
// periodic copies the opposite interior edge, reflection mirrors the first ngh interior
// cells and negates the normal velocity only, outflow copies the edge cell.  Only the
// ghosts the dimension-split step reads are written: x ghosts at interior (j,k), y ghosts
// at interior (i,k), z ghosts at interior (i,j).
__device__ __forceinline__ int ghost_source(int bc_in, int bc_out, int side, int g, int lo, int hi, bool& neg) {
  // side 0: ghost lo-NGH+g (g=0,1); side 1: ghost hi+1+g
  neg = false;
  const int bc = side ? bc_out : bc_in;
  if (bc == BC_PERIODIC) return side ? lo + g : hi - NGH + 1 + g;
  if (bc == BC_REFLECT) { neg = true; return side ? hi - g : lo + NGH - 1 - g; }
  return side ? hi : lo;
}

// nv variables; the normal velocity (P) or momentum (U) sits at index 1 + axis in both.
__global__ __launch_bounds__(BLK) void k_boundary(Geom g, int nv, int bx0, int bx1, int by0, int by1, int bz0, int bz1,
                                                  double* P) {
  const long long nxs = 4LL * g.ny * g.nz, nys = 4LL * g.nx * g.nz, nzs = 4LL * g.nx * g.ny;
  long long t = (long long)blockIdx.x * BLK + threadIdx.x;
  int dir, a, b, side, gg;
  if (t < nxs) {
    dir = 0; a = (int)(t % g.ny); t /= g.ny; b = (int)(t % g.nz); t /= g.nz; gg = (int)(t & 1); side = (int)(t >> 1);
  } else if ((t -= nxs) < nys) {
    dir = 1; a = (int)(t % g.nx); t /= g.nx; b = (int)(t % g.nz); t /= g.nz; gg = (int)(t & 1); side = (int)(t >> 1);
  } else if ((t -= nys) < nzs) {
    dir = 2; a = (int)(t % g.nx); t /= g.nx; b = (int)(t % g.ny); t /= g.ny; gg = (int)(t & 1); side = (int)(t >> 1);
  } else {
    return;
  }
  bool neg;
  int dst, src, nneg;
  if (dir == 0) {
    const int j = NGH + a, k = NGH + b;
    const int s = ghost_source(bx0, bx1, side, gg, NGH, NGH + g.nx - 1, neg);
    const int d = side ? NGH + g.nx + gg : gg;
    dst = g.cell(k, j, d); src = g.cell(k, j, s); nneg = nve1;
  } else if (dir == 1) {
    const int i = NGH + a, k = NGH + b;
    const int s = ghost_source(by0, by1, side, gg, NGH, NGH + g.ny - 1, neg);
    const int d = side ? NGH + g.ny + gg : gg;
    dst = g.cell(k, d, i); src = g.cell(k, s, i); nneg = nve2;
  } else {
    const int i = NGH + a, j = NGH + b;
    const int s = ghost_source(bz0, bz1, side, gg, NGH, NGH + g.nz - 1, neg);
    const int d = side ? NGH + g.nz + gg : gg;
    dst = g.cell(d, j, i); src = g.cell(s, j, i); nneg = nve3;
  }
  for (int n = 0; n < nv; ++n) {
    const double v = P[n * g.vstride + src];
    P[n * g.vstride + dst] = (neg && n == nneg) ? -v : v;
  }
}

// ---------------------------------------------------------------- fluxes
// GetNumericalFluxD (KH/mhd.cpp:987-1410), one thread per face.  Faces of direction D
// are indexed by the cell on their right: f in [lo, hi+1] along D, interior across.
template <int D, int EFLUX, typename R>
__global__ __launch_bounds__(BLK) void k_flux(Geom g, const double* __restrict__ P, double* __restrict__ F,
                                              double chg, const StepScalars* sc) {
  const int ni = g.nx + (D == 0), nj = g.ny + (D == 1), nk = g.nz + (D == 2);
  const int i = NGH + blockIdx.x * 64 + (threadIdx.x & 63);
  const int j = NGH + blockIdx.y * (BLK / 64) + (threadIdx.x >> 6);
  const int k = NGH + blockIdx.z;
  if (i >= NGH + ni || j >= NGH + nj || k >= NGH + nk) return;
  if (sc) chg = __longlong_as_double((long long)sc->red[sc->iter % 3][1]);
  const int c = g.cell(k, j, i);
  const int st = D == 0 ? 1 : (D == 1 ? g.pitch : g.pitch * g.jtot);
  R q[4][NPRIM];
#pragma unroll
  for (int s = 0; s < 4; ++s)
#pragma unroll
    for (int n = 0; n < NPRIM; ++n) q[s][n] = R(P[n * g.vstride + c + (s - 2) * st]);
  R ql[NPRIM], qr[NPRIM], tmp[NPRIM];
  cell_states<R>(q[0], q[1], q[2], ql, tmp);  // left state of this face from cell f-1
  cell_states<R>(q[1], q[2], q[3], tmp, qr);  // right state from cell f
  double Fo[NCONS];
  riemann_flux<D, EFLUX, R>(ql, qr, chg, 0.5 / chg, Fo);
#pragma unroll
  for (int n = 0; n < NCONS; ++n) F[n * g.vstride + c] = Fo[n];
}

// ---------------------------------------------------------------- update / damp / primitive
__device__ __forceinline__ bool loop_stopped(const StepScalars* sc) { return sc && sc->stopped; }

// UpdateConservU (KH/mhd.cpp:1412-1444), with DampPsi (KH/mhd.cpp:1613-1633) applied in the
// same pass when driven by the device loop.
__global__ __launch_bounds__(BLK) void k_update(Geom g, const double* __restrict__ x1a,
                                                const double* __restrict__ x2a, const double* __restrict__ x3a,
                                                const double* __restrict__ Fx, const double* __restrict__ Fy,
                                                const double* __restrict__ Fz, double* __restrict__ U, double dt,
                                                double cfl, double alphabp, const StepScalars* sc,
                                                double* __restrict__ save, const double* __restrict__ avg) {
  const int i = NGH + blockIdx.x * 64 + (threadIdx.x & 63);
  const int j = NGH + blockIdx.y * (BLK / 64) + (threadIdx.x >> 6);
  const int k = NGH + blockIdx.z;
  if (i >= NGH + g.nx || j >= NGH + g.ny) return;
  if (loop_stopped(sc)) return;
  double chg = 0.0;
  if (sc) {
    const unsigned long long* acc = sc->red[sc->iter % 3];
    dt = cfl * __longlong_as_double((long long)acc[0]);
    chg = __longlong_as_double((long long)acc[1]);
  }
  const double dxa = x1a[i + 1] - x1a[i], dya = x2a[j + 1] - x2a[j], dza = x3a[k + 1] - x3a[k];
  const double xinv = 1.0 / dxa, yinv = 1.0 / dya, zinv = 1.0 / dza;
  const int c = g.cell(k, j, i);
  const int sy = g.pitch, sz = g.pitch * g.jtot;
#pragma unroll
  for (int m = 0; m < NCONS; ++m) {
    const size_t o = m * g.vstride + c;
    double u = U[o];
    if (save) save[o] = u;
    u -= dt * ((Fx[o + 1] - Fx[o]) * xinv + (Fy[o + sy] - Fy[o]) * yinv + (Fz[o + sz] - Fz[o]) * zinv);
    if (m == mbps && sc) {
      const double dhl = fmin(dxa, fmin(dya, dza));
      const double taui = alphabp * chg / dhl;
      u = u * (1.0 - dt * taui);
    }
    if (avg) u = 0.5 * (avg[o] + u);
    U[o] = u;
  }
}

__global__ __launch_bounds__(BLK) void k_damp(Geom g, const double* __restrict__ x1a, const double* __restrict__ x2a,
                                              const double* __restrict__ x3a, double* __restrict__ U, double dt,
                                              double chg, double alphabp) {
  const int i = NGH + blockIdx.x * 64 + (threadIdx.x & 63);
  const int j = NGH + blockIdx.y * (BLK / 64) + (threadIdx.x >> 6);
  const int k = NGH + blockIdx.z;
  if (i >= NGH + g.nx || j >= NGH + g.ny) return;
  const double dhl = fmin(x1a[i + 1] - x1a[i], fmin(x2a[j + 1] - x2a[j], x3a[k + 1] - x3a[k]));
  const double taui = alphabp * chg / dhl;
  const size_t o = mbps * g.vstride + g.cell(k, j, i);
  U[o] = U[o] * (1.0 - dt * taui);
}

__global__ __launch_bounds__(BLK) void k_update_prim(Geom g, double gam, const double* __restrict__ U,
                                                     double* __restrict__ P, const StepScalars* sc) {
  const int i = NGH + blockIdx.x * 64 + (threadIdx.x & 63);
  const int j = NGH + blockIdx.y * (BLK / 64) + (threadIdx.x >> 6);
  const int k = NGH + blockIdx.z;
  if (i >= NGH + g.nx || j >= NGH + g.ny) return;
  if (loop_stopped(sc)) return;
  const int c = g.cell(k, j, i);
  double u[NCONS], q[NPRIM];
#pragma unroll
  for (int m = 0; m < NCONS; ++m) u[m] = U[m * g.vstride + c];
  cons_to_prim(u, gam, q);
#pragma unroll
  for (int n = 0; n < NPRIM; ++n) P[n * g.vstride + c] = q[n];
}

// interior cells of nv variables from src to dst; the ghost cells of dst keep their values
__global__ __launch_bounds__(BLK) void k_copy_interior(Geom g, const double* __restrict__ src,
                                                       double* __restrict__ dst, int nv) {
  const int i = NGH + blockIdx.x * 64 + (threadIdx.x & 63);
  const int j = NGH + blockIdx.y * (BLK / 64) + (threadIdx.x >> 6);
  const int k = NGH + blockIdx.z;
  if (i >= NGH + g.nx || j >= NGH + g.ny) return;
  const size_t c = g.cell(k, j, i);
  for (int m = 0; m < nv; ++m) dst[m * g.vstride + c] = src[m * g.vstride + c];
}

// ahead = 0: the next iteration reduces into slot iter % 3 (staged loop);
// ahead = 1: the next iteration reads slot iter % 3 and accumulates into (iter+1) % 3 (fused loop).
__global__ void k_step_end(StepScalars* sc, double cfl, double* dt_log, int ahead) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (sc->stopped) return;
  const double dt = cfl * __longlong_as_double((long long)sc->red[sc->iter % 3][0]);
  if (dt_log && sc->steps_done < sc->dt_log_cap) dt_log[sc->steps_done] = dt;
  sc->time += dt;
  sc->steps_done += 1;
  sc->iter += 1;
  unsigned long long* nxt = sc->red[(sc->iter + ahead) % 3];
  nxt[0] = DT_INIT_BITS;
  nxt[1] = CH_INIT_BITS;
  if (sc->tmax > 0.0 && sc->time > sc->tmax) sc->stopped = 1;
  // the fused loop reduces the next step's bounds in the step it applies: a bad cell in the
  // new state stops the loop after this step (the staged loop stops before the step instead)
  if (sc->halt) sc->stopped = 1;
}

// ---------------------------------------------------------------- diagnostics
// One block per interior k-plane: volume, sum X(1-X)dV, sum vy^2 dV (the sums of
// RealTimeAnalysis, src_cpp_omp_device/main.cpp:23-69) and max |divB| over cells whose six
// neighbours are interior.  Sums are reduced in a fixed tree, so they are reproducible.
__global__ __launch_bounds__(BLK) void k_diag(Geom g, const double* __restrict__ P, const double* __restrict__ x1a,
                                              const double* __restrict__ x1b, const double* __restrict__ x2a,
                                              const double* __restrict__ x2b, const double* __restrict__ x3a,
                                              const double* __restrict__ x3b, double* partial) {
  const int k = NGH + blockIdx.x;
  double sv = 0.0, sm = 0.0, sy = 0.0, md = 0.0;
  const int ie = NGH + g.nx - 1, je = NGH + g.ny - 1, ke = NGH + g.nz - 1;
  for (int t = threadIdx.x; t < g.nx * g.ny; t += BLK) {
    const int i = NGH + t % g.nx, j = NGH + t / g.nx;
    const int c = g.cell(k, j, i);
    const double dv = (x1a[i + 1] - x1a[i]) * (x2a[j + 1] - x2a[j]) * (x3a[k + 1] - x3a[k]);
    const double X = P[nst * g.vstride + c], vy = P[nve2 * g.vstride + c];
    sv += dv;
    sm += X * (1.0 - X) * dv;
    sy += vy * vy * dv;
    if (i > NGH && i < ie && j > NGH && j < je && k > NGH && k < ke) {
      const int sj = g.pitch, sk = g.pitch * g.jtot;
      const double* B1 = P + nbm1 * g.vstride;
      const double* B2 = P + nbm2 * g.vstride;
      const double* B3 = P + nbm3 * g.vstride;
      const double db = (B1[c + 1] - B1[c - 1]) / (x1b[i + 1] - x1b[i - 1]) +
                        (B2[c + sj] - B2[c - sj]) / (x2b[j + 1] - x2b[j - 1]) +
                        (B3[c + sk] - B3[c - sk]) / (x3b[k + 1] - x3b[k - 1]);
      md = fmax(md, fabs(db));
    }
  }
  __shared__ double s[4][BLK];
  s[0][threadIdx.x] = sv; s[1][threadIdx.x] = sm; s[2][threadIdx.x] = sy; s[3][threadIdx.x] = md;
  __syncthreads();
  for (int w = BLK / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      s[0][threadIdx.x] += s[0][threadIdx.x + w];
      s[1][threadIdx.x] += s[1][threadIdx.x + w];
      s[2][threadIdx.x] += s[2][threadIdx.x + w];
      s[3][threadIdx.x] = fmax(s[3][threadIdx.x], s[3][threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int q = 0; q < 4; ++q) partial[4 * blockIdx.x + q] = s[q][0];
}

}  // namespace

// ================================================================ launch wrappers
void launch_from_host_layout(const Geom& g, int nv, const double* src, double* dst, hipStream_t s) {
  dim3 grid(cdiv(g.itot, 256), g.ktot * g.jtot * nv);
  hipLaunchKernelGGL(k_from_host, grid, dim3(256), 0, s, g, nv, src, dst);
}

void launch_to_host_layout(const Geom& g, int nv, const double* src, double* dst, hipStream_t s) {
  dim3 grid(cdiv(g.itot, 256), g.ktot * g.jtot * nv);
  hipLaunchKernelGGL(k_to_host, grid, dim3(256), 0, s, g, nv, src, dst);
}

void launch_generate_grid(const Geom& g, const double* xmin, const double* xmax, int shift, double* x1a, double* x1b,
                          double* x2a, double* x2b, double* x3a, double* x3b, hipStream_t s) {
  const int off = NGH + (shift ? 1 : 0);
  hipLaunchKernelGGL(k_grid, dim3(cdiv(g.itot, 256)), dim3(256), 0, s, g.nx, g.itot, xmin[0], xmax[0], off, x1a, x1b);
  hipLaunchKernelGGL(k_grid, dim3(cdiv(g.jtot, 256)), dim3(256), 0, s, g.ny, g.jtot, xmin[1], xmax[1], off, x2a, x2b);
  hipLaunchKernelGGL(k_grid, dim3(cdiv(g.ktot, 256)), dim3(256), 0, s, g.nz, g.ktot, xmin[2], xmax[2], off, x3a, x3b);
}

void launch_problem_kh(const Geom& g, double gam, double vy_amp, double rrv, const double* xmin, const double* xmax,
                       const double* x1b, const double* x2b, const double* x3b, double* P, double* U, hipStream_t s) {
  dim3 grid(cdiv(g.itot - 1, 64), g.jtot - 1, g.ktot - 1);
  hipLaunchKernelGGL(k_problem_kh, grid, dim3(64), 0, s, g, gam, vy_amp, rrv, xmin[1], xmax[1], xmin[2], xmax[2],
                     x1b, x2b, x3b, P, U);
}

void launch_problem_shear(const Geom& g, double gam, const double* ic, uint64_t seed, const double* x2b, double* P,
                          double* U, hipStream_t s) {
  dim3 grid(cdiv(g.itot - 1, 64), g.jtot - 1, g.ktot - 1);
  hipLaunchKernelGGL(k_problem_shear, grid, dim3(64), 0, s, g, gam, ic[0], ic[1], ic[2], ic[3], ic[4], ic[5], ic[6],
                     seed, x2b, P, U);
}

void launch_reset_scalars(StepScalars* sc, int ring, hipStream_t s) {
  hipLaunchKernelGGL(k_reset_scalars, dim3(1), dim3(64), 0, s, sc, ring);
}

void launch_reduce_prim(const Geom& g, const double* P, const double* x1a, const double* x2a, const double* x3a,
                        StepScalars* sc, int ring, hipStream_t s) {
  const long long ncell = (long long)g.nx * g.ny * g.nz;
  const unsigned nb = (unsigned)std::min<long long>(4096, (ncell + BLK - 1) / BLK);
  hipLaunchKernelGGL(k_reduce_prim, dim3(nb), dim3(BLK), 0, s, g, P, x1a, x2a, x3a, sc, ring);
}

void launch_boundary(const Geom& g, const int* bc, double* A, int nv, hipStream_t s) {
  const long long n = 4LL * ((long long)g.ny * g.nz + (long long)g.nx * g.nz + (long long)g.nx * g.ny);
  hipLaunchKernelGGL(k_boundary, dim3(cdiv(n, BLK)), dim3(BLK), 0, s, g, nv, bc[0], bc[1], bc[2], bc[3], bc[4],
                     bc[5], A);
}

template <int D, int E, typename R>
static void launch_flux_t(const Geom& g, const double* P, double* F, double chg, const StepScalars* sc, hipStream_t s) {
  const int ni = g.nx + (D == 0), nj = g.ny + (D == 1), nk = g.nz + (D == 2);
  dim3 grid(cdiv(ni, 64), cdiv(nj, BLK / 64), nk);
  hipLaunchKernelGGL((k_flux<D, E, R>), grid, dim3(BLK), 0, s, g, P, F, chg, sc);
}

template <int D>
static void launch_flux_d(const Geom& g, int eflux, int mixed, const double* P, double* F, double chg,
                          const StepScalars* sc, hipStream_t s) {
  if (eflux == 0) {
    if (mixed) launch_flux_t<D, 0, float>(g, P, F, chg, sc, s);
    else launch_flux_t<D, 0, double>(g, P, F, chg, sc, s);
  } else {
    if (mixed) launch_flux_t<D, 1, float>(g, P, F, chg, sc, s);
    else launch_flux_t<D, 1, double>(g, P, F, chg, sc, s);
  }
}

void launch_flux(const Geom& g, int dir, int eflux, int mixed, const double* P, double* F, double chg,
                 const StepScalars* sc, hipStream_t s) {
  if (dir == 1) launch_flux_d<0>(g, eflux, mixed, P, F, chg, sc, s);
  else if (dir == 2) launch_flux_d<1>(g, eflux, mixed, P, F, chg, sc, s);
  else launch_flux_d<2>(g, eflux, mixed, P, F, chg, sc, s);
}

void launch_update(const Geom& g, const double* x1a, const double* x2a, const double* x3a, const double* Fx,
                   const double* Fy, const double* Fz, double* U, double dt, double cfl, double alphabp,
                   const StepScalars* sc, hipStream_t s, double* save, const double* avg) {
  dim3 grid(cdiv(g.nx, 64), cdiv(g.ny, BLK / 64), g.nz);
  hipLaunchKernelGGL(k_update, grid, dim3(BLK), 0, s, g, x1a, x2a, x3a, Fx, Fy, Fz, U, dt, cfl, alphabp, sc, save,
                     avg);
}

void launch_damp(const Geom& g, const double* x1a, const double* x2a, const double* x3a, double* U, double dt,
                 double chg, double alphabp, hipStream_t s) {
  dim3 grid(cdiv(g.nx, 64), cdiv(g.ny, BLK / 64), g.nz);
  hipLaunchKernelGGL(k_damp, grid, dim3(BLK), 0, s, g, x1a, x2a, x3a, U, dt, chg, alphabp);
}

void launch_update_prim(const Geom& g, double gam, const double* U, double* P, const StepScalars* sc, hipStream_t s) {
  dim3 grid(cdiv(g.nx, 64), cdiv(g.ny, BLK / 64), g.nz);
  hipLaunchKernelGGL(k_update_prim, grid, dim3(BLK), 0, s, g, gam, U, P, sc);
}

void launch_copy_interior(const Geom& g, const double* src, double* dst, int nv, hipStream_t s) {
  dim3 grid(cdiv(g.nx, 64), cdiv(g.ny, BLK / 64), g.nz);
  hipLaunchKernelGGL(k_copy_interior, grid, dim3(BLK), 0, s, g, src, dst, nv);
}

void launch_zero(double* p, size_t n, hipStream_t s) {
  if (n == 0) return;
  const unsigned blocks = (unsigned)std::min<size_t>(4096, (n + BLK - 1) / BLK);
  hipLaunchKernelGGL(k_zero, dim3(blocks), dim3(BLK), 0, s, p, n);
}

void launch_step_end(StepScalars* sc, double cfl, double* dt_log, int ahead, hipStream_t s) {
  hipLaunchKernelGGL(k_step_end, dim3(1), dim3(64), 0, s, sc, cfl, dt_log, ahead);
}

void launch_diagnostics(const Geom& g, const double* P, const double* x1a, const double* x1b, const double* x2a,
                        const double* x2b, const double* x3a, const double* x3b, double* partial, hipStream_t s) {
  hipLaunchKernelGGL(k_diag, dim3(g.nz), dim3(BLK), 0, s, g, P, x1a, x1b, x2a, x2b, x3a, x3b, partial);
}

}  // namespace mhd
