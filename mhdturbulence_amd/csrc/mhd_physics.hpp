// mhd_physics.hpp -- per-point arithmetic of the MHD step for CDNA4 (gfx950).
//
// Device functions shared by the staged kernels (one launch per reference function)
// and the fused step kernel.  They are templated on the reconstruction / Riemann
// precision R (double, or float for the mixed-precision path) and on the sweep
// direction D (0 = x, 1 = y, 2 = z) so each direction compiles to straight-line code
// with the component permutation resolved at compile time.
//
// Reference: src_cpp_kokkos_host/mhd.cpp (cited as KH/mhd.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include "mhd_device.hpp"

namespace mhd {

enum Prim { nden = 0, nve1, nve2, nve3, nene, npre, ncsp, nbm1, nbm2, nbm3, nbps, nst };
enum Cons { mden = 0, mrv1, mrv2, mrv3, meto, mbm1, mbm2, mbm3, mbps, mst };

// slot of Cartesian component c in the rotated (u normal, v, w) frame of direction D:
// D=0: (x,y,z)->(u,v,w), D=1: (y,z,x)->(u,v,w), D=2: (z,x,y)->(u,v,w)   (KH/mhd.cpp:1078-1118)
__host__ __device__ constexpr int slot(int c, int d) { return (c - d + 3) % 3; }
// Cartesian component sitting in rotated slot s for direction D
__host__ __device__ constexpr int comp(int s, int d) { return (s + d) % 3; }

// Division of the hot paths.  MHD_FAST_DIV=1 (default): FP64 uses the hardware reciprocal,
// one Newton step and one residual correction.  With r0 = (1/b)(1+e) the Newton step leaves
// a relative error e^2 in r, and the correction q + r (a - b q) squares that again, so for
// any reciprocal seed better than 2^-14 the quotient is within an ulp of the IEEE one (no
// special-case scaling).  FP32 uses the hardware reciprocal.  Operands are finite and
// bounded away from zero wherever the step divides (densities, wave-speed differences
// guarded by +-1e-20, sums of same-sign slopes).
#ifndef MHD_FAST_DIV
#define MHD_FAST_DIV 1
#endif
template <typename R>
__device__ __forceinline__ R dv(R a, R b) {
  return a / b;
}
#if MHD_FAST_DIV
template <>
__device__ __forceinline__ double dv<double>(double a, double b) {
  double r = __builtin_amdgcn_rcp(b);
  r = fma(fma(-b, r, 1.0), r, r);
  const double q = a * r;
  return fma(fma(-b, q, a), r, q);
}
template <>
__device__ __forceinline__ float dv<float>(float a, float b) {
  return a * __builtin_amdgcn_rcpf(b);
}
#endif

// Reciprocals that share one division (MHD_SHARED_RCP=1, default, FP64 only): 1/a = b/(ab)
// and 1/b = a/(ab), and four at once through 1/(ab) and 1/(cd).  Each result is within a
// few ulp of the IEEE reciprocal.  Used only where the solver's floors keep every product
// far inside the FP64 range: densities (taken to lie within 1e-150..1e150), wave-speed
// differences (|x| >= 1e-20 by construction) and star-region denominators (|x| >= 1e-30).
// FP32 divides each one, since its range would not hold such products.
#ifndef MHD_SHARED_RCP
#define MHD_SHARED_RCP 1
#endif
template <typename R>
__device__ __forceinline__ void recip2(R a, R b, R& ia, R& ib) {
  if constexpr (MHD_SHARED_RCP && sizeof(R) == 8) {
    const R r = dv(R(1), a * b);
    ia = b * r;
    ib = a * r;
  } else {
    ia = dv(R(1), a);
    ib = dv(R(1), b);
  }
}
template <typename R>
__device__ __forceinline__ void recip4(R a, R b, R c, R d, R& ia, R& ib, R& ic, R& id) {
  if constexpr (MHD_SHARED_RCP && sizeof(R) == 8) {
    R iab, icd;
    recip2(a * b, c * d, iab, icd);
    ia = b * iab;
    ib = a * iab;
    ic = d * icd;
    id = c * icd;
  } else {
    ia = dv(R(1), a);
    ib = dv(R(1), b);
    ic = dv(R(1), c);
    id = dv(R(1), d);
  }
}

// Square root of the hot paths.  MHD_FAST_SQRT=1 (default): FP64 from the hardware
// reciprocal square root y, one Goldschmidt step (g = x y, h = y/2, r = 1/2 - h g) and one
// residual correction g + (x - g^2) h: the error of y is squared twice, leaving the result
// within an ulp of the IEEE root.  It skips the denormal/huge-argument rescaling, which the
// step's arguments (sound and Alfven speeds squared, densities) never need; x = 0 returns 0
// and x < 0 returns NaN as sqrt does.  FP32 uses the hardware square root.
// MHD_SQRT_FLOOR=1 (default): where x may be 0, the reciprocal root is taken of x + 1e-300
// instead of selecting x = 0 afterwards.  The root itself is x y refined, which is exactly
// x (+0 or -0) when x is, and for x >= 2e-284 the floor is under half an ulp of x, so
// x + 1e-300 == x and the result is the one without the floor.
#ifndef MHD_FAST_SQRT
#define MHD_FAST_SQRT 1
#endif
#ifndef MHD_SQRT_FLOOR
#define MHD_SQRT_FLOOR 1
#endif
template <typename R>
__device__ __forceinline__ R sq(R x) {
  return sqrt(x);
}
// sq_pos: arguments known to be positive (star-region densities), no x = 0 case
template <typename R>
__device__ __forceinline__ R sq_pos(R x) {
  return sqrt(x);
}
#if MHD_FAST_SQRT
// y: the reciprocal square root of x (or of x + a floor, see above)
__device__ __forceinline__ double sq_rsq(double x, double y) {
  double g = x * y, h = 0.5 * y;
  const double r = fma(-h, g, 0.5);
  g = fma(g, r, g);
  h = fma(h, r, h);
  return fma(fma(-g, g, x), h, g);
}
__device__ __forceinline__ double sq_rsq(double x) { return sq_rsq(x, __builtin_amdgcn_rsq(x)); }
template <>
__device__ __forceinline__ double sq<double>(double x) {
  if constexpr (MHD_SQRT_FLOOR) {
    return sq_rsq(x, __builtin_amdgcn_rsq(x + 1.0e-300));
  } else {
    const double g = sq_rsq(x);
    return x == 0.0 ? x : g;
  }
}
template <>
__device__ __forceinline__ double sq_pos<double>(double x) {
  return sq_rsq(x);
}
// sqrt(x) and 1/sqrt(x) of a positive x from one reciprocal square root: the Goldschmidt
// pair (g, h) -> (sqrt x, 1/(2 sqrt x)) takes a second step for h, so b / sqrt(x) becomes
// the product b * (2h) instead of a division (both within about an ulp)
__device__ __forceinline__ void sq_rsq_pos(double x, double& s, double& rs) {
  const double y = __builtin_amdgcn_rsq(x);
  double g = x * y, h = 0.5 * y;
  double r = fma(-h, g, 0.5);
  g = fma(g, r, g);
  h = fma(h, r, h);
  r = fma(-h, g, 0.5);
  s = fma(fma(-g, g, x), h, g);
  rs = 2.0 * fma(h, r, h);
}
template <>
__device__ __forceinline__ float sq<float>(float x) {
  return __builtin_amdgcn_sqrtf(x);
}
template <>
__device__ __forceinline__ float sq_pos<float>(float x) {
  return __builtin_amdgcn_sqrtf(x);
}
#endif
#if !MHD_FAST_SQRT
__device__ __forceinline__ void sq_rsq_pos(double x, double& s, double& rs) {
  s = sqrt(x);
  rs = 1.0 / s;
}
#endif
__device__ __forceinline__ void sq_rsq_pos(float x, float& s, float& rs) {
  s = sq_pos(x);
  rs = dv(1.0f, s);
}

// Half the van Leer harmonic limiter of KH/mhd.cpp:89-95: vL(dp, dm) / 2 = dp dm / (dp + dm)
// where dp dm > 0, else 0.  Scaling by 2 is exact in binary floating point (away from the
// subnormal range) and commutes with every rounding of the quotient, so q0 +- h is the
// reference's q0 +- 0.5 * (2 dp dm / (dp + dm)) without the two scalings.
// MHD_LIM_DIV=1 (default): in FP64 the quotient takes the reciprocal and its Newton step but
// no residual correction.  The hardware reciprocal is good to 2^-24.4 on gfx950 (4.6e-8
// relative, measured over 6.7e7 arguments), so after the Newton step h is within 2.2e-15
// relative of the quotient, about 10 ulp (h enters the
// states as q0 +- h, a fraction of the cell value).
// MHD_LIM_FLOOR=1 (default, FP64): no select.  The reciprocal is taken of (dp + dm) + 1e-300,
// equal to dp + dm wherever |dp + dm| >= 2e-284, and max(dp dm, 0) replaces the product, so
// a pair of slopes of opposite sign (or a zero) gives a zero, which is -0 where dp + dm < 0:
// q0 +- h is then q0 except that a q0 of -0 may come back as +0 (and the reverse).
#ifndef MHD_LIM_DIV
#define MHD_LIM_DIV 1
#endif
#ifndef MHD_LIM_FLOOR
#define MHD_LIM_FLOOR 1
#endif
template <typename R>
__device__ __forceinline__ R half_van_leer(R dp, R dm) {
  const R prod = dp * dm;
  R h;
  if constexpr (MHD_LIM_DIV && MHD_FAST_DIV && sizeof(R) == 8) {
    const R s = dp + dm;
    if constexpr (MHD_LIM_FLOOR) {
      // where dp dm <= 0, s may be 0: the reciprocal of s + 1e-300 stays finite and the
      // product with max(dp dm, 0) = 0 is a zero
      R r = __builtin_amdgcn_rcp(s + R(1.0e-300));
      r = fma(fma(-s, r, R(1)), r, r);
      return fmax(prod, R(0)) * r;
    }
    R r = __builtin_amdgcn_rcp(s);
    r = fma(fma(-s, r, R(1)), r, r);
    h = prod * r;
  } else {
    h = dv(prod, dp + dm);
  }
  return prod > R(0) ? h : R(0);
}
// The floored FP64 limiter as the factors of h = max(dp dm, 0) r, so that the caller forms
// q0 +- h as two fused multiply-adds (one rounding each instead of two)
constexpr bool LIM_FACTORS = MHD_LIM_DIV && MHD_FAST_DIV && MHD_LIM_FLOOR;
__device__ __forceinline__ void half_van_leer_factors(double dp, double dm, double& pm, double& r) {
  const double s = dp + dm;
  r = __builtin_amdgcn_rcp(s + 1.0e-300);
  r = fma(fma(-s, r, 1.0), r, r);
  pm = fmax(dp * dm, 0.0);
}

// One interface state in the rotated frame: the quantities HLLD reads from the
// 23-entry Riemann vector of KH/mhd.cpp:997-1076, kept as named registers.
template <typename R>
struct RState {
  R rho, mu, mv, mw, E, bu, bv, bw, psi, sc;    // rotated conserved
  R fd, fu, fv, fw, fE, fbv, fbw, fsc;           // rotated physical flux (fbu = fpsi = 0)
  R cf, vn, pt;                                  // fast speed, normal velocity, total pressure
  R vv, vw, X;                                   // transverse velocities and passive scalar
};

// Prim2ConsD for direction D (KH/mhd.cpp:997-1076).  EFLUX=0: energy flux magnetic term
// |B|^2 B_n as written at KH/mhd.cpp:1049-1050; EFLUX=1: (v.B) B_n.
// rhoinv: 1/rho, computed by the caller for both states of a face
template <int D, int EFLUX, typename R>
__device__ __forceinline__ void prim_to_state(const R* q, R rhoinv, RState<R>& s) {
  constexpr int DF = (D + 1) % 3, DB = (D + 2) % 3;
  const R rho = q[nden];
  const R vsq = q[nve1] * q[nve1] + q[nve2] * q[nve2] + q[nve3] * q[nve3];
  const R bsq = q[nbm1] * q[nbm1] + q[nbm2] * q[nbm2] + q[nbm3] * q[nbm3];
  const R vn = q[nve1 + D], bn = q[nbm1 + D];
  const R vf = q[nve1 + DF], bf = q[nbm1 + DF];
  const R vb = q[nve1 + DB], bb = q[nbm1 + DB];
  s.rho = rho;
  s.mu = rho * vn;
  s.mv = rho * vf;
  s.mw = rho * vb;
  s.E = rho * q[nene] + R(0.5) * rho * vsq + R(0.5) * bsq;
  s.bu = bn;
  s.bv = bf;
  s.bw = bb;
  s.psi = q[nbps];
  s.sc = rho * q[nst];
  const R pt = q[npre] + bsq / R(2);
  s.fd = rho * vn;
  s.fu = rho * vn * vn - bn * bn + pt;
  s.fv = rho * vf * vn - bf * bn;
  s.fw = rho * vb * vn - bb * bn;
  if constexpr (EFLUX == 0) {
    s.fE = (s.E + pt) * vn - bsq * bn;
  } else {
    const R vdb = q[nve1] * q[nbm1] + q[nve2] * q[nbm2] + q[nve3] * q[nbm3];
    s.fE = (s.E + pt) * vn - vdb * bn;
  }
  s.fbv = bf * vn - vf * bn;
  s.fbw = bb * vn - vb * bn;
  s.fsc = rho * q[nst] * vn;
  const R css = q[ncsp] * q[ncsp];
  const R cts = css + bsq * rhoinv;
  s.cf = sq((cts + sq(cts * cts - R(4) * css * bn * bn * rhoinv)) / R(2));
  s.vn = vn;
  s.pt = pt;
  s.vv = vf;
  s.vw = vb;
  s.X = q[nst];
}

// Rotated-frame numerical flux from HLLD (Miyoshi & Kusano 2005, J. Comput. Phys. 208:315)
// in the branch-free form of KH/mhd.cpp:97-417; the normal-B and psi slots are left to GLM.
template <typename R>
struct RFlux {
  R d, u, v, w, E, bv, bw, sc;
};

// std::copysign(1, x) as the reference writes it (so -0.0 selects -1)
template <typename R>
__device__ __forceinline__ R copysign1(R x) {
  return copysign(R(1), x);
}

template <typename R>
__device__ __forceinline__ void hlld(const RState<R>& L, const RState<R>& Rt, RFlux<R>& nf) {
  const R eps = R(1.0e-30);
  const R tiny = R(1.0e-20);
  const R rol = L.rho, eel = L.E, rxl = L.mu, ryl = L.mv, rzl = L.mw;
  // (rho v)/rho of KH/mhd.cpp:173-175 is the reconstructed velocity up to an ulp
  const R vxl = L.vn, vyl = L.vv, vzl = L.vw;
  const R byl = L.bv, bzl = L.bw, ptl = L.pt, scl = L.sc;
  const R ror = Rt.rho, eer = Rt.E, rxr = Rt.mu, ryr = Rt.mv, rzr = Rt.mw;
  const R vxr = Rt.vn, vyr = Rt.vv, vzr = Rt.vw;
  const R byr = Rt.bv, bzr = Rt.bw, ptr = Rt.pt, scr = Rt.sc;

  const R cfmax = fmax(L.cf, Rt.cf);
  const R sl = fmin(vxl, vxr) - cfmax;
  const R sr = fmax(vxl, vxr) + cfmax;

  const R sdl = fmin(-tiny, sl - vxl);
  const R sdr = fmax(tiny, sr - vxr);
  const R rosdl = rol * sdl;
  const R rosdr = ror * sdr;
  R temp = dv(R(1), rosdr - rosdl);
  const R sm = (rosdr * vxr - rosdl * vxl - ptr + ptl) * temp;
  const R sdml = fmin(-tiny, sl - sm);
  const R sdmr = fmax(tiny, sr - sm);
  const R ptst = (rosdr * ptl - rosdl * ptr + rosdl * rosdr * (vxr - vxl)) * temp;

  // star-state denominators of both sides and their reciprocals (KH/mhd.cpp:268-276, 304-312)
  const R bxs = R(0.5) * (L.bu + Rt.bu);
  const R bxsq = bxs * bxs;
  const R tfl = rosdl * sdml - bxsq;
  const R sgl = copysign1(fabs(tfl) - eps);
  const R tfr = rosdr * sdmr - bxsq;
  const R sgr = copysign1(fabs(tfr) - eps);
  R itfl, isdml, itfr, isdmr;
  recip4(tfl + fmin(R(0), sgl), sdml, tfr + fmin(R(0), sgr), sdmr, itfl, isdml, itfr, isdmr);

  // left star state (KH/mhd.cpp:268-302)
  R maxs1 = fmax(R(0), sgl);
  R mins1 = fmin(R(0), sgl);
  R itf = itfl;
  temp = bxs * (sdl - sdml) * itf;
  const R rolst = maxs1 * (rosdl * isdml) - mins1 * rol;
  const R sclst = rolst * L.X;  // rolst/rol*scl with scl/rol = X
  const R vxlst = maxs1 * sm - mins1 * vxl;
  const R rxlst = rolst * vxlst;
  const R vylst = maxs1 * (vyl - byl * temp) - mins1 * vyl;
  const R rylst = rolst * vylst;
  const R vzlst = maxs1 * (vzl - bzl * temp) - mins1 * vzl;
  const R rzlst = rolst * vzlst;
  temp = (rosdl * sdl - bxsq) * itf;
  const R bylst = maxs1 * (byl * temp) - mins1 * byl;
  const R bzlst = maxs1 * (bzl * temp) - mins1 * bzl;
  const R vdbstl = vxlst * bxs + vylst * bylst + vzlst * bzlst;
  const R eelst =
      maxs1 * (sdl * eel - ptl * vxl + ptst * sm + bxs * (vxl * bxs + vyl * byl + vzl * bzl - vdbstl)) * isdml -
      mins1 * eel;

  // right star state (KH/mhd.cpp:304-333)
  maxs1 = fmax(R(0), sgr);
  mins1 = fmin(R(0), sgr);
  itf = itfr;
  temp = bxs * (sdr - sdmr) * itf;
  const R rorst = maxs1 * (rosdr * isdmr) - mins1 * ror;
  const R scrst = rorst * Rt.X;
  const R vxrst = maxs1 * sm - mins1 * vxr;
  const R rxrst = rorst * vxrst;
  const R vyrst = maxs1 * (vyr - byr * temp) - mins1 * vyr;
  const R ryrst = rorst * vyrst;
  const R vzrst = maxs1 * (vzr - bzr * temp) - mins1 * vzr;
  const R rzrst = rorst * vzrst;
  temp = (rosdr * sdr - bxsq) * itf;
  const R byrst = maxs1 * (byr * temp) - mins1 * byr;
  const R bzrst = maxs1 * (bzr * temp) - mins1 * bzr;
  const R vdbstr = vxrst * bxs + vyrst * byrst + vzrst * bzrst;
  const R eerst = maxs1 * ((sdr * eer - ptr * vxr + ptst * sm) * isdmr +
                           bxs * (vxr * bxs + vyr * byr + vzr * bzr - vdbstr) * isdmr) -
                  mins1 * eer;

  // double-star states (KH/mhd.cpp:337-377)
  R sqrtrol, sqrtror, rsqrtrol, rsqrtror;
  sq_rsq_pos(rolst, sqrtrol, rsqrtrol);
  sq_rsq_pos(rorst, sqrtror, rsqrtror);
  const R abbx = fabs(bxs);
  const R signbx = copysign1(bxs);
  R sign1 = copysign1(abbx - eps);
  maxs1 = fmax(R(0), sign1);
  mins1 = -fmin(R(0), sign1);
  const R invsumro = dv(maxs1, sqrtrol + sqrtror);

  temp = invsumro * (sqrtrol * vylst + sqrtror * vyrst + signbx * (byrst - bylst));
  const R vyldst = vylst * mins1 + temp;
  const R ryldst = rylst * mins1 + rolst * temp;
  const R ryrdst = ryrst * mins1 + rorst * temp;

  temp = invsumro * (sqrtrol * vzlst + sqrtror * vzrst + signbx * (bzrst - bzlst));
  const R vzldst = vzlst * mins1 + temp;
  const R rzldst = rzlst * mins1 + rolst * temp;
  const R rzrdst = rzrst * mins1 + rorst * temp;

  temp = invsumro * (sqrtrol * byrst + sqrtror * bylst + signbx * sqrtrol * sqrtror * (vyrst - vylst));
  const R byldst = bylst * mins1 + temp;
  const R byrdst = byrst * mins1 + temp;

  temp = invsumro * (sqrtrol * bzrst + sqrtror * bzlst + signbx * sqrtrol * sqrtror * (vzrst - vzlst));
  const R bzldst = bzlst * mins1 + temp;
  const R bzrdst = bzrst * mins1 + temp;

  temp = sm * bxs + vyldst * byldst + vzldst * bzldst;
  const R eeldst = eelst - sqrtrol * signbx * (vdbstl - temp) * maxs1;
  const R eerdst = eerst + sqrtror * signbx * (vdbstr - temp) * maxs1;

  // flux selection by sign(S_M) (KH/mhd.cpp:381-415)
  const R slst = (sm - abbx * rsqrtrol) * maxs1;
  const R srst = (sm + abbx * rsqrtror) * maxs1;
  sign1 = copysign1(sm);
  maxs1 = fmax(R(0), sign1);
  mins1 = -fmin(R(0), sign1);
  const R msl = fmin(sl, R(0));
  const R msr = fmax(sr, R(0));
  const R mslst = fmin(slst, R(0));
  const R msrst = fmax(srst, R(0));

  nf.d = (L.fd + (rolst - rol) * msl) * maxs1 + (Rt.fd + (rorst - ror) * msr) * mins1;
  nf.E = (L.fE + (eelst - eel) * msl + (eeldst - eelst) * mslst) * maxs1 +
         (Rt.fE + (eerst - eer) * msr + (eerdst - eerst) * msrst) * mins1;
  nf.u = (L.fu + (rxlst - rxl) * msl) * maxs1 + (Rt.fu + (rxrst - rxr) * msr) * mins1;
  nf.v = (L.fv + (rylst - ryl) * msl + (ryldst - rylst) * mslst) * maxs1 +
         (Rt.fv + (ryrst - ryr) * msr + (ryrdst - ryrst) * msrst) * mins1;
  nf.w = (L.fw + (rzlst - rzl) * msl + (rzldst - rzlst) * mslst) * maxs1 +
         (Rt.fw + (rzrst - rzr) * msr + (rzrdst - rzrst) * msrst) * mins1;
  nf.bv = (L.fbv + (bylst - byl) * msl + (byldst - bylst) * mslst) * maxs1 +
          (Rt.fbv + (byrst - byr) * msr + (byrdst - byrst) * msrst) * mins1;
  nf.bw = (L.fbw + (bzlst - bzl) * msl + (bzldst - bzlst) * mslst) * maxs1 +
          (Rt.fbw + (bzrst - bzr) * msr + (bzrdst - bzrst) * msrst) * mins1;
  nf.sc = (L.fsc + (sclst - scl) * msl) * maxs1 + (Rt.fsc + (scrst - scr) * msr) * mins1;
}

// Upwind-side HLLD (MHD_HLLD_NEAR=1, default).  The reference's branch-free flux
// (KH/mhd.cpp:381-415) is a left-side term times max(0, sign S_M) plus a right-side term
// times -min(0, sign S_M): for every face one side's term survives, the left one for
// S_M >= +0 and the right one otherwise, and the other is multiplied by an exact zero.
// This form picks that side once S_M is known and forms only its star state, double-star
// state, physical flux and flux terms, with the reference's expressions.  The other side
// contributes what the double-star average reads from both sides: its star density,
// transverse velocity and transverse field.  The result is the surviving term of the
// reference's sum; the right star energy is grouped as the left one is (KH/mhd.cpp:308-310
// against :329-331), a difference of round-off.
#ifndef MHD_HLLD_NEAR
#define MHD_HLLD_NEAR 1
#endif

// One interface state in the rotated frame, as HLLD reads it before the side is known.  The
// total energy and the energy flux's magnetic factor are formed after the pick, for the
// surviving side only (hlld_near), from e and |B|^2 kept here.
template <typename R>
struct RPrim {
  R rho, vn, vv, vw, bn, bv, bw, e, bsq, pt, X, cf, psi;
};

template <int D, int EFLUX, typename R>
__device__ __forceinline__ void prim_to_rprim(const R* q, R rhoinv, RPrim<R>& s) {
  constexpr int DF = (D + 1) % 3, DB = (D + 2) % 3;
  const R bsq = q[nbm1] * q[nbm1] + q[nbm2] * q[nbm2] + q[nbm3] * q[nbm3];
  const R bn = q[nbm1 + D];
  s.rho = q[nden];
  s.vn = q[nve1 + D];
  s.vv = q[nve1 + DF];
  s.vw = q[nve1 + DB];
  s.bn = bn;
  s.bv = q[nbm1 + DF];
  s.bw = q[nbm1 + DB];
  s.e = q[nene];
  s.bsq = bsq;
  s.pt = q[npre] + bsq / R(2);
  s.X = q[nst];
  if constexpr (sizeof(R) == 8 && MHD_FAST_SQRT && MHD_SQRT_FLOOR) {
    // c_s^2 + 1e-300 in one rounding (c_s^2 itself for c_s^2 >= 2e-284): the outer root's
    // argument is then at least css / 2 > 0 and needs no floor of its own
    const R css = fma(q[ncsp], q[ncsp], R(1.0e-300));
    const R cts = css + bsq * rhoinv;
    s.cf = sq_pos((cts + sq(cts * cts - R(4) * css * bn * bn * rhoinv)) / R(2));
  } else {
    const R css = q[ncsp] * q[ncsp];
    const R cts = css + bsq * rhoinv;
    s.cf = sq((cts + sq(cts * cts - R(4) * css * bn * bn * rhoinv)) / R(2));
  }
  s.psi = q[nbps];
}

template <typename R>
__device__ __forceinline__ R pick(bool right, R l, R r) {
  return right ? r : l;
}
// pick of two values whose low 32 bits are zero (0, +-1 in FP64): one 32-bit select
template <typename R>
__device__ __forceinline__ R pick_hi(bool right, R l, R r) {
  if constexpr (sizeof(R) == 8) {
    const int hi = right ? __double2hiint(r) : __double2hiint(l);
    return __hiloint2double(hi, 0);
  } else {
    return right ? r : l;
  }
}

// D: direction of the face (the rotated slots map back to Cartesian components for the
// energy's sums, in the order Prim2ConsD forms them); EFLUX as prim_to_state
template <int D, int EFLUX, typename R>
__device__ __forceinline__ void hlld_near(const RPrim<R>& L, const RPrim<R>& Rt, RFlux<R>& nf) {
  const R eps = R(1.0e-30);
  const R tiny = R(1.0e-20);
  const R rol = L.rho, vxl = L.vn, ptl = L.pt;
  const R ror = Rt.rho, vxr = Rt.vn, ptr = Rt.pt;

  // wave speeds, contact speed and star pressure (KH/mhd.cpp:249-266)
  const R cfmax = fmax(L.cf, Rt.cf);
  const R sl = fmin(vxl, vxr) - cfmax;
  const R sr = fmax(vxl, vxr) + cfmax;
  const R sdl = fmin(-tiny, sl - vxl);
  const R sdr = fmax(tiny, sr - vxr);
  const R rosdl = rol * sdl;
  const R rosdr = ror * sdr;
  const R temp = dv(R(1), rosdr - rosdl);
  const R sm = (rosdr * vxr - rosdl * vxl - ptr + ptl) * temp;
  const R sdml = fmin(-tiny, sl - sm);
  const R sdmr = fmax(tiny, sr - sm);
  const R ptst = (rosdr * ptl - rosdl * ptr + rosdl * rosdr * (vxr - vxl)) * temp;

  // star-state denominators of both sides and their reciprocals (KH/mhd.cpp:268-276, 304-312)
  const R bxs = R(0.5) * (L.bn + Rt.bn);
  const R bxsq = bxs * bxs;
  const R tfl = rosdl * sdml - bxsq;
  const R sgl = copysign1(fabs(tfl) - eps);
  const R tfr = rosdr * sdmr - bxsq;
  const R sgr = copysign1(fabs(tfr) - eps);
  R itfl, isdml, itfr, isdmr;
  recip4(tfl + fmin(R(0), sgl), sdml, tfr + fmin(R(0), sgr), sdmr, itfl, isdml, itfr, isdmr);

  // of both star states, what the double-star average reads (KH/mhd.cpp:268-333)
  struct Star {
    R maxs, mins, ro, vy, vz, by, bz;
  };
  auto star = [&](const RPrim<R>& P, R sd, R sdm, R rosd, R itf, R isdm, R sg) {
    Star s;
    s.maxs = fmax(R(0), sg);
    s.mins = fmin(R(0), sg);
    R t = bxs * (sd - sdm) * itf;
    s.ro = s.maxs * (rosd * isdm) - s.mins * P.rho;
    s.vy = s.maxs * (P.vv - P.bv * t) - s.mins * P.vv;
    s.vz = s.maxs * (P.vw - P.bw * t) - s.mins * P.vw;
    t = (rosd * sd - bxsq) * itf;
    s.by = s.maxs * (P.bv * t) - s.mins * P.bv;
    s.bz = s.maxs * (P.bw * t) - s.mins * P.bw;
    return s;
  };
  const Star stl = star(L, sdl, sdml, rosdl, itfl, isdml, sgl);
  const Star str = star(Rt, sdr, sdmr, rosdr, itfr, isdmr, sgr);

  // the double-star average (KH/mhd.cpp:337-360)
  R sqrtrol, sqrtror, rsqrtrol, rsqrtror;
  sq_rsq_pos(stl.ro, sqrtrol, rsqrtrol);
  sq_rsq_pos(str.ro, sqrtror, rsqrtror);
  const R abbx = fabs(bxs);
  const R signbx = copysign1(bxs);
  const R sign1 = copysign1(abbx - eps);
  const R maxs1 = fmax(R(0), sign1);
  const R mins1 = -fmin(R(0), sign1);
  const R invsumro = dv(maxs1, sqrtrol + sqrtror);
  const R tvy = invsumro * (sqrtrol * stl.vy + sqrtror * str.vy + signbx * (str.by - stl.by));
  const R tvz = invsumro * (sqrtrol * stl.vz + sqrtror * str.vz + signbx * (str.bz - stl.bz));
  const R tby = invsumro * (sqrtrol * str.by + sqrtror * stl.by + signbx * sqrtrol * sqrtror * (str.vy - stl.vy));
  const R tbz = invsumro * (sqrtrol * str.bz + sqrtror * stl.bz + signbx * sqrtrol * sqrtror * (str.vz - stl.vz));

  // the side whose term survives: right where sign(S_M) = -1 (KH/mhd.cpp:381-383)
  const bool right = copysign1(sm) < R(0);
  const R ro = pick(right, rol, ror), vx = pick(right, vxl, vxr);
  const R vy = pick(right, L.vv, Rt.vv), vz = pick(right, L.vw, Rt.vw);
  const R bn = pick(right, L.bn, Rt.bn), by = pick(right, L.bv, Rt.bv), bz = pick(right, L.bw, Rt.bw);
  const R e = pick(right, L.e, Rt.e), bsq = pick(right, L.bsq, Rt.bsq), pt = pick(right, ptl, ptr);
  const R X = pick(right, L.X, Rt.X);
  // its total energy and energy-flux factor (KH/mhd.cpp:1036-1050) from the Cartesian
  // components: slot s of the rotated frame holds component (s + D) % 3
  R vc[3], bc[3];
  vc[D] = vx;
  vc[(D + 1) % 3] = vy;
  vc[(D + 2) % 3] = vz;
  bc[D] = bn;
  bc[(D + 1) % 3] = by;
  bc[(D + 2) % 3] = bz;
  const R vsq = vc[0] * vc[0] + vc[1] * vc[1] + vc[2] * vc[2];
  const R ee = ro * e + R(0.5) * ro * vsq + R(0.5) * bsq;
  R eb;
  if constexpr (EFLUX == 0) eb = bsq;
  else eb = vc[0] * bc[0] + vc[1] * bc[1] + vc[2] * bc[2];
  const R sd = pick(right, sdl, sdr), isdm = pick(right, isdml, isdmr);
  // max(0, sign) and min(0, sign) of the side: 0 or +-1
  const R maxs = pick_hi(right, stl.maxs, str.maxs), mins = pick_hi(right, stl.mins, str.mins);
  const R rost = pick(right, stl.ro, str.ro), vyst = pick(right, stl.vy, str.vy), vzst = pick(right, stl.vz, str.vz);
  const R byst = pick(right, stl.by, str.by), bzst = pick(right, stl.bz, str.bz);
  const R sqro = pick(right, sqrtrol, sqrtror), rsqro = pick(right, rsqrtrol, rsqrtror);
  // the side's sign in the double-star energy and the Alfven speed: -1 left, +1 right
  const R sgn = right ? R(1) : R(-1);

  // its star state (KH/mhd.cpp:268-302 left, 304-333 right)
  const R vxst = maxs * sm - mins * vx;
  const R rxst = rost * vxst;
  const R ryst = rost * vyst;
  const R rzst = rost * vzst;
  const R vdbst = vxst * bxs + vyst * byst + vzst * bzst;
  const R eest = maxs * (sd * ee - pt * vx + ptst * sm + bxs * (vx * bxs + vy * by + vz * bz - vdbst)) * isdm - mins * ee;

  // its double-star state (KH/mhd.cpp:344-364)
  const R vydst = vyst * mins1 + tvy;
  const R rydst = ryst * mins1 + rost * tvy;
  const R vzdst = vzst * mins1 + tvz;
  const R rzdst = rzst * mins1 + rost * tvz;
  const R bydst = byst * mins1 + tby;
  const R bzdst = bzst * mins1 + tbz;
  const R tdot = sm * bxs + vydst * bydst + vzdst * bzdst;
  const R eedst = eest + sqro * (sgn * signbx) * (vdbst - tdot) * maxs1;

  // its outer and Alfven wave speeds, clipped to the side (KH/mhd.cpp:367-375): the Alfven
  // speed S*_side = (S_M + sgn |B_x| / sqrt(rho*_side)) maxs1 is clipped as sgn max(sgn S*, 0),
  // max(S*, 0) on the right and min(S*, 0) on the left (a multiplication by sgn is exact)
  const R ms = right ? fmax(sr, R(0)) : fmin(sl, R(0));
  const R mst = sgn * fmax((sgn * sm + abbx * rsqro) * maxs1, R(0));

  // its physical flux (KH/mhd.cpp:1036-1076) and the flux of the side (KH/mhd.cpp:384-415)
  const R rvx = ro * vx;
  const R rvy = ro * vy;
  const R rvz = ro * vz;
  nf.d = rvx + (rost - ro) * ms;
  nf.u = (ro * vx * vx - bn * bn + pt) + (rxst - rvx) * ms;
  nf.v = (ro * vy * vx - by * bn) + (ryst - rvy) * ms + (rydst - ryst) * mst;
  nf.w = (ro * vz * vx - bz * bn) + (rzst - rvz) * ms + (rzdst - rzst) * mst;
  nf.E = ((ee + pt) * vx - eb * bn) + (eest - ee) * ms + (eedst - eest) * mst;
  nf.bv = (by * vx - vy * bn) + (byst - by) * ms + (bydst - byst) * mst;
  nf.bw = (bz * vx - vz * bn) + (bzst - bz) * ms + (bzdst - bzst) * mst;
  // the scalar's flux ro X vx + (rost X - ro X) ms is X times the density flux
  nf.sc = X * nf.d;
}

// Interface flux in Cartesian conserved order (F[0..9]) from reconstructed left/right
// primitive states: HLLD for the MHD part plus the GLM (B_n, psi) flux
// F(B_n) = (psiL+psiR)/2 - c_h (B_nR-B_nL)/2,  F(psi) = c_h^2 [(B_nL+B_nR)/2 - (psiR-psiL)/(2 c_h)]
// (scatter and GLM as at KH/mhd.cpp:1225-1240, 1306-1321, 1391-1406).
// chg: the GLM speed c_h; hchg = 0.5/c_h (hoisted by the caller)
template <int D, int EFLUX, typename R>
__device__ __forceinline__ void riemann_flux(const R* ql, const R* qr, double chg, double hchg, double* F) {
  R rinvl, rinvr;
  recip2(ql[nden], qr[nden], rinvl, rinvr);
  RFlux<R> nf;
#if MHD_HLLD_NEAR
  RPrim<R> sl, sr;
  prim_to_rprim<D, EFLUX, R>(ql, rinvl, sl);
  prim_to_rprim<D, EFLUX, R>(qr, rinvr, sr);
  hlld_near<D, EFLUX, R>(sl, sr, nf);
  const double bnl = double(sl.bn), bnr = double(sr.bn);
#else
  RState<R> sl, sr;
  prim_to_state<D, EFLUX, R>(ql, rinvl, sl);
  prim_to_state<D, EFLUX, R>(qr, rinvr, sr);
  hlld<R>(sl, sr, nf);
  const double bnl = double(sl.bu), bnr = double(sr.bu);
#endif
  F[mden] = double(nf.d);
  F[meto] = double(nf.E);
  F[mrv1 + D] = double(nf.u);
  F[mrv1 + (D + 1) % 3] = double(nf.v);
  F[mrv1 + (D + 2) % 3] = double(nf.w);
  F[mbm1 + (D + 1) % 3] = double(nf.bv);
  F[mbm1 + (D + 2) % 3] = double(nf.bw);
  F[mst] = double(nf.sc);
  // GLM part in FP64 from the reconstructed (B_n, psi)
  const double psl = double(sl.psi), psr = double(sr.psi);
  F[mbm1 + D] = 0.5 * (psl + psr) - 0.5 * chg * (bnr - bnl);
  F[mbps] = (0.5 * (bnl + bnr) - hchg * (psr - psl)) * chg * chg;
}

// Left state at face f+1/2 and right state at face f-1/2 of cell f, from cells f-1, f, f+1:
//   PL(f+1/2) = P[f] + vL(P[f+1]-P[f], P[f]-P[f-1])/2,  PR(f-1/2) = P[f] - (same slope)/2
// (KH/mhd.cpp:1134-1154: the same limited slope of cell f serves both faces).
template <typename R>
__device__ __forceinline__ void cell_states(const R* qm, const R* q0, const R* qp, R* to_right, R* to_left) {
#pragma unroll
  for (int n = 0; n < NPRIM; ++n) {
    if constexpr (LIM_FACTORS && sizeof(R) == 8) {
      double pm, r;
      half_van_leer_factors(qp[n] - q0[n], q0[n] - qm[n], pm, r);
      to_right[n] = fma(pm, r, q0[n]);
      to_left[n] = fma(-pm, r, q0[n]);
    } else {
      const R h = half_van_leer(qp[n] - q0[n], q0[n] - qm[n]);
      to_right[n] = q0[n] + h;
      to_left[n] = q0[n] - h;
    }
  }
}

// UpdatePrimitvP on one cell, KH/mhd.cpp:1480-1507
__device__ __forceinline__ void cons_to_prim(const double* u, double gam, double* q) {
  const double rho = u[mden];
  const double rhoinv = dv(1.0, rho);
  q[nden] = rho;
  q[nve1] = u[mrv1] * rhoinv;
  q[nve2] = u[mrv2] * rhoinv;
  q[nve3] = u[mrv3] * rhoinv;
  const double ekin = 0.5 * (u[mrv1] * u[mrv1] + u[mrv2] * u[mrv2] + u[mrv3] * u[mrv3]) * rhoinv;
  const double emag = 0.5 * (u[mbm1] * u[mbm1] + u[mbm2] * u[mbm2] + u[mbm3] * u[mbm3]);
  const double e = (u[meto] - ekin - emag) * rhoinv;
  q[nene] = e;
  q[npre] = e * rho * (gam - 1.0);
  q[ncsp] = sq(e * gam * (gam - 1.0));
  q[nbm1] = u[mbm1];
  q[nbm2] = u[mbm2];
  q[nbm3] = u[mbm3];
  q[nbps] = u[mbps];
  q[nst] = u[mst] * rhoinv;
}

// CFL time-step bound of one cell (KH/mhd.cpp:1542-1553) with the cell widths given
__device__ __forceinline__ double cell_dt(const double* q, double dx, double dy, double dz) {
  const double ctot =
      sq(q[ncsp] * q[ncsp] + dv(q[nbm1] * q[nbm1] + q[nbm2] * q[nbm2] + q[nbm3] * q[nbm3], q[nden]));
  const double d1 = dv(dx, fabs(q[nve1]) + ctot);
  const double d2 = dv(dy, fabs(q[nve2]) + ctot);
  const double d3 = dv(dz, fabs(q[nve3]) + ctot);
  return fmin(d1, fmin(d2, d3));
}

// GLM cleaning speed bound of one cell: max_d |v_d| + c_f,d  (KH/mhd.cpp:1589-1601)
__device__ __forceinline__ double cell_ch(const double* q) {
  const double rho = q[nden];
  const double css = q[ncsp] * q[ncsp];
  const double rinv = dv(1.0, rho);
  const double ca1 = q[nbm1] * q[nbm1] * rinv;
  const double ca2 = q[nbm2] * q[nbm2] * rinv;
  const double ca3 = q[nbm3] * q[nbm3] * rinv;
  const double cts = css + ca1 + ca2 + ca3;
  const double cm1 = sq((cts + sq(cts * cts - 4.0 * css * ca1)) / 2.0);
  const double cm2 = sq((cts + sq(cts * cts - 4.0 * css * ca2)) / 2.0);
  const double cm3 = sq((cts + sq(cts * cts - 4.0 * css * ca3)) / 2.0);
  return fmax(fabs(q[nve1]) + cm1, fmax(fabs(q[nve2]) + cm2, fabs(q[nve3]) + cm3));
}

}  // namespace mhd
