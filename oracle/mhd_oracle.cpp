/*
 * mhd_oracle.cpp -- TEST INFRASTRUCTURE ONLY (the checker, never the product).
 *
 * A plain C++/OpenMP restatement of one MHDTurbulence time step as the reference's
 * src_cpp_kokkos_host variant computes it, operation for operation, so that the HIP
 * library can be compared with it on identical inputs.  Each function names the
 * reference lines it restates.  Paths are relative to the reference repository root;
 * KH/ = src_cpp_kokkos_host/.
 *
 * The reference itself cannot be compiled here (its CMake fetches Kokkos from the
 * network; the OpenMP-target variant does not compile unmodified), so this file is
 * pinned by the known-answer, invariant and convergence tests under tests/.
 */
#include "mhd_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr int NGH = 2;
constexpr int NPRIM = 12;   // KH/mhd.hpp:249-253
constexpr int NCONS = 10;   // KH/mhd.hpp:255
constexpr int NRIEM = 23;   // 2*mconsv+madd

// primitive indices (KH/mhd.hpp:251-253)
enum { nden = 0, nve1, nve2, nve3, nene, npre, ncsp, nbm1, nbm2, nbm3, nbps, nst };
// conserved indices (KH/mhd.hpp:264-268)
enum { mden = 0, mrv1, mrv2, mrv3, meto, mbm1, mbm2, mbm3, mbps, mst };
// Riemann-vector slots (KH/mhd.hpp:255-262): rotated conserved (u = normal), rotated
// physical flux, fast speed, normal velocity, total pressure.
enum {
  mudn = 0, muvu, muvv, muvw, muet, mubu, mubv, mubw, mubp, must,
  mfdn, mfvu, mfvv, mfvw, mfet, mfbu, mfbv, mfbw, mfbp, mfst,
  mcsp, mvel, mpre
};
// HLLD output slots (KH/mhd.hpp:264-266): mrvu..mrvw = 1..3, mbmu..mbmw = 5..7

enum { BC_PERIODIC = 1, BC_REFLECT = 2, BC_OUTFLOW = 3 };

}  // namespace

struct orc_ctx {
  orc_config cfg;
  int itot, jtot, ktot, is, ie, js, je, ks, ke;
  std::vector<double> x1a, x1b, x2a, x2b, x3a, x3b;
  std::vector<double> P, U, F[3];
  std::vector<double> U0;  // the state at the start of a two-stage step
  size_t idx(int nv, int n, int k, int j, int i) const {
    return ((static_cast<size_t>(k) * jtot + j) * nv + n) * itot + i;
  }
  double& p(int n, int k, int j, int i) { return P[idx(NPRIM, n, k, j, i)]; }
  double& u(int n, int k, int j, int i) { return U[idx(NCONS, n, k, j, i)]; }
  double& f(int d, int n, int k, int j, int i) { return F[d][idx(NCONS, n, k, j, i)]; }
};

namespace {

// KH/mhd.cpp:89-95
inline double van_leer(double dsvp, double dsvm) {
  if (dsvp * dsvm > 0.0) return 2.0 * dsvp * dsvm / (dsvp + dsvm);
  return 0.0;
}

// slot (0 = u normal, 1 = v, 2 = w) of Cartesian component c for sweep direction d (0-based):
// d=0 -> (x,y,z)=(u,v,w); d=1 -> (y,z,x)=(u,v,w); d=2 -> (z,x,y)=(u,v,w)  (KH/mhd.cpp:1078-1118)
inline int slot(int c, int d) { return (c - d + 3) % 3; }

// Prim2ConsD (KH/mhd.cpp:997-1076) for direction d (0-based).  energy_flux selects the
// energy-flux magnetic term: 0 = |B|^2 B_n as written at KH/mhd.cpp:1049-1050, 1 = (v.B) B_n.
void prim2riemann(const double* q, int d, int eflux, double* c) {
  const int df = (d + 1) % 3, db = (d + 2) % 3;
  const double rho = q[nden];
  const double rhoinv = 1.0 / rho;
  const double vsq = q[nve1] * q[nve1] + q[nve2] * q[nve2] + q[nve3] * q[nve3];
  const double bsq = q[nbm1] * q[nbm1] + q[nbm2] * q[nbm2] + q[nbm3] * q[nbm3];
  const double vn = q[nve1 + d], bn = q[nbm1 + d];
  c[mudn] = rho;
  for (int a = 0; a < 3; ++a) c[muvu + slot(a, d)] = rho * q[nve1 + a];
  c[muet] = rho * q[nene] + 0.5 * rho * vsq + 0.5 * bsq;
  for (int a = 0; a < 3; ++a) c[mubu + slot(a, d)] = q[nbm1 + a];
  c[mubp] = q[nbps];
  c[must] = rho * q[nst];
  const double ptl = q[npre] + bsq / 2.0;
  c[mfdn] = rho * vn;
  for (int a = 0; a < 3; ++a)
    c[mfvu + slot(a, d)] = rho * q[nve1 + a] * vn - q[nbm1 + a] * bn;
  if (eflux == 0) {
    c[mfet] = (c[muet] + ptl) * vn - bsq * bn;
  } else {
    const double vdb = q[nve1] * q[nbm1] + q[nve2] * q[nbm2] + q[nve3] * q[nbm3];
    c[mfet] = (c[muet] + ptl) * vn - vdb * bn;
  }
  c[mfvu] += ptl;
  c[mfbu] = 0.0;
  c[mfbv] = q[nbm1 + df] * vn - q[nve1 + df] * bn;
  c[mfbw] = q[nbm1 + db] * vn - q[nve1 + db] * bn;
  c[mfbp] = 0.0;
  c[mfst] = rho * q[nst] * vn;
  const double css = q[ncsp] * q[ncsp];
  const double cts = css + bsq * rhoinv;
  c[mcsp] = std::sqrt((cts + std::sqrt(cts * cts - 4.0 * css * bn * bn * rhoinv)) / 2.0);
  c[mvel] = vn;
  c[mpre] = ptl;
}

// HLLD, Miyoshi & Kusano (2005), in the branch-free form of KH/mhd.cpp:97-417.
// Writes nflux[0..9] except slot mbps (8), which the GLM flux supplies.
void hlld(const double* L, const double* R, double* nf) {
  const double eps = 1.0e-30;
  const double rol = L[mudn], eel = L[muet], rxl = L[muvu], ryl = L[muvv], rzl = L[muvw];
  const double vxl = L[muvu] / L[mudn], vyl = L[muvv] / L[mudn], vzl = L[muvw] / L[mudn];
  const double byl = L[mubv], bzl = L[mubw], ptl = L[mpre], scl = L[must];
  const double ror = R[mudn], eer = R[muet], rxr = R[muvu], ryr = R[muvv], rzr = R[muvw];
  const double vxr = R[muvu] / R[mudn], vyr = R[muvv] / R[mudn], vzr = R[muvw] / R[mudn];
  const double byr = R[mubv], bzr = R[mubw], ptr = R[mpre], scr = R[must];

  // step 1: outer wave speeds (KH/mhd.cpp:213-216)
  const double cfl = L[mcsp], cfr = R[mcsp];
  const double sl = std::min(vxl, vxr) - std::max(cfl, cfr);
  const double sr = std::max(vxl, vxr) + std::max(cfl, cfr);

  // step 2: physical fluxes (KH/mhd.cpp:220-241)
  const double frol = L[mfdn], feel = L[mfet], frxl = L[mfvu], fryl = L[mfvv], frzl = L[mfvw];
  const double fbyl = L[mfbv], fbzl = L[mfbw], fscl = L[mfst];
  const double fror = R[mfdn], feer = R[mfet], frxr = R[mfvu], fryr = R[mfvv], frzr = R[mfvw];
  const double fbyr = R[mfbv], fbzr = R[mfbw], fscr = R[mfst];

  // step 4: contact speed, Eq. (38) of Miyoshi & Kusano (KH/mhd.cpp:245-258)
  const double sdl = std::min(-1e-20, sl - vxl);
  const double sdr = std::max(1e-20, sr - vxr);
  const double rosdl = rol * sdl;
  const double rosdr = ror * sdr;
  double temp = 1.0 / (rosdr - rosdl);
  const double sm = (rosdr * vxr - rosdl * vxl - ptr + ptl) * temp;
  const double sdml = std::min(-1e-20, sl - sm);
  const double sdmr = std::max(1e-20, sr - sm);
  // total pressure in the star region (KH/mhd.cpp:263)
  const double ptst = (rosdr * ptl - rosdl * ptr + rosdl * rosdr * (vxr - vxl)) * temp;

  // step 5A: left star state (KH/mhd.cpp:268-302)
  const double bxs = 0.5 * (L[mubu] + R[mubu]);
  const double bxsq = bxs * bxs;
  double temp_fst = rosdl * sdml - bxsq;
  double sign1 = std::copysign(1.0, std::abs(temp_fst) - eps);
  double maxs1 = std::max(0.0, sign1);
  double mins1 = std::min(0.0, sign1);
  double itf = 1.0 / (temp_fst + mins1);
  const double isdml = 1.0 / sdml;
  temp = bxs * (sdl - sdml) * itf;
  const double rolst = maxs1 * (rosdl * isdml) - mins1 * rol;
  const double sclst = rolst / rol * scl;
  const double vxlst = maxs1 * sm - mins1 * vxl;
  const double rxlst = rolst * vxlst;
  const double vylst = maxs1 * (vyl - byl * temp) - mins1 * vyl;
  const double rylst = rolst * vylst;
  const double vzlst = maxs1 * (vzl - bzl * temp) - mins1 * vzl;
  const double rzlst = rolst * vzlst;
  temp = (rosdl * sdl - bxsq) * itf;
  const double bylst = maxs1 * (byl * temp) - mins1 * byl;
  const double bzlst = maxs1 * (bzl * temp) - mins1 * bzl;
  const double vdbstl = vxlst * bxs + vylst * bylst + vzlst * bzlst;
  const double eelst =
      maxs1 * (sdl * eel - ptl * vxl + ptst * sm + bxs * (vxl * bxs + vyl * byl + vzl * bzl - vdbstl)) *
          isdml -
      mins1 * eel;

  // step 5B: right star state (KH/mhd.cpp:304-333)
  temp_fst = rosdr * sdmr - bxsq;
  sign1 = std::copysign(1.0, std::abs(temp_fst) - eps);
  maxs1 = std::max(0.0, sign1);
  mins1 = std::min(0.0, sign1);
  itf = 1.0 / (temp_fst + mins1);
  const double isdmr = 1.0 / sdmr;
  temp = bxs * (sdr - sdmr) * itf;
  const double rorst = maxs1 * (rosdr * isdmr) - mins1 * ror;
  const double scrst = rorst / ror * scr;
  const double vxrst = maxs1 * sm - mins1 * vxr;
  const double rxrst = rorst * vxrst;
  const double vyrst = maxs1 * (vyr - byr * temp) - mins1 * vyr;
  const double ryrst = rorst * vyrst;
  const double vzrst = maxs1 * (vzr - bzr * temp) - mins1 * vzr;
  const double rzrst = rorst * vzrst;
  temp = (rosdr * sdr - bxsq) * itf;
  const double byrst = maxs1 * (byr * temp) - mins1 * byr;
  const double bzrst = maxs1 * (bzr * temp) - mins1 * bzr;
  const double vdbstr = vxrst * bxs + vyrst * byrst + vzrst * bzrst;
  const double eerst = maxs1 * ((sdr * eer - ptr * vxr + ptst * sm) * isdmr +
                                bxs * (vxr * bxs + vyr * byr + vzr * bzr - vdbstr) * isdmr) -
                       mins1 * eer;

  // step 5C: double-star states (KH/mhd.cpp:337-377)
  const double sqrtrol = std::sqrt(rolst);
  const double sqrtror = std::sqrt(rorst);
  const double abbx = std::abs(bxs);
  const double signbx = std::copysign(1.0, bxs);
  sign1 = std::copysign(1.0, abbx - eps);
  maxs1 = std::max(0e0, sign1);
  mins1 = -std::min(0e0, sign1);
  const double invsumro = maxs1 / (sqrtrol + sqrtror);

  temp = invsumro * (sqrtrol * vylst + sqrtror * vyrst + signbx * (byrst - bylst));
  const double vyldst = vylst * mins1 + temp;
  const double ryldst = rylst * mins1 + rolst * temp;
  const double ryrdst = ryrst * mins1 + rorst * temp;

  temp = invsumro * (sqrtrol * vzlst + sqrtror * vzrst + signbx * (bzrst - bzlst));
  const double vzldst = vzlst * mins1 + temp;
  const double rzldst = rzlst * mins1 + rolst * temp;
  const double rzrdst = rzrst * mins1 + rorst * temp;

  temp = invsumro * (sqrtrol * byrst + sqrtror * bylst + signbx * sqrtrol * sqrtror * (vyrst - vylst));
  const double byldst = bylst * mins1 + temp;
  const double byrdst = byrst * mins1 + temp;

  temp = invsumro * (sqrtrol * bzrst + sqrtror * bzlst + signbx * sqrtrol * sqrtror * (vzrst - vzlst));
  const double bzldst = bzlst * mins1 + temp;
  const double bzrdst = bzrst * mins1 + temp;

  temp = sm * bxs + vyldst * byldst + vzldst * bzldst;
  const double eeldst = eelst - sqrtrol * signbx * (vdbstl - temp) * maxs1;
  const double eerdst = eerst + sqrtror * signbx * (vdbstr - temp) * maxs1;

  // step 6: flux selection by sign(S_M) (KH/mhd.cpp:381-415)
  const double slst = (sm - abbx / sqrtrol) * maxs1;
  const double srst = (sm + abbx / sqrtror) * maxs1;
  sign1 = std::copysign(1.0, sm);
  maxs1 = std::max(0.0, sign1);
  mins1 = -std::min(0.0, sign1);
  const double msl = std::min(sl, 0.0);
  const double msr = std::max(sr, 0.0);
  const double mslst = std::min(slst, 0.0);
  const double msrst = std::max(srst, 0.0);

  nf[mden] = (frol + (rolst - rol) * msl) * maxs1 + (fror + (rorst - ror) * msr) * mins1;
  nf[meto] = (feel + (eelst - eel) * msl + (eeldst - eelst) * mslst) * maxs1 +
             (feer + (eerst - eer) * msr + (eerdst - eerst) * msrst) * mins1;
  nf[1] = (frxl + (rxlst - rxl) * msl) * maxs1 + (frxr + (rxrst - rxr) * msr) * mins1;
  nf[2] = (fryl + (rylst - ryl) * msl + (ryldst - rylst) * mslst) * maxs1 +
          (fryr + (ryrst - ryr) * msr + (ryrdst - ryrst) * msrst) * mins1;
  nf[3] = (frzl + (rzlst - rzl) * msl + (rzldst - rzlst) * mslst) * maxs1 +
          (frzr + (rzrst - rzr) * msr + (rzrdst - rzrst) * msrst) * mins1;
  nf[5] = 0.0;
  nf[6] = (fbyl + (bylst - byl) * msl + (byldst - bylst) * mslst) * maxs1 +
          (fbyr + (byrst - byr) * msr + (byrdst - byrst) * msrst) * mins1;
  nf[7] = (fbzl + (bzlst - bzl) * msl + (bzldst - bzlst) * mslst) * maxs1 +
          (fbzr + (bzrst - bzr) * msr + (bzrdst - bzrst) * msrst) * mins1;
  nf[8] = 0.0;
  nf[mst] = (fscl + (sclst - scl) * msl) * maxs1 + (fscr + (scrst - scr) * msr) * mins1;
}

// CalcFlux + scatter + GLM flux of one face (KH/mhd.cpp:1120-1158, 1217-1240).
// pm2..pp1 are the primitive vectors of the cells f-2, f-1, f, f+1 along direction d.
void face_flux(const double* pm2, const double* pm1, const double* p0, const double* pp1, int d,
               int eflux, double chg, double* F) {
  double ql[NPRIM], qr[NPRIM], cl[NRIEM], cr[NRIEM], nf[NCONS];
  for (int n = 0; n < NPRIM; ++n) ql[n] = pm1[n] + 0.5 * van_leer(p0[n] - pm1[n], pm1[n] - pm2[n]);
  prim2riemann(ql, d, eflux, cl);
  for (int n = 0; n < NPRIM; ++n) qr[n] = p0[n] - 0.5 * van_leer(pp1[n] - p0[n], p0[n] - pm1[n]);
  prim2riemann(qr, d, eflux, cr);
  hlld(cl, cr, nf);
  F[mden] = nf[mden];
  F[meto] = nf[meto];
  for (int a = 0; a < 3; ++a) F[mrv1 + a] = nf[1 + slot(a, d)];
  for (int a = 0; a < 3; ++a)
    if (a != d) F[mbm1 + a] = nf[5 + slot(a, d)];
  F[mbm1 + d] = 0.5 * (cl[mubp] + cr[mubp]) - 0.5 * chg * (cr[mubu] - cl[mubu]);
  F[mbps] = (0.5 * (cl[mubu] + cr[mubu]) - 0.5 / chg * (cr[mubp] - cl[mubp])) * chg * chg;
  F[mst] = nf[mst];
}

void load_prim(orc_ctx* c, int k, int j, int i, double* q) {
  for (int n = 0; n < NPRIM; ++n) q[n] = c->p(n, k, j, i);
}

}  // namespace

extern "C" {

orc_ctx* orc_create(const orc_config* cfg) {
  orc_ctx* c = new orc_ctx();
  c->cfg = *cfg;
  // KH/config.hpp:56-65 with ntiles = {1,1,1}
  c->itot = cfg->nx + 2 * NGH + 1;
  c->jtot = cfg->ny + 2 * NGH + 1;
  c->ktot = cfg->nz + 2 * NGH + 1;
  c->is = NGH; c->ie = NGH + cfg->nx - 1;
  c->js = NGH; c->je = NGH + cfg->ny - 1;
  c->ks = NGH; c->ke = NGH + cfg->nz - 1;
  c->x1a.assign(c->itot, 0.0); c->x1b.assign(c->itot, 0.0);
  c->x2a.assign(c->jtot, 0.0); c->x2b.assign(c->jtot, 0.0);
  c->x3a.assign(c->ktot, 0.0); c->x3b.assign(c->ktot, 0.0);
  const size_t cells = static_cast<size_t>(c->itot) * c->jtot * c->ktot;
  c->P.assign(cells * NPRIM, 0.0);
  c->U.assign(cells * NCONS, 0.0);
  for (auto& f : c->F) f.assign(cells * NCONS, 0.0);
  return c;
}

void orc_destroy(orc_ctx* c) { delete c; }

void orc_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

void orc_sizes(const orc_ctx* c, int* itot, int* jtot, int* ktot) {
  *itot = c->itot; *jtot = c->jtot; *ktot = c->ktot;
}

double* orc_prim(orc_ctx* c) { return c->P.data(); }
double* orc_cons(orc_ctx* c) { return c->U.data(); }
double* orc_flux(orc_ctx* c, int dir) { return c->F[dir - 1].data(); }
double* orc_grid(orc_ctx* c, int axis, int centre) {
  switch (axis) {
    case 1: return centre ? c->x1b.data() : c->x1a.data();
    case 2: return centre ? c->x2b.data() : c->x2a.data();
    default: return centre ? c->x3b.data() : c->x3a.data();
  }
}

// GenerateGrid, KH/main.cpp:26-65 (one tile per axis).  grid_shift=1 keeps the C++
// variants' face formula dx*(i-(ngh+1))+xmin; 0 gives dx*(i-ngh)+xmin (Fortran, FH/main.f90:76-78).
void orc_generate_grid(orc_ctx* c) {
  const int off = NGH + (c->cfg.grid_shift ? 1 : 0);
  std::vector<double>* fa[3] = {&c->x1a, &c->x2a, &c->x3a};
  std::vector<double>* fb[3] = {&c->x1b, &c->x2b, &c->x3b};
  const int ns[3] = {c->cfg.nx, c->cfg.ny, c->cfg.nz};
  const int tot[3] = {c->itot, c->jtot, c->ktot};
  for (int a = 0; a < 3; ++a) {
    const double dx = (c->cfg.xmax[a] - c->cfg.xmin[a]) / double(ns[a]);
    for (int i = 0; i < tot[a]; ++i) (*fa[a])[i] = dx * (i - off) + c->cfg.xmin[a];
    for (int i = 0; i < tot[a] - 1; ++i) (*fb[a])[i] = 0.5 * ((*fa[a])[i + 1] + (*fa[a])[i]);
  }
}

// conserved from primitive everywhere, as at the end of KH/main.cpp:160-183
void orc_prim_to_cons_all(orc_ctx* c) {
  for (int k = 0; k < c->ktot - 1; ++k)
    for (int j = 0; j < c->jtot - 1; ++j)
      for (int i = 0; i < c->itot - 1; ++i) {
        const double rho = c->p(nden, k, j, i);
        const double v1 = c->p(nve1, k, j, i), v2 = c->p(nve2, k, j, i), v3 = c->p(nve3, k, j, i);
        const double b1 = c->p(nbm1, k, j, i), b2 = c->p(nbm2, k, j, i), b3 = c->p(nbm3, k, j, i);
        c->u(mden, k, j, i) = rho;
        c->u(mrv1, k, j, i) = rho * v1;
        c->u(mrv2, k, j, i) = rho * v2;
        c->u(mrv3, k, j, i) = rho * v3;
        const double ekin = 0.5 * rho * (v1 * v1 + v2 * v2 + v3 * v3);
        const double emag = 0.5 * (b1 * b1 + b2 * b2 + b3 * b3);
        c->u(meto, k, j, i) = rho * c->p(nene, k, j, i) + ekin + emag;
        c->u(mbm1, k, j, i) = b1;
        c->u(mbm2, k, j, i) = b2;
        c->u(mbm3, k, j, i) = b3;
        c->u(mbps, k, j, i) = c->p(nbps, k, j, i);
        c->u(mst, k, j, i) = rho * c->p(nst, k, j, i);
      }
}

// The repository's Kelvin-Helmholtz initial state, KH/main.cpp:66-184 (the per-(j,k)
// random kick there is commented out; the cos*cos harmonic kick on vx, of strength rrv,
// is applied).  vy_amp is 1e-3 in KH/main.cpp:105 and 1e-2 in src_cpp_omp_device/main.cpp:151.
void orc_generate_problem_kh(orc_ctx* c, double vy_amp, double rrv) {
  const double pi = std::acos(-1.0);
  const double dv = 2.0, wid = 0.05, sig = 0.2;
  const double gam = c->cfg.gam;
  for (int k = 0; k < c->ktot - 1; ++k)
    for (int j = 0; j < c->jtot - 1; ++j)
      for (int i = 0; i < c->itot - 1; ++i) {
        const double x = c->x1b[i];
        const double y = c->x2b[j];
        const double tp = std::tanh((y + 0.5) / wid);
        const double tm = std::tanh((y - 0.5) / wid);
        c->p(nden, k, j, i) = 1.0;
        c->p(npre, k, j, i) = 1.0;
        c->p(nve1, k, j, i) = 0.5 * dv * (tp - tm - 1.0);
        c->p(nve2, k, j, i) = vy_amp * std::sin(2.0 * pi * x) *
                              (std::exp(-(y + 0.5) * (y + 0.5) / (sig * sig)) +
                               std::exp(-(y - 0.5) * (y - 0.5) / (sig * sig)));
        c->p(nve3, k, j, i) = 0.0;
        c->p(nbm1, k, j, i) = 0.0;
        c->p(nbm2, k, j, i) = 0.0;
        c->p(nbm3, k, j, i) = 0.0;
        c->p(nbps, k, j, i) = 0.0;
        c->p(nene, k, j, i) = c->p(npre, k, j, i) / (gam - 1.0);
        c->p(ncsp, k, j, i) = std::sqrt(gam * c->p(npre, k, j, i) / c->p(nden, k, j, i));
        c->p(nst, k, j, i) = 0.5 * (tp - tm);
      }
  const double ymin = c->cfg.xmin[1], ymax = c->cfg.xmax[1];
  const double zmin = c->cfg.xmin[2], zmax = c->cfg.xmax[2];
  for (int k = c->ks; k <= c->ke; ++k)
    for (int j = c->js; j <= c->je; ++j) {
      const double y = c->x2b[j], z = c->x3b[k];
      const double kick = dv * rrv * std::cos(6 * 2 * pi * (y - ymin) / (ymax - ymin)) *
                          std::cos(6 * 2 * pi * (z - zmin) / (zmax - zmin));
      for (int i = c->is; i <= c->ie; ++i) c->p(nve1, k, j, i) += kick;
    }
  orc_prim_to_cons_all(c);
}

uint64_t orc_splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ULL;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
  return x ^ (x >> 31);
}

// Single shear layer with a seeded per-cell vy kick (the ensemble/benchmark initial state):
// rho=rho0, p=p0, vx=v0*tanh(y/a), Bx=bx0, vy=amp*exp(-(y/sigma)^2)*(2u-1) with
// u = (splitmix64(seed ^ ((k*ny+j)*nx+i)) >> 11) * 2^-53 on interior (i,j,k) counted from 0,
// X = 0.5*(1+tanh(y/a)).  Ghost cells carry the same profile with no kick.
void orc_generate_problem_shear(orc_ctx* c, const orc_ic_params* ic) {
  const double gam = c->cfg.gam;
  const int nx = c->cfg.nx, ny = c->cfg.ny;
  for (int k = 0; k < c->ktot - 1; ++k)
    for (int j = 0; j < c->jtot - 1; ++j)
      for (int i = 0; i < c->itot - 1; ++i) {
        const double y = c->x2b[j];
        const double th = std::tanh(y / ic->a);
        double vy = 0.0;
        if (i >= c->is && i <= c->ie && j >= c->js && j <= c->je && k >= c->ks && k <= c->ke) {
          const uint64_t id = (static_cast<uint64_t>(k - c->ks) * ny + (j - c->js)) * nx + (i - c->is);
          const double u01 = static_cast<double>(orc_splitmix64(ic->seed ^ id) >> 11) * 0x1.0p-53;
          vy = ic->amp * std::exp(-(y / ic->sigma) * (y / ic->sigma)) * (2.0 * u01 - 1.0);
        }
        c->p(nden, k, j, i) = ic->rho0;
        c->p(npre, k, j, i) = ic->p0;
        c->p(nve1, k, j, i) = ic->v0 * th;
        c->p(nve2, k, j, i) = vy;
        c->p(nve3, k, j, i) = 0.0;
        c->p(nbm1, k, j, i) = ic->bx0;
        c->p(nbm2, k, j, i) = 0.0;
        c->p(nbm3, k, j, i) = 0.0;
        c->p(nbps, k, j, i) = 0.0;
        c->p(nene, k, j, i) = ic->p0 / ((gam - 1.0) * ic->rho0);
        c->p(ncsp, k, j, i) = std::sqrt(gam * ic->p0 / ic->rho0);
        c->p(nst, k, j, i) = 0.5 * (1.0 + th);
      }
  orc_prim_to_cons_all(c);
}

// Parallel regions only for boxes of at least 32768 interior cells: on the small boxes of the
// convergence and closed-form checks the fork/join cost outweighs the work, and a thread team
// that spins beside other processes slows the whole step loop by two orders of magnitude.
static inline bool orc_par(const orc_ctx* c) {
  return int64_t(c->ie - c->is + 1) * (c->je - c->js + 1) * (c->ke - c->ks + 1) >= 32768;
}

// ControlTimestep, KH/mhd.cpp:1511-1561 (single rank: MPIminfind is the identity)
double orc_control_timestep(orc_ctx* c) {
  double dtmin = 1.0e90;
#pragma omp parallel for collapse(2) reduction(min : dtmin) if (orc_par(c))
  for (int k = c->ks; k <= c->ke; ++k)
    for (int j = c->js; j <= c->je; ++j)
      for (int i = c->is; i <= c->ie; ++i) {
        const double b1 = c->p(nbm1, k, j, i), b2 = c->p(nbm2, k, j, i), b3 = c->p(nbm3, k, j, i);
        const double cs = c->p(ncsp, k, j, i);
        const double ctot = std::sqrt(cs * cs + (b1 * b1 + b2 * b2 + b3 * b3) / c->p(nden, k, j, i));
        const double d1 = (c->x1a[i + 1] - c->x1a[i]) / (std::abs(c->p(nve1, k, j, i)) + ctot);
        const double d2 = (c->x2a[j + 1] - c->x2a[j]) / (std::abs(c->p(nve2, k, j, i)) + ctot);
        const double d3 = (c->x3a[k + 1] - c->x3a[k]) / (std::abs(c->p(nve3, k, j, i)) + ctot);
        dtmin = std::min(dtmin, std::min(d1, std::min(d2, d3)));
      }
  return c->cfg.cfl * dtmin;
}

// EvaluateCh, KH/mhd.cpp:1563-1611
double orc_evaluate_ch(orc_ctx* c) {
  double chg = 0.0;
#pragma omp parallel for collapse(2) reduction(max : chg) if (orc_par(c))
  for (int k = c->ks; k <= c->ke; ++k)
    for (int j = c->js; j <= c->je; ++j)
      for (int i = c->is; i <= c->ie; ++i) {
        const double rho = c->p(nden, k, j, i);
        const double css = c->p(ncsp, k, j, i) * c->p(ncsp, k, j, i);
        const double ca[3] = {c->p(nbm1, k, j, i) * c->p(nbm1, k, j, i) / rho,
                              c->p(nbm2, k, j, i) * c->p(nbm2, k, j, i) / rho,
                              c->p(nbm3, k, j, i) * c->p(nbm3, k, j, i) / rho};
        const double cts = css + ca[0] + ca[1] + ca[2];
        for (int a = 0; a < 3; ++a) {
          const double cm = std::sqrt((cts + std::sqrt(cts * cts - 4.0 * css * ca[a])) / 2.0);
          chg = std::max(chg, std::abs(c->p(nve1 + a, k, j, i)) + cm);
        }
      }
  return chg;
}

// SetBoundaryCondition + SendRecvBoundary for one tile per axis
// (KH/boundary.cpp:750-893 with the ntiles==1 branches at :117-221, :321-421, :518-618):
// all six send buffers are packed from the pre-exchange P, converted by the boundary
// rule, then unpacked x, y, z in that order.
void orc_set_boundary(orc_ctx* c) {
  const int I = c->itot, J = c->jtot, K = c->ktot;
  const int ie = c->ie, je = c->je, ke = c->ke, is = c->is, js = c->js, ks = c->ks;
  // buffers indexed [n][k][j][g] (x), [n][k][g][i] (y), [n][g][j][i] (z)
  std::vector<double> sxs(NPRIM * K * J * NGH), sxe(sxs.size()), rxs(sxs.size()), rxe(sxs.size());
  std::vector<double> sys(NPRIM * K * NGH * I), sye(sys.size()), rys(sys.size()), rye(sys.size());
  std::vector<double> szs(NPRIM * NGH * J * I), sze(szs.size()), rzs(szs.size()), rze(szs.size());
  auto X = [&](int n, int k, int j, int g) { return ((static_cast<size_t>(n) * K + k) * J + j) * NGH + g; };
  auto Y = [&](int n, int k, int g, int i) { return ((static_cast<size_t>(n) * K + k) * NGH + g) * I + i; };
  auto Z = [&](int n, int g, int j, int i) { return ((static_cast<size_t>(n) * NGH + g) * J + j) * I + i; };
  for (int n = 0; n < NPRIM; ++n)
    for (int k = 0; k < K; ++k)
      for (int j = 0; j < J; ++j)
        for (int g = 0; g < NGH; ++g) {
          sxs[X(n, k, j, g)] = c->p(n, k, j, ie - NGH + 1 + g);
          sxe[X(n, k, j, g)] = c->p(n, k, j, is + g);
        }
  for (int n = 0; n < NPRIM; ++n)
    for (int k = 0; k < K; ++k)
      for (int g = 0; g < NGH; ++g)
        for (int i = 0; i < I; ++i) {
          sys[Y(n, k, g, i)] = c->p(n, k, je - NGH + 1 + g, i);
          sye[Y(n, k, g, i)] = c->p(n, k, js + g, i);
        }
  for (int n = 0; n < NPRIM; ++n)
    for (int g = 0; g < NGH; ++g)
      for (int j = 0; j < J; ++j)
        for (int i = 0; i < I; ++i) {
          szs[Z(n, g, j, i)] = c->p(n, ke - NGH + 1 + g, j, i);
          sze[Z(n, g, j, i)] = c->p(n, ks + g, j, i);
        }
  const int* bc = c->cfg.bc;
  // x (KH/boundary.cpp:152-220)
  for (int k = 0; k < K; ++k)
    for (int j = 0; j < J; ++j)
      for (int g = 0; g < NGH; ++g)
        for (int n = 0; n < NPRIM; ++n) {
          double vin = 0.0, vout = 0.0;
          if (bc[0] == BC_PERIODIC) vin = sxs[X(n, k, j, g)];
          else if (bc[0] == BC_REFLECT) vin = sxe[X(n, k, j, NGH - 1 - g)] * (n == nve1 ? -1.0 : 1.0);
          else vin = sxe[X(n, k, j, 0)];
          if (bc[1] == BC_PERIODIC) vout = sxe[X(n, k, j, g)];
          else if (bc[1] == BC_REFLECT) vout = sxs[X(n, k, j, NGH - 1 - g)] * (n == nve1 ? -1.0 : 1.0);
          else vout = sxs[X(n, k, j, NGH - 1)];
          rxs[X(n, k, j, g)] = vin;
          rxe[X(n, k, j, g)] = vout;
        }
  // y (KH/boundary.cpp:354-420)
  for (int k = 0; k < K; ++k)
    for (int g = 0; g < NGH; ++g)
      for (int i = 0; i < I; ++i)
        for (int n = 0; n < NPRIM; ++n) {
          double vin = 0.0, vout = 0.0;
          if (bc[2] == BC_PERIODIC) vin = sys[Y(n, k, g, i)];
          else if (bc[2] == BC_REFLECT) vin = sye[Y(n, k, NGH - 1 - g, i)] * (n == nve2 ? -1.0 : 1.0);
          else vin = sye[Y(n, k, 0, i)];
          if (bc[3] == BC_PERIODIC) vout = sye[Y(n, k, g, i)];
          else if (bc[3] == BC_REFLECT) vout = sys[Y(n, k, NGH - 1 - g, i)] * (n == nve2 ? -1.0 : 1.0);
          else vout = sys[Y(n, k, NGH - 1, i)];
          rys[Y(n, k, g, i)] = vin;
          rye[Y(n, k, g, i)] = vout;
        }
  // z (KH/boundary.cpp:551-617)
  for (int g = 0; g < NGH; ++g)
    for (int j = 0; j < J; ++j)
      for (int i = 0; i < I; ++i)
        for (int n = 0; n < NPRIM; ++n) {
          double vin = 0.0, vout = 0.0;
          if (bc[4] == BC_PERIODIC) vin = szs[Z(n, g, j, i)];
          else if (bc[4] == BC_REFLECT) vin = sze[Z(n, NGH - 1 - g, j, i)] * (n == nve3 ? -1.0 : 1.0);
          else vin = sze[Z(n, 0, j, i)];
          if (bc[5] == BC_PERIODIC) vout = sze[Z(n, g, j, i)];
          else if (bc[5] == BC_REFLECT) vout = szs[Z(n, NGH - 1 - g, j, i)] * (n == nve3 ? -1.0 : 1.0);
          else vout = szs[Z(n, NGH - 1, j, i)];
          rzs[Z(n, g, j, i)] = vin;
          rze[Z(n, g, j, i)] = vout;
        }
  // unpack x, y, z (KH/boundary.cpp:842-891)
  for (int n = 0; n < NPRIM; ++n)
    for (int k = 0; k < K; ++k)
      for (int j = 0; j < J; ++j)
        for (int g = 0; g < NGH; ++g) {
          c->p(n, k, j, is - NGH + g) = rxs[X(n, k, j, g)];
          c->p(n, k, j, ie + 1 + g) = rxe[X(n, k, j, g)];
        }
  for (int n = 0; n < NPRIM; ++n)
    for (int k = 0; k < K; ++k)
      for (int g = 0; g < NGH; ++g)
        for (int i = 0; i < I; ++i) {
          c->p(n, k, js - NGH + g, i) = rys[Y(n, k, g, i)];
          c->p(n, k, je + 1 + g, i) = rye[Y(n, k, g, i)];
        }
  for (int n = 0; n < NPRIM; ++n)
    for (int g = 0; g < NGH; ++g)
      for (int j = 0; j < J; ++j)
        for (int i = 0; i < I; ++i) {
          c->p(n, ks - NGH + g, j, i) = rzs[Z(n, g, j, i)];
          c->p(n, ke + 1 + g, j, i) = rze[Z(n, g, j, i)];
        }
}

// GetNumericalFluxD, KH/mhd.cpp:987-1410: faces i in [is,ie+1] (dir 1),
// j in [js,je+1] (dir 2), k in [ks,ke+1] (dir 3); the other two indices interior.
void orc_numerical_flux(orc_ctx* c, int dir, double chg) {
  const int d = dir - 1;
  const int eflux = c->cfg.energy_flux;
  const int iu = c->ie + (d == 0), ju = c->je + (d == 1), ku = c->ke + (d == 2);
#pragma omp parallel for collapse(2) if (orc_par(c))
  for (int k = c->ks; k <= ku; ++k)
    for (int j = c->js; j <= ju; ++j)
      for (int i = c->is; i <= iu; ++i) {
        double q[4][NPRIM], F[NCONS];
        for (int s = 0; s < 4; ++s) {
          const int o = s - 2;
          load_prim(c, k + (d == 2) * o, j + (d == 1) * o, i + (d == 0) * o, q[s]);
        }
        face_flux(q[0], q[1], q[2], q[3], d, eflux, chg, F);
        for (int n = 0; n < NCONS; ++n) c->f(d, n, k, j, i) = F[n];
      }
}

// UpdateConservU, KH/mhd.cpp:1412-1444.  The y and z spacings are taken per axis
// (x2a(j+1)-x2a(j), x3a(k+1)-x3a(k)); the reference indexes them with i, which is the
// same spacing up to round-off on the uniform cubic grids it runs.
void orc_update_conserv(orc_ctx* c, double dt) {
#pragma omp parallel for collapse(2) if (orc_par(c))
  for (int k = c->ks; k <= c->ke; ++k)
    for (int j = c->js; j <= c->je; ++j)
      for (int i = c->is; i <= c->ie; ++i) {
        const double xinv = 1.0 / (c->x1a[i + 1] - c->x1a[i]);
        const double yinv = 1.0 / (c->x2a[j + 1] - c->x2a[j]);
        const double zinv = 1.0 / (c->x3a[k + 1] - c->x3a[k]);
        for (int m = 0; m < NCONS; ++m) {
          c->u(m, k, j, i) -= dt * ((c->f(0, m, k, j, i + 1) - c->f(0, m, k, j, i)) * xinv +
                                    (c->f(1, m, k, j + 1, i) - c->f(1, m, k, j, i)) * yinv +
                                    (c->f(2, m, k + 1, j, i) - c->f(2, m, k, j, i)) * zinv);
        }
      }
}

// DampPsi, KH/mhd.cpp:1613-1633
void orc_damp_psi(orc_ctx* c, double dt, double chg) {
  const double alphabp = c->cfg.alphabp;
#pragma omp parallel for collapse(2) if (orc_par(c))
  for (int k = c->ks; k <= c->ke; ++k)
    for (int j = c->js; j <= c->je; ++j)
      for (int i = c->is; i <= c->ie; ++i) {
        const double dhl = std::min(c->x1a[i + 1] - c->x1a[i],
                                    std::min(c->x2a[j + 1] - c->x2a[j], c->x3a[k + 1] - c->x3a[k]));
        const double taui = alphabp * chg / dhl;
        c->u(mbps, k, j, i) = c->u(mbps, k, j, i) * (1.0 - dt * taui);
      }
}

// UpdatePrimitvP, KH/mhd.cpp:1447-1509
void orc_update_primitive(orc_ctx* c) {
  const double gam = c->cfg.gam;
#pragma omp parallel for collapse(2) if (orc_par(c))
  for (int k = c->ks; k <= c->ke; ++k)
    for (int j = c->js; j <= c->je; ++j)
      for (int i = c->is; i <= c->ie; ++i) {
        const double rho = c->u(mden, k, j, i);
        const double rhoinv = 1.0 / rho;
        const double m1 = c->u(mrv1, k, j, i), m2 = c->u(mrv2, k, j, i), m3 = c->u(mrv3, k, j, i);
        const double b1 = c->u(mbm1, k, j, i), b2 = c->u(mbm2, k, j, i), b3 = c->u(mbm3, k, j, i);
        c->p(nden, k, j, i) = rho;
        c->p(nve1, k, j, i) = m1 * rhoinv;
        c->p(nve2, k, j, i) = m2 * rhoinv;
        c->p(nve3, k, j, i) = m3 * rhoinv;
        const double ekin = 0.5 * (m1 * m1 + m2 * m2 + m3 * m3) * rhoinv;
        const double emag = 0.5 * (b1 * b1 + b2 * b2 + b3 * b3);
        const double e = (c->u(meto, k, j, i) - ekin - emag) / rho;
        c->p(nene, k, j, i) = e;
        c->p(npre, k, j, i) = e * rho * (gam - 1.0);
        c->p(ncsp, k, j, i) = std::sqrt(e * gam * (gam - 1.0));
        c->p(nbm1, k, j, i) = b1;
        c->p(nbm2, k, j, i) = b2;
        c->p(nbm3, k, j, i) = b3;
        c->p(nbps, k, j, i) = c->u(mbps, k, j, i);
        c->p(nst, k, j, i) = c->u(mst, k, j, i) * rhoinv;
      }
}

// The loop body of KH/main.cpp:229-259, nsteps times.
void orc_save_cons(orc_ctx* c) { c->U0 = c->U; }

void orc_average_saved(orc_ctx* c) {
#pragma omp parallel for collapse(2) if (orc_par(c))
  for (int k = c->ks; k <= c->ke; ++k)
    for (int j = c->js; j <= c->je; ++j)
      for (int i = c->is; i <= c->ie; ++i)
        for (int m = 0; m < NCONS; ++m) {
          const size_t x = c->idx(NCONS, m, k, j, i);
          c->U[x] = 0.5 * (c->U0[x] + c->U[x]);
        }
}

// The loop body of KH/main.cpp:229-259.  stages = 2 wraps the reference's forward-Euler
// stage E(U) = DampPsi(U - dt div F(U)) (the calls below, KH/main.cpp:246-257) in the
// two-stage strong-stability-preserving Runge-Kutta step of Shu & Osher (J. Comput. Phys.
// 77:439, 1988): U* = E(U^n), U^{n+1} = (U^n + E(U*)) / 2, both stages with the dt and c_h
// of U^n.  The reference itself has no Runge-Kutta step.
double orc_run(orc_ctx* c, int nsteps, double t0, double* dt_log) {
  double t = t0;
  const bool rk2 = c->cfg.stages == 2;
  for (int s = 0; s < nsteps; ++s) {
    const double dt = orc_control_timestep(c);
    if (dt_log) dt_log[s] = dt;
    orc_set_boundary(c);
    const double chg = orc_evaluate_ch(c);
    if (rk2) orc_save_cons(c);
    for (int stage = 0; stage < (rk2 ? 2 : 1); ++stage) {
      if (stage == 1) orc_set_boundary(c);
      orc_numerical_flux(c, 1, chg);
      orc_numerical_flux(c, 2, chg);
      orc_numerical_flux(c, 3, chg);
      orc_update_conserv(c, dt);
      orc_damp_psi(c, dt, chg);
      if (stage == 1) orc_average_saved(c);
      orc_update_primitive(c);
    }
    t += dt;
  }
  return t;
}

// RealTimeAnalysis sums (src_cpp_omp_device/main.cpp:23-69) and max|divB|.
// Sums are accumulated per k-plane, then over planes in k order, so the result does not
// depend on the thread count.  divB is the 2nd-order central difference of the
// cell-centred field, taken on interior cells whose six neighbours are interior.
void orc_diagnostics(orc_ctx* c, double* out) {
  const int nk = c->ke - c->ks + 1;
  std::vector<double> vol(nk), mix(nk), vy2(nk), dmax(nk);
#pragma omp parallel for if (orc_par(c))
  for (int kk = 0; kk < nk; ++kk) {
    const int k = c->ks + kk;
    double sv = 0.0, sm = 0.0, sy = 0.0, md = 0.0;
    for (int j = c->js; j <= c->je; ++j)
      for (int i = c->is; i <= c->ie; ++i) {
        const double dvol =
            (c->x1a[i + 1] - c->x1a[i]) * (c->x2a[j + 1] - c->x2a[j]) * (c->x3a[k + 1] - c->x3a[k]);
        sv += dvol;
        sm += c->p(nst, k, j, i) * (1.0 - c->p(nst, k, j, i)) * dvol;
        sy += c->p(nve2, k, j, i) * c->p(nve2, k, j, i) * dvol;
        if (i > c->is && i < c->ie && j > c->js && j < c->je && k > c->ks && k < c->ke) {
          const double db = (c->p(nbm1, k, j, i + 1) - c->p(nbm1, k, j, i - 1)) / (c->x1b[i + 1] - c->x1b[i - 1]) +
                            (c->p(nbm2, k, j + 1, i) - c->p(nbm2, k, j - 1, i)) / (c->x2b[j + 1] - c->x2b[j - 1]) +
                            (c->p(nbm3, k + 1, j, i) - c->p(nbm3, k - 1, j, i)) / (c->x3b[k + 1] - c->x3b[k - 1]);
          md = std::max(md, std::abs(db));
        }
      }
    vol[kk] = sv; mix[kk] = sm; vy2[kk] = sy; dmax[kk] = md;
  }
  double sv = 0.0, sm = 0.0, sy = 0.0, md = 0.0;
  for (int kk = 0; kk < nk; ++kk) {
    sv += vol[kk]; sm += mix[kk]; sy += vy2[kk]; md = std::max(md, dmax[kk]);
  }
  out[0] = sv;
  out[1] = sm / sv;
  out[2] = std::sqrt(sy / sv);
  out[3] = md;
}

double orc_vanleer(double dsvp, double dsvm) { return van_leer(dsvp, dsvm); }
void orc_prim2riemann(const double* prim12, int dir, int energy_flux, double* out23) {
  prim2riemann(prim12, dir - 1, energy_flux, out23);
}
void orc_hlld(const double* left23, const double* right23, double* nflux10) { hlld(left23, right23, nflux10); }
void orc_face_flux(const double* pm2, const double* pm1, const double* p0, const double* pp1, int dir,
                   int energy_flux, double chg, double* f10) {
  face_flux(pm2, pm1, p0, pp1, dir - 1, energy_flux, chg, f10);
}

}  // extern "C"
