/*
 * mhd_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of the MHDTurbulence (src_cpp_kokkos_host) time step, used as the
 * checker for the HIP library.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  Nothing in the product links it.
 *
 * Host array layout is the reference FieldArray layout: element (n,k,j,i) of an
 * array with nv variables lives at ((k*jtot + j)*nv + n)*itot + i
 * (Kokkos::View<double****,LayoutLeft>(i,n,j,k), src_cpp_kokkos_host/mhd.hpp:164-175).
 *
 * Pinning: tests/test_oracle.py checks every function against closed forms, invariants,
 * independent textbook solvers and the reference's own output files.  Parity unpinned:
 *  - energy_flux = 0 with B != 0: the (E+pt)vn - |B|^2 Bn energy flux is restated from
 *    src_cpp_kokkos_host/mhd.cpp:1049-1050; the tests pin only its exact difference from the
 *    physical (v.B) Bn form, since no reference run with B != 0 could be made;
 *  - max|divB| in orc_diagnostics has no counterpart in the reference.
 */
#ifndef MHD_ORACLE_H_
#define MHD_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_config {
  int nx, ny, nz;          /* interior cells per axis (ngrid1..3) */
  int bc[6];               /* xin,xout,yin,yout,zin,zout: 1 periodic, 2 reflection, 3 outflow */
  double xmin[3], xmax[3]; /* domain */
  double gam;              /* adiabatic index (reference C++: 1.4) */
  double cfl;              /* dt = cfl * dtmin (reference: 0.05) */
  double alphabp;          /* psi damping constant (reference: 0.1) */
  int energy_flux;         /* 0: (E+pt)vn - |B|^2 Bn (src_cpp_kokkos_host); 1: (E+pt)vn - (v.B) Bn */
  int grid_shift;          /* 1: x1a(i)=dx*(i-(ngh+1))+xmin (C++ variants); 0: dx*(i-ngh)+xmin */
  int stages;              /* 0 or 1: forward Euler, the reference's step; 2: two-stage SSP Runge-Kutta */
} orc_config;

typedef struct orc_ic_params {
  double rho0, p0;         /* uniform density and pressure */
  double v0;               /* vx = v0*tanh(y/a) */
  double a;                /* shear width */
  double bx0;              /* uniform Bx */
  double amp;              /* vy = amp*exp(-(y/sigma)^2)*(2u-1) */
  double sigma;
  uint64_t seed;
} orc_ic_params;

typedef struct orc_ctx orc_ctx;

orc_ctx* orc_create(const orc_config* cfg);
void     orc_destroy(orc_ctx* c);
void     orc_set_threads(int n);
void     orc_sizes(const orc_ctx* c, int* itot, int* jtot, int* ktot);
double*  orc_prim(orc_ctx* c);            /* nprim=12 vars, reference layout */
double*  orc_cons(orc_ctx* c);            /* mconsv=10 vars */
double*  orc_flux(orc_ctx* c, int dir);   /* 10 vars, dir = 1,2,3 */
double*  orc_grid(orc_ctx* c, int axis, int centre); /* axis 1..3; centre 0 = x?a, 1 = x?b */

void   orc_generate_grid(orc_ctx* c);
void   orc_generate_problem_kh(orc_ctx* c, double vy_amp, double rrv);
void   orc_generate_problem_shear(orc_ctx* c, const orc_ic_params* p);
void   orc_prim_to_cons_all(orc_ctx* c);

double orc_control_timestep(orc_ctx* c);     /* returns dt = cfl*dtmin */
void   orc_set_boundary(orc_ctx* c);
double orc_evaluate_ch(orc_ctx* c);
void   orc_numerical_flux(orc_ctx* c, int dir, double chg);
void   orc_update_conserv(orc_ctx* c, double dt);
void   orc_damp_psi(orc_ctx* c, double dt, double chg);
void   orc_update_primitive(orc_ctx* c);
/* U <- (U0 + U) / 2 on interior cells, U0 the copy orc_save_cons took (the second stage of
 * the two-stage Runge-Kutta step) */
void   orc_save_cons(orc_ctx* c);
void   orc_average_saved(orc_ctx* c);
/* nsteps of the KH loop body; returns the simulation time; dt_log (may be NULL) gets each dt */
double orc_run(orc_ctx* c, int nsteps, double t0, double* dt_log);
/* out[0]=volume, out[1]=mix, out[2]=sqrt(<vy^2>), out[3]=max|divB| */
void   orc_diagnostics(orc_ctx* c, double* out);

/* single-point kernels */
double orc_vanleer(double dsvp, double dsvm);
void   orc_prim2riemann(const double* prim12, int dir, int energy_flux, double* out23);
void   orc_hlld(const double* left23, const double* right23, double* nflux10);
void   orc_face_flux(const double* pm2, const double* pm1, const double* p0, const double* pp1,
                     int dir, int energy_flux, double chg, double* f10);
uint64_t orc_splitmix64(uint64_t x);

#ifdef __cplusplus
}
#endif
#endif
