/*
 * mhd_amd.h -- C ABI of the MI355X MHD time-step library (libmhd_amd.so).
 *
 * One context = one box of the Kelvin-Helmholtz MHD problem resident in HBM.  The
 * entry points mirror the functions the src_cpp_kokkos_host step loop calls
 * (src_cpp_kokkos_host/main.cpp:229-259); each declaration names the reference
 * function it replaces.  Plain pointers and sizes only; every function returns 0 on
 * success, a negative MHD_ERR_* on failure and never throws.  Host arrays use the
 * reference FieldArray layout: element (n,k,j,i) of an array with nv variables at
 * ((k*jtot + j)*nv + n)*itot + i (src_cpp_kokkos_host/mhd.hpp:164-175), with
 * itot = nx+5, jtot = ny+5, ktot = nz+5 (src_cpp_kokkos_host/config.hpp:56-58).
 *
 * Threading: one host thread per context.  Calls on one context are ordered on the
 * context's stream; those that return host values synchronise that stream.
 */
#ifndef MHD_AMD_H_
#define MHD_AMD_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MHD_AMD_VERSION 1

enum {
  MHD_OK = 0,
  MHD_ERR_ARG = -1,        /* invalid argument or configuration */
  MHD_ERR_HIP = -2,        /* a HIP runtime call failed */
  MHD_ERR_NOMEM = -3,      /* device allocation failed */
  MHD_ERR_UNSUPPORTED = -4,
  MHD_ERR_STATE = -5       /* call out of order (e.g. no grid set) */
};

/* physics and grid of one box; the reference fixes these at compile time
 * (src_cpp_kokkos_host/config.hpp:10-42, mhd.cpp:27,1560,1614) */
typedef struct mhd_config {
  int nx, ny, nz;          /* interior cells per axis (ngridtotal1..3, one tile) */
  int bc[6];               /* boundary_xin,xout,yin,yout,zin,zout: 1 periodic, 2 reflection, 3 outflow */
  double xmin[3], xmax[3]; /* x1min..x3max */
  double gam;              /* adiabatic index; the reference C++ uses 1.4 */
  double cfl;              /* dt = cfl*dtmin; the reference uses 0.05 */
  double alphabp;          /* psi damping constant; the reference uses 0.1 */
  int energy_flux;         /* 0: (E+pt)v_n - |B|^2 B_n, as src_cpp_kokkos_host/mhd.cpp:1049-1050;
                              1: (E+pt)v_n - (v.B) B_n, as src_cpp_omp_device/mhd.cpp:523-526 */
  int grid_shift;          /* 1: x1a(i) = dx*(i-(ngh+1))+xmin (src_cpp_kokkos_host/main.cpp:39-41);
                              0: dx*(i-ngh)+xmin */
  int precision;           /* 0: FP64 throughout; 1: reconstruction + HLLD in FP32;
                              2: as 1, and the fused step's x-y flux divergence stored in FP32
                              between its two kernels (the update itself stays FP64) */
  int device;              /* HIP device ordinal */
  int stages;              /* 0 or 1: one forward-Euler stage per step, the reference's
                              (src_cpp_kokkos_host/main.cpp:229-259); 2: the two-stage SSP
                              Runge-Kutta step over that stage, U* = E(U^n),
                              U^{n+1} = (U^n + E(U*))/2, dt and c_h from U^n (no counterpart
                              in the reference) */
} mhd_config;

/* single-shear-layer initial state (see mhd_generate_problem_shear) */
typedef struct mhd_ic_params {
  double rho0, p0, v0, a, bx0, amp, sigma;
  uint64_t seed;
} mhd_ic_params;

typedef struct mhd_ctx mhd_ctx;

int mhd_version(void);
const char* mhd_error_string(int code);
/* last error message recorded on this thread (HIP error text or argument check) */
const char* mhd_last_error(void);

/* AllocateHydroVariables (src_cpp_kokkos_host/mhd.cpp:33-77) + AllocateBoundaryVariables
 * (src_cpp_kokkos_host/boundary.cpp:36-94): allocates P, U in HBM, zero-filled. */
int mhd_create(const mhd_config* cfg, mhd_ctx** out);
/* DeallocateHydroVariables / DeallocateBoundaryVariables */
int mhd_destroy(mhd_ctx* ctx);
int mhd_get_dims(const mhd_ctx* ctx, int* itot, int* jtot, int* ktot);
/* device bytes held by the context */
int mhd_device_bytes(const mhd_ctx* ctx, uint64_t* bytes);

/* GridArray upload: x1a,x1b (itot entries), x2a,x2b (jtot), x3a,x3b (ktot), as
 * src_cpp_kokkos_host/mhd.hpp:67-110 holds them */
int mhd_set_grid(mhd_ctx* ctx, const double* x1a, const double* x1b, const double* x2a, const double* x2b,
                 const double* x3a, const double* x3b);
/* GenerateGrid (src_cpp_kokkos_host/main.cpp:26-65) computed on the device */
int mhd_generate_grid(mhd_ctx* ctx);
int mhd_get_grid(mhd_ctx* ctx, double* x1a, double* x1b, double* x2a, double* x2b, double* x3a, double* x3b);

/* FieldArray h2d/d2h (src_cpp_kokkos_device/mhd.hpp FieldArray::h2d/d2h): whole arrays,
 * ghosts included, host layout above.  P has 12 variables, U and F have 10. */
int mhd_upload_prim(mhd_ctx* ctx, const double* P);
int mhd_upload_cons(mhd_ctx* ctx, const double* U);
int mhd_download_prim(mhd_ctx* ctx, double* P);
int mhd_download_cons(mhd_ctx* ctx, double* U);
int mhd_download_flux(mhd_ctx* ctx, int dir, double* F);

/* Asynchronous snapshot of P (no counterpart in the reference, whose Output reads the host
 * mirror synchronously, src_cpp_kokkos_host/output.cpp:262-280): _begin queues a copy of the
 * current primitive state, in the host layout above, into library-owned pinned host memory
 * on a second stream and returns at once, so the transfer overlaps the steps that follow;
 * _end waits for it and points *P_host at the copy, which stays valid until the next _begin
 * or mhd_destroy.  A _begin while an earlier snapshot is in flight first waits for it. */
int mhd_snapshot_begin(mhd_ctx* ctx);
int mhd_snapshot_end(mhd_ctx* ctx, const double** P_host);

/* ---- the reference step functions, one call each ---- */
/* ControlTimestep (src_cpp_kokkos_host/mhd.cpp:1511-1561): *dt = cfl*min CFL bound of P */
int mhd_control_timestep(mhd_ctx* ctx, double* dt);
/* SetBoundaryCondition (src_cpp_kokkos_host/boundary.cpp:750-893) on P */
int mhd_set_boundary(mhd_ctx* ctx);
/* EvaluateCh (src_cpp_kokkos_host/mhd.cpp:1563-1611): *chg = max fast speed of P */
int mhd_evaluate_ch(mhd_ctx* ctx, double* chg);
/* GetNumericalFluxD (src_cpp_kokkos_host/mhd.cpp:987-1410): F_dir from P, dir = 1,2,3 */
int mhd_numerical_flux(mhd_ctx* ctx, int dir, double chg);
/* UpdateConservU (src_cpp_kokkos_host/mhd.cpp:1412-1444) */
int mhd_update_conserv(mhd_ctx* ctx, double dt);
/* DampPsi (src_cpp_kokkos_host/mhd.cpp:1613-1633) */
int mhd_damp_psi(mhd_ctx* ctx, double dt, double chg);
/* UpdatePrimitvP (src_cpp_kokkos_host/mhd.cpp:1447-1509) */
int mhd_update_primitive(mhd_ctx* ctx);

/* ---- the whole loop body on the device ----
 * nsteps iterations of src_cpp_kokkos_host/main.cpp:229-259 with dt and c_h kept in
 * HBM.  *time is read as the starting time_sim and returns the final one; with
 * tmax > 0 the loop stops after the step that takes time past tmax (main.cpp:265).
 * dt_log (may be NULL, capacity nsteps) receives each dt.  *steps_done (may be NULL)
 * returns the steps applied.  Every other entry point sees P and U of the final state.
 * A fused run (mode 0) leaves that state in the fused loop's buffers, with P not yet
 * recovered from it; the first call of another entry point that reads or writes P, U, F or
 * the tiling brings P and U up to date, and a following fused run continues from the
 * buffers and from the bounds the last step reduced (its first dt and c_h are not
 * recomputed).  A run that met a bad cell brings them up to date before it returns.
 * mode: 0 = fused step kernel (default), 1 = staged kernels (one per reference function).
 * With mhd_config.stages = 2 each iteration is the two-stage Runge-Kutta step over that
 * loop body's update (one dt and c_h per iteration, from its starting state).
 * A cell with rho <= 0, or whose CFL bound or fast speed is not finite and positive (a NaN,
 * or e < 0), stops the loop and sets status bit 0 (mhd_status); the reference checks
 * nothing and would go on with a NaN dt (src_cpp_kokkos_host/mhd.cpp:1499,1560).  The fused
 * path stops after the step that produced the cell (its dt/c_h reduction runs on the new
 * state), the staged path before the step that would read it; *steps_done counts the steps
 * applied, and the call still returns MHD_OK. */
int mhd_run(mhd_ctx* ctx, int nsteps, double tmax, int mode, double* time, double* dt_log, int* steps_done);

/* The fused loop body one reference step at a time, for a caller that issues the reference
 * loop's calls in their order (the C++ drop-in layer, src_cpp_kokkos_host/main.cpp:229-257)
 * and wants the fused kernels' speed.  _begin_step is that iteration's ControlTimestep and
 * EvaluateCh (src_cpp_kokkos_host/mhd.cpp:1511-1611): *dt = cfl*dtmin and *chg of the
 * current state, read from the bounds the previous step's fused kernels reduced (one small
 * device-to-host read, which waits for that step); *status (may be NULL) as mhd_status.
 * _step applies the rest of the iteration (SetBoundaryCondition ... UpdatePrimitvP) with
 * the fused kernels and that dt and c_h, and returns at once.  Between the calls the state
 * lives in the fused loop's buffers and P is not updated; the first call of any other entry
 * point that reads or writes P, U, F or the tiling first brings P and U up to date (so a
 * stretch of steps ends there), as after a fused mhd_run. */
int mhd_fused_begin_step(mhd_ctx* ctx, double* dt, double* chg, int* status);
int mhd_fused_step(mhd_ctx* ctx);

/* Allocation-free variant of mhd_run in mode 0 (the fused step).  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  It calls no hipMalloc,
 * hipFree or allocating library routine: the fused loop's scratch (its second copy of the
 * conserved state, the flux divergence and a dt log of 4096 entries per chunk of steps)
 * is a static device array of MHD_NOALLOC_SCRATCH_BYTES that each GPU the library runs on
 * holds until the process ends.  The caller's context holds the inputs and outputs (P, U,
 * grid, device scalars; created beforehand with mhd_create).  A box whose scratch exceeds
 * the array (about 80^3 cells in FP64) is refused with MHD_ERR_ARG and a message.
 * Arguments, results and stop rules are those of mhd_run with mode 0 (dt_log, if not NULL,
 * has room for nsteps entries).  Every call on one GPU shares that GPU's array, so calls
 * that overlap on one GPU (other host threads, other contexts' streams) race: issue them
 * one at a time.  The call returns after its steps are done. */
#define MHD_NOALLOC_SCRATCH_BYTES (128u * 1024u * 1024u)
int mhd_run_noalloc(mhd_ctx* ctx, int nsteps, double tmax, double* time, double* dt_log, int* steps_done);

/* Kelvin-Helmholtz initial state of src_cpp_kokkos_host/main.cpp:66-184 on the device
 * (vy amplitude vy_amp: 1e-3 there, 1e-2 in src_cpp_omp_device; harmonic vx kick rrv). */
int mhd_generate_problem_kh(mhd_ctx* ctx, double vy_amp, double rrv);
/* single shear layer: rho=rho0, p=p0, vx=v0*tanh(y/a), Bx=bx0, X=(1+tanh(y/a))/2,
 * vy = amp*exp(-(y/sigma)^2)*(2u-1), u=(splitmix64(seed ^ ((k*ny+j)*nx+i))>>11)*2^-53 */
int mhd_generate_problem_shear(mhd_ctx* ctx, const mhd_ic_params* ic);

/* t-prof diagnostics of RealTimeAnalysis (src_cpp_omp_device/main.cpp:23-69) on P:
 * out[0]=volume, out[1]=mix, out[2]=sqrt(<vy^2>), out[3]=max|divB| (central differences
 * on cells whose neighbours are interior; the reference has no divB diagnostic). */
int mhd_diagnostics(mhd_ctx* ctx, double* out4);
/* status word: bit 0 = a cell with rho <= 0 or a non-finite or non-positive CFL bound / fast
 * speed was met (see mhd_run); it stays set for the life of the context */
int mhd_status(mhd_ctx* ctx, int* status);

/* kernel timing of the most recent mhd_run: total ms of the step kernels on the
 * context's stream, measured with HIP events, and the launches they cover */
int mhd_last_run_timing(mhd_ctx* ctx, double* kernel_ms, int* launches);

/* Per-kernel timing (the phase timers of src_cpp_kokkos_device/main.cpp:293-373, as HIP
 * events): runs nsteps loop iterations of the given mode launch by launch, with an event
 * after every launch (recorded up front; the host waits once, at the end), and writes the
 * average ms of each launch position of an iteration into avg_ms (capacity cap); the state
 * advances by nsteps steps. */
int mhd_profile_kernels(mhd_ctx* ctx, int nsteps, int mode, double* avg_ms, int cap, int* nkernels);
/* comma-separated names of the launch positions of one iteration of the given mode */
const char* mhd_kernel_names(int mode);

/* ---- batches: many boxes advanced together ----
 * No counterpart in the reference, which runs one box per MPI job.  A batch holds up to
 * max_boxes boxes of different sizes and initial states that share base's physics
 * (gam, cfl, alphabp, bc, domain, energy_flux, grid_shift, precision, device).  Every
 * step issues one launch per fused kernel for all live boxes; each box keeps its own dt /
 * c_h reduction, time and stop flag (it stops after the step that passes its tmax).  Box
 * storage comes from one arena of arena_bytes allocated at creation, so adding and
 * removing boxes allocates nothing on the device; MHD_ERR_NOMEM means "does not fit now". */
typedef struct mhd_batch mhd_batch;
int mhd_batch_create(const mhd_config* base, int max_boxes, uint64_t arena_bytes, mhd_batch** out);
int mhd_batch_destroy(mhd_batch* batch);
/* arena bytes one box of this size takes (the admission test) */
int mhd_batch_box_bytes(const mhd_batch* batch, int nx, int ny, int nz, uint64_t* bytes);
int mhd_batch_free_bytes(const mhd_batch* batch, uint64_t* largest_free, uint64_t* total_free);
/* tile shape for boxes of size nx*ny*nz added later (sw: 16/32/64 x-strip, ly/lz: march
 * chunks; limits as mhd_set_tiling) */
int mhd_batch_set_tiling(mhd_batch* batch, int nx, int ny, int nz, int sw, int ly, int lz);
/* grid (GenerateGrid) and, if ic != NULL, the shear-layer state on the device; tmax <= 0: no stop */
int mhd_batch_add_box(mhd_batch* batch, int nx, int ny, int nz, const mhd_ic_params* ic, double tmax, int* box_id);
/* replace a box's state by host arrays (reference layout) and restart its clock at 0 */
int mhd_batch_upload_state(mhd_batch* batch, int box_id, const double* P, const double* U);
int mhd_batch_remove_box(mhd_batch* batch, int box_id);
/* nsteps loop iterations of every live box; returns when they are done */
int mhd_batch_step(mhd_batch* batch, int nsteps);
int mhd_batch_last_step_ms(const mhd_batch* batch, double* ms);
/* status: as mhd_status, for this box since it was added or its state uploaded; a box whose
 * state holds a bad cell stops on its own (stopped = 1) while the other boxes go on */
int mhd_batch_box_state(mhd_batch* batch, int box_id, double* time, int* steps_done, int* stopped, int* status);
/* ids of live boxes that have stopped (n may exceed cap; ids holds the first cap) */
int mhd_batch_poll(mhd_batch* batch, int* ids, int cap, int* n);
/* primitive recovery of the box's current state; diag4 (may be NULL) as mhd_diagnostics,
 * P_host (may be NULL) the primitive array in the reference layout */
int mhd_batch_finish_box(mhd_batch* batch, int box_id, double* diag4, double* P_host);

/* tile shape of a single box's fused kernels: sw 16/32/64 (x-strip width), 1 <= ly <= ny
 * (y chunk; also at most what the xy kernel's LDS for strip-edge faces allows per block:
 * FP64 46, 23, 11 rows (160 KiB less its window rings) and mixed 64, 32, 16 rows (64 KiB)
 * for sw 64, 32, 16; 64 rows where the FP64 march runs two wavefronts per SIMD with its
 * strip edges shared in the block, mhd_set_xy_waves_per_simd), 1 <= lz <= min(nz, 64)
 * (z chunk) */
int mhd_set_tiling(mhd_ctx* ctx, int sw, int ly, int lz);
int mhd_get_tiling(mhd_ctx* ctx, int* sw, int* ly, int* lz);

/* FP64 xy kernel of a single box's fused step (no counterpart in the reference: a tuning
 * switch beside mhd_set_tiling).  n = 2: up to two wavefronts per SIMD, the strip edges
 * shared within each block and no strip-edge pre-pass, which needs sw 64, x periodic and
 * nx = 64, 128, 192, 256 or 512 (for nx 64 / 128, nz a multiple of 4 / 2; else
 * MHD_ERR_ARG); n = 1: one wavefront per SIMD with the strip-edge pre-pass, which needs ly
 * within that kernel's LDS limit (mhd_set_tiling); n = 0: the default rule (n = 2 where the
 * box allows it and that launch has more wavefronts than the GPU has SIMDs).  The results
 * agree to round-off (the strip-edge faces are computed by different code).  Mixed precision
 * ignores it.  mhd_get_xy_waves_per_simd gives the kernel the next step takes. */
int mhd_set_xy_waves_per_simd(mhd_ctx* ctx, int n);
int mhd_get_xy_waves_per_simd(mhd_ctx* ctx, int* n);

#ifdef __cplusplus
}
#endif
#endif
