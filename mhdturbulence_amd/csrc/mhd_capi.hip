// mhd_capi.hip -- the C ABI (include/mhd_amd.h): box contexts, HBM allocation, host
// layout conversion, the reference-shaped step functions and the device-resident loop.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/mhd_amd.h"
#include "mhd_device.hpp"
#include "mhd_fused.hpp"

using mhd::Geom;
using mhd::StepScalars;

namespace mhd_capi_detail {

thread_local std::string g_last_error;

// records msg as this thread's last error and returns code (shared by every C-ABI source)
int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

}  // namespace mhd_capi_detail

namespace {

using mhd_capi_detail::fail;
using mhd_capi_detail::g_last_error;

#define HIPCHK(expr)                                                                                 \
  do {                                                                                               \
    hipError_t e_ = (expr);                                                                          \
    if (e_ != hipSuccess)                                                                            \
      return fail(MHD_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));                   \
  } while (0)

#define CHKRET(expr)            \
  do {                          \
    int r_ = (expr);            \
    if (r_ != MHD_OK) return r_; \
  } while (0)

}  // namespace

struct mhd_ctx {
  mhd_config cfg{};
  Geom g{};
  int dev = 0;
  hipStream_t stream = nullptr;
  double* P = nullptr;
  double* U = nullptr;
  double* F[3] = {nullptr, nullptr, nullptr};
  double* U2 = nullptr;  // second conserved buffer: fused loop ping-pong, or U^n of a staged two-stage step
  double* grid[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};  // x1a x1b x2a x2b x3a x3b
  double* staging = nullptr;
  double* diag = nullptr;
  double* dt_log = nullptr;
  int dt_log_cap = 0;
  StepScalars* sc = nullptr;
  bool grid_set = false;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double last_ms = 0.0;
  int last_launches = 0;
  uint64_t bytes = 0;
  // captured one-step graph of the device loop
  hipGraphExec_t graph = nullptr;
  int graph_mode = -1;
  double* graph_log = nullptr;
  // fused loop buffers: second conserved copy (U2 above), flux divergence
  double* acc = nullptr;
  mhd::FusedArgs fa{};
  bool fa_ready = false;
  // asynchronous snapshot: P in host layout on the device, its pinned host copy, the copy
  // stream and the events ordering it after the step stream
  hipStream_t copy_stream = nullptr;
  double* snap_dev = nullptr;
  double* snap_host = nullptr;
  hipEvent_t snap_ready = nullptr, snap_done = nullptr;
  bool snap_recorded = false, snap_pending = false;
  // the fused loop left live, by mhd_fused_begin_step / _step (the reference loop driven one
  // step at a time) or by a fused mhd_run / mhd_profile_kernels: the state is in U or U2 by
  // the parity of the iteration count, P is out of date and the ring holds the next step's
  // dt / c_h.  The next fused run continues from there; settle() brings U and P back first
  // for every other entry point.
  bool fz_live = false;
};

namespace {

size_t host_cells(const mhd_ctx* c) { return (size_t)c->g.itot * c->g.jtot * c->g.ktot; }

int dalloc(mhd_ctx* c, double** p, size_t n) {
  if (*p) return MHD_OK;
  hipError_t e = hipMalloc((void**)p, n * sizeof(double));
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(MHD_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  }
  c->bytes += n * sizeof(double);
  HIPCHK(hipMemsetAsync(*p, 0, n * sizeof(double), c->stream));
  return MHD_OK;
}

int ensure_flux(mhd_ctx* c) {
  for (int d = 0; d < 3; ++d) CHKRET(dalloc(c, &c->F[d], mhd::NCONS * c->g.vstride));
  return MHD_OK;
}

// the fused kernels' arguments for this box over the given second state copy and divergence
// buffer; allocates nothing.  The tiling is the one mhd_set_tiling chose, if any.
int fused_args(const mhd_ctx* c, double* U2, double* acc, mhd::FusedArgs& a) {
  a = mhd::FusedArgs{};
  a.g = c->g;
  a.mixed = c->cfg.precision != 0;
  a.acc32 = c->cfg.precision == 2;
  for (int q = 0; q < 6; ++q) a.bc[q] = c->cfg.bc[q];
  if (!mhd::fused_configure(a))
    return fail(MHD_ERR_UNSUPPORTED, "fused step: a variable column of this box spans 4 GiB or more "
                                     "(more than 2^29 cells); use the staged step (mode 1)");
  if (c->fa_ready) {
    a.sw = c->fa.sw;
    a.ns = c->fa.ns;
    a.ly = c->fa.ly;
    a.lz = c->fa.lz;
  }
  a.x1a = c->grid[0];
  a.x2a = c->grid[2];
  a.x3a = c->grid[4];
  a.U0 = c->U;
  a.U1 = U2;
  a.acc = acc;
  a.sc = c->sc;
  a.gam = c->cfg.gam;
  a.cfl = c->cfg.cfl;
  a.alphabp = c->cfg.alphabp;
  for (int q = 0; q < 6; ++q) a.bc[q] = c->cfg.bc[q];
  a.eflux = c->cfg.energy_flux;
  return MHD_OK;
}

// doubles of the divergence buffer (FP32 at precision 2: half the bytes)
size_t acc_doubles(const mhd_ctx* c) {
  const size_t n = mhd::NCONS * c->g.vstride;
  return c->cfg.precision == 2 ? (n + 1) / 2 : n;
}

int ensure_fused(mhd_ctx* c) {
  if (c->fa_ready) return MHD_OK;
  mhd::FusedArgs a;
  CHKRET(fused_args(c, nullptr, nullptr, a));
  CHKRET(dalloc(c, &c->U2, mhd::NCONS * c->g.vstride));
  CHKRET(dalloc(c, &c->acc, acc_doubles(c)));
  a.U1 = c->U2;
  a.acc = c->acc;
  c->fa = a;
  c->fa_ready = true;
  return MHD_OK;
}

int ensure_staging(mhd_ctx* c) { return dalloc(c, &c->staging, mhd::NPRIM * host_cells(c)); }

int select_dev(mhd_ctx* c) {
  HIPCHK(hipSetDevice(c->dev));
  return MHD_OK;
}

int read_scalars(mhd_ctx* c, StepScalars* h);

// Ends a live stretch of the fused loop: the state is put back in U (from U2 after an odd
// number of one-stage iterations, interior cells only) and P recovered from it.
int settle(mhd_ctx* c) {
  if (!c->fz_live) return MHD_OK;
  c->fz_live = false;
  StepScalars h;
  CHKRET(read_scalars(c, &h));
  if (c->cfg.stages != 2 && (h.iter & 1)) mhd::launch_copy_interior(c->g, c->U2, c->U, mhd::NCONS, c->stream);
  mhd::launch_update_prim(c->g, c->cfg.gam, c->U, c->P, nullptr, c->stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  return MHD_OK;
}

// the context's device made current, and P / U up to date (every entry point but the fused
// runs, the one-step fused pair and the status read goes through here)
int set_dev(mhd_ctx* c) {
  CHKRET(select_dev(c));
  return settle(c);
}

void drop_graph(mhd_ctx* c) {
  if (c->graph) (void)hipGraphExecDestroy(c->graph);
  c->graph = nullptr;
  c->graph_mode = -1;
}

int check_cfg(const mhd_config* cfg) {
  if (!cfg) return fail(MHD_ERR_ARG, "null config");
  if (cfg->nx < 4 || cfg->ny < 4 || cfg->nz < 4) return fail(MHD_ERR_ARG, "each axis needs at least 4 cells");
  if ((long long)(cfg->nx + 5) * (cfg->ny + 5) * (cfg->nz + 5) > (1LL << 30))
    return fail(MHD_ERR_ARG, "box too large for 32-bit cell indices");
  for (int q = 0; q < 6; ++q)
    if (cfg->bc[q] < 1 || cfg->bc[q] > 3) return fail(MHD_ERR_ARG, "boundary codes are 1, 2 or 3");
  for (int a = 0; a < 3; ++a)
    if (!(cfg->xmax[a] > cfg->xmin[a])) return fail(MHD_ERR_ARG, "empty domain");
  if (!(cfg->gam > 1.0)) return fail(MHD_ERR_ARG, "gam must exceed 1");
  if (!(cfg->cfl > 0.0)) return fail(MHD_ERR_ARG, "cfl must be positive");
  if (cfg->energy_flux < 0 || cfg->energy_flux > 1) return fail(MHD_ERR_ARG, "energy_flux is 0 or 1");
  if (cfg->precision < 0 || cfg->precision > 2) return fail(MHD_ERR_ARG, "precision is 0, 1 or 2");
  if (cfg->stages < 0 || cfg->stages > 2) return fail(MHD_ERR_ARG, "stages is 0 or 1 (forward Euler) or 2");
  return MHD_OK;
}

bool two_stage(const mhd_ctx* c) { return c->cfg.stages == 2; }

int read_scalars(mhd_ctx* c, StepScalars* h) {
  HIPCHK(hipMemcpyAsync(h, c->sc, sizeof(StepScalars), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return MHD_OK;
}

double bits2d(unsigned long long b) {
  double d;
  std::memcpy(&d, &b, sizeof d);
  return d;
}

}  // namespace

extern "C" {

int mhd_version(void) { return MHD_AMD_VERSION; }

const char* mhd_error_string(int code) {
  switch (code) {
    case MHD_OK: return "ok";
    case MHD_ERR_ARG: return "invalid argument";
    case MHD_ERR_HIP: return "HIP runtime error";
    case MHD_ERR_NOMEM: return "device allocation failed";
    case MHD_ERR_UNSUPPORTED: return "unsupported";
    case MHD_ERR_STATE: return "call out of order";
    default: return "unknown error";
  }
}

const char* mhd_last_error(void) { return g_last_error.c_str(); }

int mhd_destroy(mhd_ctx* c) {
  if (!c) return MHD_OK;
  // teardown: errors are not actionable here, every release is attempted
  (void)hipSetDevice(c->dev);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->copy_stream) (void)hipStreamSynchronize(c->copy_stream);
  drop_graph(c);
  double* bufs[] = {c->P, c->U, c->F[0], c->F[1], c->F[2], c->U2, c->acc, c->staging, c->diag, c->dt_log,
                    c->grid[0], c->grid[1], c->grid[2], c->grid[3], c->grid[4], c->grid[5], c->snap_dev};
  for (double* b : bufs)
    if (b) (void)hipFree(b);
  if (c->sc) (void)hipFree(c->sc);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->snap_host) (void)hipHostFree(c->snap_host);
  if (c->snap_ready) (void)hipEventDestroy(c->snap_ready);
  if (c->snap_done) (void)hipEventDestroy(c->snap_done);
  if (c->copy_stream) (void)hipStreamDestroy(c->copy_stream);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return MHD_OK;
}

int mhd_create(const mhd_config* cfg, mhd_ctx** out) {
  if (!out) return fail(MHD_ERR_ARG, "null output pointer");
  *out = nullptr;
  CHKRET(check_cfg(cfg));
  int ndev = 0;
  HIPCHK(hipGetDeviceCount(&ndev));
  if (cfg->device < 0 || cfg->device >= ndev) return fail(MHD_ERR_ARG, "no such HIP device");
  mhd_ctx* c = new (std::nothrow) mhd_ctx();
  if (!c) return fail(MHD_ERR_NOMEM, "host allocation failed");
  c->cfg = *cfg;
  c->dev = cfg->device;
  Geom& g = c->g;
  g.nx = cfg->nx; g.ny = cfg->ny; g.nz = cfg->nz;
  g.itot = cfg->nx + 5; g.jtot = cfg->ny + 5; g.ktot = cfg->nz + 5;
  g.pitch = ((g.itot + mhd::IOFF) + 15) / 16 * 16;
  g.vstride = (size_t)g.ktot * g.jtot * g.pitch;
  int rc = MHD_OK;
  auto bail = [&](int r) { mhd_destroy(c); return r; };
  if ((rc = set_dev(c)) != MHD_OK) return bail(rc);
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
    return bail(fail(MHD_ERR_HIP, "hipStreamCreate failed"));
  if (hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess)
    return bail(fail(MHD_ERR_HIP, "hipEventCreate failed"));
  if ((rc = dalloc(c, &c->P, mhd::NPRIM * g.vstride)) != MHD_OK) return bail(rc);
  if ((rc = dalloc(c, &c->U, mhd::NCONS * g.vstride)) != MHD_OK) return bail(rc);
  const int tot[3] = {g.itot, g.jtot, g.ktot};
  for (int q = 0; q < 6; ++q)
    if ((rc = dalloc(c, &c->grid[q], tot[q / 2])) != MHD_OK) return bail(rc);
  if ((rc = dalloc(c, &c->diag, 4 * (size_t)g.nz)) != MHD_OK) return bail(rc);
  if (hipMalloc((void**)&c->sc, sizeof(StepScalars)) != hipSuccess) return bail(fail(MHD_ERR_NOMEM, "hipMalloc"));
  if (hipMemsetAsync(c->sc, 0, sizeof(StepScalars), c->stream) != hipSuccess)
    return bail(fail(MHD_ERR_HIP, "hipMemsetAsync"));
  if (hipStreamSynchronize(c->stream) != hipSuccess) return bail(fail(MHD_ERR_HIP, "hipStreamSynchronize"));
  *out = c;
  return MHD_OK;
}

int mhd_get_dims(const mhd_ctx* c, int* itot, int* jtot, int* ktot) {
  if (!c || !itot || !jtot || !ktot) return fail(MHD_ERR_ARG, "null argument");
  *itot = c->g.itot; *jtot = c->g.jtot; *ktot = c->g.ktot;
  return MHD_OK;
}

int mhd_device_bytes(const mhd_ctx* c, uint64_t* bytes) {
  if (!c || !bytes) return fail(MHD_ERR_ARG, "null argument");
  *bytes = c->bytes + sizeof(StepScalars);
  return MHD_OK;
}

int mhd_set_grid(mhd_ctx* c, const double* x1a, const double* x1b, const double* x2a, const double* x2b,
                 const double* x3a, const double* x3b) {
  if (!c || !x1a || !x1b || !x2a || !x2b || !x3a || !x3b) return fail(MHD_ERR_ARG, "null argument");
  CHKRET(set_dev(c));
  const double* src[6] = {x1a, x1b, x2a, x2b, x3a, x3b};
  const int tot[3] = {c->g.itot, c->g.jtot, c->g.ktot};
  for (int q = 0; q < 6; ++q)
    HIPCHK(hipMemcpyAsync(c->grid[q], src[q], tot[q / 2] * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  c->grid_set = true;
  return MHD_OK;
}

int mhd_generate_grid(mhd_ctx* c) {
  if (!c) return fail(MHD_ERR_ARG, "null context");
  CHKRET(set_dev(c));
  mhd::launch_generate_grid(c->g, c->cfg.xmin, c->cfg.xmax, c->cfg.grid_shift, c->grid[0], c->grid[1], c->grid[2],
                            c->grid[3], c->grid[4], c->grid[5], c->stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  c->grid_set = true;
  return MHD_OK;
}

int mhd_get_grid(mhd_ctx* c, double* x1a, double* x1b, double* x2a, double* x2b, double* x3a, double* x3b) {
  if (!c || !x1a || !x1b || !x2a || !x2b || !x3a || !x3b) return fail(MHD_ERR_ARG, "null argument");
  CHKRET(set_dev(c));
  double* dst[6] = {x1a, x1b, x2a, x2b, x3a, x3b};
  const int tot[3] = {c->g.itot, c->g.jtot, c->g.ktot};
  for (int q = 0; q < 6; ++q)
    HIPCHK(hipMemcpyAsync(dst[q], c->grid[q], tot[q / 2] * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return MHD_OK;
}

static int upload(mhd_ctx* c, int nv, const double* host, double* dev) {
  if (!c || !host) return fail(MHD_ERR_ARG, "null argument");
  CHKRET(set_dev(c));
  CHKRET(ensure_staging(c));
  HIPCHK(hipMemcpyAsync(c->staging, host, nv * host_cells(c) * sizeof(double), hipMemcpyHostToDevice, c->stream));
  mhd::launch_from_host_layout(c->g, nv, c->staging, dev, c->stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  return MHD_OK;
}

static int download(mhd_ctx* c, int nv, const double* dev, double* host) {
  if (!c || !host) return fail(MHD_ERR_ARG, "null argument");
  if (!dev) return fail(MHD_ERR_STATE, "array not computed yet");
  CHKRET(set_dev(c));
  CHKRET(ensure_staging(c));
  mhd::launch_to_host_layout(c->g, nv, dev, c->staging, c->stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(host, c->staging, nv * host_cells(c) * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return MHD_OK;
}

int mhd_upload_prim(mhd_ctx* c, const double* P) { return upload(c, mhd::NPRIM, P, c ? c->P : nullptr); }
int mhd_upload_cons(mhd_ctx* c, const double* U) { return upload(c, mhd::NCONS, U, c ? c->U : nullptr); }
int mhd_download_prim(mhd_ctx* c, double* P) { return download(c, mhd::NPRIM, c ? c->P : nullptr, P); }
int mhd_download_cons(mhd_ctx* c, double* U) { return download(c, mhd::NCONS, c ? c->U : nullptr, U); }
int mhd_download_flux(mhd_ctx* c, int dir, double* F) {
  if (!c || dir < 1 || dir > 3) return fail(MHD_ERR_ARG, "dir is 1, 2 or 3");
  return download(c, mhd::NCONS, c->F[dir - 1], F);
}

int mhd_snapshot_begin(mhd_ctx* c) {
  if (!c) return fail(MHD_ERR_ARG, "null context");
  CHKRET(set_dev(c));
  const size_t n = mhd::NPRIM * host_cells(c);
  if (!c->copy_stream) {
    // default priority, as the step stream
    HIPCHK(hipStreamCreateWithFlags(&c->copy_stream, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&c->snap_ready, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&c->snap_done, hipEventDisableTiming));
  }
  CHKRET(dalloc(c, &c->snap_dev, n));
  if (!c->snap_host) {
    hipError_t e = hipHostMalloc((void**)&c->snap_host, n * sizeof(double), hipHostMallocDefault);
    if (e != hipSuccess) {
      c->snap_host = nullptr;
      return fail(MHD_ERR_NOMEM, std::string("hipHostMalloc: ") + hipGetErrorString(e));
    }
  }
  // the previous snapshot's transfer still reads snap_dev and writes snap_host
  if (c->snap_recorded) HIPCHK(hipStreamWaitEvent(c->stream, c->snap_done, 0));
  mhd::launch_to_host_layout(c->g, mhd::NPRIM, c->P, c->snap_dev, c->stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(c->snap_ready, c->stream));
  HIPCHK(hipStreamWaitEvent(c->copy_stream, c->snap_ready, 0));
  HIPCHK(hipMemcpyAsync(c->snap_host, c->snap_dev, n * sizeof(double), hipMemcpyDeviceToHost, c->copy_stream));
  HIPCHK(hipEventRecord(c->snap_done, c->copy_stream));
  c->snap_recorded = true;
  c->snap_pending = true;
  return MHD_OK;
}

int mhd_snapshot_end(mhd_ctx* c, const double** P_host) {
  if (!c || !P_host) return fail(MHD_ERR_ARG, "null argument");
  if (!c->snap_pending) return fail(MHD_ERR_STATE, "no snapshot begun (mhd_snapshot_begin first)");
  CHKRET(set_dev(c));
  HIPCHK(hipEventSynchronize(c->snap_done));
  c->snap_pending = false;
  *P_host = c->snap_host;
  return MHD_OK;
}

static int need_grid(mhd_ctx* c) {
  if (!c) return fail(MHD_ERR_ARG, "null context");
  if (!c->grid_set) return fail(MHD_ERR_STATE, "grid not set (mhd_set_grid or mhd_generate_grid first)");
  return set_dev(c);
}

static int host_reduce(mhd_ctx* c, double* dtmin, double* ch) {
  mhd::launch_reset_scalars(c->sc, 0, c->stream);
  mhd::launch_reduce_prim(c->g, c->P, c->grid[0], c->grid[2], c->grid[4], c->sc, 0, c->stream);
  HIPCHK(hipGetLastError());
  StepScalars h;
  CHKRET(read_scalars(c, &h));
  if (dtmin) *dtmin = bits2d(h.host_acc[0]);
  if (ch) *ch = bits2d(h.host_acc[1]);
  return MHD_OK;
}

int mhd_control_timestep(mhd_ctx* c, double* dt) {
  CHKRET(need_grid(c));
  if (!dt) return fail(MHD_ERR_ARG, "null dt");
  double dtmin;
  CHKRET(host_reduce(c, &dtmin, nullptr));
  *dt = c->cfg.cfl * dtmin;
  return MHD_OK;
}

int mhd_evaluate_ch(mhd_ctx* c, double* chg) {
  CHKRET(need_grid(c));
  if (!chg) return fail(MHD_ERR_ARG, "null chg");
  return host_reduce(c, nullptr, chg);
}

int mhd_set_boundary(mhd_ctx* c) {
  if (!c) return fail(MHD_ERR_ARG, "null context");
  CHKRET(set_dev(c));
  mhd::launch_boundary(c->g, c->cfg.bc, c->P, mhd::NPRIM, c->stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  return MHD_OK;
}

int mhd_numerical_flux(mhd_ctx* c, int dir, double chg) {
  if (!c || dir < 1 || dir > 3) return fail(MHD_ERR_ARG, "dir is 1, 2 or 3");
  CHKRET(set_dev(c));
  CHKRET(ensure_flux(c));
  mhd::launch_flux(c->g, dir, c->cfg.energy_flux, c->cfg.precision, c->P, c->F[dir - 1], chg, nullptr, c->stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  return MHD_OK;
}

int mhd_update_conserv(mhd_ctx* c, double dt) {
  CHKRET(need_grid(c));
  if (!c->F[0] || !c->F[1] || !c->F[2]) return fail(MHD_ERR_STATE, "fluxes not computed");
  mhd::launch_update(c->g, c->grid[0], c->grid[2], c->grid[4], c->F[0], c->F[1], c->F[2], c->U, dt, c->cfg.cfl,
                     c->cfg.alphabp, nullptr, c->stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  return MHD_OK;
}

int mhd_damp_psi(mhd_ctx* c, double dt, double chg) {
  CHKRET(need_grid(c));
  mhd::launch_damp(c->g, c->grid[0], c->grid[2], c->grid[4], c->U, dt, chg, c->cfg.alphabp, c->stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  return MHD_OK;
}

int mhd_update_primitive(mhd_ctx* c) {
  if (!c) return fail(MHD_ERR_ARG, "null context");
  CHKRET(set_dev(c));
  mhd::launch_update_prim(c->g, c->cfg.gam, c->U, c->P, nullptr, c->stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  return MHD_OK;
}

// Launches of one loop iteration; every argument is iteration-invariant (the kernels read
// the iteration's slots from c->sc), so one iteration is captured once and replayed.
// marks (may be null) receives an event after each launch, for per-kernel timing.
// stages = 2: the two-stage Runge-Kutta step, U* = E(U^n), U^{n+1} = (U^n + E(U*)) / 2 with
// E the reference's forward-Euler stage and the dt and c_h of U^n in both stages.
// fa (mode 0) and dt_log: the fused kernels' arguments and the device dt log of the run.
static int issue_step(mhd_ctx* c, int mode, hipEvent_t* marks, const mhd::FusedArgs& fa, double* dt_log) {
  const Geom& g = c->g;
  const bool rk2 = two_stage(c);
  int q = 0;
  auto mark = [&]() {
    if (marks) (void)hipEventRecord(marks[q], c->stream);
    ++q;
  };
  if (mode == 0) {
    mhd::FusedArgs a = fa;
    a.stage = rk2 ? 1 : 0;
    for (int st = 0; st < (rk2 ? 2 : 1); ++st) {
      if (st == 1) a.stage = 2;
      mhd::launch_fused_xy(a, c->stream);
      mark();
      mhd::launch_fused_z(a, c->stream);
      mark();
    }
    mhd::launch_step_end(c->sc, c->cfg.cfl, dt_log, 1, c->stream);
    mark();
    HIPCHK(hipGetLastError());
    return MHD_OK;
  }
  mhd::launch_reduce_prim(g, c->P, c->grid[0], c->grid[2], c->grid[4], c->sc, 1, c->stream);
  mark();
  for (int st = 0; st < (rk2 ? 2 : 1); ++st) {
    mhd::launch_boundary(g, c->cfg.bc, c->P, mhd::NPRIM, c->stream);
    mark();
    for (int d = 1; d <= 3; ++d) {
      mhd::launch_flux(g, d, c->cfg.energy_flux, c->cfg.precision, c->P, c->F[d - 1], 0.0, c->sc, c->stream);
      mark();
    }
    mhd::launch_update(g, c->grid[0], c->grid[2], c->grid[4], c->F[0], c->F[1], c->F[2], c->U, 0.0, c->cfg.cfl,
                       c->cfg.alphabp, c->sc, c->stream, rk2 && st == 0 ? c->U2 : nullptr,
                       rk2 && st == 1 ? c->U2 : nullptr);
    mark();
    mhd::launch_update_prim(g, c->cfg.gam, c->U, c->P, c->sc, c->stream);
    mark();
  }
  mhd::launch_step_end(c->sc, c->cfg.cfl, dt_log, 0, c->stream);
  mark();
  HIPCHK(hipGetLastError());
  return MHD_OK;
}

static const char* kernel_names(int mode) {
  if (mode == 0) return "fused_xy,fused_z_update,step_end";
  return "reduce_dt_ch,boundary,flux_x,flux_y,flux_z,update_damp,update_prim,step_end";
}

static int kernel_kinds(int mode) { return mode == 0 ? 3 : 8; }

static int launches_per_step(const mhd_ctx* c, int mode) {
  const int n = kernel_kinds(mode);
  return two_stage(c) ? (mode == 0 ? n + 2 : n + 6) : n;
}

// the entry of kernel_names(mode) that launch q of an iteration times (both Runge-Kutta
// stages of a kernel add up in one entry)
static int launch_kind(const mhd_ctx* c, int mode, int q) {
  if (!two_stage(c)) return q;
  if (mode == 0) return q < 4 ? q % 2 : 2;
  return q == 0 ? 0 : q == 13 ? 7 : 1 + (q - 1) % 6;
}

// the loop scalars of a run starting at `time` with a device dt log of dt_cap entries, and
// (mode 0) the first iteration's dt / c_h from P; allocates nothing
static int start_loop(mhd_ctx* c, int mode, double tmax, double time, int dt_cap) {
  double tt[2] = {time, tmax};
  HIPCHK(hipMemcpyAsync(&c->sc->time, tt, sizeof tt, hipMemcpyHostToDevice, c->stream));
  int ints[3] = {0, 0, 0};  // iter, steps_done, stopped
  HIPCHK(hipMemcpyAsync(&c->sc->iter, ints, sizeof ints, hipMemcpyHostToDevice, c->stream));
  const int cap_halt[2] = {dt_cap, 0};  // dt_log_cap, halt
  HIPCHK(hipMemcpyAsync(&c->sc->dt_log_cap, cap_halt, sizeof cap_halt, hipMemcpyHostToDevice, c->stream));
  mhd::launch_reset_scalars(c->sc, 1, c->stream);
  if (mode == 0) {
    // the first iteration's dt and c_h from P (ControlTimestep / EvaluateCh on the state as
    // given); the fused kernels read no ghost cell of U (they take boundary images from the
    // interior cells)
    mhd::launch_reduce_prim(c->g, c->P, c->grid[0], c->grid[2], c->grid[4], c->sc, 1, c->stream);
  }
  HIPCHK(hipGetLastError());
  return MHD_OK;
}

// the loop scalars of a fused run that continues where the previous one left the state
// live (fz_live): the caller's time and tmax and this run's counters; the iteration count,
// whose parity says which buffer holds the state, and the ring slot the last step's
// reduction filled (this run's first dt / c_h) are kept
static int continue_loop(mhd_ctx* c, double tmax, double time, int dt_cap) {
  double tt[2] = {time, tmax};
  HIPCHK(hipMemcpyAsync(&c->sc->time, tt, sizeof tt, hipMemcpyHostToDevice, c->stream));
  int ints[2] = {0, 0};  // steps_done, stopped
  HIPCHK(hipMemcpyAsync(&c->sc->steps_done, ints, sizeof ints, hipMemcpyHostToDevice, c->stream));
  const int cap_halt[2] = {dt_cap, 0};  // dt_log_cap, halt
  HIPCHK(hipMemcpyAsync(&c->sc->dt_log_cap, cap_halt, sizeof cap_halt, hipMemcpyHostToDevice, c->stream));
  return MHD_OK;
}

// loop state and buffers for nsteps iterations starting at *time.  A fused run continues a
// fused run whose state is still live; anything else settles the state first.
static int prepare_run(mhd_ctx* c, int nsteps, double tmax, int mode, double time) {
  if (!c) return fail(MHD_ERR_ARG, "null context");
  if (!c->grid_set) return fail(MHD_ERR_STATE, "grid not set (mhd_set_grid or mhd_generate_grid first)");
  if (nsteps < 0) return fail(MHD_ERR_ARG, "nsteps must be >= 0");
  if (mode != 0 && mode != 1) return fail(MHD_ERR_ARG, "mode is 0 (fused) or 1 (staged)");
  const bool cont = mode == 0 && c->fz_live;
  CHKRET(cont ? select_dev(c) : set_dev(c));
  // device dt log, grown as needed (a captured graph holds its address)
  if (nsteps > c->dt_log_cap) {
    drop_graph(c);
    if (c->dt_log) { (void)hipFree(c->dt_log); c->bytes -= c->dt_log_cap * sizeof(double); c->dt_log = nullptr; }
    CHKRET(dalloc(c, &c->dt_log, std::max(nsteps, 1)));
    c->dt_log_cap = std::max(nsteps, 1);
  }
  if (mode == 0) CHKRET(ensure_fused(c));
  else CHKRET(ensure_flux(c));
  if (mode == 1 && two_stage(c)) CHKRET(dalloc(c, &c->U2, mhd::NCONS * c->g.vstride));
  return cont ? continue_loop(c, tmax, time, c->dt_log_cap) : start_loop(c, mode, tmax, time, c->dt_log_cap);
}

// U2: the fused loop's second state copy; dev_log: the run's device dt log.  Waits for the
// stream; allocates nothing.  keep_live (fused runs of the context's own buffers): the state
// stays where the last step left it, U or (after an odd number of one-stage iterations) U2,
// with P out of date and the next step's dt / c_h in the ring, so a following fused run
// continues without recomputing them; any other entry point settles it first (settle()).
// A run that stopped (past tmax, or at a bad cell) settles at once: its stop flag would
// otherwise hold the one-step fused pair, which continues a live state as it finds it.
static int finish_loop(mhd_ctx* c, int mode, const double* U2, const double* dev_log, int nsteps, double* time,
                       double* dt_log, int* steps_done, bool keep_live) {
  StepScalars h;
  CHKRET(read_scalars(c, &h));
  if (mode == 0) {
    if (keep_live && !h.halt && !h.stopped) {
      c->fz_live = true;
    } else {
      // one-stage iterations leave the state in U2 after an odd number of them; its interior
      // goes back to U (whose ghost cells the loop leaves as they were, as the reference's
      // loop does) and P is recovered from it (two-stage iterations end in U)
      if (!two_stage(c) && (h.iter & 1)) mhd::launch_copy_interior(c->g, U2, c->U, mhd::NCONS, c->stream);
      mhd::launch_update_prim(c->g, c->cfg.gam, c->U, c->P, nullptr, c->stream);
      HIPCHK(hipGetLastError());
      HIPCHK(hipStreamSynchronize(c->stream));
      c->fz_live = false;
    }
  }
  if (dt_log && h.steps_done > 0)
    HIPCHK(hipMemcpy(dt_log, dev_log, sizeof(double) * std::min(h.steps_done, nsteps), hipMemcpyDeviceToHost));
  if (time) *time = h.time;
  if (steps_done) *steps_done = h.steps_done;
  return MHD_OK;
}

static int finish_run(mhd_ctx* c, int mode, int nsteps, double* time, double* dt_log, int* steps_done) {
  return finish_loop(c, mode, c->U2, c->dt_log, nsteps, time, dt_log, steps_done, true);
}

int mhd_run(mhd_ctx* c, int nsteps, double tmax, int mode, double* time, double* dt_log, int* steps_done) {
  if (!time) return fail(MHD_ERR_ARG, "a time pointer is required");
  CHKRET(prepare_run(c, nsteps, tmax, mode, *time));
  // capture one iteration once, then replay it
  if (!c->graph || c->graph_mode != mode || c->graph_log != c->dt_log) {
    drop_graph(c);
    HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int rc = issue_step(c, mode, nullptr, c->fa, c->dt_log);
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(c->stream, &graph);
    if (rc != MHD_OK) { if (graph) (void)hipGraphDestroy(graph); return rc; }
    if (e != hipSuccess) return fail(MHD_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    e = hipGraphInstantiate(&c->graph, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) {
      c->graph = nullptr;
      return fail(MHD_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    }
    c->graph_mode = mode;
    c->graph_log = c->dt_log;
  }
  HIPCHK(hipEventRecord(c->ev0, c->stream));
  for (int s = 0; s < nsteps; ++s) HIPCHK(hipGraphLaunch(c->graph, c->stream));
  HIPCHK(hipEventRecord(c->ev1, c->stream));
  CHKRET(finish_run(c, mode, nsteps, time, dt_log, steps_done));
  float ms = 0.0f;
  HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
  c->last_ms = ms;
  c->last_launches = launches_per_step(c, mode) * nsteps;
  return MHD_OK;
}

// ---- the fused loop body one reference step at a time ----
int mhd_fused_begin_step(mhd_ctx* c, double* dt, double* chg, int* status) {
  if (!c || !dt || !chg) return fail(MHD_ERR_ARG, "null argument");
  if (!c->grid_set) return fail(MHD_ERR_STATE, "grid not set (mhd_set_grid or mhd_generate_grid first)");
  CHKRET(select_dev(c));
  if (!c->fz_live) {
    // a new stretch: the first step's bounds from P, iteration 0
    CHKRET(ensure_fused(c));
    CHKRET(start_loop(c, 0, 0.0, 0.0, 0));
    c->fz_live = true;
  }
  StepScalars h;
  CHKRET(read_scalars(c, &h));
  *dt = c->cfg.cfl * bits2d(h.red[h.iter % 3][0]);
  *chg = bits2d(h.red[h.iter % 3][1]);
  if (status) *status = h.status;
  return MHD_OK;
}

int mhd_fused_step(mhd_ctx* c) {
  if (!c) return fail(MHD_ERR_ARG, "null context");
  CHKRET(select_dev(c));
  if (!c->fz_live) return fail(MHD_ERR_STATE, "mhd_fused_step without mhd_fused_begin_step");
  return issue_step(c, 0, nullptr, c->fa, nullptr);
}

// ---- the allocation-free run (mhd_run_noalloc) ----
// Its scratch (the fused loop's second state copy, the flux divergence and a dt log chunk)
// lives in this array, one copy per GPU the library runs on.  Host code never names the
// array itself (on the host its name is only a stand-in): it uses the device address that
// hipGetSymbolAddress returns, taken once per device with that device current.
static __device__ __align__(256) char scratch[MHD_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int NOALLOC_LOG = 4096;  // dt log entries per chunk of steps
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[MAX_DEVICES] = {};

// this device's copy of `scratch` (the device is current)
int scratch_base(int dev, char** base) {
  if (dev < 0 || dev >= MAX_DEVICES) return fail(MHD_ERR_UNSUPPORTED, "device ordinal above the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(MHD_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = p;
  }
  *base = static_cast<char*>(g_scratch_ptr[dev]);
  return MHD_OK;
}

size_t up256(size_t b) { return (b + 255) / 256 * 256; }

}  // namespace

int mhd_run_noalloc(mhd_ctx* c, int nsteps, double tmax, double* time, double* dt_log, int* steps_done) {
  if (!time) return fail(MHD_ERR_ARG, "a time pointer is required");
  CHKRET(need_grid(c));
  if (nsteps < 0) return fail(MHD_ERR_ARG, "nsteps must be >= 0");
  const size_t u2_bytes = up256(mhd::NCONS * c->g.vstride * sizeof(double));
  const size_t acc_bytes = up256(acc_doubles(c) * sizeof(double));
  const size_t need = u2_bytes + acc_bytes + NOALLOC_LOG * sizeof(double);
  if (need > (size_t)MHD_NOALLOC_SCRATCH_BYTES)
    return fail(MHD_ERR_ARG, "mhd_run_noalloc: this box needs " + std::to_string(need) + " bytes of scratch, above " +
                                 std::to_string((size_t)MHD_NOALLOC_SCRATCH_BYTES) +
                                 " (MHD_NOALLOC_SCRATCH_BYTES); use mhd_run");
  char* base = nullptr;
  CHKRET(scratch_base(c->dev, &base));
  double* U2 = reinterpret_cast<double*>(base);
  double* acc = reinterpret_cast<double*>(base + u2_bytes);
  double* dev_log = reinterpret_cast<double*>(base + u2_bytes + acc_bytes);
  mhd::FusedArgs a;
  CHKRET(fused_args(c, U2, acc, a));
  // the array keeps the previous call's contents: the second state copy and the divergence
  // start from zero, as a freshly allocated pair does (padding and the never-used last row
  // and plane are read as whole lines and must be finite)
  mhd::launch_zero(U2, u2_bytes / sizeof(double), c->stream);
  mhd::launch_zero(acc, acc_bytes / sizeof(double), c->stream);
  HIPCHK(hipGetLastError());
  double t = *time;
  int done = 0;
  // chunks of at most NOALLOC_LOG steps, each a continuation of the previous one (the next
  // chunk's first dt / c_h from the recovered P, as a run continued across mhd_run calls)
  int left = nsteps;
  do {
    const int n = std::min(left, NOALLOC_LOG);
    CHKRET(start_loop(c, 0, tmax, t, NOALLOC_LOG));
    for (int s = 0; s < n; ++s) CHKRET(issue_step(c, 0, nullptr, a, dev_log));
    int got = 0;
    CHKRET(finish_loop(c, 0, U2, dev_log, n, &t, dt_log ? dt_log + done : nullptr, &got, false));
    done += got;
    left -= n;
    if (got < n) break;  // stopped: past tmax, or a bad cell
  } while (left > 0);
  *time = t;
  if (steps_done) *steps_done = done;
  return MHD_OK;
}

int mhd_profile_kernels(mhd_ctx* c, int nsteps, int mode, double* avg_ms, int cap, int* nkernels) {
  if (!c || !avg_ms || !nkernels || nsteps < 1) return fail(MHD_ERR_ARG, "bad profile arguments");
  const int nk = launches_per_step(c, mode), nkind = kernel_kinds(mode);
  if (cap < nkind) return fail(MHD_ERR_ARG, "avg_ms too small");
  double t = 0.0;
  CHKRET(prepare_run(c, nsteps, 0.0, mode, t));
  // one event before the first launch and one after every launch of every step, all
  // recorded up front: the host waits once, at the end, so the steps run back to back.
  // The events only time: all but the last skip the system-scope release fence, whose L2
  // write-back would otherwise sit between every two launches and slow the next one.
  std::vector<hipEvent_t> ev((size_t)nsteps * nk + 1, nullptr);
  int rc = MHD_OK;
  for (size_t i = 0; i < ev.size(); ++i) {
    const unsigned flags = i + 1 < ev.size() ? hipEventDisableSystemFence : hipEventDefault;
    if (hipEventCreateWithFlags(&ev[i], flags) != hipSuccess) {
      rc = fail(MHD_ERR_HIP, "hipEventCreateWithFlags");
      break;
    }
  }
  if (rc == MHD_OK && hipEventRecord(ev[0], c->stream) != hipSuccess) rc = fail(MHD_ERR_HIP, "hipEventRecord");
  for (int s = 0; s < nsteps && rc == MHD_OK; ++s) rc = issue_step(c, mode, ev.data() + 1 + (size_t)s * nk, c->fa, c->dt_log);
  if (rc == MHD_OK && hipEventSynchronize(ev.back()) != hipSuccess) rc = fail(MHD_ERR_HIP, "hipEventSynchronize");
  std::vector<double> acc(nkind, 0.0);
  if (rc == MHD_OK)
    for (int s = 0; s < nsteps; ++s)
      for (int q = 0; q < nk; ++q) {
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, ev[(size_t)s * nk + q], ev[(size_t)s * nk + q + 1]);
        acc[launch_kind(c, mode, q)] += ms;
      }
  for (auto& e : ev)
    if (e) (void)hipEventDestroy(e);
  if (rc != MHD_OK) return rc;
  CHKRET(finish_run(c, mode, nsteps, &t, nullptr, nullptr));
  for (int q = 0; q < nkind; ++q) avg_ms[q] = acc[q] / nsteps;
  *nkernels = nkind;
  return MHD_OK;
}

const char* mhd_kernel_names(int mode) { return kernel_names(mode); }

int mhd_set_tiling(mhd_ctx* c, int sw, int ly, int lz) {
  if (!c) return fail(MHD_ERR_ARG, "null context");
  if ((sw != 16 && sw != 32 && sw != 64) || ly < 1 || lz < 1 || ly > c->g.ny || lz > c->g.nz ||
      lz > 64)
    return fail(MHD_ERR_ARG, "tiling: sw in {16,32,64}, 1 <= ly <= ny, 1 <= lz <= min(nz, 64)");
  CHKRET(set_dev(c));
  CHKRET(ensure_fused(c));
  mhd::FusedArgs a = c->fa;
  a.sw = sw;
  a.ns = (c->g.nx + sw - 1) / sw;
  if (ly > mhd::fused_ly_cap(a))
    return fail(MHD_ERR_ARG, "tiling: ly above " + std::to_string(mhd::fused_ly_cap(a)) +
                                 " (LDS of the xy kernel) for this strip width and precision");
  a.ly = ly;
  a.lz = lz;
  c->fa = a;
  drop_graph(c);
  return MHD_OK;
}

int mhd_set_xy_waves_per_simd(mhd_ctx* c, int n) {
  if (!c) return fail(MHD_ERR_ARG, "null context");
  if (n < 0 || n > 2) return fail(MHD_ERR_ARG, "xy wavefronts per SIMD: 0 (default rule), 1 or 2");
  CHKRET(set_dev(c));
  CHKRET(ensure_fused(c));
  mhd::FusedArgs a = c->fa;
  if (n == 2) {
    if (!mhd::fused_xy_two_blocks_ok(a))
      return fail(MHD_ERR_ARG, "xy wavefronts per SIMD: two need sw 64, x periodic, nx 64, 128, 192, 256 or 512 "
                               "and, for nx 64 / 128, nz a multiple of 4 / 2");
    a.duo = 1;
  } else if (n == 1) {
    a.duo = 0;
  } else {
    mhd::FusedArgs d = a;
    (void)mhd::fused_configure(d);  // the default rule's choice for this box
    a.duo = d.duo;
  }
  if (a.ly > mhd::fused_ly_cap(a))
    return fail(MHD_ERR_ARG, "xy wavefronts per SIMD: the y chunk " + std::to_string(a.ly) + " is above " +
                                 std::to_string(mhd::fused_ly_cap(a)) +
                                 " (LDS of the one-wavefront kernel); set a shorter tiling first");
  c->fa = a;
  drop_graph(c);
  return MHD_OK;
}

int mhd_get_xy_waves_per_simd(mhd_ctx* c, int* n) {
  if (!c || !n) return fail(MHD_ERR_ARG, "null argument");
  CHKRET(set_dev(c));
  CHKRET(ensure_fused(c));
  *n = (!c->fa.mixed && mhd::fused_xy_two_blocks(c->fa)) ? 2 : 1;
  return MHD_OK;
}

int mhd_get_tiling(mhd_ctx* c, int* sw, int* ly, int* lz) {
  if (!c || !sw || !ly || !lz) return fail(MHD_ERR_ARG, "null argument");
  CHKRET(set_dev(c));
  CHKRET(ensure_fused(c));
  *sw = c->fa.sw;
  *ly = c->fa.ly;
  *lz = c->fa.lz;
  return MHD_OK;
}

int mhd_generate_problem_kh(mhd_ctx* c, double vy_amp, double rrv) {
  CHKRET(need_grid(c));
  mhd::launch_problem_kh(c->g, c->cfg.gam, vy_amp, rrv, c->cfg.xmin, c->cfg.xmax, c->grid[1], c->grid[3], c->grid[5],
                         c->P, c->U, c->stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  return MHD_OK;
}

int mhd_generate_problem_shear(mhd_ctx* c, const mhd_ic_params* ic) {
  CHKRET(need_grid(c));
  if (!ic || !(ic->a > 0.0) || !(ic->rho0 > 0.0) || !(ic->p0 > 0.0) || !(ic->sigma > 0.0))
    return fail(MHD_ERR_ARG, "shear parameters need a, rho0, p0, sigma > 0");
  const double v[7] = {ic->rho0, ic->p0, ic->v0, ic->a, ic->bx0, ic->amp, ic->sigma};
  mhd::launch_problem_shear(c->g, c->cfg.gam, v, ic->seed, c->grid[3], c->P, c->U, c->stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  return MHD_OK;
}

int mhd_diagnostics(mhd_ctx* c, double* out4) {
  CHKRET(need_grid(c));
  if (!out4) return fail(MHD_ERR_ARG, "null output");
  mhd::launch_diagnostics(c->g, c->P, c->grid[0], c->grid[1], c->grid[2], c->grid[3], c->grid[4], c->grid[5], c->diag,
                          c->stream);
  HIPCHK(hipGetLastError());
  std::vector<double> part(4 * (size_t)c->g.nz);
  HIPCHK(hipMemcpyAsync(part.data(), c->diag, part.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  double sv = 0.0, sm = 0.0, sy = 0.0, md = 0.0;
  for (int k = 0; k < c->g.nz; ++k) {
    sv += part[4 * k]; sm += part[4 * k + 1]; sy += part[4 * k + 2]; md = std::max(md, part[4 * k + 3]);
  }
  out4[0] = sv;
  out4[1] = sm / sv;
  out4[2] = std::sqrt(sy / sv);
  out4[3] = md;
  return MHD_OK;
}

int mhd_status(mhd_ctx* c, int* status) {
  if (!c || !status) return fail(MHD_ERR_ARG, "null argument");
  CHKRET(select_dev(c));
  StepScalars h;
  CHKRET(read_scalars(c, &h));
  *status = h.status;
  return MHD_OK;
}

int mhd_last_run_timing(mhd_ctx* c, double* kernel_ms, int* launches) {
  if (!c || !kernel_ms || !launches) return fail(MHD_ERR_ARG, "null argument");
  *kernel_ms = c->last_ms;
  *launches = c->last_launches;
  return MHD_OK;
}

}  // extern "C"
