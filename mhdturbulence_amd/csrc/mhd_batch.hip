// mhd_batch.hip -- many boxes of different sizes and parameters advanced together.
//
// Each step issues one launch per fused kernel (xy, z) plus one step-end launch for
// all live boxes: a work table maps every block to (box, block of that box), and each box
// keeps its own dt / c_h ring (a segmented reduction: a box's bounds never mix with another
// box's), its own time and stop flag.  The per-step launches are captured in a graph that
// is rebuilt only when the set of boxes changes.  Box storage is carved from one HBM
// arena, so admitting and retiring boxes allocates nothing on the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <new>
#include <string>
#include <vector>

#include "../../include/mhd_amd.h"
#include "mhd_device.hpp"
#include "mhd_fused.hpp"

using mhd::FusedArgs;
using mhd::Geom;
using mhd::StepScalars;

namespace mhd_capi_detail {
int fail(int code, const std::string& msg);
}
using mhd_capi_detail::fail;

#define BHIP(expr)                                                                               \
  do {                                                                                           \
    hipError_t e_ = (expr);                                                                      \
    if (e_ != hipSuccess) return fail(MHD_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)
#define BCHK(expr)              \
  do {                          \
    int r_ = (expr);            \
    if (r_ != MHD_OK) return r_; \
  } while (0)

namespace {

constexpr uint64_t ALIGN = 256;

struct BatchBox {
  bool live = false;
  Geom g{};
  uint64_t off = 0, bytes = 0;
  double* P = nullptr;
  double* U0 = nullptr;
  double* U1 = nullptr;
  double* acc = nullptr;
  double* grid[6] = {};
  double* diag = nullptr;
  FusedArgs fa{};
};

struct Tiling {
  int sw, ly, lz;
};

Geom make_geom(int nx, int ny, int nz) {
  Geom g{};
  g.nx = nx; g.ny = ny; g.nz = nz;
  g.itot = nx + 5; g.jtot = ny + 5; g.ktot = nz + 5;
  g.pitch = ((g.itot + mhd::IOFF) + 15) / 16 * 16;
  g.vstride = (size_t)g.ktot * g.jtot * g.pitch;
  return g;
}

uint64_t up(uint64_t b) { return (b + ALIGN - 1) / ALIGN * ALIGN; }

// byte sizes of a box's arrays, in carve order
constexpr int NREGION = 11;
// precision: the batch's mhd_config.precision (2 stores the divergence acc in FP32)
void box_layout(const Geom& g, int precision, uint64_t* sz) {
  sz[0] = up(mhd::NPRIM * g.vstride * 8);                  // P
  sz[1] = sz[2] = up(mhd::NCONS * g.vstride * 8);          // U0 U1
  sz[3] = up(mhd::NCONS * g.vstride * (precision == 2 ? 4 : 8));  // acc
  sz[4] = sz[5] = up((uint64_t)g.itot * 8);                // x1a x1b
  sz[6] = sz[7] = up((uint64_t)g.jtot * 8);                // x2a x2b
  sz[8] = sz[9] = up((uint64_t)g.ktot * 8);                // x3a x3b
  sz[10] = up(4ull * g.nz * 8);                            // diagnostics partials
}

}  // namespace

struct mhd_batch {
  mhd_config base{};
  int dev = 0;
  hipStream_t stream = nullptr;
  int max_boxes = 0;
  char* arena = nullptr;
  uint64_t arena_bytes = 0;
  std::map<uint64_t, uint64_t> free_blocks;  // offset -> size
  std::vector<BatchBox> boxes;
  std::map<long long, Tiling> tilings;      // (nx,ny,nz) key -> tiling
  StepScalars* d_sc = nullptr;
  FusedArgs* d_args = nullptr;
  double* d_dt_last = nullptr;
  int2* d_tab[2] = {nullptr, nullptr};  // work tables of the xy and z launches, group by group
  int tab_cap[2] = {0, 0};
  int nblk[2] = {0, 0};
  int* d_live = nullptr;
  int nlive = 0;
  bool dirty = true;
  hipGraphExec_t graph = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double last_ms = 0.0;
  // box groups of the step graph: group g's blocks are tab[q][goff[g][q], +gblk[g][q]) and its
  // boxes live[lofs[g], +glive[g]); the second group runs on stream2
  int ngroups = 0;
  int goff[2][2] = {{0, 0}, {0, 0}}, gblk[2][2] = {{0, 0}, {0, 0}};
  int lofs[2] = {0, 0}, glive[2] = {0, 0};
  size_t glds[2] = {0, 0};
  hipStream_t stream2 = nullptr;
  hipEvent_t ev_xy = nullptr, ev_join = nullptr;
};

namespace {

long long shape_key(int nx, int ny, int nz) { return ((long long)nx << 40) | ((long long)ny << 20) | nz; }

FusedArgs configure(mhd_batch* b, const Geom& g, bool* ok) {
  FusedArgs a{};
  a.g = g;
  a.mixed = b->base.precision != 0;
  a.acc32 = b->base.precision == 2;
  *ok = mhd::fused_configure(a, false);
  auto it = b->tilings.find(shape_key(g.nx, g.ny, g.nz));
  if (it != b->tilings.end()) {
    a.sw = it->second.sw;
    a.ns = (g.nx + a.sw - 1) / a.sw;
    a.ly = it->second.ly;
    a.lz = it->second.lz;
  }
  return a;
}

// first-fit carve of `bytes` from the arena
bool arena_take(mhd_batch* b, uint64_t bytes, uint64_t* off) {
  for (auto it = b->free_blocks.begin(); it != b->free_blocks.end(); ++it) {
    if (it->second >= bytes) {
      *off = it->first;
      const uint64_t rest = it->second - bytes;
      const uint64_t noff = it->first + bytes;
      b->free_blocks.erase(it);
      if (rest) b->free_blocks[noff] = rest;
      return true;
    }
  }
  return false;
}

void arena_give(mhd_batch* b, uint64_t off, uint64_t bytes) {
  auto it = b->free_blocks.emplace(off, bytes).first;
  auto nx = std::next(it);
  if (nx != b->free_blocks.end() && it->first + it->second == nx->first) {
    it->second += nx->second;
    b->free_blocks.erase(nx);
  }
  if (it != b->free_blocks.begin()) {
    auto pv = std::prev(it);
    if (pv->first + pv->second == it->first) {
      pv->second += it->second;
      b->free_blocks.erase(it);
    }
  }
}

void drop_graph(mhd_batch* b) {
  if (b->graph) (void)hipGraphExecDestroy(b->graph);
  b->graph = nullptr;
}

int ensure_tab(mhd_batch* b, int q, int n) {
  if (n <= b->tab_cap[q]) return MHD_OK;
  if (b->d_tab[q]) (void)hipFree(b->d_tab[q]);
  b->d_tab[q] = nullptr;
  const int cap = std::max(n, 2 * b->tab_cap[q]);
  BHIP(hipMalloc((void**)&b->d_tab[q], (size_t)cap * sizeof(int2)));
  b->tab_cap[q] = cap;
  return MHD_OK;
}

// Box groups of one step.  With two groups the step graph runs group 0's xy kernel, then
// group 0's z kernel on the batch stream beside group 1's xy and z kernels on the second
// stream, so the CUs one launch's tail leaves idle take the other group's blocks.  Boxes go to
// the group with fewer cells so far, largest first.  FP64 batches use two groups (their
// kernels take a CU's whole LDS, one launch at a time per CU; measured 2 % faster on the
// 96-box ensemble); FP32 batches one (their z blocks fill the CUs' registers, and the
// interleaving measured 1 % slower).
#ifndef MHD_BATCH_GROUPS
#define MHD_BATCH_GROUPS 2
#endif

// work tables and the captured step graph for the current set of live boxes
int rebuild(mhd_batch* b) {
  std::vector<int2> tab[2];
  std::vector<int> live;
  std::vector<int> ids;
  for (int s = 0; s < b->max_boxes; ++s)
    if (b->boxes[s].live) ids.push_back(s);
  const auto cells = [&](int s) { const Geom& g = b->boxes[s].fa.g; return (long long)g.nx * g.ny * g.nz; };
  std::stable_sort(ids.begin(), ids.end(), [&](int p, int q) { return cells(p) > cells(q); });
  const int ngroups = (MHD_BATCH_GROUPS > 1 && ids.size() > 1 && b->base.precision == 0) ? 2 : 1;
  std::vector<int> members[2];
  long long load[2] = {0, 0};
  for (int s : ids) {
    const int g = (ngroups == 2 && load[1] < load[0]) ? 1 : 0;
    members[g].push_back(s);
    load[g] += cells(s);
  }
  b->ngroups = ngroups;
  for (int g = 0; g < ngroups; ++g) {
    std::sort(members[g].begin(), members[g].end());
    b->lofs[g] = (int)live.size();
    b->glive[g] = (int)members[g].size();
    b->glds[g] = 0;
    for (int q = 0; q < 2; ++q) b->goff[g][q] = (int)tab[q].size();
    for (int s : members[g]) {
      const BatchBox& x = b->boxes[s];
      live.push_back(s);
      const int nb[2] = {mhd::fused_blocks_xy(x.fa), mhd::fused_blocks_z(x.fa)};
      for (int q = 0; q < 2; ++q)
        for (int k = 0; k < nb[q]; ++k) tab[q].push_back(make_int2(s, k));
      b->glds[g] = std::max(b->glds[g], mhd::fused_xy_lds(x.fa));
    }
    for (int q = 0; q < 2; ++q) b->gblk[g][q] = (int)tab[q].size() - b->goff[g][q];
  }
  for (int q = 0; q < 2; ++q) {
    BCHK(ensure_tab(b, q, std::max<int>(1, (int)tab[q].size())));
    if (!tab[q].empty())
      BHIP(hipMemcpy(b->d_tab[q], tab[q].data(), tab[q].size() * sizeof(int2), hipMemcpyHostToDevice));
    b->nblk[q] = (int)tab[q].size();
  }
  if (!live.empty()) BHIP(hipMemcpy(b->d_live, live.data(), live.size() * sizeof(int), hipMemcpyHostToDevice));
  b->nlive = (int)live.size();
  drop_graph(b);
  if (b->nlive > 0) {
    const int ef = b->base.energy_flux, pr = b->base.precision;
    const bool rk2 = b->base.stages == 2;
    auto xy_at = [&](int g, hipStream_t st, int stage) {
      mhd::launch_batch_xy(b->d_args, b->d_tab[0] + b->goff[g][0], b->gblk[g][0], b->glds[g], ef, pr, st, stage);
    };
    auto z_at = [&](int g, hipStream_t st, int stage) {
      mhd::launch_batch_z(b->d_args, b->d_tab[1] + b->goff[g][1], b->gblk[g][1], ef, pr, st, stage);
    };
    // first launch of a group's iteration, then the rest of it: one Euler stage, or the two
    // Runge-Kutta stages (U0 -> U1, then U1 -> U0 averaged with U0)
    auto xy = [&](int g, hipStream_t st) { xy_at(g, st, rk2 ? 1 : 0); };
    auto z_end = [&](int g, hipStream_t st) {
      z_at(g, st, rk2 ? 1 : 0);
      if (rk2) {
        xy_at(g, st, 2);
        z_at(g, st, 2);
      }
      mhd::launch_batch_step_end(b->d_args, b->d_live + b->lofs[g], b->glive[g], b->d_dt_last, st);
    };
    BHIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    // every call below is checked after the capture has ended, so a failure never leaves the
    // stream capturing
    hipError_t ce = hipSuccess;
    auto chk = [&](hipError_t e) {
      if (ce == hipSuccess) ce = e;
    };
    xy(0, b->stream);
    if (b->ngroups == 2) {
      // fork: the second stream joins the capture after group 0's xy launch
      chk(hipEventRecord(b->ev_xy, b->stream));
      chk(hipStreamWaitEvent(b->stream2, b->ev_xy, 0));
      xy(1, b->stream2);
      z_end(1, b->stream2);
      chk(hipEventRecord(b->ev_join, b->stream2));
    }
    z_end(0, b->stream);
    if (b->ngroups == 2) chk(hipStreamWaitEvent(b->stream, b->ev_join, 0));  // join
    chk(hipGetLastError());
    hipGraph_t gr = nullptr;
    hipError_t e = hipStreamEndCapture(b->stream, &gr);
    if (ce != hipSuccess || e != hipSuccess) {
      if (gr) (void)hipGraphDestroy(gr);
      return fail(MHD_ERR_HIP, std::string("step graph capture: ") + hipGetErrorString(ce != hipSuccess ? ce : e));
    }
    e = hipGraphInstantiate(&b->graph, gr, nullptr, nullptr, 0);
    (void)hipGraphDestroy(gr);
    if (e != hipSuccess) {
      b->graph = nullptr;
      return fail(MHD_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    }
  }
  b->dirty = false;
  return MHD_OK;
}

int valid_box(mhd_batch* b, int id) {
  if (!b) return fail(MHD_ERR_ARG, "null batch");
  if (id < 0 || id >= b->max_boxes || !b->boxes[id].live) return fail(MHD_ERR_ARG, "no such box");
  return MHD_OK;
}

// dt / c_h of the box's current P into ring slot 0 (the fused kernels read no ghost cell of U0)
int box_prologue(mhd_batch* b, BatchBox& x, StepScalars* sc, double tmax) {
  StepScalars h{};
  h.tmax = tmax;
  h.red[0][0] = h.red[1][0] = h.red[2][0] = mhd::DT_INIT_BITS;
  BHIP(hipMemcpyAsync(sc, &h, sizeof h, hipMemcpyHostToDevice, b->stream));
  mhd::launch_reduce_prim(x.g, x.P, x.grid[0], x.grid[2], x.grid[4], sc, 1, b->stream);
  BHIP(hipGetLastError());
  return MHD_OK;
}

}  // namespace

extern "C" {

int mhd_batch_destroy(mhd_batch* b) {
  if (!b) return MHD_OK;
  (void)hipSetDevice(b->dev);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  drop_graph(b);
  for (int q = 0; q < 2; ++q)
    if (b->d_tab[q]) (void)hipFree(b->d_tab[q]);
  void* bufs[] = {b->arena, b->d_sc, b->d_args, b->d_dt_last, b->d_live};
  for (void* p : bufs)
    if (p) (void)hipFree(p);
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  if (b->ev_xy) (void)hipEventDestroy(b->ev_xy);
  if (b->ev_join) (void)hipEventDestroy(b->ev_join);
  if (b->stream2) (void)hipStreamDestroy(b->stream2);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return MHD_OK;
}

int mhd_batch_create(const mhd_config* base, int max_boxes, uint64_t arena_bytes, mhd_batch** out) {
  if (!base || !out || max_boxes < 1 || arena_bytes < ALIGN) return fail(MHD_ERR_ARG, "bad batch arguments");
  *out = nullptr;
  for (int q = 0; q < 6; ++q)
    if (base->bc[q] < 1 || base->bc[q] > 3) return fail(MHD_ERR_ARG, "boundary codes are 1, 2 or 3");
  if (!(base->gam > 1.0) || !(base->cfl > 0.0)) return fail(MHD_ERR_ARG, "gam > 1 and cfl > 0 required");
  if (base->stages < 0 || base->stages > 2) return fail(MHD_ERR_ARG, "stages is 0 or 1 (forward Euler) or 2");
  if (base->precision < 0 || base->precision > 2) return fail(MHD_ERR_ARG, "precision is 0, 1 or 2");
  int ndev = 0;
  BHIP(hipGetDeviceCount(&ndev));
  if (base->device < 0 || base->device >= ndev) return fail(MHD_ERR_ARG, "no such HIP device");
  mhd_batch* b = new (std::nothrow) mhd_batch();
  if (!b) return fail(MHD_ERR_NOMEM, "host allocation failed");
  b->base = *base;
  b->dev = base->device;
  b->max_boxes = max_boxes;
  b->boxes.resize(max_boxes);
  auto bail = [&](int r) { mhd_batch_destroy(b); return r; };
  if (hipSetDevice(b->dev) != hipSuccess) return bail(fail(MHD_ERR_HIP, "hipSetDevice"));
  // both streams at the default priority
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&b->stream2, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess ||
      hipEventCreateWithFlags(&b->ev_xy, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&b->ev_join, hipEventDisableTiming) != hipSuccess)
    return bail(fail(MHD_ERR_HIP, "stream/event creation failed"));
  arena_bytes = arena_bytes / ALIGN * ALIGN;
  if (hipMalloc((void**)&b->arena, arena_bytes) != hipSuccess) return bail(fail(MHD_ERR_NOMEM, "arena hipMalloc"));
  b->arena_bytes = arena_bytes;
  b->free_blocks[0] = arena_bytes;
  if (hipMalloc((void**)&b->d_sc, sizeof(StepScalars) * max_boxes) != hipSuccess ||
      hipMalloc((void**)&b->d_args, sizeof(FusedArgs) * max_boxes) != hipSuccess ||
      hipMalloc((void**)&b->d_dt_last, sizeof(double) * max_boxes) != hipSuccess ||
      hipMalloc((void**)&b->d_live, sizeof(int) * max_boxes) != hipSuccess)
    return bail(fail(MHD_ERR_NOMEM, "batch tables"));
  if (hipMemset(b->d_sc, 0, sizeof(StepScalars) * max_boxes) != hipSuccess)
    return bail(fail(MHD_ERR_HIP, "hipMemset"));
  *out = b;
  return MHD_OK;
}

int mhd_batch_box_bytes(const mhd_batch* b, int nx, int ny, int nz, uint64_t* bytes) {
  if (!b || !bytes || nx < 4 || ny < 4 || nz < 4) return fail(MHD_ERR_ARG, "bad box size");
  const Geom g = make_geom(nx, ny, nz);
  bool fits = false;  // the byte count is reported either way; add_box rejects what does not fit
  (void)configure(const_cast<mhd_batch*>(b), g, &fits);
  uint64_t sz[NREGION];
  box_layout(g, b->base.precision, sz);
  uint64_t t = 0;
  for (uint64_t v : sz) t += v;
  *bytes = t;
  return MHD_OK;
}

int mhd_batch_free_bytes(const mhd_batch* b, uint64_t* largest, uint64_t* total) {
  if (!b || !largest || !total) return fail(MHD_ERR_ARG, "null argument");
  *largest = 0;
  *total = 0;
  for (const auto& kv : b->free_blocks) {
    *largest = std::max(*largest, kv.second);
    *total += kv.second;
  }
  return MHD_OK;
}

int mhd_batch_set_tiling(mhd_batch* b, int nx, int ny, int nz, int sw, int ly, int lz) {
  if (!b || (sw != 16 && sw != 32 && sw != 64) || ly < 1 || lz < 1 || ly > ny || lz > nz || lz > 64)
    return fail(MHD_ERR_ARG, "tiling: sw in {16,32,64}, 1 <= ly <= ny, 1 <= lz <= min(nz, 64)");
  if (ly > mhd::fused_max_ly(sw, b->base.precision))
    return fail(MHD_ERR_ARG, "tiling: ly above " + std::to_string(mhd::fused_max_ly(sw, b->base.precision)) +
                                 " (LDS of the xy kernel) for this strip width and precision");
  b->tilings[shape_key(nx, ny, nz)] = Tiling{sw, ly, lz};
  return MHD_OK;
}

int mhd_batch_add_box(mhd_batch* b, int nx, int ny, int nz, const mhd_ic_params* ic, double tmax, int* box_id) {
  if (!b || !box_id) return fail(MHD_ERR_ARG, "null argument");
  *box_id = -1;
  if (nx < 4 || ny < 4 || nz < 4) return fail(MHD_ERR_ARG, "each axis needs at least 4 cells");
  if ((long long)(nx + 5) * (ny + 5) * (nz + 5) > (1LL << 30)) return fail(MHD_ERR_ARG, "box too large");
  if (ic && (!(ic->a > 0.0) || !(ic->rho0 > 0.0) || !(ic->p0 > 0.0) || !(ic->sigma > 0.0)))
    return fail(MHD_ERR_ARG, "shear parameters need a, rho0, p0, sigma > 0");
  int slot = -1;
  for (int s = 0; s < b->max_boxes; ++s)
    if (!b->boxes[s].live) { slot = s; break; }
  if (slot < 0) return fail(MHD_ERR_NOMEM, "no free box slot");
  BHIP(hipSetDevice(b->dev));
  BatchBox& x = b->boxes[slot];
  x = BatchBox{};
  x.g = make_geom(nx, ny, nz);
  bool fits = false;
  FusedArgs a = configure(b, x.g, &fits);
  if (!fits) return fail(MHD_ERR_ARG, "box too large: a variable column spans 4 GiB or more");
  uint64_t sz[NREGION];
  box_layout(x.g, b->base.precision, sz);
  uint64_t total = 0;
  for (uint64_t v : sz) total += v;
  if (!arena_take(b, total, &x.off)) return fail(MHD_ERR_NOMEM, "arena has no free region that large");
  x.bytes = total;
  char* base = b->arena + x.off;
  double** dst[NREGION] = {&x.P,       &x.U0,      &x.U1,      &x.acc,     &x.grid[0], &x.grid[1],
                           &x.grid[2], &x.grid[3], &x.grid[4], &x.grid[5], &x.diag};
  uint64_t o = 0;
  for (int q = 0; q < NREGION; ++q) {
    *dst[q] = reinterpret_cast<double*>(base + o);
    o += sz[q];
  }
  // staging for host-layout copies spans U1 and acc (at least 15 padded variables, which
  // hold the 12 unpadded host-layout ones)
  if (reinterpret_cast<char*>(x.acc) != reinterpret_cast<char*>(x.U1) + sz[2] ||
      sz[2] + sz[3] < (uint64_t)mhd::NPRIM * x.g.itot * x.g.jtot * x.g.ktot * 8)
    return fail(MHD_ERR_STATE, "box layout: staging area too small");
  // the whole region starts from zero so that the never-read padding is finite
  BHIP(hipMemsetAsync(base, 0, total, b->stream));
  mhd::launch_generate_grid(x.g, b->base.xmin, b->base.xmax, b->base.grid_shift, x.grid[0], x.grid[1], x.grid[2],
                            x.grid[3], x.grid[4], x.grid[5], b->stream);
  if (ic) {
    const double v[7] = {ic->rho0, ic->p0, ic->v0, ic->a, ic->bx0, ic->amp, ic->sigma};
    mhd::launch_problem_shear(x.g, b->base.gam, v, ic->seed, x.grid[3], x.P, x.U0, b->stream);
  }
  StepScalars* sc = b->d_sc + slot;
  x.live = true;
  a.x1a = x.grid[0];
  a.x2a = x.grid[2];
  a.x3a = x.grid[4];
  a.U0 = x.U0;
  a.U1 = x.U1;
  a.acc = x.acc;
  a.sc = sc;
  a.gam = b->base.gam;
  a.cfl = b->base.cfl;
  a.alphabp = b->base.alphabp;
  for (int q = 0; q < 6; ++q) a.bc[q] = b->base.bc[q];
  a.eflux = b->base.energy_flux;
  a.mixed = b->base.precision != 0;
  a.acc32 = b->base.precision == 2;
  x.fa = a;
  BHIP(hipMemcpyAsync(b->d_args + slot, &a, sizeof a, hipMemcpyHostToDevice, b->stream));
  BCHK(box_prologue(b, x, sc, tmax));
  BHIP(hipStreamSynchronize(b->stream));
  b->dirty = true;
  *box_id = slot;
  return MHD_OK;
}

int mhd_batch_upload_state(mhd_batch* b, int id, const double* P, const double* U) {
  BCHK(valid_box(b, id));
  if (!P || !U) return fail(MHD_ERR_ARG, "null state");
  BHIP(hipSetDevice(b->dev));
  BatchBox& x = b->boxes[id];
  const size_t cells = (size_t)x.g.itot * x.g.jtot * x.g.ktot;
  // U1 and acc are adjacent in the box region (15 or 20 variables of padded rows), which
  // hold the 12 unpadded host-layout variables; neither carries state before the restart
  double* stage = x.U1;
  BHIP(hipMemcpyAsync(stage, P, mhd::NPRIM * cells * 8, hipMemcpyHostToDevice, b->stream));
  mhd::launch_from_host_layout(x.g, mhd::NPRIM, stage, x.P, b->stream);
  BHIP(hipMemcpyAsync(stage, U, mhd::NCONS * cells * 8, hipMemcpyHostToDevice, b->stream));
  mhd::launch_from_host_layout(x.g, mhd::NCONS, stage, x.U0, b->stream);
  StepScalars h;
  BHIP(hipMemcpyAsync(&h, b->d_sc + id, sizeof h, hipMemcpyDeviceToHost, b->stream));
  BHIP(hipStreamSynchronize(b->stream));
  BCHK(box_prologue(b, x, b->d_sc + id, h.tmax));
  BHIP(hipStreamSynchronize(b->stream));
  return MHD_OK;
}

int mhd_batch_remove_box(mhd_batch* b, int id) {
  BCHK(valid_box(b, id));
  BHIP(hipSetDevice(b->dev));
  BHIP(hipStreamSynchronize(b->stream));
  BatchBox& x = b->boxes[id];
  arena_give(b, x.off, x.bytes);
  x = BatchBox{};
  b->dirty = true;
  return MHD_OK;
}

int mhd_batch_step(mhd_batch* b, int nsteps) {
  if (!b || nsteps < 0) return fail(MHD_ERR_ARG, "bad arguments");
  BHIP(hipSetDevice(b->dev));
  if (b->dirty) BCHK(rebuild(b));
  b->last_ms = 0.0;
  if (b->nlive == 0 || nsteps == 0) return MHD_OK;
  BHIP(hipEventRecord(b->ev0, b->stream));
  for (int s = 0; s < nsteps; ++s) BHIP(hipGraphLaunch(b->graph, b->stream));
  BHIP(hipEventRecord(b->ev1, b->stream));
  BHIP(hipStreamSynchronize(b->stream));
  float ms = 0.0f;
  BHIP(hipEventElapsedTime(&ms, b->ev0, b->ev1));
  b->last_ms = ms;
  return MHD_OK;
}

int mhd_batch_last_step_ms(const mhd_batch* b, double* ms) {
  if (!b || !ms) return fail(MHD_ERR_ARG, "null argument");
  *ms = b->last_ms;
  return MHD_OK;
}

int mhd_batch_box_state(mhd_batch* b, int id, double* time, int* steps, int* stopped, int* status) {
  BCHK(valid_box(b, id));
  BHIP(hipSetDevice(b->dev));
  StepScalars h;
  BHIP(hipMemcpyAsync(&h, b->d_sc + id, sizeof h, hipMemcpyDeviceToHost, b->stream));
  BHIP(hipStreamSynchronize(b->stream));
  if (time) *time = h.time;
  if (steps) *steps = h.steps_done;
  if (stopped) *stopped = h.stopped;
  if (status) *status = h.status;
  return MHD_OK;
}

int mhd_batch_poll(mhd_batch* b, int* ids, int cap, int* n) {
  if (!b || !n || (cap > 0 && !ids)) return fail(MHD_ERR_ARG, "null argument");
  BHIP(hipSetDevice(b->dev));
  std::vector<StepScalars> h(b->max_boxes);
  BHIP(hipMemcpyAsync(h.data(), b->d_sc, sizeof(StepScalars) * b->max_boxes, hipMemcpyDeviceToHost, b->stream));
  BHIP(hipStreamSynchronize(b->stream));
  int k = 0;
  for (int s = 0; s < b->max_boxes; ++s)
    if (b->boxes[s].live && h[s].stopped) {
      if (k < cap) ids[k] = s;
      ++k;
    }
  *n = k;
  return MHD_OK;
}

int mhd_batch_finish_box(mhd_batch* b, int id, double* diag4, double* P_host) {
  BCHK(valid_box(b, id));
  BHIP(hipSetDevice(b->dev));
  BatchBox& x = b->boxes[id];
  StepScalars h;
  BHIP(hipMemcpyAsync(&h, b->d_sc + id, sizeof h, hipMemcpyDeviceToHost, b->stream));
  BHIP(hipStreamSynchronize(b->stream));
  // one-stage iterations leave the state in U1 after an odd number of them (two-stage ones
  // end in U0)
  if (b->base.stages != 2 && (h.iter & 1)) {
    mhd::launch_copy_interior(x.g, x.U1, x.U0, mhd::NCONS, b->stream);
    BHIP(hipGetLastError());
    // the state is in U0 again; a later step must read U0 first
    int z = 0;
    BHIP(hipMemcpyAsync(&b->d_sc[id].iter, &z, sizeof(int), hipMemcpyHostToDevice, b->stream));
    const int ring = h.iter % 3;
    if (ring != 0) {
      // keep the ring consistent with iter = 0: slot 0 <- current slot, slot 1 cleared
      unsigned long long r[3][2];
      std::memcpy(r, h.red, sizeof r);
      unsigned long long nr[3][2] = {{r[ring][0], r[ring][1]}, {mhd::DT_INIT_BITS, mhd::CH_INIT_BITS},
                                     {mhd::DT_INIT_BITS, mhd::CH_INIT_BITS}};
      BHIP(hipMemcpyAsync(b->d_sc[id].red, nr, sizeof nr, hipMemcpyHostToDevice, b->stream));
    }
  }
  mhd::launch_update_prim(x.g, b->base.gam, x.U0, x.P, nullptr, b->stream);
  if (diag4) {
    mhd::launch_diagnostics(x.g, x.P, x.grid[0], x.grid[1], x.grid[2], x.grid[3], x.grid[4], x.grid[5], x.diag,
                            b->stream);
    std::vector<double> part(4 * (size_t)x.g.nz);
    BHIP(hipMemcpyAsync(part.data(), x.diag, part.size() * 8, hipMemcpyDeviceToHost, b->stream));
    BHIP(hipStreamSynchronize(b->stream));
    double sv = 0, sm = 0, sy = 0, md = 0;
    for (int k = 0; k < x.g.nz; ++k) {
      sv += part[4 * k]; sm += part[4 * k + 1]; sy += part[4 * k + 2]; md = std::max(md, part[4 * k + 3]);
    }
    diag4[0] = sv;
    diag4[1] = sm / sv;
    diag4[2] = std::sqrt(sy / sv);
    diag4[3] = md;
  }
  if (P_host) {
    const size_t cells = (size_t)x.g.itot * x.g.jtot * x.g.ktot;
    // U1 + acc (adjacent, 15 or 20 padded variables) stage the 12 host-layout variables; the
    // state is in U0 at this point
    mhd::launch_to_host_layout(x.g, mhd::NPRIM, x.P, x.U1, b->stream);
    BHIP(hipMemcpyAsync(P_host, x.U1, mhd::NPRIM * cells * 8, hipMemcpyDeviceToHost, b->stream));
  }
  BHIP(hipGetLastError());
  BHIP(hipStreamSynchronize(b->stream));
  return MHD_OK;
}

}  // extern "C"
