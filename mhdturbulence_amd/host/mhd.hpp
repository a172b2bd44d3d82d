// mhd.hpp -- the hydflux_mod surface of src_cpp_kokkos_host/mhd.hpp (containers, index
// constants, globals, step functions) over the MI355X library (include/mhd_amd.h).
//
// GridArray and FieldArray keep a host mirror with the reference accessors
// (x1a(i)...; operator()(n,k,j,i) on the [k][j][n][i] layout).  The state lives in HBM:
// the first step function after AllocateHydroVariables uploads G, P and U (the
// reference's GenerateProblem writes them on the host); afterwards h2d()/d2h() move
// them explicitly, as in src_cpp_kokkos_device.  Flux arrays are device scratch.
//
// Step path.  By default a loop that calls the step functions in the reference's order
// (ControlTimestep, SetBoundaryCondition, EvaluateCh, GetNumericalFluxD 1-3,
// UpdateConservU, DampPsi, UpdatePrimitvP; src_cpp_kokkos_host/main.cpp:229-257) runs on
// the fused kernels: ControlTimestep returns dt (and EvaluateCh c_h) from the bounds the
// previous step's kernels reduced, the calls in between are recorded, and UpdatePrimitvP
// issues the whole loop body as two fused launches.  A call out of that order, a dt or
// chg changed between the calls, or a transfer (h2d/d2h, e.g. of a flux array) runs the
// calls recorded so far on the staged kernels (one per reference function) and lets the
// rest of that step follow them.  SetStepPath(StepPath::staged) (or MHD_DROPIN_STAGED=1
// in the environment) makes every call run staged.
#ifndef MHD_AMD_MHD_HPP_
#define MHD_AMD_MHD_HPP_

#include <cstddef>
#include <string>
#include <vector>

#include "config.hpp"
#include "mhd_amd.h"

namespace hydflux_mod {
using index_t = int;
using size_t = std::size_t;

// the device box behind the module globals; created by AllocateHydroVariables
mhd_ctx* device_context();
// aborts with the library's message on a nonzero status (the reference fails fast)
void check(int status, const char* what);
// uploads G, P and U on the first step-function call after AllocateHydroVariables
void ensure_device_state();

template <typename T>
class GridArray {
 public:
  int n3 = 0, n2 = 0, n1 = 0;
  std::vector<T> a1, b1, a2, b2, a3, b3;
  GridArray() = default;
  void allocate(int n3_, int n2_, int n1_) {
    n3 = n3_; n2 = n2_; n1 = n1_;
    a1.assign(n1, T(0)); b1.assign(n1, T(0));
    a2.assign(n2, T(0)); b2.assign(n2, T(0));
    a3.assign(n3, T(0)); b3.assign(n3, T(0));
  }
  void deallocate() { a1.clear(); b1.clear(); a2.clear(); b2.clear(); a3.clear(); b3.clear(); }
  T& x1a(int i) noexcept { return a1[i]; }
  T& x1b(int i) noexcept { return b1[i]; }
  T& x2a(int j) noexcept { return a2[j]; }
  T& x2b(int j) noexcept { return b2[j]; }
  T& x3a(int k) noexcept { return a3[k]; }
  T& x3b(int k) noexcept { return b3[k]; }
  const T& x1a(int i) const noexcept { return a1[i]; }
  const T& x1b(int i) const noexcept { return b1[i]; }
  const T& x2a(int j) const noexcept { return a2[j]; }
  const T& x2b(int j) const noexcept { return b2[j]; }
  const T& x3a(int k) const noexcept { return a3[k]; }
  const T& x3b(int k) const noexcept { return b3[k]; }
  void h2d();
};

// which device array a FieldArray mirrors; AllocateHydroVariables binds P, U, Fx, Fy, Fz,
// any other FieldArray is host memory only (h2d / d2h do nothing)
enum class FieldKind { host, prim, cons, flux1, flux2, flux3 };

template <typename T>
class FieldArray {
 public:
  std::vector<T> h;
  T* data = nullptr;
  int nv = 0, n3 = 0, n2 = 0, n1 = 0;
  size_t size = 0;
  std::string name;
  FieldKind kind = FieldKind::host;

  FieldArray() = default;
  FieldArray(int nv_, int n3_, int n2_, int n1_) { allocate(nv_, n3_, n2_, n1_); }
  // src_cpp_kokkos_host/mhd.hpp:164: the name labels the Kokkos view there, and here only
  // this object
  void allocate(int nv_, int n3_, int n2_, int n1_, const std::string& name_ = "F") {
    nv = nv_; n3 = n3_; n2 = n2_; n1 = n1_; name = name_;
    size = static_cast<size_t>(nv) * n3 * n2 * n1;
    h.assign(size, T(0));
    data = h.data();
  }
  void deallocate() { h.clear(); data = nullptr; size = 0; }
  // reference layout: View<T****, LayoutLeft>(i, n, j, k)
  T& operator()(int n, int k, int j, int i) noexcept { return h[((static_cast<size_t>(k) * n2 + j) * nv + n) * n1 + i]; }
  const T& operator()(int n, int k, int j, int i) const noexcept {
    return h[((static_cast<size_t>(k) * n2 + j) * nv + n) * n1 + i];
  }
  T& href(int n, int k, int j, int i) const noexcept {
    return const_cast<T&>(h[((static_cast<size_t>(k) * n2 + j) * nv + n) * n1 + i]);
  }
  // dev / dref (src_cpp_kokkos_host/mhd.hpp:207-238) name the device view there.  Here the
  // device copy lives in the library's own layout in HBM, out of reach of host code, so
  // they return the host mirror's element, which h2d / d2h move to and from the device.
  T& dev(int n, int k, int j, int i) noexcept { return (*this)(n, k, j, i); }
  const T& dev(int n, int k, int j, int i) const noexcept { return (*this)(n, k, j, i); }
  T& dref(int n, int k, int j, int i) const noexcept { return href(n, k, j, i); }
  void h2d();
  void d2h();
};

template <>
void GridArray<double>::h2d();
template <>
void FieldArray<double>::h2d();
template <>
void FieldArray<double>::d2h();

inline constexpr int ncomp{1};
inline constexpr int nprim{11 + ncomp};
inline constexpr int nden{0}, nve1{1}, nve2{2}, nve3{3}, nene{4}, npre{5}, ncsp{6}, nbm1{7}, nbm2{8}, nbm3{9},
    nbps{10}, nst{nprim - ncomp}, ned{nprim - 1};
inline constexpr int mconsv{9 + ncomp}, madd{3};
// slots of the Riemann-solver vectors (conserved, flux, then c_f, v_n, p_t)
inline constexpr int mudn{0}, muvu{1}, muvv{2}, muvw{3}, muet{4}, mubu{5}, mubv{6}, mubw{7}, mubp{8},
    mfdn{mconsv}, mfvu{mconsv + 1}, mfvv{mconsv + 2}, mfvw{mconsv + 3}, mfet{mconsv + 4}, mfbu{mconsv + 5},
    mfbv{mconsv + 6}, mfbw{mconsv + 7}, mfbp{mconsv + 8}, must{mconsv - ncomp}, mued{mconsv - 1},
    mfst{2 * mconsv - ncomp}, mfed{2 * mconsv - 1}, mcsp{2 * mconsv}, mvel{2 * mconsv + 1}, mpre{2 * mconsv + 2};
inline constexpr int mden{0}, mrv1{1}, mrv2{2}, mrv3{3}, meto{4}, mbm1{5}, mbm2{6}, mbm3{7}, mbps{8},
    mrvu{1}, mrvv{2}, mrvw{3}, mbmu{5}, mbmv{6}, mbmw{7}, mst{mconsv - ncomp}, med{mconsv - 1};

extern GridArray<double> G;
extern FieldArray<double> P;
extern FieldArray<double> U;
extern FieldArray<double> Fx, Fy, Fz;
extern double gam;
extern double chg;

void AllocateHydroVariables(GridArray<double>& G, FieldArray<double>& U, FieldArray<double>& Fx,
                            FieldArray<double>& Fy, FieldArray<double>& Fz, FieldArray<double>& P);
void DeallocateHydroVariables(GridArray<double>& G, FieldArray<double>& U, FieldArray<double>& Fx,
                              FieldArray<double>& Fy, FieldArray<double>& Fz, FieldArray<double>& P);
void GetNumericalFlux1(const GridArray<double>& G, const FieldArray<double>& P, FieldArray<double>& Fx);
void GetNumericalFlux2(const GridArray<double>& G, const FieldArray<double>& P, FieldArray<double>& Fy);
void GetNumericalFlux3(const GridArray<double>& G, const FieldArray<double>& P, FieldArray<double>& Fz);
void GetNumericalFluxD(const int dir, const GridArray<double>& G, const FieldArray<double>& P, FieldArray<double>& F);
void UpdateConservU(const GridArray<double>& G, const FieldArray<double>& Fx, const FieldArray<double>& Fy,
                    const FieldArray<double>& Fz, FieldArray<double>& U);
void UpdatePrimitvP(const FieldArray<double>& U, FieldArray<double>& P);
void ControlTimestep(const GridArray<double>& G);
void EvaluateCh();
void DampPsi(const GridArray<double>& G, FieldArray<double>& U);

// Not in the reference: n iterations of the whole loop body on the device (dt and c_h
// stay in HBM), advancing resolution_mod::time_sim and dt; stops after the step that
// passes tmax when tmax > 0.  Returns the steps taken.
int AdvanceSteps(int n, double tmax);

// Not in the reference: the path of the step functions (see the top of this file).
enum class StepPath { fused, staged };
void SetStepPath(StepPath p);
StepPath GetStepPath();
// runs the step functions recorded since ControlTimestep on the staged kernels (called by
// every transfer; a caller that reaches the device context directly calls it first)
void FlushRecordedStep();
// SetBoundaryCondition's part of a recorded step: true when the call continues it
bool RecordSetBoundaryCondition();
}  // namespace hydflux_mod
#endif
