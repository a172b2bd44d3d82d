// mhd.hpp -- the hydflux_mod surface of src_cpp_kokkos_host/mhd.hpp (containers, index
// constants, globals, step functions) over the MI355X library (include/mhd_amd.h).
//
// GridArray and FieldArray keep a host mirror with the reference accessors
// (x1a(i)...; operator()(n,k,j,i) on the [k][j][n][i] layout).  The state lives in HBM:
// the first step function after AllocateHydroVariables uploads G, P and U (the
// reference's GenerateProblem writes them on the host); afterwards h2d()/d2h() move
// them explicitly, as in src_cpp_kokkos_device.  Flux arrays are device scratch.
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
  void h2d() {
    check(mhd_set_grid(device_context(), a1.data(), b1.data(), a2.data(), b2.data(), a3.data(), b3.data()),
          "mhd_set_grid");
  }
};

enum class FieldKind { prim, cons, flux1, flux2, flux3 };

template <typename T>
class FieldArray {
 public:
  std::vector<T> h;
  T* data = nullptr;
  int nv = 0, n3 = 0, n2 = 0, n1 = 0;
  size_t size = 0;
  FieldKind kind = FieldKind::prim;

  FieldArray() = default;
  void allocate(int nv_, int n3_, int n2_, int n1_, FieldKind k) {
    nv = nv_; n3 = n3_; n2 = n2_; n1 = n1_; kind = k;
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
  void h2d();
  void d2h();
};

template <>
void FieldArray<double>::h2d();
template <>
void FieldArray<double>::d2h();

inline constexpr int ncomp{1};
inline constexpr int nprim{11 + ncomp};
inline constexpr int nden{0}, nve1{1}, nve2{2}, nve3{3}, nene{4}, npre{5}, ncsp{6}, nbm1{7}, nbm2{8}, nbm3{9},
    nbps{10}, nst{nprim - ncomp}, ned{nprim - 1};
inline constexpr int mconsv{9 + ncomp}, madd{3};
inline constexpr int mden{0}, mrv1{1}, mrv2{2}, mrv3{3}, meto{4}, mbm1{5}, mbm2{6}, mbm3{7}, mbps{8},
    mst{mconsv - ncomp}, med{mconsv - 1};

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
}  // namespace hydflux_mod
#endif
