// boundary.hpp -- the boundary_mod surface of src_cpp_kokkos_host/boundary.hpp:139-144.
// The ghost fill runs on the device, so the send/receive buffers carry no data; the type
// and the Allocate/SendRecv functions remain so that callers link unchanged.
#ifndef MHD_AMD_BOUNDARY_HPP_
#define MHD_AMD_BOUNDARY_HPP_

#include "mhd.hpp"

namespace boundary_mod {
extern int boundary_xin, boundary_xout;
extern int boundary_yin, boundary_yout;
extern int boundary_zin, boundary_zout;

template <typename T>
class BoundaryArray {
 public:
  int nv = 0, ng = 0, n3 = 0, n2 = 0, n1 = 0;
  void allocate(int nv_, int ng_, int n3_, int n2_, int n1_) {
    nv = nv_; ng = ng_; n3 = n3_; n2 = n2_; n1 = n1_;
  }
  void deallocate() {}
};

extern BoundaryArray<double> Bs, Br;

void AllocateBoundaryVariables(BoundaryArray<double>& Bs, BoundaryArray<double>& Br);
void DeallocateBoundaryVariables(BoundaryArray<double>& Bs, BoundaryArray<double>& Br);
void SetBoundaryCondition(hydflux_mod::FieldArray<double>& P, BoundaryArray<double>& Bs, BoundaryArray<double>& Br);
void SendRecvBoundary(const BoundaryArray<double>& Bs, BoundaryArray<double>& Br);
}  // namespace boundary_mod
#endif
