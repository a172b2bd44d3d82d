// boundary.cpp -- boundary_mod over the MI355X library: SetBoundaryCondition fills the
// ghost cells of P in HBM (src_cpp_kokkos_host/boundary.cpp:750-893 for one tile per axis),
// or is recorded as part of a step the fused kernels run (see mhd.hpp).
#include "boundary.hpp"

#include "config.hpp"

namespace boundary_mod {

int boundary_xin = config::boundary_xin;
int boundary_xout = config::boundary_xout;
int boundary_yin = config::boundary_yin;
int boundary_yout = config::boundary_yout;
int boundary_zin = config::boundary_zin;
int boundary_zout = config::boundary_zout;

BoundaryArray<double> Bs, Br;

void AllocateBoundaryVariables(BoundaryArray<double>& Bs_, BoundaryArray<double>& Br_) {
  using namespace resolution_mod;
  Bs_.allocate(hydflux_mod::nprim, ngh, ktot, jtot, itot);
  Br_.allocate(hydflux_mod::nprim, ngh, ktot, jtot, itot);
}

void DeallocateBoundaryVariables(BoundaryArray<double>& Bs_, BoundaryArray<double>& Br_) {
  Bs_.deallocate();
  Br_.deallocate();
}

// one GPU holds the whole box: there is no neighbour to exchange with
void SendRecvBoundary(const BoundaryArray<double>&, BoundaryArray<double>&) {}

void SetBoundaryCondition(hydflux_mod::FieldArray<double>&, BoundaryArray<double>&, BoundaryArray<double>&) {
  hydflux_mod::ensure_device_state();
  // in a recorded step the fused kernels write the ghost images of the new state themselves
  if (hydflux_mod::RecordSetBoundaryCondition()) return;
  hydflux_mod::FlushRecordedStep();
  // ghost fill on the device (periodic / reflection / outflow per face)
  hydflux_mod::check(mhd_set_boundary(hydflux_mod::device_context()), "mhd_set_boundary");
}

}  // namespace boundary_mod
