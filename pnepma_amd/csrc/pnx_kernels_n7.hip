// Instantiation of the PN-CSD kernels for PN order 7.
#include "pnx_kernels_impl.hpp"

namespace pnx {

hipError_t half_step_n7(int dims, int parity, int mode, const HalfStepArgs& a, hipStream_t s) {
  return half_step_order<7>(dims, parity, mode, a, s);
}

hipError_t dens_n7(int dims, const DensArgs& a, hipStream_t s) { return dens_order<7>(dims, a, s); }

}  // namespace pnx
