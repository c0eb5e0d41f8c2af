// Instantiation of the PN-CSD kernels for PN order 9.
#include "pnx_kernels_impl.hpp"

namespace pnx {

hipError_t half_step_n9(int dims, int parity, int mode, const HalfStepArgs& a, hipStream_t s) {
  return half_step_order<9>(dims, parity, mode, a, s);
}

hipError_t rev_n9(int dims, const RevArgs& a, hipStream_t s) { return rev_order<9>(dims, a, s); }

hipError_t half_step2_n9(int dims, int mode, const HalfStepArgs& a, hipStream_t s) {
  return half_step2_order<9>(dims, mode, a, s);
}

hipError_t half_step_grouped_n9(int dims, int parity, int mode, const HalfStepArgs* args, const int2* items,
                                 const int* n_items, int grid, hipStream_t s) {
  return half_step_grouped_order<9>(dims, parity, mode, args, items, n_items, grid, s);
}

hipError_t disc_n9(int dims, int pb, int pa, const DiscArgs& a, hipStream_t s) {
  return disc_order<9>(dims, pb, pa, a, s);
}

hipError_t phase_a_grouped_n9(int dims, const HalfStepArgs* half, const RevArgs* rev, const int2* items,
                               const int* n_items, int grid, hipStream_t s) {
  return phase_a_grouped_order<9>(dims, half, rev, items, n_items, grid, s);
}

hipError_t phase_b_grouped_n9(int dims, const HalfStepArgs* half, const int2* items, const int* n_items, int grid,
                               hipStream_t s) {
  return phase_b_grouped_order<9>(dims, half, items, n_items, grid, s);
}

hipError_t pair_grouped_n9(int dims, const HalfStepArgs* half, const int2* items, const int* n_items, int grid,
                            hipStream_t s) {
  return pair_grouped_order<9>(dims, half, items, n_items, grid, s);
}

}  // namespace pnx
