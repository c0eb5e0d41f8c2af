// Launch dispatch and the small helper kernels (fill, beam sums, intensity reduction).
#include <hip/hip_runtime.h>

#include "pnx_kernels.hpp"
#include "pnx_kernels_impl.hpp"

namespace pnx {

hipError_t half_step_n1(int dims, int parity, int mode, const HalfStepArgs& a, hipStream_t s);
hipError_t half_step_n3(int dims, int parity, int mode, const HalfStepArgs& a, hipStream_t s);
hipError_t half_step_n5(int dims, int parity, int mode, const HalfStepArgs& a, hipStream_t s);
hipError_t half_step_n7(int dims, int parity, int mode, const HalfStepArgs& a, hipStream_t s);
hipError_t half_step_n9(int dims, int parity, int mode, const HalfStepArgs& a, hipStream_t s);

hipError_t half_step_grouped_n1(int dims, int parity, int mode, const HalfStepArgs* args, const int2* items,
                                 const int* n_items, int grid, hipStream_t s);
hipError_t half_step_grouped_n3(int dims, int parity, int mode, const HalfStepArgs* args, const int2* items,
                                 const int* n_items, int grid, hipStream_t s);
hipError_t half_step_grouped_n5(int dims, int parity, int mode, const HalfStepArgs* args, const int2* items,
                                 const int* n_items, int grid, hipStream_t s);
hipError_t half_step_grouped_n7(int dims, int parity, int mode, const HalfStepArgs* args, const int2* items,
                                 const int* n_items, int grid, hipStream_t s);
hipError_t half_step_grouped_n9(int dims, int parity, int mode, const HalfStepArgs* args, const int2* items,
                                 const int* n_items, int grid, hipStream_t s);

hipError_t half_step2_n1(int dims, int mode, const HalfStepArgs& a, hipStream_t s);
hipError_t half_step2_n3(int dims, int mode, const HalfStepArgs& a, hipStream_t s);
hipError_t half_step2_n5(int dims, int mode, const HalfStepArgs& a, hipStream_t s);
hipError_t half_step2_n7(int dims, int mode, const HalfStepArgs& a, hipStream_t s);
hipError_t half_step2_n9(int dims, int mode, const HalfStepArgs& a, hipStream_t s);

hipError_t rev_n1(int dims, const RevArgs& a, hipStream_t s);
hipError_t rev_n3(int dims, const RevArgs& a, hipStream_t s);
hipError_t rev_n5(int dims, const RevArgs& a, hipStream_t s);
hipError_t rev_n7(int dims, const RevArgs& a, hipStream_t s);
hipError_t rev_n9(int dims, const RevArgs& a, hipStream_t s);

hipError_t disc_n1(int dims, int pb, int pa, const DiscArgs& a, hipStream_t s);
hipError_t disc_n3(int dims, int pb, int pa, const DiscArgs& a, hipStream_t s);
hipError_t disc_n5(int dims, int pb, int pa, const DiscArgs& a, hipStream_t s);
hipError_t disc_n7(int dims, int pb, int pa, const DiscArgs& a, hipStream_t s);
hipError_t disc_n9(int dims, int pb, int pa, const DiscArgs& a, hipStream_t s);

hipError_t phase_a_grouped_n1(int dims, const HalfStepArgs* half, const RevArgs* rev, const int2* items,
                               const int* n_items, int grid, hipStream_t s);
hipError_t phase_b_grouped_n1(int dims, const HalfStepArgs* half, const int2* items, const int* n_items, int grid,
                               hipStream_t s);
hipError_t phase_a_grouped_n3(int dims, const HalfStepArgs* half, const RevArgs* rev, const int2* items,
                               const int* n_items, int grid, hipStream_t s);
hipError_t phase_b_grouped_n3(int dims, const HalfStepArgs* half, const int2* items, const int* n_items, int grid,
                               hipStream_t s);
hipError_t phase_a_grouped_n5(int dims, const HalfStepArgs* half, const RevArgs* rev, const int2* items,
                               const int* n_items, int grid, hipStream_t s);
hipError_t phase_b_grouped_n5(int dims, const HalfStepArgs* half, const int2* items, const int* n_items, int grid,
                               hipStream_t s);
hipError_t phase_a_grouped_n7(int dims, const HalfStepArgs* half, const RevArgs* rev, const int2* items,
                               const int* n_items, int grid, hipStream_t s);
hipError_t phase_b_grouped_n7(int dims, const HalfStepArgs* half, const int2* items, const int* n_items, int grid,
                               hipStream_t s);
hipError_t phase_a_grouped_n9(int dims, const HalfStepArgs* half, const RevArgs* rev, const int2* items,
                               const int* n_items, int grid, hipStream_t s);
hipError_t phase_b_grouped_n9(int dims, const HalfStepArgs* half, const int2* items, const int* n_items, int grid,
                               hipStream_t s);

hipError_t pair_grouped_n1(int dims, const HalfStepArgs* half, const int2* items, const int* n_items, int grid,
                            hipStream_t s);
hipError_t pair_grouped_n3(int dims, const HalfStepArgs* half, const int2* items, const int* n_items, int grid,
                            hipStream_t s);
hipError_t pair_grouped_n5(int dims, const HalfStepArgs* half, const int2* items, const int* n_items, int grid,
                            hipStream_t s);
hipError_t pair_grouped_n7(int dims, const HalfStepArgs* half, const int2* items, const int* n_items, int grid,
                            hipStream_t s);
hipError_t pair_grouped_n9(int dims, const HalfStepArgs* half, const int2* items, const int* n_items, int grid,
                            hipStream_t s);

hipError_t launch_half_step(int N, int dims, int parity, int mode, const HalfStepArgs& a, hipStream_t s) {
  switch (N) {
    case 1: return half_step_n1(dims, parity, mode, a, s);
    case 3: return half_step_n3(dims, parity, mode, a, s);
    case 5: return half_step_n5(dims, parity, mode, a, s);
    case 7: return half_step_n7(dims, parity, mode, a, s);
    case 9: return half_step_n9(dims, parity, mode, a, s);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_half_step2(int N, int dims, int mode, const HalfStepArgs& a, hipStream_t s) {
  switch (N) {
    case 1: return half_step2_n1(dims, mode, a, s);
    case 3: return half_step2_n3(dims, mode, a, s);
    case 5: return half_step2_n5(dims, mode, a, s);
    case 7: return half_step2_n7(dims, mode, a, s);
    case 9: return half_step2_n9(dims, mode, a, s);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_rev(int N, int dims, const RevArgs& a, hipStream_t s) {
  switch (N) {
    case 1: return rev_n1(dims, a, s);
    case 3: return rev_n3(dims, a, s);
    case 5: return rev_n5(dims, a, s);
    case 7: return rev_n7(dims, a, s);
    case 9: return rev_n9(dims, a, s);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_phase_a_grouped(int N, int dims, const HalfStepArgs* half, const RevArgs* rev, const int2* items,
                                  const int* n_items, int grid, hipStream_t s) {
  switch (N) {
    case 1: return phase_a_grouped_n1(dims, half, rev, items, n_items, grid, s);
    case 3: return phase_a_grouped_n3(dims, half, rev, items, n_items, grid, s);
    case 5: return phase_a_grouped_n5(dims, half, rev, items, n_items, grid, s);
    case 7: return phase_a_grouped_n7(dims, half, rev, items, n_items, grid, s);
    case 9: return phase_a_grouped_n9(dims, half, rev, items, n_items, grid, s);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_phase_b_grouped(int N, int dims, const HalfStepArgs* half, const int2* items, const int* n_items,
                                  int grid, hipStream_t s) {
  switch (N) {
    case 1: return phase_b_grouped_n1(dims, half, items, n_items, grid, s);
    case 3: return phase_b_grouped_n3(dims, half, items, n_items, grid, s);
    case 5: return phase_b_grouped_n5(dims, half, items, n_items, grid, s);
    case 7: return phase_b_grouped_n7(dims, half, items, n_items, grid, s);
    case 9: return phase_b_grouped_n9(dims, half, items, n_items, grid, s);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_pair_grouped(int N, int dims, const HalfStepArgs* half, const int2* items, const int* n_items,
                               int grid, hipStream_t s) {
  switch (N) {
    case 1: return pair_grouped_n1(dims, half, items, n_items, grid, s);
    case 3: return pair_grouped_n3(dims, half, items, n_items, grid, s);
    case 5: return pair_grouped_n5(dims, half, items, n_items, grid, s);
    case 7: return pair_grouped_n7(dims, half, items, n_items, grid, s);
    case 9: return pair_grouped_n9(dims, half, items, n_items, grid, s);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_half_step_grouped(int N, int dims, int parity, int mode, const HalfStepArgs* args,
                                    const int2* items, const int* n_items, int grid, hipStream_t s) {
  switch (N) {
    case 1: return half_step_grouped_n1(dims, parity, mode, args, items, n_items, grid, s);
    case 3: return half_step_grouped_n3(dims, parity, mode, args, items, n_items, grid, s);
    case 5: return half_step_grouped_n5(dims, parity, mode, args, items, n_items, grid, s);
    case 7: return half_step_grouped_n7(dims, parity, mode, args, items, n_items, grid, s);
    case 9: return half_step_grouped_n9(dims, parity, mode, args, items, n_items, grid, s);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_disc(int N, int dims, int pb, int pa, const DiscArgs& a, hipStream_t s) {
  switch (N) {
    case 1: return disc_n1(dims, pb, pa, a, s);
    case 3: return disc_n3(dims, pb, pa, a, s);
    case 5: return disc_n5(dims, pb, pa, a, s);
    case 7: return disc_n7(dims, pb, pa, a, s);
    case 9: return disc_n9(dims, pb, pa, a, s);
  }
  return hipErrorInvalidValue;
}

__global__ void fill_kernel(double* __restrict__ p, long n, double v) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long stride = (long)gridDim.x * blockDim.x;
  for (long x = i; x < n; x += stride) p[x] = v;
}

__global__ void beam_sum_kernel(const double* __restrict__ in, long stride, int n_beams, long n, double* __restrict__ out) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int b = 0; b < n_beams; ++b) s += in[b * stride + i];
  out[i] = s;
}

// One block per beam (the body is shared with the march kernel).
// Intensity reduction in two stages so that one request spreads over many CUs:
// block (g, b) sums the parts [g chunk, (g + 1) chunk) of beam b over every
// step into slot g of step 0 of the beam's partial buffer (free once the
// forward is done: the forward writes steps 1..n), then one block per beam sums
// the slots in order.  Deterministic.
__global__ __launch_bounds__(kBlock) void reduce_intensity_parts_kernel(double* __restrict__ partial, long step_stride,
                                                                        long beam_stride, int n_parts, int chunk,
                                                                        const double* __restrict__ rec, int rec_len,
                                                                        int sigma_off, int n_steps, int n_lines,
                                                                        double dE, double dvol) {
  const int g = blockIdx.x, b = blockIdx.y;
  double* pb = partial + b * beam_stride;
  const int t0 = g * chunk, t1 = t0 + chunk < n_parts ? t0 + chunk : n_parts;
  reduce_intensity_body(pb, step_stride, t0, t1, rec, rec_len, sigma_off, n_steps, n_lines, dE, dvol,
                        pb + (long)g * n_lines);
}

__global__ void reduce_intensity_final_kernel(const double* __restrict__ partial, long beam_stride, int n_chunks,
                                              int n_lines, double* __restrict__ I_out) {
  const int b = blockIdx.x, r = threadIdx.x;
  if (r >= n_lines) return;
  gcd __restrict__ pb = glb(partial) + b * beam_stride;
  double s = 0.0;
  for (int g = 0; g < n_chunks; ++g) s += pb[(long)g * n_lines + r];
  glb(I_out)[(long)b * n_lines + r] = s;
}

__global__ void copy_kernel(double* __restrict__ dst, const double* __restrict__ src, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long stride = (long)gridDim.x * blockDim.x;
  for (long x = i; x < n; x += stride) dst[x] = src[x];
}

hipError_t launch_copy(double* dst, const double* src, long n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  long blocks = (n + kBlock - 1) / kBlock;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(copy_kernel, dim3((unsigned)blocks), dim3(kBlock), 0, s, dst, src, n);
  return hipGetLastError();
}

hipError_t launch_fill(double* p, long n, double v, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  long blocks = (n + kBlock - 1) / kBlock;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(fill_kernel, dim3((unsigned)blocks), dim3(kBlock), 0, s, p, n, v);
  return hipGetLastError();
}

hipError_t launch_beam_sum(const double* in, long stride, int n_beams, long n, double* out, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const long blocks = (n + kBlock - 1) / kBlock;
  hipLaunchKernelGGL(beam_sum_kernel, dim3((unsigned)blocks), dim3(kBlock), 0, s, in, stride, n_beams, n, out);
  return hipGetLastError();
}

hipError_t launch_reduce_intensity(double* partial, long step_stride, long beam_stride, int n_parts,
                                   const double* fwd_rec, int rec_len, int sigma_off, int n_steps, int n_lines,
                                   double dE, double dvol, int n_beams, double* I_out, hipStream_t s) {
  if (n_parts < 1 || n_lines < 1 || n_lines > kMaxLines) return hipErrorInvalidValue;
  const int chunk = (n_parts + 31) / 32;  // at most 32 blocks per beam, each over chunk parts
  const int n_chunks = (n_parts + chunk - 1) / chunk;
  hipLaunchKernelGGL(reduce_intensity_parts_kernel, dim3(n_chunks, n_beams), dim3(kBlock), 0, s, partial, step_stride,
                     beam_stride, n_parts, chunk, fwd_rec, rec_len, sigma_off, n_steps, n_lines, dE, dvol);
  hipLaunchKernelGGL(reduce_intensity_final_kernel, dim3(n_beams), dim3(64), 0, s, partial, beam_stride, n_chunks,
                     n_lines, I_out);
  return hipGetLastError();
}

}  // namespace pnx
