// March kernel: one workgroup advances one 1D request through a slice of its
// forward+adjoint sequence (march_next), with the half steps, density
// pullbacks, checkpoint copies and the intensity reduction separated by
// workgroup barriers instead of kernel boundaries.  A 1D request has a few
// hundred cells (a few tiles of 256), so a workgroup holds all of it and a
// step costs a few barrier round trips rather than seven launches.  A
// workgroup of 2 x 256 threads runs two tiles of a pass side by side, as the
// grid kernels' blocks do, instead of one after the other.  The device
// functions are the ones the grid kernels use, in the same order and with the
// same tiles, so the results are bitwise those of the grid path for either
// workgroup width.
#include <hip/hip_runtime.h>

#include "pnx_kernels_impl.hpp"
#include "pnx_march.hpp"

namespace pnx {
namespace {

__device__ __forceinline__ void block_copy(double* dst_, const double* src_, long n) {
  gd __restrict__ dst = glb(dst_);
  gcd __restrict__ src = glb(src_);
  for (long x = threadIdx.x; x < n; x += blockDim.x) dst[x] = src[x];
}

// The tiles of a pass: the workgroup's kBlock-thread groups take tiles
// g, g + G, g + 2G, ... (G groups).  g is the same for the whole wavefront;
// read from the first lane it stays in a scalar register with the tile
// arithmetic that follows from it.
template <class F>
__device__ __forceinline__ void for_tiles(int n_tiles, F&& f) {
  const int G = (int)blockDim.x / kBlock;
  for (int t = __builtin_amdgcn_readfirstlane((int)threadIdx.x / kBlock); t < n_tiles; t += G) f(t);
}

template <int N, int PAR, int MODE>
__device__ __forceinline__ void half_pass(const HalfStepArgs& a, const StepView& v) {
  for_tiles(a.L.n_tiles, [&](int t) { half_step_body<N, 1, PAR, MODE, true>(a, v, 0, t); });
  __syncthreads();
}

// Fused reverse pass at step i: [last odd adjoint half step of step i + 1],
// density pullback with weight int_fac, [first odd adjoint half step of step i].
template <int N>
__device__ __forceinline__ void rev_pass(const MarchArgs& m, int i, const double* u, double int_fac, bool a3, bool a1) {
  RevArgs ra;
  ra.adj = m.adj;
  ra.adj.rec = m.rec_a + (long)i * m.rec_len;
  ra.adj.dt = m.dens.dE / 2;
  ra.dens = m.dens;
  ra.dens.u = u;
  ra.dens.rec = m.rec_d + (long)i * m.rec_len;
  ra.dens.int_fac = int_fac;
  ra.rec3 = a3 ? m.rec_a + (long)(i + 1) * m.rec_len : nullptr;
  ra.a1 = a1 ? 1 : 0;
  for_tiles(ra.adj.L.n_tiles, [&](int t) { rev_body<N, 1, true>(ra, 0, t); });
  __syncthreads();
}

__device__ __forceinline__ double* march_buffer(const MarchArgs& m, int buf, int idx, long st) {
  if (buf == kBufCkpt) return m.ckpt + (long)idx * st;
  if (buf == kBufSeg) return m.seg + (long)idx * st;
  return m.state;
}

template <int N>
__device__ __forceinline__ void march_body(const MarchArgs& m) {
  const Layout& L = m.fwd.L;
  const long st = L.beam_stride;
  const double dE = m.dens.dE, h = dE / 2;
  MarchPos pos = m.pos;
  for (int op = 0; op < m.n_ops; ++op) {
    const MarchOp o = march_next(pos, m.n, m.K, m.nck);
    const double* src = march_buffer(m, o.src_buf, o.src_idx, st);
    if (o.kind == kOpFwd || o.kind == kOpRecompute) {
      if (o.ckpt_copy >= 0) {
        block_copy(m.ckpt + (long)o.ckpt_copy * st, m.state, st);
        __syncthreads();
      }
      double* dst = march_buffer(m, o.dst_buf, o.dst_idx, st);
      const double* rec = m.rec_f + (long)o.step * m.rec_len;
      // the recompute's intensity partials go to the scratch slot of step 0
      double* part = o.kind == kOpFwd ? m.partial + (long)o.step * m.partial_step : m.partial;
      if (o.first) half_pass<N, 1, kForward>(m.fwd, StepView{src, src, dst, rec, nullptr, h});
      half_pass<N, 0, kForward>(m.fwd, StepView{dst, src, dst, rec, part, dE});
      if (o.fuse) {
        // last odd half step of this step and the first of the next, in one pass
        double* dst2 = march_buffer(m, o.dst2_buf, o.dst2_idx, st);
        const StepView v{dst, dst, dst2 == dst ? nullptr : dst, rec, nullptr, h, rec + m.rec_len, dst2, h};
        for_tiles(L.n_tiles, [&](int t) { half_step2_body<N, 1, kForward, true>(m.fwd, v, 0, t); });
        __syncthreads();
      } else {
        half_pass<N, 1, kForward>(m.fwd, StepView{dst, dst, dst, rec, nullptr, h});
      }
      if (o.reduce)
        reduce_intensity_body(m.partial, m.partial_step, 0, L.n_tiles * kWavesPerBlock, m.rec_f, m.rec_len, m.sigma_off, m.n, L.n_lines,
                              dE, L.dvol, m.I_out);
    } else if (o.kind == kOpReverse) {
      const double* rec = m.rec_a + (long)o.step * m.rec_len;
      rev_pass<N>(m, o.step, src, 1.0, o.step < m.n, true);
      half_pass<N, 0, kAdjoint>(m.adj, StepView{m.lam, m.lam, m.lam, rec, nullptr, dE});
    } else if (o.kind == kOpFinal) {
      rev_pass<N>(m, 0, src, 0.5, true, false);
    }
  }
}

// All PN orders in one launch: a tick of the 1D lane then lasts as long as
// its slowest request, not the sum of one launch per order.  The kernel
// allocates the registers of the largest order for every workgroup; the bound
// of two waves per SIMD (two workgroups of kBlock threads per CU, or one of
// kMarchWide) keeps it at 249 VGPRs without spills (unbounded it took 256
// VGPRs + 7 AGPRs, so each request held a whole CU).
__global__ __launch_bounds__(kMarchWide, 1) void march_kernel(const MarchArgs* __restrict__ args, int n) {
  for (int r = blockIdx.x; r < n; r += gridDim.x) {
    switch (args[r].N) {
      case 1: march_body<1>(args[r]); break;
      case 3: march_body<3>(args[r]); break;
      case 5: march_body<5>(args[r]); break;
      case 7: march_body<7>(args[r]); break;
      case 9: march_body<9>(args[r]); break;
      default: break;
    }
  }
}

}  // namespace

hipError_t launch_march(const MarchArgs* args, int n_requests, int threads, hipStream_t s) {
  if (n_requests <= 0) return hipSuccess;
  if (threads != kBlock && threads != kMarchWide) return hipErrorInvalidValue;
  const dim3 grid(n_requests < 8192 ? n_requests : 8192), block(threads);
  hipLaunchKernelGGL(march_kernel, grid, block, 0, s, args, n_requests);
  return hipGetLastError();
}

}  // namespace pnx
