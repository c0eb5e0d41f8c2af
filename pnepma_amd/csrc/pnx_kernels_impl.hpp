// gfx950 kernels of the PN-CSD staggered energy step (templates; instantiated per
// PN order in pnx_kernels_n<N>.hip).
//
// One fused kernel per half step (starmap/src/solver.jl:171-196): a thread owns
// one unit cell (i, j, k) of the staggered grid and updates every class of the
// stepped parity at that cell.  Per class it
//   * applies the sparse advection blocks as two-point differences along each
//     non-singleton axis (kernel!/matmuldiff!, starmap/src/solver_core.jl:59-137),
//     with the CSC traversal order of the reference: column by column;
//   * rebuilds the boundary ghost value of a neighbour in registers when the
//     difference reaches past the domain (apply_boundary_source!/_zero!,
//     starmap/src/boundary_conditions.jl:67-102): ghosts are functions of the
//     neighbour's first interior value and of this cell's own values, so they
//     are never stored;
//   * forms S and C from the per-point partial densities and the energy record
//     (weighted_sum!, pn_solver/pn_problem.jl:213-249) instead of reading
//     precomputed S, C arrays;
//   * applies the exponential integrator factor once per degree
//     (right_hand_side!/expm1div, starmap/src/solver_core.jl:38-106) and the
//     update u += dt du (solver.jl:122-124);
//   * in the forward even step, accumulates the eee zeroth moment weighted by
//     the line weights into per-block partial sums (integrate_step_and_add!,
//     pn_solver/x_ray_intensities.jl:172-191).
// The state is moment-major: each moment of each class is a plane of V
// doubles with x fastest, so a wavefront reads 64 consecutive doubles per
// moment and the stencil neighbours along x come from the same cache lines.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "pnx_kernels.hpp"
#include "pnx_tables.hpp"

namespace pnx {

constexpr int kMaxLines = 8;
constexpr int kMmsStride = 16;

// Pointers into device memory (address space 1).  Accesses through them
// compile to global_load / global_store; through a generic pointer read from
// an argument block in memory they would be flat accesses, which also count
// against the scalar-load counter and so serialise with the argument loads.
using gcd = const __attribute__((address_space(1))) double*;
using gd = __attribute__((address_space(1))) double*;
__device__ __forceinline__ gcd glb(const double* p) { return (gcd)p; }
__device__ __forceinline__ gd glb(double* p) { return (gd)p; }

// Read-only data that all lanes of a wavefront read at the same address: the
// energy records, the boundary blocks, the beam moments and the adjoint seed.
// No kernel writes them, so they are read through the constant address space
// (4), where a load at a wavefront-uniform address becomes a scalar load.
// Through a global pointer each such read would be a vector load, and after
// any store of the kernel the compiler must also wait for each one before
// the arithmetic that follows, which chains one memory latency per element
// and degree in the mixing loops.
using gkd = const __attribute__((address_space(4))) double*;
__device__ __forceinline__ gkd cst(const double* p) { return (gkd)p; }

template <int N, int C>
constexpr int ncls() {
  return Tab<N>::t.n[C];
}

template <int N, int C>
constexpr bool class_has_degree(int l) {
  for (int i = 0; i < Tab<N>::t.n[C]; ++i)
    if (Tab<N>::t.l[C][i] == l) return true;
  return false;
}

// a / b for the hot loops: the hardware reciprocal estimate refined by two
// Newton steps, r <- r + r (1 - b r), then the quotient q = a r corrected by
// one residual step, q + r (a - b q).  From a seed of about 23 correct bits
// this gave the correctly rounded quotient on 18 M samples of the arguments
// below (b = c, a = exp(c) - 1, 2e-4 < |c| < 640) in a CPU check; the
// compiler's IEEE division adds scaling and fix-up steps for the ranges
// these arguments never reach (zero, infinities, denormals).
__device__ __forceinline__ double fast_div(double a, double b) {
  double r = __builtin_amdgcn_rcp(b);
  r = fma(fma(-b, r, 1.0), r, r);
  r = fma(fma(-b, r, 1.0), r, r);
  const double q = a * r;
  return fma(fma(-b, q, a), r, q);
}

__device__ __forceinline__ double expm1div(double c) {
  if (fabs(c) <= 0.0002) return 1.0 + 0.5 * c + 1.0 / 6.0 * c * c;
  return fast_div(exp(c) - 1.0, c);
}

// du += sum over non-singleton axes of A_d[C] (u_F[R] - u_F[L]) / h_d, with
// boundary ghosts of F rebuilt from this point's own values u (length n_C).
template <int N, int DIM, int C, bool SRC>
__device__ __forceinline__ void stencil(const Layout& L, gcd __restrict__ from, gkd coef,
                                        const double* u, double* du, int i, int j, int k, long p, double gfac,
                                        gcd __restrict__ prof_b, gkd mms) {
  constexpr int NC = ncls<N, C>();
  const long V = L.V;
  static_for<0, DIM>([&](auto Dc) {
    constexpr int D = decltype(Dc)::value;
    constexpr int F = switch_axis(C, D);
    constexpr int NF = ncls<N, F>();
    if constexpr (NF > 0) {
      constexpr bool codd = odd_in(C, D) != 0;
      const long st = D == 0 ? 1L : (D == 1 ? (long)L.X : (long)L.X * L.Y);
      const long pL = codd ? p - st : p;
      const long pR = codd ? p : p + st;
      gcd __restrict__ uf = from + L.cls_off[F];
      double vL[NF], vR[NF];
#pragma unroll
      for (int J = 0; J < NF; ++J) {
        vL[J] = uf[J * V + pL];
        vR[J] = uf[J * V + pR];
      }
      if constexpr (!codd) {
        const int idx = D == 0 ? i : (D == 1 ? j : k);
        const int M = D == 0 ? L.Mx : (D == 1 ? L.My : L.Mz);
        const bool left = idx == 0, right = idx == M - 1;
        if (left || right) {
          constexpr int boff = Tab<N>::t.b_off[F][D];
          gkd Bm = coef + boff;
          double pr = 0.0;
          if constexpr (SRC && D == 0) pr = prof_b[F * (long)L.Y * L.Z + j + (long)L.Y * k];
#pragma unroll
          for (int J = 0; J < NF; ++J) {
            double bc = 0.0;
#pragma unroll
            for (int I = 0; I < NC; ++I) bc += Bm[J * NC + I] * u[I];
            if (left) {
              vL[J] = 2.0 * (-bc + 0.0) - vR[J];
            } else {
              double g = 0.0;
              if constexpr (SRC && D == 0) g = (mms[F * kMmsStride + J] * pr) * gfac;
              vR[J] = 2.0 * (bc + g) - vL[J];
            }
          }
        }
      }
      const double ih = L.inv_h[D];
      static_for<0, NF>([&](auto Jc) {
        constexpr int J = decltype(Jc)::value;
        const double d = (vR[J] - vL[J]) * ih;
        static_for<0, NC>([&](auto Ic) {
          constexpr int I = decltype(Ic)::value;
          // advection values are compile-time constants (forward sign; the adjoint problem negates A)
          if constexpr (Tab<N>::t.nz[C][D][I][J] >= 0) {
            constexpr double av = SRC ? Tab<N>::t.a[C][D][I][J] : -Tab<N>::t.a[C][D][I][J];
            du[I] += av * d;
          }
        });
      });
    }
  });
}

// Per-point loops over elements and x-ray lines run in batches of kBatch: the
// loads of a batch are issued together and waited for once (a loop whose
// bound is a runtime value would wait for each load in turn).  Lanes of a
// batch beyond the count load entry e0 again and are not used, so the sums
// keep their order and their values.
constexpr int kBatch = 4;

__device__ __forceinline__ int batch_index(int e0, int b, int n) { return e0 + b < n ? e0 + b : e0; }

// S and the per-degree C of a point (weighted_sum!, pn_problem.jl:213-220).
template <int N>
__device__ __forceinline__ void mix(gcd __restrict__ rc, long V, long p, int ne, gkd rec, double& S,
                                    double* Cl) {
  S = 0.0;
#pragma unroll
  for (int l = 0; l <= N; ++l) Cl[l] = 0.0;
  for (int e0 = 0; e0 < ne; e0 += kBatch) {
    double r[kBatch];
#pragma unroll
    for (int b = 0; b < kBatch; ++b) r[b] = rc[batch_index(e0, b, ne) * V + p];
#pragma unroll
    for (int b = 0; b < kBatch; ++b) {
      const int e = e0 + b;
      if (e < ne) {
        S += r[b] * rec[e];
        gkd ce = rec + ne + e * (N + 1);
#pragma unroll
        for (int l = 0; l <= N; ++l) Cl[l] += r[b] * ce[l];
      }
    }
  }
}

// The per-step operands of one half step: the buffers it reads and writes, the
// energy record and the step length.  Grid kernels take them from their
// HalfStepArgs; the march kernel varies them step by step.
struct StepView {
  const double* from;  // other parity (beam 0)
  const double* src;   // updated parity, read (beam 0)
  double* dst;         // updated parity, written (beam 0)
  const double* rec;   // energy record
  double* partial;     // intensity partials of this step (forward even half step)
  double dt;
  const double* rec2;  // fused odd pair: record, output and step of the second update
  double* dst2;
  double dt2;
};

__device__ __forceinline__ StepView view_of(const HalfStepArgs& a) {
  return StepView{a.from, a.to_src, a.to_dst, a.rec, a.partial, a.dt, a.rec2, a.to_dst2, a.dt2};
}

// Change of the stencil result du of class C when its own values go from u to
// u1 and the source factor from gfac to gfac1 (the second update of a fused
// odd pair).  The neighbours of the other parity are the same for both
// updates, so only the boundary ghosts change, and only at boundary cells:
// a left ghost 2(-B u) - v_R moves v_R - v_L by 2 B (u1 - u), a right ghost
// 2(B u + g) - v_L by 2 (B (u1 - u) + g1 - g).
template <int N, int DIM, int C, bool SRC>
__device__ __forceinline__ void stencil_ghost_delta(const Layout& L, gkd coef, const double* u,
                                                    const double* u1, double* du, int i, int j, int k, double gfac,
                                                    double gfac1, gcd __restrict__ prof_b, gkd mms) {
  constexpr int NC = ncls<N, C>();
  static_for<0, DIM>([&](auto Dc) {
    constexpr int D = decltype(Dc)::value;
    constexpr int F = switch_axis(C, D);
    constexpr int NF = ncls<N, F>();
    if constexpr (NF > 0 && odd_in(C, D) == 0) {
      const int idx = D == 0 ? i : (D == 1 ? j : k);
      const int M = D == 0 ? L.Mx : (D == 1 ? L.My : L.Mz);
      const bool left = idx == 0, right = idx == M - 1;
      if (left || right) {
        constexpr int boff = Tab<N>::t.b_off[F][D];
        gkd Bm = coef + boff;
        double pr = 0.0;
        if constexpr (SRC && D == 0) pr = prof_b[F * (long)L.Y * L.Z + j + (long)L.Y * k];
        const double ih = L.inv_h[D];
        double dd[NF];
#pragma unroll
        for (int J = 0; J < NF; ++J) {
          double dbc = 0.0;
#pragma unroll
          for (int I = 0; I < NC; ++I) dbc += Bm[J * NC + I] * (u1[I] - u[I]);
          double dg = 0.0;
          if constexpr (SRC && D == 0)
            if (!left) dg = mms[F * kMmsStride + J] * pr * (gfac1 - gfac);
          dd[J] = 2.0 * (dbc + dg) * ih;
        }
        static_for<0, NF>([&](auto Jc) {
          constexpr int J = decltype(Jc)::value;
          static_for<0, NC>([&](auto Ic) {
            constexpr int I = decltype(Ic)::value;
            if constexpr (Tab<N>::t.nz[C][D][I][J] >= 0) {
              constexpr double av = SRC ? Tab<N>::t.a[C][D][I][J] : -Tab<N>::t.a[C][D][I][J];
              du[I] += av * dd[J];
            }
          });
        });
      }
    }
  });
}

// The update of class C at point p from its stencil result du (the rest of
// right_hand_side! + update!): un = u + dt * phi * (q - du - C u) / S.
template <int N, int C, int MODE>
__device__ __forceinline__ void apply_update(const HalfStepArgs& a, gkd rec, double dt,
                                             const double* u, const double* du, double* un, long p, int beam);

// One update of class C at point p (right_hand_side! + update!,
// solver_core.jl:92-106, solver.jl:122-124): own values u -> un, in registers.
template <int N, int DIM, int C, int MODE>
__device__ __forceinline__ void update_core(const HalfStepArgs& a, gkd rec, double dt,
                                            gcd __restrict__ from, const double* u, double* un, int i, int j, int k,
                                            long p, int beam) {
  constexpr int NC = ncls<N, C>();
  const int ne = a.L.n_e;
  double du[NC];
#pragma unroll
  for (int m = 0; m < NC; ++m) du[m] = 0.0;
  const double gfac = rec[ne + ne * (N + 1)];
  stencil<N, DIM, C, MODE == kForward>(a.L, from, cst(a.coef), u, du, i, j, k, p, gfac,
                                       glb(a.prof) + beam * a.prof_stride, cst(a.mms));
  apply_update<N, C, MODE>(a, rec, dt, u, du, un, p, beam);
}

template <int N, int C, int MODE>
__device__ __forceinline__ void apply_update(const HalfStepArgs& a, gkd rec, double dt,
                                             const double* u, const double* du, double* un, long p, int beam) {
  constexpr int NC = ncls<N, C>();
  const Layout& L = a.L;
  const long V = L.V;
  const int ne = L.n_e;
  double S, Cl[N + 1];
  mix<N>(glb(a.rho) + a.rho_off[C], V, p, ne, rec, S, Cl);
  double phi[N + 1];
  const double invS = 1.0 / S;
  static_for<0, N + 1>([&](auto lc) {
    constexpr int l = decltype(lc)::value;
    if constexpr (class_has_degree<N, C>(l)) phi[l] = expm1div(-Cl[l] * invS * dt);
  });
  double q0 = 0.0;
  if constexpr (MODE == kAdjoint && C == 0) {
    // compute_adjoint_source! (x_ray_intensities.jl:155-170), eee m = 0 only
    if (a.ibar != nullptr) {
      gkd sig = rec + ne + ne * (N + 1) + 1;
      gkd ib = cst(a.ibar) + beam * L.n_lines;
      gcd __restrict__ Wp = glb(a.W);
      const int nl = L.n_lines;
      for (int r0 = 0; r0 < nl; r0 += kBatch) {
        double wv[kBatch];
#pragma unroll
        for (int b = 0; b < kBatch; ++b) wv[b] = Wp[batch_index(r0, b, nl) * V + p];
#pragma unroll
        for (int b = 0; b < kBatch; ++b)
          if (r0 + b < nl) q0 -= sig[r0 + b] * wv[b] * ib[r0 + b];
      }
    }
  }
  static_for<0, NC>([&](auto Ic) {
    constexpr int I = decltype(Ic)::value;
    constexpr int l = Tab<N>::t.l[C][I];
    double t;
    if constexpr (MODE == kAdjoint) {
      const double q = (C == 0 && I == 0) ? q0 : 0.0;
      t = (q - du[I]) - Cl[l] * u[I];
    } else {
      t = (-du[I]) - Cl[l] * u[I];
    }
    t = t * invS;
    t = t * phi[l];
    un[I] = u[I] + dt * t;
  });
}

// Streams that only the thread of their point reads and writes, once per
// pass: the own values of the updated class in a half step, the odd-class
// adjoint at its point, rho_bar and the generation sums in the reverse pass.
// Nontemporal accesses keep them from evicting the lines the stencils share
// (other-parity rows, densities).  On the 2D P7 160-beam even half step this
// took 210 to 200 us per launch.  In 1D the accesses stay ordinary: a 1D
// request's state is a few hundred KB, and the march kernel reads what one
// pass wrote in the next, after a barrier, from L2.
template <int DIM>
__device__ __forceinline__ double stream_load(gcd q) {
  if constexpr (DIM >= 2) return __builtin_nontemporal_load(q);
  else return *q;
}
template <int DIM>
__device__ __forceinline__ void stream_store(gd q, double v) {
  if constexpr (DIM >= 2) __builtin_nontemporal_store(v, q);
  else *q = v;
}

// u0 receives the updated zeroth moment of eee (forward even half step), from
// which the block forms the intensity contributions.
template <int N, int DIM, int C, int MODE, int PAR>
__device__ __forceinline__ void update_class(const HalfStepArgs& a, gkd rec, double dt,
                                             gcd __restrict__ from, gcd src, gd dst, int i, int j, int k, long p,
                                             int beam, double* u0) {
  // src and dst are the same buffer for in-place steps: the own values are
  // read before they are overwritten, so neither pointer is restrict.
  constexpr int NC = ncls<N, C>();
  const Layout& L = a.L;
  const long V = L.V;
  gcd uc = src + L.cls_off[C];
  double u[NC], un[NC];
#pragma unroll
  for (int m = 0; m < NC; ++m) u[m] = stream_load<DIM>(uc + m * V + p);
  update_core<N, DIM, C, MODE>(a, rec, dt, from, u, un, i, j, k, p, beam);
  gd out = dst + L.cls_off[C];
#pragma unroll
  for (int m = 0; m < NC; ++m) stream_store<DIM>(out + m * V + p, un[m]);
  if constexpr (MODE == kForward && PAR == 0 && C == 0) *u0 = un[0];
}

// Two consecutive odd-parity updates of class C at point p: the even classes
// they read do not change in between, and the intermediate values stay in
// registers (stored to dst only when dst is not null).
template <int N, int DIM, int C, int MODE>
__device__ __forceinline__ void update_class2(const HalfStepArgs& a, const StepView& v, gcd __restrict__ from, gcd src,
                                              gd dst, gd dst2, int i, int j, int k, long p, int beam) {
  constexpr int NC = ncls<N, C>();
  constexpr bool SRC = MODE == kForward;
  const Layout& L = a.L;
  const long V = L.V;
  const int ne = L.n_e;
  gcd uc = src + L.cls_off[C];
  double u[NC], u1[NC], u2[NC], du[NC];
#pragma unroll
  for (int m = 0; m < NC; ++m) {
    u[m] = stream_load<DIM>(uc + m * V + p);
    du[m] = 0.0;
  }
  gkd rec = cst(v.rec);
  gkd rec2 = cst(v.rec2);
  const double gfac = rec[ne + ne * (N + 1)], gfac2 = rec2[ne + ne * (N + 1)];
  gcd __restrict__ prof_b = glb(a.prof) + beam * a.prof_stride;
  stencil<N, DIM, C, SRC>(L, from, cst(a.coef), u, du, i, j, k, p, gfac, prof_b, cst(a.mms));
  apply_update<N, C, MODE>(a, rec, v.dt, u, du, u1, p, beam);
  // the second update's stencil differs from the first only in its boundary ghosts
  stencil_ghost_delta<N, DIM, C, SRC>(L, cst(a.coef), u, u1, du, i, j, k, gfac, gfac2, prof_b, cst(a.mms));
  apply_update<N, C, MODE>(a, rec2, v.dt2, u1, du, u2, p, beam);
  if (dst) {
    gd out = dst + L.cls_off[C];
#pragma unroll
    for (int m = 0; m < NC; ++m) stream_store<DIM>(out + m * V + p, u1[m]);
  }
  gd out2 = dst2 + L.cls_off[C];
#pragma unroll
  for (int m = 0; m < NC; ++m) stream_store<DIM>(out2 + m * V + p, u2[m]);
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Intensity sums of one beam over the parts [t0, t1) of the per-step partial
// sums (integrate_step_and_add!, x_ray_intensities.jl:172-191): thread t sums
// the steps k = 1 + t, 1 + t + 256, ..., each over the parts, then the block
// reduces and writes out[r].  Every thread of the block must call it; in a
// block wider than kBlock only the first kBlock threads sum, so the result is
// the same for every block width.
__device__ __forceinline__ void reduce_intensity_body(const double* __restrict__ partial, long step_stride, int t0,
                                                      int t1, const double* __restrict__ rec, int rec_len,
                                                      int sigma_off, int n_steps, int n_lines, double dE, double dvol,
                                                      double* __restrict__ out) {
  __shared__ double red[kBlock / 64][kMaxLines];
  double acc[kMaxLines];
#pragma unroll
  for (int r = 0; r < kMaxLines; ++r) acc[r] = 0.0;
  const int n_sum = threadIdx.x < kBlock ? n_steps : 0;
  for (int kk = 1 + threadIdx.x; kk <= n_sum; kk += kBlock) {
    gcd __restrict__ ps = glb(partial) + kk * step_stride;
    gcd __restrict__ sig = glb(rec) + (long)kk * rec_len + sigma_off;
    for (int r = 0; r < n_lines && r < kMaxLines; ++r) {
      double integral = 0.0;
      for (int t = t0; t < t1; ++t) integral += ps[(long)t * n_lines + r];
      acc[r] += dE * dvol * sig[r] * integral * 1.0;
    }
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int r = 0; r < n_lines && r < kMaxLines; ++r) {
    const double s = wave_sum(acc[r]);
    if (lane == 0 && wave < kBlock / 64) red[wave][r] = s;
  }
  __syncthreads();
  if (threadIdx.x < n_lines && threadIdx.x < kMaxLines) {
    double s = 0.0;
    for (int w = 0; w < kBlock / 64; ++w) s += red[w][threadIdx.x];
    glb(out)[threadIdx.x] = s;
  }
  __syncthreads();
}

// Maps the linear block id to (beam, tile) so that each XCD (blocks are
// dealt round-robin over 8 XCDs) works on a contiguous range of tiles and the
// stencil's neighbour rows stay in that XCD's L2.
__device__ __forceinline__ void block_coords(int n_tiles, int n_beams, int& beam, int& tile) {
  const int total = n_tiles * n_beams;
  int logical = blockIdx.x;
  if ((total & 7) == 0) logical = (blockIdx.x & 7) * (total >> 3) + (blockIdx.x >> 3);
  beam = logical / n_tiles;
  tile = logical - beam * n_tiles;
}

// Work item of this block in a grouped kernel (one item per block; a
// persistent loop over items makes the compiler keep far more registers live).
// Blocks are dealt round-robin over the 8 XCDs; with a grid that is a
// multiple of 8, XCD x takes the contiguous items x*G/8 .. (x+1)*G/8-1, so
// the consecutive tiles of one request (and their shared neighbour rows) stay
// in one XCD's L2.
__device__ __forceinline__ int grouped_first() {
  const int G = gridDim.x;
  return (G & 7) == 0 ? (blockIdx.x & 7) * (G >> 3) + (blockIdx.x >> 3) : blockIdx.x;
}

// Index of this thread within its tile's kBlock threads.  Grid kernels have
// one tile per block of kBlock threads; in a WIDE block (the 1D march kernel)
// each group of kBlock threads takes a tile of its own.
template <bool WIDE>
__device__ __forceinline__ int tile_thread() { return WIDE ? (int)threadIdx.x & (kBlock - 1) : (int)threadIdx.x; }

// Cell (i, j, k) of this thread in tile `tile` (tile_x(DIM) x 256/tile_x(DIM)
// cells of one z-plane); false for threads past the grid's edge.  Tiles are
// numbered x-fastest, then z, then y: a row of tiles marches through the
// planes, so the neighbouring planes a 3D stencil reads were read by the
// tiles just before (with the planes outermost, 3D half steps fetched
// 1.5-1.75x their algorithmic bytes).  In 1D and 2D (Mz = 1) this is the
// plain x-then-y order.
template <int DIM, bool WIDE = false>
__device__ __forceinline__ bool tile_cell(const Layout& L, int tile, int& i, int& j, int& k) {
  constexpr int TX = tile_x(DIM), TY = kBlock / TX;
  const int tx = tile % L.tiles_x;
  const int r = tile / L.tiles_x;
  k = r % L.Mz;
  const int ty = r / L.Mz;
  i = tx * TX + tile_thread<WIDE>() % TX;
  j = ty * TY + tile_thread<WIDE>() / TX;
  return i < L.Mx && j < L.My;
}

template <int N, int DIM, int PAR, int MODE, bool WIDE = false>
__device__ __forceinline__ void half_step_body(const HalfStepArgs& a, const StepView& v, int beam, int tile) {
  const Layout& L = a.L;
  int i, j, k;
  const bool valid = tile_cell<DIM, WIDE>(L, tile, i, j, k);
  if (!valid) i = j = k = 0;
  const long p = i + (long)L.X * (j + (long)L.Y * k);
  const long bo = (long)beam * L.beam_stride;
  // eee's updated zeroth moment; the line contributions W u0 are formed in the
  // reduction, so only this value stays live across the other class's update
  double u0 = 0.0;
  if (valid) {
    static_for<0, 8>([&](auto Cc) {
      constexpr int C = decltype(Cc)::value;
      if constexpr (parity_of(C) == PAR && class_active(C, DIM) && ncls<N, C>() > 0) {
        const bool interior = !((odd_in(C, 0) && i == 0) || (DIM >= 2 && odd_in(C, 1) && j == 0) ||
                                (DIM >= 3 && odd_in(C, 2) && k == 0));
        if (interior)
          update_class<N, DIM, C, MODE, PAR>(a, cst(v.rec), v.dt, glb(v.from) + bo, glb(v.src) + bo,
                                             glb(v.dst) + bo, i, j, k, p, beam, &u0);
      }
    });
  }
  if constexpr (MODE == kForward && PAR == 0) {
    // one partial sum per wavefront and line (no barrier)
    const int lane = threadIdx.x & 63, wave = tile_thread<WIDE>() >> 6;
    gcd __restrict__ Wp = glb(a.W);
    gd __restrict__ out = glb(v.partial) + beam * a.partial_stride + ((long)tile * kWavesPerBlock + wave) * L.n_lines;
    const int nl = L.n_lines < kMaxLines ? L.n_lines : kMaxLines;
    for (int r0 = 0; r0 < nl; r0 += kBatch) {
      double wv[kBatch];
#pragma unroll
      for (int b = 0; b < kBatch; ++b) wv[b] = Wp[batch_index(r0, b, nl) * L.V + p];
#pragma unroll
      for (int b = 0; b < kBatch; ++b) {
        if (r0 + b < nl) {
          const double s = wave_sum(valid ? wv[b] * u0 : 0.0);
          if (lane == 0) out[r0 + b] = s;
        }
      }
    }
  }
}

template <int N, int DIM, int PAR, int MODE>
__global__ __launch_bounds__(kBlock) void half_step_kernel(HalfStepArgs a) {
  int beam, tile;
  block_coords(a.L.n_tiles, a.n_beams, beam, tile);
  half_step_body<N, DIM, PAR, MODE>(a, view_of(a), beam, tile);
}

// Form for heterogeneous batches: block -> one (argument block, tile) item of a
// work list, each argument block describing one request.
template <int N, int DIM, int PAR, int MODE>
__global__ __launch_bounds__(kBlock) void half_step_grouped(const HalfStepArgs* __restrict__ args,
                                                            const int2* __restrict__ items,
                                                            const int* __restrict__ n_items) {
  const int it = grouped_first();
  if (it >= *n_items) return;
  const int2 w = items[it];
  const HalfStepArgs& a = args[w.x];
  half_step_body<N, DIM, PAR, MODE>(a, view_of(a), 0, w.y);
}

// Fused pair of odd half steps: the last odd half step of energy step k
// (record v.rec, step v.dt, intermediate output v.dst or none) and the first
// of step k+1 (v.rec2, v.dt2, output v.dst2).
template <int N, int DIM, int MODE, bool WIDE = false>
__device__ __forceinline__ void half_step2_body(const HalfStepArgs& a, const StepView& v, int beam, int tile) {
  const Layout& L = a.L;
  int i, j, k;
  if (!tile_cell<DIM, WIDE>(L, tile, i, j, k)) return;
  const long p = i + (long)L.X * (j + (long)L.Y * k);
  const long bo = (long)beam * L.beam_stride;
  gd dst = v.dst ? glb(v.dst) + bo : nullptr;
  static_for<0, 8>([&](auto Cc) {
    constexpr int C = decltype(Cc)::value;
    if constexpr (parity_of(C) == 1 && class_active(C, DIM) && ncls<N, C>() > 0) {
      const bool interior = !((odd_in(C, 0) && i == 0) || (DIM >= 2 && odd_in(C, 1) && j == 0) ||
                              (DIM >= 3 && odd_in(C, 2) && k == 0));
      if (interior)
        update_class2<N, DIM, C, MODE>(a, v, glb(v.from) + bo, glb(v.src) + bo, dst, glb(v.dst2) + bo, i, j, k, p,
                                       beam);
    }
  });
}

template <int N, int DIM, int MODE>
__global__ __launch_bounds__(kBlock) void half_step2_kernel(HalfStepArgs a) {
  int beam, tile;
  block_coords(a.L.n_tiles, a.n_beams, beam, tile);
  half_step2_body<N, DIM, MODE>(a, view_of(a), beam, tile);
}

// integrate_adjoint_densities! (pn_solver/pn_intensities.jl:129-187) for one
// reverse step: du of every class without the exponential factor, then
// rho_bar += lambda (s du + c u) dvol dE int_fac, plus the x-ray generation
// pullback of integrate_step_and_add_pullback! at eee points.
// lam holds the adjoint values of class C at this point (registers).
// CELL: the density gradient of this class point is returned in cell_acc[e]
// (e < n_e <= kCellE) for the block to add per cell (DensArgs::rb_cell)
// instead of being added to the class's own rho_bar.
template <int N, int DIM, int C, bool CELL = false>
__device__ __forceinline__ void dens_core(const DensArgs& a, gkd rec, double int_fac,
                                          gcd __restrict__ u_b, const double* lam, int i, int j, int k, long p,
                                          int beam, double* cell_acc = nullptr) {
  constexpr int NC = ncls<N, C>();
  const Layout& L = a.L;
  const long V = L.V;
  const int ne = L.n_e;
  gcd __restrict__ uc = u_b + L.cls_off[C];
  double u[NC], du[NC];
#pragma unroll
  for (int m = 0; m < NC; ++m) {
    u[m] = uc[m * V + p];  // shared: the other parity's stencils read it too
    du[m] = 0.0;
  }
  const double gfac = rec[ne + ne * (N + 1)];
  stencil<N, DIM, C, true>(L, u_b, cst(a.coef), u, du, i, j, k, p, gfac, glb(a.prof) + beam * a.prof_stride,
                          cst(a.mms));
  double S, Cl[N + 1];
  mix<N>(glb(a.rho) + a.rho_off[C], V, p, ne, rec, S, Cl);
  const double invS = 1.0 / S;  // one division per point, as in apply_update
  double t[NC];
  static_for<0, NC>([&](auto Ic) {
    constexpr int I = decltype(Ic)::value;
    constexpr int l = Tab<N>::t.l[C][I];
    t[I] = ((-du[I]) - Cl[l] * u[I]) * invS;
  });
  gkd s = rec;
  gkd cl = rec + ne;
  // rho_bar_e += dvol dE int_fac sum_I lam_I (s_e t_I + c_{e,l(I)} u_I)
  //            = wfac (s_e X + sum_l c_{e,l} Y_l), with X = sum_I lam_I t_I and
  //              Y_l = sum of lam_I u_I over the moments I of degree l:
  // the sums over moments are formed once for all elements.
  double X = 0.0, Y[N + 1];
  static_for<0, N + 1>([&](auto lc) { Y[decltype(lc)::value] = 0.0; });
  static_for<0, NC>([&](auto Ic) {
    constexpr int I = decltype(Ic)::value;
    constexpr int l = Tab<N>::t.l[C][I];
    X += lam[I] * t[I];
    Y[l] += lam[I] * u[I];
  });
  const double wfac = a.L.dvol * a.dE * int_fac;
  auto grad_e = [&](int e) {
    double v = s[e] * X;
    static_for<0, N + 1>([&](auto lc) {
      constexpr int l = decltype(lc)::value;
      if constexpr (class_has_degree<N, C>(l)) v += cl[e * (N + 1) + l] * Y[l];
    });
    return v * wfac;
  };
  if constexpr (CELL) {
    static_assert(kCellE <= kBatch, "one batch of elements");
#pragma unroll
    for (int e = 0; e < kCellE; ++e) cell_acc[e] = e < ne ? grad_e(e) : 0.0;
  }
  gd __restrict__ rb = glb(a.rho_bar) + beam * a.rb_stride + a.rb_off[C];
  for (int e0 = 0; e0 < (CELL ? 0 : ne); e0 += kBatch) {
    double acc[kBatch];
#pragma unroll
    for (int b = 0; b < kBatch; ++b) acc[b] = stream_load<DIM>(rb + batch_index(e0, b, ne) * V + p);
#pragma unroll
    for (int b = 0; b < kBatch; ++b) {
      const int e = e0 + b;
      if (e < ne) stream_store<DIM>(rb + e * V + p, acc[b] + grad_e(e));
    }
  }
  if constexpr (C == 0) {
    gkd sig = rec + ne + ne * (N + 1) + 1;
    gkd ib = cst(a.ibar) + beam * L.n_lines;
    gd __restrict__ g = glb(a.gen) + beam * a.gen_stride;
    const int nl = L.n_lines;
    for (int r0 = 0; r0 < nl; r0 += kBatch) {
      double gv[kBatch];
#pragma unroll
      for (int b = 0; b < kBatch; ++b) gv[b] = stream_load<DIM>(g + batch_index(r0, b, nl) * V + p);
#pragma unroll
      for (int b = 0; b < kBatch; ++b) {
        const int r = r0 + b;
        if (r < nl) {
          const double integral = a.dE * a.L.dvol * sig[r] * int_fac * ib[r];
          stream_store<DIM>(g + r * V + p, gv[b] + u[0] * integral);
        }
      }
    }
  }
}

template <int N, int DIM, int C, bool CELL = false>
__device__ __forceinline__ void dens_class(const DensArgs& a, gkd rec, double int_fac,
                                           gcd __restrict__ u_b, gcd __restrict__ lam_b, int i, int j, int k, long p,
                                           int beam, double* cell_acc = nullptr) {
  constexpr int NC = ncls<N, C>();
  gcd __restrict__ lc = lam_b + a.L.cls_off[C];
  double lam[NC];
#pragma unroll
  for (int m = 0; m < NC; ++m) lam[m] = lc[m * a.L.V + p];
  dens_core<N, DIM, C, CELL>(a, rec, int_fac, u_b, lam, i, j, k, p, beam, cell_acc);
}

// Fused reverse pass of step i (reference adjoint, pn_intensities.jl:391-401):
// the last odd adjoint half step of step i + 1, the density pullback of step
// i and the first odd adjoint half step of step i in one sweep.  The odd half
// steps read only the even classes of lambda, which none of the three
// changes, at neighbouring points; the pullback reads lambda only at its own
// point.  So each thread carries its odd-class lambda through the three in
// registers, and the sweep reads and writes odd lambda once instead of three
// times.  Even classes only take the pullback.
//   a.adj:  from = to_src = to_dst = lambda; rec = adjoint record of step i, dt = dE/2
//   a.rec3: adjoint record of step i + 1 (null for i = n)
//   a.a1:   run the first odd half step of step i (0 for the final pullback at u_0)
//   a.dens: u = u_i, rec = density record of step i, int_fac
template <int N, int DIM, int C, bool CELL = false>
__device__ __forceinline__ void rev_class(const RevArgs& a, gcd __restrict__ lam_b, gcd __restrict__ u_b, int i, int j,
                                          int k, long p, int beam, double* cell_acc = nullptr) {
  constexpr int NC = ncls<N, C>();
  const Layout& L = a.adj.L;
  const long V = L.V;
  gcd lc = lam_b + L.cls_off[C];
  gd out = glb(a.adj.to_dst) + (long)beam * L.beam_stride + L.cls_off[C];
  double l0[NC], l1[NC];
#pragma unroll
  for (int m = 0; m < NC; ++m) l0[m] = stream_load<DIM>(lc + m * V + p);
  if constexpr (N < 9 || DIM == 2) {
    if (a.rec3 != nullptr) {
      // The pullback reads lambda after the last odd half step of step i + 1
      // (l0), and the first odd half step of step i does not read what the
      // pullback writes, so the two odd half steps run back to back as a
      // fused pair (one stencil; the second adds only the change of the
      // boundary ghosts), lambda is stored, and the pullback follows with l0.
      // (2D P9: 226 VGPRs against 208 for the order below, both occupancy 2,
      // and 6 % faster; in 1D and 3D P9 keeps the order below.)
      double lin[NC], du[NC];
#pragma unroll
      for (int m = 0; m < NC; ++m) {
        lin[m] = l0[m];
        du[m] = 0.0;
      }
      stencil<N, DIM, C, false>(L, lam_b, cst(a.adj.coef), lin, du, i, j, k, p, 0.0, nullptr, nullptr);
      apply_update<N, C, kAdjoint>(a.adj, cst(a.rec3), a.adj.dt, lin, du, l0, p, beam);
      if (a.a1) {
        stencil_ghost_delta<N, DIM, C, false>(L, cst(a.adj.coef), lin, l0, du, i, j, k, 0.0, 0.0, nullptr, nullptr);
        apply_update<N, C, kAdjoint>(a.adj, cst(a.adj.rec), a.adj.dt, l0, du, l1, p, beam);
      } else {
#pragma unroll
        for (int m = 0; m < NC; ++m) l1[m] = l0[m];
      }
#pragma unroll
      for (int m = 0; m < NC; ++m) stream_store<DIM>(out + m * V + p, l1[m]);
      dens_core<N, DIM, C, CELL>(a.dens, cst(a.dens.rec), a.dens.int_fac, u_b, l0, i, j, k, p, beam, cell_acc);
      return;
    }
  }
  if (a.rec3 != nullptr) {
    update_core<N, DIM, C, kAdjoint>(a.adj, cst(a.rec3), a.adj.dt, lam_b, l0, l1, i, j, k, p, beam);
#pragma unroll
    for (int m = 0; m < NC; ++m) l0[m] = l1[m];
  }
  dens_core<N, DIM, C, CELL>(a.dens, cst(a.dens.rec), a.dens.int_fac, u_b, l0, i, j, k, p, beam, cell_acc);
  if (a.a1) {
    update_core<N, DIM, C, kAdjoint>(a.adj, cst(a.adj.rec), a.adj.dt, lam_b, l0, l1, i, j, k, p, beam);
  } else {
#pragma unroll
    for (int m = 0; m < NC; ++m) l1[m] = l0[m];
  }
#pragma unroll
  for (int m = 0; m < NC; ++m) stream_store<DIM>(out + m * V + p, l1[m]);
}

template <int N, int DIM, bool WIDE = false>
__device__ __forceinline__ void rev_body(const RevArgs& a, int beam, int tile) {
  const Layout& L = a.adj.L;
  int i, j, k;
  if (!tile_cell<DIM, WIDE>(L, tile, i, j, k)) return;
  const long p = i + (long)L.X * (j + (long)L.Y * k);
  const long bo = (long)beam * L.beam_stride;
  gcd lam_b = glb(a.adj.from) + bo;
  gcd u_b = glb(a.dens.u) + bo;
  static_for<0, 8>([&](auto Cc) {
    constexpr int C = decltype(Cc)::value;
    if constexpr (class_active(C, DIM) && ncls<N, C>() > 0) {
      const bool interior = !((odd_in(C, 0) && i == 0) || (DIM >= 2 && odd_in(C, 1) && j == 0) ||
                              (DIM >= 3 && odd_in(C, 2) && k == 0));
      if (interior) {
        if constexpr (parity_of(C) == 1)
          rev_class<N, DIM, C>(a, lam_b, u_b, i, j, k, p, beam);
        else
          dens_class<N, DIM, C>(a.dens, cst(a.dens.rec), a.dens.int_fac, u_b, lam_b, i, j, k, p, beam);
      }
    }
  });
}

// 2D form with one class per wavefront.  The plane is cut into kCwTileX x
// kCwTileX cw tiles (x-fastest), each into four 64-cell sub-tiles of
// 64 / kCwTileX rows; block `pair` takes sub-tiles 2 (pair & 1) and
// 2 (pair & 1) + 1 of cw tile pair >> 1.  A wavefront runs one class's code at
// a time, so the kernel needs the registers of one class instead of all four,
// which raises the occupancy of these register-heavy passes, and it reads
// whole row segments of kCwTileX doubles per moment.  The odd classes (eoe,
// oee) take two adjoint half steps and their pullback, the even classes
// (eee, ooe) the pullback only, so each wavefront takes one job of each kind:
// wavefronts 0 and 1 the odd class eoe / oee at the first sub-tile and the
// even class eee / ooe at the second, wavefronts 2 and 3 the same classes at
// the other sub-tiles.  (With one class per wavefront at one sub-tile the even
// wavefronts finished early while the block held its place until the odd ones
// were done.)  The jobs of a block touch disjoint data.
template <int N, class F>
__device__ __forceinline__ void class_pairs_2d(const Layout& L, int pair, F&& f) {
  constexpr int SUB_ROWS = 64 / kCwTileX;
  static_assert(64 % kCwTileX == 0 && kCwTileX == 4 * SUB_ROWS, "a sub-tile is 64 cells, a cw tile four");
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int tile = pair >> 1, s0 = 2 * (pair & 1);
  const int tx = tile % L.cw_tiles_x, ty = tile / L.cw_tiles_x;
  const int i = tx * kCwTileX + lane % kCwTileX;
  const int j0 = ty * kCwTileX + lane / kCwTileX;
  const int s_odd = s0 + (wave >> 1), s_even = s0 + 1 - (wave >> 1);
  auto job = [&](auto Cc, int sub) {
    constexpr int C = decltype(Cc)::value;
    if constexpr (ncls<N, C>() > 0) {
      const int j = j0 + sub * SUB_ROWS;
      if (i >= L.Mx || j >= L.My) return;
      const bool interior = !((odd_in(C, 0) && i == 0) || (odd_in(C, 1) && j == 0));
      if (interior) f(Cc, i, j, i + (long)L.X * j, sub - s0);
    }
  };
  if ((wave & 1) == 0) {
    job(std::integral_constant<int, 2>{}, s_odd);
    job(std::integral_constant<int, 0>{}, s_even);
  } else {
    job(std::integral_constant<int, 4>{}, s_odd);
    job(std::integral_constant<int, 6>{}, s_even);
  }
}

// With a.dens.rb_cell each job returns its class's density gradient per
// element; the wavefronts leave them in LDS (slot C / 2: eee, eoe, oee, ooe),
// and after a block barrier the first two wavefronts add the four classes of
// each cell, in that order, to the cell's rho_bar with one read and write.
template <int N, bool CELL>
__device__ __forceinline__ void rev_body_cw_impl(const RevArgs& a, int beam, int pair) {
  const Layout& L = a.adj.L;
  const long bo = (long)beam * L.beam_stride;
  gcd lam_b = glb(a.adj.from) + bo;
  gcd u_b = glb(a.dens.u) + bo;
  if constexpr (!CELL) {
    class_pairs_2d<N>(L, pair, [&](auto Cc, int i, int j, long p, int) {
      constexpr int C = decltype(Cc)::value;
      if constexpr (parity_of(C) == 1)
        rev_class<N, 2, C>(a, lam_b, u_b, i, j, 0, p, beam);
      else
        dens_class<N, 2, C>(a.dens, cst(a.dens.rec), a.dens.int_fac, u_b, lam_b, i, j, 0, p, beam);
    });
  } else {
    __shared__ double part[4][2][kCellE][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    // this wavefront's two slots start at zero (jobs outside the grid or on
    // ghost points add nothing)
    const int s_odd = wave >> 1, c_odd = (wave & 1) ? 2 : 1, c_even = (wave & 1) ? 3 : 0;
#pragma unroll
    for (int e = 0; e < kCellE; ++e) {
      part[c_odd][s_odd][e][lane] = 0.0;
      part[c_even][1 - s_odd][e][lane] = 0.0;
    }
    class_pairs_2d<N>(L, pair, [&](auto Cc, int i, int j, long p, int sub) {
      constexpr int C = decltype(Cc)::value;
      double acc[kCellE];
      if constexpr (parity_of(C) == 1)
        rev_class<N, 2, C, true>(a, lam_b, u_b, i, j, 0, p, beam, acc);
      else
        dens_class<N, 2, C, true>(a.dens, cst(a.dens.rec), a.dens.int_fac, u_b, lam_b, i, j, 0, p, beam, acc);
#pragma unroll
      for (int e = 0; e < kCellE; ++e) part[C / 2][sub][e][lane] = acc[e];
    });
    __syncthreads();
    if (threadIdx.x < 128) {
      constexpr int SUB_ROWS = 64 / kCwTileX;
      const int sub = threadIdx.x >> 6;
      const int tile = pair >> 1, s = 2 * (pair & 1) + sub;
      const int i = (tile % L.cw_tiles_x) * kCwTileX + lane % kCwTileX;
      const int j = (tile / L.cw_tiles_x) * kCwTileX + s * SUB_ROWS + lane / kCwTileX;
      if (i < L.Mx && j < L.My) {
        const long p = i + (long)L.X * j;
        gd __restrict__ rb = glb(a.dens.rho_bar) + beam * a.dens.rb_stride;
        const int ne = L.n_e;
#pragma unroll
        for (int e = 0; e < kCellE; ++e) {
          if (e < ne) {
            double v = stream_load<2>(rb + e * L.V + p);
            v += part[0][sub][e][lane];
            v += part[1][sub][e][lane];
            v += part[2][sub][e][lane];
            v += part[3][sub][e][lane];
            stream_store<2>(rb + e * L.V + p, v);
          }
        }
      }
    }
  }
}

template <int N>
__device__ __forceinline__ void rev_body_cw(const RevArgs& a, int beam, int pair) {
  if (a.dens.rb_cell)
    rev_body_cw_impl<N, true>(a, beam, pair);
  else
    rev_body_cw_impl<N, false>(a, beam, pair);
}

// Unbounded, the fused reverse pass at P7 lands two registers above the
// three-waves-per-SIMD limit (170 of 168); the bound keeps it at three.
constexpr int rev_cw_min_blocks(int N) { return N <= 7 ? 3 : 1; }

template <int N>
__global__ __launch_bounds__(kBlock, rev_cw_min_blocks(N)) void rev_kernel_cw(RevArgs a) {
  int beam, pair;
  block_coords(a.adj.L.n_cw_items / 2, a.adj.n_beams, beam, pair);
  rev_body_cw<N>(a, beam, pair);
}

// 3D form of the fused reverse pass with one class per wavefront.  Sub-tiles
// are the 64-cell sub-tiles of the 2D form in every z-plane (sub-tile
// item & 3 of cw tile item >> 2, cw tiles x-fastest, then planes, then rows of
// cw tiles, for the reuse of neighbouring planes as in tile_cell).  Block
// item2 takes half h = item2 & 1 of the classes (h = 0: eee, eeo, eoe, eoo;
// h = 1: oee, oeo, ooe, ooo; each half two odd and two even classes) at the
// two sub-tiles 2 (q & 1) and 2 (q & 1) + 1 of cw tile and plane q >> 1,
// q = item2 >> 1: n_cw_items blocks per beam.  As in 2D, wavefront w takes one
// odd-class job at one sub-tile and one even-class job at the other, so the
// four wavefronts finish together.  One thread per cell over all eight classes
// needed 179 VGPRs at P5 (occupancy 2).
template <int N>
__device__ __forceinline__ void rev_body_cw3(const RevArgs& a, int beam, int item2) {
  const Layout& L = a.adj.L;
  const int half = item2 & 1, q = item2 >> 1;
  constexpr int SUB_ROWS = 64 / kCwTileX;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int s0 = 2 * (q & 1), r = q >> 1;
  const int tx = r % L.cw_tiles_x, k = (r / L.cw_tiles_x) % L.Mz, ty = r / L.cw_tiles_x / L.Mz;
  const int i = tx * kCwTileX + lane % kCwTileX;
  const int j0 = ty * kCwTileX + lane / kCwTileX;
  const int s_odd = s0 + (wave >> 1), s_even = s0 + 1 - (wave >> 1);
  const long bo = (long)beam * L.beam_stride;
  gcd lam_b = glb(a.adj.from) + bo;
  gcd u_b = glb(a.dens.u) + bo;
  auto job = [&](auto Cc, int sub) {
    constexpr int C = decltype(Cc)::value;
    if constexpr (ncls<N, C>() > 0) {
      const int j = j0 + sub * SUB_ROWS;
      if (i >= L.Mx || j >= L.My || k >= L.Mz) return;
      const bool interior = !((odd_in(C, 0) && i == 0) || (odd_in(C, 1) && j == 0) || (odd_in(C, 2) && k == 0));
      if (!interior) return;
      const long p = i + (long)L.X * (j + (long)L.Y * k);
      if constexpr (parity_of(C) == 1)
        rev_class<N, 3, C>(a, lam_b, u_b, i, j, k, p, beam);
      else
        dens_class<N, 3, C>(a.dens, cst(a.dens.rec), a.dens.int_fac, u_b, lam_b, i, j, k, p, beam);
    }
  };
  using std::integral_constant;
  // odd classes: eeo (1), eoe (2) | oee (4), ooo (7); even: eee (0), eoo (3) | oeo (5), ooe (6)
  if (half == 0) {
    if ((wave & 1) == 0) {
      job(integral_constant<int, 1>{}, s_odd);
      job(integral_constant<int, 0>{}, s_even);
    } else {
      job(integral_constant<int, 2>{}, s_odd);
      job(integral_constant<int, 3>{}, s_even);
    }
  } else {
    if ((wave & 1) == 0) {
      job(integral_constant<int, 4>{}, s_odd);
      job(integral_constant<int, 5>{}, s_even);
    } else {
      job(integral_constant<int, 7>{}, s_odd);
      job(integral_constant<int, 6>{}, s_even);
    }
  }
}

template <int N>
__global__ __launch_bounds__(kBlock) void rev_kernel_cw3(RevArgs a) {
  int beam, item2;
  block_coords(a.adj.L.n_cw_items, a.adj.n_beams, beam, item2);
  rev_body_cw3<N>(a, beam, item2);
}

// Phase kernels of the grid lanes: the kinds of one phase share a launch, so
// a sub-tick is two launches over all its requests instead of one per kind.
// The registers are those of the phase's largest kind.
template <int N, int DIM>
__global__ __launch_bounds__(kBlock, DIM == 2 ? rev_cw_min_blocks(N) : 1) void phase_a_grouped(
    const HalfStepArgs* __restrict__ half, const RevArgs* __restrict__ rev, const int2* __restrict__ items,
    const int* __restrict__ n_items) {
  const int it = grouped_first();
  if (it >= *n_items) return;
  const int2 w = items[it];
  const int kind = w.x >> kItemKindShift, arg = w.x & ((1 << kItemKindShift) - 1);
  if (kind == kItemRev) {
    if constexpr (DIM == 2)
      rev_body_cw<N>(rev[arg], 0, w.y);
    else if constexpr (DIM == 3)
      rev_body_cw3<N>(rev[arg], 0, w.y);
    else
      rev_body<N, DIM>(rev[arg], 0, w.y);
  } else {
    const HalfStepArgs& a = half[arg];
    half_step_body<N, DIM, 0, kForward>(a, view_of(a), 0, w.y);
  }
}

// The fused odd pair has its own launch: in one kernel with the two kinds of
// phase B the compiler allocated more registers than any of them needs (2D
// P7: 209 VGPRs, P9: 255 and occupancy 1; bounded, it spilled).  At 2D P9 it
// needs 175 VGPRs, 7 over three waves per SIMD; bounded to three it keeps 36
// bytes per lane in scratch and the P9 lane serves 0.8 % faster.
constexpr int pair_min_blocks(int N, int DIM) { return N == 9 && DIM == 2 ? 3 : 1; }

template <int N, int DIM>
__global__ __launch_bounds__(kBlock, pair_min_blocks(N, DIM)) void pair_grouped(const HalfStepArgs* __restrict__ half,
                                                       const int2* __restrict__ items,
                                                       const int* __restrict__ n_items) {
  const int it = grouped_first();
  if (it >= *n_items) return;
  const int2 w = items[it];
  const HalfStepArgs& a = half[w.x & ((1 << kItemKindShift) - 1)];
  half_step2_body<N, DIM, kForward>(a, view_of(a), 0, w.y);
}

template <int N, int DIM>
__global__ __launch_bounds__(kBlock) void phase_b_grouped(const HalfStepArgs* __restrict__ half,
                                                          const int2* __restrict__ items,
                                                          const int* __restrict__ n_items) {
  const int it = grouped_first();
  if (it >= *n_items) return;
  const int2 w = items[it];
  const int kind = w.x >> kItemKindShift, arg = w.x & ((1 << kItemKindShift) - 1);
  const HalfStepArgs& a = half[arg];
  if (kind == kItemFwdOdd)
    half_step_body<N, DIM, 1, kForward>(a, view_of(a), 0, w.y);
  else
    half_step_body<N, DIM, 0, kAdjoint>(a, view_of(a), 0, w.y);
}

template <int N, int DIM>
__global__ __launch_bounds__(kBlock) void rev_kernel(RevArgs a) {
  int beam, tile;
  block_coords(a.adj.L.n_tiles, a.adj.n_beams, beam, tile);
  rev_body<N, DIM>(a, beam, tile);
}

// ----------------------------------------------------------- exact adjoint
__device__ __forceinline__ double dexpm1div(double x) {
  // derivative of expm1div: Taylor branch as the forward, else (e^x (x - 1) + 1) / x^2
  // (series below |x| < 0.05 to avoid the cancellation)
  const double ax = fabs(x);
  if (ax <= 0.0002) return 0.5 + x / 3.0;
  if (ax < 0.05) {
    double term = 1.0, sum = 0.0, fact = 1.0, xp = 1.0;
#pragma unroll
    for (int n = 1; n < 12; ++n) {
      fact *= (n + 1);
      sum += n * xp / fact;
      xp *= x;
    }
    (void)term;
    return sum;
  }
  return (exp(x) * (x - 1.0) + 1.0) / (x * x);
}

// acc[NF] += sign/h * A_D[C]^T w  for one neighbour point of class C
template <int N, int C, int D>
__device__ __forceinline__ void at_w(gkd coef, gcd __restrict__ wc, long V, long q,
                                     double scale, double* acc) {
  constexpr int NC = ncls<N, C>();
  constexpr int F = switch_axis(C, D);
  constexpr int NF = ncls<N, F>();
  double w[NC];
#pragma unroll
  for (int I = 0; I < NC; ++I) w[I] = wc[I * V + q];
  static_for<0, NF>([&](auto Jc) {
    constexpr int J = decltype(Jc)::value;
    double s = 0.0;
    static_for<0, NC>([&](auto Ic) {
      constexpr int I = decltype(Ic)::value;
      if constexpr (Tab<N>::t.nz[C][D][I][J] >= 0) s += Tab<N>::t.a[C][D][I][J] * w[I];
    });
    acc[J] += scale * s;
  });
}

// Transposed stencil into lam of class F at point p from the w of the classes
// that read F (switch_D(F) along every non-singleton axis D).
template <int N, int DIM, int F>
__device__ __forceinline__ void disc_b(const DiscArgs& a, int beam, int i, int j, int k, long p) {
  constexpr int NF = ncls<N, F>();
  const Layout& L = a.L;
  const long V = L.V;
  const long bo = (long)beam * L.beam_stride;
  double acc[NF];
#pragma unroll
  for (int J = 0; J < NF; ++J) acc[J] = 0.0;
  static_for<0, DIM>([&](auto Dc) {
    constexpr int D = decltype(Dc)::value;
    constexpr int C = switch_axis(F, D);
    if constexpr (ncls<N, C>() > 0) {
      const long st = D == 0 ? 1L : (D == 1 ? (long)L.X : (long)L.X * L.Y);
      const int iq = D == 0 ? i : (D == 1 ? j : k);
      const int M = D == 0 ? L.Mx : (D == 1 ? L.My : L.Mz);
      const double ih = L.inv_h[D];
      gcd __restrict__ wc = glb(a.w_in) + bo + L.cls_off[C];
      if constexpr (odd_in(C, D)) {  // F even along D: read by C at iq (as R) and iq+1 (as L)
        if (iq >= 1) at_w<N, C, D>(cst(a.coef), wc, V, p, -ih, acc);
        if (iq + 1 <= M - 1) at_w<N, C, D>(cst(a.coef), wc, V, p + st, ih, acc);
      } else {  // F odd along D (iq in 1..M-1): read by C at iq (as L) and iq-1 (as R); ghosts double
        const double cl = iq == M - 1 ? 2.0 : 1.0, cr = iq == 1 ? 2.0 : 1.0;
        at_w<N, C, D>(cst(a.coef), wc, V, p, cl * ih, acc);
        at_w<N, C, D>(cst(a.coef), wc, V, p - st, -cr * ih, acc);
      }
    }
  });
  gd __restrict__ lf = glb(a.lam) + bo + L.cls_off[F];
#pragma unroll
  for (int J = 0; J < NF; ++J) lf[J * V + p] += acc[J];
}

// Back-propagation of the half step of class C at point p.
template <int N, int DIM, int C>
__device__ __forceinline__ void disc_a(const DiscArgs& a, int beam, int i, int j, int k, long p) {
  constexpr int NC = ncls<N, C>();
  const Layout& L = a.L;
  const long V = L.V;
  const int ne = L.n_e;
  const long bo = (long)beam * L.beam_stride;
  gcd __restrict__ uc = glb(a.s_own) + bo + L.cls_off[C];
  double u[NC], d[NC];
#pragma unroll
  for (int m = 0; m < NC; ++m) {
    u[m] = uc[m * V + p];
    d[m] = 0.0;
  }
  gkd rec = cst(a.rec);
  const double gfac = rec[ne + ne * (N + 1)];
  stencil<N, DIM, C, true>(L, glb(a.s_from) + bo, cst(a.coef), u, d, i, j, k, p, gfac,
                          glb(a.prof) + beam * a.prof_stride, cst(a.mms));
  double S, Cl[N + 1];
  mix<N>(glb(a.rho) + a.rho_off[C], V, p, ne, rec, S, Cl);
  gd lc = glb(a.lam) + bo + L.cls_off[C];
  double lam[NC], w[NC], dS[NC], dC[NC];
  const double tau = a.tau;
  static_for<0, NC>([&](auto Ic) {
    constexpr int I = decltype(Ic)::value;
    constexpr int l = Tab<N>::t.l[C][I];
    const double x = -Cl[l] / S * tau;
    const double ph = expm1div(x), dph = dexpm1div(x);
    const double r = (-d[I]) - Cl[l] * u[I];
    const double yb = lc[I * V + p];
    w[I] = yb * tau * ph / S;
    dS[I] = yb * (tau * r / (S * S)) * (-x * dph - ph);
    dC[I] = yb * (-(tau / S) * (dph * tau * r / S + ph * u[I]));
    lam[I] = yb - w[I] * Cl[l];
  });
  // own values reach the stencil through the boundary ghosts of the faces
  static_for<0, DIM>([&](auto Dc) {
    constexpr int D = decltype(Dc)::value;
    constexpr int F = switch_axis(C, D);
    constexpr int NF = ncls<N, F>();
    if constexpr (!odd_in(C, D) && NF > 0) {
      const int idx = D == 0 ? i : (D == 1 ? j : k);
      const int M = D == 0 ? L.Mx : (D == 1 ? L.My : L.Mz);
      if (idx == 0 || idx == M - 1) {
        double at[NF];
#pragma unroll
        for (int J = 0; J < NF; ++J) at[J] = 0.0;
        static_for<0, NF>([&](auto Jc) {
          constexpr int J = decltype(Jc)::value;
          static_for<0, NC>([&](auto Ic) {
            constexpr int I = decltype(Ic)::value;
            if constexpr (Tab<N>::t.nz[C][D][I][J] >= 0) at[J] += Tab<N>::t.a[C][D][I][J] * w[I];
          });
        });
        constexpr int boff = Tab<N>::t.b_off[F][D];
        const double s2 = 2.0 * L.inv_h[D];
#pragma unroll
        for (int I = 0; I < NC; ++I) {
          double bt = 0.0;
#pragma unroll
          for (int J = 0; J < NF; ++J) bt += cst(a.coef)[boff + J * NC + I] * at[J];
          lam[I] -= s2 * bt;
        }
      }
    }
  });
  gd __restrict__ wo = glb(a.w_out) + bo + L.cls_off[C];
#pragma unroll
  for (int I = 0; I < NC; ++I) {
    lc[I * V + p] = lam[I];
    wo[I * V + p] = w[I];
  }
  gd __restrict__ rb = glb(a.rho_bar) + beam * a.rb_stride + a.rho_off[C];
  for (int e = 0; e < ne; ++e) {
    const double se = rec[e];
    gkd ce = rec + ne + e * (N + 1);
    double acc = 0.0;
    static_for<0, NC>([&](auto Ic) {
      constexpr int I = decltype(Ic)::value;
      constexpr int l = Tab<N>::t.l[C][I];
      acc += dS[I] * se + dC[I] * ce[l];
    });
    rb[e * V + p] += acc;
  }
}

template <int N, int DIM, int PB, int PA>
__device__ __forceinline__ void disc_body(const DiscArgs& a, int beam, int tile) {
  const Layout& L = a.L;
  int i, j, k;
  if (!tile_cell<DIM>(L, tile, i, j, k)) return;
  const long p = i + (long)L.X * (j + (long)L.Y * k);
  static_for<0, 8>([&](auto Cc) {
    constexpr int C = decltype(Cc)::value;
    if constexpr (PB >= 0 && parity_of(C) == PB && class_active(C, DIM) && ncls<N, C>() > 0) {
      const bool interior = !((odd_in(C, 0) && i == 0) || (DIM >= 2 && odd_in(C, 1) && j == 0) ||
                              (DIM >= 3 && odd_in(C, 2) && k == 0));
      if (interior) disc_b<N, DIM, C>(a, beam, i, j, k, p);
    }
  });
  if (a.u_direct != nullptr) {
    // d(sum_r I_bar_r I_r)/du_eee(p,0) of this step and the line-weight sensitivity
    const long V = L.V;
    const long bo = (long)beam * L.beam_stride;
    const int ne = L.n_e;
    gkd sig = cst(a.rec) + ne + ne * (N + 1) + 1;
    gkd ib = cst(a.ibar) + beam * L.n_lines;
    const double u0 = glb(a.u_direct)[bo + L.cls_off[0] + p];
    double add = 0.0;
    gd __restrict__ g = glb(a.gen) + beam * a.gen_stride;
    for (int rr = 0; rr < L.n_lines; ++rr) {
      const double wk = a.dE * L.dvol * sig[rr] * ib[rr];
      add += wk * glb(a.W)[rr * V + p];
      g[rr * V + p] += u0 * wk;
    }
    glb(a.lam)[bo + L.cls_off[0] + p] += add;
  }
  static_for<0, 8>([&](auto Cc) {
    constexpr int C = decltype(Cc)::value;
    if constexpr (PA >= 0 && parity_of(C) == PA && class_active(C, DIM) && ncls<N, C>() > 0) {
      const bool interior = !((odd_in(C, 0) && i == 0) || (DIM >= 2 && odd_in(C, 1) && j == 0) ||
                              (DIM >= 3 && odd_in(C, 2) && k == 0));
      if (interior) disc_a<N, DIM, C>(a, beam, i, j, k, p);
    }
  });
}

template <int N, int DIM, int PB, int PA>
__global__ __launch_bounds__(kBlock) void disc_kernel(DiscArgs a) {
  int beam, tile;
  block_coords(a.L.n_tiles, a.n_beams, beam, tile);
  disc_body<N, DIM, PB, PA>(a, beam, tile);
}

template <int N>
hipError_t disc_order(int dims, int pb, int pa, const DiscArgs& a, hipStream_t s) {
  const dim3 grid(a.L.n_tiles * a.n_beams), block(kBlock);
  // the three passes used by the sweep: (even, odd), (even, even), (odd, odd)
  const int key = dims * 16 + (pb + 1) * 4 + (pa + 1);
  switch (key) {
#define PNX_DCASE(D, PB_, PA_) \
  case D * 16 + (PB_ + 1) * 4 + (PA_ + 1): hipLaunchKernelGGL((disc_kernel<N, D, PB_, PA_>), grid, block, 0, s, a); break;
    PNX_DCASE(1, 0, 1) PNX_DCASE(1, 0, 0) PNX_DCASE(1, 1, 1)
    PNX_DCASE(2, 0, 1) PNX_DCASE(2, 0, 0) PNX_DCASE(2, 1, 1)
    PNX_DCASE(3, 0, 1) PNX_DCASE(3, 0, 0) PNX_DCASE(3, 1, 1)
#undef PNX_DCASE
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int N>
hipError_t half_step_order(int dims, int parity, int mode, const HalfStepArgs& a, hipStream_t s) {
  const dim3 grid(a.L.n_tiles * a.n_beams), block(kBlock);
  switch (dims * 4 + parity * 2 + mode) {
#define PNX_CASE(D, P, M) \
  case D * 4 + P * 2 + M: hipLaunchKernelGGL((half_step_kernel<N, D, P, M>), grid, block, 0, s, a); break;
    PNX_CASE(1, 0, 0) PNX_CASE(1, 0, 1) PNX_CASE(1, 1, 0) PNX_CASE(1, 1, 1)
    PNX_CASE(2, 0, 0) PNX_CASE(2, 0, 1) PNX_CASE(2, 1, 0) PNX_CASE(2, 1, 1)
    PNX_CASE(3, 0, 0) PNX_CASE(3, 0, 1) PNX_CASE(3, 1, 0) PNX_CASE(3, 1, 1)
#undef PNX_CASE
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int N>
hipError_t half_step_grouped_order(int dims, int parity, int mode, const HalfStepArgs* args, const int2* items,
                                   const int* n_items, int grid, hipStream_t s) {
  switch (dims * 4 + parity * 2 + mode) {
#define PNX_CASE(D, P, M)                                                                                  \
  case D * 4 + P * 2 + M:                                                                                  \
    hipLaunchKernelGGL((half_step_grouped<N, D, P, M>), dim3(grid), dim3(kBlock), 0, s, args, items, n_items); \
    break;
    PNX_CASE(1, 0, 0) PNX_CASE(1, 0, 1) PNX_CASE(1, 1, 0) PNX_CASE(1, 1, 1)
    PNX_CASE(2, 0, 0) PNX_CASE(2, 0, 1) PNX_CASE(2, 1, 0) PNX_CASE(2, 1, 1)
    PNX_CASE(3, 0, 0) PNX_CASE(3, 0, 1) PNX_CASE(3, 1, 0) PNX_CASE(3, 1, 1)
#undef PNX_CASE
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int N>
hipError_t half_step2_order(int dims, int mode, const HalfStepArgs& a, hipStream_t s) {
  const dim3 grid(a.L.n_tiles * a.n_beams), block(kBlock);
  switch (dims * 2 + mode) {
#define PNX_CASE(D, M) \
  case D * 2 + M: hipLaunchKernelGGL((half_step2_kernel<N, D, M>), grid, block, 0, s, a); break;
    PNX_CASE(1, 0) PNX_CASE(1, 1) PNX_CASE(2, 0) PNX_CASE(2, 1) PNX_CASE(3, 0) PNX_CASE(3, 1)
#undef PNX_CASE
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int N>
hipError_t rev_order(int dims, const RevArgs& a, hipStream_t s) {
  const dim3 grid(a.adj.L.n_tiles * a.adj.n_beams), block(kBlock);
  switch (dims) {
    case 1: hipLaunchKernelGGL((rev_kernel<N, 1>), grid, block, 0, s, a); break;
    case 2: hipLaunchKernelGGL((rev_kernel_cw<N>), dim3(a.adj.L.n_cw_items / 2 * a.adj.n_beams), block, 0, s, a); break;
    case 3:
      hipLaunchKernelGGL((rev_kernel_cw3<N>), dim3(a.adj.L.n_cw_items * a.adj.n_beams), block, 0, s, a);
      break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int N>
hipError_t phase_a_grouped_order(int dims, const HalfStepArgs* half, const RevArgs* rev, const int2* items,
                                 const int* n_items, int grid, hipStream_t s) {
  switch (dims) {
    case 1: hipLaunchKernelGGL((phase_a_grouped<N, 1>), dim3(grid), dim3(kBlock), 0, s, half, rev, items, n_items); break;
    case 2: hipLaunchKernelGGL((phase_a_grouped<N, 2>), dim3(grid), dim3(kBlock), 0, s, half, rev, items, n_items); break;
    case 3: hipLaunchKernelGGL((phase_a_grouped<N, 3>), dim3(grid), dim3(kBlock), 0, s, half, rev, items, n_items); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int N>
hipError_t pair_grouped_order(int dims, const HalfStepArgs* half, const int2* items, const int* n_items, int grid,
                              hipStream_t s) {
  switch (dims) {
    case 1: hipLaunchKernelGGL((pair_grouped<N, 1>), dim3(grid), dim3(kBlock), 0, s, half, items, n_items); break;
    case 2: hipLaunchKernelGGL((pair_grouped<N, 2>), dim3(grid), dim3(kBlock), 0, s, half, items, n_items); break;
    case 3: hipLaunchKernelGGL((pair_grouped<N, 3>), dim3(grid), dim3(kBlock), 0, s, half, items, n_items); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int N>
hipError_t phase_b_grouped_order(int dims, const HalfStepArgs* half, const int2* items, const int* n_items, int grid,
                                 hipStream_t s) {
  switch (dims) {
    case 1: hipLaunchKernelGGL((phase_b_grouped<N, 1>), dim3(grid), dim3(kBlock), 0, s, half, items, n_items); break;
    case 2: hipLaunchKernelGGL((phase_b_grouped<N, 2>), dim3(grid), dim3(kBlock), 0, s, half, items, n_items); break;
    case 3: hipLaunchKernelGGL((phase_b_grouped<N, 3>), dim3(grid), dim3(kBlock), 0, s, half, items, n_items); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace pnx
