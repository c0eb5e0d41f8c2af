// The energy-step sequence of one forward+adjoint solve with checkpoints, as a
// resumable state machine shared by the host scheduler and the device march
// kernel, and the argument block of the march kernel.
//
// The sequence is the reference pullback's (pn_solver/pn_intensities.jl:353-404)
// with the tape replaced by checkpoints every K steps:
//   forward k = 1..n (checkpoint before steps 1, K+1, 2K+1, ...; intensities
//   after step n), then for every segment from the last to the first:
//   recompute its states from the checkpoint, and run the reverse steps
//   i = k1..k0+1 (density pullback on the recomputed u_i, then one adjoint
//   step); finally the density pullback at u_0 with weight 0.5.
// Total: 3n + 1 operations.
//
// Within one run of consecutive forward steps (one checkpoint segment of the
// forward pass, or one recompute) the last odd half step of step k and the
// first of step k+1 run as one fused pass: an operation of such a run starts
// with the first odd half step only if it is the run's first (`first`), and
// ends with the fused pair if it is not the run's last (`fuse`).
//
// A reverse step i runs as one fused pass (the last odd adjoint half step of
// step i + 1 when i < n, the density pullback of step i, the first odd adjoint
// half step of step i), then the even adjoint half step of step i.  The final
// operation fuses the last odd adjoint half step of step 1 with the pullback
// at u_0.
//
// nck == 0 selects the taped sequence (the reference's own tape): every
// forward state is kept in K = n segment slots.  The forward runs once, out of
// place from the zero initial state in kBufState into slots 0..n-1, as one
// fused run; the reverse steps read their states there with no recompute, and
// the final pullback reads u_0 from kBufState.  Total: 2n + 1 operations.
#pragma once

#include "pnx_kernels.hpp"

namespace pnx {

enum MarchStage : int { kStFwd = 0, kStRecompute = 1, kStReverse = 2, kStFinal = 3, kStDone = 4 };
enum MarchKind : int { kOpNone = 0, kOpFwd = 1, kOpRecompute = 2, kOpReverse = 3, kOpFinal = 4 };
enum MarchBuf : int { kBufState = 0, kBufCkpt = 1, kBufSeg = 2 };

struct MarchPos {
  int stage, k, i, segi;
};

struct MarchOp {
  int kind;       // MarchKind
  int step;       // energy step (record index)
  int ckpt_copy;  // checkpoint slot that receives the state before a forward step, or -1
  int src_buf, src_idx;
  int dst_buf, dst_idx;
  int dst2_buf, dst2_idx;  // fused pair: output of the next step's first odd half step
  int first;      // forward: run the step's first odd half step (first step of a run)
  int fuse;       // forward: end with the fused pair instead of the last odd half step
  int reduce;     // forward finished after this step: form the intensities
};

constexpr MarchPos march_start() { return MarchPos{kStFwd, 1, 0, 0}; }

constexpr int march_ops(int n, int nck) { return nck == 0 ? 2 * n + 1 : 3 * n + 1; }

// Returns the operation at position s and advances s (n steps, checkpoints
// every K, nck of them; nck == 0: taped, K == n).
constexpr MarchOp march_next(MarchPos& s, int n, int K, int nck) {
  MarchOp o{kOpNone, 0, -1, kBufState, 0, kBufState, 0, kBufState, 0, 0, 0, 0};
  switch (s.stage) {
    case kStFwd: {
      const int k = s.k;
      o.kind = kOpFwd;
      o.step = k;
      if (nck == 0) {
        if (k > 1) {
          o.src_buf = kBufSeg;
          o.src_idx = k - 2;
        }
        o.dst_buf = kBufSeg;
        o.dst_idx = k - 1;
        o.dst2_buf = kBufSeg;
        o.dst2_idx = k;
        o.first = k == 1;
        o.fuse = k < n;
      } else {
        o.ckpt_copy = (k - 1) % K == 0 ? (k - 1) / K : -1;
        o.first = (k - 1) % K == 0;
        o.fuse = k < n && k % K != 0;
      }
      if (k == n) {
        o.reduce = 1;
        if (nck == 0) {
          s.stage = kStReverse;
          s.segi = 0;
          s.i = n;
        } else {
          s.stage = kStRecompute;
          s.segi = nck - 1;
          s.k = s.segi * K + 1;
        }
      } else {
        s.k = k + 1;
      }
      break;
    }
    case kStRecompute: {
      const int k0 = s.segi * K, k1 = n < k0 + K ? n : k0 + K, k = s.k;
      o.kind = kOpRecompute;
      o.step = k;
      if (k == k0 + 1) {
        o.src_buf = kBufCkpt;
        o.src_idx = s.segi;
      } else {
        o.src_buf = kBufSeg;
        o.src_idx = k - k0 - 2;
      }
      o.dst_buf = kBufSeg;
      o.dst_idx = k - k0 - 1;
      o.dst2_buf = kBufSeg;
      o.dst2_idx = k - k0;
      o.first = k == k0 + 1;
      o.fuse = k < k1;
      if (k == k1) {
        s.stage = kStReverse;
        s.i = k1;
      } else {
        s.k = k + 1;
      }
      break;
    }
    case kStReverse: {
      const int k0 = s.segi * K, i = s.i;
      o.kind = kOpReverse;
      o.step = i;
      o.src_buf = kBufSeg;
      o.src_idx = i - k0 - 1;
      if (i == k0 + 1) {
        if (s.segi > 0) {
          s.segi -= 1;
          s.stage = kStRecompute;
          s.k = s.segi * K + 1;
        } else {
          s.stage = kStFinal;
        }
      } else {
        s.i = i - 1;
      }
      break;
    }
    case kStFinal:
      o.kind = kOpFinal;
      o.step = 0;
      o.src_buf = nck == 0 ? kBufState : kBufCkpt;
      o.src_idx = 0;
      s.stage = kStDone;
      break;
    default:
      break;
  }
  return o;
}

// One request of the march kernel: the argument templates of its half steps
// and density pullback (the per-step pointers, records and step lengths are
// filled in by the kernel), its buffers, and the slice of the sequence to run.
struct MarchArgs {
  HalfStepArgs fwd;   // forward coefficients, no adjoint seed
  HalfStepArgs adj;   // adjoint coefficients and seed
  DensArgs dens;      // lam, rho_bar, gen, ibar
  double* state;      // forward state
  double* lam;        // adjoint state
  double* ckpt;       // nck checkpoint states
  double* seg;        // K recomputed states (taped: the n forward states)
  double* partial;    // [(n+1)][n_tiles][n_lines] intensity partials (slot 0: scratch)
  double* I_out;      // n_lines intensities
  const double* rec_f;
  const double* rec_a;
  const double* rec_d;
  long partial_step;
  int rec_len, sigma_off;
  int n, K, nck;
  int N;              // PN order
  MarchPos pos;       // first operation of this launch
  int n_ops;          // operations to run in this launch
};

// Widest march workgroup: two tiles of kBlock cells side by side.
constexpr int kMarchWide = 2 * kBlock;

// Runs args[0..n_requests) (1D, any PN orders), one workgroup of `threads`
// (kBlock or kMarchWide) per request.  The results do not depend on the width.
hipError_t launch_march(const MarchArgs* args, int n_requests, int threads, hipStream_t s);

}  // namespace pnx
