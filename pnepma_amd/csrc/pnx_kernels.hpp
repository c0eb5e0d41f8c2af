// Kernel argument blocks and launch entry points of the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace pnx {

enum Mode : int { kForward = 0, kAdjoint = 1 };

// Device layout of one batch.  Every class array uses the same padded extents
// X x Y x Z (X = x pitch >= Mx+1, Y = My+1 or 1, Z = Mz+1 or 1) so that a
// point (i, j, k) has the same linear index p = i + X*(j + Y*k) in every
// class; class c holds n_c moment planes of V = X*Y*Z doubles.  The state of
// beam b starts at state + b*beam_stride, class c at offset cls_off[c].
struct Layout {
  int Mx, My, Mz;   // grid points (1 = singleton)
  int X, Y, Z;      // padded extents
  long V;           // doubles per moment plane
  int dims;         // number of non-singleton axes
  int n_e, n_lines, N;
  long cls_off[8];  // class offsets within one beam state (-1 = inactive)
  long beam_stride; // doubles per beam state
  double inv_h[3];
  double dvol;
  int n_cells;      // Mx*My*Mz unit cells
  int tiles_x, tiles_y;  // tiles per row and per column of one z-plane
  int n_tiles;      // tiles_x * tiles_y * Mz
  int cw_tiles_x;   // reverse passes with one class per wavefront: cw tiles per row
  int n_cw_items;   // and 64-cell sub-tiles per beam (4 per cw tile and z-plane; 0 in 1D)
  int rec_len;
};

constexpr int kBlock = 256;

// A tile is the 256 cells one workgroup updates: a row segment of 256 cells in
// 1D, a 32 x 8 patch of one z-plane in 2D and 3D (x fastest within the tile,
// tiles x-fastest over the grid).  A square-ish patch fetches fewer rows of
// the other parity twice than a run of two whole rows, and a wavefront still
// covers whole 128-byte lines (two row segments of 32 doubles).
constexpr int tile_x(int dims) { return dims == 1 ? kBlock : 32; }

// The 2D passes with one class per wavefront (the fused reverse pass) tile
// the plane with their own kCwTileX x kCwTileX tiles, each taken by four
// blocks of 64 cells (kCwTileX / 4 rows).  Against 32 x 8 tiles of two-row
// blocks this took the fused reverse pass at 2D P7 from 937 to 915 us, while
// the half steps measured faster with 32 x 8.
constexpr int kCwTileX = 16;

// The forward even half step leaves one intensity partial sum per wavefront
// (no block-wide reduction, no barrier): kBlock / 64 per tile, per line.
constexpr int kWavesPerBlock = kBlock / 64;

struct HalfStepArgs {
  Layout L;
  const double* from;     // state read for the other parity (beam 0)
  const double* to_src;   // state read for the updated parity (beam 0)
  double* to_dst;         // state written for the updated parity (beam 0)
  const double* rho;      // [c][e][V] (class offsets rho_off)
  long rho_off[8];
  const double* coef;     // advection + boundary values for this order and direction
  const double* rec;      // energy-point record of this step
  const double* prof;     // [b][c][Y*Z] beam profiles of x-odd classes (forward)
  long prof_stride;       // doubles per beam in prof
  const double* mms;      // [c][16] beam moments
  const double* W;        // [r][V] line weights at eee points
  const double* ibar;     // [b][n_lines] adjoint seed (adjoint eee source), may be null
  double* partial;        // [b][n_tiles * kWavesPerBlock][n_lines] intensity partial sums of this step (forward, even)
  long partial_stride;    // doubles per beam in partial
  double dt;
  int n_beams;
  // second odd half step of a fused pair (half_step2 kernels only): the last
  // odd half step of one energy step followed by the first of the next
  const double* rec2;     // energy record of the second update
  double* to_dst2;        // state written by the second update (beam 0); to_dst then
                          // receives the intermediate odd values, or is null
  double dt2;
};

struct DensArgs {
  Layout L;
  const double* u;        // taped forward state (beam 0)
  const double* lam;      // adjoint state (beam 0)
  const double* rho;
  long rho_off[8];
  const double* coef;     // forward coefficients
  const double* rec;      // dens record of this step
  const double* prof;
  long prof_stride;
  const double* mms;
  const double* W;
  const double* ibar;     // [b][n_lines]
  double* rho_bar;        // [b][c][e][V]
  long rb_stride;         // doubles per beam
  long rb_off[8];
  double* gen;            // [b][r][V] generation-weighted pullback
  long gen_stride;
  double dE;
  double int_fac;
  int n_beams;
  // 1: rho_bar is [b][e][V] per cell, at the cell's eee index, and a 2D
  // reverse block adds the four classes' gradients at a cell, in class order,
  // with one read and write of it (densities from per-cell fractions on boxes
  // that are the grid's cells, where the material pullback sums the classes
  // of a cell anyway; n_e <= kCellE).  0: [b][c][e][V] per class point at
  // rb_off[c].
  int rb_cell;
};

// Elements a per-cell density gradient (DensArgs::rb_cell) can carry.
constexpr int kCellE = 4;

// Fused reverse pass of one reference-adjoint step (see rev_body).
// Fused reverse pass of step i: [last odd adjoint half step of step i + 1]
// (rec3 non-null), density pullback of step i, [first odd adjoint half step of
// step i] (a1 != 0).  adj carries step i's adjoint record and dt = dE/2.
struct RevArgs {
  HalfStepArgs adj;
  DensArgs dens;
  const double* rec3;  // adjoint record of step i + 1, or null
  int a1;
};

// Exact adjoint of one forward energy step, one fused pass per half step.
// In the cell of every thread the pass first applies the transposed stencil of
// the half step that was back-propagated last (parity PB points, reading its
// w at the neighbouring points of the other parity), then back-propagates the
// half step of parity PA at this cell's parity-PA points: w = lam tau phi / S
// is stored for the next pass, lam is multiplied through, and the density
// gradient of the half step is accumulated.
struct DiscArgs {
  Layout L;
  const double* s_from;   // forward input state of the PA half step, other parity (beam 0)
  const double* s_own;    // forward input state of the PA half step, parity PA (beam 0)
  double* lam;            // adjoint state (beam 0)
  const double* w_in;     // w of the PB pass (beam 0)
  double* w_out;          // w of the PA half step (beam 0)
  const double* rho;
  long rho_off[8];
  const double* coef;     // forward coefficients
  const double* rec;      // forward record of the step
  const double* prof;
  long prof_stride;
  const double* mms;
  const double* W;        // [r][V]
  const double* ibar;     // [b][n_lines]
  const double* u_direct; // state after the step (eee zeroth moment) for the intensity terms, or null
  double* rho_bar;        // [b][c][e][V]
  long rb_stride;
  double* gen;            // [b][r][V]
  long gen_stride;
  double tau;
  double dE;
  int n_beams;
};

// Launchers (defined in pnx_kernels.hip); return hipSuccess or the launch error.
hipError_t launch_half_step(int N, int dims, int parity, int mode, const HalfStepArgs& a, hipStream_t s);
// Two consecutive odd half steps in one pass (the neighbours of the even
// classes do not change between them); see HalfStepArgs::rec2.
hipError_t launch_half_step2(int N, int dims, int mode, const HalfStepArgs& a, hipStream_t s);
// Forms over a device work list of (argument index, tile) items with
// *n_items entries: one block per item; `grid` is a multiple of 8 that is at
// least *n_items.
hipError_t launch_half_step_grouped(int N, int dims, int parity, int mode, const HalfStepArgs* args,
                                    const int2* items, const int* n_items, int grid, hipStream_t s);
// Phase launches of the server's grid lanes: one block per item of a device
// work list whose entries carry their kind (item.x = kind << kItemKindShift |
// argument index, item.y = tile; a 2D reverse item's is cw tile * 4 + the
// 64-cell sub-tile of its block, a 3D one's twice that plus the half of the
// eight classes it updates).
//   phase A: kItemFwdEven (half[arg], forward even half step) and kItemRev (rev[arg]);
//   pair:    kItemPair (half[arg], fused forward odd pair);
//   phase B: kItemFwdOdd (forward odd half step) and kItemAdjEven (adjoint even half step).
enum ItemKind : int { kItemFwdOdd = 0, kItemFwdEven = 1, kItemPair = 2, kItemRev = 3, kItemAdjEven = 4 };
constexpr int kItemKindShift = 28;
hipError_t launch_phase_a_grouped(int N, int dims, const HalfStepArgs* half, const RevArgs* rev, const int2* items,
                                  const int* n_items, int grid, hipStream_t s);
hipError_t launch_pair_grouped(int N, int dims, const HalfStepArgs* half, const int2* items, const int* n_items,
                               int grid, hipStream_t s);
hipError_t launch_phase_b_grouped(int N, int dims, const HalfStepArgs* half, const int2* items, const int* n_items,
                                  int grid, hipStream_t s);
// Reverse pass of step i: density pullback of step i and the first odd
// adjoint half step of step i in one sweep.
hipError_t launch_rev(int N, int dims, const RevArgs& a, hipStream_t s);
// pb, pa in {-1 (none), 0 (even), 1 (odd)}
hipError_t launch_disc(int N, int dims, int pb, int pa, const DiscArgs& a, hipStream_t s);
// Device material layer (pnx_material.hip): densities of every active class
// from per-cell mass fractions w[cell][e], and the pullback of the per-beam
// density gradient (plus an optional eee term, [e][V]) to grad[cell][e];
// with box_rb_in_grad, grad holds on entry a per-box density gradient
// (launch_absorption_pullback) whose pullback is added too; with rb_cell the
// density gradient is per cell ([b][e][V] at the cell's eee index,
// DensArgs::rb_cell) instead of per class point.
hipError_t launch_material_densities(const Layout& L, const int class_n[8][3], const long rho_off[8],
                                     const int ncls_of[8], const double* w, const double* rho_pure, double* rho,
                                     hipStream_t s);
hipError_t launch_material_pullback(const Layout& L, const int class_n[8][3], const long rho_off[8],
                                    const int ncls_of[8], const double* w, const double* rho_pure,
                                    const double* rho_bar, long rb_stride, int n_beams, const double* extra_eee,
                                    bool box_rb_in_grad, double* grad, hipStream_t s, bool rb_cell = false);
// Absorption factors on the device (pnx_material.hip).  Device pointers:
// even points and box edges per axis, detector, the plane x = beam_x0 that
// cuts the rays; inv_span[a] = n_boxes[a] / (last edge - first edge).
struct AbsorptionGeometry {
  const double* points[3];
  const double* edges[3];
  int n_boxes[3];
  double inv_span[3];
  double detector[3];
  double beam_x0;
};
hipError_t launch_absorption(const AbsorptionGeometry& a, const Layout& L, const double* w, const double* emb,
                             const double* rho_pure, const double* mac, const double* rho, const double* atom_mass,
                             const int* line_el, double* mu_scratch, double* ac, double* W, hipStream_t s);
hipError_t launch_atoms_bar(const AbsorptionGeometry& a, const Layout& L, const double* ac, const double* gen_b,
                            long gen_stride, int n_beams, const double* atom_mass, const int* line_el, double* extra,
                            hipStream_t s);
// Adds the density gradient of the absorption pullback into box_rb[box][e]
// (zeroed by the caller), for launch_material_pullback with box_rb_in_grad.
hipError_t launch_absorption_pullback(const AbsorptionGeometry& a, const Layout& L, const double* mac,
                                      const double* ac, const double* rho, const double* atom_mass,
                                      const int* line_el, const double* gen_b, long gen_stride, int n_beams,
                                      double* box_rb, hipStream_t s);
hipError_t launch_fill(double* p, long n, double v, hipStream_t s);
// dst[i] = src[i], i < n (device to device, as a kernel: no copy-engine round trip)
hipError_t launch_copy(double* dst, const double* src, long n, hipStream_t s);
// I[b][r] = sum_k dE*dvol*sigma_r(rec_k)*w_k * sum_t partial[b][k][t][r], k = 1..n_steps,
// t over the n_parts = n_tiles * kWavesPerBlock partial sums of a step.  Step 0
// of each beam's partial buffer is overwritten (scratch of the first stage).
hipError_t launch_reduce_intensity(double* partial, long step_stride, long beam_stride, int n_parts,
                                   const double* fwd_rec, int rec_len, int sigma_off, int n_steps, int n_lines,
                                   double dE, double dvol, int n_beams, double* I_out, hipStream_t s);
// out[i] = sum_b in[b*stride + i], i < n
hipError_t launch_beam_sum(const double* in, long stride, int n_beams, long n, double* out, hipStream_t s);

}  // namespace pnx
