// Host helpers shared by the solver and the server: the device layout of one
// solve batch and the conversions between the reference's array-of-structs
// class arrays and the padded moment-major device layout.
#pragma once

#include <string>
#include <vector>

#include "../../include/pnx.h"
#include "pnx_kernels.hpp"
#include "pnx_tables.hpp"

struct pnx_context {
  int device = 0;
  hipStream_t stream = nullptr;
  double* scratch = nullptr;  // this device's copy of the allocation-free forward's scratch
};

namespace pnx {

// Records msg as the thread's last error and returns code.
int set_error(int code, const std::string& msg);

struct Geometry {
  Layout L{};
  Tables T{};
  int class_n[8][3]{};  // reference class sizes per axis
  long rho_off[8]{};    // per class [e][V] offsets in the density buffer
  long rho_total = 0;
  long prof_stride = 0;  // doubles per beam in the profile buffer ([8][Y*Z])
  int rec_len = 0;
  int sigma_off = 0;
  double dE = 0.0;
  int K = 0;             // checkpoint interval
  int nck = 0;           // checkpoints
  long partial_step = 0;  // doubles per energy step in the partial buffer
  long partial_beam = 0;  // doubles per beam in the partial buffer
  bool active(int c) const { return class_active(c, L.dims) && T.n[c] > 0; }
};

// Offsets at which the kernels read each class's densities.  When the
// densities come from per-cell mass fractions on boxes that are the grid's
// cells (pnx_problem_set_cell_fractions, served cell-fraction requests),
// point (i, j, k) of every class lies in box (i, j, k): odd point i sits on the
// lower edge of box i, and the kernels update odd points 1..M-1 only.  Every
// class then has the eee densities at the same index, so all classes read the
// eee array, and a pass that updates several classes at a point (or the four
// wavefronts of a reverse block) reads those densities from HBM once.  The
// values are the same as in each class's own array, so results do not change.
inline void density_read_offsets(const Geometry& g, bool cell_boxes, long out[8]) {
  for (int c = 0; c < 8; ++c) out[c] = cell_boxes ? g.rho_off[0] : g.rho_off[c];
}

// Validates desc and fills g; returns PNX_OK or an error code with message in err.
int make_geometry(const pnx_desc& d, Geometry& g, std::string& err);

// Coordinate of point i of axis a (odd: the M + 1 points from lo - h/2 to
// hi + h/2; even: the M points from lo to hi), computed as numpy.linspace
// does; a singleton axis has the one point lo.
double axis_point(const pnx_desc& d, int a, int odd, int i);

// Reference AoS [point][vec] of class c -> padded device layout [vec][V] at dst.
void pack_class(const Geometry& g, int c, const double* src, int vec, double* dst);
void unpack_class(const Geometry& g, int c, const double* src, int vec, double* dst);
// [eee point][line] <-> [line][V]
void pack_lines(const Geometry& g, const double* src, double* dst);
void unpack_lines(const Geometry& g, const double* src, double* dst);
// beam profiles of one beam (ppos[c], y fastest over class c's (y, z) points) -> [8][Y*Z]
void pack_profiles(const Geometry& g, const double* const* ppos, double* dst);

std::vector<double> build_coefficients(int N, bool adjoint);

}  // namespace pnx
