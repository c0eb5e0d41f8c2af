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
